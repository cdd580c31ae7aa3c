// libsesmd_fortran: the reference's routine names over its COMMON blocks.
//
// A Fortran program that links this library ahead of the reference's own
// objects gets, under each routine name of the per-timestep path, the same
// reads and writes of COMMON as the reference routine, computed on the GPU by
// the sesmd engine.  Each call
//   1. gathers the rank's molecules from COMMON storage (own 4^3 L4 boxes and
//      the W1/W2/S halo rings of 1/2/3 boxes, src/mybox.cmns:128-141) into an
//      engine context whose molecule classes and storage boxes are fixed to
//      the ones COMMON holds (sesmd_set_classes),
//   2. classifies sites ("central in box", the 3^3 blocks, the subtraction
//      masks) from the positions of the last build* call -- the reference's
//      lists are built every ifar steps and used until the next build,
//   3. runs the engine for that routine's part of the work, and
//   4. writes back what the reference routine writes: forces on the copies
//      the reference's loops reach, expansions of the rank's own boxes,
//      positions/velocities/accelerations of own molecules.
// Rank geometry comes from COMMON (own box centres rcen4, src/mybox.cmns:144)
// and the storage index maps of src/ibox.f90:1-72.
//
// This library is glue: every number it writes was computed on the GPU,
// except the Nose-Hoover scalar update of src/move_w2_a.f90:154-157 (one
// line over six all-reduced sums) and the bookkeeping of list formats.

#include "sesmd_fortran.h"
#include "ff_reference.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

// COMMON blocks (src/mybox.cmns:207-259), resolved against the Fortran side.
extern "C" {
extern char cmn01_[] __attribute__((weak));
extern char cmn02_[] __attribute__((weak));
extern char cmn03_[] __attribute__((weak));
extern char cmn04_[] __attribute__((weak));
extern char cmn05_[] __attribute__((weak));
extern char cmn06_[] __attribute__((weak));
extern char cmn07_[] __attribute__((weak));
extern char cmn08_[] __attribute__((weak));
extern char cmn09_[] __attribute__((weak));
extern char cmn99_[] __attribute__((weak));
extern char cmn10_[] __attribute__((weak));
extern char cmn11_[] __attribute__((weak));
extern char cmn12_[] __attribute__((weak));
extern char cmn13_[] __attribute__((weak));
extern char cmn14_[] __attribute__((weak));
extern char cmn15_[] __attribute__((weak));
extern char cmn16_[] __attribute__((weak));
extern char cmn17_[] __attribute__((weak));
extern char cmn18_[] __attribute__((weak));
extern char cmn19_[] __attribute__((weak));
extern char cmn20_[] __attribute__((weak));
extern char cmn21_[] __attribute__((weak));
extern char cmn22_[] __attribute__((weak));
extern char cmn23_[] __attribute__((weak));
extern char cmn24_[] __attribute__((weak));
extern char cmn25_[] __attribute__((weak));
extern char cmn26_[] __attribute__((weak));
extern char cmn27_[] __attribute__((weak));
extern char cmn28_[] __attribute__((weak));
extern char cmn29_[] __attribute__((weak));
extern char cmn30_[] __attribute__((weak));
extern char cmn100_[] __attribute__((weak));
extern char cmn31_[] __attribute__((weak));
extern char cmn48_[] __attribute__((weak));
extern char cmn49_[] __attribute__((weak));
extern char cmn50_[] __attribute__((weak));
extern char cmn51_[] __attribute__((weak));
}

namespace {

// PARAMETERs the COMMON layout is built from (src/mybox.cmns:49-99)
constexpr int LP = 25;           // (p+1)^2, p = 4
constexpr int P = 4;
constexpr int NB2 = 3;           // L2 boxes per axis (27 ranks)
constexpr int NB = 4 * NB2;      // L4 boxes per axis
constexpr int NB3 = NB * NB * NB;
constexpr int N3 = NB / 2;       // L3 boxes per axis
constexpr int KGRAPHITE = 1000;
const int GRID[3] = {6, 8, 10};  // storage grids of W1, W2, S
const int HALO[3] = {1, 2, 3};
const int NSITE[3] = {3, 3, 29};
const int NSLOT[3] = {300, 300, 580};  // nw3, nw3, ns29

std::string g_err;
int g_nerr = 0;
sesmd_allreduce_fn g_allreduce = nullptr;

bool fail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    g_nerr++;
    return false;
}

bool engine_ok(int rc, const char *what)
{
    if (rc == 0) return true;
    return fail("%s: %s", what, sesmd_last_error());
}

bool have_common()
{
    const void *blocks[] = {cmn04_, cmn05_, cmn06_, cmn07_, cmn08_, cmn09_, cmn99_, cmn10_,
                            cmn11_, cmn12_, cmn13_, cmn14_, cmn15_, cmn16_, cmn17_, cmn18_,
                            cmn19_, cmn20_, cmn21_, cmn22_, cmn23_, cmn24_, cmn25_, cmn26_,
                            cmn27_, cmn28_, cmn29_, cmn30_, cmn100_, cmn31_, cmn48_, cmn49_,
                            cmn50_, cmn51_};
    for (const void *b : blocks)
        if (!b) return fail("COMMON blocks not found: link the reference's Fortran objects");
    return true;
}

// x, y, z arrays of one COMMON block, Fortran a(slot, box) each
struct D3 {
    double *p;
    int nbox, nslot;
    double &operator()(int d, int box, int slot) const
    {
        return p[((size_t)d * nbox + box) * nslot + slot];
    }
};

D3 pos_of(int c)
{
    if (c == 0) return {(double *)cmn05_, 216, NSLOT[0]};
    if (c == 1) return {(double *)cmn06_, 512, NSLOT[1]};
    return {(double *)cmn12_, 1000, NSLOT[2]};
}
D3 frc_of(int c)
{
    if (c == 0) return {(double *)cmn09_, 216, NSLOT[0]};
    if (c == 1) return {(double *)cmn99_, 512, NSLOT[1]};
    return {(double *)cmn15_, 1000, NSLOT[2]};
}
D3 vel_of(int c)
{
    if (c == 0) return {(double *)cmn07_, 64, NSLOT[0]};
    if (c == 1) return {(double *)cmn08_, 64, NSLOT[1]};
    return {(double *)cmn13_, 64, NSLOT[2]};
}
D3 acc_of(int c)
{
    if (c == 0) return {(double *)cmn10_, 64, NSLOT[0]};
    if (c == 1) return {(double *)cmn11_, 64, NSLOT[1]};
    return {(double *)cmn14_, 64, NSLOT[2]};
}
int *count_of(int c)
{
    int *ict = (int *)cmn16_;  // ictw1(216), ictw2(512), icts(1000)
    return c == 0 ? ict : c == 1 ? ict + 216 : ict + 728;
}
double *nh_common() { return (double *)cmn50_; }  // tr pnh1 pnh2 eta sv1 sv2 w1v1 w1v2 w2v1 w2v2
double *expansions(int which)                     // complex (re, im) pairs
{
    double *o = (double *)cmn27_, *l = (double *)cmn28_;
    switch (which) {
    case 0: return o;                                // cO4(lp, 512)
    case 1: return o + 2 * LP * 512;                 // cO3(lp, 216)
    case 2: return o + 2 * LP * (512 + 216);         // cO2(lp)
    case 3: return o + 2 * LP * (512 + 216 + 1);     // cO2all(27 lp)
    case 4: return l;                                // cL4(lp, 64)
    case 5: return l + 2 * LP * 64;                  // cL3(lp, 8)
    default: return l + 2 * LP * (64 + 8);           // cL2(lp)
    }
}

// ------------------------------------------------------------------ force field

sesmd_forcefield g_ff;
bool g_ff_ok = false;

// Per-type numbers from the reference's COMMON tables (filled by its
// init_arrays, src/init_arrays.f90:266-840), on top of what its routines
// hard-code (angle/dihedral site lists, intramolecular pair rule, the
// water-carbon LJ constants, PARAMETERs).
bool load_forcefield()
{
    if (!have_common()) return false;
    sesmd_forcefield f;
    sesmd_reference_forcefield(&f);
    const double *t29 = (const double *)cmn29_;  // emw ems sgw sgs epw eps rqw rqs
    const double *emw = t29, *ems = t29 + 3, *rqw = t29 + 3 + 29 + 3 + 29 + 3 + 29,
                 *rqs = rqw + 3;
    for (int i = 0; i < 3; i++) f.q[i] = rqw[i], f.mass[i] = emw[i];
    for (int i = 0; i < 29; i++) f.q[3 + i] = rqs[i], f.mass[3 + i] = ems[i];
    const double *t100 = (const double *)cmn100_;  // sgww6 epww sgss6 epss sgws6 epws sgcs6 epcs
    const double *sgss6 = t100 + 18, *epss = sgss6 + 841, *sgws6 = epss + 841, *epws = sgws6 + 87,
                 *sgcs6 = epws + 87, *epcs = sgcs6 + 29;
    for (int n = 0; n < 29; n++) {
        // water O - surfactant n: (48 eps) s6, as the pair loops of src/evalw1s.f90 use them
        double s6 = sgws6[3 * n + 0], e = epws[3 * n + 0];
        f.lj_c[0 * 32 + 3 + n] = f.lj_c[(3 + n) * 32 + 0] = 48.0 * e * s6;
        f.lj_s6[0 * 32 + 3 + n] = f.lj_s6[(3 + n) * 32 + 0] = s6;
        for (int m = 0; m < 29; m++) {
            f.lj_c[(3 + n) * 32 + 3 + m] = 48.0 * epss[29 * m + n] * sgss6[29 * m + n];
            f.lj_s6[(3 + n) * 32 + 3 + m] = sgss6[29 * m + n];
        }
        f.gr_c[3 + n] = 48.0 * epcs[n] * sgcs6[n];
        f.gr_s6[3 + n] = sgcs6[n];
    }
    const double *t30 = (const double *)cmn30_;  // dsqw dsqs th rkth dva dvb dvc
    memcpy(f.dsqw, t30, sizeof f.dsqw);
    memcpy(f.dsqs, t30 + 3, sizeof f.dsqs);
    memcpy(f.th, t30 + 31, sizeof f.th);
    memcpy(f.rkth, t30 + 67, sizeof f.rkth);
    memcpy(f.dva, t30 + 103, sizeof f.dva);
    memcpy(f.dvb, t30 + 133, sizeof f.dvb);
    memcpy(f.dvc, t30 + 163, sizeof f.dvc);
    memcpy(f.rifact, cmn31_, sizeof f.rifact);
    const int *t48 = (const int *)cmn48_;  // iself(42) jself(42) ibond(28) jbond(28)
    memcpy(f.ibond, t48 + 84, sizeof f.ibond);
    memcpy(f.jbond, t48 + 112, sizeof f.jbond);
    if (f.rifact[0] != 1.0 || f.mass[0] <= 0.0)
        return fail("COMMON force-field tables are empty: call init_arrays first");
    g_ff = f;
    g_ff_ok = true;
    return true;
}

// ------------------------------------------------------------------ rank view

struct Rank {
    int o[3];    // global L4 index of own box 1 (src/ibox.f90:31-43: box 1 is the low corner)
    int l2[3];   // global L2 index of the rank
    double h;    // half width of the global cube (rhsx)
};

int floor_box(double x, double h, double side) { return (int)std::floor((x + h) / side); }

bool rank_geometry(Rank &R)
{
    const double *rhs = (const double *)cmn04_;
    const double *c4 = (const double *)cmn19_;  // rcen4x(64) rcen4y(64) rcen4z(64)
    double rs = g_ff.rs;
    R.h = rhs[0];
    if (std::fabs(R.h - 0.5 * NB * rs) > 1e-9 * R.h || rhs[1] != rhs[0] || rhs[2] != rhs[0])
        return fail("global cube half width %.17g does not match %d boxes of %.17g", R.h, NB, rs);
    for (int a = 0; a < 3; a++) {
        R.o[a] = floor_box(c4[64 * a], R.h, rs);
        if (R.o[a] < 0 || R.o[a] + 4 > NB || R.o[a] % 4)
            return fail("own box centres (rcen4) do not lie on the L4 grid");
        R.l2[a] = R.o[a] / 4;
    }
    return true;
}

int gidx(int x, int y, int z, int n) { return x + n * (y + n * z); }

// own box number 0..63 (src/ibox.f90: iaq-1 = x + 4y + 16z) -> storage index
int own_to_storage(int b, int c)
{
    int G = GRID[c], h = HALO[c];
    return (b % 4 + h) + G * ((b / 4 % 4 + h) + G * (b / 16 + h));
}

struct Mol {
    int c, sbox, slot0, b64;  // b64 < 0: halo copy
};

struct View {
    Rank R;
    std::vector<Mol> mol;
    std::vector<int> cls, anc;
    std::vector<double> r;  // (N, 3) current positions
    std::vector<int> sig;   // storage composition
    int nw = 0, ns = 0;
    int nsites() const { return 3 * nw + 29 * ns; }
};

// Molecules of COMMON storage: waters (W1 then W2), then surfactants, each in
// storage-box order.  own_only: the rank's own boxes only.
bool gather(View &V, bool own_only)
{
    if (!g_ff_ok && !load_forcefield()) return false;
    if (!rank_geometry(V.R)) return false;
    V.mol.clear(), V.cls.clear(), V.anc.clear(), V.r.clear(), V.sig.clear();
    V.nw = V.ns = 0;
    V.sig = {V.R.o[0], V.R.o[1], V.R.o[2], own_only};
    for (int c = 0; c < 3; c++) {
        int G = GRID[c], h = HALO[c];
        const int *cnt = count_of(c);
        D3 pos = pos_of(c);
        for (int s = 0; s < G * G * G; s++) {
            int n = cnt[s];
            if (n == 0) continue;
            if (n < 0 || n * NSITE[c] > NSLOT[c])
                return fail("storage box %d of class %d holds %d molecules", s + 1, c, n);
            int lx = s % G, ly = s / G % G, lz = s / (G * G);
            bool own = lx >= h && lx < h + 4 && ly >= h && ly < h + 4 && lz >= h && lz < h + 4;
            if (own_only && !own) continue;
            int gx = V.R.o[0] + lx - h, gy = V.R.o[1] + ly - h, gz = V.R.o[2] + lz - h;
            if (gx < 0 || gx >= NB || gy < 0 || gy >= NB || gz < 0 || gz >= NB)
                return fail("molecules stored in storage box %d (class %d) outside the domain",
                            s + 1, c);
            int b64 = own ? (lx - h) + 4 * (ly - h) + 16 * (lz - h) : -1;
            V.sig.push_back(c), V.sig.push_back(s), V.sig.push_back(n);
            for (int k = 0; k < n; k++) {
                Mol m{c, s, k * NSITE[c], b64};
                V.mol.push_back(m);
                V.cls.push_back(c);
                V.anc.push_back(gidx(gx, gy, gz, NB));
                for (int i = 0; i < NSITE[c]; i++)
                    for (int d = 0; d < 3; d++) V.r.push_back(pos(d, s, m.slot0 + i));
            }
            (c < 2 ? V.nw : V.ns) += n;
        }
    }
    return true;
}

// ------------------------------------------------------------------ engine contexts

struct Eng {
    sesmd_ctx *ctx = nullptr;
    int nw = -1, ns = -1;
};
Eng g_near, g_own;  // own + halo molecules; own molecules only

struct Frozen {
    bool valid = false;
    std::vector<int> sig;
    std::vector<double> r;
} g_frozen;

bool ensure(Eng &E, const View &V)
{
    if (!E.ctx || E.nw != V.nw || E.ns != V.ns) {
        if (E.ctx) sesmd_destroy(E.ctx);
        E.ctx = nullptr;
        E.nw = E.ns = -1;
        if (!engine_ok(sesmd_create(&E.ctx, V.nw, V.ns, NB2, P, &g_ff, SESMD_FP64), "sesmd_create"))
            return false;
        E.nw = V.nw, E.ns = V.ns;
    }
    const double *nh = nh_common();
    if (!engine_ok(sesmd_set_nh(E.ctx, nh[0], nh[1], nh[2]), "sesmd_set_nh")) return false;
    return engine_ok(sesmd_set_classes(E.ctx, V.cls.data(), V.anc.data()), "sesmd_set_classes");
}

// Engine over own + halo molecules, classified from the positions of the last
// build* call (current positions if none matches this storage).
bool prepare_near(View &V)
{
    if (!have_common() || !gather(V, false)) return false;
    if (V.mol.empty()) return true;
    if (!ensure(g_near, V)) return false;
    double eta = nh_common()[3];
    bool frozen = g_frozen.valid && g_frozen.sig == V.sig;
    const double *rc = frozen ? g_frozen.r.data() : V.r.data();
    if (!engine_ok(sesmd_set_state(g_near.ctx, rc, nullptr, nullptr, eta), "sesmd_set_state") ||
        !engine_ok(sesmd_rebuild(g_near.ctx), "sesmd_rebuild"))
        return false;
    if (frozen && !engine_ok(sesmd_set_state(g_near.ctx, V.r.data(), nullptr, nullptr, eta),
                             "sesmd_set_state"))
        return false;
    return true;
}

bool is_own(const Rank &R, int g)
{
    if (g < 0) return false;
    int x = g % NB, y = g / NB % NB, z = g / (NB * NB);
    return x >= R.o[0] && x < R.o[0] + 4 && y >= R.o[1] && y < R.o[1] + 4 && z >= R.o[2] &&
           z < R.o[2] + 4;
}

enum Rule { OWN_CENTRAL, OWN_MOLECULE };

// Add the engine's forces into COMMON for the sites this rank owns under `rule`.
bool add_forces(const View &V, Rule rule)
{
    int n = V.nsites();
    std::vector<double> F(3 * (size_t)n);
    std::vector<int> cbox;
    if (!engine_ok(sesmd_get_forces(g_near.ctx, F.data()), "sesmd_get_forces")) return false;
    if (rule == OWN_CENTRAL) {
        cbox.resize(n);
        if (!engine_ok(sesmd_get_classes(g_near.ctx, nullptr, nullptr, cbox.data(), nullptr, nullptr),
                       "sesmd_get_classes"))
            return false;
    }
    int e = 0;
    for (const Mol &m : V.mol) {
        D3 frc = frc_of(m.c);
        for (int i = 0; i < NSITE[m.c]; i++, e++) {
            bool take = rule == OWN_MOLECULE ? m.b64 >= 0 : is_own(V.R, cbox[e]);
            if (!take) continue;
            for (int d = 0; d < 3; d++) frc(d, m.sbox, m.slot0 + i) += F[3 * (size_t)e + d];
        }
    }
    return true;
}

bool run_forces(unsigned what, uint64_t mask, Rule rule)
{
    View V;
    if (!prepare_near(V)) return false;
    if (V.mol.empty()) return true;
    if (!engine_ok(sesmd_forces(g_near.ctx, what, mask), "sesmd_forces")) return false;
    return add_forces(V, rule);
}

// ------------------------------------------------------------------ near-field kernels

// (receiver kind, source kind) pairs of each reference kernel: receiver kinds
// W1@0 W2@0 W2@1 S@0 S@1 S@2, source kinds W1@0 W1@1 W2@0 W2@1 W2@2 S@0 S@1 S@2
// S@3 (class @ storage ring), bit rk*9+sk of the engine's near mask.
constexpr uint64_t bit(int rk, int sk) { return 1ull << (rk * 9 + sk); }

void near_kernel(uint64_t mask) { run_forces(SESMD_NEAR, mask, OWN_CENTRAL); }

// ------------------------------------------------------------------ lists

struct ListView {
    int *e, *ict;
    int cap;
    void clear(int i) { ict[i] = 0; }
    void add(int i, int k1, int jmol, int l1, int type)
    {
        int m = ict[i]++;
        if (m >= cap) return;
        int *b = e + (size_t)i * 4 * cap;
        b[0 * cap + m] = k1, b[1 * cap + m] = jmol, b[2 * cap + m] = l1, b[3 * cap + m] = type;
    }
};

// list pair (a, ict_a, b, ict_b) of one COMMON block, each a(cap, 4, 64)
void list_pair(char *blk, int cap, ListView &a, ListView &b)
{
    int *p = (int *)blk;
    size_t n = (size_t)cap * 4 * 64;
    a = {p, p + n, cap};
    b = {p + n + 64, p + 2 * n + 64, cap};
}

// Lists of storage ring `ring` of class c for every own box: sites central in
// the box into `central` (may be null), sites in its 3^3 block but not central
// into `near26` -- src/buildW2.f90:16-71 and its siblings, from the engine's
// classification of the frozen positions.
bool build_lists(int c, int ring, ListView *central, ListView *near26, bool central_too_in_26)
{
    View V;
    if (!prepare_near(V)) return false;
    std::vector<int> cbox(V.nsites());
    if (!V.mol.empty() &&
        !engine_ok(sesmd_get_classes(g_near.ctx, nullptr, nullptr, cbox.data(), nullptr, nullptr),
                   "sesmd_get_classes"))
        return false;
    for (int b = 0; b < 64; b++) {
        if (central) central->clear(b);
        near26->clear(b);
    }
    int G = GRID[c], h = HALO[c];
    int e = 0;
    for (const Mol &m : V.mol) {
        if (m.c != c) {
            e += NSITE[m.c];
            continue;
        }
        int sx = m.sbox % G - h, sy = m.sbox / G % G - h, sz = m.sbox / (G * G) - h;
        for (int i = 0; i < NSITE[c]; i++, e++) {
            int g = cbox[e];
            if (g < 0) continue;
            int x = g % NB - V.R.o[0], y = g / NB % NB - V.R.o[1], z = g / (NB * NB) - V.R.o[2];
            for (int bz = std::max(0, z - 1); bz <= std::min(3, z + 1); bz++)
                for (int by = std::max(0, y - 1); by <= std::min(3, y + 1); by++)
                    for (int bx = std::max(0, x - 1); bx <= std::min(3, x + 1); bx++) {
                        int ringd = std::max({std::abs(sx - bx), std::abs(sy - by), std::abs(sz - bz)});
                        if (ringd != ring) continue;
                        int b = bx + 4 * by + 16 * bz;
                        bool cen = bx == x && by == y && bz == z;
                        int slot = m.slot0 + i;
                        int jmol = m.slot0 / NSITE[c];
                        if (cen && central)
                            central->add(b, m.sbox + 1, jmol, slot + 1, i + 1);
                        else if (!cen || central_too_in_26)
                            near26->add(b, m.sbox + 1, jmol, slot + 1, i + 1);
                    }
        }
    }
    // an overflowing list keeps its first cap entries and advertises only
    // those (ict = cap), so a reference routine reading it stays in bounds
    int over = -1;
    for (int b = 0; b < 64; b++) {
        if (central && central->ict[b] > central->cap) {
            central->ict[b] = central->cap;
            if (over < 0) over = b;
        }
        if (near26->ict[b] > near26->cap) {
            near26->ict[b] = near26->cap;
            if (over < 0) over = b;
        }
    }
    if (over >= 0) return fail("list of own box %d exceeds its capacity (truncated to it)", over + 1);
    return true;
}


// ------------------------------------------------------------------ far field

using Table = std::vector<double>;  // complex (re, im) pairs, lp per box

// Global L2 / L3 / L4 indices of the rank's own boxes, from the centres in
// COMMON (rcen2 of src/mybox.cmns:142, rcen3 :145, rcen4 :144).
int own_l2(const Rank &R)
{
    const double *c = (const double *)cmn17_;
    double side = 4 * g_ff.rs;
    return gidx(floor_box(c[0], R.h, side), floor_box(c[1], R.h, side), floor_box(c[2], R.h, side),
                NB2);
}
int own_l3(const Rank &R, int kk)
{
    const double *c = (const double *)cmn20_;  // rcen3x(8) rcen3y(8) rcen3z(8)
    double side = 2 * g_ff.rs;
    return gidx(floor_box(c[kk], R.h, side), floor_box(c[8 + kk], R.h, side),
                floor_box(c[16 + kk], R.h, side), N3);
}
int own_l4(const Rank &R, int b)
{
    const double *c = (const double *)cmn19_;
    double side = g_ff.rs;
    return gidx(floor_box(c[b], R.h, side), floor_box(c[64 + b], R.h, side),
                floor_box(c[128 + b], R.h, side), NB);
}

void put(Table &t, int box, const double *src) { memcpy(&t[(size_t)box * 2 * LP], src, 2 * LP * sizeof(double)); }
void get(const Table &t, int box, double *dst) { memcpy(dst, &t[(size_t)box * 2 * LP], 2 * LP * sizeof(double)); }

// rank-local table with a halo (cO4: 8^3 around the own 4^3 L4 boxes, cO3:
// 6^3 around the own 2^3 L3 boxes; halo 2) -> global table
Table from_local(const double *loc, int grid, int halo, const int *origin, int nglob)
{
    Table t((size_t)nglob * nglob * nglob * 2 * LP, 0.0);
    for (int s = 0; s < grid * grid * grid; s++) {
        int x = origin[0] + s % grid - halo, y = origin[1] + s / grid % grid - halo,
            z = origin[2] + s / (grid * grid) - halo;
        if (x < 0 || x >= nglob || y < 0 || y >= nglob || z < 0 || z >= nglob) continue;
        put(t, gidx(x, y, z, nglob), loc + (size_t)s * 2 * LP);
    }
    return t;
}

// One engine far-field stage between global tables; null inputs are zero.
bool fmm_stage(int stage, const Table *O4, const Table *L4, const Table *O3, const Table *L3,
               const Table *O2, const Table *L2, Table *oO4, Table *oL4, Table *oO3, Table *oL3,
               Table *oO2, Table *oL2, View &V)
{
    if (!prepare_near(V)) return false;
    if (V.mol.empty() && stage == SESMD_P2M) return true;
    if (!g_near.ctx) {  // far-field stages need a context even without molecules
        if (!ensure(g_near, V)) return false;
    }
    size_t n4 = (size_t)NB3 * 2 * LP, n3 = (size_t)N3 * N3 * N3 * 2 * LP, n2 = (size_t)27 * 2 * LP;
    Table z4(n4, 0.0), z3(n3, 0.0), z2(n2, 0.0);
    auto in = [](const Table *t, Table &z) { return t ? t->data() : z.data(); };
    if (stage != SESMD_P2M &&
        !engine_ok(sesmd_set_expansions(g_near.ctx, in(O4, z4), in(L4, z4), in(O3, z3), in(L3, z3),
                                        in(O2, z2), in(L2, z2)),
                   "sesmd_set_expansions"))
        return false;
    if (!engine_ok(sesmd_fmm_stage(g_near.ctx, stage), "sesmd_fmm_stage")) return false;
    Table a(n4), b(n4), c(n3), d(n3), e(n2), f(n2);
    if (!engine_ok(sesmd_get_expansions(g_near.ctx, a.data(), b.data(), c.data(), d.data(), e.data(),
                                        f.data()),
                   "sesmd_get_expansions"))
        return false;
    if (oO4) *oO4 = std::move(a);
    if (oL4) *oL4 = std::move(b);
    if (oO3) *oO3 = std::move(c);
    if (oL3) *oL3 = std::move(d);
    if (oO2) *oO2 = std::move(e);
    if (oL2) *oL2 = std::move(f);
    return true;
}

// ------------------------------------------------------------------ integrator

// Engine over the rank's own molecules with their velocities, accelerations
// and (optionally) forces from COMMON.
bool prepare_own(View &V, bool with_forces)
{
    if (!have_common() || !gather(V, true)) return false;
    if (V.mol.empty()) return true;
    if (!ensure(g_own, V)) return false;
    size_t n = V.nsites();
    std::vector<double> v(3 * n), a(3 * n), f(with_forces ? 3 * n : 0);
    size_t e = 0;
    for (const Mol &m : V.mol) {
        D3 vel = vel_of(m.c), acc = acc_of(m.c), frc = frc_of(m.c);
        for (int i = 0; i < NSITE[m.c]; i++, e++)
            for (int d = 0; d < 3; d++) {
                v[3 * e + d] = vel(d, m.b64, m.slot0 + i);
                a[3 * e + d] = acc(d, m.b64, m.slot0 + i);
                if (with_forces) f[3 * e + d] = frc(d, m.sbox, m.slot0 + i);
            }
    }
    if (!engine_ok(sesmd_set_state(g_own.ctx, V.r.data(), v.data(), a.data(), nh_common()[3]),
                   "sesmd_set_state"))
        return false;
    return !with_forces || engine_ok(sesmd_set_forces(g_own.ctx, f.data()), "sesmd_set_forces");
}

// Write the engine's state of own molecules of the classes in cmask back to COMMON.
bool write_own(const View &V, unsigned cmask, bool r_out, bool v_out, bool a_out)
{
    size_t n = V.nsites();
    std::vector<double> r(3 * n), v(3 * n), a(3 * n);
    if (!engine_ok(sesmd_get_state(g_own.ctx, r.data(), v.data(), a.data(), nullptr), "sesmd_get_state"))
        return false;
    size_t e = 0;
    for (const Mol &m : V.mol) {
        bool mine = (cmask >> m.c) & 1u;
        D3 pos = pos_of(m.c), vel = vel_of(m.c), acc = acc_of(m.c);
        for (int i = 0; i < NSITE[m.c]; i++, e++) {
            if (!mine) continue;
            for (int d = 0; d < 3; d++) {
                if (r_out) pos(d, m.sbox, m.slot0 + i) = r[3 * e + d];
                if (v_out) vel(d, m.b64, m.slot0 + i) = v[3 * e + d];
                if (a_out) acc(d, m.b64, m.slot0 + i) = a[3 * e + d];
            }
        }
    }
    return true;
}

// mass-weighted thermostat sums of nh_water1 / nh_water2 / nh_surfactant
// (src/nh_water1.f90:38, src/nh_surfactant.f90:199) from the nine group sums
double weigh(const double *t, int species)
{
    const double *m = g_ff.mass;
    if (species == 0) return m[0] * t[0] + m[1] * t[1];
    if (species == 1) return m[0] * t[2] + m[1] * t[3];
    return m[3] * t[4] + m[4] * t[5] + m[5] * t[6] + m[3 + 17] * t[7] + m[3 + 18] * t[8];
}

void zero_forces(int c)
{
    D3 f = frc_of(c);
    memset(f.p, 0, sizeof(double) * 3 * (size_t)f.nbox * f.nslot);
}

// Part A of one species: predictor, SHAKE, sums before/after (src/move_w1_a.f90:27,140)
bool move_a(int c, double *before, double *after)
{
    View V;
    if (!prepare_own(V, false)) return false;
    double s18[18] = {0};
    if (!V.mol.empty()) {
        if (!engine_ok(sesmd_move_species(g_own.ctx, 0, 1u << c, s18), "sesmd_move_species") ||
            !write_own(V, 1u << c, true, true, false))
            return false;
    }
    *before = weigh(s18, c);
    *after = weigh(s18 + 9, c);
    zero_forces(c);
    return true;
}

bool move_b(int c)
{
    View V;
    if (!prepare_own(V, true)) return false;
    if (V.mol.empty()) return true;
    return engine_ok(sesmd_move_species(g_own.ctx, 1, 1u << c, nullptr), "sesmd_move_species") &&
           write_own(V, 1u << c, false, true, true);
}

void nh_sum(int c, double *out)
{
    View V;
    if (!out || !prepare_own(V, false)) return;
    double t[9] = {0};
    if (!V.mol.empty() && !engine_ok(sesmd_kinetic_sums(g_own.ctx, t), "sesmd_kinetic_sums")) return;
    *out = weigh(t, c);
}

}  // namespace

// ===================================================================== exports

extern "C" {

void sesmd_fortran_set_allreduce(sesmd_allreduce_fn fn) { g_allreduce = fn; }
const char *sesmd_fortran_last_error(void) { return g_err.c_str(); }
int32_t sesmd_fortran_errors(void) { return g_nerr; }

int sesmd_fortran_forcefield(sesmd_forcefield *out)
{
    if (!load_forcefield()) return 1;
    if (out) *out = g_ff;
    return 0;
}

void sesmd_fortran_release(void)
{
    for (Eng *E : {&g_near, &g_own}) {
        if (E->ctx) sesmd_destroy(E->ctx);
        E->ctx = nullptr;
        E->nw = E->ns = -1;
    }
    g_frozen = Frozen();
    g_ff_ok = false;
}

// ---- lists.  Each build call also fixes the classification positions.
static void freeze()
{
    View V;
    if (!have_common() || !gather(V, false)) return;
    g_frozen.valid = true;
    g_frozen.sig = V.sig;
    g_frozen.r = V.r;
}

void buildw2_(const int32_t *)
{
    freeze();
    ListView C, N;
    list_pair(cmn26_, 1000, C, N);  // W2Central, W226
    build_lists(1, 0, &C, &N, false);
}
void buildw2n_(const int32_t *)
{
    freeze();
    ListView C, N;
    list_pair(cmn23_, 4000, C, N);  // W2nCentral, W2n26
    build_lists(1, 1, &C, &N, false);
}
void buildw2nn_(const int32_t *)
{
    freeze();
    ListView A, B;
    list_pair(cmn21_, 1000, A, B);  // Snnn, W2nn
    build_lists(1, 2, nullptr, &B, true);
}
void builds_(const int32_t *)
{
    freeze();
    ListView C, N;
    list_pair(cmn25_, 1000, C, N);  // SCentral, S26
    build_lists(2, 0, &C, &N, false);
}
void buildsn_(const int32_t *)
{
    freeze();
    ListView C, N;
    list_pair(cmn24_, 1000, C, N);  // SnCentral, Sn26
    build_lists(2, 1, &C, &N, false);
}
void buildsnn_(const int32_t *)
{
    freeze();
    ListView C, N;
    list_pair(cmn22_, 1000, C, N);  // SnnCentral, Snn26
    build_lists(2, 2, &C, &N, false);
}
void buildsnnn_(const int32_t *)
{
    freeze();
    ListView A, B;
    list_pair(cmn21_, 1000, A, B);  // Snnn, W2nn
    build_lists(2, 3, nullptr, &A, true);
}

// isub(28, 29, 20, 64): per own surfactant site, the same-molecule sites
// outside its 3^3 block, zero-terminated (src/buildinternal.f90:11-63), from
// the engine's subtraction masks.
void buildinternal_(const int32_t *)
{
    freeze();
    View V;
    if (!prepare_near(V)) return;
    int *isub = (int *)cmn49_;
    memset(isub, 0, sizeof(int) * 28 * 29 * 20 * 64);
    if (V.mol.empty()) return;
    std::vector<unsigned> mask(V.nsites());
    if (!engine_ok(sesmd_get_classes(g_near.ctx, nullptr, nullptr, nullptr, nullptr, mask.data()),
                   "sesmd_get_classes"))
        return;
    int e = 3 * V.nw;
    for (const Mol &m : V.mol) {
        if (m.c != 2) continue;
        int k = m.slot0 / 29;
        for (int j = 0; j < 29; j++, e++) {
            if (m.b64 < 0 || k >= 20) continue;
            int *col = isub + (((size_t)m.b64 * 20 + k) * 29 + j) * 28;
            int q = 0;
            for (int s = 0; s < 29; s++)
                if ((mask[e] >> s) & 1u) col[q++] = s + 1;
        }
    }
}

// ---- the 39 pair kernels
#define NEAR(name, mask) \
    void name(const int32_t *) { near_kernel(mask); }
NEAR(evalw1w1_, bit(0, 0))
NEAR(evalw1w1n_, bit(0, 1))
NEAR(evalw1w2_, bit(0, 2) | bit(1, 0))
NEAR(evalw1w2n_, bit(0, 3) | bit(2, 0))
NEAR(evalw1w2nn_, bit(0, 4))
NEAR(evalw1s_, bit(0, 5) | bit(3, 0))
NEAR(evalw1sn_, bit(0, 6) | bit(4, 0))
NEAR(evalw1snn_, bit(0, 7) | bit(5, 0))
NEAR(evalw1snnn_, bit(0, 8))
NEAR(evalw2w1n_, bit(1, 1))
NEAR(evalw2w2_, bit(1, 2))
NEAR(evalw2w2n_, bit(1, 3) | bit(2, 2))
NEAR(evalw2w2nn_, bit(1, 4))
NEAR(evalw2s_, bit(1, 5) | bit(3, 2))
NEAR(evalw2sn_, bit(1, 6) | bit(4, 2))
NEAR(evalw2snn_, bit(1, 7) | bit(5, 2))
NEAR(evalw2snnn_, bit(1, 8))
NEAR(evalw2nw1n_, bit(2, 1))
NEAR(evalw2nw2n_, bit(2, 3))
NEAR(evalw2nw2nn_, bit(2, 4))
NEAR(evalw2ns_, bit(2, 5) | bit(3, 3))
NEAR(evalw2nsn_, bit(2, 6) | bit(4, 3))
NEAR(evalw2nsnn_, bit(2, 7) | bit(5, 3))
NEAR(evalw2nsnnn_, bit(2, 8))
NEAR(evalsw1n_, bit(3, 1))
NEAR(evalsw2nn_, bit(3, 4))
NEAR(evalss_, bit(3, 5))
NEAR(evalssn_, bit(3, 6) | bit(4, 5))
NEAR(evalssnn_, bit(3, 7) | bit(5, 5))
NEAR(evalssnnn_, bit(3, 8))
NEAR(evalsnw1n_, bit(4, 1))
NEAR(evalsnw2nn_, bit(4, 4))
NEAR(evalsnsn_, bit(4, 6))
NEAR(evalsnsnn_, bit(4, 7) | bit(5, 6))
NEAR(evalsnsnnn_, bit(4, 8))
NEAR(evalsnnw1n_, bit(5, 1))
NEAR(evalsnnw2nn_, bit(5, 4))
NEAR(evalsnnsnn_, bit(5, 7))
NEAR(evalsnnsnnn_, bit(5, 8))
#undef NEAR

// ---- bonded, intramolecular, external
void eval_subtract_fma_(const int32_t *) { run_forces(SESMD_SUBTRACT, 0, OWN_MOLECULE); }
void eval_surfactant_internal_(const int32_t *) { run_forces(SESMD_INTERNAL, 0, OWN_MOLECULE); }
void eval_valence_(const int32_t *) { run_forces(SESMD_VALENCE, 0, OWN_MOLECULE); }
void eval_dihedral_(const int32_t *) { run_forces(SESMD_DIHEDRAL, 0, OWN_MOLECULE); }
void eval_dihedral_berendsen_(const int32_t *) { run_forces(SESMD_DIHEDRAL_BE, 0, OWN_MOLECULE); }
void eval_wall_(void) { run_forces(SESMD_WALL, 0, OWN_MOLECULE); }

// Graphite: bottom-layer ranks only (src/eval_graphite.f90:13); the carbons
// are the union of the rank's per-column selections in /cmn51/, which the
// engine selects per column again with the same rule.
void eval_graphite_(const int32_t *, const int32_t *ibq)
{
    if (!ibq || *ibq != 0) return;
    View V;
    if (!prepare_near(V) || V.mol.empty()) return;
    const double *xgr = (const double *)cmn51_;
    const int *ict = (const int *)(xgr + 3 * KGRAPHITE * 16);
    std::vector<std::array<double, 3>> cs;
    for (int col = 0; col < 16; col++)
        for (int j = 0; j < std::min(ict[col], KGRAPHITE); j++)
            cs.push_back({xgr[col * KGRAPHITE + j], xgr[(16 + col) * KGRAPHITE + j],
                          xgr[(32 + col) * KGRAPHITE + j]});
    std::sort(cs.begin(), cs.end());
    cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
    std::vector<double> flat;
    for (auto &a : cs) flat.insert(flat.end(), a.begin(), a.end());
    if (!engine_ok(sesmd_set_graphite(g_near.ctx, (int)cs.size(), flat.data()), "sesmd_set_graphite"))
        return;
    if (!engine_ok(sesmd_forces(g_near.ctx, SESMD_GRAPHITE, 0), "sesmd_forces")) return;
    add_forces(V, OWN_CENTRAL);
}

// ---- far field
void computeonm4_(const int32_t *)
{
    View V;
    Table O4;
    if (!fmm_stage(SESMD_P2M, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &O4, nullptr,
                   nullptr, nullptr, nullptr, nullptr, V))
        return;
    double *cO4 = expansions(0);
    memset(cO4, 0, sizeof(double) * 2 * LP * 512);
    if (O4.empty()) return;
    for (int b = 0; b < 64; b++) get(O4, own_l4(V.R, b), cO4 + (size_t)own_to_storage(b, 1) * 2 * LP);
}

void computeonm3_(const int32_t *)
{
    View V;
    if (!have_common() || !gather(V, false)) return;
    Table in = from_local(expansions(0), 8, 2, V.R.o, NB), O3;
    if (!fmm_stage(SESMD_M2M_43, &in, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                   &O3, nullptr, nullptr, nullptr, V))
        return;
    double *cO3 = expansions(1);
    memset(cO3, 0, sizeof(double) * 2 * LP * 216);
    for (int kk = 0; kk < 8; kk++) {  // ib8to216 (src/ibox.f90:46-58): 6^3 grid, halo 2
        int s = (kk % 2 + 2) + 6 * ((kk / 2 % 2 + 2) + 6 * (kk / 4 + 2));
        get(O3, own_l3(V.R, kk), cO3 + (size_t)s * 2 * LP);
    }
}

void computeonm2_(const int32_t *)
{
    View V;
    if (!have_common() || !gather(V, false)) return;
    int o3[3] = {2 * V.R.l2[0], 2 * V.R.l2[1], 2 * V.R.l2[2]};
    Table in = from_local(expansions(1), 6, 2, o3, N3), O2;
    if (!fmm_stage(SESMD_M2M_32, nullptr, nullptr, &in, nullptr, nullptr, nullptr, nullptr, nullptr,
                   nullptr, nullptr, &O2, nullptr, V))
        return;
    get(O2, own_l2(V.R), expansions(2));
}

void computeljk2_(const int32_t *, const int32_t *)
{
    View V;
    if (!have_common() || !gather(V, false)) return;
    Table O2((size_t)27 * 2 * LP, 0.0), L2;
    const double *c2 = (const double *)cmn18_;  // rcen2x(27) rcen2y(27) rcen2z(27)
    double side = 4 * g_ff.rs;
    for (int l = 0; l < 27; l++)
        put(O2, gidx(floor_box(c2[l], V.R.h, side), floor_box(c2[27 + l], V.R.h, side),
                     floor_box(c2[54 + l], V.R.h, side), NB2),
            expansions(3) + (size_t)l * 2 * LP);
    if (!fmm_stage(SESMD_M2L_2, nullptr, nullptr, nullptr, nullptr, &O2, nullptr, nullptr, nullptr,
                   nullptr, nullptr, nullptr, &L2, V))
        return;
    get(L2, own_l2(V.R), expansions(6));
}

void shiftljk2_(const int32_t *)
{
    View V;
    if (!have_common() || !gather(V, false)) return;
    Table L2((size_t)27 * 2 * LP, 0.0), L3;
    put(L2, own_l2(V.R), expansions(6));
    if (!fmm_stage(SESMD_L2L_23, nullptr, nullptr, nullptr, nullptr, nullptr, &L2, nullptr, nullptr,
                   nullptr, &L3, nullptr, nullptr, V))
        return;
    for (int kk = 0; kk < 8; kk++) get(L3, own_l3(V.R, kk), expansions(5) + (size_t)kk * 2 * LP);
}

void computeljk3_(const int32_t *)
{
    View V;
    if (!have_common() || !gather(V, false)) return;
    int o3[3] = {2 * V.R.l2[0], 2 * V.R.l2[1], 2 * V.R.l2[2]};
    Table O3 = from_local(expansions(1), 6, 2, o3, N3);
    Table L3((size_t)N3 * N3 * N3 * 2 * LP, 0.0), out;
    for (int kk = 0; kk < 8; kk++) put(L3, own_l3(V.R, kk), expansions(5) + (size_t)kk * 2 * LP);
    if (!fmm_stage(SESMD_M2L_3, nullptr, nullptr, &O3, &L3, nullptr, nullptr, nullptr, nullptr,
                   nullptr, &out, nullptr, nullptr, V))
        return;
    for (int kk = 0; kk < 8; kk++) get(out, own_l3(V.R, kk), expansions(5) + (size_t)kk * 2 * LP);
}

void shiftljk3_(const int32_t *)
{
    View V;
    if (!have_common() || !gather(V, false)) return;
    Table L3((size_t)N3 * N3 * N3 * 2 * LP, 0.0), L4;
    for (int kk = 0; kk < 8; kk++) put(L3, own_l3(V.R, kk), expansions(5) + (size_t)kk * 2 * LP);
    if (!fmm_stage(SESMD_L2L_34, nullptr, nullptr, nullptr, &L3, nullptr, nullptr, nullptr, &L4,
                   nullptr, nullptr, nullptr, nullptr, V))
        return;
    for (int b = 0; b < 64; b++) get(L4, own_l4(V.R, b), expansions(4) + (size_t)b * 2 * LP);
}

void computeljk4_(const int32_t *)
{
    View V;
    if (!have_common() || !gather(V, false)) return;
    Table O4 = from_local(expansions(0), 8, 2, V.R.o, NB);
    Table L4((size_t)NB3 * 2 * LP, 0.0), out;
    for (int b = 0; b < 64; b++) put(L4, own_l4(V.R, b), expansions(4) + (size_t)b * 2 * LP);
    if (!fmm_stage(SESMD_M2L_4, &O4, &L4, nullptr, nullptr, nullptr, nullptr, nullptr, &out, nullptr,
                   nullptr, nullptr, nullptr, V))
        return;
    for (int b = 0; b < 64; b++) get(out, own_l4(V.R, b), expansions(4) + (size_t)b * 2 * LP);
}

// L2P of the rank's local expansions (src/evalLjk4.f90:21-284)
void evalljk4_(const int32_t *)
{
    View V;
    if (!prepare_near(V) || V.mol.empty()) return;
    Table L4((size_t)NB3 * 2 * LP, 0.0);
    for (int b = 0; b < 64; b++) put(L4, own_l4(V.R, b), expansions(4) + (size_t)b * 2 * LP);
    if (!engine_ok(sesmd_set_expansions(g_near.ctx, nullptr, L4.data(), nullptr, nullptr, nullptr,
                                        nullptr),
                   "sesmd_set_expansions") ||
        !engine_ok(sesmd_forces(g_near.ctx, SESMD_FAR, 0), "sesmd_forces"))
        return;
    add_forces(V, OWN_CENTRAL);
}

// ---- integrator and thermostat
void move_surfactant_a_(const int32_t *)
{
    double *nh = nh_common();
    if (!have_common()) return;
    move_a(2, &nh[4], &nh[5]);
}

void move_w1_a_(const int32_t *)
{
    double *nh = nh_common();
    if (!have_common()) return;
    move_a(0, &nh[6], &nh[7]);
}

void move_w2_a_(const int32_t *comm, const int32_t *)
{
    if (!have_common()) return;
    double *nh = nh_common();
    if (!move_a(1, &nh[8], &nh[9])) return;
    double tot[6] = {nh[4], nh[5], nh[6], nh[7], nh[8], nh[9]};  // sv1 sv2 w1v1 w1v2 w2v1 w2v2
    if (g_allreduce) g_allreduce(tot, 6, comm);
    double dt2 = g_ff.dt * g_ff.dt;
    double sum1 = (tot[0] + tot[2] + tot[4]) / dt2;
    double sum2 = (tot[1] + tot[3] + tot[5]) / dt2;
    nh[3] = nh[3] + nh[1] * (sum1 + nh[2] + sum2);
}

void move_surfactant_b_(const int32_t *) { move_b(2); }
void move_w1_b_(const int32_t *) { move_b(0); }
void move_w2_b_(const int32_t *) { move_b(1); }

void init_accel_(const int32_t *)
{
    View V;
    if (!prepare_own(V, true) || V.mol.empty()) return;
    if (engine_ok(sesmd_init_accel(g_own.ctx), "sesmd_init_accel")) write_own(V, 7u, false, false, true);
}

void nh_water1_(double *out) { nh_sum(0, out); }
void nh_water2_(double *out) { nh_sum(1, out); }
void nh_surfactant_(double *out) { nh_sum(2, out); }

// ---- box maps, src/ibox.f90:1-220 ----
// Plain host arithmetic in the reference's evaluation order (left to right,
// products and quotients rounded separately, int() truncating toward zero),
// over the geometry in COMMON /cmn01/ (ibdx3..), /cmn02/ (ibdx4..), /cmn03/
// (rsdx..), /cmn04/ (rhsx..) and /cmn17/ (rcen2xtmp..).  Without COMMON the
// functions that read it count an error and return 0 / .false.

static const double kRS = 10.0 / 3.166 + 2.0 / 3.166;  // rs = rc + delta (src/mybox.cmns:38-40)

static bool have_geometry(const char *who, bool l2_boxes)
{
    if (!cmn03_ || !cmn04_ || (l2_boxes && (!cmn01_ || !cmn02_)) || (!l2_boxes && !cmn17_)) {
        fail("%s: COMMON blocks not found: link the reference's Fortran objects", who);
        return false;
    }
    return true;
}

// 1 + int((x + rhs) * n / rsd) per axis, axes weighted 1, nx, nx ny
static int32_t global_box(const double *x, const double *y, const double *z, const int32_t *n)
{
    const double *rsd = (const double *)cmn03_, *rhs = (const double *)cmn04_;
    int32_t b = 1 + (int32_t)((*x + rhs[0]) * n[0] / rsd[0]);
    b = b + (int32_t)((*y + rhs[1]) * n[1] / rsd[1]) * n[0];
    b = b + (int32_t)((*z + rhs[2]) * n[2] / rsd[2]) * n[0] * n[1];
    return b;
}

// storage box of a point in a grid of g^3 boxes of side rs whose first box
// starts `off` boxes below centre (cx, cy, cz)
static int32_t grid_box(double x, double y, double z, double cx, double cy, double cz, double off, int g)
{
    double ax = x - cx + off * kRS, ay = y - cy + off * kRS, az = z - cz + off * kRS;
    return 1 + (int32_t)(ax / kRS) + (int32_t)(ay / kRS) * g + (int32_t)(az / kRS) * g * g;
}

// half-open test [c - w, c + w) on each axis
static int32_t within(const double *x, const double *y, const double *z, const double *cx,
                      const double *cy, const double *cz, double w)
{
    bool bx = *x < *cx + w && *x >= *cx - w;
    bool by = *y < *cy + w && *y >= *cy - w;
    bool bz = *z < *cz + w && *z >= *cz - w;
    return (bx && by && bz) ? 1 : 0;
}

int32_t ibox2_(const double *x, const double *y, const double *z)
{
    if (!have_geometry("ibox2", false)) return 0;
    const int32_t n[3] = {NB2, NB2, NB2};  // ibdx2 = ibdy2 = ibdz2 = 3 (PARAMETERs)
    return global_box(x, y, z, n);
}

int32_t ibox3_(const double *x, const double *y, const double *z)
{
    if (!have_geometry("ibox3", true)) return 0;
    return global_box(x, y, z, (const int32_t *)cmn01_);
}

int32_t ibox4_(const double *x, const double *y, const double *z)
{
    if (!have_geometry("ibox4", true)) return 0;
    return global_box(x, y, z, (const int32_t *)cmn02_);
}

int32_t ibox64_(const double *x, const double *y, const double *z)
{
    if (!have_geometry("ibox64", false)) return 0;
    const double *c = (const double *)cmn17_;  // rcen2xtmp, rcen2ytmp, rcen2ztmp
    return grid_box(*x, *y, *z, c[0], c[1], c[2], 2.0, 4);
}

int32_t ibox64mod_(const double *x, const double *y, const double *z, const double *cx,
                   const double *cy, const double *cz)
{
    return grid_box(*x, *y, *z, *cx, *cy, *cz, 2.0, 4);
}

int32_t ibox1000_(const double *x, const double *y, const double *z)
{
    if (!have_geometry("ibox1000", false)) return 0;
    const double *c = (const double *)cmn17_;
    return grid_box(*x, *y, *z, c[0], c[1], c[2], 5.0, 10);
}

// own box iaq = 1..64 (a 4^3 block, x fastest) -> index in a g^3 storage grid
// whose own block starts at corner (h, h, h); 8 -> 216 uses the 2^3 block of
// one L3 box
static int32_t own_to_grid(int32_t iaq, int32_t per, int32_t g, int32_t base)
{
    int32_t k = (iaq - 1) / (per * per);
    int32_t m = (iaq - per * per * k - 1) / per;
    int32_t r = iaq - per * per * k - per * m;
    return base + g * g * k + g * m + r;
}

int32_t ib64to216_(const int32_t *iaq) { return own_to_grid(*iaq, 4, 6, 43); }
int32_t ib64to512_(const int32_t *iaq) { return own_to_grid(*iaq, 4, 8, 146); }
int32_t ib64to1000_(const int32_t *iaq) { return own_to_grid(*iaq, 4, 10, 333); }
int32_t ib8to216_(const int32_t *iaq) { return own_to_grid(*iaq, 2, 6, 86); }

int32_t ib512to64_(const int32_t *iaq)
{
    int32_t k = (*iaq - 1) / 64;
    int32_t m = (*iaq - 64 * k - 1) / 8;
    return *iaq - 64 * k - 8 * m - 2 + 16 * (k - 2) + 4 * (m - 2);
}

int32_t isnearest_(const double *x, const double *y, const double *z, const double *cx,
                   const double *cy, const double *cz)
{
    return within(x, y, z, cx, cy, cz, 1.5 * kRS);
}

int32_t iscentral_(const double *x, const double *y, const double *z, const double *cx,
                   const double *cy, const double *cz)
{
    return within(x, y, z, cx, cy, cz, kRS / 2.0);
}

}  // extern "C"
