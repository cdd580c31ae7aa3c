// sesmd: MI355X (gfx950) per-timestep force and integration path of the
// surfactant-droplet MD code jhalverson/surfactant_enhanced_spreading.
//
// One translation unit: context + C ABI (include/sesmd.h), host-side
// precomputation of the FMM translation operators, and the device kernels.
//
// Device layout (all resident in HBM, sized once at create):
//   site state      SoA doubles rx ry rz | vx vy vz | ax ay az | fx fy fz (N each)
//   site statics    type[N], mol[N]; molecule first-site table first[nmol+1]
//   rebuild state   cls[nmol], anc[nmol] (storage box), cbox[N], rkind[N], isub[N]
//   box sort        perm[N] (sites sorted by cbox, stable), bstart[B+1]
//   near field      spos[N] (double4 x y z q, sorted), smeta[N] (int4), work list
//   far field       O4/L4 [B][lp], O3/L3, O2/L2 as double2; translation operators
//
// Kernels launch on the context's stream (and, inside a step, a second stream
// for the surfactant intramolecular forces); nothing synchronises the host
// inside a step, so sesmd_steps replays runs of regular steps from HIP graphs.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sesmd.h"

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;

static int fail(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return 1;
}

#define HC(x)                                                                         \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                         \
            return fail("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
    } while (0)

extern "C" const char *sesmd_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------- constants

constexpr int MAXP = 8;                       // highest supported expansion order
constexpr int MAXLP = (MAXP + 1) * (MAXP + 1);
constexpr int MAXLK = (MAXP + 1) * (MAXP + 2) / 2;  // rows with k >= 0
constexpr int NEAR_TILE = 64;                 // lanes of a near-field wavefront
constexpr int NEAR_CHUNK = 128;               // receivers per near-field work item (two per lane)
constexpr int NH_VALS = 18;                   // kinetic partial sums per block

constexpr int MAXNB = 48;  // L4 boxes per axis (nb2 <= 12; packed box coordinates hold 0..255)

struct Geo {
    int n, nw, ns, nmol, nb2, nb, nb3, n3, n33, n2, p, lp, lk;
    double rs, rsd, rhs, rc, rc2, xpi, dt, dtsq2;
    double cen4[MAXNB];  // L4 box centre per axis index (d_cen4_formula, evaluated on the device)
};

constexpr int INTRA_MAXC = 60;  // intramolecular contributions per surfactant site, at most 32 + 12 + 16

// per-site term lists of the surfactant topology (built on the host)
struct Topo {
    int npair[29];
    int pair[29][32];   // partner sites (0-based) of the internal LJ+Coulomb pairs
    int nang[29];
    int ang[29][12];    // angle index * 3 + role (0 = i, 1 = j, 2 = k)
    int ndih[29];
    int dih[29][16];    // dihedral index * 4 + role
    int pslot[29][32];  // term slot of pair[j][q]: 2 p for the pair's first site, 2 p + 1 its second
    // k_nh_intra: the contribution of term slot t (2 p / 2 p + 1 for pair p,
    // 2 ni + 3 a + role for angle a, 2 ni + 108 + 4 d + role for dihedral d)
    // goes to position dst[t] = j * INTRA_MAXC + q of the molecule's table:
    // site j's q-th contribution, pairs then angles then dihedrals
    int cnt[29];
    int dst[2 * 320 + 108 + 120];
};

// The reference surfactant's 28 bonds (src/init_arrays.f90:229-264), 0-based
// site pairs.  With the topology fixed at compile time the Gauss-Seidel loop
// unrolls and the molecule's predicted positions live in registers.
constexpr int kBI[28] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17,
                         17, 19, 20, 20, 20, 17, 24, 25, 25, 25};
constexpr int kBJ[28] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18,
                         19, 20, 21, 22, 23, 24, 25, 26, 27, 28};
struct SurfConst {
    double m[29];    // mass / m_O
    double rm[29];   // 1 / mass
    double dsq[28];  // squared bond lengths
    double k2[28];   // SHAKE: 2 (1/m_a + 1/m_b)
    double kr[28];   // RATTLE: -1 / (dt (1/m_a + 1/m_b) d^2)
    double wa[28], wb[28];  // RATTLE: dt / m_a, dt / m_b
};

struct sesmd_ctx {
    Geo g;
    int device;            // the HIP device the context was created on
    int precision;
    sesmd_forcefield ff;
    hipStream_t stream;    // every launch of the context, in order
    // a rebuild step's far field runs on far_stream beside the near field
    // (forked after the sort, joined before the site terms; default priority)
    hipStream_t far_stream;
    hipEvent_t ev_far_fork, ev_far_join;
    double *fs;            // intramolecular forces of a step, 29 ns sites (x, y, z blocks)
    // device state
    double *r, *v, *a, *f;  // each 3*N: x block, y block, z block
    int *type, *mol, *first;
    int *cls, *anc, *cbox, *rkind;
    unsigned *isub;
    int *key_in, *key_out, *val_in, *perm;
    void *cub_tmp;
    size_t cub_bytes;
    int *bstart, *nvalid;
    int *bs4;    // first sorted position of every (box, LJ class) key
    int *bflag;  // box holds a site not listed from every neighbour box (inside kcnt)
    int *kcnt;   // box sort: counts per key, then arrival counters per key
    int2 *work;                      // box-centred near-field items (masked evaluations)
    int *nwork;
    int max_work;
    bool boxlist;                    // `work` matches the last rebuild
    int2 *rwork;                     // row-centred near-field items (the full near field)
    int *nrwork;
    int4 *obox;                      // boxes holding sites, in box order: (box, first, end); then box -1
    int max_rwork;
    int2 *work_raw;                  // box-centred items in box order, before the cost sort
    unsigned *wkey_in, *wkey_out;    // 1 + estimated cost per item (0: padding)
    int *widx_in, *widx_out;
    double4 *spos;
    int4 *smeta;
    double *npart;   // near-field partial forces: 27 slots (row, source box) x (x, y, z) x N, sorted order
    int *iperm;      // site -> sorted position
    float4 *spos32;  // mixed precision: position relative to its box centre, charge
    float *lj32;     // mixed precision: lj_c (1024) then lj_s6 (1024) as float
    // far field
    double2 *O4, *L4, *O3, *L3, *O2, *L2;
    double2 *top_part;            // top-level M2L partials per (target, source)
    int *ne4, *ne3, *ne2;         // nonempty flags per box
    double2 *T4, *T3, *T2;        // M2L operators per offset
    double *AF4, *AF3;            // the same, real-expanded in MFMA operand order (k_m2l_gemm)
    unsigned short *AFB4, *AFB3;  // mixed precision: the operators as bf16 parts (k_m2l_gemm_bf3)
    // M2L lists and results, one set per level (L4: m2l[0], L3: m2l[1]), so
    // the two levels' products can run at the same time
    struct M2LBuf {
        unsigned short *ob;   // mixed precision: the source expansions as bf16 parts
        double *res;          // results, one row (M2L_RES values) per list entry
        int *cnt, *base;      // per (class, candidate) list: entries, first entry
        int *gbase, *bbase;   // per offset group: first entry, first block
        int2 *list;           // list entries (target within its class, source box)
        int *idx;             // per target box and candidate: its result row, -1 none
        int max_blocks;       // k_m2l_gemm grid (blocks of the densest possible level)
    } m2l[2];
    // the L4 products run on m2l_stream beside the L3 chain of the far field
    hipStream_t m2l_stream;
    hipEvent_t ev_m2l_fork, ev_m2l_join;
    double2 *M43, *M32, *S23, *S34; // M2M / L2L operators per child
    double *c4;                   // L4 centres (3 per box)
    double *lfac;                 // rfctr(j,|k|) table
    // graphite
    int ngr;
    double *gr;
    int *gstart, *gidx;
    // forcefield + topology on the device
    sesmd_forcefield *dff;
    Topo *dtopo;
    // integrator
    double *partials;  // [nblocks][NH_VALS]
    double *nh;        // [0] eta, [1..6] sums, [7] tr, [8] pnh1, [9] pnh2, [10] xi = integral of eta dt
    double *esum;      // energy component sums (sesmd_energies)
    int *counters;     // [0] shake failures
    int nblk_w, nblk_s;
    bool surf_fast;    // topology equals the reference's: unrolled SHAKE/RATTLE
    SurfConst sconst;
    // graphs of one regular step and of the ifar - 1 regular steps between
    // two rebuilds
    hipGraphExec_t graph, graph_run, graph_cycle;
    bool graph_ok;
    int graph_run_len;
    int graph_cycle_len;  // steps in graph_cycle: a rebuild step, then ifar - 1 regular steps
    bool built;
    bool classes_fixed;  // classes/storage boxes set by the caller (sesmd_set_classes)
};

// ------------------------------------------------------------ host math

namespace hostm {

// rLegendreP, src/rLegendreP.f90:3-48
static double legendre(int l, int m, double xx)
{
    if (l == -1 && m == 0) return 1.0;
    if (m > l) return 0.0;
    double pmm = 1.0;
    if (m > 0) {
        double somx2 = std::sqrt((1.0 - xx) * (1.0 + xx));
        double fact = 1.0;
        for (int i = 1; i <= m; i++) {
            pmm = -pmm * fact * somx2;
            fact = fact + 2.0;
        }
    }
    if (l == m) return pmm;
    double pmmp1 = xx * (2 * m + 1) * pmm;
    if (l == m + 1) return pmmp1;
    double pll = 0.0;
    for (int ll = m + 2; ll <= l; ll++) {
        pll = (xx * (2 * ll - 1) * pmmp1 - (ll + m - 1) * pmm) / (ll - m);
        pmm = pmmp1;
        pmmp1 = pll;
    }
    return pll;
}

typedef std::complex<double> C;

// Ylm, src/ifma.f90:12-26: the phase factor goes through a default-kind
// (single precision) cmplx(), so it is rounded to float.
static C ylm(const double *fact, int n, int m, double a, double b)
{
    int am = std::abs(m);
    double s = std::sqrt(fact[n - am] / fact[n + am]) * legendre(n, am, std::cos(a));
    float cr = (float)std::cos(m * b), ci = (float)std::sin(m * b);
    return C(s * (double)cr, s * (double)ci);
}

static double anm(const double *fact, int n, int m)
{
    return (double)((n & 1) ? -1 : 1) / std::sqrt(fact[n - m] * fact[n + m]);
}
static int sgnpow(int e) { return (e & 1) ? -1 : 1; }
static int ikm1(int a, int b) { return sgnpow((std::abs(a) - std::abs(b) - std::abs(a - b)) / 2); }
static int ikm2(int a, int b) { return sgnpow((std::abs(a - b) - std::abs(a) - std::abs(b)) / 2); }
static int ikm3(int a, int b) { return sgnpow((std::abs(b) - std::abs(b - a) - std::abs(a)) / 2); }
static int idx(int n, int m) { return n * (n + 1) + m; }
static int rowk(int j, int k) { return j * (j + 1) / 2 + k; }  // k >= 0 rows
static double ipow(double a, int b)
{
    int recip = b < 0;
    if (recip) b = -b;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

static void angles(double x, double y, double z, double &rho, double &al, double &be)
{
    rho = std::sqrt(x * x + y * y + z * z);
    al = std::acos(z / rho);
    be = std::atan2(y, x);
}

// M2L operator, src/computeLjk4.f90:134-160: L_jk += sum_nm T[jk][nm] O_nm
// for the source-minus-target vector (x, y, z); rows k >= 0 only.
static void m2l_op(const double *fact, int p, double x, double y, double z, C *T)
{
    int lp = (p + 1) * (p + 1);
    double rho, al, be;
    angles(x, y, z, rho, al, be);
    for (int j = 0; j <= p; j++)
        for (int k = 0; k <= j; k++) {
            double ajk = anm(fact, j, k);
            for (int n = 0; n <= p; n++)
                for (int m = -n; m <= n; m++) {
                    double rtmp = (double)ikm2(k, m) * ajk * anm(fact, n, m) /
                                  ((double)sgnpow(n) * anm(fact, j + n, m - k) * ipow(rho, j + n + 1));
                    T[rowk(j, k) * lp + idx(n, m)] = rtmp * ylm(fact, j + n, m - k, al, be);
                }
        }
}

// M2M operator, src/computeOnm3.f90:236-258: parent_jk += sum M[jk][.] child,
// vector = child centre - parent centre
static void m2m_op(const double *fact, int p, double x, double y, double z, C *M)
{
    int lp = (p + 1) * (p + 1);
    double rho, al, be;
    angles(x, y, z, rho, al, be);
    for (int i = 0; i < (p + 1) * (p + 2) / 2 * lp; i++) M[i] = 0.0;
    for (int j = 0; j <= p; j++)
        for (int k = 0; k <= j; k++) {
            double ajk = anm(fact, j, k);
            for (int n = 0; n <= j; n++)
                for (int m = -n; m <= n; m++) {
                    if (j - n - std::abs(k - m) < 0) continue;
                    double rtmp = (double)ikm1(k, m) * anm(fact, n, m) * anm(fact, j - n, k - m) *
                                  ipow(rho, n) / ajk;
                    M[rowk(j, k) * lp + idx(j - n, k - m)] = rtmp * ylm(fact, n, -m, al, be);
                }
        }
}

// L2L operator, src/shiftLjk3.f90:118-146: child_jk += sum S[jk][nm] parent_nm,
// vector = parent centre - child centre
static void l2l_op(const double *fact, int p, double x, double y, double z, C *S)
{
    int lp = (p + 1) * (p + 1);
    double rho, al, be;
    angles(x, y, z, rho, al, be);
    for (int i = 0; i < (p + 1) * (p + 2) / 2 * lp; i++) S[i] = 0.0;
    for (int j = 0; j <= p; j++)
        for (int k = 0; k <= j; k++) {
            double ajk = anm(fact, j, k);
            for (int n = 0; n <= p; n++)
                for (int m = -n; m <= n; m++) {
                    if (n - j - std::abs(m - k) < 0) continue;
                    double rtmp = (double)ikm3(k, m) * anm(fact, n - j, m - k) * ajk *
                                  ipow(rho, n - j) / ((double)sgnpow(n + j) * anm(fact, n, m));
                    S[rowk(j, k) * lp + idx(n, m)] = rtmp * ylm(fact, n - j, m - k, al, be);
                }
        }
}

}  // namespace hostm

// L2 centre along one axis, src/init_box_dimensions.f90:40-42
static double rcen2(const Geo &g, int ix)
{
    return g.rsd * (ix + 1) / (double)g.nb2 - g.rhs - g.rsd / (double)g.nb2 / 2.0;
}

// ------------------------------------------------------------ device helpers

__device__ __forceinline__ double rcube64(double x)
{
    // x^(-3/2) from v_rsq_f64's y = x^(-1/2) (1 + e), |e| < 2^-22, by one
    // Newton-type step on the cube itself: with u = 1 - x y^2 = -2e - e^2,
    // y^3 (1 + 3/2 u) = x^(-3/2) (1 - 15/2 e^2 + O(e^3)), ~4e-13 relative.
    // Written with the inline constants 1 and 1/2 only (a VOP3 operand
    // cannot be a literal here).
    double y = __builtin_amdgcn_rsq(x);
    double y2 = y * y, y3 = y2 * y;
    double u = fma(-x, y2, 1.0);
    return fma(y3, fma(u, 0.5, u), y3);
}

__device__ __forceinline__ int iabs_d(int a) { return a < 0 ? -a : a; }

__device__ __forceinline__ int src_kind(int cls, int ring)
{
    if (cls == 0) return ring <= 1 ? ring : -1;
    if (cls == 1) return ring <= 2 ? 2 + ring : -1;
    return ring <= 3 ? 5 + ring : -1;
}
__device__ __forceinline__ int recv_kind(int cls, int ring)
{
    if (cls == 0) return ring == 0 ? 0 : -1;
    if (cls == 1) return ring <= 1 ? 1 + ring : -1;
    return ring <= 2 ? 3 + ring : -1;
}
__device__ __forceinline__ int cheb(int b1, int b2, int nb)
{
    int dx = iabs_d(b1 % nb - b2 % nb), dy = iabs_d((b1 / nb) % nb - (b2 / nb) % nb),
        dz = iabs_d(b1 / (nb * nb) - b2 / (nb * nb));
    return max(dx, max(dy, dz));
}
__device__ __forceinline__ double d_rcen2(const Geo &g, int ix)
{
    return g.rsd * (ix + 1) / (double)g.nb2 - g.rhs - g.rsd / (double)g.nb2 / 2.0;
}
// L4 centre along one axis for global box coordinate gx (rcen4 of the rank)
__device__ __forceinline__ double d_cen4_formula(const Geo &g, int gx)
{
    int i = gx % 4 + 1;
    return 4.0 * g.rs * i / 4.0 - 2.0 * g.rs - 4.0 * g.rs / 8.0 + d_rcen2(g, gx / 4);
}
// the same values from the table the context filled at creation
__device__ __forceinline__ double d_cen4(const Geo &g, int gx) { return g.cen4[gx]; }

__global__ void k_cen4_table(Geo g, double *out)
{
    int gx = threadIdx.x;
    if (gx < g.nb) out[gx] = d_cen4_formula(g, gx);
}

// ------------------------------------------------------------ fill kernels

__global__ void k_fill_d(double *p, size_t n, double v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__device__ __forceinline__ void fill_i_blk(int blk,
                                             int *p,
                                             size_t n,
                                             int v)
{
    size_t i = (size_t)blk * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i(int *p, size_t n, int v)
{
    fill_i_blk(blockIdx.x, p, n, v);
}

// ------------------------------------------------------------ rebuild

// Molecule classes and storage boxes: read_config / delete_w1 logic,
// src/read_config.f90:44-175 (ibox2 + ibox64, int() truncating toward 0).
// A molecule whose filing site lies outside the cube [-rhs, rhs)^3 is filed
// in the nearest edge box, as the clamp below does, and counted in
// counters[2]: the reference's walls keep every site inside, so this flags a
// blown-up system.
__device__ __forceinline__ void classify_blk(int blk,
                                               const Geo &g,
                                               const double *__restrict__ r,
                                               const int *__restrict__ first,
                                               int *__restrict__ cls,
                                               int *__restrict__ anc,
                                               int *__restrict__ counters)
{
    int m = blk * blockDim.x + threadIdx.x;
    if (m >= g.nmol) return;
    const double *rx = r, *ry = r + g.n, *rz = r + 2 * g.n;
    int nb2 = g.nb2;
    auto ib2 = [&](int s) {
        return (int)((rx[s] + g.rhs) * nb2 / g.rsd) + (int)((ry[s] + g.rhs) * nb2 / g.rsd) * nb2 +
               (int)((rz[s] + g.rhs) * nb2 / g.rsd) * nb2 * nb2;
    };
    auto ib64 = [&](int s, int rk, int *l) {
        int qx = rk % nb2, qy = (rk / nb2) % nb2, qz = rk / (nb2 * nb2);
        l[0] = (int)((rx[s] - d_rcen2(g, qx) + 2.0 * g.rs) / g.rs);
        l[1] = (int)((ry[s] - d_rcen2(g, qy) + 2.0 * g.rs) / g.rs);
        l[2] = (int)((rz[s] - d_rcen2(g, qz) + 2.0 * g.rs) / g.rs);
        return l[0] + 4 * l[1] + 16 * l[2];
    };
    auto glob = [&](int rk, const int *l) {
        int qx = rk % nb2, qy = (rk / nb2) % nb2, qz = rk / (nb2 * nb2);
        int gx = min(max(4 * qx + l[0], 0), g.nb - 1), gy = min(max(4 * qy + l[1], 0), g.nb - 1),
            gz = min(max(4 * qz + l[2], 0), g.nb - 1);
        return gx + g.nb * (gy + g.nb * gz);
    };
    int f = first[m];
    {
        int fs = m < g.nw ? f : f + 13;
        bool out = false;
        for (int d = 0; d < 3; d++) {
            double x = r[(size_t)d * g.n + fs] + g.rhs;
            out = out || !(x >= 0.0 && x < g.rsd);
        }
        if (out) atomicAdd(&counters[2], 1);
    }
    if (m < g.nw) {
        int ibO = ib2(f), ibG = ib2(f + 1), ibH = ib2(f + 2);
        int lO[3], lG[3], lH[3];
        int bO = ib64(f, ibO, lO);
        bool w1 = false;
        if (ibO == ibG && ibO == ibH) w1 = (bO == ib64(f + 1, ibO, lG)) && (bO == ib64(f + 2, ibO, lH));
        cls[m] = w1 ? 0 : 1;
        anc[m] = glob(ibO, lO);
    } else {
        int s = f + 13;  // site 14 files the molecule
        int rk = ib2(s), l[3];
        ib64(s, rk, l);
        cls[m] = 2;
        anc[m] = glob(rk, l);
    }
}

__global__ void k_classify(Geo g, const double *__restrict__ r, const int *__restrict__ first,
                           int *__restrict__ cls, int *__restrict__ anc, int *__restrict__ counters)
{
    classify_blk(blockIdx.x, g, r, first, cls, anc, counters);
}

// Reserves one slot per active lane in counter[key] with one atomic per
// distinct key in the wavefront (keys of neighbouring lanes mostly agree, and
// contended global atomics serialise): the lanes are grouped by key first
// (register work only), then every group's first lane adds its group's size,
// all in one atomic instruction.  Returns the counter's value before the
// group's addition plus the lane's rank in its group.
__device__ __forceinline__ int wave_reserve(int *counter, int key)
{
    int lane = threadIdx.x & 63, leader_of = lane, rank = 0, size = 0;
    unsigned long long todo = __builtin_amdgcn_ballot_w64(true);
    while (todo) {
        int leader = __builtin_ctzll(todo);
        int kl = __builtin_amdgcn_readlane(key, leader);
        unsigned long long same = __builtin_amdgcn_ballot_w64(key == kl) & todo;
        if ((same >> lane) & 1ull) {
            leader_of = leader;
            rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(same >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((unsigned)same, 0u));
        }
        if (lane == leader) size = (int)__builtin_popcountll(same);
        todo &= ~same;
    }
    int base = 0;
    if (lane == leader_of) base = atomicAdd(&counter[key], size);
    return __shfl(base, leader_of, 64) + rank;
}

// Site boxes: W1 sites live in their molecule's box; W2 and surfactant sites in
// the box whose half-open cube [c - rs/2, c + rs/2) holds them (isCentral,
// src/ibox.f90:101-124, as the list builders use it).  Receiver kinds follow.
__device__ __forceinline__ void site_box_blk(int blk,
                                               const Geo &g,
                                               const double *__restrict__ r,
                                               const int *__restrict__ mol,
                                               const int *__restrict__ type,
                                               const int *__restrict__ cls,
                                               const int *__restrict__ anc,
                                               int *__restrict__ cbox,
                                               int *__restrict__ rkind,
                                               int *__restrict__ key,
                                               int *__restrict__ kcount)
{
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    int m = mol[i];
    int c = cls[m];
    int b;
    if (c == 0) {
        b = anc[m];
    } else {
        int gc[3];
        for (int a = 0; a < 3; a++) {
            double x = r[a * g.n + i];
            int q = (int)floor((x + g.rhs) / g.rs);
            q = min(max(q, 0), g.nb - 1);
            int best = q;
            for (int d = -1; d <= 1; d++) {
                int qq = q + d;
                if (qq < 0 || qq >= g.nb) continue;
                double cc = d_cen4(g, qq);
                if (x < cc + g.rs / 2.0 && x >= cc - g.rs / 2.0) {
                    best = qq;
                    break;
                }
            }
            gc[a] = best;
        }
        b = gc[0] + g.nb * (gc[1] + g.nb * gc[2]);
    }
    cbox[i] = b;
    rkind[i] = recv_kind(c, cheb(anc[m], b, g.nb));
    // sort key: box, then LJ class (water O, surfactant site, water H), so the
    // near-field loops need no per-source type test
    int t = type[i];
    int k = 4 * b + (t == 0 ? 0 : (t >= 3 ? 1 : 2));
    key[i] = k;
    wave_reserve(kcount, k);  // the box sort's key counts (cleared before this launch)
}

__global__ void k_site_box(Geo g, const double *__restrict__ r, const int *__restrict__ mol,
                           const int *__restrict__ type, const int *__restrict__ cls,
                           const int *__restrict__ anc, int *__restrict__ cbox,
                           int *__restrict__ rkind, int *__restrict__ key, int *__restrict__ kcount)
{
    site_box_blk(blockIdx.x, g, r, mol, type, cls, anc, cbox, rkind, key, kcount);
}

// buildinternal, src/buildinternal.f90:11-63: the same-molecule partners of a
// surfactant site that lie outside the 3^3 block around the site's box (box by
// ibox1000 of the molecule's rank, centre by rcen1000, test by isnearest).
// One 64-lane block per molecule: lane j < 29 tests site j against the
// molecule's other sites, read from LDS.
__device__ __forceinline__ void isub_blk(int blk,
                                           const Geo &g,
                                           const double *__restrict__ r,
                                           const int *__restrict__ anc,
                                           unsigned *__restrict__ isub)
{
    __shared__ double xs[3][29];
    int s = blk, j = threadIdx.x;
    int m = g.nw + s;
    int base = 3 * g.nw + 29 * s;
    if (j < 29)
        for (int d = 0; d < 3; d++) xs[d][j] = r[d * g.n + base + j];
    __syncthreads();
    if (j >= 29) return;
    int a = anc[m];
    int q[3] = {(a % g.nb) / 4, ((a / g.nb) % g.nb) / 4, (a / (g.nb * g.nb)) / 4};
    double cen[3];
    for (int d = 0; d < 3; d++) {
        double c2 = d_rcen2(g, q[d]);
        double xm = xs[d][j];
        int li = (int)((xm - c2 + 5.0 * g.rs) / g.rs) + 1;
        cen[d] = 10.0 * g.rs * li / 10.0 - 10.0 * g.rs / 2.0 - 10.0 * g.rs / 20.0 + c2;
    }
    double rs15 = 1.5 * g.rs;
    unsigned mask = 0;
    for (int k = 0; k < 29; k++) {
        bool near = true;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double xn = xs[d][k];
            near = near && (xn < cen[d] + rs15 && xn >= cen[d] - rs15);
        }
        if (k != j && !near) mask |= 1u << k;
    }
    isub[base + j] = mask;
}

__global__ __launch_bounds__(64) void k_isub(Geo g, const double *__restrict__ r, const int *__restrict__ anc,
                                             unsigned *__restrict__ isub)
{
    isub_blk(blockIdx.x, g, r, anc, isub);
}


// The box sort (stable by site index within each key 4 box + LJ class), a
// counting sort in four launches with no library call and no memset, so a
// rebuild can be captured in a graph: count per key, scan (key starts bs4,
// box starts bstart), scatter (positions within a key in arrival order), then
// order each key's sites by index (a rank sort per key: every key holds at
// most a few hundred sites).
// One block: exclusive scan of the key counts (bs4, K + 1 entries) and the box
// starts (bstart, every fourth).  Thread t scans `per` consecutive keys; the
// per-thread totals are scanned with wavefront prefix sums.
__device__ __forceinline__ void key_scan_block(const Geo &g, const int *__restrict__ cnt, int *__restrict__ bs4,
                                               int *__restrict__ bstart)
{
    __shared__ int wsum[16];
    int K = 4 * g.nb3, t = threadIdx.x, lane = t & 63, w = t >> 6;
    int per = (K + (int)blockDim.x - 1) / (int)blockDim.x;
    int a = min(t * per, K), b = min(a + per, K), sum = 0;
    for (int i = a; i < b; i++) sum += cnt[i];
    int incl = sum;  // inclusive prefix within the wavefront
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int base = 0, total = 0;
    for (int j = 0; j < (int)(blockDim.x >> 6); j++) {
        base += j < w ? wsum[j] : 0;
        total += wsum[j];
    }
    int run = base + incl - sum;
    for (int i = a; i < b; i++) {
        bs4[i] = run;
        if ((i & 3) == 0) bstart[i >> 2] = run;
        run += cnt[i];
    }
    if (t == 0) {
        bs4[K] = total;
        bstart[g.nb3] = total;
    }
}

// Positions within a key in arrival order (k_key_order orders them).
__device__ __forceinline__ void key_scatter_blk(int blk,
                                                  const Geo &g,
                                                  const int *__restrict__ key,
                                                  const int *__restrict__ bs4,
                                                  int *__restrict__ fill,
                                                  int *__restrict__ tmp)
{
    int i = blk * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    int k = key[i];
    tmp[bs4[k] + wave_reserve(fill, k)] = i;
}

__global__ void k_key_scatter(Geo g, const int *__restrict__ key, const int *__restrict__ bs4,
                              int *__restrict__ fill, int *__restrict__ tmp)
{
    key_scatter_blk(blockIdx.x, g, key, bs4, fill, tmp);
}

// Each key's sites in index order, by a rank sort (distinct indices; a key
// holds at most a few hundred sites), one wavefront per key 4 box + class; also
// the inverse map site -> sorted position.
__device__ __forceinline__ void key_order_blk(int blk,
                                                const Geo &g,
                                                const int *__restrict__ bs4,
                                                const int *__restrict__ tmp,
                                                int *__restrict__ perm,
                                                int *__restrict__ key_out,
                                                int *__restrict__ iperm)
{
    int kk = blk, s0 = bs4[kk], len = bs4[kk + 1] - s0;
    const int *seg = tmp + s0;
    for (int j = threadIdx.x; j < len; j += 64) {
        int v = seg[j], rank = 0, t = 0;
        for (; t + 8 <= len; t += 8) {  // eight loads in flight per round
            int w[8];
#pragma unroll
            for (int u = 0; u < 8; u++) w[u] = seg[t + u];
#pragma unroll
            for (int u = 0; u < 8; u++) rank += w[u] < v ? 1 : 0;
        }
        for (; t < len; t++) rank += seg[t] < v ? 1 : 0;
        perm[s0 + rank] = v;
        key_out[s0 + rank] = kk;
        iperm[v] = s0 + rank;
    }
}

__global__ __launch_bounds__(64) void k_key_order(Geo g, const int *__restrict__ bs4, const int *__restrict__ tmp,
                                                  int *__restrict__ perm, int *__restrict__ key_out,
                                                  int *__restrict__ iperm)
{
    key_order_blk(blockIdx.x, g, bs4, tmp, perm, key_out, iperm);
}

// Near-field work items (one block): per box, chunks of NEAR_CHUNK receivers, in box order,
// each with a cost estimate (receivers x sites in the 3^3 block) as a sort key;
// the items are then ordered by decreasing cost (k_work_gather), so the longest
// workgroups start first and the short ones fill in behind them.
__global__ void k_worklist(Geo g, const int *__restrict__ bstart, int2 *__restrict__ work,
                           unsigned *__restrict__ wkey, int *__restrict__ widx,
                           int *__restrict__ nwork, int max_work)
{
    __shared__ int s_cnt[1024];
    __shared__ int s_base;
    if (threadIdx.x == 0) s_base = 0;
    __syncthreads();
    for (int b0 = 0; b0 < g.nb3; b0 += blockDim.x) {
        int b = b0 + threadIdx.x;
        int nch = 0;
        if (b < g.nb3) nch = (bstart[b + 1] - bstart[b] + NEAR_CHUNK - 1) / NEAR_CHUNK;
        s_cnt[threadIdx.x] = nch;
        __syncthreads();
        // inclusive scan (Hillis-Steele)
        for (int off = 1; off < (int)blockDim.x; off <<= 1) {
            int v = threadIdx.x >= (unsigned)off ? s_cnt[threadIdx.x - off] : 0;
            __syncthreads();
            s_cnt[threadIdx.x] += v;
            __syncthreads();
        }
        int excl = s_base + s_cnt[threadIdx.x] - nch;
        int src = 0;
        if (nch) {
            int nb = g.nb, bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
            for (int z = max(bz - 1, 0); z <= min(bz + 1, nb - 1); z++)
                for (int y = max(by - 1, 0); y <= min(by + 1, nb - 1); y++) {
                    int r0 = max(bx - 1, 0) + nb * (y + nb * z), r1 = min(bx + 1, nb - 1) + nb * (y + nb * z);
                    src += bstart[r1 + 1] - bstart[r0];
                }
        }
        for (int c = 0; c < nch; c++)
            if (excl + c < max_work) {
                int k0 = bstart[b] + c * NEAR_CHUNK;
                int nr = min(NEAR_CHUNK, bstart[b + 1] - k0);
                work[excl + c] = make_int2(b, k0);
                wkey[excl + c] = 1u + (unsigned)min(nr * src, 1 << 30);
                widx[excl + c] = excl + c;
            }
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) s_base += s_cnt[threadIdx.x];
        __syncthreads();
    }
    for (int i = min(s_base, max_work) + threadIdx.x; i < max_work; i += blockDim.x) {
        wkey[i] = 0u;
        widx[i] = i;
    }
    if (threadIdx.x == 0) *nwork = min(s_base, max_work);
}

__global__ void k_work_gather(const int2 *__restrict__ raw, const int *__restrict__ order,
                              const int *__restrict__ nwork, int2 *__restrict__ work)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < *nwork) work[i] = raw[order[i]];
}

// static metadata of the sorted sites: (site, type | cls << 8 | split << 10 |
// (receiver kind + 1) << 11, mol, packed
// storage box) with the storage box as x | y << 8 | z << 16 and bit 24 set when
// the site is listed by the reference from every box of the 3^3 block around
// its own box (its storage ring stays within the class's list rings: W1 <= 1,
// W2 <= 2, S <= 3), which spares the near-field loop the per-box ring test.
__device__ __forceinline__ void gather_meta_blk(int blk,
                                                  const Geo &g,
                                                  const int *__restrict__ perm,
                                                  const int *__restrict__ type,
                                                  const int *__restrict__ mol,
                                                  const int *__restrict__ first,
                                                  const int *__restrict__ cls,
                                                  const int *__restrict__ anc,
                                                  const int *__restrict__ cbox,
                                                  const int *__restrict__ rkind,
                                                  int4 *__restrict__ smeta,
                                                  int *__restrict__ bflag)
{
    int k = blk * blockDim.x + threadIdx.x;
    if (k >= g.n) return;
    int i = perm[k];
    int m = mol[i];
    int a = anc[m], c = cls[m];
    int nb = g.nb;
    int ax = a % nb, ay = (a / nb) % nb, az = a / (nb * nb);
    int maxring = c == 0 ? 1 : (c == 1 ? 2 : 3);
    int always = cheb(a, cbox[i], nb) + 1 <= maxring ? 1 : 0;
    // split: the molecule has sites in more than one box, so a receiver of
    // another box can belong to it (a W1 molecule never does)
    int split = 0;
    if (c != 0)
        for (int s = first[m]; s < first[m + 1]; s++) split |= cbox[s] != cbox[i];
    smeta[k] = make_int4(i, type[i] | (c << 8) | (split << 10) | ((rkind[i] + 1) << 11), m,
                         ax | (ay << 8) | (az << 16) | (always << 24));
    if (!always) atomicOr(&bflag[cbox[i]], 1);
}

__global__ void k_gather_meta(Geo g, const int *__restrict__ perm, const int *__restrict__ type,
                              const int *__restrict__ mol, const int *__restrict__ first,
                              const int *__restrict__ cls, const int *__restrict__ anc,
                              const int *__restrict__ cbox, const int *__restrict__ rkind,
                              int4 *__restrict__ smeta, int *__restrict__ bflag)
{
    gather_meta_blk(blockIdx.x, g, perm, type, mol, first, cls, anc, cbox, rkind, smeta, bflag);
}

// sorted positions + charge, every step
__device__ __forceinline__ void gather_pos_blk(int blk,
                                                 const Geo &g,
                                                 const int *__restrict__ perm,
                                                 const double *__restrict__ r,
                                                 const int *__restrict__ type,
                                                 const sesmd_forcefield *__restrict__ ff,
                                                 double4 *__restrict__ spos)
{
    int k = blk * blockDim.x + threadIdx.x;
    if (k >= g.n) return;
    int i = perm[k];
    spos[k] = make_double4(r[i], r[g.n + i], r[2 * g.n + i], ff->q[type[i]]);
}

__global__ void k_gather_pos(Geo g, const int *__restrict__ perm, const double *__restrict__ r,
                             const int *__restrict__ type, const sesmd_forcefield *__restrict__ ff,
                             double4 *__restrict__ spos)
{
    gather_pos_blk(blockIdx.x, g, perm, r, type, ff, spos);
}

// Mixed precision: each sorted site's position relative to the centre of
// its own L4 box, in FP32 (|x - c| < rs: an absolute rounding of ~1e-7 sigma),
// and its charge.
__device__ __forceinline__ void gather_pos32_blk(int blk,
                                                   const Geo &g,
                                                   const int *__restrict__ perm,
                                                   const int *__restrict__ key,
                                                   const double *__restrict__ r,
                                                   const int *__restrict__ type,
                                                   const sesmd_forcefield *__restrict__ ff,
                                                   float4 *__restrict__ spos32)
{
    int k = blk * blockDim.x + threadIdx.x;
    if (k >= g.n) return;
    int i = perm[k], b = key[k] >> 2, nb = g.nb;
    spos32[k] = make_float4((float)(r[i] - d_cen4(g, b % nb)), (float)(r[g.n + i] - d_cen4(g, (b / nb) % nb)),
                            (float)(r[2 * g.n + i] - d_cen4(g, b / (nb * nb))), (float)ff->q[type[i]]);
}

__global__ void k_gather_pos32(Geo g, const int *__restrict__ perm, const int *__restrict__ key,
                               const double *__restrict__ r, const int *__restrict__ type,
                               const sesmd_forcefield *__restrict__ ff, float4 *__restrict__ spos32)
{
    gather_pos32_blk(blockIdx.x, g, perm, key, r, type, ff, spos32);
}

// ------------------------------------------------------------ near field

// The near field of the 39 eval[wS]* kernels (e.g. src/evalw1w1.f90:28-245,
// src/evalSS.f90:13-125): each receiver (a site central in box b at the last
// rebuild) gets, from every site of another molecule in the 3^3 block around b,
// Coulomb xpi q_i q_j r/r^3 plus LJ c r^-6 (s6 r^-6 - 1/2) r/r^2 for r <= rc.
// Each lane owns one receiver, so forces are written without atomics.  MASKED
// keeps only the (receiver kind, source kind) pairs of `mask` (one reference
// kernel or a subset of them).
constexpr int NEAR_ROWS = 9;  // (dz, dy) rows of the 3^3 neighbourhood, one wave each
// Partial-force slots per receiver: row q (0..8) and source box j (0..2: x - 1,
// x, x + 1 of the receiver's box) -> slot 3 q + j, three components each.  The
// row-centred near field writes one slot per (row, box) wave; the box-centred
// (masked) form writes a row's whole partial into its j = 1 slot and zeros
// into the other two.
constexpr int NEAR_SLOTS = 3 * NEAR_ROWS;
constexpr int SRC_PAD = 4;    // spare records after the sorted source arrays (read ahead, never used)

struct Recv {
    double x, y, z, q;       // receiver position, charge * xpi
    double qO, qH;           // q * (water O charge), q * (water H charge)
    bool ljO;                // an active receiver with a nonzero LJ constant against water O
    int ljrow, mol, rk;      // LJ table row of the receiver's type (32 * type)
    int by, bz;              // the receiver's box (y, z); every receiver of a wave has box x = bx
};

// Each lane's LJ constants (c, s6) of receiver slot r against water O, staged
// in LDS when the receivers are loaded (stage_recv_ljo): the water-O LJ term,
// the most frequent one, then reads them there instead of gathering them from
// the table in global memory at every source.  Only the lane that writes an
// entry reads it, and a near-field block is one wavefront.
static_assert(NEAR_TILE == 64, "the staged LJ constants are indexed by lane: one wavefront per block");
__device__ __forceinline__ double2 &recv_ljo(int r)
{
    __shared__ double2 s_ljo[2][64];
    return s_ljo[r][threadIdx.x & 63];
}

__device__ __forceinline__ void stage_recv_ljo(const sesmd_forcefield *__restrict__ ff, int r, int ti)
{
    recv_ljo(r) = make_double2(ff->lj_c[ti * 32], ff->lj_s6[ti * 32]);
}

// Sources t in [s0, s1) of one LJ class for NR receivers per lane (lane l
// holds receivers l and 64 + l of its chunk; their sums are independent, so
// each receiver adds the same terms in the same order either way).
// CAT 0: water O (LJ with the receiver's O parameters), 1: surfactant sites (LJ
// from the type table), 2: water H (Coulomb only).  CHECK: test the storage
// ring of every source against each receiver's box (only rows holding a site
// that some neighbour box does not list); MASKED: keep the (receiver kind,
// source kind) pairs of `mask`.
// CEN: the segment lies in the receivers' own box (every source may share a
// molecule with a receiver); elsewhere only sources of split molecules can.
// HQ: both water hydrogens carry one charge, so q_i q_j is per lane for CAT 2.
// The LJ term is evaluated only when some lane holds a pair inside rc (a
// wave-uniform branch): most LJ-class sources lie beyond rc of every receiver
// of the chunk.  Where a pair is outside rc or its LJ constant is zero the
// term is not added, as before the test.
template <int CAT, bool CHECK, bool MASKED, bool CEN, bool HQ, int NR>
__device__ __forceinline__ void near_pair(const Geo &g, const double4 &pj, const int4 &mj, int bx,
                                          const Recv (&R)[NR], const double *__restrict__ ljc,
                                          const double *__restrict__ ljs, unsigned long long mask,
                                          double (&ax)[NR], double (&ay)[NR], double (&az)[NR])
{
    int sk[NR];
    if (CHECK || MASKED) {
        int cls = (mj.y >> 8) & 3, sx = mj.w & 255, sy = (mj.w >> 8) & 255, sz = (mj.w >> 16) & 255;
        int lim = cls == 0 ? 1 : cls == 1 ? 2 : 3, first = cls == 0 ? 0 : cls == 1 ? 2 : 5;
        bool any = false;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            int ring = max(abs(sx - bx), max(abs(sy - R[r].by), abs(sz - R[r].bz)));
            sk[r] = ring <= lim ? first + ring : -1;  // -1: outside every list of the reference
            any = any || sk[r] >= 0;
        }
        if (!__builtin_amdgcn_ballot_w64(any)) return;
    }
    double dx[NR], dy[NR], dz[NR], r2[NR], rinv3[NR], ft[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        dx[r] = R[r].x - pj.x;
        dy[r] = R[r].y - pj.y;
        dz[r] = R[r].z - pj.z;
        r2[r] = dx[r] * dx[r] + dy[r] * dy[r] + dz[r] * dz[r];
        rinv3[r] = rcube64(r2[r]);
        double qq = CAT == 0 ? R[r].qO : (CAT == 2 && HQ) ? R[r].qH : R[r].q * pj.w;
        ft[r] = qq * rinv3[r];
    }
    if (CAT != 2) {
        bool in[NR], any = false;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            in[r] = r2[r] <= g.rc2 && (CAT != 0 || R[r].ljO);
            any = any || in[r];
        }
        if (__builtin_amdgcn_ballot_w64(any)) {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                // the constants are read here, only when the term is needed
                // (registers for them would cost the kernel a wavefront per SIMD)
                double c, s6;
                if (CAT == 0) {
                    double2 cs = recv_ljo(r);
                    c = cs.x;
                    s6 = cs.y;
                } else {
                    int tj = R[r].ljrow + (mj.y & 255);
                    c = ljc[tj];
                    s6 = ljs[tj];
                }
                double r6 = rinv3[r] * rinv3[r];
                double rinv2 = r6 * (r2[r] * r2[r]);  // r^-6 r^4
                double flj = c * r6 * (s6 * r6 - 0.5) * rinv2;
                ft[r] = (c != 0.0 && in[r]) ? ft[r] + flj : ft[r];
            }
        }
    }
    if (CHECK || MASKED || CEN) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            bool ok = mj.z != R[r].mol;
            if (CHECK || MASKED) ok = ok && sk[r] >= 0;
            if (MASKED) ok = ok && ((mask >> (max(R[r].rk, 0) * 9 + max(sk[r], 0))) & 1ull);
            ft[r] = ok ? ft[r] : 0.0;
        }
    } else if ((mj.y >> 10) & 1) {  // wave-uniform: a branch, not a select
        __asm__ volatile("");
#pragma unroll
        for (int r = 0; r < NR; r++) ft[r] = mj.z != R[r].mol ? ft[r] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        ax[r] = fma(dx[r], ft[r], ax[r]);
        ay[r] = fma(dy[r], ft[r], ay[r]);
        az[r] = fma(dz[r], ft[r], az[r]);
    }
}

template <int CAT, bool CHECK, bool MASKED, bool CEN, bool HQ, int NR>
__device__ __forceinline__ void near_segment(const Geo &g, const double4 *__restrict__ spos,
                                             const int4 *__restrict__ smeta, int s0, int s1,
                                             int bx, const Recv (&R)[NR],
                                             const double *__restrict__ ljc,
                                             const double *__restrict__ ljs,
                                             unsigned long long mask, double (&ax)[NR],
                                             double (&ay)[NR], double (&az)[NR])
{
    if (s0 >= s1) return;
    // Sources two at a time, the next two loaded while these are used: scalar
    // loads return out of order, so each use waits for every load in flight;
    // pairing halves the waits per source.  The loads run up to three records
    // past s1 (the sorted arrays carry SRC_PAD spare records) and use one
    // incremented pointer per array, which keeps the scalar unit's address
    // arithmetic to a few instructions per pair of sources.
    // (A loop body without the branch on the second source lets the two
    // sources' arithmetic interleave, as the FP32 form does; in FP64 the extra
    // live registers cost more than the interleaving gains.)
    const double4 *ps = spos + s0;
    const int4 *pm = smeta + s0;
    int4 mA = pm[0], mB = pm[1];
    double4 pA = ps[0], pB = ps[1];
    for (int t = s1 - s0; t > 0; t -= 2) {
        const int4 ma = mA, mb = mB;
        const double4 pa = pA, pb = pB;
        ps += 2;
        pm += 2;
        mA = pm[0];
        mB = pm[1];
        pA = ps[0];
        pB = ps[1];
        __asm__ volatile("" ::: "memory");  // keeps these loads ahead of the pair arithmetic
        near_pair<CAT, CHECK, MASKED, CEN, HQ, NR>(g, pa, ma, bx, R, ljc, ljs, mask, ax, ay, az);
        if (t > 1)
            near_pair<CAT, CHECK, MASKED, CEN, HQ, NR>(g, pb, mb, bx, R, ljc, ljs, mask, ax, ay, az);
    }
}

template <bool CHECK, bool MASKED, bool CEN, bool HQ, int NR>
__device__ __forceinline__ void near_box(const Geo &g, const double4 *__restrict__ spos,
                                         const int4 *__restrict__ smeta, const int *__restrict__ bs4,
                                         int bb, int bx, const Recv (&R)[NR],
                                         const double *__restrict__ ljc, const double *__restrict__ ljs,
                                         unsigned long long mask, double (&ax)[NR], double (&ay)[NR],
                                         double (&az)[NR])
{
    // one source box: the classes O, S, H
    int k0 = __builtin_amdgcn_readfirstlane(bs4[4 * bb]);
    int k1 = __builtin_amdgcn_readfirstlane(bs4[4 * bb + 1]);
    int k2 = __builtin_amdgcn_readfirstlane(bs4[4 * bb + 2]);
    int k3 = __builtin_amdgcn_readfirstlane(bs4[4 * bb + 3]);
    near_segment<0, CHECK, MASKED, CEN, HQ, NR>(g, spos, smeta, k0, k1, bx, R, ljc, ljs, mask, ax, ay, az);
    near_segment<1, CHECK, MASKED, CEN, HQ, NR>(g, spos, smeta, k1, k2, bx, R, ljc, ljs, mask, ax, ay, az);
    near_segment<2, CHECK, MASKED, CEN, HQ, NR>(g, spos, smeta, k2, k3, bx, R, ljc, ljs, mask, ax, ay, az);
}

template <bool CHECK, bool MASKED, bool HQ, int NR>
__device__ __forceinline__ void near_row(const Geo &g, const double4 *__restrict__ spos,
                                         const int4 *__restrict__ smeta, const int *__restrict__ bs4,
                                         int r0, int r1, int bc, int bx,
                                         const Recv (&R)[NR], const double *__restrict__ ljc,
                                         const double *__restrict__ ljs, unsigned long long mask,
                                         double (&ax)[NR], double (&ay)[NR], double (&az)[NR])
{
    // boxes r0..r1 (consecutive indices); bc: the receivers' own box
    for (int bb = r0; bb <= r1; bb++) {
        if (bb == bc)
            near_box<CHECK, MASKED, true, HQ, NR>(g, spos, smeta, bs4, bb, bx, R, ljc, ljs, mask, ax,
                                                  ay, az);
        else
            near_box<CHECK, MASKED, false, HQ, NR>(g, spos, smeta, bs4, bb, bx, R, ljc, ljs, mask, ax,
                                                   ay, az);
    }
}

// Mixed precision (SESMD_MIXED): the same pair set and order, pair arithmetic
// in FP32 on coordinates relative to the source box's centre, and every pair
// force component accumulated in fixed point with a resolution of 2^-32 force
// units, held in an FP64 register.  The accumulator starts at B = 1.5 2^20:
// the doubles of [2^20, 2^21) are exactly the multiples of 2^-32 there, so
// each pair's acc = fma(d, f, acc) rounds acc + d f (d f being the exact
// product of two FP32 factors) once, to that grid, and while every partial
// sum stays within 2^19 of B, acc - B is the sum of the terms rounded to
// multiples of 2^-32 -- a 52-bit fixed-point accumulator at the price of one
// conversion and one FMA per component (a 64-bit integer accumulator costs
// two more integer adds).  The bias comes off once at the end.  The order of
// the terms is fixed (each receiver's sources in sorted order), so results
// are deterministic; a tie (a product exactly half-way between two grid
// points) rounds to the even neighbour of the sum.  Headroom: each lane keeps
// the largest |f| (pair force over distance) it has seen; an item where that
// times the longest distance its rows allow, or a receiver's final sum,
// reaches 2^18 counts in counters[3] (a factor of two below the range, for
// partial sums that exceed the final one).
constexpr double FIX_BIAS = 1572864.0;  // 1.5 2^20
constexpr float FIX_LIMIT = 262144.0f;  // 2^18

struct Recv32 {
    float q, qO, qH, cO, s6O;
};

struct Acc3 {
    double x, y, z;  // biased fixed-point sums (see FIX_BIAS)
};

__device__ __forceinline__ void fix_add(double &acc, float d, float f)
{
    acc = fma((double)d, (double)f, acc);
}

template <int CAT, bool CHECK, bool CEN, bool HQ, int NR>
__device__ __forceinline__ void near_pair32(const float4 &pj, const int4 &mj, int bx,
                                            const float (&xr)[NR], const float (&yr)[NR],
                                            const float (&zr)[NR], const Recv (&R)[NR],
                                            const Recv32 (&Q)[NR], const float *__restrict__ lj32,
                                            float rc2, Acc3 (&acc)[NR], float &mf)
{
    bool okc[NR];
    if (CHECK) {
        int cls = (mj.y >> 8) & 3, sx = mj.w & 255, sy = (mj.w >> 8) & 255, sz = (mj.w >> 16) & 255;
        int lim = cls == 0 ? 1 : cls == 1 ? 2 : 3;
        bool any = false;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            okc[r] = max(abs(sx - bx), max(abs(sy - R[r].by), abs(sz - R[r].bz))) <= lim;
            any = any || okc[r];
        }
        if (!__builtin_amdgcn_ballot_w64(any)) return;
    }
    float dx[NR], dy[NR], dz[NR], r2[NR], rinv2[NR], ft[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        dx[r] = xr[r] - pj.x;
        dy[r] = yr[r] - pj.y;
        dz[r] = zr[r] - pj.z;
        r2[r] = dx[r] * dx[r] + dy[r] * dy[r] + dz[r] * dz[r];
        float rinv = __builtin_amdgcn_rsqf(r2[r]);
        rinv2[r] = rinv * rinv;
        // hydrogen sources of one charge (HQ): q_i q_H is the receiver's own
        float qq = CAT == 0 ? Q[r].qO : (CAT == 2 && HQ) ? Q[r].qH : Q[r].q * pj.w;
        ft[r] = qq * (rinv2[r] * rinv);
    }
    if (CAT != 2) {
        bool in[NR], any = false;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            in[r] = r2[r] <= rc2 && (CAT != 0 || R[r].ljO);
            any = any || in[r];
        }
        if (__builtin_amdgcn_ballot_w64(any)) {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                float c, s6;
                if (CAT == 0) {
                    c = Q[r].cO;
                    s6 = Q[r].s6O;
                } else {
                    int tj = R[r].ljrow + (mj.y & 255);
                    c = lj32[tj];
                    s6 = lj32[1024 + tj];
                }
                float r6 = rinv2[r] * rinv2[r] * rinv2[r];
                float flj = c * r6 * (s6 * r6 - 0.5f) * rinv2[r];
                ft[r] = (c != 0.0f && in[r]) ? ft[r] + flj : ft[r];
            }
        }
    }
    if (CHECK || CEN) {
#pragma unroll
        for (int r = 0; r < NR; r++) ft[r] = (mj.z != R[r].mol && (!CHECK || okc[r])) ? ft[r] : 0.0f;
    } else if ((mj.y >> 10) & 1) {  // wave-uniform: a branch, not a select
        __asm__ volatile("");
#pragma unroll
        for (int r = 0; r < NR; r++) ft[r] = mj.z != R[r].mol ? ft[r] : 0.0f;
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        fix_add(acc[r].x, dx[r], ft[r]);
        fix_add(acc[r].y, dy[r], ft[r]);
        fix_add(acc[r].z, dz[r], ft[r]);
        mf = fmaxf(mf, fabsf(ft[r]));
    }
}

template <int CAT, bool CHECK, bool CEN, bool HQ, int NR>
__device__ __forceinline__ void near_segment32(const float4 *__restrict__ spos32,
                                               const int4 *__restrict__ smeta, int s0, int s1,
                                               int bx, const float (&xr)[NR],
                                               const float (&yr)[NR], const float (&zr)[NR],
                                               const Recv (&R)[NR], const Recv32 (&Q)[NR],
                                               const float *__restrict__ lj32, float rc2,
                                               Acc3 (&acc)[NR], float &mf)
{
    if (s0 >= s1) return;
    // No branch between the loop body's two sources, so their pair arithmetic
    // interleaves; an odd last source runs after the loop.
    const float4 *ps = spos32 + s0;
    const int4 *pm = smeta + s0;
    int4 mA = pm[0], mB = pm[1];
    float4 pA = ps[0], pB = ps[1];
    int t = s1 - s0;
    for (; t > 1; t -= 2) {
        const int4 ma = mA, mb = mB;
        const float4 pa = pA, pb = pB;
        ps += 2;
        pm += 2;
        mA = pm[0];
        mB = pm[1];
        pA = ps[0];
        pB = ps[1];
        __asm__ volatile("" ::: "memory");  // keeps these loads ahead of the pair arithmetic
        near_pair32<CAT, CHECK, CEN, HQ, NR>(pa, ma, bx, xr, yr, zr, R, Q, lj32, rc2, acc, mf);
        near_pair32<CAT, CHECK, CEN, HQ, NR>(pb, mb, bx, xr, yr, zr, R, Q, lj32, rc2, acc, mf);
    }
    if (t == 1) near_pair32<CAT, CHECK, CEN, HQ, NR>(pA, mA, bx, xr, yr, zr, R, Q, lj32, rc2, acc, mf);
}

template <bool CHECK, bool HQ, int NR>
__device__ __forceinline__ void near_row32(const Geo &g, const float4 *__restrict__ spos32,
                                           const int4 *__restrict__ smeta, const int *__restrict__ bs4,
                                           int r0, int r1, int bc, int bx,
                                           const Recv (&R)[NR], const Recv32 (&Q)[NR],
                                           const float *__restrict__ lj32, Acc3 (&acc)[NR], float &mf)
{
    int nb = g.nb;
    float rc2 = (float)g.rc2;
    double cy = d_cen4(g, (r0 / nb) % nb), cz = d_cen4(g, r0 / (nb * nb));
    float yr[NR], zr[NR], xr[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        yr[r] = (float)(R[r].y - cy);
        zr[r] = (float)(R[r].z - cz);
    }
    for (int bb = r0; bb <= r1; bb++) {
        double cx = d_cen4(g, bb % nb);
#pragma unroll
        for (int r = 0; r < NR; r++) xr[r] = (float)(R[r].x - cx);
        int k0 = __builtin_amdgcn_readfirstlane(bs4[4 * bb]);
        int k1 = __builtin_amdgcn_readfirstlane(bs4[4 * bb + 1]);
        int k2 = __builtin_amdgcn_readfirstlane(bs4[4 * bb + 2]);
        int k3 = __builtin_amdgcn_readfirstlane(bs4[4 * bb + 3]);
        if (bb == bc) {
            near_segment32<0, CHECK, true, HQ, NR>(spos32, smeta, k0, k1, bx, xr, yr, zr, R, Q, lj32, rc2, acc, mf);
            near_segment32<1, CHECK, true, HQ, NR>(spos32, smeta, k1, k2, bx, xr, yr, zr, R, Q, lj32, rc2, acc, mf);
            near_segment32<2, CHECK, true, HQ, NR>(spos32, smeta, k2, k3, bx, xr, yr, zr, R, Q, lj32, rc2, acc, mf);
        } else {
            near_segment32<0, CHECK, false, HQ, NR>(spos32, smeta, k0, k1, bx, xr, yr, zr, R, Q, lj32, rc2, acc, mf);
            near_segment32<1, CHECK, false, HQ, NR>(spos32, smeta, k1, k2, bx, xr, yr, zr, R, Q, lj32, rc2, acc, mf);
            near_segment32<2, CHECK, false, HQ, NR>(spos32, smeta, k2, k3, bx, xr, yr, zr, R, Q, lj32, rc2, acc, mf);
        }
    }
}

// One receiver: sorted record k (active) or an empty lane.
__device__ __forceinline__ void load_recv(const Geo &g, const double4 *__restrict__ spos,
                                          const int4 *__restrict__ smeta, const int *__restrict__ rkind,
                                          const sesmd_forcefield *__restrict__ ff, int k, bool active,
                                          int by, int bz, Recv &R, int &ti)
{
    R.x = R.y = R.z = 1e18;  // an empty lane: far from every source
    R.q = 0.0;
    R.mol = -1;
    R.rk = -1;
    R.by = by;
    R.bz = bz;
    ti = 0;
    if (active) {
        double4 pi = spos[k];
        int4 me = smeta[k];
        R.x = pi.x;
        R.y = pi.y;
        R.z = pi.z;
        R.q = pi.w * g.xpi;
        ti = me.y & 255;
        R.mol = me.z;
        R.rk = rkind[me.x];
        if (R.rk < 0) active = false;
    }
    R.qO = R.q * ff->q[0];
    R.qH = R.q * ff->q[1];
    R.ljO = active && ff->lj_c[ti * 32] != 0.0;
    R.ljrow = ti * 32;
}

// The receivers of a wave against source boxes r0..r1 (one row, consecutive
// indices, every receiver in box column bx); bc: the source box that may hold
// receivers' own molecules unsplit (-1: none).
template <bool MASKED, bool HQ, bool MIX, int NR>
__device__ __forceinline__ void near_sources(const Geo &g, const double4 *__restrict__ spos,
                                             const int4 *__restrict__ smeta, const int *__restrict__ bs4,
                                             bool check, const sesmd_forcefield *__restrict__ ff,
                                             unsigned long long mask, const float4 *__restrict__ spos32,
                                             const float *__restrict__ lj32, int r0, int r1, int bc, int bx,
                                             const Recv (&R)[NR], const int (&ti)[NR], double (&ax)[NR],
                                             double (&ay)[NR], double (&az)[NR], int *__restrict__ counters)
{
    if (MIX) {
        Recv32 Q[NR];
        Acc3 acc[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            Q[r].q = (float)R[r].q;
            Q[r].qO = (float)R[r].qO;
            Q[r].qH = (float)R[r].qH;
            Q[r].cO = lj32[ti[r] * 32];
            Q[r].s6O = lj32[1024 + ti[r] * 32];
            acc[r].x = acc[r].y = acc[r].z = FIX_BIAS;
        }
        float mf = 0.0f;
        if (check)
            near_row32<true, HQ, NR>(g, spos32, smeta, bs4, r0, r1, bc, bx, R, Q, lj32, acc, mf);
        else
            near_row32<false, HQ, NR>(g, spos32, smeta, bs4, r0, r1, bc, bx, R, Q, lj32, acc, mf);
        // a receiver in the 3 x 3 boxes around the row's centre box and a
        // source in the row's three boxes are at most sqrt(17) box sides apart
        float dmax = 4.1231056f * (float)g.rs;
        bool out = mf * dmax >= FIX_LIMIT;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            double sx = acc[r].x - FIX_BIAS, sy = acc[r].y - FIX_BIAS, sz = acc[r].z - FIX_BIAS;
            out = out || fmax(fabs(sx), fmax(fabs(sy), fabs(sz))) >= (double)FIX_LIMIT;
            ax[r] += sx;
            ay[r] += sy;
            az[r] += sz;
        }
        if (__builtin_amdgcn_ballot_w64(out) && (threadIdx.x & 63) == 0) atomicAdd(&counters[3], 1);
    } else if (MASKED)
        near_row<true, true, HQ, NR>(g, spos, smeta, bs4, r0, r1, bc, bx, R, ff->lj_c, ff->lj_s6, mask, ax, ay,
                                     az);
    else if (check)
        near_row<true, false, HQ, NR>(g, spos, smeta, bs4, r0, r1, bc, bx, R, ff->lj_c, ff->lj_s6, mask, ax, ay,
                                      az);
    else
        near_row<false, false, HQ, NR>(g, spos, smeta, bs4, r0, r1, bc, bx, R, ff->lj_c, ff->lj_s6, mask, ax,
                                       ay, az);
}

// One wave's share of a box-centred work item: receivers k0 + lane (+ 64 for
// NR = 2), below kend, all in box b, against one row.
template <bool MASKED, bool HQ, bool MIX, int NR>
__device__ __forceinline__ void near_item(
    const Geo &g, const double4 *__restrict__ spos, const int4 *__restrict__ smeta,
    const int *__restrict__ bs4, const int *__restrict__ bflag, const int *__restrict__ rkind,
    const sesmd_forcefield *__restrict__ ff, unsigned long long mask, const float4 *__restrict__ spos32,
    const float *__restrict__ lj32, double *__restrict__ part, int b, int k0, int kend, int row, int lane)
{
    int nb = g.nb;
    int bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
    Recv R[NR];
    int ti[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        int k = k0 + NEAR_TILE * r + lane;
        load_recv(g, spos, smeta, rkind, ff, k, k < kend, by, bz, R[r], ti[r]);
        if (!MIX) stage_recv_ljo(ff, r, ti[r]);
    }
    double ax[NR], ay[NR], az[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) ax[r] = ay[r] = az[r] = 0.0;
    int z = bz + row / 3 - 1, y = by + row % 3 - 1;
    if (z >= 0 && z < nb && y >= 0 && y < nb) {
        int r0 = max(bx - 1, 0) + nb * (y + nb * z), r1 = min(bx + 1, nb - 1) + nb * (y + nb * z);
        bool check = MASKED;
        if (!MASKED)
            for (int bb = r0; bb <= r1; bb++) check = check || bflag[bb];
        near_sources<MASKED, HQ, MIX, NR>(g, spos, smeta, bs4, check, ff, mask, spos32, lj32, r0, r1, b, bx, R,
                                          ti, ax, ay, az, nullptr);
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
        int k = k0 + NEAR_TILE * r + lane;
        if (k < kend) {  // the row's partial in its j = 1 slot, zeros in j = 0 and 2
            size_t o = (size_t)(3 * (3 * row)) * g.n + k;
            part[o] = part[o + g.n] = part[o + 2 * g.n] = 0.0;
            part[o + 3 * g.n] = ax[r];
            part[o + 4 * g.n] = ay[r];
            part[o + 5 * g.n] = az[r];
            part[o + 6 * g.n] = part[o + 7 * g.n] = part[o + 8 * g.n] = 0.0;
        }
    }
}

template <bool MASKED, bool HQ, bool MIX>
__global__ __launch_bounds__(NEAR_TILE) void k_near(
    Geo g, const double4 *__restrict__ spos, const int4 *__restrict__ smeta,
    const int *__restrict__ bstart, const int *__restrict__ bs4, const int *__restrict__ bflag,
    const int *__restrict__ rkind, const int2 *__restrict__ work, const int *__restrict__ nwork,
    const sesmd_forcefield *__restrict__ ff, unsigned long long mask, double *__restrict__ f,
    const float4 *__restrict__ spos32, const float *__restrict__ lj32, double *__restrict__ part)
{
    // One 64-lane wavefront per (work item, row): the NEAR_CHUNK receivers of
    // an item (box b) against one (dz, dy) row of the 3^3 neighbourhood -- the
    // three x-neighbours of a row are consecutive boxes.  A chunk of more than
    // 64 receivers gives each lane two of them, so every source record loaded
    // serves two pairs per lane.  Every lane walks the same source sequence,
    // so each source record is a wave-uniform scalar load and the per-source
    // integer work runs on the scalar unit; the vector unit only does the pair
    // arithmetic.  Within a box the sources are ordered by LJ class, so the
    // loops carry no per-source type branch.  Each wave writes its partial
    // forces; the site-terms kernel adds the nine rows in order (no atomics,
    // deterministic), and no wave waits for the other rows of its item.  This
    // box-centred form serves masked evaluations (subsets of the reference's
    // 39 kernels); the full near field of a step runs k_near_rows.
    int item = blockIdx.x / NEAR_ROWS;  // items come longest first (see k_worklist)
    if (item >= *nwork) return;
    int row = blockIdx.x % NEAR_ROWS, lane = threadIdx.x;
    int2 w = work[item];
    int b = w.x, k0 = w.y;
    int kend = min(bstart[b + 1], k0 + NEAR_CHUNK);
    if (kend - k0 > NEAR_TILE)
        near_item<MASKED, HQ, MIX, 2>(g, spos, smeta, bs4, bflag, rkind, ff, mask, spos32, lj32, part, b,
                                      k0, kend, row, lane);
    else
        near_item<MASKED, HQ, MIX, 1>(g, spos, smeta, bs4, bflag, rkind, ff, mask, spos32, lj32, part, b,
                                      k0, kend, row, lane);
}

// The receiver list of source row (x0, y, z): the boxes (x0, y + dy, z + dz),
// slot 3 (dz + 1) + (dy + 1), each box's sorted sites in order; slots outside
// the grid are empty.  Slot 4 is the row's own centre box.
struct RowRecv {
    int start[9], pre[10];  // first sorted site of each slot's box, list position of each slot
};

__device__ __forceinline__ void row_receivers(const Geo &g, const int *__restrict__ bstart, int x0, int y,
                                              int z, RowRecv &L)
{
    int nb = g.nb;
    L.pre[0] = 0;
#pragma unroll
    for (int q = 0; q < 9; q++) {
        int yy = y + q % 3 - 1, zz = z + q / 3 - 1, cnt = 0;
        L.start[q] = 0;
        if (yy >= 0 && yy < nb && zz >= 0 && zz < nb) {
            int b = x0 + nb * (yy + nb * zz);
            L.start[q] = bstart[b];
            cnt = bstart[b + 1] - L.start[q];
        }
        L.pre[q + 1] = L.pre[q] + cnt;
    }
}

// Row-centred work list: one item per NEAR_CHUNK receivers of each row's
// receiver list, longest first.  Items are placed by a bucket order on their
// cost estimate (receivers x row sources; four buckets per octave), so no sort
// launch is needed; the order within a bucket is arrival order, which changes
// no result (every item writes its own receivers' row slots).
constexpr int RL_BUCKETS = 128;
constexpr int RL_ROWS_LDS = 4096;  // rows whose counts the row list keeps in LDS (16^3)

__device__ __forceinline__ int cost_bucket(unsigned cost)
{
    cost |= 1u;
    int e = 31 - __clz(cost);
    int frac = e >= 2 ? (cost >> (e - 2)) & 3 : (cost << (2 - e)) & 3;
    return min(4 * e + frac, RL_BUCKETS - 1);
}

__device__ __forceinline__ void row_item_info(const Geo &g, const int *__restrict__ bstart, int q, int &total,
                                              int &src)
{
    int nb = g.nb, x0 = q % nb, y = (q / nb) % nb, z = q / (nb * nb);
    RowRecv L;
    row_receivers(g, bstart, x0, y, z, L);
    total = L.pre[9];
    int r0 = max(x0 - 1, 0) + nb * (y + nb * z), r1 = min(x0 + 1, nb - 1) + nb * (y + nb * z);
    src = bstart[r1 + 1] - bstart[r0];
}

// The row list in one block of 1024 threads (called after the box sort's
// scan, in the same launch): tally the items per cost bucket in LDS, scan the
// buckets from the costliest down, place the items.
__device__ __forceinline__ void row_list_block(const Geo &g, const int *__restrict__ bstart,
                                               int2 *__restrict__ work, int *__restrict__ nwork, int max_work)
{
    __shared__ int s_hist[RL_BUCKETS], s_off[RL_BUCKETS];
    __shared__ int2 s_row[RL_ROWS_LDS];  // each row's (receivers, sources), kept for the second pass
    for (int i = threadIdx.x; i < RL_BUCKETS; i += blockDim.x) s_hist[i] = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < g.nb3; q += blockDim.x) {
        int total, src;
        row_item_info(g, bstart, q, total, src);
        if (q < RL_ROWS_LDS) s_row[q] = make_int2(total, src);
        for (int c0 = 0; c0 < total; c0 += NEAR_CHUNK)
            atomicAdd(&s_hist[cost_bucket((unsigned)(min(NEAR_CHUNK, total - c0) * src))], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int b = RL_BUCKETS - 1; b >= 0; b--) {
            s_off[b] = run;
            run += s_hist[b];
        }
        *nwork = min(run, max_work);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < g.nb3; q += blockDim.x) {
        int total, src;
        if (q < RL_ROWS_LDS) {
            total = s_row[q].x;
            src = s_row[q].y;
        } else {
            row_item_info(g, bstart, q, total, src);
        }
        for (int c0 = 0; c0 < total; c0 += NEAR_CHUNK) {
            int pos = atomicAdd(&s_off[cost_bucket((unsigned)(min(NEAR_CHUNK, total - c0) * src))], 1);
            if (pos < max_work) work[pos] = make_int2(q, c0);
        }
    }
}

// The box sort's scan and, from its box starts, the row list: one block.
// The boxes holding sites, in box order, as (box, first, end) sorted
// positions; every later entry is box -1.  One block (after its bstart).
__device__ __forceinline__ void occupied_block(const Geo &g, const int *__restrict__ bstart, int4 *__restrict__ obox)
{
    __shared__ int s_w[16];
    int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = (int)blockDim.x >> 6, run = 0;
    for (int b0 = 0; b0 < g.nb3; b0 += blockDim.x) {
        int b = b0 + t, k0 = 0, k1 = 0;
        if (b < g.nb3) {
            k0 = bstart[b];
            k1 = bstart[b + 1];
        }
        bool f = k1 > k0;
        unsigned long long m = __builtin_amdgcn_ballot_w64(f);
        int before = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (lane == 0) s_w[w] = __builtin_popcountll(m);
        __syncthreads();
        int off = 0, tot = 0;
        for (int j = 0; j < nw; j++) {
            off += j < w ? s_w[j] : 0;
            tot += s_w[j];
        }
        if (f) obox[run + off + before] = make_int4(b, k0, k1, 0);
        run += tot;
        __syncthreads();  // s_w is rewritten by the next round
    }
    for (int q = run + t; q < g.nb3; q += blockDim.x) obox[q] = make_int4(-1, 0, 0, 0);
}

__device__ __forceinline__ void key_scan_rows_blk(const Geo &g,
                                                    const int *__restrict__ cnt,
                                                    int *__restrict__ bs4,
                                                    int *__restrict__ bstart,
                                                    int2 *__restrict__ work,
                                                    int *__restrict__ nwork,
                                                    int max_work,
                                                    int4 *__restrict__ obox)
{
    key_scan_block(g, cnt, bs4, bstart);
    __threadfence_block();  // this block's bstart stores before its loads below
    __syncthreads();
    row_list_block(g, bstart, work, nwork, max_work);
    occupied_block(g, bstart, obox);
}

__global__ __launch_bounds__(1024) void k_key_scan_rows(Geo g, const int *__restrict__ cnt, int *__restrict__ bs4,
                                                        int *__restrict__ bstart, int2 *__restrict__ work,
                                                        int *__restrict__ nwork, int max_work,
                                                        int4 *__restrict__ obox)
{
    key_scan_rows_blk(g, cnt, bs4, bstart, work, nwork, max_work, obox);
}

// One row-centred item: positions c0 + lane (+ 64 for NR = 2) of the row's
// receiver list, below cend, against the sources of box j (0..2) of the row
// (x0 - 1 + j, y, z); each receiver's partial goes to the slot (row as its
// box sees it, j).  A box outside the grid has no slot (the site terms skip
// it) and returns at once.
template <bool HQ, bool MIX, int NR>
__device__ __forceinline__ void near_rows_item(
    const Geo &g, const double4 *__restrict__ spos, const int4 *__restrict__ smeta,
    const int *__restrict__ bs4, const int *__restrict__ bflag, const int *__restrict__ rkind,
    const sesmd_forcefield *__restrict__ ff, const float4 *__restrict__ spos32,
    const float *__restrict__ lj32, double *__restrict__ part, const RowRecv &L, int x0, int y, int z,
    int j, int c0, int cend, int lane, int *__restrict__ counters)
{
    int nb = g.nb, xs = x0 - 1 + j;
    if (xs < 0 || xs >= nb) return;  // wave-uniform
    Recv R[NR];
    int ti[NR], k[NR], slot[NR];
    bool cen = false;  // some receiver of this wave lives in box (x0, y, z)
#pragma unroll
    for (int r = 0; r < NR; r++) {
        int p = c0 + NEAR_TILE * r + lane;
        // the list slot holding position p: the last slot starting at or
        // before p (empty slots start where the next one does); selects, so
        // L stays in scalar registers
        int m = 0, st = L.start[0], pr = 0;
#pragma unroll
        for (int q = 1; q < 9; q++)
            if (p >= L.pre[q]) {
                m = q;
                st = L.start[q];
                pr = L.pre[q];
            }
        bool on = p < cend;
        int ddy = m % 3 - 1, ddz = m / 3 - 1;
        k[r] = st + p - pr;
        slot[r] = 8 - m;  // the row's (dz, dy) seen from the receiver's box: 3 (1 - ddz) + (1 - ddy)
        load_recv(g, spos, smeta, rkind, ff, k[r], on, y + ddy, z + ddz, R[r], ti[r]);
        cen = cen || (on && m == 4);
    }
    if (!MIX)
#pragma unroll
        for (int r = 0; r < NR; r++) stage_recv_ljo(ff, r, ti[r]);
    bool anycen = __builtin_amdgcn_ballot_w64(cen) != 0;
    double ax[NR], ay[NR], az[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) ax[r] = ay[r] = az[r] = 0.0;
    // one source box; a box whose sites every neighbour lists needs no ring
    // test (its sources pass it), so its own flag decides
    int bb = xs + nb * (y + nb * z);
    bool check = bflag[bb] != 0;
    int bc = anycen ? x0 + nb * (y + nb * z) : -1;
    near_sources<false, HQ, MIX, NR>(g, spos, smeta, bs4, check, ff, 0ull, spos32, lj32, bb, bb, bc, x0, R, ti,
                                     ax, ay, az, counters);
#pragma unroll
    for (int r = 0; r < NR; r++) {
        if (c0 + NEAR_TILE * r + lane < cend) {
            size_t o = (size_t)(3 * (3 * slot[r] + j)) * g.n + k[r];
            part[o] = ax[r];
            part[o + g.n] = ay[r];
            part[o + 2 * g.n] = az[r];
        }
    }
}

// FP64: at most 96 VGPRs, five wavefronts per SIMD (measured 2.5% faster than four)
template <bool HQ, bool MIX>
__global__ __launch_bounds__(NEAR_TILE, MIX ? 1 : 5) void k_near_rows(
    Geo g, const double4 *__restrict__ spos, const int4 *__restrict__ smeta,
    const int *__restrict__ bstart, const int *__restrict__ bs4, const int *__restrict__ bflag,
    const int *__restrict__ rkind, const int2 *__restrict__ work, const int *__restrict__ nwork,
    const sesmd_forcefield *__restrict__ ff, const float4 *__restrict__ spos32,
    const float *__restrict__ lj32, double *__restrict__ part, int *__restrict__ counters, int mr)
{
    // The full near field, row-centred: one 64-lane wavefront per work item,
    // the item being up to NEAR_CHUNK receivers of the nine boxes that all see
    // one source row (three x-consecutive boxes) -- the receivers of several
    // boxes fill the lanes, where a box-centred item leaves the lanes beyond
    // its box's last site empty.  Every (receiver, row) pair of the
    // box-centred form is evaluated once, with the same sources in the same
    // order, into the same row slot; slots of rows outside the grid are not
    // written (the site-terms kernel skips them).
    // block 3 i + j: item i (items come longest first, see row_list_block),
    // source box j of its row -- a row's three boxes as three waves keep the
    // waves short enough for the dispatcher to balance the SIMDs
    // The cheapest quarter of the list (items split .. nw - 1) runs as
    // halves of 64 receivers: blocks 3 i + j take the first half, and a
    // second segment of blocks after the 3 mr of the first, dispatched last,
    // takes the second halves -- short waves where SIMDs run out of work.
    int nw = *nwork, split = nw - nw / 4, item, j, sub = 0;
    int b = blockIdx.x;
    if (b < 3 * mr) {
        item = b / 3;
        j = b % 3;
    } else {
        item = split + (b - 3 * mr) / 3;
        j = (b - 3 * mr) % 3;
        sub = 1;
    }
    if (item >= nw) return;
    bool half = item >= split;
    int lane = threadIdx.x, nb = g.nb;
    int2 w = work[item];
    int q = w.x, c0 = w.y;
    int x0 = q % nb, y = (q / nb) % nb, z = q / (nb * nb);
    // the row's receiver list is the same for every lane: one copy in LDS
    // (the per-lane slot search below indexes it)
    __shared__ RowRecv L;
    {
        RowRecv Lr;
        row_receivers(g, bstart, x0, y, z, Lr);
        if (lane == 0) L = Lr;
        __syncthreads();
    }
    int cend = min(L.pre[9], c0 + NEAR_CHUNK);
    if (half) {  // receivers c0 + 64 sub .. + 63
        int s0 = c0 + NEAR_TILE * sub, s1 = min(cend, s0 + NEAR_TILE);
        if (s0 < s1)
            near_rows_item<HQ, MIX, 1>(g, spos, smeta, bs4, bflag, rkind, ff, spos32, lj32, part, L, x0, y, z, j,
                                       s0, s1, lane, counters);
    } else if (cend - c0 > NEAR_TILE) {
        near_rows_item<HQ, MIX, 2>(g, spos, smeta, bs4, bflag, rkind, ff, spos32, lj32, part, L, x0, y, z, j, c0,
                                   cend, lane, counters);
    } else {
        near_rows_item<HQ, MIX, 1>(g, spos, smeta, bs4, bflag, rkind, ff, spos32, lj32, part, L, x0, y, z, j, c0,
                                   cend, lane, counters);
    }
}


// Pairs of sorted site k for roofline accounting: np = intermolecular
// sources in the 3^3 block that some reference list holds (the near field's
// pair set), nl = those within rc with an LJ term.  Each unordered pair is
// counted from both sides.
__device__ void site_pair_count(const Geo &g, const double4 *__restrict__ spos, const int4 *__restrict__ smeta,
                                const int *__restrict__ bstart, const int *__restrict__ rkind,
                                const sesmd_forcefield *__restrict__ ff, int k, unsigned long long &np,
                                unsigned long long &nl)
{
    np = nl = 0;
    int4 me = smeta[k];
    if (rkind[me.x] < 0) return;
    double4 pi = spos[k];
    // the box of sorted position k
    int lo = 0, hi = g.nb3;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    int b = lo;
    int nb = g.nb, bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int x = bx + dx, y = by + dy, z = bz + dz;
                if (x < 0 || y < 0 || z < 0 || x >= nb || y >= nb || z >= nb) continue;
                int bb = x + nb * (y + nb * z);
                for (int t = bstart[bb]; t < bstart[bb + 1]; t++) {
                    int4 mj = smeta[t];
                    if (mj.z == me.z) continue;
                    int ring = max(abs((mj.w & 255) - bx),
                                   max(abs(((mj.w >> 8) & 255) - by), abs(((mj.w >> 16) & 255) - bz)));
                    if (src_kind((mj.y >> 8) & 3, ring) < 0) continue;
                    np++;
                    double4 pj = spos[t];
                    double ddx = pi.x - pj.x, ddy = pi.y - pj.y, ddz = pi.z - pj.z;
                    double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (r2 <= g.rc2 && ff->lj_c[(me.y & 255) * 32 + (mj.y & 255)] != 0.0) nl++;
                }
            }
}

// pair counts for roofline accounting (unordered pairs, counted from both sides / 2)
__global__ void k_pair_count(Geo g, const double4 *__restrict__ spos, const int4 *__restrict__ smeta,
                             const int *__restrict__ bstart, const int *__restrict__ rkind,
                             const sesmd_forcefield *__restrict__ ff,
                             unsigned long long *__restrict__ cnt)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.n) return;
    unsigned long long np, nl;
    site_pair_count(g, spos, smeta, bstart, rkind, ff, k, np, nl);
    if (np) atomicAdd(&cnt[0], np);
    if (nl) atomicAdd(&cnt[1], nl);
}

// ------------------------------------------------ allocation-free pair counts
// sesmd_pair_counts_noalloc keeps its partial sums in `scratch`, a file-scope
// device array that every GPU the library runs on holds until the process
// ends: two counts per block of NOALLOC_BLOCK sites, for at most
// NOALLOC_MAX_BLOCKS blocks (1 MiB; 8,388,608 sites).  Every slot a call reads
// is written by the same call first, so nothing needs clearing.  Host code
// reaches the array only through the address hipGetSymbolAddress gives on
// each device (scratch_setup), never through its name.
constexpr int NOALLOC_BLOCK = 128;
constexpr int NOALLOC_MAX_BLOCKS = 65536;
static __device__ __align__(256) unsigned long long scratch[2 * NOALLOC_MAX_BLOCKS];

__global__ __launch_bounds__(NOALLOC_BLOCK) void k_pair_count_part(Geo g, const double4 *__restrict__ spos,
                                                                   const int4 *__restrict__ smeta,
                                                                   const int *__restrict__ bstart,
                                                                   const int *__restrict__ rkind,
                                                                   const sesmd_forcefield *__restrict__ ff,
                                                                   unsigned long long *__restrict__ part)
{
    __shared__ unsigned long long s[2][NOALLOC_BLOCK];
    int t = threadIdx.x, k = blockIdx.x * NOALLOC_BLOCK + t;
    unsigned long long np = 0, nl = 0;
    if (k < g.n) site_pair_count(g, spos, smeta, bstart, rkind, ff, k, np, nl);
    s[0][t] = np;
    s[1][t] = nl;
    __syncthreads();
    for (int d = NOALLOC_BLOCK / 2; d > 0; d >>= 1) {
        if (t < d) {
            s[0][t] += s[0][t + d];
            s[1][t] += s[1][t + d];
        }
        __syncthreads();
    }
    if (t == 0) {
        part[2 * blockIdx.x] = s[0][0];
        part[2 * blockIdx.x + 1] = s[1][0];
    }
}

// one block: the block counts summed, halved (each pair was counted from both
// sides) into out[0] (pairs) and out[1] (LJ pairs)
__global__ __launch_bounds__(256) void k_pair_count_total(const unsigned long long *__restrict__ part, int nblocks,
                                                          double *__restrict__ out)
{
    __shared__ unsigned long long s[2][256];
    int t = threadIdx.x;
    unsigned long long a = 0, b = 0;
    for (int i = t; i < nblocks; i += 256) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    s[0][t] = a;
    s[1][t] = b;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (t < d) {
            s[0][t] += s[0][t + d];
            s[1][t] += s[1][t + d];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = 0.5 * (double)s[0][0];
        out[1] = 0.5 * (double)s[1][0];
    }
}

// ------------------------------------------------------------ far field

// P2M, src/computeOnm4.f90:142-162: O_n^m = sum q rho^n Y_n^{-m}(alpha, beta)
// about the box centre over the box's receivers; one block per box, one
// thread per site, deterministic tree reduction in LDS.
template <int PP>
__device__ __forceinline__ void p2m_blk(int blk,
                                          const Geo &g,
                                          const double *__restrict__ r,
                                          const int *__restrict__ perm,
                                          const int *__restrict__ bstart,
                                          const int *__restrict__ rkind,
                                          const int *__restrict__ type,
                                          const sesmd_forcefield *__restrict__ ff,
                                          double2 *__restrict__ O4,
                                          int *__restrict__ ne4)
{
    constexpr int p = PP, lk = (PP + 1) * (PP + 2) / 2, lp = (PP + 1) * (PP + 1);
    // sized by the order: 15 KB at p = 4, so every box's block is resident at once
    __shared__ double s_re[64][lk], s_im[64][lk];
    __shared__ double s_fac[lk];  // sqrt((n - m)! / (n + m)!) per row n (n + 1) / 2 + m
    int b = blk;
    int nb = g.nb;
    int s0 = bstart[b], s1 = bstart[b + 1];
    if (s0 == s1) {  // an empty box: zero expansion (block-uniform exit)
        for (int c = threadIdx.x; c < lp; c += 64) O4[(size_t)b * lp + c] = make_double2(0.0, 0.0);
        if (threadIdx.x == 0) ne4[b] = 0;
        return;
    }
    for (int c = threadIdx.x; c < lk; c += 64) {
        int n = 0;
        while ((n + 1) * (n + 2) / 2 <= c) n++;
        int m = c - n * (n + 1) / 2;
        s_fac[c] = sqrt(ff->rifact[n - m] / ff->rifact[n + m]);
    }
    __syncthreads();
    double cx = d_cen4(g, b % nb), cy = d_cen4(g, (b / nb) % nb), cz = d_cen4(g, b / (nb * nb));
    double are[MAXLK], aim[MAXLK];
#pragma unroll
    for (int c = 0; c < lk; c++) are[c] = aim[c] = 0.0;
    int any = 0;
    for (int s = s0 + threadIdx.x; s < s1; s += 64) {
        int i = perm[s];
        if (rkind[i] < 0) continue;
        any = 1;
        double dx = r[i] - cx, dy = r[g.n + i] - cy, dz = r[2 * g.n + i] - cz;
        double rho = sqrt(dx * dx + dy * dy + dz * dz);
        // cos(alpha) = dz / rho; e^{-i beta} = (dx, -dy) / rho_xy, its powers by
        // multiplication (e^{-i 0} where rho_xy = 0, as atan2(0, 0) = 0 gives)
        double ca = dz / rho;
        double rxy = sqrt(dx * dx + dy * dy);
        double2 e1 = rxy > 0.0 ? make_double2(dx / rxy, -dy / rxy) : make_double2(1.0, 0.0);
        double2 em[MAXP + 1];
        em[0] = make_double2(1.0, 0.0);
#pragma unroll
        for (int m = 1; m <= p; m++) em[m] = make_double2(em[m - 1].x * e1.x - em[m - 1].y * e1.y,
                                                          em[m - 1].x * e1.y + em[m - 1].y * e1.x);
        double q = ff->q[type[i]];
        // P_n^m(cos alpha) by the recurrence of rLegendreP, columns m = 0..p
        double P[MAXP + 1][MAXP + 1];
        double somx2 = sqrt((1.0 - ca) * (1.0 + ca));
#pragma unroll
        for (int m = 0; m <= p; m++) {
            double pmm = 1.0, fact = 1.0;
#pragma unroll
            for (int i2 = 1; i2 <= m; i2++) {
                pmm = -pmm * fact * somx2;
                fact = fact + 2.0;
            }
            P[m][m] = pmm;
            if (m + 1 <= p) {
                double pmmp1 = ca * (2 * m + 1) * pmm;
                P[m + 1][m] = pmmp1;
                double pm2 = pmm;
#pragma unroll
                for (int ll = m + 2; ll <= p; ll++) {
                    // 1 / (ll - m) as a constant: a product where rLegendreP divides
                    double pll = (ca * (2 * ll - 1) * pmmp1 - (ll + m - 1) * pm2) * (1.0 / (ll - m));
                    pm2 = pmmp1;
                    pmmp1 = pll;
                    P[ll][m] = pll;
                }
            }
        }
        double rhon = 1.0;
#pragma unroll
        for (int n = 0; n <= p; n++) {
#pragma unroll
            for (int m = 0; m <= n; m++) {
                // Y_n^{-m}: phase exp(-i m beta) rounded to single precision
                float cr = (float)em[m].x, ci = (float)em[m].y;
                double sc = s_fac[n * (n + 1) / 2 + m] * P[n][m];
                double w = q * rhon;
                int c = n * (n + 1) / 2 + m;
                are[c] += w * (sc * (double)cr);
                aim[c] += w * (sc * (double)ci);
            }
            rhon *= rho;
        }
    }
#pragma unroll
    for (int c = 0; c < lk; c++) {
        s_re[threadIdx.x][c] = are[c];
        s_im[threadIdx.x][c] = aim[c];
    }
    any = __syncthreads_or(any);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
#pragma unroll
            for (int c = 0; c < lk; c++) {
                s_re[threadIdx.x][c] += s_re[threadIdx.x + off][c];
                s_im[threadIdx.x][c] += s_im[threadIdx.x + off][c];
            }
        __syncthreads();
    }
#pragma unroll
    for (int c = threadIdx.x; c < lp; c += 64) {
        int n = (int)floor(sqrt((double)c) + 1e-9);
        int m = c - n * (n + 1);
        int am = m < 0 ? -m : m;
        int rrow = n * (n + 1) / 2 + am;
        double re = s_re[0][rrow], im = s_im[0][rrow];
        O4[(size_t)b * lp + c] = make_double2(re, m < 0 ? -im : im);
    }
    if (threadIdx.x == 0) ne4[b] = any;
}

template <int PP>
__global__ __launch_bounds__(64) void k_p2m(Geo g, const double *__restrict__ r,
                                            const int *__restrict__ perm,
                                            const int *__restrict__ bstart,
                                            const int *__restrict__ rkind,
                                            const int *__restrict__ type,
                                            const sesmd_forcefield *__restrict__ ff,
                                            double2 *__restrict__ O4, int *__restrict__ ne4)
{
    p2m_blk<PP>(blockIdx.x, g, r, perm, bstart, rkind, type, ff, O4, ne4);
}

// complex (re, im) helpers
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Dense contraction out_jk (+)= sum_c Op[jk][c] in_c for k >= 0 rows, then the
// k < 0 half by conjugation.  One block of 64 threads per target; lane jk
// (< lk) accumulates; used by M2M (8 children), M2L (interaction list) and L2L.
// acc += sum_c op(row, c) in[c], c in order; the loads of a chunk of CH terms
// are issued together (one memory round trip per chunk, not per term).
template <int CH = 8>
__device__ void contract_rows(const Geo &g, const double2 *__restrict__ op, const double2 *__restrict__ in,
                              double2 &acc, int row)
{
    int lp = g.lp, lk = g.lk;
    const double2 *o = op + row;  // column-major: element (row, c) at c * lk + row
    for (int c0 = 0; c0 < lp; c0 += CH) {
        double2 ov[CH], iv[CH];
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (c0 + k < lp) {
                ov[k] = o[(size_t)(c0 + k) * lk];
                iv[k] = in[c0 + k];
            }
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (c0 + k < lp) {
                double2 t = cmul(ov[k], iv[k]);
                acc.x += t.x;
                acc.y += t.y;
            }
    }
}

__device__ __forceinline__ void store_rows(const Geo &g, double2 *out, int row, double2 acc)
{
    // row -> (j, k)
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= row) j++;
    int k = row - j * (j + 1) / 2;
    out[j * (j + 1) + k] = acc;
    if (k > 0) out[j * (j + 1) - k] = make_double2(acc.x, -acc.y);
}

// M2M (computeOnm3 / computeOnm2): parent = sum over 8 children of M[child] O_child
// M2M: the eight children of a parent contracted side by side (64 lanes
// each), then summed in child order -- deterministic.
__device__ __forceinline__ void m2m_blk(int blk,
                                          const Geo &g,
                                          int nd_parent,
                                          const double2 *__restrict__ Oc,
                                          const int *__restrict__ nec,
                                          const double2 *__restrict__ M,
                                          double2 *__restrict__ Op,
                                          int *__restrict__ nep)
{
    __shared__ double2 s_part[8][64];
    int P = blk;
    int kc = threadIdx.x >> 6, row = threadIdx.x & 63;
    int px = P % nd_parent, py = (P / nd_parent) % nd_parent, pz = P / (nd_parent * nd_parent);
    int ndc = 2 * nd_parent;
    int C = (2 * px + (kc & 1)) + ndc * ((2 * py + ((kc >> 1) & 1)) + ndc * (2 * pz + (kc >> 2)));
    double2 acc = make_double2(0.0, 0.0);
    if (nec[C] && row < g.lk) contract_rows(g, M + (size_t)kc * g.lk * g.lp, Oc + (size_t)C * g.lp, acc, row);
    s_part[kc][row] = acc;
    __syncthreads();
    if (kc != 0) return;
    if (row < g.lk) {
        double2 sum = s_part[0][row];
        for (int q = 1; q < 8; q++) {
            sum.x += s_part[q][row].x;
            sum.y += s_part[q][row].y;
        }
        store_rows(g, Op + (size_t)P * g.lp, row, sum);
    }
    if (row == 0) {
        int any = 0;
        for (int q = 0; q < 8; q++) {
            int Cq = (2 * px + (q & 1)) + ndc * ((2 * py + ((q >> 1) & 1)) + ndc * (2 * pz + (q >> 2)));
            any |= nec[Cq];
        }
        nep[P] = any;
    }
}

__global__ __launch_bounds__(512) void k_m2m(Geo g, int nd_parent, const double2 *__restrict__ Oc,
                                             const int *__restrict__ nec, const double2 *__restrict__ M,
                                             double2 *__restrict__ Op, int *__restrict__ nep)
{
    m2m_blk(blockIdx.x, g, nd_parent, Oc, nec, M, Op, nep);
}

// M2L over the interaction list (computeLjk4 / computeLjk3): sources are the
// children of the parent's neighbours that are not neighbours of the target.
// accumulate = the table already holds the L2L part (shiftLjk zeroes, then
// computeLjk accumulates).
// The targets of one parity class (t mod 2 per axis) share their candidate
// source offsets: d in [-2 - p, 3 - p]^3 per axis (the children of the
// parent's neighbours), less the 27 near neighbours.  A wave holds TPW = 64 / R
// such targets, a 2 x 2 tile in (x, y) when R = 16, R lanes each (one per
// k >= 0 row): at every offset the lanes of all its targets read the same
// operator, so one load instruction serves them all.  The W waves of a block
// take the 216 candidate offsets in turn, and their partial sums are added in
// wave order: deterministic.
template <int R, int W>
__global__ __launch_bounds__(64 * W) void k_m2l(Geo g, int nd, const double2 *__restrict__ O,
                                               const int *__restrict__ ne, const double2 *__restrict__ T,
                                               double2 *__restrict__ L)
{
    constexpr int TPW = 64 / R, TX = TPW == 4 ? 2 : 1, TY = TPW == 4 ? 2 : 1;
    __shared__ double2 s_acc[W][64];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int row = lane % R, j = lane / R;
    int h = nd / 2, qx = (h + TX - 1) / TX, qy = (h + TY - 1) / TY;
    int b = blockIdx.x;
    int par = b & 7;
    b >>= 3;
    int bx = b % qx, by = (b / qx) % qy, hz = b / (qx * qy);
    int px = par & 1, py = (par >> 1) & 1, pz = par >> 2;
    int hx = bx * TX + j % TX, hy = by * TY + j / TX;
    bool valid = hx < h && hy < h && row < g.lk;
    int tx = 2 * hx + px, ty = 2 * hy + py, tz = 2 * hz + pz;
    double2 acc = make_double2(0.0, 0.0);
    for (int i = w; i < 216; i += W) {
        int dx = i % 6 - 2 - px, dy = (i / 6) % 6 - 2 - py, dz = i / 36 - 2 - pz;
        if (iabs_d(dx) <= 1 && iabs_d(dy) <= 1 && iabs_d(dz) <= 1) continue;  // wave-uniform
        int sx = tx + dx, sy = ty + dy, sz = tz + dz;
        if (!valid || sx < 0 || sx >= nd || sy < 0 || sy >= nd || sz < 0 || sz >= nd) continue;
        int S = sx + nd * (sy + nd * sz);
        if (!ne[S]) continue;
        int off = (dx + 3) + 7 * ((dy + 3) + 7 * (dz + 3));
        contract_rows(g, T + (size_t)off * g.lk * g.lp, O + (size_t)S * g.lp, acc, row);
    }
    s_acc[w][lane] = acc;
    __syncthreads();
    if (w != 0 || !valid) return;
    acc = s_acc[0][lane];
    for (int q = 1; q < W; q++) {
        acc.x += s_acc[q][lane].x;
        acc.y += s_acc[q][lane].y;
    }
    // add onto the L2L part (row k >= 0), rebuild the k < 0 half
    int jj = 0;
    while ((jj + 1) * (jj + 2) / 2 <= row) jj++;
    int k = row - jj * (jj + 1) / 2;
    double2 *Lt = L + (size_t)(tx + nd * (ty + nd * tz)) * g.lp;
    double2 cur = Lt[jj * (jj + 1) + k];
    cur.x += acc.x;
    cur.y += acc.y;
    Lt[jj * (jj + 1) + k] = cur;
    if (k > 0) Lt[jj * (jj + 1) - k] = make_double2(cur.x, -cur.y);
}

// M2L on the matrix cores (computeLjk3 / computeLjk4).  Every target of one
// parity class per axis has the same 216 candidate offsets (the children of
// its parent's neighbours; 189 of them are not near), and for one offset the
// translation is the same operator for every target:
//   [L_re; L_im] += [[T_re, -T_im], [T_im, T_re]] . [O_re; O_im]
// (2 lk output rows, 2 lp input columns, re / im interleaved).  So the work of
// a level is 8 x 216 matrix products, one per (class, offset), whose columns
// are the targets of the class that have a source with charge at that offset
// -- only those: a droplet fills few of the boxes, and a column per target of
// the class would mostly multiply zeros.
//
// Per rebuild, from the occupied boxes (ne): k_m2l_count counts each list's
// targets; k_m2l_scan places the lists grouped by their offset (the up to
// eight classes that share an operator side by side) and gives each group its
// first position and first block; k_m2l_fill writes the lists (targets in box
// order) and, per target and candidate, the position of its result (-1:
// none).  A GEMM block takes up to M2L_TPB x 16 consecutive entries of one
// group, so one operator load serves the targets of every class with that
// offset.  FP64 (k_m2l_gemm): the operator slab is the A operand of
// v_mfma_f64_16x16x4f64 (2 lk rows padded to MT x 16, 2 lk columns -- the
// half-space inputs of m2l_half_off -- padded to KS x 4), the 16 entries of a
// tile the N = 16 columns, the sources' expansions the B operand.  Mixed precision (k_m2l_gemm_bf3): the same
// product on v_mfma_f32_16x16x32_bf16, each FP64 operand rounded to FP32 and
// split into three bf16 parts (hi + mid + lo) whose six leading cross
// products are accumulated in FP32.  Each result column goes to its own row
// of `res`; k_m2l_accum adds a target's results in candidate order, in FP64,
// so the sums do not depend on scheduling.
constexpr int M2L_TPB = 8;        // tiles of 16 list entries per block
constexpr int M2L_LISTS = 8 * 216;
constexpr int M2L_GROUPS = 343;   // operator offsets in [-3, 3]^3
constexpr int M2L_RES = 96;       // values per result row (2 lk <= 90, padded)
typedef double m2l_d4 __attribute__((ext_vector_type(4)));
typedef float m2l_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 m2l_bf8 __attribute__((ext_vector_type(8)));
typedef unsigned m2l_u4 __attribute__((ext_vector_type(4)));  // eight bf16 bit patterns

// The products run over half of each source expansion.  Every expansion here
// holds O(n, -m) = conj(O(n, m)) (P2M and M2M write the m < 0 entries so), so
// L(j, k) = sum_nm T(jk, nm) O(n, m) is a real-linear map of the 2 lk real
// values Re, Im O(n, m), m >= 0: real input 2 c + {0, 1} is Re / Im of O(n, m)
// with c = n (n + 1) / 2 + m, at double offset 2 (n (n + 1) + m) + {0, 1} of
// the box's expansion.
__host__ __device__ __forceinline__ int m2l_half_off(int kk)
{
    int c = kk >> 1;
    int n = 0;
    while ((n + 1) * (n + 2) / 2 <= c) n++;
    return 2 * (n * (n + 1) + c - n * (n + 1) / 2) + (kk & 1);
}

// bf16 bit pattern of f, rounded to nearest even (finite f)
__host__ __device__ __forceinline__ unsigned short bf16_bits(float f)
{
    unsigned u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__host__ __device__ __forceinline__ float bf16_value(unsigned short h)
{
    unsigned u = (unsigned)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// x rounded to FP32 and split as hi + mid + lo (bf16 each); every remainder
// is exact in FP32, so the three parts carry the 24 bits of the FP32 value
__host__ __device__ __forceinline__ void bf16_split3(double x, unsigned short &h, unsigned short &m,
                                                     unsigned short &l)
{
    float f = (float)x;
    h = bf16_bits(f);
    float r = f - bf16_value(h);
    m = bf16_bits(r);
    l = bf16_bits(r - bf16_value(m));
}

// operator offset (group) of candidate ci of class par; the class's candidate
// for group grp, -1 if it has none
__device__ __forceinline__ int m2l_cand(int par, int grp)
{
    int cx = grp % 7 - 3 + 2 + (par & 1), cy = (grp / 7) % 7 - 3 + 2 + ((par >> 1) & 1),
        cz = grp / 49 - 3 + 2 + (par >> 2);
    if (cx < 0 || cx > 5 || cy < 0 || cy > 5 || cz < 0 || cz > 5) return -1;
    return cx + 6 * (cy + 6 * cz);
}

// candidate ci (0..215, x fastest over the 6^3 children of the parent's
// neighbours) of parity class par -> offset; false for the 27 near ones
__device__ __forceinline__ bool m2l_offset(int par, int ci, int &dx, int &dy, int &dz)
{
    dx = ci % 6 - 2 - (par & 1);
    dy = (ci / 6) % 6 - 2 - ((par >> 1) & 1);
    dz = ci / 36 - 2 - (par >> 2);
    return !(iabs_d(dx) <= 1 && iabs_d(dy) <= 1 && iabs_d(dz) <= 1);
}

// target q (0..h^3 - 1) of class par -> box coordinates; source box for the
// offset, or -1 (outside the grid / without charge), or -2 (with `tne`: a
// target whose flag there is 0, which nothing translates into)
__device__ __forceinline__ int m2l_src(int nd, const int *__restrict__ ne, const int *__restrict__ tne, int par,
                                       int q, int dx, int dy, int dz, int *tb)
{
    int h = nd / 2;
    int tx = 2 * (q % h) + (par & 1), ty = 2 * ((q / h) % h) + ((par >> 1) & 1), tz = 2 * (q / (h * h)) + (par >> 2);
    *tb = tx + nd * (ty + nd * tz);
    int sx = tx + dx, sy = ty + dy, sz = tz + dz;
    if (sx < 0 || sx >= nd || sy < 0 || sy >= nd || sz < 0 || sz >= nd) return -1;
    if (tne && !tne[*tb]) return -2;
    int si = sx + nd * (sy + nd * sz);
    return ne[si] ? si : -1;
}

// One block per list (class, candidate): its target count.  tne (or null):
// translate only into targets flagged there.
__device__ __forceinline__ void m2l_count_blk(int blk,
                                                int nd,
                                                const int *__restrict__ ne,
                                                const int *__restrict__ tne,
                                                int *__restrict__ cnt)
{
    int L = blk, par = L / 216, ci = L % 216, h = nd / 2, h3 = h * h * h;
    int dx, dy, dz, tb;
    bool far = m2l_offset(par, ci, dx, dy, dz);
    __shared__ int s_c;
    if (threadIdx.x == 0) s_c = 0;
    __syncthreads();
    int c = 0;
    if (far)
        for (int q = threadIdx.x; q < h3; q += blockDim.x) c += m2l_src(nd, ne, tne, par, q, dx, dy, dz, &tb) >= 0;
    if (c) atomicAdd(&s_c, c);  // integer: the total does not depend on order
    __syncthreads();
    if (threadIdx.x == 0) cnt[L] = s_c;
}

__global__ __launch_bounds__(256) void k_m2l_count(int nd, const int *__restrict__ ne, const int *__restrict__ tne,
                                                   int *__restrict__ cnt)
{
    m2l_count_blk(blockIdx.x, nd, ne, tne, cnt);
}

// One block: the lists grouped by operator offset, classes in order within a
// group.  base[list] = its first position; gbase[grp] = the group's first
// position, bbase[grp] its first block (gbase / bbase[M2L_GROUPS] = entries /
// blocks of the level).
__device__ __forceinline__ void m2l_scan_blk(const int *__restrict__ cnt,
                                               int *__restrict__ base,
                                               int *__restrict__ gbase,
                                               int *__restrict__ bbase)
{
    __shared__ int s_a[512], s_b[512];
    int t = threadIdx.x, tot = 0;
    if (t < M2L_GROUPS)
        for (int par = 0; par < 8; par++) {
            int ci = m2l_cand(par, t);
            if (ci >= 0) tot += cnt[par * 216 + ci];
        }
    int nblk = (tot + M2L_TPB * 16 - 1) / (M2L_TPB * 16);
    s_a[t] = tot;
    s_b[t] = nblk;
    __syncthreads();
    for (int d = 1; d < 512; d *= 2) {  // inclusive scan
        int va = t >= d ? s_a[t - d] : 0, vb = t >= d ? s_b[t - d] : 0;
        __syncthreads();
        s_a[t] += va;
        s_b[t] += vb;
        __syncthreads();
    }
    if (t < M2L_GROUPS) {
        int run = s_a[t] - tot;
        gbase[t] = run;
        bbase[t] = s_b[t] - nblk;
        for (int par = 0; par < 8; par++) {
            int ci = m2l_cand(par, t);
            if (ci < 0) continue;
            base[par * 216 + ci] = run;
            run += cnt[par * 216 + ci];
        }
    }
    if (t == M2L_GROUPS - 1) {
        gbase[M2L_GROUPS] = s_a[t];
        bbase[M2L_GROUPS] = s_b[t];
    }
}

__global__ __launch_bounds__(512) void k_m2l_scan(const int *__restrict__ cnt, int *__restrict__ base,
                                                  int *__restrict__ gbase, int *__restrict__ bbase)
{
    m2l_scan_blk(cnt, base, gbase, bbase);
}

// the group of GEMM block B (the last group whose first block is <= B), its
// first entry and entry count; false past the level's last block
__device__ __forceinline__ bool m2l_block(const int *__restrict__ gbase, const int *__restrict__ bbase, int B,
                                          int &grp, int &first, int &ne_)
{
    if (B >= bbase[M2L_GROUPS]) return false;
    int lo = 0, hi = M2L_GROUPS - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (bbase[mid] <= B) lo = mid;
        else hi = mid - 1;
    }
    grp = lo;
    int e0 = (B - bbase[grp]) * M2L_TPB * 16;
    ne_ = min(M2L_TPB * 16, gbase[grp + 1] - gbase[grp] - e0);
    first = gbase[grp] + e0;
    return true;
}

// One block per list: its entries (target q, source box) in q order, and for
// every target of the class the position of its result for this candidate
// (idx[target box][ci], -1 none).
__device__ __forceinline__ void m2l_fill_blk(int blk,
                                               int nd,
                                               const int *__restrict__ ne,
                                               const int *__restrict__ tne,
                                               const int *__restrict__ base,
                                               int2 *__restrict__ list,
                                               int *__restrict__ idx)
{
    int L = blk, par = L / 216, ci = L % 216, h = nd / 2, h3 = h * h * h;
    int dx, dy, dz;
    bool far = m2l_offset(par, ci, dx, dy, dz);
    __shared__ int s_w[4];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int run = base[L];
    for (int q0 = 0; q0 < h3; q0 += 256) {
        int q = q0 + (int)threadIdx.x, tb = -1, si = -1;
        if (q < h3) {
            si = m2l_src(nd, ne, tne, par, q, dx, dy, dz, &tb);
            if (!far && si != -2) si = -1;
        }
        unsigned long long m = __builtin_amdgcn_ballot_w64(si >= 0);
        int before = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (lane == 0) s_w[w] = __builtin_popcountll(m);
        __syncthreads();
        int off = 0, tot = 0;
        for (int k = 0; k < 4; k++) {
            off += k < w ? s_w[k] : 0;
            tot += s_w[k];
        }
        int pos = run + off + before;
        if (q < h3) {
            if (si >= 0) list[pos] = make_int2(q, si);
            // the sums read the result positions of flagged targets only
            if (si != -2) idx[(size_t)tb * 216 + ci] = si >= 0 ? pos : -1;
        }
        run += tot;
        __syncthreads();  // s_w is rewritten by the next round
    }
}

__global__ __launch_bounds__(256) void k_m2l_fill(int nd, const int *__restrict__ ne, const int *__restrict__ tne,
                                                  const int *__restrict__ base, int2 *__restrict__ list,
                                                  int *__restrict__ idx)
{
    m2l_fill_blk(blockIdx.x, nd, ne, tne, base, list, idx);
}

// One block per M2L_TPB x 16 consecutive entries of one list.  MT x WG
// wavefronts: wave w computes the 16-row slab w % MT of the operator times the
// expansions of tiles (w / MT) TPB / WG .. + TPB / WG - 1, so the block
// computes the (MT 16) x (TPB 16) product as slabs side by side.  A wave reads
// only its own slab of the operator, straight from global memory into
// registers (one fragment per k step serves TPB / WG tiles); the expansions --
// the B operand the waves share -- go through LDS in chunks of C k steps,
// double-buffered, loaded two chunks ahead: while chunk s's products run, chunk
// s + 1's data goes into the other buffer and chunk s + 2's loads are in flight.
template <int MT, int KS, int TPB, int WG>
__device__ __forceinline__ void m2l_gemm_blk(int blk,
                                               const Geo &g,
                                               const int *__restrict__ gbase,
                                               const int *__restrict__ bbase,
                                               const int2 *__restrict__ list,
                                               const double *__restrict__ O,
                                               const double *__restrict__ AF,
                                               double *__restrict__ res)
{
    constexpr int NT = 64 * MT * WG;            // threads
    constexpr int TW = TPB / WG;                // tiles per wave
    constexpr int NCH = (KS + 7) / 8;           // chunks
    constexpr int C = (KS + NCH - 1) / NCH;     // k steps per chunk
    constexpr int BCH = C * TPB * 64;           // doubles of one chunk of expansions
    constexpr int PER = (BCH + NT - 1) / NT;    // chunk loads per thread
    static_assert(TPB % WG == 0 && TPB == M2L_TPB, "tiles per wave");
    __shared__ double s_b[2][BCH];
    __shared__ int s_src[TPB * 16];
    int off, first, ne_;  // off: the operator index in [-3, 3]^3
    if (!m2l_block(gbase, bbase, blk, off, first, ne_)) return;
    for (int j = threadIdx.x; j < TPB * 16; j += NT) s_src[j] = j < ne_ ? list[first + j].y : -1;
    __syncthreads();
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int mt = w % MT, tw0 = (w / MT) * TW;
    int lp2 = 2 * g.lp, kh = 2 * g.lk;
    int ntile = (ne_ + 15) / 16;  // tiles with entries (block-uniform)
    m2l_d4 acc[TW];
#pragma unroll
    for (int t = 0; t < TW; t++) acc[t] = (m2l_d4){0.0, 0.0, 0.0, 0.0};
    struct Stage {
        double a[C];    // this wave's operator fragments, one per k step
        double b[PER];  // this thread's share of the chunk's expansions
    };
    Stage P0, P1;
    double acur[C];
    const double *abase = AF + (size_t)off * KS * MT * 64 + mt * 64 + lane;
    // element idx = thread + NT e of a chunk: k step idx / (64 TPB), tile
    // (idx / 64) % TPB, lane idx % 64 -- half-space input 4 k + lane / 16 of
    // the source of entry 16 tile + lane % 16.  Every load is in bounds
    // (missing entries read box 0 and are replaced by 0), so none waits on a
    // branch.
    auto load = [&](int ch, Stage &P) {
        int cs = min(C, KS - ch * C);
        const double *src = abase + (size_t)ch * C * MT * 64;
#pragma unroll
        for (int kc = 0; kc < C; kc++) {
            double v = src[(size_t)kc * MT * 64];  // the operator table carries one spare offset
            P.a[kc] = kc < cs ? v : 0.0;
        }
        int si[PER];
#pragma unroll
        for (int e = 0; e < PER; e++) {
            int idx = (int)threadIdx.x + NT * e;
            si[e] = idx < BCH ? s_src[16 * ((idx / 64) % TPB) + idx % 16] : -1;
        }
#pragma unroll
        for (int e = 0; e < PER; e++) {
            int idx = (int)threadIdx.x + NT * e;
            int kk = 4 * (ch * C + idx / (TPB * 64)) + (idx % 64) / 16;
            bool ok = si[e] >= 0 && kk < kh;
            double v = O[(size_t)max(si[e], 0) * lp2 + m2l_half_off(min(kk, kh - 1))];
            P.b[e] = ok ? v : 0.0;
        }
    };
    auto stage_in = [&](int buf, const Stage &P) {
#pragma unroll
        for (int e = 0; e < PER; e++) {
            int idx = (int)threadIdx.x + NT * e;
            if (idx < BCH) s_b[buf][idx] = P.b[e];
        }
#pragma unroll
        for (int kc = 0; kc < C; kc++) acur[kc] = P.a[kc];
    };
    auto compute = [&](int ch) {
        int cs = min(C, KS - ch * C);
        // the B operands of k step kc + 1 are read from LDS while the
        // products of step kc run (TW independent accumulator chains)
        const double *sb = &s_b[ch & 1][tw0 * 64 + lane];
        double bq[2][TW];
#pragma unroll
        for (int t = 0; t < TW; t++) bq[0][t] = sb[t * 64];
#pragma unroll
        for (int kc = 0; kc < C; kc++) {
            if (kc >= cs) break;
            if (kc + 1 < cs)
#pragma unroll
                for (int t = 0; t < TW; t++) bq[(kc + 1) & 1][t] = sb[((kc + 1) * TPB + t) * 64];
#pragma unroll
            for (int t = 0; t < TW; t++)
                if (tw0 + t < ntile)  // block-uniform
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(acur[kc], bq[kc & 1][t], acc[t], 0, 0, 0);
        }
    };
    auto step = [&](int ch, Stage &ahead, const Stage &ready) {
        if (ch + 2 < NCH) load(ch + 2, ahead);
        compute(ch);
        if (ch + 1 < NCH) stage_in((ch + 1) & 1, ready);
        __syncthreads();
    };
    load(0, P0);
    stage_in(0, P0);
    if (NCH > 1) load(1, P1);
    __syncthreads();
#pragma unroll 1
    for (int ch = 0; ch < NCH; ch += 2) {
        step(ch, P0, P1);
        if (ch + 1 < NCH) step(ch + 1, P1, P0);
    }
    // output element (16 mt + 4 r + lane / 16, lane % 16) of the wave's slab,
    // register r: row 16 mt + 4 r + lane / 16 of entry 16 tile + lane % 16
    int n = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int t = 0; t < TW; t++) {
        int j = 16 * (tw0 + t) + n;
        if (j >= ne_) continue;
        double *dst = res + (size_t)(first + j) * M2L_RES + 16 * mt + kq;
        m2l_d4 v = acc[t];
        dst[0] = v.x;
        dst[4] = v.y;
        dst[8] = v.z;
        dst[12] = v.w;
    }
}

template <int MT, int KS, int TPB, int WG>
__global__ __launch_bounds__(64 * MT * WG) void k_m2l_gemm(Geo g, const int *__restrict__ gbase,
                                                           const int *__restrict__ bbase,
                                                           const int2 *__restrict__ list,
                                                           const double *__restrict__ O,
                                                           const double *__restrict__ AF,
                                                           double *__restrict__ res)
{
    m2l_gemm_blk<MT, KS, TPB, WG>(blockIdx.x, g, gbase, bbase, list, O, AF, res);
}

// Mixed precision: each box's half-space inputs (2 lk values, m2l_half_off)
// as the B-operand parts of k_m2l_gemm_bf3: OB[box][ks][part][32], part 0..2 =
// hi, mid, lo of column 32 ks + c (zero past 2 lk).
template <int KS>
__device__ __forceinline__ void m2l_split_blk(int blk,
                                                int nbox,
                                                int lp2,
                                                int kh,
                                                const double *__restrict__ O,
                                                unsigned short *__restrict__ OB)
{
    int t = blk * blockDim.x + threadIdx.x;
    if (t >= nbox * 32 * KS) return;
    int box = t / (32 * KS), k = t % (32 * KS);
    double x = k < kh ? O[(size_t)box * lp2 + m2l_half_off(k)] : 0.0;
    unsigned short h, m, l;
    bf16_split3(x, h, m, l);
    unsigned short *dst = OB + ((size_t)box * KS + k / 32) * 96 + k % 32;
    dst[0] = h;
    dst[32] = m;
    dst[64] = l;
}

template <int KS>
__global__ __launch_bounds__(256) void k_m2l_split(int nbox, int lp2, int kh, const double *__restrict__ O,
                                                   unsigned short *__restrict__ OB)
{
    m2l_split_blk<KS>(blockIdx.x, nbox, lp2, kh, O, OB);
}

// Mixed-precision M2L GEMM: one block per M2L_TPB x 16 consecutive entries of
// one offset group, WM x WT wavefronts; wave (wm, wt) computes the 16-row
// slabs wm MTW .. + MTW - 1 of the operator against tiles wt TWW .. + TWW - 1,
// each a register-resident FP32 accumulator.  Per 32-wide k step a wave loads
// its slabs' three operator parts (AFB: [group][ks][slab][part][lane] x 8
// bf16, the A-operand layout of v_mfma_f32_16x16x32_bf16: lane l holds row
// l % 16, columns 8 (l / 16) .. + 7) and its tiles' three expansion parts
// (lane l: entry l % 16, columns 8 (l / 16) .. + 7), then runs the six
// products lo x hi, mid x mid, hi x lo, mid x hi, hi x mid, hi x hi (smallest
// first) per (slab, tile).  Output lane l holds rows 4 (l / 16) .. + 3 of
// entry l % 16.
template <int MTW, int WM, int TWW, int WT, int KS>
__device__ __forceinline__ void m2l_gemm_bf3_blk(int blk,
                                                   const int *__restrict__ gbase,
                                                   const int *__restrict__ bbase,
                                                   const int2 *__restrict__ list,
                                                   const m2l_u4 *__restrict__ OB,
                                                   const m2l_u4 *__restrict__ AFB,
                                                   float *__restrict__ res)
{
    constexpr int MT = MTW * WM, TPB = TWW * WT;
    static_assert(TPB == M2L_TPB, "block tiles");
    int grp, first, ne_;
    if (!m2l_block(gbase, bbase, blk, grp, first, ne_)) return;
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w % WM, wt = w / WM;
    int n = lane & 15, kq = lane >> 4;
    int si[TWW];
#pragma unroll
    for (int u = 0; u < TWW; u++) {
        int j = 16 * (wt * TWW + u) + n;
        si[u] = j < ne_ ? list[first + j].y : -1;  // -1: no entry (its column stays zero)
    }
    m2l_f4 acc[MTW][TWW];
#pragma unroll
    for (int i = 0; i < MTW; i++)
#pragma unroll
        for (int u = 0; u < TWW; u++) acc[i][u] = (m2l_f4){0.f, 0.f, 0.f, 0.f};
    const m2l_u4 *ab = AFB + ((size_t)grp * KS * MT + wm * MTW) * 3 * 64 + lane;
    const m2l_u4 zero = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        m2l_u4 a[MTW][3], b[TWW][3];
#pragma unroll
        for (int i = 0; i < MTW; i++)
#pragma unroll
            for (int q = 0; q < 3; q++) a[i][q] = ab[((size_t)(ks * MT + i) * 3 + q) * 64];
#pragma unroll
        for (int u = 0; u < TWW; u++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                m2l_u4 v = OB[(((size_t)max(si[u], 0) * KS + ks) * 3 + q) * 4 + kq];
                b[u][q] = si[u] >= 0 ? v : zero;
            }
#pragma unroll
        for (int i = 0; i < MTW; i++)
#pragma unroll
            for (int u = 0; u < TWW; u++) {
                m2l_bf8 a0 = __builtin_bit_cast(m2l_bf8, a[i][0]), a1 = __builtin_bit_cast(m2l_bf8, a[i][1]),
                        a2 = __builtin_bit_cast(m2l_bf8, a[i][2]);
                m2l_bf8 b0 = __builtin_bit_cast(m2l_bf8, b[u][0]), b1 = __builtin_bit_cast(m2l_bf8, b[u][1]),
                        b2 = __builtin_bit_cast(m2l_bf8, b[u][2]);
                m2l_f4 c = acc[i][u];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, b0, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b2, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, c, 0, 0, 0);
                acc[i][u] = c;
            }
    }
#pragma unroll
    for (int u = 0; u < TWW; u++) {
        int j = 16 * (wt * TWW + u) + n;
        if (j >= ne_) continue;
        float *dst = res + (size_t)(first + j) * M2L_RES + 4 * kq;
#pragma unroll
        for (int i = 0; i < MTW; i++) *(m2l_f4 *)(dst + 16 * (wm * MTW + i)) = acc[i][u];
    }
}

template <int MTW, int WM, int TWW, int WT, int KS>
__global__ __launch_bounds__(64 * WM * WT) void k_m2l_gemm_bf3(const int *__restrict__ gbase,
                                                               const int *__restrict__ bbase,
                                                               const int2 *__restrict__ list,
                                                               const m2l_u4 *__restrict__ OB,
                                                               const m2l_u4 *__restrict__ AFB,
                                                               float *__restrict__ res)
{
    m2l_gemm_bf3_blk<MTW, WM, TWW, WT, KS>(blockIdx.x, gbase, bbase, list, OB, AFB, res);
}

// Adds each target's results in candidate order onto L (the L2L part already
// there): one wavefront per target box, lane r its k >= 0 row r.  The 216
// result positions of the target are loaded at once (four per lane); the
// candidates that have a result are taken eight at a time, their rows loaded
// together and then added in candidate order.
template <typename T>
__device__ __forceinline__ void m2l_accum_blk(int blk,
                                                const Geo &g,
                                                int nd,
                                                const int *__restrict__ tne,
                                                const int *__restrict__ idx,
                                                const T *__restrict__ res,
                                                double2 *__restrict__ L)
{
    int b = blk, row = threadIdx.x, lk = g.lk;
    if (tne && !tne[b]) return;  // no translation was listed into it
    const int *ix = idx + (size_t)b * 216;
    int pv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) pv[u] = 64 * u + row < 216 ? ix[64 * u + row] : -1;
    int rr = min(row, lk - 1);
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        unsigned long long m = __builtin_amdgcn_ballot_w64(pv[u] >= 0);
        while (m) {  // wave-uniform
            int ps[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                ps[q] = -1;
                if (m) {
                    int l = __builtin_ctzll(m);
                    m &= m - 1;
                    ps[q] = __builtin_amdgcn_readlane(pv[u], l);
                }
            }
            T vr[8], vi[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const T *r = res + (size_t)max(ps[q], 0) * M2L_RES + 2 * rr;
                vr[q] = r[0];
                vi[q] = r[1];
            }
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (ps[q] >= 0) {
                    re += (double)vr[q];
                    im += (double)vi[q];
                }
        }
    }
    if (row >= lk) return;
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= row) j++;
    int k = row - j * (j + 1) / 2;
    double2 *Lt = L + (size_t)b * g.lp;
    double2 cur = Lt[j * (j + 1) + k];
    cur.x += re;
    cur.y += im;
    Lt[j * (j + 1) + k] = cur;
    if (k > 0) Lt[j * (j + 1) - k] = make_double2(cur.x, -cur.y);
}

template <typename T>
__global__ __launch_bounds__(64) void k_m2l_accum(Geo g, int nd, const int *__restrict__ tne,
                                                  const int *__restrict__ idx, const T *__restrict__ res,
                                                  double2 *__restrict__ L)
{
    m2l_accum_blk<T>(blockIdx.x, g, nd, tne, idx, res, L);
}

// top level (computeLjk2): every L2 box that is not adjacent to the target
// M2L at the top level (computeLjk2): one block per (target, source) pair of
// L2 boxes, then a per-target sum over sources in index order.
__device__ __forceinline__ void m2l_top_pair_blk(int blk,
                                                   const Geo &g,
                                                   const double2 *__restrict__ O2,
                                                   const int *__restrict__ ne2,
                                                   const double2 *__restrict__ T2,
                                                   double2 *__restrict__ part)
{
    int Tb = blk / g.n2, S = blk % g.n2, row = threadIdx.x;
    int nd = g.nb2;
    if (row >= g.lk) return;
    int tx = Tb % nd, ty = (Tb / nd) % nd, tz = Tb / (nd * nd);
    int sx = S % nd, sy = (S / nd) % nd, sz = S / (nd * nd);
    int dx = sx - tx, dy = sy - ty, dz = sz - tz;
    double2 acc = make_double2(0.0, 0.0);
    bool adjacent = iabs_d(dx) <= 1 && iabs_d(dy) <= 1 && iabs_d(dz) <= 1;
    if (!adjacent && ne2[S]) {
        int w = 2 * nd - 1;
        int off = (dx + nd - 1) + w * ((dy + nd - 1) + w * (dz + nd - 1));
        contract_rows(g, T2 + (size_t)off * g.lk * g.lp, O2 + (size_t)S * g.lp, acc, row);
    }
    part[((size_t)Tb * g.n2 + S) * g.lk + row] = acc;
}

__global__ __launch_bounds__(64) void k_m2l_top_pair(Geo g, const double2 *__restrict__ O2,
                                                     const int *__restrict__ ne2,
                                                     const double2 *__restrict__ T2,
                                                     double2 *__restrict__ part)
{
    m2l_top_pair_blk(blockIdx.x, g, O2, ne2, T2, part);
}

__device__ __forceinline__ void m2l_top_sum_blk(int blk,
                                                  const Geo &g,
                                                  const double2 *__restrict__ part,
                                                  double2 *__restrict__ L2)
{
    int Tb = blk, row = threadIdx.x;
    if (row >= g.lk) return;
    // the sources in index order, the loads of a chunk of 9 issued together
    double2 acc = make_double2(0.0, 0.0);
    for (int S0 = 0; S0 < g.n2; S0 += 9) {
        double2 v[9];
#pragma unroll
        for (int q = 0; q < 9; q++)
            if (S0 + q < g.n2) v[q] = part[((size_t)Tb * g.n2 + S0 + q) * g.lk + row];
#pragma unroll
        for (int q = 0; q < 9; q++)
            if (S0 + q < g.n2) {
                acc.x += v[q].x;
                acc.y += v[q].y;
            }
    }
    store_rows(g, L2 + (size_t)Tb * g.lp, row, acc);
}

__global__ __launch_bounds__(64) void k_m2l_top_sum(Geo g, const double2 *__restrict__ part,
                                                    double2 *__restrict__ L2)
{
    m2l_top_sum_blk(blockIdx.x, g, part, L2);
}

// The same top level with one block per target L2 box (the batched ensemble's
// form: its members' mostly empty domains would spend a (target, source)
// block on every pair), lane r its row r: each non-empty, non-adjacent
// source's product in a fresh accumulator, added in source index order as
// k_m2l_top_sum adds them -- the same result.  (Summing
// the products of the empty and adjacent sources as well -- each +0.0 -- would
// change nothing: the running sum starts at +0.0, and x + (+0.0) == x for
// every x but -0.0, which a sum that starts at +0.0 never reaches.)
__device__ __forceinline__ void m2l_top_blk(int Tb, const Geo &g, const double2 *__restrict__ O2,
                                            const int *__restrict__ ne2, const double2 *__restrict__ T2,
                                            double2 *__restrict__ L2)
{
    int row = threadIdx.x, nd = g.nb2;
    if (row >= g.lk) return;
    int tx = Tb % nd, ty = (Tb / nd) % nd, tz = Tb / (nd * nd);
    double2 acc = make_double2(0.0, 0.0);
    for (int S = 0; S < g.n2; S++) {
        int sx = S % nd, sy = (S / nd) % nd, sz = S / (nd * nd);
        int dx = sx - tx, dy = sy - ty, dz = sz - tz;
        bool adjacent = iabs_d(dx) <= 1 && iabs_d(dy) <= 1 && iabs_d(dz) <= 1;
        if (adjacent || !ne2[S]) continue;  // block-uniform
        int w = 2 * nd - 1;
        int off = (dx + nd - 1) + w * ((dy + nd - 1) + w * (dz + nd - 1));
        double2 v = make_double2(0.0, 0.0);
        contract_rows(g, T2 + (size_t)off * g.lk * g.lp, O2 + (size_t)S * g.lp, v, row);
        acc.x += v.x;
        acc.y += v.y;
    }
    store_rows(g, L2 + (size_t)Tb * g.lp, row, acc);
}


// L2L (shiftLjk2 / shiftLjk3): child table = S[child position] parent
__device__ __forceinline__ void l2l_blk(int blk,
                                          const Geo &g,
                                          int nd_child,
                                          const double2 *__restrict__ Lp,
                                          const double2 *__restrict__ S,
                                          double2 *__restrict__ Lc)
{
    int C = blk, row = threadIdx.x;
    int cx = C % nd_child, cy = (C / nd_child) % nd_child, cz = C / (nd_child * nd_child);
    int ndp = nd_child / 2;
    int P = cx / 2 + ndp * (cy / 2 + ndp * (cz / 2));
    int kc = (cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2);
    if (row >= g.lk) return;
    double2 acc = make_double2(0.0, 0.0);
    contract_rows(g, S + (size_t)kc * g.lk * g.lp, Lp + (size_t)P * g.lp, acc, row);
    store_rows(g, Lc + (size_t)C * g.lp, row, acc);
}

__global__ __launch_bounds__(64) void k_l2l(Geo g, int nd_child, const double2 *__restrict__ Lp,
                                            const double2 *__restrict__ S, double2 *__restrict__ Lc)
{
    l2l_blk(blockIdx.x, g, nd_child, Lp, S, Lc);
}

// L2L into the boxes holding sites only (the occupied list, taken blockIdx.x,
// + gridDim.x, ...): in a step, their sites' far-field terms are the only
// readers of L4
__global__ __launch_bounds__(64) void k_l2l_occupied(Geo g, const int4 *__restrict__ obox,
                                                     const double2 *__restrict__ Lp, const double2 *__restrict__ S,
                                                     double2 *__restrict__ Lc)
{
    for (int q = blockIdx.x; q < g.nb3; q += gridDim.x) {
        int box = obox[q].x;
        if (box < 0) break;  // past the occupied boxes (block-uniform)
        l2l_blk(box, g, g.nb, Lp, S, Lc);
    }
}

// ------------------------------------------------------------ per-site terms

// L2P (src/evalLjk4.f90:38-68, the same formula for every receiver class) and
// site i of a SoA (x block, y block, z block of n) array
__device__ __forceinline__ void ld3(const double *r, int n, int i, double *x)
{
    x[0] = r[i];
    x[1] = r[n + i];
    x[2] = r[2 * n + i];
}

// Legendre table P[l + 1][m] = P_l^m(x) for l = -1..p, m = 0..p (P[0][0] = 1,
// rLegendreP's l = -1 case), by the recurrence of src/rLegendreP.f90:3-48.
template <int p>
__device__ __forceinline__ void legendre_table(double rp1, double (&P)[MAXP + 2][MAXP + 1])
{
#pragma unroll
    for (int l = 0; l <= p + 1; l++)
#pragma unroll
        for (int m = 0; m <= p; m++) P[l][m] = 0.0;
    P[0][0] = 1.0;  // l = -1
    double somx2 = sqrt((1.0 - rp1) * (1.0 + rp1));
#pragma unroll
    for (int m = 0; m <= p; m++) {
        double pmm = 1.0, fact = 1.0;
#pragma unroll
        for (int i2 = 1; i2 <= m; i2++) {
            pmm = -pmm * fact * somx2;
            fact = fact + 2.0;
        }
        P[m + 1][m] = pmm;
        if (m + 1 <= p) {
            double pmmp1 = rp1 * (2 * m + 1) * pmm;
            P[m + 2][m] = pmmp1;
            double pm2 = pmm;
#pragma unroll
            for (int ll = m + 2; ll <= p; ll++) {
                // 1 / (ll - m) as a constant (the loops are unrolled): a product
                // where rLegendreP divides, within an ulp of the quotient
                double pll = (rp1 * (2 * ll - 1) * pmmp1 - (ll + m - 1) * pm2) * (1.0 / (ll - m));
                pm2 = pmmp1;
                pmmp1 = pll;
                P[ll + 1][m] = pll;
            }
        }
    }
}

// Per-site terms, and the first write of f: f = (near field: the nine row
// partials of k_near_rows, in row order) + L2P (src/evalLjk4.f90) + the 1/x^13
// walls (src/eval_wall.f90:149-188) + graphite.  The pair kernels before it
// leave their results in `part`; the kernels after it add to f.
//
// One block per box holding sites (the rebuild's list obox; block q takes
// entries q, q + grid, ...), one thread per site of the box (sorted order, so
// the reads of the sorted copies are coalesced).  The box's local expansion is
// staged in LDS by one round of loads, so every site's L2P reads its (p + 1)^2
// coefficients at LDS latency; a site-indexed thread layout, whose lanes may
// belong to different boxes, has to fetch them from memory in the middle of the
// arithmetic (21 dependent rounds at p = 8, which dominated it).  The rfctr
// factors go to LDS beside it.  Launching a block for every box instead spent
// the launch on empty ones (most of a droplet's domain; up to 32^3 boxes for
// an ensemble's largest members).
constexpr int SITE_TPB = 192;
constexpr int SITE_GRID = 2048;  // blocks at most (the headline droplet: 716 boxes hold sites)
template <int PP>
__device__ __forceinline__ void site_terms_boxes(int blk0, int nblk, const Geo &g, unsigned what, const int4 *__restrict__ obox,
                             const double2 *__restrict__ L4,
                             const double *__restrict__ lfac, const sesmd_forcefield *__restrict__ ff,
                             const double *__restrict__ part, const double4 *__restrict__ spos,
                             const int4 *__restrict__ smeta, double *__restrict__ f,
                             const double *__restrict__ gr, const int *__restrict__ gstart,
                             const int *__restrict__ gidx, int ngr)
{
    // ngr > 0: the graphite term too (the step pipeline's form)
    constexpr int p = PP, lp = (PP + 1) * (PP + 1);
    __shared__ double2 sL[lp];
    __shared__ double sF[lp];
    const int nb = g.nb;
    for (int q = blk0; q < g.nb3; q += nblk) {
        const int4 e = obox[q];
        if (e.x < 0) break;  // past the boxes holding sites (uniform over the block)
        const int b = e.x, k0 = e.y, k1 = e.z;
        const bool far = (what & SESMD_FAR) != 0;
        if (far) {
            for (int t = threadIdx.x; t < lp; t += SITE_TPB) {
                sL[t] = L4[(size_t)b * lp + t];
                sF[t] = lfac[t];
            }
            __syncthreads();
        }
        const int bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
        const double cx = d_cen4(g, bx), cy = d_cen4(g, by), cz = d_cen4(g, bz);
        for (int k = k0 + threadIdx.x; k < k1; k += SITE_TPB) {
            // the LDS coefficients are read inside each site's sum, not hoisted out of
            // this loop into hundreds of live registers per lane
            asm volatile("" ::: "memory");
            int4 me = smeta[k];
            int i = me.x, rk = ((me.y >> 11) & 7) - 1;
            double4 pk = spos[k];
            double x = pk.x, y = pk.y, z = pk.z;
            double nx = 0.0, ny = 0.0, nz = 0.0;
            if (what & SESMD_NEAR) {
                // the slots of boxes outside the grid hold stale values: only
                // the (row (dz, dy), box dx) slots of neighbour boxes inside the
                // grid are added, rows in order, each row's boxes in x order
                if (rk >= 0) {
#pragma unroll 1
                    for (int q = 0; q < NEAR_ROWS; q++) {
                        int yy = by + q % 3 - 1, zz = bz + q / 3 - 1;
                        if (yy < 0 || yy >= nb || zz < 0 || zz >= nb) continue;
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            int xx = bx + j - 1;
                            if (xx < 0 || xx >= nb) continue;
                            size_t o = (size_t)(3 * (3 * q + j)) * g.n + k;
                            nx += part[o];
                            ny += part[o + g.n];
                            nz += part[o + 2 * g.n];
                        }
                    }
                }
            }
            double fx = 0.0, fy = 0.0, fz = 0.0;
            if (far && rk >= 0) {
                double rxpb = x - cx, rypb = y - cy, rzpb = z - cz;
                double rqq = pk.w * g.xpi;
                double rho2 = rxpb * rxpb + rypb * rypb;
                double rr = sqrt(rho2 + rzpb * rzpb);
                double rp1 = rzpb / rr;
                double sr = sqrt(rho2), rsr = 1.0 / sr;
                double2 rp2 = make_double2(rxpb * rsr, rypb * rsr);
                double rp3 = 1.0 / rho2;
                // the inverse of e^{i phi}: its conjugate over |e|^2
                double rd = 1.0 / (rp2.x * rp2.x + rp2.y * rp2.y);
                double2 rp2i = make_double2(rp2.x * rd, -rp2.y * rd);
                double rrt[MAXP + 1];  // rr^(j-1)
                rrt[0] = 1.0 / rr;
#pragma unroll
                for (int j = 1; j <= p; j++) rrt[j] = rrt[j - 1] * rr;
                // The sum runs over m = |k| outermost: column m of the associated
                // Legendre functions comes from the recurrence of rLegendreP
                // (src/rLegendreP.f90:3-48) as j rises, in the same operations as
                // legendre_table, and e^{+-i m phi} by one more product per column,
                // so no table of either is held.
                //
                // Each term of src/evalLjk4.f90 adds rqq rp3 Re(rq0 (rq3 + rq4)) to
                // fx, with rq0 = sc L e^{ik phi}, sc = rfctr(j, |k|) rr^(j-1),
                // rq3 = x z rq1, rq4 = (-j x rr rq2, k y rr rq2); expanded, that is
                // x U - y V with U = Re(rq0) (z rq1 - j rr rq2), V = k rr rq2 Im(rq0);
                // fy gets y U + x V and fz -rqq Re(rq0) rq1.  So three sums carry the
                // whole expansion, and the terms k = +m and -m, which share rq1, rq2
                // and sc, enter them together: U and the fz sum through
                // Re(L+ e+) + Re(L- e-), V through m (Im(L+ e+) - Im(L- e-)).
                double2 ep = make_double2(1.0, 0.0), em = ep;  // e^{i m phi}, e^{-i m phi}
                double somx2 = sqrt((1.0 - rp1) * (1.0 + rp1));
                double pmm = 1.0;
                double su = 0.0, sv = 0.0, sw = 0.0;
#pragma unroll
                for (int m = 0; m <= p; m++) {
                    if (m > 0) {
                        pmm = -pmm * (double)(2 * m - 1) * somx2;
                        ep = cmul(ep, rp2);
                        em = cmul(em, rp2i);
                    }
                    const double mrr = m * rr;
                    // P(m-1, m) = 0; at m = 0 the l = -1 entry, 1, which meets j + m = 0
                    double plm1 = m == 0 ? 1.0 : 0.0, pl = pmm;
#pragma unroll
                    for (int j = m; j <= p; j++) {
                        double rq1 = (j + m) * plm1;  // (j + |k|) P(j-1, |k|)
                        double rq2 = pl;              // P(j, |k|)
                        double sc = sF[j * (j + 1) + m] * rrt[j];
                        double t = rzpb * rq1 - (j * rr) * rq2;
                        double a;
                        if (m == 0) {
                            a = sc * sL[j * (j + 1)].x;
                        } else {
                            double2 lp = sL[j * (j + 1) + m], lm = sL[j * (j + 1) - m];
                            double re = lp.x * ep.x - lp.y * ep.y + lm.x * em.x - lm.y * em.y;
                            double im = lp.x * ep.y + lp.y * ep.x - lm.x * em.y - lm.y * em.x;
                            a = sc * re;
                            sv = fma(sc * im, mrr * rq2, sv);
                        }
                        su = fma(a, t, su);
                        sw = fma(a, rq1, sw);
                        if (j < p) {  // P(j + 1, m)
                            double pn = j == m ? rp1 * (2 * m + 1) * pmm
                                               : (rp1 * (2 * (j + 1) - 1) * pl - (j + m) * plm1) *
                                                     (1.0 / (j + 1 - m));
                            plm1 = pl;
                            pl = pn;
                        }
                    }
                }
                double ax = rqq * rp3 * (rxpb * su - rypb * sv);
                double ay = rqq * rp3 * (rypb * su + rxpb * sv);
                double az = -rqq * sw;
                fx += ax;
                fy += ay;
                fz += az;
            }
            if (what & SESMD_WALL) {
                double h = g.rhs;
                auto p13 = [](double u) {
                    double u2 = u * u, u4 = u2 * u2, u8 = u4 * u4;
                    return u8 * u4 * u;
                };
                fx += 48.0 / p13(x - h) + 48.0 / p13(h + x);
                fy += 48.0 / p13(y - h) + 48.0 / p13(h + y);
                fz += 48.0 / p13(z - h) + 48.0 / p13(h + z);
            }
            // graphite (src/eval_graphite.f90:13-132): receivers filed in their own box
            // of the bottom layer, against that column's carbons
            if (ngr > 0 && (what & SESMD_GRAPHITE)) {
                if (b < g.nb * g.nb && (rk == 0 || rk == 1 || rk == 3)) {
                    int t = me.y & 255;
                    double c = ff->gr_c[t], s6 = ff->gr_s6[t];
                    if (c != 0.0) {
                        // the column's carbons eight at a time: their loads are issued
                        // together (a column holds ~800; one at a time, each a pair of
                        // dependent loads, the walk is latency-bound), the terms are
                        // added in list order
                        double gx = 0.0, gy = 0.0, gz = 0.0;
                        constexpr int U = 8;
                        int q0 = gstart[b], q1 = gstart[b + 1];
                        for (int q = q0; q < q1; q += U) {
                            int a[U];
#pragma unroll
                            for (int u = 0; u < U; u++) a[u] = gidx[min(q + u, q1 - 1)];
                            double cxyz[U][3];
#pragma unroll
                            for (int u = 0; u < U; u++)
#pragma unroll
                                for (int d = 0; d < 3; d++) cxyz[u][d] = gr[3 * a[u] + d];
#pragma unroll
                            for (int u = 0; u < U; u++) {
                                double xij = x - cxyz[u][0], yij = y - cxyz[u][1], zij = z - cxyz[u][2];
                                double rijsq = xij * xij + yij * yij + zij * zij;
                                if (q + u < q1 && sqrt(rijsq) <= g.rc) {
                                    double rsi = 1.0 / rijsq, r6 = rsi * rsi * rsi;
                                    double flj = c * r6 * (s6 * r6 - 0.5) * rsi;
                                    gx += xij * flj;
                                    gy += yij * flj;
                                    gz += zij * flj;
                                }
                            }
                        }
                        fx += gx;
                        fy += gy;
                        fz += gz;
                    }
                }
            }

            f[i] = nx + fx;
            f[g.n + i] = ny + fy;
            f[2 * g.n + i] = nz + fz;
        }
        __syncthreads();  // before the next box's expansion replaces this one
    }
}

// the boxes blockIdx.x, + gridDim.x, ... of the list
template <int PP>
__global__ __launch_bounds__(SITE_TPB) void k_site_terms(Geo g, unsigned what, const int4 *__restrict__ obox,
                                                         const double2 *__restrict__ L4,
                                                         const double *__restrict__ lfac,
                                                         const sesmd_forcefield *__restrict__ ff,
                                                         const double *__restrict__ part,
                                                         const double4 *__restrict__ spos,
                                                         const int4 *__restrict__ smeta, double *__restrict__ f,
                                                         const double *__restrict__ gr,
                                                         const int *__restrict__ gstart,
                                                         const int *__restrict__ gidx, int ngr)
{
    site_terms_boxes<PP>(blockIdx.x, gridDim.x, g, what, obox, L4, lfac, ff, part, spos, smeta, f, gr, gstart,
                         gidx, ngr);
}

// ------------------------------------------------------------ intramolecular


// one dihedral's forces on its four sites (src/eval_dihedral.f90:116-271).
// FAST (the step pipeline): sin 2t / sin t = 2 cos t and sin 3t / sin t =
// 4 cos^2 t - 1 replace the trigonometric calls, and the divisions by |a| |b|,
// |a|^2, |b|^2 become products with reciprocals (a few ulp from the reference's
// own evaluation, which loses up to ~1e-11 of r1 itself near trans, t -> pi).
template <bool FAST>
__device__ void dihedral_forces(const sesmd_forcefield *ff, int d, const double *Ri,
                                const double *Rj, const double *Rk, const double *Rl, double F[4][3])
{
    double rxji = Rj[0] - Ri[0], ryji = Rj[1] - Ri[1], rzji = Rj[2] - Ri[2];
    double rxkj = Rk[0] - Rj[0], rykj = Rk[1] - Rj[1], rzkj = Rk[2] - Rj[2];
    double rxnk = Rl[0] - Rk[0], rynk = Rl[1] - Rk[1], rznk = Rl[2] - Rk[2];
    double ax = rxji * rykj - ryji * rxkj, ay = rxji * rzkj - rzji * rxkj, az = ryji * rzkj - rzji * rykj;
    double bx = rxkj * rynk - rykj * rxnk, by = rxkj * rznk - rzkj * rxnk, bz = rykj * rznk - rzkj * rynk;
    double na = sqrt(ax * ax + ay * ay + az * az), nbn = sqrt(bx * bx + by * by + bz * bz);
    double rna = FAST ? 1.0 / na : 0.0, rnb = FAST ? 1.0 / nbn : 0.0;
    double arg = FAST ? (ax * bx + ay * by + az * bz) * rna * rnb : (ax * bx + ay * by + az * bz) / na / nbn;
    double clamp = d < 18 ? ff->clamp_opls : ff->clamp_be;
    if (arg <= -1.0) arg = -clamp;
    if (arg >= 1.0) arg = clamp;
    double costh, r1;
    if (FAST) {
        costh = arg;
        double s2 = 2.0 * arg, s3 = 4.0 * arg * arg - 1.0;  // sin 2t / sin t, sin 3t / sin t
        r1 = d < 18 ? -ff->dva[d] + 2.0 * ff->dvb[d] * s2 - 3.0 * ff->dvc[d] * s3 : -3.0 * ff->dva[d] * s3;
    } else {
        double th = acos(arg), sinth = sin(th);
        costh = cos(th);
        r1 = d < 18 ? -ff->dva[d] + 2.0 * ff->dvb[d] * sin(2.0 * th) / sinth -
                          3.0 * ff->dvc[d] * sin(3.0 * th) / sinth
                    : -3.0 * ff->dva[d] * sin(3.0 * th) / sinth;
    }
    double c1 = na * na, c2 = nbn * nbn, ab = na * nbn;
    double rab = rna * rnb, rc1 = rna * rna, rc2 = rnb * rnb;
    auto dab = [&](double x) { return FAST ? x * rab : x / ab; };
    auto dc1 = [&](double x) { return FAST ? x * rc1 : x / c1; };
    auto dc2 = [&](double x) { return FAST ? x * rc2 : x / c2; };
    // site i
    double fix = dab((rxkj * (-rykj * rynk - rzkj * rznk) + rxnk * (rykj * rykj + rzkj * rzkj)));
    double fiy = dab((rykj * (-rxkj * rxnk - rzkj * rznk) + rynk * (rxkj * rxkj + rzkj * rzkj)));
    double fiz = dab((rzkj * (-rxkj * rxnk - rykj * rynk) + rznk * (rxkj * rxkj + rykj * rykj)));
    double fix2 = dc1((2.0 * rxji * (-rykj * rykj - rzkj * rzkj) + 2.0 * rxkj * (ryji * rykj + rzji * rzkj)));
    double fiy2 = dc1((2.0 * ryji * (-rxkj * rxkj - rzkj * rzkj) + 2.0 * rykj * (rxji * rxkj + rzji * rzkj)));
    double fiz2 = dc1((2.0 * rzji * (-rxkj * rxkj - rykj * rykj) + 2.0 * rzkj * (rxji * rxkj + ryji * rykj)));
    F[0][0] = r1 * (fix - 0.5 * costh * fix2);
    F[0][1] = r1 * (fiy - 0.5 * costh * fiy2);
    F[0][2] = r1 * (fiz - 0.5 * costh * fiz2);
    // site j
    double fjx = rxji * (-rykj * rynk - rzkj * rznk) + rxkj * (rykj * rynk + rzkj * rznk);
    double fjy = ryji * (-rxkj * rxnk - rzkj * rznk) + rykj * (rxkj * rxnk + rzkj * rznk);
    double fjz = rzji * (-rxkj * rxnk - rykj * rynk) + rzkj * (rxkj * rxnk + rykj * rynk);
    fjx = dab((fjx + rxnk * (-ryji * rykj - rzji * rzkj - rykj * rykj - rzkj * rzkj) +
           2.0 * rxkj * (ryji * rynk + rzji * rznk)));
    fjy = dab((fjy + rynk * (-rxji * rxkj - rzji * rzkj - rxkj * rxkj - rzkj * rzkj) +
           2.0 * rykj * (rxji * rxnk + rzji * rznk)));
    fjz = dab((fjz + rznk * (-rxji * rxkj - ryji * rykj - rxkj * rxkj - rykj * rykj) +
           2.0 * rzkj * (rxji * rxnk + ryji * rynk)));
    double fjx3 = dc1((2.0 * rxji * (rykj * rykj + rzkj * rzkj + ryji * rykj + rzji * rzkj) +
                   2.0 * rxkj * (-ryji * ryji - rzji * rzji - ryji * rykj - rzji * rzkj)));
    double fjy3 = dc1((2.0 * ryji * (rxkj * rxkj + rzkj * rzkj + rxji * rxkj + rzji * rzkj) +
                   2.0 * rykj * (-rxji * rxji - rzji * rzji - rxji * rxkj - rzji * rzkj)));
    double fjz3 = dc1((2.0 * rzji * (rxkj * rxkj + rykj * rykj + rxji * rxkj + ryji * rykj) +
                   2.0 * rzkj * (-rxji * rxji - ryji * ryji - rxji * rxkj - ryji * rykj)));
    double fjx5 = dc2((2.0 * rxnk * (rykj * rynk + rzkj * rznk) + 2.0 * rxkj * (-rynk * rynk - rznk * rznk)));
    double fjy5 = dc2((2.0 * rynk * (rxkj * rxnk + rzkj * rznk) + 2.0 * rykj * (-rxnk * rxnk - rznk * rznk)));
    double fjz5 = dc2((2.0 * rznk * (rxkj * rxnk + rykj * rynk) + 2.0 * rzkj * (-rxnk * rxnk - rynk * rynk)));
    F[1][0] = r1 * (fjx - 0.5 * costh * (fjx3 + fjx5));
    F[1][1] = r1 * (fjy - 0.5 * costh * (fjy3 + fjy5));
    F[1][2] = r1 * (fjz - 0.5 * costh * (fjz3 + fjz5));
    // site k
    double fkx = dab((rxji * (rykj * rykj + rzkj * rzkj + rykj * rynk + rzkj * rznk) +
                  rxkj * (-ryji * rykj - rzji * rzkj) + rxnk * (ryji * rykj + rzji * rzkj) +
                  2.0 * rxkj * (-ryji * rynk - rzji * rznk)));
    double fky = dab((ryji * (rxkj * rxkj + rzkj * rzkj + rxkj * rxnk + rzkj * rznk) +
                  rykj * (-rxji * rxkj - rzji * rzkj) + rynk * (rxji * rxkj + rzji * rzkj) +
                  2.0 * rykj * (-rxji * rxnk - rzji * rznk)));
    double fkz = dab((rzji * (rxkj * rxkj + rykj * rykj + rxkj * rxnk + rykj * rynk) +
                  rzkj * (-rxji * rxkj - ryji * rykj) + rznk * (rxji * rxkj + ryji * rykj) +
                  2.0 * rzkj * (-rxji * rxnk - ryji * rynk)));
    double fkx4 = dc1((2.0 * rxji * (-ryji * rykj - rzji * rzkj) + 2.0 * rxkj * (ryji * ryji + rzji * rzji)));
    double fky4 = dc1((2.0 * ryji * (-rxji * rxkj - rzji * rzkj) + 2.0 * rykj * (rxji * rxji + rzji * rzji)));
    double fkz4 = dc1((2.0 * rzji * (-rxji * rxkj - ryji * rykj) + 2.0 * rzkj * (rxji * rxji + ryji * ryji)));
    double fkx5 = dc2((2.0 * rxnk * (-rykj * rykj - rzkj * rzkj - rykj * rynk - rzkj * rznk) +
                   2.0 * rxkj * (rynk * rynk + rznk * rznk + rykj * rynk + rzkj * rznk)));
    double fky5 = dc2((2.0 * rynk * (-rxkj * rxkj - rzkj * rzkj - rxkj * rxnk - rzkj * rznk) +
                   2.0 * rykj * (rxnk * rxnk + rznk * rznk + rxkj * rxnk + rzkj * rznk)));
    double fkz5 = dc2((2.0 * rznk * (-rxkj * rxkj - rykj * rykj - rxkj * rxnk - rykj * rynk) +
                   2.0 * rzkj * (rxnk * rxnk + rynk * rynk + rxkj * rxnk + rykj * rynk)));
    F[2][0] = r1 * (fkx - 0.5 * costh * (fkx4 + fkx5));
    F[2][1] = r1 * (fky - 0.5 * costh * (fky4 + fky5));
    F[2][2] = r1 * (fkz - 0.5 * costh * (fkz4 + fkz5));
    // site l
    double fnx = dab((rxji * (-rykj * rykj - rzkj * rzkj) + rxkj * (ryji * rykj + rzji * rzkj)));
    double fny = dab((ryji * (-rxkj * rxkj - rzkj * rzkj) + rykj * (rxji * rxkj + rzji * rzkj)));
    double fnz = dab((rzji * (-rxkj * rxkj - rykj * rykj) + rzkj * (rxji * rxkj + ryji * rykj)));
    double fnx2 = dc2((2.0 * rxnk * (rykj * rykj + rzkj * rzkj) + 2.0 * rxkj * (-rykj * rynk - rzkj * rznk)));
    double fny2 = dc2((2.0 * rynk * (rxkj * rxkj + rzkj * rzkj) + 2.0 * rykj * (-rxkj * rxnk - rzkj * rznk)));
    double fnz2 = dc2((2.0 * rznk * (rxkj * rxkj + rykj * rykj) + 2.0 * rzkj * (-rxkj * rxnk - rykj * rynk)));
    F[3][0] = r1 * (fnx - 0.5 * costh * fnx2);
    F[3][1] = r1 * (fny - 0.5 * costh * fny2);
    F[3][2] = r1 * (fnz - 0.5 * costh * fnz2);
}

// Intramolecular surfactant terms, one thread per surfactant site; the thread
// evaluates every term its site takes part in (from per-site term lists), so
// each site's force is written by exactly one thread:
//   eval_surfactant_internal (src/eval_surfactant_internal.f90:12-110),
//   eval_subtract_fma (src/eval_subtract_fma.f90:10-54),
//   eval_valence (src/eval_valence.f90:97-158),
//   eval_dihedral / eval_dihedral_berendsen (src/eval_dihedral.f90:102-295).
// fs: when non-null, the site's total is stored there (x, y, z blocks of 29 ns
// surfactant sites) instead of being added to f -- the step pipeline keeps the
// intramolecular forces apart so this kernel can run beside the near field.
__global__ __launch_bounds__(64) void k_intra(Geo g, unsigned what, const double *__restrict__ r,
                        const unsigned *__restrict__ isub, const sesmd_forcefield *__restrict__ ff,
                        const Topo *__restrict__ topo, double *__restrict__ f,
                        double *__restrict__ fs)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 29 * g.ns) return;
    int s = t / 29, j = t % 29;
    int base = 3 * g.nw + 29 * s;
    int n = g.n;
    double xs[3];
    ld3(r, n, base + j, xs);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (what & SESMD_INTERNAL) {
        for (int q = 0; q < topo->npair[j]; q++) {
            int k = topo->pair[j][q];
            double xk[3];
            ld3(r, n, base + k, xk);
            double x = xs[0] - xk[0], y = xs[1] - xk[1], z = xs[2] - xk[2];
            double rsq = x * x + y * y + z * z;
            double rctr = sqrt(rsq);
            int tj = 3 + j, tk = 3 + k;
            double fel = ff->q[tj] * ff->q[tk] * g.xpi / (rctr * rctr * rctr);
            double rsi = 1.0 / rsq, r6 = rsi * rsi * rsi;
            double flj = ff->lj_c[tj * 32 + tk] * r6 * (ff->lj_s6[tj * 32 + tk] * r6 - 0.5) * rsi;
            double ft = fel + flj;
            fx += x * ft;
            fy += y * ft;
            fz += z * ft;
        }
    }
    if (what & SESMD_SUBTRACT) {
        unsigned mask = isub[base + j];
        while (mask) {
            int k = __ffs(mask) - 1;
            mask &= mask - 1;
            double xk[3];
            ld3(r, n, base + k, xk);
            double xmn = xs[0] - xk[0], ymn = xs[1] - xk[1], zmn = xs[2] - xk[2];
            double rmn = sqrt(xmn * xmn + ymn * ymn + zmn * zmn);
            double fel = ff->q[3 + j] * ff->q[3 + k] * g.xpi / (rmn * rmn * rmn);
            fx -= xmn * fel;
            fy -= ymn * fel;
            fz -= zmn * fel;
        }
    }
    if (what & SESMD_VALENCE) {
        for (int q = 0; q < topo->nang[j]; q++) {
            int a = topo->ang[j][q] / 3, role = topo->ang[j][q] % 3;
            double ri[3], rj[3], rk[3];
            ld3(r, n, base + ff->ang_i[a] - 1, ri);
            ld3(r, n, base + ff->ang_j[a] - 1, rj);
            ld3(r, n, base + ff->ang_k[a] - 1, rk);
            double rxji = ri[0] - rj[0], ryji = ri[1] - rj[1], rzji = ri[2] - rj[2];
            double rxki = rk[0] - rj[0], ryki = rk[1] - rj[1], rzki = rk[2] - rj[2];
            double dp = rxji * rxki + ryji * ryki + rzji * rzki;
            double rjin = sqrt(rxji * rxji + ryji * ryji + rzji * rzji);
            double rkin = sqrt(rxki * rxki + ryki * ryki + rzki * rzki);
            double th = acos(dp / rjin / rkin);
            double costh = cos(th);
            double r1 = ff->rkth[a] * (th - ff->th[a]) / sin(th);
            double fjx = rxki / rjin / rkin - costh * rxji / (rjin * rjin);
            double fjy = ryki / rjin / rkin - costh * ryji / (rjin * rjin);
            double fjz = rzki / rjin / rkin - costh * rzji / (rjin * rjin);
            double fkx = rxji / rjin / rkin - costh * rxki / (rkin * rkin);
            double fky = ryji / rjin / rkin - costh * ryki / (rkin * rkin);
            double fkz = rzji / rjin / rkin - costh * rzki / (rkin * rkin);
            if (role == 0) {
                fx += r1 * fjx; fy += r1 * fjy; fz += r1 * fjz;
            } else if (role == 1) {
                fx -= r1 * (fjx + fkx); fy -= r1 * (fjy + fky); fz -= r1 * (fjz + fkz);
            } else {
                fx += r1 * fkx; fy += r1 * fky; fz += r1 * fkz;
            }
        }
    }
    if (what & (SESMD_DIHEDRAL | SESMD_DIHEDRAL_BE)) {
        for (int q = 0; q < topo->ndih[j]; q++) {
            int d = topo->dih[j][q] / 4, role = topo->dih[j][q] % 4;
            if (d < 18 && !(what & SESMD_DIHEDRAL)) continue;
            if (d >= 18 && !(what & SESMD_DIHEDRAL_BE)) continue;
            double Ri[3], Rj[3], Rk[3], Rl[3];
            ld3(r, n, base + ff->dih_i[d] - 1, Ri);
            ld3(r, n, base + ff->dih_j[d] - 1, Rj);
            ld3(r, n, base + ff->dih_k[d] - 1, Rk);
            ld3(r, n, base + ff->dih_l[d] - 1, Rl);
            double F[4][3];
            dihedral_forces<false>(ff, d, Ri, Rj, Rk, Rl, F);
            // select the site's row without indexing F at run time (keeps F in registers)
            double gx = role == 0 ? F[0][0] : role == 1 ? F[1][0] : role == 2 ? F[2][0] : F[3][0];
            double gy = role == 0 ? F[0][1] : role == 1 ? F[1][1] : role == 2 ? F[2][1] : F[3][1];
            double gz = role == 0 ? F[0][2] : role == 1 ? F[1][2] : role == 2 ? F[2][2] : F[3][2];
            fx += gx;
            fy += gy;
            fz += gz;
        }
    }
    if (fs) {
        int ns29 = 29 * g.ns;
        fs[t] = fx;
        fs[ns29 + t] = fy;
        fs[2 * ns29 + t] = fz;
    } else {
        f[base + j] += fx;
        f[n + base + j] += fy;
        f[2 * n + base + j] += fz;
    }
}

// The same intramolecular terms one thread per term, for the step pipeline:
// every internal pair, angle and dihedral of molecule s is evaluated once, and
// each of its sites' contributions goes to its own place in the molecule's
// table (Topo::dst) -- mol[(j * INTRA_MAXC + q) * 3 + d] -- which is then
// added per site in order (no atomics).  Terms outside `what` write zeros.
__device__ __forceinline__ void intra_term(int term, const Geo &g, unsigned what, const double *pm,
                                           const sesmd_forcefield *__restrict__ ff,
                                           const Topo *__restrict__ topo, double *mol)
{
    // pm: the molecule's positions in LDS, pm[d * 29 + site]
    int ni = ff->n_internal;
    auto ld = [&](int j, double *x) {
        x[0] = pm[j];
        x[1] = pm[29 + j];
        x[2] = pm[58 + j];
    };
    auto put = [&](int slot, double x, double y, double z) {
        double *o = mol + (size_t)topo->dst[slot] * 3;
        o[0] = x;
        o[1] = y;
        o[2] = z;
    };
    if (term < ni) {
        int j = ff->int_j[term] - 1, k = ff->int_k[term] - 1;
        double xs[3], xk[3];
        ld(j, xs);
        ld(k, xk);
        double x = xs[0] - xk[0], y = xs[1] - xk[1], z = xs[2] - xk[2];
        double ft = 0.0;
        if (what & SESMD_INTERNAL) {
            double rsq = x * x + y * y + z * z;
            double rctr = sqrt(rsq);
            int tj = 3 + j, tk = 3 + k;
            double fel = ff->q[tj] * ff->q[tk] * g.xpi / (rctr * rctr * rctr);
            double rsi = 1.0 / rsq, r6 = rsi * rsi * rsi;
            double flj = ff->lj_c[tj * 32 + tk] * r6 * (ff->lj_s6[tj * 32 + tk] * r6 - 0.5) * rsi;
            ft = fel + flj;
        }
        put(2 * term, x * ft, y * ft, z * ft);
        put(2 * term + 1, -(x * ft), -(y * ft), -(z * ft));
    } else if (term < ni + 36) {
        int a = term - ni, sl = 2 * ni + 3 * a;
        double ri[3], rj[3], rk[3];
        ld(ff->ang_i[a] - 1, ri);
        ld(ff->ang_j[a] - 1, rj);
        ld(ff->ang_k[a] - 1, rk);
        double rxji = ri[0] - rj[0], ryji = ri[1] - rj[1], rzji = ri[2] - rj[2];
        double rxki = rk[0] - rj[0], ryki = rk[1] - rj[1], rzki = rk[2] - rj[2];
        // as k_intra's angle term, with cos t = c and sin t = sqrt((1 - c)(1 + c))
        // for the calls, and reciprocals for the divisions (a few ulp apart)
        double dp = rxji * rxki + ryji * ryki + rzji * rzki;
        double rji = 1.0 / sqrt(rxji * rxji + ryji * ryji + rzji * rzji);
        double rki = 1.0 / sqrt(rxki * rxki + ryki * ryki + rzki * rzki);
        double rjk = rji * rki;
        double costh = dp * rjk;
        double th = acos(costh);
        double r1 = (what & SESMD_VALENCE)
                        ? ff->rkth[a] * (th - ff->th[a]) / sqrt((1.0 - costh) * (1.0 + costh))
                        : 0.0;
        double cj = costh * rji * rji, ck = costh * rki * rki;
        double fjx = rxki * rjk - cj * rxji, fjy = ryki * rjk - cj * ryji, fjz = rzki * rjk - cj * rzji;
        double fkx = rxji * rjk - ck * rxki, fky = ryji * rjk - ck * ryki, fkz = rzji * rjk - ck * rzki;
        put(sl, r1 * fjx, r1 * fjy, r1 * fjz);
        put(sl + 1, -(r1 * (fjx + fkx)), -(r1 * (fjy + fky)), -(r1 * (fjz + fkz)));
        put(sl + 2, r1 * fkx, r1 * fky, r1 * fkz);
    } else {
        int d = term - ni - 36, sl = 2 * ni + 108 + 4 * d;
        bool on = d < 18 ? (what & SESMD_DIHEDRAL) : (what & SESMD_DIHEDRAL_BE);
        double F[4][3] = {};
        if (on) {
            double Ri[3], Rj[3], Rk[3], Rl[3];
            ld(ff->dih_i[d] - 1, Ri);
            ld(ff->dih_j[d] - 1, Rj);
            ld(ff->dih_k[d] - 1, Rk);
            ld(ff->dih_l[d] - 1, Rl);
            dihedral_forces<true>(ff, d, Ri, Rj, Rk, Rl, F);
        }
#pragma unroll
        for (int role = 0; role < 4; role++) put(sl + role, F[role][0], F[role][1], F[role][2]);
    }
}

// eval_graphite, src/eval_graphite.f90:13-132: frozen-carbon LJ within rc on
// the receivers filed in their own box (W1/W2 oxygens, surfactant sites of
// ring 0) of the bottom L4 layer; one block per bottom-layer box.
__global__ void k_graphite(Geo g, const double *__restrict__ r, const int *__restrict__ perm,
                           const int *__restrict__ bstart, const int *__restrict__ rkind,
                           const int *__restrict__ type, const double *__restrict__ gr,
                           const int *__restrict__ gstart, const int *__restrict__ gidx,
                           const sesmd_forcefield *__restrict__ ff, double *__restrict__ f)
{
    int b = blockIdx.x;  // z = 0 layer: boxes 0 .. nb^2-1
    for (int s = bstart[b] + threadIdx.x; s < bstart[b + 1]; s += blockDim.x) {
        int i = perm[s];
        int rk = rkind[i];
        if (!(rk == 0 || rk == 1 || rk == 3)) continue;
        int t = type[i];
        double c = ff->gr_c[t], s6 = ff->gr_s6[t];
        if (c == 0.0) continue;
        double x = r[i], y = r[g.n + i], z = r[2 * g.n + i];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int q = gstart[b]; q < gstart[b + 1]; q++) {
            int a = gidx[q];
            double xij = x - gr[3 * a], yij = y - gr[3 * a + 1], zij = z - gr[3 * a + 2];
            double rijsq = xij * xij + yij * yij + zij * zij;
            if (sqrt(rijsq) <= g.rc) {
                double rsi = 1.0 / rijsq, r6 = rsi * rsi * rsi;
                double flj = c * r6 * (s6 * r6 - 0.5) * rsi;
                fx += xij * flj;
                fy += yij * flj;
                fz += zij * flj;
            }
        }
        f[i] += fx;
        f[g.n + i] += fy;
        f[2 * g.n + i] += fz;
    }
}

// ------------------------------------------------------------ integrator

// NH partial sums per block: [0..8] before, [9..17] after; groups W1-O, W1-H,
// W2-O, W2-H, S-H, S-O, S-CH2, S-Si, S-CH3 (src/nh_water1.f90:285-313,
// src/nh_surfactant.f90:1-199)
__device__ __constant__ int c_sgrp[29] = {0, 1, 2, 2, 1, 2, 2, 1, 2, 2, 1, 2, 2, 1, 2,
                                          2, 2, 3, 4, 1, 3, 4, 4, 4, 1, 3, 4, 4, 4};

template <int NV>
__device__ void block_sum_store(double (&v)[NV], double *out)
{
    __shared__ double s[256][NV];
    for (int c = 0; c < NV; c++) s[threadIdx.x][c] = v[c];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int c = 0; c < NV; c++) s[threadIdx.x][c] += s[threadIdx.x + off][c];
        __syncthreads();
    }
    if (threadIdx.x < NV) out[threadIdx.x] = s[0][threadIdx.x];
}

// Part A for waters (src/move_w1_a.f90:29-146 / move_w2_a): predictor with
// the NH friction, SHAKE over the three cyclic constraints, kinetic sums before
// and after, force zeroing.  One thread per molecule, 256 per block.
__device__ void nh_update_block(Geo g, const double *__restrict__ partials, int nblocks,
                                const sesmd_forcefield *__restrict__ ff, double *__restrict__ nh,
                                int update_eta, int *__restrict__ counters);

// Sorted-order copies of a site's new position for the next force evaluation
// (what k_gather_pos / k_gather_pos32 would write), from the integrator.
struct SortedOut {
    const int *iperm, *key;
    const double *q;  // charge per type
    double4 *spos;
    float4 *spos32;   // null unless mixed precision
};

__device__ __forceinline__ void put_sorted(const Geo &g, const SortedOut &so, int site, int type,
                                           double x, double y, double z)
{
    if (!so.spos) return;
    int k = so.iperm[site];
    double q = so.q[type];
    so.spos[k] = make_double4(x, y, z, q);
    if (so.spos32) {
        int b = so.key[k] >> 2, nb = g.nb;
        so.spos32[k] = make_float4((float)(x - d_cen4(g, b % nb)), (float)(y - d_cen4(g, (b / nb) % nb)),
                                   (float)(z - d_cen4(g, b / (nb * nb))), (float)q);
    }
}

__device__ __forceinline__ void water_a(int blk, const Geo &g, double *__restrict__ r,
                                                      double *__restrict__ v,
                                                      const double *__restrict__ a,
                                                      double *__restrict__ f,
                                                      const int *__restrict__ cls,
                                                      const sesmd_forcefield *__restrict__ ff,
                                                      const double *__restrict__ nh,
                                                      double *__restrict__ partials,
                                                      int *__restrict__ counters, unsigned clsmask,
                                                      SortedOut so = SortedOut())
{
    int m = blk * blockDim.x + threadIdx.x;
    double acc[NH_VALS];
    for (int c = 0; c < NH_VALS; c++) acc[c] = 0.0;
    int n = g.n;
    if (m < g.nw && ((clsmask >> cls[m]) & 1u)) {
        double escale = 1.0 - nh[0] * g.dt / 2.0;
        int base = 3 * m;
        int grp = cls[m] == 0 ? 0 : 2;
        double rx[3][3], px[3][3], vx[3][3];
        for (int i = 0; i < 3; i++)
            for (int d = 0; d < 3; d++) {
                double rr = r[d * n + base + i], vv = v[d * n + base + i], aa = a[d * n + base + i];
                rx[i][d] = rr;
                px[i][d] = rr + vv * escale + aa;
                vx[i][d] = vv * escale + aa;
            }
        for (int i = 0; i < 3; i++) {
            double vv = v[base + i] * v[base + i] + v[n + base + i] * v[n + base + i] +
                        v[2 * n + base + i] * v[2 * n + base + i];
            acc[grp + (i == 0 ? 0 : 1)] += vv;
        }
        double rm[3] = {1.0 / ff->mass[0], 1.0 / ff->mass[1], 1.0 / ff->mass[2]};
        double tol2 = 2.0 * 0.0001;
        int it = 0;
        bool done = false;
        while (!done && it <= 100) {
            done = true;
            for (int c = 0; c < 3; c++) {
                int ia = c, ib = c + 1 > 2 ? 0 : c + 1;
                double rabsq = ff->dsqw[c];
                double pxab = px[ia][0] - px[ib][0], pyab = px[ia][1] - px[ib][1], pzab = px[ia][2] - px[ib][2];
                double pabsq = pxab * pxab + pyab * pyab + pzab * pzab;
                double diffsq = rabsq - pabsq;
                if (fabs(diffsq) > rabsq * tol2) {
                    double rxab = rx[ia][0] - rx[ib][0], ryab = rx[ia][1] - rx[ib][1], rzab = rx[ia][2] - rx[ib][2];
                    double rpab = rxab * pxab + ryab * pyab + rzab * pzab;
                    double gab = diffsq / (2.0 * (rm[ia] + rm[ib]) * rpab);
                    double dd[3] = {rxab * gab, ryab * gab, rzab * gab};
                    for (int d = 0; d < 3; d++) {
                        px[ia][d] += rm[ia] * dd[d];
                        px[ib][d] -= rm[ib] * dd[d];
                        vx[ia][d] += rm[ia] * dd[d];
                        vx[ib][d] -= rm[ib] * dd[d];
                    }
                    done = false;
                }
            }
            it++;
        }
        if (!done) atomicAdd(&counters[0], 1);
        for (int i = 0; i < 3; i++) {
            for (int d = 0; d < 3; d++) {
                r[d * n + base + i] = px[i][d];
                v[d * n + base + i] = vx[i][d];
                f[d * n + base + i] = 0.0;
            }
            put_sorted(g, so, base + i, i, px[i][0], px[i][1], px[i][2]);
            acc[9 + grp + (i == 0 ? 0 : 1)] +=
                vx[i][0] * vx[i][0] + vx[i][1] * vx[i][1] + vx[i][2] * vx[i][2];
        }
    }
    block_sum_store<NH_VALS>(acc, partials + (size_t)blk * NH_VALS);
}

// Part A for surfactants (src/move_surfactant_a.f90:19-147): 28 bond
// constraints, Gauss-Seidel; one thread per molecule.
__global__ __launch_bounds__(64) void k_move_a_surf(Geo g, double *__restrict__ r,
                                                    double *__restrict__ v,
                                                    const double *__restrict__ a,
                                                    double *__restrict__ f,
                                                    const sesmd_forcefield *__restrict__ ff,
                                                    const double *__restrict__ nh,
                                                    double *__restrict__ partials,
                                                    int *__restrict__ counters)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[NH_VALS];
    for (int c = 0; c < NH_VALS; c++) acc[c] = 0.0;
    int n = g.n;
    if (s < g.ns) {
        double escale = 1.0 - nh[0] * g.dt / 2.0;
        int base = 3 * g.nw + 29 * s;
        double rx[29][3], px[29][3], vx[29][3];
        for (int i = 0; i < 29; i++) {
            double vv = 0.0;
            for (int d = 0; d < 3; d++) {
                double rr = r[d * n + base + i], v0 = v[d * n + base + i], aa = a[d * n + base + i];
                rx[i][d] = rr;
                px[i][d] = rr + v0 * escale + aa;
                vx[i][d] = v0 * escale + aa;
                vv += v0 * v0;
            }
            acc[4 + c_sgrp[i]] += vv;
        }
        double tol2 = 2.0 * 0.0001;
        int it = 0;
        bool done = false;
        while (!done && it <= 100) {
            done = true;
            for (int c = 0; c < 28; c++) {
                int ia = ff->ibond[c] - 1, ib = ff->jbond[c] - 1;
                double rabsq = ff->dsqs[c];
                double pxab = px[ia][0] - px[ib][0], pyab = px[ia][1] - px[ib][1], pzab = px[ia][2] - px[ib][2];
                double pabsq = pxab * pxab + pyab * pyab + pzab * pzab;
                double diffsq = rabsq - pabsq;
                if (fabs(diffsq) > rabsq * tol2) {
                    double rxab = rx[ia][0] - rx[ib][0], ryab = rx[ia][1] - rx[ib][1], rzab = rx[ia][2] - rx[ib][2];
                    double rpab = rxab * pxab + ryab * pyab + rzab * pzab;
                    double rma = 1.0 / ff->mass[3 + ia], rmb = 1.0 / ff->mass[3 + ib];
                    double gab = diffsq / (2.0 * (rma + rmb) * rpab);
                    double dd[3] = {rxab * gab, ryab * gab, rzab * gab};
                    for (int d = 0; d < 3; d++) {
                        px[ia][d] += rma * dd[d];
                        px[ib][d] -= rmb * dd[d];
                        vx[ia][d] += rma * dd[d];
                        vx[ib][d] -= rmb * dd[d];
                    }
                    done = false;
                }
            }
            it++;
        }
        if (!done) atomicAdd(&counters[0], 1);
        for (int i = 0; i < 29; i++) {
            for (int d = 0; d < 3; d++) {
                r[d * n + base + i] = px[i][d];
                v[d * n + base + i] = vx[i][d];
                f[d * n + base + i] = 0.0;
            }
            acc[13 + c_sgrp[i]] += vx[i][0] * vx[i][0] + vx[i][1] * vx[i][1] + vx[i][2] * vx[i][2];
        }
    }
    block_sum_store<NH_VALS>(acc, partials + (size_t)(g.nw > 0 ? (g.nw + 255) / 256 : 0) * NH_VALS +
                                      (size_t)blockIdx.x * NH_VALS);
}

constexpr int SURF_MPB = 4;  // molecules per 256-thread block, one wavefront each

// Pipelined Gauss-Seidel.  Bond c of sweep s runs at step kOff[c] + 4 s, one
// lane per bond.  For any two bonds j < k that share a site,
// 0 < kOff[k] - kOff[j] < 4, so (i) they never run in the same step, (ii) in a
// sweep j runs before k, and (iii) k of sweep s runs before j of sweep s + 1:
// every site receives its corrections in exactly the sequential order of
// src/move_surfactant_a.f90:70-118 (bonds that share no site commute exactly),
// while up to six sweeps are in flight.  A sweep ends at step kEnd + 4 s; if
// none of its bonds corrected, the sequential loop would stop there, and the
// bonds already run for later sweeps saw unchanged positions, so they made no
// correction either.
constexpr int kOff[28] = {0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13,
                          14, 15, 16, 17, 18, 19, 20, 21, 22, 19, 20, 21, 22, 23};
constexpr int kEnd = 23;

constexpr bool bond_schedule_ok()
{
    for (int j = 0; j < 28; j++)
        for (int k = j + 1; k < 28; k++) {
            bool share = kBI[j] == kBI[k] || kBI[j] == kBJ[k] || kBJ[j] == kBI[k] || kBJ[j] == kBJ[k];
            if (share && !(kOff[k] > kOff[j] && kOff[k] - kOff[j] < 4)) return false;
        }
    for (int k = 0; k < 28; k++)
        if (kOff[k] < 0 || kOff[k] > kEnd) return false;
    return true;
}
static_assert(bond_schedule_ok(), "bond schedule breaks the sequential update order");

__device__ __constant__ int c_bi[28] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17,
                                        17, 19, 20, 20, 20, 17, 24, 25, 25, 25};
__device__ __constant__ int c_bj[28] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18,
                                        19, 20, 21, 22, 23, 24, 25, 26, 27, 28};
__device__ __constant__ int c_boff[28] = {0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13,
                                          14, 15, 16, 17, 18, 19, 20, 21, 22, 19, 20, 21, 22, 23};

// The molecule's 29 sites (x, y, z, pad) in LDS, shared by its wavefront's lanes.
struct SurfMolLds {
    double p[SURF_MPB][29][4];  // positions being constrained (A) / velocities (B)
    double r[SURF_MPB][29][4];  // positions that define the bond vectors
};

// Runs the pipelined sweeps for one molecule (wave-uniform call).  `upd`
// applies bond c's correction to P[ia], P[ib] and returns whether it corrected;
// returns whether the iteration converged within `smax` sweeps.
// The loop runs one period of four steps per iteration: with off = 4 st + ph,
// bond c's step T = 4 it + j is its turn when j == ph, in sweep it - st.
template <typename Upd>
__device__ __forceinline__ bool pipelined_sweeps(bool bond, int off, int smax, Upd upd)
{
    static_assert(kEnd % 4 == 3, "the convergence test falls on a period's last step");
    unsigned hist = 0;  // bit (s & 7): this bond corrected in sweep s
    int ph = off & 3, st = off >> 2;
    for (int it = 0;; it++) {
        int s = it - st;
        bool on = bond && s >= 0 && s < smax;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (on && ph == j) {
                bool corr = upd();
                int bit = s & 7;
                hist = (hist & ~(1u << bit)) | ((unsigned)corr << bit);
            }
            // the next step's LDS reads must see this step's writes (one
            // wavefront: its LDS operations complete in order)
            __asm__ volatile("" ::: "memory");
        }
        // step 4 it + 3 = kEnd + 4 se ends sweep se
        int se = it - (kEnd >> 2);
        if (se >= 0) {
            bool any = __ballot(bond && ((hist >> (se & 7)) & 1u)) != 0;
            if (!any) return true;
            if (se + 1 >= smax) return false;
        }
    }
}

// Part A for surfactants (src/move_surfactant_a.f90:19-147), one wavefront per
// molecule: lane i < 29 predicts site i, lane c < 28 owns bond c (old bond
// vector and constants in registers) in the pipelined SHAKE; constraint order,
// tolerance test and position updates as the reference.  The reference applies
// every correction to p and v alike, so v = p - r_old at the end (equal up to
// the rounding of |r| ~ 20 sigma, ~1e-12 of v dt).
__device__ __forceinline__ void surf_a(
    int blk, const Geo &g, double *__restrict__ r, double *__restrict__ v, const double *__restrict__ a,
    double *__restrict__ f, SurfConst sc, const double *__restrict__ nh,
    double *__restrict__ partials, int *__restrict__ counters, SortedOut so = SortedOut())
{
    __shared__ SurfMolLds L;
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int s = blk * SURF_MPB + w;
    int n = g.n, base = 3 * g.nw + 29 * s;
    bool live = s < g.ns;  // wave-uniform
    double acc[NH_VALS];
#pragma unroll
    for (int c = 0; c < NH_VALS; c++) acc[c] = 0.0;
    bool site = live && lane < 29;
    double ro[3] = {0.0, 0.0, 0.0};
    if (site) {
        double escale = 1.0 - nh[0] * g.dt / 2.0;
        double vv = 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double v0 = v[d * n + base + lane], aa = a[d * n + base + lane];
            ro[d] = r[d * n + base + lane];
            L.r[w][lane][d] = ro[d];
            L.p[w][lane][d] = ro[d] + v0 * escale + aa;
            vv += v0 * v0;
        }
        int grp = c_sgrp[lane];
#pragma unroll
        for (int q = 0; q < 5; q++) acc[4 + q] = grp == q ? vv : 0.0;  // no dynamic register index
    }
    __asm__ volatile("" ::: "memory");
    if (live) {
        bool bond = lane < 28;
        int ia = bond ? c_bi[lane] : 0, ib = bond ? c_bj[lane] : 1;
        int off = bond ? c_boff[lane] : 0;
        double rbx = L.r[w][ia][0] - L.r[w][ib][0], rby = L.r[w][ia][1] - L.r[w][ib][1],
               rbz = L.r[w][ia][2] - L.r[w][ib][2];
        int cb = bond ? lane : 0;
        double dsq = sc.dsq[cb], tolc = sc.dsq[cb] * (2.0 * 0.0001), k2 = sc.k2[cb];
        double rma = sc.rm[ia], rmb = sc.rm[ib];
        double *Pa = &L.p[w][ia][0], *Pb = &L.p[w][ib][0];
        bool conv = pipelined_sweeps(bond, off, 101, [&]() {
            double ax = Pa[0], ay = Pa[1], az = Pa[2], bx = Pb[0], by = Pb[1], bz = Pb[2];
            double pxab = ax - bx, pyab = ay - by, pzab = az - bz;
            double pabsq = pxab * pxab + pyab * pyab + pzab * pzab;
            double diffsq = dsq - pabsq;
            bool corr = fabs(diffsq) > tolc;
            if (corr) {
                double rpab = rbx * pxab + rby * pyab + rbz * pzab;
                double gab = diffsq / (k2 * rpab);
                double dx = rbx * gab, dy = rby * gab, dz = rbz * gab;
                Pa[0] = ax + rma * dx;
                Pa[1] = ay + rma * dy;
                Pa[2] = az + rma * dz;
                Pb[0] = bx - rmb * dx;
                Pb[1] = by - rmb * dy;
                Pb[2] = bz - rmb * dz;
            }
            return corr;
        });
        if (!conv && lane == 0) atomicAdd(&counters[0], 1);
    }
    if (site) {
        double pn[3], vv = 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            pn[d] = L.p[w][lane][d];
            double vn = pn[d] - ro[d];
            r[d * n + base + lane] = pn[d];
            v[d * n + base + lane] = vn;
            f[d * n + base + lane] = 0.0;
            vv += vn * vn;
        }
        int grp = c_sgrp[lane];
#pragma unroll
        for (int q = 0; q < 5; q++) acc[13 + q] = grp == q ? vv : 0.0;
        put_sorted(g, so, base + lane, 3 + lane, pn[0], pn[1], pn[2]);
    }
    block_sum_store<NH_VALS>(acc, partials + (size_t)(g.nw > 0 ? (g.nw + 255) / 256 : 0) * NH_VALS +
                                      (size_t)blk * NH_VALS);
}

// Part B for surfactants (src/move_surfactant_b.f90:23-122): lane i < 29 forms
// site i's new acceleration and velocity, lane c < 28 owns bond c in the
// pipelined RATTLE (reference order and tolerance).  The two divisions of the
// correction are folded into one factor per bond (within 2 ulp).
// Intramolecular force on a surfactant site in part B: fs, the per-site
// totals (of k_intra, or of k_nh_intra's terms), then with isub set the
// eval_subtract_fma terms from the molecule's positions in LDS.
struct IntraIn {
    const double *fs = nullptr;
    const unsigned *isub = nullptr;
    const sesmd_forcefield *ff = nullptr;
};

__device__ __forceinline__ void surf_b(
    int blk, const Geo &g, const double *__restrict__ r, double *__restrict__ v, double *__restrict__ a,
    const double *__restrict__ f, IntraIn in, SurfConst sc,
    const double *__restrict__ nh, int *__restrict__ counters)
{
    __shared__ SurfMolLds L;
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int s = blk * SURF_MPB + w;
    if (s >= g.ns) return;  // wave-uniform; no block-wide barrier follows
    int n = g.n, base = 3 * g.nw + 29 * s;
    bool site = lane < 29;
    if (site)
#pragma unroll
        for (int d = 0; d < 3; d++) L.r[w][lane][d] = r[d * n + base + lane];
    __asm__ volatile("" ::: "memory");
    if (site) {
        double fi[3] = {0.0, 0.0, 0.0};
        if (in.fs) {
#pragma unroll
            for (int d = 0; d < 3; d++) fi[d] = in.fs[d * 29 * g.ns + 29 * s + lane];
        }
        if (in.isub) {
            unsigned mask = in.isub[base + lane];
            while (mask) {
                int k = __ffs(mask) - 1;
                mask &= mask - 1;
                double xmn = L.r[w][lane][0] - L.r[w][k][0], ymn = L.r[w][lane][1] - L.r[w][k][1],
                       zmn = L.r[w][lane][2] - L.r[w][k][2];
                double rmn = sqrt(xmn * xmn + ymn * ymn + zmn * zmn);
                double fel = in.ff->q[3 + lane] * in.ff->q[3 + k] * g.xpi / (rmn * rmn * rmn);
                fi[0] -= xmn * fel;
                fi[1] -= ymn * fel;
                fi[2] -= zmn * fel;
            }
        }
        double escale = 2.0 / (2.0 + nh[0] * g.dt);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double fv = f[d * n + base + lane];
            if (in.fs) fv += fi[d];
            double ac = fv * g.dtsq2 / sc.m[lane];
            a[d * n + base + lane] = ac;
            L.p[w][lane][d] = escale * (v[d * n + base + lane] + ac);
        }
    }
    __asm__ volatile("" ::: "memory");
    bool bond = lane < 28;
    int ia = bond ? c_bi[lane] : 0, ib = bond ? c_bj[lane] : 1;
    int off = bond ? c_boff[lane] : 0;
    double rbx = L.r[w][ia][0] - L.r[w][ib][0], rby = L.r[w][ia][1] - L.r[w][ib][1],
           rbz = L.r[w][ia][2] - L.r[w][ib][2];
    int cb = bond ? lane : 0;
    double kc = sc.kr[cb], wa = sc.wa[cb], wb = sc.wb[cb];
    double *Va = &L.p[w][ia][0], *Vb = &L.p[w][ib][0];
    bool conv = pipelined_sweeps(bond, off, 101, [&]() {
        // gab = -(r_ab . v_ab / dt) / ((1/m_a + 1/m_b) d^2)
        double ax = Va[0], ay = Va[1], az = Va[2], bx = Vb[0], by = Vb[1], bz = Vb[2];
        double vxab = ax - bx, vyab = ay - by, vzab = az - bz;
        double gab = (rbx * vxab + rby * vyab + rbz * vzab) * kc;
        bool app = fabs(gab) > 0.0001;
        if (app) {
            double dx = rbx * gab, dy = rby * gab, dz = rbz * gab;
            Va[0] = ax + wa * dx;
            Va[1] = ay + wa * dy;
            Va[2] = az + wa * dz;
            Vb[0] = bx - wb * dx;
            Vb[1] = by - wb * dy;
            Vb[2] = bz - wb * dz;
        }
        return app;
    });
    if (!conv && lane == 0) atomicAdd(&counters[0], 1);
    if (site) {
#pragma unroll
        for (int d = 0; d < 3; d++) v[d * n + base + lane] = L.p[w][lane][d];
    }
}

// kinetic sums -> the eta update of src/move_w2_a.f90:145-157 (single block);
// a step whose kinetic sums are not finite (a NaN or an infinity reached some
// velocity) is counted in counters[1]
__device__ void nh_update_block(Geo g, const double *__restrict__ partials, int nblocks,
                                const sesmd_forcefield *__restrict__ ff, double *__restrict__ nh,
                                int update_eta, int *__restrict__ counters)
{
    double acc[NH_VALS];
    for (int c = 0; c < NH_VALS; c++) acc[c] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int c = 0; c < NH_VALS; c++) acc[c] += partials[(size_t)b * NH_VALS + c];
    __shared__ double tot[NH_VALS];
    block_sum_store<NH_VALS>(acc, tot);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double *m = ff->mass;
        double S[6];
        for (int ph = 0; ph < 2; ph++) {
            const double *t = tot + 9 * ph;
            S[0 + ph] = m[3] * t[4] + m[4] * t[5] + m[5] * t[6] + m[3 + 17] * t[7] + m[3 + 18] * t[8];
            S[2 + ph] = m[0] * t[0] + m[1] * t[1];
            S[4 + ph] = m[0] * t[2] + m[1] * t[3];
        }
        double dt2 = g.dt * g.dt;
        double sum1 = (S[0] + S[2] + S[4]) / dt2;
        double sum2 = (S[1] + S[3] + S[5]) / dt2;
        if (!isfinite(sum1 + sum2)) counters[1] += 1;  // one thread: a plain store
        if (update_eta) {
            double eta_old = nh[0];
            nh[0] = nh[0] + nh[8] * (sum1 + nh[9] + sum2);
            nh[10] += 0.5 * g.dt * (eta_old + nh[0]);  // for the thermostat energy
        }
        for (int c = 0; c < 6; c++) nh[1 + c] = S[c];
        for (int c = 0; c < NH_VALS; c++) nh[16 + c] = tot[c];
    }
}

__global__ __launch_bounds__(256) void k_nh_update(Geo g, const double *__restrict__ partials,
                                                   int nblocks, const sesmd_forcefield *__restrict__ ff,
                                                   double *__restrict__ nh, int update_eta,
                                                   int *__restrict__ counters)
{
    nh_update_block(g, partials, nblocks, ff, nh, update_eta, counters);
}

// Step pipeline: block 0 updates eta from part A's kinetic sums (k_nh_update);
// block 1 + s evaluates molecule s's intramolecular terms, one thread per term,
// into its table in LDS, then adds each site's contributions in table order
// (internal pairs, angles, dihedrals) into fs[d][29 ns]; part B adds the
// eval_subtract_fma terms with the lists of the step's own rebuild (surf_b).
__device__ __forceinline__ void nh_intra_block(int blk, const Geo &g, const double *__restrict__ partials,
                                               int nblocks, const sesmd_forcefield *__restrict__ ff,
                                               double *__restrict__ nh, const double *__restrict__ r,
                                               const Topo *__restrict__ topo, double *__restrict__ fs,
                                               int *__restrict__ counters)
{
    if (blk == 0) {
        nh_update_block(g, partials, nblocks, ff, nh, 1, counters);
        return;
    }
    __shared__ double tab[29 * INTRA_MAXC * 3];
    __shared__ double pm[3 * 29];
    int s = blk - 1, nterm = ff->n_internal + 66;
    if (threadIdx.x < 3 * 29) {
        int d = threadIdx.x / 29, j = threadIdx.x - 29 * d;
        pm[threadIdx.x] = r[(size_t)d * g.n + 3 * g.nw + 29 * s + j];
    }
    __syncthreads();
    // last term first: the dihedrals and angles (the long ones) take the first
    // pass of the block's threads, the internal pairs the second
    for (int t = threadIdx.x; t < nterm; t += blockDim.x)
        intra_term(nterm - 1 - t, g, SESMD_INTERNAL | SESMD_VALENCE | SESMD_DIHEDRAL | SESMD_DIHEDRAL_BE, pm,
                   ff, topo, tab);
    __syncthreads();
    for (int t = threadIdx.x; t < 29 * 3; t += blockDim.x) {
        int j = t / 3, d = t - 3 * j, cnt = topo->cnt[j];
        double acc = 0.0;
        for (int q = 0; q < cnt; q++) acc += tab[(j * INTRA_MAXC + q) * 3 + d];
        fs[(size_t)d * 29 * g.ns + 29 * s + j] = acc;
    }
}

__global__ __launch_bounds__(256) void k_nh_intra(Geo g, const double *__restrict__ partials, int nblocks,
                                                  const sesmd_forcefield *__restrict__ ff,
                                                  double *__restrict__ nh, const double *__restrict__ r,
                                                  const Topo *__restrict__ topo, double *__restrict__ fs,
                                                  int *__restrict__ counters)
{
    nh_intra_block(blockIdx.x, g, partials, nblocks, ff, nh, r, topo, fs, counters);
}

// Part B for waters (src/move_w1_b.f90:23-123): a = F dt^2/2m, velocity with
// the new eta, RATTLE.
__device__ __forceinline__ void water_b(int blk, const Geo &g, const double *__restrict__ r,
                                                      double *__restrict__ v, double *__restrict__ a,
                                                      const double *__restrict__ f,
                                                      const sesmd_forcefield *__restrict__ ff,
                                                      const double *__restrict__ nh,
                                                      int *__restrict__ counters,
                                                      const int *__restrict__ cls, unsigned clsmask)
{
    int m = blk * blockDim.x + threadIdx.x;
    if (m >= g.nw || !((clsmask >> cls[m]) & 1u)) return;
    int n = g.n, base = 3 * m;
    double escale = 2.0 / (2.0 + nh[0] * g.dt);
    double rx[3][3], vx[3][3], rm[3];
    for (int i = 0; i < 3; i++) {
        double mi = ff->mass[i];
        rm[i] = 1.0 / mi;
        for (int d = 0; d < 3; d++) {
            rx[i][d] = r[d * n + base + i];
            double ac = f[d * n + base + i] * g.dtsq2 / mi;
            a[d * n + base + i] = ac;
            vx[i][d] = escale * (v[d * n + base + i] + ac);
        }
    }
    // gab = -(r_ab . v_ab / dt) / ((1/m_a + 1/m_b) d^2): the two divisions
    // folded into one factor per constraint (within 2 ulp), as for surfactants
    double kc[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        int ia = c, ib = c + 1 > 2 ? 0 : c + 1;
        kc[c] = -1.0 / (g.dt * ((rm[ia] + rm[ib]) * ff->dsqw[c]));
    }
    int it = 0;
    bool done = false;
    while (!done && it <= 100) {
        done = true;
        for (int c = 0; c < 3; c++) {
            int ia = c, ib = c + 1 > 2 ? 0 : c + 1;
            double vxab = vx[ia][0] - vx[ib][0], vyab = vx[ia][1] - vx[ib][1], vzab = vx[ia][2] - vx[ib][2];
            double rxab = rx[ia][0] - rx[ib][0], ryab = rx[ia][1] - rx[ib][1], rzab = rx[ia][2] - rx[ib][2];
            double gab = (rxab * vxab + ryab * vyab + rzab * vzab) * kc[c];
            if (fabs(gab) > 0.0001) {
                double dd[3] = {rxab * gab, ryab * gab, rzab * gab};
                for (int d = 0; d < 3; d++) {
                    vx[ia][d] += rm[ia] * dd[d] * g.dt;
                    vx[ib][d] -= rm[ib] * dd[d] * g.dt;
                }
                done = false;
            }
        }
        it++;
    }
    if (!done) atomicAdd(&counters[0], 1);
    for (int i = 0; i < 3; i++)
        for (int d = 0; d < 3; d++) v[d * n + base + i] = vx[i][d];
}

// Kernels over the integrator bodies above.  The fused forms put the
// surfactant blocks (first) and the water blocks of one half-step in a single
// launch so the two run side by side on the device.
__global__ __launch_bounds__(256) void k_move_a_water(Geo g, double *__restrict__ r, double *__restrict__ v,
                                                      const double *__restrict__ a, double *__restrict__ f,
                                                      const int *__restrict__ cls,
                                                      const sesmd_forcefield *__restrict__ ff,
                                                      const double *__restrict__ nh,
                                                      double *__restrict__ partials,
                                                      int *__restrict__ counters, unsigned clsmask)
{
    water_a(blockIdx.x, g, r, v, a, f, cls, ff, nh, partials, counters, clsmask);
}

__global__ __launch_bounds__(256) void k_move_a_surf_fast(Geo g, double *__restrict__ r,
                                                              double *__restrict__ v,
                                                              const double *__restrict__ a,
                                                              double *__restrict__ f, SurfConst sc,
                                                              const double *__restrict__ nh,
                                                              double *__restrict__ partials,
                                                              int *__restrict__ counters)
{
    surf_a(blockIdx.x, g, r, v, a, f, sc, nh, partials, counters);
}

// Part A of a regular step in one launch: surfactant blocks and water blocks,
// the new positions also written in sorted order for the force evaluation.
__global__ __launch_bounds__(256) void k_move_a_fused(Geo g, double *__restrict__ r, double *__restrict__ v,
                                                      const double *__restrict__ a, double *__restrict__ f,
                                                      const int *__restrict__ cls,
                                                      const sesmd_forcefield *__restrict__ ff,
                                                      SurfConst sc, double *__restrict__ nh,
                                                      double *__restrict__ partials,
                                                      int *__restrict__ counters, int nblk_s, SortedOut so)
{
    if ((int)blockIdx.x < nblk_s)
        surf_a(blockIdx.x, g, r, v, a, f, sc, nh, partials, counters, so);
    else
        water_a(blockIdx.x - nblk_s, g, r, v, a, f, cls, ff, nh, partials, counters, 7u, so);
}

__global__ __launch_bounds__(256) void k_move_b_water(Geo g, const double *__restrict__ r,
                                                      double *__restrict__ v, double *__restrict__ a,
                                                      const double *__restrict__ f,
                                                      const sesmd_forcefield *__restrict__ ff,
                                                      const double *__restrict__ nh,
                                                      int *__restrict__ counters,
                                                      const int *__restrict__ cls, unsigned clsmask)
{
    water_b(blockIdx.x, g, r, v, a, f, ff, nh, counters, cls, clsmask);
}

__global__ __launch_bounds__(256) void k_move_b_surf_fast(Geo g, const double *__restrict__ r,
                                                              double *__restrict__ v,
                                                              double *__restrict__ a,
                                                              const double *__restrict__ f,
                                                              const double *__restrict__ fs,
                                                              SurfConst sc,
                                                              const double *__restrict__ nh,
                                                              int *__restrict__ counters)
{
    IntraIn in;
    in.fs = fs;
    surf_b(blockIdx.x, g, r, v, a, f, in, sc, nh, counters);
}

__global__ __launch_bounds__(256) void k_move_b_fused(Geo g, const double *__restrict__ r,
                                                      double *__restrict__ v, double *__restrict__ a,
                                                      const double *__restrict__ f, IntraIn in,
                                                      const sesmd_forcefield *__restrict__ ff,
                                                      SurfConst sc, const double *__restrict__ nh,
                                                      int *__restrict__ counters,
                                                      const int *__restrict__ cls, int nblk_s)
{
    if ((int)blockIdx.x < nblk_s)
        surf_b(blockIdx.x, g, r, v, a, f, in, sc, nh, counters);
    else
        water_b(blockIdx.x - nblk_s, g, r, v, a, f, ff, nh, counters, cls, 7u);
}

// Part B for surfactants (src/move_surfactant_b.f90:23-122)
__global__ __launch_bounds__(64) void k_move_b_surf(Geo g, const double *__restrict__ r,
                                                    double *__restrict__ v, double *__restrict__ a,
                                                    const double *__restrict__ f,
                                                    const double *__restrict__ fs,
                                                    const sesmd_forcefield *__restrict__ ff,
                                                    const double *__restrict__ nh,
                                                    int *__restrict__ counters)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.ns) return;
    int n = g.n, base = 3 * g.nw + 29 * s;
    double escale = 2.0 / (2.0 + nh[0] * g.dt);
    double rx[29][3], vx[29][3];
    for (int i = 0; i < 29; i++) {
        double mi = ff->mass[3 + i];
        for (int d = 0; d < 3; d++) {
            rx[i][d] = r[d * n + base + i];
            double fv = f[d * n + base + i];
            if (fs) fv += fs[d * 29 * g.ns + 29 * s + i];
            double ac = fv * g.dtsq2 / mi;
            a[d * n + base + i] = ac;
            vx[i][d] = escale * (v[d * n + base + i] + ac);
        }
    }
    int it = 0;
    bool done = false;
    while (!done && it <= 100) {
        done = true;
        for (int c = 0; c < 28; c++) {
            int ia = ff->ibond[c] - 1, ib = ff->jbond[c] - 1;
            double rma = 1.0 / ff->mass[3 + ia], rmb = 1.0 / ff->mass[3 + ib];
            double vxab = vx[ia][0] - vx[ib][0], vyab = vx[ia][1] - vx[ib][1], vzab = vx[ia][2] - vx[ib][2];
            double rxab = rx[ia][0] - rx[ib][0], ryab = rx[ia][1] - rx[ib][1], rzab = rx[ia][2] - rx[ib][2];
            double rvab = (rxab * vxab + ryab * vyab + rzab * vzab) / g.dt;
            double gab = -rvab / ((rma + rmb) * ff->dsqs[c]);
            if (fabs(gab) > 0.0001) {
                double dd[3] = {rxab * gab, ryab * gab, rzab * gab};
                for (int d = 0; d < 3; d++) {
                    vx[ia][d] += rma * dd[d] * g.dt;
                    vx[ib][d] -= rmb * dd[d] * g.dt;
                }
                done = false;
            }
        }
        it++;
    }
    if (!done) atomicAdd(&counters[0], 1);
    for (int i = 0; i < 29; i++)
        for (int d = 0; d < 3; d++) v[d * n + base + i] = vx[i][d];
}

// init_accel, src/init_accel.f90:1-49
__global__ void k_init_accel(Geo g, const int *__restrict__ type, const double *__restrict__ f,
                             const sesmd_forcefield *__restrict__ ff, double *__restrict__ a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    double mi = ff->mass[type[i]];
    for (int d = 0; d < 3; d++) a[d * g.n + i] = f[d * g.n + i] * g.dtsq2 / mi;
}

// ------------------------------------------------------------ host side

static inline unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
static int dalloc(T **p, size_t n)
{
    HC(hipMalloc((void **)p, sizeof(T) * (n ? n : 1)));
    HC(hipMemset(*p, 0, sizeof(T) * (n ? n : 1)));
    return 0;
}

#define CK(x)                  \
    do {                       \
        if ((x) != 0) return 1; \
    } while (0)

static void build_topology(const sesmd_forcefield &ff, Topo &t)
{
    memset(&t, 0, sizeof t);
    for (int p = 0; p < ff.n_internal; p++) {
        int j = ff.int_j[p] - 1, k = ff.int_k[p] - 1;
        t.pslot[j][t.npair[j]] = 2 * p;
        t.pslot[k][t.npair[k]] = 2 * p + 1;
        t.pair[j][t.npair[j]++] = k;
        t.pair[k][t.npair[k]++] = j;
    }
    for (int a = 0; a < 36; a++) {
        int s[3] = {ff.ang_i[a] - 1, ff.ang_j[a] - 1, ff.ang_k[a] - 1};
        for (int role = 0; role < 3; role++) t.ang[s[role]][t.nang[s[role]]++] = 3 * a + role;
    }
    for (int d = 0; d < 30; d++) {
        int s[4] = {ff.dih_i[d] - 1, ff.dih_j[d] - 1, ff.dih_k[d] - 1, ff.dih_l[d] - 1};
        for (int role = 0; role < 4; role++) t.dih[s[role]][t.ndih[s[role]]++] = 4 * d + role;
    }
    int ni = ff.n_internal;
    for (int j = 0; j < 29; j++) {
        int q = 0;
        for (int k = 0; k < t.npair[j]; k++) t.dst[t.pslot[j][k]] = j * INTRA_MAXC + q++;
        for (int k = 0; k < t.nang[j]; k++) t.dst[2 * ni + t.ang[j][k]] = j * INTRA_MAXC + q++;
        for (int k = 0; k < t.ndih[j]; k++) t.dst[2 * ni + 108 + t.dih[j][k]] = j * INTRA_MAXC + q++;
        t.cnt[j] = q;
    }
}

static int check_topology(const sesmd_forcefield &ff)
{
    int np[29] = {0}, na[29] = {0}, nd[29] = {0};
    if (ff.n_internal < 0 || ff.n_internal > 320) return fail("n_internal out of range");
    for (int p = 0; p < ff.n_internal; p++) {
        int j = ff.int_j[p], k = ff.int_k[p];
        if (j < 1 || j > 29 || k < 1 || k > 29 || j == k) return fail("bad internal pair %d", p);
        if (++np[j - 1] > 32 || ++np[k - 1] > 32) return fail("too many internal pairs per site");
    }
    for (int a = 0; a < 36; a++)
        for (int s : {ff.ang_i[a], ff.ang_j[a], ff.ang_k[a]}) {
            if (s < 1 || s > 29) return fail("bad angle site");
            if (++na[s - 1] > 12) return fail("too many angles per site");
        }
    for (int d = 0; d < 30; d++)
        for (int s : {ff.dih_i[d], ff.dih_j[d], ff.dih_k[d], ff.dih_l[d]}) {
            if (s < 1 || s > 29) return fail("bad dihedral site");
            if (++nd[s - 1] > 16) return fail("too many dihedrals per site");
        }
    for (int b = 0; b < 28; b++)
        if (ff.ibond[b] < 1 || ff.ibond[b] > 29 || ff.jbond[b] < 1 || ff.jbond[b] > 29)
            return fail("bad bond");
    return 0;
}

static int upload_ops(sesmd_ctx *c)
{
    const Geo &g = c->g;
    int lp = g.lp, lk = g.lk, p = g.p;
    const double *fact = c->ff.rifact;
    std::vector<hostm::C> T((size_t)lk * lp);
    std::vector<double2> buf;
    auto put = [&](double2 *dst, size_t count) -> int {
        HC(hipMemcpy(dst, buf.data(), sizeof(double2) * count, hipMemcpyHostToDevice));
        return 0;
    };
    // operators are stored column-major (element (row, c) at c * lk + row), so
    // the lanes of a contraction (one row each) read consecutive elements
    auto append = [&]() {
        for (int cc = 0; cc < lp; cc++)
            for (int rr = 0; rr < lk; rr++) {
                const hostm::C &z = T[(size_t)rr * lp + cc];
                buf.push_back(make_double2(z.real(), z.imag()));
            }
    };
    // Real operator of the half-space products (m2l_half_off): output row
    // 2r + {re, im} of target row r, input 2c + {re, im} of O(n, m), m >= 0.
    // With A = T(r, (n, m)), B = T(r, (n, -m)) (zero for m = 0) and
    // O(n, -m) = conj(O(n, m)) = a - ib, the output gains A (a + ib) + B (a - ib):
    // Re: (A.re + B.re) a + (B.im - A.im) b, Im: (A.im + B.im) a + (A.re - B.re) b.
    auto half_op = [&](int mm, int kk) -> double {
        int r = mm >> 1, ro = mm & 1, ci = kk & 1, cc = kk >> 1;
        int n = 0;
        while ((n + 1) * (n + 2) / 2 <= cc) n++;
        int m = cc - n * (n + 1) / 2;
        const hostm::C &A = T[(size_t)r * lp + n * (n + 1) + m];
        hostm::C B = m > 0 ? T[(size_t)r * lp + n * (n + 1) - m] : hostm::C(0.0);
        if (ro == 0) return ci == 0 ? A.real() + B.real() : B.imag() - A.imag();
        return ci == 0 ? A.imag() + B.imag() : A.real() - B.real();
    };
    // MFMA operand order of the real operator: padded to MT x 16 rows and
    // KS x 4 columns; fragments ordered (ks, mt), so the fragments of
    // consecutive k steps are contiguous; lane l of fragment (ks, mt) holds
    // element (16 mt + l % 16, 4 ks + l / 16) -- the A operand of
    // v_mfma_f64_16x16x4f64.
    const int MT = (2 * lk + 15) / 16, KS = (2 * lk + 3) / 4;
    std::vector<double> af;
    auto append_frag = [&]() {
        for (int ks = 0; ks < KS; ks++)
            for (int mt = 0; mt < MT; mt++)
                for (int l = 0; l < 64; l++) {
                    int mm = 16 * mt + l % 16, kk = 4 * ks + l / 16;
                    af.push_back(mm < 2 * lk && kk < 2 * lk ? half_op(mm, kk) : 0.0);
                }
    };
    // mixed precision: the same operators as bf16 parts in the A-operand
    // layout of v_mfma_f32_16x16x32_bf16, fragments [ks][slab][part] of 64
    // lanes x 8 values; lane l of fragment (ks, slab) holds row 16 slab + l % 16,
    // columns 32 ks + 8 (l / 16) .. + 7
    const int KS32 = (2 * lk + 31) / 32;
    std::vector<unsigned short> afb;
    auto append_bf3 = [&]() {
        for (int ks = 0; ks < KS32; ks++)
            for (int mt = 0; mt < MT; mt++) {
                size_t at = afb.size();
                afb.resize(at + (size_t)3 * 64 * 8, 0);
                for (int l = 0; l < 64; l++)
                    for (int e = 0; e < 8; e++) {
                        int mm = 16 * mt + l % 16, kk = 32 * ks + 8 * (l / 16) + e;
                        double v = mm < 2 * lk && kk < 2 * lk ? half_op(mm, kk) : 0.0;
                        unsigned short h, m, lo;
                        bf16_split3(v, h, m, lo);
                        afb[at + (size_t)(0 * 64 + l) * 8 + e] = h;
                        afb[at + (size_t)(1 * 64 + l) * 8 + e] = m;
                        afb[at + (size_t)(2 * 64 + l) * 8 + e] = lo;
                    }
            }
    };
    bool mixed = c->precision == SESMD_MIXED;
    // M2L at L4 and L3: offsets in [-3, 3]^3 (spacing rs and 2 rs)
    for (int lvl = 0; lvl < 2; lvl++) {
        buf.clear();
        af.clear();
        afb.clear();
        double sp = lvl == 0 ? g.rs : 2.0 * g.rs;
        for (int dz = -3; dz <= 3; dz++)
            for (int dy = -3; dy <= 3; dy++)
                for (int dx = -3; dx <= 3; dx++) {
                    if (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1) {
                        std::fill(T.begin(), T.end(), hostm::C(0.0));
                    } else {
                        hostm::m2l_op(fact, p, dx * sp, dy * sp, dz * sp, T.data());
                    }
                    append();
                    append_frag();
                    if (mixed) append_bf3();
                }
        CK(put(lvl == 0 ? c->T4 : c->T3, buf.size()));
        if (mixed)
            HC(hipMemcpy(lvl == 0 ? c->AFB4 : c->AFB3, afb.data(), sizeof(unsigned short) * afb.size(),
                         hipMemcpyHostToDevice));
        af.resize(af.size() + (size_t)MT * KS * 64, 0.0);  // the spare offset
        HC(hipMemcpy(lvl == 0 ? c->AF4 : c->AF3, af.data(), sizeof(double) * af.size(),
                     hipMemcpyHostToDevice));
    }
    // M2L at the top level (spacing 4 rs), offsets in [-(nb2-1), nb2-1]^3
    {
        buf.clear();
        int w = g.nb2 - 1;
        double sp = 4.0 * g.rs;
        for (int dz = -w; dz <= w; dz++)
            for (int dy = -w; dy <= w; dy++)
                for (int dx = -w; dx <= w; dx++) {
                    if (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1)
                        std::fill(T.begin(), T.end(), hostm::C(0.0));
                    else
                        hostm::m2l_op(fact, p, dx * sp, dy * sp, dz * sp, T.data());
                    append();
                }
        CK(put(c->T2, buf.size()));
    }
    // M2M 4->3 and 3->2 (child - parent = +-half the child spacing... per axis),
    // L2L 2->3 and 3->4 (parent - child)
    for (int which = 0; which < 4; which++) {
        buf.clear();
        double h = (which == 0 || which == 3) ? g.rs / 2.0 : g.rs;
        for (int kc = 0; kc < 8; kc++) {
            double ox = (kc & 1) ? h : -h, oy = ((kc >> 1) & 1) ? h : -h, oz = (kc >> 2) ? h : -h;
            if (which < 2)
                hostm::m2m_op(fact, p, ox, oy, oz, T.data());
            else
                hostm::l2l_op(fact, p, -ox, -oy, -oz, T.data());
            append();
        }
        double2 *dst = which == 0 ? c->M43 : which == 1 ? c->M32 : which == 2 ? c->S23 : c->S34;
        CK(put(dst, buf.size()));
    }
    // rfctr(j, k) table for L2P
    std::vector<double> lf(lp);
    for (int j = 0; j <= p; j++)
        for (int k = -j; k <= j; k++) {
            int ak = std::abs(k);
            lf[j * (j + 1) + k] = std::sqrt(fact[j - ak] / fact[j + ak]);
        }
    HC(hipMemcpy(c->lfac, lf.data(), sizeof(double) * lp, hipMemcpyHostToDevice));
    return 0;
}

static hipError_t upload_lj32(sesmd_ctx *c, const sesmd_forcefield *ff)
{
    std::vector<float> t(2048);
    for (int k = 0; k < 1024; k++) {
        t[k] = (float)ff->lj_c[k];
        t[1024 + k] = (float)ff->lj_s6[k];
    }
    return hipMemcpy(c->lj32, t.data(), sizeof(float) * 2048, hipMemcpyHostToDevice);
}

// Device address of `scratch` on every device a context was created on,
// looked up once per device with that device current.
constexpr int SCRATCH_DEVICES = 64;
static void *g_scratch[SCRATCH_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_setup(int dev)
{
    if (dev < 0 || dev >= SCRATCH_DEVICES) return fail("device %d: the library tracks devices 0..%d", dev,
                                                       SCRATCH_DEVICES - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[dev]) return 0;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch[dev] = ptr;
    return 0;
}

static void *scratch_of(int dev)
{
    if (dev < 0 || dev >= SCRATCH_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[dev];
}

extern "C" int sesmd_create(sesmd_ctx **out, int n_water, int n_surf, int nb2, int p,
                            const sesmd_forcefield *ff, int precision)
{
    if (!out || !ff) return fail("null argument");
    *out = nullptr;
    if (n_water < 0 || n_surf < 0 || n_water + n_surf == 0) return fail("empty system");
    if (nb2 < 1 || 4 * nb2 > MAXNB) return fail("nb2 must be 1..%d", MAXNB / 4);
    if (p != 4 && p != 8) return fail("p must be 4 (the reference) or 8");
    if (precision != SESMD_FP64 && precision != SESMD_MIXED) return fail("bad precision");
    if (check_topology(*ff)) return 1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device");
    int dev = 0;
    HC(hipGetDevice(&dev));
    if (scratch_setup(dev)) return 1;
    sesmd_ctx *c = new sesmd_ctx();
    memset((void *)c, 0, sizeof(*c));
    c->device = dev;
    Geo &g = c->g;
    g.nw = n_water;
    g.ns = n_surf;
    g.n = 3 * n_water + 29 * n_surf;
    g.nmol = n_water + n_surf;
    g.nb2 = nb2;
    g.nb = 4 * nb2;
    g.nb3 = g.nb * g.nb * g.nb;
    g.n3 = g.nb / 2;
    g.n33 = g.n3 * g.n3 * g.n3;
    g.n2 = nb2 * nb2 * nb2;
    g.p = p;
    g.lp = (p + 1) * (p + 1);
    g.lk = (p + 1) * (p + 2) / 2;
    g.rs = ff->rs;
    g.rsd = g.nb * g.rs;
    g.rhs = g.rsd / 2.0;
    g.rc = ff->rc;
    g.rc2 = ff->rc * ff->rc;
    g.xpi = ff->xpi;
    g.dt = ff->dt;
    g.dtsq2 = ff->dtsq2;
    c->precision = precision;
    c->ff = *ff;
    int rc = 0;
    auto A = [&](int r2) {
        if (r2) rc = 1;
    };
    A(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ? fail("stream create failed")
                                                                                 : 0);
    A(hipStreamCreateWithFlags(&c->m2l_stream, hipStreamNonBlocking) != hipSuccess ||
              hipEventCreateWithFlags(&c->ev_m2l_fork, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&c->ev_m2l_join, hipEventDisableTiming) != hipSuccess
          ? fail("M2L stream create failed")
          : 0);
    A(hipStreamCreateWithFlags(&c->far_stream, hipStreamNonBlocking) != hipSuccess ||
              hipEventCreateWithFlags(&c->ev_far_fork, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&c->ev_far_join, hipEventDisableTiming) != hipSuccess
          ? fail("far-field stream create failed")
          : 0);
    size_t n = g.n;
    {  // box centres, computed on the device with the classification's own formula
        double *tab = nullptr;
        if (!rc && hipMalloc(&tab, sizeof(double) * MAXNB) == hipSuccess) {
            k_cen4_table<<<1, MAXNB>>>(g, tab);
            A(hipMemcpy(g.cen4, tab, sizeof(double) * g.nb, hipMemcpyDeviceToHost) != hipSuccess
                  ? fail("box centre table failed") : 0);
            hipFree(tab);
        } else {
            A(fail("box centre table alloc failed"));
        }
    }
    A(dalloc(&c->r, 3 * n));
    A(dalloc(&c->v, 3 * n));
    A(dalloc(&c->a, 3 * n));
    A(dalloc(&c->f, 3 * n));
    A(dalloc(&c->type, n));
    A(dalloc(&c->mol, n));
    A(dalloc(&c->first, g.nmol + 1));
    A(dalloc(&c->cls, g.nmol));
    A(dalloc(&c->anc, g.nmol));
    A(dalloc(&c->cbox, n));
    A(dalloc(&c->rkind, n));
    A(dalloc(&c->isub, n));
    A(dalloc(&c->key_in, n));
    A(dalloc(&c->key_out, n));
    A(dalloc(&c->val_in, n));
    A(dalloc(&c->perm, n));
    A(dalloc(&c->bstart, g.nb3 + 1));
    A(dalloc(&c->bs4, 4 * (size_t)g.nb3 + 1));
    // key counts (4 nb3), arrival counters (4 nb3) and bflag (nb3): one allocation, one fill per rebuild
    A(dalloc(&c->kcnt, 8 * (size_t)g.nb3 + g.nb3));
    c->bflag = c->kcnt + 8 * (size_t)g.nb3;
    A(dalloc(&c->nvalid, 1));
    c->max_work = g.nb3 + (int)(n / NEAR_CHUNK) + 1;
    // a row's list holds the receivers of up to 9 boxes: sum over rows of
    // ceil(receivers / NEAR_CHUNK) <= 9 n / NEAR_CHUNK + (rows with receivers)
    c->max_rwork = g.nb3 + (int)(9 * n / NEAR_CHUNK) + 1;
    A(dalloc(&c->work, c->max_work));
    A(dalloc(&c->rwork, c->max_rwork));
    A(dalloc(&c->work_raw, c->max_work));
    A(dalloc(&c->wkey_in, c->max_work));
    A(dalloc(&c->wkey_out, c->max_work));
    A(dalloc(&c->widx_in, c->max_work));
    A(dalloc(&c->widx_out, c->max_work));
    A(dalloc(&c->nwork, 1));
    A(dalloc(&c->nrwork, 1));
    A(dalloc(&c->obox, g.nb3));
    A(dalloc(&c->spos, n + SRC_PAD));
    A(dalloc(&c->npart, (size_t)3 * NEAR_SLOTS * n));
    A(dalloc(&c->iperm, n));
    A(dalloc(&c->smeta, n + SRC_PAD));
    if (precision == SESMD_MIXED) {
        A(dalloc(&c->spos32, n + SRC_PAD));
        A(dalloc(&c->lj32, 2048));
    }
    A(dalloc(&c->O4, (size_t)g.nb3 * g.lp));
    A(dalloc(&c->L4, (size_t)g.nb3 * g.lp));
    A(dalloc(&c->O3, (size_t)g.n33 * g.lp));
    A(dalloc(&c->L3, (size_t)g.n33 * g.lp));
    A(dalloc(&c->O2, (size_t)g.n2 * g.lp));
    A(dalloc(&c->top_part, (size_t)g.n2 * g.n2 * g.lk));
    A(dalloc(&c->L2, (size_t)g.n2 * g.lp));
    A(dalloc(&c->ne4, g.nb3));
    A(dalloc(&c->ne3, g.n33));
    A(dalloc(&c->ne2, g.n2));
    A(dalloc(&c->T4, (size_t)343 * g.lk * g.lp));
    {
        size_t mt = (2 * g.lk + 15) / 16, ks = (2 * g.lk + 3) / 4;
        // one spare offset of zeros at the end: k_m2l_gemm's operator loads
        // run unconditionally up to a chunk past an offset's last k step
        A(dalloc(&c->AF4, (size_t)344 * mt * ks * 64));
        A(dalloc(&c->AF3, (size_t)344 * mt * ks * 64));
        // M2L lists per level: every (target, candidate) pair at most (8
        // classes x 216 candidates x (nd/2)^3 targets)
        for (int lv = 0; lv < 2; lv++) {
            size_t nbox = lv == 0 ? (size_t)g.nb3 : (size_t)g.n33;
            size_t pairs = (size_t)216 * nbox;
            sesmd_ctx::M2LBuf &B = c->m2l[lv];
            // result rows: M2L_RES doubles, or floats for the bf16x3 products
            A(dalloc(&B.res, pairs * M2L_RES / (precision == SESMD_MIXED ? 2 : 1)));
            A(dalloc(&B.list, pairs));
            A(dalloc(&B.idx, pairs));
            A(dalloc(&B.cnt, M2L_LISTS));
            A(dalloc(&B.base, M2L_LISTS));
            A(dalloc(&B.gbase, M2L_GROUPS + 1));
            A(dalloc(&B.bbase, M2L_GROUPS + 1));
            B.max_blocks = (int)(pairs / (M2L_TPB * 16)) + M2L_GROUPS;
            if (precision == SESMD_MIXED) A(dalloc(&B.ob, nbox * ((2 * g.lk + 31) / 32) * 96));
        }
        if (precision == SESMD_MIXED) {
            size_t ks32 = (2 * g.lk + 31) / 32;
            A(dalloc(&c->AFB4, (size_t)M2L_GROUPS * ks32 * mt * 3 * 64 * 8));
            A(dalloc(&c->AFB3, (size_t)M2L_GROUPS * ks32 * mt * 3 * 64 * 8));
        }
    }
    A(dalloc(&c->T3, (size_t)343 * g.lk * g.lp));
    int w2 = 2 * nb2 - 1;
    A(dalloc(&c->T2, (size_t)w2 * w2 * w2 * g.lk * g.lp));
    A(dalloc(&c->M43, (size_t)8 * g.lk * g.lp));
    A(dalloc(&c->M32, (size_t)8 * g.lk * g.lp));
    A(dalloc(&c->S23, (size_t)8 * g.lk * g.lp));
    A(dalloc(&c->S34, (size_t)8 * g.lk * g.lp));
    A(dalloc(&c->lfac, g.lp));
    A(dalloc(&c->gstart, g.nb * g.nb + 1));
    A(dalloc(&c->gidx, 1));
    A(dalloc(&c->gr, 3));
    A(dalloc(&c->dff, 1));
    A(dalloc(&c->dtopo, 1));
    c->surf_fast = true;
    for (int b = 0; b < 28; b++)
        if (ff->ibond[b] - 1 != kBI[b] || ff->jbond[b] - 1 != kBJ[b]) c->surf_fast = false;
    for (int i = 0; i < 29; i++) {
        c->sconst.m[i] = ff->mass[3 + i];
        c->sconst.rm[i] = 1.0 / ff->mass[3 + i];
    }
    for (int b = 0; b < 28; b++) {
        double rma = c->sconst.rm[kBI[b]], rmb = c->sconst.rm[kBJ[b]];
        c->sconst.dsq[b] = ff->dsqs[b];
        c->sconst.k2[b] = 2.0 * (rma + rmb);
        c->sconst.kr[b] = -1.0 / (ff->dt * ((rma + rmb) * ff->dsqs[b]));
        c->sconst.wa[b] = rma * ff->dt;
        c->sconst.wb[b] = rmb * ff->dt;
    }
    c->nblk_w = g.nw > 0 ? (int)nblk(g.nw, 256) : 0;
    c->nblk_s = g.ns > 0 ? (int)nblk(g.ns, c->surf_fast ? SURF_MPB : 64) : 0;
    A(dalloc(&c->partials, (size_t)(c->nblk_w + c->nblk_s + 1) * NH_VALS));
    A(dalloc(&c->fs, 3 * 29 * (size_t)g.ns + 1));
    A(dalloc(&c->nh, 16 + NH_VALS));
    A(dalloc(&c->esum, 16));
    A(dalloc(&c->counters, 4));
    if (rc) {
        std::string e = g_err;
        sesmd_destroy(c);
        g_err = e;
        return 1;
    }
    // statics
    std::vector<int> type(n), mol(n), first(g.nmol + 1);
    for (int w = 0; w < n_water; w++) {
        first[w] = 3 * w;
        for (int k = 0; k < 3; k++) {
            type[3 * w + k] = k;
            mol[3 * w + k] = w;
        }
    }
    for (int s = 0; s < n_surf; s++) {
        int b = 3 * n_water + 29 * s;
        first[n_water + s] = b;
        for (int k = 0; k < 29; k++) {
            type[b + k] = 3 + k;
            mol[b + k] = n_water + s;
        }
    }
    first[g.nmol] = (int)n;
    Topo topo;
    build_topology(*ff, topo);
    if (hipMemcpy(c->type, type.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->mol, mol.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->first, first.data(), sizeof(int) * (g.nmol + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->dff, ff, sizeof(*ff), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->dtopo, &topo, sizeof topo, hipMemcpyHostToDevice) != hipSuccess ||
        (c->lj32 && upload_lj32(c, ff) != hipSuccess)) {
        sesmd_destroy(c);
        return fail("upload of statics failed");
    }
    if (upload_ops(c)) {
        std::string e = g_err;
        sesmd_destroy(c);
        g_err = e;
        return 1;
    }
    // radix-sort workspace
    // (the box-centred list of masked evaluations only; the rebuild itself calls no library sort)
    size_t bytes = 0;
    hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, c->wkey_in, c->wkey_out, c->widx_in,
                                                 c->widx_out, c->max_work, 0, 32, c->stream);
    c->cub_bytes = bytes;
    if (hipMalloc(&c->cub_tmp, bytes ? bytes : 16) != hipSuccess) {
        sesmd_destroy(c);
        return fail("workspace alloc failed");
    }
    *out = c;
    return 0;
}

extern "C" int sesmd_destroy(sesmd_ctx *c)
{
    if (!c) return 0;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->graph_ok) hipGraphExecDestroy(c->graph);
    if (c->graph_run_len) hipGraphExecDestroy(c->graph_run);
    if (c->graph_cycle_len) hipGraphExecDestroy(c->graph_cycle);
    void *ptrs[] = {c->r, c->v, c->a, c->f, c->type, c->mol, c->first, c->cls, c->anc, c->cbox,
                    c->rkind, c->isub, c->key_in, c->key_out, c->val_in, c->perm, c->cub_tmp,
                    c->bstart, c->bs4, c->kcnt, c->nvalid, c->work, c->nwork, c->rwork, c->nrwork, c->obox, c->work_raw, c->npart, c->iperm, c->wkey_in,
                    c->wkey_out, c->widx_in, c->widx_out, c->spos, c->smeta, c->O4, c->L4,
                    c->O3, c->L3, c->O2, c->L2, c->ne4, c->ne3, c->ne2, c->T4, c->T3, c->T2, c->top_part,
                    c->M43, c->M32, c->S23, c->S34, c->lfac, c->gstart, c->gidx, c->gr, c->dff,
                    c->AF4, c->AF3, c->m2l[0].res, c->m2l[0].list, c->m2l[0].idx, c->m2l[0].cnt, c->m2l[0].base,
                    c->m2l[0].gbase, c->m2l[0].bbase, c->m2l[0].ob, c->m2l[1].res, c->m2l[1].list,
                    c->m2l[1].idx, c->m2l[1].cnt, c->m2l[1].base, c->m2l[1].gbase, c->m2l[1].bbase,
                    c->m2l[1].ob, c->AFB4, c->AFB3,
                    c->dtopo, c->partials, c->nh, c->esum, c->counters, c->spos32, c->lj32, c->fs};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (c->far_stream) {
        hipStreamSynchronize(c->far_stream);
        hipStreamDestroy(c->far_stream);
    }
    if (c->m2l_stream) {
        hipStreamSynchronize(c->m2l_stream);
        hipStreamDestroy(c->m2l_stream);
    }
    if (c->ev_m2l_fork) hipEventDestroy(c->ev_m2l_fork);
    if (c->ev_m2l_join) hipEventDestroy(c->ev_m2l_join);
    if (c->ev_far_fork) hipEventDestroy(c->ev_far_fork);
    if (c->ev_far_join) hipEventDestroy(c->ev_far_join);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

// host AoS (N,3) <-> device SoA
static int up_aos(sesmd_ctx *c, double *dst, const double *src)
{
    size_t n = c->g.n;
    std::vector<double> t(3 * n);
    for (size_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) t[d * n + i] = src[3 * i + d];
    HC(hipMemcpyAsync(dst, t.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return 0;
}
static int down_aos(sesmd_ctx *c, double *dst, const double *src)
{
    size_t n = c->g.n;
    std::vector<double> t(3 * n);
    HC(hipMemcpyAsync(t.data(), src, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) dst[3 * i + d] = t[d * n + i];
    return 0;
}

extern "C" int sesmd_set_state(sesmd_ctx *c, const double *r, const double *v, const double *a,
                               double eta)
{
    if (!c) return fail("null context");
    if (r) CK(up_aos(c, c->r, r));
    if (v) CK(up_aos(c, c->v, v));
    if (a) CK(up_aos(c, c->a, a));
    HC(hipMemcpy(c->nh, &eta, sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int sesmd_get_state(sesmd_ctx *c, double *r, double *v, double *a, double *eta)
{
    if (!c) return fail("null context");
    if (r) CK(down_aos(c, r, c->r));
    if (v) CK(down_aos(c, v, c->v));
    if (a) CK(down_aos(c, a, c->a));
    if (eta) HC(hipMemcpy(eta, c->nh, sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int sesmd_set_nh(sesmd_ctx *c, double tr, double pnh1, double pnh2)
{
    if (!c) return fail("null context");
    double t[3] = {tr, pnh1, pnh2};
    HC(hipMemcpy(c->nh + 7, t, sizeof t, hipMemcpyHostToDevice));
    return 0;
}

// Captured step graphs bake in host-side choices (classification on or off,
// graphite arrays): drop them when those change.
static void drop_graphs(sesmd_ctx *c)
{
    if (c->graph_ok) hipGraphExecDestroy(c->graph);
    c->graph_ok = false;
    if (c->graph_run_len) hipGraphExecDestroy(c->graph_run);
    c->graph_run_len = 0;
    if (c->graph_cycle_len) hipGraphExecDestroy(c->graph_cycle);
    c->graph_cycle_len = 0;
}

extern "C" int sesmd_set_graphite(sesmd_ctx *c, int n, const double *xyz)
{
    if (!c) return fail("null context");
    if (n < 0 || (n > 0 && !xyz)) return fail("bad graphite");
    const Geo &g = c->g;
    int nc = g.nb * g.nb;
    double rcut = g.rc + std::sqrt(2.0) * g.rs / 2.0;
    rcut = rcut * rcut;
    std::vector<int> start(nc + 1, 0), idx;
    std::vector<double> cen(g.nb);
    for (int q = 0; q < g.nb; q++) {
        int i = q % 4 + 1;
        cen[q] = 4.0 * g.rs * i / 4.0 - 2.0 * g.rs - 4.0 * g.rs / 8.0 + rcen2(g, q / 4);
    }
    for (int col = 0; col < nc; col++) {
        start[col] = (int)idx.size();
        double cx = cen[col % g.nb], cy = cen[col / g.nb];
        for (int a = 0; a < n; a++) {
            double dx = xyz[3 * a] - cx, dy = xyz[3 * a + 1] - cy;
            if (dx * dx + dy * dy <= rcut) idx.push_back(a);
        }
    }
    start[nc] = (int)idx.size();
    hipFree(c->gidx);
    hipFree(c->gr);
    c->gidx = nullptr;
    c->gr = nullptr;
    CK(dalloc(&c->gidx, idx.size()));
    CK(dalloc(&c->gr, 3 * (size_t)n));
    if (!idx.empty()) HC(hipMemcpy(c->gidx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
    if (n) HC(hipMemcpy(c->gr, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    HC(hipMemcpy(c->gstart, start.data(), sizeof(int) * (nc + 1), hipMemcpyHostToDevice));
    c->ngr = n;
    drop_graphs(c);
    return 0;
}

// the sorted metadata (smeta, bflag) the near field and the site terms read
static void launch_gather_meta(sesmd_ctx *c)
{
    const Geo &g = c->g;
    k_gather_meta<<<nblk(g.n, 256), 256, 0, c->stream>>>(g, c->perm, c->type, c->mol, c->first, c->cls, c->anc,
                                                         c->cbox, c->rkind, c->smeta, c->bflag);
}

// meta: also the sorted metadata (the step pipeline issues it itself, after P2M)
static int launch_rebuild(sesmd_ctx *c, bool meta = true)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    if (!c->classes_fixed)
        k_classify<<<nblk(g.nmol, 256), 256, 0, st>>>(g, c->r, c->first, c->cls, c->anc, c->counters);
    // counting sort by key (kcnt: counts, then arrival counters; then bflag), cleared by one fill
    int K = 4 * g.nb3;
    k_fill_i<<<nblk(2 * K + g.nb3, 256), 256, 0, st>>>(c->kcnt, 2 * K + g.nb3, 0);
    k_site_box<<<nblk(g.n, 256), 256, 0, st>>>(g, c->r, c->mol, c->type, c->cls, c->anc, c->cbox,
                                               c->rkind, c->key_in, c->kcnt);
    if (g.ns) k_isub<<<g.ns, 64, 0, st>>>(g, c->r, c->anc, c->isub);
    k_key_scan_rows<<<1, 1024, 0, st>>>(g, c->kcnt, c->bs4, c->bstart, c->rwork, c->nrwork, c->max_rwork,
                                        c->obox);
    k_key_scatter<<<nblk(g.n, 256), 256, 0, st>>>(g, c->key_in, c->bs4, c->kcnt + K, c->val_in);
    k_key_order<<<K, 64, 0, st>>>(g, c->bs4, c->val_in, c->perm, c->key_out, c->iperm);
    c->boxlist = false;  // built on the first masked evaluation after this rebuild
    if (meta) launch_gather_meta(c);
    HC(hipGetLastError());
    c->built = true;
    return 0;
}

extern "C" int sesmd_rebuild(sesmd_ctx *c)
{
    if (!c) return fail("null context");
    CK(launch_rebuild(c));
    HC(hipStreamSynchronize(c->stream));
    return 0;
}

// M2L over the interaction lists of one level (computeLjk3 / computeLjk4);
// nd is even at every level (each level halves the one below it).  Matrix
// cores for p = 4 and p = 8; the vector kernel k_m2l stays for other orders.
// The products of one level (lists, then the GEMMs into the level's result
// rows): they read only the level's multipoles, so the L4 products can run
// before the L4 locals exist.  M2L_accum adds them in; the vector kernel k_m2l
// (orders other than 4 and 8) does both at once.
// occ: translate only into the boxes flagged in ne (those holding a site of a
// receiver kind): the step pipeline's form, whose only reader of a local
// expansion is the L2P of the box's own sites (at L3, the L2L into such boxes).
// Without it every box gets its full local expansion, as computeLjk3 /
// computeLjk4 write it (the sesmd_fmm and sesmd_fmm_stage form).
static void launch_m2l_lists(sesmd_ctx *c, int nd, const int *ne, bool occ)
{
    const Geo &g = c->g;
    if (g.p != 4 && g.p != 8) return;
    const sesmd_ctx::M2LBuf &B = c->m2l[nd == g.nb ? 0 : 1];
    hipStream_t st = c->stream;
    const int *tne = occ ? ne : nullptr;
    k_m2l_count<<<M2L_LISTS, 256, 0, st>>>(nd, ne, tne, B.cnt);
    k_m2l_scan<<<1, 512, 0, st>>>(B.cnt, B.base, B.gbase, B.bbase);
    k_m2l_fill<<<M2L_LISTS, 256, 0, st>>>(nd, ne, tne, B.base, B.list, B.idx);
}

// the products of one level over the lists launch_m2l_lists wrote
static void launch_m2l_gemms(sesmd_ctx *c, int nd, const double2 *O)
{
    const Geo &g = c->g;
    if (g.p != 4 && g.p != 8) return;
    const sesmd_ctx::M2LBuf &B = c->m2l[nd == g.nb ? 0 : 1];
    hipStream_t st = c->stream;
    const double *Od = (const double *)O;
    if (c->precision == SESMD_MIXED) {
        const m2l_u4 *AFB = (const m2l_u4 *)(nd == g.nb ? c->AFB4 : c->AFB3);
        const m2l_u4 *OB = (const m2l_u4 *)B.ob;
        float *res = (float *)B.res;
        int nbox = nd * nd * nd;
        if (g.p == 4) {  // two slabs, one 32-wide k step: four waves of two tiles
            k_m2l_split<1><<<nblk(nbox * 32, 256), 256, 0, st>>>(nbox, 2 * g.lp, 2 * g.lk, Od, B.ob);
            k_m2l_gemm_bf3<2, 1, 2, 4, 1><<<B.max_blocks, 256, 0, st>>>(B.gbase, B.bbase, B.list, OB, AFB, res);
        } else {  // six slabs, three k steps: four waves of 3 slabs x 4 tiles
            k_m2l_split<3><<<nblk(nbox * 96, 256), 256, 0, st>>>(nbox, 2 * g.lp, 2 * g.lk, Od, B.ob);
            k_m2l_gemm_bf3<3, 2, 4, 2, 3><<<B.max_blocks, 256, 0, st>>>(B.gbase, B.bbase, B.list, OB, AFB, res);
        }
        return;
    }
    const double *AF = nd == g.nb ? c->AF4 : c->AF3;
    if (g.p == 4)  // two slabs of rows: two tile halves per block, four waves
        k_m2l_gemm<2, 8, M2L_TPB, 2><<<B.max_blocks, 256, 0, st>>>(g, B.gbase, B.bbase, B.list, Od, AF, B.res);
    else  // six slabs x two tile halves: twelve waves (12 waves per CU, against 6
          // for a six-wave form at 214 VGPRs)
        k_m2l_gemm<6, 23, M2L_TPB, 2><<<B.max_blocks, 768, 0, st>>>(g, B.gbase, B.bbase, B.list, Od, AF, B.res);
}

static void launch_m2l_products(sesmd_ctx *c, int nd, const double2 *O, const int *ne, bool occ)
{
    launch_m2l_lists(c, nd, ne, occ);
    launch_m2l_gemms(c, nd, O);
}

// The level's results added into its locals L (after its L2L wrote them).
static void launch_m2l_accum(sesmd_ctx *c, int nd, const double2 *O, const int *ne, const double2 *T, double2 *L,
                             bool occ)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    if (g.p == 4 || g.p == 8) {
        const sesmd_ctx::M2LBuf &B = c->m2l[nd == g.nb ? 0 : 1];
        if (c->precision == SESMD_MIXED)
            k_m2l_accum<float><<<nd * nd * nd, 64, 0, st>>>(g, nd, occ ? ne : nullptr, B.idx,
                                                             (const float *)B.res, L);
        else
            k_m2l_accum<double><<<nd * nd * nd, 64, 0, st>>>(g, nd, occ ? ne : nullptr, B.idx, B.res, L);
        return;
    }
    int h = nd / 2;
    if (g.lk <= 16) {
        int nblk_ = 8 * ((h + 1) / 2) * ((h + 1) / 2) * h;
        k_m2l<16, 16><<<nblk_, 64 * 16, 0, st>>>(g, nd, O, ne, T, L);
    } else {
        k_m2l<64, 8><<<8 * h * h * h, 64 * 8, 0, st>>>(g, nd, O, ne, T, L);
    }
}

// M2L over the interaction lists of one level (computeLjk3 / computeLjk4).
static void launch_m2l(sesmd_ctx *c, int nd, const double2 *O, const int *ne, const double2 *T, double2 *L,
                       bool occ)
{
    launch_m2l_products(c, nd, O, ne, occ);
    launch_m2l_accum(c, nd, O, ne, T, L, occ);
}

// The far field in two parts: the head, P2M, runs alone in a rebuild step
// (beside the near field it took ~4x as long, and it heads the longer branch);
// the tail runs beside the near field.  Moving M2M to L3 and the M2L lists
// into the head as well left the L4 products sharing the CUs with the near
// field and starved the chain of short kernels at L3 instead (rebuild step
// ~565 -> ~590 us).
static void launch_fmm_head(sesmd_ctx *c)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    if (g.p == 4)
        k_p2m<4><<<g.nb3, 64, 0, st>>>(g, c->r, c->perm, c->bstart, c->rkind, c->type, c->dff, c->O4, c->ne4);
    else
        k_p2m<8><<<g.nb3, 64, 0, st>>>(g, c->r, c->perm, c->bstart, c->rkind, c->type, c->dff, c->O4, c->ne4);
}

// fork: the L4 products on m2l_stream (the step pipeline of one context);
// otherwise everything on the context's stream (an ensemble's members already
// run side by side on the graph's branches, and one more fork and join per
// member cost more than they hid)
static int launch_fmm_tail(sesmd_ctx *c, bool occ, bool fork)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    // the L4 products need only O4: on m2l_stream, beside the L3 chain
    // (M2M, top level, L2L, L3 M2L, L2L), joined before their sums go into
    // L4.  The two levels' lists and results live in separate buffers.
    if (fork) {
        HC(hipEventRecord(c->ev_m2l_fork, st));
        HC(hipStreamWaitEvent(c->m2l_stream, c->ev_m2l_fork, 0));
        c->stream = c->m2l_stream;
    }
    launch_m2l_products(c, g.nb, c->O4, c->ne4, occ);
    c->stream = st;
    if (fork) HC(hipEventRecord(c->ev_m2l_join, c->m2l_stream));
    k_m2m<<<g.n33, 512, 0, st>>>(g, g.n3, c->O4, c->ne4, c->M43, c->O3, c->ne3);
    k_m2m<<<g.n2, 512, 0, st>>>(g, g.nb2, c->O3, c->ne3, c->M32, c->O2, c->ne2);
    k_m2l_top_pair<<<g.n2 * g.n2, 64, 0, st>>>(g, c->O2, c->ne2, c->T2, c->top_part);
    k_m2l_top_sum<<<g.n2, 64, 0, st>>>(g, c->top_part, c->L2);
    k_l2l<<<g.n33, 64, 0, st>>>(g, g.n3, c->L2, c->S23, c->L3);
    launch_m2l(c, g.n3, c->O3, c->ne3, c->T3, c->L3, occ);
    if (occ)  // the step pipeline: L4 only where sites read it
        k_l2l_occupied<<<std::min(g.nb3, SITE_GRID), 64, 0, st>>>(g, c->obox, c->L3, c->S34, c->L4);
    else
        k_l2l<<<g.nb3, 64, 0, st>>>(g, g.nb, c->L3, c->S34, c->L4);
    if (fork) HC(hipStreamWaitEvent(st, c->ev_m2l_join, 0));
    launch_m2l_accum(c, g.nb, c->O4, c->ne4, c->T4, c->L4, occ);
    HC(hipGetLastError());
    return 0;
}

static int launch_fmm(sesmd_ctx *c, bool occ, bool fork)
{
    launch_fmm_head(c);
    return launch_fmm_tail(c, occ, fork);
}

extern "C" int sesmd_fmm(sesmd_ctx *c)
{
    if (!c) return fail("null context");
    if (!c->built) return fail("sesmd_fmm before sesmd_rebuild");
    CK(launch_fmm(c, false, true));
    HC(hipStreamSynchronize(c->stream));
    return 0;
}

// The box-centred work list for masked evaluations, once per rebuild.
static void ensure_boxlist(sesmd_ctx *c)
{
    if (c->boxlist) return;
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    k_worklist<<<1, 1024, 0, st>>>(g, c->bstart, c->work_raw, c->wkey_in, c->widx_in, c->nwork,
                                   c->max_work);
    size_t wbytes = c->cub_bytes;
    hipcub::DeviceRadixSort::SortPairsDescending(c->cub_tmp, wbytes, c->wkey_in, c->wkey_out, c->widx_in,
                                                 c->widx_out, c->max_work, 0, 32, st);
    k_work_gather<<<nblk(c->max_work, 256), 256, 0, st>>>(c->work_raw, c->widx_out, c->nwork, c->work);
    c->boxlist = true;
}

// The near-field kernel for `mask` (~0: all 39 reference kernels, the
// row-centred form; otherwise the box-centred form with the mask).
static void launch_near(sesmd_ctx *c, uint64_t mask)
{
    const Geo &g = c->g;
    bool hq = c->ff.q[1] == c->ff.q[2];
    if (mask == ~0ull) {
        bool mix = c->precision == SESMD_MIXED;
        // three waves per item, and the second halves of the last quarter
        dim3 grid(3 * c->max_rwork + 3 * (c->max_rwork / 4 + 1)), block(NEAR_TILE);
#define ROWS_ARGS \
    g, c->spos, c->smeta, c->bstart, c->bs4, c->bflag, c->rkind, c->rwork, c->nrwork, c->dff, c->spos32, \
        c->lj32, c->npart, c->counters, c->max_rwork
        if (mix && hq) k_near_rows<true, true><<<grid, block, 0, c->stream>>>(ROWS_ARGS);
        else if (mix) k_near_rows<false, true><<<grid, block, 0, c->stream>>>(ROWS_ARGS);
        else if (hq) k_near_rows<true, false><<<grid, block, 0, c->stream>>>(ROWS_ARGS);
        else k_near_rows<false, false><<<grid, block, 0, c->stream>>>(ROWS_ARGS);
#undef ROWS_ARGS
        return;
    }
    ensure_boxlist(c);
    dim3 grid(c->max_work * NEAR_ROWS), block(NEAR_TILE);
#define NEAR_ARGS \
    g, c->spos, c->smeta, c->bstart, c->bs4, c->bflag, c->rkind, c->work, c->nwork, c->dff, mask, c->f, \
        c->spos32, c->lj32, c->npart
    if (hq) k_near<true, true, false><<<grid, block, 0, c->stream>>>(NEAR_ARGS);
    else k_near<true, false, false><<<grid, block, 0, c->stream>>>(NEAR_ARGS);
#undef NEAR_ARGS
}

// gathered: the sorted copies of the positions are already current (written
// by the fused part A of a regular step)
// fold: graphite evaluated inside k_site_terms (the step pipeline's form)
// The forces around the near field: first the sorted copies of the
// positions that the near field and the site terms read (unless part A wrote
// them), then, after the near field, the per-site terms.
static void forces_gather(sesmd_ctx *c, unsigned what, uint64_t mask, bool gathered)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    if (!gathered)  // the site terms read the sorted copies too
        k_gather_pos<<<nblk(g.n, 256), 256, 0, st>>>(g, c->perm, c->r, c->type, c->dff, c->spos);
    if ((what & SESMD_NEAR) && c->precision == SESMD_MIXED && mask == ~0ull && !gathered)
        k_gather_pos32<<<nblk(g.n, 256), 256, 0, st>>>(g, c->perm, c->key_out, c->r, c->type, c->dff,
                                                       c->spos32);
}

static int forces_sites(sesmd_ctx *c, unsigned what, bool fold);

static int launch_forces(sesmd_ctx *c, unsigned what, uint64_t mask, bool gathered = false, bool fold = false)
{
    forces_gather(c, what, mask, gathered);
    if (what & SESMD_NEAR) launch_near(c, mask);
    return forces_sites(c, what, fold);
}

static int forces_sites(sesmd_ctx *c, unsigned what, bool fold)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    const unsigned intra = SESMD_INTERNAL | SESMD_SUBTRACT | SESMD_VALENCE | SESMD_DIHEDRAL | SESMD_DIHEDRAL_BE;
    (g.p == 4 ? k_site_terms<4> : k_site_terms<8>)<<<std::min(g.nb3, SITE_GRID), SITE_TPB, 0, st>>>(
        g, what, c->obox, c->L4, c->lfac, c->dff, c->npart, c->spos, c->smeta, c->f, c->gr,
        c->gstart, c->gidx, fold ? c->ngr : 0);
    if (g.ns && (what & intra))
        k_intra<<<nblk(29 * g.ns, 64), 64, 0, st>>>(g, what, c->r, c->isub, c->dff, c->dtopo, c->f,
                                                    nullptr);
    if (!fold && (what & SESMD_GRAPHITE) && c->ngr > 0)
        k_graphite<<<g.nb * g.nb, 64, 0, st>>>(g, c->r, c->perm, c->bstart, c->rkind, c->type, c->gr,
                                               c->gstart, c->gidx, c->dff, c->f);
    HC(hipGetLastError());
    return 0;
}

extern "C" int sesmd_forces(sesmd_ctx *c, unsigned what, uint64_t near_mask)
{
    if (!c) return fail("null context");
    if (!c->built) return fail("sesmd_forces before sesmd_rebuild");
    CK(launch_forces(c, what, near_mask));
    HC(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int sesmd_near_only(sesmd_ctx *c)
{
    if (!c || !c->built) return fail("not built");
    const Geo &g = c->g;
    launch_near(c, ~0ull);
    HC(hipGetLastError());
    return 0;
}

extern "C" int sesmd_get_forces(sesmd_ctx *c, double *f)
{
    if (!c || !f) return fail("null argument");
    return down_aos(c, f, c->f);
}

extern "C" int sesmd_set_forces(sesmd_ctx *c, const double *f)
{
    if (!c || !f) return fail("null argument");
    return up_aos(c, c->f, f);
}

static int launch_move_a(sesmd_ctx *c, unsigned clsmask = 7u, int update_eta = 1)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    if (c->nblk_w)
        k_move_a_water<<<c->nblk_w, 256, 0, st>>>(g, c->r, c->v, c->a, c->f, c->cls, c->dff, c->nh,
                                                   c->partials, c->counters, clsmask);
    if (c->nblk_s && (clsmask & 4u) && c->surf_fast)
        k_move_a_surf_fast<<<c->nblk_s, 256, 0, st>>>(g, c->r, c->v, c->a, c->f, c->sconst, c->nh,
                                                         c->partials, c->counters);
    else if (c->nblk_s && (clsmask & 4u))
        k_move_a_surf<<<c->nblk_s, 64, 0, st>>>(g, c->r, c->v, c->a, c->f, c->dff, c->nh, c->partials,
                                                c->counters);
    else if (c->nblk_s)
        k_fill_d<<<nblk((size_t)c->nblk_s * NH_VALS, 256), 256, 0, st>>>(
            c->partials + (size_t)c->nblk_w * NH_VALS, (size_t)c->nblk_s * NH_VALS, 0.0);
    k_nh_update<<<1, 256, 0, st>>>(g, c->partials, c->nblk_w + c->nblk_s, c->dff, c->nh, update_eta,
                                   c->counters);
    HC(hipGetLastError());
    return 0;
}

extern "C" int sesmd_move_a(sesmd_ctx *c, double *sums6)
{
    if (!c) return fail("null context");
    if (!c->built) return fail("sesmd_move_a before sesmd_rebuild (classes unknown)");
    CK(launch_move_a(c));
    if (sums6) HC(hipMemcpyAsync(sums6, c->nh + 1, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return 0;
}

static int launch_move_b(sesmd_ctx *c, unsigned clsmask = 7u)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    if (c->nblk_w)
        k_move_b_water<<<c->nblk_w, 256, 0, st>>>(g, c->r, c->v, c->a, c->f, c->dff, c->nh, c->counters,
                                                   c->cls, clsmask);
    if (!(clsmask & 4u)) {
        HC(hipGetLastError());
        return 0;
    }
    if (c->nblk_s && c->surf_fast)
        k_move_b_surf_fast<<<c->nblk_s, 256, 0, st>>>(g, c->r, c->v, c->a, c->f, nullptr, c->sconst,
                                                         c->nh, c->counters);
    else if (c->nblk_s)
        k_move_b_surf<<<c->nblk_s, 64, 0, st>>>(g, c->r, c->v, c->a, c->f, nullptr, c->dff, c->nh,
                                                c->counters);
    HC(hipGetLastError());
    return 0;
}

extern "C" int sesmd_move_b(sesmd_ctx *c)
{
    if (!c) return fail("null context");
    CK(launch_move_b(c));
    HC(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int sesmd_init_accel(sesmd_ctx *c)
{
    if (!c) return fail("null context");
    const Geo &g = c->g;
    k_init_accel<<<nblk(g.n, 256), 256, 0, c->stream>>>(g, c->type, c->f, c->dff, c->a);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(c->stream));
    return 0;
}

// One full step (src/mybox.f90:130-227), every launch on the context's one
// stream.  Part A and part B each run the surfactant and water molecules in
// one launch (the few, latency-bound surfactant blocks first); the eta update
// and the intramolecular surfactant forces share one launch, writing the
// latter into their own buffer fs, which part B adds:
//   A(surf+water) - eta + intra -> fs - [rebuild, far field] - near - site terms - B(surf+water)
// (A second stream for the eta/intra launch, beside the forces, measured
// slower: the join before part B costs more than the launch it hides.)
static int step_part_a(sesmd_ctx *c, bool rebuild, bool &gathered, bool far_aside = false);
static int step_part_b(sesmd_ctx *c);
static const unsigned STEP_INTRA =
    SESMD_INTERNAL | SESMD_SUBTRACT | SESMD_VALENCE | SESMD_DIHEDRAL | SESMD_DIHEDRAL_BE;

// the force components a step evaluates in launch_forces (with the reference
// topology the intramolecular terms run beside the eta update)
static unsigned step_what(const sesmd_ctx *c) { return c->surf_fast ? SESMD_ALL & ~STEP_INTRA : SESMD_ALL; }

static int launch_step(sesmd_ctx *c, bool rebuild)
{
    bool gathered = false;
    // A rebuild step runs the head of the far field (P2M) alone right after
    // the sort, then forks the tail to far_stream, beside the sorted
    // metadata, the position gathers and the near field: the branches touch
    // disjoint buffers (the far field reads the multipoles, the near field
    // the sorted copies and writes only its partials); the site terms need
    // both, so the join comes before them.  Regular steps stay on one
    // stream: a fork and join between queues costs more than the ~10 us they
    // could hide.  The near-field launches are issued before the far field's,
    // so a captured graph lists them first.
    CK(step_part_a(c, rebuild, gathered, rebuild));
    if (rebuild) {
        launch_fmm_head(c);
        HC(hipEventRecord(c->ev_far_fork, c->stream));
        HC(hipStreamWaitEvent(c->far_stream, c->ev_far_fork, 0));
        launch_gather_meta(c);
    }
    // with the reference topology the intramolecular terms were evaluated
    // with the eta update and part B adds them per site (with the
    // subtraction); graphite runs inside k_site_terms
    unsigned what = step_what(c);
    forces_gather(c, what, ~0ull, gathered);
    if (what & SESMD_NEAR) launch_near(c, ~0ull);
    if (rebuild) {
        hipStream_t own = c->stream;
        c->stream = c->far_stream;
        int r = launch_fmm_tail(c, true, true);
        c->stream = own;
        CK(r);
        HC(hipEventRecord(c->ev_far_join, c->far_stream));
        HC(hipStreamWaitEvent(c->stream, c->ev_far_join, 0));
    }
    CK(forces_sites(c, what, true));
    return step_part_b(c);
}

// part A of a step (with the rebuild and far field when `rebuild`); gathered:
// part A wrote the sorted position copies
static int step_part_a(sesmd_ctx *c, bool rebuild, bool &gathered, bool far_aside)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    gathered = false;
    if (c->surf_fast) {
        SortedOut so;
        so.iperm = c->iperm;
        so.key = c->key_out;
        so.q = reinterpret_cast<const double *>(reinterpret_cast<const char *>(c->dff) +
                                                offsetof(sesmd_forcefield, q));  // device table
        so.spos = rebuild ? nullptr : c->spos;  // a rebuild re-sorts: gathered afterwards
        so.spos32 = rebuild ? nullptr : c->spos32;
        gathered = !rebuild;
        k_move_a_fused<<<c->nblk_s + c->nblk_w, 256, 0, st>>>(g, c->r, c->v, c->a, c->f, c->cls, c->dff,
                                                              c->sconst, c->nh, c->partials, c->counters,
                                                              c->nblk_s, so);
        if (g.ns) {  // eta update and the intramolecular terms in one launch
            k_nh_intra<<<1 + g.ns, 256, 0, st>>>(g, c->partials, c->nblk_w + c->nblk_s, c->dff, c->nh, c->r,
                                                 c->dtopo, c->fs, c->counters);
        } else {
            k_nh_update<<<1, 256, 0, st>>>(g, c->partials, c->nblk_w + c->nblk_s, c->dff, c->nh, 1, c->counters);
        }
    } else {
        if (c->nblk_w)
            k_move_a_water<<<c->nblk_w, 256, 0, st>>>(g, c->r, c->v, c->a, c->f, c->cls, c->dff, c->nh,
                                                       c->partials, c->counters, 7u);
        if (c->nblk_s)
            k_move_a_surf<<<c->nblk_s, 64, 0, st>>>(g, c->r, c->v, c->a, c->f, c->dff, c->nh, c->partials,
                                                    c->counters);
        k_nh_update<<<1, 256, 0, st>>>(g, c->partials, c->nblk_w + c->nblk_s, c->dff, c->nh, 1, c->counters);
    }
    HC(hipGetLastError());
    if (rebuild) {
        // far_aside: the caller issues the sorted metadata and the far field
        CK(launch_rebuild(c, !far_aside));
        if (!far_aside) CK(launch_fmm(c, true, false));
    }
    return 0;
}

static int step_part_b(sesmd_ctx *c)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    if (c->surf_fast) {
        IntraIn in;
        if (g.ns) {
            in.fs = c->fs;
            in.isub = c->isub;
            in.ff = c->dff;
        }
        k_move_b_fused<<<c->nblk_s + c->nblk_w, 256, 0, st>>>(g, c->r, c->v, c->a, c->f, in, c->dff,
                                                              c->sconst, c->nh, c->counters, c->cls,
                                                              c->nblk_s);
    } else {
        if (c->nblk_w)
            k_move_b_water<<<c->nblk_w, 256, 0, st>>>(g, c->r, c->v, c->a, c->f, c->dff, c->nh, c->counters,
                                                       c->cls, 7u);
        if (c->nblk_s)
            k_move_b_surf<<<c->nblk_s, 64, 0, st>>>(g, c->r, c->v, c->a, c->f, nullptr, c->dff, c->nh,
                                                    c->counters);
    }
    HC(hipGetLastError());
    return 0;
}

// Captures nsteps steps, the first a rebuild step if `rebuild`.  Every launch of
// a step is a kernel (the rebuild's sort included), with no memset node.
static int capture_steps(sesmd_ctx *c, int nsteps, hipGraphExec_t *out, bool rebuild = false)
{
    hipGraph_t gr;
    HC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = 0;
    for (int k = 0; k < nsteps && !r; k++) r = launch_step(c, rebuild && k == 0);
    hipError_t e = hipStreamEndCapture(c->stream, &gr);
    if (r) return 1;
    HC(e);
    hipError_t ei = hipGraphInstantiate(out, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    HC(ei);
    return 0;
}

// Captures (without running) every graph sesmd_steps replays for this ifar:
// the whole cycle, the run of ifar - 1 regular steps and the single regular
// step.  Capture and instantiation cost milliseconds, so they happen once, at
// the first sesmd_steps call for an ifar, before any of its steps run.
static int ensure_graphs(sesmd_ctx *c, int ifar)
{
    int run = ifar - 1;
    if (c->graph_cycle_len != ifar) {
        if (c->graph_cycle_len) HC(hipGraphExecDestroy(c->graph_cycle));
        c->graph_cycle_len = 0;
        CK(capture_steps(c, ifar, &c->graph_cycle, true));
        c->graph_cycle_len = ifar;
    }
    if (run > 1 && c->graph_run_len != run) {
        if (c->graph_run_len) HC(hipGraphExecDestroy(c->graph_run));
        c->graph_run_len = 0;
        CK(capture_steps(c, run, &c->graph_run));
        c->graph_run_len = run;
    }
    if (!c->graph_ok) {
        CK(capture_steps(c, 1, &c->graph));
        c->graph_ok = true;
    }
    return 0;
}

extern "C" int sesmd_steps(sesmd_ctx *c, int nsteps, int ifar, int *kb)
{
    if (!c || !kb) return fail("null argument");
    if (!c->built) return fail("sesmd_steps before sesmd_rebuild");
    if (ifar < 1) return fail("ifar must be >= 1");
    if (nsteps < 0) return fail("nsteps must be >= 0");
    int run = ifar - 1;  // regular steps between two rebuild steps
    if (nsteps > 0) CK(ensure_graphs(c, ifar));
    for (int s = 0; s < nsteps;) {
        if ((*kb + 1) % ifar == 0 && nsteps - s >= ifar) {
            // a whole cycle, the rebuild step and its run of regular steps, as one graph
            HC(hipGraphLaunch(c->graph_cycle, c->stream));
            c->boxlist = false;  // the replayed rebuild made a new box sort
            *kb += ifar;
            s += ifar;
            continue;
        }
        if ((*kb + 1) % ifar == 0) {
            (*kb)++;
            s++;
            CK(launch_step(c, true));
            continue;
        }
        if (run > 1 && (*kb + 1) % ifar == 1 && nsteps - s >= run) {
            // a whole run of regular steps as one graph
            if (c->graph_run_len != run) {
                if (c->graph_run_len) HC(hipGraphExecDestroy(c->graph_run));
                c->graph_run_len = 0;
                CK(capture_steps(c, run, &c->graph_run));
                c->graph_run_len = run;
            }
            HC(hipGraphLaunch(c->graph_run, c->stream));
            *kb += run;
            s += run;
            continue;
        }
        if (!c->graph_ok) {
            CK(capture_steps(c, 1, &c->graph));
            c->graph_ok = true;
        }
        HC(hipGraphLaunch(c->graph, c->stream));
        (*kb)++;
        s++;
    }
    HC(hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// Ensembles: independent systems, one context each, advanced together.  A
// step of the ensemble is every member's part A (with its rebuild and far
// field in the first step of a cycle) in member order, then ONE near-field
// launch over the flattened row-item lists of all members, then every
// member's site terms and part B -- the whole of it one captured graph on the
// ensemble's stream.  The near-field launch is persistent: a fixed grid whose
// wavefronts stride over the flattened items.  An item (member m, row q,
// receivers c0 ..) runs the arithmetic of k_near_rows on member m's arrays
// for each receiver -- the same sources in the same order -- so a member's
// forces are bit-for-bit those of its context stepped alone, whatever the
// launch shape.  The launch shape (wavefronts per workgroup, receivers per
// wavefront, workgroups) is chosen by sesmd_ensemble_tune from timed runs.
struct EnsMember {
    Geo g;
    const double4 *spos;
    const int4 *smeta;
    const int *bstart, *bs4, *bflag, *rkind;
    const int2 *rwork;
    const int *nrwork;
    const sesmd_forcefield *ff;
    const float4 *spos32;
    const float *lj32;
    double *part;
    int *counters;
};

constexpr int ENS_BRANCHES = 4;  // streams the members' launches spread over (graph branches)

struct sesmd_ensemble {
    int n;
    std::vector<sesmd_ctx *> m;
    std::vector<hipStream_t> own;  // the members' own streams (restored after each use)
    hipStream_t stream;
    hipStream_t branch[ENS_BRANCHES];
    hipEvent_t ev_fork, ev_join[ENS_BRANCHES];
    EnsMember *dm;                 // device copy of the member table
    int *prefix;                   // first flattened item of each member; [n] = total
    bool mix;
    int waves, chunk, grid;        // launch shape: wavefronts per workgroup, receivers per wavefront, workgroups
    // batched integrator and site-term launches (see EnsStepArgs)
    bool batched;
    int p;                         // the members' expansion order
    SurfConst sc;                  // the members' shared constraint constants
    struct EnsStepArgs *da;        // device copy of the members' arguments
    int *pre;                      // block prefixes per launch kind: ENS_KINDS x (n + 1)
    int tot[19];                   // blocks per launch kind (ENS_KINDS)
    hipGraphExec_t g_cycle, g_step;
    int cycle_len;
    bool step_ok;
};

// flattened item counts: member i contributes its row items x (NEAR_CHUNK /
// chunk); one block of 1024 threads, each member's count loaded by its own
// thread, then a scan in LDS
constexpr int ENS_MAX = 1024;
__global__ __launch_bounds__(ENS_MAX) void k_ens_prefix(const EnsMember *__restrict__ dm, int n, int per,
                                                        int *__restrict__ prefix)
{
    __shared__ int s_c[ENS_MAX];
    int t = threadIdx.x;
    int c = t < n ? *dm[t].nrwork * per : 0;
    s_c[t] = c;
    __syncthreads();
    for (int d = 1; d < ENS_MAX; d *= 2) {  // inclusive scan
        int v = t >= d ? s_c[t - d] : 0;
        __syncthreads();
        s_c[t] += v;
        __syncthreads();
    }
    if (t < n) prefix[t] = s_c[t] - c;
    if (t == n - 1) prefix[n] = s_c[t];
}

template <bool MIX>
__global__ __launch_bounds__(256) void k_near_rows_ens(const EnsMember *__restrict__ dm,
                                                       const int *__restrict__ prefix, int n, int chunk)
{
    int lane = threadIdx.x & 63;
    int W = blockDim.x >> 6;
    int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int per = NEAR_CHUNK / chunk;
    int total = prefix[n];
    int stride = gridDim.x * W;
    for (int i = blockIdx.x * W + wv; i < total; i += stride) {  // wave-uniform
        int lo = 0, hi = n - 1;  // the member: the last whose first item is <= i
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        const EnsMember &M = dm[lo];
        int it = i - prefix[lo];
        // unit 3 (per item + sub) + j: receiver sub-chunk `sub` of item
        // `item`, source box j of its row (as k_near_rows' waves)
        int item = it / (3 * per), sub = (it / 3) % per, j = it % 3;
        const Geo &g = M.g;
        int nb = g.nb;
        int2 w = M.rwork[item];
        int q = w.x;
        int x0 = q % nb, y = (q / nb) % nb, z = q / (nb * nb);
        // the row's receiver list in the wave's LDS slot, as in k_near_rows;
        // every lane stores the same values, so each lane reads back the
        // list of this item whichever store lands last
        __shared__ RowRecv s_L[4];
        RowRecv &L = s_L[wv];
        {
            RowRecv Lr;
            row_receivers(g, M.bstart, x0, y, z, Lr);
            L = Lr;
        }
        int cend = min(L.pre[9], w.y + NEAR_CHUNK);
        int s0 = w.y + sub * chunk, s1 = min(cend, s0 + chunk);
        if (s0 >= s1) continue;
        if (s1 - s0 > NEAR_TILE)
            near_rows_item<true, MIX, 2>(g, M.spos, M.smeta, M.bs4, M.bflag, M.rkind, M.ff, M.spos32, M.lj32,
                                         M.part, L, x0, y, z, j, s0, s1, lane, M.counters);
        else
            near_rows_item<true, MIX, 1>(g, M.spos, M.smeta, M.bs4, M.bflag, M.rkind, M.ff, M.spos32, M.lj32,
                                         M.part, L, x0, y, z, j, s0, s1, lane, M.counters);
    }
}

// while one of these lives, member i launches on the ensemble's branch
// stream i % ENS_BRANCHES (or on the ensemble's stream: one = true)
struct EnsStreams {
    sesmd_ensemble *e;
    EnsStreams(sesmd_ensemble *e_, bool one = false) : e(e_)
    {
        for (int i = 0; i < e->n; i++) e->m[i]->stream = one ? e->stream : e->branch[i % ENS_BRANCHES];
    }
    ~EnsStreams()
    {
        for (int i = 0; i < e->n; i++) e->m[i]->stream = e->own[i];
    }
};

// ------------------------------------------------ batched ensemble launches
// A step's integrator and site-term launches of all members as one launch
// each.  Block B of such a launch belongs to member m, the last whose first
// block pre[m] (a prefix over the members' block counts, fixed when the
// ensemble is created) is at or below B, and runs block B - pre[m] of that
// member's own launch with the member's arguments -- the same code on the
// same data, so every member's results are those it has stepped alone.
// Used when every member runs the reference topology's fused integrator, at
// one expansion order, with one set of constraint constants.
struct EnsStepArgs {
    Geo g;
    double *r, *v, *a, *f;
    const int *cls;
    const sesmd_forcefield *ff;
    double *nh, *partials;
    int *counters;
    int nblk_s, nblk_w;
    SortedOut so;  // the sorted copies part A writes on a regular step
    const Topo *topo;
    double *fs;
    const unsigned *isub;
    const int4 *obox;
    const double2 *L4;
    const double *lfac, *npart;
    const double4 *spos;
    const int4 *smeta;
    const double *gr;
    const int *gstart, *gidx;
    int ngr;
    // the rebuild and far field
    int classes_fixed, max_rwork;
    const int *first, *mol, *type;
    int *cls_w, *anc, *cbox, *rkind, *key_in, *key_out, *val_in, *perm, *iperm, *kcnt, *bs4, *bstart, *bflag,
        *nrwork;
    unsigned *isub_w;
    int2 *rwork;
    int4 *obox_w, *smeta_w;
    double4 *spos_w;
    float4 *spos32_w;
    double2 *O4, *O3, *O2, *L2, *L3, *L4_w;
    int *ne4, *ne3, *ne2;
    const double2 *M43, *M32, *T2, *S23, *S34;
    const double *AF4, *AF3;
    const m2l_u4 *AFB4, *AFB3;
    struct Level {
        unsigned short *ob;
        double *res;
        int *cnt, *base, *gbase, *bbase;
        int2 *list;
        int *idx;
    } m2l[2];  // L4, L3
};

// launch kinds, each with its block-count prefix over the members: parts A
// and B, eta + intra, site terms; then the rebuild's and far field's grids
enum {
    ENS_AB, ENS_NH, ENS_ST,
    ENS_CLS, ENS_FILL, ENS_SITE, ENS_ISUB, ENS_ONE, ENS_KEY, ENS_NB3, ENS_N33, ENS_N2, ENS_N2SQ, ENS_LISTS,
    ENS_SPLIT4, ENS_SPLIT3, ENS_GEMM4, ENS_GEMM3, ENS_NEF,
    ENS_KINDS
};
static_assert(ENS_KINDS == 19, "sesmd_ensemble::tot holds one count per launch kind");

__device__ __forceinline__ int ens_owner(const int *__restrict__ pre, int n, int B)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= B) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// part A (k_move_a_fused); rebuild: no sorted copies (the rebuild re-sorts)
__global__ __launch_bounds__(256) void k_ens_move_a(const EnsStepArgs *__restrict__ da, const int *__restrict__ pre,
                                                    int n, SurfConst sc, int rebuild)
{
    int m = ens_owner(pre, n, blockIdx.x), blk = blockIdx.x - pre[m];
    const EnsStepArgs &A = da[m];
    SortedOut so = A.so;
    if (rebuild) {
        so.spos = nullptr;
        so.spos32 = nullptr;
    }
    if (blk < A.nblk_s)
        surf_a(blk, A.g, A.r, A.v, A.a, A.f, sc, A.nh, A.partials, A.counters, so);
    else
        water_a(blk - A.nblk_s, A.g, A.r, A.v, A.a, A.f, A.cls, A.ff, A.nh, A.partials, A.counters, 7u, so);
}

// the eta update and the intramolecular terms (k_nh_intra; a member without
// surfactants has block 0 only, k_nh_update's work)
__global__ __launch_bounds__(256) void k_ens_nh_intra(const EnsStepArgs *__restrict__ da,
                                                      const int *__restrict__ pre, int n)
{
    int m = ens_owner(pre, n, blockIdx.x), blk = blockIdx.x - pre[m];
    const EnsStepArgs &A = da[m];
    nh_intra_block(blk, A.g, A.partials, A.nblk_w + A.nblk_s, A.ff, A.nh, A.r, A.topo, A.fs, A.counters);
}

// the site terms (k_site_terms): member m's pre[m + 1] - pre[m] blocks take
// its occupied boxes in turn
template <int PP>
__global__ __launch_bounds__(SITE_TPB) void k_ens_site_terms(const EnsStepArgs *__restrict__ da,
                                                             const int *__restrict__ pre, int n, unsigned what)
{
    int m = ens_owner(pre, n, blockIdx.x);
    const EnsStepArgs &A = da[m];
    site_terms_boxes<PP>(blockIdx.x - pre[m], pre[m + 1] - pre[m], A.g, what, A.obox, A.L4, A.lfac, A.ff,
                         A.npart, A.spos, A.smeta, A.f, A.gr, A.gstart, A.gidx, A.ngr);
}

// part B (k_move_b_fused)
__global__ __launch_bounds__(256) void k_ens_move_b(const EnsStepArgs *__restrict__ da, const int *__restrict__ pre,
                                                    int n, SurfConst sc)
{
    int m = ens_owner(pre, n, blockIdx.x), blk = blockIdx.x - pre[m];
    const EnsStepArgs &A = da[m];
    if (blk < A.nblk_s) {
        IntraIn in;
        in.fs = A.fs;
        in.isub = A.isub;
        in.ff = A.ff;
        surf_b(blk, A.g, A.r, A.v, A.a, A.f, in, sc, A.nh, A.counters);
    } else {
        water_b(blk - A.nblk_s, A.g, A.r, A.v, A.a, A.f, A.ff, A.nh, A.counters, A.cls, 7u);
    }
}


// ---- the rebuild and far field of every member, one launch per kernel
// kind (the per-member sequence of launch_rebuild, launch_fmm and the
// position gathers; each batched kernel runs the per-block body of the
// member's own kernel)
struct EnsBlock {
    int m, blk;
};
__device__ __forceinline__ EnsBlock ens_block(const int *__restrict__ pre, int n)
{
    int m = ens_owner(pre, n, blockIdx.x);
    return EnsBlock{m, (int)blockIdx.x - pre[m]};
}
#define ENS_KERNEL(NAME, TPB) \
    __global__ __launch_bounds__(TPB) void NAME(const EnsStepArgs *__restrict__ da, const int *__restrict__ pre, int n)
#define ENS_MEMBER                     \
    const EnsBlock b = ens_block(pre, n); \
    const EnsStepArgs &A = da[b.m];

ENS_KERNEL(k_ens_classify, 256)
{
    ENS_MEMBER
    classify_blk(b.blk, A.g, A.r, A.first, A.cls_w, A.anc, A.counters);
}
ENS_KERNEL(k_ens_fill_keys, 256)
{
    ENS_MEMBER
    fill_i_blk(b.blk, A.kcnt, (size_t)9 * A.g.nb3, 0);  // counts, arrival counters (4 nb3 each), bflag (nb3)
}
ENS_KERNEL(k_ens_site_box, 256)
{
    ENS_MEMBER
    site_box_blk(b.blk, A.g, A.r, A.mol, A.type, A.cls, A.anc, A.cbox, A.rkind, A.key_in, A.kcnt);
}
ENS_KERNEL(k_ens_isub, 64)
{
    ENS_MEMBER
    isub_blk(b.blk, A.g, A.r, A.anc, A.isub_w);
}
ENS_KERNEL(k_ens_key_scan_rows, 1024)
{
    ENS_MEMBER
    key_scan_rows_blk(A.g, A.kcnt, A.bs4, A.bstart, A.rwork, A.nrwork, A.max_rwork, A.obox_w);
}
ENS_KERNEL(k_ens_key_scatter, 256)
{
    ENS_MEMBER
    key_scatter_blk(b.blk, A.g, A.key_in, A.bs4, A.kcnt + 4 * A.g.nb3, A.val_in);
}
// the keys blk, + stride, ... (most keys of a droplet's domain are empty)
ENS_KERNEL(k_ens_key_order, 64)
{
    ENS_MEMBER
    int stride = pre[b.m + 1] - pre[b.m];
    for (int kk = b.blk; kk < 4 * A.g.nb3; kk += stride)
        key_order_blk(kk, A.g, A.bs4, A.val_in, A.perm, A.key_out, A.iperm);
}
ENS_KERNEL(k_ens_gather_meta, 256)
{
    ENS_MEMBER
    gather_meta_blk(b.blk, A.g, A.perm, A.type, A.mol, A.first, A.cls, A.anc, A.cbox, A.rkind, A.smeta_w,
                    A.bflag);
}
ENS_KERNEL(k_ens_gather_pos, 256)
{
    ENS_MEMBER
    gather_pos_blk(b.blk, A.g, A.perm, A.r, A.type, A.ff, A.spos_w);
}
ENS_KERNEL(k_ens_gather_pos32, 256)
{
    ENS_MEMBER
    gather_pos32_blk(b.blk, A.g, A.perm, A.key_out, A.r, A.type, A.ff, A.spos32_w);
}
// P2M over the boxes holding sites (the occupied list of this rebuild, taken
// blk, + stride, ...), after k_ens_ne4_clear has cleared every box's flag:
// the multipoles of the other boxes are left as they were, and no reader
// takes a multipole whose flag is clear
ENS_KERNEL(k_ens_ne4_clear, 256)
{
    ENS_MEMBER
    fill_i_blk(b.blk, A.ne4, (size_t)A.g.nb3, 0);
}
template <int PP>
ENS_KERNEL(k_ens_p2m, 64)
{
    ENS_MEMBER
    int stride = pre[b.m + 1] - pre[b.m];
    for (int q = b.blk; q < A.g.nb3; q += stride) {
        int box = A.obox[q].x;
        if (box < 0) break;  // past the occupied boxes (block-uniform)
        p2m_blk<PP>(box, A.g, A.r, A.perm, A.bstart, A.rkind, A.type, A.ff, A.O4, A.ne4);
        __syncthreads();  // the block's LDS serves the next box
    }
}
// lv 0: L4 -> L3; lv 1: L3 -> L2
ENS_KERNEL(k_ens_m2m_4, 512)
{
    ENS_MEMBER
    m2m_blk(b.blk, A.g, A.g.n3, A.O4, A.ne4, A.M43, A.O3, A.ne3);
}
ENS_KERNEL(k_ens_m2m_3, 512)
{
    ENS_MEMBER
    m2m_blk(b.blk, A.g, A.g.nb2, A.O3, A.ne3, A.M32, A.O2, A.ne2);
}
// one block per target L2 box (m2l_top_blk); bound by the operator reads (a
// 58 KB operator per product at p = 8), not by the per-target chain
ENS_KERNEL(k_ens_top, 64)
{
    ENS_MEMBER
    m2l_top_blk(b.blk, A.g, A.O2, A.ne2, A.T2, A.L2);
}
ENS_KERNEL(k_ens_l2l_3, 64)
{
    ENS_MEMBER
    l2l_blk(b.blk, A.g, A.g.n3, A.L2, A.S23, A.L3);
}
// L2L into the boxes holding sites (the only readers of L4 here are their
// sites' far-field terms), taken blk, + stride, ...
ENS_KERNEL(k_ens_l2l_4, 64)
{
    ENS_MEMBER
    int stride = pre[b.m + 1] - pre[b.m];
    for (int q = b.blk; q < A.g.nb3; q += stride) {
        int box = A.obox[q].x;
        if (box < 0) break;
        l2l_blk(box, A.g, A.g.nb, A.L3, A.S34, A.L4_w);
    }
}
// the M2L lists of level lv (0: L4, nd = nb; 1: L3, nd = n3), occupied targets only
__global__ __launch_bounds__(256) void k_ens_m2l_count(const EnsStepArgs *__restrict__ da, const int *__restrict__ pre,
                                                       int n, int lv)
{
    ENS_MEMBER
    const int *ne = lv == 0 ? A.ne4 : A.ne3;
    m2l_count_blk(b.blk, lv == 0 ? A.g.nb : A.g.n3, ne, ne, A.m2l[lv].cnt);
}
__global__ __launch_bounds__(512) void k_ens_m2l_scan(const EnsStepArgs *__restrict__ da, const int *__restrict__ pre,
                                                      int n, int lv)
{
    ENS_MEMBER
    const EnsStepArgs::Level &B = A.m2l[lv];
    m2l_scan_blk(B.cnt, B.base, B.gbase, B.bbase);
}
__global__ __launch_bounds__(256) void k_ens_m2l_fill(const EnsStepArgs *__restrict__ da, const int *__restrict__ pre,
                                                      int n, int lv)
{
    ENS_MEMBER
    const int *ne = lv == 0 ? A.ne4 : A.ne3;
    const EnsStepArgs::Level &B = A.m2l[lv];
    m2l_fill_blk(b.blk, lv == 0 ? A.g.nb : A.g.n3, ne, ne, B.base, B.list, B.idx);
}
template <int KS>
__global__ __launch_bounds__(256) void k_ens_m2l_split(const EnsStepArgs *__restrict__ da,
                                                       const int *__restrict__ pre, int n, int lv)
{
    ENS_MEMBER
    int nd = lv == 0 ? A.g.nb : A.g.n3;
    m2l_split_blk<KS>(b.blk, nd * nd * nd, 2 * A.g.lp, 2 * A.g.lk, (const double *)(lv == 0 ? A.O4 : A.O3),
                      A.m2l[lv].ob);
}
// the GEMMs: member m's pre[m + 1] - pre[m] blocks take its list blocks in turn
template <int MTW, int WM, int TWW, int WT, int KS>
__global__ __launch_bounds__(64 * WM * WT) void k_ens_m2l_gemm_bf3(const EnsStepArgs *__restrict__ da,
                                                                   const int *__restrict__ pre, int n, int lv)
{
    ENS_MEMBER
    const EnsStepArgs::Level &B = A.m2l[lv];
    int nblocks = B.bbase[M2L_GROUPS], stride = pre[b.m + 1] - pre[b.m];
    for (int bk = b.blk; bk < nblocks; bk += stride)
        m2l_gemm_bf3_blk<MTW, WM, TWW, WT, KS>(bk, B.gbase, B.bbase, B.list, (const m2l_u4 *)B.ob,
                                               lv == 0 ? A.AFB4 : A.AFB3, (float *)B.res);
}
template <int MT, int KS, int TPB, int WG>
__global__ __launch_bounds__(64 * MT * WG) void k_ens_m2l_gemm(const EnsStepArgs *__restrict__ da,
                                                               const int *__restrict__ pre, int n, int lv)
{
    ENS_MEMBER
    const EnsStepArgs::Level &B = A.m2l[lv];
    int nblocks = B.bbase[M2L_GROUPS], stride = pre[b.m + 1] - pre[b.m];
    for (int bk = b.blk; bk < nblocks; bk += stride) {
        m2l_gemm_blk<MT, KS, TPB, WG>(bk, A.g, B.gbase, B.bbase, B.list, (const double *)(lv == 0 ? A.O4 : A.O3),
                                      lv == 0 ? A.AF4 : A.AF3, B.res);
        __syncthreads();  // the block's LDS is reused by its next list block
    }
}
template <typename T>
__global__ __launch_bounds__(64) void k_ens_m2l_accum(const EnsStepArgs *__restrict__ da, const int *__restrict__ pre,
                                                      int n, int lv)
{
    ENS_MEMBER
    const EnsStepArgs::Level &B = A.m2l[lv];
    if (lv == 0)
        m2l_accum_blk<T>(b.blk, A.g, A.g.nb, A.ne4, B.idx, (const T *)B.res, A.L4_w);
    else
        m2l_accum_blk<T>(b.blk, A.g, A.g.n3, A.ne3, B.idx, (const T *)B.res, A.L3);
}
#undef ENS_MEMBER
#undef ENS_KERNEL

// the branch streams start after the work queued on the ensemble's stream ...
static int ens_fork(sesmd_ensemble *e)
{
    HC(hipEventRecord(e->ev_fork, e->stream));
    for (int b = 0; b < ENS_BRANCHES; b++) HC(hipStreamWaitEvent(e->branch[b], e->ev_fork, 0));
    return 0;
}

// ... and the ensemble's stream continues after all of theirs
static int ens_join(sesmd_ensemble *e)
{
    for (int b = 0; b < ENS_BRANCHES; b++) {
        HC(hipEventRecord(e->ev_join[b], e->branch[b]));
        HC(hipStreamWaitEvent(e->stream, e->ev_join[b], 0));
    }
    return 0;
}

static void ens_near(sesmd_ensemble *e)
{
    k_ens_prefix<<<1, ENS_MAX, 0, e->stream>>>(e->dm, e->n, 3 * (NEAR_CHUNK / e->chunk), e->prefix);
    if (e->mix)
        k_near_rows_ens<true><<<e->grid, 64 * e->waves, 0, e->stream>>>(e->dm, e->prefix, e->n, e->chunk);
    else
        k_near_rows_ens<false><<<e->grid, 64 * e->waves, 0, e->stream>>>(e->dm, e->prefix, e->n, e->chunk);
}

// One step.  Batched (ens_step_batched): parts A, the eta updates with the
// intramolecular terms, the near field, the site terms and parts B of all
// members as one launch each on the ensemble's stream, and on a rebuild step
// each member's sort, lists and far field on the branch streams between the
// eta updates and the near field.  Otherwise the members' parts A spread over
// the branch streams, joined; the near field of all members on the ensemble's
// stream; the members' site terms and parts B on the branches, joined again.
// Captured, the branches are independent paths of the graph.
// every member's rebuild, far field and position gathers (launch_rebuild,
// launch_fmm with occupied targets, forces_gather), one launch per kernel
template <int PP>
static int ens_rebuild_batched(sesmd_ensemble *e)
{
    int n = e->n;
    hipStream_t st = e->stream;
    const EnsStepArgs *da = e->da;
    auto P = [&](int k) { return e->pre + (size_t)k * (n + 1); };
    auto T = [&](int k) { return e->tot[k]; };
    if (T(ENS_CLS)) k_ens_classify<<<T(ENS_CLS), 256, 0, st>>>(da, P(ENS_CLS), n);
    k_ens_fill_keys<<<T(ENS_FILL), 256, 0, st>>>(da, P(ENS_FILL), n);
    k_ens_site_box<<<T(ENS_SITE), 256, 0, st>>>(da, P(ENS_SITE), n);
    if (T(ENS_ISUB)) k_ens_isub<<<T(ENS_ISUB), 64, 0, st>>>(da, P(ENS_ISUB), n);
    k_ens_key_scan_rows<<<T(ENS_ONE), 1024, 0, st>>>(da, P(ENS_ONE), n);
    k_ens_key_scatter<<<T(ENS_SITE), 256, 0, st>>>(da, P(ENS_SITE), n);
    k_ens_key_order<<<T(ENS_ST), 64, 0, st>>>(da, P(ENS_ST), n);
    k_ens_gather_meta<<<T(ENS_SITE), 256, 0, st>>>(da, P(ENS_SITE), n);
    k_ens_ne4_clear<<<T(ENS_NEF), 256, 0, st>>>(da, P(ENS_NEF), n);
    k_ens_p2m<PP><<<T(ENS_ST), 64, 0, st>>>(da, P(ENS_ST), n);
    // one level's M2L products (lists, then the GEMMs into its result rows)
    auto products = [&](int lv) {
        k_ens_m2l_count<<<T(ENS_LISTS), 256, 0, st>>>(da, P(ENS_LISTS), n, lv);
        k_ens_m2l_scan<<<T(ENS_ONE), 512, 0, st>>>(da, P(ENS_ONE), n, lv);
        k_ens_m2l_fill<<<T(ENS_LISTS), 256, 0, st>>>(da, P(ENS_LISTS), n, lv);
        int kg = lv == 0 ? ENS_GEMM4 : ENS_GEMM3, ksp = lv == 0 ? ENS_SPLIT4 : ENS_SPLIT3;
        if (e->mix && PP == 4) {
            k_ens_m2l_split<1><<<T(ksp), 256, 0, st>>>(da, P(ksp), n, lv);
            k_ens_m2l_gemm_bf3<2, 1, 2, 4, 1><<<T(kg), 256, 0, st>>>(da, P(kg), n, lv);
        } else if (e->mix) {
            k_ens_m2l_split<3><<<T(ksp), 256, 0, st>>>(da, P(ksp), n, lv);
            k_ens_m2l_gemm_bf3<3, 2, 4, 2, 3><<<T(kg), 256, 0, st>>>(da, P(kg), n, lv);
        } else if (PP == 4) {
            k_ens_m2l_gemm<2, 8, M2L_TPB, 2><<<T(kg), 256, 0, st>>>(da, P(kg), n, lv);
        } else {
            k_ens_m2l_gemm<6, 23, M2L_TPB, 2><<<T(kg), 768, 0, st>>>(da, P(kg), n, lv);
        }
    };
    auto accum = [&](int lv) {
        int k = lv == 0 ? ENS_NB3 : ENS_N33;
        if (e->mix) k_ens_m2l_accum<float><<<T(k), 64, 0, st>>>(da, P(k), n, lv);
        else k_ens_m2l_accum<double><<<T(k), 64, 0, st>>>(da, P(k), n, lv);
    };
    products(0);
    k_ens_m2m_4<<<T(ENS_N33), 512, 0, st>>>(da, P(ENS_N33), n);
    k_ens_m2m_3<<<T(ENS_N2), 512, 0, st>>>(da, P(ENS_N2), n);
    k_ens_top<<<T(ENS_N2), 64, 0, st>>>(da, P(ENS_N2), n);
    k_ens_l2l_3<<<T(ENS_N33), 64, 0, st>>>(da, P(ENS_N33), n);
    products(1);
    accum(1);
    k_ens_l2l_4<<<T(ENS_ST), 64, 0, st>>>(da, P(ENS_ST), n);
    accum(0);
    k_ens_gather_pos<<<T(ENS_SITE), 256, 0, st>>>(da, P(ENS_SITE), n);
    if (e->mix) k_ens_gather_pos32<<<T(ENS_SITE), 256, 0, st>>>(da, P(ENS_SITE), n);
    for (int i = 0; i < n; i++) {
        e->m[i]->boxlist = false;  // the masked evaluations' box list, rebuilt on demand
        e->m[i]->built = true;
    }
    HC(hipGetLastError());
    return 0;
}

static int ens_step_batched(sesmd_ensemble *e, bool rebuild)
{
    int n = e->n;
    hipStream_t st = e->stream;
    const int *pre = e->pre;
    k_ens_move_a<<<e->tot[ENS_AB], 256, 0, st>>>(e->da, pre + ENS_AB * (n + 1), n, e->sc, rebuild ? 1 : 0);
    k_ens_nh_intra<<<e->tot[ENS_NH], 256, 0, st>>>(e->da, pre + ENS_NH * (n + 1), n);
    if (rebuild) CK(e->p == 4 ? ens_rebuild_batched<4>(e) : ens_rebuild_batched<8>(e));
    ens_near(e);
    unsigned what = step_what(e->m[0]);
    if (e->p == 4)
        k_ens_site_terms<4><<<e->tot[ENS_ST], SITE_TPB, 0, st>>>(e->da, pre + ENS_ST * (n + 1), n, what);
    else
        k_ens_site_terms<8><<<e->tot[ENS_ST], SITE_TPB, 0, st>>>(e->da, pre + ENS_ST * (n + 1), n, what);
    k_ens_move_b<<<e->tot[ENS_AB], 256, 0, st>>>(e->da, pre + ENS_AB * (n + 1), n, e->sc);
    HC(hipGetLastError());
    return 0;
}

static int ens_step(sesmd_ensemble *e, bool rebuild)
{
    if (e->batched) return ens_step_batched(e, rebuild);
    EnsStreams on(e);
    CK(ens_fork(e));
    for (int i = 0; i < e->n; i++) {
        sesmd_ctx *c = e->m[i];
        bool gathered = false;
        CK(step_part_a(c, rebuild, gathered));
        forces_gather(c, step_what(c), ~0ull, gathered);
    }
    CK(ens_join(e));
    ens_near(e);
    CK(ens_fork(e));
    for (int i = 0; i < e->n; i++) {
        sesmd_ctx *c = e->m[i];
        CK(forces_sites(c, step_what(c), true));
        CK(step_part_b(c));
    }
    CK(ens_join(e));
    HC(hipGetLastError());
    return 0;
}

static int ens_capture(sesmd_ensemble *e, int nsteps, bool rebuild, hipGraphExec_t *out)
{
    hipGraph_t gr;
    HC(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int r = 0;
    for (int k = 0; k < nsteps && !r; k++) r = ens_step(e, rebuild && k == 0);
    hipError_t er = hipStreamEndCapture(e->stream, &gr);
    if (r) return 1;
    HC(er);
    hipError_t ei = hipGraphInstantiate(out, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    HC(ei);
    return 0;
}

static void ens_drop_graphs(sesmd_ensemble *e)
{
    if (e->cycle_len) hipGraphExecDestroy(e->g_cycle);
    if (e->step_ok) hipGraphExecDestroy(e->g_step);
    e->cycle_len = 0;
    e->step_ok = false;
}

extern "C" int sesmd_ensemble_create(sesmd_ensemble **out, sesmd_ctx *const *members, int n)
{
    if (!out || !members) return fail("null argument");
    *out = nullptr;
    if (n < 1 || n > ENS_MAX) return fail("an ensemble has 1..%d members", ENS_MAX);
    for (int i = 0; i < n; i++) {
        const sesmd_ctx *c = members[i];
        if (!c) return fail("null member %d", i);
        if (!c->built) return fail("member %d not built (sesmd_rebuild first)", i);
        if (c->precision != members[0]->precision) return fail("members differ in precision");
        if (c->device != members[0]->device) return fail("members live on different devices");
        if (c->ff.q[1] != c->ff.q[2]) return fail("member %d: the ensemble needs equal hydrogen charges", i);
        for (int j = 0; j < i; j++)
            if (members[j] == c) return fail("member %d repeated", i);
    }
    sesmd_ensemble *e = new sesmd_ensemble();
    e->n = n;
    e->m.assign(members, members + n);
    for (int i = 0; i < n; i++) e->own.push_back(members[i]->stream);
    e->mix = members[0]->precision == SESMD_MIXED;
    e->waves = 1;
    e->chunk = NEAR_CHUNK;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, members[0]->device) != hipSuccess ||
        cus < 1)
        cus = 256;
    e->grid = 8 * 4 * cus;  // eight wavefronts per SIMD
    std::vector<EnsMember> hm(n);
    for (int i = 0; i < n; i++) {
        sesmd_ctx *c = members[i];
        EnsMember &M = hm[i];
        M.g = c->g;
        M.spos = c->spos;
        M.smeta = c->smeta;
        M.bstart = c->bstart;
        M.bs4 = c->bs4;
        M.bflag = c->bflag;
        M.rkind = c->rkind;
        M.rwork = c->rwork;
        M.nrwork = c->nrwork;
        M.ff = c->dff;
        M.spos32 = c->spos32;
        M.lj32 = c->lj32;
        M.part = c->npart;
        M.counters = c->counters;
    }
    // batched launches: every member on the fused integrator path, at one
    // expansion order, with the same constraint constants
    e->p = members[0]->g.p;
    e->sc = members[0]->sconst;
    e->batched = true;
    for (int i = 0; i < n; i++) {
        const sesmd_ctx *c = members[i];
        if (!c->surf_fast || c->g.p != e->p || (e->p != 4 && e->p != 8) ||
            memcmp(&c->sconst, &e->sc, sizeof(SurfConst)) != 0)
            e->batched = false;
    }
    std::vector<EnsStepArgs> ha(n);
    std::vector<int> hpre((size_t)ENS_KINDS * (n + 1));
    for (int k = 0; k < ENS_KINDS; k++) e->tot[k] = 0;
    for (int i = 0; i < n; i++) {
        sesmd_ctx *c = members[i];
        EnsStepArgs &A = ha[i];
        A.g = c->g;
        A.r = c->r;
        A.v = c->v;
        A.a = c->a;
        A.f = c->f;
        A.cls = c->cls;
        A.ff = c->dff;
        A.nh = c->nh;
        A.partials = c->partials;
        A.counters = c->counters;
        A.nblk_s = c->nblk_s;
        A.nblk_w = c->nblk_w;
        A.so.iperm = c->iperm;
        A.so.key = c->key_out;
        A.so.q = reinterpret_cast<const double *>(reinterpret_cast<const char *>(c->dff) +
                                                  offsetof(sesmd_forcefield, q));
        A.so.spos = c->spos;
        A.so.spos32 = c->spos32;
        A.topo = c->dtopo;
        A.fs = c->fs;
        A.isub = c->isub;
        A.obox = c->obox;
        A.L4 = c->L4;
        A.lfac = c->lfac;
        A.npart = c->npart;
        A.spos = c->spos;
        A.smeta = c->smeta;
        A.gr = c->gr;
        A.gstart = c->gstart;
        A.gidx = c->gidx;
        A.ngr = c->ngr;
        A.classes_fixed = c->classes_fixed ? 1 : 0;
        A.max_rwork = c->max_rwork;
        A.first = c->first;
        A.mol = c->mol;
        A.type = c->type;
        A.cls_w = c->cls;
        A.anc = c->anc;
        A.cbox = c->cbox;
        A.rkind = c->rkind;
        A.key_in = c->key_in;
        A.key_out = c->key_out;
        A.val_in = c->val_in;
        A.perm = c->perm;
        A.iperm = c->iperm;
        A.kcnt = c->kcnt;
        A.bs4 = c->bs4;
        A.bstart = c->bstart;
        A.bflag = c->bflag;
        A.nrwork = c->nrwork;
        A.isub_w = c->isub;
        A.rwork = c->rwork;
        A.obox_w = c->obox;
        A.smeta_w = c->smeta;
        A.spos_w = c->spos;
        A.spos32_w = c->spos32;
        A.O4 = c->O4;
        A.O3 = c->O3;
        A.O2 = c->O2;
        A.L2 = c->L2;
        A.L3 = c->L3;
        A.L4_w = c->L4;
        A.ne4 = c->ne4;
        A.ne3 = c->ne3;
        A.ne2 = c->ne2;
        A.M43 = c->M43;
        A.M32 = c->M32;
        A.T2 = c->T2;
        A.S23 = c->S23;
        A.S34 = c->S34;
        A.AF4 = c->AF4;
        A.AF3 = c->AF3;
        A.AFB4 = (const m2l_u4 *)c->AFB4;
        A.AFB3 = (const m2l_u4 *)c->AFB3;
        for (int lv = 0; lv < 2; lv++) {
            const sesmd_ctx::M2LBuf &B = c->m2l[lv];
            A.m2l[lv] = EnsStepArgs::Level{B.ob, B.res, B.cnt, B.base, B.gbase, B.bbase, B.list, B.idx};
        }
        // site terms: about one block per 64 sites (a droplet's boxes hold
        // ~100 sites each), at most the single launch's grid
        int nst = std::max(1, std::min(std::min(c->g.nb3, SITE_GRID), (c->g.n + 63) / 64));
        const Geo &g = c->g;
        int ks = (2 * g.lk + 31) / 32;  // k steps of the bf16 products
        // the GEMMs: about one block per 256 (L4) / 1,024 (L3) sites, each
        // taking the member's list blocks in turn
        int gm4 = std::max(1, std::min(c->m2l[0].max_blocks, std::max(8, g.n / 256)));
        int gm3 = std::max(1, std::min(c->m2l[1].max_blocks, std::max(4, g.n / 1024)));
        auto nb256 = [](int m) { return (int)nblk(m, 256); };
        int cnt[ENS_KINDS] = {
            c->nblk_s + c->nblk_w, 1 + g.ns, nst,
            c->classes_fixed ? 0 : nb256(g.nmol), nb256(9 * g.nb3), nb256(g.n), g.ns, 1, 4 * g.nb3,
            g.nb3, g.n33, g.n2, g.n2 * g.n2, M2L_LISTS,
            nb256(g.nb3 * 32 * ks), nb256(g.n33 * 32 * ks), gm4, gm3, nb256(g.nb3)};
        for (int k = 0; k < ENS_KINDS; k++) {
            hpre[(size_t)k * (n + 1) + i] = e->tot[k];
            e->tot[k] += cnt[k];
        }
    }
    for (int k = 0; k < ENS_KINDS; k++) hpre[(size_t)k * (n + 1) + n] = e->tot[k];
    int rc = 0;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) rc = fail("stream create failed");
    for (int b = 0; b < ENS_BRANCHES && !rc; b++)
        if (hipStreamCreateWithFlags(&e->branch[b], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e->ev_join[b], hipEventDisableTiming) != hipSuccess)
            rc = fail("branch stream create failed");
    if (!rc && hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess)
        rc = fail("event create failed");
    if (!rc && dalloc(&e->dm, (size_t)n)) rc = 1;
    if (!rc && dalloc(&e->prefix, (size_t)n + 1)) rc = 1;
    if (!rc && hipMemcpy(e->dm, hm.data(), sizeof(EnsMember) * n, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail("member table upload failed");
    if (!rc && dalloc(&e->da, (size_t)n)) rc = 1;
    if (!rc && dalloc(&e->pre, hpre.size())) rc = 1;
    if (!rc && (hipMemcpy(e->da, ha.data(), sizeof(EnsStepArgs) * n, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(e->pre, hpre.data(), sizeof(int) * hpre.size(), hipMemcpyHostToDevice) != hipSuccess))
        rc = fail("member argument upload failed");
    for (int i = 0; i < n && !rc; i++)  // the members' queued work before the ensemble's
        if (hipStreamSynchronize(e->own[i]) != hipSuccess) rc = fail("member %d stream failed", i);
    if (rc) {
        if (e->dm) hipFree(e->dm);
        if (e->prefix) hipFree(e->prefix);
        if (e->da) hipFree(e->da);
        if (e->pre) hipFree(e->pre);
        if (e->stream) hipStreamDestroy(e->stream);
        for (int b = 0; b < ENS_BRANCHES; b++) {
            if (e->branch[b]) hipStreamDestroy(e->branch[b]);
            if (e->ev_join[b]) hipEventDestroy(e->ev_join[b]);
        }
        if (e->ev_fork) hipEventDestroy(e->ev_fork);
        delete e;
        return 1;
    }
    *out = e;
    return 0;
}

extern "C" int sesmd_ensemble_destroy(sesmd_ensemble *e)
{
    if (!e) return 0;
    hipStreamSynchronize(e->stream);
    ens_drop_graphs(e);
    hipFree(e->dm);
    hipFree(e->prefix);
    hipFree(e->da);
    hipFree(e->pre);
    hipStreamDestroy(e->stream);
    for (int b = 0; b < ENS_BRANCHES; b++) {
        hipStreamDestroy(e->branch[b]);
        hipEventDestroy(e->ev_join[b]);
    }
    hipEventDestroy(e->ev_fork);
    delete e;
    return 0;
}

extern "C" int sesmd_ensemble_set_shape(sesmd_ensemble *e, int waves, int chunk, int grid)
{
    if (!e) return fail("null ensemble");
    if (waves != 1 && waves != 2 && waves != 4) return fail("waves per workgroup must be 1, 2 or 4");
    if (!e->mix && waves != 1)  // the FP64 pair loop stages per-lane constants in LDS by lane
        return fail("an FP64 ensemble runs one wavefront per workgroup");
    if (chunk != 64 && chunk != 128) return fail("receivers per wavefront must be 64 or 128");
    if (grid < 1) return fail("grid must be positive");
    if (waves != e->waves || chunk != e->chunk || grid != e->grid) ens_drop_graphs(e);
    e->waves = waves;
    e->chunk = chunk;
    e->grid = grid;
    return 0;
}

extern "C" int sesmd_ensemble_batched(sesmd_ensemble *e)
{
    if (!e) return -1;
    return e->batched ? 1 : 0;
}

extern "C" int sesmd_ensemble_get_shape(sesmd_ensemble *e, int *shape3)
{
    if (!e || !shape3) return fail("null argument");
    shape3[0] = e->waves;
    shape3[1] = e->chunk;
    shape3[2] = e->grid;
    return 0;
}

extern "C" void *sesmd_ensemble_stream(sesmd_ensemble *e) { return e ? (void *)e->stream : nullptr; }

// the ensemble's near field once, at the members' current sorted positions
extern "C" int sesmd_ensemble_near_only(sesmd_ensemble *e)
{
    if (!e) return fail("null ensemble");
    ens_near(e);
    HC(hipGetLastError());
    return 0;
}

// Times every launch shape (wavefronts per workgroup 1, 2, 4 -- 1 for FP64;
// receivers per wavefront 64, 128; 2, 4, 8 or 16 wavefronts per SIMD) on the members'
// current state -- `reps` launches each after one untimed -- and keeps the
// fastest.  The shapes give identical forces, so the choice changes speed only.
extern "C" int sesmd_ensemble_tune(sesmd_ensemble *e, int reps, int *shape3, float *best_ms)
{
    if (!e) return fail("null ensemble");
    if (reps < 1) return fail("reps must be >= 1");
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->m[0]->device) != hipSuccess ||
        cus < 1)
        cus = 256;
    {
        EnsStreams on(e, true);
        for (int i = 0; i < e->n; i++) forces_gather(e->m[i], step_what(e->m[i]), ~0ull, false);
    }
    hipEvent_t t0, t1;
    HC(hipEventCreate(&t0));
    HC(hipEventCreate(&t1));
    int bw = e->waves, bc = e->chunk, bg = e->grid;
    float best = -1.f;
    for (int waves : {1, 2, 4})
        for (int chunk : {64, 128})
            for (int per_simd : {2, 4, 8, 16, 32, 64}) {
                if (!e->mix && waves != 1) continue;
                e->waves = waves;
                e->chunk = chunk;
                e->grid = per_simd * 4 * cus / waves;
                ens_near(e);
                HC(hipEventRecord(t0, e->stream));
                for (int r = 0; r < reps; r++) ens_near(e);
                HC(hipEventRecord(t1, e->stream));
                HC(hipEventSynchronize(t1));
                float ms = 0.f;
                HC(hipEventElapsedTime(&ms, t0, t1));
                ms /= reps;
                if (best < 0.f || ms < best) {
                    best = ms;
                    bw = waves;
                    bc = chunk;
                    bg = e->grid;
                }
            }
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    e->waves = e->chunk = e->grid = 0;  // force the graphs to be captured again
    CK(sesmd_ensemble_set_shape(e, bw, bc, bg));
    if (shape3) {
        shape3[0] = bw;
        shape3[1] = bc;
        shape3[2] = bg;
    }
    if (best_ms) *best_ms = best;
    return 0;
}

// nsteps steps of every member (kb: the shared step counter, as in
// sesmd_steps): whole cycles replay one graph, other steps one step graph or,
// for a rebuild step outside a whole cycle, eager launches.
extern "C" int sesmd_ensemble_steps(sesmd_ensemble *e, int nsteps, int ifar, int *kb)
{
    if (!e || !kb) return fail("null argument");
    if (ifar < 1) return fail("ifar must be >= 1");
    if (nsteps < 0) return fail("nsteps must be >= 0");
    if (nsteps > 0) {
        if (e->cycle_len != ifar) {
            if (e->cycle_len) HC(hipGraphExecDestroy(e->g_cycle));
            e->cycle_len = 0;
            CK(ens_capture(e, ifar, true, &e->g_cycle));
            e->cycle_len = ifar;
        }
        if (!e->step_ok) {
            CK(ens_capture(e, 1, false, &e->g_step));
            e->step_ok = true;
        }
    }
    for (int s = 0; s < nsteps;) {
        if ((*kb + 1) % ifar == 0 && nsteps - s >= ifar) {
            HC(hipGraphLaunch(e->g_cycle, e->stream));
            for (int i = 0; i < e->n; i++) e->m[i]->boxlist = false;  // the replayed rebuild re-sorted
            *kb += ifar;
            s += ifar;
        } else if ((*kb + 1) % ifar == 0) {
            CK(ens_step(e, true));
            for (int i = 0; i < e->n; i++) e->m[i]->boxlist = false;
            (*kb)++;
            s++;
        } else {
            HC(hipGraphLaunch(e->g_step, e->stream));
            (*kb)++;
            s++;
        }
    }
    HC(hipStreamSynchronize(e->stream));
    return 0;
}

extern "C" int sesmd_get_classes(sesmd_ctx *c, int *cls, int *anc, int *cbox, int *rkind,
                                 unsigned *isub)
{
    if (!c) return fail("null context");
    const Geo &g = c->g;
    if (cls) HC(hipMemcpy(cls, c->cls, sizeof(int) * g.nmol, hipMemcpyDeviceToHost));
    if (anc) HC(hipMemcpy(anc, c->anc, sizeof(int) * g.nmol, hipMemcpyDeviceToHost));
    if (cbox) HC(hipMemcpy(cbox, c->cbox, sizeof(int) * g.n, hipMemcpyDeviceToHost));
    if (rkind) HC(hipMemcpy(rkind, c->rkind, sizeof(int) * g.n, hipMemcpyDeviceToHost));
    if (isub) HC(hipMemcpy(isub, c->isub, sizeof(unsigned) * g.n, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int sesmd_get_expansions(sesmd_ctx *c, double *O4, double *L4, double *O3, double *L3,
                                    double *O2, double *L2)
{
    if (!c) return fail("null context");
    const Geo &g = c->g;
    size_t s4 = sizeof(double2) * g.nb3 * g.lp, s3 = sizeof(double2) * g.n33 * g.lp,
           s2 = sizeof(double2) * g.n2 * g.lp;
    if (O4) HC(hipMemcpy(O4, c->O4, s4, hipMemcpyDeviceToHost));
    if (L4) HC(hipMemcpy(L4, c->L4, s4, hipMemcpyDeviceToHost));
    if (O3) HC(hipMemcpy(O3, c->O3, s3, hipMemcpyDeviceToHost));
    if (L3) HC(hipMemcpy(L3, c->L3, s3, hipMemcpyDeviceToHost));
    if (O2) HC(hipMemcpy(O2, c->O2, s2, hipMemcpyDeviceToHost));
    if (L2) HC(hipMemcpy(L2, c->L2, s2, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int sesmd_get_counters(sesmd_ctx *c, int *shake_fail)
{
    if (!c) return fail("null context");
    int t[4];
    HC(hipMemcpy(t, c->counters, sizeof t, hipMemcpyDeviceToHost));
    if (shake_fail) *shake_fail = t[0];
    return 0;
}

extern "C" int sesmd_device_arrays(sesmd_ctx *c, void **r, void **v, void **a, void **f)
{
    if (!c) return fail("null context");
    if (r) *r = c->r;
    if (v) *v = c->v;
    if (a) *a = c->a;
    if (f) *f = c->f;
    return 0;
}

extern "C" void *sesmd_stream(sesmd_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int sesmd_pair_counts(sesmd_ctx *c, double *pairs, double *lj_pairs)
{
    if (!c || !c->built) return fail("not built");
    const Geo &g = c->g;
    unsigned long long *cnt;
    HC(hipMalloc(&cnt, 2 * sizeof(unsigned long long)));
    HC(hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned long long), c->stream));
    k_gather_pos<<<nblk(g.n, 256), 256, 0, c->stream>>>(g, c->perm, c->r, c->type, c->dff, c->spos);
    k_pair_count<<<nblk(g.n, 128), 128, 0, c->stream>>>(g, c->spos, c->smeta, c->bstart, c->rkind,
                                                         c->dff, cnt);
    unsigned long long h[2];
    HC(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    HC(hipFree(cnt));
    if (pairs) *pairs = 0.5 * (double)h[0];
    if (lj_pairs) *lj_pairs = 0.5 * (double)h[1];
    return 0;
}

extern "C" int sesmd_pair_counts_noalloc(sesmd_ctx *c, double *counts)
{
    if (!c || !c->built) return fail("not built");
    if (!counts) return fail("null output");
    const Geo &g = c->g;
    int nb = (g.n + NOALLOC_BLOCK - 1) / NOALLOC_BLOCK;
    if (nb > NOALLOC_MAX_BLOCKS)
        return fail("sesmd_pair_counts_noalloc: %d sites exceed its maximum of %d", g.n,
                    NOALLOC_MAX_BLOCKS * NOALLOC_BLOCK);
    unsigned long long *part = (unsigned long long *)scratch_of(c->device);
    if (!part) return fail("sesmd_pair_counts_noalloc: no scratch on device %d", c->device);
    k_gather_pos<<<nblk(g.n, 256), 256, 0, c->stream>>>(g, c->perm, c->r, c->type, c->dff, c->spos);
    k_pair_count_part<<<nb, NOALLOC_BLOCK, 0, c->stream>>>(g, c->spos, c->smeta, c->bstart, c->rkind, c->dff,
                                                          part);
    k_pair_count_total<<<1, 256, 0, c->stream>>>(part, nb, counts);
    HC(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------ finer-grained API

extern "C" int sesmd_set_classes(sesmd_ctx *c, const int *cls, const int *anc)
{
    if (!c) return fail("null context");
    const Geo &g = c->g;
    if (!cls || !anc) {
        if (c->classes_fixed) drop_graphs(c);
        c->classes_fixed = false;
        return 0;
    }
    for (int m = 0; m < g.nmol; m++) {
        if (anc[m] < 0 || anc[m] >= g.nb3) return fail("storage box of molecule %d out of range", m);
        int want = m < g.nw ? (cls[m] == 0 || cls[m] == 1) : (cls[m] == 2);
        if (!want) return fail("class %d of molecule %d does not fit its kind", cls[m], m);
    }
    HC(hipMemcpy(c->cls, cls, sizeof(int) * g.nmol, hipMemcpyHostToDevice));
    HC(hipMemcpy(c->anc, anc, sizeof(int) * g.nmol, hipMemcpyHostToDevice));
    if (!c->classes_fixed) drop_graphs(c);  // a rebuild no longer classifies
    c->classes_fixed = true;
    return 0;
}

extern "C" int sesmd_get_thermostat(sesmd_ctx *c, double *eta, double *xi)
{
    if (!c) return fail("null context");
    double t[11];
    HC(hipMemcpy(t, c->nh, sizeof t, hipMemcpyDeviceToHost));
    if (eta) *eta = t[0];
    if (xi) *xi = t[10];
    return 0;
}

extern "C" int sesmd_set_thermostat(sesmd_ctx *c, double eta, double xi)
{
    if (!c) return fail("null context");
    HC(hipMemcpy(c->nh, &eta, sizeof(double), hipMemcpyHostToDevice));
    HC(hipMemcpy(c->nh + 10, &xi, sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int sesmd_get_flags(sesmd_ctx *c, int *flags, int n)
{
    if (!c) return fail("null context");
    if (n < 0 || (n > 0 && !flags)) return fail("bad flag buffer");
    int t[4];
    HC(hipMemcpy(t, c->counters, sizeof t, hipMemcpyDeviceToHost));
    for (int k = 0; k < n; k++) flags[k] = k < 4 ? t[k] : 0;
    return 0;
}

extern "C" int sesmd_set_eta(sesmd_ctx *c, double eta)
{
    if (!c) return fail("null context");
    HC(hipMemcpy(c->nh, &eta, sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

__global__ void k_nonempty(const double2 *__restrict__ O, int nbox, int lp, int *__restrict__ ne)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    int any = 0;
    for (int q = 0; q < lp; q++) any |= (O[(size_t)b * lp + q].x != 0.0 || O[(size_t)b * lp + q].y != 0.0);
    ne[b] = any;
}

extern "C" int sesmd_set_expansions(sesmd_ctx *c, const double *O4, const double *L4,
                                    const double *O3, const double *L3, const double *O2,
                                    const double *L2)
{
    if (!c) return fail("null context");
    const Geo &g = c->g;
    size_t s4 = sizeof(double2) * g.nb3 * g.lp, s3 = sizeof(double2) * g.n33 * g.lp,
           s2 = sizeof(double2) * g.n2 * g.lp;
    if (O4) HC(hipMemcpy(c->O4, O4, s4, hipMemcpyHostToDevice));
    if (L4) HC(hipMemcpy(c->L4, L4, s4, hipMemcpyHostToDevice));
    if (O3) HC(hipMemcpy(c->O3, O3, s3, hipMemcpyHostToDevice));
    if (L3) HC(hipMemcpy(c->L3, L3, s3, hipMemcpyHostToDevice));
    if (O2) HC(hipMemcpy(c->O2, O2, s2, hipMemcpyHostToDevice));
    if (L2) HC(hipMemcpy(c->L2, L2, s2, hipMemcpyHostToDevice));
    if (O4) k_nonempty<<<nblk(g.nb3, 128), 128, 0, c->stream>>>(c->O4, g.nb3, g.lp, c->ne4);
    if (O3) k_nonempty<<<nblk(g.n33, 128), 128, 0, c->stream>>>(c->O3, g.n33, g.lp, c->ne3);
    if (O2) k_nonempty<<<nblk(g.n2, 128), 128, 0, c->stream>>>(c->O2, g.n2, g.lp, c->ne2);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(c->stream));
    return 0;
}

// Enqueues one far-field stage on the context's stream without waiting.
static int enqueue_fmm_stage(sesmd_ctx *c, int stage)
{
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    switch (stage) {
    case SESMD_P2M:
        if (!c->built) return fail("P2M before sesmd_rebuild");
        if (g.p == 4)
            k_p2m<4><<<g.nb3, 64, 0, st>>>(g, c->r, c->perm, c->bstart, c->rkind, c->type, c->dff, c->O4,
                                           c->ne4);
        else
            k_p2m<8><<<g.nb3, 64, 0, st>>>(g, c->r, c->perm, c->bstart, c->rkind, c->type, c->dff, c->O4,
                                           c->ne4);
        break;
    case SESMD_M2M_43: k_m2m<<<g.n33, 512, 0, st>>>(g, g.n3, c->O4, c->ne4, c->M43, c->O3, c->ne3); break;
    case SESMD_M2M_32: k_m2m<<<g.n2, 512, 0, st>>>(g, g.nb2, c->O3, c->ne3, c->M32, c->O2, c->ne2); break;
    case SESMD_M2L_2:
        k_m2l_top_pair<<<g.n2 * g.n2, 64, 0, st>>>(g, c->O2, c->ne2, c->T2, c->top_part);
        k_m2l_top_sum<<<g.n2, 64, 0, st>>>(g, c->top_part, c->L2);
        break;
    case SESMD_L2L_23: k_l2l<<<g.n33, 64, 0, st>>>(g, g.n3, c->L2, c->S23, c->L3); break;
    case SESMD_M2L_3: launch_m2l(c, g.n3, c->O3, c->ne3, c->T3, c->L3, false); break;
    case SESMD_L2L_34: k_l2l<<<g.nb3, 64, 0, st>>>(g, g.nb, c->L3, c->S34, c->L4); break;
    case SESMD_M2L_4: launch_m2l(c, g.nb, c->O4, c->ne4, c->T4, c->L4, false); break;
    default: return fail("unknown far-field stage %d", stage);
    }
    HC(hipGetLastError());
    return 0;
}

extern "C" int sesmd_fmm_stage(sesmd_ctx *c, int stage)
{
    if (!c) return fail("null context");
    CK(enqueue_fmm_stage(c, stage));
    HC(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int sesmd_fmm_stage_enqueue(sesmd_ctx *c, int stage)
{
    if (!c) return fail("null context");
    return enqueue_fmm_stage(c, stage);
}

extern "C" int sesmd_move_species(sesmd_ctx *c, int part, unsigned cls_mask, double *sums18)
{
    if (!c) return fail("null context");
    if (!c->built && !c->classes_fixed) return fail("molecule classes unknown: rebuild or set_classes");
    if (part == 0) {
        CK(launch_move_a(c, cls_mask, 0));
        if (sums18) HC(hipMemcpyAsync(sums18, c->nh + 16, NH_VALS * sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream));
    } else {
        CK(launch_move_b(c, cls_mask));
    }
    HC(hipStreamSynchronize(c->stream));
    return 0;
}

// Unweighted |v dt|^2 per thermostat site group (nh_water1, nh_water2,
// nh_surfactant: src/nh_water1.f90:1-43, src/nh_surfactant.f90:1-201), one
// thread per molecule; groups as in the part-A kernels' first nine sums.
__global__ __launch_bounds__(256) void k_kinetic(Geo g, const double *__restrict__ v,
                                                 const int *__restrict__ cls,
                                                 double *__restrict__ partials)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[NH_VALS];
    for (int c = 0; c < NH_VALS; c++) acc[c] = 0.0;
    int n = g.n;
    if (m < g.nw) {
        int grp = cls[m] == 0 ? 0 : 2;
        for (int i = 0; i < 3; i++) {
            int s = 3 * m + i;
            acc[grp + (i == 0 ? 0 : 1)] += v[s] * v[s] + v[n + s] * v[n + s] + v[2 * n + s] * v[2 * n + s];
        }
    } else if (m < g.nmol) {
        int base = 3 * g.nw + 29 * (m - g.nw);
        for (int i = 0; i < 29; i++) {
            int s = base + i;
            acc[4 + c_sgrp[i]] += v[s] * v[s] + v[n + s] * v[n + s] + v[2 * n + s] * v[2 * n + s];
        }
    }
    block_sum_store<NH_VALS>(acc, partials + (size_t)blockIdx.x * NH_VALS);
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ partials, int nblocks,
                                                      double *__restrict__ out)
{
    double acc[NH_VALS];
    for (int c = 0; c < NH_VALS; c++) acc[c] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int c = 0; c < NH_VALS; c++) acc[c] += partials[(size_t)b * NH_VALS + c];
    block_sum_store<NH_VALS>(acc, out);
}

// ------------------------------------------------------------ energies
//
// Not computed by the reference: the potentials whose gradients are the force
// terms, for energy and drift reporting.  Each kernel writes one value per
// site and component; k_sum_cols adds them in a fixed order (deterministic).
// Pair terms are seen from both ends and weighted 1/2.

constexpr int E_SITE_COLS = 6;   // near Coulomb, near LJ, virial, far, wall, graphite (sorted sites)
constexpr int E_INTRA_COLS = 6;  // internal Coulomb, internal LJ, subtract, angle, dihedral, Berendsen

// LJ energy of the force c r^-6 (s6 r^-6 - 1/2) r / r^2
__device__ __forceinline__ double lj_energy(double c, double s6, double rsq)
{
    double rsi = 1.0 / rsq, r6 = rsi * rsi * rsi;
    return c / 12.0 * r6 * (s6 * r6 - 1.0);
}

template <int PP>
__global__ __launch_bounds__(128) void k_energy_sites(
    Geo g, const double4 *__restrict__ spos, const int4 *__restrict__ smeta, const int *__restrict__ bstart,
    const int *__restrict__ rkind, const int *__restrict__ cbox, const double2 *__restrict__ L4,
    const double *__restrict__ lfac, const sesmd_forcefield *__restrict__ ff, const double *__restrict__ gr,
    const int *__restrict__ gstart, const int *__restrict__ gidx, int ngr, double *__restrict__ E)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.n) return;
    double4 pi = spos[k];
    int4 me = smeta[k];
    int i = me.x, ti = me.y & 255, rk = rkind[i];
    double ec = 0.0, el = 0.0, vir = 0.0, far = 0.0, wl = 0.0, egr = 0.0;
    int nb = g.nb;
    if (rk >= 0) {
        int b = cbox[i], bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
        double qi = pi.w * g.xpi;
        // near field: the pair set of k_near (same ring rule, other molecules)
        for (int z = max(bz - 1, 0); z <= min(bz + 1, nb - 1); z++)
            for (int y = max(by - 1, 0); y <= min(by + 1, nb - 1); y++)
                for (int x = max(bx - 1, 0); x <= min(bx + 1, nb - 1); x++) {
                    int bb = x + nb * (y + nb * z);
                    for (int t = bstart[bb]; t < bstart[bb + 1]; t++) {
                        int4 mj = smeta[t];
                        if (mj.z == me.z) continue;
                        int cls = (mj.y >> 8) & 3;
                        int ring = max(abs((mj.w & 255) - bx),
                                       max(abs(((mj.w >> 8) & 255) - by), abs(((mj.w >> 16) & 255) - bz)));
                        int lim = cls == 0 ? 1 : cls == 1 ? 2 : 3;
                        if (ring > lim) continue;
                        double4 pj = spos[t];
                        double u = pi.x - pj.x, v = pi.y - pj.y, w = pi.z - pj.z;
                        double rsq = u * u + v * v + w * w, r = sqrt(rsq);
                        double qq = qi * pj.w;
                        ec += 0.5 * qq / r;
                        vir += 0.5 * qq / r;
                        int tj = mj.y & 255;
                        double c = ff->lj_c[ti * 32 + tj];
                        if (c != 0.0 && r <= g.rc) {
                            double s6 = ff->lj_s6[ti * 32 + tj];
                            double rsi = 1.0 / rsq, r6 = rsi * rsi * rsi;
                            el += 0.5 * lj_energy(c, s6, rsq);
                            vir += 0.5 * c * r6 * (s6 * r6 - 0.5);
                        }
                    }
                }
        // far field: 1/2 q_i Phi(x_i), Phi = Re sum L_jk lfac_jk e^{ik phi} r^j P_j^|k|
        constexpr int p = PP;
        double rxpb = pi.x - d_cen4(g, bx), rypb = pi.y - d_cen4(g, by), rzpb = pi.z - d_cen4(g, bz);
        double rho2 = rxpb * rxpb + rypb * rypb;
        double rr = sqrt(rho2 + rzpb * rzpb);
        double sr = sqrt(rho2);
        double2 rp2 = make_double2(rxpb / sr, rypb / sr);
        double P[MAXP + 2][MAXP + 1];
        legendre_table<p>(rzpb / rr, P);
        double2 pw[MAXP + 1];
        pw[0] = make_double2(1.0, 0.0);
#pragma unroll
        for (int k2 = 1; k2 <= p; k2++) pw[k2] = cmul(pw[k2 - 1], rp2);
        const double2 *L = L4 + (size_t)b * g.lp;
        double phi = 0.0, rj = 1.0;
#pragma unroll
        for (int j = 0; j <= p; j++) {
#pragma unroll
            for (int kk = -j; kk <= j; kk++) {
                int ak = kk < 0 ? -kk : kk;
                double2 e = pw[ak];
                if (kk < 0) {
                    double d = e.x * e.x + e.y * e.y;
                    e = make_double2(e.x / d, -e.y / d);
                }
                double2 l = L[j * (j + 1) + kk];
                double sc = lfac[j * (j + 1) + kk];
                phi += (l.x * sc * e.x - l.y * sc * e.y) * rj * P[j + 1][ak];
            }
            rj *= rr;
        }
        far = 0.5 * qi * phi;
        // graphite: receivers filed in their own box of the bottom layer
        if (ngr > 0 && b < nb * nb && (rk == 0 || rk == 1 || rk == 3)) {
            double c = ff->gr_c[ti], s6 = ff->gr_s6[ti];
            if (c != 0.0)
                for (int q = gstart[b]; q < gstart[b + 1]; q++) {
                    int a = gidx[q];
                    double u = pi.x - gr[3 * a], v = pi.y - gr[3 * a + 1], w = pi.z - gr[3 * a + 2];
                    double rsq = u * u + v * v + w * w;
                    if (sqrt(rsq) <= g.rc) egr += lj_energy(c, s6, rsq);
                }
        }
    }
    {
        double h = g.rhs, xs[3] = {pi.x, pi.y, pi.z};
        for (int a = 0; a < 3; a++) {
            double h1 = xs[a] - h, h2 = xs[a] + h;
            double u1 = 1.0 / (h1 * h1), u2 = 1.0 / (h2 * h2);
            double u16 = u1 * u1 * u1, u26 = u2 * u2 * u2;
            wl += 4.0 * u16 * u16 + 4.0 * u26 * u26;
        }
    }
    size_t n = g.n;
    E[k] = ec;
    E[n + k] = el;
    E[2 * n + k] = vir;
    E[3 * n + k] = far;
    E[4 * n + k] = wl;
    E[5 * n + k] = egr;
}

// Intramolecular energies, one thread per surfactant site (the term lists of
// k_intra): pair terms weighted 1/2 (listed at both sites), each angle and
// dihedral counted at its first site.
__global__ __launch_bounds__(64) void k_energy_intra(Geo g, const double *__restrict__ r,
                                                     const unsigned *__restrict__ isub,
                                                     const sesmd_forcefield *__restrict__ ff,
                                                     const Topo *__restrict__ topo, double *__restrict__ E)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int m = 29 * g.ns;
    if (t >= m) return;
    int s = t / 29, j = t % 29;
    int base = 3 * g.nw + 29 * s, n = g.n;
    double xs[3];
    ld3(r, n, base + j, xs);
    double ic = 0.0, il = 0.0, sb = 0.0, an = 0.0, dh = 0.0, db = 0.0;
    for (int q = 0; q < topo->npair[j]; q++) {
        int k = topo->pair[j][q];
        double xk[3];
        ld3(r, n, base + k, xk);
        double u = xs[0] - xk[0], v = xs[1] - xk[1], w = xs[2] - xk[2];
        double rsq = u * u + v * v + w * w;
        int tj = 3 + j, tk = 3 + k;
        ic += 0.5 * ff->q[tj] * ff->q[tk] * g.xpi / sqrt(rsq);
        il += 0.5 * lj_energy(ff->lj_c[tj * 32 + tk], ff->lj_s6[tj * 32 + tk], rsq);
    }
    unsigned mask = isub[base + j];
    while (mask) {
        int k = __ffs(mask) - 1;
        mask &= mask - 1;
        double xk[3];
        ld3(r, n, base + k, xk);
        double u = xs[0] - xk[0], v = xs[1] - xk[1], w = xs[2] - xk[2];
        sb -= 0.5 * ff->q[3 + j] * ff->q[3 + k] * g.xpi / sqrt(u * u + v * v + w * w);
    }
    for (int q = 0; q < topo->nang[j]; q++) {
        if (topo->ang[j][q] % 3 != 0) continue;
        int a = topo->ang[j][q] / 3;
        double ri[3], rjv[3], rk[3];
        ld3(r, n, base + ff->ang_i[a] - 1, ri);
        ld3(r, n, base + ff->ang_j[a] - 1, rjv);
        ld3(r, n, base + ff->ang_k[a] - 1, rk);
        double rxji = ri[0] - rjv[0], ryji = ri[1] - rjv[1], rzji = ri[2] - rjv[2];
        double rxki = rk[0] - rjv[0], ryki = rk[1] - rjv[1], rzki = rk[2] - rjv[2];
        double dp = rxji * rxki + ryji * ryki + rzji * rzki;
        double rjin = sqrt(rxji * rxji + ryji * ryji + rzji * rzji);
        double rkin = sqrt(rxki * rxki + ryki * ryki + rzki * rzki);
        double d = acos(dp / rjin / rkin) - ff->th[a];
        an += 0.5 * ff->rkth[a] * d * d;
    }
    for (int q = 0; q < topo->ndih[j]; q++) {
        if (topo->dih[j][q] % 4 != 0) continue;
        int d = topo->dih[j][q] / 4;
        double Ri[3], Rj[3], Rk[3], Rl[3];
        ld3(r, n, base + ff->dih_i[d] - 1, Ri);
        ld3(r, n, base + ff->dih_j[d] - 1, Rj);
        ld3(r, n, base + ff->dih_k[d] - 1, Rk);
        ld3(r, n, base + ff->dih_l[d] - 1, Rl);
        double rxji = Rj[0] - Ri[0], ryji = Rj[1] - Ri[1], rzji = Rj[2] - Ri[2];
        double rxkj = Rk[0] - Rj[0], rykj = Rk[1] - Rj[1], rzkj = Rk[2] - Rj[2];
        double rxnk = Rl[0] - Rk[0], rynk = Rl[1] - Rk[1], rznk = Rl[2] - Rk[2];
        double ax = rxji * rykj - ryji * rxkj, ay = rxji * rzkj - rzji * rxkj, az = ryji * rzkj - rzji * rykj;
        double bx = rxkj * rynk - rykj * rxnk, by = rxkj * rznk - rzkj * rxnk, bz = rykj * rznk - rzkj * rynk;
        double na = sqrt(ax * ax + ay * ay + az * az), nbn = sqrt(bx * bx + by * by + bz * bz);
        double arg = (ax * bx + ay * by + az * bz) / na / nbn;
        double clamp = d < 18 ? ff->clamp_opls : ff->clamp_be;
        if (arg <= -1.0) arg = -clamp;
        if (arg >= 1.0) arg = clamp;
        double th = acos(arg);
        if (d < 18)
            dh += ff->dva[d] * (1.0 + cos(th)) + ff->dvb[d] * (1.0 - cos(2.0 * th)) + ff->dvc[d] * (1.0 + cos(3.0 * th));
        else
            db += ff->dva[d] * (1.0 + cos(3.0 * th));
    }
    E[t] = ic;
    E[m + t] = il;
    E[2 * m + t] = sb;
    E[3 * m + t] = an;
    E[4 * m + t] = dh;
    E[5 * m + t] = db;
}

// out[c] = sum over k of E[c * n + k] for c < NC: one block, each thread a
// fixed stride, then a fixed tree (deterministic)
template <int NC>
__global__ __launch_bounds__(256) void k_sum_cols(const double *__restrict__ E, int n, double *__restrict__ out)
{
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c] += E[(size_t)c * n + k];
    block_sum_store<NC>(acc, out);
}

extern "C" int sesmd_energies(sesmd_ctx *c, double *e)
{
    if (!c || !e) return fail("null argument");
    if (!c->built) return fail("sesmd_energies before sesmd_rebuild");
    const Geo &g = c->g;
    hipStream_t st = c->stream;
    // the near-field partials are free between steps: per-site values go there
    double *E = c->npart, *EI = c->npart + (size_t)E_SITE_COLS * g.n;
    k_gather_pos<<<nblk(g.n, 256), 256, 0, st>>>(g, c->perm, c->r, c->type, c->dff, c->spos);
    (g.p == 4 ? k_energy_sites<4> : k_energy_sites<8>)<<<nblk(g.n, 128), 128, 0, st>>>(
        g, c->spos, c->smeta, c->bstart, c->rkind, c->cbox, c->L4, c->lfac, c->dff, c->gr, c->gstart,
        c->gidx, c->ngr, E);
    k_sum_cols<E_SITE_COLS><<<1, 256, 0, st>>>(E, g.n, c->esum);
    if (g.ns) {
        k_energy_intra<<<nblk(29 * g.ns, 64), 64, 0, st>>>(g, c->r, c->isub, c->dff, c->dtopo, EI);
        k_sum_cols<E_INTRA_COLS><<<1, 256, 0, st>>>(EI, 29 * g.ns, c->esum + E_SITE_COLS);
    } else {
        k_fill_d<<<1, 64, 0, st>>>(c->esum + E_SITE_COLS, E_INTRA_COLS, 0.0);
    }
    int nbk = nblk(g.nmol, 256);
    if (nbk > c->nblk_w + c->nblk_s + 1) return fail("energies: partial buffer too small");
    k_kinetic<<<nbk, 256, 0, st>>>(g, c->v, c->cls, c->partials);
    k_sum_partials<<<1, 256, 0, st>>>(c->partials, nbk, c->nh + 16);
    HC(hipGetLastError());
    double es[E_SITE_COLS + E_INTRA_COLS], nh[16 + NH_VALS];
    HC(hipMemcpyAsync(es, c->esum, sizeof es, hipMemcpyDeviceToHost, st));
    HC(hipMemcpyAsync(nh, c->nh, sizeof nh, hipMemcpyDeviceToHost, st));
    HC(hipStreamSynchronize(st));
    for (int k = 0; k < SESMD_E_COUNT; k++) e[k] = 0.0;
    e[SESMD_E_NEAR_COUL] = es[0];
    e[SESMD_E_NEAR_LJ] = es[1];
    e[SESMD_E_VIRIAL] = es[2];
    e[SESMD_E_FAR] = es[3];
    e[SESMD_E_WALL] = es[4];
    e[SESMD_E_GRAPHITE] = es[5];
    e[SESMD_E_INTERNAL_COUL] = es[6];
    e[SESMD_E_INTERNAL_LJ] = es[7];
    e[SESMD_E_SUBTRACT] = es[8];
    e[SESMD_E_ANGLE] = es[9];
    e[SESMD_E_DIHEDRAL] = es[10];
    e[SESMD_E_DIHEDRAL_BE] = es[11];
    double pot = 0.0;
    for (int k = 0; k < SESMD_E_POTENTIAL; k++) pot += e[k];
    e[SESMD_E_POTENTIAL] = pot;
    // kinetic: the nine group sums of |v dt|^2 weighted by the group masses
    const double *t = nh + 16, *m = c->ff.mass;
    double mv2 = m[0] * (t[0] + t[2]) + m[1] * (t[1] + t[3]) + m[3] * t[4] + m[4] * t[5] + m[5] * t[6] +
                 m[3 + 17] * t[7] + m[3 + 18] * t[8];
    double dt = g.dt, kin = 0.5 * mv2 / (dt * dt);
    e[SESMD_E_KINETIC] = kin;
    int df = 3 * g.n - (3 * g.nw + 28 * g.ns);
    e[SESMD_E_TEMPERATURE] = df > 0 ? 2.0 * kin / df : 0.0;
    double eta = nh[0], pnh1 = nh[8], pnh2 = nh[9], xi = nh[10];
    e[SESMD_E_THERMOSTAT] = pnh1 != 0.0 ? 0.25 * dt / pnh1 * eta * eta - 0.5 * pnh2 * xi : 0.0;
    e[SESMD_E_TOTAL] = pot + kin + e[SESMD_E_THERMOSTAT];
    return 0;
}

extern "C" int sesmd_kinetic_sums(sesmd_ctx *c, double *sums9)
{
    if (!c) return fail("null context");
    if (!c->built && !c->classes_fixed) return fail("molecule classes unknown: rebuild or set_classes");
    const Geo &g = c->g;
    int nb = nblk(g.nmol, 256);
    if (nb > c->nblk_w + c->nblk_s + 1) return fail("kinetic sums: partial buffer too small");
    k_kinetic<<<nb, 256, 0, c->stream>>>(g, c->v, c->cls, c->partials);
    k_sum_partials<<<1, 256, 0, c->stream>>>(c->partials, nb, c->nh + 16);
    HC(hipGetLastError());
    double t[NH_VALS];
    HC(hipMemcpyAsync(t, c->nh + 16, sizeof t, hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    if (sums9) memcpy(sums9, t, 9 * sizeof(double));
    return 0;
}
