/* CPU restatement of the reference's per-timestep force and integration path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle the HIP path is checked against and the
 * CPU baseline bench.py times).  Compiled with -ffp-contract=off so that every
 * expression is evaluated in the order written, like the reference's.  All
 * citations are to files of the reference (src/...).
 */
#include "sesmd_oracle.h"

#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

typedef struct {
    int nw, ns, n, nmol, nb2, nb, nb3, n3, n33, n2, p, lp;
    double rs, rsd, rhs;
    orc_ff ff;
    double *r, *v, *a;
    double eta, tr, pnh1, pnh2;
    double xi;               /* time integral of eta (thermostat energy) */
    int ngr;
    double *gr;
    int *type, *mol, *first; /* per site: type, molecule; per molecule: first site */
    int *cls, *anc;          /* per molecule */
    int *cbox, *rkind;       /* per site */
    unsigned *isub;          /* per site */
    int *bstart, *bsite;     /* sites sorted by cbox */
    double *c4, *c3, *c2;    /* box centres (3 per box) */
    cplx *O4, *L4, *O3, *L3, *O2, *L2;
    int *gstart, *gidx;      /* graphite atoms per bottom-layer column */
    int shake_fail;
} orc;

/* ------------------------------------------------------------------ helpers */

static int imax(int a, int b) { return a > b ? a : b; }
static int iabs_(int a) { return a < 0 ? -a : a; }

/* x**n for integer n, binary exponentiation (how an integer power with a
 * run-time exponent is evaluated) */
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

static cplx cipow(cplx a, int b)
{
    int recip = b < 0;
    if (recip) b = -b;
    cplx r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

/* rLegendreP, src/rLegendreP.f90:3-48 (associated Legendre function with the
 * Condon-Shortley phase, Numerical Recipes recurrence) */
static double legendre(int l, int m, double xx)
{
    if (l == -1 && m == 0) return 1.0;
    if (m > l) return 0.0;
    double pmm = 1.0;
    if (m > 0) {
        double somx2 = sqrt((1.0 - xx) * (1.0 + xx));
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

/* Ylm, src/ifma.f90:12-26.  `cmplx(cos(m b), sin(m b))` has no kind argument,
 * so the phase factor is rounded to single precision. */
static cplx ylm(const orc *o, int n, int m, double a, double b)
{
    int am = iabs_(m);
    double ri1 = o->ff.rifact[n - am], ri2 = o->ff.rifact[n + am];
    double rlg = legendre(n, am, cos(a));
    float cr = (float)cos(m * b), ci = (float)sin(m * b);
    double s = sqrt(ri1 / ri2) * rlg;
    return s * (double)cr + I * (s * (double)ci);
}

void orc_ylm(void *h, int n, int m, double a, double b, double *out)
{
    cplx y = ylm((orc *)h, n, m, a, b);
    out[0] = creal(y);
    out[1] = cimag(y);
}

double orc_legendre(int l, int m, double x) { return legendre(l, m, x); }

/* rAnm, src/ifma.f90:41-52 */
static double anm(const orc *o, int n, int m)
{
    double ri1 = o->ff.rifact[n - m], ri2 = o->ff.rifact[n + m];
    return (double)((n & 1) ? -1 : 1) / sqrt(ri1 * ri2);
}
static int sgnpow(int e) { return (e & 1) ? -1 : 1; } /* (-1)**e, any sign of e */
static int ikm1(int a, int b) { return sgnpow((iabs_(a) - iabs_(b) - iabs_(a - b)) / 2); }
static int ikm2(int a, int b) { return sgnpow((iabs_(a - b) - iabs_(a) - iabs_(b)) / 2); }
static int ikm3(int a, int b) { return sgnpow((iabs_(b) - iabs_(b - a) - iabs_(a)) / 2); }
static int idx(int n, int m) { return n * (n + 1) + m; } /* igetIndex - 1 */

/* --------------------------------------------------------------- geometry */

/* L2 centre along one axis (get_box_center, src/init_box_dimensions.f90:40-42) */
static double rcen2(const orc *o, int ix)
{
    return o->rsd * (ix + 1) / (double)o->nb2 - o->rhs - o->rsd / (double)o->nb2 / 2.0;
}

static void centres(orc *o)
{
    double rs = o->rs;
    for (int g = 0; g < o->nb; g++) {
        /* rcen4 of src/init_box_dimensions.f90:66-73 for local box i of rank g/4 */
        int i = g % 4 + 1;
        double c = 4.0 * rs * i / 4.0 - 2.0 * rs - 4.0 * rs / 8.0 + rcen2(o, g / 4);
        for (int a = 0; a < 3; a++) (void)a;
        o->c4[3 * g] = c; /* 1-D table, expanded below */
    }
    double *c1 = malloc(sizeof(double) * o->nb);
    for (int g = 0; g < o->nb; g++) c1[g] = o->c4[3 * g];
    for (int z = 0; z < o->nb; z++)
        for (int y = 0; y < o->nb; y++)
            for (int x = 0; x < o->nb; x++) {
                int b = x + o->nb * (y + o->nb * z);
                o->c4[3 * b] = c1[x];
                o->c4[3 * b + 1] = c1[y];
                o->c4[3 * b + 2] = c1[z];
            }
    free(c1);
    int n3 = o->nb / 2;
    double *d1 = malloc(sizeof(double) * n3);
    for (int g = 0; g < n3; g++) {
        int i = g % 2 + 1; /* rcen3, src/init_box_dimensions.f90:55-61 */
        d1[g] = 4.0 * rs * i / 2.0 - 2.0 * rs - 4.0 * rs / 4.0 + rcen2(o, g / 2);
    }
    for (int z = 0; z < n3; z++)
        for (int y = 0; y < n3; y++)
            for (int x = 0; x < n3; x++) {
                int b = x + n3 * (y + n3 * z);
                o->c3[3 * b] = d1[x];
                o->c3[3 * b + 1] = d1[y];
                o->c3[3 * b + 2] = d1[z];
            }
    free(d1);
    for (int z = 0; z < o->nb2; z++)
        for (int y = 0; y < o->nb2; y++)
            for (int x = 0; x < o->nb2; x++) {
                int b = x + o->nb2 * (y + o->nb2 * z);
                o->c2[3 * b] = rcen2(o, x);
                o->c2[3 * b + 1] = rcen2(o, y);
                o->c2[3 * b + 2] = rcen2(o, z);
            }
}

/* ibox2, src/ibox.f90:127-139 (int() truncates toward zero) */
static int ibox2(const orc *o, const double *x)
{
    int nb2 = o->nb2;
    return (int)((x[0] + o->rhs) * nb2 / o->rsd) + (int)((x[1] + o->rhs) * nb2 / o->rsd) * nb2 +
           (int)((x[2] + o->rhs) * nb2 / o->rsd) * nb2 * nb2;
}

/* ibox64 relative to the L2 centre of rank rk, src/ibox.f90:172-186; returns
 * local (i, j, k) in 0..3 through l */
static int ibox64(const orc *o, const double *x, int rk, int *l)
{
    int rx = rk % o->nb2, ry = (rk / o->nb2) % o->nb2, rz = rk / (o->nb2 * o->nb2);
    double rap = x[0] - rcen2(o, rx) + 2.0 * o->rs;
    double rbp = x[1] - rcen2(o, ry) + 2.0 * o->rs;
    double rcp = x[2] - rcen2(o, rz) + 2.0 * o->rs;
    l[0] = (int)(rap / o->rs);
    l[1] = (int)(rbp / o->rs);
    l[2] = (int)(rcp / o->rs);
    return l[0] + l[1] * 4 + l[2] * 16;
}

static int glob_of(const orc *o, int rk, const int *l)
{
    int rx = rk % o->nb2, ry = (rk / o->nb2) % o->nb2, rz = rk / (o->nb2 * o->nb2);
    int gx = 4 * rx + l[0], gy = 4 * ry + l[1], gz = 4 * rz + l[2];
    if (gx < 0) gx = 0;
    if (gy < 0) gy = 0;
    if (gz < 0) gz = 0;
    if (gx >= o->nb) gx = o->nb - 1;
    if (gy >= o->nb) gy = o->nb - 1;
    if (gz >= o->nb) gz = o->nb - 1;
    return gx + o->nb * (gy + o->nb * gz);
}

static int cheb(const orc *o, int b1, int b2)
{
    int nb = o->nb;
    int dx = iabs_(b1 % nb - b2 % nb), dy = iabs_((b1 / nb) % nb - (b2 / nb) % nb),
        dz = iabs_(b1 / (nb * nb) - b2 / (nb * nb));
    return imax(dx, imax(dy, dz));
}

/* isCentral along one axis, src/ibox.f90:101-124 */
static int central1(double x, double c, double rs) { return x < c + rs / 2.0 && x >= c - rs / 2.0; }

/* receiver / source kinds: receivers W1@0 W2@0 W2@1 S@0 S@1 S@2 (0..5), sources
 * W1@0 W1@1 W2@0 W2@1 W2@2 S@0 S@1 S@2 S@3 (0..8); -1 = not in any list */
static int recv_kind(int cls, int ring)
{
    if (cls == 0) return ring == 0 ? 0 : -1;
    if (cls == 1) return ring <= 1 ? 1 + ring : -1;
    return ring <= 2 ? 3 + ring : -1;
}
static int src_kind(int cls, int ring)
{
    if (cls == 0) return ring <= 1 ? ring : -1;
    if (cls == 1) return ring <= 2 ? 2 + ring : -1;
    return ring <= 3 ? 5 + ring : -1;
}

/* ----------------------------------------------------------------- create */

void *orc_create(int n_water, int n_surf, int nb2, int p, const orc_ff *ff)
{
    orc *o = calloc(1, sizeof(orc));
    o->nw = n_water;
    o->ns = n_surf;
    o->n = 3 * n_water + 29 * n_surf;
    o->nmol = n_water + n_surf;
    o->nb2 = nb2;
    o->nb = 4 * nb2;
    o->nb3 = o->nb * o->nb * o->nb;
    o->n3 = o->nb / 2;
    o->n33 = o->n3 * o->n3 * o->n3;
    o->n2 = nb2 * nb2 * nb2;
    o->p = p;
    o->lp = (p + 1) * (p + 1);
    o->ff = *ff;
    o->rs = ff->rs;
    o->rsd = o->nb * o->rs;
    o->rhs = o->rsd / 2.0;
    int n = o->n;
    o->r = calloc(3 * (size_t)n, sizeof(double));
    o->v = calloc(3 * (size_t)n, sizeof(double));
    o->a = calloc(3 * (size_t)n, sizeof(double));
    o->type = malloc(sizeof(int) * n);
    o->mol = malloc(sizeof(int) * n);
    o->first = malloc(sizeof(int) * (o->nmol + 1));
    for (int w = 0; w < n_water; w++) {
        o->first[w] = 3 * w;
        for (int k = 0; k < 3; k++) {
            o->type[3 * w + k] = k;
            o->mol[3 * w + k] = w;
        }
    }
    for (int s = 0; s < n_surf; s++) {
        int b = 3 * n_water + 29 * s;
        o->first[n_water + s] = b;
        for (int k = 0; k < 29; k++) {
            o->type[b + k] = 3 + k;
            o->mol[b + k] = n_water + s;
        }
    }
    o->first[o->nmol] = n;
    o->cls = malloc(sizeof(int) * o->nmol);
    o->anc = malloc(sizeof(int) * o->nmol);
    o->cbox = malloc(sizeof(int) * n);
    o->rkind = malloc(sizeof(int) * n);
    o->isub = calloc(n, sizeof(unsigned));
    o->bstart = malloc(sizeof(int) * (o->nb3 + 1));
    o->bsite = malloc(sizeof(int) * n);
    o->c4 = malloc(sizeof(double) * 3 * o->nb3);
    o->c3 = malloc(sizeof(double) * 3 * o->n33);
    o->c2 = malloc(sizeof(double) * 3 * o->n2);
    o->O4 = calloc((size_t)o->nb3 * o->lp, sizeof(cplx));
    o->L4 = calloc((size_t)o->nb3 * o->lp, sizeof(cplx));
    o->O3 = calloc((size_t)o->n33 * o->lp, sizeof(cplx));
    o->L3 = calloc((size_t)o->n33 * o->lp, sizeof(cplx));
    o->O2 = calloc((size_t)o->n2 * o->lp, sizeof(cplx));
    o->L2 = calloc((size_t)o->n2 * o->lp, sizeof(cplx));
    o->gstart = calloc(o->nb * o->nb + 1, sizeof(int));
    o->gidx = NULL;
    centres(o);
    return o;
}

void orc_destroy(void *h)
{
    orc *o = h;
    if (!o) return;
    free(o->r); free(o->v); free(o->a); free(o->type); free(o->mol); free(o->first);
    free(o->cls); free(o->anc); free(o->cbox); free(o->rkind); free(o->isub);
    free(o->bstart); free(o->bsite); free(o->c4); free(o->c3); free(o->c2);
    free(o->O4); free(o->L4); free(o->O3); free(o->L3); free(o->O2); free(o->L2);
    free(o->gstart); free(o->gidx); free(o->gr);
    free(o);
}

void orc_set_threads(int n) { omp_set_num_threads(n); }

void orc_set_state(void *h, const double *r, const double *v, const double *a, double eta)
{
    orc *o = h;
    size_t nb = 3 * (size_t)o->n * sizeof(double);
    if (r) memcpy(o->r, r, nb);
    if (v) memcpy(o->v, v, nb);
    if (a) memcpy(o->a, a, nb);
    o->eta = eta;
}

void orc_get_state(void *h, double *r, double *v, double *a, double *eta)
{
    orc *o = h;
    size_t nb = 3 * (size_t)o->n * sizeof(double);
    if (r) memcpy(r, o->r, nb);
    if (v) memcpy(v, o->v, nb);
    if (a) memcpy(a, o->a, nb);
    if (eta) *eta = o->eta;
}

void orc_set_nh(void *h, double tr, double pnh1, double pnh2)
{
    orc *o = h;
    o->tr = tr;
    o->pnh1 = pnh1;
    o->pnh2 = pnh2;
}

void orc_set_graphite(void *h, int n, const double *xyz)
{
    orc *o = h;
    free(o->gr);
    o->ngr = n;
    o->gr = malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    if (n) memcpy(o->gr, xyz, sizeof(double) * 3 * (size_t)n);
    /* per bottom-layer column: atoms within rc + sqrt(2) rs / 2 in xy of the
     * column centre, src/read_graphite.f90:203-229 */
    int nc = o->nb * o->nb;
    double rcut = (o->ff.rc + sqrt(2.0) * o->rs / 2.0);
    rcut = rcut * rcut;
    int *cnt = calloc(nc, sizeof(int));
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            o->gstart[0] = 0;
            for (int c = 0; c < nc; c++) o->gstart[c + 1] = o->gstart[c] + cnt[c];
            free(o->gidx);
            o->gidx = malloc(sizeof(int) * (o->gstart[nc] + 1));
            memset(cnt, 0, sizeof(int) * nc);
        }
        for (int c = 0; c < nc; c++) {
            double cx = o->c4[3 * c], cy = o->c4[3 * c + 1];
            for (int g = 0; g < n; g++) {
                double dx = o->gr[3 * g] - cx, dy = o->gr[3 * g + 1] - cy;
                double rsq = dx * dx + dy * dy;
                if (rsq <= rcut) {
                    if (pass == 1) o->gidx[o->gstart[c] + cnt[c]] = g;
                    cnt[c]++;
                }
            }
        }
    }
    free(cnt);
}

/* ----------------------------------------------------------------- rebuild */

/* Molecule classes and storage boxes (src/read_config.f90:44-175,
 * src/delete_w1.f90:31-120), site boxes (isCentral against the L4 centres,
 * the test the list builders make, src/buildW2.f90:38, src/buildS.f90:269),
 * and the intramolecular subtraction lists (src/buildinternal.f90:11-63). */
void orc_rebuild(void *h)
{
    orc *o = h;
    int nb = o->nb;
    for (int m = 0; m < o->nmol; m++) {
        int f = o->first[m];
        if (m < o->nw) {
            const double *O = o->r + 3 * f, *G = O + 3, *H = O + 6;
            int ibO = ibox2(o, O), ibG = ibox2(o, G), ibH = ibox2(o, H);
            int lO[3], lG[3], lH[3];
            int bO = ibox64(o, O, ibO, lO);
            int w1 = 0;
            if (ibO == ibG && ibO == ibH) {
                int bG = ibox64(o, G, ibO, lG), bH = ibox64(o, H, ibO, lH);
                w1 = (bO == bG && bO == bH);
            }
            o->cls[m] = w1 ? 0 : 1;
            o->anc[m] = glob_of(o, ibO, lO);
        } else {
            const double *X = o->r + 3 * (f + 13); /* site 14 */
            int rk = ibox2(o, X), l[3];
            ibox64(o, X, rk, l);
            o->cls[m] = 2;
            o->anc[m] = glob_of(o, rk, l);
        }
    }
    for (int i = 0; i < o->n; i++) {
        int m = o->mol[i];
        int b;
        if (o->cls[m] == 0) {
            b = o->anc[m];
        } else {
            int g[3];
            for (int a = 0; a < 3; a++) {
                double x = o->r[3 * i + a];
                int c = (int)floor((x + o->rhs) / o->rs);
                if (c < 0) c = 0;
                if (c >= nb) c = nb - 1;
                /* settle on the box whose half-open cube holds x */
                int best = c;
                for (int d = -1; d <= 1; d++) {
                    int cc = c + d;
                    if (cc < 0 || cc >= nb) continue;
                    if (central1(x, o->c4[3 * cc], o->rs)) { best = cc; break; }
                }
                g[a] = best;
            }
            b = g[0] + nb * (g[1] + nb * g[2]);
        }
        o->cbox[i] = b;
        o->rkind[i] = recv_kind(o->cls[m], cheb(o, o->anc[m], b));
    }
    /* sites grouped by box, in site order */
    memset(o->bstart, 0, sizeof(int) * (o->nb3 + 1));
    for (int i = 0; i < o->n; i++) o->bstart[o->cbox[i] + 1]++;
    for (int b = 0; b < o->nb3; b++) o->bstart[b + 1] += o->bstart[b];
    int *fill = malloc(sizeof(int) * o->nb3);
    memcpy(fill, o->bstart, sizeof(int) * o->nb3);
    for (int i = 0; i < o->n; i++) o->bsite[fill[o->cbox[i]]++] = i;
    free(fill);
    /* isub: partners outside the 3^3 block around the site's box, where the
     * box comes from ibox1000 relative to the molecule's rank and its centre
     * from rcen1000 (src/buildinternal.f90:24-55, src/ibox.f90:206-220) */
    for (int s = 0; s < o->ns; s++) {
        int m = o->nw + s, f = o->first[m];
        int anc = o->anc[m];
        int rx = (anc % nb) / 4, ry = ((anc / nb) % nb) / 4, rz = (anc / (nb * nb)) / 4;
        double c2[3] = {rcen2(o, rx), rcen2(o, ry), rcen2(o, rz)};
        for (int j = 0; j < 29; j++) {
            const double *xm = o->r + 3 * (f + j);
            double cen[3];
            for (int a = 0; a < 3; a++) {
                int li = (int)((xm[a] - c2[a] + 5.0 * o->rs) / o->rs) + 1;
                cen[a] = 10.0 * o->rs * li / 10.0 - 10.0 * o->rs / 2.0 - 10.0 * o->rs / 20.0 + c2[a];
            }
            unsigned mask = 0;
            double rs15 = 1.5 * o->rs;
            for (int k = 0; k < 29; k++) {
                if (k == j) continue;
                const double *xn = o->r + 3 * (f + k);
                int near = 1;
                for (int a = 0; a < 3; a++)
                    near &= (xn[a] < cen[a] + rs15 && xn[a] >= cen[a] - rs15);
                if (!near) mask |= 1u << k;
            }
            o->isub[f + j] = mask;
        }
    }
}

void orc_get_classes(void *h, int *cls, int *anc, int *cbox, int *rkind, unsigned *isub)
{
    orc *o = h;
    if (cls) memcpy(cls, o->cls, sizeof(int) * o->nmol);
    if (anc) memcpy(anc, o->anc, sizeof(int) * o->nmol);
    if (cbox) memcpy(cbox, o->cbox, sizeof(int) * o->n);
    if (rkind) memcpy(rkind, o->rkind, sizeof(int) * o->n);
    if (isub) memcpy(isub, o->isub, sizeof(unsigned) * o->n);
}

/* -------------------------------------------------------------- far field */

/* P2M, src/computeOnm4.f90:24-162: O_n^m = sum q rho^n Y_n^{-m} about the L4
 * centre over the sites the lists make central to the box (the receivers) */
static void p2m(orc *o)
{
    int lp = o->lp, p = o->p;
#pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < o->nb3; b++) {
        cplx *O = o->O4 + (size_t)b * lp;
        for (int c = 0; c < lp; c++) O[c] = 0.0;
        const double *c = o->c4 + 3 * b;
        for (int n = 0; n <= p; n++)
            for (int m = 0; m <= n; m++) {
                cplx acc = 0.0;
                for (int s = o->bstart[b]; s < o->bstart[b + 1]; s++) {
                    int i = o->bsite[s];
                    if (o->rkind[i] < 0) continue;
                    const double *x = o->r + 3 * i;
                    double dx = x[0] - c[0], dy = x[1] - c[1], dz = x[2] - c[2];
                    double rho = sqrt(dx * dx + dy * dy + dz * dz);
                    double al = acos(dz / rho), be = atan2(dy, dx);
                    acc = acc + (o->ff.q[o->type[i]] * ipow(rho, n)) * ylm(o, n, -m, al, be);
                }
                O[idx(n, m)] = acc;
                if (m > 0) O[idx(n, -m)] = conj(acc);
            }
    }
}

/* M2M, src/computeOnm3.f90:16-98 / src/computeOnm2.f90:14-63: the children's
 * multipoles shifted to the parent centre */
static void m2m(orc *o, const cplx *child, const double *cc, int nc_dim, cplx *parent,
                const double *pc, int np_dim)
{
    int lp = o->lp, p = o->p;
#pragma omp parallel for
    for (int P = 0; P < np_dim * np_dim * np_dim; P++) {
        int px = P % np_dim, py = (P / np_dim) % np_dim, pz = P / (np_dim * np_dim);
        cplx *out = parent + (size_t)P * lp;
        for (int c = 0; c < lp; c++) out[c] = 0.0;
        for (int kc = 0; kc < 8; kc++) {
            int cx = 2 * px + (kc & 1), cy = 2 * py + ((kc >> 1) & 1), cz = 2 * pz + (kc >> 2);
            int C = cx + nc_dim * (cy + nc_dim * cz);
            const cplx *O = child + (size_t)C * lp;
            double rxpc = -pc[3 * P] + cc[3 * C], rypc = -pc[3 * P + 1] + cc[3 * C + 1],
                   rzpc = -pc[3 * P + 2] + cc[3 * C + 2];
            double rho = sqrt(rxpc * rxpc + rypc * rypc + rzpc * rzpc);
            double al = acos(rzpc / rho), be = atan2(rypc, rxpc);
            for (int j = 0; j <= p; j++)
                for (int k = 0; k <= j; k++) {
                    double ajk = anm(o, j, k);
                    for (int n = 0; n <= j; n++)
                        for (int m = -n; m <= n; m++) {
                            if (j - n - iabs_(k - m) < 0) continue;
                            cplx ctmp = O[idx(j - n, k - m)] * ylm(o, n, -m, al, be);
                            double rtmp = (double)ikm1(k, m) * anm(o, n, m) * anm(o, j - n, k - m) *
                                          ipow(rho, n) / ajk;
                            out[idx(j, k)] = out[idx(j, k)] + rtmp * ctmp;
                        }
                    if (k > 0) out[idx(j, -k)] = conj(out[idx(j, k)]);
                }
        }
    }
}

/* one M2L translation (src/computeLjk4.f90:126-160): L += T(vec) O, vec =
 * source centre - target centre */
static void m2l_one(const orc *o, double rx, double ry, double rz, const cplx *O, cplx *L)
{
    int p = o->p;
    double rho = sqrt(rx * rx + ry * ry + rz * rz);
    double al = acos(rz / rho), be = atan2(ry, rx);
    cplx Y[17][33];
    for (int J = 0; J <= 2 * p; J++)
        for (int M = -J; M <= J; M++) Y[J][M + 16] = ylm(o, J, M, al, be);
    for (int j = 0; j <= p; j++)
        for (int k = 0; k <= j; k++) {
            double ajk = anm(o, j, k);
            cplx acc = L[idx(j, k)];
            for (int n = 0; n <= p; n++)
                for (int m = -n; m <= n; m++) {
                    cplx ctmp = O[idx(n, m)] * Y[j + n][m - k + 16];
                    double rtmp = (double)ikm2(k, m) * ajk * anm(o, n, m) /
                                  ((double)sgnpow(n) * anm(o, j + n, m - k) * ipow(rho, j + n + 1));
                    acc = acc + rtmp * ctmp;
                }
            L[idx(j, k)] = acc;
            if (k > 0) L[idx(j, -k)] = conj(acc);
        }
}

static int nonzero(const cplx *O, int lp)
{
    for (int c = 0; c < lp; c++)
        if (creal(O[c]) != 0.0 || cimag(O[c]) != 0.0) return 1;
    return 0;
}

/* M2L over the interaction list at one level: sources are children of the
 * parent's neighbours that are not neighbours of the target
 * (src/computeLjk4.f90:114-124, src/computeLjk3.f90:116-122) */
static void m2l_level(orc *o, const cplx *Osrc, cplx *L, const double *cen, int nd)
{
    int lp = o->lp;
#pragma omp parallel for schedule(dynamic, 2)
    for (int T = 0; T < nd * nd * nd; T++) {
        int tx = T % nd, ty = (T / nd) % nd, tz = T / (nd * nd);
        int lo[3], hi[3], t[3] = {tx, ty, tz};
        for (int a = 0; a < 3; a++) {
            int par = t[a] / 2;
            lo[a] = 2 * (par - 1);
            hi[a] = 2 * (par + 1) + 1;
            if (lo[a] < 0) lo[a] = 0;
            if (hi[a] > nd - 1) hi[a] = nd - 1;
        }
        for (int sz = lo[2]; sz <= hi[2]; sz++)
            for (int sy = lo[1]; sy <= hi[1]; sy++)
                for (int sx = lo[0]; sx <= hi[0]; sx++) {
                    if (iabs_(sx - tx) <= 1 && iabs_(sy - ty) <= 1 && iabs_(sz - tz) <= 1) continue;
                    int S = sx + nd * (sy + nd * sz);
                    const cplx *O = Osrc + (size_t)S * lp;
                    if (!nonzero(O, lp)) continue;
                    m2l_one(o, cen[3 * S] - cen[3 * T], cen[3 * S + 1] - cen[3 * T + 1],
                            cen[3 * S + 2] - cen[3 * T + 2], O, L + (size_t)T * lp);
                }
    }
}

/* M2L at the top level: every L2 box not adjacent to the target
 * (src/computeLjk2.f90:187-243) */
static void m2l_top(orc *o)
{
    int lp = o->lp, nd = o->nb2;
    for (int T = 0; T < o->n2; T++) {
        cplx *L = o->L2 + (size_t)T * lp;
        for (int c = 0; c < lp; c++) L[c] = 0.0;
        int tx = T % nd, ty = (T / nd) % nd, tz = T / (nd * nd);
        for (int S = 0; S < o->n2; S++) {
            int sx = S % nd, sy = (S / nd) % nd, sz = S / (nd * nd);
            if (iabs_(sx - tx) <= 1 && iabs_(sy - ty) <= 1 && iabs_(sz - tz) <= 1) continue;
            m2l_one(o, o->c2[3 * S] - o->c2[3 * T], o->c2[3 * S + 1] - o->c2[3 * T + 1],
                    o->c2[3 * S + 2] - o->c2[3 * T + 2], o->O2 + (size_t)S * lp, L);
        }
    }
}

/* L2L, src/shiftLjk2.f90:14-63 / src/shiftLjk3.f90:64-157: the parent's local
 * expansion re-expanded about each child centre; the child table is zeroed
 * first */
static void l2l(orc *o, const cplx *Lp, const double *pc, int np_dim, cplx *Lc,
                const double *cc, int nc_dim)
{
    int lp = o->lp, p = o->p;
#pragma omp parallel for
    for (int C = 0; C < nc_dim * nc_dim * nc_dim; C++) {
        int cx = C % nc_dim, cy = (C / nc_dim) % nc_dim, cz = C / (nc_dim * nc_dim);
        int P = cx / 2 + np_dim * (cy / 2 + np_dim * (cz / 2));
        const cplx *L = Lp + (size_t)P * lp;
        cplx *out = Lc + (size_t)C * lp;
        for (int c = 0; c < lp; c++) out[c] = 0.0;
        double rxpc = pc[3 * P] - cc[3 * C], rypc = pc[3 * P + 1] - cc[3 * C + 1],
               rzpc = pc[3 * P + 2] - cc[3 * C + 2];
        double rho = sqrt(rxpc * rxpc + rypc * rypc + rzpc * rzpc);
        double al = acos(rzpc / rho), be = atan2(rypc, rxpc);
        for (int j = 0; j <= p; j++)
            for (int k = 0; k <= j; k++) {
                double ajk = anm(o, j, k);
                for (int n = 0; n <= p; n++)
                    for (int m = -n; m <= n; m++) {
                        if (n - j - iabs_(m - k) < 0) continue;
                        cplx ctmp = L[idx(n, m)] * ylm(o, n - j, m - k, al, be);
                        double rtmp = (double)ikm3(k, m) * anm(o, n - j, m - k) * ajk *
                                      ipow(rho, n - j) / ((double)sgnpow(n + j) * anm(o, n, m));
                        out[idx(j, k)] = out[idx(j, k)] + rtmp * ctmp;
                    }
                if (k > 0) out[idx(j, -k)] = conj(out[idx(j, k)]);
            }
    }
}

/* the far-field chain of a rebuild step (src/mybox.f90:77-81 and the nested
 * calls of src/up_down_Onm4_isr.f90:45-50, src/up_down_Onm3_isr.f90:201-204):
 * P2M, M2M 4->3->2, M2L at L2, L2L 2->3 then M2L at L3, L2L 3->4 then M2L at L4 */
void orc_fmm(void *h)
{
    orc *o = h;
    p2m(o);
    m2m(o, o->O4, o->c4, o->nb, o->O3, o->c3, o->n3);
    m2m(o, o->O3, o->c3, o->n3, o->O2, o->c2, o->nb2);
    m2l_top(o);
    l2l(o, o->L2, o->c2, o->nb2, o->L3, o->c3, o->n3);
    m2l_level(o, o->O3, o->L3, o->c3, o->n3);
    l2l(o, o->L3, o->c3, o->n3, o->L4, o->c4, o->nb);
    m2l_level(o, o->O4, o->L4, o->c4, o->nb);
}

void orc_get_expansions(void *h, double *O4, double *L4, double *O3, double *L3, double *O2,
                        double *L2)
{
    orc *o = h;
    size_t s4 = sizeof(cplx) * o->nb3 * o->lp, s3 = sizeof(cplx) * o->n33 * o->lp,
           s2 = sizeof(cplx) * o->n2 * o->lp;
    if (O4) memcpy(O4, o->O4, s4);
    if (L4) memcpy(L4, o->L4, s4);
    if (O3) memcpy(O3, o->O3, s3);
    if (L3) memcpy(L3, o->L3, s3);
    if (O2) memcpy(O2, o->O2, s2);
    if (L2) memcpy(L2, o->L2, s2);
}

/* L2P, src/evalLjk4.f90:21-284: gradient of the frozen local expansion of the
 * receiver's box at the current position */
static void l2p_site(const orc *o, int i, double *f)
{
    int b = o->cbox[i], p = o->p;
    const double *c = o->c4 + 3 * b;
    const cplx *L = o->L4 + (size_t)b * o->lp;
    const double *x = o->r + 3 * i;
    double rqq = o->ff.q[o->type[i]] * o->ff.xpi;
    double rxpb = x[0] - c[0], rypb = x[1] - c[1], rzpb = x[2] - c[2];
    double rr = sqrt(rxpb * rxpb + rypb * rypb + rzpb * rzpb);
    double rp1 = rzpb / rr;
    cplx rp2 = (rxpb + rypb * I) / sqrt(rxpb * rxpb + rypb * rypb);
    double rp3 = 1.0 / (rxpb * rxpb + rypb * rypb);
    double fx = f[0], fy = f[1], fz = f[2];
    for (int j = 0; j <= p; j++)
        for (int k = -j; k <= j; k++) {
            int ak = iabs_(k);
            double rfctr = sqrt(o->ff.rifact[j - ak] / o->ff.rifact[j + ak]);
            cplx rq0 = L[idx(j, k)] * rfctr * cipow(rp2, k) * ipow(rr, j - 1);
            double rq1 = (j + ak) * legendre(j - 1, ak, rp1);
            double rq2 = legendre(j, ak, rp1);
            double rq3 = rxpb * rzpb * rq1;
            cplx rq4 = (I * k * rypb - j * rxpb) * rr * rq2;
            double rq5 = rypb * rzpb * rq1;
            cplx rq6 = -(I * k * rxpb + j * rypb) * rr * rq2;
            fx = fx + creal(rqq * rp3 * rq0 * (rq3 + rq4));
            fy = fy + creal(rqq * rp3 * rq0 * (rq5 + rq6));
            fz = fz - creal(rqq * rq0 * rq1);
        }
    f[0] = fx;
    f[1] = fy;
    f[2] = fz;
}

/* ------------------------------------------------------------- near field */

/* Pair force on receiver i from source j: Coulomb xpi q_i q_j r/r^3 plus
 * LJ 48 eps s6 r^-6 (s6 r^-6 - 1/2) r/r^2 for r <= rc, the expression every
 * eval[wS]* kernel evaluates (e.g. src/evalSS.f90:38-56). */
static void pair_force(const orc *o, int i, int j, double *acc)
{
    const double *xi = o->r + 3 * i, *xj = o->r + 3 * j;
    int ti = o->type[i], tj = o->type[j];
    double xij = xi[0] - xj[0], yij = xi[1] - xj[1], zij = xi[2] - xj[2];
    double rijsq = xij * xij + yij * yij + zij * zij;
    double rij = sqrt(rijsq);
    double rqq = o->ff.q[ti] * o->ff.xpi;
    double fel = rqq * o->ff.q[tj] / (rij * rij * rij);
    double flj = 0.0;
    double c = o->ff.lj_c[ti * 32 + tj];
    if (c != 0.0 && rij <= o->ff.rc) {
        double s6 = o->ff.lj_s6[ti * 32 + tj];
        double rsi = 1.0 / rijsq;
        double r6 = rsi * rsi * rsi;
        double rpl = c * r6 * (s6 * r6 - 0.5);
        flj = rpl * rsi;
    }
    double ft = fel + flj;
    acc[0] += xij * ft;
    acc[1] += yij * ft;
    acc[2] += zij * ft;
}

/* The near field the 39 eval[wS]* kernels compute together: every receiver
 * (site central in box b at the last rebuild) gets the pair force of every
 * site of another molecule in the 3^3 block around b.  The kinds
 * (class, storage ring) of receiver and source pick the reference kernel a
 * pair belongs to; near_mask keeps a subset of them. */
static void near_field(const orc *o, unsigned long long mask, double *f)
{
    int nb = o->nb;
#pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < o->nb3; b++) {
        int bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
        for (int s = o->bstart[b]; s < o->bstart[b + 1]; s++) {
            int i = o->bsite[s];
            int rk = o->rkind[i];
            if (rk < 0) continue;
            int mi = o->mol[i];
            double acc[3] = {0, 0, 0};
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        int x = bx + dx, y = by + dy, z = bz + dz;
                        if (x < 0 || y < 0 || z < 0 || x >= nb || y >= nb || z >= nb) continue;
                        int nbx = x + nb * (y + nb * z);
                        for (int t = o->bstart[nbx]; t < o->bstart[nbx + 1]; t++) {
                            int j = o->bsite[t];
                            int mj = o->mol[j];
                            if (mj == mi) continue;
                            int sk = src_kind(o->cls[mj], cheb(o, o->anc[mj], b));
                            if (sk < 0) continue;
                            if (!((mask >> (rk * 9 + sk)) & 1ull)) continue;
                            pair_force(o, i, j, acc);
                        }
                    }
            f[3 * i] += acc[0];
            f[3 * i + 1] += acc[1];
            f[3 * i + 2] += acc[2];
        }
    }
}

/* ----------------------------------------------------------- intramolecular */

/* eval_surfactant_internal, src/eval_surfactant_internal.f90:12-110: LJ (no
 * cutoff) + Coulomb for sites (j <= 15, k >= j + 4) and the 42 listed pairs */
static void internal(const orc *o, double *f)
{
    for (int s = 0; s < o->ns; s++) {
        int base = o->first[o->nw + s];
        for (int p = 0; p < o->ff.n_internal; p++) {
            int j = o->ff.int_j[p] - 1, k = o->ff.int_k[p] - 1;
            const double *xj = o->r + 3 * (base + j), *xk = o->r + 3 * (base + k);
            double x = xj[0] - xk[0], y = xj[1] - xk[1], z = xj[2] - xk[2];
            double rsq = x * x + y * y + z * z;
            double rctr = sqrt(rsq);
            int tj = 3 + j, tk = 3 + k;
            double fel = o->ff.q[tj] * o->ff.q[tk] * o->ff.xpi / (rctr * rctr * rctr);
            double rsi = 1.0 / rsq;
            double r6 = rsi * rsi * rsi;
            double rpl = o->ff.lj_c[tj * 32 + tk] * r6 * (o->ff.lj_s6[tj * 32 + tk] * r6 - 0.5);
            double flj = rpl * rsi;
            double ft = fel + flj;
            double *fj = f + 3 * (base + j), *fk = f + 3 * (base + k);
            fj[0] += x * ft; fj[1] += y * ft; fj[2] += z * ft;
            fk[0] -= x * ft; fk[1] -= y * ft; fk[2] -= z * ft;
        }
    }
}

/* eval_subtract_fma, src/eval_subtract_fma.f90:10-54: remove the Coulomb term
 * the far field counted for same-molecule partners outside the site's block */
static void subtract_fma(const orc *o, double *f)
{
    for (int s = 0; s < o->ns; s++) {
        int base = o->first[o->nw + s];
        for (int j = 0; j < 29; j++) {
            const double *xm = o->r + 3 * (base + j);
            unsigned mask = o->isub[base + j];
            double *fm = f + 3 * (base + j);
            for (int k = 0; k < 29; k++) {
                if (!((mask >> k) & 1u)) continue;
                const double *xn = o->r + 3 * (base + k);
                double xmn = xm[0] - xn[0], ymn = xm[1] - xn[1], zmn = xm[2] - xn[2];
                double rmnsq = xmn * xmn + ymn * ymn + zmn * zmn;
                double rmn = sqrt(rmnsq);
                double fel = o->ff.q[3 + j] * o->ff.q[3 + k] * o->ff.xpi / (rmn * rmn * rmn);
                fm[0] -= xmn * fel;
                fm[1] -= ymn * fel;
                fm[2] -= zmn * fel;
            }
        }
    }
}

/* eval_valence, src/eval_valence.f90:97-158: harmonic angles, force
 * k (theta - theta0) / sin(theta) times the gradient of cos(theta) */
static void valence(const orc *o, double *f)
{
    for (int s = 0; s < o->ns; s++) {
        int base = o->first[o->nw + s];
        for (int a = 0; a < 36; a++) {
            int ig = base + o->ff.ang_i[a] - 1, jg = base + o->ff.ang_j[a] - 1,
                kg = base + o->ff.ang_k[a] - 1;
            const double *ri = o->r + 3 * ig, *rj = o->r + 3 * jg, *rk = o->r + 3 * kg;
            double rxji = ri[0] - rj[0], ryji = ri[1] - rj[1], rzji = ri[2] - rj[2];
            double rxki = rk[0] - rj[0], ryki = rk[1] - rj[1], rzki = rk[2] - rj[2];
            double dp = rxji * rxki + ryji * ryki + rzji * rzki;
            double rjin = sqrt(rxji * rxji + ryji * ryji + rzji * rzji);
            double rkin = sqrt(rxki * rxki + ryki * ryki + rzki * rzki);
            double thjik = acos(dp / rjin / rkin);
            double costh = cos(thjik);
            double r1 = o->ff.rkth[a] * (thjik - o->ff.th[a]) / sin(thjik);
            double fjx = rxki / rjin / rkin - costh * rxji / (rjin * rjin);
            double fjy = ryki / rjin / rkin - costh * ryji / (rjin * rjin);
            double fjz = rzki / rjin / rkin - costh * rzji / (rjin * rjin);
            double fkx = rxji / rjin / rkin - costh * rxki / (rkin * rkin);
            double fky = ryji / rjin / rkin - costh * ryki / (rkin * rkin);
            double fkz = rzji / rjin / rkin - costh * rzki / (rkin * rkin);
            double *fi = f + 3 * ig, *fj = f + 3 * jg, *fk = f + 3 * kg;
            fi[0] += r1 * fjx; fi[1] += r1 * fjy; fi[2] += r1 * fjz;
            fj[0] -= r1 * (fjx + fkx); fj[1] -= r1 * (fjy + fky); fj[2] -= r1 * (fjz + fkz);
            fk[0] += r1 * fkx; fk[1] += r1 * fky; fk[2] += r1 * fkz;
        }
    }
}

/* eval_dihedral / eval_dihedral_berendsen, src/eval_dihedral.f90:102-295 and
 * the Berendsen variant: V = a(1+cos phi) + b(1-cos 2phi) + c(1+cos 3phi) for
 * dihedrals 1..18, V = a(1+cos 3phi) for 19..30 */
static void dihedral(const orc *o, int lo, int hi, double *f)
{
    for (int s = 0; s < o->ns; s++) {
        int base = o->first[o->nw + s];
        for (int d = lo; d < hi; d++) {
            int iv = base + o->ff.dih_i[d] - 1, jv = base + o->ff.dih_j[d] - 1,
                kv = base + o->ff.dih_k[d] - 1, lv = base + o->ff.dih_l[d] - 1;
            const double *Ri = o->r + 3 * iv, *Rj = o->r + 3 * jv, *Rk = o->r + 3 * kv,
                         *Rl = o->r + 3 * lv;
            double rxji = Rj[0] - Ri[0], ryji = Rj[1] - Ri[1], rzji = Rj[2] - Ri[2];
            double rxkj = Rk[0] - Rj[0], rykj = Rk[1] - Rj[1], rzkj = Rk[2] - Rj[2];
            double rxnk = Rl[0] - Rk[0], rynk = Rl[1] - Rk[1], rznk = Rl[2] - Rk[2];
            double cpjikjx = rxji * rykj - ryji * rxkj;
            double cpjikjy = rxji * rzkj - rzji * rxkj;
            double cpjikjz = ryji * rzkj - rzji * rykj;
            double cpkjnkx = rxkj * rynk - rykj * rxnk;
            double cpkjnky = rxkj * rznk - rzkj * rxnk;
            double cpkjnkz = rykj * rznk - rzkj * rynk;
            double cpjikjn = sqrt(cpjikjx * cpjikjx + cpjikjy * cpjikjy + cpjikjz * cpjikjz);
            double cpkjnkn = sqrt(cpkjnkx * cpkjnkx + cpkjnky * cpkjnky + cpkjnkz * cpkjnkz);
            double dp = cpjikjx * cpkjnkx + cpjikjy * cpkjnky + cpjikjz * cpkjnkz;
            double arg = dp / cpjikjn / cpkjnkn;
            double clamp = d < 18 ? o->ff.clamp_opls : o->ff.clamp_be;
            if (arg <= -1.0) arg = -clamp;
            if (arg >= 1.0) arg = clamp;
            double th = acos(arg);
            double costh = cos(th), sinth = sin(th);
            double r1;
            if (d < 18)
                r1 = -o->ff.dva[d] + 2.0 * o->ff.dvb[d] * sin(2.0 * th) / sinth -
                     3.0 * o->ff.dvc[d] * sin(3.0 * th) / sinth;
            else
                r1 = -3.0 * o->ff.dva[d] * sin(3.0 * th) / sinth;
            double n1n2 = cpjikjn * cpkjnkn;
            (void)n1n2;
            double c1 = cpjikjn * cpjikjn, c2 = cpkjnkn * cpkjnkn;
            /* particle i */
            double fix = rxkj * (-rykj * rynk - rzkj * rznk) + rxnk * (rykj * rykj + rzkj * rzkj);
            double fiy = rykj * (-rxkj * rxnk - rzkj * rznk) + rynk * (rxkj * rxkj + rzkj * rzkj);
            double fiz = rzkj * (-rxkj * rxnk - rykj * rynk) + rznk * (rxkj * rxkj + rykj * rykj);
            fix = fix / cpjikjn / cpkjnkn;
            fiy = fiy / cpjikjn / cpkjnkn;
            fiz = fiz / cpjikjn / cpkjnkn;
            double fix2 = 2.0 * rxji * (-rykj * rykj - rzkj * rzkj) + 2.0 * rxkj * (ryji * rykj + rzji * rzkj);
            double fiy2 = 2.0 * ryji * (-rxkj * rxkj - rzkj * rzkj) + 2.0 * rykj * (rxji * rxkj + rzji * rzkj);
            double fiz2 = 2.0 * rzji * (-rxkj * rxkj - rykj * rykj) + 2.0 * rzkj * (rxji * rxkj + ryji * rykj);
            fix2 = fix2 / c1;
            fiy2 = fiy2 / c1;
            fiz2 = fiz2 / c1;
            double fxi = r1 * (fix - 0.5 * costh * fix2);
            double fyi = r1 * (fiy - 0.5 * costh * fiy2);
            double fzi = r1 * (fiz - 0.5 * costh * fiz2);
            /* particle j */
            double fjx = rxji * (-rykj * rynk - rzkj * rznk) + rxkj * (rykj * rynk + rzkj * rznk);
            double fjy = ryji * (-rxkj * rxnk - rzkj * rznk) + rykj * (rxkj * rxnk + rzkj * rznk);
            double fjz = rzji * (-rxkj * rxnk - rykj * rynk) + rzkj * (rxkj * rxnk + rykj * rynk);
            double fjx1 = rxnk * (-ryji * rykj - rzji * rzkj - rykj * rykj - rzkj * rzkj);
            double fjy1 = rynk * (-rxji * rxkj - rzji * rzkj - rxkj * rxkj - rzkj * rzkj);
            double fjz1 = rznk * (-rxji * rxkj - ryji * rykj - rxkj * rxkj - rykj * rykj);
            double fjx2 = 2.0 * rxkj * (ryji * rynk + rzji * rznk);
            double fjy2 = 2.0 * rykj * (rxji * rxnk + rzji * rznk);
            double fjz2 = 2.0 * rzkj * (rxji * rxnk + ryji * rynk);
            fjx = (fjx + fjx1 + fjx2) / cpjikjn / cpkjnkn;
            fjy = (fjy + fjy1 + fjy2) / cpjikjn / cpkjnkn;
            fjz = (fjz + fjz1 + fjz2) / cpjikjn / cpkjnkn;
            double fjx3 = 2.0 * rxji * (rykj * rykj + rzkj * rzkj + ryji * rykj + rzji * rzkj);
            double fjy3 = 2.0 * ryji * (rxkj * rxkj + rzkj * rzkj + rxji * rxkj + rzji * rzkj);
            double fjz3 = 2.0 * rzji * (rxkj * rxkj + rykj * rykj + rxji * rxkj + ryji * rykj);
            double fjx4 = 2.0 * rxkj * (-ryji * ryji - rzji * rzji - ryji * rykj - rzji * rzkj);
            double fjy4 = 2.0 * rykj * (-rxji * rxji - rzji * rzji - rxji * rxkj - rzji * rzkj);
            double fjz4 = 2.0 * rzkj * (-rxji * rxji - ryji * ryji - rxji * rxkj - ryji * rykj);
            fjx3 = (fjx3 + fjx4) / c1;
            fjy3 = (fjy3 + fjy4) / c1;
            fjz3 = (fjz3 + fjz4) / c1;
            double fjx5 = 2.0 * rxnk * (rykj * rynk + rzkj * rznk) + 2.0 * rxkj * (-rynk * rynk - rznk * rznk);
            double fjy5 = 2.0 * rynk * (rxkj * rxnk + rzkj * rznk) + 2.0 * rykj * (-rxnk * rxnk - rznk * rznk);
            double fjz5 = 2.0 * rznk * (rxkj * rxnk + rykj * rynk) + 2.0 * rzkj * (-rxnk * rxnk - rynk * rynk);
            fjx5 = fjx5 / c2;
            fjy5 = fjy5 / c2;
            fjz5 = fjz5 / c2;
            double fxj = r1 * (fjx - 0.5 * costh * (fjx3 + fjx5));
            double fyj = r1 * (fjy - 0.5 * costh * (fjy3 + fjy5));
            double fzj = r1 * (fjz - 0.5 * costh * (fjz3 + fjz5));
            /* particle k */
            double fkx = rxji * (rykj * rykj + rzkj * rzkj + rykj * rynk + rzkj * rznk);
            double fky = ryji * (rxkj * rxkj + rzkj * rzkj + rxkj * rxnk + rzkj * rznk);
            double fkz = rzji * (rxkj * rxkj + rykj * rykj + rxkj * rxnk + rykj * rynk);
            double fkx2 = rxkj * (-ryji * rykj - rzji * rzkj);
            double fky2 = rykj * (-rxji * rxkj - rzji * rzkj);
            double fkz2 = rzkj * (-rxji * rxkj - ryji * rykj);
            double fkx3 = rxnk * (ryji * rykj + rzji * rzkj) + 2.0 * rxkj * (-ryji * rynk - rzji * rznk);
            double fky3 = rynk * (rxji * rxkj + rzji * rzkj) + 2.0 * rykj * (-rxji * rxnk - rzji * rznk);
            double fkz3 = rznk * (rxji * rxkj + ryji * rykj) + 2.0 * rzkj * (-rxji * rxnk - ryji * rynk);
            fkx = (fkx + fkx2 + fkx3) / cpjikjn / cpkjnkn;
            fky = (fky + fky2 + fky3) / cpjikjn / cpkjnkn;
            fkz = (fkz + fkz2 + fkz3) / cpjikjn / cpkjnkn;
            double fkx4 = 2.0 * rxji * (-ryji * rykj - rzji * rzkj) + 2.0 * rxkj * (ryji * ryji + rzji * rzji);
            double fky4 = 2.0 * ryji * (-rxji * rxkj - rzji * rzkj) + 2.0 * rykj * (rxji * rxji + rzji * rzji);
            double fkz4 = 2.0 * rzji * (-rxji * rxkj - ryji * rykj) + 2.0 * rzkj * (rxji * rxji + ryji * ryji);
            fkx4 = fkx4 / c1;
            fky4 = fky4 / c1;
            fkz4 = fkz4 / c1;
            double fkx5 = 2.0 * rxnk * (-rykj * rykj - rzkj * rzkj - rykj * rynk - rzkj * rznk);
            double fky5 = 2.0 * rynk * (-rxkj * rxkj - rzkj * rzkj - rxkj * rxnk - rzkj * rznk);
            double fkz5 = 2.0 * rznk * (-rxkj * rxkj - rykj * rykj - rxkj * rxnk - rykj * rynk);
            double fkx6 = 2.0 * rxkj * (rynk * rynk + rznk * rznk + rykj * rynk + rzkj * rznk);
            double fky6 = 2.0 * rykj * (rxnk * rxnk + rznk * rznk + rxkj * rxnk + rzkj * rznk);
            double fkz6 = 2.0 * rzkj * (rxnk * rxnk + rynk * rynk + rxkj * rxnk + rykj * rynk);
            fkx5 = (fkx5 + fkx6) / c2;
            fky5 = (fky5 + fky6) / c2;
            fkz5 = (fkz5 + fkz6) / c2;
            double fxk = r1 * (fkx - 0.5 * costh * (fkx4 + fkx5));
            double fyk = r1 * (fky - 0.5 * costh * (fky4 + fky5));
            double fzk = r1 * (fkz - 0.5 * costh * (fkz4 + fkz5));
            /* particle n */
            double fnx = rxji * (-rykj * rykj - rzkj * rzkj) + rxkj * (ryji * rykj + rzji * rzkj);
            double fny = ryji * (-rxkj * rxkj - rzkj * rzkj) + rykj * (rxji * rxkj + rzji * rzkj);
            double fnz = rzji * (-rxkj * rxkj - rykj * rykj) + rzkj * (rxji * rxkj + ryji * rykj);
            fnx = fnx / cpjikjn / cpkjnkn;
            fny = fny / cpjikjn / cpkjnkn;
            fnz = fnz / cpjikjn / cpkjnkn;
            double fnx2 = 2.0 * rxnk * (rykj * rykj + rzkj * rzkj) + 2.0 * rxkj * (-rykj * rynk - rzkj * rznk);
            double fny2 = 2.0 * rynk * (rxkj * rxkj + rzkj * rzkj) + 2.0 * rykj * (-rxkj * rxnk - rzkj * rznk);
            double fnz2 = 2.0 * rznk * (rxkj * rxkj + rykj * rykj) + 2.0 * rzkj * (-rxkj * rxnk - rykj * rynk);
            fnx2 = fnx2 / c2;
            fny2 = fny2 / c2;
            fnz2 = fnz2 / c2;
            double fxn = r1 * (fnx - 0.5 * costh * fnx2);
            double fyn = r1 * (fny - 0.5 * costh * fny2);
            double fzn = r1 * (fnz - 0.5 * costh * fnz2);
            double *fi = f + 3 * iv, *fj = f + 3 * jv, *fk = f + 3 * kv, *fl = f + 3 * lv;
            fi[0] += fxi; fi[1] += fyi; fi[2] += fzi;
            fj[0] += fxj; fj[1] += fyj; fj[2] += fzj;
            fk[0] += fxk; fk[1] += fyk; fk[2] += fzk;
            fl[0] += fxn; fl[1] += fyn; fl[2] += fzn;
        }
    }
}

/* -------------------------------------------------------------- external */

/* eval_wall, src/eval_wall.f90:8-59: 48/(x-h)^13 + 48/(x+h)^13 on each axis */
static double p13(double x)
{
    double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4;
    return x8 * x4 * x;
}
static void wall(const orc *o, double *f)
{
    double h = o->rhs;
    for (int i = 0; i < o->n; i++)
        for (int a = 0; a < 3; a++) {
            double x = o->r[3 * i + a];
            f[3 * i + a] = f[3 * i + a] + 48.0 / p13(x - h);
            f[3 * i + a] = f[3 * i + a] + 48.0 / p13(h + x);
        }
}

/* eval_graphite, src/eval_graphite.f90:10-135: frozen-carbon LJ (r <= rc) on
 * the receivers filed in their own box (W1 and W2 oxygens, surfactant sites
 * of ring 0) of the bottom L4 layer, against that column's carbon list */
static void graphite(const orc *o, double *f)
{
    int nb = o->nb;
    for (int b = 0; b < nb * nb; b++) { /* z = 0 layer */
        for (int s = o->bstart[b]; s < o->bstart[b + 1]; s++) {
            int i = o->bsite[s];
            int rk = o->rkind[i];
            if (!(rk == 0 || rk == 1 || rk == 3)) continue;
            int t = o->type[i];
            double c = o->ff.gr_c[t], s6 = o->ff.gr_s6[t];
            if (c == 0.0) continue;
            const double *x = o->r + 3 * i;
            for (int g = o->gstart[b]; g < o->gstart[b + 1]; g++) {
                const double *cg = o->gr + 3 * o->gidx[g];
                double xij = x[0] - cg[0], yij = x[1] - cg[1], zij = x[2] - cg[2];
                double rijsq = xij * xij + yij * yij + zij * zij;
                double rij = sqrt(rijsq);
                if (rij <= o->ff.rc) {
                    double rsi = 1.0 / rijsq;
                    double r6 = rsi * rsi * rsi;
                    double rpl = c * r6 * (s6 * r6 - 0.5);
                    double flj = rpl * rsi;
                    f[3 * i] += xij * flj;
                    f[3 * i + 1] += yij * flj;
                    f[3 * i + 2] += zij * flj;
                }
            }
        }
    }
}

void orc_forces(void *h, unsigned what, unsigned long long near_mask, double *f)
{
    orc *o = h;
    memset(f, 0, sizeof(double) * 3 * (size_t)o->n);
    if (what & ORC_NEAR) near_field(o, near_mask, f);
    if (what & ORC_FAR)
        for (int i = 0; i < o->n; i++)
            if (o->rkind[i] >= 0) l2p_site(o, i, f + 3 * i);
    if (what & ORC_INTERNAL) internal(o, f);
    if (what & ORC_SUBTRACT) subtract_fma(o, f);
    if (what & ORC_VALENCE) valence(o, f);
    if (what & ORC_DIHEDRAL) dihedral(o, 0, 18, f);
    if (what & ORC_DIHEDRAL_BE) dihedral(o, 18, 30, f);
    if (what & ORC_WALL) wall(o, f);
    if (what & ORC_GRAPHITE) graphite(o, f);
}

/* ------------------------------------------------------------- integrator */

/* kinetic sums of nh_water1/2 (src/nh_water1.f90:1-43) and nh_surfactant
 * (src/nh_surfactant.f90:1-201) over one class of molecules */
static double kin_sum(const orc *o, int cls)
{
    double g[5] = {0, 0, 0, 0, 0};
    /* groups: water O, water H; surfactant H, O, CH2, Si, CH3 */
    static const int sgrp[29] = {0, 1, 2, 2, 1, 2, 2, 1, 2, 2, 1, 2, 2, 1, 2, 2, 2, 3, 4, 1,
                                 3, 4, 4, 4, 1, 3, 4, 4, 4};
    for (int m = 0; m < o->nmol; m++) {
        if (o->cls[m] != cls) continue;
        int f = o->first[m];
        int ns = m < o->nw ? 3 : 29;
        for (int k = 0; k < ns; k++) {
            const double *v = o->v + 3 * (f + k);
            double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (ns == 3) g[k == 0 ? 0 : 1] += s;
            else g[sgrp[k]] += s;
        }
    }
    if (cls < 2) return o->ff.mass[0] * g[0] + o->ff.mass[1] * g[1];
    return o->ff.mass[3 + 0] * g[0] + o->ff.mass[3 + 1] * g[1] + o->ff.mass[3 + 2] * g[2] +
           o->ff.mass[3 + 17] * g[3] + o->ff.mass[3 + 18] * g[4];
}

/* part A for one molecule (src/move_w1_a.f90:29-138, src/move_surfactant_a.f90:29-139):
 * p = r + v(1 - eta dt/2) + a, then SHAKE (Gauss-Seidel over the constraints,
 * tolerance 1e-4 on |d^2 - p^2| > 2 tol d^2, at most 101 sweeps) */
static int shake_mol(orc *o, int m, double escale)
{
    int f = o->first[m];
    int water = m < o->nw;
    int ns = water ? 3 : 29, nb = water ? 3 : 28;
    double rx[29][3], px[29][3], vx[29][3], rm[29];
    for (int i = 0; i < ns; i++) {
        for (int a = 0; a < 3; a++) {
            double r = o->r[3 * (f + i) + a], v = o->v[3 * (f + i) + a], ac = o->a[3 * (f + i) + a];
            rx[i][a] = r;
            px[i][a] = r + v * escale + ac;
            vx[i][a] = v * escale + ac;
        }
        rm[i] = 1.0 / o->ff.mass[o->type[f + i]];
    }
    double tol2 = 2.0 * 0.0001;
    int it = 0, done = 0;
    while (!done && it <= 100) {
        done = 1;
        for (int c = 0; c < nb; c++) {
            int ia, ib;
            double rabsq;
            if (water) {
                ia = c;
                ib = c + 1 > 2 ? 0 : c + 1;
                rabsq = o->ff.dsqw[c];
            } else {
                ia = o->ff.ibond[c] - 1;
                ib = o->ff.jbond[c] - 1;
                rabsq = o->ff.dsqs[c];
            }
            double pxab = px[ia][0] - px[ib][0], pyab = px[ia][1] - px[ib][1],
                   pzab = px[ia][2] - px[ib][2];
            double pabsq = pxab * pxab + pyab * pyab + pzab * pzab;
            double diffsq = rabsq - pabsq;
            if (fabs(diffsq) > rabsq * tol2) {
                double rxab = rx[ia][0] - rx[ib][0], ryab = rx[ia][1] - rx[ib][1],
                       rzab = rx[ia][2] - rx[ib][2];
                double rpab = rxab * pxab + ryab * pyab + rzab * pzab;
                double rma = rm[ia], rmb = rm[ib];
                double gab = diffsq / (2.0 * (rma + rmb) * rpab);
                double dx = rxab * gab, dy = ryab * gab, dz = rzab * gab;
                px[ia][0] = px[ia][0] + rma * dx; px[ia][1] = px[ia][1] + rma * dy;
                px[ia][2] = px[ia][2] + rma * dz;
                px[ib][0] = px[ib][0] - rmb * dx; px[ib][1] = px[ib][1] - rmb * dy;
                px[ib][2] = px[ib][2] - rmb * dz;
                vx[ia][0] = vx[ia][0] + rma * dx; vx[ia][1] = vx[ia][1] + rma * dy;
                vx[ia][2] = vx[ia][2] + rma * dz;
                vx[ib][0] = vx[ib][0] - rmb * dx; vx[ib][1] = vx[ib][1] - rmb * dy;
                vx[ib][2] = vx[ib][2] - rmb * dz;
                done = 0;
            }
        }
        it++;
    }
    for (int i = 0; i < ns; i++)
        for (int a = 0; a < 3; a++) {
            o->r[3 * (f + i) + a] = px[i][a];
            o->v[3 * (f + i) + a] = vx[i][a];
        }
    return !done;
}

/* Part A of a step for all species (src/mybox.f90:136-138): S, W1, W2 all
 * move with the old eta; then eta += pnh1 (sum1 + pnh2 + sum2) from the six
 * kinetic sums (src/move_w2_a.f90:145-157).  sums6 = sv1 sv2 w1v1 w1v2 w2v1 w2v2. */
void orc_move_a(void *h, double *sums6)
{
    orc *o = h;
    double dt = o->ff.dt;
    double escale = 1.0 - o->eta * dt / 2.0;
    double s[6];
    s[0] = kin_sum(o, 2);
    s[2] = kin_sum(o, 0);
    s[4] = kin_sum(o, 1);
    int fails = 0;
#pragma omp parallel for reduction(+ : fails) schedule(dynamic, 64)
    for (int m = 0; m < o->nmol; m++) fails += shake_mol(o, m, escale);
    o->shake_fail += fails;
    s[1] = kin_sum(o, 2);
    s[3] = kin_sum(o, 0);
    s[5] = kin_sum(o, 1);
    double sum1 = (s[0] + s[2] + s[4]) / (dt * dt);
    double sum2 = (s[1] + s[3] + s[5]) / (dt * dt);
    double eta_old = o->eta;
    o->eta = o->eta + o->pnh1 * (sum1 + o->pnh2 + sum2);
    o->xi += 0.5 * dt * (eta_old + o->eta);
    if (sums6) memcpy(sums6, s, sizeof(s));
}

/* part B for one molecule (src/move_w1_b.f90:23-123, src/move_surfactant_b.f90:23-122):
 * a = F dt^2/2m, v = 2/(2 + eta dt) (v + a), then RATTLE on the velocities */
static int rattle_mol(orc *o, int m, double escale, const double *F)
{
    int f = o->first[m];
    int water = m < o->nw;
    int ns = water ? 3 : 29, nb = water ? 3 : 28;
    double rx[29][3], vx[29][3], rm[29];
    double dt = o->ff.dt, dtsq2 = o->ff.dtsq2;
    for (int i = 0; i < ns; i++) {
        double mi = o->ff.mass[o->type[f + i]];
        for (int a = 0; a < 3; a++) {
            rx[i][a] = o->r[3 * (f + i) + a];
            double ac = F[3 * (f + i) + a] * dtsq2 / mi;
            o->a[3 * (f + i) + a] = ac;
            vx[i][a] = escale * (o->v[3 * (f + i) + a] + ac);
        }
        rm[i] = 1.0 / mi;
    }
    double tol = 0.0001;
    int it = 0, done = 0;
    while (!done && it <= 100) {
        done = 1;
        for (int c = 0; c < nb; c++) {
            int ia, ib;
            double dsq;
            if (water) {
                ia = c;
                ib = c + 1 > 2 ? 0 : c + 1;
                dsq = o->ff.dsqw[c];
            } else {
                ia = o->ff.ibond[c] - 1;
                ib = o->ff.jbond[c] - 1;
                dsq = o->ff.dsqs[c];
            }
            double vxab = vx[ia][0] - vx[ib][0], vyab = vx[ia][1] - vx[ib][1],
                   vzab = vx[ia][2] - vx[ib][2];
            double rxab = rx[ia][0] - rx[ib][0], ryab = rx[ia][1] - rx[ib][1],
                   rzab = rx[ia][2] - rx[ib][2];
            double rvab = rxab * vxab + ryab * vyab + rzab * vzab;
            rvab = rvab / dt;
            double rma = rm[ia], rmb = rm[ib];
            double gab = -rvab / ((rma + rmb) * dsq);
            if (fabs(gab) > tol) {
                double dx = rxab * gab, dy = ryab * gab, dz = rzab * gab;
                vx[ia][0] = vx[ia][0] + rma * dx * dt; vx[ia][1] = vx[ia][1] + rma * dy * dt;
                vx[ia][2] = vx[ia][2] + rma * dz * dt;
                vx[ib][0] = vx[ib][0] - rmb * dx * dt; vx[ib][1] = vx[ib][1] - rmb * dy * dt;
                vx[ib][2] = vx[ib][2] - rmb * dz * dt;
                done = 0;
            }
        }
        it++;
    }
    for (int i = 0; i < ns; i++)
        for (int a = 0; a < 3; a++) o->v[3 * (f + i) + a] = vx[i][a];
    return !done;
}

void orc_move_b(void *h, const double *F)
{
    orc *o = h;
    double escale = 2.0 / (2.0 + o->eta * o->ff.dt);
    int fails = 0;
#pragma omp parallel for reduction(+ : fails) schedule(dynamic, 64)
    for (int m = 0; m < o->nmol; m++) fails += rattle_mol(o, m, escale, F);
    o->shake_fail += fails;
}

/* init_accel, src/init_accel.f90:1-49 */
void orc_init_accel(void *h, const double *F)
{
    orc *o = h;
    for (int i = 0; i < o->n; i++) {
        double mi = o->ff.mass[o->type[i]];
        for (int a = 0; a < 3; a++) o->a[3 * i + a] = F[3 * i + a] * o->ff.dtsq2 / mi;
    }
}

int orc_shake_failures(void *h) { return ((orc *)h)->shake_fail; }

double orc_xi(void *h) { return ((orc *)h)->xi; }
void orc_set_xi(void *h, double xi) { ((orc *)h)->xi = xi; }

/* ------------------------------------------------------------- energies
 *
 * Not computed by the reference: the potential whose gradient is each force
 * term above, for energy and drift reporting.  Pair terms are summed over
 * receivers with weight 1/2 (every pair is seen from both ends); the far-field
 * energy is 1/2 sum q_i Phi(x_i) with Phi the frozen local expansion; the
 * thermostat adds Q eta^2/2 + df tr xi with Q = df tr tau^2/2 = dt/(2 pnh1). */

/* LJ energy of the force c r^-6 (s6 r^-6 - 1/2) r/r^2: (c/12) r^-6 (s6 r^-6 - 1) */
static double lj_energy(double c, double s6, double rsq)
{
    double rsi = 1.0 / rsq, r6 = rsi * rsi * rsi;
    return c / 12.0 * r6 * (s6 * r6 - 1.0);
}

/* rqq * Phi(x_i), Phi = Re sum_jk rfctr L_jk e^{ik phi} r^j P_j^|k|(cos theta):
 * the potential whose gradient l2p_site evaluates */
static double l2p_energy_site(const orc *o, int i)
{
    int b = o->cbox[i], p = o->p;
    const double *c = o->c4 + 3 * b;
    const cplx *L = o->L4 + (size_t)b * o->lp;
    const double *x = o->r + 3 * i;
    double rqq = o->ff.q[o->type[i]] * o->ff.xpi;
    double rxpb = x[0] - c[0], rypb = x[1] - c[1], rzpb = x[2] - c[2];
    double rr = sqrt(rxpb * rxpb + rypb * rypb + rzpb * rzpb);
    double rp1 = rzpb / rr;
    cplx rp2 = (rxpb + rypb * I) / sqrt(rxpb * rxpb + rypb * rypb);
    double phi = 0.0;
    for (int j = 0; j <= p; j++)
        for (int k = -j; k <= j; k++) {
            int ak = iabs_(k);
            double rfctr = sqrt(o->ff.rifact[j - ak] / o->ff.rifact[j + ak]);
            phi += creal(L[idx(j, k)] * rfctr * cipow(rp2, k)) * ipow(rr, j) * legendre(j, ak, rp1);
        }
    return rqq * phi;
}

static double angle_energy(const orc *o, int base, int a)
{
    const double *ri = o->r + 3 * (base + o->ff.ang_i[a] - 1), *rj = o->r + 3 * (base + o->ff.ang_j[a] - 1),
                 *rk = o->r + 3 * (base + o->ff.ang_k[a] - 1);
    double rxji = ri[0] - rj[0], ryji = ri[1] - rj[1], rzji = ri[2] - rj[2];
    double rxki = rk[0] - rj[0], ryki = rk[1] - rj[1], rzki = rk[2] - rj[2];
    double dp = rxji * rxki + ryji * ryki + rzji * rzki;
    double rjin = sqrt(rxji * rxji + ryji * ryji + rzji * rzji);
    double rkin = sqrt(rxki * rxki + ryki * ryki + rzki * rzki);
    double d = acos(dp / rjin / rkin) - o->ff.th[a];
    return 0.5 * o->ff.rkth[a] * d * d;
}

static double dihedral_energy(const orc *o, int base, int d)
{
    const double *Ri = o->r + 3 * (base + o->ff.dih_i[d] - 1), *Rj = o->r + 3 * (base + o->ff.dih_j[d] - 1),
                 *Rk = o->r + 3 * (base + o->ff.dih_k[d] - 1), *Rl = o->r + 3 * (base + o->ff.dih_l[d] - 1);
    double rxji = Rj[0] - Ri[0], ryji = Rj[1] - Ri[1], rzji = Rj[2] - Ri[2];
    double rxkj = Rk[0] - Rj[0], rykj = Rk[1] - Rj[1], rzkj = Rk[2] - Rj[2];
    double rxnk = Rl[0] - Rk[0], rynk = Rl[1] - Rk[1], rznk = Rl[2] - Rk[2];
    double ax = rxji * rykj - ryji * rxkj, ay = rxji * rzkj - rzji * rxkj, az = ryji * rzkj - rzji * rykj;
    double bx = rxkj * rynk - rykj * rxnk, by = rxkj * rznk - rzkj * rxnk, bz = rykj * rznk - rzkj * rynk;
    double na = sqrt(ax * ax + ay * ay + az * az), nbn = sqrt(bx * bx + by * by + bz * bz);
    double arg = (ax * bx + ay * by + az * bz) / na / nbn;
    double clamp = d < 18 ? o->ff.clamp_opls : o->ff.clamp_be;
    if (arg <= -1.0) arg = -clamp;
    if (arg >= 1.0) arg = clamp;
    double th = acos(arg);
    if (d < 18)
        return o->ff.dva[d] * (1.0 + cos(th)) + o->ff.dvb[d] * (1.0 - cos(2.0 * th)) +
               o->ff.dvc[d] * (1.0 + cos(3.0 * th));
    return o->ff.dva[d] * (1.0 + cos(3.0 * th));
}

void orc_energies(void *h, double *e)
{
    orc *o = h;
    memset(e, 0, sizeof(double) * ORC_E_COUNT);
    int nb = o->nb;
    double ec = 0.0, el = 0.0, vir = 0.0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : ec, el, vir)
    for (int b = 0; b < o->nb3; b++) {
        int bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
        for (int s = o->bstart[b]; s < o->bstart[b + 1]; s++) {
            int i = o->bsite[s];
            if (o->rkind[i] < 0) continue;
            int mi = o->mol[i], ti = o->type[i];
            const double *xi = o->r + 3 * i;
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        int x = bx + dx, y = by + dy, z = bz + dz;
                        if (x < 0 || y < 0 || z < 0 || x >= nb || y >= nb || z >= nb) continue;
                        int nbx = x + nb * (y + nb * z);
                        for (int t = o->bstart[nbx]; t < o->bstart[nbx + 1]; t++) {
                            int j = o->bsite[t], mj = o->mol[j], tj = o->type[j];
                            if (mj == mi) continue;
                            if (src_kind(o->cls[mj], cheb(o, o->anc[mj], b)) < 0) continue;
                            const double *xj = o->r + 3 * j;
                            double u = xi[0] - xj[0], v = xi[1] - xj[1], w = xi[2] - xj[2];
                            double rsq = u * u + v * v + w * w, r = sqrt(rsq);
                            double qq = o->ff.q[ti] * o->ff.xpi * o->ff.q[tj];
                            ec += 0.5 * qq / r;
                            vir += 0.5 * qq / r;
                            double c = o->ff.lj_c[ti * 32 + tj];
                            if (c != 0.0 && r <= o->ff.rc) {
                                double s6 = o->ff.lj_s6[ti * 32 + tj];
                                double rsi = 1.0 / rsq, r6 = rsi * rsi * rsi;
                                el += 0.5 * lj_energy(c, s6, rsq);
                                vir += 0.5 * c * r6 * (s6 * r6 - 0.5);
                            }
                        }
                    }
        }
    }
    e[ORC_E_NEAR_COUL] = ec;
    e[ORC_E_NEAR_LJ] = el;
    e[ORC_E_VIRIAL] = vir;
    double far = 0.0, wl = 0.0;
    for (int i = 0; i < o->n; i++) {
        if (o->rkind[i] >= 0) far += 0.5 * l2p_energy_site(o, i);
        for (int a = 0; a < 3; a++) {
            double x = o->r[3 * i + a], h1 = x - o->rhs, h2 = x + o->rhs;
            double u1 = 1.0 / (h1 * h1), u2 = 1.0 / (h2 * h2);
            double u16 = u1 * u1 * u1, u26 = u2 * u2 * u2;
            wl += 4.0 * u16 * u16 + 4.0 * u26 * u26;
        }
    }
    e[ORC_E_FAR] = far;
    e[ORC_E_WALL] = wl;
    double gr = 0.0;
    for (int b = 0; b < nb * nb; b++)
        for (int s = o->bstart[b]; s < o->bstart[b + 1]; s++) {
            int i = o->bsite[s], rk = o->rkind[i];
            if (!(rk == 0 || rk == 1 || rk == 3)) continue;
            int t = o->type[i];
            double c = o->ff.gr_c[t], s6 = o->ff.gr_s6[t];
            if (c == 0.0) continue;
            const double *x = o->r + 3 * i;
            for (int g = o->gstart[b]; g < o->gstart[b + 1]; g++) {
                const double *cg = o->gr + 3 * o->gidx[g];
                double u = x[0] - cg[0], v = x[1] - cg[1], w = x[2] - cg[2];
                double rsq = u * u + v * v + w * w;
                if (sqrt(rsq) <= o->ff.rc) gr += lj_energy(c, s6, rsq);
            }
        }
    e[ORC_E_GRAPHITE] = gr;
    double ic = 0.0, il = 0.0, sb = 0.0, an = 0.0, dh = 0.0, db = 0.0;
    for (int s = 0; s < o->ns; s++) {
        int base = o->first[o->nw + s];
        for (int p = 0; p < o->ff.n_internal; p++) {
            int j = o->ff.int_j[p] - 1, k = o->ff.int_k[p] - 1;
            const double *xj = o->r + 3 * (base + j), *xk = o->r + 3 * (base + k);
            double u = xj[0] - xk[0], v = xj[1] - xk[1], w = xj[2] - xk[2];
            double rsq = u * u + v * v + w * w;
            int tj = 3 + j, tk = 3 + k;
            ic += o->ff.q[tj] * o->ff.q[tk] * o->ff.xpi / sqrt(rsq);
            il += lj_energy(o->ff.lj_c[tj * 32 + tk], o->ff.lj_s6[tj * 32 + tk], rsq);
        }
        for (int j = 0; j < 29; j++) {
            unsigned mask = o->isub[base + j];
            const double *xm = o->r + 3 * (base + j);
            for (int k = 0; k < 29; k++) {
                if (!((mask >> k) & 1u)) continue;
                const double *xn = o->r + 3 * (base + k);
                double u = xm[0] - xn[0], v = xm[1] - xn[1], w = xm[2] - xn[2];
                sb -= 0.5 * o->ff.q[3 + j] * o->ff.q[3 + k] * o->ff.xpi / sqrt(u * u + v * v + w * w);
            }
        }
        for (int a = 0; a < 36; a++) an += angle_energy(o, base, a);
        for (int d = 0; d < 18; d++) dh += dihedral_energy(o, base, d);
        for (int d = 18; d < 30; d++) db += dihedral_energy(o, base, d);
    }
    e[ORC_E_INTERNAL_COUL] = ic;
    e[ORC_E_INTERNAL_LJ] = il;
    e[ORC_E_SUBTRACT] = sb;
    e[ORC_E_ANGLE] = an;
    e[ORC_E_DIHEDRAL] = dh;
    e[ORC_E_DIHEDRAL_BE] = db;
    double pot = 0.0;
    for (int k = 0; k < ORC_E_POTENTIAL; k++) pot += e[k];
    e[ORC_E_POTENTIAL] = pot;
    double dt = o->ff.dt, kin = 0.0;
    for (int i = 0; i < o->n; i++) {
        const double *v = o->v + 3 * i;
        kin += 0.5 * o->ff.mass[o->type[i]] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / (dt * dt);
    }
    e[ORC_E_KINETIC] = kin;
    int df = 3 * o->n - (3 * o->nw + 28 * o->ns);
    e[ORC_E_TEMPERATURE] = 2.0 * kin / df;
    e[ORC_E_THERMOSTAT] = o->pnh1 != 0.0 ? 0.25 * dt / o->pnh1 * o->eta * o->eta - 0.5 * o->pnh2 * o->xi : 0.0;
    e[ORC_E_TOTAL] = pot + kin + e[ORC_E_THERMOSTAT];
}

/* Full steps in the reference's order (src/mybox.f90:130-227): part A; on
 * every ifar-th step the rebuild and the far-field chain; the forces; part B. */
void orc_steps(void *h, int nsteps, int ifar, int *kb, double *f)
{
    orc *o = h;
    for (int s = 0; s < nsteps; s++) {
        (*kb)++;
        orc_move_a(o, NULL);
        if (*kb % ifar == 0) {
            orc_rebuild(o);
            orc_fmm(o);
        }
        orc_forces(o, ORC_ALL, ~0ull, f);
        orc_move_b(o, f);
    }
}
