/* CPU restatement of the reference's per-timestep force and integration path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the oracle the HIP path is checked against
 * (and the CPU baseline bench.py times).  It restates, for the whole system on
 * one address space, what the reference computes rank by rank; every function
 * cites the reference routine it follows.  Reduced units throughout; positions,
 * velocities, accelerations and forces are (N, 3) row-major double arrays with
 * all waters (O, H1, H2) first, then the 29-site surfactants.
 */
#ifndef SESMD_ORACLE_H
#define SESMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ff {
    double q[32], mass[32];
    double lj_c[32 * 32], lj_s6[32 * 32];
    double gr_c[32], gr_s6[32];
    double dsqw[3], dsqs[28];
    int ibond[28], jbond[28];
    double th[36], rkth[36];
    int ang_i[36], ang_j[36], ang_k[36];
    double dva[30], dvb[30], dvc[30];
    int dih_i[30], dih_j[30], dih_k[30], dih_l[30];
    int n_internal;
    int int_j[320], int_k[320];
    double rifact[33];
    double xpi, rc, rs, dt, dtsq2;
    double clamp_opls, clamp_be;
} orc_ff;

/* force-component bits for orc_forces */
enum {
    ORC_NEAR = 1, ORC_FAR = 2, ORC_INTERNAL = 4, ORC_SUBTRACT = 8, ORC_VALENCE = 16,
    ORC_DIHEDRAL = 32, ORC_DIHEDRAL_BE = 64, ORC_WALL = 128, ORC_GRAPHITE = 256,
    ORC_ALL = 511
};

void *orc_create(int n_water, int n_surf, int nb2, int p, const orc_ff *ff);
void orc_destroy(void *h);
void orc_set_state(void *h, const double *r, const double *v, const double *a, double eta);
void orc_get_state(void *h, double *r, double *v, double *a, double *eta);
void orc_set_nh(void *h, double tr, double pnh1, double pnh2);
void orc_set_graphite(void *h, int n, const double *xyz);
void orc_set_threads(int n);
/* the expansion primitives Ylm (src/ifma.f90:12-26) and rLegendreP */
void orc_ylm(void *h, int n, int m, double a, double b, double *out);
double orc_legendre(int l, int m, double x);

/* every-10-steps work: molecule classes, storage boxes, site boxes, the
 * intramolecular subtraction masks and the far-field coefficients */
void orc_rebuild(void *h);
void orc_fmm(void *h);
void orc_get_classes(void *h, int *cls, int *anc, int *cbox, int *rkind, unsigned *isub);
/* multipole / local tables, complex as (re, im) pairs: O4/L4 nb^3*lp,
 * O3/L3 (nb/2)^3*lp, O2/L2 nb2^3*lp */
void orc_get_expansions(void *h, double *O4, double *L4, double *O3, double *L3,
                        double *O2, double *L2);

/* forces at the current positions: f (N,3) is overwritten.  near_mask selects
 * (receiver kind, source kind) combinations, bit rk*9+sk (all ones = all). */
void orc_forces(void *h, unsigned what, unsigned long long near_mask, double *f);

/* integrator: part A (predict + SHAKE + NH sums + eta update), part B
 * (accelerations, velocity, RATTLE), and a = F dt^2/2m */
void orc_move_a(void *h, double *sums6);
void orc_move_b(void *h, const double *f);
void orc_init_accel(void *h, const double *f);
int orc_shake_failures(void *h);

/* energies (not in the reference; the potentials of the force terms above):
 * components in this order, then their sum, kinetic, thermostat, total,
 * near-field pair virial sum r.F, temperature (2K/df).  orc_xi: the time
 * integral of eta accumulated by orc_move_a. */
enum {
    ORC_E_NEAR_COUL, ORC_E_NEAR_LJ, ORC_E_FAR, ORC_E_SUBTRACT, ORC_E_INTERNAL_COUL,
    ORC_E_INTERNAL_LJ, ORC_E_ANGLE, ORC_E_DIHEDRAL, ORC_E_DIHEDRAL_BE, ORC_E_WALL,
    ORC_E_GRAPHITE, ORC_E_POTENTIAL, ORC_E_KINETIC, ORC_E_THERMOSTAT, ORC_E_TOTAL,
    ORC_E_VIRIAL, ORC_E_TEMPERATURE, ORC_E_COUNT
};
void orc_energies(void *h, double *e);
double orc_xi(void *h);
void orc_set_xi(void *h, double xi);

/* full steps in the reference's cadence (rebuild + far field every ifar
 * steps): kb is the step counter, advanced by nsteps */
void orc_steps(void *h, int nsteps, int ifar, int *kb, double *f_scratch);

#ifdef __cplusplus
}
#endif
#endif
