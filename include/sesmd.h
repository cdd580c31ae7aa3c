/* sesmd -- MI355X (gfx950) per-timestep force and integration path of the
 * surfactant-droplet MD code jhalverson/surfactant_enhanced_spreading.
 *
 * Plain C ABI: opaque context, host pointers and sizes, int status codes
 * (0 = success, nonzero = error; the text is in sesmd_last_error()).  All
 * quantities are in the reference's reduced units (sigma = 3.166 A,
 * epsilon = 650.2 J/mol, m_O); host arrays are (N, 3) row-major doubles with
 * all waters (O, H1, H2) first, then 29-site surfactants -- the order
 * src/read_config.f90:19-34 reads them in.  Velocities are stored as v*dt and
 * accelerations as F dt^2 / (2m), as in src/init_accel.f90:21.
 *
 * State stays resident on the GPU between calls; the host arrays are only
 * read or written by the set_/get_ calls.  A context is not thread-safe: call
 * it from one thread, like the reference calls its routines from the rank's
 * main thread.
 */
#ifndef SESMD_H
#define SESMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sesmd_ctx sesmd_ctx;

/* Force field in a 32-type space (0..2 water O/H1/H2, 3..31 surfactant sites
 * 1..29).  The reference holds the same numbers in COMMON /cmn29/, /cmn30/,
 * /cmn100/, /cmn31/, /cmn48/ (src/mybox.cmns:192-201,236-239,256) plus the
 * site lists hard-coded in src/eval_valence.f90:13-95 and
 * src/eval_dihedral.f90:23-100. */
typedef struct sesmd_forcefield {
    double q[32], mass[32];              /* charge (e), mass / m_O            */
    double lj_c[32 * 32], lj_s6[32 * 32];/* (48 eps) s6 and s6 = sigma^6       */
    double gr_c[32], gr_s6[32];          /* the same against graphite carbon  */
    double dsqw[3], dsqs[28];            /* squared constraint lengths        */
    int ibond[28], jbond[28];            /* surfactant bonds (1-based sites)  */
    double th[36], rkth[36];             /* angle equilibria / force consts   */
    int ang_i[36], ang_j[36], ang_k[36];
    double dva[30], dvb[30], dvc[30];    /* dihedral coefficients             */
    int dih_i[30], dih_j[30], dih_k[30], dih_l[30];
    int n_internal;                      /* intramolecular LJ+Coulomb pairs   */
    int int_j[320], int_k[320];
    double rifact[33];                   /* 0! .. 32!                         */
    double xpi, rc, rs, dt, dtsq2;       /* Coulomb const, cutoff, box, dt    */
    double clamp_opls, clamp_be;         /* dihedral cos clamps               */
} sesmd_forcefield;

/* precision of the pair forces: SESMD_MIXED evaluates the near-field pairs in
 * FP32 and accumulates each pair force component in fixed point (resolution
 * 2^-32 force units, held in a biased FP64 register: range +-2^19 per sum;
 * see sesmd_get_flags),
 * and runs the M2L contractions as split-BF16x3 products with FP32
 * accumulation; expansions, integration, constraints and the thermostat stay
 * FP64 */
enum { SESMD_FP64 = 0, SESMD_MIXED = 1 };

/* force components, bits of `what` in sesmd_forces (one per reference
 * routine family) */
enum {
    SESMD_NEAR = 1,          /* the 39 eval[wS]* box-pair kernels             */
    SESMD_FAR = 2,           /* evalLjk4 (L2P of the frozen local expansions) */
    SESMD_INTERNAL = 4,      /* eval_surfactant_internal                      */
    SESMD_SUBTRACT = 8,      /* eval_subtract_fma                             */
    SESMD_VALENCE = 16,      /* eval_valence                                  */
    SESMD_DIHEDRAL = 32,     /* eval_dihedral                                 */
    SESMD_DIHEDRAL_BE = 64,  /* eval_dihedral_berendsen                       */
    SESMD_WALL = 128,        /* eval_wall                                     */
    SESMD_GRAPHITE = 256,    /* eval_graphite                                 */
    SESMD_ALL = 511
};

/* Create a context for n_water waters and n_surf surfactants in a cube of
 * (4 nb2)^3 L4 boxes of side ff->rs (the reference: nb2 = 3, i.e. its 27
 * ranks' L2 boxes, src/mybox.cmns:55-57), multipole order p (reference 4,
 * src/mybox.cmns:49).  Replaces the reference's start-up of
 * init_box_dimensions/get_box_center/init_arrays (src/mybox.f90:22,54-55). */
int sesmd_create(sesmd_ctx **out, int n_water, int n_surf, int nb2, int p,
                 const sesmd_forcefield *ff, int precision);
int sesmd_destroy(sesmd_ctx *ctx);
const char *sesmd_last_error(void);

/* positions, velocities (v dt), accelerations (F dt^2/2m) and the Nose-Hoover
 * friction eta; NULL skips an array */
int sesmd_set_state(sesmd_ctx *ctx, const double *r, const double *v, const double *a,
                    double eta);
int sesmd_get_state(sesmd_ctx *ctx, double *r, double *v, double *a, double *eta);
/* The Nose-Hoover friction eta and its time integral xi = integral of eta dt
 * (accumulated by every eta update; it enters only the thermostat energy
 * SESMD_E_THERMOSTAT).  Together with sesmd_get_state/sesmd_set_state this is
 * the whole dynamical state: a run restarted from them at a step that
 * rebuilds (kb % ifar == ifar - 1 before the call) continues bit for bit.
 * The reference restarts from data_cont2.dat positions and velocities with
 * eta and xi at zero (src/read_config.f90:19-34, src/mybox.f90:44-51). */
int sesmd_get_thermostat(sesmd_ctx *ctx, double *eta, double *xi);
int sesmd_set_thermostat(sesmd_ctx *ctx, double eta, double xi);
/* Nose-Hoover target tr and the constants pnh1, pnh2 of src/mybox.f90:44-51 */
int sesmd_set_nh(sesmd_ctx *ctx, double tr, double pnh1, double pnh2);
/* frozen graphite carbons (reduced units); the per-column selection of
 * src/read_graphite.f90:203-229 is made here */
int sesmd_set_graphite(sesmd_ctx *ctx, int n, const double *xyz);

/* Every-ifar-steps work of src/mybox.f90:142-175 on the GPU: molecule classes
 * and storage boxes (read_config/delete_*, ibox*), the per-box site sort that
 * replaces the build* lists, the intramolecular subtraction lists
 * (buildinternal) -- and, in sesmd_fmm, the far-field chain computeOnm4 ->
 * computeOnm3 -> computeOnm2 -> computeLjk2 -> shiftLjk2 -> computeLjk3 ->
 * shiftLjk3 -> computeLjk4. */
int sesmd_rebuild(sesmd_ctx *ctx);
int sesmd_fmm(sesmd_ctx *ctx);

/* Forces at the current positions into the device force array (overwritten).
 * near_mask selects (receiver kind, source kind) pairs of the near field, bit
 * rk*9+sk; ~0 = all.  Receiver kinds 0..5: W1@0 W2@0 W2@1 S@0 S@1 S@2, source
 * kinds 0..8: W1@0 W1@1 W2@0 W2@1 W2@2 S@0 S@1 S@2 S@3 (class @ storage ring);
 * each of the 39 eval[wS]* kernels is a fixed subset. */
int sesmd_forces(sesmd_ctx *ctx, unsigned what, uint64_t near_mask);
int sesmd_get_forces(sesmd_ctx *ctx, double *f);
int sesmd_set_forces(sesmd_ctx *ctx, const double *f);

/* Integrator.  sesmd_move_a: part A for every species (move_surfactant_a,
 * move_w1_a, move_w2_a of src/mybox.f90:136-138: predictor, SHAKE, the six
 * Nose-Hoover kinetic sums, the eta update of src/move_w2_a.f90:154-157, force
 * zeroing); sums6 (host, may be NULL) receives sv1 sv2 w1v1 w1v2 w2v1 w2v2.
 * sesmd_move_b: part B (move_*_b, src/mybox.f90:225-227) from the device
 * forces.  sesmd_init_accel: src/init_accel.f90. */
int sesmd_move_a(sesmd_ctx *ctx, double *sums6);
int sesmd_move_b(sesmd_ctx *ctx);
int sesmd_init_accel(sesmd_ctx *ctx);

/* nsteps full steps in the order of src/mybox.f90:130-227 (part A; rebuild +
 * far field when kb % ifar == 0; forces; part B), all on the GPU, the regular
 * steps replayed from a captured HIP graph.  *kb is the step counter.  The
 * first call for an ifar captures every graph it replays (the whole cycle of
 * ifar steps, the run of ifar - 1 regular steps, one regular step) before it
 * runs any step. */
int sesmd_steps(sesmd_ctx *ctx, int nsteps, int ifar, int *kb);

/* Inspection: classes (per molecule 0 W1, 1 W2, 2 S), storage box (global L4
 * index x + nb*(y + nb*z)), per-site central box, receiver kind (-1 none) and
 * subtraction mask (bit k: site k+1 of the same surfactant); any may be NULL */
int sesmd_get_classes(sesmd_ctx *ctx, int *cls, int *anc, int *cbox, int *rkind,
                      unsigned *isub);
/* expansions as complex (re, im) pairs: O4/L4 nb^3*lp, O3/L3 (nb/2)^3*lp, O2/L2
 * nb2^3*lp.  sesmd_fmm gives every box its full expansions; after
 * sesmd_steps the local expansions hold the M2L terms only for boxes with a
 * receiver site, and L4 is evaluated only for boxes holding sites (the only
 * ones a step reads). */
int sesmd_get_expansions(sesmd_ctx *ctx, double *O4, double *L4, double *O3, double *L3,
                         double *O2, double *L2);
/* counters: SHAKE/RATTLE molecules that hit the iteration limit, cumulative */
int sesmd_get_counters(sesmd_ctx *ctx, int *shake_fail);
/* Error flags counted on the device, cumulative over the context's life:
 * flags[0] SHAKE/RATTLE molecules that hit the iteration limit (the
 * reference's "Did not finish" messages, src/move_w1_a.f90:113-121);
 * flags[1] steps whose kinetic sums were not finite (a NaN or an infinity in
 * some velocity); flags[2] molecule filings whose filing site lay outside the
 * (4 nb2)^3-box cube (filed in the nearest edge box instead); flags[3]
 * near-field work items of a SESMD_MIXED context where a pair force component
 * or a receiver's sum reached 2^18 force units, half the range of its
 * fixed-point accumulation (2^-32 resolution, +-2^19).  n <= 4 entries are
 * written (more are zero). */
int sesmd_get_flags(sesmd_ctx *ctx, int *flags, int n);
/* device pointers of the resident SoA arrays (x,y,z blocks of N doubles) */
int sesmd_device_arrays(sesmd_ctx *ctx, void **r, void **v, void **a, void **f);
/* the HIP stream every kernel of this context is launched on */
void *sesmd_stream(sesmd_ctx *ctx);
/* Launch the near-field kernel alone (for timing) on the current sorted data. */
int sesmd_near_only(sesmd_ctx *ctx);
/* pair count of the current near field (unordered intermolecular pairs in the
 * 3^3 blocks, and those within rc with an LJ term), for roofline accounting */
int sesmd_pair_counts(sesmd_ctx *ctx, double *pairs, double *lj_pairs);
/* The same counts without allocating anything.  No counterpart in the
 * reference: this is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  counts[0] = pairs,
 * counts[1] = lj_pairs, equal to sesmd_pair_counts's; counts must be memory
 * the GPU can write (device or pinned host memory).  The work is only
 * enqueued on sesmd_stream(ctx): synchronise that stream before reading.
 * The partial sums live in a static device array of 1 MiB that each GPU holds
 * until the process ends, which limits the system to 8,388,608 sites (larger
 * ones are rejected); calls that overlap on one GPU share that array and
 * race. */
int sesmd_pair_counts_noalloc(sesmd_ctx *ctx, double *counts);

/* ---- finer-grained calls, used by the Fortran drop-in shim (sesmd_fortran.h)
 * to run one reference routine at a time ---- */

/* Fix molecule classes (0 W1, 1 W2, 2 S) and storage boxes instead of
 * classifying in sesmd_rebuild -- for a caller that files molecules itself,
 * as the reference's delete_w1/add_w1 migration does (src/delete_w1.f90:31-120).
 * NULL, NULL returns to automatic classification. */
int sesmd_set_classes(sesmd_ctx *ctx, const int *cls, const int *anc);
int sesmd_set_eta(sesmd_ctx *ctx, double eta);
/* load expansion tables (layout of sesmd_get_expansions; NULL leaves a table) */
int sesmd_set_expansions(sesmd_ctx *ctx, const double *O4, const double *L4, const double *O3,
                         const double *L3, const double *O2, const double *L2);
/* one stage of the far-field chain; each matches one reference routine */
enum {
    SESMD_P2M = 0,     /* computeOnm4 (src/computeOnm4.f90:22-162)  */
    SESMD_M2M_43 = 1,  /* computeOnm3 (src/computeOnm3.f90:16-98)   */
    SESMD_M2M_32 = 2,  /* computeOnm2 (src/computeOnm2.f90:14-63)   */
    SESMD_M2L_2 = 3,   /* computeLjk2 (src/computeLjk2.f90:17-73)   */
    SESMD_L2L_23 = 4,  /* shiftLjk2 (src/shiftLjk2.f90:14-41)       */
    SESMD_M2L_3 = 5,   /* computeLjk3 (src/computeLjk3.f90:45-102)  */
    SESMD_L2L_34 = 6,  /* shiftLjk3 (src/shiftLjk3.f90:16-70)       */
    SESMD_M2L_4 = 7    /* computeLjk4 (src/computeLjk4.f90:58-170)  */
};
int sesmd_fmm_stage(sesmd_ctx *ctx, int stage);
/* The same stage enqueued on sesmd_stream(ctx) without waiting for it (for
 * timing one far-field kernel with events on that stream). */
int sesmd_fmm_stage_enqueue(sesmd_ctx *ctx, int stage);
/* One integrator half for the molecule classes in cls_mask (bit 0 W1, 1 W2,
 * 2 S) without the eta update: part 0 = A (move_*_a without the all-reduce),
 * part 1 = B (move_*_b).  For part A, sums18 (may be NULL) receives the
 * unweighted sums of |v dt|^2 per site group, before SHAKE [0..8] and after
 * [9..17]: W1 O, W1 H, W2 O, W2 H, S H, S O, S CH2, S Si, S CH3 (each group
 * shares one mass; nh_water1/2 and nh_surfactant weight them). */
int sesmd_move_species(sesmd_ctx *ctx, int part, unsigned cls_mask, double *sums18);
/* The nine group sums of |v dt|^2 (order of sums18[0..8]) at the current
 * velocities: what nh_water1, nh_water2 and nh_surfactant weight by mass
 * (src/nh_water1.f90:1-43, src/nh_surfactant.f90:1-201). */
int sesmd_kinetic_sums(sesmd_ctx *ctx, double *sums9);

/* ---- energies (no counterpart in the reference, which reports none: the
 * potentials whose gradients are the force terms above, for energy and drift
 * reporting) ---- */
enum {
    SESMD_E_NEAR_COUL = 0,     /* 1/2 sum over receivers of xpi q_i q_j / r (the 39 kernels' pairs) */
    SESMD_E_NEAR_LJ = 1,       /* 1/2 sum of (c/12) r^-6 (s6 r^-6 - 1), r <= rc */
    SESMD_E_FAR = 2,           /* 1/2 sum q_i Phi(x_i), Phi the frozen local expansion (evalLjk4) */
    SESMD_E_SUBTRACT = 3,      /* minus the intramolecular Coulomb eval_subtract_fma removes */
    SESMD_E_INTERNAL_COUL = 4, /* eval_surfactant_internal, Coulomb */
    SESMD_E_INTERNAL_LJ = 5,   /* eval_surfactant_internal, LJ (no cutoff) */
    SESMD_E_ANGLE = 6,         /* 1/2 k (theta - theta0)^2 (eval_valence) */
    SESMD_E_DIHEDRAL = 7,      /* a(1+cos phi) + b(1-cos 2phi) + c(1+cos 3phi) (eval_dihedral) */
    SESMD_E_DIHEDRAL_BE = 8,   /* a(1+cos 3phi) (eval_dihedral_berendsen) */
    SESMD_E_WALL = 9,          /* 4/(x-h)^12 + 4/(x+h)^12 per axis (eval_wall) */
    SESMD_E_GRAPHITE = 10,     /* carbon LJ within rc (eval_graphite) */
    SESMD_E_POTENTIAL = 11,    /* sum of 0..10 */
    SESMD_E_KINETIC = 12,      /* 1/2 sum m v^2 (v = stored v dt / dt) */
    SESMD_E_THERMOSTAT = 13,   /* Q eta^2 / 2 + df tr xi, Q = dt / (2 pnh1), xi = integral of eta dt */
    SESMD_E_TOTAL = 14,        /* potential + kinetic + thermostat: conserved by the dynamics */
    SESMD_E_VIRIAL = 15,       /* near-field pair virial, 1/2 sum r . F */
    SESMD_E_TEMPERATURE = 16,  /* 2 K / df (src/compute_temperature.f90:242) */
    SESMD_E_COUNT = 17
};
/* All components at the current state, in FP64 whatever the context's
 * precision (reduced units: energies in epsilon).  xi accumulates over the
 * steps run since the context was created. */
int sesmd_energies(sesmd_ctx *ctx, double *e);

/* Ensembles: n independent systems (one built context each, same precision
 * and device, equal hydrogen charges) advanced together -- the reference's
 * step loop (src/mybox.f90:130-241) for every member in one captured graph,
 * with the near field of all members in one persistent launch over their
 * flattened row-item lists, and -- when every member has the reference
 * surfactant topology, one expansion order and the same constraint
 * constants -- the members' integrator and site-term launches of a step
 * batched into one launch each (a batched rebuild evaluates multipoles and
 * L4 local expansions only for the boxes holding sites, the only ones a step
 * reads: sesmd_get_expansions on a member then gives unspecified values for
 * the other boxes).  Each member's state after sesmd_ensemble_steps
 * is bit-for-bit what sesmd_steps on that context alone gives.  Members stay
 * owned by the caller and must outlive the ensemble; while an ensemble steps
 * them, their own calls must not run.  No counterpart in the reference (it
 * runs one system per MPI job). */
typedef struct sesmd_ensemble sesmd_ensemble;
int sesmd_ensemble_create(sesmd_ensemble **out, sesmd_ctx *const *members, int n);
int sesmd_ensemble_destroy(sesmd_ensemble *ens);
/* launch shape of the ensemble's near field: wavefronts per workgroup (1, 2,
 * 4), receivers per wavefront (64, 128), workgroups (the persistent grid) */
int sesmd_ensemble_set_shape(sesmd_ensemble *ens, int waves, int chunk, int grid);
int sesmd_ensemble_get_shape(sesmd_ensemble *ens, int *shape3);
/* the autotuner: times every shape (reps launches each) at the members'
 * current state, keeps and returns the fastest and its time per launch */
int sesmd_ensemble_tune(sesmd_ensemble *ens, int reps, int *shape3, float *best_ms);
/* the near field of every member once, enqueued on the ensemble's stream
 * (timing) */
int sesmd_ensemble_near_only(sesmd_ensemble *ens);
/* the ensemble's stream (a hipStream_t) */
void *sesmd_ensemble_stream(sesmd_ensemble *ens);
/* 1: the members' integrator and site-term launches are batched (one launch
 * each per step); 0: they run per member on the branch streams; -1: null */
int sesmd_ensemble_batched(sesmd_ensemble *ens);
/* nsteps steps of every member; kb is the shared step counter (sesmd_steps) */
int sesmd_ensemble_steps(sesmd_ensemble *ens, int nsteps, int ifar, int *kb);

#ifdef __cplusplus
}
#endif
#endif
