/* sesmd_fortran -- the reference's own routine names, for linking in place of
 * the reference's Fortran objects (libsesmd_fortran.so, built over libsesmd).
 *
 * Every entry point has the external name and argument list of the Fortran
 * routine it replaces (lower case, one trailing underscore, all arguments by
 * reference), reads its inputs from and writes its outputs to the reference's
 * COMMON blocks in place (src/mybox.cmns:123-259), and computes on the GPU
 * through the sesmd engine.  Ownership rules are the reference's:
 *   - forces are accumulated (+=) into fxw1/fxw2/fxs: pair and far-field
 *     forces on the storage copies of the sites central in the rank's own
 *     4^3 L4 boxes, bonded/intramolecular and wall forces on the molecules
 *     stored in own boxes;
 *   - move_*_a zero their species' force arrays, computeOnm4 zeroes cO4,
 *     computeOnm3 cO3, computeLjk2 cL2; shiftLjk2/shiftLjk3 overwrite
 *     cL3/cL4 and computeLjk3/computeLjk4 accumulate into them.
 * The pair-list routines build* write the reference's lists and also fix the
 * positions that define "central in box" until the next build*, as the
 * reference's lists do between rebuilds.  No routine returns an error:
 * failures are counted and described by sesmd_fortran_last_error(), and the
 * COMMON outputs are left untouched.
 *
 * COMMON blocks are weak references resolved against the Fortran program (or
 * library) that defines them; called without them, every routine fails
 * softly.  One GPU context per process (per rank), used from one thread.
 */
#ifndef SESMD_FORTRAN_H
#define SESMD_FORTRAN_H

#include <stdint.h>
#include "sesmd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- per-box pair lists and intramolecular subtraction lists ---- */
void buildw2_(const int32_t *iaq);     /* src/buildW2.f90:1-71      */
void buildw2n_(const int32_t *iaq);    /* src/buildW2n.f90          */
void buildw2nn_(const int32_t *iaq);   /* src/buildW2nn.f90         */
void builds_(const int32_t *iaq);      /* src/buildS.f90            */
void buildsn_(const int32_t *iaq);     /* src/buildSn.f90           */
void buildsnn_(const int32_t *iaq);    /* src/buildSnn.f90          */
void buildsnnn_(const int32_t *iaq);   /* src/buildSnnn.f90         */
void buildinternal_(const int32_t *iaq); /* src/buildinternal.f90:1-63 */

/* ---- the 39 near-field box-pair kernels (src/eval*.f90), all
 * `subroutine evalXX(iaq)`; each adds its (receiver, source) subset ---- */
void evalw1w1_(const int32_t *iaq);    /* src/evalw1w1.f90:24-245   */
void evalw1w2_(const int32_t *iaq);
void evalw1s_(const int32_t *iaq);     /* src/evalw1s.f90:13-98     */
void evalw2w2_(const int32_t *iaq);    /* src/evalw2w2.f90          */
void evalw2s_(const int32_t *iaq);
void evalss_(const int32_t *iaq);      /* src/evalSS.f90:13-125     */
void evalw1w1n_(const int32_t *iaq);   /* src/evalw1w1n.f90:33-229  */
void evalw1w2n_(const int32_t *iaq);
void evalw1w2nn_(const int32_t *iaq);
void evalw1sn_(const int32_t *iaq);
void evalw1snn_(const int32_t *iaq);
void evalw1snnn_(const int32_t *iaq);
void evalw2w1n_(const int32_t *iaq);
void evalw2w2n_(const int32_t *iaq);
void evalw2w2nn_(const int32_t *iaq);
void evalw2sn_(const int32_t *iaq);
void evalw2snn_(const int32_t *iaq);
void evalw2snnn_(const int32_t *iaq);
void evalw2nw1n_(const int32_t *iaq);
void evalw2nw2n_(const int32_t *iaq);  /* src/evalw2nw2n.f90        */
void evalw2nw2nn_(const int32_t *iaq);
void evalw2ns_(const int32_t *iaq);
void evalw2nsn_(const int32_t *iaq);
void evalw2nsnn_(const int32_t *iaq);
void evalw2nsnnn_(const int32_t *iaq);
void evalsw1n_(const int32_t *iaq);    /* src/evalSw1n.f90:20-112   */
void evalsw2nn_(const int32_t *iaq);
void evalssn_(const int32_t *iaq);     /* src/evalSSn.f90:13-147    */
void evalssnn_(const int32_t *iaq);
void evalssnnn_(const int32_t *iaq);
void evalsnw1n_(const int32_t *iaq);
void evalsnw2nn_(const int32_t *iaq);
void evalsnsn_(const int32_t *iaq);
void evalsnsnn_(const int32_t *iaq);
void evalsnsnnn_(const int32_t *iaq);
void evalsnnw1n_(const int32_t *iaq);
void evalsnnw2nn_(const int32_t *iaq);
void evalsnnsnn_(const int32_t *iaq);
void evalsnnsnnn_(const int32_t *iaq);

/* ---- far field ---- */
void computeonm4_(const int32_t *iaq); /* src/computeOnm4.f90:22-162 (P2M; the
                                          diagnostic charge mpi_reduce of
                                          :203-207 is not repeated) */
void computeonm3_(const int32_t *iaq); /* src/computeOnm3.f90:16-98 (M2M 4->3) */
void computeonm2_(const int32_t *iaq); /* src/computeOnm2.f90:14-63 (M2M 3->2) */
void computeljk2_(const int32_t *iaq, const int32_t ibq[3]); /* src/computeLjk2.f90:17-73 */
void shiftljk2_(const int32_t *iaq);   /* src/shiftLjk2.f90:14-41 (L2L 2->3)   */
void computeljk3_(const int32_t *iaq); /* src/computeLjk3.f90:45-102 (M2L L3)  */
void shiftljk3_(const int32_t *iaq);   /* src/shiftLjk3.f90:16-70 (L2L 3->4)   */
void computeljk4_(const int32_t *iaq); /* src/computeLjk4.f90:58-170 (M2L L4)  */
void evalljk4_(const int32_t *iaq);    /* src/evalLjk4.f90:21-284 (L2P)        */

/* ---- bonded, intramolecular, external ---- */
void eval_subtract_fma_(const int32_t *iaq);        /* src/eval_subtract_fma.f90:10-54 */
void eval_surfactant_internal_(const int32_t *iaq); /* src/eval_surfactant_internal.f90:12-110 */
void eval_valence_(const int32_t *iaq);             /* src/eval_valence.f90:13-158 */
void eval_dihedral_(const int32_t *iaq);            /* src/eval_dihedral.f90:23-295 */
void eval_dihedral_berendsen_(const int32_t *iaq);  /* src/eval_dihedral_berendsen.f90 */
void eval_wall_(void);                              /* src/eval_wall.f90:8-59 */
void eval_graphite_(const int32_t *iaq, const int32_t *ibq); /* src/eval_graphite.f90:10-135 */

/* ---- integrator and thermostat ---- */
void move_surfactant_a_(const int32_t *iaq);        /* src/move_surfactant_a.f90:19-147 */
void move_w1_a_(const int32_t *iaq);                /* src/move_w1_a.f90:19-146 */
/* src/move_w2_a.f90:22-161: the six thermostat sums are summed over ranks by
 * the function registered with sesmd_fortran_set_allreduce (none: this
 * process is the only rank), then eta is updated as at :154-157. */
void move_w2_a_(const int32_t *comm, const int32_t *iaq);
void move_surfactant_b_(const int32_t *iaq);        /* src/move_surfactant_b.f90:23-122 */
void move_w1_b_(const int32_t *iaq);                /* src/move_w1_b.f90:23-123 */
void move_w2_b_(const int32_t *iaq);                /* src/move_w2_b.f90 */
void init_accel_(const int32_t *iaq);               /* src/init_accel.f90:1-49 */
void nh_water1_(double *out);                       /* src/nh_water1.f90:1-43 */
void nh_water2_(double *out);                       /* src/nh_water2.f90 */
void nh_surfactant_(double *out);                   /* src/nh_surfactant.f90:1-201 */

/* ---- box maps (src/ibox.f90:1-220): integer functions of positions in
 * reduced units; int() truncates toward zero.  ibox2/3/4: global L2/L3/L4
 * box (1-based) from /cmn03/ /cmn04/ and ibdx3.. /cmn01/, ibdx4.. /cmn02/;
 * ibox64/ibox1000: box in the rank's 4^3 / 10^3 grid around rcen2*tmp
 * (/cmn17/); ibox64mod: the 4^3 grid around a given centre; ib64to216/512/1000,
 * ib8to216, ib512to64: storage index maps.  isnearest / iscentral are Fortran
 * LOGICAL(4) functions (1 = .true., 0 = .false.): the half-open tests
 * [c - 1.5 rs, c + 1.5 rs) and [c - rs/2, c + rs/2) on every axis. ---- */
int32_t ibox2_(const double *x, const double *y, const double *z);      /* src/ibox.f90:127-139 */
int32_t ibox3_(const double *x, const double *y, const double *z);      /* src/ibox.f90:142-154 */
int32_t ibox4_(const double *x, const double *y, const double *z);      /* src/ibox.f90:157-169 */
int32_t ibox64_(const double *x, const double *y, const double *z);     /* src/ibox.f90:172-186 */
int32_t ibox64mod_(const double *x, const double *y, const double *z,
                   const double *cx, const double *cy, const double *cz); /* src/ibox.f90:189-203 */
int32_t ibox1000_(const double *x, const double *y, const double *z);   /* src/ibox.f90:206-220 */
int32_t ib64to216_(const int32_t *iaq);   /* src/ibox.f90:31-43 */
int32_t ib64to512_(const int32_t *iaq);   /* src/ibox.f90:16-28 */
int32_t ib64to1000_(const int32_t *iaq);  /* src/ibox.f90:1-13  */
int32_t ib8to216_(const int32_t *iaq);    /* src/ibox.f90:46-58 */
int32_t ib512to64_(const int32_t *iaq);   /* src/ibox.f90:61-72 */
int32_t isnearest_(const double *x, const double *y, const double *z,
                   const double *cx, const double *cy, const double *cz); /* src/ibox.f90:75-98  */
int32_t iscentral_(const double *x, const double *y, const double *z,
                   const double *cx, const double *cy, const double *cz); /* src/ibox.f90:101-124 */

/* ---- library calls with no Fortran counterpart ---- */
/* Sum n doubles over all ranks of communicator *comm (the Fortran handle
 * move_w2_a receives), in place; the reference's mpi_allreduce calls of
 * src/move_w2_a.f90:145-152.  NULL: single rank. */
typedef void (*sesmd_allreduce_fn)(double *vals, int32_t n, const int32_t *comm);
void sesmd_fortran_set_allreduce(sesmd_allreduce_fn fn);
/* the force field the library assembled from COMMON /cmn29/ /cmn30/ /cmn100/
 * /cmn31/ /cmn48/ and the reference's fixed tables; nonzero if COMMON is
 * absent.  Needs no GPU. */
int sesmd_fortran_forcefield(sesmd_forcefield *out);
const char *sesmd_fortran_last_error(void);
int32_t sesmd_fortran_errors(void);        /* failed calls so far */
void sesmd_fortran_release(void);          /* free the GPU contexts */

#ifdef __cplusplus
}
#endif
#endif
