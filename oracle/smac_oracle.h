/*
 * CPU restatement of miniSMAC2D's implicit iteration -- TEST INFRASTRUCTURE ONLY.
 *
 * The oracle is the checker for the gfx950 path: tests, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it; the product never does.  It restates the
 * reference routines one to one (file:line citations in smac_oracle.c) in plain C99
 * built with -O2 -ffp-contract=off, keeping the reference's operation order.  The
 * reference itself cannot be built in this image (its sources include mpif.h; there is no
 * MPI), so the restatement is pinned by tests/test_oracle_pinning.py against what the
 * mathematics fixes (dense solves of the factorised systems, factor identities, the
 * wake-cut mean, freestream preservation and the block-sum identity, brute-force wall
 * distance, first-index argmax), by tests/test_oracle_closed_forms.py (the face split as the
 * positive part of the flux Jacobian, third-order flux convergence with second-order ends,
 * closed-form viscous terms, the far-field vortex and characteristic rows, the SA source in
 * closed form and its linearisation against differences, second-order convergence of the
 * wall force), each with a planted-slip negative control, and by tests/test_grid_metric.py
 * (exact metrics of affine grids).
 *
 * Layout: the reference's Fortran arrays (j fastest; see include/smac_amd.h).
 */
#ifndef SMAC_ORACLE_H
#define SMAC_ORACLE_H

#include "../include/smac_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_case orc_case;

/* rtxy/dj may be NULL: then they are formed from x, y as ref/metric.f does. */
orc_case *orc_create(const smac_params *p, int jmax, int kmax, const double *x, const double *y,
                     const double *rtxy, const double *dj, int nreg, const smac_bc_region *regs);
void orc_destroy(orc_case *c);
/* Free-stream initial condition with u=v=0 on walls (ref/geom.f:88-131). */
void orc_ic(orc_case *c);
/* One pass of solve's loop body: nt+1, step, force + clift (ref/smac2d.f:1386-1407). */
int orc_step(orc_case *c, smac_iter_stats *st);
/* The parts of step, for kernel-level checks:
 *   orc_assemble: zero fill, SA (with its own solve and update), conup, viscterms, blanking,
 *                 diagonal, bcmain, resid and rhsnorm (ref/step.f:104-519)
 *   orc_solve:    lrelax or lusgs (ref/step.f:537-560)
 *   orc_finish:   update and the force/clift part of iterout (ref/step.f:581,
 *                 ref/smac2d.f:1709-1825) */
int orc_assemble(orc_case *c, smac_iter_stats *st);
int orc_solve(orc_case *c);
int orc_finish(orc_case *c, smac_iter_stats *st);
/* Named arrays (live, Fortran layout): q dq s anut vnut btc bjm bjp bkm bkp rtxy dj x y
 * smin, and "lhs_*" copies of the five blocks as assembled (before the solver factors
 * them).  Returns NULL for an unknown name. */
double *orc_array(orc_case *c, const char *name);
/* Set nt, resmax0, clift (restart from a saved state). */
void orc_set_counters(orc_case *c, int nt, double resmax0, double clift);

/* The face split of conup on its own (ref/conup.f:127-185): A+ and A- of one face,
 * entry (m,n) at m + 3n. */
void orc_face_split(double rt, double rx, double ry, double q2, double q3, double beta,
                    double *ap, double *am);

/* ref/metric.f:49-278 on its own. */
void orc_metric(int jmax, int kmax, const double *x, const double *y, double *rtxy, double *dj);

/* Stand-alone line relaxation / LU-SGS on caller arrays (same contract as smac_lrelax /
 * smac_lusgs in include/smac_amd.h; btc/bjp/bkp are modified in place as the reference
 * modifies them). */
int orc_lrelax(const smac_params *p, int jmax, int kmax, int npair, const int *wake,
               const double *q, double *dq, double *s, double *btc, double *bjm, double *bjp,
               double *bkm, double *bkp);
int orc_lusgs(const smac_params *p, int jmax, int kmax, int neqs, int npair, const int *wake,
              const double *q, double *dq, double *rhs, double *btc, double *bjm, double *bjp,
              double *bkm, double *bkp, double *scr1);

#ifdef __cplusplus
}
#endif
#endif
