/*
 * smac_amd.h -- C ABI of the MI355X (gfx950) implementation of miniSMAC2D's implicit
 * pseudo-time iteration (one call of the reference `step`, ref/step.f:53-609).
 *
 * Conventions shared by every entry point
 *  - Arrays are the reference's: Fortran order, j fastest.  A 3-vector field q(j,k,n) is
 *    3 planes of jmax*kmax doubles; a 3x3 block field b(j,k,m,n) is 9 planes ordered
 *    (m,n) = (1,1),(2,1),(3,1),(1,2),...,(3,3) (plane index (m-1) + 3*(n-1)), m = equation
 *    (row), n = variable (column); rtxy(j,k,a,b) is 6 planes (a-1) + 2*(b-1).
 *  - Indices handed to or returned from the library (BC regions, jres/kres) are 1-based.
 *  - Every call returns an smac_status; on failure smac_last_error() names the cause.
 *  - A handle is used from one host thread at a time; handles are independent (no global
 *    mutable state besides the last-error text of each thread).
 */
#ifndef SMAC_AMD_H
#define SMAC_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum smac_status {
    SMAC_OK = 0,
    SMAC_ERR_ARG = 1,          /* null pointer, bad size, inconsistent arrays            */
    SMAC_ERR_IMPSCH = 2,       /* impsch not 1 (line relaxation) or 2 (LU-SGS)           */
    SMAC_ERR_IFLXO = 3,        /* iflxo not 3 (only the third-order flux is reachable)   */
    SMAC_ERR_NEQS = 4,         /* neqs not 1 or 3                                         */
    SMAC_ERR_SIZE = 5,         /* jmax/kmax beyond the supported range                    */
    SMAC_ERR_BC = 6,           /* unsupported or malformed boundary region                */
    SMAC_ERR_NONFINITE = 7,    /* residual or update produced Inf/NaN                     */
    SMAC_ERR_HIP = 8,          /* a HIP runtime call failed                               */
    SMAC_ERR_NOGPU = 9,        /* no gfx950 device visible                                */
    SMAC_ERR_STALL = 10,       /* an LU-SGS band hand-over wait timed out (bounded spin)  */
    SMAC_ERR_NOCONV = 11       /* smac_adjoint: GMRES stopped at maxit above 10 x tol     */
} smac_status;

/* Solver settings: the reference's namelist $input (ref/initia.f:85-89) plus the values its
 * inizone/initia hard-code (ref/initia.f:104-112, 418-437; ref/geom.f:72-76). */
typedef struct smac_params {
    double alpha;      /* angle of attack, degrees                                  */
    double beta;       /* artificial compressibility                                */
    double reynum;     /* Reynolds number; vnu = 1/reynum (ref/initia.f:324)        */
    double dtau;       /* pseudo-time step                                          */
    double dt;         /* SA time step (1e12 in the shipped input)                  */
    double underr;     /* line-relaxation under-relaxation, 0 < underr <= 1          */
    double cdis;       /* flux dissipation factor (1)                               */
    double pin, pout;  /* inflow / outflow pressure (1.5, 1.0)                      */
    double chord, xmom, ymom;   /* force normalisation and moment centre (1, .25, 0)  */
    int impsch;        /* 1 = line relaxation (lrelax), 2 = LU-SGS (lusgs)           */
    int ivis;          /* 1 = viscous                                                */
    int iturb;         /* 1 = Spalart-Allmaras (needs ivis = 1)                      */
    int iflxo;         /* flux order, must be 3                                      */
    int njsp, nksp;    /* xi / eta line sweeps per iteration (2, 2)                  */
    int nbcimp;        /* wake-injection frequency in sweeps (1)                     */
    int ntjsp, ntksp;  /* SA j / k line sweeps (2, 0)                                */
    int precision;     /* SMAC_FP64 or SMAC_FP32_LHS                                  */
    int line_order;    /* SMAC_ORDER_REFERENCE or SMAC_ORDER_ZEBRA                    */
} smac_params;

enum { SMAC_FP64 = 0, SMAC_FP32_LHS = 1 };
/* SMAC_ORDER_REFERENCE: lines in the reference's Gauss-Seidel order (per-iteration parity).
 * SMAC_ORDER_ZEBRA (line relaxation only): lines 1, 3, 5, ... (1-based), then 2, 4, ...,
 * every line of a colour at once, in each sweep direction; the SA's k lines follow the same
 * order.  A different iteration with the same fixed point, judged at convergence.  FP32-LHS
 * and zebra line relaxation use the parallel line solves (option "lr_solve" = 1). */
enum { SMAC_ORDER_REFERENCE = 0, SMAC_ORDER_ZEBRA = 1 };

/* One BC region of bcmain.dat with indices already resolved to 1..jmax / 1..kmax
 * (ref/rbc.f:384-396).  ibcval: 0 no-slip wall, 25 C-grid far field with point vortex,
 * 31 outflow (fixed p), 60 wake cut; a region is a j-line (jbeg == jend) or a k-line. */
typedef struct smac_bc_region {
    int ibcval, jbeg, jend, kbeg, kend;
} smac_bc_region;

/* Per-iteration results: what the reference's solve/iterout keep in COMMON
 * (/rmssr/, /rmssi/, /airfr/clift; ref/step.f:613-756, ref/smac2d.f:1709-1825). */
typedef struct smac_iter_stats {
    int nt;            /* iteration counter after this call                         */
    int jres, kres;    /* location of resmax                                        */
    int nonfinite;     /* 1 if the residual or q went non-finite                    */
    double resmax, resmax0, divmax, turres;
    double rhsnorm;    /* ||s||_2 entering the solver (ref/step.f:513-519)           */
    double cl, cd, cm, clp, cdp;
    double fxp, fyp, fxs, fys, fmom;
} smac_iter_stats;

typedef struct smac_handle smac_handle;

/* ---- Tier 2: device-resident iteration ------------------------------------------------- */

/* Set up one case on the current HIP device: grid, metrics (as ref/metric.f forms them),
 * BC regions (wall list and wake pairs are derived as ref/rbc.f:972-990, 1160-1470 do),
 * and the free-stream initial condition of ref/geom.f:88-131.
 * Replaces the state solve() builds before its time loop (ref/smac2d.f:1239-1332). */
smac_status smac_create(const smac_params *p, int jmax, int kmax, const double *x,
                        const double *y, const double *rtxy, const double *dj, int nreg,
                        const smac_bc_region *regs, smac_handle **out);
smac_status smac_destroy(smac_handle *h);

/* Overwrite q (3 planes) and anut (1 plane); either may be NULL to keep it.  nt, resmax0
 * and clift are set as given (a restart from a saved state). */
smac_status smac_set_state(smac_handle *h, const double *q, const double *anut, int nt,
                           double resmax0, double clift);
/* Copy q (3 planes), anut, vnut (1 plane each) to the host; any pointer may be NULL. */
smac_status smac_get_state(smac_handle *h, double *q, double *anut, double *vnut);

/* One pass of the reference's time loop body (ref/smac2d.f:1386-1407): nt += 1, step
 * (SA, assembly, BC rows, residual, line relaxation or LU-SGS, update), then the wall
 * force integration that sets clift for the next pass.  Replaces step() at
 * ref/smac2d.f:1395 together with the force/clift part of iterout at ref/smac2d.f:1407. */
smac_status smac_step(smac_handle *h, smac_iter_stats *st);

/* niter passes of smac_step.  With drop > 0 it stops early once rhsnorm <= drop * r0, r0 the
 * rhsnorm of the case's first pass (nt = 1) when this handle made it, else of the first pass
 * of this call; the stats are then read back after every pass (one small copy and a stream
 * synchronisation).  With drop = 0 (and timing off) the passes are enqueued without host
 * round trips and the stats are read once, after the last.  *done = passes made. */
smac_status smac_run(smac_handle *h, int niter, double drop, smac_iter_stats *last, int *done);

/* Stats of the last pass made on h (by smac_step, smac_run, smac_batch_run or
 * smac_batch_run_until). */
smac_status smac_get_stats(smac_handle *h, smac_iter_stats *st);

/* ---- Batches of cases (no counterpart in the reference: its cases are separate runs) ----
 * A fixed set of cases advanced together on one stream of their own.  The LU-SGS band
 * sweeps -- a few wavefronts per case, bound by the latency of their hyperplane steps -- run
 * for all cases in one launch per sweep direction (one for the mean flow of each precision,
 * one for the SA model), so the bands of every case sweep side by side; the fully parallel
 * kernels of each case are enqueued between them.  Each case gets exactly the kernels and
 * operands of its own pass: its results are those of smac_step on it alone.  The cases may
 * differ in size, scheme, settings and precision; options "timing" and "keep_lhs" must be
 * off; all must live on the device that is current when the batch is created.  A case must
 * not be stepped on its own while a smac_batch_run on it is in progress. */
typedef struct smac_batch smac_batch;
smac_status smac_batch_create(smac_handle **cases, int n, smac_batch **out);
/* npass passes of every case; returns after the device has finished them, having checked
 * each case's non-finite and stall flags. */
smac_status smac_batch_run(smac_batch *b, int npass);
/* Up to max_pass passes with a stop test per case: a case stops after the pass whose ||s||_2
 * is at most drop times its first pass's, the criterion of smac_run and of the reference's
 * time loop (ref/smac2d.f:1376-1431), or after a non-finite or stalled pass; it gets no
 * further passes, so its state and pass count equal smac_run(h, max_pass, drop) on it alone.
 * With every case on LU-SGS (impsch 2) the test runs on the device and the host enqueues the
 * passes without reading statistics (it looks at the cases' flags every 16 passes and returns
 * once all have stopped); otherwise the host reads the flags after every pass.  passes_out
 * (n ints, may be NULL) receives each case's passes; drop 0 runs every case max_pass passes. */
smac_status smac_batch_run_until(smac_batch *b, int max_pass, double drop, int *passes_out);
/* Device time (ms) of the forward mean-flow band sweep launch of the last pass of the last
 * smac_batch_run (the launch covering the most points: out[1] grid points, precision
 * out[2]: 0 FP64, 1 FP32-LHS), from HIP events on the batch's stream. */
smac_status smac_batch_timings(smac_batch *b, double out[3]);
smac_status smac_batch_destroy(smac_batch *b);

/* Copy a named device array to the host (diagnostics and kernel-level tests):
 * "q","dq","s","anut","vnut","smin","sf","stc","ds","vort" and, with option "keep_lhs", the
 * assembled LHS "btc","bjm","bjp","bkm","bkp" (FP64 planes as assembled, before the solver
 * factors them). */
smac_status smac_get_array(smac_handle *h, const char *name, double *out);

/* Averaged device time (ms) per pass of the last smac_run (option "timing" on), per kernel
 * family: out[0] SA, [1] assembly + BC rows, [2] residual norms, [3] line relaxation or
 * LU-SGS, [4] update + wall forces, [5] the forward band sweep of the mean-flow LU-SGS alone
 * (0 for line relaxation), [6] whole pass.  With "timing" on, every family runs on one
 * stream so that the events bracket it. */
smac_status smac_timings(smac_handle *h, double out[7]);

/* Options: "keep_lhs" (1: keep a copy of the five assembled blocks of each pass for
 * smac_get_array), "timing" (1: time kernel families with HIP events on the handle's
 * stream), "lusgs_rows" (1..64, default 64: grid rows per wavefront band of the LU-SGS
 * sweeps), "lusgs_prefetch" / "lusgs_prefetch_sa" (1, 2, 3, 4, 6 or 8; defaults 4 and 8:
 * steps of operand prefetch in the 3x3 / scalar LU-SGS sweeps), "lusgs_ring" (0 or 1, default 1:
 * the 3x3 sweeps stream their operands through an LDS ring fed by loader wavefronts),
 * "lr_lines" (1..64, default 64: lines per workgroup of the line-relaxation
 * factorisations), "adj_restart" (4..400, default 100: GMRES restart length of
 * smac_adjoint), "adj_prec_steps" (1..8, default 3: steps of the transposed update per
 * application of its GMRES preconditioner, with the operator between them; more steps, fewer
 * iterations), "adj_reorth" (0..1000, default 100: its Gram-Schmidt projects a vector a
 * second time when less than this many thousandths of its square norm survive the first
 * projection; 500 is the DGKS test, 0 never), "lr_refactor" (1..64, default 1: the
 * scan and zebra line-relaxation forms factor their lines every this many passes and keep
 * the factored lines and neighbour blocks in between, with each pass's own residual: a
 * lagged implicit operator with the same fixed point, not the reference's iteration).  All
 * but the first two and the last are tuning knobs; results do not depend on them beyond the
 * solver tolerances.
 * No counterpart in the reference. */
smac_status smac_set_option(smac_handle *h, const char *name, int value);

/* ---- Discrete adjoint (no counterpart in the reference, which computes no sensitivities) --
 * Gradients of CL and CD at the case's current state, which must be a steady state of the
 * iteration (run it to a tight drop first).  The steady equations are the assembled residual
 * s = 0 on every row but the wake-cut targets, the wake-cut relation
 * q(jt,1) = (q(jt,2) + q(jb,2)) / 2 there (what the injection ref/bc.f:571-591 enforces), and
 * clift = CL (the far field reads the CL of the pass before).  The adjoint of that system is
 * solved once for CL and once for CD with restarted GMRES (option "adj_restart", default
 * 100), right-preconditioned by the transpose of the iteration's own update: the LU-SGS solve
 * for impsch = 2, the zebra line-relaxation sweeps for impsch = 1 (whatever the line order),
 * with a Sherman-Morrison correction for the rank-one lift coupling of the far field.  The
 * mean flow does not read the SA variables, so the SA part of the adjoint vanishes.
 *  - dcl[0], dcd[0]: d/d(alpha), alpha in degrees;
 *  - dcl[1+i], dcd[1+i]: along shape direction i, given as grid displacements dx, dy
 *    (nshape x jmax*kmax each, Fortran order; the metrics follow them as the reference forms
 *    them); nshape may be 0.
 * maxit bounds the GMRES iterations of each solve, tol is the relative residual to reach.
 * A solve that stops at maxit with a residual above 10 x tol makes the call return
 * SMAC_ERR_NOCONV after filling dcl, dcd and st (the gradients are then not reliable: a short
 * restart can stall on the lift mode). */
typedef struct smac_adjoint_stats {
    int iters[2];      /* GMRES iterations of the CL and the CD adjoint                        */
    double resid[2];   /* their final residuals relative to the right-hand side             */
    double rhsnorm;    /* ||s||_2 of the state the gradients are taken at                   */
    double seconds;    /* wall time of the call                                              */
} smac_adjoint_stats;
smac_status smac_adjoint(smac_handle *h, int nshape, const double *dx, const double *dy,
                         int maxit, double tol, double *dcl, double *dcd,
                         smac_adjoint_stats *st);
/* The operators of the last smac_adjoint call applied to a host vector (3 planes):
 * op 0: F_q^T in; 1: the adjoint operator -F_q^T in - (in . F_c) CL_q; 2: P^T in, the
 * transposed update; 3: P in, the iteration's own update (LU-SGS solve and wake injection).
 * For checks of the adjoint (dot-product tests); arrays "adj_jac" (81 planes: 9 offsets x 9
 * entries), "adj_lam_cl", "adj_lam_cd", "adj_gcl", "adj_gcd" come from smac_get_array. */
smac_status smac_adjoint_apply(smac_handle *h, int op, const double *in, double *out);

/* ---- Tier 1: drop-ins at the reference's own Fortran call sites (host arrays) --------- */

/* lrelax(jmax,kmax,q,dq,s,btc,bjm,bjp,bkm,bkp), ref/lrelax.f:51-221, called at
 * ref/step.f:538.  Writes dq; writes s at the wake-cut points as bcimpds does
 * (ref/bc.f:571-591).  btc/bjp/bkp are read only (the reference factors them in place and
 * recomposes them; its caller never reads them afterwards).  wake holds npair (jt, jb)
 * pairs: dq(jt,1) is injected from the average of points jt and jb at k=2. */
smac_status smac_lrelax(const smac_params *p, int jmax, int kmax, int npair, const int *wake,
                        const double *q, double *dq, double *s, const double *btc,
                        const double *bjm, const double *bjp, const double *bkm,
                        const double *bkp);

/* lusgs(jmax,kmax,neqs,q,dq,rhs,btc,bjm,bjp,bkm,bkp,scr1), ref/lusgs.f:51-181, called at
 * ref/step.f:552 (neqs=3) and ref/tursa.f:252 (neqs=1).  Writes dq, scr1 (the LU of btc)
 * and rhs (destroyed as in the reference: the forward-sweep result, with the wake-cut
 * points overwritten). */
smac_status smac_lusgs(const smac_params *p, int jmax, int kmax, int neqs, int npair,
                       const int *wake, const double *q, double *dq, double *rhs,
                       const double *btc, const double *bjm, const double *bjp,
                       const double *bkm, const double *bkp, double *scr1);

/* ---- Allocation-free LU-SGS ------------------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory: code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc.  The solve of
 * smac_lusgs (the same kernels, bit-identical results) on arrays that are already in device
 * memory, enqueued on `stream` (a hipStream_t; NULL: the null stream) and not waited for.
 * It calls no hipMalloc/hipFree: its workspace (the sheared planes and the band hand-over
 * slots) lives in one static device array of SMAC_NOALLOC_SCRATCH_BYTES on each GPU it runs
 * on, held until the process ends.  A grid whose workspace does not fit is rejected with
 * SMAC_ERR_SIZE (neqs = 3: up to about 477x121 points).
 *  - q, dq, rhs, btc, bjm, bjp, bkm, bkp, scr1: device arrays, the layouts of smac_lusgs.
 *  - wake: npair (jt, jb) pairs, 1-based, in device memory (a pair outside 1..jmax is skipped).
 *  - stall: NULL, or a device int set to 0 by the call and to 1 if a band hand-over timed out.
 *  - As for smac_lusgs, bjm(1,k) and bjp(jmax,k) must be zero (the C-grid rows); device
 *    arrays are not inspected on the host, so this is the caller's to ensure.
 * Every call on a GPU shares that GPU's workspace: calls that overlap on one GPU race, so
 * issue them on one stream (or order them with events). */
#define SMAC_NOALLOC_SCRATCH_BYTES (64ull << 20)
smac_status smac_lusgs_noalloc(const smac_params *p, int jmax, int kmax, int neqs, int npair,
                               const int *wake, const double *q, double *dq, double *rhs,
                               const double *btc, const double *bjm, const double *bjp,
                               const double *bkm, const double *bkp, double *scr1, int *stall,
                               void *stream);

/* Text of the last error on this thread ("" if none). */
const char *smac_last_error(void);
/* Library version string. */
const char *smac_version(void);

#ifdef __cplusplus
}
#endif
#endif
