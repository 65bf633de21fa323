// Host-side launchers of the gfx950 kernels (one translation unit per kernel family).
#pragma once

#include "smac_dev.h"

#include <vector>

namespace smac {

struct BcRegion {
    int ibcval, jbeg, jend, kbeg, kend;   // 0-based, resolved
};

// device scratch owned by a handle
struct Scratch {
    int nred;            // blocks of the residual reduction
    double *red_v;       // 2*nred
    long *red_i;         // 2*nred
    double *red_s;       // nred
    double *sa_v;        // nred: SA max(ds) partials
    int *flag;           // non-finite flag
};

struct WakePairs {
    int n;
    const int *pairs;   // device, 2n ints (jt, jb)
    int base = 0;       // index base of `pairs` (0, or 1 for caller-owned Fortran indices)
};

struct Walls {
    int n;              // k-line walls only
    int k[8], j1[8], j2[8];
};

// assemble.hip
template <typename LT>
void launch_assemble(hipStream_t, const Dims &, const DevParams &, const double *q,
                     const double *met, const double *dj, double *s, LT *btc, LT *bjm, LT *bjp,
                     LT *bkm, LT *bkp);
template <typename LT>
void launch_bc(hipStream_t, const Dims &, const DevParams &, const BcRegion *regs, int nreg,
               const DevStats *, const double *x, const double *y, const double *met,
               const double *q, double *s, LT *btc, LT *bjm, LT *bjp, LT *bkm, LT *bkp);
void launch_resid(hipStream_t, const Dims &, int nt, double beta, const double *s,
                  const double *dj, Scratch &, DevStats *);

// linerelax.hip: lrelax in the reference's line order (ref/lrelax.f:51-221)
template <typename LT>
void launch_lrelax_ref(hipStream_t, const Dims &, const DevParams &, int njsp, int nksp,
                       int nbcimp, const WakePairs &, const double *q, double *dq, double *s,
                       LT *btc, LT *bjm, LT *bjp, LT *bkm, LT *bkp, int lines_per_block = 64);
void launch_wake(hipStream_t, const Dims &, int neqs, const WakePairs &, const double *sq,
                 double *ds, double *rhs);

// linescan.hip: line relaxation with each line's solve parallel inside the line (prefix scans
// of the substitutions' affine maps), in the reference's line order or the zebra order
struct LrScan {
    void *opx, *ope;     // 45 line-major operand planes (LT) of the xi and of the eta lines
    double *s_t, *dq_t;  // eta-major s and dq (3 planes each)
    double *sa_op;       // 5 planes of the SA's scalar line system
    bool zebra;          // the SA's lines in the zebra order too
};
long lr_scan_bytes(const Dims &);
// the line factorisations of both directions (either may be skipped) into opx / ope,
// twisted (from both ends) or plain
template <typename LT>
void launch_lr_factor3(hipStream_t, const Dims &, bool xi, bool eta, bool twist, const LT *btc,
                       const LT *bjm, const LT *bjp, const LT *bkm, const LT *bkp, LT *opx,
                       LT *ope);
// factor = false: the operand planes in the workspace are already those of these blocks.
// twist: 0 the plain factorisation, else (default) the twisted one for the zebra order
template <typename LT>
void launch_lrelax_scan(hipStream_t, const Dims &, const DevParams &, int njsp, int nksp,
                        int nbcimp, bool zebra, const WakePairs &, const double *q, double *dq,
                        double *s, const LT *btc, const LT *bjm, const LT *bjp, const LT *bkm,
                        const LT *bkp, const LrScan &, bool factor = true, int twist = 1);
// out = L^T in, L the zebra sweeps (no wake injection) of the blocks factored in wt as their
// transposed and shifted system; v: work vectors of 3 planes each
struct LrTWork {
    double *a, *z, *at, *st, *zt;
};
template <typename LT>
void launch_lrelax_scan_t(hipStream_t, const Dims &, int njsp, int nksp, double underr,
                          const double *in, double *out, const LrScan &wt, const LrTWork &v);

// lusgs.hip: LU-SGS with hyperplane (j+k = const) sweeps (ref/lusgs.f, ref/ilusol.f).
// Workspace in the sheared layout: point (j,k) at (j+k+16)*ks + k of every plane, planes
// stored in pairs for the 3x3 system.
struct LusgsWs {
    int rows;                    // rows per band (one wavefront each), 1..64
    int nband;                   // bands covering k = 0..km-1
    int ks;                      // pitch of one hyperplane
    int H;                       // last hyperplane, jm + km - 2
    long ns;                     // positions per plane, (H+33)*ks
    int epitch;                  // slots per band region: (jm + 2*128 padding)*4
    long nedge;                  // (2*nband + 1)*epitch slots
    void *lhs;                   // 5 block groups (factors, bjm, bkm, bjp, bkp), LT
    void *r;                     // neqs planes: rhs, forward result, middle result
    void *dq;                    // neqs planes (the solve's type: FP64, or FP32 for FP32-LHS)
    unsigned long long *edge;    // band hand-over slots (forward half, backward half)
    int *stall;                  // raised when a hand-over wait times out
    int pf3, pf1;                // prefetch distance (steps) of the 3x3 and scalar sweeps
    int ring;                    // 1: the 3x3 sweeps stream their operands through LDS
    hipEvent_t ev_fwd[2];        // when set: recorded around the forward band sweep
};
// one case's part of a multi-case band launch
struct BandJob {
    Dims d;
    LusgsWs w;
    int bcase;   // the case's index in its batch (selects its entry of the batch's active flags)
};
// the solve in phases, for callers that put the band sweeps of several cases in one launch:
// pre (transpose and factor), band (forward / backward), mid, post (back to the natural
// layout, wake rows); launch_lusgs3 / launch_lusgs1 run all of them for one case
template <typename LT, int NEQ>
void lusgs_pre(hipStream_t, const Dims &, const LusgsWs &, const double *rhs, const LT *btc,
               const LT *bjm, const LT *bjp, const LT *bkm, const LT *bkp, LT *fac);
template <typename LT, int NEQ>
void lusgs_band(hipStream_t, const Dims &, const LusgsWs &, bool fwd);
// jobs and map in device memory; map[i] = (job, band in dispatch order), nblk entries
template <typename LT, int NEQ>
void lusgs_band_multi(hipStream_t, const BandJob *jobs, const int2 *map, int nblk, int pf,
                      bool fwd, const int *active);
template <typename LT, int NEQ>
void lusgs_mid(hipStream_t, const Dims &, const LusgsWs &);
template <typename LT, int NEQ>
void lusgs_post(hipStream_t, const Dims &, const LusgsWs &, const WakePairs &, const double *q,
                double *dq, double *rhs, bool rhs_out);
void lusgs_ws_shape(int jm, int km, int rows, LusgsWs &);
long lusgs_ws_lhs_bytes(const LusgsWs &, int neqs);   // the five block groups
long lusgs_ws_vec_bytes(const LusgsWs &, int neqs);   // one vector (rhs or dq)
bool lusgs_ws_fits(const LusgsWs &);                  // 32-bit buffer offsets suffice
// fac (nullable) receives the diagonal factors in the natural layout (the reference's scr1);
// rhs_out also returns the middle-step rhs in the natural layout, as the reference leaves it
template <typename LT>
void launch_lusgs3(hipStream_t, const Dims &, const LusgsWs &, const WakePairs &, const double *q,
                   double *dq, double *rhs, const LT *btc, const LT *bjm, const LT *bjp,
                   const LT *bkm, const LT *bkp, LT *fac, bool rhs_out);
void launch_lusgs1(hipStream_t, const Dims &, const LusgsWs &, const WakePairs &, const double *q,
                   double *dq, double *rhs, const double *btc, const double *bjm,
                   const double *bjp, const double *bkm, const double *bkp, double *fac,
                   bool rhs_out);

// sa.hip: Spalart-Allmaras (ref/tursa.f, ref/lreltur.f)
struct SaArrays {
    double *anut, *vnut, *smin, *sjm, *sjp, *skm, *skp, *stc, *sf, *ds, *vort, *cu, *ck, *cv,
        *fr, *bprm;
};
void launch_sa_first(hipStream_t, const Dims &, const Walls &, const double *x, const double *y,
                     SaArrays &);
void launch_sa_assemble(hipStream_t, const Dims &, const DevParams &, int first_pass,
                        const double *q, const double *met, const double *dj, int nreg,
                        const BcRegion *regs, SaArrays &);
void launch_sa_solve(hipStream_t, const Dims &, const DevParams &, int ntjsp, const WakePairs &,
                     const LusgsWs *, SaArrays &, const LrScan *scan = nullptr);
void launch_sa_lreltur_scan(hipStream_t, const Dims &, int ntjsp, const WakePairs &, SaArrays &,
                            const LrScan &);
void launch_sa_update(hipStream_t, const Dims &, const DevParams &, Scratch &, DevStats *,
                      SaArrays &);

// update + wall forces (ref/step.f:760-856; ref/geom.f:135-213, ref/fm.f:156-324)
void launch_update_force(hipStream_t, const Dims &, const DevParams &, const Walls &,
                         const double *x, const double *y, const double *met, const double *dj,
                         double *q, const double *dq, Scratch &, DevStats *);
// the two halves of launch_update_force: q += dq (raising the non-finite flag), then the
// wall forces, CL/CD and clift (clearing the flag into st->nonfinite)
void launch_update(hipStream_t, const Dims &, double *q, const double *dq, Scratch &);
void launch_batch_gate(hipStream_t, int n, DevStats *const *st, double *first, double drop,
                       int *active, int *passes);
void launch_force(hipStream_t, const Dims &, const DevParams &, const Walls &, const double *x,
                  const double *y, const double *met, const double *dj, const double *q,
                  Scratch &, DevStats *);
// st -> *host (pinned, device-mapped), then *seq = val with a system-scope release
void launch_publish(hipStream_t, const DevStats *st, DevStats *host, unsigned *seq, unsigned val);

// adjoint.hip: discrete adjoint of the converged mean flow (CL, CD gradients)
struct AdjCtx {
    Dims d;
    hipStream_t sm;
    const double *x, *y, *met, *dj, *q;
    double *blk[5];              // FP64 blocks of the state (the primal's LU-SGS operator M)
    LusgsWs ws;                  // LU-SGS workspace for M^-T and M^-1
    Walls walls;
    double beta, cdis, vnu, pout, alpha, chord, clift, cl, cd;
    int ivis;
    int m;                       // GMRES restart length
    double reorth;               // second Gram-Schmidt pass when less than this fraction of
                                 // the vector's square norm survives the first (0: never)
    int psteps;                  // preconditioner: this many steps of the transposed update
                                 // (with the operator between them)
    double *pk1, *pk2;           // its work vectors (3 planes each)
    const signed char *rtype;    // row type per point
    int npair;
    const int *pairs, *inv;      // wake pairs (0-based) and, per pair, the pair whose partner
                                 // is its target
    int lr;                      // preconditioner: 1 the zebra line relaxation, 0 LU-SGS
    int njsp, nksp;
    DevParams dp;
    LrScan lr_f, lr_t;           // factored lines of M and of M^T (lr = 1)
    LrTWork lrv;
    double *jac, *tb[5], *lam[2], *g[2], *fc, *fa, *V, *w, *z, *t1, *t2, *t1b;
    double *smu, sm_binv;        // the preconditioner's rank-one correction (see adj_build)
    int sm_on;
    double *dmet, *ddj, *dxy, *fb, *fpart;
    double *part, *red, *red2;
    int *gate;                   // second Gram-Schmidt pass wanted (device flag)
};
void adj_build(AdjCtx &);
void adj_apply_op(AdjCtx &, const double *in, double *out);
void adj_apply_prec_t(AdjCtx &, const double *in, double *out);
void adj_apply_prec(AdjCtx &, const double *in, double *out);
void adj_apply_jt(AdjCtx &, const double *in, double *out);
int adj_gmres(AdjCtx &, const double *b, double *x, int maxit, double tol, double *rel);
void adj_gradients(AdjCtx &, int nshape, const double *dx, const double *dy, double *dcl,
                   double *dcd);

// ---- multi-case launches (the batch pass): per-item arguments, one table per kernel ------
// A table lives on the device: args[n] and start[n + 1] (see k_multi).
template <typename A>
struct MultiTab {
    A *args = nullptr;
    int *start = nullptr;
    int *item_case = nullptr;     // device: the batch case of each item
    const int *active = nullptr;  // device: the batch's per-case flags; 0 = item skipped
    int n = 0, nblk = 0;
};

constexpr int BC_THREADS = 256;
constexpr int BC_MAX = 16;
struct BcTable {
    int n;
    BcRegion r[BC_MAX];
};

struct AsmArgs {      // k_assemble
    unsigned gx, gy;
    Dims d;
    DevParams prm;
    const double *q, *met, *dj;
    double *s;
    void *btc, *bjm, *bjp, *bkm, *bkp;   // LT
};
struct BcArgs {       // k_bc
    unsigned gx, gy;
    Dims d;
    BcTable t;
    double pout, alpha, beta;
    const DevStats *st;
    const double *x, *y, *met, *q;
    double *s;
    void *btc, *bjm, *bjp, *bkm, *bkp;   // LT
};
struct ResidPArgs {   // k_resid_partial
    unsigned gx, gy;
    Dims d;
    const double *s, *dj;
    double *part_v;
    long *part_i;
    double *part_s;
};
struct ResidFArgs {   // k_resid_final
    unsigned gx, gy;
    Dims d;
    int nparts, nt;
    double beta;
    const double *s, *dj, *part_v;
    const long *part_i;
    const double *part_s;
    DevStats *st;
};
struct SaDerivArgs {  // k_sa_deriv
    unsigned gx, gy;
    Dims d;
    const double *q, *met, *dj;
    double *vort, *cu, *ck, *cv, *fr;
};
struct SaAsmArgs {    // k_sa_assemble
    unsigned gx, gy;
    Dims d;
    DevParams prm;
    const double *q, *met, *dj, *anut, *smin;
    int use_smin;
    const double *vort, *cu, *ck, *cv, *fr;
    double *sjm, *sjp, *skm, *skp, *stc, *sf;
};
struct SaBcArgs {     // k_sa_bc (one item per outflow region)
    unsigned gx, gy;
    Dims d;
    BcRegion r;
    const double *anut;
    double *sjm, *sjp, *skm, *skp, *stc, *sf;
};
struct SaUpdArgs {    // k_sa_update
    unsigned gx, gy;
    long N;
    double vnu;
    double *anut;
    const double *ds;
    double *vnut, *part;
};
struct SaTurArgs {    // k_sa_turres
    unsigned gx, gy;
    int n;
    const double *part;
    DevStats *st;
};
struct UpdArgs {      // k_update
    unsigned gx, gy;
    long n3;
    double *q;
    const double *dq;
    int *flag;
};
struct ForceArgs {    // k_force
    unsigned gx, gy;
    Dims d;
    DevParams prm;
    Walls wl;
    const double *x, *y, *met, *dj, *q;
    int *flag;
    DevStats *st;
};
struct WakeArgs {     // k_wake
    unsigned gx, gy;
    long N;
    int jm, neqs, npair, base;
    const int *pairs;
    const double *sq;
    double *ds, *rhs;
};
struct ShearArgs {    // k_shear
    unsigned gx, gy;
    Dims d;
    LusgsWs w;
    const double *rhs;
    const void *btc, *bjm, *bkm, *bjp, *bkp;   // LT
    void *fac;                                 // LT
};
struct MidArgs {      // k_mid
    unsigned gx, gy;
    Dims d;
    LusgsWs w;
};
struct UnshearArgs {  // k_unshear
    unsigned gx, gy;
    Dims d;
    LusgsWs w;
    const void *src;  // the solve's type
    double *dst;
};

// the arguments of one case's launches, as the single-case launchers form them
template <typename LT>
AsmArgs asm_args(const Dims &, const DevParams &, const double *q, const double *met,
                 const double *dj, double *s, LT *btc, LT *bjm, LT *bjp, LT *bkm, LT *bkp);
ResidPArgs residp_args(const Dims &, const double *s, const double *dj, Scratch &);
ResidFArgs residf_args(const Dims &, int nt, double beta, const double *s, const double *dj,
                       Scratch &, DevStats *);
template <typename LT>
ShearArgs shear_args(const Dims &, const LusgsWs &, const double *rhs, const LT *btc,
                     const LT *bjm, const LT *bjp, const LT *bkm, const LT *bkp, LT *fac);
template <typename LT, int NEQ>
MidArgs mid_args(const Dims &, const LusgsWs &);   // gx = 0: no launch (scaled FP32 form)
UnshearArgs unshear_args(const Dims &, const LusgsWs &, const void *src, double *dst);
WakeArgs wake_args(const Dims &, int neqs, const WakePairs &, const double *sq, double *ds,
                   double *rhs);          // gx = 0 when there are no pairs
UpdArgs upd_args(const Dims &, double *q, const double *dq, Scratch &);
ForceArgs force_args(const Dims &, const DevParams &, const Walls &, const double *x,
                     const double *y, const double *met, const double *dj, const double *q,
                     Scratch &, DevStats *);
SaDerivArgs sa_deriv_args(const Dims &, const double *q, const double *met, const double *dj,
                          SaArrays &);
SaAsmArgs sa_asm_args(const Dims &, const DevParams &, int first_pass, const double *q,
                      const double *met, const double *dj, SaArrays &);
// the outflow rows of the SA: one item per region (appended to out)
void sa_bc_args(const Dims &, int nreg, const BcRegion *regs, SaArrays &,
                std::vector<SaBcArgs> &out);
SaUpdArgs sa_upd_args(const Dims &, const DevParams &, Scratch &, SaArrays &);
SaTurArgs sa_tur_args(Scratch &, DevStats *);
// the BC rows of a case: one item per launch the single-case launcher would make
template <typename LT>
void bc_args(const Dims &, const DevParams &, const BcRegion *regs, int nreg, const DevStats *,
             const double *x, const double *y, const double *met, const double *q, double *s,
             LT *btc, LT *bjm, LT *bjp, LT *bkm, LT *bkp, std::vector<BcArgs> &out);

// one launch over all items of a table
template <typename LT> void multi_assemble(hipStream_t, const MultiTab<AsmArgs> &);
template <typename LT> void multi_bc(hipStream_t, const MultiTab<BcArgs> &);
void multi_resid_partial(hipStream_t, const MultiTab<ResidPArgs> &);
void multi_resid_final(hipStream_t, const MultiTab<ResidFArgs> &);
template <typename LT, int NEQ> void multi_shear(hipStream_t, const MultiTab<ShearArgs> &);
template <typename LT, int NEQ> void multi_mid(hipStream_t, const MultiTab<MidArgs> &);
template <typename LT, int NEQ> void multi_unshear(hipStream_t, const MultiTab<UnshearArgs> &);
void multi_wake(hipStream_t, const MultiTab<WakeArgs> &);
void multi_update(hipStream_t, const MultiTab<UpdArgs> &);
void multi_force(hipStream_t, const MultiTab<ForceArgs> &);
void multi_sa_deriv(hipStream_t, const MultiTab<SaDerivArgs> &);
void multi_sa_assemble(hipStream_t, const MultiTab<SaAsmArgs> &);
void multi_sa_bc(hipStream_t, const MultiTab<SaBcArgs> &);
void multi_sa_update(hipStream_t, const MultiTab<SaUpdArgs> &);
void multi_sa_turres(hipStream_t, const MultiTab<SaTurArgs> &);


}  // namespace smac
