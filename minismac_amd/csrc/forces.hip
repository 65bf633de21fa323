// Solution update and wall forces: update (ref/step.f:760-856), force/fmk
// (ref/geom.f:135-213, ref/fm.f:156-324) and the force coefficients of iterout with the
// circulation fed back to the far field (ref/smac2d.f:1709-1741, 1825).
//
// The force sums are sequential in the reference (fxp = fxp - fx face after face), and CL
// enters the next pass's far-field rows, so they are kept in the reference order: the
// per-face terms are formed in parallel into LDS, then five lanes run the five running sums
// (fxp, fyp, fxs, fys, fmom) in the reference's order, side by side.
#include "smac_dev.h"
#include "smac_kernels.h"

namespace smac {

SMAC_DEV void b_update(const Blk &bk, long n3, double *__restrict__ q,
                       const double *__restrict__ dq, int *__restrict__ flag)
{
    int bad = 0;
    for (long i = (long)bk.x * blockDim.x + threadIdx.x; i < n3;
         i += (long)bk.nx * blockDim.x) {
        double v = q[i] + dq[i];
        q[i] = v;
        bad |= !(fabs(v) <= 1.0e300);
    }
    if (bad) atomicOr(flag, 1);
}

__global__ void k_update(long n3, double *__restrict__ q, const double *__restrict__ dq,
                         int *__restrict__ flag)
{
    b_update(hw_blk(), n3, q, dq, flag);
}

constexpr int FORCE_THREADS = 256;

SMAC_DEV void b_force(const Dims &d, const DevParams &prm, const Walls &wl,
                      const double *__restrict__ x, const double *__restrict__ y,
                      const double *__restrict__ met, const double *__restrict__ dj,
                      const double *__restrict__ q, int *__restrict__ flag, DevStats *st)
{
    __shared__ double t_fx[FORCE_THREADS], t_fy[FORCE_THREADS], t_m1[FORCE_THREADS],
        t_m2[FORCE_THREADS], t_zero[FORCE_THREADS];
    const long N = d.N;
    const double *r22 = met + 2 * N, *r23 = met + 3 * N;
    const double x0 = prm.xmom, y0 = prm.ymom, vnu = prm.vnu;
    double fxp = 0.0, fyp = 0.0, fxs = 0.0, fys = 0.0, fmom = 0.0;   // lanes 0..4 own one each
    const int lane = threadIdx.x;
    t_zero[lane] = 0.0;
    for (int w = 0; w < wl.n; ++w) {
        const int k = wl.k[w], jbeg = wl.j1[w], jend = wl.j2[w];
        fxp = 0.; fyp = 0.; fxs = 0.; fys = 0.;   // fmk starts its sums afresh (ref/fm.f:196-199)
        for (int pass = 0; pass < (prm.ivis != 0 ? 2 : 1); ++pass) {
            for (int j0 = jbeg; j0 <= jend - 1; j0 += FORCE_THREADS) {
                const int j = j0 + lane;
                if (j <= jend - 1) {
                    const long p = P(d, j, k), p1 = p + 1;
                    double da1 = 0.5 * (r22[p1] + r22[p]);
                    double da2 = 0.5 * (r23[p1] + r23[p]);
                    double xave = 0.5 * (x[p1] + x[p]) - x0;
                    double yave = 0.5 * (y[p1] + y[p]) - y0;
                    if (pass == 0) {
                        // pressure (ref/fm.f:210-233)
                        double pave = 0.5 * (q[p1] + q[p]);
                        double fx = pave * da1, fy = pave * da2;
                        t_fx[lane] = fx;
                        t_fy[lane] = fy;
                        t_m1[lane] = xave * fy - yave * fx;
                    } else {
                        // shear from one-sided eta derivatives (ref/fm.f:277-309)
                        double tt[2][3];
                        for (int a = 0; a < 2; ++a) {
                            const long pa = p + a, s = d.jm;
                            double ueta = 0.5 * (-3. * q[N + pa] + 4. * q[N + pa + s] - q[N + pa + 2 * s]);
                            double veta = 0.5 * (-3. * q[2 * N + pa] + 4. * q[2 * N + pa + s] -
                                                 q[2 * N + pa + 2 * s]);
                            tt[a][0] = 2. * vnu * r22[pa] * dj[pa] * ueta;
                            tt[a][1] = 2. * vnu * r23[pa] * dj[pa] * veta;
                            tt[a][2] = vnu * (r23[pa] * ueta + r22[pa] * veta) * dj[pa];
                        }
                        double t11ave = 0.5 * (tt[1][0] + tt[0][0]);
                        double t22ave = 0.5 * (tt[1][1] + tt[0][1]);
                        double t12ave = 0.5 * (tt[1][2] + tt[0][2]);
                        double t21ave = t12ave;
                        double fx = t11ave * da1 + t12ave * da2;
                        double fy = t21ave * da1 + t22ave * da2;
                        t_fx[lane] = fx;
                        t_fy[lane] = fy;
                        t_m1[lane] = xave * fy;
                        t_m2[lane] = yave * fx;
                    }
                }
                __syncthreads();
                const int cnt = (jend - 1) - j0 + 1 < FORCE_THREADS ? (jend - 1) - j0 + 1 : FORCE_THREADS;
                // lanes 0..4 run their sums with one instruction stream: lane-selected
                // operands, acc + (-1)*a == acc - a and acc - (+0) == acc exactly
                if (lane < 5) {
                    const bool act = pass == 0 ? (lane == 0 || lane == 1 || lane == 4) : lane >= 2;
                    const double *A = (lane == 0 || lane == 2) ? t_fx
                                      : (lane == 1 || lane == 3) ? t_fy : t_m1;
                    const double *B = (pass == 1 && lane == 4) ? t_m2 : t_zero;
                    const double sgn = pass == 0 ? -1.0 : 1.0;
                    double acc = lane == 0 ? fxp : lane == 1 ? fyp : lane == 2 ? fxs
                               : lane == 3 ? fys : fmom;
                    if (act) {
                        if (pass == 0) {
                            // (acc + (-1)*a) - (+0) == acc - a exactly: one dependent op a face
#pragma unroll 8
                            for (int i = 0; i < cnt; ++i) acc = acc - A[i];
                        } else {
#pragma unroll 8
                            for (int i = 0; i < cnt; ++i) acc = (acc + sgn * A[i]) - B[i];
                        }
                    }
                    if (lane == 0) fxp = acc;
                    else if (lane == 1) fyp = acc;
                    else if (lane == 2) fxs = acc;
                    else if (lane == 3) fys = acc;
                    else fmom = acc;
                }
                __syncthreads();
            }
        }
    }
    __shared__ double res[5];
    if (lane < 5) res[lane] = lane == 0 ? fxp : lane == 1 ? fyp : lane == 2 ? fxs : lane == 3 ? fys : fmom;
    __syncthreads();
    if (lane == 0) {
        fxp = res[0]; fyp = res[1]; fxs = res[2]; fys = res[3]; fmom = res[4];
        double pi = 4.0 * atan(1.0);
        double ca = cos(prm.alpha * pi / 180.0), sa = sin(prm.alpha * pi / 180.0);
        double chord = prm.chord;
        double cd = (2. * (fxp + fxs) * ca + 2. * (fyp + fys) * sa) / chord;
        double cl = (2. * (fyp + fys) * ca - 2. * (fxp + fxs) * sa) / chord;
        st->cd = cd;
        st->cl = cl;
        st->cm = -(2. * fmom / (chord * chord));
        st->cdp = (2. * fxp * ca + 2. * fyp * sa) / chord;
        st->clp = (2. * fyp * ca - 2. * fxp * sa) / chord;
        st->fxp = fxp; st->fyp = fyp; st->fxs = fxs; st->fys = fys; st->fmom = fmom;
        st->clift = cl;
        if (*flag) st->nonfinite = 1;
        *flag = 0;
    }
}

__global__ void __launch_bounds__(FORCE_THREADS)
k_force(Dims d, DevParams prm, Walls wl, const double *__restrict__ x, const double *__restrict__ y,
        const double *__restrict__ met, const double *__restrict__ dj, const double *__restrict__ q,
        int *__restrict__ flag, DevStats *st)
{
    b_force(d, prm, wl, x, y, met, dj, q, flag, st);
}

void launch_update(hipStream_t sm, const Dims &d, double *q, const double *dq, Scratch &w)
{
    int nb = (int)((3 * d.N + 255) / 256);
    nb = nb > 4096 ? 4096 : nb;
    hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, sm, 3 * d.N, q, dq, w.flag);
}

void launch_force(hipStream_t sm, const Dims &d, const DevParams &prm, const Walls &wl,
                  const double *x, const double *y, const double *met, const double *dj,
                  const double *q, Scratch &w, DevStats *st)
{
    hipLaunchKernelGGL(k_force, dim3(1), dim3(FORCE_THREADS), 0, sm, d, prm, wl, x, y, met, dj, q,
                       w.flag, st);
}

void launch_update_force(hipStream_t sm, const Dims &d, const DevParams &prm, const Walls &wl,
                         const double *x, const double *y, const double *met, const double *dj,
                         double *q, const double *dq, Scratch &w, DevStats *st)
{
    launch_update(sm, d, q, dq, w);
    launch_force(sm, d, prm, wl, x, y, met, dj, q, w, st);
}

// Copies a pass's statistics into pinned host memory, then writes `seq` there with a
// system-scope release: a host thread that reads `seq` == val (acquire) then reads this pass's
// statistics, with no copy engine or stream synchronisation in between.
static constexpr int PUB_WORDS = (int)(sizeof(DevStats) / 4);
static_assert(sizeof(DevStats) % 4 == 0 && PUB_WORDS <= 64, "one wavefront copies the statistics");

__global__ void __launch_bounds__(64)
k_publish(const DevStats *__restrict__ st, DevStats *host, unsigned *seq, unsigned val)
{
    const int t = (int)threadIdx.x;
    if (t < PUB_WORDS)
        __hip_atomic_store((unsigned *)host + t, ((const unsigned *)st)[t], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
    __syncthreads();
    if (t == 0) __hip_atomic_store(seq, val, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void launch_publish(hipStream_t sm, const DevStats *st, DevStats *host, unsigned *seq, unsigned val)
{
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, sm, st, host, seq, val);
}

// ---- multi-case launches ------------------------------------------------------------------

UpdArgs upd_args(const Dims &d, double *q, const double *dq, Scratch &w)
{
    int nb = (int)((3 * d.N + 255) / 256);
    nb = nb > 4096 ? 4096 : nb;
    return UpdArgs{(unsigned)nb, 1u, 3 * d.N, q, dq, w.flag};
}

ForceArgs force_args(const Dims &d, const DevParams &prm, const Walls &wl, const double *x,
                     const double *y, const double *met, const double *dj, const double *q,
                     Scratch &w, DevStats *st)
{
    return ForceArgs{1u, 1u, d, prm, wl, x, y, met, dj, q, w.flag, st};
}

struct CUpdate {
    using Args = UpdArgs;
    static constexpr int THREADS = 256;
    static SMAC_DEV void run(const Blk &b, const Args &a) { b_update(b, a.n3, a.q, a.dq, a.flag); }
};
struct CForce {
    using Args = ForceArgs;
    static constexpr int THREADS = FORCE_THREADS;
    static SMAC_DEV void run(const Blk &, const Args &a)
    {
        b_force(a.d, a.prm, a.wl, a.x, a.y, a.met, a.dj, a.q, a.flag, a.st);
    }
};

// Device-side stop of a batch: after each pass, every case still running counts the pass
// and stops once its residual norm has dropped to `drop` times its reference norm (the norm
// of its first pass, as smac_run's stop criterion), or once a pass of it went non-finite or
// stalled.  A stopped case's items in later passes return at once (k_multi, the multi-case
// band sweeps), so its state stays that of the pass that met the criterion.
__global__ void k_batch_gate(int n, DevStats *const *__restrict__ st, double *__restrict__ first,
                             double drop, int *__restrict__ active, int *__restrict__ passes)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n || active[i] == 0) return;
    passes[i] += 1;
    const DevStats *s = st[i];
    const double r = s->rhsnorm;
    if (!(first[i] > 0.0)) first[i] = r;
    if (s->nonfinite || s->stall || (drop > 0.0 && r <= drop * first[i])) active[i] = 0;
}

void launch_batch_gate(hipStream_t sm, int n, DevStats *const *st, double *first, double drop,
                       int *active, int *passes)
{
    hipLaunchKernelGGL(k_batch_gate, dim3((n + 63) / 64), dim3(64), 0, sm, n, st, first, drop,
                       active, passes);
}

void multi_update(hipStream_t sm, const MultiTab<UpdArgs> &t)
{
    if (t.n > 0 && t.nblk > 0)
        hipLaunchKernelGGL(k_multi<CUpdate>, dim3(t.nblk), dim3(256), 0, sm, t.args, t.start, t.n,
                           t.item_case, t.active);
}
void multi_force(hipStream_t sm, const MultiTab<ForceArgs> &t)
{
    if (t.n > 0 && t.nblk > 0)
        hipLaunchKernelGGL(k_multi<CForce>, dim3(t.nblk), dim3(FORCE_THREADS), 0, sm, t.args,
                           t.start, t.n, t.item_case, t.active);
}

}  // namespace smac
