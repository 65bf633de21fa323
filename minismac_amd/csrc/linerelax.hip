// Block line relaxation of the implicit iteration: lrelax / ljsweep / lksweep
// (ref/lrelax.f:51-561), the 3x3 block-tridiagonal Crout factor and its undo (btlu3j,
// btlu3k, btund3j; ref/btlu3.f:52-838) and the wake-cut injection after every sweep
// (bcimpds single-rank branch, ref/bc.f:571-591).
//
// Reference order ("parity mode"): the xi sweep visits k-lines in Gauss-Seidel order
// (line k reads dq of line k-1 from the same sweep), so lines are strictly sequential.  One
// workgroup walks the lines: its threads build the line's right-hand side in LDS in
// parallel, one lane runs the block forward/back substitution with the factored blocks,
// and the threads apply the under-relaxed update.  The factorisations are independent per
// line: one thread per line, eta lines coalesced along j.
#include "smac_dev.h"
#include "smac_kernels.h"

namespace smac {

// 3x3 helpers on registers, entry (m,n) at index m + 3n ------------------------------------

template <typename T>
SMAC_DEV void lu3(T b[9])
{
    // ref/btlu3.f:148-155
    b[0] = (T)1. / b[0];
    b[3] = b[3] * b[0];
    b[6] = b[6] * b[0];
    b[4] = (T)1. / (b[4] - b[1] * b[3]);
    b[7] = (b[7] - b[1] * b[6]) * b[4];
    b[5] = b[5] - b[2] * b[3];
    b[8] = (T)1. / (b[8] - b[2] * b[6] - b[7] * b[5]);
}

template <typename T>
SMAC_DEV void csolve3(const T b[9], T c[9])
{
    // c <- b^-1 c, column by column (ref/btlu3.f:157-179)
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        T d1 = b[0] * c[3 * n];
        T d2 = b[4] * (c[1 + 3 * n] - b[1] * d1);
        T d3 = b[8] * (c[2 + 3 * n] - b[2] * d1 - b[5] * d2);
        c[2 + 3 * n] = d3;
        c[1 + 3 * n] = d2 - b[7] * c[2 + 3 * n];
        c[3 * n] = d1 - b[6] * c[2 + 3 * n] - b[3] * c[1 + 3 * n];
    }
}

template <typename T>
SMAC_DEV void bminus3(T b[9], const T a[9], const T c[9])
{
    // b <- b - a c (ref/btlu3.f:245-273)
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int m = 0; m < 3; ++m)
            b[m + 3 * n] = b[m + 3 * n] - a[m] * c[3 * n] - a[m + 3] * c[1 + 3 * n] -
                           a[m + 6] * c[2 + 3 * n];
}

template <typename T>
SMAC_DEV void bplus3(T b[9], const T a[9], const T c[9])
{
    // b <- b + a c (ref/btlu3.f:660-690)
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int m = 0; m < 3; ++m)
            b[m + 3 * n] = b[m + 3 * n] + a[m] * c[3 * n] + a[m + 3] * c[1 + 3 * n] +
                           a[m + 6] * c[2 + 3 * n];
}

template <typename T>
SMAC_DEV void unlu3(T b[9])
{
    // recompose a factored block (ref/btlu3.f:640-650)
    b[8] = (T)1. / b[8] + b[2] * b[6] + b[5] * b[7];
    b[5] = b[5] + b[2] * b[3];
    b[4] = (T)1. / b[4];
    b[7] = b[7] * b[4] + b[1] * b[6];
    b[4] = b[4] + b[1] * b[3];
    b[0] = (T)1. / b[0];
    b[6] = b[6] * b[0];
    b[3] = b[3] * b[0];
}

template <typename T>
SMAC_DEV void ctimes3(const T b[9], T c[9])
{
    // c <- b c (ref/btlu3.f:700-735)
    T t[9];
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int m = 0; m < 3; ++m)
            t[m + 3 * n] = b[m] * c[3 * n] + b[m + 3] * c[1 + 3 * n] + b[m + 6] * c[2 + 3 * n];
#pragma unroll
    for (int e = 0; e < 9; ++e) c[e] = t[e];
}

template <typename LT>
SMAC_DEV void ld9(const LT *__restrict__ a, long N, long p, LT r[9])
{
#pragma unroll
    for (int e = 0; e < 9; ++e) r[e] = a[e * N + p];
}
template <typename LT>
SMAC_DEV void st9(LT *__restrict__ a, long N, long p, const LT r[9])
{
#pragma unroll
    for (int e = 0; e < 9; ++e) a[e * N + p] = r[e];
}

// ------------------------------------------------------------------------------------------
// factorisation along lines: btlu3j (lines of constant k, walking j) and btlu3k.
// `stride` is the point step along the line, `lstride` the step between lines.

template <typename LT>
__global__ void __launch_bounds__(64) k_btlu3(long N, int nline, int len, long stride,
                                             long lstride, const LT *__restrict__ a,
                                             LT *__restrict__ b, LT *__restrict__ c)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nline) return;
    LT cp[9];
    for (int i = 0; i < len; ++i) {
        const long p = lstride * l + stride * i;
        LT bb[9];
        ld9(b, N, p, bb);
        if (i > 0) {
            LT aa[9];
            ld9(a, N, p, aa);
            bminus3(bb, aa, cp);
        }
        lu3(bb);
        if (i < len - 1) {
            LT cc[9];
            ld9(c, N, p, cc);
            csolve3(bb, cc);
            st9(c, N, p, cc);
#pragma unroll
            for (int e = 0; e < 9; ++e) cp[e] = cc[e];
        }
        st9(b, N, p, bb);
    }
}

// btund3j: recompose btc and bjp from their xi factors (ref/btlu3.f:628-838), walking each
// line backwards so that c(i-1) is still factored when b(i) reads it
template <typename LT>
__global__ void __launch_bounds__(64) k_btund3(long N, int nline, int len, long stride,
                                              long lstride, const LT *__restrict__ a,
                                              LT *__restrict__ b, LT *__restrict__ c)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nline) return;
    for (int i = len - 1; i >= 0; --i) {
        const long p = lstride * l + stride * i;
        LT bb[9];
        ld9(b, N, p, bb);
        unlu3(bb);
        if (i < len - 1) {
            LT cc[9];
            ld9(c, N, p, cc);
            ctimes3(bb, cc);
            st9(c, N, p, cc);
        }
        if (i > 0) {
            LT aa[9], cp[9];
            ld9(a, N, p, aa);
            ld9(c, N, p - stride, cp);
            bplus3(bb, aa, cp);
        }
        st9(b, N, p, bb);
    }
}

// ------------------------------------------------------------------------------------------
// one sweep in reference line order (ljsweep / lksweep).  Lines run one after another in one
// workgroup; `dir` chooses xi lines (lines of constant k) or eta lines (constant j).

constexpr int SWEEP_THREADS = 256;
constexpr int SWEEP_CHUNK = 256;   // points of the band staged per chunk (two chunks in LDS)

template <typename LT>
__global__ void __launch_bounds__(SWEEP_THREADS)
k_sweep_ref(Dims d, int eta, int nsweep, double underr, double *__restrict__ dq,
            const double *__restrict__ s, const LT *__restrict__ btc, const LT *__restrict__ lm,
            const LT *__restrict__ lp, const LT *__restrict__ om, const LT *__restrict__ op)
{
    // lm/lp: blocks along the line (bjm/bjp for xi, bkm/bkp for eta);
    // om/op: blocks to the neighbouring lines (bkm/bkp for xi, bjm/bjp for eta)
    extern __shared__ double lds[];   // 3 * len rhs, then two chunks of 18 band planes
    double *rhs = lds, *band = lds + 3 * (eta ? d.km : d.jm);
    const long N = d.N;
    const int nline = eta ? d.jm : d.km;
    const int len = eta ? d.km : d.jm;
    const long stride = eta ? d.jm : 1;          // along a line
    const long lstride = eta ? 1 : d.jm;         // between lines
    const bool rev = (nsweep / 2) * 2 == nsweep;
    const double half = 0.5;
    for (int ll = 0; ll < nline; ++ll) {
        const int l = rev ? nline - 1 - ll : ll;
        const int lmn = l - 1 > 0 ? l - 1 : 0;
        const int lpl = l + 1 < nline - 1 ? l + 1 : nline - 1;
        const double rkm = (double)(float)(l - lmn - 1);
        const double rkp = (double)(float)(lpl - l - 1);
        const double swm = half + copysign(half, rkm);
        const double swp = half + copysign(half, rkp);
        for (int i = threadIdx.x; i < len; i += SWEEP_THREADS) {
            const long p = lstride * l + stride * i;
            const long pm = lstride * lmn + stride * i, pp = lstride * lpl + stride * i;
            double x1 = dq[pm], x2 = dq[N + pm], x3 = dq[2 * N + pm];
            double y1 = dq[pp], y2 = dq[N + pp], y3 = dq[2 * N + pp];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                double bm = (double)om[m * N + p] * x1 + (double)om[(m + 3) * N + p] * x2 +
                            (double)om[(m + 6) * N + p] * x3;
                double bp = (double)op[m * N + p] * y1 + (double)op[(m + 3) * N + p] * y2 +
                            (double)op[(m + 6) * N + p] * y3;
                rhs[3 * i + m] = s[m * N + p] - bm * swm - bp * swp;
            }
        }
        __syncthreads();
        // The solve along the line is one dependent chain, run by thread 0 from LDS: the
        // band is staged in chunks of SWEEP_CHUNK points (the reference copies it to /btri/,
        // ref/lrelax.f:310-321), the next chunk loaded by waves 1.. while thread 0 works
        // through the current one.
        // forward with the factored diagonal blocks and the lower band (ref/lrelax.f:331-351)
        {
            const int nch = (len + SWEEP_CHUNK - 1) / SWEEP_CHUNK;
            auto stage = [&](int c, double *buf) {
                for (int ii = (int)threadIdx.x - 64; ii < SWEEP_CHUNK; ii += SWEEP_THREADS - 64) {
                    if (ii < 0) continue;
                    const int i = c * SWEEP_CHUNK + ii;
                    if (i >= len) break;
                    const long p = lstride * l + stride * i;
#pragma unroll
                    for (int e = 0; e < 9; ++e) {
                        buf[e * SWEEP_CHUNK + ii] = (double)lm[e * N + p];
                        buf[(9 + e) * SWEEP_CHUNK + ii] = (double)btc[e * N + p];
                    }
                }
            };
            if (threadIdx.x >= 64) stage(0, band);
            __syncthreads();
            double r1 = 0, r2 = 0, r3 = 0;
            for (int c = 0; c < nch; ++c) {
                double *cur = band + (c & 1) * 18 * SWEEP_CHUNK;
                if (threadIdx.x >= 64) {
                    if (c + 1 < nch) stage(c + 1, band + ((c + 1) & 1) * 18 * SWEEP_CHUNK);
                } else if (threadIdx.x == 0) {
                    const int i1 = (c + 1) * SWEEP_CHUNK < len ? (c + 1) * SWEEP_CHUNK : len;
#pragma unroll 4
                    for (int i = c * SWEEP_CHUNK; i < i1; ++i) {
                        const int ii = i - c * SWEEP_CHUNK;
                        double t1 = rhs[3 * i], t2 = rhs[3 * i + 1], t3 = rhs[3 * i + 2];
                        if (i > 0) {
                            const double *a = cur + ii;
                            t1 = t1 - a[0] * r1 - a[3 * SWEEP_CHUNK] * r2 - a[6 * SWEEP_CHUNK] * r3;
                            t2 = t2 - a[1 * SWEEP_CHUNK] * r1 - a[4 * SWEEP_CHUNK] * r2 -
                                 a[7 * SWEEP_CHUNK] * r3;
                            t3 = t3 - a[2 * SWEEP_CHUNK] * r1 - a[5 * SWEEP_CHUNK] * r2 -
                                 a[8 * SWEEP_CHUNK] * r3;
                        }
                        const double *bb = cur + 9 * SWEEP_CHUNK + ii;
                        double b0 = bb[0], b1 = bb[SWEEP_CHUNK], b2 = bb[2 * SWEEP_CHUNK],
                               b3 = bb[3 * SWEEP_CHUNK], b4 = bb[4 * SWEEP_CHUNK],
                               b5 = bb[5 * SWEEP_CHUNK], b6 = bb[6 * SWEEP_CHUNK],
                               b7 = bb[7 * SWEEP_CHUNK], b8 = bb[8 * SWEEP_CHUNK];
                        double d1 = b0 * t1;
                        double d2 = b4 * (t2 - b1 * d1);
                        double d3 = b8 * (t3 - b2 * d1 - b5 * d2);
                        r3 = d3;
                        r2 = d2 - b7 * r3;
                        r1 = d1 - b6 * r3 - b3 * r2;
                        rhs[3 * i] = r1;
                        rhs[3 * i + 1] = r2;
                        rhs[3 * i + 2] = r3;
                    }
                }
                __syncthreads();
            }
            // back substitution with the upper band (ref/lrelax.f:367-373), chunks from the end
            auto stage_c = [&](int c, double *buf) {
                for (int ii = (int)threadIdx.x - 64; ii < SWEEP_CHUNK; ii += SWEEP_THREADS - 64) {
                    if (ii < 0) continue;
                    const int i = c * SWEEP_CHUNK + ii;
                    if (i > len - 2) break;
                    const long p = lstride * l + stride * i;
#pragma unroll
                    for (int e = 0; e < 9; ++e) buf[e * SWEEP_CHUNK + ii] = (double)lp[e * N + p];
                }
            };
            const int nb = len >= 2 ? (len - 2) / SWEEP_CHUNK + 1 : 0;   // chunks holding 0..len-2
            if (nb > 0 && threadIdx.x >= 64) stage_c(nb - 1, band);
            __syncthreads();
            for (int q = 0; q < nb; ++q) {
                const int c = nb - 1 - q;
                double *cur = band + (q & 1) * 18 * SWEEP_CHUNK;
                if (threadIdx.x >= 64) {
                    if (c > 0) stage_c(c - 1, band + ((q + 1) & 1) * 18 * SWEEP_CHUNK);
                } else if (threadIdx.x == 0) {
                    const int ihi = (c + 1) * SWEEP_CHUNK - 1 < len - 2 ? (c + 1) * SWEEP_CHUNK - 1
                                                                      : len - 2;
#pragma unroll 4
                    for (int i = ihi; i >= c * SWEEP_CHUNK; --i) {
                        const double *cc = cur + (i - c * SWEEP_CHUNK);
                        double t1 = rhs[3 * i] - cc[0] * r1 - cc[3 * SWEEP_CHUNK] * r2 -
                                    cc[6 * SWEEP_CHUNK] * r3;
                        double t2 = rhs[3 * i + 1] - cc[1 * SWEEP_CHUNK] * r1 -
                                    cc[4 * SWEEP_CHUNK] * r2 - cc[7 * SWEEP_CHUNK] * r3;
                        double t3 = rhs[3 * i + 2] - cc[2 * SWEEP_CHUNK] * r1 -
                                    cc[5 * SWEEP_CHUNK] * r2 - cc[8 * SWEEP_CHUNK] * r3;
                        r1 = t1; r2 = t2; r3 = t3;
                        rhs[3 * i] = r1;
                        rhs[3 * i + 1] = r2;
                        rhs[3 * i + 2] = r3;
                    }
                }
                __syncthreads();
            }
        }
        for (int i = threadIdx.x; i < len; i += SWEEP_THREADS) {
            const long p = lstride * l + stride * i;
#pragma unroll
            for (int m = 0; m < 3; ++m)
                dq[m * N + p] = dq[m * N + p] + underr * (rhs[3 * i + m] - dq[m * N + p]);
        }
        __syncthreads();
    }
}

// wake-cut injection, bcimpds single-rank branch (ref/bc.f:571-591): k = 0 row of the cut.
// Pairs whose indices fall outside the row are skipped.
SMAC_DEV void b_wake(const Blk &bk, long N, int jm, int neqs, int npair, int base,
                     const int *__restrict__ pairs, const double *__restrict__ sq,
                     double *__restrict__ ds, double *__restrict__ rhs)
{
    const int i = bk.x * blockDim.x + threadIdx.x;
    if (i >= npair) return;
    const long jt = pairs[2 * i] - base, jb = pairs[2 * i + 1] - base;
    if (jt < 0 || jt >= jm || jb < 0 || jb >= jm) return;
    for (int n = 0; n < neqs; ++n) {
        double snew = 0.5 * (sq[n * N + jt + jm] + sq[n * N + jb + jm]);
        double dsnew = 0.5 * (ds[n * N + jt + jm] + ds[n * N + jb + jm]);
        ds[n * N + jt] = snew + dsnew - sq[n * N + jt];
        rhs[n * N + jt] = ds[n * N + jt];
    }
}

__global__ void k_wake(long N, int jm, int neqs, int npair, int base,
                       const int *__restrict__ pairs, const double *__restrict__ sq,
                       double *__restrict__ ds, double *__restrict__ rhs)
{
    b_wake(hw_blk(), N, jm, neqs, npair, base, pairs, sq, ds, rhs);
}

__global__ void k_fill(double *__restrict__ a, long n, double v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        a[i] = v;
}

WakeArgs wake_args(const Dims &d, int neqs, const WakePairs &w, const double *sq, double *ds,
                   double *rhs)
{
    return WakeArgs{(unsigned)(w.n > 0 ? (w.n + 63) / 64 : 0), 1u, d.N, d.jm, neqs, w.n, w.base,
                    w.pairs, sq, ds, rhs};
}

struct CWake {
    using Args = WakeArgs;
    static constexpr int THREADS = 64;
    static SMAC_DEV void run(const Blk &b, const Args &a)
    {
        b_wake(b, a.N, a.jm, a.neqs, a.npair, a.base, a.pairs, a.sq, a.ds, a.rhs);
    }
};

void multi_wake(hipStream_t sm, const MultiTab<WakeArgs> &t)
{
    if (t.n > 0 && t.nblk > 0)
        hipLaunchKernelGGL(k_multi<CWake>, dim3(t.nblk), dim3(64), 0, sm, t.args, t.start, t.n,
                           t.item_case, t.active);
}

void launch_wake(hipStream_t sm, const Dims &d, int neqs, const WakePairs &w, const double *sq,
                 double *ds, double *rhs)
{
    if (w.n <= 0) return;
    hipLaunchKernelGGL(k_wake, dim3((w.n + 63) / 64), dim3(64), 0, sm, d.N, d.jm, neqs, w.n,
                       w.base, w.pairs, sq, ds, rhs);
}

template <typename LT>
void launch_lrelax_ref(hipStream_t sm, const Dims &d, const DevParams &prm, int njsp, int nksp,
                       int nbcimp, const WakePairs &w, const double *q, double *dq, double *s,
                       LT *btc, LT *bjm, LT *bjp, LT *bkm, LT *bkp, int lpb)
{
    const long N = d.N;
    // lines per workgroup of the factor kernels (one lane per line)
    lpb = lpb < 1 ? 1 : (lpb > 64 ? 64 : lpb);
    hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, sm, dq, 3 * N, 0.0);
    if (njsp > 0)
        hipLaunchKernelGGL(k_btlu3<LT>, dim3((d.km + lpb - 1) / lpb), dim3(lpb), 0, sm, N, d.km, d.jm,
                           1L, (long)d.jm, bjm, btc, bjp);
    for (int nj = 1; nj <= njsp; ++nj) {
        hipLaunchKernelGGL(k_sweep_ref<LT>, dim3(1), dim3(SWEEP_THREADS),
                           sizeof(double) * (3 * d.jm + 36 * SWEEP_CHUNK), sm, d, 0, nj, prm.underr, dq, s, btc, bjm,
                           bjp, bkm, bkp);
        if ((nj / nbcimp) * nbcimp == nj) launch_wake(sm, d, 3, w, q, dq, s);
    }
    if (njsp > 0)
        hipLaunchKernelGGL(k_btund3<LT>, dim3((d.km + lpb - 1) / lpb), dim3(lpb), 0, sm, N, d.km, d.jm,
                           1L, (long)d.jm, bjm, btc, bjp);
    if (nksp > 0)
        hipLaunchKernelGGL(k_btlu3<LT>, dim3((d.jm + lpb - 1) / lpb), dim3(lpb), 0, sm, N, d.jm, d.km,
                           (long)d.jm, 1L, bkm, btc, bkp);
    for (int nk = 1; nk <= nksp; ++nk) {
        hipLaunchKernelGGL(k_sweep_ref<LT>, dim3(1), dim3(SWEEP_THREADS),
                           sizeof(double) * (3 * d.km + 36 * SWEEP_CHUNK), sm, d, 1, nk, prm.underr, dq, s, btc, bkm,
                           bkp, bjm, bjp);
        if ((nk / nbcimp) * nbcimp == nk) launch_wake(sm, d, 3, w, q, dq, s);
    }
}

template void launch_lrelax_ref<double>(hipStream_t, const Dims &, const DevParams &, int, int,
                                        int, const WakePairs &, const double *, double *,
                                        double *, double *, double *, double *, double *,
                                        double *, int);

}  // namespace smac
