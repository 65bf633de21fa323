// LU-SGS of the implicit iteration: lusgs / lusini3 / lusini1 (ref/lusgs.f:51-236) and the
// approximate factorisation solve (D+L) D^-1 (D+U) of ilusol3 / ilusol1 (ref/ilusol.f).
//
// The reference sweeps in natural order (j fastest).  Point (j,k) of the forward sweep reads
// only (j-1,k) and (j,k-1) -- plus (jmax,k-1) at j = 1 and (jmax,kmax)/(1,kmax) at j = jmax
// in the backward sweep, through blocks that the C-grid boundary rows leave exactly zero
// (bjm at j = 1, bjp at j = jmax; ref/bc.f:1622-1648 after the blanking of ref/step.f:386-413)
// -- so visiting the hyperplanes j+k = const in order performs the reference's arithmetic
// operation for operation.
#include "smac_dev.h"
#include "smac_kernels.h"

namespace smac {

constexpr int HP_THREADS = 1024;

// ------------------------------------------------------------------------------------------
// Column pipeline: one workgroup, each thread owns C consecutive columns j.  At hyperplane h
// a thread updates the points (j, h-j) of its columns: the (j,k-1) neighbour is its own result
// of step h-1 (registers), the (j-1,k) neighbour is the result of the column to the left at
// step h-1 (registers, or the left thread's last column through LDS).  One barrier per
// hyperplane.  The diagonal factor (lusini3) is formed in registers from btc, and the D
// multiply of the middle step is applied as soon as a point's forward value is known (it is
// pointwise), so the forward pass writes the middle result directly.  The arithmetic of every
// point is the reference's (same expressions, same order), so the result is bit-identical.
constexpr int CP_MAXC = 4;

template <typename LT>
SMAC_DEV void lu_factor_reg(const LT *__restrict__ btc, long N, long p, double f[9])
{
    // lusini3 (ref/lusgs.f:196-209) in registers, entry (m,n) at m + 3n
    double t11 = btc[p], t21 = btc[N + p], t31 = btc[2 * N + p];
    double t12 = btc[3 * N + p], t22 = btc[4 * N + p], t32 = btc[5 * N + p];
    double t13 = btc[6 * N + p], t23 = btc[7 * N + p], t33 = btc[8 * N + p];
    f[0] = 1. / t11;
    f[3] = t12 * f[0];
    f[6] = t13 * f[0];
    f[1] = t21;
    f[4] = 1. / (t22 - f[1] * f[3]);
    f[7] = (t23 - f[1] * f[6]) * f[4];
    f[2] = t31;
    f[5] = t32 - f[2] * f[3];
    f[8] = 1. / (t33 - f[2] * f[6] - f[5] * f[7]);
}

SMAC_DEV void lu_solve_reg(const double f[9], double r1, double r2, double r3, double o[3])
{
    double z1 = r1 * f[0];
    double z2 = (r2 - f[1] * z1) * f[4];
    double z3 = (r3 - f[2] * z1 - f[5] * z2) * f[8];
    o[2] = z3;
    o[1] = z2 - f[7] * o[2];
    o[0] = z1 - f[6] * o[2] - f[3] * o[1];
}

template <typename LT, int C>
__global__ void __launch_bounds__(HP_THREADS)
k_lusgs3_fwd_cp(Dims d, double *__restrict__ rhs, const LT *__restrict__ btc,
                const LT *__restrict__ bjm, const LT *__restrict__ bkm, LT *__restrict__ lu_out)
{
    __shared__ double xch[2][HP_THREADS][3];
    const long N = d.N;
    const int t = threadIdx.x, j0 = t * C;
    const int H = d.jm + d.km - 2;
    double res[C][3];
#pragma unroll
    for (int c = 0; c < C; ++c) res[c][0] = res[c][1] = res[c][2] = 0.0;
    for (int h = 0; h <= H; ++h) {
        double left[3] = {0.0, 0.0, 0.0};
        if (t > 0 && h > 0) {
            left[0] = xch[(h - 1) & 1][t - 1][0];
            left[1] = xch[(h - 1) & 1][t - 1][1];
            left[2] = xch[(h - 1) & 1][t - 1][2];
        }
#pragma unroll
        for (int c = C - 1; c >= 0; --c) {
            const int j = j0 + c, k = h - j;
            if (j >= d.jm || k < 0 || k >= d.km) continue;
            const long p = P(d, j, k);
            double r[3] = {rhs[p], rhs[N + p], rhs[2 * N + p]};
            if (j > 0) {
                const double a1 = c > 0 ? res[c - 1][0] : left[0];
                const double a2 = c > 0 ? res[c - 1][1] : left[1];
                const double a3 = c > 0 ? res[c - 1][2] : left[2];
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    r[m] = r[m] - bjm[m * N + p] * a1 - bjm[(m + 3) * N + p] * a2 -
                           bjm[(m + 6) * N + p] * a3;
            }
            if (k > 0) {
                // the j = 1 term through bjm(1,k) is an exact zero and is not formed
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    r[m] = r[m] - bkm[m * N + p] * res[c][0] - bkm[(m + 3) * N + p] * res[c][1] -
                           bkm[(m + 6) * N + p] * res[c][2];
            }
            double f[9];
            lu_factor_reg(btc, N, p, f);
            double o[3];
            lu_solve_reg(f, r[0], r[1], r[2], o);
            res[c][0] = o[0]; res[c][1] = o[1]; res[c][2] = o[2];
            // middle step: D recomposed from its factors (ref/ilusol.f:307-341)
            double b11 = 1. / f[0];
            double b12 = f[3] * b11;
            double b13 = f[6] * b11;
            double b21 = f[1];
            double b22 = 1. / f[4] + f[1] * f[3];
            double b23 = f[7] / f[4] + f[1] * f[6];
            double b31 = f[2];
            double b32 = f[5] + f[2] * f[3];
            double b33 = 1. / f[8] + f[2] * f[6] + f[5] * f[7];
            rhs[p] = b11 * o[0] + b12 * o[1] + b13 * o[2];
            rhs[N + p] = b21 * o[0] + b22 * o[1] + b23 * o[2];
            rhs[2 * N + p] = b31 * o[0] + b32 * o[1] + b33 * o[2];
            if (lu_out)
#pragma unroll
                for (int e = 0; e < 9; ++e) lu_out[e * N + p] = (LT)f[e];
        }
        xch[h & 1][t][0] = res[C - 1][0];
        xch[h & 1][t][1] = res[C - 1][1];
        xch[h & 1][t][2] = res[C - 1][2];
        __syncthreads();
    }
}

template <typename LT, int C>
__global__ void __launch_bounds__(HP_THREADS)
k_lusgs3_bwd_cp(Dims d, double *__restrict__ dq, const double *__restrict__ rhs,
                const LT *__restrict__ btc, const LT *__restrict__ bjp, const LT *__restrict__ bkp)
{
    __shared__ double xch[2][HP_THREADS][3];
    const long N = d.N;
    const int t = threadIdx.x, j0 = t * C;
    const int H = d.jm + d.km - 2;
    double res[C][3];
#pragma unroll
    for (int c = 0; c < C; ++c) res[c][0] = res[c][1] = res[c][2] = 0.0;
    for (int h = H; h >= 0; --h) {
        double right[3] = {0.0, 0.0, 0.0};
        if (h < H && t + 1 < (int)blockDim.x) {
            right[0] = xch[(h + 1) & 1][t + 1][0];
            right[1] = xch[(h + 1) & 1][t + 1][1];
            right[2] = xch[(h + 1) & 1][t + 1][2];
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int j = j0 + c, k = h - j;
            if (j >= d.jm || k < 0 || k >= d.km) continue;
            const long p = P(d, j, k);
            double r[3] = {rhs[p], rhs[N + p], rhs[2 * N + p]};
            if (k < d.km - 1) {
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    r[m] = r[m] - bkp[m * N + p] * res[c][0] - bkp[(m + 3) * N + p] * res[c][1] -
                           bkp[(m + 6) * N + p] * res[c][2];
            }
            if (j < d.jm - 1) {
                // the j = jmax term through bjp(jmax,k) is an exact zero and is not formed
                const double a1 = c + 1 < C ? res[c + 1 < C ? c + 1 : 0][0] : right[0];
                const double a2 = c + 1 < C ? res[c + 1 < C ? c + 1 : 0][1] : right[1];
                const double a3 = c + 1 < C ? res[c + 1 < C ? c + 1 : 0][2] : right[2];
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    r[m] = r[m] - bjp[m * N + p] * a1 - bjp[(m + 3) * N + p] * a2 -
                           bjp[(m + 6) * N + p] * a3;
            }
            double f[9];
            lu_factor_reg(btc, N, p, f);
            double o[3];
            lu_solve_reg(f, r[0], r[1], r[2], o);
            res[c][0] = o[0]; res[c][1] = o[1]; res[c][2] = o[2];
            dq[p] = o[0]; dq[N + p] = o[1]; dq[2 * N + p] = o[2];
        }
        xch[h & 1][t][0] = res[0][0];
        xch[h & 1][t][1] = res[0][1];
        xch[h & 1][t][2] = res[0][2];
        __syncthreads();
    }
}

static void cp_shape(int jm, int &threads, int &C)
{
    C = (jm + HP_THREADS - 1) / HP_THREADS;
    threads = (jm + C - 1) / C;
    threads = (threads + 63) / 64 * 64;
}

template <typename LT>
void launch_lusgs3(hipStream_t sm, const Dims &d, const WakePairs &w, const double *q,
                   double *dq, double *rhs, const LT *btc, const LT *bjm, const LT *bjp,
                   const LT *bkm, const LT *bkp, LT *scr1)
{
    // scr1 receives the diagonal factors (the reference's output of lusini3) when non-null
    int threads, C;
    cp_shape(d.jm, threads, C);
#define CP_LAUNCH(CC)                                                                           \
    do {                                                                                        \
        hipLaunchKernelGGL((k_lusgs3_fwd_cp<LT, CC>), dim3(1), dim3(threads), 0, sm, d, rhs,   \
                           btc, bjm, bkm, scr1);                                                \
        hipLaunchKernelGGL((k_lusgs3_bwd_cp<LT, CC>), dim3(1), dim3(threads), 0, sm, d, dq,    \
                           rhs, btc, bjp, bkp);                                                 \
    } while (0)
    switch (C) {   // jmax <= CP_MAXC * HP_THREADS is checked at set-up
    case 1: CP_LAUNCH(1); break;
    case 2: CP_LAUNCH(2); break;
    case 3: CP_LAUNCH(3); break;
    default: CP_LAUNCH(4); break;
    }
#undef CP_LAUNCH
    launch_wake(sm, d, 3, w, q, dq, rhs);
}

// ------------------------------------------------------------------------------------------
// scalar (SA) LU-SGS: lusini1 + ilusol1 (ref/lusgs.f:218-236, ref/ilusol.f:486-591)

__global__ void k_lusini1(long N, const double *__restrict__ btc, double *__restrict__ bp)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x)
        bp[i] = 1. / btc[i];
}

__global__ void __launch_bounds__(HP_THREADS)
k_lusgs1_fwd(Dims d, double *__restrict__ rhs, const double *__restrict__ b,
             const double *__restrict__ bjm, const double *__restrict__ bkm)
{
    const int H = d.jm + d.km - 2;
    for (int h = 0; h <= H; ++h) {
        const int jlo = h - (d.km - 1) > 0 ? h - (d.km - 1) : 0;
        const int jhi = h < d.jm - 1 ? h : d.jm - 1;
        for (int j = jlo + threadIdx.x; j <= jhi; j += HP_THREADS) {
            const int k = h - j;
            const long p = P(d, j, k);
            double v;
            if (k == 0) {
                v = j == 0 ? rhs[p] * b[p] : b[p] * (rhs[p] - bjm[p] * rhs[p - 1]);
            } else if (j == 0) {
                // bjm(1,k) * rhs(jmax,k-1) is an exact zero and is not formed
                v = b[p] * (rhs[p] - bkm[p] * rhs[p - d.jm]);
            } else {
                v = b[p] * (rhs[p] - bjm[p] * rhs[p - 1] - bkm[p] * rhs[p - d.jm]);
            }
            rhs[p] = v;
        }
        __syncthreads();
    }
}

__global__ void k_lusgs1_mid(long N, double *__restrict__ rhs, const double *__restrict__ b)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x)
        rhs[i] = rhs[i] / b[i];
}

__global__ void __launch_bounds__(HP_THREADS)
k_lusgs1_bwd(Dims d, double *__restrict__ dq, const double *__restrict__ rhs,
             const double *__restrict__ b, const double *__restrict__ bjp,
             const double *__restrict__ bkp)
{
    const int H = d.jm + d.km - 2;
    for (int h = H; h >= 0; --h) {
        const int jlo = h - (d.km - 1) > 0 ? h - (d.km - 1) : 0;
        const int jhi = h < d.jm - 1 ? h : d.jm - 1;
        for (int j = jlo + threadIdx.x; j <= jhi; j += HP_THREADS) {
            const int k = h - j;
            const long p = P(d, j, k);
            double v;
            if (k == d.km - 1) {
                v = j == d.jm - 1 ? rhs[p] * b[p] : b[p] * (rhs[p] - bjp[p] * dq[p + 1]);
            } else if (j == d.jm - 1) {
                // bjp(jmax,k) * dq(1,k+1) is an exact zero and is not formed
                v = b[p] * (rhs[p] - bkp[p] * dq[p + d.jm]);
            } else {
                v = b[p] * (rhs[p] - bkp[p] * dq[p + d.jm] - bjp[p] * dq[p + 1]);
            }
            dq[p] = v;
        }
        __syncthreads();
    }
}

void launch_lusgs1(hipStream_t sm, const Dims &d, const WakePairs &w, const double *q,
                   double *dq, double *rhs, const double *btc, const double *bjm,
                   const double *bjp, const double *bkm, const double *bkp, double *scr1)
{
    const long N = d.N;
    int nb = (int)((N + 255) / 256);
    nb = nb > 2048 ? 2048 : nb;
    hipLaunchKernelGGL(k_lusini1, dim3(nb), dim3(256), 0, sm, N, btc, scr1);
    hipLaunchKernelGGL(k_lusgs1_fwd, dim3(1), dim3(HP_THREADS), 0, sm, d, rhs, scr1, bjm, bkm);
    hipLaunchKernelGGL(k_lusgs1_mid, dim3(nb), dim3(256), 0, sm, N, rhs, scr1);
    hipLaunchKernelGGL(k_lusgs1_bwd, dim3(1), dim3(HP_THREADS), 0, sm, d, dq, rhs, scr1, bjp, bkp);
    launch_wake(sm, d, 1, w, q, dq, rhs);
}

template void launch_lusgs3<double>(hipStream_t, const Dims &, const WakePairs &, const double *,
                                    double *, double *, const double *, const double *,
                                    const double *, const double *, const double *, double *);

}  // namespace smac
