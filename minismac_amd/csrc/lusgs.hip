// LU-SGS of the implicit iteration: lusgs / lusini3 / lusini1 (ref/lusgs.f:51-236) and the
// approximate factorisation solve (D+L) D^-1 (D+U) of ilusol3 / ilusol1 (ref/ilusol.f).
//
// The reference sweeps in natural order (j fastest).  Point (j,k) of the forward sweep reads
// only (j-1,k) and (j,k-1) -- plus (jmax,k-1) at j = 1 and (jmax,kmax)/(1,kmax) at j = jmax
// in the backward sweep, through blocks that the C-grid boundary rows leave exactly zero
// (bjm at j = 1, bjp at j = jmax; ref/bc.f:1622-1648 after the blanking of ref/step.f:386-413)
// -- so visiting the hyperplanes j+k = const in order performs the reference's arithmetic
// operation for operation.  One workgroup walks the hyperplanes with a barrier between them;
// each point's update is the reference's expression, so the result is bit-identical.
#include "smac_dev.h"
#include "smac_kernels.h"

namespace smac {

constexpr int HP_THREADS = 1024;

template <typename LT>
__global__ void k_lusini3(long N, const LT *__restrict__ btc, LT *__restrict__ bp)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        // ref/lusgs.f:196-209; entry (m,n) at plane m + 3n
        LT b11 = (LT)1. / btc[i];
        LT b12 = btc[3 * N + i] * b11;
        LT b13 = btc[6 * N + i] * b11;
        LT b21 = btc[N + i];
        LT b22 = (LT)1. / (btc[4 * N + i] - b21 * b12);
        LT b23 = (btc[7 * N + i] - b21 * b13) * b22;
        LT b31 = btc[2 * N + i];
        LT b32 = btc[5 * N + i] - b31 * b12;
        LT b33 = (LT)1. / (btc[8 * N + i] - b31 * b13 - b32 * b23);
        bp[i] = b11; bp[3 * N + i] = b12; bp[6 * N + i] = b13;
        bp[N + i] = b21; bp[4 * N + i] = b22; bp[7 * N + i] = b23;
        bp[2 * N + i] = b31; bp[5 * N + i] = b32; bp[8 * N + i] = b33;
    }
}

// z = LU-solve(r) with the factors of lusini3 (ref/ilusol.f:80-88)
template <typename LT>
SMAC_DEV void lu_apply(const LT *__restrict__ b, long N, long p, double r1, double r2, double r3,
                       double &o1, double &o2, double &o3)
{
    double z1 = r1 * b[p];
    double z2 = (r2 - b[N + p] * z1) * b[4 * N + p];
    double z3 = (r3 - b[2 * N + p] * z1 - b[5 * N + p] * z2) * b[8 * N + p];
    o3 = z3;
    o2 = z2 - b[7 * N + p] * o3;
    o1 = z1 - b[6 * N + p] * o3 - b[3 * N + p] * o2;
}

template <typename LT>
__global__ void __launch_bounds__(HP_THREADS)
k_lusgs3_fwd(Dims d, double *__restrict__ rhs, const LT *__restrict__ b,
             const LT *__restrict__ bjm, const LT *__restrict__ bkm)
{
    const long N = d.N;
    const int H = d.jm + d.km - 2;
    for (int h = 0; h <= H; ++h) {
        const int jlo = h - (d.km - 1) > 0 ? h - (d.km - 1) : 0;
        const int jhi = h < d.jm - 1 ? h : d.jm - 1;
        for (int j = jlo + threadIdx.x; j <= jhi; j += HP_THREADS) {
            const int k = h - j;
            const long p = P(d, j, k);
            double r[3] = {rhs[p], rhs[N + p], rhs[2 * N + p]};
            if (j > 0) {
                // ref/ilusol.f:150-158 (k = 1) and :252-270
                const long q = p - 1;
                double a1 = rhs[q], a2 = rhs[N + q], a3 = rhs[2 * N + q];
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    r[m] = r[m] - bjm[m * N + p] * a1 - bjm[(m + 3) * N + p] * a2 -
                           bjm[(m + 6) * N + p] * a3;
            }
            if (k > 0) {
                // the j = 1 term through bjm(1,k) is an exact zero and is not formed
                const long q = p - d.jm;
                double a1 = rhs[q], a2 = rhs[N + q], a3 = rhs[2 * N + q];
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    r[m] = r[m] - bkm[m * N + p] * a1 - bkm[(m + 3) * N + p] * a2 -
                           bkm[(m + 6) * N + p] * a3;
            }
            double o1, o2, o3;
            lu_apply(b, N, p, r[0], r[1], r[2], o1, o2, o3);
            rhs[p] = o1; rhs[N + p] = o2; rhs[2 * N + p] = o3;
        }
        __syncthreads();
    }
}

template <typename LT>
__global__ void k_lusgs3_mid(long N, double *__restrict__ rhs, const LT *__restrict__ b)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        // D recomposed from its factors (ref/ilusol.f:307-341)
        double f11 = b[i], f12 = b[3 * N + i], f13 = b[6 * N + i];
        double f21 = b[N + i], f22 = b[4 * N + i], f23 = b[7 * N + i];
        double f31 = b[2 * N + i], f32 = b[5 * N + i], f33 = b[8 * N + i];
        double b11 = 1. / f11;
        double b12 = f12 * b11;
        double b13 = f13 * b11;
        double b21 = f21;
        double b22 = 1. / f22 + f21 * f12;
        double b23 = f23 / f22 + f21 * f13;
        double b31 = f31;
        double b32 = f32 + f31 * f12;
        double b33 = 1. / f33 + f31 * f13 + f32 * f23;
        double r1 = rhs[i], r2 = rhs[N + i], r3 = rhs[2 * N + i];
        rhs[i] = b11 * r1 + b12 * r2 + b13 * r3;
        rhs[N + i] = b21 * r1 + b22 * r2 + b23 * r3;
        rhs[2 * N + i] = b31 * r1 + b32 * r2 + b33 * r3;
    }
}

template <typename LT>
__global__ void __launch_bounds__(HP_THREADS)
k_lusgs3_bwd(Dims d, double *__restrict__ dq, const double *__restrict__ rhs,
             const LT *__restrict__ b, const LT *__restrict__ bjp, const LT *__restrict__ bkp)
{
    const long N = d.N;
    const int H = d.jm + d.km - 2;
    for (int h = H; h >= 0; --h) {
        const int jlo = h - (d.km - 1) > 0 ? h - (d.km - 1) : 0;
        const int jhi = h < d.jm - 1 ? h : d.jm - 1;
        for (int j = jlo + threadIdx.x; j <= jhi; j += HP_THREADS) {
            const int k = h - j;
            const long p = P(d, j, k);
            double r[3] = {rhs[p], rhs[N + p], rhs[2 * N + p]};
            if (k < d.km - 1) {
                // ref/ilusol.f:413-469: the eta neighbour first, then the xi neighbour
                const long q = p + d.jm;
                double a1 = dq[q], a2 = dq[N + q], a3 = dq[2 * N + q];
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    r[m] = r[m] - bkp[m * N + p] * a1 - bkp[(m + 3) * N + p] * a2 -
                           bkp[(m + 6) * N + p] * a3;
            }
            if (j < d.jm - 1) {
                // the j = jmax term through bjp(jmax,k) is an exact zero and is not formed
                const long q = p + 1;
                double a1 = dq[q], a2 = dq[N + q], a3 = dq[2 * N + q];
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    r[m] = r[m] - bjp[m * N + p] * a1 - bjp[(m + 3) * N + p] * a2 -
                           bjp[(m + 6) * N + p] * a3;
            }
            double o1, o2, o3;
            lu_apply(b, N, p, r[0], r[1], r[2], o1, o2, o3);
            dq[p] = o1; dq[N + p] = o2; dq[2 * N + p] = o3;
        }
        __syncthreads();
    }
}

template <typename LT>
void launch_lusgs3(hipStream_t sm, const Dims &d, const WakePairs &w, const double *q,
                   double *dq, double *rhs, const LT *btc, const LT *bjm, const LT *bjp,
                   const LT *bkm, const LT *bkp, LT *scr1)
{
    const long N = d.N;
    int nb = (int)((N + 255) / 256);
    nb = nb > 2048 ? 2048 : nb;
    hipLaunchKernelGGL(k_lusini3<LT>, dim3(nb), dim3(256), 0, sm, N, btc, scr1);
    hipLaunchKernelGGL(k_lusgs3_fwd<LT>, dim3(1), dim3(HP_THREADS), 0, sm, d, rhs, scr1, bjm, bkm);
    hipLaunchKernelGGL(k_lusgs3_mid<LT>, dim3(nb), dim3(256), 0, sm, N, rhs, scr1);
    hipLaunchKernelGGL(k_lusgs3_bwd<LT>, dim3(1), dim3(HP_THREADS), 0, sm, d, dq, rhs, scr1, bjp,
                       bkp);
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
