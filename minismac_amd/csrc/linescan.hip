// Block line relaxation with the line solves parallel inside each line (option lr_solve = 1,
// and the only form for FP32-LHS and for the zebra line order).
//
// The reference (ref/lrelax.f:225-561, ref/btlu3.f:52-624) solves each line's 3x3
// block-tridiagonal system by block Thomas: the factorisation B_i = b_i - a_i C_(i-1),
// C_i = B_i^-1 c_i, then y_i = B_i^-1 (r_i - a_i y_(i-1)) forward and x_i = y_i - C_i x_(i+1)
// backward.  The factorisation does not depend on the right-hand side, so it is done once per
// direction for all lines at once (k_lr_factor, one thread per line); it also forms
// M_i = -B_i^-1 a_i, so that both substitutions are affine recurrences,
//     y_i = M_i y_(i-1) + B_i^-1 r_i,        x_i = (-C_i) x_(i+1) + y_i,
// and each is a prefix scan of affine maps (3x3 matrix, 3-vector) along the line.  One
// workgroup solves one line: every thread composes the maps of its P consecutive points,
// the waves scan their lanes' maps with shuffles (six doubling steps), the waves' totals are
// chained through LDS (a 3x3 matrix-vector product per wave), and every thread re-applies
// its own maps from the value entering its segment.  Two barriers per line.
//
// Lines keep the reference's Gauss-Seidel order (one workgroup walks them, line l reading
// the dq of line l-1 that it has just produced -- kept in registers, since a thread owns
// the same points of every line), or the zebra order (every line of one parity at once,
// one workgroup per line).  The arithmetic is reassociated, so results agree with the
// reference to rounding, not bit for bit.
//
// Operands are staged line-major: point i of line l at l*len + i in every plane, which is
// the natural layout for xi lines; for eta lines the factor kernel writes its planes
// transposed and s, dq are transposed around the eta sweeps.  The SA's line relaxation
// (lreltur / scatri, ref/lreltur.f:52-359) uses the same sweep with 1x1 blocks.
#include "smac_dev.h"
#include "smac_kernels.h"

namespace smac {

// operand planes per point: factored diagonal (B, or 1/b for the scalar system), M, -C,
// the block to the previous line and the block to the next line
template <int NEQ> constexpr int lr_planes() { return 5 * NEQ * NEQ; }

template <typename T, int NEQ>
struct Aff {
    T m[NEQ * NEQ];   // entry (r, c) at r + NEQ*c
    T v[NEQ];
};

// a <- a o b (b applied first)
template <typename T, int NEQ>
SMAC_DEV void aff_after(Aff<T, NEQ> &a, const Aff<T, NEQ> &b)
{
    Aff<T, NEQ> o;
#pragma unroll
    for (int r = 0; r < NEQ; ++r) {
#pragma unroll
        for (int c = 0; c < NEQ; ++c) {
            T acc = a.m[r] * b.m[NEQ * c];
#pragma unroll
            for (int k = 1; k < NEQ; ++k) acc = acc + a.m[r + NEQ * k] * b.m[k + NEQ * c];
            o.m[r + NEQ * c] = acc;
        }
        T acc = a.v[r];
#pragma unroll
        for (int k = 0; k < NEQ; ++k) acc = acc + a.m[r + NEQ * k] * b.v[k];
        o.v[r] = acc;
    }
    a = o;
}

template <typename T, int NEQ>
SMAC_DEV void aff_apply(const T *m, const T *v, const T *x, T *out)
{
#pragma unroll
    for (int r = 0; r < NEQ; ++r) {
        T acc = v[r];
#pragma unroll
        for (int k = 0; k < NEQ; ++k) acc = acc + m[r + NEQ * k] * x[k];
        out[r] = acc;
    }
}

template <typename T, int NEQ>
SMAC_DEV void aff_shfl_up(Aff<T, NEQ> &o, const Aff<T, NEQ> &a, int d)
{
#pragma unroll
    for (int e = 0; e < NEQ * NEQ; ++e) o.m[e] = __shfl_up(a.m[e], d, 64);
#pragma unroll
    for (int e = 0; e < NEQ; ++e) o.v[e] = __shfl_up(a.v[e], d, 64);
}

template <typename T, int NEQ>
SMAC_DEV void aff_shfl_down(Aff<T, NEQ> &o, const Aff<T, NEQ> &a, int d)
{
#pragma unroll
    for (int e = 0; e < NEQ * NEQ; ++e) o.m[e] = __shfl_down(a.m[e], d, 64);
#pragma unroll
    for (int e = 0; e < NEQ; ++e) o.v[e] = __shfl_down(a.v[e], d, 64);
}

// ------------------------------------------------------------------------------------------
// factorisation (3x3): btlu3j / btlu3k (ref/btlu3.f:52-624) for every line at once, one
// thread per line, writing the line-major operand planes.  lm / lp: blocks along the line
// (bjm/bjp for xi lines, bkm/bkp for eta lines); om / op: blocks to the neighbouring lines.

template <typename T>
SMAC_DEV void lu3s(T b[9])
{
    b[0] = (T)1. / b[0];
    b[3] = b[3] * b[0];
    b[6] = b[6] * b[0];
    b[4] = (T)1. / (b[4] - b[1] * b[3]);
    b[7] = (b[7] - b[1] * b[6]) * b[4];
    b[5] = b[5] - b[2] * b[3];
    b[8] = (T)1. / (b[8] - b[2] * b[6] - b[7] * b[5]);
}

template <typename T>
SMAC_DEV void csolve3s(const T b[9], T c[9])
{
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

template <typename LT>
__global__ void __launch_bounds__(64)
k_lr_factor3(Dims d, int eta, const LT *__restrict__ btc, const LT *__restrict__ lm,
             const LT *__restrict__ lp, const LT *__restrict__ om, const LT *__restrict__ op,
             LT *__restrict__ out)
{
    const long N = d.N;
    const int nline = eta ? d.jm : d.km, len = eta ? d.km : d.jm;
    const long stride = eta ? d.jm : 1, lstride = eta ? 1 : d.jm;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nline) return;
    LT cprev[9];
    for (int i = 0; i < len; ++i) {
        const long p = lstride * l + stride * i;     // natural layout
        const long o = (long)l * len + i;            // line-major
        LT bb[9], mm[9], cc[9], t[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) bb[e] = btc[e * N + p];
        if (i > 0) {
            LT aa[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) aa[e] = lm[e * N + p];
            // b <- b - a C(i-1)
#pragma unroll
            for (int n = 0; n < 3; ++n)
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    bb[m + 3 * n] = bb[m + 3 * n] - aa[m] * cprev[3 * n] -
                                    aa[m + 3] * cprev[1 + 3 * n] - aa[m + 6] * cprev[2 + 3 * n];
            lu3s(bb);
#pragma unroll
            for (int e = 0; e < 9; ++e) mm[e] = -aa[e];
            csolve3s(bb, mm);                        // M = -B^-1 a
        } else {
            lu3s(bb);
#pragma unroll
            for (int e = 0; e < 9; ++e) mm[e] = (LT)0;
        }
        if (i < len - 1) {
#pragma unroll
            for (int e = 0; e < 9; ++e) cc[e] = lp[e * N + p];
            csolve3s(bb, cc);                        // C = B^-1 c
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                cprev[e] = cc[e];
                cc[e] = -cc[e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < 9; ++e) cc[e] = (LT)0;
        }
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            out[e * N + o] = bb[e];
            out[(9 + e) * N + o] = mm[e];
            out[(18 + e) * N + o] = cc[e];
        }
#pragma unroll
        for (int e = 0; e < 9; ++e) t[e] = om[e * N + p];
#pragma unroll
        for (int e = 0; e < 9; ++e) out[(27 + e) * N + o] = t[e];
#pragma unroll
        for (int e = 0; e < 9; ++e) t[e] = op[e * N + p];
#pragma unroll
        for (int e = 0; e < 9; ++e) out[(36 + e) * N + o] = t[e];
    }
}

// factorisation of the SA's scalar xi lines (scatri, ref/lreltur.f:337-359): 1/b, -a/b, -C,
// and the k-neighbour coefficients skm/skp, all in the natural layout (= line-major for xi)
__global__ void __launch_bounds__(64)
k_lr_factor1(Dims d, const double *__restrict__ sjm, const double *__restrict__ stc,
             const double *__restrict__ sjp, const double *__restrict__ skm,
             const double *__restrict__ skp, double *__restrict__ out)
{
    const long N = d.N;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.km) return;
    double cprev = 0.0;
    for (int j = 0; j < d.jm; ++j) {
        const long p = P(d, j, k);
        const double a = sjm[p], b = stc[p];
        const double z = j > 0 ? 1. / (b - a * cprev) : 1. / b;
        double c = 0.0;
        if (j < d.jm - 1) {
            c = sjp[p] * z;
            cprev = c;
        }
        out[p] = z;
        out[N + p] = j > 0 ? -a * z : 0.0;
        out[2 * N + p] = -c;
        out[3 * N + p] = skm[p];
        out[4 * N + p] = skp[p];
    }
}

// ------------------------------------------------------------------------------------------
// the sweep: lines count of them, line l = (first + idx*step) mod nline.  SEQ: one
// workgroup walks them in order; otherwise one workgroup per line (lines of one colour).
// s, dq: NEQ planes, line-major; op: the factor kernel's planes.  assign: dq = x (the SA's
// lreltur); otherwise dq += underr (x - dq) (ref/lrelax.f:377-381).

template <typename LT, int NEQ, int NT, int PP, bool SEQ>
__global__ void __launch_bounds__(NT)
k_lr_sweep(int nline, int len, long N, int count, int first, int step, double underr,
           int assign, double *__restrict__ dq, const double *__restrict__ s,
           const LT *__restrict__ op)
{
    using AT = LT;
    constexpr int NB = NEQ * NEQ, NW = NT / 64;
    __shared__ AT lf[NW][NB + NEQ], lb[NW][NB + NEQ];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const LT *pB = op, *pM = op + NB * N, *pC = op + 2 * NB * N, *pOm = op + 3 * NB * N,
             *pOp = op + 4 * NB * N;
    double dqprev[PP][NEQ];
    int lprev = -2;
    for (int idx = SEQ ? 0 : (int)blockIdx.x; idx < count; idx += SEQ ? 1 : (int)gridDim.x) {
        int l = (first + idx * step) % nline;
        if (l < 0) l += nline;
        const bool hm = l > 0, hp = l < nline - 1;
        const bool regm = SEQ && lprev == l - 1, regp = SEQ && lprev == l + 1;
        Aff<AT, NEQ> mp[PP], agg;
        AT cm[PP][NB];
        double dold[PP][NEQ];
        // ---- maps of this thread's points: y_i = M_i y_(i-1) + B_i^-1 r_i
#pragma unroll
        for (int u = 0; u < PP; ++u) {
            const int i = t * PP + u;
            if (i < len) {
                const long o = (long)l * len + i;
                double r[NEQ];
#pragma unroll
                for (int n = 0; n < NEQ; ++n) {
                    r[n] = s[n * N + o];
                    dold[u][n] = dq[n * N + o];
                }
                if (hm) {
                    double x[NEQ];
#pragma unroll
                    for (int n = 0; n < NEQ; ++n) x[n] = regm ? dqprev[u][n] : dq[n * N + o - len];
#pragma unroll
                    for (int m = 0; m < NEQ; ++m) {
                        double b = (double)pOm[m * N + o] * x[0];
#pragma unroll
                        for (int n = 1; n < NEQ; ++n) b = b + (double)pOm[(m + NEQ * n) * N + o] * x[n];
                        r[m] = r[m] - b;
                    }
                }
                if (hp) {
                    double x[NEQ];
#pragma unroll
                    for (int n = 0; n < NEQ; ++n) x[n] = regp ? dqprev[u][n] : dq[n * N + o + len];
#pragma unroll
                    for (int m = 0; m < NEQ; ++m) {
                        double b = (double)pOp[m * N + o] * x[0];
#pragma unroll
                        for (int n = 1; n < NEQ; ++n) b = b + (double)pOp[(m + NEQ * n) * N + o] * x[n];
                        r[m] = r[m] - b;
                    }
                }
                if constexpr (NEQ == 3) {
                    AT bb[9];
#pragma unroll
                    for (int e = 0; e < 9; ++e) bb[e] = pB[e * N + o];
                    const AT t1 = (AT)r[0], t2 = (AT)r[1], t3 = (AT)r[2];
                    const AT d1 = bb[0] * t1;
                    const AT d2 = bb[4] * (t2 - bb[1] * d1);
                    const AT d3 = bb[8] * (t3 - bb[2] * d1 - bb[5] * d2);
                    mp[u].v[2] = d3;
                    mp[u].v[1] = d2 - bb[7] * d3;
                    mp[u].v[0] = d1 - bb[6] * d3 - bb[3] * mp[u].v[1];
                } else {
                    mp[u].v[0] = pB[o] * (AT)r[0];
                }
#pragma unroll
                for (int e = 0; e < NB; ++e) {
                    mp[u].m[e] = pM[e * N + o];
                    cm[u][e] = pC[e * N + o];
                }
            } else {
#pragma unroll
                for (int e = 0; e < NB; ++e) {
                    mp[u].m[e] = (AT)0;
                    cm[u][e] = (AT)0;
                }
#pragma unroll
                for (int n = 0; n < NEQ; ++n) {
                    mp[u].v[n] = (AT)0;
                    dold[u][n] = 0.0;
                }
            }
            if (u == 0) agg = mp[0];
            else {
                Aff<AT, NEQ> a = mp[u];
                aff_after(a, agg);
                agg = a;
            }
        }
        // ---- forward scan over the lanes, then over the waves
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            Aff<AT, NEQ> o;
            aff_shfl_up(o, agg, off);
            if (lane >= off) aff_after(agg, o);
        }
        if (lane == 63) {
#pragma unroll
            for (int e = 0; e < NB; ++e) lf[wv][e] = agg.m[e];
#pragma unroll
            for (int n = 0; n < NEQ; ++n) lf[wv][NB + n] = agg.v[n];
        }
        __syncthreads();
        AT Y[NEQ];
#pragma unroll
        for (int n = 0; n < NEQ; ++n) Y[n] = (AT)0;
        for (int w = 0; w < wv; ++w) {
            AT nY[NEQ];
            aff_apply<AT, NEQ>(lf[w], lf[w] + NB, Y, nY);
#pragma unroll
            for (int n = 0; n < NEQ; ++n) Y[n] = nY[n];
        }
        AT ye[NEQ], cur[NEQ];
        aff_apply<AT, NEQ>(agg.m, agg.v, Y, ye);
#pragma unroll
        for (int n = 0; n < NEQ; ++n) {
            const AT up = __shfl_up(ye[n], 1, 64);
            cur[n] = lane == 0 ? Y[n] : up;
        }
        AT y[PP][NEQ];
#pragma unroll
        for (int u = 0; u < PP; ++u) {
            AT nx[NEQ];
            aff_apply<AT, NEQ>(mp[u].m, mp[u].v, cur, nx);
#pragma unroll
            for (int n = 0; n < NEQ; ++n) y[u][n] = cur[n] = nx[n];
        }
        // ---- backward: x_i = (-C_i) x_(i+1) + y_i, scanned from the end of the line
        Aff<AT, NEQ> bagg;
#pragma unroll
        for (int u = PP - 1; u >= 0; --u) {
            Aff<AT, NEQ> a;
#pragma unroll
            for (int e = 0; e < NB; ++e) a.m[e] = cm[u][e];
#pragma unroll
            for (int n = 0; n < NEQ; ++n) a.v[n] = y[u][n];
            if (u == PP - 1) bagg = a;
            else {
                aff_after(a, bagg);
                bagg = a;
            }
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            Aff<AT, NEQ> o;
            aff_shfl_down(o, bagg, off);
            if (lane + off < 64) aff_after(bagg, o);
        }
        if (lane == 0) {
#pragma unroll
            for (int e = 0; e < NB; ++e) lb[wv][e] = bagg.m[e];
#pragma unroll
            for (int n = 0; n < NEQ; ++n) lb[wv][NB + n] = bagg.v[n];
        }
        __syncthreads();
        AT X[NEQ];
#pragma unroll
        for (int n = 0; n < NEQ; ++n) X[n] = (AT)0;
        for (int w = NW - 1; w > wv; --w) {
            AT nX[NEQ];
            aff_apply<AT, NEQ>(lb[w], lb[w] + NB, X, nX);
#pragma unroll
            for (int n = 0; n < NEQ; ++n) X[n] = nX[n];
        }
        AT xs[NEQ];
        aff_apply<AT, NEQ>(bagg.m, bagg.v, X, xs);
#pragma unroll
        for (int n = 0; n < NEQ; ++n) {
            const AT dn = __shfl_down(xs[n], 1, 64);
            cur[n] = lane == 63 ? X[n] : dn;
        }
#pragma unroll
        for (int u = PP - 1; u >= 0; --u) {
            AT nx[NEQ];
            aff_apply<AT, NEQ>(cm[u], y[u], cur, nx);
#pragma unroll
            for (int n = 0; n < NEQ; ++n) cur[n] = nx[n];
            const int i = t * PP + u;
            if (i < len) {
                const long o = (long)l * len + i;
#pragma unroll
                for (int n = 0; n < NEQ; ++n) {
                    const double xd = (double)cur[n];
                    const double nv = assign ? xd : dold[u][n] + underr * (xd - dold[u][n]);
                    dq[n * N + o] = nv;
                    dqprev[u][n] = nv;
                }
            }
        }
        lprev = l;
    }
}

// ------------------------------------------------------------------------------------------
// natural (j fastest) <-> eta-major (k fastest) transposition of `planes` planes of N

constexpr int TT = 32;

__global__ void __launch_bounds__(256)
k_transpose(const double *__restrict__ src, double *__restrict__ dst, int rows, int cols, long N)
{
    // src plane: rows x cols, cols fastest; dst plane: cols x rows, rows fastest
    __shared__ double tile[TT][TT + 1];
    const long base = (long)blockIdx.z * N;
    const int c0 = blockIdx.x * TT, r0 = blockIdx.y * TT;
    for (int rr = threadIdx.y; rr < TT; rr += 8) {
        const int r = r0 + rr, c = c0 + threadIdx.x;
        if (r < rows && c < cols) tile[rr][threadIdx.x] = src[base + (long)r * cols + c];
    }
    __syncthreads();
    for (int cc = threadIdx.y; cc < TT; cc += 8) {
        const int c = c0 + cc, r = r0 + threadIdx.x;
        if (r < rows && c < cols) dst[base + (long)c * rows + r] = tile[threadIdx.x][cc];
    }
}

static void transpose(hipStream_t sm, const double *src, double *dst, int rows, int cols, long N,
                      int planes)
{
    dim3 g((cols + TT - 1) / TT, (rows + TT - 1) / TT, planes);
    hipLaunchKernelGGL(k_transpose, g, dim3(TT, 8), 0, sm, src, dst, rows, cols, N);
}

// wake-cut injection (bcimpds single-rank branch, ref/bc.f:571-591) with ds and rhs in the
// eta-major layout; q stays natural.  Pairs outside the row are skipped.
__global__ void k_wake_eta(long N, int jm, int km, int neqs, int npair, int base,
                           const int *__restrict__ pairs, const double *__restrict__ sq,
                           double *__restrict__ ds, double *__restrict__ rhs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npair) return;
    const long jt = pairs[2 * i] - base, jb = pairs[2 * i + 1] - base;
    if (jt < 0 || jt >= jm || jb < 0 || jb >= jm) return;
    const long t0 = jt * km, t1 = jt * km + 1, b1 = jb * km + 1;   // (jt,0), (jt,1), (jb,1)
    for (int n = 0; n < neqs; ++n) {
        double snew = 0.5 * (sq[n * N + jt + jm] + sq[n * N + jb + jm]);
        double dsnew = 0.5 * (ds[n * N + t1] + ds[n * N + b1]);
        ds[n * N + t0] = snew + dsnew - sq[n * N + jt];
        rhs[n * N + t0] = ds[n * N + t0];
    }
}

__global__ void k_zero(double *__restrict__ a, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        a[i] = 0.0;
}

// ------------------------------------------------------------------------------------------

template <typename LT, int NEQ, bool SEQ>
static void sweep_launch(hipStream_t sm, int nline, int len, long N, int count, int first,
                         int step, double underr, int assign, double *dq, const double *s,
                         const LT *op)
{
    const dim3 g(SEQ ? 1 : count);
#define SW(NT, PPV)                                                                            \
    hipLaunchKernelGGL((k_lr_sweep<LT, NEQ, NT, PPV, SEQ>), g, dim3(NT), 0, sm, nline, len, N, \
                       count, first, step, underr, assign, dq, s, op)
    if (len <= 256) SW(256, 1);
    else if (len <= 512) SW(512, 1);
    else if (len <= 1024) SW(512, 2);
    else SW(1024, 2);
#undef SW
}

// one sweep of lines (reference order: odd sweeps ascending, even descending, as
// ref/lrelax.f:249-275; zebra: the lines of even index, then those of odd index)
template <typename LT, int NEQ>
static void sweep(hipStream_t sm, int nline, int len, long N, int nsweep, bool zebra,
                  double underr, double *dq, const double *s, const LT *op)
{
    if (zebra) {
        sweep_launch<LT, NEQ, false>(sm, nline, len, N, (nline + 1) / 2, 0, 2, underr, 0, dq, s, op);
        sweep_launch<LT, NEQ, false>(sm, nline, len, N, nline / 2, 1, 2, underr, 0, dq, s, op);
        return;
    }
    const bool rev = (nsweep / 2) * 2 == nsweep;
    sweep_launch<LT, NEQ, true>(sm, nline, len, N, nline, rev ? nline - 1 : 0, rev ? -1 : 1,
                                underr, 0, dq, s, op);
}

long lr_scan_bytes(const Dims &d)
{
    // 45 operand planes (FP64 size) + eta-major s and dq + the SA's 5 scalar planes
    return (long)(lr_planes<3>() + 6 + lr_planes<1>()) * d.N * 8 + 4 * 256;
}

template <typename LT>
void launch_lrelax_scan(hipStream_t sm, const Dims &d, const DevParams &prm, int njsp, int nksp,
                        int nbcimp, bool zebra, const WakePairs &w, const double *q, double *dq,
                        double *s, const LT *btc, const LT *bjm, const LT *bjp, const LT *bkm,
                        const LT *bkp, const LrScan &ws)
{
    const long N = d.N;
    LT *op = (LT *)ws.op;
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, sm, dq, 3 * N);
    if (njsp > 0) {
        hipLaunchKernelGGL(k_lr_factor3<LT>, dim3((d.km + 63) / 64), dim3(64), 0, sm, d, 0, btc,
                           bjm, bjp, bkm, bkp, op);
        for (int nj = 1; nj <= njsp; ++nj) {
            sweep<LT, 3>(sm, d.km, d.jm, N, nj, zebra, prm.underr, dq, s, op);
            if ((nj / nbcimp) * nbcimp == nj) launch_wake(sm, d, 3, w, q, dq, s);
        }
    }
    if (nksp > 0) {
        hipLaunchKernelGGL(k_lr_factor3<LT>, dim3((d.jm + 63) / 64), dim3(64), 0, sm, d, 1, btc,
                           bkm, bkp, bjm, bjp, op);
        transpose(sm, s, ws.s_t, d.km, d.jm, N, 3);
        transpose(sm, dq, ws.dq_t, d.km, d.jm, N, 3);
        for (int nk = 1; nk <= nksp; ++nk) {
            sweep<LT, 3>(sm, d.jm, d.km, N, nk, zebra, prm.underr, ws.dq_t, ws.s_t, op);
            if ((nk / nbcimp) * nbcimp == nk && w.n > 0)
                hipLaunchKernelGGL(k_wake_eta, dim3((w.n + 63) / 64), dim3(64), 0, sm, N, d.jm,
                                   d.km, 3, w.n, w.base, w.pairs, q, ws.dq_t, ws.s_t);
        }
        transpose(sm, ws.dq_t, dq, d.jm, d.km, N, 3);
        transpose(sm, ws.s_t, s, d.jm, d.km, N, 3);
    }
}

void launch_sa_lreltur_scan(hipStream_t sm, const Dims &d, int ntjsp, const WakePairs &w,
                            SaArrays &a, const LrScan &ws)
{
    // lreltur (ref/lreltur.f:52-252): ds = 0; ntjsp passes of the k lines in ascending order
    // followed by a re-solve of line k = 1, each pass ending with the wake injection
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, sm, a.ds, d.N);
    hipLaunchKernelGGL(k_lr_factor1, dim3((d.km + 63) / 64), dim3(64), 0, sm, d, a.sjm, a.stc,
                       a.sjp, a.skm, a.skp, ws.sa_op);
    for (int pass = 1; pass <= ntjsp; ++pass) {
        const dim3 g(1);
#define SW1(NT, PPV)                                                                           \
    hipLaunchKernelGGL((k_lr_sweep<double, 1, NT, PPV, true>), g, dim3(NT), 0, sm, d.km, d.jm,  \
                       d.N, d.km + 1, 0, 1, 1.0, 1, a.ds, a.sf, ws.sa_op)
        if (d.jm <= 256) SW1(256, 1);
        else if (d.jm <= 512) SW1(512, 1);
        else if (d.jm <= 1024) SW1(512, 2);
        else SW1(1024, 2);
#undef SW1
        launch_wake(sm, d, 1, w, a.anut, a.ds, a.sf);
    }
}

template void launch_lrelax_scan<double>(hipStream_t, const Dims &, const DevParams &, int, int,
                                         int, bool, const WakePairs &, const double *, double *,
                                         double *, const double *, const double *,
                                         const double *, const double *, const double *,
                                         const LrScan &);
template void launch_lrelax_scan<float>(hipStream_t, const Dims &, const DevParams &, int, int,
                                        int, bool, const WakePairs &, const double *, double *,
                                        double *, const float *, const float *, const float *,
                                        const float *, const float *, const LrScan &);

}  // namespace smac
