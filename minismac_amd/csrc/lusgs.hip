// LU-SGS of the implicit iteration: lusgs / lusini3 / lusini1 (ref/lusgs.f:51-236) and the
// approximate factorisation solve (D+L) D^-1 (D+U) of ilusol3 / ilusol1 (ref/ilusol.f).
//
// The reference sweeps in natural order (j fastest).  Point (j,k) of the forward sweep reads
// only (j-1,k) and (j,k-1) -- plus (jmax,k-1) at j = 1 and (jmax,kmax)/(1,kmax) at j = jmax
// in the backward sweep, through blocks that the C-grid boundary rows leave exactly zero
// (bjm at j = 1, bjp at j = jmax; ref/bc.f:1622-1648 after the blanking of ref/step.f:386-413)
// -- so visiting the hyperplanes h = j+k in order performs the reference's arithmetic
// operation for operation.
//
// Schedule on gfx950:
//  1. k_shear: the five block planes and the right-hand side are transposed, through LDS
//     tiles, into the sheared layout of the workspace, where point (j,k) sits at h*ks + k.
//     The diagonal factors (lusini3 / lusini1) are formed on the way (fully parallel).
//  2. k_fwd_band: the rows are cut into bands of `rows` <= 64 rows, one wavefront each.  Lane
//     l of band b owns row k = b*rows + l and walks h = j+k upward: (j-1,k) is its own result
//     of the previous step (registers), (j,k-1) is lane l-1's result of the previous step (a
//     cross-lane shuffle), so a step needs no barrier.  Loads of step h+PF are issued before
//     step h is computed.  The last row of band b hands its results to band b+1 through an
//     edge buffer in HBM whose slots start as a signalling-NaN sentinel (no arithmetic
//     produces it); the consumer spins on a relaxed agent-scope load until the slot is filled,
//     with a bounded spin count that raises a stall flag instead of hanging.
//  3. k_mid: the D multiply of the middle step (pointwise, fully parallel).
//  4. k_bwd_band: the mirror image, h downward, bands handing their first row down.
//  5. k_unshear: dq (and, for callers that want it, the middle-step rhs) back to the natural
//     layout; then the wake-cut rows (bcimpds) as in the reference.
// Every point performs the reference's expressions in the reference's order.
#include "smac_dev.h"
#include "smac_kernels.h"

namespace smac {

constexpr unsigned long long EDGE_EMPTY = 0x7ff4dead0000beefull;   // signalling NaN
constexpr int EDGE_SPIN_MAX = 200000;
constexpr int BAND_LANES = 64;

SMAC_DEV long SX(const LusgsWs &w, int j, int k) { return (long)(j + k) * w.ks + k; }

// ------------------------------------------------------------------------------------------
// diagonal factors

template <typename LT>
SMAC_DEV void lu_factor3(const LT t[9], LT f[9])
{
    // lusini3 (ref/lusgs.f:196-209), entry (m,n) at m + 3n
    f[0] = (LT)1. / t[0];
    f[3] = t[3] * f[0];
    f[6] = t[6] * f[0];
    f[1] = t[1];
    f[4] = (LT)1. / (t[4] - f[1] * f[3]);
    f[7] = (t[7] - f[1] * f[6]) * f[4];
    f[2] = t[2];
    f[5] = t[5] - f[2] * f[3];
    f[8] = (LT)1. / (t[8] - f[2] * f[6] - f[5] * f[7]);
}

template <typename AT>
SMAC_DEV void lu_solve3(const AT f[9], AT r1, AT r2, AT r3, AT o[3])
{
    // forward/back substitution with the lusini3 factors (ref/ilusol.f:146-170)
    AT z1 = r1 * f[0];
    AT z2 = (r2 - f[1] * z1) * f[4];
    AT z3 = (r3 - f[2] * z1 - f[5] * z2) * f[8];
    o[2] = z3;
    o[1] = z2 - f[7] * o[2];
    o[0] = z1 - f[6] * o[2] - f[3] * o[1];
}

// ------------------------------------------------------------------------------------------
// natural <-> sheared transposition through LDS tiles of TK rows x TJ columns

constexpr int TJ = 32, TK = 16, TPAD = TJ + 2, TILE_THREADS = 256;

template <int NP>
SMAC_DEV void tile_to_sheared(const LusgsWs &w, int jm, int km, int j0, int k0,
                              double (*lds)[TK][TPAD], double *dst, long stride)
{
    for (int e = threadIdx.x; e < (TJ + TK - 1) * TK; e += TILE_THREADS) {
        const int kk = e % TK, hh = e / TK, jj = hh - kk;
        const int j = j0 + jj, k = k0 + kk;
        if (jj < 0 || jj >= TJ || j >= jm || k >= km) continue;
        const long o = SX(w, j, k);
#pragma unroll
        for (int m = 0; m < NP; ++m) dst[m * stride + o] = lds[m][kk][jj];
    }
}

template <typename LT, int NP>
SMAC_DEV void tile_to_sheared_lt(const LusgsWs &w, int jm, int km, int j0, int k0,
                                 double (*lds)[TK][TPAD], LT *dst, long stride)
{
    for (int e = threadIdx.x; e < (TJ + TK - 1) * TK; e += TILE_THREADS) {
        const int kk = e % TK, hh = e / TK, jj = hh - kk;
        const int j = j0 + jj, k = k0 + kk;
        if (jj < 0 || jj >= TJ || j >= jm || k >= km) continue;
        const long o = SX(w, j, k);
#pragma unroll
        for (int m = 0; m < NP; ++m) dst[m * stride + o] = (LT)lds[m][kk][jj];
    }
}

template <typename LT, int NP>
SMAC_DEV void tile_from_natural(const Dims &d, int j0, int k0, double (*lds)[TK][TPAD],
                                const LT *src)
{
    for (int e = threadIdx.x; e < TK * TJ; e += TILE_THREADS) {
        const int kk = e / TJ, jj = e % TJ, j = j0 + jj, k = k0 + kk;
        if (j >= d.jm || k >= d.km) continue;
        const long p = P(d, j, k);
#pragma unroll
        for (int m = 0; m < NP; ++m) lds[m][kk][jj] = (double)src[m * d.N + p];
    }
}

// Transposes one tile of the system into the sheared workspace.  Group 0 is the diagonal,
// factored on the way (and written to `fac` in the natural layout when non-null); then
// bjm, bkm, bjp, bkp and the right-hand side.  Block 0 also resets the edge buffers.
template <typename LT, int NEQ>
__global__ void __launch_bounds__(TILE_THREADS)
k_shear(Dims d, LusgsWs w, const double *__restrict__ rhs, const LT *__restrict__ btc,
        const LT *__restrict__ bjm, const LT *__restrict__ bkm, const LT *__restrict__ bjp,
        const LT *__restrict__ bkp, LT *__restrict__ fac)
{
    constexpr int NB = NEQ * NEQ;
    __shared__ double lds[NB][TK][TPAD];
    const int j0 = blockIdx.x * TJ, k0 = blockIdx.y * TK;
    LT *wl = (LT *)w.lhs;
    const long lstride = w.ns;

    if (blockIdx.x == 0 && blockIdx.y == 0)
        for (long i = threadIdx.x; i < w.nedge; i += TILE_THREADS) w.edge[i] = EDGE_EMPTY;

    // group 0: diagonal factors
    for (int e = threadIdx.x; e < TK * TJ; e += TILE_THREADS) {
        const int kk = e / TJ, jj = e % TJ, j = j0 + jj, k = k0 + kk;
        if (j >= d.jm || k >= d.km) continue;
        const long p = P(d, j, k);
        LT t[NB], f[NB];
#pragma unroll
        for (int m = 0; m < NB; ++m) t[m] = btc[m * d.N + p];
        if constexpr (NEQ == 3) {
            lu_factor3<LT>(t, f);
        } else {
            f[0] = (LT)1. / t[0];   // lusini1 (ref/lusgs.f:230-234)
        }
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            lds[m][kk][jj] = (double)f[m];
            if (fac) fac[m * d.N + p] = f[m];
        }
    }
    __syncthreads();
    tile_to_sheared_lt<LT, NB>(w, d.jm, d.km, j0, k0, lds, wl, lstride);

    const LT *src[4] = {bjm, bkm, bjp, bkp};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        __syncthreads();
        tile_from_natural<LT, NB>(d, j0, k0, lds, src[g]);
        __syncthreads();
        tile_to_sheared_lt<LT, NB>(w, d.jm, d.km, j0, k0, lds, wl + (long)(g + 1) * NB * lstride,
                                   lstride);
    }
    __syncthreads();
    tile_from_natural<double, NEQ>(d, j0, k0, lds, rhs);
    __syncthreads();
    tile_to_sheared<NEQ>(w, d.jm, d.km, j0, k0, lds, w.r, lstride);
}

// sheared -> natural for NP planes
template <int NP>
__global__ void __launch_bounds__(TILE_THREADS)
k_unshear(Dims d, LusgsWs w, const double *__restrict__ src, double *__restrict__ dst)
{
    __shared__ double lds[NP][TK][TPAD];
    const int j0 = blockIdx.x * TJ, k0 = blockIdx.y * TK;
    for (int e = threadIdx.x; e < (TJ + TK - 1) * TK; e += TILE_THREADS) {
        const int kk = e % TK, hh = e / TK, jj = hh - kk;
        const int j = j0 + jj, k = k0 + kk;
        if (jj < 0 || jj >= TJ || j >= d.jm || k >= d.km) continue;
        const long o = SX(w, j, k);
#pragma unroll
        for (int m = 0; m < NP; ++m) lds[m][kk][jj] = src[m * w.ns + o];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < TK * TJ; e += TILE_THREADS) {
        const int kk = e / TJ, jj = e % TJ, j = j0 + jj, k = k0 + kk;
        if (j >= d.jm || k >= d.km) continue;
        const long p = P(d, j, k);
#pragma unroll
        for (int m = 0; m < NP; ++m) dst[m * d.N + p] = lds[m][kk][jj];
    }
}

// ------------------------------------------------------------------------------------------
// middle step: rhs <- D rhs, pointwise (ref/ilusol.f:307-341 for neqs = 3; ref/ilusol.f:548-556
// for neqs = 1, where D = 1/b)

template <typename LT, int NEQ>
__global__ void k_mid(Dims d, LusgsWs w)
{
    constexpr int NB = NEQ * NEQ;
    const LT *F = (const LT *)w.lhs;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < w.ns;
         i += (long)gridDim.x * blockDim.x) {
        const int h = (int)(i / w.ks), k = (int)(i % w.ks), j = h - k;
        if (k >= d.km || j < 0 || j >= d.jm) continue;
        if constexpr (NEQ == 3) {
            double f[NB];
#pragma unroll
            for (int m = 0; m < NB; ++m) f[m] = (double)F[m * w.ns + i];
            const double o0 = w.r[i], o1 = w.r[w.ns + i], o2 = w.r[2 * w.ns + i];
            double b11 = 1. / f[0];
            double b12 = f[3] * b11;
            double b13 = f[6] * b11;
            double b21 = f[1];
            double b22 = 1. / f[4] + f[1] * f[3];
            double b23 = f[7] / f[4] + f[1] * f[6];
            double b31 = f[2];
            double b32 = f[5] + f[2] * f[3];
            double b33 = 1. / f[8] + f[2] * f[6] + f[5] * f[7];
            w.r[i] = b11 * o0 + b12 * o1 + b13 * o2;
            w.r[w.ns + i] = b21 * o0 + b22 * o1 + b23 * o2;
            w.r[2 * w.ns + i] = b31 * o0 + b32 * o1 + b33 * o2;
        } else {
            w.r[i] = w.r[i] / (double)F[i];
        }
    }
}

// ------------------------------------------------------------------------------------------
// band sweeps

SMAC_DEV unsigned long long edge_ld(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

SMAC_DEV void edge_st(unsigned long long *p, double v)
{
    __hip_atomic_store(p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// waits (bounded) until the NEQ slots at p are filled; on time-out raises the stall flag and
// stops waiting for the rest of the sweep
template <int NEQ>
SMAC_DEV void edge_wait(const unsigned long long *p, unsigned long long e[NEQ], bool &broken,
                        int *stall)
{
    if (broken) return;
    int n = 0;
    for (;;) {
        bool ok = true;
#pragma unroll
        for (int m = 0; m < NEQ; ++m) ok = ok && e[m] != EDGE_EMPTY;
        if (ok) return;
        if (++n > EDGE_SPIN_MAX) {
            broken = true;
            atomicOr(stall, 1);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
#pragma unroll
        for (int m = 0; m < NEQ; ++m) e[m] = edge_ld(p + m);
    }
}

// Per-step operands of one lane, loaded PF steps ahead.
template <typename LT, int NEQ>
struct Slot {
    LT f[NEQ * NEQ], a[NEQ * NEQ], b[NEQ * NEQ];
    double r[NEQ];
    unsigned long long e[NEQ];
};

template <typename LT, int NEQ>
SMAC_DEV void slot_load(Slot<LT, NEQ> &s, const LusgsWs &w, const LT *F, const LT *A, const LT *B,
                        const double *R, int h, int k, bool edge_lane, const unsigned long long *ep,
                        int j, int jm)
{
    constexpr int NB = NEQ * NEQ;
    const int hc = h < w.H ? (h > 0 ? h : 0) : w.H;
    const long i = (long)hc * w.ks + k;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        s.f[m] = F[m * w.ns + i];
        s.a[m] = A[m * w.ns + i];
        s.b[m] = B[m * w.ns + i];
    }
#pragma unroll
    for (int m = 0; m < NEQ; ++m) s.r[m] = R[m * w.ns + i];
    if (edge_lane && j >= 0 && j < jm) {
#pragma unroll
        for (int m = 0; m < NEQ; ++m) s.e[m] = edge_ld(ep + (long)j * 4 + m);
    } else {
#pragma unroll
        for (int m = 0; m < NEQ; ++m) s.e[m] = 0ull;
    }
}

SMAC_DEV double shfl_prev(double v, int lane)   // value of lane-1
{
    return __shfl_up(v, 1, BAND_LANES);
}

SMAC_DEV double shfl_next(double v, int lane)   // value of lane+1
{
    return __shfl_down(v, 1, BAND_LANES);
}

// Forward sweep (ilusol3 / ilusol1 forward part, ref/ilusol.f:127-202 and 502-540): writes the
// forward result into the workspace rhs planes.
template <typename LT, int NEQ, int PF>
__global__ void __launch_bounds__(BAND_LANES)
k_fwd_band(Dims d, LusgsWs w)
{
    constexpr int NB = NEQ * NEQ;
    using AT = LT;   // arithmetic type of the sweep
    const int lane = threadIdx.x, b = blockIdx.x, nband = gridDim.x;
    const int k0 = b * w.rows, k = k0 + lane;
    const bool inband = lane < w.rows && k < d.km;
    const int klast = (k0 + w.rows < d.km ? k0 + w.rows : d.km) - 1;
    const int hbeg = k0, hend = d.jm - 1 + klast;
    const LT *wl = (const LT *)w.lhs;
    const LT *F = wl, *A = wl + (long)NB * w.ns, *B = wl + 2L * NB * w.ns;
    double *R = w.r;
    const bool consumer = b > 0 && lane == 0;
    const bool producer = b + 1 < nband && k == klast;
    const unsigned long long *ein = w.edge + (long)(b > 0 ? b - 1 : 0) * d.jm * 4;
    unsigned long long *eout = w.edge + (long)b * d.jm * 4;
    bool broken = false;

    AT own[NEQ];
#pragma unroll
    for (int m = 0; m < NEQ; ++m) own[m] = (AT)0;

    Slot<LT, NEQ> s[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u)
        slot_load<LT, NEQ>(s[u], w, F, A, B, R, hbeg + u, k, consumer, ein, hbeg + u - k0, d.jm);

    for (int h0 = hbeg; h0 <= hend; h0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int h = h0 + u;
            if (h <= hend) {
                const int j = h - k;
                const bool act = inband && j >= 0 && j < d.jm;
                AT nb[NEQ];
#pragma unroll
                for (int m = 0; m < NEQ; ++m) nb[m] = (AT)shfl_prev((double)own[m], lane);
                if (consumer && act) {
                    edge_wait<NEQ>(ein + (long)j * 4, s[u].e, broken, w.stall);
#pragma unroll
                    for (int m = 0; m < NEQ; ++m) nb[m] = (AT)__longlong_as_double((long long)s[u].e[m]);
                }
                AT o[NEQ];
                if constexpr (NEQ == 3) {
                    AT r[3] = {(AT)s[u].r[0], (AT)s[u].r[1], (AT)s[u].r[2]};
                    if (j > 0) {
#pragma unroll
                        for (int m = 0; m < 3; ++m)
                            r[m] = r[m] - s[u].a[m] * own[0] - s[u].a[m + 3] * own[1] -
                                   s[u].a[m + 6] * own[2];
                    }
                    if (k > 0) {
                        // at j = 1 the reference's bjm(1,k)*rhs(jmax,k-1) is an exact zero
#pragma unroll
                        for (int m = 0; m < 3; ++m)
                            r[m] = r[m] - s[u].b[m] * nb[0] - s[u].b[m + 3] * nb[1] -
                                   s[u].b[m + 6] * nb[2];
                    }
                    lu_solve3<AT>(s[u].f, r[0], r[1], r[2], o);
                } else {
                    const AT rr = (AT)s[u].r[0], f = s[u].f[0];
                    if (k == 0)
                        o[0] = j == 0 ? rr * f : f * (rr - s[u].a[0] * own[0]);
                    else if (j == 0)
                        o[0] = f * (rr - s[u].b[0] * nb[0]);
                    else
                        o[0] = f * (rr - s[u].a[0] * own[0] - s[u].b[0] * nb[0]);
                }
                if (act) {
                    const long i = (long)h * w.ks + k;
#pragma unroll
                    for (int m = 0; m < NEQ; ++m) R[m * w.ns + i] = (double)o[m];
                    if (producer)
#pragma unroll
                        for (int m = 0; m < NEQ; ++m) edge_st(eout + (long)j * 4 + m, (double)o[m]);
                }
#pragma unroll
                for (int m = 0; m < NEQ; ++m) own[m] = act ? o[m] : (AT)0;
            }
            slot_load<LT, NEQ>(s[u], w, F, A, B, R, h + PF, k, consumer, ein, h + PF - k0, d.jm);
        }
    }
}

// Backward sweep (ref/ilusol.f:353-469 and 560-591): reads the middle-step rhs, writes dq.
template <typename LT, int NEQ, int PF>
__global__ void __launch_bounds__(BAND_LANES)
k_bwd_band(Dims d, LusgsWs w)
{
    constexpr int NB = NEQ * NEQ;
    using AT = LT;
    // bands are numbered top down so that workgroups are dispatched in dependency order
    const int lane = threadIdx.x, nband = gridDim.x, b = nband - 1 - (int)blockIdx.x;
    const int k0 = b * w.rows, k = k0 + lane;
    const bool inband = lane < w.rows && k < d.km;
    const int klast = (k0 + w.rows < d.km ? k0 + w.rows : d.km) - 1;
    const int hbeg = k0, hend = d.jm - 1 + klast;
    const LT *wl = (const LT *)w.lhs;
    const LT *F = wl, *C = wl + 3L * NB * w.ns, *D = wl + 4L * NB * w.ns;
    const double *R = w.r;
    double *DQ = w.dq;
    const bool consumer = b + 1 < nband && k == klast;
    const bool producer = b > 0 && lane == 0;
    unsigned long long *eb = w.edge + (long)nband * d.jm * 4;   // backward half
    const unsigned long long *ein = eb + (long)(b + 1 < nband ? b + 1 : b) * d.jm * 4;
    unsigned long long *eout = eb + (long)b * d.jm * 4;
    bool broken = false;

    AT own[NEQ];
#pragma unroll
    for (int m = 0; m < NEQ; ++m) own[m] = (AT)0;

    Slot<LT, NEQ> s[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u)
        slot_load<LT, NEQ>(s[u], w, F, D, C, R, hend - u, k, consumer, ein, hend - u - k, d.jm);

    for (int h0 = hend; h0 >= hbeg; h0 -= PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int h = h0 - u;
            if (h >= hbeg) {
                const int j = h - k;
                const bool act = inband && j >= 0 && j < d.jm;
                AT nb[NEQ];
#pragma unroll
                for (int m = 0; m < NEQ; ++m) nb[m] = (AT)shfl_next((double)own[m], lane);
                if (consumer && act) {
                    edge_wait<NEQ>(ein + (long)j * 4, s[u].e, broken, w.stall);
#pragma unroll
                    for (int m = 0; m < NEQ; ++m) nb[m] = (AT)__longlong_as_double((long long)s[u].e[m]);
                }
                AT o[NEQ];
                // slot fields: a = bkp (k-neighbour), b = bjp (j-neighbour)
                if constexpr (NEQ == 3) {
                    AT r[3] = {(AT)s[u].r[0], (AT)s[u].r[1], (AT)s[u].r[2]};
                    if (k < d.km - 1) {
#pragma unroll
                        for (int m = 0; m < 3; ++m)
                            r[m] = r[m] - s[u].a[m] * nb[0] - s[u].a[m + 3] * nb[1] -
                                   s[u].a[m + 6] * nb[2];
                    }
                    if (j < d.jm - 1) {
                        // at j = jmax the reference's bjp(jmax,k)*dq(1,k+1) is an exact zero
#pragma unroll
                        for (int m = 0; m < 3; ++m)
                            r[m] = r[m] - s[u].b[m] * own[0] - s[u].b[m + 3] * own[1] -
                                   s[u].b[m + 6] * own[2];
                    }
                    lu_solve3<AT>(s[u].f, r[0], r[1], r[2], o);
                } else {
                    const AT rr = (AT)s[u].r[0], f = s[u].f[0];
                    if (k == d.km - 1)
                        o[0] = j == d.jm - 1 ? rr * f : f * (rr - s[u].b[0] * own[0]);
                    else if (j == d.jm - 1)
                        o[0] = f * (rr - s[u].a[0] * nb[0]);
                    else
                        o[0] = f * (rr - s[u].a[0] * nb[0] - s[u].b[0] * own[0]);
                }
                if (act) {
                    const long i = (long)h * w.ks + k;
#pragma unroll
                    for (int m = 0; m < NEQ; ++m) DQ[m * w.ns + i] = (double)o[m];
                    if (producer)
#pragma unroll
                        for (int m = 0; m < NEQ; ++m) edge_st(eout + (long)j * 4 + m, (double)o[m]);
                }
#pragma unroll
                for (int m = 0; m < NEQ; ++m) own[m] = act ? o[m] : (AT)0;
            }
            slot_load<LT, NEQ>(s[u], w, F, D, C, R, h - PF, k, consumer, ein, h - PF - k, d.jm);
        }
    }
}

// ------------------------------------------------------------------------------------------

void lusgs_ws_shape(int jm, int km, int rows, LusgsWs &w)
{
    w.rows = rows < 1 ? 1 : (rows > BAND_LANES ? BAND_LANES : rows);
    w.nband = (km + w.rows - 1) / w.rows;
    const int need = (w.nband - 1) * w.rows + BAND_LANES;   // every lane's row index fits
    w.ks = (need + 63) / 64 * 64;
    w.H = jm + km - 2;
    w.ns = (long)(w.H + 1) * w.ks;
    w.nedge = 2L * w.nband * jm * 4;
}

long lusgs_ws_lhs_bytes(const LusgsWs &w, int neqs, int lt_bytes)
{
    return 5L * neqs * neqs * w.ns * lt_bytes;
}

template <typename LT, int NEQ>
static void launch_lusgs(hipStream_t sm, const Dims &d, const LusgsWs &w, const WakePairs &wk,
                         const double *q, double *dq, double *rhs, const LT *btc, const LT *bjm,
                         const LT *bjp, const LT *bkm, const LT *bkp, LT *fac, bool rhs_out)
{
    dim3 tg((d.jm + TJ - 1) / TJ, (d.km + TK - 1) / TK);
    hipLaunchKernelGGL((k_shear<LT, NEQ>), tg, dim3(TILE_THREADS), 0, sm, d, w, rhs, btc, bjm,
                       bkm, bjp, bkp, fac);
    constexpr int PF = NEQ == 3 ? 3 : 6;
    hipLaunchKernelGGL((k_fwd_band<LT, NEQ, PF>), dim3(w.nband), dim3(BAND_LANES), 0, sm, d, w);
    long nb = (w.ns + 255) / 256;
    hipLaunchKernelGGL((k_mid<LT, NEQ>), dim3((unsigned)(nb > 4096 ? 4096 : nb)), dim3(256), 0, sm,
                       d, w);
    hipLaunchKernelGGL((k_bwd_band<LT, NEQ, PF>), dim3(w.nband), dim3(BAND_LANES), 0, sm, d, w);
    hipLaunchKernelGGL((k_unshear<NEQ>), tg, dim3(TILE_THREADS), 0, sm, d, w, w.dq, dq);
    if (rhs_out)
        hipLaunchKernelGGL((k_unshear<NEQ>), tg, dim3(TILE_THREADS), 0, sm, d, w, w.r, rhs);
    launch_wake(sm, d, NEQ, wk, q, dq, rhs);
}

template <typename LT>
void launch_lusgs3(hipStream_t sm, const Dims &d, const LusgsWs &w, const WakePairs &wk,
                   const double *q, double *dq, double *rhs, const LT *btc, const LT *bjm,
                   const LT *bjp, const LT *bkm, const LT *bkp, LT *fac, bool rhs_out)
{
    launch_lusgs<LT, 3>(sm, d, w, wk, q, dq, rhs, btc, bjm, bjp, bkm, bkp, fac, rhs_out);
}

void launch_lusgs1(hipStream_t sm, const Dims &d, const LusgsWs &w, const WakePairs &wk,
                   const double *q, double *dq, double *rhs, const double *btc, const double *bjm,
                   const double *bjp, const double *bkm, const double *bkp, double *fac,
                   bool rhs_out)
{
    launch_lusgs<double, 1>(sm, d, w, wk, q, dq, rhs, btc, bjm, bjp, bkm, bkp, fac, rhs_out);
}

template void launch_lusgs3<double>(hipStream_t, const Dims &, const LusgsWs &, const WakePairs &,
                                    const double *, double *, double *, const double *,
                                    const double *, const double *, const double *,
                                    const double *, double *, bool);

}  // namespace smac
