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
//     tiles, into the sheared layout of the workspace: point (j,k) sits at position
//     (h + HP0)*ks + k, with HP0 empty hyperplanes on either side so that a sweep can run past
//     its ends without bounds tests.  Planes are stored in pairs (16 bytes per position) so
//     that one load fetches two of them.  The diagonal factors (lusini3 / lusini1) are formed
//     on the way (fully parallel).
//  2. k_band (forward): the rows are cut into bands of `rows` <= 64 rows, one wavefront each.
//     Lane l of band b owns row k = b*rows + l and walks h = j+k upward: (j-1,k) is its own
//     result of the previous step (registers), (j,k-1) is lane l-1's result of the previous
//     step (a DPP wave shift), so a step has no barrier and no LDS traffic.  Operands are
//     fetched with buffer loads (per-lane offset in one VGPR, plane pair in SGPRs) PF steps
//     ahead.  The last row of band b hands its results to band b+1 through an edge buffer in
//     HBM whose slots start as a signalling-NaN sentinel (no arithmetic produces it); the
//     consumer spins on a relaxed agent-scope load until the slot is filled, with a bounded
//     spin count that raises a stall flag instead of hanging.
//  3. k_mid: the D multiply of the middle step (pointwise, fully parallel).
//  4. k_band (backward): the mirror image, h downward, bands handing their first row down.
//  5. k_unshear: dq (and, for callers that want it, the middle-step rhs) back to the natural
//     layout; then the wake-cut rows (bcimpds) as in the reference.
// Every point performs the reference's expressions in the reference's order; terms the
// reference does not form at the first/last row or column are dropped by selects.
#include "smac_dev.h"
#include "smac_kernels.h"

namespace smac {

constexpr unsigned long long EDGE_EMPTY = 0x7ff4dead0000beefull;   // signalling NaN

// A slot word counts as not yet written when its high half is the sentinel's: a hand-over
// value is finite, so its high half never is (0x7ff4dead is a NaN exponent), and a word read
// while it was being written, with the sentinel's high half and the value's low half, is
// then waited for and read again instead of taken as a NaN.
SMAC_DEV bool edge_empty(unsigned long long v)
{
    return (unsigned)(v >> 32) == (unsigned)(EDGE_EMPTY >> 32);
}

constexpr int EDGE_SPIN_MAX = 200000;
constexpr int BAND_LANES = 64;
constexpr int HP0 = 16;          // empty hyperplanes before h = 0 and after h = H
constexpr int PF_MAX = 8;        // 2*PF_MAX - 1 <= HP0
constexpr int EPAD = 128;        // padding slots on either side of a band's edge row

// position of point (j,k) in the sheared planes
SMAC_DEV long SX(const LusgsWs &w, int j, int k) { return (long)(j + k + HP0) * w.ks + k; }

// Planes per 16-byte group for the 3x3 system (2 doubles or 4 floats: one load fetches a
// group), 1 for the scalar one; element index of plane m at position pos.
template <int NEQ, typename T> constexpr int pair_width() { return NEQ == 3 ? 16 / (int)sizeof(T) : 1; }
template <int NEQ, typename T> constexpr int pairs_of(int planes)
{
    return (planes + pair_width<NEQ, T>() - 1) / pair_width<NEQ, T>();
}
template <int NEQ, typename T>
SMAC_DEV long EI(long ns, int m, long pos)
{
    constexpr int PW = pair_width<NEQ, T>();
    return ((long)(m / PW) * ns + pos) * PW + (m % PW);
}

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

// FP32-LHS form of the 3x3 substitution (fused multiply-adds; not the reference's rounding)
SMAC_DEV void lu_solve3_fma(const float f[9], float r1, float r2, float r3, float o[3])
{
    const float z1 = r1 * f[0];
    const float z2 = __builtin_fmaf(-f[1], z1, r2) * f[4];
    const float z3 = __builtin_fmaf(-f[5], z2, __builtin_fmaf(-f[2], z1, r3)) * f[8];
    o[2] = z3;
    o[1] = __builtin_fmaf(-f[7], o[2], z2);
    o[0] = __builtin_fmaf(-f[3], o[1], __builtin_fmaf(-f[6], o[2], z1));
}

// ------------------------------------------------------------------------------------------
// natural <-> sheared transposition through LDS tiles of TK rows x TJ columns

constexpr int TJ = 32, TK = 16, TPAD = TJ + 2, TILE_THREADS = 256;

template <typename LT, int NEQ, int NP>
SMAC_DEV void tile_to_sheared(const LusgsWs &w, int jm, int km, int j0, int k0,
                              double (*lds)[TK][TPAD], LT *dst)
{
    for (int e = threadIdx.x; e < (TJ + TK - 1) * TK; e += TILE_THREADS) {
        const int kk = e % TK, hh = e / TK, jj = hh - kk;
        const int j = j0 + jj, k = k0 + kk;
        if (jj < 0 || jj >= TJ || j >= jm || k >= km) continue;
        const long o = SX(w, j, k);
#pragma unroll
        for (int m = 0; m < NP; ++m) dst[EI<NEQ, LT>(w.ns, m, o)] = (LT)lds[m][kk][jj];
    }
}

// one LDS plane of the tile into plane m of the sheared block at dst
template <typename LT, int NEQ>
SMAC_DEV void tile_plane_to_sheared(const LusgsWs &w, int jm, int km, int j0, int k0,
                                    const double (*plane)[TPAD], LT *dst, int m)
{
    for (int e = threadIdx.x; e < (TJ + TK - 1) * TK; e += TILE_THREADS) {
        const int kk = e % TK, hh = e / TK, jj = hh - kk;
        const int j = j0 + jj, k = k0 + kk;
        if (jj < 0 || jj >= TJ || j >= jm || k >= km) continue;
        dst[EI<NEQ, LT>(w.ns, m, SX(w, j, k))] = (LT)plane[kk][jj];
    }
}

// FP64 3x3 blocks fill 5 groups of 2, the last one half: its spare slot (plane 9) carries
// what the LDS-ring sweeps need beside the block's ninth plane, so that a step reads 15 full
// groups instead of 17 part-filled ones: f[8] beside bjm[8] and bjp[8], the rhs's third plane
// beside bkm[8] (forward), the middle step's beside bkp[8] (backward).  The factors' fifth
// group and the rhs's second stay written for the other readers.
constexpr int SPARE = 9;

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
SMAC_DEV void b_shear(const Blk &bk, const Dims &d, const LusgsWs &w,
                      const double *__restrict__ rhs, const LT *__restrict__ btc,
                      const LT *__restrict__ bjm, const LT *__restrict__ bkm,
                      const LT *__restrict__ bjp, const LT *__restrict__ bkp, LT *__restrict__ fac)
{
    constexpr int NB = NEQ * NEQ, PW = pair_width<NEQ, LT>();
    __shared__ double lds[NB][TK][TPAD];
    const int j0 = bk.x * TJ, k0 = bk.y * TK;
    LT *wl = (LT *)w.lhs;
    const long gstride = (long)pairs_of<NEQ, LT>(NB) * w.ns * PW;   // elements per block group

    {
        // hand-over slots of both sweeps: the jm real ones become the sentinel, the padding
        // on either side stays a finite value (the dummy region after them is not touched)
        const long nthr = (long)bk.nx * bk.ny * TILE_THREADS;
        const long t0 = ((long)bk.y * bk.nx + bk.x) * TILE_THREADS + threadIdx.x;
        const long nreal = 2L * w.nband * w.epitch;
        for (long i = t0; i < nreal; i += nthr) {
            const long pos = (i % w.epitch) / 4;
            w.edge[i] = (pos >= EPAD && pos < EPAD + d.jm) ? EDGE_EMPTY : 0ull;
        }
    }

    // The FP32-LHS sweeps use the scaled form of the same operator: with D = LU(btc),
    // (D+L) y = r is y = D^-1 r - (D^-1 L) y, and (D+U) x = D y is x = y - (D^-1 U) x, so the
    // four neighbour blocks and the rhs are stored premultiplied by D^-1 (three substitutions
    // per block), and neither the sweeps nor the middle step need the factors.
    constexpr bool SCALED = NEQ == 3 && sizeof(LT) == 4;
    constexpr int NS = TK * TJ / TILE_THREADS;   // points of the tile per thread
    static_assert(TK * TJ % TILE_THREADS == 0, "tile points per thread");
    float fs[SCALED ? NS : 1][9];

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
        if constexpr (SCALED) {
#pragma unroll
            for (int m = 0; m < 9; ++m) fs[(e - (int)threadIdx.x) / TILE_THREADS][m] = (float)f[m];
        }
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            lds[m][kk][jj] = (double)f[m];
            if (fac) fac[m * d.N + p] = f[m];
        }
    }
    __syncthreads();
    if constexpr (!SCALED) tile_to_sheared<LT, NEQ, NB>(w, d.jm, d.km, j0, k0, lds, wl);
    constexpr bool PACK = NEQ == 3 && sizeof(LT) == 8;
    if constexpr (PACK) {
        tile_plane_to_sheared<LT, NEQ>(w, d.jm, d.km, j0, k0, lds[8], wl + 1 * gstride, SPARE);
        tile_plane_to_sheared<LT, NEQ>(w, d.jm, d.km, j0, k0, lds[8], wl + 3 * gstride, SPARE);
    }

    // scaled form: this thread's points of the tile in LDS, NP planes (blocks: 3 columns)
    auto scale = [&](int np) {
        for (int e = threadIdx.x; e < TK * TJ; e += TILE_THREADS) {
            const int kk = e / TJ, jj = e % TJ, j = j0 + jj, k = k0 + kk;
            if (j >= d.jm || k >= d.km) continue;
            const float *f = fs[(e - (int)threadIdx.x) / TILE_THREADS];
            for (int c = 0; c < np / 3; ++c) {
                float o[3];
                lu_solve3_fma(f, (float)lds[3 * c][kk][jj], (float)lds[3 * c + 1][kk][jj],
                              (float)lds[3 * c + 2][kk][jj], o);
                lds[3 * c][kk][jj] = (double)o[0];
                lds[3 * c + 1][kk][jj] = (double)o[1];
                lds[3 * c + 2][kk][jj] = (double)o[2];
            }
        }
    };

    const LT *src[4] = {bjm, bkm, bjp, bkp};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        __syncthreads();
        tile_from_natural<LT, NB>(d, j0, k0, lds, src[g]);
        __syncthreads();
        if constexpr (SCALED) {
            scale(9);
            __syncthreads();
        }
        tile_to_sheared<LT, NEQ, NB>(w, d.jm, d.km, j0, k0, lds, wl + (g + 1) * gstride);
    }
    __syncthreads();
    tile_from_natural<double, NEQ>(d, j0, k0, lds, rhs);
    __syncthreads();
    if constexpr (SCALED) {
        scale(3);
        __syncthreads();
    }
    tile_to_sheared<LT, NEQ, NEQ>(w, d.jm, d.km, j0, k0, lds, (LT *)w.r);   // the solve's type
    if constexpr (PACK)
        tile_plane_to_sheared<LT, NEQ>(w, d.jm, d.km, j0, k0, lds[2], wl + 2 * gstride, SPARE);
}

template <typename LT, int NEQ>
__global__ void __launch_bounds__(TILE_THREADS)
k_shear(Dims d, LusgsWs w, const double *__restrict__ rhs, const LT *__restrict__ btc,
        const LT *__restrict__ bjm, const LT *__restrict__ bkm, const LT *__restrict__ bjp,
        const LT *__restrict__ bkp, LT *__restrict__ fac)
{
    b_shear<LT, NEQ>(xcd_blk(), d, w, rhs, btc, bjm, bkm, bjp, bkp, fac);
}

// sheared -> natural (FP64) for the NEQ planes of a vector (dq, or the middle-step rhs)
template <int NEQ, typename T>
SMAC_DEV void b_unshear(const Blk &bk, const Dims &d, const LusgsWs &w, const T *__restrict__ src,
                        double *__restrict__ dst)
{
    constexpr int NP = NEQ;
    __shared__ double lds[NP][TK][TPAD];
    const int j0 = bk.x * TJ, k0 = bk.y * TK;
    for (int e = threadIdx.x; e < (TJ + TK - 1) * TK; e += TILE_THREADS) {
        const int kk = e % TK, hh = e / TK, jj = hh - kk;
        const int j = j0 + jj, k = k0 + kk;
        if (jj < 0 || jj >= TJ || j >= d.jm || k >= d.km) continue;
        const long o = SX(w, j, k);
#pragma unroll
        for (int m = 0; m < NP; ++m) lds[m][kk][jj] = (double)src[EI<NEQ, T>(w.ns, m, o)];
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

template <int NEQ, typename T>
__global__ void __launch_bounds__(TILE_THREADS)
k_unshear(Dims d, LusgsWs w, const T *__restrict__ src, double *__restrict__ dst)
{
    b_unshear<NEQ, T>(xcd_blk(), d, w, src, dst);
}

// ------------------------------------------------------------------------------------------
// middle step: rhs <- D rhs, pointwise (ref/ilusol.f:307-341 for neqs = 3; ref/ilusol.f:548-556
// for neqs = 1, where D = 1/b)

template <typename LT, int NEQ>
SMAC_DEV void b_mid(const Blk &bk, const Dims &d, const LusgsWs &w)
{
    constexpr int NB = NEQ * NEQ;
    const LT *F = (const LT *)w.lhs;
    LT *R = (LT *)w.r;
    for (long i = (long)bk.x * blockDim.x + threadIdx.x; i < w.ns;
         i += (long)bk.nx * blockDim.x) {
        const int h = (int)(i / w.ks) - HP0, k = (int)(i % w.ks), j = h - k;
        if (k >= d.km || j < 0 || j >= d.jm) continue;
        if constexpr (NEQ == 3) {
            LT f[NB];
#pragma unroll
            for (int m = 0; m < NB; ++m) f[m] = F[EI<NEQ, LT>(w.ns, m, i)];
            const long i0 = EI<NEQ, LT>(w.ns, 0, i), i1 = EI<NEQ, LT>(w.ns, 1, i),
                       i2 = EI<NEQ, LT>(w.ns, 2, i);
            const LT o0 = R[i0], o1 = R[i1], o2 = R[i2];
            const LT one = (LT)1.;
            LT b11 = one / f[0];
            LT b12 = f[3] * b11;
            LT b13 = f[6] * b11;
            LT b21 = f[1];
            LT b22 = one / f[4] + f[1] * f[3];
            LT b23 = f[7] / f[4] + f[1] * f[6];
            LT b31 = f[2];
            LT b32 = f[5] + f[2] * f[3];
            LT b33 = one / f[8] + f[2] * f[6] + f[5] * f[7];
            R[i0] = b11 * o0 + b12 * o1 + b13 * o2;
            R[i1] = b21 * o0 + b22 * o1 + b23 * o2;
            const LT r2 = b31 * o0 + b32 * o1 + b33 * o2;
            R[i2] = r2;
            if constexpr (sizeof(LT) == 8) {
                constexpr int PW = pair_width<NEQ, LT>();
                const long gstride = (long)pairs_of<NEQ, LT>(NB) * w.ns * PW;
                ((LT *)w.lhs)[4 * gstride + EI<NEQ, LT>(w.ns, SPARE, i)] = r2;   // beside bkp[8]
            }
        } else {
            R[i] = R[i] / F[i];
        }
    }
}

template <typename LT, int NEQ>
__global__ void k_mid(Dims d, LusgsWs w)
{
    b_mid<LT, NEQ>(hw_blk(), d, w);
}

// ------------------------------------------------------------------------------------------
// band sweeps

using rsrc_t = __amdgpu_buffer_rsrc_t;
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));
struct D2 {
    double x, y;
};

// raw buffer over `bytes` bytes (gfx9 resource word 3); accesses beyond it read 0 and drop
SMAC_DEV rsrc_t mk_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                             0x00020000);
}
SMAC_DEV D2 bld2(rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(D2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)vo, (int)so, 0));
}
SMAC_DEV double bld1(rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0));
}
SMAC_DEV void bst2(double x, double y, rsrc_t r, unsigned vo, unsigned so)
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, (D2{x, y})), r, (int)vo,
                                           (int)so, 0);
}
SMAC_DEV void bst1(double x, rsrc_t r, unsigned vo, unsigned so)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, x), r, (int)vo, (int)so, 0);
}
struct F4 {
    float x, y, z, w;
};
SMAC_DEV F4 bld4f(rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(F4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)vo, (int)so, 0));
}
SMAC_DEV float bld1f(rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0));
}
SMAC_DEV void bst4f(F4 v, rsrc_t r, unsigned vo, unsigned so)
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, v), r, (int)vo, (int)so, 0);
}

// NP planes of one group at the current position (vo), group p of planes at soffset p*pstride
template <int NEQ, int NP>
SMAC_DEV void ld_group(float *out, rsrc_t r, unsigned vo, unsigned pstride)
{
    static_assert(NEQ == 3, "the scalar sweep is FP64");
#pragma unroll
    for (int p = 0; p < NP / 4; ++p) {
        F4 t = bld4f(r, vo, p * pstride);
        out[4 * p] = t.x;
        out[4 * p + 1] = t.y;
        out[4 * p + 2] = t.z;
        out[4 * p + 3] = t.w;
    }
    if constexpr (NP % 4 == 1) {
        out[NP - 1] = bld1f(r, vo, (NP / 4) * pstride);
    } else if constexpr (NP % 4 > 1) {
        F4 t = bld4f(r, vo, (NP / 4) * pstride);   // the group's unused lanes are padding
        const float v[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int m = 0; m < NP % 4; ++m) out[(NP / 4) * 4 + m] = v[m];
    }
}

template <int NEQ, int NP>
SMAC_DEV void ld_group(double *out, rsrc_t r, unsigned vo, unsigned pstride)
{
    if constexpr (pair_width<NEQ, double>() == 2) {
#pragma unroll
        for (int p = 0; p < NP / 2; ++p) {
            D2 t = bld2(r, vo, p * pstride);
            out[2 * p] = t.x;
            out[2 * p + 1] = t.y;
        }
        if constexpr (NP % 2) out[NP - 1] = bld1(r, vo, (NP / 2) * pstride);
    } else {
#pragma unroll
        for (int p = 0; p < NP; ++p) out[p] = bld1(r, vo, p * pstride);
    }
}

template <int NEQ>
SMAC_DEV void st_vec(const float *v, rsrc_t r, unsigned vo, unsigned)
{
    static_assert(NEQ == 3, "the scalar sweep is FP64");
    bst4f(F4{v[0], v[1], v[2], 0.f}, r, vo, 0);
}

template <int NEQ>
SMAC_DEV void st_vec(const double *v, rsrc_t r, unsigned vo, unsigned pstride)
{
    if constexpr (NEQ == 3) {
        bst2(v[0], v[1], r, vo, 0);
        bst1(v[2], r, vo, pstride);
    } else {
        bst1(v[0], r, vo, 0);
    }
}

// Hand-over slots are read and written with the agent-coherent cache policy (sc1), the form
// the compiler gives relaxed agent-scope atomics; every slot is an aligned 8-byte word.
constexpr int EDGE_POLICY = 0x10;

SMAC_DEV unsigned long long edge_ld(rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(unsigned long long,
                              __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, EDGE_POLICY));
}

// Slots are written one 8-byte word per store.  A buffer store of more than 8 bytes whose
// offset comes from an SGPR can read its data registers after the next instruction has
// issued, and the compiler does not keep that instruction off them: with a 16-byte store of
// words 0-1 followed by the FP32-to-FP64 conversion of word 2 into the same registers, the
// slot sometimes received word 2 in place of word 0 (seen a few times in a thousand runs of
// the FP32 sweeps, never in the FP64 ones, whose data registers are not reused there).
SMAC_DEV void edge_st(double v, rsrc_t r, unsigned vo, unsigned so)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, (int)vo, (int)so,
                                          EDGE_POLICY);
}



// Re-reads one hand-over slot and waits for it.  Written as one asm block so that the
// compiler's wait-count bookkeeping never sees a load on the (rare) polling path: a load
// there, with the full wait it needs, would otherwise degrade the waits the prefetch ring of
// the sweep relies on.  Same instruction and cache policy as a relaxed agent-scope atomic load.
SMAC_DEV unsigned long long edge_reload(const unsigned long long *p)
{
    unsigned long long v;
    asm volatile("global_load_dwordx2 %0, %1, off sc1\n\ts_waitcnt vmcnt(0)"
                 : "=v"(v)
                 : "v"(p)
                 : "memory");
    return v;
}

// Waits (bounded) until the first word of the slot at p is filled.  Every slot up to the
// producer band's last step is written and the padding after it is finite, so the wait ends.
SMAC_DEV void edge_lag(const unsigned long long *p, bool &broken, int *stall)
{
    int n = 0;
    while (edge_empty(edge_reload(p))) {
        if (++n > EDGE_SPIN_MAX) {
            broken = true;
            atomicOr(stall, 1);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
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
        for (int m = 0; m < NEQ; ++m) ok = ok && !edge_empty(e[m]);
        if (ok) return;
        if (++n > EDGE_SPIN_MAX) {
            broken = true;
            atomicOr(stall, 1);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
#pragma unroll
        for (int m = 0; m < NEQ; ++m) e[m] = edge_reload(p + m);
    }
}

// lane l receives lane l-1's value (FWD) or lane l+1's (backward); the end lane receives 0
template <bool FWD>
SMAC_DEV float wave_shift(float v)
{
    constexpr int ctrl = FWD ? 0x138 : 0x130;   // DPP wave_shr:1 / wave_shl:1, bound_ctrl
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), ctrl,
                                                              0xf, 0xf, true));
}

template <bool FWD>
SMAC_DEV double wave_shift(double v)
{
    constexpr int ctrl = FWD ? 0x138 : 0x130;   // DPP wave_shr:1 / wave_shl:1, bound_ctrl
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, ctrl, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), ctrl, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}

// As wave_shift, but the end lane (the one without a source) receives `in`: the incoming
// hand-over value of a band's boundary row, which is that lane when the band has one.
template <bool FWD>
SMAC_DEV float wave_shift_in(float v, float in)
{
    constexpr int ctrl = FWD ? 0x138 : 0x130;
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, in),
                                                                 __builtin_bit_cast(int, v), ctrl,
                                                                 0xf, 0xf, false));
}

template <bool FWD>
SMAC_DEV double wave_shift_in(double v, double in)
{
    constexpr int ctrl = FWD ? 0x138 : 0x130;
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned long long w = (unsigned long long)__double_as_longlong(in);
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)w, (int)(unsigned)u, ctrl, 0xf, 0xf,
                                               false);
    const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(w >> 32), (int)(unsigned)(u >> 32),
                                               ctrl, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}

// Forces x to be computed where it stands: a term that is formed and then selected stays a
// select instead of being sunk into a divergent branch around its computation.
template <typename T>
SMAC_DEV T keep(T x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// ---- LDS ring for the 3x3 sweeps ----------------------------------------------------------
// The band's operands can be streamed into an LDS ring by RING_LOADERS loader wavefronts of
// the same workgroup, RING_D steps ahead of the computing wavefront, with LDS DMA
// (buffer_load ... lds): the memory-level parallelism of three wavefronts instead of one, and
// no operand registers in flight.  Each loader issues every RING_LOADERS-th step and keeps
// RING_M of its steps in flight beyond the two the computing wave is about to read, so one
// constant vmcnt wait per step suffices; one s_barrier per step paces the waves.  A slot holds
// the step's operand groups (64 lanes x 16 bytes each) and the 32-byte hand-over slot.
// The computing wave holds RING_PF steps' operands in registers: at the start of step h it
// passes the barrier and issues the LDS reads of step h + 2, which then run under step h's
// arithmetic (the reads are 17 instructions, more than the LDS counter can track, so issuing
// them costs the wave a wait for the first ones; here that wait overlaps the arithmetic).
constexpr int RING_LOADERS = 3;
constexpr int RING_PF = 3;
constexpr int RING_GROUP = BAND_LANES * 16;
// FP32-LHS sweeps run in the scaled form (the blocks and the rhs premultiplied by the inverse
// diagonal): they read no diagonal factors
template <typename T> constexpr bool scaled_form() { return sizeof(T) == 4; }
// block groups a step reads: factors (FP64 only), j-neighbour and k-neighbour blocks
template <typename T> constexpr int ring_blocks() { return scaled_form<T>() ? 2 : 3; }
// FP64: the packed step of 15 full groups (factors 0-3, j-block 0-4, k-block 0-4, rhs 0)
template <typename T> constexpr bool ring_packed() { return sizeof(T) == 8; }
template <typename T> constexpr int ring_groups()
{
    return ring_packed<T>() ? 4 + 5 + 5 + 1 : ring_blocks<T>() * pairs_of<3, T>(9) + pairs_of<3, T>(3);
}
template <typename T> constexpr int ring_m() { return sizeof(T) == 8 ? 2 : 3; }
template <typename T> constexpr int ring_d() { return RING_LOADERS * ring_m<T>() + 2; }
template <typename T> constexpr int ring_slots() { return ring_d<T>() + 1; }
template <typename T> constexpr int ring_slot_bytes() { return ring_groups<T>() * RING_GROUP + 32; }
template <typename T> constexpr int ring_bytes() { return ring_slots<T>() * ring_slot_bytes<T>(); }
// VMEM instructions per issued step: the groups and two for the hand-over slot
template <typename T> constexpr int ring_vm_per_step() { return ring_groups<T>() + 2; }
static_assert((ring_m<double>() + 1) * ring_vm_per_step<double>() <= 63, "vmcnt budget");
static_assert((ring_m<float>() + 1) * ring_vm_per_step<float>() <= 63, "vmcnt budget");
// the loaders run ring_d steps past the last chunk, which ends up to RING_PF - 1 steps past the band
static_assert(ring_d<float>() + RING_PF <= HP0 && ring_d<double>() + RING_PF <= HP0,
              "ring runs past padding");
static_assert(ring_bytes<double>() <= 160 * 1024 && ring_bytes<float>() <= 160 * 1024, "LDS");

// gfx9 s_waitcnt immediate: wait until at most n vector-memory operations are outstanding
constexpr int vmcnt_imm(int n) { return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14); }
// ... and until no LDS read is outstanding.  Issued once before each of a sweep's step loops:
// the wait analysis merges a loop's entry state into every iteration, and reads left in flight
// by the code before the loop would otherwise cost a full LDS wait in every iteration.
constexpr int LGKM_DRAIN = 15 | (7 << 4) | (0 << 8) | (3 << 14);

template <int AUX>
SMAC_DEV void ring_dma(rsrc_t r, unsigned char *lds, unsigned vo, unsigned so)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void *)lds, 16,
                                             (int)vo, (int)so, 0, AUX);
}

SMAC_DEV void ring_barrier() { asm volatile("s_barrier" ::: "memory"); }

// Per-step operands of one lane, loaded PF steps ahead.
template <int NEQ, typename T>
struct Slot {
    T f[NEQ * NEQ], jn[NEQ * NEQ], kn[NEQ * NEQ], r[NEQ];
    unsigned long long e[NEQ];
};

// One wavefront band of one sweep of ilusol3 / ilusol1.  FWD: the forward part
// (ref/ilusol.f:127-202 and 502-540) from the rhs planes into the rhs planes.  Backward:
// ref/ilusol.f:353-469 and 560-591, from the middle-step rhs into the dq planes.
// j-neighbour block: bjm (FWD) / bjp; k-neighbour block: bkm (FWD) / bkp.
// T = double: the reference's arithmetic.  T = float (3x3 only): the FP32-LHS mode.
template <int NEQ, int PF, bool FWD, typename T, bool RING = false>
struct Band {
    static constexpr int NB = NEQ * NEQ;
    Dims d;
    const unsigned char *ring;   // RING: the LDS ring (this lane's 16 bytes at lane * 16)
    unsigned rnext;              // RING: byte offset of the slot the next refill reads
    int k, kin, kout;
    bool lane_in, kterm, broken;
    unsigned vout;      // producer lane: 0; others: out of range, so their stores drop
    const unsigned long long *ein_ptr;
    int *stall;
    rsrc_t rF, rJ, rK, rR, rO, rEin, rEout;
    unsigned pstride;
    int dstep;
    // running state
    T own[NEQ];
    Slot<NEQ, T> s[PF];
    int h;
    unsigned vo, vp;       // byte offset of this lane's position: current / prefetch step
    unsigned ein_pf, eout; // soffset of the hand-over slots: prefetch step / current step

    // RING: the first n elements of this lane's 16 bytes of group q.  A full group is one
    // ds_read_b128; a partial one (a block's last group, the rhs group) reads only up to its
    // last element, so no register receives a padding element: a register whose loaded value
    // is dead gets reused at once, and writing it makes the wave wait for every LDS read in
    // flight (a write-after-write hazard), twice a step.
    template <int N>
    SMAC_DEV void ring_elems(const unsigned char *b, int q, T *dst)
    {
        constexpr int PW = pair_width<3, T>();
        const unsigned char *a = b + q * RING_GROUP;
        if constexpr (N == PW) {
            typedef T vec_t __attribute__((ext_vector_type(PW)));
            const vec_t v = *(const vec_t *)a;
#pragma unroll
            for (int e = 0; e < PW; ++e) dst[e] = v[e];
        } else {
            static_assert(N >= 1 && N < PW, "partial group");
            int e = 0;
            if constexpr (N >= 2) {
                typedef T vec2_t __attribute__((ext_vector_type(2)));
                const vec2_t v = *(const vec2_t *)a;
                dst[0] = v[0];
                dst[1] = v[1];
                e = 2;
            }
            if constexpr (N % 2 == 1) dst[e] = *(const T *)(a + e * sizeof(T));
        }
    }

    // RING: one step's operands from LDS (the hand-over slot is read as a broadcast)
    SMAC_DEV void ring_read(Slot<NEQ, T> &sl, unsigned off)
    {
        static_assert(NEQ == 3, "the ring serves the 3x3 sweeps");
        constexpr int GB = pairs_of<3, T>(9), GR = pairs_of<3, T>(3), PW = pair_width<3, T>();
        // reads from one base with immediate offsets
        const unsigned char *b = ring + off + (threadIdx.x & 63) * 16;
        if constexpr (ring_packed<T>()) {
            // factors 0-7, j-block 0-7, (j8, f8), k-block 0-7, (k8, r2), (r0, r1)
#pragma unroll
            for (int p = 0; p < 4; ++p) ring_elems<2>(b, p, sl.f + 2 * p);
#pragma unroll
            for (int p = 0; p < 4; ++p) ring_elems<2>(b, 4 + p, sl.jn + 2 * p);
            T t[2];
            ring_elems<2>(b, 8, t);
            sl.jn[8] = t[0];
            sl.f[8] = t[1];
#pragma unroll
            for (int p = 0; p < 4; ++p) ring_elems<2>(b, 9 + p, sl.kn + 2 * p);
            ring_elems<2>(b, 13, t);
            sl.kn[8] = t[0];
            sl.r[2] = t[1];
            ring_elems<2>(b, 14, sl.r);
        } else {
            constexpr int NBK = ring_blocks<T>();
            T *dst[3] = {sl.f, sl.jn, sl.kn};
            constexpr int LB = 9 - (GB - 1) * PW, LR = 3 - (GR - 1) * PW;   // last groups' elements
#pragma unroll
            for (int g = 0; g < NBK; ++g) {
#pragma unroll
                for (int p = 0; p < GB - 1; ++p) ring_elems<PW>(b, g * GB + p, dst[g + 3 - NBK] + p * PW);
                ring_elems<LB>(b, g * GB + GB - 1, dst[g + 3 - NBK] + (GB - 1) * PW);
            }
#pragma unroll
            for (int p = 0; p < GR - 1; ++p) ring_elems<PW>(b, NBK * GB + p, sl.r + p * PW);
            ring_elems<LR>(b, NBK * GB + GR - 1, sl.r + (GR - 1) * PW);
        }
        const unsigned long long *e = (const unsigned long long *)(ring + off + ring_groups<T>() * RING_GROUP);
        sl.e[0] = e[0];
        sl.e[1] = e[1];
        sl.e[2] = e[2];
    }

    SMAC_DEV void load(Slot<NEQ, T> &sl, unsigned v, unsigned so)
    {
        if constexpr (!(NEQ == 3 && scaled_form<T>())) ld_group<NEQ, NB>(sl.f, rF, v, pstride);
        ld_group<NEQ, NB>(sl.jn, rJ, v, pstride);
        ld_group<NEQ, NB>(sl.kn, rK, v, pstride);
        ld_group<NEQ, NEQ>(sl.r, rR, v, pstride);
        // padded: rows outside 0..jm-1 read finite padding, bands without a neighbour read
        // a region of zeros
        if constexpr (NEQ == 3) {
            // one 8-byte load per word: a wider load is not guaranteed to see each word
            // whole, and a word read half written (the sentinel's high half, a NaN, with the
            // value's low half) would pass the sentinel test
            sl.e[0] = edge_ld(rEin, 0, so);
            sl.e[1] = edge_ld(rEin, 8, so);
            sl.e[2] = edge_ld(rEin, 16, so);
        } else {
            sl.e[0] = edge_ld(rEin, 0, so);
        }
    }

    // JSEL: some lane may sit on the first (FWD) / last column, where the reference does not
    // form the j-neighbour term; KSEL: the band holds the first (FWD) / last grid row, where it
    // does not form the k-neighbour term.  Elsewhere both terms are formed unconditionally.
    // One step on the operands in c.  RING: fill receives the operands of the step two ahead;
    // unless SEL the band's incoming row is the end lane of the shift, which then takes the
    // hand-over value directly instead of through a select.
    template <bool JSEL, bool KSEL, bool SEL = true>
    SMAC_DEV void step(Slot<NEQ, T> &c, Slot<NEQ, T> &fill)
    {
        if constexpr (RING) {
            // Every LDS read issued so far has landed (the slot a loader refills after this
            // barrier was read two steps ago); past the barrier the loaders have the step two
            // ahead in LDS.
            __builtin_amdgcn_s_waitcnt(LGKM_DRAIN);
            ring_barrier();
            ring_read(fill, rnext);
            rnext = rnext + ring_slot_bytes<T>() == ring_bytes<T>() ? 0u : rnext + ring_slot_bytes<T>();
        }
        T nb[NEQ];
        bool empty = false;
#pragma unroll
        for (int m = 0; m < NEQ; ++m) empty = empty || edge_empty(c.e[m]);
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(empty) != 0, 0))   // uniform; rare
            edge_wait<NEQ>(ein_ptr + (long)(h - kin + EPAD) * 4, c.e, broken, stall);
        if constexpr (RING && !SEL) {
#pragma unroll
            for (int m = 0; m < NEQ; ++m)
                nb[m] = wave_shift_in<FWD>(own[m], (T)__longlong_as_double((long long)c.e[m]));
        } else {
#pragma unroll
            for (int m = 0; m < NEQ; ++m) nb[m] = wave_shift<FWD>(own[m]);
#pragma unroll
            for (int m = 0; m < NEQ; ++m)
                nb[m] = lane_in ? (T)__longlong_as_double((long long)c.e[m]) : nb[m];
        }
        const int j = h - k;
        // FWD: no j-term at j = 1 (bjm(1,k)*rhs(jmax,k-1) is an exact zero); backward: none at
        // j = jmax (bjp(jmax,k)*dq(1,k+1), likewise)
        const bool jterm = FWD ? j > 0 : j < d.jm - 1;
        T o[NEQ];
        if constexpr (NEQ == 3 && sizeof(T) == 4) {
            T r[3] = {c.r[0], c.r[1], c.r[2]};
            auto jpart = [&]() {
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const T t = __builtin_fmaf(-c.jn[m + 6], own[2],
                                __builtin_fmaf(-c.jn[m + 3], own[1],
                                __builtin_fmaf(-c.jn[m], own[0], r[m])));
                    if constexpr (JSEL) {
                        const T tk = keep(t);
                        r[m] = jterm ? tk : r[m];
                    } else {
                        r[m] = t;
                    }
                }
            };
            auto kpart = [&]() {
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const T t = __builtin_fmaf(-c.kn[m + 6], nb[2],
                                __builtin_fmaf(-c.kn[m + 3], nb[1],
                                __builtin_fmaf(-c.kn[m], nb[0], r[m])));
                    if constexpr (KSEL) {
                        const T tk = keep(t);
                        r[m] = kterm ? tk : r[m];
                    } else {
                        r[m] = t;
                    }
                }
            };
            if constexpr (FWD) {
                jpart();
                kpart();
            } else {
                kpart();
                jpart();
            }
            // scaled form: the blocks and the rhs already carry the inverse diagonal
            o[0] = r[0];
            o[1] = r[1];
            o[2] = r[2];
        } else if constexpr (NEQ == 3) {
            double r[3] = {c.r[0], c.r[1], c.r[2]};
            auto jpart = [&]() {
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const double t = r[m] - c.jn[m] * own[0] - c.jn[m + 3] * own[1] -
                                     c.jn[m + 6] * own[2];
                    if constexpr (JSEL) {
                        const double tk = keep(t);
                        r[m] = jterm ? tk : r[m];
                    } else {
                        r[m] = t;
                    }
                }
            };
            auto kpart = [&]() {
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const double t = r[m] - c.kn[m] * nb[0] - c.kn[m + 3] * nb[1] -
                                     c.kn[m + 6] * nb[2];
                    if constexpr (KSEL) {
                        const double tk = keep(t);
                        r[m] = kterm ? tk : r[m];
                    } else {
                        r[m] = t;
                    }
                }
            };
            if constexpr (FWD) {
                jpart();
                kpart();
            } else {
                kpart();
                jpart();
            }
            lu_solve3<double>(c.f, r[0], r[1], r[2], o);
        } else {
            // f*(r - a*x - b*y) with the terms of the reference's branches (f*r == r*f)
            double t = c.r[0];
            auto jpart = [&]() {
                if constexpr (JSEL) {
                    const double u = keep(t - c.jn[0] * own[0]);
                    t = jterm ? u : t;
                } else {
                    t = t - c.jn[0] * own[0];
                }
            };
            auto kpart = [&]() {
                if constexpr (KSEL) {
                    const double u = keep(t - c.kn[0] * nb[0]);
                    t = kterm ? u : t;
                } else {
                    t = t - c.kn[0] * nb[0];
                }
            };
            if constexpr (FWD) {
                jpart();
                kpart();
            } else {
                kpart();
                jpart();
            }
            o[0] = c.f[0] * t;
        }
        st_vec<NEQ>(o, rO, vo, pstride);   // inactive positions are padding
        if constexpr (NEQ == 3) {
#pragma unroll
            for (int m = 0; m < 3; ++m) edge_st((double)o[m], rEout, vout + 8 * m, eout);
        } else {
            edge_st((double)o[0], rEout, vout, eout);
        }
#pragma unroll
        for (int m = 0; m < NEQ; ++m) own[m] = o[m];
        if constexpr (!RING) load(c, vp, ein_pf);
        constexpr int ED = FWD ? 32 : -32;
        h += FWD ? 1 : -1;
        vo += dstep;
        vp += dstep;
        ein_pf += ED;
        eout += ED;
    }

    template <bool JSEL, bool KSEL, bool SEL = true>
    SMAC_DEV void chunks(int n)
    {
        for (int i = 0; i < n; ++i) {
            if constexpr (RING) {
                static_assert(PF == 3, "three register slots: in use, next, being read");
                step<JSEL, KSEL, SEL>(s[0], s[2]);
                step<JSEL, KSEL, SEL>(s[1], s[0]);
                step<JSEL, KSEL, SEL>(s[2], s[1]);
            } else {
#pragma unroll
                for (int u = 0; u < PF; ++u) step<JSEL, KSEL>(s[u], s[u]);
            }
        }
    }
};

// One band of one case: blk is the band's place in dispatch order (forward bands bottom up,
// backward bands top down, so that every band is dispatched after the band it waits for).
template <int NEQ, int PF, bool FWD, typename T, bool RING = false>
SMAC_DEV void band_run(const Dims &d, const LusgsWs &w, int blk)
{
    constexpr int NB = NEQ * NEQ, PW = pair_width<NEQ, T>();
    const int lane = threadIdx.x & 63, nband = w.nband;
    const int b = FWD ? blk : nband - 1 - blk;
    const int k0 = b * w.rows;
    const int klast = (k0 + w.rows < d.km ? k0 + w.rows : d.km) - 1;
    const bool real = lane < w.rows;
    const int hbeg = k0, hend = d.jm - 1 + klast;
    const int nsteps = hend - hbeg + 1;
    const int dir = FWD ? 1 : -1;
    const int h0 = FWD ? hbeg : hend;

    Band<NEQ, PF, FWD, T, RING> B;
    B.d = d;
    B.k = real ? k0 + lane : w.ks - 1;          // idle lanes run on a padding row
    // hand-over rows: the consumer row takes its k-neighbour from the next band over
    const bool has_in = FWD ? b > 0 : b + 1 < nband;
    const bool has_out = FWD ? b + 1 < nband : b > 0;
    B.kin = FWD ? k0 : klast;
    B.kout = FWD ? klast : k0;
    B.lane_in = has_in && real && B.k == B.kin;
    B.vout = (has_out && real && B.k == B.kout) ? 0u : 0x80000000u;
    // edge regions: forward sweep bands 0..nband-1, backward sweep nband..2*nband-1, then a
    // region of zeros for bands with no neighbour on the consuming side
    const unsigned long long *ebase = w.edge + (FWD ? 0 : (long)nband * w.epitch);
    const unsigned long long *ein = has_in ? ebase + (long)(b - dir) * w.epitch
                                           : w.edge + 2L * nband * w.epitch;
    const unsigned ebytes = (unsigned)w.epitch * 8u;
    B.rEin = mk_rsrc(ein, ebytes);
    B.ein_ptr = ein;
    B.rEout = mk_rsrc(ebase + (long)b * w.epitch, ebytes - 8);
    B.stall = w.stall;
    B.broken = false;
    // the k-neighbour term is not formed on the first (FWD) / last grid row
    B.kterm = FWD ? B.k > 0 : B.k < d.km - 1;

    B.pstride = (unsigned)(w.ns * PW * sizeof(T));                      // bytes per plane group
    const unsigned gbytes = (unsigned)pairs_of<NEQ, T>(NB) * B.pstride;  // bytes per block group
    const unsigned vbytes = (unsigned)pairs_of<NEQ, T>(NEQ) * B.pstride;
    const char *lhs = (const char *)w.lhs;
    // The last hyperplane of every plane is never touched (sweeps reach HP0 - 1 past
    // their ends), so each range may stop a few bytes short of its group; distinct ranges
    // keep the compiler from merging descriptor words and re-assembling descriptors per load.
    B.rF = mk_rsrc(lhs, gbytes - 16);
    B.rJ = mk_rsrc(lhs + (FWD ? 1 : 3) * (size_t)gbytes, gbytes - 32);
    B.rK = mk_rsrc(lhs + (FWD ? 2 : 4) * (size_t)gbytes, gbytes - 48);
    B.rR = mk_rsrc(w.r, vbytes - 64);
    B.rO = FWD ? B.rR : mk_rsrc(w.dq, vbytes - 80);
    B.dstep = dir * w.ks * PW * (int)sizeof(T);
    B.vo = (unsigned)(((long)(h0 + HP0) * w.ks + B.k) * PW * sizeof(T));   // current step
    B.vp = B.vo + PF * B.dstep;                                     // prefetch step
    B.h = h0;
    B.eout = (unsigned)(h0 - B.kout + EPAD) * 32u;
    const unsigned ein0 = (unsigned)(h0 - B.kin + EPAD) * 32u;
    B.ein_pf = ein0 + PF * dir * 32;
#pragma unroll
    for (int m = 0; m < NEQ; ++m) B.own[m] = (T)0;
    // the first klast-k0+1 steps are the only ones where a lane meets the j boundary
    const int nch = (nsteps + PF - 1) / PF;
    if constexpr (RING) {
        static_assert(PF == RING_PF && NEQ == 3, "the ring feeds the register slots");
        __shared__ __attribute__((aligned(16))) unsigned char ring_lds[ring_bytes<T>()];
        const int wave = threadIdx.x >> 6;
        if (wave > 0) {
            // loader wavefront: steps s = wave - 1 (mod RING_LOADERS), RING_D ahead
            constexpr int D = ring_d<T>(), SB = ring_slot_bytes<T>(), R = ring_slots<T>();
            constexpr int GB = pairs_of<3, T>(9), GR = pairs_of<3, T>(3);
            const int wl = wave - 1;
            constexpr int NBK = ring_blocks<T>();
            const rsrc_t rs[3] = {B.rF, B.rJ, B.rK};
            auto issue = [&](int st) {
                if (st % RING_LOADERS != wl) return;
                unsigned char *sb = ring_lds + (st % R) * SB;
                const unsigned v = B.vo + (unsigned)(st * B.dstep);
                if constexpr (ring_packed<T>()) {
                    // factors' groups 0-3, both neighbour blocks' 0-4 (group 4 carries f8 / the
                    // rhs's third plane in its spare slot), the rhs's group 0
#pragma unroll
                    for (int p = 0; p < 4; ++p)
                        ring_dma<0>(B.rF, sb + p * RING_GROUP, v, p * B.pstride);
#pragma unroll
                    for (int p = 0; p < 5; ++p)
                        ring_dma<0>(B.rJ, sb + (4 + p) * RING_GROUP, v, p * B.pstride);
#pragma unroll
                    for (int p = 0; p < 5; ++p)
                        ring_dma<0>(B.rK, sb + (9 + p) * RING_GROUP, v, p * B.pstride);
                    ring_dma<0>(B.rR, sb + 14 * RING_GROUP, v, 0);
                } else {
#pragma unroll
                    for (int g = 0; g < NBK; ++g)
#pragma unroll
                        for (int p = 0; p < GB; ++p)
                            ring_dma<0>(rs[g + 3 - NBK], sb + (g * GB + p) * RING_GROUP, v,
                                        p * B.pstride);
#pragma unroll
                    for (int p = 0; p < GR; ++p)
                        ring_dma<0>(B.rR, sb + (NBK * GB + p) * RING_GROUP, v, p * B.pstride);
                }
                // the hand-over slot: lane 0 alone, so 32 bytes land at the slot's end
                const unsigned es = ein0 + (unsigned)(st * dir * 32);
                unsigned char *eb = sb + ring_groups<T>() * RING_GROUP;
                if (lane == 0) {
                    ring_dma<EDGE_POLICY>(B.rEin, eb, 0, es);
                    ring_dma<EDGE_POLICY>(B.rEin, eb + 16, 16, es);
                }
            };
            constexpr int WAIT = vmcnt_imm(ring_m<T>() * ring_vm_per_step<T>());
            ring_barrier();   // the computing wave has let its producer band get ahead
            for (int st = 0; st < D; ++st) issue(st);
            __builtin_amdgcn_s_waitcnt(WAIT);
            ring_barrier();
            const int nexec = nch * PF;
            for (int i = 0; i < nexec; ++i) {
                issue(i + D);
                __builtin_amdgcn_s_waitcnt(WAIT);
                ring_barrier();
            }
            __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
            return;
        }
        B.ring = ring_lds;
        // A band that reads a neighbour's hand-over slots starts RING_D + 4 steps behind it:
        // the loaders fetch each slot RING_D steps ahead, and a slot fetched before its
        // producer wrote it costs the computing wave a device-scope round trip to re-read.
        {
            constexpr int LAG = ring_d<T>() + 4;
            edge_lag(B.ein_ptr + (long)(h0 + dir * LAG - B.kin + EPAD) * 4, B.broken, B.stall);
        }
        ring_barrier();
        ring_barrier();
        B.ring_read(B.s[0], 0);
        B.ring_read(B.s[1], ring_slot_bytes<T>());
        B.rnext = 2 * ring_slot_bytes<T>();
    } else {
#pragma unroll
        for (int u = 0; u < PF; ++u) B.load(B.s[u], B.vo + u * B.dstep, ein0 + u * dir * 32);
    }
    int nramp = (klast - k0 + 1 + PF - 1) / PF;
    nramp = nramp < nch ? nramp : nch;
    const bool krow = FWD ? k0 == 0 : klast == d.km - 1;
    if constexpr (RING) {
        // The band's incoming row is the shift's end lane (lane 0 forward; lane 63 backward,
        // where only the head band, which has no incoming row, can be narrower than the wave)
        // unless a backward band has fewer than 64 rows: then it selects the value (SEL).  A
        // band without an incoming row takes the zeros of the dummy slots at its end lane, as
        // the plain shift does.
        const bool sel = has_in && !FWD && klast - k0 != BAND_LANES - 1;
        __builtin_amdgcn_s_waitcnt(LGKM_DRAIN);
        if (krow) {
            B.template chunks<true, true, false>(nramp);
            __builtin_amdgcn_s_waitcnt(LGKM_DRAIN);
            B.template chunks<false, true, false>(nch - nramp);
        } else if (!sel) {
            B.template chunks<true, false, false>(nramp);
            __builtin_amdgcn_s_waitcnt(LGKM_DRAIN);
            B.template chunks<false, false, false>(nch - nramp);
        } else {
            B.template chunks<true, false, true>(nramp);
            __builtin_amdgcn_s_waitcnt(LGKM_DRAIN);
            B.template chunks<false, false, true>(nch - nramp);
        }
    } else if (krow) {
        B.template chunks<true, true>(nramp);
        B.template chunks<false, true>(nch - nramp);
    } else {
        B.template chunks<true, false>(nramp);
        B.template chunks<false, false>(nch - nramp);
    }
}

template <int NEQ, int PF, bool FWD, typename T>
__global__ void __launch_bounds__(BAND_LANES)
k_band(Dims d, LusgsWs w)
{
    band_run<NEQ, PF, FWD, T>(d, w, (int)blockIdx.x);
}

// The bands of several cases in one launch: workgroup i runs band map[i].y of case map[i].x.
// The map lists each case's bands contiguously in dispatch order, so every wait is on a
// workgroup dispatched earlier.
template <int NEQ, int PF, bool FWD, typename T>
__global__ void __launch_bounds__(BAND_LANES)
k_band_multi(const BandJob *__restrict__ jobs, const int2 *__restrict__ map,
             const int *__restrict__ active)
{
    const int2 m = map[blockIdx.x];
    // a stopped case: all of its bands leave, so none waits on another
    if (active && active[jobs[m.x].bcase] == 0) return;
    const Dims d = jobs[m.x].d;
    const LusgsWs w = jobs[m.x].w;
    band_run<NEQ, PF, FWD, T>(d, w, m.y);
}

// the 3x3 sweeps fed through the LDS ring: one computing and RING_LOADERS loader wavefronts
template <bool FWD, typename T>
__global__ void __launch_bounds__(BAND_LANES * (1 + RING_LOADERS))
k_band_ring(Dims d, LusgsWs w)
{
    band_run<3, RING_PF, FWD, T, true>(d, w, (int)blockIdx.x);
}

template <bool FWD, typename T>
__global__ void __launch_bounds__(BAND_LANES * (1 + RING_LOADERS))
k_band_ring_multi(const BandJob *__restrict__ jobs, const int2 *__restrict__ map,
                  const int *__restrict__ active)
{
    const int2 m = map[blockIdx.x];
    if (active && active[jobs[m.x].bcase] == 0) return;
    const Dims d = jobs[m.x].d;
    const LusgsWs w = jobs[m.x].w;
    band_run<3, RING_PF, FWD, T, true>(d, w, m.y);
}

// ------------------------------------------------------------------------------------------

void lusgs_ws_shape(int jm, int km, int rows, LusgsWs &w)
{
    if (w.pf3 == 0) w.pf3 = 4;
    if (w.pf1 == 0) w.pf1 = 8;
    w.rows = rows < 1 ? 1 : (rows > BAND_LANES ? BAND_LANES : rows);
    w.nband = (km + w.rows - 1) / w.rows;
    // every lane's row fits, and row ks-1 is padding (idle lanes work there)
    int need = (w.nband - 1) * w.rows + BAND_LANES;
    if (need < km + 1) need = km + 1;
    w.ks = (need + 63) / 64 * 64;
    w.H = jm + km - 2;
    w.ns = (long)(w.H + 1 + 2 * HP0) * w.ks;
    w.epitch = (jm + 2 * EPAD) * 4;
    w.nedge = (2L * w.nband + 1) * w.epitch;
}

// sizes for the FP64 system (the FP32-LHS one needs less and fits the same allocation)
long lusgs_ws_lhs_bytes(const LusgsWs &w, int neqs)
{
    return neqs == 3 ? 5L * pairs_of<3, double>(9) * w.ns * 16 : 5L * w.ns * 8;
}

long lusgs_ws_vec_bytes(const LusgsWs &w, int neqs)
{
    return neqs == 3 ? (long)pairs_of<3, double>(3) * w.ns * 16 : w.ns * 8;
}

bool lusgs_ws_fits(const LusgsWs &w)
{
    // buffer offsets are 32-bit: one block group (5 plane pairs) must stay below 2 GiB
    return (long)pairs_of<3, double>(9) * w.ns * 16 < (1L << 31);
}

template <typename LT, int NEQ>
void lusgs_pre(hipStream_t sm, const Dims &d, const LusgsWs &w, const double *rhs, const LT *btc,
               const LT *bjm, const LT *bjp, const LT *bkm, const LT *bkp, LT *fac)
{
    dim3 tg((d.jm + TJ - 1) / TJ, (d.km + TK - 1) / TK);
    hipLaunchKernelGGL((k_shear<LT, NEQ>), tg, dim3(TILE_THREADS), 0, sm, d, w, rhs, btc, bjm,
                       bkm, bjp, bkp, fac);
}

#define BAND_SWITCH(PFSEL, LAUNCH)                                                             \
    do {                                                                                       \
        switch (PFSEL) {                                                                       \
        case 1: LAUNCH(1); break;                                                              \
        case 2: LAUNCH(2); break;                                                              \
        case 4: LAUNCH(4); break;                                                              \
        case 6: LAUNCH(6); break;                                                              \
        case 8: LAUNCH(8); break;                                                              \
        default: LAUNCH(3); break;                                                             \
        }                                                                                      \
    } while (0)

template <typename LT, int NEQ>
void lusgs_band(hipStream_t sm, const Dims &d, const LusgsWs &w, bool fwd)
{
    if (NEQ == 3 && w.ring) {
        const dim3 blk(BAND_LANES * (1 + RING_LOADERS));
        if (fwd)
            hipLaunchKernelGGL((k_band_ring<true, LT>), dim3(w.nband), blk, 0, sm, d, w);
        else
            hipLaunchKernelGGL((k_band_ring<false, LT>), dim3(w.nband), blk, 0, sm, d, w);
        return;
    }
    const int pf = NEQ == 3 ? w.pf3 : w.pf1;
#define ONE_F(PFV) hipLaunchKernelGGL((k_band<NEQ, PFV, true, LT>), dim3(w.nband), dim3(BAND_LANES), 0, sm, d, w)
#define ONE_B(PFV) hipLaunchKernelGGL((k_band<NEQ, PFV, false, LT>), dim3(w.nband), dim3(BAND_LANES), 0, sm, d, w)
    if (fwd)
        BAND_SWITCH(pf, ONE_F);
    else
        BAND_SWITCH(pf, ONE_B);
#undef ONE_F
#undef ONE_B
}

template <typename LT, int NEQ>
void lusgs_band_multi(hipStream_t sm, const BandJob *jobs, const int2 *map, int nblk, int pf,
                      bool fwd, const int *active)
{
    if (nblk <= 0) return;
    if (NEQ == 3 && pf == 0) {   // pf 0 selects the LDS-ring sweeps
        const dim3 blk(BAND_LANES * (1 + RING_LOADERS));
        if (fwd)
            hipLaunchKernelGGL((k_band_ring_multi<true, LT>), dim3(nblk), blk, 0, sm, jobs, map,
                               active);
        else
            hipLaunchKernelGGL((k_band_ring_multi<false, LT>), dim3(nblk), blk, 0, sm, jobs, map,
                               active);
        return;
    }
#define MUL_F(PFV) hipLaunchKernelGGL((k_band_multi<NEQ, PFV, true, LT>), dim3(nblk), dim3(BAND_LANES), 0, sm, jobs, map, active)
#define MUL_B(PFV) hipLaunchKernelGGL((k_band_multi<NEQ, PFV, false, LT>), dim3(nblk), dim3(BAND_LANES), 0, sm, jobs, map, active)
    if (fwd)
        BAND_SWITCH(pf, MUL_F);
    else
        BAND_SWITCH(pf, MUL_B);
#undef MUL_F
#undef MUL_B
}
#undef BAND_SWITCH

template <typename LT, int NEQ>
void lusgs_mid(hipStream_t sm, const Dims &d, const LusgsWs &w)
{
    if (NEQ == 3 && sizeof(LT) == 4) return;   // scaled form: D^-1 (D y) = y
    long nb = (w.ns + 255) / 256;
    hipLaunchKernelGGL((k_mid<LT, NEQ>), dim3((unsigned)(nb > 4096 ? 4096 : nb)), dim3(256), 0, sm,
                       d, w);
}

template <typename LT, int NEQ>
void lusgs_post(hipStream_t sm, const Dims &d, const LusgsWs &w, const WakePairs &wk,
                const double *q, double *dq, double *rhs, bool rhs_out)
{
    dim3 tg((d.jm + TJ - 1) / TJ, (d.km + TK - 1) / TK);
    hipLaunchKernelGGL((k_unshear<NEQ, LT>), tg, dim3(TILE_THREADS), 0, sm, d, w,
                       (const LT *)w.dq, dq);
    if (rhs_out)
        hipLaunchKernelGGL((k_unshear<NEQ, LT>), tg, dim3(TILE_THREADS), 0, sm, d, w,
                           (const LT *)w.r, rhs);
    launch_wake(sm, d, NEQ, wk, q, dq, rhs);
}

// ---- multi-case launches ------------------------------------------------------------------

template <typename LT>
ShearArgs shear_args(const Dims &d, const LusgsWs &w, const double *rhs, const LT *btc,
                     const LT *bjm, const LT *bjp, const LT *bkm, const LT *bkp, LT *fac)
{
    return ShearArgs{(unsigned)((d.jm + TJ - 1) / TJ), (unsigned)((d.km + TK - 1) / TK), d, w, rhs,
                     btc, bjm, bkm, bjp, bkp, fac};
}

template <typename LT, int NEQ>
MidArgs mid_args(const Dims &d, const LusgsWs &w)
{
    if (NEQ == 3 && sizeof(LT) == 4) return MidArgs{0u, 1u, d, w};   // scaled form
    long nb = (w.ns + 255) / 256;
    return MidArgs{(unsigned)(nb > 4096 ? 4096 : nb), 1u, d, w};
}

UnshearArgs unshear_args(const Dims &d, const LusgsWs &w, const void *src, double *dst)
{
    return UnshearArgs{(unsigned)((d.jm + TJ - 1) / TJ), (unsigned)((d.km + TK - 1) / TK), d, w,
                       src, dst};
}

template <typename LT, int NEQ>
struct CShear {
    using Args = ShearArgs;
    static constexpr int THREADS = TILE_THREADS;
    static SMAC_DEV void run(const Blk &b, const Args &a)
    {
        b_shear<LT, NEQ>(b, a.d, a.w, a.rhs, (const LT *)a.btc, (const LT *)a.bjm,
                         (const LT *)a.bkm, (const LT *)a.bjp, (const LT *)a.bkp, (LT *)a.fac);
    }
};
template <typename LT, int NEQ>
struct CMid {
    using Args = MidArgs;
    static constexpr int THREADS = 256;
    static SMAC_DEV void run(const Blk &b, const Args &a) { b_mid<LT, NEQ>(b, a.d, a.w); }
};
template <typename LT, int NEQ>
struct CUnshear {
    using Args = UnshearArgs;
    static constexpr int THREADS = TILE_THREADS;
    static SMAC_DEV void run(const Blk &b, const Args &a)
    {
        b_unshear<NEQ, LT>(b, a.d, a.w, (const LT *)a.src, a.dst);
    }
};

template <class C>
static void multi_launch(hipStream_t sm, const MultiTab<typename C::Args> &t)
{
    if (t.n > 0 && t.nblk > 0)
        hipLaunchKernelGGL(k_multi<C>, dim3(t.nblk), dim3(C::THREADS), 0, sm, t.args, t.start, t.n,
                           t.item_case, t.active);
}

template <typename LT, int NEQ> void multi_shear(hipStream_t sm, const MultiTab<ShearArgs> &t)
{
    multi_launch<CShear<LT, NEQ>>(sm, t);
}
template <typename LT, int NEQ> void multi_mid(hipStream_t sm, const MultiTab<MidArgs> &t)
{
    multi_launch<CMid<LT, NEQ>>(sm, t);
}
template <typename LT, int NEQ> void multi_unshear(hipStream_t sm, const MultiTab<UnshearArgs> &t)
{
    multi_launch<CUnshear<LT, NEQ>>(sm, t);
}

template ShearArgs shear_args<double>(const Dims &, const LusgsWs &, const double *,
                                      const double *, const double *, const double *,
                                      const double *, const double *, double *);
template ShearArgs shear_args<float>(const Dims &, const LusgsWs &, const double *, const float *,
                                     const float *, const float *, const float *, const float *,
                                     float *);
template MidArgs mid_args<double, 1>(const Dims &, const LusgsWs &);
template MidArgs mid_args<double, 3>(const Dims &, const LusgsWs &);
template MidArgs mid_args<float, 3>(const Dims &, const LusgsWs &);
template void multi_shear<double, 1>(hipStream_t, const MultiTab<ShearArgs> &);
template void multi_shear<double, 3>(hipStream_t, const MultiTab<ShearArgs> &);
template void multi_shear<float, 3>(hipStream_t, const MultiTab<ShearArgs> &);
template void multi_mid<double, 1>(hipStream_t, const MultiTab<MidArgs> &);
template void multi_mid<double, 3>(hipStream_t, const MultiTab<MidArgs> &);
template void multi_mid<float, 3>(hipStream_t, const MultiTab<MidArgs> &);
template void multi_unshear<double, 1>(hipStream_t, const MultiTab<UnshearArgs> &);
template void multi_unshear<double, 3>(hipStream_t, const MultiTab<UnshearArgs> &);
template void multi_unshear<float, 3>(hipStream_t, const MultiTab<UnshearArgs> &);

template <typename LT, int NEQ>
static void launch_lusgs(hipStream_t sm, const Dims &d, const LusgsWs &w, const WakePairs &wk,
                         const double *q, double *dq, double *rhs, const LT *btc, const LT *bjm,
                         const LT *bjp, const LT *bkm, const LT *bkp, LT *fac, bool rhs_out)
{
    lusgs_pre<LT, NEQ>(sm, d, w, rhs, btc, bjm, bjp, bkm, bkp, fac);
    if (w.ev_fwd[0]) (void)hipEventRecord(w.ev_fwd[0], sm);
    lusgs_band<LT, NEQ>(sm, d, w, true);
    if (w.ev_fwd[1]) (void)hipEventRecord(w.ev_fwd[1], sm);
    lusgs_mid<LT, NEQ>(sm, d, w);
    lusgs_band<LT, NEQ>(sm, d, w, false);
    lusgs_post<LT, NEQ>(sm, d, w, wk, q, dq, rhs, rhs_out);
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

#define PHASES(LT, NEQ)                                                                        \
    template void lusgs_pre<LT, NEQ>(hipStream_t, const Dims &, const LusgsWs &, const double *, \
                                     const LT *, const LT *, const LT *, const LT *, const LT *, \
                                     LT *);                                                    \
    template void lusgs_band<LT, NEQ>(hipStream_t, const Dims &, const LusgsWs &, bool);       \
    template void lusgs_band_multi<LT, NEQ>(hipStream_t, const BandJob *, const int2 *, int,   \
                                            int, bool, const int *);                           \
    template void lusgs_mid<LT, NEQ>(hipStream_t, const Dims &, const LusgsWs &);              \
    template void lusgs_post<LT, NEQ>(hipStream_t, const Dims &, const LusgsWs &,              \
                                      const WakePairs &, const double *, double *, double *,   \
                                      bool);
PHASES(double, 3)
PHASES(float, 3)
PHASES(double, 1)
#undef PHASES

template void launch_lusgs3<double>(hipStream_t, const Dims &, const LusgsWs &, const WakePairs &,
                                    const double *, double *, double *, const double *,
                                    const double *, const double *, const double *,
                                    const double *, double *, bool);
template void launch_lusgs3<float>(hipStream_t, const Dims &, const LusgsWs &, const WakePairs &,
                                   const double *, double *, double *, const float *,
                                   const float *, const float *, const float *, const float *,
                                   float *, bool);

}  // namespace smac
