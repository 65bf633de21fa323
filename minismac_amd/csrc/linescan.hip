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

// This file's arithmetic is reassociated anyway, so multiply-adds may fuse here (the library
// is otherwise built without contraction, which keeps the reference-order paths bit-exact).
#pragma clang fp contract(fast)

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
// factorisation: the block Thomas recurrence of btlu3j / btlu3k (ref/btlu3.f:52-624) and of
// scatri (ref/lreltur.f:337-359), for every line of both directions in one launch, one
// wavefront per line.  Waves [0, nx) take the xi lines (k = const: contiguous along the line),
// waves [nx, nx + ne) the eta lines (j = const: stride jm).  The recurrence
//     T_i = b_i - a_i C_(i-1),   C_i = T_i^-1 c_i            (a_0 = 0, c_(len-1) = 0)
// is sequential along the line, so the wave walks it in chunks of 64 points: lane u loads
// point c0+u's blocks (coalesced for xi lines; for eta lines each lane reads its own cache
// line, which the neighbouring lines' waves share through L2), the blocks are staged in LDS,
// the next chunk's loads are issued, and then every lane steps through the chunk's points
// reading each point's blocks from one LDS address (a broadcast) -- the recurrence runs
// redundantly in all lanes, which costs nothing in a wave that has no other work.  A step
// forms T_i^-1 c_i by the adjugate (one reciprocal on the chain) and writes C_i to LDS; after
// the chunk lane u redoes its own point's pivot from C_(u-1), forms T_i^-1 and
// M_i = -T_i^-1 a_i, and stores the operand planes line-major with coalesced stores.  The blocks to the neighbouring lines are copied
// through unchanged.  (The reference keeps T_i as LU factors; the explicit inverse makes
// the sweeps' T_i^-1 r_i a parallel product.)

// 1/x: the hardware reciprocal refined by Newton-Raphson steps r <- r + r (1 - x r), each of
// which doubles the number of correct bits (two for FP64, one for FP32)
SMAC_DEV double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = 1.0 - x * r;
    r = r + r * e;
    e = 1.0 - x * r;
    return r + r * e;
}
SMAC_DEV float rcp_nr(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    return r + r * (1.0f - x * r);
}

// adj(T) (column-major, entry (r, c) at r + 3c) and det(T) of a 3x3 T
template <typename T>
SMAC_DEV T adj3(const T t[9], T adj[9])
{
    adj[0] = t[4] * t[8] - t[7] * t[5];
    adj[1] = t[7] * t[2] - t[1] * t[8];
    adj[2] = t[1] * t[5] - t[4] * t[2];
    adj[3] = t[6] * t[5] - t[3] * t[8];
    adj[4] = t[0] * t[8] - t[6] * t[2];
    adj[5] = t[3] * t[2] - t[0] * t[5];
    adj[6] = t[3] * t[7] - t[6] * t[4];
    adj[7] = t[6] * t[1] - t[0] * t[7];
    adj[8] = t[0] * t[4] - t[3] * t[1];
    return t[0] * adj[0] + t[3] * adj[1] + t[6] * adj[2];
}

// out = x y (NEQ x NEQ, column-major)
template <typename T, int NEQ>
SMAC_DEV void mm(const T *x, const T *y, T *out)
{
#pragma unroll
    for (int r = 0; r < NEQ; ++r)
#pragma unroll
        for (int c = 0; c < NEQ; ++c) {
            T acc = x[r] * y[NEQ * c];
#pragma unroll
            for (int k = 1; k < NEQ; ++k) acc = acc + x[r + NEQ * k] * y[k + NEQ * c];
            out[r + NEQ * c] = acc;
        }
}

constexpr int FW = 64;   // points per chunk = lanes of the wave

// The line is split at m = len/2 and eliminated from both ends at once (a twisted
// factorisation): wave 0 runs the recurrence forward over points 0..m-1,
//     T_i = b_i - a_i C_(i-1),   C_i = T_i^-1 c_i,      M_i = -T_i^-1 a_i,
// wave 1 runs it backward over points len-1..m+1,
//     T_i = b_i - c_i D_(i+1),   D_i = T_i^-1 a_i,      M_i = -T_i^-1 c_i,
// and the middle point takes both: T_m = b_m - a_m C_(m-1) - c_m D_(m+1), with
// Ma = -T_m^-1 a_m and Mc = -T_m^-1 c_m.  Each half's chain is half the line.  Planes per
// point: T^-1; M (Ma at m); -C or -D (Mc at m); the blocks to the previous and next line.
template <typename LT, int NEQ>
__global__ void __launch_bounds__(2 * FW)
k_lr_factor(Dims d, int nx, int twist, const LT *__restrict__ btc, const LT *__restrict__ bjm,
            const LT *__restrict__ bjp, const LT *__restrict__ bkm, const LT *__restrict__ bkp,
            LT *__restrict__ opx, LT *__restrict__ ope)
{
    constexpr int NB = NEQ * NEQ;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool eta = (int)blockIdx.x >= nx;
    // the xi lines are the longer chains: their waves win issue slots on a shared SIMD
    if (!eta) __builtin_amdgcn_s_setprio(2);
    // eta lines: a point's nine planes are read by lanes jm apart, and the lines of
    // neighbouring j share those cache lines; workgroups reach the 8 XCDs round-robin, so
    // each XCD takes a contiguous run of j (its own L2 then holds the shared lines)
    int l = (int)blockIdx.x;
    if (eta) {
        const int t = (int)blockIdx.x - nx, ne = (int)gridDim.x - nx;
        const int r = t % 8, q = ne / 8, rem = ne % 8;
        l = r * q + (r < rem ? r : rem) + t / 8;
    }
    const int len = eta ? d.km : d.jm;
    // twist = 0: the plain (one-sided) factorisation, forward over the whole line (wave 0;
    // launched with one wave)
    const int mid = twist ? len / 2 : len;
    const long N = d.N;
    const long stride = eta ? d.jm : 1, base = eta ? (long)l : (long)l * d.jm;
    const LT *lm = eta ? bkm : bjm, *lp = eta ? bkp : bjp, *om = eta ? bjm : bkm,
             *opn = eta ? bjp : bkp;
    LT *out = eta ? ope : opx;
    __shared__ LT sh[2][FW][3 * NB + 1];   // per wave: carried-side, b, solve-side blocks
    __shared__ LT res[2][FW][NB + 1];      // per wave: C (or D) of the chunk's points
    __shared__ LT cinl[2][NB];             // per wave: the C entering the chunk
    __shared__ LT edge[2][NB];             // C_(m-1) and D_(m+1)
    // wave 0 walks points 0, 1, .., mid-1; wave 1 walks len-1, len-2, .., mid+1
    const bool fwd = wv == 0;
    const int first = fwd ? 0 : len - 1, count = fwd ? mid : (len - 1 - mid > 0 ? len - 1 - mid : 0),
              step = fwd ? 1 : -1;
    // forward: the carried side is a (times C of the point before), the solved side c;
    // backward the other way round
    const LT *carried = fwd ? lm : lp, *solved = fwd ? lp : lm;
    LT rcar[NB], rb[NB], rsol[NB], rom[NB], rop[NB];
    // a chunk's blocks of this direction (prefetched over the previous chunk's recurrence) ...
    auto load = [&](int c0) {
        const int u = c0 + lane;
        if (u < count) {
            const int i = first + step * u;
            const long p = base + stride * i;
            const bool has_a = i > 0, has_c = i < len - 1;
            const bool has_car = fwd ? has_a : has_c, has_sol = fwd ? has_c : has_a;
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                rb[e] = btc[e * N + p];
                rcar[e] = has_car ? carried[e * N + p] : (LT)0;
                rsol[e] = has_sol ? solved[e * N + p] : (LT)0;
            }
        }
    };
    // ... and the other direction's two blocks, which pass through to the operand planes
    // (loaded after the recurrence, so that they hold no registers during it)
    auto load_o = [&](int c0) {
        const int u = c0 + lane;
        if (u < count) {
            const long p = base + stride * (first + step * u);
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                rom[e] = om[e * N + p];
                rop[e] = opn[e * N + p];
            }
        }
    };
    // T = b - x C;  det(T) and adj(T) (T^-1 = adj / det)
    auto pivot = [&](const LT *x, LT *t, const LT *C, LT *ad) -> LT {
#pragma unroll
        for (int cc = 0; cc < NEQ; ++cc)
#pragma unroll
            for (int r = 0; r < NEQ; ++r)
#pragma unroll
                for (int k = 0; k < NEQ; ++k)
                    t[r + NEQ * cc] = t[r + NEQ * cc] - x[r + NEQ * k] * C[k + NEQ * cc];
        if constexpr (NEQ == 3) return adj3(t, ad);
        ad[0] = (LT)1;
        return t[0];
    };
    LT C[NB];
#pragma unroll
    for (int e = 0; e < NB; ++e) C[e] = (LT)0;
    load(0);
    load_o(0);
    // both waves make the same number of chunk steps (their barriers pair up); the shorter
    // half idles through its last one
    const int nchunk = twist && len - 1 - mid > mid ? len - 1 - mid : mid;
    for (int c0 = 0; c0 < nchunk; c0 += FW) {
        const int u0 = c0 + lane;
        const long o = (long)l * len + (first + step * u0);
        if (lane == 0) {
#pragma unroll
            for (int e = 0; e < NB; ++e) cinl[wv][e] = C[e];
        }
        if (u0 < count) {
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                sh[wv][lane][e] = rcar[e];
                sh[wv][lane][NB + e] = rb[e];
                sh[wv][lane][2 * NB + e] = rsol[e];
                out[(3 * NB + e) * N + o] = rom[e];
                out[(4 * NB + e) * N + o] = rop[e];
            }
        }
        __syncthreads();
        if (c0 + FW < count) load(c0 + FW);
        const int nv = count - c0 < FW ? (count - c0 > 0 ? count - c0 : 0) : FW;
        // the recurrence, in every lane of the wave: only C goes on to the next point.  Two
        // register sets alternate: the next point's blocks are read from LDS while this
        // one's are used, so that their latency is off the chain; one lane keeps C for the
        // per-lane pass below.
        LT xa[NB], ta[NB], ca[NB], xb[NB], tb[NB], cb[NB];
        auto rd3 = [&](int u, LT *x, LT *t, LT *c) {
            const int q = u < FW ? u : FW - 1;   // past the chunk: read, never used
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                x[e] = sh[wv][q][e];
                t[e] = sh[wv][q][NB + e];
                c[e] = sh[wv][q][2 * NB + e];
            }
        };
        auto rec = [&](const LT *x, LT *t, const LT *c, int u) {
            LT ad[NB], ac[NB];
            const LT det = pivot(x, t, C, ad);
            mm<LT, NEQ>(ad, c, ac);
            const LT rd = rcp_nr(det);
#pragma unroll
            for (int e = 0; e < NB; ++e) C[e] = ac[e] * rd;
            if (lane == 0) {
#pragma unroll
                for (int e = 0; e < NB; ++e) res[wv][u][e] = C[e];
            }
        };
        rd3(0, xa, ta, ca);
        for (int u = 0; u < nv; u += 2) {
            rd3(u + 1, xb, tb, cb);
            rec(xa, ta, ca, u);
            if (u + 1 < nv) {
                rd3(u + 2, xa, ta, ca);
                rec(xb, tb, cb, u + 1);
            }
        }
        __syncthreads();
        if (c0 + FW < count) load_o(c0 + FW);
        // each lane redoes its own point's pivot from the C entering it (the same operations
        // as the recurrence, so the same values) and forms T^-1 and M = -T^-1 x; its blocks
        // are still in the chunk's LDS
        if (u0 < count) {
            LT cp[NB], ad[NB], bi[NB], m[NB], ocar[NB], ob[NB];
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                cp[e] = lane > 0 ? res[wv][lane - 1][e] : cinl[wv][e];
                ocar[e] = sh[wv][lane][e];
                ob[e] = sh[wv][lane][NB + e];
            }
            const LT rd = rcp_nr(pivot(ocar, ob, cp, ad));
#pragma unroll
            for (int e = 0; e < NB; ++e) bi[e] = ad[e] * rd;
            mm<LT, NEQ>(bi, ocar, m);
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                out[e * N + o] = bi[e];
                out[(NB + e) * N + o] = -m[e];
                out[(2 * NB + e) * N + o] = -res[wv][lane][e];
            }
        }
        __syncthreads();
    }
    if (!twist) return;
    if (lane == 0) {
#pragma unroll
        for (int e = 0; e < NB; ++e) edge[wv][e] = C[e];   // C_(m-1) / D_(m+1) (0 if empty)
    }
    __syncthreads();
    // the middle point
    if (threadIdx.x == 0) {
        const long p = base + stride * mid, o = (long)l * len + mid;
        LT a[NB], b[NB], c[NB], t[NB], ad[NB], u[NB];
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            b[e] = btc[e * N + p];
            a[e] = mid > 0 ? lm[e * N + p] : (LT)0;
            c[e] = mid < len - 1 ? lp[e * N + p] : (LT)0;
            out[(3 * NB + e) * N + o] = om[e * N + p];
            out[(4 * NB + e) * N + o] = opn[e * N + p];
            t[e] = b[e];
        }
        pivot(a, t, edge[0], ad);        // b - a C_(m-1)
        const LT det = pivot(c, t, edge[1], ad);   // - c D_(m+1)
        const LT rd = rcp_nr(det);
        LT bi[NB];
#pragma unroll
        for (int e = 0; e < NB; ++e) bi[e] = ad[e] * rd;
        mm<LT, NEQ>(bi, a, u);
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            out[e * N + o] = bi[e];
            out[(NB + e) * N + o] = -u[e];
        }
        mm<LT, NEQ>(bi, c, u);
#pragma unroll
        for (int e = 0; e < NB; ++e) out[(2 * NB + e) * N + o] = -u[e];
    }
}

// ------------------------------------------------------------------------------------------
// the sweep: lines count of them, line l = (first + idx*step) mod nline.  SEQ: one
// workgroup walks them in order; otherwise one workgroup per line (lines of one colour).
// s, dq: NEQ planes, line-major; op: the factor kernel's planes.  assign: dq = x (the SA's
// lreltur); otherwise dq += underr (x - dq) (ref/lrelax.f:377-381).  couple = 0 leaves out
// the neighbouring lines' terms: a plain solve of each line (the transposed sweeps).
//
// The solve follows the twisted factorisation (mid = len/2):
//   phase 1  y_i = M_i y_(i-1) + T_i^-1 r_i  forward over i < mid, and
//            z_i = M_i z_(i+1) + T_i^-1 r_i  backward over i > mid, both scans at once
//            (each point's map is the identity in the scan of the other half); then
//            x_mid = T^-1 r + Ma y_(mid-1) + Mc z_(mid+1) in the thread owning mid;
//   phase 2  x_i = -C_i x_(i+1) + y_i  backward over i < mid, and
//            x_i = -D_i x_(i-1) + z_i  forward over i > mid, the point mid entering both
//            scans as the constant map x_mid.

// one scan of the thread's PP points along the line (FWD: ascending, else descending) with
// the value 0 entering the line: get(u, a) gives point u's map, put(u, x) receives the value
// after it.  tot: the waves' totals in LDS (the caller synchronises before it is written
// again).
template <typename AT, int NEQ, int NT, int PP, bool FWD, class Get, class Put>
SMAC_DEV void line_scan(Get get, Put put, AT (*tot)[NEQ * NEQ + NEQ], int lane, int wv)
{
    constexpr int NB = NEQ * NEQ, NW = NT / 64;
    Aff<AT, NEQ> agg;
#pragma unroll
    for (int k = 0; k < PP; ++k) {
        const int u = FWD ? k : PP - 1 - k;
        Aff<AT, NEQ> a;
        get(u, a);
        if (k == 0) agg = a;
        else {
            aff_after(a, agg);
            agg = a;
        }
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        Aff<AT, NEQ> o;
        if constexpr (FWD) {
            aff_shfl_up(o, agg, off);
            if (lane >= off) aff_after(agg, o);
        } else {
            aff_shfl_down(o, agg, off);
            if (lane + off < 64) aff_after(agg, o);
        }
    }
    if (lane == (FWD ? 63 : 0)) {
#pragma unroll
        for (int e = 0; e < NB; ++e) tot[wv][e] = agg.m[e];
#pragma unroll
        for (int n = 0; n < NEQ; ++n) tot[wv][NB + n] = agg.v[n];
    }
    __syncthreads();
    // the value entering this wave, then this thread
    AT Y[NEQ], cur[NEQ];
#pragma unroll
    for (int n = 0; n < NEQ; ++n) Y[n] = (AT)0;
    for (int k = 0; k < (FWD ? wv : NW - 1 - wv); ++k) {
        const int w = FWD ? k : NW - 1 - k;
        AT nY[NEQ];
        aff_apply<AT, NEQ>(tot[w], tot[w] + NB, Y, nY);
#pragma unroll
        for (int n = 0; n < NEQ; ++n) Y[n] = nY[n];
    }
    AT ye[NEQ];
    aff_apply<AT, NEQ>(agg.m, agg.v, Y, ye);
#pragma unroll
    for (int n = 0; n < NEQ; ++n) {
        const AT nb = FWD ? __shfl_up(ye[n], 1, 64) : __shfl_down(ye[n], 1, 64);
        cur[n] = lane == (FWD ? 0 : 63) ? Y[n] : nb;
    }
#pragma unroll
    for (int k = 0; k < PP; ++k) {
        const int u = FWD ? k : PP - 1 - k;
        Aff<AT, NEQ> a;
        get(u, a);
        AT nx[NEQ];
        aff_apply<AT, NEQ>(a.m, a.v, cur, nx);
#pragma unroll
        for (int n = 0; n < NEQ; ++n) cur[n] = nx[n];
        put(u, cur);
    }
}

template <typename AT, int NEQ>
SMAC_DEV void aff_identity(Aff<AT, NEQ> &a)
{
#pragma unroll
    for (int r = 0; r < NEQ; ++r) {
#pragma unroll
        for (int c = 0; c < NEQ; ++c) a.m[r + NEQ * c] = r == c ? (AT)1 : (AT)0;
        a.v[r] = (AT)0;
    }
}

template <typename LT, int NEQ, int NT, int PP, bool SEQ, bool TWIST>
__global__ void __launch_bounds__(NT)
k_lr_sweep(int nline, int len, long N, int count, int first, int step, double underr,
           int assign, int couple, double *__restrict__ dq, const double *__restrict__ s,
           const LT *__restrict__ op)
{
    using AT = LT;
    constexpr int NB = NEQ * NEQ, NW = NT / 64;
    __shared__ AT lf[NW][NB + NEQ], lb[NW][NB + NEQ];
    __shared__ AT xch[2][NEQ];   // twisted: y_(mid-1) to the middle point, x_mid back
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    // The plain factorisation: the forward half is the whole line.  The twisted one: the
    // scans run over the line's points renumbered v = 0..len-1 as the first half in order
    // (i = v < mid), the second half reversed (i = len - 1 - (v - mid)) and the middle point
    // last, so that each phase is one scan: phase 1 forward, restarted where the second half
    // begins (i = len - 1, whose map ignores what comes before it), and phase 2 backward from
    // the middle point, the first half restarted from x_mid.
    const int mid = TWIST ? len / 2 : len;
    auto phys = [&](int v) { return (!TWIST || v < mid) ? v : (v < len - 1 ? len - 1 - v + mid : mid); };
    const LT *pB = op, *pM = op + NB * N, *pC = op + 2 * NB * N, *pOm = op + 3 * NB * N,
             *pOp = op + 4 * NB * N;
    double dqprev[PP][NEQ];
    int lprev = -2;
    for (int idx = SEQ ? 0 : (int)blockIdx.x; idx < count; idx += SEQ ? 1 : (int)gridDim.x) {
        int l = (first + idx * step) % nline;
        if (l < 0) l += nline;
        const bool hm = couple && l > 0, hp = couple && l < nline - 1;
        const bool regm = SEQ && lprev == l - 1, regp = SEQ && lprev == l + 1;
        AT mmat[PP][NB];               // M (Ma at mid)
        AT bv[PP][NEQ];                // T^-1 r
        AT w[PP][NEQ];                 // y (i < mid) or z (i > mid); z_(mid+1), then x_mid, at
                                       // the middle point; then x
        // ---- this thread's points: T^-1 r, M, -C / -D
#pragma unroll
        for (int u = 0; u < PP; ++u) {
            const int v = t * PP + u;
            if (v < len) {
                const long o = (long)l * len + phys(v);
                double r[NEQ];
#pragma unroll
                for (int n = 0; n < NEQ; ++n) r[n] = s[n * N + o];
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
                AT bi[NB], rr[NEQ];
#pragma unroll
                for (int e = 0; e < NB; ++e) bi[e] = pB[e * N + o];
#pragma unroll
                for (int n = 0; n < NEQ; ++n) rr[n] = (AT)r[n];
#pragma unroll
                for (int m = 0; m < NEQ; ++m) {
                    AT acc = bi[m] * rr[0];
#pragma unroll
                    for (int n = 1; n < NEQ; ++n) acc = acc + bi[m + NEQ * n] * rr[n];
                    bv[u][m] = acc;
                }
#pragma unroll
                for (int e = 0; e < NB; ++e) mmat[u][e] = pM[e * N + o];
            } else {
#pragma unroll
                for (int e = 0; e < NB; ++e) mmat[u][e] = (AT)0;
#pragma unroll
                for (int n = 0; n < NEQ; ++n) bv[u][n] = (AT)0;
            }
        }
        // ---- phase 1: y forward over i < mid, z backward over i > mid, in one forward scan
        line_scan<AT, NEQ, NT, PP, true>(
            [&](int u, Aff<AT, NEQ> &a) {
                // every field written once, through selects (branches on the struct put it
                // in scratch memory): past the line and at the middle point (which passes
                // z_(mid+1) through) the identity; at the second half's first point a map
                // that ignores what comes before it
                const int v = t * PP + u;
                const bool pass = v >= len || (TWIST && v == len - 1), rst = TWIST && v == mid;
#pragma unroll
                for (int r = 0; r < NEQ; ++r) {
#pragma unroll
                    for (int c = 0; c < NEQ; ++c)
                        a.m[r + NEQ * c] = pass ? (r == c ? (AT)1 : (AT)0)
                                                : rst ? (AT)0 : mmat[u][r + NEQ * c];
                    a.v[r] = pass ? (AT)0 : bv[u][r];
                }
            },
            [&](int u, const AT *x) {
                const int v = t * PP + u;
#pragma unroll
                for (int n = 0; n < NEQ; ++n) {
                    if (v < len) w[u][n] = x[n];
                    if (TWIST && v == mid - 1) xch[0][n] = x[n];
                }
            },
            lf, lane, wv);
        // the -C / -D planes (Mc at mid), read only now: phase 1 does not need them
        AT cm[PP][NB];
#pragma unroll
        for (int u = 0; u < PP; ++u) {
            const int v = t * PP + u;
#pragma unroll
            for (int e = 0; e < NB; ++e)
                cm[u][e] = v < len ? pC[e * N + (long)l * len + phys(v)] : (AT)0;
        }
        __syncthreads();   // phase 1's wave totals have been read; y_(mid-1) is in xch
        // ---- the middle point: x_mid = T^-1 r + Ma y_(mid-1) + Mc z_(mid+1) (the plain
        // factorisation has none: there mid = len, past the line)
        if constexpr (TWIST) {
#pragma unroll
            for (int u = 0; u < PP; ++u)
                if (t * PP + u == len - 1) {
                    AT ymid[NEQ], zmid[NEQ];
#pragma unroll
                    for (int n = 0; n < NEQ; ++n) {
                        ymid[n] = mid > 0 ? xch[0][n] : (AT)0;
                        zmid[n] = len - 1 > mid ? w[u][n] : (AT)0;
                    }
#pragma unroll
                    for (int r = 0; r < NEQ; ++r) {
                        AT acc = bv[u][r];
#pragma unroll
                        for (int k = 0; k < NEQ; ++k)
                            acc = acc + mmat[u][r + NEQ * k] * ymid[k] + cm[u][r + NEQ * k] * zmid[k];
                        w[u][r] = acc;
                        xch[1][r] = acc;
                    }
                }
            __syncthreads();   // x_mid is in xch
        }
        // ---- phase 2: x backward over i < mid and forward over i > mid, from x_mid, in one
        // backward scan: the middle point (last) gives x_mid, the second half follows it, and
        // the first half's last point (v = mid - 1) restarts from x_mid
        line_scan<AT, NEQ, NT, PP, false>(
            [&](int u, Aff<AT, NEQ> &a) {
                // as in phase 1, every field once: past the line the identity; at the middle
                // point (0, x_mid); at the first half's last point (0, -C x_mid + y), a
                // restart from x_mid; elsewhere (-C or -D, y or z)
                const int v = t * PP + u;
                const bool past = v >= len, midp = TWIST && v == len - 1;
                const bool rst = TWIST && v == mid - 1;
                AT xm[NEQ];
#pragma unroll
                for (int k = 0; k < NEQ; ++k) xm[k] = TWIST ? xch[1][k] : (AT)0;
#pragma unroll
                for (int r = 0; r < NEQ; ++r) {
#pragma unroll
                    for (int c = 0; c < NEQ; ++c)
                        a.m[r + NEQ * c] = past ? (r == c ? (AT)1 : (AT)0)
                                                : (midp || rst) ? (AT)0 : cm[u][r + NEQ * c];
                    AT acc = w[u][r];
#pragma unroll
                    for (int k = 0; k < NEQ; ++k) acc = acc + cm[u][r + NEQ * k] * xm[k];
                    a.v[r] = past ? (AT)0 : rst ? acc : w[u][r];
                }
            },
            [&](int u, const AT *x) {
                const int v = t * PP + u;
#pragma unroll
                for (int n = 0; n < NEQ; ++n)
                    if (v < len) w[u][n] = x[n];
            },
            lb, lane, wv);
#pragma unroll
        for (int u = 0; u < PP; ++u) {
            const int v = t * PP + u;
            if (v < len) {
                const long o = (long)l * len + phys(v);
#pragma unroll
                for (int n = 0; n < NEQ; ++n) {
                    const double xd = (double)w[u][n];
                    const double dold = assign ? 0.0 : dq[n * N + o];
                    const double nv = assign ? xd : dold + underr * (xd - dold);
                    dq[n * N + o] = nv;
                    dqprev[u][n] = nv;
                }
            }
        }
        __syncthreads();   // phase 2's wave totals and xch have been read before the next line
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

// twist: the operand planes hold the twisted factorisation (the reference-order walk always
// takes the plain one)
template <typename LT, int NEQ, bool SEQ>
static void sweep_launch(hipStream_t sm, int nline, int len, long N, int count, int first,
                         int step, double underr, int assign, double *dq, const double *s,
                         const LT *op, int couple, bool twist)
{
    const dim3 g(SEQ ? 1 : count);
#define SW(NT, PPV, TW)                                                                         \
    hipLaunchKernelGGL((k_lr_sweep<LT, NEQ, NT, PPV, SEQ, TW>), g, dim3(NT), 0, sm, nline, len, \
                       N, count, first, step, underr, assign, couple, dq, s, op)
#define SWT(TW)                                                                                 \
    do {                                                                                       \
        if (len <= 256) SW(256, 1, TW);                                                        \
        else if (len <= 512) SW(512, 1, TW);                                                   \
        else if (len <= 1024) SW(512, 2, TW);                                                  \
        else SW(1024, 2, TW);                                                                  \
    } while (0)
    if constexpr (SEQ) SWT(false);
    else if (twist) SWT(true);
    else SWT(false);
#undef SWT
#undef SW
}

// one sweep of lines (reference order: odd sweeps ascending, even descending, as
// ref/lrelax.f:249-275; zebra: the lines of even index, then those of odd index)
template <typename LT, int NEQ>
static void sweep(hipStream_t sm, int nline, int len, long N, int nsweep, bool zebra, bool twist,
                  double underr, double *dq, const double *s, const LT *op)
{
    if (zebra) {
        sweep_launch<LT, NEQ, false>(sm, nline, len, N, (nline + 1) / 2, 0, 2, underr, 0, dq, s,
                                     op, 1, twist);
        sweep_launch<LT, NEQ, false>(sm, nline, len, N, nline / 2, 1, 2, underr, 0, dq, s, op, 1,
                                     twist);
        return;
    }
    const bool rev = (nsweep / 2) * 2 == nsweep;
    sweep_launch<LT, NEQ, true>(sm, nline, len, N, nline, rev ? nline - 1 : 0, rev ? -1 : 1,
                                underr, 0, dq, s, op, 1, false);
}

long lr_scan_bytes(const Dims &d)
{
    // 45 operand planes (FP64 size) per direction + eta-major s and dq + the SA's 5 planes
    return (long)(2 * lr_planes<3>() + 6 + lr_planes<1>()) * d.N * 8 + 4 * 256;
}

template <typename LT>
void launch_lr_factor3(hipStream_t sm, const Dims &d, bool xi, bool eta, bool twist,
                       const LT *btc, const LT *bjm, const LT *bjp, const LT *bkm, const LT *bkp,
                       LT *opx, LT *ope)
{
    const int nx = xi ? d.km : 0, ne = eta ? d.jm : 0;
    if (nx + ne > 0)
        hipLaunchKernelGGL((k_lr_factor<LT, 3>), dim3(nx + ne), dim3(twist ? 2 * FW : FW), 0, sm, d,
                           nx, twist ? 1 : 0, btc, bjm, bjp, bkm, bkp, opx, ope);
}

template <typename LT>
void launch_lrelax_scan(hipStream_t sm, const Dims &d, const DevParams &prm, int njsp, int nksp,
                        int nbcimp, bool zebra, const WakePairs &w, const double *q, double *dq,
                        double *s, const LT *btc, const LT *bjm, const LT *bjp, const LT *bkm,
                        const LT *bkp, const LrScan &ws, bool factor, int twist)
{
    const long N = d.N;
    LT *opx = (LT *)ws.opx, *ope = (LT *)ws.ope;
    // the twisted factorisation for the zebra passes (unless the caller says otherwise);
    // the plain one for the reference-order walk
    const bool tw = zebra && twist != 0;
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, sm, dq, 3 * N);
    // both directions' factorisations depend on the blocks only: one launch
    if (factor)
        launch_lr_factor3<LT>(sm, d, njsp > 0, nksp > 0, tw, btc, bjm, bjp, bkm, bkp, opx, ope);
    for (int nj = 1; nj <= njsp; ++nj) {
        sweep<LT, 3>(sm, d.km, d.jm, N, nj, zebra, tw, prm.underr, dq, s, opx);
        if ((nj / nbcimp) * nbcimp == nj) launch_wake(sm, d, 3, w, q, dq, s);
    }
    if (nksp > 0) {
        transpose(sm, s, ws.s_t, d.km, d.jm, N, 3);
        transpose(sm, dq, ws.dq_t, d.km, d.jm, N, 3);
        for (int nk = 1; nk <= nksp; ++nk) {
            sweep<LT, 3>(sm, d.jm, d.km, N, nk, zebra, tw, prm.underr, ws.dq_t, ws.s_t, ope);
            if ((nk / nbcimp) * nbcimp == nk && w.n > 0)
                hipLaunchKernelGGL(k_wake_eta, dim3((w.n + 63) / 64), dim3(64), 0, sm, N, d.jm,
                                   d.km, 3, w.n, w.base, w.pairs, q, ws.dq_t, ws.s_t);
        }
        transpose(sm, ws.dq_t, dq, d.jm, d.km, N, 3);
        transpose(sm, ws.s_t, s, d.jm, d.km, N, 3);
    }
}

// ------------------------------------------------------------------------------------------
// the transpose of the zebra sweeps (the adjoint's preconditioner).  The sweeps map the
// right-hand side s to dq linearly (dq starts at zero; one colour step replaces dq_l on the
// lines of its colour by (1-w) dq_l + w T_l^-1 (s_l - Om_l dq_(l-1) - Op_l dq_(l+1))).  Their
// transpose runs the colour steps in reverse order, each transposed: with a the adjoint of
// dq and z_l = T_l^-T a_l on the colour's lines (a plain line solve of the transposed line
// system, whose factorisation is that of the transposed and shifted blocks),
//     sbar_l += w z_l,   a_l <- (1-w) a_l                 (lines of the colour)
//     a_m -= w (Om'_m z_(m-1) + Op'_m z_(m+1))            (the other lines)
// where Om'_m = (Op_(m-1))^T and Op'_m = (Om_(m+1))^T are again the transposed system's
// blocks to the neighbouring lines at the point.

template <typename LT>
__global__ void __launch_bounds__(256)
k_lr_tspread(int nline, int len, long N, int par, double w, double *__restrict__ a,
             double *__restrict__ sbar, const double *__restrict__ z, const LT *__restrict__ op)
{
    const LT *pOm = op + 27 * N, *pOp = op + 36 * N;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < N;
         o += (long)gridDim.x * blockDim.x) {
        const int l = (int)o / len;   // o < N < 2^31: a 32-bit division
        if ((l & 1) == par) {
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                sbar[n * N + o] += w * z[n * N + o];
                a[n * N + o] *= 1.0 - w;
            }
            continue;
        }
        double acc[3] = {0.0, 0.0, 0.0};
        if (l > 0) {
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    acc[m] += (double)pOm[(m + 3 * k) * N + o] * z[k * N + o - len];
        }
        if (l < nline - 1) {
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    acc[m] += (double)pOp[(m + 3 * k) * N + o] * z[k * N + o + len];
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) a[m * N + o] -= w * acc[m];
    }
}

__global__ void k_copy(const double *__restrict__ src, double *__restrict__ dst, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <typename LT>
static void tsweep(hipStream_t sm, int nline, int len, long N, int nsweep, double w, double *a,
                   double *sbar, double *z, const LT *op)
{
    const unsigned g = (unsigned)((N + 255) / 256 < 2048 ? (N + 255) / 256 : 2048);
    for (int sw = 0; sw < nsweep; ++sw)
        for (int par = 1; par >= 0; --par) {
            const int count = par ? nline / 2 : (nline + 1) / 2;
            if (count > 0)
                sweep_launch<LT, 3, false>(sm, nline, len, N, count, par, 2, 1.0, 1, z, a, op, 0,
                                           false);
            hipLaunchKernelGGL(k_lr_tspread<LT>, dim3(g), dim3(256), 0, sm, nline, len, N, par, w,
                               a, sbar, (const double *)z, op);
        }
}

template <typename LT>
void launch_lrelax_scan_t(hipStream_t sm, const Dims &d, int njsp, int nksp, double underr,
                          const double *in, double *out, const LrScan &wt, const LrTWork &v)
{
    const long N = d.N;
    const unsigned g = (unsigned)((3 * N + 255) / 256 < 2048 ? (3 * N + 255) / 256 : 2048);
    hipLaunchKernelGGL(k_copy, dim3(g), dim3(256), 0, sm, in, v.a, 3 * N);
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, sm, out, 3 * N);
    // the eta sweeps came last in the primal: their transposes come first
    if (nksp > 0) {
        transpose(sm, v.a, v.at, d.km, d.jm, N, 3);
        hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, sm, v.st, 3 * N);
        tsweep<LT>(sm, d.jm, d.km, N, nksp, underr, v.at, v.st, v.zt, (const LT *)wt.ope);
        transpose(sm, v.st, out, d.jm, d.km, N, 3);
        transpose(sm, v.at, v.a, d.jm, d.km, N, 3);
    }
    if (njsp > 0) tsweep<LT>(sm, d.km, d.jm, N, njsp, underr, v.a, out, v.z, (const LT *)wt.opx);
}

void launch_sa_lreltur_scan(hipStream_t sm, const Dims &d, int ntjsp, const WakePairs &w,
                            SaArrays &a, const LrScan &ws)
{
    // lreltur (ref/lreltur.f:52-252): ds = 0; ntjsp passes of the k lines in ascending order
    // followed by a re-solve of line k = 1, each pass ending with the wake injection.  With
    // the zebra order (that of the mean flow) a pass takes the even lines, the odd lines, and
    // then line k = 1 again.
    hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, sm, a.ds, d.N);
    hipLaunchKernelGGL((k_lr_factor<double, 1>), dim3(d.km), dim3(ws.zebra ? 2 * FW : FW), 0, sm, d,
                       d.km, ws.zebra ? 1 : 0, a.stc,
                       a.sjm, a.sjp, a.skm, a.skp, ws.sa_op, (double *)nullptr);
    for (int pass = 1; pass <= ntjsp; ++pass) {
        if (ws.zebra) {
            sweep_launch<double, 1, false>(sm, d.km, d.jm, d.N, (d.km + 1) / 2, 0, 2, 1.0, 1,
                                           a.ds, a.sf, ws.sa_op, 1, true);
            sweep_launch<double, 1, false>(sm, d.km, d.jm, d.N, d.km / 2, 1, 2, 1.0, 1, a.ds,
                                           a.sf, ws.sa_op, 1, true);
            sweep_launch<double, 1, false>(sm, d.km, d.jm, d.N, 1, 0, 1, 1.0, 1, a.ds, a.sf,
                                           ws.sa_op, 1, true);
        } else {
            sweep_launch<double, 1, true>(sm, d.km, d.jm, d.N, d.km + 1, 0, 1, 1.0, 1, a.ds, a.sf,
                                          ws.sa_op, 1, false);
        }
        launch_wake(sm, d, 1, w, a.anut, a.ds, a.sf);
    }
}

template void launch_lr_factor3<double>(hipStream_t, const Dims &, bool, bool, bool, const double *,
                                        const double *, const double *, const double *,
                                        const double *, double *, double *);
template void launch_lrelax_scan_t<double>(hipStream_t, const Dims &, int, int, double,
                                           const double *, double *, const LrScan &,
                                           const LrTWork &);
template void launch_lrelax_scan<double>(hipStream_t, const Dims &, const DevParams &, int, int,
                                         int, bool, const WakePairs &, const double *, double *,
                                         double *, const double *, const double *,
                                         const double *, const double *, const double *,
                                         const LrScan &, bool, int);
template void launch_lrelax_scan<float>(hipStream_t, const Dims &, const DevParams &, int, int,
                                        int, bool, const WakePairs &, const double *, double *,
                                        double *, const float *, const float *, const float *,
                                        const float *, const float *, const LrScan &, bool, int);

}  // namespace smac
