// Residual and implicit-operator assembly of one pass of the implicit iteration:
// conup + fluxj3/fluxk3 (ref/conup.f:49-478, ref/flux.f:160-372), viscterms
// (ref/visc.f:50-160), boundary blanking and the pseudo-time diagonal (ref/step.f:386-437),
// the BC rows of bcmain (ref/bc.f:60-224 and the routines it calls), and resid + rhsnorm
// (ref/step.f:613-756, 513-519).
//
// One thread per grid point computes everything its point needs: the flux-split face
// Jacobians A+/A- of its four xi faces and four eta faces (the third-order flux of a face
// reads dfp/dfm of the faces either side), the point Jacobians of its neighbours, and the
// constant-viscosity face coefficients, then writes s (3 planes) and the five 3x3 blocks
// (45 planes) once.  Faces are recomputed by neighbouring threads instead of being staged
// through memory: the kernel is store-bound (48 words out per point against 8 in), so the
// extra FP64 work is cheaper than a second pass over HBM.
#include "smac_dev.h"
#include "smac_kernels.h"

namespace smac {

// ---------------------------------------------------------------------------------------
// face split, ref/conup.f:127-185 (xi) and :306-355 (eta); rt (time metric) is zero for the
// static grids of this path (ref/metric.f:244-248), kept in the expressions as 0.0
struct Face {
    double ap[9], am[9];   // (m + 3n)
};

SMAC_DEV void face_split(double rx, double ry, double q2, double q3, double beta, Face &f)
{
    const double rt = 0.0;
    double cvel = rx * q2 + ry * q3;
    double t = cvel + 0.5 * rt;
    double cc = sqrt(t * t + beta * (rx * rx + ry * ry));
    double eig = cvel + rt;
    double eig1 = 0.5 * (eig + fabs(eig));
    double eig2 = cvel + rt * 0.5 + cc;
    double x21 = -ry * eig1 / (cc * cc - 0.25 * rt * rt);
    double x31 = rx * eig1 / (cc * cc - 0.25 * rt * rt);
    double x12 = 0.5 * eig2 / cc;
    double x22 = 0.5 * eig2 * (q2 * eig2 + beta * rx) / (beta * cc * (cc - 0.5 * rt));
    double x32 = 0.5 * eig2 * (q3 * eig2 + beta * ry) / (beta * cc * (cc - 0.5 * rt));
    double xi11 = ry * q2 - rx * q3;
    double xi12 = -q3 * (cvel + rt) - ry * beta;
    double xi13 = q2 * (cvel + rt) + rx * beta;
    double xi21 = cc - cvel - 0.5 * rt;
    double xi22 = rx * beta;
    double xi23 = ry * beta;
    f.ap[0] = x12 * xi21;
    f.ap[1] = x21 * xi11 + x22 * xi21;
    f.ap[2] = x31 * xi11 + x32 * xi21;
    f.ap[3] = x12 * xi22;
    f.ap[4] = x21 * xi12 + x22 * xi22;
    f.ap[5] = x31 * xi12 + x32 * xi22;
    f.ap[6] = x12 * xi23;
    f.ap[7] = x21 * xi13 + x22 * xi23;
    f.ap[8] = x31 * xi13 + x32 * xi23;
    f.am[0] = -f.ap[0];
    f.am[1] = -f.ap[1] + rx;
    f.am[2] = -f.ap[2] + ry;
    f.am[3] = -f.ap[3] + rx * beta;
    f.am[4] = -f.ap[4] + rt + 2. * rx * q2 + ry * q3;
    f.am[5] = -f.ap[5] + rx * q3;
    f.am[6] = -f.ap[6] + ry * beta;
    f.am[7] = -f.ap[7] + ry * q2;
    f.am[8] = -f.ap[8] + rt + rx * q2 + 2. * ry * q3;
}

// point flux Jacobian, ref/conup.f:89-105
SMAC_DEV void point_jac(double rx, double ry, double u, double v, double beta, double at[9])
{
    const double rt = 0.0;
    double cvel = rt + rx * u + ry * v;
    at[0] = rt;
    at[1] = rx;
    at[2] = ry;
    at[3] = rx * beta;
    at[4] = rx * u + cvel;
    at[5] = rx * v;
    at[6] = ry * beta;
    at[7] = ry * u;
    at[8] = ry * v + cvel;
}

// physical flux of a point, ref/flux.f:177-181
SMAC_DEV void point_flux(double p, double u, double v, double rx, double ry, double beta,
                         double fv[3])
{
    const double rt = 0.0;
    double cvel = rt + rx * u + ry * v;
    fv[0] = rt * p + beta * (cvel - rt);
    fv[1] = u * cvel + rx * p;
    fv[2] = v * cvel + ry * p;
}

// One direction of the stencil of point c (index t = 0 in a line of n points whose local
// index is `i`): values of the 5 points i-2..i+2 that exist.
struct Line5 {
    double p[5], u[5], v[5], rx[5], ry[5], dj[5];
};

SMAC_DEV void load_line(const double *__restrict__ q, const double *__restrict__ mx,
                        const double *__restrict__ my, const double *__restrict__ dj, long N,
                        long base, long stride, int i, int n, Line5 &L)
{
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        int ii = i + t - 2;
        ii = ii < 0 ? 0 : (ii > n - 1 ? n - 1 : ii);   // clamped copies are never used
        long a = base + stride * ii;
        L.p[t] = q[a];
        L.u[t] = q[N + a];
        L.v[t] = q[2 * N + a];
        L.rx[t] = mx[a];
        L.ry[t] = my[a];
        L.dj[t] = dj[a];
    }
}

// The convective part of one direction at interior point i of its line (flux.f + conup.f):
// returns the flux difference f(i) - f(i-1) and the three blocks m (i-1), p (i+1), c.
SMAC_DEV void convect_dir(const Line5 &L, int i, int n, double beta, double cdis, double dflux[3],
                          double bm[9], double bp[9], double bc[9])
{
    // faces i-2 .. i+1 (face f lies between points f and f+1); index s = f - (i-2).  The
    // blocks are formed as faces i-1 and i appear (same expressions, same order), so only
    // face i-1's split is held until face i's.
    double dfp[4][3], dfm[4][3];
    double ap0[9], am0[9];   // face i-1
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        int f = i - 2 + s;
        bool exists = f >= 0 && f <= n - 2;
        if (!exists) {
            // a face the reference reads only as 0*stale (cp = 0 or cm = 0)
            dfp[s][0] = dfp[s][1] = dfp[s][2] = 0.0;
            dfm[s][0] = dfm[s][1] = dfm[s][2] = 0.0;
            continue;
        }
        int a = s, b = s + 1;   // local points of the face
        double rx = 0.5 * (L.rx[b] + L.rx[a]);
        double ry = 0.5 * (L.ry[b] + L.ry[a]);
        double q2 = 0.5 * (L.u[b] + L.u[a]);
        double q3 = 0.5 * (L.v[b] + L.v[a]);
        Face F;
        face_split(rx, ry, q2, q3, beta, F);
        double dq1 = L.p[b] - L.p[a];
        double dq2 = L.u[b] - L.u[a];
        double dq3 = L.v[b] - L.v[a];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            dfp[s][m] = F.ap[m] * dq1 + F.ap[m + 3] * dq2 + F.ap[m + 6] * dq3;
            dfm[s][m] = F.am[m] * dq1 + F.am[m + 3] * dq2 + F.am[m + 6] * dq3;
        }
        if (s == 1) {
            double atm[9];
            point_jac(L.rx[1], L.ry[1], L.u[1], L.v[1], beta, atm);
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                ap0[e] = F.ap[e];
                am0[e] = F.am[e];
                bm[e] = 0.5 * (-atm[e] - F.ap[e] + F.am[e]);
            }
        }
        if (s == 2) {
            double atp[9];
            point_jac(L.rx[3], L.ry[3], L.u[3], L.v[3], beta, atp);
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                bp[e] = 0.5 * (atp[e] - F.ap[e] + F.am[e]);
                bc[e] = 0.5 * (F.ap[e] + ap0[e] - F.am[e] - am0[e]);
            }
        }
    }
    const double r6 = (double)(1.0f / 6.0f);   // a default-REAL quotient in the reference
    double ff[2][3];
    // Point i is interior (1 <= i <= n-2), so for faces f = i-1 and i the neighbour faces
    // f+1 and f+2 (clamped at n-1) and the points f, f+1 are the local entries sc = 1+w,
    // sp = 2+w (no clamp reaches them), and face f-1, clamped at 0, is entry w except for
    // i = 1, w = 0, where it is face 0 itself (entry 1).  Constant indices keep dfp/dfm in
    // registers.
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        int f = i - 1 + w;             // face index; its neighbours are faces f-1 (clamped) and f+1
        int fm = f - 1 > 0 ? f - 1 : 0;
        int fp = f + 1 < n - 1 ? f + 1 : n - 1;
        int fpp = f + 2 < n - 1 ? f + 2 : n - 1;
        float cp = (float)(fpp - fp);
        float cm = (float)(f - fm);
        constexpr int sc = 1, sp = 2;  // + w below
        const bool sm1 = w == 0 && i == 1;
        const int pf = 1 + w, pp = 2 + w;   // local point indices f, f+1
        double fv0[3], fv1[3];
        point_flux(L.p[pf], L.u[pf], L.v[pf], L.rx[pf], L.ry[pf], beta, fv0);
        point_flux(L.p[pp], L.u[pp], L.v[pp], L.rx[pp], L.ry[pp], beta, fv1);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const double pc = dfp[sc + w][m], mc = dfm[sc + w][m];
            const double pn = dfp[sp + w][m], mn = dfm[sp + w][m];
            const double pm = sm1 ? dfp[1][m] : dfp[w][m], mm = sm1 ? dfm[1][m] : dfm[w][m];
            double fmn = (pc + mc - pn - mn) * r6;
            double fpl = (pm + mm - pc - mc) * r6;
            double f3 = (pm - pc + mc - mn) * r6;
            ff[w][m] = 0.5 * (fv1[m] + fv0[m]) +
                       cdis * ((double)(cp * cm) * f3 + (double)(cm * (1.f - cp)) * fpl +
                               (double)(cp * (1.f - cm)) * fmn);
        }
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) dflux[m] = ff[1][m] - ff[0][m];
}

// viscous face coefficients of one direction at interior point i (ref/visc.f:95-113):
// vf, udif, vdif of faces i-1 (index 0) and i (index 1), all default REAL
SMAC_DEV void visc_dir(const Line5 &L, double vnu, float vf[2], float ud[2], float vd[2])
{
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        int a = 1 + w, b = 2 + w;
        double djv = 0.5 * (L.dj[a] * vnu + L.dj[b] * vnu);
        ud[w] = (float)(L.u[b] - L.u[a]);
        vd[w] = (float)(L.v[b] - L.v[a]);
        double rx = 0.5 * (L.rx[a] + L.rx[b]);
        double ry = 0.5 * (L.ry[a] + L.ry[b]);
        vf[w] = (float)(djv * (rx * rx + ry * ry));
    }
}

template <typename LT>
SMAC_DEV void b_assemble(const Blk &bk, Dims d, DevParams prm, const double *__restrict__ q,
                         const double *__restrict__ met, const double *__restrict__ dj,
                         double *__restrict__ s, LT *__restrict__ btc, LT *__restrict__ bjm,
                         LT *__restrict__ bjp, LT *__restrict__ bkm, LT *__restrict__ bkp)
{
    const int j = bk.x * blockDim.x + threadIdx.x;
    const int k = bk.y;
    if (j >= d.jm || k >= d.km) return;
    const long N = d.N, p = P(d, j, k);
    const bool interior = j > 0 && j < d.jm - 1 && k > 0 && k < d.km - 1;
    if (!interior) {
        // blanked boundary row: all blocks and s zero, btc = identity (ref/step.f:386-437)
#pragma unroll
        for (int m = 0; m < 3; ++m) s[m * N + p] = 0.0;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            btc[e * N + p] = (LT)((e == 0 || e == 4 || e == 8) ? 1.0 : 0.0);
            bjm[e * N + p] = (LT)0.0;
            bjp[e * N + p] = (LT)0.0;
            bkm[e * N + p] = (LT)0.0;
            bkp[e * N + p] = (LT)0.0;
        }
        return;
    }
    // Each direction's neighbour blocks are final once that direction is done and are stored
    // then; the centre block and s take their terms in the reference's order: xi convective,
    // eta convective, xi viscous, eta viscous, pseudo-time diagonal.
    const double *xix = met, *xiy = met + N, *etx = met + 2 * N, *ety = met + 3 * N;
    double vs[3], vc[9], df[3];
    float vfx[2], udx[2], vdx[2];
    Line5 L, Le;
    // xi: assign (ref/conup.f:216-271); the eta line's loads are issued first, so that their
    // latency runs under the xi arithmetic
    load_line(q, etx, ety, dj, N, (long)j, d.jm, k, d.km, Le);
    load_line(q, xix, xiy, dj, N, (long)d.jm * k, 1, j, d.jm, L);
    {
        double bm[9], bp[9];
        convect_dir(L, j, d.jm, prm.beta, prm.cdis, df, bm, bp, vc);
#pragma unroll
        for (int m = 0; m < 3; ++m) vs[m] = -df[m];
        if (prm.ivis == 1) {
            visc_dir(L, prm.vnu, vfx, udx, vdx);
            // xi viscous blocks (ref/visc.f:95-113)
            const double a = vfx[0], b = vfx[1];
            bm[4] = bm[4] - a; bm[8] = bm[8] - a;
            bp[4] = bp[4] - b; bp[8] = bp[8] - b;
        }
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            bjm[e * N + p] = (LT)bm[e];
            bjp[e * N + p] = (LT)bp[e];
        }
    }
    // eta: add (ref/conup.f:377-457)
    {
        double bm[9], bp[9], bc[9];
        convect_dir(Le, k, d.km, prm.beta, prm.cdis, df, bm, bp, bc);
#pragma unroll
        for (int e = 0; e < 9; ++e) vc[e] = vc[e] + bc[e];
#pragma unroll
        for (int m = 0; m < 3; ++m) vs[m] = vs[m] - df[m];
        if (prm.ivis == 1) {
            float vfy[2], udy[2], vdy[2];
            visc_dir(Le, prm.vnu, vfy, udy, vdy);
            // xi viscous terms of the centre and s (ref/visc.f:95-113), then eta (:121-140)
            double a = vfx[0], b = vfx[1];
            vc[4] = vc[4] + a + b; vc[8] = vc[8] + a + b;
            float t2 = vfx[1] * udx[1] - vfx[0] * udx[0];
            float t3 = vfx[1] * vdx[1] - vfx[0] * vdx[0];
            vs[1] = vs[1] + (double)t2;
            vs[2] = vs[2] + (double)t3;
            a = vfy[0]; b = vfy[1];
            bm[4] = bm[4] - a; bm[8] = bm[8] - a;
            vc[4] = vc[4] + a + b; vc[8] = vc[8] + a + b;
            bp[4] = bp[4] - b; bp[8] = bp[8] - b;
            t2 = vfy[1] * udy[1] - vfy[0] * udy[0];
            t3 = vfy[1] * vdy[1] - vfy[0] * vdy[0];
            vs[1] = vs[1] + (double)t2;
            vs[2] = vs[2] + (double)t3;
        }
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            bkm[e * N + p] = (LT)bm[e];
            bkp[e * N + p] = (LT)bp[e];
        }
    }
    // pseudo-time diagonal (ref/step.f:419-437), fib = 1
    double rdjdtau = 1.0 / (dj[p] * prm.dtau) + 1. - 1.0;
    vc[0] = vc[0] + rdjdtau;
    vc[4] = vc[4] + rdjdtau;
    vc[8] = vc[8] + rdjdtau;
#pragma unroll
    for (int m = 0; m < 3; ++m) s[m * N + p] = vs[m];
#pragma unroll
    for (int e = 0; e < 9; ++e) btc[e * N + p] = (LT)vc[e];
}

template <typename LT>
__global__ void __launch_bounds__(256)
k_assemble(Dims d, DevParams prm, const double *__restrict__ q, const double *__restrict__ met,
           const double *__restrict__ dj, double *__restrict__ s, LT *__restrict__ btc,
           LT *__restrict__ bjm, LT *__restrict__ bjp, LT *__restrict__ bkm, LT *__restrict__ bkp)
{
    b_assemble<LT>(xcd_blk(), d, prm, q, met, dj, s, btc, bjm, bjp, bkm, bkp);
}

// ---------------------------------------------------------------------------------------
// BC rows (ref/bc.f): every region of the case in one single-workgroup launch, region after
// region in file order (a barrier between regions, so a later region's rows win where two
// regions share a point, as in the reference's sequence of calls)

// bcoutj, extrapolation branch of ibval 31 plus the fixed-pressure row (ref/bc.f:1627-1676)
template <typename LT>
SMAC_DEV void bc_outj(const Dims &d, int j, int k, double pout, const double *__restrict__ q,
                      double *__restrict__ s, LT *btc, LT *bjm, LT *bjp)
{
    const long N = d.N, p = P(d, j, k);
    const int jn = j == 0 ? j + 1 : j - 1;
    LT *bn = j == 0 ? bjp : bjm;
    const long pn = P(d, jn, k);
    btc[4 * N + p] = (LT)1.0;
    btc[8 * N + p] = (LT)1.0;
    bn[4 * N + p] = (LT)-1.0;
    bn[8 * N + p] = (LT)-1.0;
    s[N + p] = q[N + pn] - q[N + p];
    s[2 * N + p] = q[2 * N + pn] - q[2 * N + p];
    btc[p] = (LT)1.0;
    s[p] = pout - q[p];
}

// bcnoslk, ibval 0 branch (ref/bc.f:1147-1158)
template <typename LT>
SMAC_DEV void bc_noslk(const Dims &d, int j, int k, const double *__restrict__ q,
                       double *__restrict__ s, LT *btc, LT *bkp)
{
    const long N = d.N, p = P(d, j, k), pu = P(d, j, k + 1);
    btc[p] = (LT)1.0;
    btc[4 * N + p] = (LT)1.0;
    btc[8 * N + p] = (LT)1.0;
    bkp[p] = (LT)-1.0;
    s[p] = q[pu] - q[p];
    s[N + p] = -q[N + p];
    s[2 * N + p] = -q[2 * N + p];
}

// bccgridk, ibval 25 at k = kmax: free stream plus a point vortex of strength clift for
// u, v and the outgoing characteristic for p (ref/bc.f:1334-1431)
template <typename LT>
SMAC_DEV void bc_cgridk(const Dims &d, int j, int k, double alpha, double beta, double clift,
                        const double *__restrict__ x, const double *__restrict__ y,
                        const double *__restrict__ met, const double *__restrict__ q,
                        double *__restrict__ s, LT *btc, LT *bkm)
{
    const long N = d.N, p = P(d, j, k), pd = P(d, j, k - 1);
    double pi = 4.0 * atan(1.0);
    double uinf = cos(alpha * pi / 180.);
    double vinf = sin(alpha * pi / 180.);
    double dx = x[p] - 0.25;
    double dy = y[p];
    double ucirc = clift * dy / (4. * pi * (dx * dx + dy * dy));
    double vcirc = -clift * dx / (4. * pi * (dx * dx + dy * dy));
    btc[4 * N + p] = (LT)1.0;
    btc[8 * N + p] = (LT)1.0;
    s[N + p] = uinf + ucirc - q[N + p];
    s[2 * N + p] = vinf + vcirc - q[2 * N + p];
    double r2 = met[2 * N + p], r3 = met[3 * N + p];
    double qq = r2 * q[N + p] + r3 * q[2 * N + p];
    double cc = sqrt(qq * qq + beta * (r2 * r2 + r3 * r3));
    double b11 = -(r2 * r2 + r3 * r3);
    double b12 = -r2 * (qq + cc);
    double b13 = -r3 * (qq + cc);
    bkm[p] = (LT)b11;
    bkm[3 * N + p] = (LT)b12;
    bkm[6 * N + p] = (LT)b13;
    btc[p] = (LT)(-b11);
    btc[3 * N + p] = (LT)(-b12);
    btc[6 * N + p] = (LT)(-b13);
    double dq1 = q[p] - q[pd];
    double dq2 = q[N + p] - q[N + pd];
    double dq3 = q[2 * N + p] - q[2 * N + pd];
    s[p] = b11 * dq1 + b12 * dq2 + b13 * dq3;
}

template <typename LT>
SMAC_DEV void b_bc(const Dims &d, const BcTable &t, double pout, double alpha, double beta,
                   const DevStats *__restrict__ st, const double *__restrict__ x,
                   const double *__restrict__ y, const double *__restrict__ met,
                   const double *__restrict__ q, double *__restrict__ s, LT *btc, LT *bjm,
                   LT *bjp, LT *bkm, LT *bkp)
{
    const double clift = st->clift;
    for (int i = 0; i < t.n; ++i) {
        const BcRegion r = t.r[i];
        if (r.ibcval == 31) {
            for (int k = r.kbeg + (int)threadIdx.x; k <= r.kend; k += BC_THREADS)
                bc_outj<LT>(d, r.jbeg, k, pout, q, s, btc, bjm, bjp);
        } else if (r.ibcval == 0) {
            for (int j = r.jbeg + (int)threadIdx.x; j <= r.jend; j += BC_THREADS)
                bc_noslk<LT>(d, j, r.kbeg, q, s, btc, bkp);
        } else if (r.ibcval == 25) {
            for (int j = r.jbeg + (int)threadIdx.x; j <= r.jend; j += BC_THREADS)
                bc_cgridk<LT>(d, j, r.kbeg, alpha, beta, clift, x, y, met, q, s, btc, bkm);
        }
        __syncthreads();
    }
}

template <typename LT>
__global__ void __launch_bounds__(BC_THREADS)
k_bc(Dims d, BcTable t, double pout, double alpha, double beta, const DevStats *__restrict__ st,
     const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ met,
     const double *__restrict__ q, double *__restrict__ s, LT *btc, LT *bjm, LT *bjp, LT *bkm,
     LT *bkp)
{
    b_bc<LT>(d, t, pout, alpha, beta, st, x, y, met, q, s, btc, bjm, bjp, bkm, bkp);
}

// ---------------------------------------------------------------------------------------
// resid + rhsnorm: isamax over s*dj (ties to the first index, ref/blas.f:48-83) for all 3N
// entries and for the first N, and sum of s^2; partials per block, then one final block.
// The partition is fixed, so the result is the same on every run.

struct ArgMax {
    double v;
    long i;
};

SMAC_DEV ArgMax amax_join(ArgMax a, ArgMax b)
{
    if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
    return a;
}

constexpr int RED_THREADS = 256;

SMAC_DEV void b_resid_partial(const Blk &bk, Dims d, const double *__restrict__ s,
                              const double *__restrict__ dj, double *__restrict__ part_v,
                              long *__restrict__ part_i, double *__restrict__ part_s)
{
    const long N = d.N, n3 = 3 * N;
    ArgMax all = {-1.0, 0x7fffffffffffffffL}, first = {-1.0, 0x7fffffffffffffffL};
    double ss = 0.0;
    for (long i = (long)bk.x * blockDim.x + threadIdx.x; i < n3;
         i += (long)bk.nx * blockDim.x) {
        long p = i % N;
        double sv = s[i];
        double a = fabs(sv * dj[p]);
        all = amax_join(all, {a, i});
        if (i < N) first = amax_join(first, {a, i});
        ss += sv * sv;
    }
    __shared__ double sv_[RED_THREADS], fv_[RED_THREADS], sq_[RED_THREADS];
    __shared__ long si_[RED_THREADS], fi_[RED_THREADS];
    const int t = threadIdx.x;
    sv_[t] = all.v; si_[t] = all.i; fv_[t] = first.v; fi_[t] = first.i; sq_[t] = ss;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if (t < w) {
            ArgMax a = amax_join({sv_[t], si_[t]}, {sv_[t + w], si_[t + w]});
            sv_[t] = a.v; si_[t] = a.i;
            ArgMax b = amax_join({fv_[t], fi_[t]}, {fv_[t + w], fi_[t + w]});
            fv_[t] = b.v; fi_[t] = b.i;
            sq_[t] = sq_[t] + sq_[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        part_v[2 * bk.x] = sv_[0]; part_i[2 * bk.x] = si_[0];
        part_v[2 * bk.x + 1] = fv_[0]; part_i[2 * bk.x + 1] = fi_[0];
        part_s[bk.x] = sq_[0];
    }
}

__global__ void __launch_bounds__(RED_THREADS)
k_resid_partial(Dims d, const double *__restrict__ s, const double *__restrict__ dj,
                double *__restrict__ part_v, long *__restrict__ part_i, double *__restrict__ part_s)
{
    b_resid_partial(hw_blk(), d, s, dj, part_v, part_i, part_s);
}

SMAC_DEV void b_resid_final(const Dims &d, int nparts, int nt, double beta,
                            const double *__restrict__ s, const double *__restrict__ dj,
                            const double *__restrict__ part_v, const long *__restrict__ part_i,
                            const double *__restrict__ part_s, DevStats *st)
{
    const int t = threadIdx.x;
    ArgMax all = {-1.0, 0x7fffffffffffffffL}, first = {-1.0, 0x7fffffffffffffffL};
    double ss = 0.0;
    for (int b = t; b < nparts; b += RED_THREADS) {
        all = amax_join(all, {part_v[2 * b], part_i[2 * b]});
        first = amax_join(first, {part_v[2 * b + 1], part_i[2 * b + 1]});
        ss += part_s[b];
    }
    __shared__ double sv_[RED_THREADS], fv_[RED_THREADS], sq_[RED_THREADS];
    __shared__ long si_[RED_THREADS], fi_[RED_THREADS];
    sv_[t] = all.v; si_[t] = all.i; fv_[t] = first.v; fi_[t] = first.i; sq_[t] = ss;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
        if (t < w) {
            ArgMax a = amax_join({sv_[t], si_[t]}, {sv_[t + w], si_[t + w]});
            sv_[t] = a.v; si_[t] = a.i;
            ArgMax b = amax_join({fv_[t], fi_[t]}, {fv_[t + w], fi_[t + w]});
            fv_[t] = b.v; fi_[t] = b.i;
            sq_[t] = sq_[t] + sq_[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const long N = d.N;
        long irmax = si_[0];                 // 0-based index into the 3N array
        long id = fi_[0];
        int bad = 0;
        // an all-NaN residual leaves no candidate: report it instead of indexing with it
        if (irmax < 0 || irmax >= 3 * N) { irmax = 0; bad = 1; }
        if (id < 0 || id >= N) { id = 0; bad = 1; }
        long p = irmax % N;
        st->kres = (int)(p / d.jm) + 1;
        st->jres = (int)(p % d.jm) + 1;
        st->resmax = fabs(s[irmax] * dj[p]);
        if (nt == 1) {
            st->resmax0 = st->resmax;
            if (st->resmax0 < 1.) st->resmax0 = 1.0;
        }
        st->divmax = fabs(s[id] * dj[id] / beta);
        st->rhsnorm = sqrt(sq_[0]);
        st->nonfinite = bad || !(st->rhsnorm <= 1.0e300);
    }
}

__global__ void __launch_bounds__(RED_THREADS)
k_resid_final(Dims d, int nparts, int nt, double beta, const double *__restrict__ s,
              const double *__restrict__ dj, const double *__restrict__ part_v,
              const long *__restrict__ part_i, const double *__restrict__ part_s, DevStats *st)
{
    b_resid_final(d, nparts, nt, beta, s, dj, part_v, part_i, part_s, st);
}

// ---------------------------------------------------------------------------------------
// launchers

template <typename LT>
void launch_assemble(hipStream_t sm, const Dims &d, const DevParams &prm, const double *q,
                     const double *met, const double *dj, double *s, LT *btc, LT *bjm, LT *bjp,
                     LT *bkm, LT *bkp)
{
    dim3 blk(64), grd((d.jm + 63) / 64, d.km);
    hipLaunchKernelGGL(k_assemble<LT>, grd, blk, 0, sm, d, prm, q, met, dj, s, btc, bjm, bjp, bkm,
                       bkp);
}

template <typename LT>
void launch_bc(hipStream_t sm, const Dims &d, const DevParams &prm, const BcRegion *regs,
               int nreg, const DevStats *st, const double *x, const double *y, const double *met,
               const double *q, double *s, LT *btc, LT *bjm, LT *bjp, LT *bkm, LT *bkp)
{
    // up to BC_MAX regions per launch; more go in further launches, still in file order
    BcTable t{};
    for (int i = 0; i <= nreg; ++i) {
        if (i == nreg || t.n == BC_MAX) {
            if (t.n)
                hipLaunchKernelGGL(k_bc<LT>, dim3(1), dim3(BC_THREADS), 0, sm, d, t, prm.pout,
                                   prm.alpha, prm.beta, st, x, y, met, q, s, btc, bjm, bjp, bkm,
                                   bkp);
            t.n = 0;
            if (i == nreg) break;
        }
        const int ib = regs[i].ibcval;
        if (ib != 31 && ib != 0 && ib != 25) continue;   // wake cuts: bcimpds in the solve
        t.r[t.n++] = regs[i];
    }
}

void launch_resid(hipStream_t sm, const Dims &d, int nt, double beta, const double *s,
                  const double *dj, Scratch &w, DevStats *st)
{
    int nb = w.nred;
    hipLaunchKernelGGL(k_resid_partial, dim3(nb), dim3(RED_THREADS), 0, sm, d, s, dj, w.red_v,
                       w.red_i, w.red_s);
    hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(RED_THREADS), 0, sm, d, nb, nt, beta, s, dj,
                       w.red_v, w.red_i, w.red_s, st);
}

// ---- multi-case launches ------------------------------------------------------------------

template <typename LT>
AsmArgs asm_args(const Dims &d, const DevParams &prm, const double *q, const double *met,
                 const double *dj, double *s, LT *btc, LT *bjm, LT *bjp, LT *bkm, LT *bkp)
{
    return AsmArgs{(unsigned)((d.jm + 63) / 64), (unsigned)d.km, d, prm, q, met, dj, s, btc, bjm,
                   bjp, bkm, bkp};
}

template <typename LT>
void bc_args(const Dims &d, const DevParams &prm, const BcRegion *regs, int nreg,
             const DevStats *st, const double *x, const double *y, const double *met,
             const double *q, double *s, LT *btc, LT *bjm, LT *bjp, LT *bkm, LT *bkp,
             std::vector<BcArgs> &out)
{
    BcTable t{};
    for (int i = 0; i <= nreg; ++i) {
        if (i == nreg || t.n == BC_MAX) {
            if (t.n)
                out.push_back(BcArgs{1u, 1u, d, t, prm.pout, prm.alpha, prm.beta, st, x, y, met,
                                     q, s, btc, bjm, bjp, bkm, bkp});
            t.n = 0;
            if (i == nreg) break;
        }
        const int ib = regs[i].ibcval;
        if (ib != 31 && ib != 0 && ib != 25) continue;
        t.r[t.n++] = regs[i];
    }
}

ResidPArgs residp_args(const Dims &d, const double *s, const double *dj, Scratch &w)
{
    return ResidPArgs{(unsigned)w.nred, 1u, d, s, dj, w.red_v, w.red_i, w.red_s};
}

ResidFArgs residf_args(const Dims &d, int nt, double beta, const double *s, const double *dj,
                       Scratch &w, DevStats *st)
{
    return ResidFArgs{1u, 1u, d, w.nred, nt, beta, s, dj, w.red_v, w.red_i, w.red_s, st};
}

template <typename LT>
struct CAssemble {
    using Args = AsmArgs;
    static constexpr int THREADS = 64;
    static SMAC_DEV void run(const Blk &b, const Args &a)
    {
        b_assemble<LT>(b, a.d, a.prm, a.q, a.met, a.dj, a.s, (LT *)a.btc, (LT *)a.bjm,
                       (LT *)a.bjp, (LT *)a.bkm, (LT *)a.bkp);
    }
};

template <typename LT>
struct CBc {
    using Args = BcArgs;
    static constexpr int THREADS = BC_THREADS;
    static SMAC_DEV void run(const Blk &, const Args &a)
    {
        b_bc<LT>(a.d, a.t, a.pout, a.alpha, a.beta, a.st, a.x, a.y, a.met, a.q, a.s, (LT *)a.btc,
                 (LT *)a.bjm, (LT *)a.bjp, (LT *)a.bkm, (LT *)a.bkp);
    }
};

struct CResidP {
    using Args = ResidPArgs;
    static constexpr int THREADS = RED_THREADS;
    static SMAC_DEV void run(const Blk &b, const Args &a)
    {
        b_resid_partial(b, a.d, a.s, a.dj, a.part_v, a.part_i, a.part_s);
    }
};

struct CResidF {
    using Args = ResidFArgs;
    static constexpr int THREADS = RED_THREADS;
    static SMAC_DEV void run(const Blk &, const Args &a)
    {
        b_resid_final(a.d, a.nparts, a.nt, a.beta, a.s, a.dj, a.part_v, a.part_i, a.part_s,
                      a.st);
    }
};

template <class C>
static void multi_launch(hipStream_t sm, const MultiTab<typename C::Args> &t)
{
    if (t.n > 0 && t.nblk > 0)
        hipLaunchKernelGGL(k_multi<C>, dim3(t.nblk), dim3(C::THREADS), 0, sm, t.args, t.start, t.n,
                           t.item_case, t.active);
}

template <typename LT> void multi_assemble(hipStream_t sm, const MultiTab<AsmArgs> &t)
{
    multi_launch<CAssemble<LT>>(sm, t);
}
template <typename LT> void multi_bc(hipStream_t sm, const MultiTab<BcArgs> &t)
{
    multi_launch<CBc<LT>>(sm, t);
}
void multi_resid_partial(hipStream_t sm, const MultiTab<ResidPArgs> &t) { multi_launch<CResidP>(sm, t); }
void multi_resid_final(hipStream_t sm, const MultiTab<ResidFArgs> &t) { multi_launch<CResidF>(sm, t); }

template AsmArgs asm_args<double>(const Dims &, const DevParams &, const double *, const double *,
                                  const double *, double *, double *, double *, double *,
                                  double *, double *);
template AsmArgs asm_args<float>(const Dims &, const DevParams &, const double *, const double *,
                                 const double *, double *, float *, float *, float *, float *,
                                 float *);
template void bc_args<double>(const Dims &, const DevParams &, const BcRegion *, int,
                              const DevStats *, const double *, const double *, const double *,
                              const double *, double *, double *, double *, double *, double *,
                              double *, std::vector<BcArgs> &);
template void bc_args<float>(const Dims &, const DevParams &, const BcRegion *, int,
                             const DevStats *, const double *, const double *, const double *,
                             const double *, double *, float *, float *, float *, float *,
                             float *, std::vector<BcArgs> &);
template void multi_assemble<double>(hipStream_t, const MultiTab<AsmArgs> &);
template void multi_assemble<float>(hipStream_t, const MultiTab<AsmArgs> &);
template void multi_bc<double>(hipStream_t, const MultiTab<BcArgs> &);
template void multi_bc<float>(hipStream_t, const MultiTab<BcArgs> &);

template void launch_assemble<double>(hipStream_t, const Dims &, const DevParams &,
                                      const double *, const double *, const double *, double *,
                                      double *, double *, double *, double *, double *);
template void launch_assemble<float>(hipStream_t, const Dims &, const DevParams &,
                                     const double *, const double *, const double *, double *,
                                     float *, float *, float *, float *, float *);
template void launch_bc<double>(hipStream_t, const Dims &, const DevParams &, const BcRegion *,
                                int, const DevStats *, const double *, const double *,
                                const double *, const double *, double *, double *, double *,
                                double *, double *, double *);
template void launch_bc<float>(hipStream_t, const Dims &, const DevParams &, const BcRegion *, int,
                               const DevStats *, const double *, const double *, const double *,
                               const double *, double *, float *, float *, float *, float *,
                               float *);

}  // namespace smac
