// Discrete adjoint of the converged mean flow (no counterpart in the reference, which has no
// sensitivities): gradients of CL and CD with respect to the angle of attack and to shape
// directions (grid displacements), at a steady state of the implicit iteration.
//
// The steady equations.  A converged pass changes nothing: the assembled right-hand side s
// (interior flux-split residual plus viscous terms, outflow / wall / far-field BC rows) is
// zero at every row except the wake-cut targets, whose s is blanked and whose update is the
// injection ref/bc.f:571-591, so there the equation is q(jt,1) = (q(jt,2) + q(jb,2)) / 2.
// The far-field rows read clift, the previous pass's CL, which at a steady state is CL(q).
// So F(q, c) = 0 with c = CL(q).  With the Lagrangian J + lambda.F + mu (c - CL):
//     -F_q^T lambda - (lambda.F_c) CL_q^T = J_q^T                (the adjoint equation)
//     dJ/dp = J_p + lambda.F_p + (lambda.F_c) CL_p
// for J = CL and J = CD, p = alpha or a shape direction.
//
// Kernels:
//  - k_adj_jac: the exact Jacobian F_q, 9 cross-stencil 3x3 blocks per row, by forward-mode
//    dual numbers: the interior row is the sum of a xi part (points j-2..j+2) and an eta part
//    (k-2..k+2), each evaluated once per stencil point with the point's three variables
//    seeded (the same expressions as the assembly kernel, FP32 islands rounded in value);
//    BC rows are written directly.
//  - k_adj_jtv + k_adj_wake_t: lambda -> F_q^T lambda, a gather over the 9 offsets plus the
//    wake cut's far partner (jb,2).
//  - the preconditioner: the transpose of the primal iteration's own solve
//    P = W M^-1 (on non-target rows) + identity (on targets).  For LU-SGS (impsch 2),
//    M = (D+L) D^-1 (D+U) of the converged state; M^T = (D+U)^T D^-T (D+L)^T is again an
//    LU-SGS operator, with D^T and the neighbour blocks transposed and shifted
//    (bjm'(p) = bjp(p-1)^T, bkm'(p) = bkp(p-jm)^T, bjp'(p) = bjm(p+1)^T, bkp'(p) = bkm(p+jm)^T),
//    so M^-T runs on the primal's own hyperplane band kernels (k_adj_transpose_blocks).  For
//    line relaxation (impsch 1), M^-1 is the zebra sweeps of the blocks (xi then eta, wake
//    injection only at the end, in W) and M^-T their exact transpose: the colour steps in
//    reverse order, each a line solve of the transposed line system (the factorisation of
//    the same transposed and shifted blocks) and a spread to the neighbouring lines.
//  - the force gradient (k_adj_force_grad: CL_q, CD_q by dual numbers on the wall faces),
//    the metric tangent of a shape direction (k_adj_metric_tangent, as the metric is formed:
//    second-order differences, one-sided at the edges), the residual tangent at fixed q
//    (k_adj_resid_tangent) and the force tangent (k_adj_force_tangent).
//  - GMRES(m) with right preconditioning, classical Gram-Schmidt applied twice; its inner
//    products are one multi-vector kernel per pass.  The adjoint operator is
//    Op = K - g f^T with K = -F_q^T, g = CL_q^T and f = F_c (the far field's lift coupling);
//    P^T approximates K^-1 only, and GMRES with P^T alone stagnates on the rank-one part
//    (restarted, it loses that direction).  So the preconditioner is P^T with the
//    Sherman-Morrison correction for the rank-one term:
//        Pc v = P^T v + u (f . P^T v) / (1 - f . u),     u = P^T g  (formed once),
//    one extra dot product per application.
#include "smac_dev.h"
#include "smac_kernels.h"

#include <cmath>
#include <vector>

namespace smac {

// ------------------------------------------------------------------------------------------
// forward-mode dual numbers

template <int K>
struct Dn {
    double v;
    double d[K];
};

template <int K> SMAC_DEV Dn<K> dn(double v)
{
    Dn<K> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = 0.0;
    return r;
}
template <int K> SMAC_DEV Dn<K> operator+(const Dn<K> &a, const Dn<K> &b)
{
    Dn<K> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int K> SMAC_DEV Dn<K> operator-(const Dn<K> &a, const Dn<K> &b)
{
    Dn<K> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int K> SMAC_DEV Dn<K> operator-(const Dn<K> &a)
{
    Dn<K> r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int K> SMAC_DEV Dn<K> operator*(const Dn<K> &a, const Dn<K> &b)
{
    Dn<K> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int K> SMAC_DEV Dn<K> operator/(const Dn<K> &a, const Dn<K> &b)
{
    Dn<K> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <int K> SMAC_DEV Dn<K> operator+(const Dn<K> &a, double b) { return a + dn<K>(b); }
template <int K> SMAC_DEV Dn<K> operator+(double a, const Dn<K> &b) { return dn<K>(a) + b; }
template <int K> SMAC_DEV Dn<K> operator-(const Dn<K> &a, double b) { return a - dn<K>(b); }
template <int K> SMAC_DEV Dn<K> operator-(double a, const Dn<K> &b) { return dn<K>(a) - b; }
template <int K> SMAC_DEV Dn<K> operator*(const Dn<K> &a, double b)
{
    Dn<K> r;
    r.v = a.v * b;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <int K> SMAC_DEV Dn<K> operator*(double a, const Dn<K> &b) { return b * a; }
template <int K> SMAC_DEV Dn<K> operator/(const Dn<K> &a, double b) { return a / dn<K>(b); }
template <int K> SMAC_DEV Dn<K> operator/(double a, const Dn<K> &b) { return dn<K>(a) / b; }

template <int K> SMAC_DEV Dn<K> sqrt_s(const Dn<K> &a)
{
    Dn<K> r;
    r.v = sqrt(a.v);
    const double h = 0.5 / r.v;
#pragma unroll
    for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * h;
    return r;
}
template <int K> SMAC_DEV Dn<K> fabs_s(const Dn<K> &a) { return a.v < 0.0 ? -a : a; }
// a default-REAL (FP32) temporary of the reference: the value rounded, the derivative kept
template <int K> SMAC_DEV Dn<K> r32(const Dn<K> &a)
{
    Dn<K> r = a;
    r.v = (double)(float)a.v;
    return r;
}
SMAC_DEV double sqrt_s(double a) { return sqrt(a); }
SMAC_DEV double fabs_s(double a) { return fabs(a); }
SMAC_DEV double r32(double a) { return (double)(float)a; }
template <int K> SMAC_DEV double val(const Dn<K> &a) { return a.v; }
SMAC_DEV double val(double a) { return a; }

// ------------------------------------------------------------------------------------------
// one direction of the interior residual: -(f(i) - f(i-1)) + viscous (conup/flux3 and
// viscterms in the expressions of the assembly kernel), on the 5 points i-2..i+2 of a line

template <typename S>
struct LineS {
    S p[5], u[5], v[5], rx[5], ry[5], dj[5];
};

template <typename S>
SMAC_DEV void face_dq(const S &rx, const S &ry, const S &q2, const S &q3, double beta,
                      const S &dq1, const S &dq2, const S &dq3, S dfp[3], S dfm[3])
{
    // A+ / A- of the face times the jump (face_split of the assembly, rt = 0)
    S cvel = rx * q2 + ry * q3;
    S cc = sqrt_s(cvel * cvel + beta * (rx * rx + ry * ry));
    S eig1 = 0.5 * (cvel + fabs_s(cvel));
    S eig2 = cvel + cc;
    S x21 = -ry * eig1 / (cc * cc);
    S x31 = rx * eig1 / (cc * cc);
    S x12 = 0.5 * eig2 / cc;
    S x22 = 0.5 * eig2 * (q2 * eig2 + beta * rx) / (beta * cc * cc);
    S x32 = 0.5 * eig2 * (q3 * eig2 + beta * ry) / (beta * cc * cc);
    S xi11 = ry * q2 - rx * q3;
    S xi12 = -q3 * cvel - ry * beta;
    S xi13 = q2 * cvel + rx * beta;
    S xi21 = cc - cvel;
    S xi22 = rx * beta;
    S xi23 = ry * beta;
    S ap[9], am[9];
    ap[0] = x12 * xi21;
    ap[1] = x21 * xi11 + x22 * xi21;
    ap[2] = x31 * xi11 + x32 * xi21;
    ap[3] = x12 * xi22;
    ap[4] = x21 * xi12 + x22 * xi22;
    ap[5] = x31 * xi12 + x32 * xi22;
    ap[6] = x12 * xi23;
    ap[7] = x21 * xi13 + x22 * xi23;
    ap[8] = x31 * xi13 + x32 * xi23;
    am[0] = -ap[0];
    am[1] = -ap[1] + rx;
    am[2] = -ap[2] + ry;
    am[3] = -ap[3] + rx * beta;
    am[4] = -ap[4] + 2. * rx * q2 + ry * q3;
    am[5] = -ap[5] + rx * q3;
    am[6] = -ap[6] + ry * beta;
    am[7] = -ap[7] + ry * q2;
    am[8] = -ap[8] + rx * q2 + 2. * ry * q3;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        dfp[m] = ap[m] * dq1 + ap[m + 3] * dq2 + ap[m + 6] * dq3;
        dfm[m] = am[m] * dq1 + am[m + 3] * dq2 + am[m + 6] * dq3;
    }
}

template <typename S>
SMAC_DEV void dir_resid(const LineS<S> &L, int i, int n, double beta, double cdis, int ivis,
                        double vnu, S out[3])
{
    S dfp[4][3], dfm[4][3];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int f = i - 2 + s;
        if (!(f >= 0 && f <= n - 2)) {
#pragma unroll
            for (int m = 0; m < 3; ++m) dfp[s][m] = dfm[s][m] = S(L.p[0] * 0.0);
            continue;
        }
        const int a = s, b = s + 1;
        S rx = 0.5 * (L.rx[b] + L.rx[a]);
        S ry = 0.5 * (L.ry[b] + L.ry[a]);
        S q2 = 0.5 * (L.u[b] + L.u[a]);
        S q3 = 0.5 * (L.v[b] + L.v[a]);
        face_dq(rx, ry, q2, q3, beta, L.p[b] - L.p[a], L.u[b] - L.u[a], L.v[b] - L.v[a],
                dfp[s], dfm[s]);
    }
    const double r6 = (double)(1.0f / 6.0f);
    S ff[2][3];
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const int f = i - 1 + w;
        const int fm = f - 1 > 0 ? f - 1 : 0;
        const int fp = f + 1 < n - 1 ? f + 1 : n - 1;
        const int fpp = f + 2 < n - 1 ? f + 2 : n - 1;
        const float cp = (float)(fpp - fp), cm = (float)(f - fm);
        const int sm = fm - (i - 2), sc = f - (i - 2), sp = fp - (i - 2);
        const int pf = f - (i - 2), pp = fp - (i - 2);
        S cv0 = L.rx[pf] * L.u[pf] + L.ry[pf] * L.v[pf];
        S cv1 = L.rx[pp] * L.u[pp] + L.ry[pp] * L.v[pp];
        S fv0[3] = {beta * cv0, L.u[pf] * cv0 + L.rx[pf] * L.p[pf], L.v[pf] * cv0 + L.ry[pf] * L.p[pf]};
        S fv1[3] = {beta * cv1, L.u[pp] * cv1 + L.rx[pp] * L.p[pp], L.v[pp] * cv1 + L.ry[pp] * L.p[pp]};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            S fmn = (dfp[sc][m] + dfm[sc][m] - dfp[sp][m] - dfm[sp][m]) * r6;
            S fpl = (dfp[sm][m] + dfm[sm][m] - dfp[sc][m] - dfm[sc][m]) * r6;
            S f3 = (dfp[sm][m] - dfp[sc][m] + dfm[sc][m] - dfm[sp][m]) * r6;
            ff[w][m] = 0.5 * (fv1[m] + fv0[m]) +
                       cdis * ((double)(cp * cm) * f3 + (double)(cm * (1.f - cp)) * fpl +
                               (double)(cp * (1.f - cm)) * fmn);
        }
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) out[m] = ff[0][m] - ff[1][m];
    if (ivis == 1) {
        S vf[2], ud[2], vd[2];
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const int a = 1 + w, b = 2 + w;
            S djv = 0.5 * (L.dj[a] * vnu + L.dj[b] * vnu);
            ud[w] = r32(L.u[b] - L.u[a]);
            vd[w] = r32(L.v[b] - L.v[a]);
            S rx = 0.5 * (L.rx[a] + L.rx[b]);
            S ry = 0.5 * (L.ry[a] + L.ry[b]);
            vf[w] = r32(djv * (rx * rx + ry * ry));
        }
        out[1] = out[1] + r32(r32(vf[1] * ud[1]) - r32(vf[0] * ud[0]));
        out[2] = out[2] + r32(r32(vf[1] * vd[1]) - r32(vf[0] * vd[0]));
    }
}

// row types of the steady equations (built on the host from the BC regions in file order)
enum { RT_INTERIOR = 0, RT_OUT_LO = 1, RT_OUT_HI = 2, RT_WALL = 3, RT_FAR = 4, RT_WAKE = 5 };

// the 9 offsets of the cross stencil: (dj, dk)
__constant__ int c_off[9][2] = {{0, 0}, {-1, 0}, {1, 0}, {-2, 0}, {2, 0},
                                {0, -1}, {0, 1}, {0, -2}, {0, 2}};
SMAC_DEV int off_index(int dj, int dk)
{
    if (dk == 0) return dj == 0 ? 0 : (dj == -1 ? 1 : dj == 1 ? 2 : dj == -2 ? 3 : 4);
    return dk == -1 ? 5 : dk == 1 ? 6 : dk == -2 ? 7 : 8;
}

struct AdjParams {
    double beta, cdis, vnu, pout, alpha, clift;
    int ivis;
};

// the line of point (j,k) in one direction, with the variables of grid point g seeded
SMAC_DEV void load_line_seeded(const Dims &d, const double *q, const double *mx,
                               const double *my, const double *dj, int j, int k, int eta,
                               int g, LineS<Dn<3>> &L)
{
    const long N = d.N;
    const int n = eta ? d.km : d.jm, i = eta ? k : j;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        int ii = i + t - 2;
        ii = ii < 0 ? 0 : (ii > n - 1 ? n - 1 : ii);
        const long a = eta ? P(d, j, ii) : P(d, ii, k);
        L.p[t] = dn<3>(q[a]);
        L.u[t] = dn<3>(q[N + a]);
        L.v[t] = dn<3>(q[2 * N + a]);
        L.rx[t] = dn<3>(mx[a]);
        L.ry[t] = dn<3>(my[a]);
        L.dj[t] = dn<3>(dj[a]);
        if (ii == g) {
            L.p[t].d[0] = 1.0;
            L.u[t].d[1] = 1.0;
            L.v[t].d[2] = 1.0;
        }
    }
}

// Jacobian blocks of the steady equations: jac[(o*9 + m + 3n)*N + p] = dF_(p,m) / dq_(p+o,n)
__global__ void __launch_bounds__(128)
k_adj_jac(Dims d, AdjParams prm, const signed char *__restrict__ rtype,
          const double *__restrict__ x, const double *__restrict__ y,
          const double *__restrict__ met, const double *__restrict__ dj,
          const double *__restrict__ q, double *__restrict__ jac)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (j >= d.jm || k >= d.km) return;
    const long N = d.N, p = P(d, j, k);
    for (int e = 0; e < 81; ++e) jac[e * N + p] = 0.0;
    auto put = [&](int o, int m, int n, double v) { jac[(o * 9 + m + 3 * n) * N + p] += v; };
    const int t = rtype[p];
    if (t == RT_INTERIOR) {
        for (int eta = 0; eta < 2; ++eta) {
            const double *mx = met + (eta ? 2 : 0) * N, *my = met + (eta ? 3 : 1) * N;
            const int n = eta ? d.km : d.jm, i = eta ? k : j;
            for (int o = -2; o <= 2; ++o) {
                const int g = i + o;
                if (g < 0 || g > n - 1) continue;
                LineS<Dn<3>> L;
                load_line_seeded(d, q, mx, my, dj, j, k, eta, g, L);
                Dn<3> r[3];
                dir_resid(L, i, n, prm.beta, prm.cdis, prm.ivis, prm.vnu, r);
                const int oi = eta ? off_index(0, o) : off_index(o, 0);
                for (int m = 0; m < 3; ++m)
                    for (int nn = 0; nn < 3; ++nn) put(oi, m, nn, r[m].d[nn]);
            }
        }
    } else if (t == RT_OUT_LO || t == RT_OUT_HI) {
        // pout - p, u(nb) - u, v(nb) - v
        const int on = t == RT_OUT_LO ? off_index(1, 0) : off_index(-1, 0);
        put(0, 0, 0, -1.0);
        put(0, 1, 1, -1.0);
        put(0, 2, 2, -1.0);
        put(on, 1, 1, 1.0);
        put(on, 2, 2, 1.0);
    } else if (t == RT_WALL) {
        // p(k+1) - p, -u, -v
        put(0, 0, 0, -1.0);
        put(off_index(0, 1), 0, 0, 1.0);
        put(0, 1, 1, -1.0);
        put(0, 2, 2, -1.0);
    } else if (t == RT_FAR) {
        // u, v: free stream + vortex(clift) - u; p: the outgoing characteristic row
        put(0, 1, 1, -1.0);
        put(0, 2, 2, -1.0);
        const long pd = P(d, j, k - 1);
        const double r2 = met[2 * N + p], r3 = met[3 * N + p];
        Dn<3> qu = dn<3>(q[N + p]), qv = dn<3>(q[2 * N + p]);
        qu.d[1] = 1.0;
        qv.d[2] = 1.0;
        Dn<3> qq = r2 * qu + r3 * qv;
        Dn<3> cc = sqrt_s(qq * qq + prm.beta * (r2 * r2 + r3 * r3));
        Dn<3> b11 = dn<3>(-(r2 * r2 + r3 * r3));
        Dn<3> b12 = -r2 * (qq + cc);
        Dn<3> b13 = -r3 * (qq + cc);
        Dn<3> dq1 = dn<3>(q[p] - q[pd]);
        Dn<3> dq2 = qu - q[N + pd];
        Dn<3> dq3 = qv - q[2 * N + pd];
        dq1.d[0] = 1.0;
        Dn<3> s1 = b11 * dq1 + b12 * dq2 + b13 * dq3;
        for (int nn = 0; nn < 3; ++nn) put(0, 0, nn, s1.d[nn]);
        const int od = off_index(0, -1);
        put(od, 0, 0, -b11.v);
        put(od, 0, 1, -b12.v);
        put(od, 0, 2, -b13.v);
    } else if (t == RT_WAKE) {
        // (q(jt,2) + q(jb,2)) / 2 - q(jt,1); the (jb,2) part is applied by k_adj_wake_t
        for (int m = 0; m < 3; ++m) {
            put(0, m, m, -1.0);
            put(off_index(0, 1), m, m, 0.5);
        }
    }
}


// out = F_q^T lam over the 9 offsets (the wake cut's far partner is added by k_adj_wake_t)
__global__ void __launch_bounds__(256)
k_adj_jtv(Dims d, const double *__restrict__ jac, const double *__restrict__ lam,
          double *__restrict__ out)
{
    const Blk bk = xcd_blk();   // the rows of the 9 offsets in this XCD's L2
    const int j = bk.x * blockDim.x + threadIdx.x, k = bk.y;
    if (j >= d.jm || k >= d.km) return;
    const long N = d.N, p = P(d, j, k);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int o = 0; o < 9; ++o) {
        const int jr = j - c_off[o][0], kr = k - c_off[o][1];   // the row whose offset o hits p
        if (jr < 0 || jr >= d.jm || kr < 0 || kr >= d.km) continue;
        const long r = P(d, jr, kr);
        const double l0 = lam[r], l1 = lam[N + r], l2 = lam[2 * N + r];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const long b = (long)(o * 9 + 3 * n) * N + r;
            acc[n] += jac[b] * l0 + jac[b + N] * l1 + jac[b + 2 * N] * l2;
        }
    }
#pragma unroll
    for (int n = 0; n < 3; ++n) out[n * N + p] = acc[n];
}

// the wake rows' dependence on their partner across the cut: out(jb,2) += lam(jt,1)/2.  Every
// point of the cut is the partner of exactly one pair, so no two threads write one entry.
__global__ void k_adj_wake_t(Dims d, int npair, const int *__restrict__ pairs,
                             const double *__restrict__ lam, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npair) return;
    const long jt = pairs[2 * i], jb = pairs[2 * i + 1];
    const long N = d.N;
    if (jt < 0 || jt >= d.jm || jb < 0 || jb >= d.jm) return;
    for (int n = 0; n < 3; ++n) out[n * N + jb + d.jm] += 0.5 * lam[n * N + jt];
}

// W^T r, given v = r with the targets cleared: a target's entry moves, halved, to (jt,2) and
// (jb,2).  Every cut point is the target of one pair and the partner of one other (inv[i]:
// the pair whose partner is jt of pair i), so thread i gathers both halves for (jt_i, 2).
__global__ void k_adj_wake_wt(Dims d, int npair, const int *__restrict__ pairs,
                              const int *__restrict__ inv, const double *__restrict__ r,
                              double *__restrict__ v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npair) return;
    const long jt = pairs[2 * i], jo = pairs[2 * inv[i]];
    const long N = d.N;
    for (int n = 0; n < 3; ++n)
        v[n * N + jt + d.jm] += 0.5 * r[n * N + jt] + 0.5 * r[n * N + jo];
}

// P's wake rows: out(target) = (z(jt,2) + z(jb,2)) / 2 + u(target), z elsewhere
__global__ void k_adj_wake_w(Dims d, int npair, const int *__restrict__ pairs,
                             const double *__restrict__ z, const double *__restrict__ u,
                             double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npair) return;
    const long jt = pairs[2 * i], jb = pairs[2 * i + 1];
    const long N = d.N;
    for (int n = 0; n < 3; ++n)
        out[n * N + jt] = 0.5 * (z[n * N + jt + d.jm] + z[n * N + jb + d.jm]) + u[n * N + jt];
}

// F_c (far-field u, v rows: the vortex velocity per unit clift) and F_alpha (per degree)
__global__ void k_adj_fc(Dims d, double alpha, const signed char *__restrict__ rtype,
                         const double *__restrict__ x, const double *__restrict__ y,
                         double *__restrict__ fc, double *__restrict__ fa)
{
    const long N = d.N;
    const double pi = 4.0 * atan(1.0);
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
         p += (long)gridDim.x * blockDim.x) {
        double c1 = 0.0, c2 = 0.0, a1 = 0.0, a2 = 0.0;
        if (rtype[p] == RT_FAR) {
            const double dx = x[p] - 0.25, dy = y[p];
            const double rr = 4. * pi * (dx * dx + dy * dy);
            c1 = dy / rr;
            c2 = -dx / rr;
            a1 = -sin(alpha * pi / 180.) * (pi / 180.);
            a2 = cos(alpha * pi / 180.) * (pi / 180.);
        }
        fc[p] = 0.0;
        fc[N + p] = c1;
        fc[2 * N + p] = c2;
        fa[p] = 0.0;
        fa[N + p] = a1;
        fa[2 * N + p] = a2;
    }
}

// out = -out - s[0] * z  (the rank-one clift coupling of the adjoint operator)
__global__ void k_adj_negate_rank1(long len, const double *__restrict__ s,
                                   const double *__restrict__ z, double *__restrict__ out)
{
    const double c = s[0];
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len;
         e += (long)gridDim.x * blockDim.x)
        out[e] = -out[e] - c * z[e];
}

// out = target ? u : z
__global__ void k_adj_merge_targets(long N, const signed char *__restrict__ rtype,
                                    const double *__restrict__ u, const double *__restrict__ z,
                                    double *__restrict__ out)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < 3 * N;
         e += (long)gridDim.x * blockDim.x)
        out[e] = rtype[e % N] == RT_WAKE ? u[e] : z[e];
}

// the block arrays of M^T's LU-SGS form (see the file header)
__global__ void __launch_bounds__(256)
k_adj_transpose_blocks(Dims d, const double *__restrict__ btc, const double *__restrict__ bjm,
                       const double *__restrict__ bjp, const double *__restrict__ bkm,
                       const double *__restrict__ bkp, double *__restrict__ tc,
                       double *__restrict__ tjm, double *__restrict__ tjp,
                       double *__restrict__ tkm, double *__restrict__ tkp)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (j >= d.jm || k >= d.km) return;
    const long N = d.N, p = P(d, j, k);
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const int e = m + 3 * n, et = n + 3 * m;   // (m,n) of the transpose = (n,m)
            tc[e * N + p] = btc[et * N + p];
            tjm[e * N + p] = j > 0 ? bjp[et * N + p - 1] : 0.0;
            tkm[e * N + p] = k > 0 ? bkp[et * N + p - d.jm] : 0.0;
            tjp[e * N + p] = j < d.jm - 1 ? bjm[et * N + p + 1] : 0.0;
            tkp[e * N + p] = k < d.km - 1 ? bkm[et * N + p + d.jm] : 0.0;
        }
}

// ------------------------------------------------------------------------------------------
// forces: one wall face's contribution to Fx, Fy (pressure, then laminar shear when ivis),
// as the force kernel forms it (ref/fm.f:210-309)

template <typename S>
struct FaceIn {
    S p[2], u[2][3], v[2][3];      // point a of the face (0, 1), k = wall, wall+1, wall+2
    S r22[2], r23[2], dj[2];
};

template <typename S>
SMAC_DEV void face_force(const FaceIn<S> &F, int ivis, double vnu, S &fx, S &fy)
{
    S da1 = 0.5 * (F.r22[1] + F.r22[0]);
    S da2 = 0.5 * (F.r23[1] + F.r23[0]);
    S pave = 0.5 * (F.p[1] + F.p[0]);
    fx = -(pave * da1);
    fy = -(pave * da2);
    if (ivis != 0) {
        S tt[2][3];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            S ueta = 0.5 * (-3. * F.u[a][0] + 4. * F.u[a][1] - F.u[a][2]);
            S veta = 0.5 * (-3. * F.v[a][0] + 4. * F.v[a][1] - F.v[a][2]);
            tt[a][0] = 2. * vnu * F.r22[a] * F.dj[a] * ueta;
            tt[a][1] = 2. * vnu * F.r23[a] * F.dj[a] * veta;
            tt[a][2] = vnu * (F.r23[a] * ueta + F.r22[a] * veta) * F.dj[a];
        }
        S t11 = 0.5 * (tt[1][0] + tt[0][0]);
        S t22 = 0.5 * (tt[1][1] + tt[0][1]);
        S t12 = 0.5 * (tt[1][2] + tt[0][2]);
        fx = fx + (t11 * da1 + t12 * da2);
        fy = fy + (t12 * da1 + t22 * da2);
    }
}

// CL_q and CD_q: one thread per wall point, its nine variables (p, u, v at the wall and the two
// points above) seeded, the two faces it belongs to summed
__global__ void k_adj_force_grad(Dims d, int ivis, double vnu, double alpha, double chord,
                                 Walls wl, const double *__restrict__ met,
                                 const double *__restrict__ dj, const double *__restrict__ q,
                                 double *__restrict__ gcl, double *__restrict__ gcd)
{
    const long N = d.N;
    const double pi = 4.0 * atan(1.0);
    const double ca = cos(alpha * pi / 180.0), sa = sin(alpha * pi / 180.0);
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    for (int w = 0; w < wl.n; ++w) {
        const int np = wl.j2[w] - wl.j1[w] + 1;
        if (idx >= np) {
            idx -= np;
            continue;
        }
        const int j = wl.j1[w] + idx, k = wl.k[w];
        Dn<9> fxs = dn<9>(0.0), fys = dn<9>(0.0);
        for (int face = j - 1; face <= j; ++face) {
            if (face < wl.j1[w] || face > wl.j2[w] - 1) continue;
            FaceIn<Dn<9>> F;
            for (int a = 0; a < 2; ++a) {
                const int ja = face + a;
                const long pa = P(d, ja, k);
                F.p[a] = dn<9>(q[pa]);
                for (int kk = 0; kk < 3; ++kk) {
                    F.u[a][kk] = dn<9>(q[N + pa + (long)kk * d.jm]);
                    F.v[a][kk] = dn<9>(q[2 * N + pa + (long)kk * d.jm]);
                }
                F.r22[a] = dn<9>(met[2 * N + pa]);
                F.r23[a] = dn<9>(met[3 * N + pa]);
                F.dj[a] = dn<9>(dj[pa]);
                if (ja == j) {
                    F.p[a].d[0] = 1.0;
                    for (int kk = 0; kk < 3; ++kk) {
                        F.u[a][kk].d[1 + 3 * kk] = 1.0;
                        F.v[a][kk].d[2 + 3 * kk] = 1.0;
                    }
                }
            }
            Dn<9> fx, fy;
            face_force(F, ivis, vnu, fx, fy);
            fxs = fxs + fx;
            fys = fys + fy;
        }
        // variable v of the seed list: (kk, n) = (v / 3, v % 3), kk the height above the wall
        for (int v = 0; v < 9; ++v) {
            const int kk = v / 3, n = v % 3;
            const long pv = n * N + P(d, j, k + kk);
            gcl[pv] += (2. * fys.d[v] * ca - 2. * fxs.d[v] * sa) / chord;
            gcd[pv] += (2. * fxs.d[v] * ca + 2. * fys.d[v] * sa) / chord;
        }
        return;
    }
}

// tangent of the metric planes (xi_x, xi_y, eta_x, eta_y) and of dj along a grid displacement
// (dx, dy), as the metric is formed: second-order central differences, second-order one-sided
// differences at the edges, dj = 1 / (x_xi y_eta - x_eta y_xi)
__global__ void __launch_bounds__(256)
k_adj_metric_tangent(Dims d, const double *__restrict__ x, const double *__restrict__ y,
                     const double *__restrict__ dx, const double *__restrict__ dy,
                     double *__restrict__ dmet, double *__restrict__ ddj)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (j >= d.jm || k >= d.km) return;
    const long N = d.N, p = P(d, j, k);
    auto dxi = [&](const double *a, int jj) -> double {
        const long c = P(d, jj, k);
        if (jj == 0) return 0.5 * ((-a[c + 2] + 4.0 * a[c + 1]) - 3.0 * a[c]);
        if (jj == d.jm - 1) return 0.5 * ((3.0 * a[c] - 4.0 * a[c - 1]) + a[c - 2]);
        return 0.5 * (a[c + 1] - a[c - 1]);
    };
    auto deta = [&](const double *a, int kk) -> double {
        const long c = P(d, j, kk), s = d.jm;
        if (kk == 0) return 0.5 * ((-a[c + 2 * s] + 4.0 * a[c + s]) - 3.0 * a[c]);
        if (kk == d.km - 1) return 0.5 * ((3.0 * a[c] - 4.0 * a[c - s]) + a[c - 2 * s]);
        return 0.5 * (a[c + s] - a[c - s]);
    };
    const double xj = dxi(x, j), yj = dxi(y, j), xk = deta(x, k), yk = deta(y, k);
    const double dxj = dxi(dx, j), dyj = dxi(dy, j), dxk = deta(dx, k), dyk = deta(dy, k);
    const double rdj = xj * yk - xk * yj;
    const double drdj = dxj * yk + xj * dyk - dxk * yj - xk * dyj;
    dmet[p] = dyk;           // xi_x  = y_eta
    dmet[N + p] = -dxk;      // xi_y  = -x_eta
    dmet[2 * N + p] = -dyj;  // eta_x = -y_xi
    dmet[3 * N + p] = dxj;   // eta_y = x_xi
    ddj[p] = -drdj / (rdj * rdj);
}

// tangent of F at fixed q along (metric, dj, x, y) tangents
__global__ void __launch_bounds__(128)
k_adj_resid_tangent(Dims d, AdjParams prm, const signed char *__restrict__ rtype,
                    const double *__restrict__ x, const double *__restrict__ y,
                    const double *__restrict__ dx, const double *__restrict__ dy,
                    const double *__restrict__ met, const double *__restrict__ dmet,
                    const double *__restrict__ dj, const double *__restrict__ ddj,
                    const double *__restrict__ q, double *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (j >= d.jm || k >= d.km) return;
    const long N = d.N, p = P(d, j, k);
    double r[3] = {0.0, 0.0, 0.0};
    const int t = rtype[p];
    if (t == RT_INTERIOR) {
        for (int eta = 0; eta < 2; ++eta) {
            const int n = eta ? d.km : d.jm, i = eta ? k : j;
            const int cx = eta ? 2 : 0, cy = eta ? 3 : 1;
            LineS<Dn<1>> L;
#pragma unroll
            for (int tt = 0; tt < 5; ++tt) {
                int ii = i + tt - 2;
                ii = ii < 0 ? 0 : (ii > n - 1 ? n - 1 : ii);
                const long a = eta ? P(d, j, ii) : P(d, ii, k);
                L.p[tt] = dn<1>(q[a]);
                L.u[tt] = dn<1>(q[N + a]);
                L.v[tt] = dn<1>(q[2 * N + a]);
                L.rx[tt] = dn<1>(met[cx * N + a]);
                L.rx[tt].d[0] = dmet[cx * N + a];
                L.ry[tt] = dn<1>(met[cy * N + a]);
                L.ry[tt].d[0] = dmet[cy * N + a];
                L.dj[tt] = dn<1>(dj[a]);
                L.dj[tt].d[0] = ddj[a];
            }
            Dn<1> o[3];
            dir_resid(L, i, n, prm.beta, prm.cdis, prm.ivis, prm.vnu, o);
            for (int m = 0; m < 3; ++m) r[m] += o[m].d[0];
        }
    } else if (t == RT_FAR) {
        const double pi = 4.0 * atan(1.0);
        Dn<1> X = dn<1>(x[p] - 0.25), Y = dn<1>(y[p]);
        X.d[0] = dx[p];
        Y.d[0] = dy[p];
        Dn<1> rr = 4. * pi * (X * X + Y * Y);
        r[1] = (prm.clift * Y / rr).d[0];
        r[2] = (-prm.clift * X / rr).d[0];
        const long pd = P(d, j, k - 1);
        Dn<1> r2 = dn<1>(met[2 * N + p]), r3 = dn<1>(met[3 * N + p]);
        r2.d[0] = dmet[2 * N + p];
        r3.d[0] = dmet[3 * N + p];
        Dn<1> qq = r2 * q[N + p] + r3 * q[2 * N + p];
        Dn<1> cc = sqrt_s(qq * qq + prm.beta * (r2 * r2 + r3 * r3));
        Dn<1> s1 = -(r2 * r2 + r3 * r3) * (q[p] - q[pd]) - r2 * (qq + cc) * (q[N + p] - q[N + pd]) -
                   r3 * (qq + cc) * (q[2 * N + p] - q[2 * N + pd]);
        r[0] = s1.d[0];
    }
    for (int m = 0; m < 3; ++m) out[m * N + p] = r[m];
}

// tangent of the wall force sums along (metric, dj) tangents at fixed q: one thread per face,
// the face's (dFx, dFy) to part[face] (summed on the host)
__global__ void k_adj_force_tangent(Dims d, int ivis, double vnu, Walls wl,
                                    const double *__restrict__ met, const double *__restrict__ dmet,
                                    const double *__restrict__ dj, const double *__restrict__ ddj,
                                    const double *__restrict__ q, double *__restrict__ part)
{
    const long N = d.N;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int base = 0;
    for (int w = 0; w < wl.n; ++w) {
        const int nf = wl.j2[w] - wl.j1[w];
        if (idx >= nf) {
            idx -= nf;
            base += nf;
            continue;
        }
        const int face = wl.j1[w] + idx, k = wl.k[w];
        FaceIn<Dn<1>> F;
        for (int a = 0; a < 2; ++a) {
            const long pa = P(d, face + a, k);
            F.p[a] = dn<1>(q[pa]);
            for (int kk = 0; kk < 3; ++kk) {
                F.u[a][kk] = dn<1>(q[N + pa + (long)kk * d.jm]);
                F.v[a][kk] = dn<1>(q[2 * N + pa + (long)kk * d.jm]);
            }
            F.r22[a] = dn<1>(met[2 * N + pa]);
            F.r22[a].d[0] = dmet[2 * N + pa];
            F.r23[a] = dn<1>(met[3 * N + pa]);
            F.r23[a].d[0] = dmet[3 * N + pa];
            F.dj[a] = dn<1>(dj[pa]);
            F.dj[a].d[0] = ddj[pa];
        }
        Dn<1> fx, fy;
        face_force(F, ivis, vnu, fx, fy);
        part[2 * (base + idx)] = fx.d[0];
        part[2 * (base + idx) + 1] = fy.d[0];
        return;
    }
}

// ------------------------------------------------------------------------------------------
// vector kernels of the GMRES driver

// partial dot products of w with each of the nv vectors V[i] (stride len), per block
// nv dot products V_i . w: each pass over the elements takes MDG vectors at once (w read once
// per group); wavefront shuffles and one LDS step reduce each block's sums to per-block
// partials, which k_adj_msum adds in a fixed order (deterministic).  A single product takes
// the plain one-vector form.
constexpr int MDG = 8;

SMAC_DEV double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ void __launch_bounds__(256)
k_adj_mdot(long len, int nv, const double *__restrict__ V, const double *__restrict__ w,
           double *__restrict__ part, const int *__restrict__ gate)
{
    if (gate && *gate == 0) return;
    __shared__ double sh[4][MDG];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nb = gridDim.x;
    for (int i0 = 0; i0 < nv; i0 += MDG) {
        const int g = nv - i0 < MDG ? nv - i0 : MDG;
        double acc[MDG];
#pragma unroll
        for (int k = 0; k < MDG; ++k) acc[k] = 0.0;
        const double *v0 = V + (long)i0 * len;
        if (g == MDG) {
            for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len;
                 e += (long)gridDim.x * blockDim.x) {
                const double we = w[e];
#pragma unroll
                for (int k = 0; k < MDG; ++k) acc[k] += v0[(long)k * len + e] * we;
            }
        } else {
            for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len;
                 e += (long)gridDim.x * blockDim.x) {
                const double we = w[e];
#pragma unroll
                for (int k = 0; k < MDG; ++k)
                    if (k < g) acc[k] += v0[(long)k * len + e] * we;
            }
        }
#pragma unroll
        for (int k = 0; k < MDG; ++k) {
            const double t = wave_sum(acc[k]);
            if (lane == 0) sh[wv][k] = t;
        }
        __syncthreads();
        if ((int)threadIdx.x < g)
            part[(long)(i0 + threadIdx.x) * nb + blockIdx.x] =
                sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256)
k_adj_dot1(long len, const double *__restrict__ v, const double *__restrict__ w,
           double *__restrict__ part)
{
    __shared__ double sh[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double acc = 0.0;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len;
         e += (long)gridDim.x * blockDim.x)
        acc += v[e] * w[e];
    acc = wave_sum(acc);
    if (lane == 0) sh[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// out[i] = the sum of product i's nblk partials, one wavefront per product (four per block)
__global__ void __launch_bounds__(256)
k_adj_msum(int nv, int nblk, const double *__restrict__ part, double *__restrict__ out,
           const int *__restrict__ gate)
{
    const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nv) return;
    if (gate && *gate == 0) {
        if (lane == 0) out[i] = 0.0;
        return;
    }
    double a = 0.0;
    for (int b = lane; b < nblk; b += 64) a += part[(long)i * nblk + b];
    a = wave_sum(a);
    if (lane == 0) out[i] = a;
}

__global__ void __launch_bounds__(256)
k_adj_maxpy(long len, int nv, double a, double cs, const double *__restrict__ V,
            const double *__restrict__ c, double *__restrict__ w, const int *__restrict__ gate)
{
    if (gate && *gate == 0) return;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len;
         e += (long)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int i = 0; i < nv; ++i) acc += c[i] * V[(long)i * len + e];
        w[e] = a * w[e] + cs * acc;
    }
}

// y = a*x + b*y + g*z (any of x, z may be null)
__global__ void __launch_bounds__(256)
k_adj_axpby(long len, double a, const double *__restrict__ x, double b, double *__restrict__ y,
            double g, const double *__restrict__ z)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len;
         e += (long)gridDim.x * blockDim.x) {
        double v = b * y[e];
        if (x) v += a * x[e];
        if (z) v += g * z[e];
        y[e] = v;
    }
}

// out = in with the wake targets' entries cleared (keep = 0) or with only them kept (keep = 1)
__global__ void k_adj_mask_targets(long N, const signed char *__restrict__ rtype, int keep,
                                   const double *__restrict__ in, double *__restrict__ out)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < 3 * N;
         e += (long)gridDim.x * blockDim.x) {
        const bool tg = rtype[e % N] == RT_WAKE;
        out[e] = (tg == (keep != 0)) ? in[e] : 0.0;
    }
}


// ------------------------------------------------------------------------------------------
// host side

static unsigned nblk(long n, long cap = 2048)
{
    long b = (n + 255) / 256;
    return (unsigned)(b > cap ? cap : b);
}

constexpr int ADJ_DOT_BLOCKS = 512;

// nv dot products of w with V[0..nv-1] into out[0..nv-1] (device); with a gate, nothing is
// computed (and zeros are written) when *gate == 0
static void mdot(AdjCtx &c, const double *V, int nv, const double *w, double *out,
                 const int *gate = nullptr)
{
    const long len = 3 * c.d.N;
    if (nv == 1)
        hipLaunchKernelGGL(k_adj_dot1, dim3(ADJ_DOT_BLOCKS), dim3(256), 0, c.sm, len, V, w, c.part);
    else
        hipLaunchKernelGGL(k_adj_mdot, dim3(ADJ_DOT_BLOCKS), dim3(256), 0, c.sm, len, nv, V, w,
                           c.part, gate);
    hipLaunchKernelGGL(k_adj_msum, dim3((nv + 3) / 4), dim3(256), 0, c.sm, nv, ADJ_DOT_BLOCKS,
                       c.part, out, gate);
}

void adj_build(AdjCtx &c)
{
    const Dims &d = c.d;
    const long N = d.N;
    AdjParams prm{c.beta, c.cdis, c.vnu, c.pout, c.alpha, c.clift, c.ivis};
    const dim3 g2((d.jm + 127) / 128, d.km), g3((d.jm + 255) / 256, d.km);
    hipLaunchKernelGGL(k_adj_jac, g2, dim3(128), 0, c.sm, d, prm, c.rtype, c.x, c.y, c.met, c.dj,
                       c.q, c.jac);
    hipLaunchKernelGGL(k_adj_transpose_blocks, g3, dim3(256), 0, c.sm, d, c.blk[0], c.blk[1],
                       c.blk[2], c.blk[3], c.blk[4], c.tb[0], c.tb[1], c.tb[2], c.tb[3], c.tb[4]);
    if (c.lr) {
        // the plain factorisation: factored once and swept thousands of times, the one-sided
        // substitution (one scan each way) is the cheaper one here
        launch_lr_factor3<double>(c.sm, d, c.njsp > 0, c.nksp > 0, false, c.blk[0], c.blk[1],
                                  c.blk[2], c.blk[3], c.blk[4], (double *)c.lr_f.opx,
                                  (double *)c.lr_f.ope);
        launch_lr_factor3<double>(c.sm, d, c.njsp > 0, c.nksp > 0, false, c.tb[0], c.tb[1],
                                  c.tb[2], c.tb[3], c.tb[4], (double *)c.lr_t.opx,
                                  (double *)c.lr_t.ope);
    }
    for (int o = 0; o < 2; ++o)
        hipLaunchKernelGGL(k_adj_axpby, dim3(nblk(3 * N)), dim3(256), 0, c.sm, 3 * N, 0.0,
                           (const double *)nullptr, 0.0, c.g[o], 0.0, (const double *)nullptr);
    int nw = 0;
    for (int w = 0; w < c.walls.n; ++w) nw += c.walls.j2[w] - c.walls.j1[w] + 1;
    if (nw > 0)
        hipLaunchKernelGGL(k_adj_force_grad, dim3((nw + 63) / 64), dim3(64), 0, c.sm, d, c.ivis,
                           c.vnu, c.alpha, c.chord, c.walls, c.met, c.dj, c.q, c.g[0], c.g[1]);
    hipLaunchKernelGGL(k_adj_fc, dim3(nblk(N)), dim3(256), 0, c.sm, d, c.alpha, c.rtype, c.x, c.y,
                       c.fc, c.fa);
    // the rank-one correction of the preconditioner: u = P^T g and 1 - f . u (skipped when
    // that is too close to zero to divide by)
    adj_apply_prec_t(c, c.g[0], c.smu);
    mdot(c, c.fc, 1, c.smu, c.red);
    double fu = 0.0;
    (void)hipMemcpyAsync(&fu, c.red, sizeof fu, hipMemcpyDeviceToHost, c.sm);
    (void)hipStreamSynchronize(c.sm);
    const double beta = 1.0 - fu;
    c.sm_on = std::fabs(beta) > 1e-8 ? 1 : 0;
    c.sm_binv = c.sm_on ? 1.0 / beta : 0.0;
}

// out = F_q^T in
void adj_apply_jt(AdjCtx &c, const double *in, double *out)
{
    const Dims &d = c.d;
    hipLaunchKernelGGL(k_adj_jtv, dim3((d.jm + 255) / 256, d.km), dim3(256), 0, c.sm, d, c.jac,
                       in, out);
    if (c.npair > 0)
        hipLaunchKernelGGL(k_adj_wake_t, dim3((c.npair + 63) / 64), dim3(64), 0, c.sm, d, c.npair,
                           c.pairs, in, out);
}

// out = -F_q^T in - (in . F_c) CL_q
void adj_apply_op(AdjCtx &c, const double *in, double *out)
{
    const Dims &d = c.d;
    hipLaunchKernelGGL(k_adj_jtv, dim3((d.jm + 255) / 256, d.km), dim3(256), 0, c.sm, d, c.jac,
                       in, out);
    if (c.npair > 0)
        hipLaunchKernelGGL(k_adj_wake_t, dim3((c.npair + 63) / 64), dim3(64), 0, c.sm, d, c.npair,
                           c.pairs, in, out);
    mdot(c, c.fc, 1, in, c.red);
    hipLaunchKernelGGL(k_adj_negate_rank1, dim3(nblk(3 * d.N)), dim3(256), 0, c.sm, 3 * d.N,
                       (const double *)c.red, (const double *)c.g[0], out);
}

// out = P^T in = (targets cleared) M^-T W^T in  +  (targets only) in
void adj_apply_prec_t(AdjCtx &c, const double *in, double *out)
{
    const Dims &d = c.d;
    const long N = d.N;
    hipLaunchKernelGGL(k_adj_mask_targets, dim3(nblk(3 * N)), dim3(256), 0, c.sm, N, c.rtype, 0,
                       in, c.t1);
    if (c.npair > 0)
        hipLaunchKernelGGL(k_adj_wake_wt, dim3((c.npair + 63) / 64), dim3(64), 0, c.sm, d,
                           c.npair, c.pairs, c.inv, in, c.t1);
    const WakePairs none{0, nullptr};
    if (c.lr)
        launch_lrelax_scan_t<double>(c.sm, d, c.njsp, c.nksp, c.dp.underr, c.t1, c.t2, c.lr_t,
                                     c.lrv);
    else
        launch_lusgs3<double>(c.sm, d, c.ws, none, c.q, c.t2, c.t1, c.tb[0], c.tb[1], c.tb[2],
                              c.tb[3], c.tb[4], nullptr, false);
    hipLaunchKernelGGL(k_adj_merge_targets, dim3(nblk(3 * N)), dim3(256), 0, c.sm, N, c.rtype,
                       in, c.t2, out);
}

// out = P in = W M^-1 (targets cleared) in  +  (targets only) in: the primal's own update
void adj_apply_prec(AdjCtx &c, const double *in, double *out)
{
    const Dims &d = c.d;
    const long N = d.N;
    hipLaunchKernelGGL(k_adj_mask_targets, dim3(nblk(3 * N)), dim3(256), 0, c.sm, N, c.rtype, 0,
                       in, c.t1);
    const WakePairs none{0, nullptr};
    if (c.lr)
        launch_lrelax_scan<double>(c.sm, d, c.dp, c.njsp, c.nksp, 1, true, none, c.q, out, c.t1,
                                   c.blk[0], c.blk[1], c.blk[2], c.blk[3], c.blk[4], c.lr_f,
                                   false, 0);
    else
        launch_lusgs3<double>(c.sm, d, c.ws, none, c.q, out, c.t1, c.blk[0], c.blk[1], c.blk[2],
                              c.blk[3], c.blk[4], nullptr, false);
    if (c.npair > 0)
        hipLaunchKernelGGL(k_adj_wake_w, dim3((c.npair + 63) / 64), dim3(64), 0, c.sm, d, c.npair,
                           c.pairs, out, in, out);
}

// gate = 1 when less than the fraction keep of the square norm red[n] of the vector survives
// the projections red[0..n-1] (then it is projected a second time; keep = 1/2 is the DGKS
// test, 0 never projects twice)
__global__ void k_adj_dgks(int n, const double *__restrict__ red, double keep, int *__restrict__ gate)
{
    double s2 = 0.0;
    for (int i = threadIdx.x; i < n; i += 64) s2 += red[i] * red[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s2 += __shfl_xor(s2, off, 64);
    if (threadIdx.x == 0) *gate = (red[n] - s2 < keep * red[n]) ? 1 : 0;
}

// out += u * red[0] * binv
__global__ void __launch_bounds__(256)
k_adj_sm_add(long len, const double *__restrict__ red, double binv, const double *__restrict__ u,
             double *__restrict__ out)
{
    const double a = red[0] * binv;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len;
         e += (long)gridDim.x * blockDim.x)
        out[e] += a * u[e];
}

// the preconditioner GMRES uses: P^T with the rank-one correction (see the file header)
static void adj_apply_prec_gm(AdjCtx &c, const double *in, double *out)
{
    adj_apply_prec_t(c, in, out);
    if (!c.sm_on) return;
    const long len = 3 * c.d.N;
    mdot(c, c.fc, 1, out, c.red);
    hipLaunchKernelGGL(k_adj_sm_add, dim3(nblk(len)), dim3(256), 0, c.sm, len,
                       (const double *)c.red, c.sm_binv, (const double *)c.smu, out);
}

// The GMRES preconditioner: psteps steps of the transposed update Q as a stationary
// iteration on Op, z = Q in, then z += Q (in - Op z) psteps - 1 times -- the polynomial
// sum_i (I - Q Op)^i Q, a fixed linear operator.  Each step costs one more Q and Op and cuts
// the GMRES iterations about as much (945x241: 4140 / 3455 with one step, 1961 / 1761 with
// two, 1301 / 1254 with three).
static void adj_apply_prec_poly(AdjCtx &c, const double *in, double *out)
{
    adj_apply_prec_gm(c, in, out);
    const long len = 3 * c.d.N;
    for (int st = 1; st < c.psteps; ++st) {
        adj_apply_op(c, out, c.pk1);
        hipLaunchKernelGGL(k_adj_axpby, dim3(nblk(len)), dim3(256), 0, c.sm, len, 1.0, in, -1.0,
                           c.pk1, 0.0, (const double *)nullptr);
        adj_apply_prec_gm(c, c.pk1, c.pk2);
        hipLaunchKernelGGL(k_adj_axpby, dim3(nblk(len)), dim3(256), 0, c.sm, len, 1.0, c.pk2, 1.0,
                           out, 0.0, (const double *)nullptr);
    }
}

// ||v||_2 on the host
static double norm2(AdjCtx &c, const double *v)
{
    mdot(c, v, 1, v, c.red);
    double h = 0.0;
    (void)hipMemcpyAsync(&h, c.red, sizeof h, hipMemcpyDeviceToHost, c.sm);
    (void)hipStreamSynchronize(c.sm);
    return std::sqrt(h);
}

// GMRES(m), right-preconditioned: solve Op x = b with x = P^T y.  Returns the iterations made;
// *rel receives the final true residual relative to ||b||.
int adj_gmres(AdjCtx &c, const double *b, double *x, int maxit, double tol, double *rel)
{
    const long len = 3 * c.d.N;
    const int m = c.m;
    const unsigned nb = nblk(len);
    hipLaunchKernelGGL(k_adj_axpby, dim3(nb), dim3(256), 0, c.sm, len, 0.0, (const double *)nullptr,
                       0.0, x, 0.0, (const double *)nullptr);
    const double bn = norm2(c, b);
    if (!(bn > 0.0)) {
        if (rel) *rel = 0.0;
        return 0;
    }
    // r = b - Op x  (x = 0 at the start)
    hipLaunchKernelGGL(k_adj_axpby, dim3(nb), dim3(256), 0, c.sm, len, 1.0, b, 0.0, c.w, 0.0,
                       (const double *)nullptr);
    int it = 0;
    double res = bn;
    std::vector<double> H((m + 1) * m), cs(m), sn(m), gv(m + 1), h(m + 1), h2(m + 1), y(m);
    while (it < maxit) {
        res = norm2(c, c.w);
        if (res <= tol * bn) break;
        // V0 = r / ||r||
        hipLaunchKernelGGL(k_adj_axpby, dim3(nb), dim3(256), 0, c.sm, len, 1.0 / res, c.w, 0.0,
                           c.V, 0.0, (const double *)nullptr);
        std::fill(gv.begin(), gv.end(), 0.0);
        gv[0] = res;
        int jl = -1;
        for (int j = 0; j < m && it < maxit; ++j) {
            double *vj = c.V + (long)j * len, *vn = c.V + (long)(j + 1) * len;
            adj_apply_prec_poly(c, vj, c.z);
            adj_apply_op(c, c.z, vn);
            // classical Gram-Schmidt; its products with the basis and vn . vn in one pass (vn
            // is V_(j+1)), then a second pass only where the first cancelled more than half of
            // vn's square norm (the DGKS test, decided on the device)
            mdot(c, c.V, j + 2, vn, c.red);
            hipLaunchKernelGGL(k_adj_dgks, dim3(1), dim3(64), 0, c.sm, j + 1,
                               (const double *)c.red, c.reorth, c.gate);
            hipLaunchKernelGGL(k_adj_maxpy, dim3(nb), dim3(256), 0, c.sm, len, j + 1, 1.0, -1.0,
                               (const double *)c.V, (const double *)c.red, vn,
                               (const int *)nullptr);
            (void)hipMemcpyAsync(h.data(), c.red, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, c.sm);
            mdot(c, c.V, j + 1, vn, c.red2, c.gate);
            hipLaunchKernelGGL(k_adj_maxpy, dim3(nb), dim3(256), 0, c.sm, len, j + 1, 1.0, -1.0,
                               (const double *)c.V, (const double *)c.red2, vn,
                               (const int *)c.gate);
            (void)hipMemcpyAsync(h2.data(), c.red2, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, c.sm);
            const double hn = norm2(c, vn);   // synchronises
            for (int i = 0; i <= j; ++i) h[i] += h2[i];
            h[j + 1] = hn;
            if (hn > 0.0)
                hipLaunchKernelGGL(k_adj_axpby, dim3(nb), dim3(256), 0, c.sm, len, 0.0,
                                   (const double *)nullptr, 1.0 / hn, vn, 0.0, (const double *)nullptr);
            // Givens rotations of the new column
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * h[i] + sn[i] * h[i + 1];
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                h[i] = t;
            }
            const double rr = std::sqrt(h[j] * h[j] + h[j + 1] * h[j + 1]);
            cs[j] = rr > 0.0 ? h[j] / rr : 1.0;
            sn[j] = rr > 0.0 ? h[j + 1] / rr : 0.0;
            h[j] = rr;
            h[j + 1] = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            for (int i = 0; i <= j; ++i) H[i * m + j] = h[i];
            ++it;
            jl = j;
            if (std::fabs(gv[j + 1]) <= tol * bn || hn == 0.0) break;
        }
        if (jl < 0) break;
        // y = H^-1 g (upper triangular), u = sum y_i V_i, x += Pc u
        for (int i = jl; i >= 0; --i) {
            double t = gv[i];
            for (int k = i + 1; k <= jl; ++k) t -= H[i * m + k] * y[k];
            y[i] = t / H[i * m + i];
        }
        (void)hipMemcpyAsync(c.red, y.data(), sizeof(double) * (jl + 1), hipMemcpyHostToDevice, c.sm);
        hipLaunchKernelGGL(k_adj_maxpy, dim3(nb), dim3(256), 0, c.sm, len, jl + 1, 0.0, 1.0,
                           (const double *)c.V, (const double *)c.red, c.t1b,
                           (const int *)nullptr);
        adj_apply_prec_poly(c, c.t1b, c.z);
        hipLaunchKernelGGL(k_adj_axpby, dim3(nb), dim3(256), 0, c.sm, len, 1.0, c.z, 1.0, x, 0.0,
                           (const double *)nullptr);
        // true residual r = b - Op x
        adj_apply_op(c, x, c.w);
        hipLaunchKernelGGL(k_adj_axpby, dim3(nb), dim3(256), 0, c.sm, len, 1.0, b, -1.0, c.w, 0.0,
                           (const double *)nullptr);
    }
    res = norm2(c, c.w);
    if (rel) *rel = res / bn;
    return it;
}

// dJ/dp = J_p + lam.F_p + (lam.F_c) CL_p for J = CL, CD.  out_cl[0], out_cd[0]: alpha (per
// degree); then one entry per shape direction (dx, dy host arrays of N each).
void adj_gradients(AdjCtx &c, int nshape, const double *dx, const double *dy, double *out_cl,
                   double *out_cd)
{
    const Dims &d = c.d;
    const long N = d.N;
    const double pi = 4.0 * atan(1.0);
    const double ca = cos(c.alpha * pi / 180.0), sa = sin(c.alpha * pi / 180.0);
    double dots[2][3];   // lam_o . (F_c, F_alpha, F_b)
    for (int o = 0; o < 2; ++o) {
        mdot(c, c.fc, 1, c.lam[o], c.red);
        mdot(c, c.fa, 1, c.lam[o], c.red + 1);
        (void)hipMemcpyAsync(dots[o], c.red, sizeof(double) * 2, hipMemcpyDeviceToHost, c.sm);
        (void)hipStreamSynchronize(c.sm);
    }
    // alpha: CL = 2 (Fy ca - Fx sa) / c, CD = 2 (Fx ca + Fy sa) / c
    const double cl_a = -c.cd * pi / 180.0, cd_a = c.cl * pi / 180.0;
    out_cl[0] = cl_a + dots[0][1] + dots[0][0] * cl_a;
    out_cd[0] = cd_a + dots[1][1] + dots[1][0] * cl_a;
    AdjParams prm{c.beta, c.cdis, c.vnu, c.pout, c.alpha, c.clift, c.ivis};
    int nf = 0;
    for (int w = 0; w < c.walls.n; ++w) nf += c.walls.j2[w] - c.walls.j1[w];
    std::vector<double> fpart(2 * (nf > 0 ? nf : 1));
    for (int i = 0; i < nshape; ++i) {
        (void)hipMemcpyAsync(c.dxy, dx + (long)i * N, N * 8, hipMemcpyHostToDevice, c.sm);
        (void)hipMemcpyAsync(c.dxy + N, dy + (long)i * N, N * 8, hipMemcpyHostToDevice, c.sm);
        hipLaunchKernelGGL(k_adj_metric_tangent, dim3((d.jm + 255) / 256, d.km), dim3(256), 0, c.sm,
                           d, c.x, c.y, (const double *)c.dxy, (const double *)(c.dxy + N), c.dmet,
                           c.ddj);
        hipLaunchKernelGGL(k_adj_resid_tangent, dim3((d.jm + 127) / 128, d.km), dim3(128), 0, c.sm,
                           d, prm, c.rtype, c.x, c.y, (const double *)c.dxy,
                           (const double *)(c.dxy + N), c.met, (const double *)c.dmet, c.dj,
                           (const double *)c.ddj, c.q, c.fb);
        if (nf > 0)
            hipLaunchKernelGGL(k_adj_force_tangent, dim3((nf + 63) / 64), dim3(64), 0, c.sm, d,
                               c.ivis, c.vnu, c.walls, c.met, (const double *)c.dmet, c.dj,
                               (const double *)c.ddj, c.q, c.fpart);
        double lf[2];
        for (int o = 0; o < 2; ++o) {
            mdot(c, c.fb, 1, c.lam[o], c.red + o);
        }
        (void)hipMemcpyAsync(lf, c.red, sizeof lf, hipMemcpyDeviceToHost, c.sm);
        if (nf > 0)
            (void)hipMemcpyAsync(fpart.data(), c.fpart, sizeof(double) * 2 * nf,
                                 hipMemcpyDeviceToHost, c.sm);
        (void)hipStreamSynchronize(c.sm);
        double dfx = 0.0, dfy = 0.0;
        for (int f = 0; f < nf; ++f) {
            dfx += fpart[2 * f];
            dfy += fpart[2 * f + 1];
        }
        const double cl_b = (2. * dfy * ca - 2. * dfx * sa) / c.chord;
        const double cd_b = (2. * dfx * ca + 2. * dfy * sa) / c.chord;
        out_cl[1 + i] = cl_b + lf[0] + dots[0][0] * cl_b;
        out_cd[1 + i] = cd_b + lf[1] + dots[1][0] * cl_b;
    }
}

}  // namespace smac
