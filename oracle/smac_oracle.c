/*
 * CPU restatement of miniSMAC2D's implicit iteration -- TEST INFRASTRUCTURE ONLY
 * (see smac_oracle.h).  Each function cites the reference routine it restates; the
 * operation order inside every expression follows the reference (Fortran evaluates
 * a-b-c as (a-b)-c, and -a*b as -(a*b), which is bit-identical to (-a)*b).
 *
 * Rounding facts of the reference build that this file reproduces on purpose:
 *  - default REAL is 4 bytes, so some literals and temporaries are float
 *    (r6 = 1./6., the SA constants, epss..., viscterms' vf/udif/vdif);
 *  - tursa.f is compiled at -O0 (flang crashes at -O1 and up in tsafill), where
 *    x**n for n >= 4 is binary powering (x**4 = (x*x)**2, x**6 = x*x * (x*x)**2);
 *    every other file is -O2, where x**2 = x*x (the only power used there);
 *  - no FMA contraction anywhere (-ffp-contract=off here, no -march there);
 *  - fixed-form source ends at column 72: ref/tursa.f:460 sets epss3 to 1, not 1e-20.
 */
#include "smac_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define I2(j, k) ((long)((j)-1) + (long)jm * ((k)-1))
#define Q3(a, j, k, n) (a)[I2(j, k) + N * ((n)-1)]
#define B9(a, j, k, m, n) (a)[I2(j, k) + N * ((m)-1) + 3 * N * ((n)-1)]
#define RT(j, k, a, b) rtxy[I2(j, k) + N * ((a)-1) + 2 * N * ((b)-1)]
#define S2(a, j, k) (a)[I2(j, k)]

#define MAXREG 64

/* Planted slips, for the negative controls of tests/test_oracle_pinning.py only: built with
 * -DORC_SLIPS into a separate library (_ref/libsmac_oracle_slips.so), where orc_set_slip(n)
 * turns on one plausible transcription mistake.  In the oracle itself SLIP(n) is the
 * constant 0 and every slip site compiles to the reference's expression. */
#ifdef ORC_SLIPS
static int g_slip = 0;
void orc_set_slip(int n) { g_slip = n; }
#define SLIP(n) (g_slip == (n))
#else
#define SLIP(n) 0
#endif
enum {
    SLIP_FACE_X21 = 1,    /* face_split: sign of x21                                    */
    SLIP_FACE_XI13 = 2,   /* face_split: xi13 built from q3 instead of q2                */
    SLIP_FLUX_R6 = 3,     /* flux3: r6 = 1/3                                             */
    SLIP_FLUX_DFM = 4,    /* flux3: sign of the A- differences in the third-order term   */
    SLIP_FLUX_ENDS = 5,   /* flux3: the two one-sided end forms exchanged                */
    SLIP_VISC_UV = 6,     /* viscterms: u equation fed with the v difference (xi)        */
    SLIP_VISC_RY = 7,     /* viscterms: ry^2 missing from the xi face coefficient        */
    SLIP_BC25_VORTEX = 8, /* far field: point-vortex velocity of the wrong sense         */
    SLIP_BC25_CHAR = 9,   /* far field: pressure row along the U-c characteristic        */
    SLIP_SA_FV2 = 10,     /* tsafill: fv2 = (1+chi/cv2)^-2                               */
    SLIP_SA_LIN = 11,     /* tsafill: linearisation without the fw derivative            */
    SLIP_FORCE_NORMAL = 12 /* force: x component of the wall normal taken from eta_y     */
};

struct orc_case {
    smac_params p;
    double vnu;
    int jm, km;
    long N;
    double *x, *y, *rtxy, *dj;
    double *q, *dq, *s, *anut, *vnut;
    double *btc, *bjm, *bjp, *bkm, *bkp, *scr1;
    double *lhs[5];
    /* SA scratch (separate arrays; the reference aliases step's scratch planes, which
     * only ever leaves finite values that the blanking multiplies by zero) */
    double *vort, *diff, *conv, *prod, *dest, *fr1, *sjm, *sjp, *skm, *skp, *stc, *ds, *sf,
        *smin, *sabprm;
    int nreg;
    smac_bc_region reg[MAXREG];
    int jkbc[MAXREG];
    int nwall, wjk[MAXREG], wj1[MAXREG], wj2[MAXREG], wk1[MAXREG], wk2[MAXREG], winc[MAXREG];
    int npair;
    int *pairs; /* (jt, jb) */
    int nt, jres, kres, first_tursa;
    double resmax0, resmax, divmax, turres, clift, rhsnorm;
    /* COMMON /dflx/, /flxs/ line scratch: persistent and zero-initialised like COMMON */
    long L;
    double *dfp, *dfm, *f, *fv;
};

/* ------------------------------------------------------------------------------------ */
/* small helpers                                                                          */

static double pw_sq(double x, int n) /* x**n as flang -O0 evaluates it (binary powering) */
{
    double r = 1.0;
    for (;;) {
        if (n & 1) r = r * x;
        n /= 2;
        if (n == 0) break;
        x = x * x;
    }
    return r;
}

static double fsign(double a, double b) { return copysign(fabs(a), b); }
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* isamax: first index (1-based) of max |x|, ties to the first (ref/blas.f:48-83) */
static long isamax(long n, const double *x)
{
    if (n < 1) return 0;
    long im = 1;
    double smax = fabs(x[0]);
    for (long i = 2; i <= n; ++i) {
        if (fabs(x[i - 1]) <= smax) continue;
        im = i;
        smax = fabs(x[i - 1]);
    }
    return im;
}

/* ------------------------------------------------------------------------------------ */
/* metrics: ref/metric.f:49-278                                                           */

void orc_metric(int jm, int km, const double *x, const double *y, double *rtxy, double *dj)
{
    long N = (long)jm * km;
    double *xj = malloc(sizeof(double) * N), *yj = malloc(sizeof(double) * N);
    double *xk = malloc(sizeof(double) * N), *yk = malloc(sizeof(double) * N);
    for (int j = 2; j <= jm - 1; ++j)
        for (int k = 1; k <= km; ++k) {
            S2(xj, j, k) = 0.5 * (S2(x, j + 1, k) - S2(x, j - 1, k));
            S2(yj, j, k) = 0.5 * (S2(y, j + 1, k) - S2(y, j - 1, k));
        }
    for (int k = 1; k <= km; ++k) {
        int j = 1;
        S2(xj, j, k) = 0.5 * (-S2(x, j + 2, k) + 4. * S2(x, j + 1, k) - 3. * S2(x, j, k));
        S2(yj, j, k) = 0.5 * (-S2(y, j + 2, k) + 4. * S2(y, j + 1, k) - 3. * S2(y, j, k));
        j = jm;
        S2(xj, j, k) = 0.5 * (3. * S2(x, j, k) - 4. * S2(x, j - 1, k) + S2(x, j - 2, k));
        S2(yj, j, k) = 0.5 * (3. * S2(y, j, k) - 4. * S2(y, j - 1, k) + S2(y, j - 2, k));
    }
    for (int k = 2; k <= km - 1; ++k)
        for (int j = 1; j <= jm; ++j) {
            S2(xk, j, k) = 0.5 * (S2(x, j, k + 1) - S2(x, j, k - 1));
            S2(yk, j, k) = 0.5 * (S2(y, j, k + 1) - S2(y, j, k - 1));
        }
    for (int j = 1; j <= jm; ++j) {
        int k = 1;
        S2(xk, j, k) = 0.5 * (-S2(x, j, k + 2) + 4. * S2(x, j, k + 1) - 3. * S2(x, j, k));
        S2(yk, j, k) = 0.5 * (-S2(y, j, k + 2) + 4. * S2(y, j, k + 1) - 3. * S2(y, j, k));
        k = km;
        S2(xk, j, k) = 0.5 * (3. * S2(x, j, k) - 4. * S2(x, j, k - 1) + S2(x, j, k - 2));
        S2(yk, j, k) = 0.5 * (3. * S2(y, j, k) - 4. * S2(y, j, k - 1) + S2(y, j, k - 2));
    }
    for (int k = 1; k <= km; ++k)
        for (int j = 1; j <= jm; ++j) {
            double rdj = S2(xj, j, k) * S2(yk, j, k) - S2(xk, j, k) * S2(yj, j, k);
            S2(dj, j, k) = 1. / rdj;
            RT(j, k, 1, 2) = S2(yk, j, k);
            RT(j, k, 1, 3) = -S2(xk, j, k);
            RT(j, k, 2, 2) = -S2(yj, j, k);
            RT(j, k, 2, 3) = S2(xj, j, k);
            RT(j, k, 1, 1) = 0.0;
            RT(j, k, 2, 1) = 0.0;
        }
    free(xj); free(yj); free(xk); free(yk);
}

/* ------------------------------------------------------------------------------------ */
/* conup + fluxj3/fluxk3: ref/conup.f:49-478, ref/flux.f:160-372                          */

static void face_split(double rt, double rx, double ry, double q2, double q3, double beta,
                       double ap[3][3], double am[3][3])
{
    /* ref/conup.f:127-185 (xi faces), :306-355 (eta faces) */
    double cvel = rx * q2 + ry * q3;
    double t = cvel + 0.5 * rt;
    double cc = sqrt(t * t + beta * (rx * rx + ry * ry));
    double eig = cvel + rt;
    double eig1 = 0.5 * (eig + fabs(eig));
    double eig2 = cvel + rt * 0.5 + cc;
    double x21 = (SLIP(SLIP_FACE_X21) ? ry : -ry) * eig1 / (cc * cc - 0.25 * rt * rt);
    double x31 = rx * eig1 / (cc * cc - 0.25 * rt * rt);
    double x12 = 0.5 * eig2 / cc;
    double x22 = 0.5 * eig2 * (q2 * eig2 + beta * rx) / (beta * cc * (cc - 0.5 * rt));
    double x32 = 0.5 * eig2 * (q3 * eig2 + beta * ry) / (beta * cc * (cc - 0.5 * rt));
    double xi11 = ry * q2 - rx * q3;
    double xi12 = -q3 * (cvel + rt) - ry * beta;
    double xi13 = (SLIP(SLIP_FACE_XI13) ? q3 : q2) * (cvel + rt) + rx * beta;
    double xi21 = cc - cvel - 0.5 * rt;
    double xi22 = rx * beta;
    double xi23 = ry * beta;
    ap[0][0] = x12 * xi21;
    ap[1][0] = x21 * xi11 + x22 * xi21;
    ap[2][0] = x31 * xi11 + x32 * xi21;
    ap[0][1] = x12 * xi22;
    ap[1][1] = x21 * xi12 + x22 * xi22;
    ap[2][1] = x31 * xi12 + x32 * xi22;
    ap[0][2] = x12 * xi23;
    ap[1][2] = x21 * xi13 + x22 * xi23;
    ap[2][2] = x31 * xi13 + x32 * xi23;
    am[0][0] = -ap[0][0];
    am[1][0] = -ap[1][0] + rx;
    am[2][0] = -ap[2][0] + ry;
    am[0][1] = -ap[0][1] + rx * beta;
    am[1][1] = -ap[1][1] + rt + 2. * rx * q2 + ry * q3;
    am[2][1] = -ap[2][1] + rx * q3;
    am[0][2] = -ap[0][2] + ry * beta;
    am[1][2] = -ap[1][2] + ry * q2;
    am[2][2] = -ap[2][2] + rt + rx * q2 + 2. * ry * q3;
}

void orc_face_split(double rt, double rx, double ry, double q2, double q3, double beta,
                    double *ap, double *am)
{
    double a[3][3], b[3][3];
    face_split(rt, rx, ry, q2, q3, beta, a, b);
    for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) {
            ap[m + 3 * n] = a[m][n];
            am[m + 3 * n] = b[m][n];
        }
}

static void point_jac(double rt, double rx, double ry, double u, double v, double beta,
                      double at[3][3])
{
    /* ref/conup.f:89-105, :288-300 */
    double cvel = rt + rx * u + ry * v;
    at[0][0] = rt;
    at[1][0] = rx;
    at[2][0] = ry;
    at[0][1] = rx * beta;
    at[1][1] = rx * u + cvel;
    at[2][1] = rx * v;
    at[0][2] = ry * beta;
    at[1][2] = ry * u;
    at[2][2] = ry * v + cvel;
}

/* face fluxes f(i), i=1..n-1, of one line (ref/flux.f:160-266 / 270-372).  Inputs per
 * point i=1..n: p,u,v,rt,rx,ry; faces: ap/am (n-1).  Uses the persistent dfp/dfm arrays
 * (COMMON /dflx/), which keep stale entries at i=n from earlier lines exactly as the
 * reference's do; those entries are only ever multiplied by cp = 0. */
static void flux3_line(orc_case *c, int n, const double *pq, const double *uq, const double *vq,
                       const double *rt, const double *rx, const double *ry,
                       double (*ap)[3][3], double (*am)[3][3])
{
    const double beta = c->p.beta, cdis = c->p.cdis;
    const long L = c->L;
#define DFP(i, m) c->dfp[(i)-1 + L * ((m)-1)]
#define DFM(i, m) c->dfm[(i)-1 + L * ((m)-1)]
#define FF(i, m) c->f[(i)-1 + L * ((m)-1)]
#define FV(i, m) c->fv[(i)-1 + L * ((m)-1)]
    for (int i = 1; i <= n; ++i) {
        double cvel = rt[i - 1] + rx[i - 1] * uq[i - 1] + ry[i - 1] * vq[i - 1];
        FV(i, 1) = rt[i - 1] * pq[i - 1] + beta * (cvel - rt[i - 1]);
        FV(i, 2) = uq[i - 1] * cvel + rx[i - 1] * pq[i - 1];
        FV(i, 3) = vq[i - 1] * cvel + ry[i - 1] * pq[i - 1];
    }
    for (int i = 1; i <= n - 1; ++i) {
        double dq1 = pq[i] - pq[i - 1];
        double dq2 = uq[i] - uq[i - 1];
        double dq3 = vq[i] - vq[i - 1];
        for (int m = 1; m <= 3; ++m) {
            DFP(i, m) = ap[i - 1][m - 1][0] * dq1 + ap[i - 1][m - 1][1] * dq2 + ap[i - 1][m - 1][2] * dq3;
            DFM(i, m) = am[i - 1][m - 1][0] * dq1 + am[i - 1][m - 1][1] * dq2 + am[i - 1][m - 1][2] * dq3;
        }
    }
    const double r6 = SLIP(SLIP_FLUX_R6) ? 1.0 / 3.0
                                         : (double)(1.0f / 6.0f); /* r6 = 1./6.: a default-REAL quotient */
    for (int i = 1; i <= n - 1; ++i) {
        int im = i - 1 > 1 ? i - 1 : 1;
        int ip = i + 1 < n ? i + 1 : n;
        int ipp = i + 2 < n ? i + 2 : n;
        float cp = (float)(ipp - ip);
        float cm = (float)(i - im);
        for (int m = 1; m <= 3; ++m) {
            double fm = (DFP(i, m) + DFM(i, m) - DFP(ip, m) - DFM(ip, m)) * r6;
            double fp = (DFP(im, m) + DFM(im, m) - DFP(i, m) - DFM(i, m)) * r6;
            double f3 = SLIP(SLIP_FLUX_DFM) ? (DFP(im, m) - DFP(i, m) - DFM(i, m) + DFM(ip, m)) * r6
                                            : (DFP(im, m) - DFP(i, m) + DFM(i, m) - DFM(ip, m)) * r6;
            if (SLIP(SLIP_FLUX_ENDS)) {
                double t = fm;
                fm = fp;
                fp = t;
            }
            FF(i, m) = 0.5 * (FV(ip, m) + FV(i, m)) +
                       cdis * ((double)(cp * cm) * f3 + (double)(cm * (1.f - cp)) * fp +
                               (double)(cp * (1.f - cm)) * fm);
        }
    }
#undef DFP
#undef DFM
#undef FV
}

static void conup(orc_case *c)
{
    const int jm = c->jm, km = c->km;
    const long N = c->N;
    const double beta = c->p.beta;
    double *q = c->q, *rtxy = c->rtxy, *s = c->s;
    int nmax = jm > km ? jm : km;
    double (*at)[3][3] = malloc(sizeof(double[3][3]) * nmax);
    double (*ap)[3][3] = malloc(sizeof(double[3][3]) * nmax);
    double (*am)[3][3] = malloc(sizeof(double[3][3]) * nmax);
    double *pq = malloc(sizeof(double) * nmax * 6);
    double *uq = pq + nmax, *vq = uq + nmax, *rt = vq + nmax, *rx = rt + nmax, *ry = rx + nmax;
    const long L = c->L;
    /* xi direction: ref/conup.f:67-275 */
    for (int k = 2; k <= km - 1; ++k) {
        for (int j = 1; j <= jm; ++j) {
            point_jac(RT(j, k, 1, 1), RT(j, k, 1, 2), RT(j, k, 1, 3), Q3(q, j, k, 2), Q3(q, j, k, 3),
                      beta, at[j - 1]);
            pq[j - 1] = Q3(q, j, k, 1); uq[j - 1] = Q3(q, j, k, 2); vq[j - 1] = Q3(q, j, k, 3);
            rt[j - 1] = RT(j, k, 1, 1); rx[j - 1] = RT(j, k, 1, 2); ry[j - 1] = RT(j, k, 1, 3);
        }
        for (int j = 1; j <= jm - 1; ++j) {
            double frt = 0.5 * (RT(j + 1, k, 1, 1) + RT(j, k, 1, 1));
            double frx = 0.5 * (RT(j + 1, k, 1, 2) + RT(j, k, 1, 2));
            double fry = 0.5 * (RT(j + 1, k, 1, 3) + RT(j, k, 1, 3));
            double q2 = 0.5 * (Q3(q, j + 1, k, 2) + Q3(q, j, k, 2));
            double q3 = 0.5 * (Q3(q, j + 1, k, 3) + Q3(q, j, k, 3));
            face_split(frt, frx, fry, q2, q3, beta, ap[j - 1], am[j - 1]);
        }
        flux3_line(c, jm, pq, uq, vq, rt, rx, ry, ap, am);
        for (int j = 2; j <= jm - 1; ++j)
            for (int n = 1; n <= 3; ++n)
                for (int m = 1; m <= 3; ++m) {
                    B9(c->bjm, j, k, m, n) =
                        0.5 * (-at[j - 2][m - 1][n - 1] - ap[j - 2][m - 1][n - 1] + am[j - 2][m - 1][n - 1]);
                    B9(c->bjp, j, k, m, n) =
                        0.5 * (at[j][m - 1][n - 1] - ap[j - 1][m - 1][n - 1] + am[j - 1][m - 1][n - 1]);
                    B9(c->btc, j, k, m, n) = 0.5 * (ap[j - 1][m - 1][n - 1] + ap[j - 2][m - 1][n - 1] -
                                                    am[j - 1][m - 1][n - 1] - am[j - 2][m - 1][n - 1]);
                }
        for (int j = 2; j <= jm - 1; ++j)
            for (int n = 1; n <= 3; ++n)
                Q3(s, j, k, n) = -(FF(j, n) - FF(j - 1, n));
    }
    /* eta direction: ref/conup.f:282-461 */
    for (int j = 2; j <= jm - 1; ++j) {
        for (int k = 1; k <= km; ++k) {
            point_jac(RT(j, k, 2, 1), RT(j, k, 2, 2), RT(j, k, 2, 3), Q3(q, j, k, 2), Q3(q, j, k, 3),
                      beta, at[k - 1]);
            pq[k - 1] = Q3(q, j, k, 1); uq[k - 1] = Q3(q, j, k, 2); vq[k - 1] = Q3(q, j, k, 3);
            rt[k - 1] = RT(j, k, 2, 1); rx[k - 1] = RT(j, k, 2, 2); ry[k - 1] = RT(j, k, 2, 3);
        }
        for (int k = 1; k <= km - 1; ++k) {
            double frt = 0.5 * (RT(j, k + 1, 2, 1) + RT(j, k, 2, 1));
            double frx = 0.5 * (RT(j, k + 1, 2, 2) + RT(j, k, 2, 2));
            double fry = 0.5 * (RT(j, k + 1, 2, 3) + RT(j, k, 2, 3));
            double q2 = 0.5 * (Q3(q, j, k + 1, 2) + Q3(q, j, k, 2));
            double q3 = 0.5 * (Q3(q, j, k + 1, 3) + Q3(q, j, k, 3));
            face_split(frt, frx, fry, q2, q3, beta, ap[k - 1], am[k - 1]);
        }
        flux3_line(c, km, pq, uq, vq, rt, rx, ry, ap, am);
        for (int k = 2; k <= km - 1; ++k)
            for (int n = 1; n <= 3; ++n)
                for (int m = 1; m <= 3; ++m) {
                    B9(c->bkm, j, k, m, n) =
                        0.5 * (-at[k - 2][m - 1][n - 1] - ap[k - 2][m - 1][n - 1] + am[k - 2][m - 1][n - 1]);
                    B9(c->bkp, j, k, m, n) =
                        0.5 * (at[k][m - 1][n - 1] - ap[k - 1][m - 1][n - 1] + am[k - 1][m - 1][n - 1]);
                    B9(c->btc, j, k, m, n) =
                        B9(c->btc, j, k, m, n) + .5 * (ap[k - 1][m - 1][n - 1] + ap[k - 2][m - 1][n - 1] -
                                                       am[k - 1][m - 1][n - 1] - am[k - 2][m - 1][n - 1]);
                }
        for (int k = 2; k <= km - 1; ++k)
            for (int n = 1; n <= 3; ++n)
                Q3(s, j, k, n) = Q3(s, j, k, n) - (FF(k, n) - FF(k - 1, n));
    }
    (void)L;
    free(at); free(ap); free(am); free(pq);
}
#undef FF

/* ------------------------------------------------------------------------------------ */
/* viscterms: ref/visc.f:50-160 (udif, vdif, vf are default REAL)                         */

/* The default-REAL temporaries of viscterms.  A build with -DORC_FP64_ISLANDS
 * (_ref/libsmac_oracle_fp64.so) keeps them in double: a smooth residual whose central
 * differences carry no FP32 rounding, the reference for the FP64 tangent checks of the
 * adjoint.  The oracle itself is the reference's FP32. */
#ifdef ORC_FP64_ISLANDS
typedef double visl_t;
#else
typedef float visl_t;
#endif

static void viscterms(orc_case *c)
{
    const int jm = c->jm, km = c->km;
    const long N = c->N;
    const double vnu = c->vnu;
    double *q = c->q, *rtxy = c->rtxy, *dj = c->dj, *s = c->s;
    int nmax = jm > km ? jm : km;
    visl_t *udif = malloc(sizeof(visl_t) * nmax * 3), *vdif = udif + nmax, *vf = vdif + nmax;
    for (int k = 2; k <= km - 1; ++k) {
        for (int j = 1; j <= jm - 1; ++j) {
            double djv = 0.5 * (S2(dj, j, k) * vnu + S2(dj, j + 1, k) * vnu);
            udif[j - 1] = (visl_t)(Q3(q, j + 1, k, 2) - Q3(q, j, k, 2));
            vdif[j - 1] = (visl_t)(Q3(q, j + 1, k, 3) - Q3(q, j, k, 3));
            double rx = 0.5 * (RT(j, k, 1, 2) + RT(j + 1, k, 1, 2));
            double ry = 0.5 * (RT(j, k, 1, 3) + RT(j + 1, k, 1, 3));
            vf[j - 1] = SLIP(SLIP_VISC_RY) ? (visl_t)(djv * (rx * rx)) : (visl_t)(djv * (rx * rx + ry * ry));
        }
        for (int j = 2; j <= jm - 1; ++j) {
            double a = vf[j - 2], b = vf[j - 1];
            B9(c->bjm, j, k, 2, 2) = B9(c->bjm, j, k, 2, 2) - a;
            B9(c->bjm, j, k, 3, 3) = B9(c->bjm, j, k, 3, 3) - a;
            B9(c->btc, j, k, 2, 2) = B9(c->btc, j, k, 2, 2) + a + b;
            B9(c->btc, j, k, 3, 3) = B9(c->btc, j, k, 3, 3) + a + b;
            B9(c->bjp, j, k, 2, 2) = B9(c->bjp, j, k, 2, 2) - b;
            B9(c->bjp, j, k, 3, 3) = B9(c->bjp, j, k, 3, 3) - b;
            visl_t t2 = SLIP(SLIP_VISC_UV) ? vf[j - 1] * vdif[j - 1] - vf[j - 2] * vdif[j - 2]
                                           : vf[j - 1] * udif[j - 1] - vf[j - 2] * udif[j - 2];
            visl_t t3 = vf[j - 1] * vdif[j - 1] - vf[j - 2] * vdif[j - 2];
            Q3(s, j, k, 2) = Q3(s, j, k, 2) + (double)t2;
            Q3(s, j, k, 3) = Q3(s, j, k, 3) + (double)t3;
        }
    }
    for (int j = 2; j <= jm - 1; ++j) {
        for (int k = 1; k <= km - 1; ++k) {
            double djv = 0.5 * (S2(dj, j, k) * vnu + S2(dj, j, k + 1) * vnu);
            udif[k - 1] = (visl_t)(Q3(q, j, k + 1, 2) - Q3(q, j, k, 2));
            vdif[k - 1] = (visl_t)(Q3(q, j, k + 1, 3) - Q3(q, j, k, 3));
            double rx = 0.5 * (RT(j, k, 2, 2) + RT(j, k + 1, 2, 2));
            double ry = 0.5 * (RT(j, k, 2, 3) + RT(j, k + 1, 2, 3));
            vf[k - 1] = (visl_t)(djv * (rx * rx + ry * ry));
        }
        for (int k = 2; k <= km - 1; ++k) {
            double a = vf[k - 2], b = vf[k - 1];
            B9(c->bkm, j, k, 2, 2) = B9(c->bkm, j, k, 2, 2) - a;
            B9(c->bkm, j, k, 3, 3) = B9(c->bkm, j, k, 3, 3) - a;
            B9(c->btc, j, k, 2, 2) = B9(c->btc, j, k, 2, 2) + a + b;
            B9(c->btc, j, k, 3, 3) = B9(c->btc, j, k, 3, 3) + a + b;
            B9(c->bkp, j, k, 2, 2) = B9(c->bkp, j, k, 2, 2) - b;
            B9(c->bkp, j, k, 3, 3) = B9(c->bkp, j, k, 3, 3) - b;
            visl_t t2 = vf[k - 1] * udif[k - 1] - vf[k - 2] * udif[k - 2];
            visl_t t3 = vf[k - 1] * vdif[k - 1] - vf[k - 2] * vdif[k - 2];
            Q3(s, j, k, 2) = Q3(s, j, k, 2) + (double)t2;
            Q3(s, j, k, 3) = Q3(s, j, k, 3) + (double)t3;
        }
    }
    free(udif);
}

/* ------------------------------------------------------------------------------------ */
/* boundary rows: bcmain ref/bc.f:60-224 and the routines it calls                        */

static void bcnoslk(orc_case *c, int k, int jbeg, int jend)
{
    /* ref/bc.f:1134-1186, ibval 0 branch (the only one reachable from bcmain) */
    const int jm = c->jm;
    const long N = c->N;
    double *q = c->q, *s = c->s;
    for (int j = jbeg; j <= jend; ++j) {
        B9(c->btc, j, k, 1, 1) = 1.0;
        B9(c->btc, j, k, 2, 2) = 1.0;
        B9(c->btc, j, k, 3, 3) = 1.0;
        B9(c->bkp, j, k, 1, 1) = -1.0;
        Q3(s, j, k, 1) = Q3(q, j, k + 1, 1) - Q3(q, j, k, 1);
        Q3(s, j, k, 2) = -Q3(q, j, k, 2);
        Q3(s, j, k, 3) = -Q3(q, j, k, 3);
    }
}

static int bccgridk(orc_case *c, int ibval, int k, int jbeg, int jend)
{
    /* ref/bc.f:1334-1490 */
    const int jm = c->jm, km = c->km;
    const long N = c->N;
    double *q = c->q, *s = c->s, *rtxy = c->rtxy, *x = c->x, *y = c->y;
    const double beta = c->p.beta, clift = c->clift;
    const int jk = 2;
    double pi = 4.0 * atan(1.0);
    double uinf = cos(c->p.alpha * pi / 180.);
    double vinf = sin(c->p.alpha * pi / 180.);
    for (int j = jbeg; j <= jend; ++j) {
        double dx = S2(x, j, k) - 0.25;
        double dy = S2(y, j, k);
        double ucirc = clift * dy / (4. * pi * (dx * dx + dy * dy));
        double vcirc = -clift * dx / (4. * pi * (dx * dx + dy * dy));
        if (SLIP(SLIP_BC25_VORTEX)) {
            ucirc = -ucirc;
            vcirc = -vcirc;
        }
        B9(c->btc, j, k, 2, 2) = 1.0;
        B9(c->btc, j, k, 3, 3) = 1.0;
        Q3(s, j, k, 2) = uinf + ucirc - Q3(q, j, k, 2);
        Q3(s, j, k, 3) = vinf + vcirc - Q3(q, j, k, 3);
    }
    if (ibval != 25) return SMAC_ERR_BC;
    if (k == 1) {
        for (int j = jbeg; j <= jend; ++j) {
            double r2 = RT(j, k, jk, 2), r3 = RT(j, k, jk, 3);
            double qq = r2 * Q3(q, j, k, 2) + r3 * Q3(q, j, k, 3);
            double cc = sqrt(qq * qq + beta * (r2 * r2 + r3 * r3));
            B9(c->bkp, j, k, 1, 1) = r2 * r2 + r3 * r3;
            B9(c->bkp, j, k, 1, 2) = r2 * (qq - cc);
            B9(c->bkp, j, k, 1, 3) = r3 * (qq - cc);
            B9(c->btc, j, k, 1, 1) = -B9(c->bkp, j, k, 1, 1);
            B9(c->btc, j, k, 1, 2) = -B9(c->bkp, j, k, 1, 2);
            B9(c->btc, j, k, 1, 3) = -B9(c->bkp, j, k, 1, 3);
            double dq1 = Q3(q, j, k + 1, 1) - Q3(q, j, k, 1);
            double dq2 = Q3(q, j, k + 1, 2) - Q3(q, j, k, 2);
            double dq3 = Q3(q, j, k + 1, 3) - Q3(q, j, k, 3);
            Q3(s, j, k, 1) = -B9(c->bkp, j, k, 1, 1) * dq1 - B9(c->bkp, j, k, 1, 2) * dq2 -
                             B9(c->bkp, j, k, 1, 3) * dq3;
        }
    } else if (k == km) {
        for (int j = jbeg; j <= jend; ++j) {
            double r2 = RT(j, k, jk, 2), r3 = RT(j, k, jk, 3);
            double qq = r2 * Q3(q, j, k, 2) + r3 * Q3(q, j, k, 3);
            double cc = sqrt(qq * qq + beta * (r2 * r2 + r3 * r3));
            if (SLIP(SLIP_BC25_CHAR)) cc = -cc;
            B9(c->bkm, j, k, 1, 1) = -(r2 * r2 + r3 * r3);
            B9(c->bkm, j, k, 1, 2) = -r2 * (qq + cc);
            B9(c->bkm, j, k, 1, 3) = -r3 * (qq + cc);
            B9(c->btc, j, k, 1, 1) = -B9(c->bkm, j, k, 1, 1);
            B9(c->btc, j, k, 1, 2) = -B9(c->bkm, j, k, 1, 2);
            B9(c->btc, j, k, 1, 3) = -B9(c->bkm, j, k, 1, 3);
            double dq1 = Q3(q, j, k, 1) - Q3(q, j, k - 1, 1);
            double dq2 = Q3(q, j, k, 2) - Q3(q, j, k - 1, 2);
            double dq3 = Q3(q, j, k, 3) - Q3(q, j, k - 1, 3);
            Q3(s, j, k, 1) = B9(c->bkm, j, k, 1, 1) * dq1 + B9(c->bkm, j, k, 1, 2) * dq2 +
                             B9(c->bkm, j, k, 1, 3) * dq3;
        }
    } else {
        return SMAC_ERR_BC;
    }
    return SMAC_OK;
}

static int bcoutj(orc_case *c, int ibval, int j, int kbeg, int kend)
{
    /* ref/bc.f:1494-1680, extrapolation branches (ibval 31/33) and the fixed-pressure
     * row (30/31); the characteristic branches (30/32) are not reachable from the C-grid
     * BC files this library accepts */
    const int jm = c->jm;
    const long N = c->N;
    double *q = c->q, *s = c->s;
    if ((ibval == 31 || ibval == 33) && j == 1) {
        for (int k = kbeg; k <= kend; ++k) {
            B9(c->btc, j, k, 2, 2) = 1.0;
            B9(c->btc, j, k, 3, 3) = 1.0;
            B9(c->bjp, j, k, 2, 2) = -1.0;
            B9(c->bjp, j, k, 3, 3) = -1.0;
            Q3(s, j, k, 2) = Q3(q, j + 1, k, 2) - Q3(q, j, k, 2);
            Q3(s, j, k, 3) = Q3(q, j + 1, k, 3) - Q3(q, j, k, 3);
        }
    } else if ((ibval == 31 || ibval == 33) && j == jm) {
        for (int k = kbeg; k <= kend; ++k) {
            B9(c->btc, j, k, 2, 2) = 1.0;
            B9(c->btc, j, k, 3, 3) = 1.0;
            B9(c->bjm, j, k, 2, 2) = -1.0;
            B9(c->bjm, j, k, 3, 3) = -1.0;
            Q3(s, j, k, 2) = Q3(q, j - 1, k, 2) - Q3(q, j, k, 2);
            Q3(s, j, k, 3) = Q3(q, j - 1, k, 3) - Q3(q, j, k, 3);
        }
    } else {
        return SMAC_ERR_BC;
    }
    if (ibval == 30 || ibval == 31) {
        for (int k = kbeg; k <= kend; ++k) {
            B9(c->btc, j, k, 1, 1) = 1.0;
            Q3(s, j, k, 1) = c->p.pout - Q3(q, j, k, 1);
        }
    }
    return SMAC_OK;
}

static int bcmain(orc_case *c)
{
    for (int nr = 0; nr < c->nreg; ++nr) {
        const smac_bc_region *r = &c->reg[nr];
        int rc = SMAC_OK;
        if (r->ibcval == 0) {
            if (c->jkbc[nr] == 2) bcnoslk(c, r->kbeg, r->jbeg, r->jend);
            else rc = SMAC_ERR_BC; /* j-line walls: bcnoslj, not needed by C-grids */
        } else if (r->ibcval == 25) {
            if (c->jkbc[nr] == 2) rc = bccgridk(c, 25, r->kbeg, r->jbeg, r->jend);
            else rc = SMAC_ERR_BC;
        } else if (r->ibcval == 31) {
            if (c->jkbc[nr] == 1) rc = bcoutj(c, 31, r->jbeg, r->kbeg, r->kend);
            else rc = SMAC_ERR_BC;
        } else if (r->ibcval != 60) {
            rc = SMAC_ERR_BC;
        }
        if (rc) return rc;
    }
    return SMAC_OK;
}

/* wake-cut injection, bcimpds single-rank branch: ref/bc.f:571-591 */
static void bcimpds(int jm, int km, int neqs, int npair, const int *pairs, const double *sq,
                    double *ds, double *rhs)
{
    long N = (long)jm * km;
    (void)km;
    for (int ip = 0; ip < npair; ++ip) {
        int jt = pairs[2 * ip], jb = pairs[2 * ip + 1];
        for (int neq = 1; neq <= neqs; ++neq) {
            double snew = 0.5 * (Q3(sq, jt, 2, neq) + Q3(sq, jb, 2, neq));
            double dsnew = 0.5 * (Q3(ds, jt, 2, neq) + Q3(ds, jb, 2, neq));
            Q3(ds, jt, 1, neq) = snew + dsnew - Q3(sq, jt, 1, neq);
            Q3(rhs, jt, 1, neq) = Q3(ds, jt, 1, neq);
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* block tridiagonal LU: ref/btlu3.f                                                      */

/* 3x3 Crout factor in place at one point (ref/btlu3.f:148-155 and repeats) */
#define LU3(B, J, K)                                                                      \
    do {                                                                                  \
        B(J, K, 1, 1) = 1. / B(J, K, 1, 1);                                               \
        B(J, K, 1, 2) = B(J, K, 1, 2) * B(J, K, 1, 1);                                    \
        B(J, K, 1, 3) = B(J, K, 1, 3) * B(J, K, 1, 1);                                    \
        B(J, K, 2, 2) = 1. / (B(J, K, 2, 2) - B(J, K, 2, 1) * B(J, K, 1, 2));             \
        B(J, K, 2, 3) = (B(J, K, 2, 3) - B(J, K, 2, 1) * B(J, K, 1, 3)) * B(J, K, 2, 2);  \
        B(J, K, 3, 2) = B(J, K, 3, 2) - B(J, K, 3, 1) * B(J, K, 1, 2);                    \
        B(J, K, 3, 3) =                                                                   \
            1. / (B(J, K, 3, 3) - B(J, K, 3, 1) * B(J, K, 1, 3) - B(J, K, 2, 3) * B(J, K, 3, 2)); \
    } while (0)

/* c <- b^-1 c with b factored (ref/btlu3.f:157-179) */
#define CSOLVE(B, C, J, K)                                                                \
    do {                                                                                  \
        for (int n_ = 1; n_ <= 3; ++n_) {                                                 \
            double d1 = B(J, K, 1, 1) * C(J, K, 1, n_);                                   \
            double d2 = B(J, K, 2, 2) * (C(J, K, 2, n_) - B(J, K, 2, 1) * d1);            \
            double d3 = B(J, K, 3, 3) * (C(J, K, 3, n_) - B(J, K, 3, 1) * d1 - B(J, K, 3, 2) * d2); \
            C(J, K, 3, n_) = d3;                                                          \
            C(J, K, 2, n_) = d2 - B(J, K, 2, 3) * C(J, K, 3, n_);                         \
            C(J, K, 1, n_) = d1 - B(J, K, 1, 3) * C(J, K, 3, n_) - B(J, K, 1, 2) * C(J, K, 2, n_); \
        }                                                                                 \
    } while (0)

/* b <- b - a * cprev (ref/btlu3.f:245-273) */
#define BMINUS(A, B, C, J, K, JP, KP)                                                     \
    do {                                                                                  \
        for (int n_ = 1; n_ <= 3; ++n_)                                                   \
            for (int m_ = 1; m_ <= 3; ++m_)                                               \
                B(J, K, m_, n_) = B(J, K, m_, n_) - A(J, K, m_, 1) * C(JP, KP, 1, n_) -   \
                                  A(J, K, m_, 2) * C(JP, KP, 2, n_) -                     \
                                  A(J, K, m_, 3) * C(JP, KP, 3, n_);                      \
    } while (0)

/* ref/btlu3.f loops keep the order (2,1),(3,1)... i.e. column n outer, row m inner for the
 * b - a*c update: b11,b21,b31,b12,... -- each entry depends only on the old b and on c of
 * the previous point, so the order inside the update does not change any value. */

static void btlu3j(int jm, int km, double *a, double *b, double *c, int jbeg, int jend, int kbeg,
                   int kend)
{
    long N = (long)jm * km;
#define A_(j, k, m, n) B9(a, j, k, m, n)
#define B_(j, k, m, n) B9(b, j, k, m, n)
#define C_(j, k, m, n) B9(c, j, k, m, n)
    for (int k = kbeg; k <= kend; ++k) {
        LU3(B_, jbeg, k);
        CSOLVE(B_, C_, jbeg, k);
    }
    for (int j = jbeg + 1; j <= jend - 1; ++j)
        for (int k = kbeg; k <= kend; ++k) {
            BMINUS(A_, B_, C_, j, k, j - 1, k);
            LU3(B_, j, k);
            CSOLVE(B_, C_, j, k);
        }
    for (int k = kbeg; k <= kend; ++k) {
        BMINUS(A_, B_, C_, jend, k, jend - 1, k);
        LU3(B_, jend, k);
    }
}

static void btlu3k(int jm, int km, double *a, double *b, double *c, int jbeg, int jend, int kbeg,
                   int kend)
{
    long N = (long)jm * km;
    for (int j = jbeg; j <= jend; ++j) {
        LU3(B_, j, kbeg);
        CSOLVE(B_, C_, j, kbeg);
    }
    for (int k = kbeg + 1; k <= kend - 1; ++k)
        for (int j = jbeg; j <= jend; ++j) {
            BMINUS(A_, B_, C_, j, k, j, k - 1);
            LU3(B_, j, k);
            CSOLVE(B_, C_, j, k);
        }
    for (int j = jbeg; j <= jend; ++j) {
        BMINUS(A_, B_, C_, j, kend, j, kend - 1);
        LU3(B_, j, kend);
    }
}

/* recompose a point's factor (ref/btlu3.f:628-650 loop 10) */
#define UNLU3(B, J, K)                                                                    \
    do {                                                                                  \
        B(J, K, 3, 3) = 1. / B(J, K, 3, 3) + B(J, K, 3, 1) * B(J, K, 1, 3) + B(J, K, 3, 2) * B(J, K, 2, 3); \
        B(J, K, 3, 2) = B(J, K, 3, 2) + B(J, K, 3, 1) * B(J, K, 1, 2);                    \
        B(J, K, 2, 2) = 1. / B(J, K, 2, 2);                                               \
        B(J, K, 2, 3) = B(J, K, 2, 3) * B(J, K, 2, 2) + B(J, K, 2, 1) * B(J, K, 1, 3);    \
        B(J, K, 2, 2) = B(J, K, 2, 2) + B(J, K, 2, 1) * B(J, K, 1, 2);                    \
        B(J, K, 1, 1) = 1. / B(J, K, 1, 1);                                               \
        B(J, K, 1, 3) = B(J, K, 1, 3) * B(J, K, 1, 1);                                    \
        B(J, K, 1, 2) = B(J, K, 1, 2) * B(J, K, 1, 1);                                    \
    } while (0)

#define BPLUS(A, B, C, J, K, JP, KP)                                                      \
    do {                                                                                  \
        for (int n_ = 1; n_ <= 3; ++n_)                                                   \
            for (int m_ = 1; m_ <= 3; ++m_)                                               \
                B(J, K, m_, n_) = B(J, K, m_, n_) + A(J, K, m_, 1) * C(JP, KP, 1, n_) +   \
                                  A(J, K, m_, 2) * C(JP, KP, 2, n_) +                     \
                                  A(J, K, m_, 3) * C(JP, KP, 3, n_);                      \
    } while (0)

/* c <- b * c with b recomposed (ref/btlu3.f:700-735) */
#define CTIMES(B, C, J, K)                                                                \
    do {                                                                                  \
        double t_[3][3];                                                                  \
        for (int n_ = 1; n_ <= 3; ++n_)                                                   \
            for (int m_ = 1; m_ <= 3; ++m_)                                               \
                t_[m_ - 1][n_ - 1] = B(J, K, m_, 1) * C(J, K, 1, n_) + B(J, K, m_, 2) * C(J, K, 2, n_) + \
                                     B(J, K, m_, 3) * C(J, K, 3, n_);                     \
        for (int n_ = 1; n_ <= 3; ++n_)                                                   \
            for (int m_ = 1; m_ <= 3; ++m_) C(J, K, m_, n_) = t_[m_ - 1][n_ - 1];         \
    } while (0)

static void btund3j(int jm, int km, double *a, double *b, double *c, int jbeg, int jend, int kbeg,
                    int kend)
{
    /* ref/btlu3.f:628-838 */
    long N = (long)jm * km;
    for (int j = jbeg; j <= jend; ++j)
        for (int k = kbeg; k <= kend; ++k) UNLU3(B_, j, k);
    for (int k = kbeg; k <= kend; ++k) BPLUS(A_, B_, C_, jend, k, jend - 1, k);
    for (int j = jend - 1; j >= jbeg + 1; --j) {
        for (int k = kbeg; k <= kend; ++k) CTIMES(B_, C_, j, k);
        for (int k = kbeg; k <= kend; ++k) BPLUS(A_, B_, C_, j, k, j - 1, k);
    }
    for (int k = kbeg; k <= kend; ++k) CTIMES(B_, C_, jbeg, k);
}
#undef A_
#undef B_
#undef C_

/* ------------------------------------------------------------------------------------ */
/* line relaxation: ref/lrelax.f:51-561                                                   */

typedef struct {
    double *a, *b, *c, *rhs; /* /btri/ line copies, length n x 9 / n x 3 */
} btri;

static void line_solve(int n, btri *t)
{
    /* forward with the factored b and a, back with c (ref/lrelax.f:331-373) */
#define TA(i, m, nn) t->a[((i)-1) * 9 + ((m)-1) + 3 * ((nn)-1)]
#define TB(i, m, nn) t->b[((i)-1) * 9 + ((m)-1) + 3 * ((nn)-1)]
#define TC(i, m, nn) t->c[((i)-1) * 9 + ((m)-1) + 3 * ((nn)-1)]
#define TR(i, m) t->rhs[((i)-1) * 3 + ((m)-1)]
    for (int i = 1; i <= n; ++i) {
        if (i > 1)
            for (int m = 1; m <= 3; ++m)
                TR(i, m) = TR(i, m) - TA(i, m, 1) * TR(i - 1, 1) - TA(i, m, 2) * TR(i - 1, 2) -
                           TA(i, m, 3) * TR(i - 1, 3);
        double d1 = TB(i, 1, 1) * TR(i, 1);
        double d2 = TB(i, 2, 2) * (TR(i, 2) - TB(i, 2, 1) * d1);
        double d3 = TB(i, 3, 3) * (TR(i, 3) - TB(i, 3, 1) * d1 - TB(i, 3, 2) * d2);
        TR(i, 3) = d3;
        TR(i, 2) = d2 - TB(i, 2, 3) * TR(i, 3);
        TR(i, 1) = d1 - TB(i, 1, 3) * TR(i, 3) - TB(i, 1, 2) * TR(i, 2);
    }
    for (int i = n - 1; i >= 1; --i)
        for (int m = 1; m <= 3; ++m)
            TR(i, m) = TR(i, m) - TC(i, m, 1) * TR(i + 1, 1) - TC(i, m, 2) * TR(i + 1, 2) -
                       TC(i, m, 3) * TR(i + 1, 3);
#undef TA
#undef TB
#undef TC
#undef TR
}

static void ljsweep(int jm, int km, int nj, double underr, double *dq, const double *s,
                    const double *btc, const double *bjm, const double *bjp, const double *bkm,
                    const double *bkp, btri *t)
{
    long N = (long)jm * km;
    int kb, ke, ki;
    if ((nj / 2) * 2 == nj) { kb = km; ke = 1; ki = -1; }
    else { kb = 1; ke = km; ki = 1; }
    const double half = 0.5;
    for (int k = kb; ki > 0 ? k <= ke : k >= ke; k += ki) {
        int km_ = k - 1 > 1 ? k - 1 : 1;
        int kp_ = k + 1 < km ? k + 1 : km;
        double rkm = (double)(float)(k - 1 - km_);
        double rkp = (double)(float)(kp_ - k - 1);
        double swkm = half + fsign(half, rkm);
        double swkp = half + fsign(half, rkp);
        for (int j = 1; j <= jm; ++j)
            for (int m = 1; m <= 3; ++m)
                t->rhs[(j - 1) * 3 + m - 1] =
                    Q3(s, j, k, m) -
                    (B9(bkm, j, k, m, 1) * Q3(dq, j, km_, 1) + B9(bkm, j, k, m, 2) * Q3(dq, j, km_, 2) +
                     B9(bkm, j, k, m, 3) * Q3(dq, j, km_, 3)) * swkm -
                    (B9(bkp, j, k, m, 1) * Q3(dq, j, kp_, 1) + B9(bkp, j, k, m, 2) * Q3(dq, j, kp_, 2) +
                     B9(bkp, j, k, m, 3) * Q3(dq, j, kp_, 3)) * swkp;
        for (int j = 1; j <= jm; ++j)
            for (int n = 1; n <= 3; ++n)
                for (int m = 1; m <= 3; ++m) {
                    t->a[(j - 1) * 9 + (m - 1) + 3 * (n - 1)] = B9(bjm, j, k, m, n);
                    t->b[(j - 1) * 9 + (m - 1) + 3 * (n - 1)] = B9(btc, j, k, m, n);
                    t->c[(j - 1) * 9 + (m - 1) + 3 * (n - 1)] = B9(bjp, j, k, m, n);
                }
        line_solve(jm, t);
        for (int j = 1; j <= jm; ++j)
            for (int m = 1; m <= 3; ++m)
                Q3(dq, j, k, m) = Q3(dq, j, k, m) + underr * (t->rhs[(j - 1) * 3 + m - 1] - Q3(dq, j, k, m));
    }
}

static void lksweep(int jm, int km, int nk, double underr, double *dq, const double *s,
                    const double *btc, const double *bjm, const double *bjp, const double *bkm,
                    const double *bkp, btri *t)
{
    long N = (long)jm * km;
    int jb, je, ji;
    if ((nk / 2) * 2 == nk) { jb = jm; je = 1; ji = -1; }
    else { jb = 1; je = jm; ji = 1; }
    const double half = 0.5;
    for (int j = jb; ji > 0 ? j <= je : j >= je; j += ji) {
        int jm_ = j - 1 > 1 ? j - 1 : 1;
        int jp_ = j + 1 < jm ? j + 1 : jm;
        double rjm = (double)(float)(j - 1 - jm_);
        double rjp = (double)(float)(jp_ - j - 1);
        double swjm = half + fsign(half, rjm);
        double swjp = half + fsign(half, rjp);
        for (int k = 1; k <= km; ++k)
            for (int m = 1; m <= 3; ++m)
                t->rhs[(k - 1) * 3 + m - 1] =
                    Q3(s, j, k, m) -
                    (B9(bjm, j, k, m, 1) * Q3(dq, jm_, k, 1) + B9(bjm, j, k, m, 2) * Q3(dq, jm_, k, 2) +
                     B9(bjm, j, k, m, 3) * Q3(dq, jm_, k, 3)) * swjm -
                    (B9(bjp, j, k, m, 1) * Q3(dq, jp_, k, 1) + B9(bjp, j, k, m, 2) * Q3(dq, jp_, k, 2) +
                     B9(bjp, j, k, m, 3) * Q3(dq, jp_, k, 3)) * swjp;
        for (int k = 1; k <= km; ++k)
            for (int n = 1; n <= 3; ++n)
                for (int m = 1; m <= 3; ++m) {
                    t->a[(k - 1) * 9 + (m - 1) + 3 * (n - 1)] = B9(bkm, j, k, m, n);
                    t->b[(k - 1) * 9 + (m - 1) + 3 * (n - 1)] = B9(btc, j, k, m, n);
                    t->c[(k - 1) * 9 + (m - 1) + 3 * (n - 1)] = B9(bkp, j, k, m, n);
                }
        line_solve(km, t);
        for (int k = 1; k <= km; ++k)
            for (int m = 1; m <= 3; ++m)
                Q3(dq, j, k, m) = Q3(dq, j, k, m) + underr * (t->rhs[(k - 1) * 3 + m - 1] - Q3(dq, j, k, m));
    }
}

int orc_lrelax(const smac_params *p, int jm, int km, int npair, const int *wake, const double *q,
               double *dq, double *s, double *btc, double *bjm, double *bjp, double *bkm,
               double *bkp)
{
    long N = (long)jm * km;
    int nmax = jm > km ? jm : km;
    btri t;
    t.a = malloc(sizeof(double) * nmax * 30);
    t.b = t.a + 9 * nmax;
    t.c = t.b + 9 * nmax;
    t.rhs = t.c + 9 * nmax;
    for (long i = 0; i < 3 * N; ++i) dq[i] = 0.0;
    if (p->njsp > 0) btlu3j(jm, km, bjm, btc, bjp, 1, jm, 1, km);
    for (int nj = 1; nj <= p->njsp; ++nj) {
        ljsweep(jm, km, nj, p->underr, dq, s, btc, bjm, bjp, bkm, bkp, &t);
        if ((nj / p->nbcimp) * p->nbcimp == nj) bcimpds(jm, km, 3, npair, wake, q, dq, s);
    }
    if (p->njsp > 0) btund3j(jm, km, bjm, btc, bjp, 1, jm, 1, km);
    if (p->nksp > 0) btlu3k(jm, km, bkm, btc, bkp, 1, jm, 1, km);
    for (int nk = 1; nk <= p->nksp; ++nk) {
        lksweep(jm, km, nk, p->underr, dq, s, btc, bjm, bjp, bkm, bkp, &t);
        if ((nk / p->nbcimp) * p->nbcimp == nk) bcimpds(jm, km, 3, npair, wake, q, dq, s);
    }
    /* btund3k (ref/lrelax.f:215) recomposes btc/bkp; nothing reads them afterwards */
    free(t.a);
    return SMAC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* LU-SGS: ref/lusgs.f:51-236, ref/ilusol.f:50-591                                        */

static void lusini3(long N, const double *btc, double *bp)
{
    for (long i = 0; i < N; ++i) {
#define BT(m, n) btc[i + N * ((m)-1) + 3 * N * ((n)-1)]
#define BP(m, n) bp[i + N * ((m)-1) + 3 * N * ((n)-1)]
        BP(1, 1) = 1. / BT(1, 1);
        BP(1, 2) = BT(1, 2) * BP(1, 1);
        BP(1, 3) = BT(1, 3) * BP(1, 1);
        BP(2, 1) = BT(2, 1);
        BP(2, 2) = 1. / (BT(2, 2) - BP(2, 1) * BP(1, 2));
        BP(2, 3) = (BT(2, 3) - BP(2, 1) * BP(1, 3)) * BP(2, 2);
        BP(3, 1) = BT(3, 1);
        BP(3, 2) = BT(3, 2) - BP(3, 1) * BP(1, 2);
        BP(3, 3) = 1. / (BT(3, 3) - BP(3, 1) * BP(1, 3) - BP(3, 2) * BP(2, 3));
#undef BT
#undef BP
    }
}

static void ilusol3(int jm, int km, double *dq, double *rhs, const double *b, const double *bjm,
                    const double *bjp, const double *bkm, const double *bkp)
{
    long N = (long)jm * km;
#define LB(j, k, m, n) B9(b, j, k, m, n)
#define LUSOLVE(OUT, J, K, R1, R2, R3)                                                    \
    do {                                                                                  \
        double z1 = (R1) * LB(J, K, 1, 1);                                                \
        double z2 = ((R2) - LB(J, K, 2, 1) * z1) * LB(J, K, 2, 2);                        \
        double z3 = ((R3) - LB(J, K, 3, 1) * z1 - LB(J, K, 3, 2) * z2) * LB(J, K, 3, 3);  \
        Q3(OUT, J, K, 3) = z3;                                                            \
        Q3(OUT, J, K, 2) = z2 - LB(J, K, 2, 3) * Q3(OUT, J, K, 3);                        \
        Q3(OUT, J, K, 1) = z1 - LB(J, K, 1, 3) * Q3(OUT, J, K, 3) - LB(J, K, 1, 2) * Q3(OUT, J, K, 2); \
    } while (0)
    /* forward sweep, natural order (ref/ilusol.f:80-300) */
    LUSOLVE(rhs, 1, 1, Q3(rhs, 1, 1, 1), Q3(rhs, 1, 1, 2), Q3(rhs, 1, 1, 3));
    for (int j = 2; j <= jm; ++j) {
        double rr[3];
        for (int m = 1; m <= 3; ++m)
            rr[m - 1] = Q3(rhs, j, 1, m) - B9(bjm, j, 1, m, 1) * Q3(rhs, j - 1, 1, 1) -
                        B9(bjm, j, 1, m, 2) * Q3(rhs, j - 1, 1, 2) - B9(bjm, j, 1, m, 3) * Q3(rhs, j - 1, 1, 3);
        LUSOLVE(rhs, j, 1, rr[0], rr[1], rr[2]);
    }
    for (int k = 2; k <= km; ++k)
        for (int j = 1; j <= jm; ++j) {
            int jw = j == 1 ? jm : j - 1, kw = j == 1 ? k - 1 : k;
            double rr[3];
            for (int m = 1; m <= 3; ++m)
                rr[m - 1] = Q3(rhs, j, k, m) - B9(bjm, j, k, m, 1) * Q3(rhs, jw, kw, 1) -
                            B9(bjm, j, k, m, 2) * Q3(rhs, jw, kw, 2) - B9(bjm, j, k, m, 3) * Q3(rhs, jw, kw, 3) -
                            B9(bkm, j, k, m, 1) * Q3(rhs, j, k - 1, 1) - B9(bkm, j, k, m, 2) * Q3(rhs, j, k - 1, 2) -
                            B9(bkm, j, k, m, 3) * Q3(rhs, j, k - 1, 3);
            LUSOLVE(rhs, j, k, rr[0], rr[1], rr[2]);
        }
    /* multiply by D recomposed from its factors (ref/ilusol.f:307-341) */
    for (int k = 1; k <= km; ++k)
        for (int j = 1; j <= jm; ++j) {
            double b11 = 1. / LB(j, k, 1, 1);
            double b12 = LB(j, k, 1, 2) * b11;
            double b13 = LB(j, k, 1, 3) * b11;
            double b21 = LB(j, k, 2, 1);
            double b22 = 1. / LB(j, k, 2, 2) + LB(j, k, 2, 1) * LB(j, k, 1, 2);
            double b23 = LB(j, k, 2, 3) / LB(j, k, 2, 2) + LB(j, k, 2, 1) * LB(j, k, 1, 3);
            double b31 = LB(j, k, 3, 1);
            double b32 = LB(j, k, 3, 2) + LB(j, k, 3, 1) * LB(j, k, 1, 2);
            double b33 = 1. / LB(j, k, 3, 3) + LB(j, k, 3, 1) * LB(j, k, 1, 3) + LB(j, k, 3, 2) * LB(j, k, 2, 3);
            double r1 = Q3(rhs, j, k, 1), r2 = Q3(rhs, j, k, 2), r3 = Q3(rhs, j, k, 3);
            Q3(rhs, j, k, 1) = b11 * r1 + b12 * r2 + b13 * r3;
            Q3(rhs, j, k, 2) = b21 * r1 + b22 * r2 + b23 * r3;
            Q3(rhs, j, k, 3) = b31 * r1 + b32 * r2 + b33 * r3;
        }
    /* backward sweep (ref/ilusol.f:353-469) */
    LUSOLVE(dq, jm, km, Q3(rhs, jm, km, 1), Q3(rhs, jm, km, 2), Q3(rhs, jm, km, 3));
    for (int j = jm - 1; j >= 1; --j) {
        double rr[3];
        for (int m = 1; m <= 3; ++m)
            rr[m - 1] = Q3(rhs, j, km, m) - B9(bjp, j, km, m, 1) * Q3(dq, j + 1, km, 1) -
                        B9(bjp, j, km, m, 2) * Q3(dq, j + 1, km, 2) - B9(bjp, j, km, m, 3) * Q3(dq, j + 1, km, 3);
        LUSOLVE(dq, j, km, rr[0], rr[1], rr[2]);
    }
    for (int k = km - 1; k >= 1; --k)
        for (int j = jm; j >= 1; --j) {
            double rr[3];
            for (int m = 1; m <= 3; ++m) {
                /* j = jmax reads dq(jmax,kmax) for rows 1-2 and dq(1,kmax) for row 3
                 * through bjp(jmax,k), as the reference does (ref/ilusol.f:415-436) */
                int jx = j == jm ? (m == 3 ? 1 : jm) : j + 1;
                int kx = j == jm ? km : k;
                rr[m - 1] = Q3(rhs, j, k, m) - B9(bkp, j, k, m, 1) * Q3(dq, j, k + 1, 1) -
                            B9(bkp, j, k, m, 2) * Q3(dq, j, k + 1, 2) - B9(bkp, j, k, m, 3) * Q3(dq, j, k + 1, 3) -
                            B9(bjp, j, k, m, 1) * Q3(dq, jx, kx, 1) - B9(bjp, j, k, m, 2) * Q3(dq, jx, kx, 2) -
                            B9(bjp, j, k, m, 3) * Q3(dq, jx, kx, 3);
            }
            LUSOLVE(dq, j, k, rr[0], rr[1], rr[2]);
        }
#undef LB
#undef LUSOLVE
}

static void ilusol1(int jm, int km, double *dq, double *rhs, const double *b, const double *bjm,
                    const double *bjp, const double *bkm, const double *bkp)
{
    /* ref/ilusol.f:486-591 */
    long N = (long)jm * km;
    (void)N;
    S2(rhs, 1, 1) = S2(rhs, 1, 1) * S2(b, 1, 1);
    for (int i = 2; i <= jm; ++i)
        S2(rhs, i, 1) = S2(b, i, 1) * (S2(rhs, i, 1) - S2(bjm, i, 1) * S2(rhs, i - 1, 1));
    for (int k = 2; k <= km; ++k)
        for (int j = 1; j <= jm; ++j) {
            if (j == 1)
                S2(rhs, j, k) = S2(b, j, k) * (S2(rhs, j, k) - S2(bjm, j, k) * S2(rhs, jm, k - 1) -
                                               S2(bkm, j, k) * S2(rhs, j, k - 1));
            else
                S2(rhs, j, k) = S2(b, j, k) * (S2(rhs, j, k) - S2(bjm, j, k) * S2(rhs, j - 1, k) -
                                               S2(bkm, j, k) * S2(rhs, j, k - 1));
        }
    for (int k = 1; k <= km; ++k)
        for (int j = 1; j <= jm; ++j) S2(rhs, j, k) = S2(rhs, j, k) / S2(b, j, k);
    S2(dq, jm, km) = S2(rhs, jm, km) * S2(b, jm, km);
    for (int j = jm - 1; j >= 1; --j)
        S2(dq, j, km) = S2(b, j, km) * (S2(rhs, j, km) - S2(bjp, j, km) * S2(dq, j + 1, km));
    for (int k = km - 1; k >= 1; --k)
        for (int j = jm; j >= 1; --j) {
            if (j == jm)
                S2(dq, j, k) = S2(b, j, k) * (S2(rhs, j, k) - S2(bkp, j, k) * S2(dq, j, k + 1) -
                                              S2(bjp, j, k) * S2(dq, 1, k + 1));
            else
                S2(dq, j, k) = S2(b, j, k) * (S2(rhs, j, k) - S2(bkp, j, k) * S2(dq, j, k + 1) -
                                              S2(bjp, j, k) * S2(dq, j + 1, k));
        }
}

int orc_lusgs(const smac_params *p, int jm, int km, int neqs, int npair, const int *wake,
              const double *q, double *dq, double *rhs, double *btc, double *bjm, double *bjp,
              double *bkm, double *bkp, double *scr1)
{
    (void)p;
    long N = (long)jm * km;
    if (neqs == 3) {
        lusini3(N, btc, scr1);
        ilusol3(jm, km, dq, rhs, scr1, bjm, bjp, bkm, bkp);
    } else if (neqs == 1) {
        for (long i = 0; i < N; ++i) scr1[i] = 1. / btc[i];
        ilusol1(jm, km, dq, rhs, scr1, bjm, bjp, bkm, bkp);
    } else {
        return SMAC_ERR_NEQS;
    }
    bcimpds(jm, km, neqs, npair, wake, q, dq, rhs);
    return SMAC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* Spalart-Allmaras: ref/tursa.f (compiled at -O0 in the reference build: see top)        */

#define F32(v) ((double)(v##f))

static void vortsa(orc_case *c)
{
    /* ref/tursa.f:949-1028; derivative planes go to sjm, sjp, skm, skp as the reference
     * passes them (uj, vj, uk, vk) */
    const int jm = c->jm, km = c->km;
    const long N = c->N;
    double *q = c->q, *rtxy = c->rtxy, *dj = c->dj;
    double *uj = c->sjm, *vj = c->sjp, *uk = c->skm, *vk = c->skp;
    for (int j = 2; j <= jm - 1; ++j)
        for (int k = 1; k <= km; ++k) {
            S2(uj, j, k) = 0.5 * (Q3(q, j + 1, k, 2) - Q3(q, j - 1, k, 2));
            S2(vj, j, k) = 0.5 * (Q3(q, j + 1, k, 3) - Q3(q, j - 1, k, 3));
        }
    for (int k = 1; k <= km; ++k) {
        int j = 1;
        S2(uj, j, k) = 0.5 * (-Q3(q, j + 2, k, 2) + 4. * Q3(q, j + 1, k, 2) - 3. * Q3(q, j, k, 2));
        S2(vj, j, k) = 0.5 * (-Q3(q, j + 2, k, 3) + 4. * Q3(q, j + 1, k, 3) - 3. * Q3(q, j, k, 3));
        j = jm;
        S2(uj, j, k) = 0.5 * (3. * Q3(q, j, k, 2) - 4. * Q3(q, j - 1, k, 2) + Q3(q, j - 2, k, 2));
        S2(vj, j, k) = 0.5 * (3. * Q3(q, j, k, 3) - 4. * Q3(q, j - 1, k, 3) + Q3(q, j - 2, k, 3));
    }
    for (int k = 2; k <= km - 1; ++k)
        for (int j = 1; j <= jm; ++j) {
            S2(uk, j, k) = 0.5 * (Q3(q, j, k + 1, 2) - Q3(q, j, k - 1, 2));
            S2(vk, j, k) = 0.5 * (Q3(q, j, k + 1, 3) - Q3(q, j, k - 1, 3));
        }
    for (int j = 1; j <= jm; ++j) {
        int k = 1;
        S2(uk, j, k) = 0.5 * (-Q3(q, j, k + 2, 2) + 4. * Q3(q, j, k + 1, 2) - 3. * Q3(q, j, k, 2));
        S2(vk, j, k) = 0.5 * (-Q3(q, j, k + 2, 3) + 4. * Q3(q, j, k + 1, 3) - 3. * Q3(q, j, k, 3));
        k = km;
        S2(uk, j, k) = 0.5 * (3. * Q3(q, j, k, 2) - 4. * Q3(q, j, k - 1, 2) + Q3(q, j, k - 2, 2));
        S2(vk, j, k) = 0.5 * (3. * Q3(q, j, k, 3) - 4. * Q3(q, j, k - 1, 3) + Q3(q, j, k - 2, 3));
    }
    for (int k = 1; k <= km; ++k)
        for (int j = 1; j <= jm; ++j)
            S2(c->vort, j, k) = fabs(S2(dj, j, k) * (RT(j, k, 1, 2) * S2(vj, j, k) + RT(j, k, 2, 2) * S2(vk, j, k) -
                                                     RT(j, k, 1, 3) * S2(uj, j, k) - RT(j, k, 2, 3) * S2(uk, j, k)));
}

static void curvature(orc_case *c, double cr1, double cr2, double cr3)
{
    /* ref/tursa.f:1032-1169 */
    const int jm = c->jm, km = c->km;
    const long N = c->N;
    double *q = c->q, *rtxy = c->rtxy, *dj = c->dj, *fr1 = c->fr1;
    double *uj = c->sjm, *vj = c->sjp, *uk = c->skm, *vk = c->skp;
    for (int j = 2; j <= jm - 1; ++j)
        for (int k = 1; k <= km; ++k) {
            double ujc = Q3(q, j + 1, k, 2) - Q3(q, j - 1, k, 2);
            double vjc = Q3(q, j + 1, k, 3) - Q3(q, j - 1, k, 3);
            S2(uj, j, k) = .5 * ujc;
            S2(vj, j, k) = .5 * vjc;
        }
    for (int k = 1; k <= km; ++k) {
        int j = 1;
        S2(uj, j, k) = 0.5 * (-Q3(q, j + 2, k, 2) + 4. * Q3(q, j + 1, k, 2) - 3. * Q3(q, j, k, 2));
        S2(vj, j, k) = 0.5 * (-Q3(q, j + 2, k, 3) + 4. * Q3(q, j + 1, k, 3) - 3. * Q3(q, j, k, 3));
        j = jm;
        S2(uj, j, k) = 0.5 * (3. * Q3(q, j, k, 2) - 4. * Q3(q, j - 1, k, 2) + Q3(q, j - 2, k, 2));
        S2(vj, j, k) = 0.5 * (3. * Q3(q, j, k, 3) - 4. * Q3(q, j - 1, k, 3) + Q3(q, j - 2, k, 3));
    }
    for (int k = 2; k <= km - 1; ++k)
        for (int j = 1; j <= jm; ++j) {
            double ukc = Q3(q, j, k + 1, 2) - Q3(q, j, k - 1, 2);
            double vkc = Q3(q, j, k + 1, 3) - Q3(q, j, k - 1, 3);
            S2(uk, j, k) = .5 * ukc;
            S2(vk, j, k) = .5 * vkc;
        }
    for (int j = 1; j <= jm; ++j) {
        int k = 1;
        S2(uk, j, k) = 0.5 * (-Q3(q, j, k + 2, 2) + 4. * Q3(q, j, k + 1, 2) - 3. * Q3(q, j, k, 2));
        S2(vk, j, k) = 0.5 * (-Q3(q, j, k + 2, 3) + 4. * Q3(q, j, k + 1, 3) - 3. * Q3(q, j, k, 3));
        k = km;
        S2(uk, j, k) = 0.5 * (3. * Q3(q, j, k, 2) - 4. * Q3(q, j, k - 1, 2) + Q3(q, j, k - 2, 2));
        S2(vk, j, k) = 0.5 * (3. * Q3(q, j, k, 3) - 4. * Q3(q, j, k - 1, 3) + Q3(q, j, k - 2, 3));
    }
    for (int k = 1; k <= km; ++k)
        for (int j = 1; j <= jm; ++j) {
            double d = S2(dj, j, k);
            double ux = d * (RT(j, k, 1, 2) * S2(uj, j, k) + RT(j, k, 2, 2) * S2(uk, j, k));
            double vx = d * (RT(j, k, 1, 2) * S2(vj, j, k) + RT(j, k, 2, 2) * S2(vk, j, k));
            double uy = d * (RT(j, k, 1, 3) * S2(uj, j, k) + RT(j, k, 2, 3) * S2(uk, j, k));
            double vy = d * (RT(j, k, 1, 3) * S2(vj, j, k) + RT(j, k, 2, 3) * S2(vk, j, k));
            S2(vj, j, k) = sqrt(2. * (ux * ux + vx * uy + vy * vy) + uy * uy + vx * vx);
            if (vx - uy != 0.) S2(vj, j, k) = S2(vj, j, k) / fabs(vx - uy);
            S2(uj, j, k) = (ux - vy) / 2;
            S2(uk, j, k) = (uy + vx) / 2;
            S2(fr1, j, k) = 2 * (vx - uy);
            double denom = (2. * (ux * ux + vx * uy + vy * vy) + uy * uy + vx * vx) *
                           (ux * ux + vx * vx + uy * uy + vy * vy);
            if (denom != 0.) S2(fr1, j, k) = S2(fr1, j, k) / denom;
        }
    for (int k = 1; k <= km; ++k) {
        int kmn = k - 1 > 1 ? k - 1 : 1;
        int kpl = k + 1 < km ? k + 1 : km;
        double dk = (double)(kpl - kmn);
        for (int j = 2; j <= jm - 1; ++j) {
            double uu = (RT(j, k, 2, 2) * Q3(q, j, k, 2) + RT(j, k, 2, 3) * Q3(q, j, k, 3)) * S2(dj, j, k);
            double vv = (RT(j, k, 1, 2) * Q3(q, j, k, 2) + RT(j, k, 1, 3) * Q3(q, j, k, 3)) * S2(dj, j, k);
            double ds11dt = uu * (S2(uj, j, kpl) - S2(uj, j, kmn)) / dk + vv * (S2(uj, j + 1, k) - S2(uj, j - 1, k)) / 2;
            double ds12dt = uu * (S2(uk, j, kpl) - S2(uk, j, kmn)) / dk + vv * (S2(uk, j + 1, k) - S2(uk, j - 1, k)) / 2;
            double rtilde = (S2(uj, j, k) * ds12dt - S2(uk, j, k) * ds11dt) * S2(fr1, j, k);
            S2(fr1, j, k) = (1 + cr1) * 2 * S2(vj, j, k) / (1 + S2(vj, j, k)) * (1 - cr3 * atan(cr2 * rtilde)) - cr1;
        }
        {
            int j = 1;
            double uu = (RT(j, k, 2, 2) * Q3(q, j, k, 2) + RT(j, k, 2, 3) * Q3(q, j, k, 3)) * S2(dj, j, k);
            double vv = (RT(j, k, 1, 2) * Q3(q, j, k, 2) + RT(j, k, 1, 3) * Q3(q, j, k, 3)) * S2(dj, j, k);
            double ds11dt = uu * (S2(uj, j, kpl) - S2(uj, j, kmn)) / dk + vv * (S2(uj, 2, k) - S2(uj, 1, k));
            double ds12dt = uu * (S2(uk, j, kpl) - S2(uk, j, kmn)) / dk + vv * (S2(uk, 2, k) - S2(uk, 1, k));
            double rtilde = (S2(uj, j, k) * ds12dt - S2(uk, j, k) * ds11dt) * S2(fr1, j, k);
            S2(fr1, j, k) = (1 + cr1) * 2 * S2(vj, j, k) / (1 + S2(vj, j, k)) * (1 - cr3 * atan(cr2 * rtilde)) - cr1;
        }
        {
            int j = jm;
            double uu = (RT(j, k, 2, 2) * Q3(q, j, k, 2) + RT(j, k, 2, 3) * Q3(q, j, k, 3)) * S2(dj, j, k);
            double vv = (RT(j, k, 1, 2) * Q3(q, j, k, 2) + RT(j, k, 1, 3) * Q3(q, j, k, 3)) * S2(dj, j, k);
            double ds11dt = uu * (S2(uj, j, kpl) - S2(uj, j, kmn)) / dk + vv * (S2(uj, jm, k) - S2(uj, jm - 1, k));
            double ds12dt = uu * (S2(uk, j, kpl) - S2(uk, j, kmn)) / dk + vv * (S2(uk, jm, k) - S2(uk, jm - 1, k));
            double rtilde = (S2(uj, j, k) * ds12dt - S2(uk, j, k) * ds11dt) * S2(fr1, j, k);
            S2(fr1, j, k) = (1 + cr1) * 2 * S2(vj, j, k) / (1 + S2(vj, j, k)) * (1 - cr3 * atan(cr2 * rtilde)) - cr1;
        }
    }
}

static void tsafill(orc_case *c)
{
    /* ref/tursa.f:437-945 (itrans reads 0 in the reference build: ct3tmp = 0) */
    const int jm = c->jm, km = c->km;
    const long N = c->N;
    double *q = c->q, *rtxy = c->rtxy, *dj = c->dj, *anut = c->anut;
    const double vnu = c->vnu;
    const double akarman = F32(0.41), cb1 = F32(0.1355), cb2 = F32(0.622);
    const double two = 2.0, three = 3.0, sigma = two / three;
    const double cw1 = cb1 / pw_sq(akarman, 2) + (1. + cb2) / sigma;
    const double cw2 = F32(0.3), cw3 = 2.0, cv1 = F32(7.1), cv2 = 5.0;
    const double ct3 = F32(1.2), ct4 = F32(0.5);
    const double cr1 = 1, cr2 = 12, cr3 = 1;
    const double cw3r6 = pw_sq(cw3, 6);
    /* ref/tursa.f:460 reads `epss3=1.e-20` but runs past column 72 of fixed-form source, so
     * the compiler sees `epss3=1`: sst is clamped at 1, not 1e-20 */
    const double ten = 10.0, epss = F32(1.e-12), epss2 = F32(0.00001), epss3 = 1.0;
    const double ct3tmp = 0.0;
    (void)ct3;
    vortsa(c);
    curvature(c, cr1, cr2, cr3);
    double *sjm = c->sjm, *sjp = c->sjp, *skm = c->skm, *skp = c->skp, *stc = c->stc;
    double *diff = c->diff, *conv = c->conv, *sf = c->sf;
    /* sjm..skp held the derivative planes; the reference overwrites them below only at
     * interior points, so the boundary entries keep those values until tsaibt zeroes them */
    for (int k = 2; k <= km - 1; ++k) {
        int km1 = k - 1, kp1 = k + 1;
        for (int j = 2; j <= jm - 1; ++j) {
            double xy3p = .5 * (RT(j, k, 2, 2) * S2(dj, j, k) + RT(j, kp1, 2, 2) * S2(dj, j, kp1));
            double xy4p = .5 * (RT(j, k, 2, 3) * S2(dj, j, k) + RT(j, kp1, 2, 3) * S2(dj, j, kp1));
            double ttp = (xy3p * RT(j, k, 2, 2) * S2(dj, j, k) + xy4p * RT(j, k, 2, 3) * S2(dj, j, k));
            double xy3m = .5 * (RT(j, k, 2, 2) * S2(dj, j, k) + RT(j, km1, 2, 2) * S2(dj, j, km1));
            double xy4m = .5 * (RT(j, k, 2, 3) * S2(dj, j, k) + RT(j, km1, 2, 3) * S2(dj, j, km1));
            double ttm = (xy3m * RT(j, k, 2, 2) * S2(dj, j, k) + xy4m * RT(j, k, 2, 3) * S2(dj, j, k));
            double coef = -(cb2 * vnu * (1. + S2(anut, j, k)) / sigma);
            S2(diff, j, k) = coef * (ttp * (S2(anut, j, kp1) - S2(anut, j, k)) - ttm * (S2(anut, j, k) - S2(anut, j, km1)));
            S2(skm, j, k) = -(coef * ttm);
            S2(skp, j, k) = -(coef * ttp);
            S2(stc, j, k) = coef * (ttm + ttp);
            coef = vnu * (1. + cb2) / sigma;
            double vnup = 1.0 + 0.5 * (S2(anut, j, kp1) + S2(anut, j, k));
            double vnum = 1.0 + 0.5 * (S2(anut, j, km1) + S2(anut, j, k));
            ttp = ttp * vnup;
            ttm = ttm * vnum;
            S2(diff, j, k) = S2(diff, j, k) + coef * (ttp * (S2(anut, j, kp1) - S2(anut, j, k)) -
                                                      ttm * (S2(anut, j, k) - S2(anut, j, km1)));
            S2(skm, j, k) = S2(skm, j, k) - coef * ttm;
            S2(skp, j, k) = S2(skp, j, k) - coef * ttp;
            S2(stc, j, k) = S2(stc, j, k) + coef * (ttm + ttp);
        }
    }
    const double one = 1.0;
    for (int k = 2; k <= km - 1; ++k) {
        int km1 = k - 1, kp1 = k + 1;
        for (int j = 2; j <= jm - 1; ++j) {
            double uu = (RT(j, k, 2, 2) * Q3(q, j, k, 2) + RT(j, k, 2, 3) * Q3(q, j, k, 3)) * S2(dj, j, k);
            double sgnu = fsign(one, uu);
            double app = .5 * (1. + sgnu), apm = .5 * (1. - sgnu);
            S2(conv, j, k) = uu * (app * (S2(anut, j, k) - S2(anut, j, km1)) + apm * (S2(anut, j, kp1) - S2(anut, j, k)));
            S2(skm, j, k) = S2(skm, j, k) - uu * app;
            S2(stc, j, k) = S2(stc, j, k) + uu * (app - apm);
            S2(skp, j, k) = S2(skp, j, k) + uu * apm;
        }
    }
    for (int k = 2; k <= km - 1; ++k)
        for (int j = 2; j <= jm - 1; ++j) {
            int jp1 = j + 1, jm1 = j - 1;
            double xy1p = .5 * (RT(j, k, 1, 2) * S2(dj, j, k) + RT(jp1, k, 1, 2) * S2(dj, jp1, k));
            double xy2p = .5 * (RT(j, k, 1, 3) * S2(dj, j, k) + RT(jp1, k, 1, 3) * S2(dj, jp1, k));
            double ttp = (xy1p * RT(j, k, 1, 2) * S2(dj, j, k) + xy2p * RT(j, k, 1, 3) * S2(dj, j, k));
            double xy1m = .5 * (RT(j, k, 1, 2) * S2(dj, j, k) + RT(jm1, k, 1, 2) * S2(dj, jm1, k));
            double xy2m = .5 * (RT(j, k, 1, 3) * S2(dj, j, k) + RT(jm1, k, 1, 3) * S2(dj, jm1, k));
            double ttm = (xy1m * RT(j, k, 1, 2) * S2(dj, j, k) + xy2m * RT(j, k, 1, 3) * S2(dj, j, k));
            double coef = -(cb2 * vnu * (1. + S2(anut, j, k)) / sigma);
            S2(diff, j, k) = S2(diff, j, k) + coef * (ttp * (S2(anut, j + 1, k) - S2(anut, j, k)) -
                                                      ttm * (S2(anut, j, k) - S2(anut, j - 1, k)));
            S2(sjm, j, k) = -(coef * ttm);
            S2(sjp, j, k) = -(coef * ttp);
            S2(stc, j, k) = coef * (ttm + ttp) + S2(stc, j, k);
            coef = vnu * (1. + cb2) / sigma;
            double vnup = 1.0 + 0.5 * (S2(anut, j + 1, k) + S2(anut, j, k));
            double vnum = 1.0 + 0.5 * (S2(anut, j - 1, k) + S2(anut, j, k));
            ttp = ttp * vnup;
            ttm = ttm * vnum;
            S2(diff, j, k) = S2(diff, j, k) + coef * (ttp * (S2(anut, j + 1, k) - S2(anut, j, k)) -
                                                      ttm * (S2(anut, j, k) - S2(anut, j - 1, k)));
            S2(sjm, j, k) = S2(sjm, j, k) - coef * ttm;
            S2(sjp, j, k) = S2(sjp, j, k) - coef * ttp;
            S2(stc, j, k) = S2(stc, j, k) + coef * (ttm + ttp);
        }
    for (int k = 2; k <= km - 1; ++k)
        for (int j = 2; j <= jm - 1; ++j) {
            double uu = (RT(j, k, 1, 2) * Q3(q, j, k, 2) + RT(j, k, 1, 3) * Q3(q, j, k, 3)) * S2(dj, j, k);
            double sgnu = fsign(one, uu);
            double app = .5 * (1. + sgnu), apm = .5 * (1. - sgnu);
            S2(conv, j, k) = S2(conv, j, k) + uu * (app * (S2(anut, j, k) - S2(anut, j - 1, k)) +
                                                    apm * (S2(anut, j + 1, k) - S2(anut, j, k)));
            S2(sjm, j, k) = S2(sjm, j, k) - uu * app;
            S2(stc, j, k) = S2(stc, j, k) + uu * (app - apm);
            S2(sjp, j, k) = S2(sjp, j, k) + uu * apm;
        }
    for (int k = 2; k <= km - 1; ++k)
        for (int j = 2; j <= jm - 1; ++j) S2(sf, j, k) = -S2(conv, j, k) + S2(diff, j, k);
    for (int k = 2; k <= km - 1; ++k)
        for (int j = 2; j <= jm - 1; ++j) {
            const double fib = 1;
            double an = S2(anut, j, k);
            double sminlim = dmax(S2(c->smin, j, k), epss);
            double rkd2 = 1. / pw_sq(akarman * sminlim, 2);
            double ss = S2(c->vort, j, k);
            double chi = dmax(an, epss2);
            double fv1 = pw_sq(chi, 3) / (pw_sq(chi, 3) + pw_sq(cv1, 3));
            double fv2 = SLIP(SLIP_SA_FV2) ? 1 / pw_sq(1. + chi / cv2, 2) : 1 / pw_sq(1. + chi / cv2, 3);
            double fv3 = (1 + chi * fv1) / chi * (1 - fv2);
            double ft2 = ct3tmp * exp(-(ct4 * pw_sq(chi, 2)));
            double sst = fv3 * ss + vnu * chi * fv2 * rkd2;
            sst = dmax(sst, epss3);
            double rr = dmin(ten, vnu * an * rkd2 / sst);
            double gg = rr + cw2 * (pw_sq(rr, 6) - rr);
            double gg6 = pw_sq(gg, 6);
            double fg = (1. + cw3r6) / (gg6 + cw3r6);
            double fwdg = pow(fg, (double)(1.0f / 6.0f));
            double fw = gg * fwdg;
            double pp = cb1 * sst * (1. - ft2);
            double dd = vnu * (cw1 * fw - cb1 * ft2 / pw_sq(akarman, 2)) * an / pw_sq(sminlim, 2);
            double ft2p = -(2. * ct4 * ct3tmp * chi * exp(-(ct4 * pw_sq(chi, 2))));
            double fv1p = 3. * pw_sq(chi, 2) * pw_sq(cv1, 3) / pw_sq(pw_sq(chi, 3) + pw_sq(cv1, 3), 2);
            double fv2p = -(3 / cv2 / pw_sq(1. + chi / cv2, 4));
            double fv3p = ((fv1 + chi * fv1p) * (1 - fv2) - (1 + chi * fv1) * ((1 - fv2) / chi + fv2p)) / chi;
            double sstp = fv3p * ss + vnu * (chi * fv2p + fv2) * rkd2;
            double rrp = vnu * rkd2 * (sst - chi * sstp) / pw_sq(sst, 2);
            double ggp = rrp * (1. + cw2 * (6. * pw_sq(rr, 5) - 1.));
            double fwp = SLIP(SLIP_SA_LIN) ? 0.0 : ggp * fwdg * (1. - gg6 / (gg6 + cw3r6));
            double ppp = cb1 * (sstp - ft2p * sst - ft2 * sstp);
            double ddp = vnu * (cw1 * fw - cb1 * ft2 / pw_sq(akarman, 2) +
                                (cw1 * fwp - cb1 * ft2p / pw_sq(akarman, 2)) * chi) / pw_sq(sminlim, 2);
            S2(c->prod, j, k) = fib * pp * an * S2(c->fr1, j, k);
            S2(c->dest, j, k) = fib * dd * an;
            S2(sf, j, k) = S2(sf, j, k) + S2(c->prod, j, k) - S2(c->dest, j, k);
            const double zero = 0.;
            double addpp = dmax(dd - pp, zero);
            double addpppp = dmax(ddp - ppp, zero);
            S2(stc, j, k) = S2(stc, j, k) + fib * addpp;
            S2(stc, j, k) = S2(stc, j, k) + fib * addpppp * an;
        }
}

static void tsaibt(orc_case *c)
{
    /* ref/tursa.f:1327-1372 */
    const int jm = c->jm, km = c->km;
    double rterm = 1. / c->p.dt;
    for (int j = 1; j <= jm; ++j)
        for (int k = 1; k <= km; ++k) {
            double fib = (j == 1 || j == jm || k == 1 || k == km) ? 0 : 1;
            S2(c->sjm, j, k) = fib * S2(c->sjm, j, k);
            S2(c->sjp, j, k) = fib * S2(c->sjp, j, k);
            S2(c->skm, j, k) = fib * S2(c->skm, j, k);
            S2(c->skp, j, k) = fib * S2(c->skp, j, k);
            S2(c->stc, j, k) = fib * (S2(c->stc, j, k) + rterm) + (1. - fib);
            S2(c->sf, j, k) = fib * S2(c->sf, j, k);
        }
}

static void tsabc(orc_case *c)
{
    /* ref/tursa.f:1173-1259 */
    const int jm = c->jm, km = c->km;
    double *anut = c->anut;
    for (int nr = 0; nr < c->nreg; ++nr) {
        const smac_bc_region *r = &c->reg[nr];
        int iv = r->ibcval;
        if (!(iv == 30 || iv == 31 || iv == 32 || iv == 33 || iv == 40)) continue;
        if (c->jkbc[nr] == 1 && r->jbeg == 1) {
            int j = r->jbeg;
            for (int k = r->kbeg; k <= r->kend; ++k) {
                S2(c->stc, j, k) = 1.0;
                S2(c->sjp, j, k) = -1.0;
                S2(c->sf, j, k) = S2(anut, j + 1, k) - S2(anut, j, k);
            }
        } else if (c->jkbc[nr] == 1 && r->jbeg == jm) {
            int j = r->jbeg;
            for (int k = r->kbeg; k <= r->kend; ++k) {
                S2(c->stc, j, k) = 1.0;
                S2(c->sjm, j, k) = -1.0;
                S2(c->sf, j, k) = S2(anut, j - 1, k) - S2(anut, j, k);
            }
        } else if (c->jkbc[nr] == 2 && r->kbeg == 1) {
            int k = r->kbeg;
            for (int j = r->jbeg; j <= r->jend; ++j) {
                S2(c->stc, j, k) = 1.0;
                S2(c->skp, j, k) = -1.0;
                S2(c->sf, j, k) = S2(anut, j, k + 1) - S2(anut, j, k);
            }
        } else if (c->jkbc[nr] == 2 && r->kbeg == km) {
            int k = r->kbeg;
            for (int j = r->jbeg; j <= r->jend; ++j) {
                S2(c->stc, j, k) = 1.0;
                S2(c->skm, j, k) = -1.0;
                S2(c->sf, j, k) = S2(anut, j, k - 1) - S2(anut, j, k);
            }
        }
    }
}

static void scatri(int n, double *a, double *b, double *cc, double *f)
{
    /* ref/lreltur.f:337-359 with ibeg = 1, iend = n */
    f[0] = f[0] / b[0];
    cc[0] = cc[0] / b[0];
    for (int i = 2; i <= n - 1; ++i) {
        double z = 1. / (b[i - 1] - a[i - 1] * cc[i - 2]);
        cc[i - 1] = cc[i - 1] * z;
        f[i - 1] = (f[i - 1] - a[i - 1] * f[i - 2]) * z;
    }
    int i = n;
    f[i - 1] = (f[i - 1] - a[i - 1] * f[i - 2]) / (b[i - 1] - a[i - 1] * cc[i - 2]);
    for (i = n - 1; i >= 1; --i) f[i - 1] = f[i - 1] - cc[i - 1] * f[i];
}

static void turswpj_line(orc_case *c, int k, double *aaa, double *bbb, double *ccc, double *ff)
{
    const int jm = c->jm, km = c->km;
    double *rhs = c->sf, *ds = c->ds;
    for (int j = 1; j <= jm; ++j) {
        aaa[j - 1] = S2(c->sjm, j, k);
        bbb[j - 1] = S2(c->stc, j, k);
        ccc[j - 1] = S2(c->sjp, j, k);
        if (k == 1) ff[j - 1] = S2(rhs, j, k) - S2(c->skp, j, k) * S2(ds, j, k + 1);
        else if (k == km) ff[j - 1] = S2(rhs, j, k) - S2(c->skm, j, k) * S2(ds, j, k - 1);
        else ff[j - 1] = S2(rhs, j, k) - S2(c->skm, j, k) * S2(ds, j, k - 1) - S2(c->skp, j, k) * S2(ds, j, k + 1);
    }
    scatri(jm, aaa, bbb, ccc, ff);
    for (int j = 1; j <= jm; ++j) S2(ds, j, k) = ff[j - 1];
}

static void lreltur(orc_case *c)
{
    /* ref/lreltur.f:52-252: ntksp = 0 in the reference setup, so only j-line passes */
    const int jm = c->jm, km = c->km;
    for (long i = 0; i < c->N; ++i) c->ds[i] = 0.0;
    int nsmax = c->p.ntjsp > c->p.ntksp ? c->p.ntjsp : c->p.ntksp;
    double *tmp = malloc(sizeof(double) * 4 * jm);
    for (int njk = 1; njk <= nsmax; ++njk) {
        if (c->p.ntjsp >= njk) {
            for (int k = 1; k <= km; ++k) turswpj_line(c, k, tmp, tmp + jm, tmp + 2 * jm, tmp + 3 * jm);
            turswpj_line(c, 1, tmp, tmp + jm, tmp + 2 * jm, tmp + 3 * jm);
            bcimpds(jm, km, 1, c->npair, c->pairs, c->anut, c->ds, c->sf);
        }
    }
    free(tmp);
}

static void tursa_assemble(orc_case *c)
{
    /* ref/tursa.f:66-210 (zero fill, wall distance and initial value on the first call,
     * source assembly, boundary rows); tursa_solve does :240-292 */
    const int jm = c->jm, km = c->km;
    const long N = c->N;
    const double anutinf = F32(0.001);
    for (long i = 0; i < N; ++i) {
        c->sjm[i] = c->sjp[i] = c->skm[i] = c->skp[i] = 0.0;
        c->stc[i] = c->sf[i] = c->ds[i] = c->smin[i] = 0.0;
        c->vort[i] = c->diff[i] = c->conv[i] = c->prod[i] = c->dest[i] = c->fr1[i] = 0.0;
    }
    if (c->first_tursa) {
        c->first_tursa = 0;
        /* dfunsa: brute-force wall distance (ref/tursa.f:351-433) */
        for (long i = 0; i < N; ++i) c->smin[i] = 1.e6;
        for (int nw = 0; nw < c->nwall; ++nw) {
            if (c->wjk[nw] != 2) continue;
            int kw = c->wk2[nw];
            for (int jw = c->wj1[nw]; jw <= c->wj2[nw]; jw += c->winc[nw]) {
                double xw = S2(c->x, jw, kw), yw = S2(c->y, jw, kw);
                for (int k = 1; k <= km; ++k)
                    for (int j = 1; j <= jm; ++j) {
                        double dxx = S2(c->x, j, k) - xw, dyy = S2(c->y, j, k) - yw;
                        S2(c->smin, j, k) = dmin(S2(c->smin, j, k), sqrt(dxx * dxx + dyy * dyy));
                    }
            }
        }
        /* tsaic (ref/tursa.f:1263-1323) */
        for (int k = 1; k <= km; ++k)
            for (int j = 1; j <= jm; ++j) S2(c->anut, j, k) = anutinf + 10.0 * exp(-10. * S2(c->smin, j, k));
        for (int nw = 0; nw < c->nwall; ++nw)
            for (int k = c->wk1[nw]; k <= c->wk2[nw]; ++k)
                for (int j = c->wj1[nw]; j <= c->wj2[nw]; ++j) S2(c->anut, j, k) = 0.0;
    }
    tsafill(c);
    tsaibt(c);
    tsabc(c);
}

static void tursa_solve(orc_case *c)
{
    const int jm = c->jm, km = c->km;
    const long N = c->N;
    if (c->p.impsch == 1) {
        lreltur(c);
    } else {
        for (long i = 0; i < N; ++i) c->sabprm[i] = 1. / c->stc[i];
        ilusol1(jm, km, c->ds, c->sf, c->sabprm, c->sjm, c->sjp, c->skm, c->skp);
        bcimpds(jm, km, 1, c->npair, c->pairs, c->anut, c->ds, c->sf);
    }
    /* tsaupd (ref/tursa.f:314-347) */
    double dsmax = 0.0;
    for (int k = 1; k <= km; ++k)
        for (int j = 1; j <= jm; ++j) {
            S2(c->anut, j, k) = dmax(0.0, S2(c->anut, j, k) + S2(c->ds, j, k));
            dsmax = dmax(S2(c->ds, j, k), dsmax);
        }
    c->turres = dsmax;
    /* eddy viscosity (ref/tursa.f:287-292) */
    const double cv1 = F32(7.1);
    for (int k = 1; k <= km; ++k)
        for (int j = 1; j <= jm; ++j) {
            double chi = S2(c->anut, j, k);
            double fv1 = pw_sq(chi, 3) / (pw_sq(chi, 3) + pw_sq(cv1, 3));
            S2(c->vnut, j, k) = fv1 * S2(c->anut, j, k) * c->vnu;
        }
}

/* ------------------------------------------------------------------------------------ */
/* step: ref/step.f:53-609                                                                */

int orc_assemble(orc_case *c, smac_iter_stats *st)
{
    const int jm = c->jm, km = c->km;
    const long N = c->N;
    double *btc = c->btc, *bjm = c->bjm, *bjp = c->bjp, *bkm = c->bkm, *bkp = c->bkp, *s = c->s;
    for (long i = 0; i < 9 * N; ++i) btc[i] = bjp[i] = bjm[i] = bkm[i] = bkp[i] = c->scr1[i] = 0.0;
    if (c->p.ivis == 1 && c->p.iturb == 1) {
        tursa_assemble(c);
        tursa_solve(c);
    }
    conup(c);
    if (c->p.ivis == 1) viscterms(c);
    /* blank boundary rows (ref/step.f:386-413) */
    for (int m = 1; m <= 3; ++m)
        for (int k = 1; k <= km; ++k)
            for (int j = 1; j <= jm; ++j) {
                double fib = (j == 1 || j == jm || k == 1 || k == km) ? 0 : 1;
                for (int n = 1; n <= 3; ++n) {
                    B9(btc, j, k, m, n) = fib * B9(btc, j, k, m, n);
                    B9(bjp, j, k, m, n) = fib * B9(bjp, j, k, m, n);
                    B9(bjm, j, k, m, n) = fib * B9(bjm, j, k, m, n);
                    B9(bkp, j, k, m, n) = fib * B9(bkp, j, k, m, n);
                    B9(bkm, j, k, m, n) = fib * B9(bkm, j, k, m, n);
                }
                Q3(s, j, k, m) = fib * Q3(s, j, k, m);
            }
    /* diagonal (ref/step.f:419-437) */
    for (int k = 1; k <= km; ++k)
        for (int j = 1; j <= jm; ++j) {
            double fib = (j == 1 || j == jm || k == 1 || k == km) ? 0 : 1;
            double rdjdtau = fib / (S2(c->dj, j, k) * c->p.dtau) + 1. - fib;
            B9(btc, j, k, 1, 1) = B9(btc, j, k, 1, 1) + rdjdtau;
            B9(btc, j, k, 2, 2) = B9(btc, j, k, 2, 2) + rdjdtau;
            B9(btc, j, k, 3, 3) = B9(btc, j, k, 3, 3) + rdjdtau;
        }
    int rc = bcmain(c);
    if (rc) return rc;
    /* resid (ref/step.f:613-756) */
    double *r = c->scr1;
    for (int k = 1; k <= km; ++k)
        for (int j = 1; j <= jm; ++j)
            for (int n = 1; n <= 3; ++n) Q3(r, j, k, n) = Q3(s, j, k, n) * S2(c->dj, j, k);
    long irmax = isamax(3 * N, r);
    long jk = (irmax - 1) / N + 1;
    irmax = irmax - (jk - 1) * N;
    c->kres = (int)((irmax - 1) / jm + 1);
    c->jres = (int)(irmax - (long)(c->kres - 1) * jm);
    c->resmax = fabs(Q3(r, c->jres, c->kres, jk));
    if (c->nt == 1) {
        c->resmax0 = c->resmax;
        if (c->resmax0 < 1.) c->resmax0 = 1.0;
    }
    long idmax = isamax(N, r);
    int kd = (int)((idmax - 1) / jm + 1);
    int jd = (int)(idmax - (long)(kd - 1) * jm);
    c->divmax = fabs(Q3(r, jd, kd, 1) / c->p.beta);
    /* rhsnorm (ref/step.f:513-519) */
    double acc = 0.0;
    for (long i = 0; i < 3 * N; ++i) acc = acc + s[i] * s[i];
    c->rhsnorm = sqrt(acc);
    for (int b = 0; b < 5; ++b)
        memcpy(c->lhs[b], (double *[]){btc, bjm, bjp, bkm, bkp}[b], sizeof(double) * 9 * N);
    if (st) {
        st->nt = c->nt;
        st->resmax = c->resmax;
        st->resmax0 = c->resmax0;
        st->jres = c->jres;
        st->kres = c->kres;
        st->divmax = c->divmax;
        st->turres = c->turres;
        st->rhsnorm = c->rhsnorm;
    }
    return SMAC_OK;
}

int orc_solve(orc_case *c)
{
    if (c->p.impsch == 1)
        return orc_lrelax(&c->p, c->jm, c->km, c->npair, c->pairs, c->q, c->dq, c->s, c->btc,
                          c->bjm, c->bjp, c->bkm, c->bkp);
    if (c->p.impsch == 2)
        return orc_lusgs(&c->p, c->jm, c->km, 3, c->npair, c->pairs, c->q, c->dq, c->s, c->btc,
                         c->bjm, c->bjp, c->bkm, c->bkp, c->scr1);
    return SMAC_ERR_IMPSCH;
}

/* wall forces: force/fmk (ref/geom.f:135-213, ref/fm.f:156-324) and the coefficient
 * formulas of iterout (ref/smac2d.f:1709-1739) */
int orc_finish(orc_case *c, smac_iter_stats *st)
{
    const int jm = c->jm;
    const long N = c->N;
    double *q = c->q, *rtxy = c->rtxy, *x = c->x, *y = c->y, *dj = c->dj;
    for (long i = 0; i < 3 * N; ++i) c->q[i] = c->q[i] + c->dq[i];
    double fxp = 0.0, fyp = 0.0, fxs = 0.0, fys = 0.0, fmom = 0.0;
    const double x0 = c->p.xmom, y0 = c->p.ymom, vnu = c->vnu;
    for (int nw = 0; nw < c->nwall; ++nw) {
        if (c->wjk[nw] != 2) return SMAC_ERR_BC; /* fmj walls not needed by C-grids */
        int k = c->wk1[nw], jbeg = c->wj1[nw], jend = c->wj2[nw];
        int jstart = jbeg; /* ioverlap offset only applies to walls after the first */
        fxp = 0.; fyp = 0.; fxs = 0.; fys = 0.;
        for (int j = jstart; j <= jend - 1; ++j) {
            double pave = 0.5 * (Q3(q, j + 1, k, 1) + Q3(q, j, k, 1));
            double da1 = SLIP(SLIP_FORCE_NORMAL) ? 0.5 * (RT(j + 1, k, 2, 3) + RT(j, k, 2, 3))
                                                 : 0.5 * (RT(j + 1, k, 2, 2) + RT(j, k, 2, 2));
            double da2 = 0.5 * (RT(j + 1, k, 2, 3) + RT(j, k, 2, 3));
            double xave = 0.5 * (S2(x, j + 1, k) + S2(x, j, k)) - x0;
            double yave = 0.5 * (S2(y, j + 1, k) + S2(y, j, k)) - y0;
            double fx = pave * da1, fy = pave * da2;
            fxp = fxp - fx;
            fyp = fyp - fy;
            fmom = fmom - (xave * fy - yave * fx);
        }
        if (c->p.ivis != 0) {
            int n = jend - jstart + 1;
            double *t = malloc(sizeof(double) * 3 * n);
            for (int j = jstart; j <= jend; ++j) {
                double ueta = 0.5 * (-3. * Q3(q, j, k, 2) + 4. * Q3(q, j, k + 1, 2) - Q3(q, j, k + 2, 2));
                double veta = 0.5 * (-3. * Q3(q, j, k, 3) + 4. * Q3(q, j, k + 1, 3) - Q3(q, j, k + 2, 3));
                t[j - jstart] = 2. * vnu * RT(j, k, 2, 2) * S2(dj, j, k) * ueta;
                t[n + j - jstart] = 2. * vnu * RT(j, k, 2, 3) * S2(dj, j, k) * veta;
                t[2 * n + j - jstart] = vnu * (RT(j, k, 2, 3) * ueta + RT(j, k, 2, 2) * veta) * S2(dj, j, k);
            }
            for (int j = jstart; j <= jend - 1; ++j) {
                int i = j - jstart;
                double t11ave = 0.5 * (t[i + 1] + t[i]);
                double t22ave = 0.5 * (t[n + i + 1] + t[n + i]);
                double t12ave = 0.5 * (t[2 * n + i + 1] + t[2 * n + i]);
                double t21ave = t12ave;
                double da1 = 0.5 * (RT(j + 1, k, 2, 2) + RT(j, k, 2, 2));
                double da2 = 0.5 * (RT(j + 1, k, 2, 3) + RT(j, k, 2, 3));
                double xave = 0.5 * (S2(x, j + 1, k) + S2(x, j, k)) - x0;
                double yave = 0.5 * (S2(y, j + 1, k) + S2(y, j, k)) - y0;
                double fx = t11ave * da1 + t12ave * da2;
                double fy = t21ave * da1 + t22ave * da2;
                fxs = fxs + fx;
                fys = fys + fy;
                fmom = fmom + xave * fy - yave * fx;
            }
            free(t);
        }
    }
    double pi = 4.0 * atan(1.0);
    double ca = cos(c->p.alpha * pi / 180.0), sa = sin(c->p.alpha * pi / 180.0);
    double chord = c->p.chord;
    double cd = (2. * (fxp + fxs) * ca + 2. * (fyp + fys) * sa) / chord;
    double cl = (2. * (fyp + fys) * ca - 2. * (fxp + fxs) * sa) / chord;
    double cm = -(2. * fmom / (chord * chord));
    double cdp = (2. * fxp * ca + 2. * fyp * sa) / chord;
    double clp = (2. * fyp * ca - 2. * fxp * sa) / chord;
    c->clift = cl;
    if (st) {
        st->cl = cl; st->cd = cd; st->cm = cm; st->cdp = cdp; st->clp = clp;
        st->fxp = fxp; st->fyp = fyp; st->fxs = fxs; st->fys = fys; st->fmom = fmom;
        int bad = 0;
        for (long i = 0; i < 3 * N; ++i) if (!isfinite(c->q[i])) { bad = 1; break; }
        st->nonfinite = bad || !isfinite(c->rhsnorm);
    }
    return SMAC_OK;
}

int orc_step(orc_case *c, smac_iter_stats *st)
{
    c->nt = c->nt + 1;
    int rc = orc_assemble(c, st);
    if (rc) return rc;
    rc = orc_solve(c);
    if (rc) return rc;
    return orc_finish(c, st);
}

/* ------------------------------------------------------------------------------------ */
/* set-up                                                                                 */

orc_case *orc_create(const smac_params *p, int jm, int km, const double *x, const double *y,
                     const double *rtxy, const double *dj, int nreg, const smac_bc_region *regs)
{
    if (jm < 5 || km < 5 || nreg < 1 || nreg > MAXREG) return NULL;
    orc_case *c = calloc(1, sizeof *c);
    c->p = *p;
    c->vnu = 1.0 / p->reynum;
    c->jm = jm;
    c->km = km;
    long N = c->N = (long)jm * km;
    c->L = (jm > km ? jm : km) + 2;
    /* x y | rtxy | dj | q dq s | anut vnut | 5 blocks + scr1 | 5 lhs copies | 15 SA planes */
    double *big = calloc((size_t)(N * (2 + 6 + 1 + 9 + 2 + 54 + 45 + 15) + 4 * 3 * c->L),
                         sizeof(double));
    double *w = big;
#define TAKE(ptr, n) (ptr = w, w += (n))
    TAKE(c->x, N); TAKE(c->y, N); TAKE(c->rtxy, 6 * N); TAKE(c->dj, N);
    TAKE(c->q, 3 * N); TAKE(c->dq, 3 * N); TAKE(c->s, 3 * N); TAKE(c->anut, N); TAKE(c->vnut, N);
    TAKE(c->btc, 9 * N); TAKE(c->bjm, 9 * N); TAKE(c->bjp, 9 * N); TAKE(c->bkm, 9 * N);
    TAKE(c->bkp, 9 * N); TAKE(c->scr1, 9 * N);
    for (int b = 0; b < 5; ++b) TAKE(c->lhs[b], 9 * N);
    TAKE(c->vort, N); TAKE(c->diff, N); TAKE(c->conv, N); TAKE(c->prod, N); TAKE(c->dest, N);
    TAKE(c->fr1, N); TAKE(c->sjm, N); TAKE(c->sjp, N); TAKE(c->skm, N); TAKE(c->skp, N);
    TAKE(c->stc, N); TAKE(c->ds, N); TAKE(c->sf, N); TAKE(c->smin, N); TAKE(c->sabprm, N);
    TAKE(c->dfp, 3 * c->L); TAKE(c->dfm, 3 * c->L); TAKE(c->f, 3 * c->L); TAKE(c->fv, 3 * c->L);
#undef TAKE
    memcpy(c->x, x, sizeof(double) * N);
    memcpy(c->y, y, sizeof(double) * N);
    if (rtxy && dj) {
        memcpy(c->rtxy, rtxy, sizeof(double) * 6 * N);
        memcpy(c->dj, dj, sizeof(double) * N);
    } else {
        orc_metric(jm, km, x, y, c->rtxy, c->dj);
    }
    c->nreg = nreg;
    int w0 = -1, w1 = -1;
    for (int i = 0; i < nreg; ++i) {
        c->reg[i] = regs[i];
        c->jkbc[i] = regs[i].jbeg == regs[i].jend ? 1 : (regs[i].kbeg == regs[i].kend ? 2 : 0);
        if (!c->jkbc[i]) { free(big); free(c); return NULL; }
        if (regs[i].ibcval == 0 || regs[i].ibcval == 1 || regs[i].ibcval == 4 || regs[i].ibcval == 5) {
            int nw = c->nwall++;
            c->wjk[nw] = c->jkbc[i];
            c->wj1[nw] = regs[i].jbeg; c->wj2[nw] = regs[i].jend;
            c->wk1[nw] = regs[i].kbeg; c->wk2[nw] = regs[i].kend;
            c->winc[nw] = (regs[i].ibcval == 0 || regs[i].ibcval == 4) ? 1 : -1;
        }
        if (regs[i].ibcval == 60) {
            if (w0 < 0) w0 = i; else if (w1 < 0) w1 = i;
        }
    }
    /* wake pairs as ref/rbc.f:1160-1270 forms them for a pair of equal cut segments on
     * k = 1: segment A's points pair with B's in reverse order, then B's with A's */
    if (w0 >= 0 && w1 >= 0) {
        const smac_bc_region *a = &c->reg[w0], *b = &c->reg[w1];
        int na = a->jend - a->jbeg + 1, nb = b->jend - b->jbeg + 1;
        int n = na < nb ? na : nb;
        c->npair = 2 * n;
        c->pairs = malloc(sizeof(int) * 4 * n);
        int ip = 0;
        for (int i = 0; i < n; ++i) { c->pairs[2 * ip] = a->jbeg + i; c->pairs[2 * ip + 1] = b->jend - i; ++ip; }
        for (int i = 0; i < n; ++i) { c->pairs[2 * ip] = b->jbeg + i; c->pairs[2 * ip + 1] = a->jend - i; ++ip; }
    }
    c->first_tursa = 1;
    c->nt = 0;
    c->clift = 0.0;
    orc_ic(c);
    return c;
}

void orc_ic(orc_case *c)
{
    /* ref/geom.f:88-131 */
    const int jm = c->jm, km = c->km;
    const long N = c->N;
    double pi = 4.0 * atan(1.0);
    double ca = cos(c->p.alpha * pi / 180.), sa = sin(c->p.alpha * pi / 180.);
    for (int j = 1; j <= jm; ++j)
        for (int k = 1; k <= km; ++k) {
            Q3(c->q, j, k, 1) = 1.0;
            Q3(c->q, j, k, 2) = ca;
            Q3(c->q, j, k, 3) = sa;
        }
    for (int nw = 0; nw < c->nwall; ++nw)
        for (int j = c->wj1[nw]; j <= c->wj2[nw]; ++j)
            for (int k = c->wk1[nw]; k <= c->wk2[nw]; ++k) {
                Q3(c->q, j, k, 2) = 0.0;
                Q3(c->q, j, k, 3) = 0.0;
            }
}

void orc_destroy(orc_case *c)
{
    if (!c) return;
    free(c->x);
    free(c->pairs);
    free(c);
}

void orc_set_counters(orc_case *c, int nt, double resmax0, double clift)
{
    c->nt = nt;
    c->resmax0 = resmax0;
    c->clift = clift;
    if (nt > 0) c->first_tursa = 0;
}

double *orc_array(orc_case *c, const char *name)
{
    struct { const char *n; double *p; } t[] = {
        {"q", c->q}, {"dq", c->dq}, {"s", c->s}, {"anut", c->anut}, {"vnut", c->vnut},
        {"btc", c->btc}, {"bjm", c->bjm}, {"bjp", c->bjp}, {"bkm", c->bkm}, {"bkp", c->bkp},
        {"rtxy", c->rtxy}, {"dj", c->dj}, {"x", c->x}, {"y", c->y}, {"smin", c->smin},
        {"lhs_btc", c->lhs[0]}, {"lhs_bjm", c->lhs[1]}, {"lhs_bjp", c->lhs[2]},
        {"lhs_bkm", c->lhs[3]}, {"lhs_bkp", c->lhs[4]}, {"sf", c->sf}, {"stc", c->stc},
        {"ds", c->ds}, {"scr1", c->scr1}, {"sjm", c->sjm}, {"sjp", c->sjp},
        {"skm", c->skm}, {"skp", c->skp}, {"vort", c->vort}, {"fr1", c->fr1},
    };
    for (size_t i = 0; i < sizeof t / sizeof t[0]; ++i)
        if (!strcmp(t[i].n, name)) return t[i].p;
    return NULL;
}

int orc_npairs(orc_case *c) { return c->npair; }
const int *orc_pairs(orc_case *c) { return c->pairs; }

/* SA assembly of one step on its own (ref/tursa.f:66-210), leaving sf/stc/sjm.. for
 * inspection; used by the kernel-level checks only. */
int orc_sa_assemble(orc_case *c)
{
    tursa_assemble(c);
    return SMAC_OK;
}
