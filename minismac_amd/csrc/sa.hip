// Spalart-Allmaras one-equation model of the implicit iteration: tursa and what it calls
// (ref/tursa.f:66-1372: dfunsa, tsaic, tsafill with vortsa and curvature, tsaibt, tsabc,
// tsaupd, eddy viscosity) and its scalar line relaxation lreltur/turswpj/scatri
// (ref/lreltur.f:52-359).
//
// Numerics the reference build fixes and these kernels reproduce: the model constants are
// default-REAL literals widened to double; tursa.f is built at -O0, where x**n for n >= 4 is
// binary powering; ref/tursa.f:460 runs past column 72, so the floor on S~ (epss3) is 1, not
// 1e-20.  exp/pow/atan come from the device math library, so SA values agree with the CPU to
// within a few ulp rather than bit for bit; the mean flow does not read them (the viscous
// terms use the laminar viscosity, ref/visc.f:97).
#include "smac_dev.h"
#include "smac_kernels.h"

namespace smac {

SMAC_DEV double pw(double x, int n)
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

#define F32(v) ((double)(v##f))

// dfunsa + tsaic (first call only): brute-force wall distance and the initial nu~/nu
__global__ void k_sa_first(Dims d, Walls wl, const double *__restrict__ x,
                           const double *__restrict__ y, double *__restrict__ smin,
                           double *__restrict__ anut)
{
    const long N = d.N;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N;
         p += (long)gridDim.x * blockDim.x) {
        double sm = 1.e6;
        const double xp = x[p], yp = y[p];
        for (int w = 0; w < wl.n; ++w)
            for (int jw = wl.j1[w]; jw <= wl.j2[w]; ++jw) {
                long pw_ = (long)jw + (long)d.jm * wl.k[w];
                double dxx = xp - x[pw_], dyy = yp - y[pw_];
                sm = dmin_(sm, sqrt(dxx * dxx + dyy * dyy));
            }
        smin[p] = sm;
        double a = F32(0.001) + 10.0 * exp(-10. * sm);
        const int j = (int)(p % d.jm), k = (int)(p / d.jm);
        for (int w = 0; w < wl.n; ++w)
            if (k == wl.k[w] && j >= wl.j1[w] && j <= wl.j2[w]) a = 0.0;
        anut[p] = a;
    }
}

// first derivatives of u, v (vortsa / curvature loops 10-80), the vorticity, and the
// point-local part of the curvature correction (curvature loop 100)
SMAC_DEV void b_sa_deriv(const Blk &bk, const Dims &d, const double *__restrict__ q,
                         const double *__restrict__ met, const double *__restrict__ dj,
                         double *__restrict__ vort, double *__restrict__ cu,
                         double *__restrict__ ck, double *__restrict__ cv, double *__restrict__ fr)
{
    const int j = bk.x * blockDim.x + threadIdx.x;
    const int k = bk.y;
    if (j >= d.jm) return;
    const long N = d.N, p = P(d, j, k);
    const double *u = q + N, *v = q + 2 * N;
    double uj, vj, uk, vk;
    if (j == 0) {
        uj = 0.5 * (-u[p + 2] + 4. * u[p + 1] - 3. * u[p]);
        vj = 0.5 * (-v[p + 2] + 4. * v[p + 1] - 3. * v[p]);
    } else if (j == d.jm - 1) {
        uj = 0.5 * (3. * u[p] - 4. * u[p - 1] + u[p - 2]);
        vj = 0.5 * (3. * v[p] - 4. * v[p - 1] + v[p - 2]);
    } else {
        uj = 0.5 * (u[p + 1] - u[p - 1]);
        vj = 0.5 * (v[p + 1] - v[p - 1]);
    }
    const long s = d.jm;
    if (k == 0) {
        uk = 0.5 * (-u[p + 2 * s] + 4. * u[p + s] - 3. * u[p]);
        vk = 0.5 * (-v[p + 2 * s] + 4. * v[p + s] - 3. * v[p]);
    } else if (k == d.km - 1) {
        uk = 0.5 * (3. * u[p] - 4. * u[p - s] + u[p - 2 * s]);
        vk = 0.5 * (3. * v[p] - 4. * v[p - s] + v[p - 2 * s]);
    } else {
        uk = 0.5 * (u[p + s] - u[p - s]);
        vk = 0.5 * (v[p + s] - v[p - s]);
    }
    const double r12 = met[p], r13 = met[N + p], r22 = met[2 * N + p], r23 = met[3 * N + p];
    const double dd = dj[p];
    vort[p] = fabs(dd * (r12 * vj + r22 * vk - r13 * uj - r23 * uk));
    double ux = dd * (r12 * uj + r22 * uk);
    double vx = dd * (r12 * vj + r22 * vk);
    double uy = dd * (r13 * uj + r23 * uk);
    double vy = dd * (r13 * vj + r23 * vk);
    double m = sqrt(2. * (ux * ux + vx * uy + vy * vy) + uy * uy + vx * vx);
    if (vx - uy != 0.) m = m / fabs(vx - uy);
    cv[p] = m;
    cu[p] = (ux - vy) / 2;
    ck[p] = (uy + vx) / 2;
    double f = 2 * (vx - uy);
    double denom = (2. * (ux * ux + vx * uy + vy * vy) + uy * uy + vx * vx) *
                   (ux * ux + vx * vx + uy * uy + vy * vy);
    if (denom != 0.) f = f / denom;
    fr[p] = f;
}

__global__ void k_sa_deriv(Dims d, const double *__restrict__ q, const double *__restrict__ met,
                           const double *__restrict__ dj, double *__restrict__ vort,
                           double *__restrict__ cu, double *__restrict__ ck,
                           double *__restrict__ cv, double *__restrict__ fr)
{
    b_sa_deriv(xcd_blk(), d, q, met, dj, vort, cu, ck, cv, fr);
}

// tsafill loops 180-260 + tsaibt at one point (ref/tursa.f:597-917, 1353-1367)
SMAC_DEV void b_sa_assemble(const Blk &bk, const Dims &d, const DevParams &prm,
                            const double *__restrict__ q, const double *__restrict__ met,
                            const double *__restrict__ dj, const double *__restrict__ anut,
                            const double *__restrict__ smin, int use_smin,
                            const double *__restrict__ vort, const double *__restrict__ cu,
                            const double *__restrict__ ck, const double *__restrict__ cv,
                            const double *__restrict__ fr, double *__restrict__ sjm,
                            double *__restrict__ sjp, double *__restrict__ skm,
                            double *__restrict__ skp, double *__restrict__ stc,
                            double *__restrict__ sf)
{
    const int j = bk.x * blockDim.x + threadIdx.x;
    const int k = bk.y;
    if (j >= d.jm) return;
    const long N = d.N, p = P(d, j, k), s = d.jm;
    if (j == 0 || j == d.jm - 1 || k == 0 || k == d.km - 1) {
        sjm[p] = sjp[p] = skm[p] = skp[p] = 0.0;
        stc[p] = 1.0;
        sf[p] = 0.0;
        return;
    }
    const double vnu = prm.vnu;
    const double akarman = F32(0.41), cb1 = F32(0.1355), cb2 = F32(0.622);
    const double two = 2.0, three = 3.0, sigma = two / three;
    const double cw1 = cb1 / pw(akarman, 2) + (1. + cb2) / sigma;
    const double cw2 = F32(0.3), cw3 = 2.0, cv1 = F32(7.1), cv2 = 5.0, ct4 = F32(0.5);
    const double cr1 = 1, cr2 = 12, cr3 = 1;
    const double cw3r6 = pw(cw3, 6);
    const double ten = 10.0, epss = F32(1.e-12), epss2 = F32(0.00001), epss3 = 1.0;
    const double ct3tmp = 0.0;
    const double *r12 = met, *r13 = met + N, *r22 = met + 2 * N, *r23 = met + 3 * N;
    const double an = anut[p];
    const double uq = q[N + p], vq = q[2 * N + p];
    // curvature loop 101/102 at an interior point
    double fr1;
    {
        double uu = (r22[p] * uq + r23[p] * vq) * dj[p];
        double vv = (r12[p] * uq + r13[p] * vq) * dj[p];
        double dk = 2.0;
        double ds11dt = uu * (cu[p + s] - cu[p - s]) / dk + vv * (cu[p + 1] - cu[p - 1]) / 2;
        double ds12dt = uu * (ck[p + s] - ck[p - s]) / dk + vv * (ck[p + 1] - ck[p - 1]) / 2;
        double rtilde = (cu[p] * ds12dt - ck[p] * ds11dt) * fr[p];
        fr1 = (1 + cr1) * 2 * cv[p] / (1 + cv[p]) * (1 - cr3 * atan(cr2 * rtilde)) - cr1;
    }
    double vskm, vskp, vsjm, vsjp, vstc, diff, conv;
    // eta diffusion (loop 180)
    {
        const long pp = p + s, pm = p - s;
        double xy3p = .5 * (r22[p] * dj[p] + r22[pp] * dj[pp]);
        double xy4p = .5 * (r23[p] * dj[p] + r23[pp] * dj[pp]);
        double ttp = (xy3p * r22[p] * dj[p] + xy4p * r23[p] * dj[p]);
        double xy3m = .5 * (r22[p] * dj[p] + r22[pm] * dj[pm]);
        double xy4m = .5 * (r23[p] * dj[p] + r23[pm] * dj[pm]);
        double ttm = (xy3m * r22[p] * dj[p] + xy4m * r23[p] * dj[p]);
        double coef = -(cb2 * vnu * (1. + an) / sigma);
        diff = coef * (ttp * (anut[pp] - an) - ttm * (an - anut[pm]));
        vskm = -(coef * ttm);
        vskp = -(coef * ttp);
        vstc = coef * (ttm + ttp);
        coef = vnu * (1. + cb2) / sigma;
        double vnup = 1.0 + 0.5 * (anut[pp] + an);
        double vnum = 1.0 + 0.5 * (anut[pm] + an);
        ttp = ttp * vnup;
        ttm = ttm * vnum;
        diff = diff + coef * (ttp * (anut[pp] - an) - ttm * (an - anut[pm]));
        vskm = vskm - coef * ttm;
        vskp = vskp - coef * ttp;
        vstc = vstc + coef * (ttm + ttp);
        // eta advection (loop 190)
        double uu = (r22[p] * uq + r23[p] * vq) * dj[p];
        double sgnu = copysign(1.0, uu);
        double app = .5 * (1. + sgnu), apm = .5 * (1. - sgnu);
        conv = uu * (app * (an - anut[pm]) + apm * (anut[pp] - an));
        vskm = vskm - uu * app;
        vstc = vstc + uu * (app - apm);
        vskp = vskp + uu * apm;
    }
    // xi diffusion (loop 220)
    {
        const long pp = p + 1, pm = p - 1;
        double xy1p = .5 * (r12[p] * dj[p] + r12[pp] * dj[pp]);
        double xy2p = .5 * (r13[p] * dj[p] + r13[pp] * dj[pp]);
        double ttp = (xy1p * r12[p] * dj[p] + xy2p * r13[p] * dj[p]);
        double xy1m = .5 * (r12[p] * dj[p] + r12[pm] * dj[pm]);
        double xy2m = .5 * (r13[p] * dj[p] + r13[pm] * dj[pm]);
        double ttm = (xy1m * r12[p] * dj[p] + xy2m * r13[p] * dj[p]);
        double coef = -(cb2 * vnu * (1. + an) / sigma);
        diff = diff + coef * (ttp * (anut[pp] - an) - ttm * (an - anut[pm]));
        vsjm = -(coef * ttm);
        vsjp = -(coef * ttp);
        vstc = coef * (ttm + ttp) + vstc;
        coef = vnu * (1. + cb2) / sigma;
        double vnup = 1.0 + 0.5 * (anut[pp] + an);
        double vnum = 1.0 + 0.5 * (anut[pm] + an);
        ttp = ttp * vnup;
        ttm = ttm * vnum;
        diff = diff + coef * (ttp * (anut[pp] - an) - ttm * (an - anut[pm]));
        vsjm = vsjm - coef * ttm;
        vsjp = vsjp - coef * ttp;
        vstc = vstc + coef * (ttm + ttp);
        // xi advection (loop 230)
        double uu = (r12[p] * uq + r13[p] * vq) * dj[p];
        double sgnu = copysign(1.0, uu);
        double app = .5 * (1. + sgnu), apm = .5 * (1. - sgnu);
        conv = conv + uu * (app * (an - anut[pm]) + apm * (anut[pp] - an));
        vsjm = vsjm - uu * app;
        vstc = vstc + uu * (app - apm);
        vsjp = vsjp + uu * apm;
    }
    double vsf = -conv + diff;   // loop 235
    // production / destruction with its positive-part linearisation (loop 260)
    {
        const double fib = 1;
        double sminlim = dmax_(use_smin ? smin[p] : 0.0, epss);
        double rkd2 = 1. / pw(akarman * sminlim, 2);
        double ss = vort[p];
        double chi = dmax_(an, epss2);
        double fv1 = pw(chi, 3) / (pw(chi, 3) + pw(cv1, 3));
        double fv2 = 1 / pw(1. + chi / cv2, 3);
        double fv3 = (1 + chi * fv1) / chi * (1 - fv2);
        double ft2 = ct3tmp * exp(-(ct4 * pw(chi, 2)));
        double sst = fv3 * ss + vnu * chi * fv2 * rkd2;
        sst = dmax_(sst, epss3);
        double rr = dmin_(ten, vnu * an * rkd2 / sst);
        double gg = rr + cw2 * (pw(rr, 6) - rr);
        double gg6 = pw(gg, 6);
        double fg = (1. + cw3r6) / (gg6 + cw3r6);
        double fwdg = pow(fg, (double)(1.0f / 6.0f));
        double fw = gg * fwdg;
        double pp = cb1 * sst * (1. - ft2);
        double ddd = vnu * (cw1 * fw - cb1 * ft2 / pw(akarman, 2)) * an / pw(sminlim, 2);
        double ft2p = -(2. * ct4 * ct3tmp * chi * exp(-(ct4 * pw(chi, 2))));
        double fv1p = 3. * pw(chi, 2) * pw(cv1, 3) / pw(pw(chi, 3) + pw(cv1, 3), 2);
        double fv2p = -(3 / cv2 / pw(1. + chi / cv2, 4));
        double fv3p = ((fv1 + chi * fv1p) * (1 - fv2) - (1 + chi * fv1) * ((1 - fv2) / chi + fv2p)) / chi;
        double sstp = fv3p * ss + vnu * (chi * fv2p + fv2) * rkd2;
        double rrp = vnu * rkd2 * (sst - chi * sstp) / pw(sst, 2);
        double ggp = rrp * (1. + cw2 * (6. * pw(rr, 5) - 1.));
        double fwp = ggp * fwdg * (1. - gg6 / (gg6 + cw3r6));
        double ppp = cb1 * (sstp - ft2p * sst - ft2 * sstp);
        double ddp = vnu * (cw1 * fw - cb1 * ft2 / pw(akarman, 2) +
                            (cw1 * fwp - cb1 * ft2p / pw(akarman, 2)) * chi) / pw(sminlim, 2);
        double prod = fib * pp * an * fr1;
        double dest = fib * ddd * an;
        vsf = vsf + prod - dest;
        const double zero = 0.;
        double addpp = dmax_(ddd - pp, zero);
        double addpppp = dmax_(ddp - ppp, zero);
        vstc = vstc + fib * addpp;
        vstc = vstc + fib * addpppp * an;
    }
    // tsaibt at an interior point: + 1/dt on the diagonal
    sjm[p] = vsjm;
    sjp[p] = vsjp;
    skm[p] = vskm;
    skp[p] = vskp;
    stc[p] = vstc + 1. / prm.dt;
    sf[p] = vsf;
}

__global__ void k_sa_assemble(Dims d, DevParams prm, const double *__restrict__ q,
                              const double *__restrict__ met, const double *__restrict__ dj,
                              const double *__restrict__ anut, const double *__restrict__ smin,
                              int use_smin, const double *__restrict__ vort,
                              const double *__restrict__ cu, const double *__restrict__ ck,
                              const double *__restrict__ cv, const double *__restrict__ fr,
                              double *__restrict__ sjm, double *__restrict__ sjp,
                              double *__restrict__ skm, double *__restrict__ skp,
                              double *__restrict__ stc, double *__restrict__ sf)
{
    b_sa_assemble(xcd_blk(), d, prm, q, met, dj, anut, smin, use_smin, vort, cu, ck, cv, fr, sjm,
                  sjp, skm, skp, stc, sf);
}

// tsabc: zero-gradient rows of the outflow regions (ref/tursa.f:1197-1254)
SMAC_DEV void b_sa_bc(const Blk &bk, const Dims &d, const BcRegion &r,
                      const double *__restrict__ anut, double *sjm, double *sjp, double *skm,
                      double *skp, double *stc, double *sf)
{
    const int t = bk.x * blockDim.x + threadIdx.x;
    if (r.jbeg == r.jend) {
        const int k = r.kbeg + t, j = r.jbeg;
        if (k > r.kend) return;
        const long p = P(d, j, k);
        stc[p] = 1.0;
        if (j == 0) { sjp[p] = -1.0; sf[p] = anut[p + 1] - anut[p]; }
        else { sjm[p] = -1.0; sf[p] = anut[p - 1] - anut[p]; }
    } else {
        const int j = r.jbeg + t, k = r.kbeg;
        if (j > r.jend) return;
        const long p = P(d, j, k);
        stc[p] = 1.0;
        if (k == 0) { skp[p] = -1.0; sf[p] = anut[p + d.jm] - anut[p]; }
        else { skm[p] = -1.0; sf[p] = anut[p - d.jm] - anut[p]; }
    }
}

__global__ void k_sa_bc(Dims d, BcRegion r, const double *__restrict__ anut, double *sjm,
                        double *sjp, double *skm, double *skp, double *stc, double *sf)
{
    b_sa_bc(hw_blk(), d, r, anut, sjm, sjp, skm, skp, stc, sf);
}

// lreltur in reference order: ntjsp passes of Gauss-Seidel over k lines with a Thomas solve
// along j (turswpj + scatri), the re-solve of line k = 1, then the wake injection
// (ref/lreltur.f:52-252, ref/bc.f:571-591).  One workgroup; lines are sequential.
constexpr int SA_THREADS = 256;

__global__ void __launch_bounds__(SA_THREADS)
k_sa_lreltur(Dims d, int ntjsp, int npair, const int *__restrict__ pairs,
             const double *__restrict__ anut, double *__restrict__ ds, double *__restrict__ sf,
             const double *__restrict__ sjm, const double *__restrict__ sjp,
             const double *__restrict__ skm, const double *__restrict__ skp,
             const double *__restrict__ stc)
{
    extern __shared__ double sh[];
    const int jm = d.jm, km = d.km;
    double *aa = sh, *bb = sh + jm, *cc = sh + 2 * jm, *ff = sh + 3 * jm;
    for (long i = threadIdx.x; i < d.N; i += SA_THREADS) ds[i] = 0.0;
    __syncthreads();
    for (int pass = 1; pass <= ntjsp; ++pass) {
        for (int kk = 0; kk <= km; ++kk) {
            const int k = kk < km ? kk : 0;   // the last line is the re-solve of k = 1
            for (int j = threadIdx.x; j < jm; j += SA_THREADS) {
                const long p = P(d, j, k);
                aa[j] = sjm[p];
                bb[j] = stc[p];
                cc[j] = sjp[p];
                if (k == 0) ff[j] = sf[p] - skp[p] * ds[p + jm];
                else if (k == km - 1) ff[j] = sf[p] - skm[p] * ds[p - jm];
                else ff[j] = sf[p] - skm[p] * ds[p - jm] - skp[p] * ds[p + jm];
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                ff[0] = ff[0] / bb[0];
                cc[0] = cc[0] / bb[0];
                for (int i = 1; i < jm - 1; ++i) {
                    double z = 1. / (bb[i] - aa[i] * cc[i - 1]);
                    cc[i] = cc[i] * z;
                    ff[i] = (ff[i] - aa[i] * ff[i - 1]) * z;
                }
                int i = jm - 1;
                ff[i] = (ff[i] - aa[i] * ff[i - 1]) / (bb[i] - aa[i] * cc[i - 1]);
                for (i = jm - 2; i >= 0; --i) ff[i] = ff[i] - cc[i] * ff[i + 1];
            }
            __syncthreads();
            for (int j = threadIdx.x; j < jm; j += SA_THREADS) ds[P(d, j, k)] = ff[j];
            __syncthreads();
        }
        for (int i = threadIdx.x; i < npair; i += SA_THREADS) {
            const long jt = pairs[2 * i], jb = pairs[2 * i + 1];
            double snew = 0.5 * (anut[jt + jm] + anut[jb + jm]);
            double dsnew = 0.5 * (ds[jt + jm] + ds[jb + jm]);
            ds[jt] = snew + dsnew - anut[jt];
            sf[jt] = ds[jt];
        }
        __syncthreads();
    }
}

// tsaupd + eddy viscosity (ref/tursa.f:314-347, 287-292)
SMAC_DEV void b_sa_update(const Blk &bk, long N, double vnu, double *__restrict__ anut,
                          const double *__restrict__ ds, double *__restrict__ vnut,
                          double *__restrict__ part)
{
    const double cv1 = F32(7.1);
    double mx = 0.0;
    for (long i = (long)bk.x * blockDim.x + threadIdx.x; i < N;
         i += (long)bk.nx * blockDim.x) {
        double a = dmax_(0.0, anut[i] + ds[i]);
        anut[i] = a;
        mx = dmax_(ds[i], mx);
        double chi = a;
        double fv1 = pw(chi, 3) / (pw(chi, 3) + pw(cv1, 3));
        vnut[i] = fv1 * a * vnu;
    }
    __shared__ double sm[256];
    sm[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] = dmax_(sm[threadIdx.x + w], sm[threadIdx.x]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[bk.x] = sm[0];
}

__global__ void __launch_bounds__(256)
k_sa_update(long N, double vnu, double *__restrict__ anut, const double *__restrict__ ds,
            double *__restrict__ vnut, double *__restrict__ part)
{
    b_sa_update(hw_blk(), N, vnu, anut, ds, vnut, part);
}

SMAC_DEV void b_sa_turres(int n, const double *__restrict__ part, DevStats *st)
{
    __shared__ double sm[256];
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) mx = dmax_(part[i], mx);
    sm[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] = dmax_(sm[threadIdx.x + w], sm[threadIdx.x]);
        __syncthreads();
    }
    if (threadIdx.x == 0) st->turres = sm[0];
}

__global__ void k_sa_turres(int n, const double *__restrict__ part, DevStats *st)
{
    b_sa_turres(n, part, st);
}

// ------------------------------------------------------------------------------------------

void launch_sa_first(hipStream_t sm, const Dims &d, const Walls &w, const double *x,
                     const double *y, SaArrays &a)
{
    int nb = (int)((d.N + 255) / 256);
    hipLaunchKernelGGL(k_sa_first, dim3(nb), dim3(256), 0, sm, d, w, x, y, a.smin, a.anut);
}

void launch_sa_assemble(hipStream_t sm, const Dims &d, const DevParams &prm, int first_pass,
                        const double *q, const double *met, const double *dj, int nreg,
                        const BcRegion *regs, SaArrays &a)
{
    dim3 blk(64), grd((d.jm + 63) / 64, d.km);
    hipLaunchKernelGGL(k_sa_deriv, grd, blk, 0, sm, d, q, met, dj, a.vort, a.cu, a.ck, a.cv, a.fr);
    hipLaunchKernelGGL(k_sa_assemble, grd, blk, 0, sm, d, prm, q, met, dj, a.anut, a.smin,
                       first_pass, a.vort, a.cu, a.ck, a.cv, a.fr, a.sjm, a.sjp, a.skm, a.skp,
                       a.stc, a.sf);
    for (int i = 0; i < nreg; ++i) {
        const BcRegion &r = regs[i];
        int iv = r.ibcval;
        if (!(iv == 30 || iv == 31 || iv == 32 || iv == 33 || iv == 40)) continue;
        int n = r.jbeg == r.jend ? r.kend - r.kbeg + 1 : r.jend - r.jbeg + 1;
        hipLaunchKernelGGL(k_sa_bc, dim3((n + 63) / 64), dim3(64), 0, sm, d, r, a.anut, a.sjm,
                           a.sjp, a.skm, a.skp, a.stc, a.sf);
    }
}

void launch_sa_solve(hipStream_t sm, const Dims &d, const DevParams &prm, int ntjsp,
                     const WakePairs &w, const LusgsWs *ws, SaArrays &a, const LrScan *scan)
{
    if (prm.impsch == 1 && scan) {
        launch_sa_lreltur_scan(sm, d, ntjsp, w, a, *scan);
    } else if (prm.impsch == 1) {
        hipLaunchKernelGGL(k_sa_lreltur, dim3(1), dim3(SA_THREADS), sizeof(double) * 4 * d.jm, sm,
                           d, ntjsp, w.n, w.pairs, a.anut, a.ds, a.sf, a.sjm, a.sjp, a.skm, a.skp,
                           a.stc);
    } else {
        launch_lusgs1(sm, d, *ws, w, a.anut, a.ds, a.sf, a.stc, a.sjm, a.sjp, a.skm, a.skp,
                      nullptr, false);
    }
}

void launch_sa_update(hipStream_t sm, const Dims &d, const DevParams &prm, Scratch &w,
                      DevStats *st, SaArrays &a)
{
    hipLaunchKernelGGL(k_sa_update, dim3(w.nred), dim3(256), 0, sm, d.N, prm.vnu, a.anut, a.ds,
                       a.vnut, w.sa_v);
    hipLaunchKernelGGL(k_sa_turres, dim3(1), dim3(256), 0, sm, w.nred, w.sa_v, st);
}

// ---- multi-case launches ------------------------------------------------------------------

SaDerivArgs sa_deriv_args(const Dims &d, const double *q, const double *met, const double *dj,
                          SaArrays &a)
{
    return SaDerivArgs{(unsigned)((d.jm + 63) / 64), (unsigned)d.km, d, q, met, dj, a.vort, a.cu,
                       a.ck, a.cv, a.fr};
}

SaAsmArgs sa_asm_args(const Dims &d, const DevParams &prm, int first_pass, const double *q,
                      const double *met, const double *dj, SaArrays &a)
{
    return SaAsmArgs{(unsigned)((d.jm + 63) / 64), (unsigned)d.km, d, prm, q, met, dj, a.anut,
                     a.smin, first_pass, a.vort, a.cu, a.ck, a.cv, a.fr, a.sjm, a.sjp, a.skm,
                     a.skp, a.stc, a.sf};
}

void sa_bc_args(const Dims &d, int nreg, const BcRegion *regs, SaArrays &a,
                std::vector<SaBcArgs> &out)
{
    for (int i = 0; i < nreg; ++i) {
        const BcRegion &r = regs[i];
        const int iv = r.ibcval;
        if (!(iv == 30 || iv == 31 || iv == 32 || iv == 33 || iv == 40)) continue;
        const int n = r.jbeg == r.jend ? r.kend - r.kbeg + 1 : r.jend - r.jbeg + 1;
        out.push_back(SaBcArgs{(unsigned)((n + 63) / 64), 1u, d, r, a.anut, a.sjm, a.sjp, a.skm,
                               a.skp, a.stc, a.sf});
    }
}

SaUpdArgs sa_upd_args(const Dims &d, const DevParams &prm, Scratch &w, SaArrays &a)
{
    return SaUpdArgs{(unsigned)w.nred, 1u, d.N, prm.vnu, a.anut, a.ds, a.vnut, w.sa_v};
}

SaTurArgs sa_tur_args(Scratch &w, DevStats *st) { return SaTurArgs{1u, 1u, w.nred, w.sa_v, st}; }

struct CSaDeriv {
    using Args = SaDerivArgs;
    static constexpr int THREADS = 64;
    static SMAC_DEV void run(const Blk &b, const Args &a)
    {
        b_sa_deriv(b, a.d, a.q, a.met, a.dj, a.vort, a.cu, a.ck, a.cv, a.fr);
    }
};
struct CSaAsm {
    using Args = SaAsmArgs;
    static constexpr int THREADS = 64;
    static SMAC_DEV void run(const Blk &b, const Args &a)
    {
        b_sa_assemble(b, a.d, a.prm, a.q, a.met, a.dj, a.anut, a.smin, a.use_smin, a.vort, a.cu,
                      a.ck, a.cv, a.fr, a.sjm, a.sjp, a.skm, a.skp, a.stc, a.sf);
    }
};
struct CSaBc {
    using Args = SaBcArgs;
    static constexpr int THREADS = 64;
    static SMAC_DEV void run(const Blk &b, const Args &a)
    {
        b_sa_bc(b, a.d, a.r, a.anut, a.sjm, a.sjp, a.skm, a.skp, a.stc, a.sf);
    }
};
struct CSaUpd {
    using Args = SaUpdArgs;
    static constexpr int THREADS = 256;
    static SMAC_DEV void run(const Blk &b, const Args &a)
    {
        b_sa_update(b, a.N, a.vnu, a.anut, a.ds, a.vnut, a.part);
    }
};
struct CSaTur {
    using Args = SaTurArgs;
    static constexpr int THREADS = 256;
    static SMAC_DEV void run(const Blk &, const Args &a) { b_sa_turres(a.n, a.part, a.st); }
};

template <class C>
static void multi_launch(hipStream_t sm, const MultiTab<typename C::Args> &t)
{
    if (t.n > 0 && t.nblk > 0)
        hipLaunchKernelGGL(k_multi<C>, dim3(t.nblk), dim3(C::THREADS), 0, sm, t.args, t.start, t.n,
                           t.item_case, t.active);
}

void multi_sa_deriv(hipStream_t sm, const MultiTab<SaDerivArgs> &t) { multi_launch<CSaDeriv>(sm, t); }
void multi_sa_assemble(hipStream_t sm, const MultiTab<SaAsmArgs> &t) { multi_launch<CSaAsm>(sm, t); }
void multi_sa_bc(hipStream_t sm, const MultiTab<SaBcArgs> &t) { multi_launch<CSaBc>(sm, t); }
void multi_sa_update(hipStream_t sm, const MultiTab<SaUpdArgs> &t) { multi_launch<CSaUpd>(sm, t); }
void multi_sa_turres(hipStream_t sm, const MultiTab<SaTurArgs> &t) { multi_launch<CSaTur>(sm, t); }

}  // namespace smac
