/*
 * Recorder behind the Fortran seams of oracle/ref_seams.F (TEST
 * INFRASTRUCTURE ONLY).  It writes what the reference computed, at full
 * precision, so tests can pin the C restatement and the GPU path against
 * the compiled reference.
 *
 * Environment:
 *   SMAC_TRACE     text trace file (default smac_trace.txt); one line per
 *                  record, doubles as %.17e:
 *                    STEP  nt resmax0 resmax jres kres divmax turres clift_in
 *                    NORM  nt rhsnorm          (= ||s||_2, reference order)
 *                    FORCE nt alpha fxp fyp fxs fys fmom
 *   SMAC_DUMP      comma-separated list of nt at which to dump arrays
 *   SMAC_DUMP_DIR  directory for the dumps (default .); raw little-endian
 *                  float64 in Fortran order, files named <what>_<nt>.bin:
 *                    qin   q, anut entering step          (4 planes)
 *                    lhs   q, s, btc, bjm, bjp, bkm, bkp  (51 planes)
 *                    dq    dq and s after the solve       (6 planes)
 *                    qout  q, anut, vnut, dq after step   (6 planes)
 *                    sa    anut, sf, sjm, sjp, skm, skp, stc entering the SA
 *                          line relaxation                  (7 planes)
 *                    sads  ds after the SA line relaxation   (1 plane)
 *                    sascr vort diff conv prod dest fr1 tran of the SA
 *                          assembly                          (7 planes)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static FILE *trace_fp(void)
{
    static FILE *fp;
    if (!fp) {
        const char *path = getenv("SMAC_TRACE");
        fp = fopen(path && *path ? path : "smac_trace.txt", "w");
        if (!fp) { perror("ref_trace: open"); abort(); }
    }
    return fp;
}

static int want_dump(int nt)
{
    const char *list = getenv("SMAC_DUMP");
    if (!list) return 0;
    char buf[4096];
    strncpy(buf, list, sizeof buf - 1);
    buf[sizeof buf - 1] = 0;
    for (char *t = strtok(buf, ","); t; t = strtok(NULL, ","))
        if (atoi(t) == nt) return 1;
    return 0;
}

static void dump(const char *what, int nt, const double *const *planes, const long *counts,
                 int n)
{
    const char *dir = getenv("SMAC_DUMP_DIR");
    char path[4096];
    snprintf(path, sizeof path, "%s/%s_%d.bin", dir && *dir ? dir : ".", what, nt);
    FILE *fp = fopen(path, "wb");
    if (!fp) { perror("ref_trace: dump"); abort(); }
    for (int i = 0; i < n; ++i) fwrite(planes[i], sizeof(double), (size_t)counts[i], fp);
    fclose(fp);
}

void trace_step_(int *nt, double *resmax0, double *resmax, int *jres, int *kres, double *divmax,
                 double *turres, double *clift_in)
{
    fprintf(trace_fp(), "STEP %d %.17e %.17e %d %d %.17e %.17e %.17e\n", *nt, *resmax0, *resmax,
            *jres, *kres, *divmax, *turres, *clift_in);
    fflush(trace_fp());
}

void trace_norm_(int *nt, int *jmax, int *kmax, double *s)
{
    /* step.f:513-519 accumulates s(j,k,i)**2 with i outermost, j innermost,
     * which is plain linear order of the Fortran array. */
    long n = 3L * *jmax * *kmax;
    double acc = 0.0;
    for (long i = 0; i < n; ++i) acc = acc + s[i] * s[i];
    fprintf(trace_fp(), "NORM %d %.17e\n", *nt, sqrt(acc));
}

void trace_force_(int *nt, double *alpha, double *fxp, double *fyp, double *fxs, double *fys,
                  double *fmom)
{
    fprintf(trace_fp(), "FORCE %d %.17e %.17e %.17e %.17e %.17e %.17e\n", *nt, *alpha, *fxp, *fyp,
            *fxs, *fys, *fmom);
}

void trace_state_in_(int *nt, int *jmax, int *kmax, double *q, double *turvar)
{
    if (!want_dump(*nt)) return;
    long n = (long)*jmax * *kmax;
    const double *p[2] = {q, turvar};
    long c[2] = {3 * n, n};
    dump("qin", *nt, p, c, 2);
}

void trace_state_out_(int *nt, int *jmax, int *kmax, double *q, double *turvar, double *vnut,
                      double *dq)
{
    if (!want_dump(*nt)) return;
    long n = (long)*jmax * *kmax;
    const double *p[4] = {q, turvar, vnut, dq};
    long c[4] = {3 * n, n, n, 3 * n};
    dump("qout", *nt, p, c, 4);
}

void trace_lhs_(int *nt, int *jmax, int *kmax, double *q, double *s, double *btc, double *bjm,
                double *bjp, double *bkm, double *bkp)
{
    if (!want_dump(*nt)) return;
    long n = (long)*jmax * *kmax;
    const double *p[7] = {q, s, btc, bjm, bjp, bkm, bkp};
    long c[7] = {3 * n, 3 * n, 9 * n, 9 * n, 9 * n, 9 * n, 9 * n};
    dump("lhs", *nt, p, c, 7);
}

void trace_solve_(int *nt, int *jmax, int *kmax, double *dq, double *s)
{
    if (!want_dump(*nt)) return;
    long n = (long)*jmax * *kmax;
    const double *p[2] = {dq, s};
    long c[2] = {3 * n, 3 * n};
    dump("dq", *nt, p, c, 2);
}

void trace_sa_(int *nt, int *jmax, int *kmax, double *anut, double *sf, double *sjm, double *sjp,
               double *skm, double *skp, double *stc)
{
    if (!want_dump(*nt)) return;
    long n = (long)*jmax * *kmax;
    const double *p[7] = {anut, sf, sjm, sjp, skm, skp, stc};
    long c[7] = {n, n, n, n, n, n, n};
    dump("sa", *nt, p, c, 7);
    /* turvis hands tursa nine consecutive planes of step's scr1 as vort, diff, conv, prod,
     * dest, fr1, tran, sjm, sjp (ref/visc.f:464-469), so sjm - 7n is vort */
    const double *base = sjm - 7 * n;
    const double *q[7] = {base, base + n, base + 2 * n, base + 3 * n, base + 4 * n, base + 5 * n,
                          base + 6 * n};
    long cq[7] = {n, n, n, n, n, n, n};
    dump("sascr", *nt, q, cq, 7);
}

void trace_sads_(int *nt, int *jmax, int *kmax, double *ds)
{
    if (!want_dump(*nt)) return;
    long n = (long)*jmax * *kmax;
    const double *p[1] = {ds};
    long c[1] = {n};
    dump("sads", *nt, p, c, 1);
}
