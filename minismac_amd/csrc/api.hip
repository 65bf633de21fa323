// C ABI of the gfx950 implicit iteration (include/smac_amd.h): per-case device state, the
// launch sequence of one pass of the reference's time loop, and the host-array drop-ins for
// the reference's lrelax / lusgs call sites.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/smac_amd.h"
#include "smac_kernels.h"

using namespace smac;

static thread_local std::string g_err;

static smac_status fail(smac_status st, const std::string &msg)
{
    g_err = msg;
    return st;
}

// sets the calling thread's error text for the entry points of other translation units
smac_status smac_fail_text(smac_status st, const char *msg) { return fail(st, msg); }

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(SMAC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

struct smac_handle {
    smac_params p;
    DevParams dp;
    Dims d;
    int device = 0;
    hipStream_t stream = nullptr;
    // The SA model does not feed the mean flow within a pass (the viscous terms use the
    // laminar viscosity, ref/visc.f:97), so it runs on a second stream beside it.
    hipStream_t stream_sa = nullptr;
    hipEvent_t ev_q = nullptr, ev_sa_read = nullptr, ev_sa_done = nullptr;
    // In runs without a stop criterion a pass's wall forces run on stream_sa after the
    // update, beside the next pass's assembly; the next pass's BC rows (which read clift)
    // wait for ev_force.  force_side: the last pass left its forces there.
    hipEvent_t ev_force = nullptr;
    bool force_side = false;
    LusgsWs lws_sa = {};       // the SA's own LU-SGS workspace (neqs = 1)
    // grid (device)
    double *x = nullptr, *y = nullptr, *met = nullptr, *dj = nullptr;
    // flow state and LHS (device, FP64)
    double *q = nullptr, *dq = nullptr, *s = nullptr;
    double *blk[6] = {};   // btc bjm bjp bkm bkp scr1
    double *lhs_copy[5] = {};
    SaArrays sa = {};
    DevStats *st = nullptr;
    // Pinned host copy of the statistics and its sequence word, written by k_publish at the end
    // of each pass whose statistics the host reads (host pointer; pin_dev is the device's).
    struct Pinned {
        DevStats s;
        unsigned seq;
    } *pin = nullptr, *pin_dev = nullptr;
    unsigned pin_seq = 0;
    Scratch w = {};
    int *pairs = nullptr;
    WakePairs wake = {};
    LusgsWs lws = {};          // LU-SGS sheared workspace (impsch = 2)
    int lusgs_rows = 64;
    int lusgs_ring = 1;    // the 3x3 sweeps stream their operands through an LDS ring
    int lr_lines = 64;     // lines per workgroup of the line-relaxation factor kernels
    int lr_refactor = 1;   // scan/zebra forms: factor the lines every this many passes
    long lr_factored = -1; // the pass whose blocks the line operands hold (-1: none)
    // line relaxation with each line solved by prefix scans inside the line (1), or by the
    // reference's sequential substitution (0, bit-identical to the reference)
    int lr_solve = 0;
    LrScan lrs = {};
    // discrete adjoint (allocated by the first smac_adjoint call)
    bool adj_ready = false;
    int adj_m = 0;
    AdjCtx adj = {};
    LusgsWs adj_ws = {};
    LusgsWs adj_ws2 = {};      // the second adjoint solve's (LU-SGS preconditioner)
    AdjCtx adj2 = {};          // the second adjoint solve's work vectors (the rest is adj's)
    std::vector<void *> lws_allocs;
    Walls walls = {};
    std::vector<BcRegion> regs;
    int nt = 0;
    double rhs0 = 0.0;            // rhsnorm of pass nt = 1 when this handle made it (else 0)
    int adj_restart = 100;        // GMRES restart length of smac_adjoint
    int adj_prec_steps = 3;       // steps of the transposed update per preconditioner call
    int adj_reorth = 100;         // per mille: when its Gram-Schmidt projects a second time
    bool first_sa = true;
    bool keep_lhs = false;
    bool timing = false;
    hipEvent_t ev[7] = {};
    hipEvent_t ev_band[2] = {};   // forward band sweep of the mean flow (timing mode)
    double t_acc[7] = {};
    int t_n = 0;
    std::vector<void *> allocs;
};

static bool check_gpu(std::string &why)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        why = "no HIP device visible";
        return false;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, dev) != hipSuccess) {
        why = "hipGetDeviceProperties failed";
        return false;
    }
    if (std::strncmp(pr.gcnArchName, "gfx950", 6) != 0) {
        why = std::string("this library is built for gfx950; device is ") + pr.gcnArchName;
        return false;
    }
    return true;
}

template <typename T>
static smac_status dalloc(smac_handle *h, T **ptr, size_t n)
{
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, n * sizeof(T) + 256));
    HIP_TRY(hipMemset(p, 0, n * sizeof(T) + 256));
    h->allocs.push_back(p);
    *ptr = (T *)p;
    return SMAC_OK;
}

static void free_lusgs_ws(smac_handle *h)
{
    for (void *p : h->lws_allocs) (void)hipFree(p);
    h->lws_allocs.clear();
    h->lws = LusgsWs{};
    h->lws_sa = LusgsWs{};
    h->adj_ws = LusgsWs{};
    h->adj_ws2 = LusgsWs{};
}

static smac_status alloc_one_ws(smac_handle *h, int rows, int neqs, LusgsWs &w)
{
    const int pf3 = w.pf3, pf1 = w.pf1, ring = w.ring;
    hipEvent_t e0 = w.ev_fwd[0], e1 = w.ev_fwd[1];
    w = LusgsWs{};
    w.pf3 = pf3;
    w.pf1 = pf1;
    w.ring = ring;
    w.ev_fwd[0] = e0;
    w.ev_fwd[1] = e1;
    lusgs_ws_shape(h->d.jm, h->d.km, rows, w);
    if (!lusgs_ws_fits(w)) return fail(SMAC_ERR_SIZE, "grid too large for the LU-SGS workspace");
    void *p[4] = {};
    size_t sz[4] = {(size_t)lusgs_ws_lhs_bytes(w, neqs), (size_t)lusgs_ws_vec_bytes(w, neqs),
                    (size_t)lusgs_ws_vec_bytes(w, neqs), (size_t)(w.nedge * 8)};
    for (int i = 0; i < 4; ++i) {
        HIP_TRY(hipMalloc(&p[i], sz[i] + 256));
        h->lws_allocs.push_back(p[i]);
        // the padding of the sheared planes is read (and ignored) by idle lanes, and the
        // hand-over padding must not hold the sentinel: start from zeros
        HIP_TRY(hipMemset(p[i], 0, sz[i] + 256));
    }
    w.lhs = p[0];
    w.r = p[1];
    w.dq = p[2];
    w.edge = (unsigned long long *)p[3];
    w.stall = &h->st->stall;
    return SMAC_OK;
}

// LU-SGS workspaces: the mean flow (neqs = 3) and the SA (neqs = 1), which may run at once
static smac_status alloc_lusgs_ws(smac_handle *h, int rows)
{
    const LusgsWs keep = h->lws, keep_sa = h->lws_sa;
    free_lusgs_ws(h);
    h->lws.ev_fwd[0] = keep.ev_fwd[0];
    h->lws.ev_fwd[1] = keep.ev_fwd[1];
    h->lws.pf3 = keep.pf3;
    h->lws.pf1 = keep.pf1;
    h->lws.ring = h->lusgs_ring;
    h->lws_sa.pf3 = keep_sa.pf3;
    h->lws_sa.pf1 = keep_sa.pf1;
    smac_status rc = alloc_one_ws(h, rows, 3, h->lws);
    if (rc) return rc;
    rc = alloc_one_ws(h, rows, 1, h->lws_sa);
    if (rc) return rc;
    h->lusgs_rows = h->lws.rows;
    return SMAC_OK;
}

static void free_handle(smac_handle *h)
{
    if (!h) return;
    free_lusgs_ws(h);
    for (void *p : h->allocs) (void)hipFree(p);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : h->ev_band)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    if (h->stream_sa) (void)hipStreamDestroy(h->stream_sa);
    for (hipEvent_t e : {h->ev_q, h->ev_sa_read, h->ev_sa_done, h->ev_force})
        if (e) (void)hipEventDestroy(e);
    if (h->pin) (void)hipHostFree(h->pin);
    delete h;
}

static smac_status check_params(const smac_params *p)
{
    if (!p) return fail(SMAC_ERR_ARG, "params is NULL");
    if (p->impsch != 1 && p->impsch != 2)
        return fail(SMAC_ERR_IMPSCH, "impsch must be 1 (line relaxation) or 2 (LU-SGS)");
    if (p->iflxo != 3) return fail(SMAC_ERR_IFLXO, "iflxo must be 3 (third-order upwind flux)");
    if (p->ivis != 0 && p->ivis != 1) return fail(SMAC_ERR_ARG, "ivis must be 0 or 1");
    if (p->iturb != 0 && p->iturb != 1) return fail(SMAC_ERR_ARG, "iturb must be 0 or 1");
    if (!(p->underr > 0.0 && p->underr <= 1.0))
        return fail(SMAC_ERR_ARG, "underr must be in (0, 1] (ref/initia.f:243-247)");
    if (p->nbcimp < 1) return fail(SMAC_ERR_ARG, "nbcimp must be >= 1");
    if (p->njsp < 0 || p->nksp < 0 || p->ntjsp < 0) return fail(SMAC_ERR_ARG, "negative sweep count");
    if (p->ntksp != 0)
        return fail(SMAC_ERR_ARG, "ntksp > 0 (SA eta sweeps) is not supported; the reference uses 0");
    if (p->precision != SMAC_FP64 && p->precision != SMAC_FP32_LHS)
        return fail(SMAC_ERR_ARG, "precision must be SMAC_FP64 or SMAC_FP32_LHS");
    if (p->line_order != SMAC_ORDER_REFERENCE && p->line_order != SMAC_ORDER_ZEBRA)
        return fail(SMAC_ERR_ARG, "line_order must be SMAC_ORDER_REFERENCE or SMAC_ORDER_ZEBRA");
    if (p->line_order == SMAC_ORDER_ZEBRA && p->impsch != 1)
        return fail(SMAC_ERR_ARG, "SMAC_ORDER_ZEBRA orders the lines of line relaxation (impsch = 1)");
    if (!(p->dtau > 0.0) || !(p->reynum > 0.0) || !(p->beta > 0.0) || !(p->dt > 0.0))
        return fail(SMAC_ERR_ARG, "dtau, reynum, beta and dt must be positive");
    return SMAC_OK;
}

extern "C" {

const char *smac_last_error(void) { return g_err.c_str(); }
const char *smac_version(void) { return "smac_amd 0.1 (gfx950)"; }

smac_status smac_create(const smac_params *p, int jmax, int kmax, const double *x,
                        const double *y, const double *rtxy, const double *dj, int nreg,
                        const smac_bc_region *regs, smac_handle **out)
{
    g_err.clear();
    if (!out) return fail(SMAC_ERR_ARG, "out is NULL");
    *out = nullptr;
    smac_status rc = check_params(p);
    if (rc) return rc;
    // FP32-LHS line relaxation and the zebra order exist only in the scan form of the solve
    const bool scan_only = p->impsch == 1 && (p->precision == SMAC_FP32_LHS ||
                                              p->line_order == SMAC_ORDER_ZEBRA);
    if (!x || !y || !rtxy || !dj || !regs || nreg < 1)
        return fail(SMAC_ERR_ARG, "grid, metric or BC arrays missing");
    if (jmax < 5 || kmax < 5 || jmax > 8100 || kmax > 8100)
        return fail(SMAC_ERR_SIZE, "jmax and kmax must be in [5, 8100] (ref/include/common.f:45)");
    if ((p->impsch == 1 && (jmax > 2700 || kmax > 2700)) || (p->iturb && jmax > 2000))
        return fail(SMAC_ERR_SIZE, "reference-order line sweeps stage a line in LDS: "
                                   "lines longer than 2700 (2000 with SA) are not supported");
    if (scan_only && (jmax > 2048 || kmax > 2048))
        return fail(SMAC_ERR_SIZE, "the scan form of the line solves takes lines of up to 2048 points");
    std::string why;
    if (!check_gpu(why)) return fail(SMAC_ERR_NOGPU, why);
    const long N = (long)jmax * kmax;
    for (long i = 0; i < N; ++i)
        if (rtxy[i] != 0.0 || rtxy[N + i] != 0.0)
            return fail(SMAC_ERR_ARG, "time metrics rtxy(:,:,:,1) must be zero (static grid)");
    smac_handle *h = new smac_handle();
    h->p = *p;
    h->d = Dims{jmax, kmax, N};
    DevParams &dp = h->dp;
    dp.beta = p->beta; dp.vnu = 1.0 / p->reynum; dp.dtau = p->dtau; dp.dt = p->dt;
    dp.underr = p->underr; dp.cdis = p->cdis; dp.pout = p->pout; dp.alpha = p->alpha;
    dp.chord = p->chord; dp.xmom = p->xmom; dp.ymom = p->ymom;
    dp.impsch = p->impsch; dp.ivis = p->ivis; dp.iturb = p->iturb;
    // BC regions -> 0-based, with the structure the kernels rely on
    int wake_idx[2] = {-1, -1}, nwake = 0;
    for (int i = 0; i < nreg; ++i) {
        const smac_bc_region &r = regs[i];
        BcRegion b{r.ibcval, r.jbeg - 1, r.jend - 1, r.kbeg - 1, r.kend - 1};
        bool jline = b.jbeg == b.jend, kline = b.kbeg == b.kend;
        if (b.jbeg < 0 || b.jend >= jmax || b.kbeg < 0 || b.kend >= kmax || b.jbeg > b.jend ||
            b.kbeg > b.kend || !(jline || kline)) {
            free_handle(h);
            return fail(SMAC_ERR_BC, "BC region " + std::to_string(i) + " out of range or not a line");
        }
        bool ok = (r.ibcval == 31 && jline && (b.jbeg == 0 || b.jbeg == jmax - 1)) ||
                  (r.ibcval == 0 && kline && b.kbeg == 0) ||
                  (r.ibcval == 25 && kline && b.kbeg == kmax - 1) ||
                  (r.ibcval == 60 && kline && b.kbeg == 0);
        if (!ok) {
            free_handle(h);
            return fail(SMAC_ERR_BC, "BC region " + std::to_string(i) + " (ibcval " +
                                         std::to_string(r.ibcval) +
                                         ") is not one of the C-grid rows this library implements");
        }
        if (r.ibcval == 0) {
            if (h->walls.n >= 8) { free_handle(h); return fail(SMAC_ERR_BC, "more than 8 walls"); }
            h->walls.k[h->walls.n] = b.kbeg;
            h->walls.j1[h->walls.n] = b.jbeg;
            h->walls.j2[h->walls.n] = b.jend;
            h->walls.n++;
        }
        if (r.ibcval == 60 && nwake < 2) wake_idx[nwake++] = i;
        h->regs.push_back(b);
    }
    std::vector<int> pairs;
    if (nwake == 2) {
        const smac_bc_region &a = regs[wake_idx[0]], &b = regs[wake_idx[1]];
        int na = a.jend - a.jbeg + 1, nb = b.jend - b.jbeg + 1, n = na < nb ? na : nb;
        for (int i = 0; i < n; ++i) { pairs.push_back(a.jbeg + i - 1); pairs.push_back(b.jend - i - 1); }
        for (int i = 0; i < n; ++i) { pairs.push_back(b.jbeg + i - 1); pairs.push_back(a.jend - i - 1); }
    }
#define TRY(x)                     \
    do {                           \
        smac_status r_ = (x);      \
        if (r_) {                  \
            std::string e_ = g_err; \
            free_handle(h);        \
            g_err = e_;            \
            return r_;             \
        }                          \
    } while (0)
    (void)hipGetDevice(&h->device);
    if (hipStreamCreate(&h->stream) != hipSuccess || hipStreamCreate(&h->stream_sa) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_q, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_sa_read, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_sa_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_force, hipEventDisableTiming) != hipSuccess) {
        free_handle(h);
        return fail(SMAC_ERR_HIP, "hipStreamCreate / hipEventCreate failed");
    }
    TRY(dalloc(h, &h->x, N));
    TRY(dalloc(h, &h->y, N));
    TRY(dalloc(h, &h->met, 4 * N));
    TRY(dalloc(h, &h->dj, N));
    TRY(dalloc(h, &h->q, 3 * N));
    TRY(dalloc(h, &h->dq, 3 * N));
    TRY(dalloc(h, &h->s, 3 * N));
    for (int b = 0; b < 6; ++b) TRY(dalloc(h, &h->blk[b], 9 * N));
    double **sa_planes[] = {&h->sa.anut, &h->sa.vnut, &h->sa.smin, &h->sa.sjm, &h->sa.sjp,
                            &h->sa.skm, &h->sa.skp, &h->sa.stc, &h->sa.sf, &h->sa.ds,
                            &h->sa.vort, &h->sa.cu, &h->sa.ck, &h->sa.cv, &h->sa.fr, &h->sa.bprm};
    for (double **pp : sa_planes) TRY(dalloc(h, pp, N));
    TRY(dalloc(h, &h->st, 1));
    if (hipHostMalloc((void **)&h->pin, sizeof *h->pin, hipHostMallocCoherent | hipHostMallocMapped) !=
            hipSuccess ||
        hipHostGetDevicePointer((void **)&h->pin_dev, h->pin, 0) != hipSuccess) {
        free_handle(h);
        return fail(SMAC_ERR_HIP, "hipHostMalloc / hipHostGetDevicePointer failed (statistics)");
    }
    std::memset((void *)h->pin, 0, sizeof *h->pin);
    if (p->impsch == 2) TRY(alloc_lusgs_ws(h, h->lusgs_rows));
    if (p->impsch == 1) {
        char *lb = nullptr;
        TRY(dalloc(h, &lb, (size_t)lr_scan_bytes(h->d)));
        h->lrs.opx = lb;
        h->lrs.ope = lb + 45 * N * 8 + 256;
        h->lrs.s_t = (double *)(lb + 2 * (45 * N * 8 + 256));
        h->lrs.dq_t = h->lrs.s_t + 3 * N;
        h->lrs.sa_op = (double *)((char *)(h->lrs.dq_t + 3 * N) + 256);
        h->lrs.zebra = p->line_order == SMAC_ORDER_ZEBRA;
        h->lr_solve = scan_only ? 1 : 0;
    }
    h->w.nred = 512;
    TRY(dalloc(h, &h->w.red_v, 2 * h->w.nred));
    TRY(dalloc(h, &h->w.red_i, 2 * h->w.nred));
    TRY(dalloc(h, &h->w.red_s, h->w.nred));
    TRY(dalloc(h, &h->w.sa_v, h->w.nred));
    TRY(dalloc(h, &h->w.flag, 1));
    if (!pairs.empty()) TRY(dalloc(h, &h->pairs, pairs.size()));
    h->wake = WakePairs{(int)pairs.size() / 2, h->pairs};
    // uploads
    std::vector<double> met(4 * N), q0(3 * N);
    std::memcpy(met.data(), rtxy + 2 * N, N * sizeof(double));          // xi_x  = rtxy(.,.,1,2)
    std::memcpy(met.data() + N, rtxy + 4 * N, N * sizeof(double));      // xi_y  = rtxy(.,.,1,3)
    std::memcpy(met.data() + 2 * N, rtxy + 3 * N, N * sizeof(double));  // eta_x = rtxy(.,.,2,2)
    std::memcpy(met.data() + 3 * N, rtxy + 5 * N, N * sizeof(double));  // eta_y = rtxy(.,.,2,3)
    // free-stream initial condition, u = v = 0 on walls (ref/geom.f:88-131)
    double pi = 4.0 * std::atan(1.0);
    double ca = std::cos(p->alpha * pi / 180.), sa = std::sin(p->alpha * pi / 180.);
    for (long i = 0; i < N; ++i) { q0[i] = 1.0; q0[N + i] = ca; q0[2 * N + i] = sa; }
    for (int wv = 0; wv < h->walls.n; ++wv)
        for (int j = h->walls.j1[wv]; j <= h->walls.j2[wv]; ++j) {
            long pnt = (long)j + (long)jmax * h->walls.k[wv];
            q0[N + pnt] = 0.0;
            q0[2 * N + pnt] = 0.0;
        }
    if (hipMemcpy(h->x, x, N * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->y, y, N * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->dj, dj, N * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->met, met.data(), 4 * N * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->q, q0.data(), 3 * N * 8, hipMemcpyHostToDevice) != hipSuccess ||
        (!pairs.empty() && hipMemcpy(h->pairs, pairs.data(), pairs.size() * 4,
                                     hipMemcpyHostToDevice) != hipSuccess)) {
        free_handle(h);
        return fail(SMAC_ERR_HIP, "upload failed");
    }
    for (auto &e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) { free_handle(h); return fail(SMAC_ERR_HIP, "hipEventCreate failed"); }
    for (auto &e : h->ev_band)
        if (hipEventCreate(&e) != hipSuccess) { free_handle(h); return fail(SMAC_ERR_HIP, "hipEventCreate failed"); }
#undef TRY
    *out = h;
    return SMAC_OK;
}

smac_status smac_destroy(smac_handle *h)
{
    if (!h) return fail(SMAC_ERR_ARG, "handle is NULL");
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamSynchronize(h->stream_sa);
    free_handle(h);
    return SMAC_OK;
}

smac_status smac_set_option(smac_handle *h, const char *name, int value)
{
    if (!h || !name) return fail(SMAC_ERR_ARG, "handle or name is NULL");
    if (!std::strcmp(name, "keep_lhs")) {
        if (value && h->p.precision != SMAC_FP64)
            return fail(SMAC_ERR_ARG, "keep_lhs keeps FP64 blocks; the handle is FP32-LHS");
        h->keep_lhs = value != 0;
        if (h->keep_lhs && !h->lhs_copy[0])
            for (int b = 0; b < 5; ++b) {
                smac_status rc = dalloc(h, &h->lhs_copy[b], 9 * h->d.N);
                if (rc) return rc;
            }
        return SMAC_OK;
    }
    if (!std::strcmp(name, "timing")) {
        h->timing = value != 0;
        h->lws.ev_fwd[0] = h->timing ? h->ev_band[0] : nullptr;
        h->lws.ev_fwd[1] = h->timing ? h->ev_band[1] : nullptr;
        return SMAC_OK;
    }
    if (!std::strcmp(name, "lusgs_prefetch") || !std::strcmp(name, "lusgs_prefetch_sa")) {
        const bool mean = !std::strcmp(name, "lusgs_prefetch");
        const bool ok = value == 1 || value == 2 || value == 3 || value == 4 || value == 6 ||
                        value == 8;
        if (!ok) return fail(SMAC_ERR_ARG, std::string(name) + ": unsupported prefetch distance");
        if (mean) h->lws.pf3 = value;
        else h->lws_sa.pf1 = value;
        return SMAC_OK;
    }
    if (!std::strcmp(name, "lusgs_ring")) {
        if (value != 0 && value != 1) return fail(SMAC_ERR_ARG, "lusgs_ring must be 0 or 1");
        h->lusgs_ring = value;
        h->lws.ring = value;
        return SMAC_OK;
    }
    if (!std::strcmp(name, "lr_solve")) {
        if (value != 0 && value != 1) return fail(SMAC_ERR_ARG, "lr_solve must be 0 or 1");
        if (h->p.impsch != 1) return fail(SMAC_ERR_ARG, "lr_solve applies to line relaxation (impsch = 1)");
        if (!value && (h->p.precision == SMAC_FP32_LHS || h->p.line_order == SMAC_ORDER_ZEBRA))
            return fail(SMAC_ERR_ARG, "FP32-LHS and zebra line relaxation exist only with lr_solve = 1");
        if (value && (h->d.jm > 2048 || h->d.km > 2048))
            return fail(SMAC_ERR_SIZE, "the scan form of the line solves takes lines of up to 2048 points");
        h->lr_solve = value;
        h->lr_factored = -1;   // the scan workspace holds no operands of this case yet
        return SMAC_OK;
    }
    if (!std::strcmp(name, "lr_refactor")) {
        if (value < 1 || value > 64) return fail(SMAC_ERR_ARG, "lr_refactor must be in [1, 64]");
        h->lr_refactor = value;
        h->lr_factored = -1;
        return SMAC_OK;
    }
    if (!std::strcmp(name, "lr_lines")) {
        if (value < 1 || value > 64) return fail(SMAC_ERR_ARG, "lr_lines must be in [1, 64]");
        h->lr_lines = value;
        return SMAC_OK;
    }
    if (!std::strcmp(name, "adj_reorth")) {
        if (value < 0 || value > 1000) return fail(SMAC_ERR_ARG, "adj_reorth must be in [0, 1000]");
        h->adj_reorth = value;
        return SMAC_OK;
    }
    if (!std::strcmp(name, "adj_prec_steps")) {
        if (value < 1 || value > 8) return fail(SMAC_ERR_ARG, "adj_prec_steps must be in [1, 8]");
        h->adj_prec_steps = value;
        return SMAC_OK;
    }
    if (!std::strcmp(name, "adj_restart")) {
        if (value < 4 || value > 400) return fail(SMAC_ERR_ARG, "adj_restart must be in [4, 400]");
        h->adj_restart = value;
        return SMAC_OK;
    }
    if (!std::strcmp(name, "lusgs_rows")) {
        if (value < 1 || value > 64) return fail(SMAC_ERR_ARG, "lusgs_rows must be in [1, 64]");
        h->lusgs_rows = value;
        if (h->p.impsch != 2) return SMAC_OK;
        HIP_TRY(hipStreamSynchronize(h->stream));
        return alloc_lusgs_ws(h, value);
    }
    return fail(SMAC_ERR_ARG, std::string("unknown option ") + name);
}

smac_status smac_set_state(smac_handle *h, const double *q, const double *anut, int nt,
                           double resmax0, double clift)
{
    if (!h) return fail(SMAC_ERR_ARG, "handle is NULL");
    const long N = h->d.N;
    if (q) HIP_TRY(hipMemcpy(h->q, q, 3 * N * 8, hipMemcpyHostToDevice));
    if (anut) HIP_TRY(hipMemcpy(h->sa.anut, anut, N * 8, hipMemcpyHostToDevice));
    DevStats st;
    HIP_TRY(hipMemcpy(&st, h->st, sizeof st, hipMemcpyDeviceToHost));
    st.resmax0 = resmax0;
    st.clift = clift;
    HIP_TRY(hipMemcpy(h->st, &st, sizeof st, hipMemcpyHostToDevice));
    h->nt = nt;
    h->rhs0 = 0.0;
    h->lr_factored = -1;
    if (nt > 0) h->first_sa = false;
    return SMAC_OK;
}

smac_status smac_get_state(smac_handle *h, double *q, double *anut, double *vnut)
{
    if (!h) return fail(SMAC_ERR_ARG, "handle is NULL");
    const long N = h->d.N;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (q) HIP_TRY(hipMemcpy(q, h->q, 3 * N * 8, hipMemcpyDeviceToHost));
    if (anut) HIP_TRY(hipMemcpy(anut, h->sa.anut, N * 8, hipMemcpyDeviceToHost));
    if (vnut) HIP_TRY(hipMemcpy(vnut, h->sa.vnut, N * 8, hipMemcpyDeviceToHost));
    return SMAC_OK;
}

static void mark(smac_handle *h, int i)
{
    if (h->timing) (void)hipEventRecord(h->ev[i], h->stream);
}

// the line-relaxation solve of a pass: the scan form (FP32 or FP64 blocks, reference or zebra
// line order) or the reference's sequential substitution
static void enqueue_lrelax(smac_handle *h, hipStream_t sm)
{
    const Dims &d = h->d;
    double *btc = h->blk[0], *bjm = h->blk[1], *bjp = h->blk[2], *bkm = h->blk[3],
           *bkp = h->blk[4];
    const bool zebra = h->p.line_order == SMAC_ORDER_ZEBRA;
    // lr_refactor > 1: the lines keep the operands (factored line blocks and the neighbour
    // blocks) of an earlier pass for up to that many passes, with this pass's residual; the
    // fixed point s = 0 is unchanged (a lagged implicit operator, not the reference's)
    const bool factor = !h->lr_solve || h->lr_refactor <= 1 || h->lr_factored < 0 ||
                        h->nt - h->lr_factored >= h->lr_refactor;
    h->lr_factored = !h->lr_solve ? -1 : (factor ? h->nt : h->lr_factored);
    if (!h->lr_solve)
        launch_lrelax_ref<double>(sm, d, h->dp, h->p.njsp, h->p.nksp, h->p.nbcimp, h->wake, h->q,
                                  h->dq, h->s, btc, bjm, bjp, bkm, bkp, h->lr_lines);
    else if (h->p.precision == SMAC_FP32_LHS)
        launch_lrelax_scan<float>(sm, d, h->dp, h->p.njsp, h->p.nksp, h->p.nbcimp, zebra, h->wake,
                                  h->q, h->dq, h->s, (const float *)btc, (const float *)bjm,
                                  (const float *)bjp, (const float *)bkm, (const float *)bkp,
                                  h->lrs, factor);
    else
        launch_lrelax_scan<double>(sm, d, h->dp, h->p.njsp, h->p.nksp, h->p.nbcimp, zebra,
                                   h->wake, h->q, h->dq, h->s, btc, bjm, bjp, bkm, bkp, h->lrs,
                                   factor);
}

// one pass of the time loop body: SA, assembly, BC rows, residual, solve, update + forces
// defer_force: leave this pass's wall forces on stream_sa, unjoined (smac_run between the
// passes of a run without a stop criterion); the next pass joins them before its BC rows.
static smac_status enqueue_pass(smac_handle *h, bool defer_force = false)
{
    const Dims &d = h->d;
    const DevParams &dp = h->dp;
    hipStream_t sm = h->stream;
    double *btc = h->blk[0], *bjm = h->blk[1], *bjp = h->blk[2], *bkm = h->blk[3],
           *bkp = h->blk[4], *scr1 = h->blk[5];
    h->nt += 1;
    mark(h, 0);
    const bool sa = dp.ivis == 1 && dp.iturb == 1;
    // timing mode keeps every family on one stream so that the events bracket them
    const bool side = sa && !h->timing;
    hipStream_t ss = side ? h->stream_sa : sm;
    if (sa) {
        if (side) {
            HIP_TRY(hipEventRecord(h->ev_q, sm));             // q of the previous pass
            HIP_TRY(hipStreamWaitEvent(ss, h->ev_q, 0));
        }
        if (h->first_sa) launch_sa_first(ss, d, h->walls, h->x, h->y, h->sa);
        launch_sa_assemble(ss, d, dp, h->first_sa ? 1 : 0, h->q, h->met, h->dj,
                           (int)h->regs.size(), h->regs.data(), h->sa);
        if (side) HIP_TRY(hipEventRecord(h->ev_sa_read, ss));  // done reading q
        launch_sa_solve(ss, d, dp, h->p.ntjsp, h->wake, &h->lws_sa, h->sa,
                        h->lr_solve ? &h->lrs : nullptr);
        launch_sa_update(ss, d, dp, h->w, h->st, h->sa);
        if (side) HIP_TRY(hipEventRecord(h->ev_sa_done, ss));
        h->first_sa = false;
    }
    mark(h, 1);
    const bool f32 = h->p.precision == SMAC_FP32_LHS;
    // FP32-LHS: blocks stored (and factored, solved) in FP32 in the same allocations
    float *fb[5] = {(float *)btc, (float *)bjm, (float *)bjp, (float *)bkm, (float *)bkp};
    if (f32)
        launch_assemble<float>(sm, d, dp, h->q, h->met, h->dj, h->s, fb[0], fb[1], fb[2], fb[3],
                               fb[4]);
    else
        launch_assemble<double>(sm, d, dp, h->q, h->met, h->dj, h->s, btc, bjm, bjp, bkm, bkp);
    if (h->force_side) {   // the far-field rows read the previous pass's clift
        HIP_TRY(hipStreamWaitEvent(sm, h->ev_force, 0));
        h->force_side = false;
    }
    if (f32)
        launch_bc<float>(sm, d, dp, h->regs.data(), (int)h->regs.size(), h->st, h->x, h->y,
                         h->met, h->q, h->s, fb[0], fb[1], fb[2], fb[3], fb[4]);
    else
        launch_bc<double>(sm, d, dp, h->regs.data(), (int)h->regs.size(), h->st, h->x, h->y,
                          h->met, h->q, h->s, btc, bjm, bjp, bkm, bkp);
    mark(h, 2);
    launch_resid(sm, d, h->nt, dp.beta, h->s, h->dj, h->w, h->st);
    if (h->keep_lhs)
        for (int b = 0; b < 5; ++b)
            HIP_TRY(hipMemcpyAsync(h->lhs_copy[b], h->blk[b], 9 * d.N * 8, hipMemcpyDeviceToDevice, sm));
    mark(h, 3);
    if (h->p.impsch == 1)
        enqueue_lrelax(h, sm);
    else if (f32)
        launch_lusgs3<float>(sm, d, h->lws, h->wake, h->q, h->dq, h->s, (const float *)btc,
                             (const float *)bjm, (const float *)bjp, (const float *)bkm,
                             (const float *)bkp, nullptr, false);
    else
        launch_lusgs3<double>(sm, d, h->lws, h->wake, h->q, h->dq, h->s, btc, bjm, bjp, bkm,
                              bkp, nullptr, false);
    (void)scr1;
    mark(h, 4);
    if (side) HIP_TRY(hipStreamWaitEvent(sm, h->ev_sa_read, 0));   // q is overwritten next
    if (defer_force && !h->timing) {
        // the forces read the updated q on stream_sa (after this pass's SA), unjoined
        launch_update(sm, d, h->q, h->dq, h->w);
        HIP_TRY(hipEventRecord(h->ev_q, sm));
        HIP_TRY(hipStreamWaitEvent(h->stream_sa, h->ev_q, 0));
        launch_force(h->stream_sa, d, dp, h->walls, h->x, h->y, h->met, h->dj, h->q, h->w, h->st);
        HIP_TRY(hipEventRecord(h->ev_force, h->stream_sa));
        h->force_side = true;
    } else {
        launch_update_force(sm, d, dp, h->walls, h->x, h->y, h->met, h->dj, h->q, h->dq, h->w,
                            h->st);
        if (side) HIP_TRY(hipStreamWaitEvent(sm, h->ev_sa_done, 0));   // the pass ends with both
    }
    mark(h, 5);
    HIP_TRY(hipGetLastError());
    return SMAC_OK;
}

static void fill_stats(const smac_handle *h, const DevStats &s, smac_iter_stats *o)
{
    o->nt = h->nt;
    o->jres = s.jres; o->kres = s.kres; o->nonfinite = s.nonfinite;
    o->resmax = s.resmax; o->resmax0 = s.resmax0; o->divmax = s.divmax; o->turres = s.turres;
    o->rhsnorm = s.rhsnorm; o->cl = s.cl; o->cd = s.cd; o->cm = s.cm; o->clp = s.clp;
    o->cdp = s.cdp; o->fxp = s.fxp; o->fyp = s.fyp; o->fxs = s.fxs; o->fys = s.fys;
    o->fmom = s.fmom;
}

static void accumulate_timing(smac_handle *h)
{
    if (!h->timing) return;
    float ms;
    for (int i = 0; i < 5; ++i) {
        (void)hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]);
        h->t_acc[i] += ms;
    }
    if (h->p.impsch == 2 && hipEventElapsedTime(&ms, h->ev_band[0], h->ev_band[1]) == hipSuccess)
        h->t_acc[5] += ms;
    (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[5]);
    h->t_acc[6] += ms;
    h->t_n += 1;
}

// The statistics of the pass just enqueued, once it has run.  k_publish copies them into
// pinned memory behind the pass and the host spins on its sequence word, so the device waits
// for the host only as long as the host takes to see the word and enqueue the next pass (a
// device-to-host copy and a stream synchronisation cost about 60 us a pass here).  The spin
// asks the stream every 1024 turns, so a stream that failed or ended without the word ends it.
// Timing mode reads them after a stream synchronisation (its events are read then too).
static smac_status read_stats(smac_handle *h, DevStats &s)
{
    if (h->timing) {
        HIP_TRY(hipMemcpyAsync(&s, h->st, sizeof s, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        return SMAC_OK;
    }
    const unsigned want = ++h->pin_seq;
    launch_publish(h->stream, h->st, &h->pin_dev->s, &h->pin_dev->seq, want);
    HIP_TRY(hipGetLastError());
    const unsigned *seq = &h->pin->seq;
    for (unsigned turn = 1; __atomic_load_n(seq, __ATOMIC_ACQUIRE) != want; ++turn) {
        if ((turn & 1023) == 0) {
            const hipError_t e = hipStreamQuery(h->stream);
            if (e == hipSuccess) {
                if (__atomic_load_n(seq, __ATOMIC_ACQUIRE) == want) break;
                return fail(SMAC_ERR_HIP, "the pass statistics were not published");
            }
            if (e != hipErrorNotReady)
                return fail(SMAC_ERR_HIP, std::string("pass statistics: ") + hipGetErrorString(e));
        }
        __builtin_ia32_pause();
    }
    std::memcpy(&s, (const void *)&h->pin->s, sizeof s);
    return SMAC_OK;
}

smac_status smac_step(smac_handle *h, smac_iter_stats *st)
{
    if (!h) return fail(SMAC_ERR_ARG, "handle is NULL");
    smac_status rc = enqueue_pass(h);
    if (rc) return rc;
    DevStats s;
    rc = read_stats(h, s);
    if (rc) return rc;
    accumulate_timing(h);
    if (h->nt == 1) h->rhs0 = s.rhsnorm;
    if (st) fill_stats(h, s, st);
    if (s.stall) return fail(SMAC_ERR_STALL, "LU-SGS band hand-over timed out at nt=" + std::to_string(h->nt));
    if (s.nonfinite) return fail(SMAC_ERR_NONFINITE, "residual or solution went non-finite at nt=" + std::to_string(h->nt));
    return SMAC_OK;
}

smac_status smac_run(smac_handle *h, int niter, double drop, smac_iter_stats *last, int *done)
{
    if (!h) return fail(SMAC_ERR_ARG, "handle is NULL");
    if (niter < 0) return fail(SMAC_ERR_ARG, "niter < 0");
    if (h->timing) {
        for (double &t : h->t_acc) t = 0.0;
        h->t_n = 0;
    }
    double first = -1.0;
    DevStats s = {};
    int n = 0;
    for (; n < niter; ++n) {
        // a pass whose statistics nobody reads leaves its wall forces beside the next pass
        const bool read = drop > 0.0 || h->timing || n == niter - 1;
        smac_status rc = enqueue_pass(h, !read);
        if (rc) {
            if (h->force_side) {
                (void)hipStreamWaitEvent(h->stream, h->ev_force, 0);
                h->force_side = false;
            }
            return rc;
        }
        if (read) {
            rc = read_stats(h, s);
            if (rc) return rc;
            accumulate_timing(h);
            if (s.stall) {
                if (done) *done = n + 1;
                return fail(SMAC_ERR_STALL, "LU-SGS band hand-over timed out at nt=" + std::to_string(h->nt));
            }
            if (s.nonfinite) {
                if (done) *done = n + 1;
                if (last) fill_stats(h, s, last);
                return fail(SMAC_ERR_NONFINITE, "non-finite at nt=" + std::to_string(h->nt));
            }
            if (h->nt == 1) h->rhs0 = s.rhsnorm;
            if (first < 0.0) first = h->rhs0 > 0.0 ? h->rhs0 : s.rhsnorm;
            if (drop > 0.0 && s.rhsnorm <= drop * first) { ++n; break; }
        }
    }
    if (n == 0) {
        HIP_TRY(hipMemcpy(&s, h->st, sizeof s, hipMemcpyDeviceToHost));
    }
    if (done) *done = n;
    if (last) fill_stats(h, s, last);
    return SMAC_OK;
}

smac_status smac_get_stats(smac_handle *h, smac_iter_stats *st)
{
    if (!h || !st) return fail(SMAC_ERR_ARG, "NULL argument");
    DevStats s;
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(&s, h->st, sizeof s, hipMemcpyDeviceToHost));
    fill_stats(h, s, st);
    return SMAC_OK;
}

smac_status smac_timings(smac_handle *h, double out[7])
{
    if (!h || !out) return fail(SMAC_ERR_ARG, "NULL argument");
    for (int i = 0; i < 7; ++i) out[i] = h->t_n ? h->t_acc[i] / h->t_n : 0.0;
    return SMAC_OK;
}

smac_status smac_get_array(smac_handle *h, const char *name, double *out)
{
    if (!h || !name || !out) return fail(SMAC_ERR_ARG, "NULL argument");
    const long N = h->d.N;
    HIP_TRY(hipStreamSynchronize(h->stream));
    struct { const char *n; const double *p; long cnt; } t[] = {
        {"q", h->q, 3 * N}, {"dq", h->dq, 3 * N}, {"s", h->s, 3 * N}, {"anut", h->sa.anut, N},
        {"vnut", h->sa.vnut, N}, {"smin", h->sa.smin, N}, {"sf", h->sa.sf, N},
        {"stc", h->sa.stc, N}, {"ds", h->sa.ds, N}, {"vort", h->sa.vort, N},
        {"btc", h->lhs_copy[0], 9 * N}, {"bjm", h->lhs_copy[1], 9 * N},
        {"bjp", h->lhs_copy[2], 9 * N}, {"bkm", h->lhs_copy[3], 9 * N},
        {"bkp", h->lhs_copy[4], 9 * N},
        {"adj_jac", h->adj_ready ? h->adj.jac : nullptr, 81 * N},
        {"adj_lam_cl", h->adj_ready ? h->adj.lam[0] : nullptr, 3 * N},
        {"adj_lam_cd", h->adj_ready ? h->adj.lam[1] : nullptr, 3 * N},
        {"adj_gcl", h->adj_ready ? h->adj.g[0] : nullptr, 3 * N},
        {"adj_gcd", h->adj_ready ? h->adj.g[1] : nullptr, 3 * N},
    };
    for (auto &e : t)
        if (!std::strcmp(e.n, name)) {
            if (!e.p)
                return fail(SMAC_ERR_ARG, std::string(name) + (name[0] == 'a'
                                                   ? " exists after a smac_adjoint call"
                                                   : " is kept only with option keep_lhs"));
            HIP_TRY(hipMemcpy(out, e.p, e.cnt * 8, hipMemcpyDeviceToHost));
            return SMAC_OK;
        }
    return fail(SMAC_ERR_ARG, std::string("unknown array ") + name);
}

}  // extern "C"

// ---- discrete adjoint ------------------------------------------------------------------

// row types of the steady equations (the BC rows in file order, then the wake-cut targets)
static smac_status adj_row_types(smac_handle *h, std::vector<signed char> &rt,
                                 std::vector<int> &inv)
{
    const int jm = h->d.jm, km = h->d.km;
    rt.assign(h->d.N, (signed char)-1);
    for (int k = 1; k < km - 1; ++k)
        for (int j = 1; j < jm - 1; ++j) rt[(long)j + (long)jm * k] = 0;
    for (const BcRegion &r : h->regs)
        for (int k = r.kbeg; k <= r.kend; ++k)
            for (int j = r.jbeg; j <= r.jend; ++j) {
                const long p = (long)j + (long)jm * k;
                if (r.ibcval == 31) rt[p] = j == 0 ? 1 : 2;
                else if (r.ibcval == 0) rt[p] = 3;
                else if (r.ibcval == 25) rt[p] = 4;
            }
    const int np = h->wake.n;
    std::vector<int> pr(2 * np);
    if (np > 0 && hipMemcpy(pr.data(), h->pairs, 2 * np * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(SMAC_ERR_HIP, "reading the wake pairs failed");
    inv.assign(np, -1);
    for (int i = 0; i < np; ++i) {
        const long p = pr[2 * i];
        if (rt[p] != -1)
            return fail(SMAC_ERR_BC, "a wake-cut point also carries a BC row: not a steady system the adjoint covers");
        rt[p] = 5;
        for (int i2 = 0; i2 < np; ++i2)
            if (pr[2 * i2 + 1] == pr[2 * i]) inv[i] = i2;
        if (inv[i] < 0) return fail(SMAC_ERR_BC, "wake cut: a target that is no pair's partner");
    }
    for (long p = 0; p < h->d.N; ++p)
        if (rt[p] < 0) return fail(SMAC_ERR_BC, "a boundary point without a BC row or wake-cut relation");
    return SMAC_OK;
}

static smac_status adj_prepare(smac_handle *h, int m)
{
    const long N = h->d.N;
    if (!h->adj_ready || h->adj_m != m) {
        AdjCtx &c = h->adj;
        c = AdjCtx{};
        std::vector<signed char> rt;
        std::vector<int> inv;
        smac_status rc = adj_row_types(h, rt, inv);
        if (rc) return rc;
        signed char *rtd = nullptr;
        int *invd = nullptr;
        if ((rc = dalloc(h, &rtd, N)) || (rc = dalloc(h, &invd, inv.size() + 1))) return rc;
        HIP_TRY(hipMemcpy(rtd, rt.data(), N, hipMemcpyHostToDevice));
        if (!inv.empty()) HIP_TRY(hipMemcpy(invd, inv.data(), inv.size() * sizeof(int), hipMemcpyHostToDevice));
        c.rtype = rtd;
        c.inv = invd;
        double **v3[] = {&c.lam[0], &c.lam[1], &c.g[0], &c.g[1], &c.fc, &c.fa, &c.w, &c.z,
                         &c.t1, &c.t2, &c.t1b, &c.fb, &c.smu, &c.pk1, &c.pk2};
        if ((rc = dalloc(h, &c.jac, 81 * N))) return rc;
        for (int b = 0; b < 5; ++b)
            if ((rc = dalloc(h, &c.tb[b], 9 * N))) return rc;
        for (double **pp : v3)
            if ((rc = dalloc(h, pp, 3 * N))) return rc;
        if ((rc = dalloc(h, &c.V, 3 * N * (m + 1)))) return rc;
        // the CD adjoint runs beside the CL adjoint with its own work vectors
        AdjCtx &c2 = h->adj2;
        c2 = AdjCtx{};
        double **w3[] = {&c2.w, &c2.z, &c2.t1, &c2.t2, &c2.t1b, &c2.pk1, &c2.pk2};
        for (double **pp : w3)
            if ((rc = dalloc(h, pp, 3 * N))) return rc;
        if ((rc = dalloc(h, &c2.V, 3 * N * (m + 1))) || (rc = dalloc(h, &c2.part, 512 * (m + 2))) ||
            (rc = dalloc(h, &c2.red, m + 2)) || (rc = dalloc(h, &c2.red2, m + 2)) ||
            (rc = dalloc(h, &c2.gate, 1)))
            return rc;
        if ((rc = dalloc(h, &c.dmet, 4 * N)) || (rc = dalloc(h, &c.ddj, N)) ||
            (rc = dalloc(h, &c.dxy, 2 * N)) || (rc = dalloc(h, &c.fpart, 2 * h->d.jm)) ||
            (rc = dalloc(h, &c.part, 512 * (m + 2))) || (rc = dalloc(h, &c.red, m + 2)) ||
            (rc = dalloc(h, &c.red2, m + 2)) || (rc = dalloc(h, &c.gate, 1)))
            return rc;
        c.lr = h->p.impsch == 1;
        if (c.lr) {
            // the factored lines of M and M^T (FP64) and the transposed sweeps' vectors
            char *f = nullptr, *t = nullptr;
            if ((rc = dalloc(h, &f, 2 * 45 * N * 8)) || (rc = dalloc(h, &t, 2 * 45 * N * 8)))
                return rc;
            c.lr_f.opx = f;
            c.lr_f.ope = f + 45 * N * 8;
            c.lr_t.opx = t;
            c.lr_t.ope = t + 45 * N * 8;
            double *sv = nullptr;
            if ((rc = dalloc(h, &sv, 6 * N))) return rc;
            c.lr_f.s_t = c.lr_t.s_t = sv;
            c.lr_f.dq_t = c.lr_t.dq_t = sv + 3 * N;
            double **lv[] = {&c.lrv.a, &c.lrv.z, &c.lrv.at, &c.lrv.st, &c.lrv.zt,
                             &c2.lrv.a, &c2.lrv.z, &c2.lrv.at, &c2.lrv.st, &c2.lrv.zt};
            for (double **pp : lv)
                if ((rc = dalloc(h, pp, 3 * N))) return rc;
        }
        h->adj_m = m;
        h->adj_ready = true;
    }
    // the LU-SGS preconditioner's workspaces (one per solve; a change of lusgs_rows frees them)
    if (h->p.impsch != 1)
        for (LusgsWs *ws : {&h->adj_ws, &h->adj_ws2})
            if (!ws->lhs) {
                ws->ring = 1;
                smac_status rc = alloc_one_ws(h, 64, 3, *ws);
                if (rc) return rc;
            }
    AdjCtx &c = h->adj;
    c.d = h->d;
    c.sm = h->stream;
    c.x = h->x; c.y = h->y; c.met = h->met; c.dj = h->dj; c.q = h->q;
    for (int b = 0; b < 5; ++b) c.blk[b] = h->blk[b];
    c.ws = h->adj_ws;
    c.walls = h->walls;
    c.beta = h->dp.beta; c.cdis = h->dp.cdis; c.vnu = h->dp.vnu; c.pout = h->dp.pout;
    c.alpha = h->dp.alpha; c.chord = h->dp.chord; c.ivis = h->dp.ivis;
    c.m = m;
    c.njsp = h->p.njsp;
    c.nksp = h->p.nksp;
    c.dp = h->dp;
    c.npair = h->wake.n;
    c.pairs = h->wake.pairs;
    return SMAC_OK;
}

extern "C" {

smac_status smac_adjoint(smac_handle *h, int nshape, const double *dx, const double *dy,
                         int maxit, double tol, double *dcl, double *dcd, smac_adjoint_stats *st)
{
    if (!h || !dcl || !dcd) return fail(SMAC_ERR_ARG, "NULL argument");
    if (nshape < 0 || (nshape > 0 && (!dx || !dy))) return fail(SMAC_ERR_ARG, "shape directions missing");
    if (maxit < 1 || !(tol > 0.0)) return fail(SMAC_ERR_ARG, "maxit must be >= 1 and tol > 0");
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream_sa));
    const int m = h->adj_restart;
    smac_status rc = adj_prepare(h, m);
    if (rc) return rc;
    const Dims &d = h->d;
    // CL, CD and the far field's circulation of this state (after a run, the same values its
    // last pass left; after smac_set_state, none has been computed for it yet)
    launch_force(h->stream, d, h->dp, h->walls, h->x, h->y, h->met, h->dj, h->q, h->w, h->st);
    // the FP64 residual and blocks of the state (the iteration's own operator M)
    launch_assemble<double>(h->stream, d, h->dp, h->q, h->met, h->dj, h->s, h->blk[0], h->blk[1],
                            h->blk[2], h->blk[3], h->blk[4]);
    launch_bc<double>(h->stream, d, h->dp, h->regs.data(), (int)h->regs.size(), h->st, h->x, h->y,
                      h->met, h->q, h->s, h->blk[0], h->blk[1], h->blk[2], h->blk[3], h->blk[4]);
    launch_resid(h->stream, d, h->nt > 1 ? h->nt : 2, h->dp.beta, h->s, h->dj, h->w, h->st);   // no resmax0 latch
    DevStats ds;
    HIP_TRY(hipMemcpyAsync(&ds, h->st, sizeof ds, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    AdjCtx &c = h->adj;
    c.reorth = h->adj_reorth * 1e-3;
    c.psteps = h->adj_prec_steps;
    c.clift = ds.clift;
    c.cl = ds.cl;
    c.cd = ds.cd;
    adj_build(c);
    double rel[2] = {0.0, 0.0};
    int its[2] = {0, 0};
    // The two solves run side by side, CL's on the handle's stream and CD's on its second
    // stream from a host thread of its own: each GMRES iteration waits for its Hessenberg
    // column on the host, and the other solve's kernels fill that gap.  They share the
    // operator and the preconditioner's factors (read only) and have their own work vectors.
    AdjCtx c2 = c;
    {
        const AdjCtx &b2 = h->adj2;
        c2.sm = h->stream_sa;
        c2.V = b2.V; c2.w = b2.w; c2.z = b2.z; c2.t1 = b2.t1; c2.t2 = b2.t2; c2.t1b = b2.t1b;
        c2.pk1 = b2.pk1; c2.pk2 = b2.pk2; c2.part = b2.part; c2.red = b2.red;
        c2.red2 = b2.red2; c2.gate = b2.gate; c2.lrv = b2.lrv;
        c2.ws = h->adj_ws2;
    }
    HIP_TRY(hipStreamSynchronize(h->stream));   // the operator and factors are built
    std::thread cd([&]() { its[1] = adj_gmres(c2, c.g[1], c.lam[1], maxit, tol, &rel[1]); });
    its[0] = adj_gmres(c, c.g[0], c.lam[0], maxit, tol, &rel[0]);
    cd.join();
    HIP_TRY(hipStreamSynchronize(h->stream_sa));
    adj_gradients(c, nshape, dx, dy, dcl, dcd);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (st) {
        st->iters[0] = its[0];
        st->iters[1] = its[1];
        st->resid[0] = rel[0];
        st->resid[1] = rel[1];
        st->rhsnorm = ds.rhsnorm;
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (!(std::isfinite(dcl[0]) && std::isfinite(dcd[0])))
        return fail(SMAC_ERR_NONFINITE, "adjoint gradient is not finite");
    for (int o = 0; o < 2; ++o)
        if (its[o] >= maxit && rel[o] > 10.0 * tol)
            return fail(SMAC_ERR_NOCONV, std::string(o ? "CD" : "CL") + " adjoint: GMRES stopped at " +
                                             std::to_string(maxit) + " iterations, residual " +
                                             std::to_string(rel[o]));
    return SMAC_OK;
}

smac_status smac_adjoint_apply(smac_handle *h, int op, const double *in, double *out)
{
    if (!h || !in || !out) return fail(SMAC_ERR_ARG, "NULL argument");
    if (!h->adj_ready) return fail(SMAC_ERR_ARG, "no adjoint context: call smac_adjoint first");
    if (op < 0 || op > 3) return fail(SMAC_ERR_ARG, "op must be 0..3");
    AdjCtx &c = h->adj;
    const long n3 = 3 * h->d.N;
    HIP_TRY(hipMemcpy(c.w, in, n3 * 8, hipMemcpyHostToDevice));
    if (op == 0) adj_apply_jt(c, c.w, c.z);
    else if (op == 1) adj_apply_op(c, c.w, c.z);
    else if (op == 2) adj_apply_prec_t(c, c.w, c.z);
    else adj_apply_prec(c, c.w, c.z);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c.sm));
    HIP_TRY(hipMemcpy(out, c.z, n3 * 8, hipMemcpyDeviceToHost));
    return SMAC_OK;
}

}  // extern "C"

// ---- batches ------------------------------------------------------------------------------
//
// A batch advances its cases together on one stream.  The LU-SGS band sweeps, the part of a
// pass that occupies only a few wavefronts per case and is bound by the latency of its
// hyperplane steps, run for every case of the batch in one launch (k_band_multi), so that the
// bands of all cases sweep side by side on the device; the fully parallel kernels of each case
// are enqueued in between.  Every case sees exactly the kernels and operands of its own pass,
// so its results are those of smac_step on it alone.

struct BandGroup {
    std::vector<int> members;        // indices into cases
    BandJob *jobs = nullptr;         // device
    int2 *map = nullptr;             // device
    int nblk = 0, pf = 3;
};

// streams over which the per-case kernels of a batch phase are spread
constexpr int BATCH_LANES = 8;

struct smac_batch {
    std::vector<smac_handle *> cases;
    hipStream_t stream = nullptr;                 // band launches; phases fork from and join it
    hipStream_t lane[BATCH_LANES] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_band[2] = {};      // around the forward mean-flow band launch of a run's last pass
    int band_group = -1;             // the group timed (the one with the most points)
    double band_ms = 0.0, band_pts = 0.0;
    hipEvent_t ev_join[BATCH_LANES] = {};
    // multi-case passes: the SA launches run on lane[0] beside the mean flow's on `stream`
    hipEvent_t ev_sa_read = nullptr, ev_sa_done = nullptr;
    BandGroup sa;                    // SA LU-SGS (neqs 1, FP64) of the LU-SGS cases
    BandGroup mean[2];               // mean-flow LU-SGS: FP64, FP32-LHS
    std::vector<LusgsWs> snap_mean, snap_sa;   // workspaces the tables were built from
    // the multi-case form of the per-case kernels (every case LU-SGS; passes after the first)
    bool multi_ok = false;
    MultiTab<SaDerivArgs> sa_deriv;
    MultiTab<SaAsmArgs> sa_asm;
    MultiTab<SaBcArgs> sa_bc;
    MultiTab<ShearArgs> sa_shear;
    MultiTab<MidArgs> sa_mid;
    MultiTab<UnshearArgs> sa_unshear;
    MultiTab<WakeArgs> sa_wake;
    MultiTab<SaUpdArgs> sa_upd;
    MultiTab<SaTurArgs> sa_tur;
    MultiTab<AsmArgs> asm_[2];       // [0] FP64 cases, [1] FP32-LHS cases
    MultiTab<BcArgs> bc[2];
    MultiTab<ResidPArgs> resp;
    MultiTab<ResidFArgs> resf;
    MultiTab<ShearArgs> shear[2];
    MultiTab<MidArgs> mid[2];
    MultiTab<UnshearArgs> unshear[2];
    MultiTab<WakeArgs> wake;
    MultiTab<UpdArgs> upd;
    MultiTab<ForceArgs> force;
    // device-side stop (smac_batch_run_until): per case, whether it still runs, the passes it
    // made, its reference norm, and its statistics record
    int *active = nullptr;
    int *passes = nullptr;
    double *first = nullptr;
    DevStats **stp = nullptr;
};

template <typename A>
static void mt_free(MultiTab<A> &t)
{
    if (t.args) (void)hipFree(t.args);
    if (t.start) (void)hipFree(t.start);
    if (t.item_case) (void)hipFree(t.item_case);
    t = MultiTab<A>{};
}

// the items of one multi-case kernel and the batch case each belongs to
template <typename A>
struct Items {
    std::vector<A> v;
    std::vector<int> c;
    void tag(int i) { c.resize(v.size(), i); }   // items added since the last tag are case i's
};

// a device table of the items with a non-empty grid; `active` is the batch's per-case flags
template <typename A>
static smac_status mt_build(MultiTab<A> &t, const Items<A> &items, const int *active)
{
    mt_free(t);
    std::vector<A> keep;
    std::vector<int> kcase;
    for (size_t i = 0; i < items.v.size(); ++i)
        if (items.v[i].gx > 0 && items.v[i].gy > 0) {
            keep.push_back(items.v[i]);
            kcase.push_back(items.c[i]);
        }
    std::vector<int> start(keep.size() + 1, 0);
    for (size_t i = 0; i < keep.size(); ++i)
        start[i + 1] = start[i] + (int)(keep[i].gx * keep[i].gy);
    t.n = (int)keep.size();
    t.nblk = start.back();
    if (t.n == 0) return SMAC_OK;
    if (hipMalloc(&t.args, keep.size() * sizeof(A)) != hipSuccess ||
        hipMalloc(&t.start, start.size() * sizeof(int)) != hipSuccess ||
        hipMalloc(&t.item_case, kcase.size() * sizeof(int)) != hipSuccess)
        return fail(SMAC_ERR_HIP, "hipMalloc failed (batch kernel tables)");
    HIP_TRY(hipMemcpy(t.args, keep.data(), keep.size() * sizeof(A), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.start, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.item_case, kcase.data(), kcase.size() * sizeof(int), hipMemcpyHostToDevice));
    t.active = active;
    return SMAC_OK;
}

static void free_multi(smac_batch *b)
{
    mt_free(b->sa_deriv); mt_free(b->sa_asm); mt_free(b->sa_bc); mt_free(b->sa_shear);
    mt_free(b->sa_mid); mt_free(b->sa_unshear); mt_free(b->sa_wake); mt_free(b->sa_upd);
    mt_free(b->sa_tur); mt_free(b->resp); mt_free(b->resf); mt_free(b->wake); mt_free(b->upd);
    mt_free(b->force);
    for (int g = 0; g < 2; ++g) {
        mt_free(b->asm_[g]); mt_free(b->bc[g]); mt_free(b->shear[g]); mt_free(b->mid[g]);
        mt_free(b->unshear[g]);
    }
}

#define B_TRY(x)                   \
    do {                           \
        smac_status r_ = (x);      \
        if (r_) return r_;         \
    } while (0)

static bool is_lusgs(const smac_handle *h) { return h->p.impsch == 2; }
static bool has_sa(const smac_handle *h) { return h->dp.ivis == 1 && h->dp.iturb == 1; }

// frees the device tables (the member list stays)
static void free_group(BandGroup &g)
{
    if (g.jobs) (void)hipFree(g.jobs);
    if (g.map) (void)hipFree(g.map);
    g.jobs = nullptr;
    g.map = nullptr;
    g.nblk = 0;
}

static void free_batch(smac_batch *b)
{
    if (!b) return;
    free_multi(b);
    free_group(b->sa);
    free_group(b->mean[0]);
    free_group(b->mean[1]);
    for (int k = 0; k < BATCH_LANES; ++k) {
        if (b->lane[k]) (void)hipStreamDestroy(b->lane[k]);
        if (b->ev_join[k]) (void)hipEventDestroy(b->ev_join[k]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    for (hipEvent_t e : {b->ev_sa_read, b->ev_sa_done})
        if (e) (void)hipEventDestroy(e);
    for (auto &e : b->ev_band)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (void *p : {(void *)b->active, (void *)b->passes, (void *)b->first, (void *)b->stp})
        if (p) (void)hipFree(p);
    delete b;
}

static bool same_ws(const LusgsWs &a, const LusgsWs &c)
{
    return a.lhs == c.lhs && a.r == c.r && a.dq == c.dq && a.edge == c.edge &&
           a.nband == c.nband && a.rows == c.rows && a.pf3 == c.pf3 && a.pf1 == c.pf1 &&
           a.ring == c.ring;
}

static smac_status build_group(BandGroup &g, const std::vector<BandJob> &jobs, int pf)
{
    free_group(g);
    if (g.members.size() != jobs.size()) return fail(SMAC_ERR_ARG, "batch table mismatch");
    std::vector<int2> map;
    for (size_t i = 0; i < jobs.size(); ++i)
        for (int blk = 0; blk < jobs[i].w.nband; ++blk) map.push_back(make_int2((int)i, blk));
    g.nblk = (int)map.size();
    g.pf = pf;
    if (jobs.empty()) return SMAC_OK;
    if (hipMalloc(&g.jobs, jobs.size() * sizeof(BandJob)) != hipSuccess ||
        hipMalloc(&g.map, map.size() * sizeof(int2)) != hipSuccess)
        return fail(SMAC_ERR_HIP, "hipMalloc failed (batch band tables)");
    HIP_TRY(hipMemcpy(g.jobs, jobs.data(), jobs.size() * sizeof(BandJob), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g.map, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice));
    return SMAC_OK;
}

// The multi-case kernel tables: each item's arguments as its single-case launcher forms them
// for a pass after the first (nt > 1, the SA's wall distance known).
static smac_status build_multi(smac_batch *b)
{
    free_multi(b);
    b->multi_ok = true;
    for (smac_handle *h : b->cases) b->multi_ok = b->multi_ok && is_lusgs(h);
    if (!b->multi_ok) return SMAC_OK;
    Items<SaDerivArgs> sd;
    Items<SaAsmArgs> sas;
    Items<SaBcArgs> sbc;
    Items<ShearArgs> ssh, sh[2];
    Items<MidArgs> smid, mid[2];
    Items<UnshearArgs> sun, un[2];
    Items<WakeArgs> swk, wk;
    Items<SaUpdArgs> sup;
    Items<SaTurArgs> stu;
    Items<AsmArgs> as[2];
    Items<BcArgs> bc[2];
    Items<ResidPArgs> rp;
    Items<ResidFArgs> rf;
    Items<UpdArgs> up;
    Items<ForceArgs> fo;
    for (size_t ci = 0; ci < b->cases.size(); ++ci) {
        smac_handle *h = b->cases[ci];
        const Dims &d = h->d;
        const DevParams &dp = h->dp;
        if (has_sa(h)) {
            SaArrays &a = h->sa;
            sd.v.push_back(sa_deriv_args(d, h->q, h->met, h->dj, a));
            sas.v.push_back(sa_asm_args(d, dp, 0, h->q, h->met, h->dj, a));
            sa_bc_args(d, (int)h->regs.size(), h->regs.data(), a, sbc.v);
            ssh.v.push_back(shear_args<double>(d, h->lws_sa, a.sf, a.stc, a.sjm, a.sjp, a.skm,
                                             a.skp, (double *)nullptr));
            smid.v.push_back(mid_args<double, 1>(d, h->lws_sa));
            sun.v.push_back(unshear_args(d, h->lws_sa, h->lws_sa.dq, a.ds));
            swk.v.push_back(wake_args(d, 1, h->wake, a.anut, a.ds, a.sf));
            sup.v.push_back(sa_upd_args(d, dp, h->w, a));
            stu.v.push_back(sa_tur_args(h->w, h->st));
        }
        const int g = h->p.precision == SMAC_FP32_LHS ? 1 : 0;
        if (g) {
            float *fb[5] = {(float *)h->blk[0], (float *)h->blk[1], (float *)h->blk[2],
                            (float *)h->blk[3], (float *)h->blk[4]};
            as[g].v.push_back(asm_args<float>(d, dp, h->q, h->met, h->dj, h->s, fb[0], fb[1], fb[2],
                                            fb[3], fb[4]));
            bc_args<float>(d, dp, h->regs.data(), (int)h->regs.size(), h->st, h->x, h->y, h->met,
                           h->q, h->s, fb[0], fb[1], fb[2], fb[3], fb[4], bc[g].v);
            sh[g].v.push_back(shear_args<float>(d, h->lws, h->s, fb[0], fb[1], fb[2], fb[3], fb[4],
                                              (float *)nullptr));
            mid[g].v.push_back(mid_args<float, 3>(d, h->lws));
        } else {
            double **db = h->blk;
            as[g].v.push_back(asm_args<double>(d, dp, h->q, h->met, h->dj, h->s, db[0], db[1],
                                             db[2], db[3], db[4]));
            bc_args<double>(d, dp, h->regs.data(), (int)h->regs.size(), h->st, h->x, h->y,
                            h->met, h->q, h->s, db[0], db[1], db[2], db[3], db[4], bc[g].v);
            sh[g].v.push_back(shear_args<double>(d, h->lws, h->s, db[0], db[1], db[2], db[3], db[4],
                                               (double *)nullptr));
            mid[g].v.push_back(mid_args<double, 3>(d, h->lws));
        }
        un[g].v.push_back(unshear_args(d, h->lws, h->lws.dq, h->dq));
        rp.v.push_back(residp_args(d, h->s, h->dj, h->w));
        rf.v.push_back(residf_args(d, 2, dp.beta, h->s, h->dj, h->w, h->st));   // nt > 1
        wk.v.push_back(wake_args(d, 3, h->wake, h->q, h->dq, h->s));
        up.v.push_back(upd_args(d, h->q, h->dq, h->w));
        fo.v.push_back(force_args(d, dp, h->walls, h->x, h->y, h->met, h->dj, h->q, h->w, h->st));
        const int i = (int)ci;
        sd.tag(i); sas.tag(i); sbc.tag(i); ssh.tag(i); smid.tag(i); sun.tag(i); swk.tag(i);
        sup.tag(i); stu.tag(i); rp.tag(i); rf.tag(i); wk.tag(i); up.tag(i); fo.tag(i);
        for (int gg = 0; gg < 2; ++gg) {
            as[gg].tag(i); bc[gg].tag(i); sh[gg].tag(i); mid[gg].tag(i); un[gg].tag(i);
        }
    }
    B_TRY(mt_build(b->sa_deriv, sd, b->active));
    B_TRY(mt_build(b->sa_asm, sas, b->active));
    B_TRY(mt_build(b->sa_bc, sbc, b->active));
    B_TRY(mt_build(b->sa_shear, ssh, b->active));
    B_TRY(mt_build(b->sa_mid, smid, b->active));
    B_TRY(mt_build(b->sa_unshear, sun, b->active));
    B_TRY(mt_build(b->sa_wake, swk, b->active));
    B_TRY(mt_build(b->sa_upd, sup, b->active));
    B_TRY(mt_build(b->sa_tur, stu, b->active));
    for (int g = 0; g < 2; ++g) {
        B_TRY(mt_build(b->asm_[g], as[g], b->active));
        B_TRY(mt_build(b->bc[g], bc[g], b->active));
        B_TRY(mt_build(b->shear[g], sh[g], b->active));
        B_TRY(mt_build(b->mid[g], mid[g], b->active));
        B_TRY(mt_build(b->unshear[g], un[g], b->active));
    }
    B_TRY(mt_build(b->resp, rp, b->active));
    B_TRY(mt_build(b->resf, rf, b->active));
    B_TRY(mt_build(b->wake, wk, b->active));
    B_TRY(mt_build(b->upd, up, b->active));
    B_TRY(mt_build(b->force, fo, b->active));
    return SMAC_OK;
}

// (Re)builds the multi-case band tables from the cases' current workspaces.
static smac_status build_tables(smac_batch *b)
{
    std::vector<BandJob> jsa, jm[2];
    b->sa.members.clear();
    b->mean[0].members.clear();
    b->mean[1].members.clear();
    b->snap_mean.clear();
    b->snap_sa.clear();
    int pf_sa = 0, pf_m[2] = {0, 0};
    for (size_t i = 0; i < b->cases.size(); ++i) {
        smac_handle *h = b->cases[i];
        b->snap_mean.push_back(h->lws);
        b->snap_sa.push_back(h->lws_sa);
        if (!is_lusgs(h)) continue;
        LusgsWs w = h->lws;
        w.ev_fwd[0] = w.ev_fwd[1] = nullptr;
        const int g = h->p.precision == SMAC_FP32_LHS ? 1 : 0;
        b->mean[g].members.push_back((int)i);
        jm[g].push_back(BandJob{h->d, w, (int)i});
        if (!pf_m[g]) pf_m[g] = w.ring ? -1 : w.pf3;
        if (has_sa(h)) {
            LusgsWs ws = h->lws_sa;
            ws.ev_fwd[0] = ws.ev_fwd[1] = nullptr;
            b->sa.members.push_back((int)i);
            jsa.push_back(BandJob{h->d, ws, (int)i});
            if (!pf_sa) pf_sa = ws.pf1;
        }
    }
    // the sweep form (LDS ring or register prefetch) and prefetch distance of a multi-case
    // launch are those of its first case (every form and distance gives the same results)
    B_TRY(build_group(b->sa, jsa, pf_sa ? pf_sa : 8));
    // pf 0 selects the LDS-ring sweeps
    B_TRY(build_group(b->mean[0], jm[0], pf_m[0] < 0 ? 0 : (pf_m[0] ? pf_m[0] : 4)));
    B_TRY(build_group(b->mean[1], jm[1], pf_m[1] < 0 ? 0 : (pf_m[1] ? pf_m[1] : 4)));
    return build_multi(b);
}

static smac_status check_tables(smac_batch *b)
{
    for (size_t i = 0; i < b->cases.size(); ++i)
        if (!same_ws(b->snap_mean[i], b->cases[i]->lws) ||
            !same_ws(b->snap_sa[i], b->cases[i]->lws_sa)) {
            HIP_TRY(hipStreamSynchronize(b->stream));
            return build_tables(b);
        }
    return SMAC_OK;
}

// Phase boundaries: the lanes start after the batch stream's work so far (fork), and the
// batch stream continues after all of theirs (join).
static smac_status fork_lanes(smac_batch *b)
{
    HIP_TRY(hipEventRecord(b->ev_fork, b->stream));
    for (int k = 0; k < BATCH_LANES; ++k) HIP_TRY(hipStreamWaitEvent(b->lane[k], b->ev_fork, 0));
    return SMAC_OK;
}

static smac_status join_lanes(smac_batch *b)
{
    for (int k = 0; k < BATCH_LANES; ++k) {
        HIP_TRY(hipEventRecord(b->ev_join[k], b->lane[k]));
        HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_join[k], 0));
    }
    return SMAC_OK;
}

// One pass of every case with each kernel of the pass launched once for all cases (k_multi),
// in the order of enqueue_pass, on the batch stream: every case gets the kernels and
// operands of its own pass, so it stays bit-identical to being stepped alone.  For batches of
// LU-SGS cases past their first pass.
static smac_status enqueue_batch_pass_multi(smac_batch *b, bool time_band)
{
    hipStream_t sm = b->stream;
    for (smac_handle *h : b->cases) h->nt += 1;
    // The SA runs on a second stream beside the mean flow, as in a single case's pass (the
    // mean flow does not read the SA's results within a pass): it starts from the previous
    // pass's q, and the mean flow's update, which overwrites q, waits until the SA has read it.
    hipStream_t ss = b->lane[0];
    HIP_TRY(hipEventRecord(b->ev_fork, sm));
    HIP_TRY(hipStreamWaitEvent(ss, b->ev_fork, 0));
    multi_sa_deriv(ss, b->sa_deriv);
    multi_sa_assemble(ss, b->sa_asm);
    multi_sa_bc(ss, b->sa_bc);
    HIP_TRY(hipEventRecord(b->ev_sa_read, ss));   // the SA's last read of q
    multi_shear<double, 1>(ss, b->sa_shear);
    BandGroup &gs = b->sa;
    if (gs.nblk) {
        lusgs_band_multi<double, 1>(ss, gs.jobs, gs.map, gs.nblk, gs.pf, true, b->active);
        multi_mid<double, 1>(ss, b->sa_mid);
        lusgs_band_multi<double, 1>(ss, gs.jobs, gs.map, gs.nblk, gs.pf, false, b->active);
    }
    multi_unshear<double, 1>(ss, b->sa_unshear);
    multi_wake(ss, b->sa_wake);
    multi_sa_update(ss, b->sa_upd);
    multi_sa_turres(ss, b->sa_tur);
    HIP_TRY(hipEventRecord(b->ev_sa_done, ss));
    // mean flow up to its band sweeps
    multi_assemble<double>(sm, b->asm_[0]);
    multi_assemble<float>(sm, b->asm_[1]);
    multi_bc<double>(sm, b->bc[0]);
    multi_bc<float>(sm, b->bc[1]);
    multi_resid_partial(sm, b->resp);
    multi_resid_final(sm, b->resf);
    multi_shear<double, 3>(sm, b->shear[0]);
    multi_shear<float, 3>(sm, b->shear[1]);
    for (int g = 0; g < 2; ++g) {
        BandGroup &gm = b->mean[g];
        if (!gm.nblk) continue;
        const bool rec = time_band && g == b->band_group;
        if (rec) HIP_TRY(hipEventRecord(b->ev_band[0], sm));
        if (g == 0)
            lusgs_band_multi<double, 3>(sm, gm.jobs, gm.map, gm.nblk, gm.pf, true, b->active);
        else
            lusgs_band_multi<float, 3>(sm, gm.jobs, gm.map, gm.nblk, gm.pf, true, b->active);
        if (rec) HIP_TRY(hipEventRecord(b->ev_band[1], sm));
        if (g == 0) {
            multi_mid<double, 3>(sm, b->mid[0]);
            lusgs_band_multi<double, 3>(sm, gm.jobs, gm.map, gm.nblk, gm.pf, false, b->active);
        } else {
            lusgs_band_multi<float, 3>(sm, gm.jobs, gm.map, gm.nblk, gm.pf, false, b->active);
        }
    }
    multi_unshear<double, 3>(sm, b->unshear[0]);
    multi_unshear<float, 3>(sm, b->unshear[1]);
    multi_wake(sm, b->wake);
    HIP_TRY(hipStreamWaitEvent(sm, b->ev_sa_read, 0));   // q is overwritten next
    multi_update(sm, b->upd);
    multi_force(sm, b->force);
    HIP_TRY(hipStreamWaitEvent(sm, b->ev_sa_done, 0));   // the pass ends with both
    HIP_TRY(hipGetLastError());
    return SMAC_OK;
}

// one pass of every case, in the order of enqueue_pass for each case.  Case i's kernels of a
// phase run on lane i % BATCH_LANES; the band sweeps of all cases on the batch stream.
static smac_status enqueue_batch_pass(smac_batch *b, bool time_band,
                                      const std::vector<char> *on = nullptr)
{
    if (b->multi_ok) {
        bool later = true;
        for (smac_handle *h : b->cases) later = later && h->nt >= 1 && !(has_sa(h) && h->first_sa);
        if (later && !on) return enqueue_batch_pass_multi(b, time_band);
    }
    hipStream_t sm = b->stream;
    auto lane_of = [&](int i) { return b->lane[i % BATCH_LANES]; };
    // cases that stopped (smac_batch_run_until with a host-side test) get no kernels; their
    // bands leave the multi-case sweeps at once on the device flags, which mirror `on`
    auto runs = [&](int i) { return !on || (*on)[i] != 0; };
    const int *act = on ? b->active : nullptr;
    const int nc = (int)b->cases.size();
    // SA up to its band sweeps (line-relaxation cases: the whole SA step)
    B_TRY(fork_lanes(b));
    for (int i = 0; i < nc; ++i) {
        if (!runs(i)) continue;
        smac_handle *h = b->cases[i];
        hipStream_t ls = lane_of(i);
        h->nt += 1;
        if (!has_sa(h)) continue;
        const Dims &d = h->d;
        SaArrays &a = h->sa;
        if (h->first_sa) launch_sa_first(ls, d, h->walls, h->x, h->y, a);
        launch_sa_assemble(ls, d, h->dp, h->first_sa ? 1 : 0, h->q, h->met, h->dj,
                           (int)h->regs.size(), h->regs.data(), a);
        h->first_sa = false;
        if (is_lusgs(h)) {
            lusgs_pre<double, 1>(ls, d, h->lws_sa, a.sf, a.stc, a.sjm, a.sjp, a.skm, a.skp, nullptr);
        } else {
            launch_sa_solve(ls, d, h->dp, h->p.ntjsp, h->wake, &h->lws_sa, a,
                            h->lr_solve ? &h->lrs : nullptr);
            launch_sa_update(ls, d, h->dp, h->w, h->st, a);
        }
    }
    B_TRY(join_lanes(b));
    BandGroup &gs = b->sa;
    if (gs.nblk) {
        lusgs_band_multi<double, 1>(sm, gs.jobs, gs.map, gs.nblk, gs.pf, true, act);
        B_TRY(fork_lanes(b));
        for (size_t n = 0; n < gs.members.size(); ++n) {
            if (!runs(gs.members[n])) continue;
            smac_handle *h = b->cases[gs.members[n]];
            lusgs_mid<double, 1>(lane_of(gs.members[n]), h->d, h->lws_sa);
        }
        B_TRY(join_lanes(b));
        lusgs_band_multi<double, 1>(sm, gs.jobs, gs.map, gs.nblk, gs.pf, false, act);
    }
    // SA after its sweeps; mean flow up to its band sweeps (line relaxation: the whole solve)
    B_TRY(fork_lanes(b));
    for (int i = 0; i < nc; ++i) {
        if (!runs(i)) continue;
        smac_handle *h = b->cases[i];
        hipStream_t ls = lane_of(i);
        const Dims &d = h->d;
        const DevParams &dp = h->dp;
        if (has_sa(h) && is_lusgs(h)) {
            SaArrays &a = h->sa;
            lusgs_post<double, 1>(ls, d, h->lws_sa, h->wake, a.anut, a.ds, a.sf, false);
            launch_sa_update(ls, d, dp, h->w, h->st, a);
        }
        double *btc = h->blk[0], *bjm = h->blk[1], *bjp = h->blk[2], *bkm = h->blk[3],
               *bkp = h->blk[4];
        const bool f32 = h->p.precision == SMAC_FP32_LHS;
        float *fb[5] = {(float *)btc, (float *)bjm, (float *)bjp, (float *)bkm, (float *)bkp};
        if (f32) {
            launch_assemble<float>(ls, d, dp, h->q, h->met, h->dj, h->s, fb[0], fb[1], fb[2],
                                   fb[3], fb[4]);
            launch_bc<float>(ls, d, dp, h->regs.data(), (int)h->regs.size(), h->st, h->x,
                             h->y, h->met, h->q, h->s, fb[0], fb[1], fb[2], fb[3], fb[4]);
        } else {
            launch_assemble<double>(ls, d, dp, h->q, h->met, h->dj, h->s, btc, bjm, bjp, bkm,
                                    bkp);
            launch_bc<double>(ls, d, dp, h->regs.data(), (int)h->regs.size(), h->st, h->x,
                              h->y, h->met, h->q, h->s, btc, bjm, bjp, bkm, bkp);
        }
        launch_resid(ls, d, h->nt, dp.beta, h->s, h->dj, h->w, h->st);
        if (!is_lusgs(h))
            enqueue_lrelax(h, ls);
        else if (f32)
            lusgs_pre<float, 3>(ls, d, h->lws, h->s, fb[0], fb[1], fb[2], fb[3], fb[4], nullptr);
        else
            lusgs_pre<double, 3>(ls, d, h->lws, h->s, btc, bjm, bjp, bkm, bkp, nullptr);
    }
    B_TRY(join_lanes(b));
    for (int g = 0; g < 2; ++g) {
        BandGroup &gm = b->mean[g];
        if (!gm.nblk) continue;
        const bool rec = time_band && g == b->band_group;
        if (rec) HIP_TRY(hipEventRecord(b->ev_band[0], sm));
        if (g == 0)
            lusgs_band_multi<double, 3>(sm, gm.jobs, gm.map, gm.nblk, gm.pf, true, act);
        else
            lusgs_band_multi<float, 3>(sm, gm.jobs, gm.map, gm.nblk, gm.pf, true, act);
        if (rec) HIP_TRY(hipEventRecord(b->ev_band[1], sm));
        B_TRY(fork_lanes(b));
        for (int i : gm.members) {
            if (!runs(i)) continue;
            if (g == 0)
                lusgs_mid<double, 3>(lane_of(i), b->cases[i]->d, b->cases[i]->lws);
            else
                lusgs_mid<float, 3>(lane_of(i), b->cases[i]->d, b->cases[i]->lws);
        }
        B_TRY(join_lanes(b));
        if (g == 0)
            lusgs_band_multi<double, 3>(sm, gm.jobs, gm.map, gm.nblk, gm.pf, false, act);
        else
            lusgs_band_multi<float, 3>(sm, gm.jobs, gm.map, gm.nblk, gm.pf, false, act);
    }
    B_TRY(fork_lanes(b));
    for (int i = 0; i < nc; ++i) {
        if (!runs(i)) continue;
        smac_handle *h = b->cases[i];
        hipStream_t ls = lane_of(i);
        const Dims &d = h->d;
        if (is_lusgs(h)) {
            if (h->p.precision == SMAC_FP32_LHS)
                lusgs_post<float, 3>(ls, d, h->lws, h->wake, h->q, h->dq, h->s, false);
            else
                lusgs_post<double, 3>(ls, d, h->lws, h->wake, h->q, h->dq, h->s, false);
        }
        launch_update_force(ls, d, h->dp, h->walls, h->x, h->y, h->met, h->dj, h->q, h->dq,
                            h->w, h->st);
    }
    B_TRY(join_lanes(b));
    HIP_TRY(hipGetLastError());
    return SMAC_OK;
}

extern "C" {

smac_status smac_batch_create(smac_handle **cases, int n, smac_batch **out)
{
    g_err.clear();
    if (!cases || n < 1 || !out) return fail(SMAC_ERR_ARG, "cases, n or out invalid");
    *out = nullptr;
    smac_batch *b = new smac_batch();
    for (int i = 0; i < n; ++i) {
        smac_handle *h = cases[i];
        if (!h || h->timing || h->keep_lhs) {
            free_batch(b);
            return fail(SMAC_ERR_ARG, "batch case " + std::to_string(i) +
                                          " is NULL or has option timing / keep_lhs on");
        }
        for (int k = 0; k < i; ++k)
            if (cases[k] == h) {
                free_batch(b);
                return fail(SMAC_ERR_ARG, "a case appears twice in the batch");
            }
        b->cases.push_back(h);
    }
    bool ok = hipStreamCreate(&b->stream) == hipSuccess &&
              hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->ev_sa_read, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->ev_sa_done, hipEventDisableTiming) == hipSuccess &&
              hipEventCreate(&b->ev_band[0]) == hipSuccess &&
              hipEventCreate(&b->ev_band[1]) == hipSuccess;
    for (int k = 0; k < BATCH_LANES && ok; ++k)
        ok = hipStreamCreate(&b->lane[k]) == hipSuccess &&
             hipEventCreateWithFlags(&b->ev_join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        free_batch(b);
        return fail(SMAC_ERR_HIP, "hipStreamCreate / hipEventCreate failed");
    }
    {
        std::vector<int> ones(n, 1);
        std::vector<DevStats *> stp(n);
        for (int i = 0; i < n; ++i) stp[i] = cases[i]->st;
        if (hipMalloc(&b->active, n * sizeof(int)) != hipSuccess ||
            hipMalloc(&b->passes, n * sizeof(int)) != hipSuccess ||
            hipMalloc(&b->first, n * sizeof(double)) != hipSuccess ||
            hipMalloc(&b->stp, n * sizeof(DevStats *)) != hipSuccess ||
            hipMemcpy(b->active, ones.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->stp, stp.data(), n * sizeof(DevStats *), hipMemcpyHostToDevice) != hipSuccess) {
            free_batch(b);
            return fail(SMAC_ERR_HIP, "hipMalloc / hipMemcpy failed (batch flags)");
        }
    }
    smac_status rc = build_tables(b);
    if (rc) {
        free_batch(b);
        return rc;
    }
    *out = b;
    return SMAC_OK;
}

smac_status smac_batch_run(smac_batch *b, int npass)
{
    if (!b) return fail(SMAC_ERR_ARG, "batch is NULL");
    if (npass < 0) return fail(SMAC_ERR_ARG, "npass < 0");
    // the batch's stream starts after whatever the cases' own streams hold
    for (smac_handle *h : b->cases) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream_sa));
    }
    B_TRY(check_tables(b));
    {
        std::vector<int> ones(b->cases.size(), 1);   // every case runs every pass
        HIP_TRY(hipMemcpy(b->active, ones.data(), ones.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    // the mean-flow group with the most points is timed in the last pass
    double pts[2] = {0.0, 0.0};
    for (int g = 0; g < 2; ++g)
        for (int i : b->mean[g].members) pts[g] += (double)b->cases[i]->d.N;
    b->band_group = pts[0] == 0.0 && pts[1] == 0.0 ? -1 : (pts[1] > pts[0] ? 1 : 0);
    for (int it = 0; it < npass; ++it) B_TRY(enqueue_batch_pass(b, it == npass - 1));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (npass > 0 && b->band_group >= 0) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, b->ev_band[0], b->ev_band[1]));
        b->band_ms = ms;
        b->band_pts = pts[b->band_group];
    }
    for (size_t i = 0; i < b->cases.size(); ++i) {
        smac_handle *h = b->cases[i];
        DevStats s;
        HIP_TRY(hipMemcpy(&s, h->st, sizeof s, hipMemcpyDeviceToHost));
        if (s.stall)
            return fail(SMAC_ERR_STALL, "batch case " + std::to_string(i) +
                                            ": LU-SGS band hand-over timed out");
        if (s.nonfinite)
            return fail(SMAC_ERR_NONFINITE, "batch case " + std::to_string(i) +
                                                ": non-finite at nt=" + std::to_string(h->nt));
    }
    return SMAC_OK;
}

// Runs up to max_pass passes; each case stops after the pass whose residual norm is at most
// drop times its first pass's (smac_run's criterion).  With every case on LU-SGS the test is
// on the device: the host enqueues passes without reading any case's statistics, looks at
// the cases' flags every 16 passes and returns once every case has stopped.  Otherwise (the
// per-case form of the pass, line-relaxation cases) the host reads the flags after every
// pass and gives stopped cases no further kernels.
smac_status smac_batch_run_until(smac_batch *b, int max_pass, double drop, int *passes_out)
{
    if (!b) return fail(SMAC_ERR_ARG, "batch is NULL");
    if (max_pass < 0) return fail(SMAC_ERR_ARG, "max_pass < 0");
    if (!(drop >= 0.0)) return fail(SMAC_ERR_ARG, "drop must be >= 0");
    for (smac_handle *h : b->cases) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream_sa));
    }
    B_TRY(check_tables(b));
    const int n = (int)b->cases.size();
    std::vector<int> flag(n, 1), cnt(n, 0), nt0(n);
    std::vector<double> first(n);
    for (int i = 0; i < n; ++i) {
        nt0[i] = b->cases[i]->nt;
        first[i] = b->cases[i]->rhs0 > 0.0 ? b->cases[i]->rhs0 : 0.0;   // 0: this run's first
    }
    HIP_TRY(hipMemcpy(b->active, flag.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->passes, cnt.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->first, first.data(), n * sizeof(double), hipMemcpyHostToDevice));
    // each case's pass count and reference norm from the device (also after a failed pass,
    // so that the handles count what ran)
    auto settle = [&]() -> hipError_t {
        hipError_t e = hipStreamSynchronize(b->stream);
        if (e == hipSuccess) e = hipMemcpy(cnt.data(), b->passes, n * sizeof(int), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(first.data(), b->first, n * sizeof(double), hipMemcpyDeviceToHost);
        for (int i = 0; i < n; ++i) {
            smac_handle *h = b->cases[i];
            h->nt = nt0[i] + cnt[i];   // the host counted every enqueued pass
            if (nt0[i] == 0 && cnt[i] > 0) h->rhs0 = first[i];
            if (passes_out) passes_out[i] = cnt[i];
        }
        return e;
    };
    std::vector<char> on(n, 1);
    for (int it = 0; it < max_pass; ++it) {
        smac_status rc = enqueue_batch_pass(b, false, b->multi_ok ? nullptr : &on);
        if (rc) {
            (void)settle();
            return rc;
        }
        launch_batch_gate(b->stream, n, b->stp, b->first, drop, b->active, b->passes);
        HIP_TRY(hipGetLastError());
        if (!b->multi_ok) {
            HIP_TRY(hipMemcpy(flag.data(), b->active, n * sizeof(int), hipMemcpyDeviceToHost));
            bool any = false;
            for (int i = 0; i < n; ++i) {
                on[i] = flag[i] != 0;
                any = any || on[i];
            }
            if (!any) break;
        } else if ((it + 1) % 16 == 0 && it + 1 < max_pass) {
            HIP_TRY(hipMemcpy(flag.data(), b->active, n * sizeof(int), hipMemcpyDeviceToHost));
            bool any = false;
            for (int v : flag) any = any || v != 0;
            if (!any) break;
        }
    }
    HIP_TRY(settle());
    b->band_ms = 0.0;
    b->band_pts = 0.0;
    b->band_group = -1;
    for (int i = 0; i < n; ++i) {
        DevStats s;
        HIP_TRY(hipMemcpy(&s, b->cases[i]->st, sizeof s, hipMemcpyDeviceToHost));
        if (s.stall)
            return fail(SMAC_ERR_STALL, "batch case " + std::to_string(i) +
                                            ": LU-SGS band hand-over timed out");
        if (s.nonfinite)
            return fail(SMAC_ERR_NONFINITE, "batch case " + std::to_string(i) +
                                                ": non-finite at nt=" + std::to_string(b->cases[i]->nt));
    }
    return SMAC_OK;
}

smac_status smac_batch_timings(smac_batch *b, double out[3])
{
    if (!b || !out) return fail(SMAC_ERR_ARG, "NULL argument");
    out[0] = b->band_ms;
    out[1] = b->band_pts;
    out[2] = b->band_group == 1 ? 1.0 : 0.0;
    return SMAC_OK;
}

smac_status smac_batch_destroy(smac_batch *b)
{
    if (!b) return fail(SMAC_ERR_ARG, "batch is NULL");
    (void)hipStreamSynchronize(b->stream);
    free_batch(b);
    return SMAC_OK;
}

}  // extern "C"

// ---- Tier 1 -------------------------------------------------------------------------------

struct DevBuf {
    std::vector<void *> ptrs;
    ~DevBuf()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T *get(size_t n)
    {
        void *p = nullptr;
        if (hipMalloc(&p, n * sizeof(T) + 256) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T *)p;
    }
};

static smac_status wake_to_device(DevBuf &b, int jmax, int kmax, int npair, const int *wake,
                                  WakePairs &w)
{
    w = WakePairs{0, nullptr};
    if (npair <= 0) return SMAC_OK;
    if (!wake) return fail(SMAC_ERR_ARG, "wake pairs missing");
    std::vector<int> z(2 * npair);
    for (int i = 0; i < 2 * npair; ++i) {
        if (wake[i] < 1 || wake[i] > jmax) return fail(SMAC_ERR_ARG, "wake pair index out of range");
        z[i] = wake[i] - 1;
    }
    (void)kmax;
    int *d = b.get<int>(2 * npair);
    if (!d) return fail(SMAC_ERR_HIP, "hipMalloc failed");
    HIP_TRY(hipMemcpy(d, z.data(), z.size() * 4, hipMemcpyHostToDevice));
    w = WakePairs{npair, d};
    return SMAC_OK;
}

extern "C" {

smac_status smac_lrelax(const smac_params *p, int jmax, int kmax, int npair, const int *wake,
                        const double *q, double *dq, double *s, const double *btc,
                        const double *bjm, const double *bjp, const double *bkm,
                        const double *bkp)
{
    g_err.clear();
    smac_status rc = check_params(p);
    if (rc) return rc;
    if (p->precision != SMAC_FP64)
        return fail(SMAC_ERR_ARG, "the host-array drop-ins solve the reference's FP64 system");
    if (!q || !dq || !s || !btc || !bjm || !bjp || !bkm || !bkp)
        return fail(SMAC_ERR_ARG, "array argument is NULL");
    if (jmax < 3 || kmax < 3 || jmax > 2700 || kmax > 2700)
        return fail(SMAC_ERR_SIZE, "jmax, kmax must be in [3, 2700]");
    std::string why;
    if (!check_gpu(why)) return fail(SMAC_ERR_NOGPU, why);
    const long N = (long)jmax * kmax;
    Dims d{jmax, kmax, N};
    DevParams dp = {};
    dp.underr = p->underr;
    DevBuf b;
    double *dq_ = b.get<double>(3 * N), *s_ = b.get<double>(3 * N), *q_ = b.get<double>(3 * N);
    double *bl[5];
    const double *hb[5] = {btc, bjm, bjp, bkm, bkp};
    for (int i = 0; i < 5; ++i) bl[i] = b.get<double>(9 * N);
    if (!dq_ || !s_ || !q_ || !bl[0] || !bl[1] || !bl[2] || !bl[3] || !bl[4])
        return fail(SMAC_ERR_HIP, "hipMalloc failed");
    WakePairs w;
    rc = wake_to_device(b, jmax, kmax, npair, wake, w);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(q_, q, 3 * N * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s_, s, 3 * N * 8, hipMemcpyHostToDevice));
    for (int i = 0; i < 5; ++i) HIP_TRY(hipMemcpy(bl[i], hb[i], 9 * N * 8, hipMemcpyHostToDevice));
    launch_lrelax_ref<double>(0, d, dp, p->njsp, p->nksp, p->nbcimp, w, q_, dq_, s_, bl[0], bl[1],
                              bl[2], bl[3], bl[4]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(dq, dq_, 3 * N * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(s, s_, 3 * N * 8, hipMemcpyDeviceToHost));
    return SMAC_OK;
}

smac_status smac_lusgs(const smac_params *p, int jmax, int kmax, int neqs, int npair,
                       const int *wake, const double *q, double *dq, double *rhs,
                       const double *btc, const double *bjm, const double *bjp,
                       const double *bkm, const double *bkp, double *scr1)
{
    g_err.clear();
    if (!p) return fail(SMAC_ERR_ARG, "params is NULL");
    if (p->precision != SMAC_FP64)
        return fail(SMAC_ERR_ARG, "the host-array drop-ins solve the reference's FP64 system");
    if (neqs != 1 && neqs != 3) return fail(SMAC_ERR_NEQS, "neqs must be 1 or 3 (ref/lusgs.f:108-115)");
    if (!q || !dq || !rhs || !btc || !bjm || !bjp || !bkm || !bkp || !scr1)
        return fail(SMAC_ERR_ARG, "array argument is NULL");
    if (jmax < 2 || kmax < 2 || jmax > 8100 || kmax > 8100)
        return fail(SMAC_ERR_SIZE, "jmax and kmax must be in [2, 8100] (ref/include/common.f:45)");
    const long N = (long)jmax * kmax;
    const int nb = neqs * neqs;
    // the hyperplane schedule needs the couplings the natural ordering reads across rows
    // to be zero: bjm at j = 1 and bjp at j = jmax
    for (int e = 0; e < nb; ++e)
        for (int k = 0; k < kmax; ++k)
            if (bjm[e * N + (long)jmax * k] != 0.0 || bjp[e * N + (long)jmax * k + jmax - 1] != 0.0)
                return fail(SMAC_ERR_ARG, "bjm(1,k) and bjp(jmax,k) must be zero (C-grid rows)");
    std::string why;
    if (!check_gpu(why)) return fail(SMAC_ERR_NOGPU, why);
    Dims d{jmax, kmax, N};
    DevBuf b;
    double *dq_ = b.get<double>(neqs * N), *r_ = b.get<double>(neqs * N), *q_ = b.get<double>(neqs * N);
    double *bl[6];
    const double *hb[5] = {btc, bjm, bjp, bkm, bkp};
    for (int i = 0; i < 6; ++i) bl[i] = b.get<double>(nb * N);
    for (int i = 0; i < 6; ++i)
        if (!bl[i]) return fail(SMAC_ERR_HIP, "hipMalloc failed");
    if (!dq_ || !r_ || !q_) return fail(SMAC_ERR_HIP, "hipMalloc failed");
    WakePairs w;
    smac_status rc = wake_to_device(b, jmax, kmax, npair, wake, w);
    if (rc) return rc;
    LusgsWs ws{};
    lusgs_ws_shape(jmax, kmax, 64, ws);
    ws.ring = 1;
    if (!lusgs_ws_fits(ws)) return fail(SMAC_ERR_SIZE, "grid too large for the LU-SGS workspace");
    const long lb = lusgs_ws_lhs_bytes(ws, neqs), vb = lusgs_ws_vec_bytes(ws, neqs);
    ws.lhs = b.get<char>(lb);
    ws.r = b.get<char>(vb);
    ws.dq = b.get<char>(vb);
    ws.edge = b.get<unsigned long long>(ws.nedge);
    ws.stall = b.get<int>(1);
    if (!ws.lhs || !ws.r || !ws.dq || !ws.edge || !ws.stall) return fail(SMAC_ERR_HIP, "hipMalloc failed");
    // padding positions are read by idle lanes and must hold finite values
    HIP_TRY(hipMemset(ws.lhs, 0, lb));
    HIP_TRY(hipMemset(ws.r, 0, vb));
    HIP_TRY(hipMemset(ws.dq, 0, vb));
    HIP_TRY(hipMemset(ws.edge, 0, ws.nedge * 8));
    HIP_TRY(hipMemset(ws.stall, 0, sizeof(int)));
    HIP_TRY(hipMemcpy(q_, q, neqs * N * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(r_, rhs, neqs * N * 8, hipMemcpyHostToDevice));
    for (int i = 0; i < 5; ++i) HIP_TRY(hipMemcpy(bl[i], hb[i], nb * N * 8, hipMemcpyHostToDevice));
    if (neqs == 3)
        launch_lusgs3<double>(0, d, ws, w, q_, dq_, r_, bl[0], bl[1], bl[2], bl[3], bl[4], bl[5],
                              true);
    else
        launch_lusgs1(0, d, ws, w, q_, dq_, r_, bl[0], bl[1], bl[2], bl[3], bl[4], bl[5], true);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    int stall = 0;
    HIP_TRY(hipMemcpy(&stall, ws.stall, sizeof(int), hipMemcpyDeviceToHost));
    if (stall) return fail(SMAC_ERR_STALL, "LU-SGS band hand-over timed out");
    HIP_TRY(hipMemcpy(dq, dq_, neqs * N * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(rhs, r_, neqs * N * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(scr1, bl[5], nb * N * 8, hipMemcpyDeviceToHost));
    return SMAC_OK;
}

}  // extern "C"
