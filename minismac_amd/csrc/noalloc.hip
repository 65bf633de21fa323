// Allocation-free variant of the LU-SGS drop-in (smac_lusgs_noalloc, include/smac_amd.h): the
// same kernels as smac_lusgs, on caller-owned device arrays, with the workspace in one static
// device array per GPU instead of hipMalloc'd buffers.
//
// `scratch` below is a device symbol.  Host code never uses its name as an address (on the
// host it is only a stand-in for the symbol): the device address of each GPU's copy is looked
// up once per device with hipGetSymbolAddress, with that device current, and kept in
// g_scratch[device]; kernel arguments use that address.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/smac_amd.h"
#include "smac_kernels.h"

using namespace smac;

// error text of the calling thread (api.hip)
smac_status smac_fail_text(smac_status st, const char *msg);

static __device__ __align__(256) unsigned char scratch[SMAC_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char *g_scratch[MAX_DEVICES] = {};

// the device address of this GPU's scratch array (looked up on the first call on the device)
smac_status scratch_of(int dev, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, dev) != hipSuccess)
            return smac_fail_text(SMAC_ERR_HIP, "hipGetDeviceProperties failed");
        if (std::string(pr.gcnArchName).compare(0, 6, "gfx950") != 0)
            return smac_fail_text(SMAC_ERR_NOGPU, "this library is built for gfx950");
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return smac_fail_text(SMAC_ERR_HIP, (std::string("hipGetSymbolAddress(scratch): ") +
                                                 hipGetErrorString(e)).c_str());
        g_scratch[dev] = (unsigned char *)p;
    }
    *out = g_scratch[dev];
    return SMAC_OK;
}

// zeroes the workspace part of the scratch array (the sheared planes' padding is read by idle
// lanes and must be finite; the hand-over slots are reset again by the shear kernel) and the
// stall flag.  A kernel rather than hipMemsetAsync, so that the call can be graph-captured.
__global__ void k_noalloc_init(uint4 *__restrict__ base, long n16, int *__restrict__ stall)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        base[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && threadIdx.x == 0) *stall = 0;
}

long align256(long n) { return (n + 255) / 256 * 256; }

}  // namespace

extern "C" smac_status smac_lusgs_noalloc(const smac_params *p, int jmax, int kmax, int neqs,
                                          int npair, const int *wake, const double *q,
                                          double *dq, double *rhs, const double *btc,
                                          const double *bjm, const double *bjp,
                                          const double *bkm, const double *bkp, double *scr1,
                                          int *stall, void *stream)
{
    if (!p) return smac_fail_text(SMAC_ERR_ARG, "params is NULL");
    if (p->precision != SMAC_FP64)
        return smac_fail_text(SMAC_ERR_ARG, "the LU-SGS drop-in solves the reference's FP64 system");
    if (neqs != 1 && neqs != 3)
        return smac_fail_text(SMAC_ERR_NEQS, "neqs must be 1 or 3 (ref/lusgs.f:108-115)");
    if (!q || !dq || !rhs || !btc || !bjm || !bjp || !bkm || !bkp || !scr1)
        return smac_fail_text(SMAC_ERR_ARG, "array argument is NULL");
    if (npair < 0 || (npair > 0 && !wake))
        return smac_fail_text(SMAC_ERR_ARG, "wake pairs missing");
    if (jmax < 2 || kmax < 2 || jmax > 8100 || kmax > 8100)
        return smac_fail_text(SMAC_ERR_SIZE, "jmax and kmax must be in [2, 8100] (ref/include/common.f:45)");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
        return smac_fail_text(SMAC_ERR_NOGPU, "no usable HIP device");
    unsigned char *base = nullptr;
    smac_status rc = scratch_of(dev, &base);
    if (rc) return rc;

    LusgsWs ws{};
    lusgs_ws_shape(jmax, kmax, 64, ws);
    ws.ring = 1;
    if (!lusgs_ws_fits(ws)) return smac_fail_text(SMAC_ERR_SIZE, "grid too large for the LU-SGS workspace");
    // workspace regions, each with the 256 bytes of slack the kernels' buffers are given
    const long lb = align256(lusgs_ws_lhs_bytes(ws, neqs) + 256);
    const long vb = align256(lusgs_ws_vec_bytes(ws, neqs) + 256);
    const long eb = align256(ws.nedge * 8 + 256);
    const long need = lb + 2 * vb + eb + 256;
    if (need > (long)SMAC_NOALLOC_SCRATCH_BYTES)
        return smac_fail_text(SMAC_ERR_SIZE,
                              ("the LU-SGS workspace of a " + std::to_string(jmax) + "x" +
                               std::to_string(kmax) + " grid needs " + std::to_string(need) +
                               " bytes; the allocation-free scratch holds " +
                               std::to_string((long)SMAC_NOALLOC_SCRATCH_BYTES))
                                  .c_str());
    ws.lhs = base;
    ws.r = base + lb;
    ws.dq = base + lb + vb;
    ws.edge = (unsigned long long *)(base + lb + 2 * vb);
    ws.stall = stall ? stall : (int *)(base + lb + 2 * vb + eb);

    const hipStream_t sm = (hipStream_t)stream;
    const long n16 = (lb + 2 * vb + eb) / 16;
    hipLaunchKernelGGL(k_noalloc_init, dim3(1024), dim3(256), 0, sm, (uint4 *)base, n16, ws.stall);
    const Dims d{jmax, kmax, (long)jmax * kmax};
    WakePairs w{npair, wake, 1};
    if (neqs == 3)
        launch_lusgs3<double>(sm, d, ws, w, q, dq, rhs, btc, bjm, bjp, bkm, bkp, scr1, true);
    else
        launch_lusgs1(sm, d, ws, w, q, dq, rhs, btc, bjm, bjp, bkm, bkp, scr1, true);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return smac_fail_text(SMAC_ERR_HIP, (std::string("kernel launch: ") + hipGetErrorString(e)).c_str());
    return SMAC_OK;
}
