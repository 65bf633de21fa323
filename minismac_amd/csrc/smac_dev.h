// Device-side definitions shared by the gfx950 kernels of the implicit iteration.
//
// Layout in HBM (one case): the reference's Fortran arrays, j fastest, 0-based here:
//   point (j,k) -> p = j + jm*k;  component n of a 3-vector field -> a[n*N + p];
//   block entry (m,n) of a 3x3 field (m = equation/row, n = variable/column) -> b[(m+3n)*N + p]
// so every plane is contiguous and a wavefront along j reads 64 consecutive doubles per
// plane (coalesced), which is what the point kernels and the eta-line kernels do; xi-line
// kernels walk j inside one thread and lean on L2 for reuse.
//
// Arithmetic: compiled with -ffp-contract=off; every expression keeps the operation order of
// the reference routine it restates (cited at each kernel), so FP64 results are the
// reference's bit for bit where only + - * / sqrt are involved.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SMAC_DEV __device__ __forceinline__

namespace smac {

struct Dims {
    int jm, km;
    long N;
};

SMAC_DEV long P(const Dims &d, int j, int k) { return (long)j + (long)d.jm * k; }

// device-resident per-iteration state and results (one per case)
struct DevStats {
    double resmax, resmax0, divmax, turres, rhsnorm;
    double cl, cd, cm, clp, cdp, fxp, fyp, fxs, fys, fmom;
    double clift;   // enters the far-field rows of the next pass
    int jres, kres, nonfinite;
    int stall;      // a band hand-over of the LU-SGS sweeps timed out
};

// solver constants the kernels read (copied from smac_params at set-up)
struct DevParams {
    double beta, vnu, dtau, dt, underr, cdis, pout, alpha, chord, xmom, ymom;
    int impsch, ivis, iturb;
};

SMAC_DEV double dmax_(double a, double b) { return a > b ? a : b; }
SMAC_DEV double dmin_(double a, double b) { return a < b ? a : b; }

// The block index and grid a kernel body works with: the hardware's in a plain launch
// (hw_blk), an item's own in a multi-case launch (k_multi).
struct Blk {
    unsigned x, y, nx, ny;
};
SMAC_DEV Blk hw_blk() { return Blk{blockIdx.x, blockIdx.y, gridDim.x, gridDim.y}; }

// The hardware's block index remapped so that each of the 8 XCDs (each with its own L2)
// takes a contiguous run of the grid in row-major order: workgroup L of a launch runs on XCD
// L % 8 as its (L / 8)-th, and gets logical block start(L % 8) + L / 8, a bijection onto the
// grid.  Kernels whose points read neighbouring rows then find the rows shared with the
// blocks beside them in their own XCD's L2.
SMAC_DEV unsigned xcd_lin()
{
    constexpr unsigned NX = 8;
    const unsigned nb = gridDim.x * gridDim.y, L = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned q = nb / NX, r = nb % NX, x = L % NX;
    return x * q + (x < r ? x : r) + L / NX;
}
SMAC_DEV Blk xcd_blk()
{
    const unsigned lg = xcd_lin();
    return Blk{lg % gridDim.x, lg / gridDim.x, gridDim.x, gridDim.y};
}

// One launch for many items (cases, or a case's BC regions) of one kernel: item i's own grid
// is args[i].gx x args[i].gy blocks, and the launch's blocks are those grids one after the
// other (start[i] = the first block of item i, start[n] = the total).  Each block finds its
// item and runs the kernel body C::run with the item's arguments and its block index in the
// item's grid, so every item gets exactly the work (and the arithmetic) of its own launch.
// Blocks are taken in the XCD-aware order (xcd_lin), so that an item's neighbouring blocks
// share an L2; no body waits on another block, so the order is free.
template <class C>
__global__ void __launch_bounds__(C::THREADS)
k_multi(const typename C::Args *__restrict__ args, const int *__restrict__ start, int n,
        const int *__restrict__ item_case, const int *__restrict__ active)
{
    const int b = (int)xcd_lin();
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    if (active[item_case[lo]] == 0) return;   // the case has stopped (device-side test)
    const typename C::Args &a = args[lo];
    const unsigned lb = (unsigned)(b - start[lo]);
    C::run(Blk{lb % a.gx, lb / a.gx, a.gx, a.gy}, a);
}

}  // namespace smac
