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

}  // namespace smac
