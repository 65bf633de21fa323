"""Benchmark of the implicit iteration on one MI355X.

One step = one pass of the reference's time-loop body (SA, residual and implicit-operator
assembly, BC rows, residual norms, LU-SGS or line relaxation, update, wall forces), on the
945x241 NACA 4412 C-grid (wall spacing 2e-6 c, Re = 3e6, alpha = 8 deg), FP64, with the
reference's arithmetic.  value = whole-job Mcell-iterations/s (cells x passes / wall time)
with the case resident in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic HBM bytes per grid point of each kernel family, FP64: every array a family
# must read or write counted once per pass; zero fills and other reference artefacts excluded
BYTES_PER_POINT = {
    "solve_lusgs": 144 + 264 + 120 + 264,      # lusini3, forward, D multiply, backward
    "solve_lr": 360 + 864 + 360 + 864,         # xi factor, 2 xi sweeps, eta factor, 2 eta sweeps
    "assemble": 448,                            # q, metrics, dj in; s and five blocks out
    "resid": 32,
    "update_force": 72,
    "sa_lusgs": 272 - 32,                       # SA assembly + scalar LU-SGS (update counted in sa)
    "sa_lr": 120 + 128,
}
PEAK_HBM_GBS = 8000.0


def cpu_baseline(case, iters: int) -> dict:
    """The compiled reference (oracle/_ref/smac2d_ref_fast: flang -O3 -march=native, one
    rank, one core) on the same case for a bounded number of iterations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    try:
        r = pyoracle.run_reference(case, iters, fast=True, timeout=600)
        sec = r["solver_seconds"]
        return {"value": case.npts * iters / sec / 1e6, "unit": "Mcell-iters/s", "cores": 1,
                "kind": "reference",
                "sample": f"{case.jmax}x{case.kmax}, {iters} iterations of the reference "
                          f"(flang -O3 -march=native, 1 rank), solver time {sec:.2f} s "
                          f"incl. first-iteration wall distance"}
    except Exception as e:  # fall back to the C restatement
        o = pyoracle.Oracle(case)
        t0 = time.perf_counter()
        for _ in range(iters):
            o.step()
        sec = time.perf_counter() - t0
        return {"value": case.npts * iters / sec / 1e6, "unit": "Mcell-iters/s", "cores": 1,
                "kind": "port", "sample": f"{case.jmax}x{case.kmax}, {iters} iterations of the "
                                          f"C restatement ({sec:.2f} s); reference run failed: {e}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--impsch", type=int, default=2, help="1 line relaxation, 2 LU-SGS")
    ap.add_argument("--config", type=int, default=1, help="0: 119x31, 1: 945x241, 2: 1889x481")
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU decomposition is not part of this library)", file=sys.stderr)
        return 2

    from minismac_amd.case import config_case
    from minismac_amd.solver import Solver

    case = config_case(args.config, impsch=args.impsch)
    print(f"bench: case {case.name}, impsch={args.impsch}, {case.npts} points", file=sys.stderr,
          flush=True)
    with Solver(case) as s:
        s.run(args.warmup)
        t0 = time.perf_counter()
        st, done = s.run(args.steps)          # ends with a device synchronisation
        wall = time.perf_counter() - t0
    value = case.npts * done / wall / 1e6
    ms_per_step = wall / done * 1e3
    # kernel-family device times from HIP events on the solver's stream, on the same state
    with Solver(case, timing=True) as s:
        s.run(args.warmup)
        s.run(args.steps)
        tm = s.timings()
    fam = {"sa": tm["sa"], "assemble": tm["assemble"], "resid": tm["resid"],
           "solve": tm["solve"], "update_force": tm["update_force"]}
    dom = max(fam, key=fam.get)
    key = {"solve": "solve_lusgs" if args.impsch == 2 else "solve_lr",
           "sa": "sa_lusgs" if args.impsch == 2 else "sa_lr"}.get(dom, dom)
    alg_bytes = BYTES_PER_POINT[key] * case.npts
    achieved = alg_bytes / (fam[dom] * 1e-3) / 1e9
    line = {
        "metric": "Mcell-iters/s FP64 vs FP32-LHS at equal CL/CD (<=1e-6); adjoint/primal time",
        "value": round(value, 4),
        "unit": "Mcell-iters/s",
        "n_gpus": 1,
        "steps": done,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: generated NACA 4412 C-grid, free-stream start",
        "config": {
            "workload": f"{case.jmax}x{case.kmax} NACA 4412 C-grid (wall 2e-6c, far field 15c), "
                        f"Re=3e6, alpha=8, beta=10, SA, "
                        f"{'LU-SGS' if args.impsch == 2 else 'line relaxation'} in reference order, "
                        f"FP64, bit-identical to the CPU oracle",
            "impsch": args.impsch,
            "precision": "fp64",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 3),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 6),
            "traffic": None,
            "alg_bytes_per_launch": alg_bytes,
            "avg_ms_per_pass": round(fam[dom], 5),
        },
        "kernel_ms_per_pass": {k: round(v, 5) for k, v in fam.items()},
        "last_iter": {"cl": st["cl"], "cd": st["cd"], "rhsnorm": st["rhsnorm"]},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(case, args.cpu_iters)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
