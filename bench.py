"""Benchmark of the implicit iteration on one MI355X.

One step = one pass of the reference's time-loop body (SA, residual and implicit-operator
assembly, BC rows, residual norms, LU-SGS or line relaxation, update, wall forces) on the
945x241 NACA 4412 C-grid (wall spacing 2e-6 c, Re = 3e6, alpha = 8 deg).  value = whole-job
Mcell-iterations/s (cells x passes / wall time) of the FP64 path, which performs the
reference's arithmetic, with the case resident in HBM.  Beside it:
  - fp32_lhs: the same per pass with FP32 Jacobians, factors and LU-SGS sweeps (FP64
    residual and update);
  - line_relaxation: the line-relaxation pass in its three forms (reference order with the
    sequential substitution, bit-identical; reference order with the line solves as prefix
    scans; zebra order), FP64 and FP32-LHS;
  - converged: the metric itself -- wall-clock to a 1e-10 drop of ||s||_2 and the CL/CD it
    converges to, FP64 against FP32-LHS, on the steady variant of the grid (Re = 1000: the
    reference's mean flow is laminar, and at Re = 3e6 it never settles);
  - adjoint: gradients of CL and CD in alpha and in eight Hicks-Henne bump amplitudes at that
    converged FP64 state, and the adjoint/primal time;
  - batched: the 64-case seeded batch;
  - cpu_baseline: the oracle on one core.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# Algorithmic HBM bytes per grid point of the forward band sweep of the mean-flow LU-SGS
# (ilusol3 forward part): diagonal factors, bjm, bkm (9 words each) and rhs (3) read, the
# forward result (3) written -- 30 + 3 words.  FP32-LHS runs the scaled form (bjm and bkm
# premultiplied by the inverse diagonal, no factors): 18 + 3 words read and 3 written, all at
# 4 bytes.
BAND_FWD_BYTES = {"fp64": 33 * 8, "fp32": 24 * 4}
# HBM bytes per grid point that the same kernel moves: a fixed figure, not measured by this
# run -- rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE at 945x241 (scaled by 2.0 and 1.0, the
# factors that make the two counters read 256 MiB for a plain 256 MiB device copy), recorded
# in profiles/lusgs_band_945x241_counters.md
BAND_FWD_TRAFFIC = {"fp64": (2.0 * 30497.0 + 8318.0) * 1024 / 227745,
                    "fp32": (2.0 * 14324.0 + 4300.0) * 1024 / 227745}
TRAFFIC_SOURCE = "fixed figure from profiles/lusgs_band_945x241_counters.md (rocprofv3 FETCH_SIZE + WRITE_SIZE)"
# Algorithmic bytes per grid point of the line-relaxation solve family (two xi and two eta
# sweeps with their factorisations).  The reference-order chain factors in place and reads
# the blocks in its sweeps: 360 + 2*432 per direction.  The scan forms factor into 45
# line-major operand planes (45 read + 45 written per direction), and each sweep reads those
# 45 planes, s, dq and the neighbouring lines' dq and writes dq (45 LT + 13 FP64 words), plus
# the eta-major transposes of s and dq (4 x 3 planes read and written): 2*720 + 4*480 + 192 in
# FP64 and 2*360 + 4*300 + 192 with the blocks in FP32.
LR_SOLVE_BYTES = {"chain": 2448, "fp64": 3552, "fp32": 2112}


def cpu_baseline(case, iters: int) -> dict:
    """The oracle -- the C restatement of the reference's arithmetic, one core -- on the same
    case for a bounded number of passes.  (The reference itself cannot be built in this
    image: it needs MPI.)"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(case)
    o.step()                      # the first pass includes the O(N * N_wall) wall distance
    t0 = time.perf_counter()
    for _ in range(iters):
        o.step()
    sec = time.perf_counter() - t0
    return {"value": case.npts * iters / sec / 1e6, "unit": "Mcell-iters/s", "cores": 1,
            "kind": "port",
            "sample": f"{case.jmax}x{case.kmax}, {iters} passes of the C restatement "
                      f"(gcc -O2 -ffp-contract=off, 1 thread) after one untimed pass, {sec:.2f} s"}


def timed_run(case, steps: int, warmup: int):
    from minismac_amd.solver import Solver
    with Solver(case) as s:
        s.run(warmup)
        t0 = time.perf_counter()
        st, done = s.run(steps)          # ends with a device synchronisation
        wall = time.perf_counter() - t0
    return st, done, wall


def kernel_times(case, steps: int, warmup: int) -> dict:
    """Per-family device times (ms per pass) from HIP events on the solver's stream."""
    from minismac_amd.solver import Solver
    with Solver(case, timing=True) as s:
        s.run(warmup)
        s.run(steps)
        return s.timings()


def line_relaxation_run(cfg: int, steps: int, only=None) -> dict:
    """The line-relaxation pass at the workload size in its three forms (or those named in
    `only`)."""
    from minismac_amd._abi import SMAC_FP32_LHS, SMAC_ORDER_ZEBRA
    from minismac_amd.case import config_case
    from minismac_amd.solver import Solver
    out = {}
    forms = [("chain_fp64", {}, {}, 2), ("scan_fp64", {}, {"lr_solve": 1}, steps),
             ("scan_fp32", {"precision": SMAC_FP32_LHS}, {}, steps),
             ("zebra_fp64", {"line_order": SMAC_ORDER_ZEBRA}, {}, steps),
             ("zebra_fp32", {"line_order": SMAC_ORDER_ZEBRA, "precision": SMAC_FP32_LHS}, {},
              steps)]
    for name, kw, opts, n in forms:
        if only and name not in only:
            continue
        case = config_case(cfg, impsch=1, **kw)
        with Solver(case, options=opts, timing=True) as s:
            s.run(1)
            t0 = time.perf_counter()
            st, done = s.run(n)
            wall = time.perf_counter() - t0
            tm = s.timings()
        prec = "chain" if name == "chain_fp64" else ("fp32" if "fp32" in name else "fp64")
        out[name] = {"ms_per_pass": round(wall / done * 1e3, 4),
                     "solve_ms": round(tm["solve"], 4), "sa_ms": round(tm["sa"], 4),
                     "mcell_iters_s": round(case.npts * done / wall / 1e6, 4),
                     "solve_achieved_GBs": round(LR_SOLVE_BYTES[prec] * case.npts /
                                                 (tm["solve"] * 1e-3) / 1e9, 3)}
    out["note"] = ("chain = the reference's sequential block substitution (bit-identical to "
                   "the oracle); scan = the same line order with each line solved by prefix "
                   "scans (per-pass norms within 1e-10 of the oracle); zebra = even then odd "
                   "lines, each colour at once; one stream with HIP-event timing of the "
                   "families")
    return out


STEADY = dict(reynum=1000.0, dtau=10.0, underr=1.0)


def converged_run(cfg: int, maxit: int, drop: float, adjoint: bool) -> dict:
    """Wall-clock to a `drop` fall of ||s||_2 from the free stream, FP64 and FP32-LHS, zebra
    line relaxation, on the steady variant of the workload grid; then the adjoint gradients
    at the FP64 state."""
    from minismac_amd import grid
    from minismac_amd._abi import SMAC_FP32_LHS, SMAC_ORDER_ZEBRA
    from minismac_amd.case import config_case
    from minismac_amd.solver import Solver
    out = {}
    res = {}
    for prec, name in ((0, "fp64"), (SMAC_FP32_LHS, "fp32_lhs")):
        case = config_case(cfg, impsch=1, line_order=SMAC_ORDER_ZEBRA, precision=prec, **STEADY)
        s = Solver(case)
        try:
            s.step()                              # the first pass (wall distance)
            t0 = time.perf_counter()
            st, done = s.run(maxit, drop=drop)
            wall = time.perf_counter() - t0
            res[name] = {"passes": done + 1, "seconds": round(wall, 3),
                         "mcell_iters_s": round(case.npts * done / wall / 1e6, 4),
                         "converged": bool(done < maxit), "cl": st["cl"], "cd": st["cd"],
                         "rhsnorm": st["rhsnorm"]}
            if name == "fp64" and adjoint:
                jl, ju = [(r[1], r[2]) for r in case.resolved_regions() if r[0] == 0][0]
                modes = grid.bump_modes(case.x, case.y, jl, ju)
                dcl, dcd, ast = s.adjoint(modes, maxit=8000, tol=1e-10)
                out["adjoint"] = {
                    "seconds": round(ast["seconds"], 3), "gmres_iters": ast["iters"],
                    "resid": ast["resid"],
                    "adjoint_over_primal": round(ast["seconds"] / wall, 4),
                    "dcl_dalpha_per_deg": dcl[0], "dcd_dalpha_per_deg": dcd[0],
                    "dcl_dbumps": list(dcl[1:]), "dcd_dbumps": list(dcd[1:]),
                    "note": "both adjoints (CL, CD) by GMRES(100) with the transposed zebra "
                            "line-relaxation update (the primal's own, exactly transposed) as "
                            "right preconditioner, plus the gradients of 8 "
                            "Hicks-Henne bump modes (peaks 0.1, 0.3, 0.5, 0.7 on each side, "
                            "algebraic mesh deformation); primal = FP64 zebra passes from the "
                            "free stream to the same state"}
        finally:
            s.close()
    out.update(res)
    a, b = res["fp64"], res["fp32_lhs"]
    out["speedup_fp32_lhs"] = round(a["seconds"] / b["seconds"], 4)
    out["cl_rel_diff"] = abs(b["cl"] - a["cl"]) / abs(a["cl"])
    out["cd_rel_diff"] = abs(b["cd"] - a["cd"]) / abs(a["cd"])
    out["workload"] = (f"945x241 NACA 4412 C-grid, Re=1000 (steady), alpha=8, beta=10, SA, "
                       f"dtau=10, under-relaxation 1, zebra line relaxation, to a {drop:g} drop "
                       f"of ||s||_2 from the free stream")
    # the same runs with the line operands factored every LAG passes (option lr_refactor)
    LAG = 8
    lag = {"lr_refactor": LAG}
    for prec, name in ((0, "fp64"), (SMAC_FP32_LHS, "fp32_lhs")):
        case = config_case(cfg, impsch=1, line_order=SMAC_ORDER_ZEBRA, precision=prec, **STEADY)
        with Solver(case, options={"lr_refactor": LAG}) as s:
            s.step()
            t0 = time.perf_counter()
            st, done = s.run(maxit, drop=drop)
            wall = time.perf_counter() - t0
        lag[name] = {"passes": done + 1, "seconds": round(wall, 3),
                     "converged": bool(done < maxit),
                     "cl_rel_diff_vs_every_pass": abs(st["cl"] / res[name]["cl"] - 1.0),
                     "cd_rel_diff_vs_every_pass": abs(st["cd"] / res[name]["cd"] - 1.0)}
    lag["speedup_fp32_lhs"] = round(lag["fp64"]["seconds"] / lag["fp32_lhs"]["seconds"], 4)
    lag["note"] = ("the zebra lines factored every lr_refactor passes, their operands kept in "
                   "between with each pass's own residual: a lagged implicit operator with the "
                   "same converged state (not the reference's iteration)")
    out["lagged_line_operands"] = lag
    return out


def history_run(cfg: int, impsch: int, npass: int, every: int) -> dict:
    """Residual drop and forces every `every` passes over `npass` passes of the workload case
    from the free stream (the record behind 'never settles' for the Re = 3e6 flow)."""
    from minismac_amd.case import config_case
    from minismac_amd.solver import Solver
    from minismac_amd.solver import SmacError
    case = config_case(cfg, impsch=impsch)
    rows, end = [], None
    with Solver(case) as s:
        st = s.step()
        r0 = st["rhsnorm"]
        done = 1
        while done < npass:
            try:
                st, n = s.run(min(every, npass - done))
            except SmacError as e:
                end = f"stopped in the passes after {done}: {e}"
                break
            done += n
            rows.append([done, st["rhsnorm"] / r0, st["cl"], st["cd"]])
            print(f"bench: history {done} passes, drop {rows[-1][1]:.3e}, cl {st['cl']:.6f}",
                  file=sys.stderr, flush=True)
    return {"workload": case.name, "impsch": impsch,
            "columns": ["pass", "rhsnorm/rhsnorm(pass 1)", "cl", "cd"], "rows": rows,
            "end": end}


def batched_run(n: int, steps: int, warmup: int, seed: int = 0) -> dict:
    """The batched workload: n seeded NACA 4-digit cases of three sizes, FP32-LHS LU-SGS,
    advanced together (smac_batch_run), and the same cases advanced one after another for
    comparison."""
    from minismac_amd import autotune
    from minismac_amd._abi import SMAC_FP32_LHS
    from minismac_amd.case import batch_cases
    from minismac_amd.solver import Batch, Solver
    cases = batch_cases(n, seed=seed, impsch=2, precision=SMAC_FP32_LHS)
    npts = sum(c.npts for c in cases)
    t0 = time.perf_counter()
    opts = autotune.tune_cases(cases)
    t_tune = time.perf_counter() - t0
    tuned = {f"{c.jmax}x{c.kmax}": o for c, o in zip(cases, opts)}
    sol = [Solver(c, options=o) for c, o in zip(cases, opts)]
    try:
        with Batch(sol) as b:
            b.run(1 + warmup)             # the first pass (wall distance) and warm-up
            t0 = time.perf_counter()
            b.run(steps)                  # ends with every case's stream synchronised
            wall = time.perf_counter() - t0
            bt = b.timings()
        t0 = time.perf_counter()
        for s in sol:
            s.run(steps)
        wall_seq = time.perf_counter() - t0
        finite = all(s.stats()["nonfinite"] == 0 for s in sol)
    finally:
        for s in sol:
            s.close()
    sizes = {}
    for c in cases:
        sizes[f"{c.jmax}x{c.kmax}"] = sizes.get(f"{c.jmax}x{c.kmax}", 0) + 1
    return {"value": round(npts * steps / wall / 1e6, 4), "unit": "Mcell-iters/s",
            "cases": n, "points": npts, "sizes": sizes, "seed": seed, "steps": steps,
            "ms_per_pass_of_all": round(wall / steps * 1e3, 4),
            "sequential_value": round(npts * steps / wall_seq / 1e6, 4),
            "speedup_vs_sequential": round(wall_seq / wall, 4), "finite": finite,
            "precision": "fp32-lhs", "tuned_options": tuned,
            "band_fwd_multi_ms": round(bt["band_fwd_ms"], 5),
            "band_fwd_multi_achieved_GBs": round(
                BAND_FWD_BYTES["fp32" if bt["precision"] else "fp64"] * bt["points"]
                / (bt["band_fwd_ms"] * 1e-3) / 1e9, 3) if bt["band_fwd_ms"] > 0 else None,
            "autotune_seconds": round(t_tune, 2),
            "note": "one launch per kernel for all cases (the LU-SGS band sweeps of all cases "
                    "in one launch per direction), the SA's launches on a second stream beside "
                    "the mean flow's; sequential = the same cases run one after another with "
                    "smac_run"}


def batched_drop_run(n: int, max_pass: int, drop: float, seed: int = 1) -> dict:
    """n seeded steady cases (the batched workload's generator at Re = 1000, dtau = 10,
    FP32-LHS LU-SGS) run to a residual drop: together with the stop test on the device
    (smac_batch_run_until, each case stopping after its own pass that meets the drop), and one
    after another with smac_run's host-side test; the pass counts and states must agree."""
    import numpy as np
    from minismac_amd._abi import SMAC_FP32_LHS
    from minismac_amd.case import batch_cases
    from minismac_amd.solver import Batch, Solver
    cases = batch_cases(n, seed=seed, impsch=2, precision=SMAC_FP32_LHS, reynum=1000.0,
                        dtau=10.0)
    sol = [Solver(c) for c in cases]
    alone = []
    try:
        with Batch(sol) as b:
            t0 = time.perf_counter()
            passes = b.run_until(max_pass, drop)
            wall = time.perf_counter() - t0
        alone = [Solver(c) for c in cases]
        t0 = time.perf_counter()
        done = [s.run(max_pass, drop)[1] for s in alone]
        wall_seq = time.perf_counter() - t0
        same = passes == done and all(
            all(np.array_equal(a, c) for a, c in zip(s0.get_state(), s1.get_state()))
            for s0, s1 in zip(alone, sol))
    finally:
        for s in sol + alone:
            s.close()
    cell_passes = sum(c.npts * p for c, p in zip(cases, passes))
    stopped = sum(1 for p in passes if p < max_pass)
    return {"value": round(cell_passes / wall / 1e6, 4), "unit": "Mcell-iters/s",
            "cases": n, "seed": seed, "drop": drop, "max_pass": max_pass,
            "stopped_early": stopped, "passes_min_median_max": [
                int(min(passes)), int(sorted(passes)[len(passes) // 2]), int(max(passes))],
            "seconds": round(wall, 3), "sequential_seconds": round(wall_seq, 3),
            "speedup_vs_sequential": round(wall_seq / wall, 4),
            "same_as_sequential": same, "precision": "fp32-lhs",
            "note": "value counts the passes each case made (a stopped case's later passes do "
                    "nothing); the host reads no statistics between passes, only the cases' "
                    "flags every 16 passes"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--impsch", type=int, default=2, help="1 line relaxation, 2 LU-SGS")
    ap.add_argument("--config", type=int, default=1, help="0: 119x31, 1: 945x241, 2: 1889x481")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-LHS measurement")
    ap.add_argument("--batch-cases", type=int, default=64,
                    help="cases in the batched measurement (0 skips it)")
    ap.add_argument("--drop-cases", type=int, default=32,
                    help="cases in the batched run to a residual drop (0 skips it)")
    ap.add_argument("--no-lr", action="store_true", help="skip the line-relaxation forms")
    ap.add_argument("--lr-forms", default="",
                    help="comma list of the line-relaxation forms to time (default: all of "
                         "chain_fp64, scan_fp64, scan_fp32, zebra_fp64, zebra_fp32)")
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the converged FP64 / FP32-LHS comparison and the adjoint")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint gradients")
    ap.add_argument("--converge-maxit", type=int, default=12000)
    ap.add_argument("--history", type=int, default=0,
                    help="also record the residual drop and CL/CD of this many passes of the "
                         "workload case (every 1/20th), under 'history'")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 256 MiB device to device (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU decomposition is not part of this library)", file=sys.stderr)
        return 2

    from minismac_amd._abi import SMAC_FP32_LHS
    from minismac_amd.case import config_case

    if args.calibrate_copy:
        import torch
        a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()

    case = config_case(args.config, impsch=args.impsch)
    print(f"bench: case {case.name}, impsch={args.impsch}, {case.npts} points", file=sys.stderr,
          flush=True)
    st, done, wall = timed_run(case, args.steps, args.warmup)
    value = case.npts * done / wall / 1e6
    ms_per_step = wall / done * 1e3
    tm = kernel_times(case, args.steps, args.warmup)
    fam = {k: tm[k] for k in ("sa", "assemble", "resid", "solve", "update_force")}
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
        "kernel_ms_per_pass": {k: round(v, 5) for k, v in fam.items()},
        "last_iter": {"cl": st["cl"], "cd": st["cd"], "rhsnorm": st["rhsnorm"]},
    }
    if args.impsch == 2:
        t = tm["band_fwd"]
        alg = BAND_FWD_BYTES["fp64"] * case.npts
        achieved = alg / (t * 1e-3) / 1e9
        line["roofline"] = {
            "bound": "hbm", "kernel": "k_band_ring<fwd,double> (LU-SGS forward band sweep, LDS-ring form)",
            "achieved": round(achieved, 3), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 6),
            "traffic": round(BAND_FWD_TRAFFIC["fp64"] * case.npts),
            "traffic_source": TRAFFIC_SOURCE,
            "alg_bytes_per_launch": alg, "avg_ms_per_launch": round(t, 5),
        }
    else:
        from_bytes = LR_SOLVE_BYTES["chain"] * case.npts   # the solve family, reference order
        achieved = from_bytes / (fam["solve"] * 1e-3) / 1e9
        line["roofline"] = {
            "bound": "hbm", "kernel": "line relaxation (factor + sweeps, reference order)",
            "achieved": round(achieved, 3), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 6), "traffic": None,
            "alg_bytes_per_launch": from_bytes, "avg_ms_per_launch": round(fam["solve"], 5),
        }
    if args.impsch == 2 and not args.no_fp32:
        c32 = config_case(args.config, impsch=args.impsch, precision=SMAC_FP32_LHS)
        st32, done32, wall32 = timed_run(c32, args.steps, args.warmup)
        tm32 = kernel_times(c32, args.steps, args.warmup)
        v32 = c32.npts * done32 / wall32 / 1e6
        line["fp32_lhs"] = {
            "value": round(v32, 4), "unit": "Mcell-iters/s", "ms_per_step": round(wall32 / done32 * 1e3, 4),
            "speedup_vs_fp64": round(v32 / value, 4),
            "band_fwd_ms": round(tm32["band_fwd"], 5),
            "band_fwd_achieved_GBs": round(BAND_FWD_BYTES["fp32"] * c32.npts / (tm32["band_fwd"] * 1e-3) / 1e9, 3),
            "band_fwd_traffic": round(BAND_FWD_TRAFFIC["fp32"] * c32.npts),
            "note": "per pass on the Re=3e6 flow, which has no steady state (the reference's "
                    "mean flow is laminar; LU-SGS runs away after 181 passes, as in the "
                    "oracle); time to the same converged CL/CD is under 'converged'",
        }
    if not args.no_lr:
        print("bench: line relaxation forms", file=sys.stderr, flush=True)
        line["line_relaxation"] = line_relaxation_run(
            args.config, args.steps // 4 or 1, [f for f in args.lr_forms.split(",") if f] or None)
    if not args.no_converge:
        print("bench: converged FP64 / FP32-LHS and adjoint", file=sys.stderr, flush=True)
        line["converged"] = converged_run(args.config, args.converge_maxit, 1e-10,
                                          not args.no_adjoint)
    if args.history > 0:
        line["history"] = history_run(args.config, args.impsch, args.history,
                                      max(args.history // 20, 1))
    if args.batch_cases > 0:
        print(f"bench: batched workload, {args.batch_cases} cases", file=sys.stderr, flush=True)
        line["batched"] = batched_run(args.batch_cases, args.steps, args.warmup)
    if args.drop_cases > 0:
        print(f"bench: batched run to a drop, {args.drop_cases} cases", file=sys.stderr,
              flush=True)
        line["batched_to_drop"] = batched_drop_run(args.drop_cases, 600, 5e-4)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(case, args.cpu_iters)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
