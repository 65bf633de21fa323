"""Launch-shape autotuner for the implicit iteration.

The tunable knobs change how the work is cut into workgroups, never the arithmetic (results
are bit-identical for every setting):

* ``lusgs_rows``        grid rows per wavefront band of the LU-SGS hyperplane sweeps;
* ``lusgs_prefetch``    steps of operand prefetch in the 3x3 (mean-flow) band sweeps;
* ``lusgs_prefetch_sa`` the same for the scalar (Spalart-Allmaras) band sweeps;
* ``lusgs_ring``        1: the 3x3 sweeps stream their operands through an LDS ring fed by
                        loader wavefronts (then ``lusgs_prefetch`` is not used);
* ``lr_lines``          lines per workgroup of the line-relaxation factorisations.

``tune(case)`` times candidate settings on a private copy of the case (its own device
state, so the caller's solver is untouched), by wall clock around ``smac_run`` (which ends
with a device synchronisation), coordinate by coordinate, and caches the winner per device,
grid size, scheme and precision in a JSON file.  ``tune_cases`` tunes each distinct
(size, scheme, precision) of a list once; ``apply`` sets the options on a solver.
"""
from __future__ import annotations

import json
import os
import time

from .case import Case
from .solver import Solver

CACHE_ENV = "MINISMAC_AMD_TUNE_CACHE"

LUSGS_ROWS = (16, 24, 32, 48, 64)
LUSGS_PREFETCH = (2, 3, 4, 6, 8)
LUSGS_PREFETCH_SA = (2, 4, 6, 8)
LUSGS_RING = (0, 1)
LR_LINES = (8, 16, 32, 64)
# a setting replaces the current one only when it is faster by more than this fraction
# (wall-clock noise of a few passes)
MIN_GAIN = 0.02


def cache_path() -> str:
    return os.environ.get(CACHE_ENV) or os.path.join(
        os.path.expanduser("~"), ".cache", "minismac_amd", "autotune.json")


def _device_name() -> str:
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:          # noqa: BLE001 -- the name only keys the cache
        return "unknown"


def key(case: Case) -> str:
    p = case.params
    sa = int(p.ivis == 1 and p.iturb == 1)
    return f"{_device_name()}|{case.jmax}x{case.kmax}|impsch{p.impsch}|prec{p.precision}|sa{sa}"


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path: str, table: dict) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


class _Probe:
    """A private solver of the case that every timing restarts from the state after its
    first pass, so that long tuning runs never drift into a different flow."""

    def __init__(self, case: Case):
        self.s = Solver(case)
        st = self.s.run(1)[0]                  # the first pass (wall distance) is not timed
        q, a, _ = self.s.get_state()
        self.start = dict(q=q, anut=a, nt=st["nt"], resmax0=st["resmax0"], clift=st["cl"])

    def time(self, opts: dict, passes: int, repeats: int) -> float:
        for k, v in opts.items():
            self.s.set_option(k, v)
        best = float("inf")
        for _ in range(repeats):
            self.s.set_state(**self.start)
            self.s.run(2)                      # warm the launch shape
            t0 = time.perf_counter()
            self.s.run(passes)
            best = min(best, (time.perf_counter() - t0) / passes)
        return best

    def close(self) -> None:
        self.s.close()


def _space(case: Case) -> list[tuple[str, tuple]]:
    p = case.params
    sa = p.ivis == 1 and p.iturb == 1
    if p.impsch == 2:
        dims = [("lusgs_ring", LUSGS_RING), ("lusgs_rows", LUSGS_ROWS),
                ("lusgs_prefetch", LUSGS_PREFETCH)]
        if sa:
            dims.append(("lusgs_prefetch_sa", LUSGS_PREFETCH_SA))
        return dims
    return [("lr_lines", LR_LINES)]


def tune(case: Case, passes: int = 10, repeats: int = 2, sweeps: int = 2,
         use_cache: bool = True, path: str | None = None) -> dict:
    """Best options for this case's shape.  Coordinate descent over the knobs: each knob in
    turn takes its fastest value with the others held, `sweeps` times over."""
    path = path or cache_path()
    k = key(case)
    if use_cache:
        hit = _load(path).get(k)
        if hit is not None:
            return dict(hit["options"])
    dims = _space(case)
    defaults = {"lusgs_ring": 1, "lusgs_rows": 64, "lusgs_prefetch": 4, "lusgs_prefetch_sa": 8,
                "lr_lines": 64}
    cur = {name: defaults[name] for name, _ in dims}
    probe = _Probe(case)
    try:
        t_cur = t_default = probe.time(cur, passes, repeats)
        for _ in range(sweeps):
            changed = False
            for name, vals in dims:
                for v in vals:
                    if v == cur[name]:
                        continue
                    trial = dict(cur, **{name: v})
                    t = probe.time(trial, passes, repeats)
                    if t < t_cur * (1.0 - MIN_GAIN):
                        cur, t_cur, changed = trial, t, True
            if not changed:
                break
    finally:
        probe.close()
    best = cur
    if use_cache:
        table = _load(path)
        table[k] = {"options": best, "ms_per_pass": t_cur * 1e3,
                    "default_ms_per_pass": t_default * 1e3}
        _store(path, table)
    return dict(best)


def tune_cases(cases, **kw) -> list[dict]:
    """Options for every case of a list, tuning each distinct shape once."""
    done: dict[str, dict] = {}
    out = []
    for c in cases:
        k = key(c)
        if k not in done:
            done[k] = tune(c, **kw)
        out.append(dict(done[k]))
    return out


def apply(solver: Solver, options: dict) -> None:
    for name, value in options.items():
        solver.set_option(name, value)
