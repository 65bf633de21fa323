"""The reference's output side for a one-zone run: the per-iteration ``XX`` line of ``iterout``
(``ref/smac2d.f:1827-1840``, format 110 at ``:1999-2003``), the YAML run record
(``ref/smac2d.f:707-740, 1080-1108``) and the PLOT3D solution file ``Q2D.p3d``
(``ref/smac2d.f:2010-2080``, formats 205 and 206 at ``:2188-2189``).

A driver that replaces the reference's loop with ``Solver.step`` / ``Solver.run`` writes
these to keep the reference's outputs: the same columns and number formats, so the tools
that read the reference's logs, YAML and q files read these.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["fortran_e", "fortran_f", "xx_header", "xx_line", "write_q2d", "read_q2d",
           "write_yaml"]


def fortran_e(v: float, w: int, d: int) -> str:
    """Fortran ``1PEw.d``: one digit before the point, d after, a two-digit exponent with its
    sign (``E+05``), or a three-digit one without the letter (``+100``) as the standard
    requires beyond 99; right-justified in w columns, all stars if it does not fit."""
    if not math.isfinite(v):
        s = "NaN" if math.isnan(v) else ("-Infinity" if v < 0 else "Infinity")
        return s.rjust(w) if len(s) <= w else "*" * w
    m = f"{v:.{d}E}"
    mant, exp = m.split("E")
    e = int(exp)
    s = f"{mant}E{e:+03d}" if abs(e) <= 99 else f"{mant}{e:+04d}"
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_f(v: float, w: int, d: int) -> str:
    """Fortran ``Fw.d`` (no scale factor), right-justified in w columns."""
    s = f"{v:.{d}f}"
    if s.startswith("0.") and len(s) > w:
        s = s[1:]
    elif s.startswith("-0.") and len(s) > w:
        s = "-" + s[2:]
    return s.rjust(w) if len(s) <= w else "*" * w


def xx_header() -> str:
    """Format 100 of ``iterout``: the column titles above the ``XX`` lines."""
    return ("\n     nt   resmax     divmax     turres      j     k    nz     cputime       cd"
            "          cl\n"
            "   ----  ---------  ---------  ---------  ----  ----  ----  ----------  "
            "----------  ----------")


def xx_line(st: dict, cputime: float = 0.0) -> str:
    """One ``XX`` line (format 110: ``'XX',i5,1p3e11.3,3i6,1p3e12.4``) from a pass's stats:
    nt, resmax/resmax0 (resmax0 taken as 1 when zero), divmax, turres, the location of
    resmax, zone 0 of a one-rank run, the CPU time of the pass, CD and CL."""
    r0 = st["resmax0"] if st["resmax0"] != 0.0 else 1.0
    rmax = st["resmax"] / r0
    dmax = max(0.0, st["divmax"])
    tmax = max(0.0, st["turres"])
    return ("XX" + f"{st['nt']:5d}" + fortran_e(rmax, 11, 3) + fortran_e(dmax, 11, 3) +
            fortran_e(tmax, 11, 3) + f"{st['jres']:6d}{st['kres']:6d}{0:6d}" +
            fortran_e(cputime, 12, 4) + fortran_e(st["cd"], 12, 4) + fortran_e(st["cl"], 12, 4))


def _q_planes(q: np.ndarray) -> np.ndarray:
    """The four PLOT3D planes the reference writes: 1, u, v and p/0.4 + (u^2 + v^2)/2 (0.4 is
    a default-REAL literal there)."""
    p, u, v = q[:, :, 0], q[:, :, 1], q[:, :, 2]
    e = p / float(np.float32(0.4)) + 0.5 * (u * u + v * v)
    return np.stack([np.ones_like(p), u, v, e], axis=2)


def write_q2d(path: str, q: np.ndarray, alpha: float, reynum: float, cputime: float = 0.0):
    """``Q2D.p3d`` of a one-zone run: the zone count, ``jmax kmax`` (list-directed), the
    header ``0.0 alpha reynum cputime`` (format 205: ``1pf5.2,1x,1pe13.5,...``) and the four
    planes in ``6f10.6`` records, j fastest."""
    jmax, kmax = q.shape[:2]
    planes = _q_planes(np.asarray(q, dtype=np.float64))
    flat = planes.ravel(order="F")
    with open(path, "w") as f:
        f.write(f"{1:12d}\n")
        f.write(f"{jmax:12d}{kmax:12d}\n")
        f.write(fortran_f(0.0, 5, 2) + " " + fortran_e(alpha, 13, 5) + " " +
                fortran_e(reynum, 13, 5) + " " + fortran_e(cputime, 13, 5) + "\n")
        for i in range(0, flat.size, 6):
            f.write("".join(fortran_f(v, 10, 6) for v in flat[i:i + 6]) + "\n")


def read_q2d(path: str):
    """Reads a one-zone ``Q2D.p3d`` back: (header values, planes (jmax, kmax, 4))."""
    with open(path) as f:
        toks = f.read().split()
    nz, jmax, kmax = int(toks[0]), int(toks[1]), int(toks[2])
    if nz != 1:
        raise ValueError("one zone expected")
    head = [float(t) for t in toks[3:7]]
    vals = np.array([float(t) for t in toks[7:7 + 4 * jmax * kmax]])
    return head, vals.reshape((jmax, kmax, 4), order="F")


def write_yaml(path: str, case, niter: int, solve_sec: float, fp_precision: str = "double",
               extra: dict | None = None) -> None:
    """The run record the reference writes (the keys of ``ref/smac2d.f:709-740`` and
    ``:1086-1108`` that describe the case and its timing), one key per line."""
    p = case.params
    scheme = {1: "line relaxation", 2: "lusgs factorization"}.get(p.impsch, "unknown")
    cputiter = solve_sec / max(niter, 1)
    lines = [
        "code: smac2d",
        f"title: {case.name}",
        "num_grids: 1",
        "mpi_ranks: 1",
        f"grid_file: {case.name}",
        f"largest_dim: {max(case.jmax, case.kmax)}",
        f"j-dim: {case.jmax}",
        f"k-dim: {case.kmax}",
        f"total_points: {case.npts}",
        f"beta: {p.beta!r}",
        f"fp_precision: {fp_precision}",
        "dtau: " + fortran_e(p.dtau, 12, 4).strip(),
        f"reynolds_number: {p.reynum!r}",
        f"angle_of_attack_degs: {p.alpha!r}",
        f"implicit_scheme: {scheme}",
        "sec_per_iteration: " + fortran_e(cputiter, 12, 5).strip(),
        "microsec_per_iter_per_point: " + fortran_e(cputiter / case.npts * 1e6, 12, 5).strip(),
        f"number_of_iterations: {niter}",
        f"alpha_deg: {p.alpha!r}",
        "solve_sec: " + fortran_e(solve_sec, 12, 5).strip(),
    ]
    for k, v in (extra or {}).items():
        lines.append(f"{k}: {v}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
