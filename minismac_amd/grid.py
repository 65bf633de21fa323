"""C-grids around NACA 4-digit airfoils, the reference's input files, and grid metrics.

This is the on-disk input side of the implicit iteration: the reference reads a PLOT3D
grid (``ref/preproc.f:62-95``: zone count, ``jmax kmax``, then all x then all y,
list-directed), a BC-region file (``ref/rbc.f:300-345``: ``ibcval zone jbeg jend kbeg kend``
with ``-1``/``-2`` meaning ``max``/``max-1``, ``ref/rbc.f:384-396``) and the ``$input``
namelist (``ref/initia.f:85-89``).  The reference ships only one 119x31 grid, so every other
case is generated here.

Topology (same as the shipped grid ``ref/xy.dat_original_119x31``): ``j`` runs from the
downstream end of the lower wake cut, along the cut to the trailing edge, round the lower
surface to the leading edge and back along the upper surface and upper wake cut; ``k=1`` is
the surface/cut and ``k=kmax`` the C-shaped far field.  Wake point ``j`` coincides with
``jmax+1-j`` at ``k=1``.

Arrays are numpy float64 of shape ``(jmax, kmax)`` in Fortran order, so ``x.ravel('F')`` is
the reference's memory order (``j`` fastest).
"""
from __future__ import annotations

import dataclasses
import math
import os

import numpy as np

__all__ = [
    "CGrid", "naca4_surface", "make_cgrid", "cgrid_bc_regions", "read_plot3d", "write_plot3d",
    "read_bcmain", "write_bcmain", "write_namelist", "metric", "hicks_henne", "bump_modes",
    "deform",
]


@dataclasses.dataclass
class CGrid:
    x: np.ndarray           # (jmax, kmax) Fortran order
    y: np.ndarray
    jte_lower: int          # 1-based j of the trailing edge on the lower surface
    jte_upper: int          # 1-based j of the trailing edge on the upper surface

    @property
    def jmax(self) -> int:
        return self.x.shape[0]

    @property
    def kmax(self) -> int:
        return self.x.shape[1]


def naca4_surface(m: float, p: float, t: float, xc: np.ndarray):
    """Thickness and camber of a NACA 4-digit section with a closed trailing edge.

    Published NACA 4-digit definition (Abbott & von Doenhoff), with the last thickness
    coefficient -0.1036 so that the section closes at x/c=1.  Returns (yt, yc, dyc/dx).
    """
    xc = np.asarray(xc, dtype=np.float64)
    yt = 5.0 * t * (0.2969 * np.sqrt(np.maximum(xc, 0.0)) - 0.1260 * xc - 0.3516 * xc**2
                    + 0.2843 * xc**3 - 0.1036 * xc**4)
    yc = np.zeros_like(xc)
    dyc = np.zeros_like(xc)
    if m > 0.0 and p > 0.0:
        fwd = xc < p
        yc[fwd] = m / p**2 * (2 * p * xc[fwd] - xc[fwd] ** 2)
        dyc[fwd] = 2 * m / p**2 * (p - xc[fwd])
        aft = ~fwd
        yc[aft] = m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * xc[aft] - xc[aft] ** 2)
        dyc[aft] = 2 * m / (1 - p) ** 2 * (p - xc[aft])
    return yt, yc, dyc


def hicks_henne(xc: np.ndarray, peaks, amps, width: float = 3.0) -> np.ndarray:
    """Sum of Hicks-Henne bumps a_i * sin(pi * x^(ln 0.5 / ln h_i))^w on x in [0, 1]."""
    xc = np.clip(np.asarray(xc, dtype=np.float64), 0.0, 1.0)
    out = np.zeros_like(xc)
    for h, a in zip(peaks, amps):
        e = math.log(0.5) / math.log(h)
        out += a * np.sin(math.pi * xc**e) ** width
    return out


def _airfoil_points(m, p, t, nhalf, bumps_upper=None, bumps_lower=None):
    """Lower TE -> LE -> upper TE, 2*nhalf-1 points, cosine clustered at both ends."""
    beta = np.linspace(0.0, math.pi, nhalf)
    xc = 0.5 * (1.0 - np.cos(beta))                     # 0 (LE) .. 1 (TE)
    yt, yc, dyc = naca4_surface(m, p, t, xc)
    th = np.arctan(dyc)
    xu, yu = xc - yt * np.sin(th), yc + yt * np.cos(th)
    xl, yl = xc + yt * np.sin(th), yc - yt * np.cos(th)
    if bumps_upper is not None:
        yu = yu + hicks_henne(xu, *bumps_upper)
    if bumps_lower is not None:
        yl = yl - hicks_henne(xl, *bumps_lower)
    # exact LE at the origin and closed TE at (1, 0), as in the shipped grid
    xu[0] = xl[0] = 0.0
    yu[0] = yl[0] = 0.0
    xu[-1] = xl[-1] = 1.0
    yu[-1] = yl[-1] = 0.0
    xs = np.concatenate([xl[::-1], xu[1:]])
    ys = np.concatenate([yl[::-1], yu[1:]])
    return xs, ys


def _geometric(n: int, first: float, total: float) -> np.ndarray:
    """n points from 0 to total with first spacing `first` and constant growth ratio."""
    if n < 2:
        raise ValueError("need at least two points")
    if first * (n - 1) >= total:
        return np.linspace(0.0, total, n)
    lo, hi = 1.0 + 1e-12, min(10.0, math.exp(600.0 / (n - 1)))
    for _ in range(200):
        r = 0.5 * (lo + hi)
        s = first * (r ** (n - 1) - 1.0) / (r - 1.0)
        if s > total:
            hi = r
        else:
            lo = r
    r = 0.5 * (lo + hi)
    d = first * r ** np.arange(n - 1)
    s = np.concatenate([[0.0], np.cumsum(d)])
    return s * (total / s[-1])


def make_cgrid(jmax: int, kmax: int, m: float = 0.04, p: float = 0.4, t: float = 0.12,
               wall_spacing: float = 1e-5, farfield: float = 15.0, wake_fraction: float = 0.17,
               wake_spacing_out: float = 0.01, bumps_upper=None, bumps_lower=None,
               smooth_normals: float = 0.02) -> CGrid:
    """Algebraic C-grid: surface/cut at k=1, far field of radius `farfield` chords.

    Each j-line is a cubic Hermite curve leaving the surface along its normal and meeting the
    outer boundary, with points distributed by geometric stretching from `wall_spacing`
    (the first spacing grows linearly along the wake cut to `wake_spacing_out`).
    """
    if jmax < 11 or kmax < 5 or jmax % 2 == 0:
        raise ValueError("jmax must be odd and >= 11, kmax >= 5")
    nwake = max(2, int(round(wake_fraction * jmax)))
    naf = jmax - 2 * nwake                      # airfoil points incl. both TE copies
    if naf % 2 == 0:
        nwake -= 1
        naf += 2
    nhalf = (naf + 1) // 2
    xs, ys = _airfoil_points(m, p, t, nhalf, bumps_upper, bumps_lower)
    # wake cut: from TE along y=0, spacing growing from the TE chordwise spacing
    dte = math.hypot(xs[1] - xs[0], ys[1] - ys[0])
    xw = 1.0 + _geometric(nwake + 1, dte, farfield)[1:]
    xin = np.concatenate([xw[::-1], xs, xw])
    yin = np.concatenate([np.zeros(nwake), ys, np.zeros(nwake)])
    jte_l, jte_u = nwake + 1, nwake + naf
    # j-lines are straight rays along a smoothed surface normal, out to the C-shaped far
    # field (lines y=-R / y=+R behind x=1 joined by a semicircle of radius R round (1, 0));
    # the normal angle is smoothed along j and kept monotone so that no two rays cross
    R = farfield
    tx = np.gradient(xin)
    ty = np.gradient(yin)
    # the outward normal (-ty, tx) turns clockwise from -pi/2 (lower cut) through -pi (LE)
    # to -3pi/2 (upper cut)
    th = np.unwrap(np.arctan2(tx, -ty))
    th[:nwake] = -0.5 * math.pi
    th[jmax - nwake:] = -1.5 * math.pi
    sig = max(1.0, smooth_normals * naf)
    half = 4 * int(math.ceil(sig))
    w = np.exp(-0.5 * (np.arange(-half, half + 1) / sig) ** 2)
    w /= w.sum()
    for _ in range(3):
        pad = np.concatenate([np.full(half, th[0]), th, np.full(half, th[-1])])
        sm = np.convolve(pad, w, mode="valid")
        sm[:nwake] = -0.5 * math.pi
        sm[jmax - nwake:] = -1.5 * math.pi
        th = sm
    th = np.minimum.accumulate(th)
    dxr, dyr = np.cos(th), np.sin(th)
    x = np.empty((jmax, kmax), order="F")
    y = np.empty((jmax, kmax), order="F")
    for j in range(jmax):
        px, py, ux, uy = xin[j], yin[j], dxr[j], dyr[j]
        tq = None
        if abs(uy) > 1e-14:
            for yl in (-R, R):
                tl = (yl - py) / uy
                if tl > 0 and px + tl * ux >= 1.0:
                    tq = tl if tq is None else min(tq, tl)
        # circle |p + t u - (1,0)| = R
        bx, by = px - 1.0, py
        bb = bx * ux + by * uy
        cc = bx * bx + by * by - R * R
        tc = -bb + math.sqrt(bb * bb - cc)
        if px + tc * ux < 1.0:
            tq = tc if tq is None else min(tq, tc)
        if tq is None:
            raise ValueError("ray does not reach the far field")
        ds0 = wall_spacing
        if j < nwake or j >= jmax - nwake:
            frac = (xin[j] - 1.0) / max(farfield, 1e-12)
            ds0 = wall_spacing + (wake_spacing_out - wall_spacing) * frac
        sk = _geometric(kmax, ds0, tq)
        x[j, :] = px + sk * ux
        y[j, :] = py + sk * uy
    # the cut: lower and upper wake lines share their k=1 points exactly
    for j in range(nwake):
        x[jmax - 1 - j, 0] = x[j, 0]
        y[jmax - 1 - j, 0] = y[j, 0]
    return CGrid(x=x, y=y, jte_lower=jte_l, jte_upper=jte_u)


def cgrid_bc_regions(jte_lower: int, jte_upper: int):
    """BC regions of a one-zone C-grid, in the order and index convention of the shipped
    ``ref/bcmain.dat_original_119x31`` (31 outflow, 0 no-slip wall, 25 far field with point
    vortex, 60 wake cut)."""
    return [
        (31, 1, 1, 1, 1, -1),
        (31, 1, -1, -1, 1, -1),
        (0, 1, jte_lower, jte_upper, 1, 1),
        (25, 1, 2, -2, -1, -1),
        (60, 1, 2, jte_lower - 1, 1, 1),
        (60, 1, jte_upper + 1, -2, 1, 1),
    ]


def write_plot3d(path: str, x: np.ndarray, y: np.ndarray) -> None:
    """Formatted single-zone PLOT3D at full precision (round-trips float64 exactly)."""
    jmax, kmax = x.shape
    with open(path, "w") as f:
        f.write("  1\n")
        f.write(f"  {jmax}    {kmax}\n")
        for arr in (x, y):
            flat = np.asarray(arr).ravel(order="F")
            for i in range(0, flat.size, 4):
                f.write(" ".join(f"{v:.17e}" for v in flat[i:i + 4]) + "\n")


def read_plot3d(path: str):
    tok = open(path).read().split()
    nz = int(tok[0])
    if nz != 1:
        raise ValueError("only single-zone grids are supported")
    jmax, kmax = int(tok[1]), int(tok[2])
    v = np.array([float(a.replace("D", "E").replace("d", "e")) for a in tok[3:3 + 2 * jmax * kmax]])
    x = v[: jmax * kmax].reshape((jmax, kmax), order="F")
    y = v[jmax * kmax:].reshape((jmax, kmax), order="F")
    return np.asfortranarray(x), np.asfortranarray(y)


def write_bcmain(path: str, regions, title: str = "generated C-grid") -> None:
    with open(path, "w") as f:
        f.write("c" + "*" * 70 + "\nc\n")
        f.write(f"c  {title}\n")
        f.write("c  ibcval    zone	jbeg jend	kbeg kend\nc\n")
        for r in regions:
            f.write("\t" + "\t".join(str(v) for v in r) + "\n")


def read_bcmain(path: str):
    """Regions as (ibcval, zone, jbeg, jend, kbeg, kend); lines starting with c/C are
    comments (``ref/rbc.f:326-327``)."""
    regs = []
    for line in open(path):
        if not line.strip() or line[0] in "cC":
            continue
        vals = line.replace("\t", " ").split()
        if len(vals) >= 6:
            regs.append(tuple(int(v) for v in vals[:6]))
    return regs


def write_namelist(path: str, *, alpha: float, beta: float, reynum: float, ntmax: int,
                   impsch: int = 1, iturb: int = 1, ivis: int = 1, dtau: float = 0.1,
                   dt: float = 1.0e12, underr: float = 0.01, grid: str = "xy.dat",
                   bcfile: str = "bcmain.dat", title: str = "miniSMAC2D case",
                   geometry: str = "NACA 4412 Airfoil") -> None:
    def d(v: float) -> str:
        return repr(float(v)).replace("e", "d") if "e" in repr(float(v)) else repr(float(v)) + "d0"
    with open(path, "w") as f:
        f.write(" $input\n")
        f.write(f"  alpha={d(alpha)}\n  beta={d(beta)}\n  dcoef2=0.0d0\n  dcoef4=0.0d0\n")
        f.write(f"  dt={d(dt)}\n  dtau={d(dtau)}\n  geometry='{geometry}'\n")
        f.write(f"  impsch={impsch}\n  iturb={iturb}\n  ivis={ivis}\n  kadd=0\n")
        f.write(f"  ntmax={ntmax}\n  reynum={d(reynum)}\n  title='{title}'\n")
        f.write(f"  underr={d(underr)}\n  zz_run_default_input=.false.\n")
        f.write("  write_plot3d_q_file=.false.\n")
        f.write(f"  bcmain_directory='.'\n  bcmain_filename='{os.path.basename(bcfile)}'\n")
        f.write(f"  xygrid_directory='.'\n  xygrid_filename='{os.path.basename(grid)}'\n")
        f.write(" $end\n")


def metric(x: np.ndarray, y: np.ndarray):
    """Grid metrics as ``ref/metric.f:49-278`` forms them.

    Returns ``rtxy`` of shape (jmax, kmax, 2, 3) and ``dj`` (jmax, kmax), Fortran order:
    ``rtxy[:,:,0,1:] = (y_eta, -x_eta)`` (xi_x, xi_y), ``rtxy[:,:,1,1:] = (-y_xi, x_xi)``
    (eta_x, eta_y), time metrics zero, ``dj = 1/(x_xi*y_eta - x_eta*y_xi)``.  Second-order
    central differences inside, second-order one-sided at the edges, with the reference's
    operation order so the values are bit-identical.
    """
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    jmax, kmax = x.shape

    def dxi(a):
        d = np.empty_like(a)
        d[1:-1] = 0.5 * (a[2:] - a[:-2])
        d[0] = 0.5 * ((-a[2] + 4.0 * a[1]) - 3.0 * a[0])
        d[-1] = 0.5 * ((3.0 * a[-1] - 4.0 * a[-2]) + a[-3])
        return d

    def deta(a):
        d = np.empty_like(a)
        d[:, 1:-1] = 0.5 * (a[:, 2:] - a[:, :-2])
        d[:, 0] = 0.5 * ((-a[:, 2] + 4.0 * a[:, 1]) - 3.0 * a[:, 0])
        d[:, -1] = 0.5 * ((3.0 * a[:, -1] - 4.0 * a[:, -2]) + a[:, -3])
        return d

    xj, yj, xk, yk = dxi(x), dxi(y), deta(x), deta(y)
    rdj = xj * yk - xk * yj
    if np.any(rdj == 0.0):
        raise ValueError("zero grid Jacobian")
    dj = np.asfortranarray(1.0 / rdj)
    rtxy = np.zeros((jmax, kmax, 2, 3), order="F")
    rtxy[:, :, 0, 1] = yk
    rtxy[:, :, 0, 2] = -xk
    rtxy[:, :, 1, 1] = -yj
    rtxy[:, :, 1, 2] = xj
    return rtxy, dj


BUMP_PEAKS = (0.1, 0.3, 0.5, 0.7)


def bump_modes(x: np.ndarray, y: np.ndarray, jte_lower: int, jte_upper: int,
               peaks=BUMP_PEAKS, width: float = 3.0):
    """Shape modes of a C-grid: Hicks-Henne bumps on the upper surface (one per peak), then on
    the lower surface, each carried into the volume by algebraic mesh deformation.

    A bump moves the surface points of its side normal to the chord by
    ``sin(pi x^(ln 0.5/ln h))^width`` (upward on the upper side, downward on the lower side:
    positive amplitudes thicken the section, as in ``make_cgrid``'s bumps); a grid point on
    the j-line of a surface point moves by the same amount times ``(1 - s/S)^3``, s being its
    distance along the line from the surface and S that of the far-field point, so the far
    field, the wake cut and the trailing and leading edges stay fixed.  The deformation is
    linear in the amplitudes: the grid of amplitudes ``a`` is ``(x, y + sum_i a_i dy_i)``.
    Returns a list of ``(dx, dy)`` pairs, ``2 * len(peaks)`` of them.
    """
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    jmax, kmax = x.shape
    jl, ju = jte_lower - 1, jte_upper - 1          # 0-based trailing-edge points
    jle = (jl + ju) // 2                            # leading edge (middle of the wall)
    s = np.sqrt((x - x[:, :1]) ** 2 + (y - y[:, :1]) ** 2)
    w = (1.0 - s / s[:, -1:]) ** 3
    modes = []
    for side in (+1, -1):
        js = np.arange(jle, ju + 1) if side > 0 else np.arange(jl, jle + 1)
        for h in peaks:
            e = math.log(0.5) / math.log(h)
            xs = np.clip(x[js, 0], 0.0, 1.0)
            dyw = np.zeros(jmax)
            dyw[js] = side * np.sin(math.pi * xs ** e) ** width
            dy = np.asfortranarray(dyw[:, None] * w)
            modes.append((np.zeros_like(x, order="F"), dy))
    return modes


def deform(x: np.ndarray, y: np.ndarray, modes, amps):
    """The grid of shape amplitudes ``amps`` (one per mode of ``bump_modes``)."""
    xn = np.array(x, dtype=np.float64, order="F")
    yn = np.array(y, dtype=np.float64, order="F")
    for (dx, dy), a in zip(modes, amps):
        if a:
            xn += a * dx
            yn += a * dy
    return xn, yn
