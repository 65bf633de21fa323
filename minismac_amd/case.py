"""A case = grid + metrics + BC regions + solver settings, and its reference input files.

The configurations the benchmarks and parity tests use are built here:
  * ``shipped_case``: the reference's own 119x31 NACA 4412 grid and BC file (when a copy of
    the reference inputs is given), with the shipped namelist values;
  * ``naca_case``: a generated NACA 4-digit C-grid at any size with the config values.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from . import grid as _grid
from ._abi import SmacBcRegion, SmacParams, make_params, params_dict


@dataclasses.dataclass
class Case:
    x: np.ndarray                 # (jmax, kmax) Fortran order
    y: np.ndarray
    regions: list                 # raw (ibcval, zone, jbeg, jend, kbeg, kend), file convention
    params: SmacParams
    name: str = "case"

    @property
    def jmax(self) -> int:
        return self.x.shape[0]

    @property
    def kmax(self) -> int:
        return self.x.shape[1]

    @property
    def npts(self) -> int:
        return self.jmax * self.kmax

    def metrics(self):
        return _grid.metric(self.x, self.y)

    def resolved_regions(self):
        """Indices with -1/-2 resolved to max/max-1 (ref/rbc.f:384-396)."""
        out = []
        for (ib, _zone, jb, je, kb, ke) in self.regions:
            def res(v, vmax):
                return vmax if v == -1 else (vmax - 1 if v == -2 else v)
            out.append((ib, res(jb, self.jmax), res(je, self.jmax), res(kb, self.kmax),
                        res(ke, self.kmax)))
        return out

    def c_regions(self):
        regs = self.resolved_regions()
        arr = (SmacBcRegion * len(regs))()
        for i, (ib, jb, je, kb, ke) in enumerate(regs):
            arr[i].ibcval, arr[i].jbeg, arr[i].jend, arr[i].kbeg, arr[i].kend = ib, jb, je, kb, ke
        return arr

    def wake_pairs(self) -> np.ndarray:
        """(jt, jb) pairs of the wake cut in the order the reference injects them
        (ref/rbc.f:1160-1270 for two equal cut segments on k=1)."""
        segs = [r for r in self.resolved_regions() if r[0] == 60]
        if len(segs) < 2:
            return np.zeros((0, 2), dtype=np.int32)
        (_, a0, a1, _, _), (_, b0, b1, _, _) = segs[0], segs[1]
        n = min(a1 - a0 + 1, b1 - b0 + 1)
        pairs = [(a0 + i, b1 - i) for i in range(n)] + [(b0 + i, a1 - i) for i in range(n)]
        return np.asarray(pairs, dtype=np.int32)

    def write_reference_inputs(self, directory: str, ntmax: int) -> None:
        """xy.dat, bcmain.dat and smac2d.in for the reference executable."""
        os.makedirs(directory, exist_ok=True)
        _grid.write_plot3d(os.path.join(directory, "xy.dat"), self.x, self.y)
        _grid.write_bcmain(os.path.join(directory, "bcmain.dat"), self.regions, self.name)
        p = params_dict(self.params)
        _grid.write_namelist(os.path.join(directory, "smac2d.in"), alpha=p["alpha"],
                             beta=p["beta"], reynum=p["reynum"], ntmax=ntmax,
                             impsch=p["impsch"], iturb=p["iturb"], ivis=p["ivis"],
                             dtau=p["dtau"], dt=p["dt"], underr=p["underr"],
                             title=self.name[:60])


def naca_case(jmax: int, kmax: int, *, m=0.04, p=0.4, t=0.12, wall_spacing=1e-5,
              farfield=15.0, name=None, bumps_upper=None, bumps_lower=None, **params) -> Case:
    g = _grid.make_cgrid(jmax, kmax, m=m, p=p, t=t, wall_spacing=wall_spacing,
                         farfield=farfield, bumps_upper=bumps_upper, bumps_lower=bumps_lower)
    regs = _grid.cgrid_bc_regions(g.jte_lower, g.jte_upper)
    nm = name or f"NACA {int(round(m*100))}{int(round(p*10))}{int(round(t*100)):02d} {jmax}x{kmax}"
    return Case(x=g.x, y=g.y, regions=regs, params=make_params(**params), name=nm)


def config_case(index: int, **overrides) -> Case:
    """The workload configurations of the benchmark definition:
    0: 119x31, 1e-5c wall, 15c far field, beta=10, Re=1e6, alpha=4, 500 iterations;
    1: 945x241, 2e-6c wall, Re=3e6, alpha=8;
    2: 1889x481, 1e-6c wall, Re=6e6, alpha=12."""
    table = {
        0: dict(jmax=119, kmax=31, wall_spacing=1e-5, alpha=4.0, reynum=1.0e6),
        1: dict(jmax=945, kmax=241, wall_spacing=2e-6, alpha=8.0, reynum=3.0e6),
        2: dict(jmax=1889, kmax=481, wall_spacing=1e-6, alpha=12.0, reynum=6.0e6),
    }
    c = dict(table[index])
    c.update(overrides)
    jmax, kmax, ws = c.pop("jmax"), c.pop("kmax"), c.pop("wall_spacing")
    c.setdefault("beta", 10.0)
    return naca_case(jmax, kmax, wall_spacing=ws, **c)


BATCH_SIZES = ((239, 61, 1.0e-5), (477, 121, 5.0e-6), (945, 241, 2.0e-6))


def batch_cases(n: int = 64, seed: int = 0, **overrides) -> list[Case]:
    """The batched workload: n NACA 4-digit C-grid cases, seeded (numpy PCG64).
    m ~ U[0, 0.06], p ~ U[0.2, 0.6], t ~ U[0.08, 0.18], size uniform over 239x61, 477x121 and
    945x241 (wall spacing 1e-5c, 5e-6c, 2e-6c), alpha ~ U[-4, 16] deg, log10 Re ~ U[5, 7];
    beta = 10 and far field 15c as in the single-case workloads."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        m, pp, t = rng.uniform(0.0, 0.06), rng.uniform(0.2, 0.6), rng.uniform(0.08, 0.18)
        jm, km, ws = BATCH_SIZES[int(rng.integers(0, len(BATCH_SIZES)))]
        alpha, lre = rng.uniform(-4.0, 16.0), rng.uniform(5.0, 7.0)
        c = dict(alpha=alpha, reynum=10.0 ** lre, beta=10.0)
        c.update(overrides)
        out.append(naca_case(jm, km, m=m, p=pp, t=t, wall_spacing=ws,
                             name=f"batch {i}: m={m:.3f} p={pp:.2f} t={t:.3f} {jm}x{km} "
                                  f"a={alpha:.2f} Re={10.0 ** lre:.3g}", **c))
    return out


def shipped_case(xy_path: str, bc_path: str, **params) -> Case:
    """The reference's shipped case from copies of its two input files."""
    x, y = _grid.read_plot3d(xy_path)
    regs = _grid.read_bcmain(bc_path)
    base = dict(alpha=10.0, beta=5.0, reynum=1.5e6, dtau=0.1, dt=1.0e12, underr=0.01)
    base.update(params)
    return Case(x=x, y=y, regions=regs, params=make_params(**base), name="NACA 4412 119x31")


def fortran(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(C.POINTER(C.c_double))
