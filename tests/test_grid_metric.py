"""Grid generation, input writers and metrics (the input side of the implicit iteration)."""
import os

import numpy as np
import pytest

import pyoracle
from minismac_amd import grid
from minismac_amd.case import config_case, shipped_case

from conftest import GOLDEN


def test_metric_affine_grid_is_exact():
    """On x = a*j + b*k, y = c*j + d*k every difference formula is exact (ref/metric.f)."""
    jm, km = 9, 7
    J, K = np.meshgrid(np.arange(jm, dtype=float), np.arange(km, dtype=float), indexing="ij")
    a, b, c, d = 0.5, 0.25, -0.125, 2.0
    x = np.asfortranarray(a * J + b * K)
    y = np.asfortranarray(c * J + d * K)
    rtxy, dj = grid.metric(x, y)
    assert np.all(rtxy[:, :, 0, 1] == d) and np.all(rtxy[:, :, 0, 2] == -b)
    assert np.all(rtxy[:, :, 1, 1] == -c) and np.all(rtxy[:, :, 1, 2] == a)
    assert np.all(dj == 1.0 / (a * d - b * c))
    assert np.all(rtxy[:, :, :, 0] == 0.0)


def test_metric_matches_oracle_bitwise():
    case = config_case(0)
    r1, d1 = grid.metric(case.x, case.y)
    r2, d2 = pyoracle.metric(case.x, case.y)
    assert np.array_equal(r1, r2) and np.array_equal(d1, d2)


@pytest.mark.parametrize("jm,km,ws", [(119, 31, 1e-5), (239, 61, 1e-5), (477, 121, 1e-5),
                                      (945, 241, 2e-6), (1889, 481, 1e-6)])
def test_generated_cgrids_are_valid(jm, km, ws):
    g = grid.make_cgrid(jm, km, wall_spacing=ws)
    _, dj = grid.metric(g.x, g.y)
    assert np.all(dj > 0)
    jle = (jm - 1) // 2
    d0 = np.hypot(g.x[jle, 1] - g.x[jle, 0], g.y[jle, 1] - g.y[jle, 0])
    assert d0 == pytest.approx(ws, rel=1e-6)
    # far field 15 chords from the trailing edge, cut lines coincide
    assert np.allclose(np.hypot(g.x[:, -1] - 1.0, g.y[:, -1])[jle], 15.0, rtol=1e-9)
    nw = g.jte_lower - 1
    assert np.array_equal(g.x[:nw, 0], g.x[::-1][:nw, 0]) and np.all(g.y[:nw, 0] == 0.0)
    ratio = np.hypot(np.diff(g.x[jle]), np.diff(g.y[jle]))
    # the 945x241 benchmark grid asks for stretching <= 1.15; coarser ones are looser
    assert (ratio[1:] / ratio[:-1]).max() < (1.15 if km >= 200 else 1.7)


def test_plot3d_and_bc_roundtrip(tmp_path):
    g = grid.make_cgrid(61, 21)
    p = tmp_path / "xy.dat"
    grid.write_plot3d(str(p), g.x, g.y)
    x, y = grid.read_plot3d(str(p))
    assert np.array_equal(x, g.x) and np.array_equal(y, g.y)
    regs = grid.cgrid_bc_regions(g.jte_lower, g.jte_upper)
    b = tmp_path / "bc.dat"
    grid.write_bcmain(str(b), regs)
    assert grid.read_bcmain(str(b)) == regs


def test_shipped_case_regions_resolve_like_rbc():
    case = shipped_case(os.path.join(GOLDEN, "shipped_xy_119x31.dat"),
                        os.path.join(GOLDEN, "shipped_bcmain_119x31.dat"))
    # ref/rbc.f:384-396 re-indexing as the reference prints it for this file
    assert case.resolved_regions() == [(31, 1, 1, 1, 31), (31, 119, 119, 1, 31),
                                       (0, 21, 99, 1, 1), (25, 2, 118, 31, 31),
                                       (60, 2, 20, 1, 1), (60, 100, 118, 1, 1)]
