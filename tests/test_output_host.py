"""The output side (minismac_amd/output.py) on the CPU: the Fortran edit descriptors the
reference's XX line and q file use, the XX line's columns, and a Q2D.p3d round trip."""
import numpy as np
import yaml

from minismac_amd import output
from minismac_amd.case import config_case


def test_fortran_edit_descriptors():
    # 1PE11.3 / 1PE12.4 as the reference's format 110 prints them
    assert output.fortran_e(1.0, 11, 3) == "  1.000E+00"
    assert output.fortran_e(-2.5e-7, 11, 3) == " -2.500E-07"
    assert output.fortran_e(1.234567e123, 12, 4) == "  1.2346+123"
    assert output.fortran_e(0.0, 12, 4) == "  0.0000E+00"
    # F10.6 and F5.2
    assert output.fortran_f(0.5, 10, 6) == "  0.500000"
    assert output.fortran_f(-1.25, 10, 6) == " -1.250000"
    assert output.fortran_f(0.0, 5, 2) == " 0.00"
    assert output.fortran_f(12345.0, 10, 6) == "**********"


def test_xx_line_columns():
    st = {"nt": 17, "resmax": 2.0e-3, "resmax0": 4.0, "divmax": 1.5e-2, "turres": -1.0,
          "jres": 21, "kres": 3, "cd": 0.0123, "cl": 1.25}
    line = output.xx_line(st, cputime=0.5)
    assert line.startswith("XX   17")
    assert len(line) == 2 + 5 + 3 * 11 + 3 * 6 + 3 * 12
    fields = line[2:].split()
    assert int(fields[0]) == 17
    assert float(fields[1]) == 5.0e-4                  # resmax / resmax0
    assert float(fields[2]) == 1.5e-2
    assert float(fields[3]) == 0.0                      # turres floored at 0 as max(0, turres)
    assert [int(v) for v in fields[4:7]] == [21, 3, 0]
    assert float(fields[8]) == 0.0123 and float(fields[9]) == 1.25


def test_q2d_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    q = rng.uniform(-1.0, 1.5, (13, 7, 3))
    path = tmp_path / "Q2D.p3d"
    output.write_q2d(str(path), q, alpha=4.0, reynum=1.0e6, cputime=3.5)
    head, planes = output.read_q2d(str(path))
    assert head == [0.0, 4.0, 1.0e6, 3.5]
    np.testing.assert_array_equal(planes[:, :, 0], 1.0)
    np.testing.assert_allclose(planes[:, :, 1], q[:, :, 1], atol=5e-7)
    np.testing.assert_allclose(planes[:, :, 2], q[:, :, 2], atol=5e-7)
    e = q[:, :, 0] / float(np.float32(0.4)) + 0.5 * (q[:, :, 1] ** 2 + q[:, :, 2] ** 2)
    np.testing.assert_allclose(planes[:, :, 3], e, atol=5e-7)


def test_yaml_record_parses(tmp_path):
    case = config_case(0, impsch=2)
    path = tmp_path / "run.yaml"
    output.write_yaml(str(path), case, niter=500, solve_sec=2.0)
    d = yaml.safe_load(open(path))
    assert d["j-dim"] == case.jmax and d["k-dim"] == case.kmax
    assert d["implicit_scheme"] == "lusgs factorization"
    assert abs(d["microsec_per_iter_per_point"] - 2.0 / 500 / case.npts * 1e6) < 1e-4
