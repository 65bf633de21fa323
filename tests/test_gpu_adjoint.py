"""The discrete adjoint (smac_adjoint): gradients of CL and CD in alpha and in Hicks-Henne
shape modes (algebraic mesh deformation) at a steady state.

The reference evaluates the viscous terms in default REAL (FP32, ref/visc.f:68), so its
residual is a smooth function rounded to FP32 in those terms: differences of it carry that
rounding, and every iteration stalls in a noise band (a drop near 1e-11 here).  The dual
numbers of the GPU Jacobian differentiate the unrounded function, which is what the
FP64 build of the oracle (oracle/_ref/libsmac_oracle_fp64.so: the same restatement with
those temporaries in FP64) evaluates.  So:
  - the Jacobian (9 cross-stencil blocks per row) against central differences of the FP64
    oracle's residual, to 1e-6 (1e-8 inviscid), and the residual itself bit for bit against
    the oracle's;
  - its products: F_q^T against the transpose of those blocks, P^T against the primal
    update P in dot products <P v, w> = <v, P^T w>, and P applied to the steady residual
    against the oracle's own update of the same state;
  - its gradients against the FP64 CPU tangent (the FP64 oracle differentiated by differences
    and solved directly, tests/adjoint_cpu.py) to 2e-6 of the largest component, and against
    central differences of converged solutions: alpha with a step of 0.02 deg to 2e-5, two
    bump amplitudes by Richardson extrapolation of steps 1e-3 and 5e-4 chord to 1e-4 of the
    largest component (the FP32 rounding moves converged CL/CD by about 5e-8 when the grid
    moves, which bounds what a difference can resolve: about 1e-4 in a derivative here);
  - a torch.autograd.Function around the converged solve whose backward is the adjoint.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import pyoracle
from minismac_amd import grid
from minismac_amd._abi import SMAC_ORDER_ZEBRA
from minismac_amd.case import config_case
from minismac_amd.solver import SmacError, Solver

import adjoint_cpu

pytestmark = pytest.mark.gpu

OFFS = [(0, 0), (-1, 0), (1, 0), (-2, 0), (2, 0), (0, -1), (0, 1), (0, -2), (0, 2)]


def _sparse_from_blocks(jac, pairs, jm, km):
    """F_q of the fetched blocks (jm, km, 81) plus the wake cut's partner terms."""
    N = jm * km
    jj, kk = np.meshgrid(np.arange(jm), np.arange(km), indexing="ij")
    rows, cols, vals = [], [], []
    for o, (dj, dk) in enumerate(OFFS):
        pj, pk = jj + dj, kk + dk
        ok = (pj >= 0) & (pj < jm) & (pk >= 0) & (pk < km)
        for m in range(3):
            for n in range(3):
                v = jac[:, :, o * 9 + m + 3 * n]
                sel = ok & (v != 0.0)
                rows.append((m * N + jj + jm * kk)[sel])
                cols.append((n * N + pj + jm * pk)[sel])
                vals.append(v[sel])
    for jt, jb in pairs:
        for m in range(3):
            rows.append(np.array([m * N + jt]))
            cols.append(np.array([m * N + jb + jm]))
            vals.append(np.array([0.5]))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(3 * N, 3 * N))


def _unsteady_state(ivis, npass=30):
    case = config_case(0, impsch=2, reynum=1000.0, dtau=10.0, iturb=0, ivis=ivis)
    o = pyoracle.Oracle(case)
    for _ in range(npass):
        st = o.step()
    return case, o, st


@pytest.mark.parametrize("ivis,tol", [(0, 1e-8), (1, 1e-6)])
def test_adjoint_jacobian_matches_oracle_differences(ivis, tol):
    case, o, ost = _unsteady_state(ivis)
    q = o.array("q")
    pairs = o.wake_pairs() - 1
    with Solver(case) as g:
        g.run(30)                                    # bit-identical to the oracle's state
        assert np.array_equal(g.get_state()[0], q)
        g.adjoint((), maxit=2, tol=1e-3, check=False)  # builds the operators at this state
        jac = g.array("adj_jac")
        s_gpu = g.array("s")
        jt_w = [g.adjoint_apply(0, w) for w in np.random.default_rng(1).standard_normal((2,) + q.shape)]
    S = adjoint_cpu.Steady(case)
    Jc = adjoint_cpu.jacobian(adjoint_cpu.Steady(case, pyoracle.fp64_lib()), q, ost["cl"],
                              rel=1e-6)
    Jg = _sparse_from_blocks(jac, pairs, case.jmax, case.kmax)
    rng = np.random.default_rng(0)
    for _ in range(3):
        v = rng.standard_normal(3 * case.npts)
        a, b = Jg @ v, Jc @ v
        assert np.abs(a - b).max() <= tol * np.abs(b).max()
    # the residual the blocks belong to is the oracle's (bit for bit, away from the cut rows)
    s_orc = S.F(q, ost["cl"])
    keep = np.ones(q.shape, dtype=bool)
    keep[pairs[:, 0], 0, :] = False
    assert np.array_equal(s_gpu[keep], s_orc[keep])
    # F_q^T on the GPU is the transpose of the blocks
    for i, w in enumerate(np.random.default_rng(1).standard_normal((2,) + q.shape)):
        want = Jg.T @ w.ravel(order="F")
        assert np.abs(jt_w[i].ravel(order="F") - want).max() <= 1e-13 * np.abs(want).max()


def test_lusgs_adjoint_solves_side_by_side_agree_with_one_step():
    """The LU-SGS-preconditioned adjoint (each of the two side-by-side solves with its own
    band workspace): the CL and CD solves at a 119x31 LU-SGS state give the same dCL/dalpha
    and dCD/dalpha with one preconditioner step and with three, to the solve tolerance."""
    case = config_case(0, impsch=2, reynum=1000.0, dtau=10.0)
    out = []
    for k in (1, 3):
        with Solver(case, options={"adj_prec_steps": k}) as g:
            g.run(40)
            dcl, dcd, ast = g.adjoint((), maxit=6000, tol=1e-10)
        print(f"LU-SGS adjoint, {k} step(s): iterations {ast['iters']}, resid {ast['resid']}",
              flush=True)
        assert max(ast["resid"]) < 1e-9
        out.append((dcl[0], dcd[0], ast["iters"]))
    (cl1, cd1, it1), (cl3, cd3, it3) = out
    assert cl3 == pytest.approx(cl1, rel=1e-7) and cd3 == pytest.approx(cd1, rel=1e-7)
    assert all(a < b for a, b in zip(it3, it1))


def test_adjoint_update_is_the_primal_update_and_its_transpose():
    case, o, ost = _unsteady_state(1)
    q = o.array("q")
    S = adjoint_cpu.Steady(case)
    F = S.F(q, ost["cl"])
    o.assemble()
    o.solve()
    dq_orc = o.array("dq")
    rng = np.random.default_rng(2)
    with Solver(case) as g:
        g.run(30)
        g.adjoint((), maxit=2, tol=1e-3, check=False)
        # P F(q) is the iteration's own update of the state (LU-SGS solve + wake injection)
        pf = g.adjoint_apply(3, F)
        assert np.abs(pf - dq_orc).max() <= 1e-13 * np.abs(dq_orc).max()
        for _ in range(3):
            v, w = rng.standard_normal((2,) + q.shape)
            lhs = np.vdot(g.adjoint_apply(3, v), w)
            rhs = np.vdot(v, g.adjoint_apply(2, w))
            assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_adjoint_line_relaxation_update_and_its_transpose():
    """With line relaxation the preconditioner is the zebra sweeps of the state's blocks with
    the wake injection at the end, and its exact transpose.  With one xi sweep and no eta
    sweep that is the primal's own pass, so P F(q) is the pass's dq; with the workload's
    sweeps (two each way, under-relaxed) P and P^T agree in dot products."""
    rng = np.random.default_rng(5)
    case = _steady_case(njsp=1, nksp=0)
    S = adjoint_cpu.Steady(case)
    with Solver(case) as g:
        st, _ = g.run(40)
        q, _, _ = g.get_state()
        F = S.F(q, st["cl"])
        g.adjoint((), maxit=2, tol=1e-3, check=False)
        pf = g.adjoint_apply(3, F)
        g.set_state(q=q, nt=41, resmax0=1.0, clift=st["cl"])
        g.step()
        dq = g.array("dq")
    assert np.abs(pf - dq).max() <= 1e-12 * np.abs(dq).max()
    with Solver(_steady_case(underr=0.8)) as g:
        g.run(40)
        with pytest.raises(SmacError) as ei:          # a solve stopped short is an error
            g.adjoint((), maxit=2, tol=1e-12)
        assert ei.value.status == 11
        g.adjoint((), maxit=2, tol=1e-3, check=False)
        for _ in range(3):
            v, w = rng.standard_normal((2,) + q.shape)
            lhs = np.vdot(g.adjoint_apply(3, v), w)
            rhs = np.vdot(v, g.adjoint_apply(2, w))
            assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


STEADY = dict(reynum=1000.0, dtau=10.0, underr=1.0, iturb=0, line_order=SMAC_ORDER_ZEBRA)


def _steady_case(**kw):
    return config_case(0, impsch=1, **{**STEADY, **kw})


def _modes(case):
    jl, ju = [(r[1], r[2]) for r in case.resolved_regions() if r[0] == 0][0]
    return grid.bump_modes(case.x, case.y, jl, ju)


def _converged(case, q0=None, cl0=0.0):
    """From the free stream to a 1e-11 drop (or the FP32 noise band); from a nearby state
    (q0), 4000 passes, which take a perturbation of the state into that band."""
    with Solver(case) as g:
        if q0 is not None:
            g.set_state(q=q0, nt=5, resmax0=1.0, clift=cl0)
            st, n = g.run(4000)
        else:
            st, n = g.run(8000, drop=1e-11)
        q, _, _ = g.get_state()
    return st, q


@pytest.fixture(scope="module")
def steady_gradients():
    case = _steady_case()
    modes = _modes(case)
    with Solver(case) as g:
        st, n = g.run(8000, drop=1e-11)
        print(f"steady 119x31: {n} passes, ||s|| {st['rhsnorm']:.3e}, cl {st['cl']!r}", flush=True)
        dcl, dcd, ast = g.adjoint(modes, maxit=4000, tol=1e-12)
        q, _, _ = g.get_state()
    print(f"adjoint: {ast}", flush=True)
    return case, modes, st, q, dcl, dcd, ast


def test_adjoint_gradients_match_the_cpu_tangent(steady_gradients):
    case, modes, st, q, dcl, dcd, ast = steady_gradients
    assert max(ast["resid"]) < 1e-10
    tg, _, _ = adjoint_cpu.tangent_gradients(case, q, st["cl"], shape_dirs=modes, rel=1e-6,
                                             L=pyoracle.fp64_lib())
    print("cpu tangent dcl", tg[0], "\nadjoint dcl", dcl, flush=True)
    np.testing.assert_allclose(dcl, tg[0], rtol=0, atol=2e-6 * np.abs(tg[0]).max())
    np.testing.assert_allclose(dcd, tg[1], rtol=0, atol=2e-6 * np.abs(tg[1]).max())


def test_adjoint_preconditioner_steps_agree(steady_gradients):
    """adj_prec_steps: one step of the transposed update per preconditioner application (the
    plain form) and the default three give the same gradients to the solve tolerance, the
    latter in fewer GMRES iterations."""
    case, modes, st, q, dcl, dcd, ast = steady_gradients
    with Solver(case, options={"adj_prec_steps": 1}) as g:
        g.set_state(q=q, nt=5, resmax0=1.0, clift=st["cl"])
        dcl1, dcd1, ast1 = g.adjoint(modes, maxit=8000, tol=1e-12)
    print(f"iterations: 1 step {ast1['iters']}, 3 steps {ast['iters']}", flush=True)
    assert all(a < b for a, b in zip(ast["iters"], ast1["iters"]))
    np.testing.assert_allclose(dcl, dcl1, rtol=0, atol=1e-8 * np.abs(dcl1).max())
    np.testing.assert_allclose(dcd, dcd1, rtol=0, atol=1e-8 * np.abs(dcd1).max())


def test_adjoint_gradients_match_central_differences(steady_gradients):
    case, modes, st, q, dcl, dcd, ast = steady_gradients
    h = 2e-2
    p = [_converged(_steady_case(alpha=case.params.alpha + sgn * h), q, st["cl"])[0]
         for sgn in (1, -1)]
    fcl, fcd = (p[0]["cl"] - p[1]["cl"]) / (2 * h), (p[0]["cd"] - p[1]["cd"]) / (2 * h)
    print(f"alpha: adjoint {dcl[0]!r} {dcd[0]!r}  fd {fcl!r} {fcd!r}", flush=True)
    assert dcl[0] == pytest.approx(fcl, rel=2e-5)
    assert dcd[0] == pytest.approx(fcd, rel=2e-5)

    def central(i, h):
        dx, dy = modes[i - 1]
        p = []
        for sgn in (1, -1):
            c2 = _steady_case()
            c2.x, c2.y = case.x + sgn * h * dx, case.y + sgn * h * dy
            p.append(_converged(c2, q, st["cl"])[0])
        return np.array([(p[0]["cl"] - p[1]["cl"]) / (2 * h), (p[0]["cd"] - p[1]["cd"]) / (2 * h)])

    for i in (2, 6):        # upper-surface bump at x/c = 0.3, lower-surface bump at x/c = 0.3
        f1, f2 = central(i, 1e-3), central(i, 5e-4)
        fd = (4.0 * f2 - f1) / 3.0                    # Richardson: the h^2 term removed
        print(f"bump {i}: adjoint {dcl[i]!r} {dcd[i]!r}  fd {fd[0]!r} {fd[1]!r}", flush=True)
        assert abs(dcl[i] - fd[0]) <= 1e-4 * np.abs(dcl).max()
        assert abs(dcd[i] - fd[1]) <= 1e-4 * np.abs(dcd).max()


def test_torch_autograd_function_wraps_the_adjoint(steady_gradients):
    """A torch.autograd.Function whose forward converges the flow at (alpha, bump amplitudes)
    and whose backward is the adjoint: torch's gradient of a loss of CL and CD equals the
    adjoint gradient, and agrees with a central difference of the loss along a direction."""
    torch = pytest.importorskip("torch")
    case, modes, st, q, dcl, dcd, ast = steady_gradients

    class SteadyForces(torch.autograd.Function):
        @staticmethod
        def forward(ctx, alpha, amps):
            a = amps.detach().numpy()
            c2 = _steady_case(alpha=float(alpha))
            c2.x, c2.y = grid.deform(case.x, case.y, modes, a)
            # the grid moves linearly in the amplitudes: its derivative in amplitude i is
            # mode i of the undeformed grid
            with Solver(c2) as g:
                g.set_state(q=q, nt=5, resmax0=1.0, clift=st["cl"])
                s2, _ = g.run(4000)
                gcl, gcd, _ = g.adjoint(modes, maxit=4000, tol=1e-12)
            ctx.save_for_backward(torch.tensor(gcl), torch.tensor(gcd))
            return torch.tensor([s2["cl"], s2["cd"]], dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            gcl, gcd = ctx.saved_tensors
            gr = gout[0] * gcl + gout[1] * gcd
            return gr[0], gr[1:]

    alpha = torch.tensor(case.params.alpha, dtype=torch.float64, requires_grad=True)
    amps = torch.zeros(len(modes), dtype=torch.float64, requires_grad=True)
    out = SteadyForces.apply(alpha, amps)
    loss = out[0] - 3.0 * out[1]
    loss.backward()
    # the forward's state is another iterate in the same noise band as the fixture's, so the
    # two adjoint gradients agree to that band, not bit for bit
    gl = dcl - 3.0 * dcd
    assert alpha.grad.item() == pytest.approx(gl[0], rel=1e-6)
    np.testing.assert_allclose(amps.grad.numpy(), gl[1:], rtol=0, atol=1e-6 * np.abs(gl).max())
    # central difference of the loss along a mixed direction of alpha and the amplitudes
    rng = np.random.default_rng(4)
    da, db = 1.0, rng.uniform(-1.0, 1.0, len(modes)) * 0.5
    h = 2e-3
    vals = []
    for sgn in (1, -1):
        with torch.no_grad():
            o2 = SteadyForces.apply(alpha + sgn * h * da, amps + sgn * h * torch.tensor(db))
        vals.append((o2[0] - 3.0 * o2[1]).item())
    fd = (vals[0] - vals[1]) / (2 * h)
    want = alpha.grad.item() * da + float(amps.grad.numpy() @ db)
    assert fd == pytest.approx(want, rel=1e-3)
