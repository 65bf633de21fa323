"""The FP64 CPU tangent of the steady mean flow, from the oracle (test infrastructure).

The steady equations are the oracle's assembled residual s (the reference's arithmetic) on
every row but the wake-cut targets, the cut relation q(jt,1) = (q(jt,2) + q(jb,2))/2 on the
targets, and c = CL(q) for the circulation the far field reads.  Their derivatives are taken
by central differences of the oracle (its residual with a colouring of the 5x5 stencil
boxes, its force integration point by point), and the tangent system

    [ F_q    F_c ] [dq]     [ F_p  ]
    [ -CL_q   1  ] [dc] = - [ -CL_p]

is solved directly (scipy), giving dJ/dp = J_p + J_q dq for J = CL, CD and p = alpha or a
shape direction of the grid.  Independent of the GPU adjoint: other differentiation (finite
differences of the oracle rather than dual numbers), other linear algebra (a sparse direct
solve of the tangent rather than GMRES on the adjoint).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import pyoracle
from minismac_amd.case import Case


def _case_with(case: Case, x=None, y=None, **params) -> Case:
    p = pyoracle.SmacParams()
    for name, _ in p._fields_:
        setattr(p, name, getattr(case.params, name))
    for k, v in params.items():
        setattr(p, k, v)
    return dataclasses.replace(case, x=case.x if x is None else x, y=case.y if y is None else y,
                               params=p)


class Steady:
    """F(q, c) and CL, CD of one case through the oracle (or another build of it, L)."""

    def __init__(self, case: Case, L=None):
        self.case = case
        self.o = pyoracle.Oracle(case, L=L)
        self.pairs = self.o.wake_pairs() - 1          # 0-based (jt, jb)

    def F(self, q, c):
        o = self.o
        o.set_array("q", q)
        o.set_counters(5, 1.0, c)
        o.assemble()
        s = o.array("s")
        for jt, jb in self.pairs:
            s[jt, 0, :] = 0.5 * (q[jt, 1, :] + q[jb, 1, :]) - q[jt, 0, :]
        return s

    def forces(self, q):
        o = self.o
        o.set_array("q", q)
        o.set_array("dq", np.zeros_like(q))
        st = o.finish()
        return np.array([st["cl"], st["cd"]])


def _vec(a):
    """(jm, km, 3) Fortran field -> vector, index n*N + j + jm*k (the device layout)."""
    return np.asarray(a).ravel(order="F")


def jacobian(S: Steady, q, c, rel=1e-6):
    """F_q as a sparse matrix over the device layout, by central differences with the 25
    colours (j mod 5, k mod 5) per variable; the cut rows exactly."""
    jm, km = q.shape[:2]
    N = jm * km
    targets = {int(jt) for jt, _ in S.pairs}
    rows, cols, vals = [], [], []
    jj, kk = np.meshgrid(np.arange(jm), np.arange(km), indexing="ij")
    keep = np.ones((jm, km), dtype=bool)
    for jt in targets:
        keep[jt, 0] = False
    for n in range(3):
        h = rel * max(1.0, float(np.abs(q[:, :, n]).max()))
        for cj in range(5):
            for ck in range(5):
                mask = (jj % 5 == cj) & (kk % 5 == ck)
                qp, qm = q.copy(order="F"), q.copy(order="F")
                qp[:, :, n][mask] += h
                qm[:, :, n][mask] -= h
                dF = (S.F(qp, c) - S.F(qm, c)) / (2 * h)
                # the point of this colour in the 5x5 box around each row's point
                pj = jj + (cj - jj + 2) % 5 - 2
                pk = kk + (ck - kk + 2) % 5 - 2
                ok = keep & (pj >= 0) & (pj < jm) & (pk >= 0) & (pk < km)
                for m in range(3):
                    sel = ok & (dF[:, :, m] != 0.0)
                    rows.append((m * N + jj + jm * kk)[sel])
                    cols.append((n * N + pj + jm * pk)[sel])
                    vals.append(dF[:, :, m][sel])
    rows, cols, vals = list(np.concatenate(rows)), list(np.concatenate(cols)), list(np.concatenate(vals))
    for jt, jb in S.pairs:
        for m in range(3):
            r = m * N + jt
            rows += [r, r, r]
            cols += [m * N + jt, m * N + jt + jm, m * N + jb + jm]
            vals += [-1.0, 0.5, 0.5]
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * N, 3 * N))


def force_gradient(S: Steady, q, rel=1e-6):
    """d(CL, CD)/dq over the wall rows k = 0, 1, 2 (the force reads no others)."""
    jm, km = q.shape[:2]
    N = jm * km
    g = np.zeros((2, 3 * N))
    walls = [r for r in S.case.resolved_regions() if r[0] == 0]
    for (_, j1, j2, k1, _) in walls:
        for k in range(k1 - 1, k1 + 2):
            for j in range(j1 - 1, j2):
                for n in range(3):
                    h = rel * max(1.0, abs(q[j, k, n]))
                    qp, qm = q.copy(order="F"), q.copy(order="F")
                    qp[j, k, n] += h
                    qm[j, k, n] -= h
                    g[:, n * N + j + jm * k] = (S.forces(qp) - S.forces(qm)) / (2 * h)
    return g


def tangent_gradients(case: Case, q, c, shape_dirs=(), rel=1e-6, alpha_step=1e-6, L=None):
    """dCL/dp and dCD/dp (p = alpha in degrees, then each shape direction) at the steady
    state q of `case`, c its CL, through the oracle build L (default: the oracle)."""
    S = Steady(case, L)
    jm, km = q.shape[:2]
    N = jm * km
    Fq = jacobian(S, q, c, rel)
    hc = rel * max(1.0, abs(c))
    Fc = _vec((S.F(q, c + hc) - S.F(q, c - hc)) / (2 * hc))
    Jq = force_gradient(S, q, rel)
    # the bordered tangent matrix
    A = sp.bmat([[Fq, sp.csr_matrix(Fc[:, None])],
                 [sp.csr_matrix(-Jq[0][None, :]), sp.csr_matrix(np.ones((1, 1)))]], format="csc")
    lu = spla.splu(A)
    params = []
    # alpha
    a0 = case.params.alpha
    Sp = Steady(_case_with(case, alpha=a0 + alpha_step), L)
    Sm = Steady(_case_with(case, alpha=a0 - alpha_step), L)
    Fp = _vec((Sp.F(q, c) - Sm.F(q, c)) / (2 * alpha_step))
    Jp = (Sp.forces(q) - Sm.forces(q)) / (2 * alpha_step)
    params.append((Fp, Jp))
    for dx, dy in shape_dirs:
        h = rel
        Sp = Steady(_case_with(case, x=case.x + h * dx, y=case.y + h * dy), L)
        Sm = Steady(_case_with(case, x=case.x - h * dx, y=case.y - h * dy), L)
        Fp = _vec((Sp.F(q, c) - Sm.F(q, c)) / (2 * h))
        Jp = (Sp.forces(q) - Sm.forces(q)) / (2 * h)
        params.append((Fp, Jp))
    out = np.zeros((2, len(params)))
    for i, (Fp, Jp) in enumerate(params):
        rhs = -np.concatenate([Fp, [-Jp[0]]])
        sol = lu.solve(rhs)
        dq = sol[:-1]
        out[:, i] = Jp + Jq @ dq
    return out, Fq, Jq
