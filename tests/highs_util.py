"""Exact LP solutions from scipy's HiGHS (a checker for tests and fixture
scripts; never used by the product).

``highs_xy(lp)`` returns (x, y, objective) with y in the PDHG dual sign
convention d = c + A'y (src/pdhg/pdhg.cu:231-248): HiGHS reports marginals
as d obj / d rhs, and for every binding row d obj / d b_i = -y_i.  With
``refine`` (default) a nondegenerate vertex is recomputed from its basis by
dense FP64 LU, which gives the near-machine-precision solutions that
finite differences with h = 1e-4 need.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.optimize import linprog


def highs_xy(lp, refine=True):
    A = sp.csr_matrix((lp.ax, lp.ai, lp.ap), shape=(lp.m, lp.n))
    eq = lp.rmin == lp.rmax
    hi = np.flatnonzero(np.isfinite(lp.rmax) & ~eq)
    lo = np.flatnonzero(np.isfinite(lp.rmin) & ~eq)
    kw = {}
    if hi.size or lo.size:
        kw.update(A_ub=sp.vstack([A[hi], -A[lo]]).tocsr(),
                  b_ub=np.concatenate([lp.rmax[hi], -lp.rmin[lo]]))
    eqi = np.flatnonzero(eq)
    if eqi.size:
        kw.update(A_eq=A[eqi], b_eq=lp.rmin[eqi])
    bounds = np.column_stack([np.where(np.isfinite(lp.cl), lp.cl, -np.inf),
                              np.where(np.isfinite(lp.cu), lp.cu, np.inf)])
    r = linprog(lp.c, bounds=bounds, method="highs",
                options=dict(primal_feasibility_tolerance=1e-10,
                             dual_feasibility_tolerance=1e-10), **kw)
    if r.status != 0:
        raise RuntimeError(f"HiGHS status {r.status}: {r.message}")
    y = np.zeros(lp.m)
    if eqi.size:
        y[eqi] = -r.eqlin.marginals
    if hi.size or lo.size:
        mu = r.ineqlin.marginals
        y[hi] += -mu[:hi.size]           # rows  A x <= rmax
        y[lo] += mu[hi.size:]            # rows -A x <= -rmin
    x = r.x
    if refine:
        x, y = _vertex(lp, x, y)
    return x, y, float(lp.c @ x) + lp.obj_const


def _vertex(lp, x, y, margin=1e-7):
    """Recompute a nondegenerate vertex from its basis with dense FP64 LU
    (more digits than the solver's tolerances leave): x_N snapped to its
    bound, x_B = A_RB^-1 (b_R - A_RN x_N), y_R = -A_RB^-T c_B, y = 0 off R.
    Returns the input unchanged when the basis is not square."""
    A = lp.to_scipy().toarray()
    B = basis(lp, x, margin)
    ax = A @ x
    eq = lp.rmin == lp.rmax
    at_hi = np.abs(ax - lp.rmax) <= 1e-9 * (1 + np.abs(lp.rmax))
    at_lo = np.abs(ax - lp.rmin) <= 1e-9 * (1 + np.abs(lp.rmin))
    R = eq | at_hi | at_lo
    if B.sum() != R.sum():
        return x, y
    b = np.where(at_hi & ~eq, lp.rmax, lp.rmin)
    xn = x.copy()
    lower = np.abs(x - lp.cl) <= np.abs(lp.cu - x)
    xn[~B] = np.where(lower, lp.cl, lp.cu)[~B]
    ARB = A[np.ix_(R, B)]
    rhs = b[R] - A[np.ix_(R, ~B)] @ xn[~B]
    xn[B] = np.linalg.solve(ARB, rhs)
    yn = np.zeros(lp.m)
    yn[R] = np.linalg.solve(ARB.T, -lp.c[B])
    return xn, yn


def basis(lp, x, margin=1e-7):
    """Columns strictly inside their bounds by more than ``margin``."""
    return (x > lp.cl + margin) & (x < lp.cu - margin)
