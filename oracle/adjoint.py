"""TEST INFRASTRUCTURE ONLY — dense numpy restatement of the batched
adjoint (the product kernel lives in igaos_amd/csrc/adjoint.hip).

No reference counterpart: the reference (src/pdhg/pdhg.cu) has no
gradients.  The algorithm is implicit differentiation of the LP solution
map at a nondegenerate solution:

  active sets from complementarity in Ruiz-scaled units (the scaling of
  pdhg.cu:469-492, from ``oracle.ruiz``):
    B = columns whose distance to the nearer bound exceeds |d~_j|,
        d~ = c~ + A~'y~ (dual sign convention d = c + A'y, pdhg.cu:231-248)
    R = equality rows, and inequality rows with |y~_i| above the row's
        distance to the nearer bound;
  stationarity and feasibility restricted to them,
    c_B + A_RB' y_R = 0,   A_RB x_B + A_RN x_N = b_R,
  differentiated, and transposed for upstream gradients (gx, gy, gobj):
    A_RB p = gy_R,   A_RB' q = w_B,   w = gx + gobj c,
  solved here by dense least squares (minimum norm, like CGLS from zero).

Pinned in tests/test_adjoint_oracle.py by central finite differences of
exact simplex solutions (scipy HiGHS) and by the closed forms of the
objective's gradients (dobj/dc = x, dobj/db = -y, dobj/dA = y x').
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import ruiz


@dataclass
class Grad:
    gc: np.ndarray
    grmin: np.ndarray
    grmax: np.ndarray
    gcl: np.ndarray
    gcu: np.ndarray
    gax: np.ndarray      # CSR order of the LP
    basic: np.ndarray    # bool mask of B
    active: np.ndarray   # bool mask of R


def masks(lp, x, y, dr=None, dc=None):
    """Active sets (B, R) by the complementarity rule in scaled units."""
    if dr is None:
        dr, dc = ruiz(lp)
    A = lp.to_scipy()
    xt = x / dc
    yt = y * dr
    # A~ = D_r^-1 A D_c
    at = (A @ (xt * dc)) / dr
    eq = lp.rmin == lp.rmax
    with np.errstate(invalid="ignore"):
        slr = np.minimum(at - lp.rmin / dr, lp.rmax / dr - at)
    R = eq | (np.abs(yt) > slr)
    dt = np.abs(lp.c * dc + dc * (A.T @ (yt / dr)))
    with np.errstate(invalid="ignore", divide="ignore"):
        slc = np.minimum(xt - lp.cl / dc, lp.cu / dc - xt)
        score = np.where(dt > 0, slc / np.where(dt > 0, dt, 1.0),
                         np.where(slc > 0, np.inf, 0.0))
    B = score > 1.0
    # a nondegenerate vertex has |B| = |R|: move the lowest-scoring columns
    # out of B (or the highest-scoring ones in), smallest index on ties
    for _ in range(64):
        nb, nr = int(B.sum()), int(R.sum())
        if nb == nr:
            break
        if nb > nr:
            cand = np.flatnonzero(B)
            j = cand[np.argmin(score[cand])]
            B[j] = False
        else:
            cand = np.flatnonzero(~B)
            j = cand[np.argmax(score[cand])]
            B[j] = True
    return B, R


def adjoint(lp, x, y, gx=None, gy=None, gobj=0.0) -> Grad:
    """Gradients of L(x*, y*, obj) w.r.t. c, the active bounds and A's
    values (CSR order), at the solution (x, y) of ``lp``."""
    m, n = lp.m, lp.n
    gx = np.zeros(n) if gx is None else np.asarray(gx, float)
    gy = np.zeros(m) if gy is None else np.asarray(gy, float)
    B, R = masks(lp, x, y)
    A = lp.to_scipy().toarray()
    ARB = A[np.ix_(R, B)]
    w = gx + gobj * lp.c
    p = np.zeros(n)
    q = np.zeros(m)
    if ARB.size:
        p[B] = np.linalg.lstsq(ARB, gy[R], rcond=None)[0]
        q[R] = np.linalg.lstsq(ARB.T, w[B], rcond=None)[0]
    gc = gobj * x - p
    up = y >= 0.0
    grmin = np.where(up, 0.0, q)
    grmax = np.where(up, q, 0.0)
    aq = A.T @ q
    gb = np.where(B, 0.0, w - aq)
    dist_lo = x - lp.cl
    dist_hi = lp.cu - x
    lower = dist_lo <= dist_hi
    gcl = np.where(lower, gb, 0.0)
    gcu = np.where(lower, 0.0, gb)
    rows = np.repeat(np.arange(m), np.diff(lp.ap))
    cols = lp.ai
    yR = np.where(R, y, 0.0)
    gax = -(p[cols] * yR[rows] + q[rows] * x[cols])
    return Grad(gc, grmin, grmax, gcl, gcu, gax, B, R)
