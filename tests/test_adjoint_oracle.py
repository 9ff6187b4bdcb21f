"""Pins the adjoint oracle (oracle/adjoint.py) on the CPU.

Checks that a plausible mistake in the adjoint (a swapped system, a lost
sign, A instead of A', the wrong bound side) would fail:
  * closed forms of the objective's gradients at exact simplex solutions:
    dobj/dc = x*, dobj/db = -y* (PDHG sign convention), dobj/dA_ij = y*_i x*_j;
  * central finite differences (h = 1e-4, unit-norm random directions, FP64)
    of L = w.x* + 0.5 |y*|^2 through exact HiGHS solutions, for c, b and A's
    values, on small nondegenerate equality LPs and on LPs with inequality
    rows (gradients of rmax);
  * the committed fixtures of the 512-LP differentiable batch
    (tests/golden/adjoint_fd.json) on their first instances.
"""
import json
import os

import numpy as np
import pytest

from igaos_amd import lpgen
from oracle.adjoint import adjoint, masks
from tests.highs_util import basis, highs_xy

HERE = os.path.dirname(os.path.abspath(__file__))
H = 1e-4


def _small(index, ineq=False):
    lp = lpgen.gen_batch_instance(5, index, m_lo=20, m_hi=60, ratio_lo=2.0,
                                  ratio_hi=2.5, k_lo=3, k_hi=5)
    lp = lpgen.perturb_nondegenerate(lp, seed=index, scale=1e-6)
    if ineq:
        # every other row becomes  A_i x <= b_i + slack_i
        rng = np.random.default_rng(index)
        sel = np.arange(lp.m) % 2 == 1
        lp.rmax = np.where(sel, lp.rmax + rng.uniform(0.0, 1.0, lp.m), lp.rmax)
        lp.rmin = np.where(sel, -np.inf, lp.rmin)
    return lp


def _nondegenerate(lp, x, y):
    B = basis(lp, x)
    A = lp.to_scipy()
    d = lp.c + A.T @ y
    ax = A @ x
    eq = lp.rmin == lp.rmax
    tight = eq | (np.abs(ax - lp.rmax) < 1e-9) | (np.abs(ax - lp.rmin) < 1e-9)
    ok_rows = np.all(np.abs(y[tight & ~eq]) > 1e-7)
    return (B.sum() == tight.sum() and np.abs(d[~B]).min() > 1e-7 and ok_rows)


def _loss(lp, w):
    x, y, _ = highs_xy(lp)
    return w @ x + 0.5 * y @ y, basis(lp, x)


def _fd(lp, w, group, u):
    B0 = basis(lp, highs_xy(lp)[0])
    vals = []
    for s in (H, -H):
        if group == "c":
            q = lp.with_values(c=lp.c + s * u)
        elif group == "b":
            q = lp.with_values(rmin=lp.rmin + s * u, rmax=lp.rmax + s * u)
        elif group == "rmax":
            q = lp.with_values(rmax=lp.rmax + s * u)
        else:
            q = lp.with_values(ax=lp.ax + s * u)
        v, B = _loss(q, w)
        if not (B == B0).all():
            return None
        vals.append(v)
    return (vals[0] - vals[1]) / (2 * H)


def _rows(lp):
    return np.repeat(np.arange(lp.m), np.diff(lp.ap))


@pytest.mark.parametrize("index", range(4))
def test_closed_forms_of_objective_gradients(index):
    lp = _small(index)
    x, y, _ = highs_xy(lp)
    g = adjoint(lp, x, y, gobj=1.0)
    scale = 1 + np.abs(x).max() * (1 + np.abs(y).max())
    assert np.array_equal(g.gc, x)
    assert np.abs(g.grmin + g.grmax + y).max() <= 1e-9 * (1 + np.abs(y).max())
    assert np.abs(g.gax - y[_rows(lp)] * x[lp.ai]).max() <= 1e-9 * scale


@pytest.mark.parametrize("index,ineq", [(0, False), (1, False), (2, False),
                                        (3, True), (4, True)])
def test_finite_differences(index, ineq):
    lp = _small(index, ineq)
    x, y, _ = highs_xy(lp)
    if not _nondegenerate(lp, x, y):
        pytest.skip("instance degenerate at the exact solution")
    B, R = masks(lp, x, y)
    assert (B == basis(lp, x)).all() and B.sum() == R.sum()
    rng = np.random.default_rng(100 + index)
    w = rng.standard_normal(lp.n)
    g = adjoint(lp, x, y, gx=w, gy=y)
    groups = {"c": (lp.n, g.gc), "A": (lp.nnz, g.gax)}
    groups["rmax" if ineq else "b"] = (lp.m, g.grmax if ineq else g.grmin + g.grmax)
    checked = 0
    for name, (size, grad) in groups.items():
        for _ in range(4):
            u = rng.standard_normal(size)
            if name == "rmax":
                u = np.where(np.isfinite(lp.rmin), 0.0, u)  # inequality rows only
            u /= np.linalg.norm(u)
            fd = _fd(lp, w, name, u)
            if fd is None:
                continue
            an = grad @ u
            assert abs(an - fd) <= 1e-5 * max(1.0, abs(fd)), (name, an, fd)
            checked += 1
            break
    assert checked == len(groups)


def test_mutations_fail_finite_differences():
    # the same checks catch a transposed or sign-flipped adjoint
    lp = _small(0)
    x, y, _ = highs_xy(lp)
    rng = np.random.default_rng(7)
    w = rng.standard_normal(lp.n)
    g = adjoint(lp, x, y, gx=w, gy=y)
    u = rng.standard_normal(lp.nnz)
    u /= np.linalg.norm(u)
    fd = _fd(lp, w, "A", u)
    assert fd is not None
    rows, cols = _rows(lp), lp.ai
    p = -(g.gc - 0.0)  # gobj = 0: gc = -p
    q = g.grmin + g.grmax
    good = -(p[cols] * y[rows] + q[rows] * x[cols])
    assert abs(good @ u - fd) <= 1e-5 * max(1, abs(fd))
    for bad in (-(p[cols] * y[rows] - q[rows] * x[cols]),   # lost sign
                -(q[rows] * x[cols]),                          # dropped term
                (p[cols] * y[rows] + q[rows] * x[cols])):      # flipped sign
        assert abs(bad @ u - fd) > 1e-3 * max(1, abs(fd))


def test_committed_fixtures_first_instances():
    with open(os.path.join(HERE, "golden", "adjoint_fd.json")) as f:
        fx = json.load(f)
    from tests.golden.make_adjoint_fixtures import direction, instance, weights
    for ent in fx["instances"][:2]:
        lp = instance(ent["index"])
        assert (lp.m, lp.n, lp.nnz) == (ent["m"], ent["n"], ent["nnz"])
        x, y, obj = highs_xy(lp)
        assert abs(obj - ent["objective"]) <= 1e-9 * (1 + abs(obj))
        w = weights(ent["index"], lp.n)
        g = adjoint(lp, x, y, gx=w, gy=y)
        grads = {"c": g.gc, "b": g.grmin + g.grmax, "A": g.gax}
        for name, rec in ent["fd"].items():
            u = direction(ent["index"], name, rec["attempt"], grads[name].size)
            an = grads[name] @ u
            assert abs(an - rec["value"]) <= 1e-5 * max(1.0, abs(rec["value"])), name


def test_masks_balance_a_misclassified_column():
    # an approximate solution whose nonbasic column sits off its bound by
    # more than its reduced cost: the raw rule puts it in B, the count
    # balance (|B| = |R| at a nondegenerate vertex) takes it out again
    from oracle import ruiz
    lp = _small(1)
    x, y, _ = highs_xy(lp)
    B0 = basis(lp, x)
    dr, dc = ruiz(lp)
    A = lp.to_scipy()
    dt = np.abs(lp.c + A.T @ y) * dc          # |d~|
    j = np.flatnonzero(~B0)[np.argmin(dt[~B0])]
    xs = x.copy()
    step = 1.5 * dt[j] * dc[j]                # scaled distance 1.5 |d~_j|
    xs[j] = lp.cl[j] + step if x[j] <= lp.cl[j] + 1e-12 else lp.cu[j] - step
    B, R = masks(lp, xs, y)
    assert (B == B0).all() and B.sum() == R.sum()


def _enumerate(lp):
    """Optimal vertex of a tiny equality LP with box bounds by brute force:
    every choice of m basic columns and bound for each other column."""
    import itertools
    A = lp.to_scipy().toarray()
    best = None
    for B in itertools.combinations(range(lp.n), lp.m):
        AB = A[:, B]
        if abs(np.linalg.det(AB)) < 1e-9:
            continue
        N = [j for j in range(lp.n) if j not in B]
        for side in itertools.product((0, 1), repeat=len(N)):
            x = np.zeros(lp.n)
            for j, sd in zip(N, side):
                x[j] = lp.cu[j] if sd else lp.cl[j]
            x[list(B)] = np.linalg.solve(AB, lp.rmin - A[:, N] @ x[N])
            if np.any(x < lp.cl - 1e-9) or np.any(x > lp.cu + 1e-9):
                continue
            obj = lp.c @ x
            if best is None or obj < best[0] - 1e-12:
                y = np.linalg.solve(AB.T, -lp.c[list(B)])
                best = (obj, x, y)
    return best


def test_adjoint_oracle_against_vertex_enumeration():
    # tiny LPs solved by brute force (no LP solver involved): the oracle's
    # gradients equal central differences of the enumerated optimum
    rng = np.random.default_rng(11)
    checked = 0
    for trial in range(12):
        m, n = 2, 5
        A = rng.standard_normal((m, n))
        x0 = rng.uniform(0, 10, n)
        lp = lpgen.LP.from_triplets(m, n, *np.nonzero(A), A[np.nonzero(A)],
                                    rng.standard_normal(n), np.zeros(n),
                                    np.full(n, 10.0), A @ x0, A @ x0)
        base = _enumerate(lp)
        if base is None:
            continue
        _, x, y = base
        inside = (x > 1e-6) & (x < 10 - 1e-6)
        if inside.sum() != m:
            continue  # degenerate vertex
        w = rng.standard_normal(n)
        g = adjoint(lp, x, y, gx=w, gy=y)
        L = lambda q: (lambda r: w @ r[1] + 0.5 * r[2] @ r[2])(_enumerate(q))  # noqa
        for name, grad, size in (("c", g.gc, n), ("b", g.grmin + g.grmax, m),
                                 ("A", g.gax, lp.nnz)):
            u = rng.standard_normal(size)
            u /= np.linalg.norm(u)
            if name == "c":
                qp, qm = lp.with_values(c=lp.c + H * u), lp.with_values(c=lp.c - H * u)
            elif name == "b":
                qp = lp.with_values(rmin=lp.rmin + H * u, rmax=lp.rmax + H * u)
                qm = lp.with_values(rmin=lp.rmin - H * u, rmax=lp.rmax - H * u)
            else:
                qp, qm = lp.with_values(ax=lp.ax + H * u), lp.with_values(ax=lp.ax - H * u)
            fd = (L(qp) - L(qm)) / (2 * H)
            assert abs(grad @ u - fd) <= 1e-6 * max(1.0, abs(fd)), (trial, name)
        checked += 1
    assert checked >= 4
