"""GPU tests of the batched adjoint kernel and the autograd surface
(igaos_batch_adjoint, igaos_amd.autograd).

Bars:
  * directional derivatives of L = w.x* + 0.5 |y*|^2 with respect to c, b
    and A's values, from FP32 requires_grad tensors through the FP32+refine
    batched solve and the adjoint kernel, within 1e-4 relative of the FP64
    central finite differences of exact solutions committed in
    tests/golden/adjoint_fd.json (32 nondegenerate instances of the
    512-LP differentiable batch);
  * the kernel's gradients equal the dense adjoint oracle's at the same
    (x*, y*) to 1e-8 relative;
  * the objective's closed forms: dobj/dc = x*, dobj/db = -y*,
    dobj/dA = y* x*';
  * active-set polishing recovers the exact vertex (x*, y*) of every
    instance to 1e-9 relative (the FP64 dense refinement of the HiGHS basis).
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle.adjoint import adjoint as oracle_adjoint
from tests.golden.make_adjoint_fixtures import direction, instance, weights
from tests.highs_util import highs_xy

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fixtures():
    with open(os.path.join(HERE, "golden", "adjoint_fd.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def layer(fixtures):
    from igaos_amd.autograd import DiffLPBatch
    lps = [instance(e["index"]) for e in fixtures["instances"]]
    prob = DiffLPBatch(lps, precision=32, tolerance=1e-8, time_limit=120)
    yield prob
    prob.close()


def _rel(a, b):
    return np.abs(a - b).max() / max(1e-300, np.abs(b).max())


def test_gradients_match_finite_differences(fixtures, layer):
    c, rmin, rmax, ax = layer.data(torch.float32)
    assert torch.equal(rmin, rmax)
    b = rmin.clone().requires_grad_(True)
    c.requires_grad_(True)
    ax.requires_grad_(True)
    x, y, obj = layer.solve_eq(c, b, ax)
    sol = layer.last_solution
    assert sol.solved == len(layer.lps), sol.status
    assert layer.last_polish_info["accepted"].all(), layer.last_polish_info
    w = torch.from_numpy(np.concatenate(
        [weights(e["index"], e["n"]) for e in fixtures["instances"]])).float()
    loss = (w * x).sum() + 0.5 * (y * y).sum()
    loss.backward()
    info = layer.last_adjoint_info
    grads = {"c": layer.split(c.grad.double().numpy(), "n"),
             "b": layer.split(b.grad.double().numpy(), "m"),
             "A": layer.split(ax.grad.double().numpy(), "nnz")}
    worst = 0.0
    for i, ent in enumerate(fixtures["instances"]):
        assert info["basic_cols"][i] == info["active_rows"][i] == ent["m"], i
        for name, rec in ent["fd"].items():
            g = grads[name][i]
            u = direction(ent["index"], name, rec["attempt"], g.size)
            err = abs(g @ u - rec["value"]) / max(1.0, abs(rec["value"]))
            worst = max(worst, err)
            assert err <= 1e-4, (i, name, g @ u, rec["value"])
    print(f"worst relative error vs finite differences: {worst:.2e}")


def test_kernel_matches_dense_oracle(fixtures, layer):
    c, rmin, rmax, ax = (t.double().numpy() for t in layer.data(torch.float64))
    layer.solver.update(c=c, rmin=rmin, rmax=rmax, ax=ax)
    sol = layer.solver.run(tolerance=1e-6, time_limit=120)
    rng = np.random.default_rng(3)
    gx = rng.standard_normal(int(layer.coff[-1]))
    gy = rng.standard_normal(int(layer.roff[-1]))
    gobj = rng.standard_normal(len(layer.lps))
    g = layer.solver.adjoint(gx=gx, gy=gy, gobj=gobj, tolerance=1e-12)
    for i, lp in enumerate(layer.lps[:8]):
        cs, rs, zs = (slice(int(o[i]), int(o[i + 1])) for o in (layer.coff, layer.roff, layer.zoff))
        ref = oracle_adjoint(lp, sol.x[i], sol.y[i], gx=gx[cs], gy=gy[rs], gobj=gobj[i])
        assert _rel(g["gc"][cs], ref.gc) <= 1e-8, i
        assert _rel(g["grmin"][rs] + g["grmax"][rs], ref.grmin + ref.grmax) <= 1e-8, i
        assert _rel(g["gax"][zs], ref.gax) <= 1e-8, i
        assert _rel(g["gcl"][cs] + g["gcu"][cs], ref.gcl + ref.gcu) <= 1e-8, i
        assert g["info"]["res_p"][i] <= 1e-9 and g["info"]["res_q"][i] <= 1e-9


def test_objective_closed_forms(layer):
    c, rmin, rmax, ax = (t.double() for t in layer.data(torch.float64))
    for t in (c, rmin, ax):
        t.requires_grad_(True)
    x, y, obj = layer.solve_eq(c, rmin, ax)
    obj.sum().backward()
    xs = x.detach().numpy()
    ys = y.detach().numpy()
    assert np.array_equal(c.grad.numpy(), xs)
    assert _rel(rmin.grad.numpy(), -ys) <= 1e-8
    for i, lp in enumerate(layer.lps[:8]):
        zs = slice(int(layer.zoff[i]), int(layer.zoff[i + 1]))
        xi = layer.split(xs, "n")[i]
        yi = layer.split(ys, "m")[i]
        rows = np.repeat(np.arange(lp.m), np.diff(lp.ap))
        assert _rel(ax.grad.numpy()[zs], yi[rows] * xi[lp.ai]) <= 1e-8, i


def test_stale_forward_is_solved_again(layer):
    c, rmin, rmax, ax = layer.data(torch.float64)
    c1 = c.clone().requires_grad_(True)
    x1, _, _ = layer.solve_eq(c1, rmin, ax)
    c2 = (c * 1.01).requires_grad_(True)
    layer.solve_eq(c2, rmin, ax)   # replaces the device state
    x1.sum().backward()            # must differentiate the first problem
    ref = layer.solver.adjoint(gx=np.ones(int(layer.coff[-1])))
    assert np.allclose(c1.grad.numpy(), ref["gc"])


def test_polish_recovers_exact_vertex(layer):
    c, rmin, rmax, ax = (t.double().numpy() for t in layer.data(torch.float64))
    layer.solver.update(c=c, rmin=rmin, rmax=rmax, ax=ax)
    sol = layer.solver.run(tolerance=1e-8, time_limit=120)
    xs, ys, obj, info = layer.solver.polish(tolerance=1e-12)
    assert info["accepted"].all(), info
    assert info["viol"].max() <= 1e-12
    for i, lp in enumerate(layer.lps[:6]):
        xr, yr, objr = highs_xy(lp)
        assert _rel(xs[i], xr) <= 1e-9, i
        assert _rel(ys[i], yr) <= 1e-9, i
        assert abs(obj[i] - objr) <= 1e-12 * (1 + abs(objr)), i
        # the unpolished PDHG pair meets the solve tolerance only
        assert abs(sol.objective[i] - objr) <= 1e-8 * (1 + abs(objr))


def test_rejected_polish_is_solved_again(layer):
    # at 1e-6 some instance's active set is misread (a swapped pair); the
    # forward solves it again 100x tighter before polishing
    c, rmin, rmax, ax = layer.data(torch.float64)
    tol = layer.tolerance
    layer.tolerance = 1e-6
    try:
        x, y, obj = layer.solve_eq(c, rmin, ax)
    finally:
        layer.tolerance = tol
    assert layer.last_polish_info["accepted"].all(), layer.last_polish_info
