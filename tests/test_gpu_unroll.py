"""GPU tests of the unrolled mode (igaos_batch_unroll / _backward).

Bars: the device's K-step iterate and its gradients equal the oracle's
(oracle/unrolled.py: the same map in torch on the CPU, reverse mode by torch
autograd) to 1e-10 / 1e-8 relative, for K on and off the 64-step
checkpoint grid; central differences of the device forward agree with the
device gradients; the autograd layer in unrolled mode returns the same.
"""
import numpy as np
import pytest
import torch

from igaos_amd import lpgen
from igaos_amd.batch import BatchSolver
from oracle import unrolled as U

pytestmark = pytest.mark.gpu


def _lps(count=4, seed=7):
    out = []
    for i in range(count):
        lp = lpgen.gen_batch_instance(seed, i, m_lo=20, m_hi=60, ratio_lo=1.8,
                                      ratio_hi=2.4, k_lo=3, k_hi=5)
        rng = np.random.default_rng(100 + i)
        if i % 2:  # inequality rows, a free column bound
            sel = np.arange(lp.m) % 3 == 1
            lp.rmax = np.where(sel, lp.rmax + rng.uniform(0.0, 1.0, lp.m), lp.rmax)
            lp.rmin = np.where(sel, -np.inf, lp.rmin)
            lp.cu = np.where(np.arange(lp.n) % 5 == 2, np.inf, lp.cu)
        out.append(lp)
    return out


def _rel(a, b):
    return np.abs(a - b).max() / max(1.0, np.abs(b).max())


def _split(v, sizes):
    return np.split(v, np.cumsum(sizes)[:-1])


@pytest.mark.parametrize("K", [0, 10, 64, 200])
def test_unrolled_forward_and_gradients_match_oracle(K):
    lps = _lps()
    rng = np.random.default_rng(K)
    gx = [rng.standard_normal(lp.n) for lp in lps]
    gy = [rng.standard_normal(lp.m) for lp in lps]
    gobj = rng.standard_normal(len(lps))
    b = BatchSolver(lps, precision=32)
    try:
        f = b.unroll(K)
        g = b.unroll_backward(gx=np.concatenate(gx), gy=np.concatenate(gy), gobj=gobj)
    finally:
        b.close()
    xs = _split(f["x"], [lp.n for lp in lps])
    ys = _split(f["y"], [lp.m for lp in lps])
    parts = {k: _split(g[k], [getattr(lp, s) for lp in lps])
             for k, s in (("gc", "n"), ("gcl", "n"), ("gcu", "n"), ("grmin", "m"),
                          ("grmax", "m"), ("gax", "nnz"))}
    for i, lp in enumerate(lps):
        r = U.unrolled(lp, K, gx=gx[i], gy=gy[i], gobj=gobj[i])
        assert abs(f["tau"][i] - r["tau"]) <= 1e-12 * r["tau"]
        assert abs(f["sigma"][i] - r["sigma"]) <= 1e-12 * r["sigma"]
        assert _rel(xs[i], r["x"]) <= 1e-10 and _rel(ys[i], r["y"]) <= 1e-10, i
        assert abs(f["objective"][i] - r["objective"]) <= 1e-10 * (1 + abs(r["objective"]))
        for k in parts:
            assert _rel(parts[k][i], r[k]) <= 1e-8, (i, k, _rel(parts[k][i], r[k]))


def test_unrolled_gradients_against_device_differences():
    lps = _lps(2, seed=9)
    K, h = 128, 1e-6
    rng = np.random.default_rng(1)
    b = BatchSolver(lps, precision=32)
    try:
        base = {k: np.concatenate([getattr(lp, k) for lp in lps])
                for k in ("c", "rmin", "rmax", "ax")}
        w = rng.standard_normal(base["c"].size)
        f = b.unroll(K)
        steps = (f["tau"], f["sigma"])  # the map's constants, held for the differences
        g = b.unroll_backward(gx=w)
        for name, grad in (("c", g["gc"]), ("ax", g["gax"]), ("rmax", g["grmax"])):
            u = rng.standard_normal(base[name].size)
            u = np.where(np.isfinite(base[name]), u, 0.0)
            vals = []
            for s in (h, -h):
                b.update(**{name: base[name] + s * u})
                vals.append(w @ b.unroll(K, steps=steps)["x"])
            b.update(**{name: base[name]})
            b.unroll(K, steps=steps)  # back to the base data for the next group
            g2 = b.unroll_backward(gx=w)
            assert np.array_equal(g2[{"c": "gc", "ax": "gax", "rmax": "grmax"}[name]], grad)
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(grad @ u - fd) <= 1e-5 * max(1.0, abs(fd)), (name, grad @ u, fd)
    finally:
        b.close()


def test_autograd_layer_in_unrolled_mode():
    from igaos_amd.autograd import DiffLPBatch
    lps = _lps(3, seed=11)
    layer = DiffLPBatch(lps, differentiation="unrolled", unroll_iterations=96)
    try:
        c, rmin, rmax, ax = (v.double().requires_grad_(True)
                             for v in layer.data(torch.float64))
        x, y, obj = layer(c, rmin, rmax, ax)
        w = torch.from_numpy(np.random.default_rng(3).standard_normal(x.numel()))
        loss = (w * x).sum() + 0.5 * (y * y).sum() + obj.sum()
        loss.backward()
        xs = np.split(x.detach().numpy(), np.cumsum([lp.n for lp in lps])[:-1])
        ws = np.split(w.numpy(), np.cumsum([lp.n for lp in lps])[:-1])
        gcs = np.split(c.grad.numpy(), np.cumsum([lp.n for lp in lps])[:-1])
        ys = np.split(y.detach().numpy(), np.cumsum([lp.m for lp in lps])[:-1])
        for i, lp in enumerate(lps):
            # gy = y_K: the oracle's loss uses the device iterate as the
            # (constant) upstream gradient of y, as autograd does here
            r = U.unrolled(lp, 96, gx=ws[i], gy=ys[i], gobj=1.0)
            assert _rel(xs[i], r["x"]) <= 1e-10
            assert _rel(gcs[i], r["gc"]) <= 1e-8
    finally:
        layer.close()
