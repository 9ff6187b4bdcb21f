"""Pins the unrolled-mode oracle (oracle/unrolled.py) on the CPU.

  * one step from x = y = 0 against its closed form: x1 = clamp(-tau c),
    y1 = t - sigma clamp(t / sigma), t = sigma A (2 x1);
  * tau, sigma against the SVD: tau sigma |A|^2 just above 0.81 (the power
    iteration's estimate is a lower bound), sigma / tau = omega^2;
  * the autograd gradients against central differences of the same forward
    (h = 1e-6, FP64) for c, rmin, rmax, the column bounds and A's values,
    with inequality rows and infinite bounds present;
  * a planted mistake (the dual clamp's derivative flipped) fails that check.
"""
import numpy as np
import pytest
import torch

from igaos_amd import lpgen
from oracle import unrolled as U

H = 1e-6


def _lp(seed, m=12, n=26, ineq=True):
    lp = lpgen.gen_batch_instance(seed, 0, m_lo=m, m_hi=m, ratio_lo=n / m, ratio_hi=n / m,
                                  k_lo=3, k_hi=4)
    if ineq:
        rng = np.random.default_rng(seed)
        sel = np.arange(lp.m) % 3 == 1
        lp.rmax = np.where(sel, lp.rmax + rng.uniform(0.0, 1.0, lp.m), lp.rmax)
        lp.rmin = np.where(sel, -np.inf, lp.rmin)
        up = np.arange(lp.n) % 5 == 2
        lp.cu = np.where(up, np.inf, lp.cu)
    return lp


def test_steps_against_svd():
    lp = _lp(1)
    tau, sigma = U.steps(lp)
    nrm = np.linalg.norm(lp.to_scipy().toarray(), 2)
    # 30 power iterations approach |A|^2 from below (a Rayleigh quotient):
    # the product sits just above 0.81 and far below the PDHG bound 1
    assert 0.81 - 1e-12 <= tau * sigma * nrm ** 2 <= 0.82
    q = np.where(np.isfinite(lp.rmax), lp.rmax, np.where(np.isfinite(lp.rmin), lp.rmin, 0))
    om = min(max(np.linalg.norm(lp.c) / np.linalg.norm(q), 0.3), 3.0)
    assert abs(sigma / tau - om ** 2) <= 1e-12 * om ** 2


def test_one_step_closed_form():
    lp = _lp(2)
    tau, sigma = U.steps(lp)
    r = U.unrolled(lp, 1)
    A = lp.to_scipy().toarray()
    x1 = np.clip(-tau * lp.c, lp.cl, lp.cu)
    t = sigma * A @ (2 * x1)
    y1 = t - sigma * np.clip(t / sigma, lp.rmin, lp.rmax)
    assert np.allclose(r["x"], x1, rtol=0, atol=1e-15)
    assert np.allclose(r["y"], y1, rtol=0, atol=1e-13)


def _fd(lp, K, w, v, name, u):
    tau, sigma = U.steps(lp)
    vals = []
    for s in (H, -H):
        d = {k: torch.from_numpy(np.asarray(getattr(lp, k), dtype=np.float64).copy())
             for k in ("c", "rmin", "rmax", "cl", "cu", "ax")}
        d[name] = d[name] + s * torch.from_numpy(u)
        x, y, obj = U.forward(lp, K, d["c"], d["rmin"], d["rmax"], d["cl"], d["cu"],
                              d["ax"], tau, sigma)
        vals.append(float(w @ x.numpy() + v @ y.numpy() + 0.7 * float(obj)))
    return (vals[0] - vals[1]) / (2 * H)


@pytest.mark.parametrize("K", [5, 40])
def test_gradients_against_central_differences(K):
    lp = _lp(3)
    rng = np.random.default_rng(K)
    w, v = rng.standard_normal(lp.n), rng.standard_normal(lp.m)
    r = U.unrolled(lp, K, gx=w, gy=v, gobj=0.7)
    for name, g, size in (("c", r["gc"], lp.n), ("rmin", r["grmin"], lp.m),
                          ("rmax", r["grmax"], lp.m), ("cl", r["gcl"], lp.n),
                          ("cu", r["gcu"], lp.n), ("ax", r["gax"], lp.nnz)):
        u = rng.standard_normal(size)
        base = np.asarray(getattr(lp, name), dtype=np.float64)
        u = np.where(np.isfinite(base), u, 0.0)  # no steps on infinite bounds
        fd = _fd(lp, K, w, v, name, u)
        assert abs(g @ u - fd) <= 1e-6 * max(1.0, abs(fd)), (name, g @ u, fd)


def test_planted_mistake_fails(monkeypatch):
    # the dual clamp passing the gradient inside the box instead of at the
    # bounds: the finite-difference check must catch it
    lp = _lp(4, ineq=False)

    def bad_clamp(v, lo, hi):
        inner = torch.where(v <= lo, lo, torch.where(v >= hi, hi, v))
        return inner.detach() + (v - v.detach()) * (v > lo) * (v < hi)

    K = 20
    rng = np.random.default_rng(0)
    w, v = rng.standard_normal(lp.n), rng.standard_normal(lp.m)
    u = rng.standard_normal(lp.m)
    fd = _fd(lp, K, w, v, "rmin", u)
    good = U.unrolled(lp, K, gx=w, gy=v, gobj=0.7)["grmin"] @ u
    assert abs(good - fd) <= 1e-6 * max(1.0, abs(fd))
    monkeypatch.setattr(U, "_clamp", bad_clamp)
    bad = U.unrolled(lp, K, gx=w, gy=v, gobj=0.7)["grmin"] @ u
    assert abs(bad - fd) > 1e-3 * max(1.0, abs(fd))
