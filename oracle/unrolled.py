"""TEST INFRASTRUCTURE ONLY — CPU restatement of the unrolled mode (the
product kernels live in igaos_amd/csrc/unroll.hip), with torch autograd
(CPU, float64) as the reverse-mode reference.

No reference counterpart: the reference (src/pdhg/pdhg.cu) has no
gradients.  The map is K steps of fixed-step PDHG in original units from
x = y = 0, the step of pdhg.cu:32-61 (dual sign convention d = c + A'y)
without line search, restarts or averaging:

  x' = clamp(x - tau (c + A'y), cl, cu),   xb = 2 x' - x,
  t  = y + sigma A xb,                      y' = t - sigma clamp(t / sigma, rmin, rmax),

clamp(v, lo, hi) = lo if v <= lo, else hi if v >= hi, else v (the lower
bound first at a tie), tau = 0.9 / (omega |A|), sigma = 0.9 omega / |A|,
|A|^2 from 30 power iterations of A'A from v = 1/sqrt(n) (pdhg.cu:605-631)
and omega = clamp(|c| / |q|, 0.3, 3) with q_i = rmax_i if finite, else
rmin_i if finite, else 0 (pdhg.cu:647-663); tau and sigma are constants
(no gradient flows through them).  Objective c'x_K + obj_const.

Pinned in tests/test_unrolled_oracle.py by the closed form of one step and
by central differences of this forward.
"""
from __future__ import annotations

import numpy as np
import torch


def steps(lp):
    """(tau, sigma) of the unrolled map for ``lp`` (numpy, float64)."""
    rows = np.repeat(np.arange(lp.m), np.diff(lp.ap))
    cols = np.asarray(lp.ai, dtype=np.int64)
    ax = np.asarray(lp.ax, dtype=np.float64)

    def Av(v):
        return np.bincount(rows, weights=ax * v[cols], minlength=lp.m)

    def ATy(y):
        return np.bincount(cols, weights=ax * y[rows], minlength=lp.n)

    v = np.full(lp.n, 1.0 / np.sqrt(lp.n)) if lp.n else np.zeros(0)
    for _ in range(30):
        w = ATy(Av(v))
        nw = np.sqrt(w @ w)
        if nw < 1e-300:
            break
        v = w / nw
    t = Av(v)
    lam2 = float(t @ t)
    if not (lam2 > 0) or not np.isfinite(lam2):
        lam2 = 1.0
    c = np.asarray(lp.c, dtype=np.float64)
    q = np.where(np.isfinite(lp.rmax), lp.rmax, np.where(np.isfinite(lp.rmin), lp.rmin, 0.0))
    cn, qn = c @ c, q @ q
    om = 1.0
    if cn > 1e-24 and qn > 1e-24:
        om = min(max(np.sqrt(cn / qn), 0.3), 3.0)
    na = np.sqrt(lam2)
    return 0.9 / (om * na), 0.9 * om / na


def _clamp(v, lo, hi):
    return torch.where(v <= lo, lo, torch.where(v >= hi, hi, v))


def forward(lp, K, c, rmin, rmax, cl, cu, ax, tau, sigma):
    """x_K, y_K, objective of the K-step map (torch tensors; the data
    tensors may require grad)."""
    rows = torch.from_numpy(np.repeat(np.arange(lp.m), np.diff(lp.ap)).astype(np.int64))
    cols = torch.from_numpy(np.asarray(lp.ai, dtype=np.int64))
    x = torch.zeros(lp.n, dtype=torch.float64)
    y = torch.zeros(lp.m, dtype=torch.float64)
    for _ in range(K):
        aty = torch.zeros(lp.n, dtype=torch.float64).index_add(0, cols, ax * y[rows])
        xn = _clamp(x - tau * (c + aty), cl, cu)
        xb = 2.0 * xn - x
        axb = torch.zeros(lp.m, dtype=torch.float64).index_add(0, rows, ax * xb[cols])
        t = y + sigma * axb
        y = t - sigma * _clamp(t / sigma, rmin, rmax)
        x = xn
    return x, y, c @ x + float(getattr(lp, "obj_const", 0.0))


def unrolled(lp, K, gx=None, gy=None, gobj=0.0):
    """Forward outputs and the gradients of L = gx.x_K + gy.y_K +
    gobj * objective with respect to c, rmin, rmax, cl, cu and A's values
    (CSR order), by torch autograd."""
    tau, sigma = steps(lp)
    t = lambda a: torch.tensor(np.asarray(a, dtype=np.float64), requires_grad=True)  # noqa
    c, rmin, rmax, cl, cu, ax = (t(lp.c), t(lp.rmin), t(lp.rmax), t(lp.cl), t(lp.cu),
                                 t(lp.ax))
    x, y, obj = forward(lp, K, c, rmin, rmax, cl, cu, ax, tau, sigma)
    loss = obj * float(gobj)
    if gx is not None:
        loss = loss + torch.from_numpy(np.asarray(gx, dtype=np.float64)) @ x
    if gy is not None:
        loss = loss + torch.from_numpy(np.asarray(gy, dtype=np.float64)) @ y
    grads = torch.autograd.grad(loss, (c, rmin, rmax, cl, cu, ax), allow_unused=True)
    g = [np.zeros(v.numel()) if gr is None else gr.numpy() for gr, v in
         zip(grads, (c, rmin, rmax, cl, cu, ax))]
    return dict(x=x.detach().numpy(), y=y.detach().numpy(), objective=float(obj.detach()),
                tau=tau, sigma=sigma, gc=g[0], grmin=g[1], grmax=g[2], gcl=g[3],
                gcu=g[4], gax=g[5])
