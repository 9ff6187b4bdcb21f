"""Differentiable batched LP solve (no reference counterpart).

``DiffLPBatch(lps)`` fixes the instances' sparsity patterns and column
bounds; calling it with tensors ``c`` (sum n), ``rmin``, ``rmax`` (sum m)
and ``a_values`` (sum nnz, CSR order of each instance) -- concatenated over
the instances, any of them ``requires_grad`` -- solves every instance on the
GPU with the batched persistent PDHG engine and returns ``x*`` (sum n),
``y*`` (sum m, dual sign convention d = c + A'y) and the objectives
(count).  With ``polish`` (default) each PDHG solution is first polished on
its active set (igaos_batch_polish: the exact vertex of the identified
basis, kept only if feasible to 1e-9), so x*, y* and the gradients carry
FP64 accuracy rather than the solve tolerance.  ``backward`` runs the batched adjoint kernel
(igaos_batch_adjoint: masked CGLS on the transposed fixed-point system of
the PDHG map linearised at the solution) and returns the gradients of c,
rmin, rmax and A's values.  For equality rows pass the same tensor ``b`` as
rmin and rmax (``solve_eq``): autograd adds the two halves to dL/db.

The column bounds are fixed data; their gradients are available from
``BatchSolver.adjoint`` (gcl, gcu) but not through this function.

``differentiation="unrolled"`` replaces the solve by ``unroll_iterations``
steps of fixed-step PDHG from x = y = 0 (igaos_batch_unroll, FP64, iterates
checkpointed on the device every 64 steps) and returns that iterate; the
backward is reverse mode through those steps (igaos_batch_unroll_backward,
each 64-step segment recomputed from its checkpoint).  The step sizes are
``unroll_steps = (tau, sigma)`` or, by default, those of the first call's
data, held fixed for later calls, so the layer is one fixed K-step map and
its gradients are that map's own, exact to rounding, whether or not the
iterate has converged.
"""
from __future__ import annotations

import numpy as np
import torch

from .batch import BatchSolver


def _host(t):
    return t.detach().to("cpu", torch.float64).numpy()


class _SolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, prob, c, rmin, rmax, a_values):
        prob._forward(_host(c), _host(rmin), _host(rmax), _host(a_values))
        ctx.prob = prob
        ctx.generation = prob._generation
        ctx.save_for_backward(c, rmin, rmax, a_values)
        dev, dt = c.device, c.dtype
        x = torch.from_numpy(prob._x).to(dev, dt)
        y = torch.from_numpy(prob._y).to(dev, dt)
        obj = torch.from_numpy(prob._obj).to(dev, dt)
        return x, y, obj

    @staticmethod
    def backward(ctx, gx, gy, gobj):
        prob = ctx.prob
        c, rmin, rmax, a_values = ctx.saved_tensors
        if prob._generation != ctx.generation:
            # a later forward replaced the device state: solve this one again
            prob._forward(_host(c), _host(rmin), _host(rmax), _host(a_values))
        t = lambda a, like: torch.from_numpy(a).to(like.device, like.dtype)  # noqa
        if prob.differentiation == "unrolled":
            g = prob.solver.unroll_backward(
                gx=None if gx is None else _host(gx),
                gy=None if gy is None else _host(gy),
                gobj=None if gobj is None else _host(gobj))
            return (None, t(g["gc"], c), t(g["grmin"], rmin), t(g["grmax"], rmax),
                    t(g["gax"], a_values))
        g = prob.solver.adjoint(
            gx=None if gx is None else _host(gx),
            gy=None if gy is None else _host(gy),
            gobj=None if gobj is None else _host(gobj),
            tolerance=prob.adjoint_tolerance,
            max_iterations=prob.adjoint_max_iterations)
        prob.last_adjoint_info = g["info"]
        return (None, t(g["gc"], c), t(g["grmin"], rmin), t(g["grmax"], rmax),
                t(g["gax"], a_values))


class DiffLPBatch:
    """Batched LP layer: patterns and column bounds from ``lps`` (a list of
    ``igaos_amd.lpgen.LP``); numeric data per call."""

    def __init__(self, lps, precision=32, tolerance=1e-6, time_limit=60.0,
                 max_iterations=500000, adjoint_tolerance=1e-12,
                 adjoint_max_iterations=None, polish=True, mode="persistent",
                 differentiation="implicit", unroll_iterations=1024, unroll_steps=None):
        if differentiation not in ("implicit", "unrolled"):
            raise ValueError("differentiation must be 'implicit' or 'unrolled'")
        self.differentiation = differentiation
        self.unroll_iterations = int(unroll_iterations)
        self.unroll_steps = unroll_steps
        self.last_unroll = None
        self.lps = list(lps)
        self.solver = BatchSolver(self.lps, precision=precision, mode=mode)
        self.polish = polish
        self.last_polish_info = None
        self.tolerance = tolerance
        self.time_limit = time_limit
        self.max_iterations = max_iterations
        self.adjoint_tolerance = adjoint_tolerance
        self.adjoint_max_iterations = adjoint_max_iterations
        self._generation = 0
        self.last_solution = None
        self.last_adjoint_info = None
        self.coff = self.solver._coff
        self.roff = self.solver._roff
        self.zoff = self.solver._zoff

    def _forward(self, c, rmin, rmax, ax):
        self.solver.update(c=c, rmin=rmin, rmax=rmax, ax=ax)
        if self.differentiation == "unrolled":
            u = self.solver.unroll(self.unroll_iterations, steps=self.unroll_steps)
            if self.unroll_steps is None:  # from the first call's data, then held
                self.unroll_steps = (u["tau"].copy(), u["sigma"].copy())
            self._generation += 1
            self.last_unroll = u
            self._x, self._y = u["x"], u["y"]
            self._obj = u["objective"]
            return
        sol = self.solver.run(tolerance=self.tolerance, time_limit=self.time_limit,
                              max_iterations=self.max_iterations)
        self._generation += 1
        self.last_solution = sol
        xs, ys, obj = sol.x, sol.y, sol.objective
        if self.polish:
            xs, ys, obj, info = self.solver.polish(
                tolerance=self.adjoint_tolerance,
                max_iterations=self.adjoint_max_iterations)
            certified = np.array([st == "near-optimal" for st in sol.status])
            rej = np.flatnonzero((info["accepted"] == 0) & certified)
            if rej.size:
                # certified, but the active set was misread at this
                # tolerance: solve those instances again 100x tighter and
                # polish once more (uncertified instances keep their PDHG
                # pair; their gradients are not meaningful)
                cat = lambda v, off, i: v[int(off[i]):int(off[i + 1])]  # noqa
                sub_lps = [self.lps[i].with_values(
                    c=cat(c, self.coff, i), rmin=cat(rmin, self.roff, i),
                    rmax=cat(rmax, self.roff, i), ax=cat(ax, self.zoff, i))
                    for i in rej]
                sub = BatchSolver(sub_lps, precision=self.solver.precision,
                                  mode=self.solver.mode)
                try:
                    sub.run(tolerance=max(self.tolerance * 1e-2, 1e-11),
                            time_limit=self.time_limit,
                            max_iterations=self.max_iterations, want_xy=False)
                    sx, sy, sobj, sinfo = sub.polish(
                        tolerance=self.adjoint_tolerance,
                        max_iterations=self.adjoint_max_iterations)
                finally:
                    sub.close()
                xs, ys, obj = list(xs), list(ys), np.array(obj)
                for k, i in enumerate(rej):
                    if sinfo["accepted"][k]:
                        xs[i], ys[i], obj[i] = sx[k], sy[k], sobj[k]
                        for key in info:
                            info[key][i] = sinfo[key][k]
                self.solver.set_solution(xs, ys)
            self.last_polish_info = info
        self._x = np.concatenate(xs) if xs else np.zeros(0)
        self._y = np.concatenate(ys) if ys else np.zeros(0)
        self._obj = np.asarray(obj, dtype=np.float64)

    def __call__(self, c, rmin, rmax, a_values):
        return _SolveFn.apply(self, c, rmin, rmax, a_values)

    def solve_eq(self, c, b, a_values):
        """Equality rows: rmin = rmax = b."""
        return _SolveFn.apply(self, c, b, b, a_values)

    def data(self, dtype=torch.float32, device="cpu"):
        """The instances' own (c, rmin, rmax, a_values) as tensors."""
        cat = lambda k: torch.from_numpy(np.concatenate([getattr(lp, k) for lp in self.lps]))  # noqa
        return tuple(cat(k).to(device, dtype) for k in ("c", "rmin", "rmax", "ax"))

    def split(self, v, kind):
        """Per-instance views of a concatenated tensor ('n', 'm' or 'nnz')."""
        off = {"n": self.coff, "m": self.roff, "nnz": self.zoff}[kind]
        return [v[int(off[i]):int(off[i + 1])] for i in range(len(self.lps))]

    def close(self):
        self.solver.close()
