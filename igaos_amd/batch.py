"""Batched persistent PDHG over the C-ABI (igaos_batch_* in
include/igaos_pdhg.h): thousands of independent LPs of uneven sizes solved
in one launch, each with its own line search, restarts and termination.
No reference counterpart (the reference solves one LP per call); every
instance follows the single-LP engine's algorithm, except that the dense
host dual repair is not run.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _core
from ._batch_sigs import BatchResult
from .api import STATUS_NAMES, _options


@dataclass
class BatchSolution:
    status: list
    objective: np.ndarray
    pinf: np.ndarray
    dinf: np.ndarray
    rel_gap: np.ndarray
    iterations: np.ndarray
    attempts: np.ndarray
    restarts: np.ndarray
    solve_ms: np.ndarray
    prow: np.ndarray = None
    x: list = field(default_factory=list)
    y: list = field(default_factory=list)
    kernel_ms: float = 0.0

    @property
    def solved(self) -> int:
        return sum(s == "near-optimal" for s in self.status)


def _cat(arrs, dtype):
    return np.ascontiguousarray(np.concatenate(arrs) if arrs else
                                np.zeros(0), dtype=dtype)


class BatchSolver:
    """Uploads a list of ``lpgen.LP`` once; ``run`` solves all of them."""

    def __init__(self, lps, precision=32, mode="persistent"):
        L = _core.lib()
        self.count = len(lps)
        self.precision = precision
        self.m = np.array([lp.m for lp in lps], np.int32)
        self.n = np.array([lp.n for lp in lps], np.int32)
        self.nnz = np.array([lp.nnz for lp in lps], np.int32)
        d = lambda key: _cat([getattr(lp, key) for lp in lps], np.float64)  # noqa
        i = lambda key: _cat([getattr(lp, key) for lp in lps], np.int32)  # noqa
        self._keep = dict(c=d("c"), cl=d("cl"), cu=d("cu"), rmin=d("rmin"),
                          rmax=d("rmax"), ap=i("ap"), ai=i("ai"), ax=d("ax"),
                          cp=i("cp"), ci=i("ci"), acx=d("acx"),
                          oc=np.array([lp.obj_const for lp in lps], np.float64))
        k = self._keep
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa
        h = L.igaos_batch_create(self.count, ip(self.m), ip(self.n),
                                 ip(self.nnz), dp(k["c"]), dp(k["cl"]),
                                 dp(k["cu"]), dp(k["rmin"]), dp(k["rmax"]),
                                 ip(k["ap"]), ip(k["ai"]), dp(k["ax"]),
                                 ip(k["cp"]), ip(k["ci"]), dp(k["acx"]),
                                 dp(k["oc"]), int(precision))
        if not h:
            raise RuntimeError(L.igaos_batch_last_error().decode())
        self._h = C.c_void_p(h)
        self._roff = np.concatenate([[0], np.cumsum(self.m)])
        self._coff = np.concatenate([[0], np.cumsum(self.n)])
        self._zoff = np.concatenate([[0], np.cumsum(self.nnz.astype(np.int64))])
        self._lps = list(lps)
        self.set_mode(mode)

    MODES = ("persistent", "lockstep")

    def set_mode(self, mode):
        """'persistent': one launch, a workgroup per instance at a time;
        'lockstep': all active instances advance one attempt per round."""
        if mode not in self.MODES:
            raise ValueError(f"mode must be one of {self.MODES}, got {mode!r}")
        L = _core.lib()
        if L.igaos_batch_set_mode(self._h, self.MODES.index(mode)) != 0:
            raise RuntimeError(L.igaos_batch_last_error().decode())
        self.mode = mode

    def schedule(self):
        """(lanes per row of each instance, persistent grid, waves/SIMD)."""
        tpr = np.zeros(self.count, np.int32)
        g, w = C.c_int(), C.c_int()
        _core.lib().igaos_batch_get_schedule(
            self._h, tpr.ctypes.data_as(C.POINTER(C.c_int)), C.byref(g),
            C.byref(w))
        return tpr, g.value, w.value

    def set_schedule(self, tpr=None, grid=0, waves=0):
        L = _core.lib()
        t = None if tpr is None else np.ascontiguousarray(tpr, np.int32)
        rc = L.igaos_batch_set_schedule(
            self._h, None if t is None else t.ctypes.data_as(C.POINTER(C.c_int)),
            int(grid), int(waves))
        if rc != 0:
            raise ValueError(L.igaos_batch_last_error().decode())

    def run(self, tolerance=1e-6, time_limit=120.0, max_iterations=500000,
            want_xy=True, progress=None, poll_s=2.0) -> BatchSolution:
        """Solve every instance.  ``progress(done, count)`` is called about
        every ``poll_s`` seconds while the persistent kernel runs (the
        completion itself is noticed within 50 ms)."""
        import time
        L = _core.lib()
        opt = _options(time_limit, tolerance, max_iterations, True, 0)
        res = (BatchResult * self.count)()
        x = np.zeros(int(self._coff[-1])) if want_xy else None
        y = np.zeros(int(self._roff[-1])) if want_xy else None
        if L.igaos_batch_launch(self._h, C.byref(opt)) != 0:
            raise RuntimeError(L.igaos_batch_last_error().decode())
        if progress is not None:
            try:
                last = time.perf_counter()
                progress(0, self.count)
                while True:
                    done = L.igaos_batch_progress(self._h)
                    if done >= self.count:
                        progress(done, self.count)
                        break
                    if time.perf_counter() - last >= poll_s:
                        progress(done, self.count)
                        last = time.perf_counter()
                    time.sleep(0.05)
            except BaseException:
                # the launched run still ends before the batch is usable again
                L.igaos_batch_wait(self._h, None, None, None)
                raise
        rc = L.igaos_batch_wait(
            self._h,
            x.ctypes.data_as(C.POINTER(C.c_double)) if want_xy else None,
            y.ctypes.data_as(C.POINTER(C.c_double)) if want_xy else None, res)
        if rc != 0:
            raise RuntimeError(L.igaos_batch_last_error().decode())
        f = lambda k: np.array([getattr(r, k) for r in res])  # noqa
        sol = BatchSolution(
            status=[STATUS_NAMES[r.status] for r in res],
            objective=f("objective"), pinf=f("pinf"), dinf=f("dinf"),
            rel_gap=f("rel_gap"), iterations=f("iterations"),
            attempts=f("attempts"), restarts=f("restarts"),
            solve_ms=f("solve_ms"), prow=f("prow"),
            kernel_ms=L.igaos_batch_kernel_ms(self._h))
        if want_xy:
            sol.x = [x[self._coff[i]:self._coff[i + 1]] for i in range(self.count)]
            sol.y = [y[self._roff[i]:self._roff[i + 1]] for i in range(self.count)]
        return sol

    def repair(self, sol: BatchSolution, tolerance=1e-6) -> list:
        """Certification fallback at exit for the batched engine, which
        does not run the in-loop dual repair: every instance that stopped
        uncertified with its primal residuals at tolerance (pinf, prow <=
        tol, dinf <= 10 tol, the trigger of src/pdhg/pdhg.cu:882-887) but
        its gap not certified gets the engine's device dual repair
        (igaos_dual_repair_device, both column tiers); where the repaired
        gap is certified the instance becomes near-optimal with the
        repaired y (x is unchanged).  ``sol`` needs x and y (run with
        want_xy=True) and is updated in place; returns the repaired
        instances."""
        from .api import dual_repair_device, model_from_lp
        if not sol.x or not sol.y:
            raise ValueError("repair needs the run's x and y (want_xy=True)")
        done = []
        for i, st in enumerate(sol.status):
            if st == "near-optimal":
                continue
            if not (sol.pinf[i] <= tolerance and sol.prow[i] <= tolerance and
                    sol.dinf[i] <= 10 * tolerance):
                continue
            md = model_from_lp(self._lps[i])
            for col_tol in (1e-6, 1e-1):
                ok, y, info = dual_repair_device(md, sol.x[i], sol.y[i], col_tol)
                if ok and info["gap"] <= tolerance:
                    sol.y[i] = y
                    sol.status[i] = "near-optimal"
                    sol.rel_gap[i] = info["gap"]
                    sol.dinf[i] = info["dinf"]
                    done.append(i)
                    break
        return done

    def update(self, c=None, rmin=None, rmax=None, ax=None, cl=None, cu=None):
        """Replace numeric data in place (concatenated over instances; ax in
        CSR order, the CSC copy follows).  The next ``run`` re-scales."""
        L = _core.lib()
        if ax is not None and not hasattr(self, "_perm"):
            self._perm = np.concatenate(
                [lp.csc_perm() + z for lp, z in zip(self._lps, self._zoff[:-1])])
        arr = {}
        for key, v, size in (("c", c, self._coff[-1]), ("cl", cl, self._coff[-1]),
                             ("cu", cu, self._coff[-1]), ("rmin", rmin, self._roff[-1]),
                             ("rmax", rmax, self._roff[-1]), ("ax", ax, self._zoff[-1])):
            if v is not None:
                v = np.ascontiguousarray(v, dtype=np.float64)
                if v.shape != (size,):
                    raise ValueError(f"{key} must have shape ({size},), got {v.shape}")
                arr[key] = v
        if "ax" in arr:
            arr["acx"] = np.ascontiguousarray(arr["ax"][self._perm])
        dp = lambda k: arr[k].ctypes.data_as(C.POINTER(C.c_double)) if k in arr else None  # noqa
        rc = L.igaos_batch_update(self._h, dp("c"), dp("cl"), dp("cu"), dp("rmin"),
                                  dp("rmax"), dp("ax"), dp("acx"))
        if rc != 0:
            raise RuntimeError(L.igaos_batch_last_error().decode())

    def set_solution(self, x=None, y=None):
        """Replace the last run's solutions (concatenated) that polish and
        adjoint start from (igaos_batch_set_solution)."""
        L = _core.lib()
        conv = lambda v: None if v is None else np.ascontiguousarray(np.concatenate(v) if isinstance(v, list) else v, np.float64)  # noqa
        xa, ya = conv(x), conv(y)
        for key, a, size in (("x", xa, self._coff[-1]), ("y", ya, self._roff[-1])):
            if a is not None and a.shape != (int(size),):
                raise ValueError(f"{key} must have {int(size)} entries in all, got {a.shape}")
        dp = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))  # noqa
        if L.igaos_batch_set_solution(self._h, dp(xa), dp(ya)) != 0:
            raise RuntimeError(L.igaos_batch_last_error().decode())

    def lockstep_timing(self) -> dict:
        """HIP-event timing of sampled lockstep rounds of the last run
        (igaos_batch_lockstep_timing): mean ms and algorithmic bytes per
        launch of the primal and dual attempt kernels."""
        L = _core.lib()
        v = [C.c_double() for _ in range(4)]
        n = C.c_long()
        L.igaos_batch_lockstep_timing(self._h, *[C.byref(t) for t in v], C.byref(n))
        k = max(1, n.value)
        return {"samples": n.value,
                "primal_ms": v[0].value / k, "primal_bytes": v[1].value / k,
                "dual_ms": v[2].value / k, "dual_bytes": v[3].value / k}

    def _info(self, info):
        f = lambda k: np.array([getattr(r, k) for r in info])  # noqa
        return {k: f(k) for k in ("basic_cols", "active_rows", "iters_p",
                                  "iters_q", "accepted", "flags", "res_p",
                                  "res_q", "viol")}

    def polish(self, tolerance=1e-12, max_iterations=None):
        """Active-set polishing of the last run's solutions
        (igaos_batch_polish): returns (x list, y list, objectives, info);
        info["accepted"] marks the instances whose polished pair was kept."""
        from ._batch_sigs import AdjointInfo
        L = _core.lib()
        n, m = int(self._coff[-1]), int(self._roff[-1])
        if max_iterations is None:
            max_iterations = int(min(200000, 20 * (self.m.max() + self.n.max())))
        x, y, obj = np.zeros(n), np.zeros(m), np.zeros(self.count)
        info = (AdjointInfo * self.count)()
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa
        rc = L.igaos_batch_polish(self._h, float(tolerance), int(max_iterations),
                                  dp(x), dp(y), dp(obj), info)
        if rc != 0:
            raise RuntimeError(L.igaos_batch_last_error().decode())
        xs = [x[self._coff[i]:self._coff[i + 1]] for i in range(self.count)]
        ys = [y[self._roff[i]:self._roff[i + 1]] for i in range(self.count)]
        return xs, ys, obj, self._info(info)

    def adjoint(self, gx=None, gy=None, gobj=None, tolerance=1e-10,
                max_iterations=None) -> dict:
        """Gradients through the last run's solutions (include/igaos_pdhg.h,
        igaos_batch_adjoint): returns gc, grmin, grmax, gcl, gcu, gax
        (concatenated, gax in CSR order) and per-instance CGLS info."""
        from ._batch_sigs import AdjointInfo
        L = _core.lib()
        n, m, z = int(self._coff[-1]), int(self._roff[-1]), int(self._zoff[-1])
        if max_iterations is None:
            max_iterations = int(min(200000, 20 * (self.m.max() + self.n.max())))
        ins = {}
        for key, v, size in (("gx", gx, n), ("gy", gy, m), ("gobj", gobj, self.count)):
            if v is not None:
                ins[key] = np.ascontiguousarray(np.broadcast_to(
                    np.asarray(v, dtype=np.float64), (size,)))
        out = dict(gc=np.zeros(n), grmin=np.zeros(m), grmax=np.zeros(m),
                   gcl=np.zeros(n), gcu=np.zeros(n), gax=np.zeros(z))
        info = (AdjointInfo * self.count)()
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa
        rc = L.igaos_batch_adjoint(
            self._h, *[dp(ins[k]) if k in ins else None for k in ("gx", "gy", "gobj")],
            float(tolerance), int(max_iterations),
            *[dp(out[k]) for k in ("gc", "grmin", "grmax", "gcl", "gcu", "gax")], info)
        if rc != 0:
            raise RuntimeError(L.igaos_batch_last_error().decode())
        out["info"] = self._info(info)
        return out

    def unroll(self, iterations, steps=None) -> dict:
        """Fixed-step PDHG unrolled for ``iterations`` steps on the current
        data (include/igaos_pdhg.h, igaos_batch_unroll), with step sizes
        ``steps = (tau, sigma)`` (per instance) or, if None, the ones the
        data give: returns x, y (concatenated), objective, tau and sigma."""
        L = _core.lib()
        n, m = int(self._coff[-1]), int(self._roff[-1])
        out = dict(x=np.zeros(n), y=np.zeros(m), objective=np.zeros(self.count),
                   tau=np.zeros(self.count), sigma=np.zeros(self.count))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa
        st = None
        if steps is not None:
            st = np.ascontiguousarray(np.concatenate([
                np.broadcast_to(np.asarray(s, dtype=np.float64), (self.count,))
                for s in steps]))
        rc = L.igaos_batch_unroll(self._h, int(iterations), None if st is None else dp(st),
                                  *[dp(out[k]) for k in ("x", "y", "objective", "tau",
                                                         "sigma")])
        if rc != 0:
            raise RuntimeError(L.igaos_batch_last_error().decode())
        return out

    def unroll_backward(self, gx=None, gy=None, gobj=None) -> dict:
        """Reverse mode through the last ``unroll`` (igaos_batch_unroll_backward):
        returns gc, grmin, grmax, gcl, gcu, gax (gax in CSR order)."""
        L = _core.lib()
        n, m, z = int(self._coff[-1]), int(self._roff[-1]), int(self._zoff[-1])
        ins = {}
        for key, v, size in (("gx", gx, n), ("gy", gy, m), ("gobj", gobj, self.count)):
            if v is not None:
                ins[key] = np.ascontiguousarray(np.broadcast_to(
                    np.asarray(v, dtype=np.float64), (size,)))
        out = dict(gc=np.zeros(n), grmin=np.zeros(m), grmax=np.zeros(m),
                   gcl=np.zeros(n), gcu=np.zeros(n), gax=np.zeros(z))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa
        rc = L.igaos_batch_unroll_backward(
            self._h, *[dp(ins[k]) if k in ins else None for k in ("gx", "gy", "gobj")],
            *[dp(out[k]) for k in ("gc", "grmin", "grmax", "gcl", "gcu", "gax")])
        if rc != 0:
            raise RuntimeError(L.igaos_batch_last_error().decode())
        return out

    def close(self):
        h = getattr(self, "_h", None)
        lib = getattr(_core, "_lib", None) if _core is not None else None
        if h and lib is not None:
            lib.igaos_batch_destroy(h)
        self._h = None

    def __del__(self):
        self.close()


def solve_batch(lps, precision=32, mode="persistent", **kw) -> BatchSolution:
    """One-shot batched solve of a list of LPs."""
    b = BatchSolver(lps, precision=precision, mode=mode)
    try:
        return b.run(**kw)
    finally:
        b.close()


def solve_node_relaxations(lp, nodes, precision=32, mode="lockstep", repair=True, **kw):
    """LP relaxations of branch-and-bound nodes in one batched solve.

    The reference's MILP solves each node's LP on its own with the simplex
    (src/milp/milp.cpp:399-421); a node's LP is the root LP with the
    node's column bounds (``lp.cl = nd.cl; lp.cu = nd.cu``).  Here
    ``nodes`` is a list of (cl, cu) arrays (n each; +-inf for open bounds)
    and all of them are solved together by the batched engine (one
    instance per node, the root's pattern and values shared by copy).  A
    node whose bounds cross (cl_j > cu_j) is infeasible without a solve and
    reported as such.  Returns the BatchSolution over the nodes (solved
    nodes only carry x, y; infeasible ones have status "infeasible").
    """
    import dataclasses
    keep, out_status = [], []
    insts = []
    for k, (cl, cu) in enumerate(nodes):
        cl = np.ascontiguousarray(cl, dtype=np.float64)
        cu = np.ascontiguousarray(cu, dtype=np.float64)
        if cl.shape != (lp.n,) or cu.shape != (lp.n,):
            raise ValueError(f"node {k}: bounds must have {lp.n} entries")
        if np.any(cl > cu):
            out_status.append("infeasible")
            continue
        keep.append(k)
        out_status.append(None)
        insts.append(dataclasses.replace(lp, cl=cl, cu=cu))
    kw.setdefault("want_xy", True)
    sol = None
    if insts:
        b = BatchSolver(insts, precision=precision, mode=mode)
        try:
            sol = b.run(**kw)
            if repair and kw["want_xy"]:
                b.repair(sol, kw.get("tolerance", 1e-6))
        finally:
            b.close()
    count = len(nodes)
    res = BatchSolution(status=[None] * count, objective=np.full(count, np.nan),
                        pinf=np.full(count, np.nan), dinf=np.full(count, np.nan),
                        rel_gap=np.full(count, np.nan), iterations=np.zeros(count, np.int64),
                        attempts=np.zeros(count, np.int64), restarts=np.zeros(count, np.int64),
                        solve_ms=np.zeros(count), prow=np.full(count, np.nan),
                        x=[None] * count, y=[None] * count,
                        kernel_ms=sol.kernel_ms if sol is not None else 0.0)
    for k, st in enumerate(out_status):
        if st is not None:
            res.status[k] = st
    for j, k in enumerate(keep):
        res.status[k] = sol.status[j]
        for key in ("objective", "pinf", "dinf", "rel_gap", "iterations", "attempts",
                    "restarts", "solve_ms", "prow"):
            getattr(res, key)[k] = getattr(sol, key)[j]
        if sol.x:
            res.x[k], res.y[k] = sol.x[j], sol.y[j]
    return res
