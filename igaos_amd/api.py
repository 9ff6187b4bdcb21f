"""Python surface of the pdhg engine slot (see package docstring)."""
from __future__ import annotations

import ctypes as C
import enum
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _core
from ._core import EngineConfig, Options, SolveInfo

_ENGINES = ("auto", "simplex", "pdhg", "milp", "qp")


class Status(enum.IntEnum):
    """Same enumerators as igaos::Status (include/igaos/status.h:5-14)."""
    Optimal = 0
    NearOptimal = 1
    Feasible = 2
    Infeasible = 3
    Unbounded = 4
    IterationLimit = 5
    TimeLimit = 6
    Error = 7


STATUS_NAMES = ["optimal", "near-optimal", "feasible", "infeasible",
                "unbounded", "iteration-limit", "time-limit", "error"]


@dataclass
class Solution:
    """Fields of igaos::Solution (include/igaos/solution.h:10-22).

    ``status`` is the status string (python/igaos_bindings.cpp:107-109),
    ``status_enum`` the enum; ``y`` follows the pdhg convention
    d = c + A'y.  ``info`` carries engine diagnostics.
    """
    status: str = "error"
    status_enum: Status = Status.Error
    objective: float = 0.0
    x: np.ndarray = field(default_factory=lambda: np.zeros(0))
    y: np.ndarray = field(default_factory=lambda: np.zeros(0))
    row_activity: np.ndarray = field(default_factory=lambda: np.zeros(0))
    pinf: float = 0.0
    dinf: float = 0.0
    rel_gap: float = 0.0
    iterations: int = 0
    solve_time_ms: float = 0.0
    message: str = ""
    info: dict = field(default_factory=dict)

    def __repr__(self):
        return (f"<igaos.Solution status='{self.status}' "
                f"objective={self.objective:.6f}>")


class Model:
    """An LP held by the C library (the reference's io::Model)."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError(_core.last_error() or "model construction failed")
        self._h = C.c_void_p(handle)
        m, n, nnz = C.c_int(), C.c_int(), C.c_int()
        oc = C.c_double()
        fr, rg, ni = C.c_int(), C.c_int(), C.c_int()
        _core.lib().igaos_model_dims(self._h, C.byref(m), C.byref(n),
                                     C.byref(nnz), C.byref(oc), C.byref(fr),
                                     C.byref(rg), C.byref(ni))
        self.m, self.n, self._nnz = m.value, n.value, nnz.value
        self.obj_const = oc.value
        self.n_fr_parsed, self.n_ranges_parsed = fr.value, rg.value
        self.n_int = ni.value

    def __del__(self):
        h = getattr(self, "_h", None)
        lib = getattr(_core, "_lib", None) if _core is not None else None
        if h and lib is not None:
            lib.igaos_model_free(h)
        self._h = None

    def nnz(self):
        return self._nnz

    def arrays(self) -> dict:
        """Copy of the model's arrays (rmin, rmax, cl, cu, c, CSR, CSC)."""
        m, n, z = self.m, self.n, self._nnz
        out = dict(rmin=np.zeros(m), rmax=np.zeros(m), cl=np.zeros(n),
                   cu=np.zeros(n), c=np.zeros(n),
                   ap=np.zeros(m + 1, np.int32), ai=np.zeros(z, np.int32),
                   ax=np.zeros(z), cp=np.zeros(n + 1, np.int32),
                   ci=np.zeros(z, np.int32), acx=np.zeros(z))
        ptr = {k: (v.ctypes.data_as(C.POINTER(C.c_int)) if v.dtype == np.int32
                   else v.ctypes.data_as(C.POINTER(C.c_double)))
               for k, v in out.items()}
        _core.lib().igaos_model_get(self._h, ptr["rmin"], ptr["rmax"],
                                    ptr["cl"], ptr["cu"], ptr["c"], ptr["ap"],
                                    ptr["ai"], ptr["ax"], ptr["cp"],
                                    ptr["ci"], ptr["acx"])
        return out

    def to_lp(self, name=""):
        from .lpgen import LP
        a = self.arrays()
        return LP(m=self.m, n=self.n, obj_const=self.obj_const, name=name, **a)

    @property
    def stats(self) -> dict:
        """Row/column class counts (python/igaos_bindings.cpp:51-88)."""
        a = self.arrays()
        e = l = g = ranged = 0
        for lo, hi in zip(a["rmin"], a["rmax"]):
            if lo == hi:
                e += 1
                continue
            fl, fu = math.isfinite(lo), math.isfinite(hi)
            g += fl
            l += fu
            ranged += fl and fu
        nfree = nfixed = nboxed = nones = 0
        for lo, hi in zip(a["cl"], a["cu"]):
            fl, fu = math.isfinite(lo), math.isfinite(hi)
            if not fl and not fu:
                nfree += 1
            elif fl and fu and lo == hi:
                nfixed += 1
            elif fl and fu:
                nboxed += 1
            else:
                nones += 1
        return dict(rows_e=e, rows_l=max(l - ranged, 0), rows_g=g,
                    rows_ranged=ranged, cols_free=nfree, cols_fixed=nfixed,
                    cols_boxed=nboxed, cols_one_sided=nones, n_int=self.n_int,
                    n_fr_parsed=self.n_fr_parsed,
                    n_ranges_parsed=self.n_ranges_parsed,
                    obj_const=self.obj_const)

    def __repr__(self):
        return f"<igaos.Model m={self.m} n={self.n} nnz={self._nnz}>"


def _has_pdhg() -> bool:
    # False when the library is not built yet (e.g. while building it) or
    # no HIP device is visible; every solve still goes through the library
    # and fails loudly if it is missing.
    try:
        return _core.device_count() > 0
    except (ImportError, OSError, AttributeError):
        _core._lib = None
        return False


def read_mps(path: str) -> Model:
    """Parse a free-format MPS file (src/io/model.hpp:69-218)."""
    h = _core.lib().igaos_model_read_mps(os.fsencode(path))
    if not h:
        raise RuntimeError(_core.last_error())
    return Model(h)


def model_from_lp(lp) -> Model:
    """Wrap an ``igaos_amd.lpgen.LP`` (CSR + CSC arrays) as a Model."""
    d = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    i = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    arrs = [d(lp.c), d(lp.cl), d(lp.cu), d(lp.rmin), d(lp.rmax), i(lp.ap),
            i(lp.ai), d(lp.ax), i(lp.cp), i(lp.ci), d(lp.acx)]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_int)) if a.dtype == np.int32
            else a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
    h = _core.lib().igaos_model_from_csr_csc(lp.m, lp.n, *ptrs,
                                             float(lp.obj_const))
    if not h:
        raise ValueError(_core.last_error())
    return Model(h)


def model_from_csr(m, n, c, cl, cu, rmin, rmax, ap, ai, ax, obj_const=0.0) -> Model:
    """A Model from CSR arrays alone; the CSC is built on the GPU
    (igaos_model_from_csr), each column's entries in row order."""
    d = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    i = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    arrs = [d(c), d(cl), d(cu), d(rmin), d(rmax), i(ap), i(ai), d(ax)]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_int)) if a.dtype == np.int32
            else a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
    h = _core.lib().igaos_model_from_csr(int(m), int(n), *ptrs, float(obj_const))
    if not h:
        raise ValueError(_core.last_error())
    return Model(h)


def polish(model: Model, x, y, tolerance=1e-12, max_iterations=None):
    """Active-set polishing of one LP's solution (igaos_pdhg_polish):
    returns (x, y, objective, info); info["accepted"] says whether the
    polished vertex replaced the input pair."""
    from ._batch_sigs import AdjointInfo
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    if x.shape != (model.n,) or y.shape != (model.m,):
        raise ValueError(f"x must have {model.n} entries and y {model.m}")
    if max_iterations is None:
        max_iterations = int(min(200000, 20 * (model.m + model.n)))
    xo, yo, ob = np.zeros(model.n), np.zeros(model.m), C.c_double()
    info = AdjointInfo()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    if _core.lib().igaos_pdhg_polish(model._h, dp(x), dp(y), float(tolerance),
                                     int(max_iterations), dp(xo), dp(yo), C.byref(ob),
                                     C.byref(info)) != 0:
        raise RuntimeError(_core.last_error())
    return xo, yo, ob.value, {k: getattr(info, k) for k, _ in AdjointInfo._fields_}


def dual_repair_device(model: Model, x, y_it, col_tol=1e-6):
    """The engine's device dual repair on a point (igaos_dual_repair_device):
    returns (found, y, info)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y_it = np.ascontiguousarray(y_it, dtype=np.float64)
    y = np.zeros(model.m)
    out = np.zeros(6)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    rc = _core.lib().igaos_dual_repair_device(model._h, dp(x), dp(y_it), float(col_tol),
                                              dp(y), dp(out))
    if rc < 0:
        raise RuntimeError(_core.last_error())
    keys = ("gap", "dinf", "dual_bound", "iterations", "cg_iterations", "columns_in_J")
    return rc == 1, y, dict(zip(keys, out.tolist()))


def crossover_basis(model: Model, x, y, tol=1e-6, fill=True):
    """Starting basis for the reference's revised simplex from a
    near-optimal (x, y) (igaos_crossover_basis): returns (basis, nb_state,
    count) in the simplex's WarmStart layout."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    basis = np.zeros(model.m, np.int32)
    st = np.zeros(model.n + model.m, np.uint8)
    k = _core.lib().igaos_crossover_basis(
        model._h, x.ctypes.data_as(C.POINTER(C.c_double)),
        y.ctypes.data_as(C.POINTER(C.c_double)), float(tol), int(bool(fill)),
        basis.ctypes.data_as(C.POINTER(C.c_int)),
        st.ctypes.data_as(C.POINTER(C.c_ubyte)))
    if k < 0:
        raise ValueError(_core.last_error())
    return basis, st, k


def _run(model: Model, opt: Options, cfg: EngineConfig | None,
         engine=None) -> Solution:
    L = _core.lib()
    x = np.zeros(model.n)
    y = np.zeros(model.m)
    info = SolveInfo()
    msg = C.create_string_buffer(512)
    xp = x.ctypes.data_as(C.POINTER(C.c_double))
    yp = y.ctypes.data_as(C.POINTER(C.c_double))
    if engine is None:
        rc = L.igaos_pdhg_solve(model._h, C.byref(opt),
                                C.byref(cfg) if cfg is not None else None,
                                xp, yp, C.byref(info), msg, 512)
    else:
        rc = L.igaos_pdhg_run(engine, C.byref(opt), xp, yp, C.byref(info),
                              msg, 512)
    if rc != 0:
        raise ValueError(_core.last_error())
    act = np.zeros(0)
    if Status(info.status) != Status.Error:
        # A x as the reference simplex fills it (igaos_model_row_activity)
        act = np.zeros(model.m)
        if L.igaos_model_row_activity(model._h, xp, act.ctypes.data_as(
                C.POINTER(C.c_double))) != 0:
            raise RuntimeError(_core.last_error())
    return _solution(info, x, y, msg, act)


class Engine:
    """A prepared engine (upload + scaling + power iteration done once).

    ``run(...)`` solves afresh from x = y = 0 on the HBM-resident data and
    returns the same Solution ``solve_lp`` would.
    """

    def __init__(self, lp, precision=64, graph_slots=None, max_repair_m=None,
                 column_slices=None, exact=False, xcd_pinned=None, scaling="ruiz",
                 small_runner=None, stream_loads=None):
        self.model = lp if isinstance(lp, Model) else model_from_lp(lp)
        self.precision = precision
        cfg = _config(precision, graph_slots, max_repair_m, column_slices, exact,
                      xcd_pinned, scaling, small_runner, stream_loads)
        h = _core.lib().igaos_pdhg_create(self.model._h, C.byref(cfg))
        if not h:
            raise RuntimeError(_core.last_error())
        self._h = C.c_void_p(h)
        self.setup_ms = _core.lib().igaos_pdhg_setup_ms(self._h)

    def run(self, time_limit=120.0, tolerance=1e-4, max_iterations=500000,
            verbosity=0) -> Solution:
        return _run(self.model, _options(time_limit, tolerance,
                                         max_iterations, True, 0, verbosity),
                    None, engine=self._h)

    def profile_attempts(self, attempts=20):
        """Mean device ms per launch of (k_primal, k_dual) and grid sizes."""
        a, b = C.c_double(), C.c_double()
        ga, gb = C.c_int(), C.c_int()
        rc = _core.lib().igaos_pdhg_profile_attempts(
            self._h, int(attempts), C.byref(a), C.byref(b), C.byref(ga),
            C.byref(gb))
        if rc != 0:
            raise RuntimeError(_core.last_error())
        return dict(ms_primal=a.value, ms_dual=b.value,
                    blocks_primal=ga.value, blocks_dual=gb.value)

    def column_slices(self):
        """Column slices of the primal-step (A'y) and dual-step (A xbar)
        products; 0 means the fused row-block kernels."""
        p, d = C.c_int(), C.c_int()
        if _core.lib().igaos_pdhg_column_slices(self._h, C.byref(p), C.byref(d)) != 0:
            raise RuntimeError(_core.last_error())
        return dict(primal=p.value, dual=d.value)

    def close(self):
        h = getattr(self, "_h", None)
        lib = getattr(_core, "_lib", None) if _core is not None else None
        if h and lib is not None:
            lib.igaos_pdhg_destroy(h)
        self._h = None

    def __del__(self):
        self.close()


def _solution(info, x, y, msg, row_activity) -> Solution:
    st = Status(info.status)
    extra = {k: getattr(info, k) for k, _ in SolveInfo._fields_}
    return Solution(status=STATUS_NAMES[st], status_enum=st,
                    objective=info.objective, x=x, y=y,
                    row_activity=row_activity, pinf=info.pinf, dinf=info.dinf,
                    rel_gap=info.rel_gap, iterations=info.iterations,
                    solve_time_ms=info.solve_time_ms,
                    message=msg.value.decode(errors="replace"), info=extra)


def _options(time_limit, tolerance, max_iterations, presolve, seed,
             verbosity=0) -> Options:
    o = Options()
    _core.lib().igaos_options_default(C.byref(o))
    o.time_limit_s = float(time_limit)
    o.tolerance = float(tolerance)
    o.max_iterations = int(max_iterations)
    o.presolve = int(bool(presolve))
    o.seed = int(seed)
    o.verbosity = int(verbosity)
    return o


# diagonal preconditioning: Ruiz (the reference's) or Ruiz + Pock-Chambolle
SCALINGS = {"ruiz": 0, "ruiz+pock-chambolle": 1}


def _config(precision=64, graph_slots=None, max_repair_m=None,
            column_slices=None, exact=False, xcd_pinned=None,
            scaling="ruiz", small_runner=None, stream_loads=None) -> EngineConfig:
    c = EngineConfig()
    _core.lib().igaos_engine_config_default(C.byref(c))
    c.precision = int(precision)
    if scaling not in SCALINGS:
        raise ValueError(f"scaling must be one of {sorted(SCALINGS)}, got {scaling!r}")
    c.scaling = SCALINGS[scaling]
    c.exact = int(bool(exact))
    if xcd_pinned is not None:
        c.xcd_pinned = int(bool(xcd_pinned))
    if graph_slots is not None:
        c.graph_slots = int(graph_slots)
    if max_repair_m is not None:
        c.max_repair_m = int(max_repair_m)
    if column_slices is not None:
        c.column_slices = int(column_slices)
    if small_runner is not None:
        c.small_runner = int(small_runner)
    if stream_loads is not None:
        c.stream_loads = int(stream_loads)
    return c


def solve(path, time_limit=120.0, tolerance=1e-4, max_iterations=500000,
          presolve=True, seed=0, engine="auto") -> Solution:
    """``igaos.solve`` (python/igaos_bindings.cpp:26-46, 144-147).

    ``engine="pdhg"`` runs the MI355X engine.  The other valid names
    (auto, simplex, milp, qp) belong to engines outside the pdhg slot and
    return an ``error`` Solution; unknown names raise ValueError.
    """
    if engine not in _ENGINES:
        raise ValueError(f"engine must be one of auto|simplex|pdhg|milp|qp, "
                         f"got '{engine}'")
    model = read_mps(path)
    if engine != "pdhg":
        return Solution(message=f"engine '{engine}' is not part of this build "
                                f"(pdhg engine slot only)")
    return _run(model, _options(time_limit, tolerance, max_iterations,
                                presolve, seed), None)


def solve_lp(lp, time_limit=120.0, tolerance=1e-4, max_iterations=500000,
             precision=64, verbosity=0, graph_slots=None,
             max_repair_m=None, column_slices=None, exact=False,
             xcd_pinned=None, scaling="ruiz", small_runner=None,
             stream_loads=None) -> Solution:
    """Solve an in-memory LP (``lpgen.LP`` or ``Model``) with the engine.

    ``precision=64`` is the plain FP64 engine; ``precision=32`` runs FP32
    iterations on an FP64-refined correction problem.  ``exact=True``
    (FP64, iteration-bounded runs) fixes every summation order to the
    reference's sequential one and disables fused multiply-adds, so the run
    is bit-reproducible against a sequential FP64 host replay.
    ``small_runner`` (-1 or 1: wherever the LP fits, 0: never): small LPs
    run a check period of attempts per launch in one workgroup (exact
    mode: the exact runner; default mode: the register / LDS runner for
    max(n, m) <= 3072 whose products, x and y fit in LDS).  ``stream_loads`` (-1: nontemporal when both matrices
    exceed the last-level cache, 0: cached, 1: nontemporal) picks how the
    attempt kernels load the matrix streams; results do not depend on it.
    """
    model = lp if isinstance(lp, Model) else model_from_lp(lp)
    return _run(model, _options(time_limit, tolerance, max_iterations, True,
                                0, verbosity),
                _config(precision, graph_slots, max_repair_m, column_slices, exact,
                        xcd_pinned, scaling, small_runner, stream_loads))


has_pdhg = _has_pdhg()
