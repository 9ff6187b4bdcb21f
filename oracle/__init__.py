"""TEST INFRASTRUCTURE ONLY — ctypes front end of the oracle.

The oracle (``pdhg_oracle.c``) is a plain-C FP64 host restatement of the
reference PDHG engine (src/pdhg/pdhg.cu).  It is the checker for the HIP
engine and is imported only by ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg; the product (``igaos_amd``) never imports
it.  ``_ref/libref_simplex.so`` is the reference's own revised simplex
(src/simplex) compiled from its sources by ``Makefile``; it supplies
objective fixtures and a CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_simplex.so")

STATUS_NAMES = ["optimal", "near-optimal", "feasible", "infeasible",
                "unbounded", "iteration-limit", "time-limit", "error"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _Model(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int),
                ("rmin", _dp), ("rmax", _dp), ("cl", _dp), ("cu", _dp),
                ("c", _dp), ("ap", _ip), ("ai", _ip), ("ax", _dp),
                ("cp", _ip), ("ci", _ip), ("acx", _dp),
                ("obj_const", C.c_double)]


class _Options(C.Structure):
    _fields_ = [("time_limit_s", C.c_double), ("tolerance", C.c_double),
                ("max_iterations", C.c_long), ("verbosity", C.c_int),
                ("scaling", C.c_int)]


class Metrics(C.Structure):
    _fields_ = [("pinf", C.c_double), ("dinf", C.c_double),
                ("gap", C.c_double), ("op", C.c_double), ("od", C.c_double),
                ("prow", C.c_double)]


class _Result(C.Structure):
    _fields_ = [("status", C.c_int), ("objective", C.c_double),
                ("pinf", C.c_double), ("dinf", C.c_double),
                ("rel_gap", C.c_double), ("iterations", C.c_long),
                ("solve_time_ms", C.c_double), ("attempts", C.c_long),
                ("restarts", C.c_long), ("eta", C.c_double),
                ("omega", C.c_double), ("lam2", C.c_double),
                ("eta0", C.c_double), ("omega0", C.c_double),
                ("last_is_avg", C.c_int), ("used_incumbent", C.c_int),
                ("repaired", C.c_int), ("rollbacks", C.c_long)]


_lib = None
_ref = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.orc_pdhg_solve.argtypes = [C.POINTER(_Model), C.POINTER(_Options),
                                       _dp, _dp, C.POINTER(_Result), _dp,
                                       C.c_long]
        lib.orc_evaluate.restype = Metrics
        lib.orc_evaluate.argtypes = [C.c_int, C.c_int] + [_dp] * 9
        lib.orc_power.restype = C.c_double
        lib.orc_power.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _ip, _ip,
                                  _dp]
        lib.orc_omega0.restype = C.c_double
        lib.orc_omega0.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        lib.orc_spmv.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        lib.orc_ruiz.argtypes = [C.POINTER(_Model), _dp, _dp]
        lib.orc_pock_chambolle.argtypes = [C.POINTER(_Model), _dp, _dp]
        lib.orc_scale.argtypes = [C.POINTER(_Model), _dp, _dp] + [_dp] * 7
        lib.orc_dots.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        lib.orc_dual_repair.argtypes = [C.POINTER(_Model), _dp, _dp, _dp,
                                        C.c_double, _dp]
        lib.orc_set_mutation.argtypes = [C.c_int]
        lib.orc_set_threads.argtypes = [C.c_int]
        lib.orc_avg_weight.restype = C.c_double
        lib.orc_avg_weight.argtypes = [C.c_long]
        lib.orc_mix_weights.argtypes = [C.c_int, _dp]
        lib.orc_ls_decide.restype = C.c_int
        lib.orc_ls_decide.argtypes = [C.c_long, _dp, C.c_double, _dp,
                                      C.POINTER(C.c_int)]
        _lib = lib
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pd(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


class _Held:
    """Keeps the numpy buffers behind an orc_model alive."""

    def __init__(self, lp):
        self.arrs = dict(rmin=_d(lp.rmin), rmax=_d(lp.rmax), cl=_d(lp.cl),
                         cu=_d(lp.cu), c=_d(lp.c), ap=_i(lp.ap), ai=_i(lp.ai),
                         ax=_d(lp.ax), cp=_i(lp.cp), ci=_i(lp.ci),
                         acx=_d(lp.acx))
        a = self.arrs
        self.model = _Model(lp.m, lp.n, _pd(a["rmin"]), _pd(a["rmax"]),
                            _pd(a["cl"]), _pd(a["cu"]), _pd(a["c"]),
                            _pi(a["ap"]), _pi(a["ai"]), _pd(a["ax"]),
                            _pi(a["cp"]), _pi(a["ci"]), _pd(a["acx"]),
                            float(lp.obj_const))


@dataclass
class Result:
    status: str
    objective: float
    pinf: float
    dinf: float
    rel_gap: float
    iterations: int
    solve_time_ms: float
    attempts: int
    restarts: int
    eta: float
    omega: float
    lam2: float
    eta0: float
    omega0: float
    last_is_avg: bool
    used_incumbent: bool
    repaired: bool
    x: np.ndarray
    y: np.ndarray
    trace_eta: np.ndarray | None = None
    rollbacks: int = 0


def set_mutation(code: int):
    load().orc_set_mutation(int(code))


def set_threads(k: int):
    """Host threads of the oracle's SpMV (rows split; results unchanged)."""
    load().orc_set_threads(int(k))


def solve(lp, time_limit=120.0, tolerance=1e-4, max_iterations=500000,
          verbosity=0, trace=0, scaling=0) -> Result:
    """Host FP64 replay of igaos::pdhg::solve (src/pdhg/pdhg.cu:457-1156);
    ``scaling=1`` adds Pock-Chambolle after Ruiz (orc_pock_chambolle)."""
    lib = load()
    h = _Held(lp)
    opt = _Options(float(time_limit), float(tolerance), int(max_iterations),
                   int(verbosity), int(scaling))
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    r = _Result()
    tr = np.zeros(max(trace, 1))
    lib.orc_pdhg_solve(C.byref(h.model), C.byref(opt), _pd(x), _pd(y),
                       C.byref(r), _pd(tr) if trace else None, int(trace))
    return Result(STATUS_NAMES[r.status], r.objective, r.pinf, r.dinf,
                  r.rel_gap, r.iterations, r.solve_time_ms, r.attempts,
                  r.restarts, r.eta, r.omega, r.lam2, r.eta0, r.omega0,
                  bool(r.last_is_avg), bool(r.used_incumbent),
                  bool(r.repaired), x, y,
                  tr[:min(trace, r.iterations)].copy() if trace else None,
                  r.rollbacks)


def spmv(ptr, idx, val, x):
    lib = load()
    ptr, idx, val, x = _i(ptr), _i(idx), _d(val), _d(x)
    out = np.zeros(ptr.shape[0] - 1)
    lib.orc_spmv(ptr.shape[0] - 1, _pi(ptr), _pi(idx), _pd(val), _pd(x),
                 _pd(out))
    return out


def ruiz(lp):
    lib = load()
    h = _Held(lp)
    dr = np.zeros(lp.m)
    dc = np.zeros(lp.n)
    lib.orc_ruiz(C.byref(h.model), _pd(dr), _pd(dc))
    return dr, dc


def pock_chambolle(lp, dr, dc):
    """Pock-Chambolle (alpha = 1) factors on top of (dr, dc)."""
    lib = load()
    h = _Held(lp)
    dr = np.ascontiguousarray(dr, dtype=np.float64).copy()
    dc = np.ascontiguousarray(dc, dtype=np.float64).copy()
    lib.orc_pock_chambolle(C.byref(h.model), _pd(dr), _pd(dc))
    return dr, dc


def scale(lp, dr, dc):
    lib = load()
    h = _Held(lp)
    dr, dc = _d(dr), _d(dc)
    out = dict(sax=np.zeros(lp.nnz), sacx=np.zeros(lp.nnz), sc=np.zeros(lp.n),
               scl=np.zeros(lp.n), scu=np.zeros(lp.n), srmin=np.zeros(lp.m),
               srmax=np.zeros(lp.m))
    lib.orc_scale(C.byref(h.model), _pd(dr), _pd(dc),
                  *[_pd(out[k]) for k in ("sax", "sacx", "sc", "scl", "scu",
                                          "srmin", "srmax")])
    return out


def power(m, n, ap, ai, sax, cp, ci, sacx):
    lib = load()
    ap, ai, cp, ci, sax, sacx = _i(ap), _i(ai), _i(cp), _i(ci), _d(sax), _d(sacx)
    return lib.orc_power(m, n, _pi(ap), _pi(ai), _pd(sax), _pi(cp), _pi(ci),
                         _pd(sacx))


def omega0(sc, srmin, srmax):
    lib = load()
    sc, srmin, srmax = _d(sc), _d(srmin), _d(srmax)
    return lib.orc_omega0(srmin.shape[0], sc.shape[0], _pd(sc), _pd(srmin),
                          _pd(srmax))


def dots(dx, dy, adx):
    lib = load()
    dx, dy, adx = _d(dx), _d(dy), _d(adx)
    out = np.zeros(3)
    lib.orc_dots(dx.shape[0], dy.shape[0], _pd(dx), _pd(dy), _pd(adx),
                 _pd(out))
    return out


def evaluate(rmin, rmax, cl, cu, c, x, y, Ax, d) -> Metrics:
    lib = load()
    arrs = [_d(a) for a in (rmin, rmax, cl, cu, c, x, y, Ax, d)]
    return lib.orc_evaluate(arrs[0].shape[0], arrs[2].shape[0],
                            *[_pd(a) for a in arrs])


def mix_weights(count):
    """Weight of each of `count` accepted steps in the running average."""
    out = np.zeros(count)
    load().orc_mix_weights(int(count), _pd(out))
    return out


def ls_decide(k, eta, omega, dots, retries=0):
    """One line-search decision: returns (accept, new_eta, new_retries)."""
    e = C.c_double(eta)
    r = C.c_int(retries)
    d = _d(dots)
    acc = load().orc_ls_decide(int(k), C.byref(e), float(omega), _pd(d),
                               C.byref(r))
    return bool(acc), e.value, r.value


def dual_repair(lp, x, Ax, y_it, col_tol):
    lib = load()
    h = _Held(lp)
    x, Ax, y_it = _d(x), _d(Ax), _d(y_it)
    y = np.zeros(lp.m)
    ok = lib.orc_dual_repair(C.byref(h.model), _pd(x), _pd(Ax), _pd(y_it),
                             float(col_tol), _pd(y))
    return bool(ok), y


# ---- the reference simplex, built from its sources into _ref/ ----------

def have_ref_simplex() -> bool:
    return os.path.exists(REF_LIB)


def ref_simplex(lp, time_limit=3600.0, warm=None):
    """Solve with the reference revised simplex (src/simplex/simplex.cpp).

    Returns (status, objective, x, y, iterations, solve_ms).  The simplex
    reports y with d = c - A'y, i.e. y_simplex = -y_pdhg.  ``warm`` =
    (basis (m ints), states (n + m bytes)) is passed as its WarmStart.
    """
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_LIB)
        _ref.ref_simplex_solve.restype = C.c_int
        _ref.ref_simplex_solve.argtypes = (
            [C.c_int, C.c_int] + [_dp] * 5 + [_ip, _ip, _dp, _ip, _ip, _dp] +
            [C.c_double, C.c_double, _dp, _dp, _dp, C.POINTER(C.c_long), _dp])
        _ref.ref_simplex_solve_warm.restype = C.c_int
        _ref.ref_simplex_solve_warm.argtypes = (
            _ref.ref_simplex_solve.argtypes + [_ip, C.POINTER(C.c_ubyte)])
    h = _Held(lp)
    a = h.arrs
    obj = C.c_double(0)
    its = C.c_long(0)
    ms = C.c_double(0)
    x = np.zeros(lp.n)
    y = np.zeros(lp.m)
    args = (lp.m, lp.n, _pd(a["rmin"]), _pd(a["rmax"]), _pd(a["cl"]), _pd(a["cu"]),
            _pd(a["c"]), _pi(a["ap"]), _pi(a["ai"]), _pd(a["ax"]), _pi(a["cp"]),
            _pi(a["ci"]), _pd(a["acx"]), float(lp.obj_const), float(time_limit),
            C.byref(obj), _pd(x), _pd(y), C.byref(its), C.byref(ms))
    if warm is None:
        st = _ref.ref_simplex_solve(*args)
    else:
        wb = np.ascontiguousarray(warm[0], dtype=np.int32)
        ws = np.ascontiguousarray(warm[1], dtype=np.uint8)
        st = _ref.ref_simplex_solve_warm(*args, _pi(wb),
                                         ws.ctypes.data_as(C.POINTER(C.c_ubyte)))
    return STATUS_NAMES[st], obj.value, x, y, its.value, ms.value
