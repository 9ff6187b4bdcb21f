"""ctypes signatures of the batched and adjoint entry points of
include/igaos_pdhg.h."""
import ctypes as C

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class BatchResult(C.Structure):
    _fields_ = [("status", C.c_int), ("last_is_avg", C.c_int),
                ("objective", C.c_double), ("pinf", C.c_double),
                ("dinf", C.c_double), ("rel_gap", C.c_double),
                ("prow", C.c_double), ("iterations", C.c_longlong),
                ("attempts", C.c_longlong), ("restarts", C.c_longlong),
                ("solve_ms", C.c_double), ("eta", C.c_double),
                ("omega", C.c_double), ("lam2", C.c_double)]


class AdjointInfo(C.Structure):
    _fields_ = [("basic_cols", C.c_int), ("active_rows", C.c_int),
                ("iters_p", C.c_int), ("iters_q", C.c_int),
                ("accepted", C.c_int), ("flags", C.c_int),
                ("res_p", C.c_double), ("res_q", C.c_double),
                ("viol", C.c_double)]


SIGNATURES = {
    "igaos_batch_create": (_vp, [C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _dp,
                                 _dp, _ip, _ip, _dp, _ip, _ip, _dp, _dp,
                                 C.c_int]),
    "igaos_batch_run": (C.c_int, [_vp, _vp, _dp, _dp,
                                  C.POINTER(BatchResult)]),
    "igaos_batch_launch": (C.c_int, [_vp, _vp]),
    "igaos_batch_progress": (C.c_int, [_vp]),
    "igaos_batch_wait": (C.c_int, [_vp, _dp, _dp, C.POINTER(BatchResult)]),
    "igaos_batch_set_schedule": (C.c_int, [_vp, _ip, C.c_int, C.c_int]),
    "igaos_batch_get_schedule": (C.c_int, [_vp, _ip, _ip, _ip]),
    "igaos_batch_kernel_ms": (C.c_double, [_vp]),
    "igaos_batch_last_error": (C.c_char_p, []),
    "igaos_batch_destroy": (None, [_vp]),
    "igaos_batch_set_mode": (C.c_int, [_vp, C.c_int]),
    "igaos_batch_get_mode": (C.c_int, [_vp]),
    "igaos_batch_lockstep_timing": (C.c_int, [_vp, _dp, _dp, _dp, _dp,
                                              C.POINTER(C.c_long)]),
    "igaos_batch_update": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "igaos_batch_set_solution": (C.c_int, [_vp, _dp, _dp]),
    "igaos_batch_polish": (C.c_int, [_vp, C.c_double, C.c_int, _dp, _dp, _dp,
                                     C.POINTER(AdjointInfo)]),
    "igaos_batch_adjoint": (C.c_int, [_vp, _dp, _dp, _dp, C.c_double, C.c_int,
                                      _dp, _dp, _dp, _dp, _dp, _dp,
                                      C.POINTER(AdjointInfo)]),
    "igaos_batch_unroll": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
    "igaos_gather_order_build": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), C.POINTER(C.c_ushort),
                                           C.POINTER(C.c_ushort)]),
    "igaos_batch_unroll_backward": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp,
                                              _dp, _dp, _dp]),
}
