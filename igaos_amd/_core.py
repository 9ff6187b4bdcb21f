"""ctypes binding of libigaos_pdhg.so (include/igaos_pdhg.h).

This is the Python side of the C-ABI boundary: the same role the
reference's pybind11 ``_core`` module plays (python/igaos_bindings.cpp),
but over plain C entry points.  Loading fails loudly when the library has
not been built; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libigaos_pdhg.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class Options(C.Structure):
    _fields_ = [("time_limit_s", C.c_double), ("tolerance", C.c_double),
                ("max_iterations", C.c_long), ("seed", C.c_int),
                ("presolve", C.c_int), ("restarts", C.c_int),
                ("verbosity", C.c_int)]


class EngineConfig(C.Structure):
    _fields_ = [("precision", C.c_int), ("graph_slots", C.c_int),
                ("max_repair_m", C.c_int), ("device", C.c_int),
                ("column_slices", C.c_int), ("xcd_pinned", C.c_int),
                ("scaling", C.c_int), ("exact", C.c_int),
                ("small_runner", C.c_int), ("stream_loads", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("objective", C.c_double),
                ("pinf", C.c_double), ("dinf", C.c_double),
                ("rel_gap", C.c_double), ("iterations", C.c_long),
                ("solve_time_ms", C.c_double), ("setup_ms", C.c_double),
                ("attempts", C.c_long), ("restarts", C.c_long),
                ("checks", C.c_long), ("refinements", C.c_long),
                ("eta", C.c_double), ("omega", C.c_double),
                ("lam2", C.c_double), ("eta0", C.c_double),
                ("omega0", C.c_double), ("last_is_avg", C.c_int),
                ("used_incumbent", C.c_int), ("repaired", C.c_int),
                ("kernel_ms", C.c_double), ("rollbacks", C.c_long)]


# Every symbol include/igaos_pdhg.h declares: (name, restype, argtypes).
SIGNATURES = {
    "igaos_pdhg_abi_version": (C.c_int, []),
    "igaos_pdhg_device_count": (C.c_int, []),
    "igaos_last_error": (C.c_char_p, []),
    "igaos_status_name": (C.c_char_p, [C.c_int]),
    "igaos_model_read_mps": (_vp, [C.c_char_p]),
    "igaos_model_from_triplets": (_vp, [C.c_int, C.c_int, C.c_int, _ip, _ip,
                                        _dp, _dp, _dp, _dp, _dp, _dp,
                                        C.c_double]),
    "igaos_model_from_csr_csc": (_vp, [C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                       _dp, _ip, _ip, _dp, _ip, _ip, _dp,
                                       C.c_double]),
    "igaos_model_from_csr": (_vp, [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip,
                                   _dp, C.c_double]),
    "igaos_model_free": (None, [_vp]),
    "igaos_model_row_activity": (C.c_int, [_vp, _dp, _dp]),
    "igaos_dual_repair": (C.c_int, [_vp, _dp, _dp, _dp, C.c_double, _dp]),
    "igaos_kkt_evaluate": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "igaos_dual_repair_device": (C.c_int, [_vp, _dp, _dp, C.c_double, _dp, _dp]),
    "igaos_pdhg_polish": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_int, _dp, _dp, _dp,
                                    _vp]),
    "igaos_crossover_basis": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_int, _ip,
                                        C.POINTER(C.c_ubyte)]),
    "igaos_model_dims": (None, [_vp, _ip, _ip, _ip, _dp, _ip, _ip, _ip]),
    "igaos_model_get": (None, [_vp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp,
                               _ip, _ip, _dp]),
    "igaos_options_default": (None, [C.POINTER(Options)]),
    "igaos_engine_config_default": (None, [C.POINTER(EngineConfig)]),
    "igaos_pdhg_solve": (C.c_int, [_vp, C.POINTER(Options),
                                   C.POINTER(EngineConfig), _dp, _dp,
                                   C.POINTER(SolveInfo), C.c_char_p,
                                   C.c_int]),
    "igaos_pdhg_create": (_vp, [_vp, C.POINTER(EngineConfig)]),
    "igaos_pdhg_run": (C.c_int, [_vp, C.POINTER(Options), _dp, _dp,
                                 C.POINTER(SolveInfo), C.c_char_p, C.c_int]),
    "igaos_pdhg_setup_ms": (C.c_double, [_vp]),
    "igaos_pdhg_column_slices": (C.c_int, [_vp, _ip, _ip]),
    "igaos_pdhg_destroy": (None, [_vp]),
    "igaos_pdhg_profile_attempts": (C.c_int, [_vp, C.c_int, _dp, _dp, _ip,
                                              _ip]),
    "igaos_rowblocks_build": (C.c_int, [C.c_int, _ip, _ip, C.c_int]),
    "igaos_slices_build": (C.c_int, [C.c_int, C.c_int, _ip, _ip, C.c_int, _ip, _ip, _ip,
                                     _ip, _ip, C.c_int]),
    "igaos_small_ell_layout": (C.c_long, [C.c_int, _ip, C.c_int, _ip, _ip, _ip, C.c_long]),
    "igaos_slices_gather_order": (C.c_int, [C.c_int, C.c_int, _ip, _ip, C.c_int, _ip, _ip,
                                            _vp]),
    "igaos_spmv_csr_f64": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _vp]),
    "igaos_spmv_csr_f32": (C.c_int, [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _vp]),
    "igaos_ruiz_f64": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp,
                                 _vp, _vp, _vp, _vp]),
    "igaos_csr_to_csc_f64": (C.c_int, [C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp,
                                       _vp, _vp]),
    "igaos_ls_dots_work": (C.c_long, [C.c_int, C.c_int]),
    "igaos_ls_dots_f64": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "igaos_ls_dots_f64_noalloc": (C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None


def lib():
    """Load libigaos_pdhg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with igaos_amd.build.build()")
        # One HIP runtime per process: torch ships its own libamdhip64.so.7
        # with the same soname as /opt/rocm's.  Loading torch first makes
        # the library bind to that runtime, so device pointers and streams
        # handed over from torch tensors are valid for the library.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _extra_signatures(L)
        _lib = L
    return _lib


def _extra_signatures(L):
    # batched and adjoint entry points, when the library provides them
    from . import _batch_sigs
    for name, (res, args) in _batch_sigs.SIGNATURES.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args


def last_error() -> str:
    return lib().igaos_last_error().decode()


def device_count() -> int:
    return int(lib().igaos_pdhg_device_count())
