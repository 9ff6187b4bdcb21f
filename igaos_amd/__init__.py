"""igaos_amd — MI355X-native drop-in for the IGAOS ``pdhg`` engine slot.

Mirrors the reference's Python surface (python/igaos/__init__.py:14-23,
python/igaos_bindings.cpp:90-156): ``solve(path, ..., engine="pdhg")``,
``read_mps``, ``Model``, ``Solution``, ``Status`` and ``has_pdhg``.  Every
solve runs the HIP engine in ``libigaos_pdhg.so`` through its C-ABI; every
solving entry point raises ImportError when that library has not been built
(there is no CPU fallback).

Additions with no reference counterpart: ``solve_lp`` (in-memory LPs,
``precision=32`` for FP32 + FP64 refinement), and the batched and
differentiable solvers in ``igaos_amd.batch`` and ``igaos_amd.autograd``.
"""
from .api import (  # noqa: F401
    Model,
    Solution,
    Status,
    has_pdhg,
    read_mps,
    solve,
    solve_lp,
    model_from_csr,
    model_from_lp,
    Engine,
)

__version__ = "0.1.0"
