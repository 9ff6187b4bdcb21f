"""ctypes signatures of the batched and adjoint entry points of
include/igaos_pdhg.h (filled in as those entry points are added)."""
import ctypes as C

SIGNATURES = {}
