"""igaos_ls_dots_f64_noalloc: the line-search dots with the library's own
scratch (no allocation, no workspace argument) against igaos_ls_dots_f64
with a caller workspace, through the C-ABI, bit for bit, on two calls in a
row (the second smaller, so stale partials of the first would show), and
against the oracle's restatement of the reference layout
(src/pdhg/pdhg.cu:63-125)."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _lib():
    from igaos_amd import _core
    return _core.lib(), _core


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float64)).cuda()


def _p(t):
    return C.c_void_p(t.data_ptr())


def _both(L, core, dx, dy, adx):
    n, m = len(dx), len(dy)
    tdx, tdy, tadx = _dev(dx), _dev(dy), _dev(adx)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    work = torch.full((int(L.igaos_ls_dots_work(n, m)),), float("nan"),
                      dtype=torch.float64, device="cuda")
    out_ws = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda")
    out_na = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda")
    assert L.igaos_ls_dots_f64(n, m, _p(tdx), _p(tdy), _p(tadx), _p(work), _p(out_ws),
                               stream) == 0, core.last_error()
    assert L.igaos_ls_dots_f64_noalloc(n, m, _p(tdx), _p(tdy), _p(tadx), _p(out_na),
                                       stream) == 0, core.last_error()
    torch.cuda.synchronize()
    return out_ws.cpu().numpy(), out_na.cpu().numpy()


def test_noalloc_dots_equal_workspace_dots_twice_shrinking():
    L, core = _lib()
    rng = np.random.default_rng(5)
    for n, m in ((1_200_000, 700_000), (3000, 900)):  # second call smaller
        dx = rng.standard_normal(n)
        dy = rng.standard_normal(m) * 1e3
        adx = rng.standard_normal(m)
        ws, na = _both(L, core, dx, dy, adx)
        assert np.array_equal(ws, na), (n, m, ws, na)
        assert np.array_equal(na, np.array(oracle.dots(dx, dy, adx))), (n, m)


def test_noalloc_dots_rejects_bad_arguments():
    L, core = _lib()
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.igaos_ls_dots_f64_noalloc(-1, 3, None, None, None, _p(out), stream) != 0
    assert "negative size" in core.last_error()
    assert L.igaos_ls_dots_f64_noalloc(5, 3, None, None, None, _p(out), stream) != 0
    assert "null device array" in core.last_error()
