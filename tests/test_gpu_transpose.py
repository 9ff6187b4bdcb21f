"""Device CSR -> CSC (igaos_csr_to_csc_f64, igaos_model_from_csr) against
the host's stable transpose, which is what the reference's counting_sort
(src/io/model.hpp:39-67) produces for triplets given row by row: each
column's entries in row order.  Integer and copy work: bit-exact."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import igaos_amd
from igaos_amd import _core, lpgen

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _host_transpose(m, n, ap, ai, ax):
    order = np.argsort(ai, kind="stable")
    rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(ap))
    cp = np.zeros(n + 1, np.int64)
    cp[1:] = np.cumsum(np.bincount(ai, minlength=n))
    return cp.astype(np.int32), rows[order], ax[order]


def _dev(a, dt):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype=dt, device="cuda")


def _device_transpose(m, n, ap, ai, ax):
    nnz = len(ai)
    d_ap, d_ai, d_ax = _dev(ap, torch.int32), _dev(ai, torch.int32), _dev(ax, torch.float64)
    d_cp = torch.full((n + 1,), -1, dtype=torch.int32, device="cuda")
    d_ci = torch.full((max(nnz, 1),), -1, dtype=torch.int32, device="cuda")
    d_acx = torch.full((max(nnz, 1),), float("nan"), dtype=torch.float64, device="cuda")
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = _core.lib().igaos_csr_to_csc_f64(m, n, nnz, p(d_ap), p(d_ai), p(d_ax), p(d_cp),
                                          p(d_ci), p(d_acx), stream)
    assert rc == 0, _core.last_error()
    return d_cp.cpu().numpy(), d_ci.cpu().numpy()[:nnz], d_acx.cpu().numpy()[:nnz]


@pytest.mark.parametrize("m,n,density", [(1, 1, 1.0), (50, 30, 0.2), (3000, 7000, 0.002),
                                         (200_000, 150_000, 2e-5), (40, 60, 0.0)])
def test_device_transpose_equals_stable_host_transpose(m, n, density):
    rng = np.random.default_rng(m + n)
    A = sp.random(m, n, density=density, random_state=rng, format="csr")
    A.data = rng.standard_normal(A.nnz)
    ap, ai, ax = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    if m == 3000:  # a few empty rows and one very long row
        A = A.tolil()
        A[0, :] = 0
        A[7, :] = rng.standard_normal(n)
        A = A.tocsr()
        ap, ai, ax = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    got = _device_transpose(m, n, ap, ai, ax)
    ref = _host_transpose(m, n, ap, ai, ax)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


def test_model_from_csr_equals_the_triplet_model_and_solves_alike():
    lp = lpgen.gen_random_eq(seed=6, m=200, n=500)
    a = igaos_amd.model_from_csr(lp.m, lp.n, lp.c, lp.cl, lp.cu, lp.rmin, lp.rmax,
                                 lp.ap, lp.ai, lp.ax).arrays()
    b = igaos_amd.model_from_lp(lp).arrays()
    for k in b:
        assert np.array_equal(a[k], b[k]), k
    m1 = igaos_amd.model_from_csr(lp.m, lp.n, lp.c, lp.cl, lp.cu, lp.rmin, lp.rmax,
                                  lp.ap, lp.ai, lp.ax)
    s1 = igaos_amd.solve_lp(m1, time_limit=1e7, tolerance=0.0, max_iterations=200)
    s2 = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=200)
    assert np.array_equal(s1.x, s2.x) and np.array_equal(s1.y, s2.y)
