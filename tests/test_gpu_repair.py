"""The device dual repair (igaos_dual_repair_device, the engine's repair for
m > max_repair_m; no reference counterpart above the dense repair's reach,
src/pdhg/pdhg.cu:288-448).  It is pinned by what the mathematics fixes --
closed forms where every column is interior (the unique y with
c + A'y = 0), weak duality (the repaired bound never exceeds the known
optimum) and a certified gap on the points where the reference's dense
repair certifies -- and by the engine reaching near-optimal through it on
an LP where the oracle (and the engine's host path) certify by repair."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.sparse as sp

import igaos_amd
import oracle
from igaos_amd import _core, lpgen
from igaos_amd.lpgen import INF
from tests.conftest import netlib_path

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _repair(md, x, y_it, col_tol=1e-6):
    x, y_it = np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y_it, np.float64)
    y = np.zeros(md.m)
    out = np.zeros(6)
    rc = _core.lib().igaos_dual_repair_device(md._h, _dp(x), _dp(y_it), col_tol, _dp(y),
                                              _dp(out))
    assert rc >= 0, _core.last_error()
    return rc, y, out


def test_device_repair_closed_forms():
    # the two cases the host repair is pinned by, and a random one: every
    # column strictly inside its bounds, A of full row rank, c = -A'y*, so
    # y* is the only y with d = c + A'y = 0
    cases = [
        (lpgen.LP.from_triplets(1, 2, [0, 0], [0, 1], [1.0, 1.0], np.ones(2), np.zeros(2),
                                np.full(2, 10.0), np.array([2.0]), np.array([2.0])),
         np.array([1.0, 1.0]), np.array([0.5]), np.array([-1.0])),
        (lpgen.LP.from_triplets(1, 1, [0], [0], [2.0], np.array([3.0]), np.zeros(1),
                                np.full(1, INF), np.array([2.0]), np.array([INF])),
         np.array([1.0]), np.array([0.7]), np.array([-1.5])),
    ]
    rng = np.random.default_rng(3)
    m, n = 120, 300
    A = sp.random(m, n, density=0.05, random_state=rng, format="csr")
    A = (A + sp.hstack([sp.eye(m), sp.csr_matrix((m, n - m))])).tocsr()  # full row rank
    ys = rng.standard_normal(m)
    x0 = rng.uniform(1.0, 9.0, n)
    b = A @ x0
    coo = A.tocoo()
    lp = lpgen.LP.from_triplets(m, n, coo.row, coo.col, coo.data, -(A.T @ ys), np.zeros(n),
                                np.full(n, 10.0), b, b)
    cases.append((lp, x0, rng.standard_normal(m), ys))
    for lp, x, y_it, want in cases:
        md = igaos_amd.model_from_lp(lp)
        rc, y, out = _repair(md, x, y_it)
        assert rc == 1
        assert np.abs(y - want).max() <= 1e-8 * (1 + np.abs(want).max()), (y, want)
        assert out[5] == lp.n  # every column in J


@pytest.mark.parametrize("name", ["afiro", "sc50a", "kb2", "adlittle", "sc205"])
def test_device_repair_certifies_where_the_dense_repair_does(name, objectives):
    # at the oracle's 1e-4 point: the dense host repair (bit-identical to
    # the reference's restatement) certifies the gap; the device repair
    # must too, with a finite bound below the known optimum (weak duality)
    # and close to it
    md = igaos_amd.read_mps(netlib_path(name))
    lp = md.to_lp()
    r = oracle.solve(lp, tolerance=1e-4, time_limit=60)
    x = np.ascontiguousarray(r.x)
    A = lp.to_scipy()
    Ax = A @ x
    yh = np.zeros(lp.m)
    assert _core.lib().igaos_dual_repair(md._h, _dp(x), _dp(Ax), _dp(np.ascontiguousarray(r.y)),
                                         1e-6, _dp(yh)) == 1
    rc, y, out = _repair(md, x, r.y)
    assert rc == 1
    mt = oracle.evaluate(lp.rmin, lp.rmax, lp.cl, lp.cu, lp.c, x, y, Ax, lp.c + A.T @ y)
    mh = oracle.evaluate(lp.rmin, lp.rmax, lp.cl, lp.cu, lp.c, x, yh, Ax, lp.c + A.T @ yh)
    ref = objectives[f"netlib/{name}"]["highs_ref"]
    bound = mt.od + lp.obj_const
    assert math.isfinite(mt.gap) and abs(mt.gap - out[0]) <= 1e-9 * (1 + abs(out[0]))
    assert bound <= ref + 1e-7 * (1 + abs(ref))
    assert mt.gap <= max(1e-4, 2 * mh.gap), (mt.gap, mh.gap)


def test_engine_certifies_through_the_device_repair():
    # sc105 at 1e-4: the oracle certifies by its dense dual repair at the
    # check of k = 7,104.  The exact engine follows the oracle's trajectory
    # bit for bit up to that check; with the dense repair switched off
    # (max_repair_m = 0, as for every LP above 4,096 rows) the device repair
    # runs there instead and must certify at the same iteration
    lp = igaos_amd.read_mps(netlib_path("sc105")).to_lp()
    ref = oracle.solve(lp, time_limit=1e7, tolerance=1e-4, max_iterations=60000)
    assert ref.repaired and ref.status == "near-optimal"
    s = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=1e-4, max_iterations=60000,
                           exact=True, max_repair_m=0)
    assert s.status == "near-optimal" and s.info["repaired"] == 1, s.message
    assert s.iterations == ref.iterations
    assert s.rel_gap <= 1e-4 and s.pinf <= 1e-4
    assert np.array_equal(s.x, ref.x)  # the primal point is not touched
    assert abs(s.objective - ref.objective) <= 1e-12 * (1 + abs(ref.objective))
    # the default kernels and FP32 + refinement certify too
    for prec in (64, 32):
        s = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=1e-4, max_iterations=60000,
                               precision=prec, max_repair_m=0)
        assert s.status == "near-optimal", (prec, s.message)
