"""Exact FP64 mode against the oracle, bit for bit, over long runs.

The exact mode (igaos_engine_config.exact = 1, igaos_amd/csrc/exact.hip)
sums every sparse row in storage order, forms the line-search dots in the
reference's two-stage layout (src/pdhg/pdhg.cu:63-125), adds norms and the
KKT sums in index order, takes k^-0.3 / k^-0.6 from the host's pow and
never fuses a multiply-add: the arithmetic the oracle (oracle/pdhg_oracle.c,
a sequential FP64 replay of pdhg.cu built with -ffp-contract=off) performs.
The engine's results must then equal the oracle's exactly -- not to a
tolerance -- however long the run, including the paths that only fire at
check time: restarts and the primal-weight update (pdhg.cu:989-1075), the
incumbent and its rollback (933-988), the dual repair (854-931) and the
exit's incumbent fallback (1090-1111).  The cases were picked (by running
the oracle on the CPU) so that each of those paths fires:

  sc105,   tol 1e-4: dual repair at k >= 5000 certifies (7,104 iterations)
  beaconfd, tol 1e-6: one rollback, then near-optimal (16,448)
  sc205,   tol 1e-4: one rollback, near-optimal after 34,496
  kb2,     tol 1e-4: iteration limit, the exit returns the incumbent (60,000)
  vtp.base, tol 1e-8: two rollbacks, iteration limit (60,000)

and every one restarts 20+ times, so omega is updated 20+ times.
"""
import numpy as np
import pytest

import igaos_amd
import oracle
from igaos_amd import lpgen
from tests.conftest import netlib_path

pytestmark = pytest.mark.gpu


def _netlib(name):
    return igaos_amd.read_mps(netlib_path(name)).to_lp(name)


def _engine(lp, tol, maxit, runner=-1):
    # runner -1 / 1: the one-workgroup runner where the LP fits it (every
    # case here); 0: six kernels per attempt.  Both sum alike.
    return igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=tol, max_iterations=maxit,
                              exact=True, small_runner=runner)


def _pair(lp, tol, maxit, runner=-1):
    ref = oracle.solve(lp, time_limit=1e7, tolerance=tol, max_iterations=maxit)
    return ref, _engine(lp, tol, maxit, runner)


def _assert_identical(ref, got):
    i = got.info
    assert got.status == ref.status, (got.status, ref.status, got.message)
    assert got.iterations == ref.iterations
    assert i["attempts"] == ref.attempts
    assert i["restarts"] == ref.restarts
    assert i["rollbacks"] == ref.rollbacks
    assert bool(i["repaired"]) == ref.repaired
    assert bool(i["used_incumbent"]) == ref.used_incumbent
    assert bool(i["last_is_avg"]) == ref.last_is_avg
    for k in ("lam2", "eta0", "omega0", "eta", "omega"):
        assert i[k] == getattr(ref, k), (k, i[k], getattr(ref, k))
    for k in ("objective", "pinf", "dinf", "rel_gap"):
        a, b = getattr(got, k), getattr(ref, k)
        assert a == b or (np.isnan(a) and np.isnan(b)), (k, a, b)
    assert np.array_equal(got.x, ref.x), np.abs(got.x - ref.x).max()
    assert np.array_equal(got.y, ref.y), np.abs(got.y - ref.y).max()


@pytest.mark.parametrize("runner", [0, 1])
@pytest.mark.parametrize("iters", [1, 2, 63, 64, 65, 128, 1000])
def test_exact_trajectory_random_lp(iters, runner):
    # tolerance 0: the checks run but never stop the solve, so the run
    # ends at `iters` with restarts, weight updates and the incumbent in it
    lp = lpgen.gen_random_eq(seed=3, m=120, n=300)
    ref, got = _pair(lp, 0.0, iters, runner)
    _assert_identical(ref, got)


@pytest.mark.parametrize("name", ["afiro", "sc50a", "adlittle", "share2b"])
def test_exact_kkt_values_at_fixed_iterations(name):
    # the device KKT evaluation (pdhg.cu:173-254) equals the host's exactly
    lp = _netlib(name)
    for iters in (64, 128):
        ref, got = _pair(lp, 0.0, iters)
        _assert_identical(ref, got)


@pytest.mark.parametrize("name,tol,maxit,event", [
    ("sc105", 1e-4, 60000, "repaired"),
    ("beaconfd", 1e-6, 60000, "rollbacks"),
    ("sc205", 1e-4, 60000, "rollbacks"),
    ("kb2", 1e-4, 60000, "used_incumbent"),
    ("vtp.base", 1e-8, 60000, "rollbacks"),
])
def test_exact_long_runs_with_check_time_paths(name, tol, maxit, event):
    lp = _netlib(name)
    ref, got = _pair(lp, tol, maxit)
    assert getattr(ref, event), f"the oracle no longer takes the {event} path on {name}"
    assert ref.restarts >= 20 and ref.omega != ref.omega0
    _assert_identical(ref, got)
    _assert_identical(ref, _engine(lp, tol, maxit, 0))


@pytest.mark.parametrize("name", ["afiro", "sc105", "share2b"])
def test_exact_pock_chambolle_scaling(name):
    # the opt-in Ruiz + Pock-Chambolle preconditioning (device factors from
    # sequential row / column sums) against the oracle's, bit for bit
    lp = _netlib(name)
    ref = oracle.solve(lp, time_limit=1e7, tolerance=1e-6, max_iterations=20000, scaling=1)
    got = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=1e-6, max_iterations=20000,
                             exact=True, scaling="ruiz+pock-chambolle")
    _assert_identical(ref, got)


@pytest.mark.parametrize("name", ["afiro", "sc50a", "adlittle", "sc205"])
def test_pock_chambolle_solves_certify_to_highs(name, objectives):
    # the default (fast) kernels with the opt-in scaling: certified at 1e-6
    # and within 1e-6 of the reference's recorded HiGHS optimum
    ref = objectives[f"netlib/{name}"]["highs_ref"]
    for prec in (64, 32):
        s = igaos_amd.solve_lp(_netlib(name), time_limit=120, tolerance=1e-6, precision=prec,
                               scaling="ruiz+pock-chambolle")
        assert s.status == "near-optimal", (prec, s.message)
        assert abs(s.objective - ref) <= 2e-6 * (1 + abs(ref)), (prec, s.objective, ref)


def test_exact_mode_refuses_time_limited_runs():
    # a wall-clock budget makes MAXIT 2e7 (pdhg.cu:669-671), beyond the
    # exponent table: the engine says so instead of running
    lp = lpgen.gen_random_eq(seed=3, m=20, n=50)
    s = igaos_amd.solve_lp(lp, exact=True, time_limit=10.0)
    assert s.status == "error" and "exact mode runs at most" in s.message


def _dots_gpu(dx, dy, adx):
    import ctypes as C
    import torch
    from igaos_amd import _core
    L = _core.lib()
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64)).cuda()  # noqa
    tdx, tdy, tadx = dev(dx), dev(dy), dev(adx)
    work = torch.full((int(L.igaos_ls_dots_work(len(dx), len(dy))),), float("nan"),
                      dtype=torch.float64, device="cuda")
    out = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda")
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.igaos_ls_dots_f64(len(dx), len(dy), p(tdx), p(tdy), p(tadx), p(work), p(out),
                             stream)
    assert rc == 0, _core.last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("n,m", [(1, 1), (300, 120), (100, 5000), (70000, 257),
                                 (1_300_000, 400_000), (0, 10)])
def test_ls_dots_entry_point_equals_reference_layout(n, m):
    # igaos_ls_dots_f64 (kern_dots + kern_dotfin, pdhg.cu:63-125) against the
    # oracle's restatement of the same two-stage order, bit for bit; the
    # 1.3M case runs at the 4096-workgroup cap with grid-stride loops
    rng = np.random.default_rng(n + m)
    dx = rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))
    dy = rng.standard_normal(m)
    adx = rng.standard_normal(m) * 1e-3
    got = _dots_gpu(dx, dy, adx)
    ref = np.array(oracle.dots(dx, dy, adx))
    assert np.array_equal(got, ref), (got, ref)
