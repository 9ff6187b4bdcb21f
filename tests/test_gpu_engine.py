"""GPU parity of the MI355X PDHG engine against the oracle and fixtures.

Every case calls the HIP library through its C-ABI (libigaos_pdhg.so).
Bars: integer/index work and the Ruiz factors are bit-exact; FP64 SpMV
agrees with the oracle's sequential row sums to 1e-14 relative; the plain
FP64 engine matches the oracle's trajectory to 1e-9 relative over one
check period (64 iterations; beyond that the line search amplifies
last-bit rounding differences by about 10x every 8 iterations, see
tests/test_oracle.py and the parity-mode tests); every solve to
tolerance 1e-6 certifies (near-optimal) and lands within 1e-6 relative
of the reference simplex / HiGHS objective.
"""
import ctypes as C
import json
import math
import subprocess

import numpy as np
import pytest

import igaos_amd
import oracle
from igaos_amd import _core, lpgen
from tests.conftest import ROOT, netlib_path

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype=dtype, device="cuda")


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _rowblocks(ptr):
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    L = _core.lib()
    k = L.igaos_rowblocks_build(len(ptr) - 1, ptr.ctypes.data_as(C.POINTER(C.c_int)), None, 0)
    out = np.zeros(k, np.int32)
    L.igaos_rowblocks_build(len(ptr) - 1, ptr.ctypes.data_as(C.POINTER(C.c_int)),
                            out.ctypes.data_as(C.POINTER(C.c_int)), k)
    return out


def _gpu_spmv(ptr, idx, val, x, f32=False):
    rb = _rowblocks(ptr)
    fdt = torch.float32 if f32 else torch.float64
    d_rb, d_ptr, d_idx = (_dev(rb, torch.int32), _dev(ptr, torch.int32),
                          _dev(idx, torch.int32))
    d_val, d_x = _dev(val, fdt), _dev(x, fdt)
    d_y = torch.full((len(ptr) - 1,), float("nan"), dtype=fdt, device="cuda")
    fn = _core.lib().igaos_spmv_csr_f32 if f32 else _core.lib().igaos_spmv_csr_f64
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(len(rb) // 2 - 1, _ptr(d_rb), _ptr(d_ptr), _ptr(d_idx), _ptr(d_val),
            _ptr(d_x), _ptr(d_y), stream)
    assert rc == 0, _core.last_error()
    torch.cuda.synchronize()
    return d_y.cpu().numpy().astype(np.float64)


def _long_row_matrix(rng):
    # rows of 0, 1, 5, 300, 2048, 2049 and 9000 nonzeros (vector mode > 2048)
    lens = [0, 1, 5, 300, 2048, 2049, 9000, 3, 0, 64]
    n = 12000
    ptr = np.zeros(len(lens) + 1, np.int64)
    ptr[1:] = np.cumsum(lens)
    idx = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens]).astype(np.int32)
    return ptr.astype(np.int32), idx, rng.standard_normal(ptr[-1]), n


@pytest.mark.parametrize("case", ["random_eq", "zipf", "long_rows"])
def test_spmv_f64_matches_oracle(case):
    rng = np.random.default_rng(11)
    if case == "long_rows":
        ptr, idx, val, n = _long_row_matrix(rng)
        mats = [(ptr, idx, val, n)]
    else:
        lp = (lpgen.gen_random_eq(seed=3) if case == "random_eq"
              else lpgen.gen_zipf_badly_scaled(seed=3, m=2000, n=5000, kmax=500))
        mats = [(lp.ap, lp.ai, lp.ax, lp.n), (lp.cp, lp.ci, lp.acx, lp.m)]
    for ptr, idx, val, ncols in mats:
        x = rng.standard_normal(ncols)
        ref = oracle.spmv(ptr, idx, val, x)
        got = _gpu_spmv(ptr, idx, val, x)
        scale = np.abs(val).max() * np.abs(x).max() * np.maximum(np.diff(ptr), 1)
        assert np.all(np.abs(got - ref) <= 1e-14 * scale), case


def test_spmv_f32_matches_oracle():
    rng = np.random.default_rng(12)
    lp = lpgen.gen_random_eq(seed=4)
    x = rng.standard_normal(lp.n).astype(np.float32).astype(np.float64)
    val = lp.ax.astype(np.float32).astype(np.float64)
    ref = oracle.spmv(lp.ap, lp.ai, val, x)
    got = _gpu_spmv(lp.ap, lp.ai, val, x, f32=True)
    bound = 2e-7 * np.abs(val).max() * np.abs(x).max() * np.diff(lp.ap)
    assert np.all(np.abs(got - ref) <= bound)


def test_ruiz_bit_exact():
    lp = lpgen.gen_zipf_badly_scaled(seed=5, m=500, n=1500)
    dr_ref, dc_ref = oracle.ruiz(lp)
    t = {k: _dev(getattr(lp, k), torch.int32 if k in ("ap", "ai", "cp", "ci") else torch.float64)
         for k in ("ap", "ai", "ax", "cp", "ci", "acx")}
    dr = torch.zeros(lp.m, dtype=torch.float64, device="cuda")
    dc = torch.zeros(lp.n, dtype=torch.float64, device="cuda")
    rc = _core.lib().igaos_ruiz_f64(lp.m, lp.n, _ptr(t["ap"]), _ptr(t["ai"]),
                                    _ptr(t["ax"]), _ptr(t["cp"]), _ptr(t["ci"]),
                                    _ptr(t["acx"]), _ptr(dr), _ptr(dc),
                                    C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(dr.cpu().numpy(), dr_ref)
    assert np.array_equal(dc.cpu().numpy(), dc_ref)


def _rel(a, b):
    return np.abs(a - b).max() / max(1.0, np.abs(b).max())


# runner 0: three launches per attempt (row-block products); runner 1: the
# default mode's one-workgroup runner (the LP's state in registers and LDS,
# rows summed in storage order); the LP fits both
@pytest.mark.parametrize("runner", [0, 1])
@pytest.mark.parametrize("K", [16, 64])
def test_fixed_iteration_trajectory_parity(K, runner):
    lp = lpgen.gen_random_eq(seed=0)
    r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=K)
    s = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=K,
                           small_runner=runner)
    assert s.iterations == r.iterations == K
    assert s.info["attempts"] == r.attempts
    assert s.info["restarts"] == r.restarts
    # a run that stops before the first check keeps the zero-initialised
    # metrics and reports near-optimal (pdhg.cu:734,1096); both replicate it
    assert s.status == r.status == ("near-optimal" if K < 64 else "iteration-limit")
    assert _rel(s.x, r.x) <= 1e-9 and _rel(s.y, r.y) <= 1e-9
    assert math.isclose(s.info["lam2"], r.lam2, rel_tol=1e-12)
    assert math.isclose(s.info["omega0"], r.omega0, rel_tol=1e-14)


@pytest.mark.parametrize("runner", [0, 1])
def test_trajectory_stays_close_through_two_checks(runner):
    # past one check period the adaptive line search amplifies rounding
    # differences about 10x every 8 iterations (measured on the oracle
    # alone: a 1e-16 perturbation grows to 1.7e-9 by k = 128; engine vs
    # oracle here measured 2.6e-6), so the bar over two periods is 1e-4:
    # the same restart decisions, and no divergence of the trajectory
    lp = lpgen.gen_random_eq(seed=0)
    K = 128
    r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=K)
    s = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=K,
                           small_runner=runner)
    assert s.iterations == r.iterations == K
    assert s.info["restarts"] == r.restarts
    assert _rel(s.x, r.x) <= 1e-4 and _rel(s.y, r.y) <= 1e-4


@pytest.mark.parametrize("m,n", [(300, 900), (600, 1800), (1000, 3000), (1500, 900),
                                 (2500, 1500), (2500, 2100)])
def test_small_runner_register_layouts(m, n):
    # the runner keeps 1, 2 or 3 columns and, separately, 1, 2 or 3 rows per
    # thread: 64 iterations of each layout follow the oracle like the three-kernel
    # path; in FP32 the two paths sum rows in different orders, and the
    # line search amplifies that rounding (about 10x every 8 iterations),
    # so they are compared over 16 iterations
    lp = lpgen.gen_random_eq(seed=7, m=m, n=n)
    K = 64
    r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=K)
    for runner in (0, 1):
        s = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=K,
                               small_runner=runner)
        assert s.iterations == r.iterations == K
        assert s.info["attempts"] == r.attempts and s.info["restarts"] == r.restarts
        assert _rel(s.x, r.x) <= 1e-9 and _rel(s.y, r.y) <= 1e-9
    a, b = (igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=16,
                               precision=32, small_runner=runner) for runner in (0, 1))
    assert a.iterations == b.iterations == 16
    assert _rel(b.x, a.x) <= 1e-3 and _rel(b.y, a.y) <= 1e-3


def test_small_runner_on_ragged_lp(objectives):
    # empty rows and columns, one long row and column, every bound kind: 64
    # runner iterations follow the oracle like the three-kernel path does,
    # and the runner's solve to 1e-6 certifies at HiGHS's objective
    lp = lpgen.gen_ragged(seed=5)
    r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=64)
    for runner in (0, 1):
        s = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=64,
                               small_runner=runner)
        assert s.iterations == r.iterations == 64
        assert s.info["attempts"] == r.attempts and s.info["restarts"] == r.restarts
        assert _rel(s.x, r.x) <= 1e-9 and _rel(s.y, r.y) <= 1e-9
    a, b = (igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=16,
                               precision=32, small_runner=runner) for runner in (0, 1))
    assert _rel(b.x, a.x) <= 1e-3 and _rel(b.y, a.y) <= 1e-3
    ref = objectives["gen_ragged(seed=5)"]["scipy_highs"]
    for prec in (64, 32):
        s = igaos_amd.solve_lp(lp, tolerance=1e-6, time_limit=60, precision=prec,
                               small_runner=1)
        assert s.status == "near-optimal", (prec, s.status, s.message)
        assert abs(s.objective - ref) <= 2e-6 * (1 + abs(ref))


@pytest.mark.parametrize("precision", [64, 32])
def test_runner_check_graph_matches_direct_launches(precision):
    # the runner's check periods replayed as one captured graph (runner,
    # flush, KKT metrics, copies to the host) against the same work launched
    # directly (graph_slots = 0): the same run, bit for bit, over 40 checks
    # with their restarts, and the same solve to 1e-6
    lp = lpgen.gen_random_eq(seed=0)
    a, b = (igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=2560,
                               precision=precision, small_runner=1, graph_slots=g)
            for g in (16, 0))
    assert a.iterations == b.iterations == 2560
    assert a.info["restarts"] == b.info["restarts"] and a.info["checks"] == b.info["checks"]
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    a, b = (igaos_amd.solve_lp(lp, time_limit=60, tolerance=1e-6, precision=precision,
                               small_runner=1, graph_slots=g) for g in (16, 0))
    assert a.status == b.status == "near-optimal"
    assert a.iterations == b.iterations and a.objective == b.objective


def _check_solution(s, ref_obj, tol=1e-6):
    assert s.status == "near-optimal", (s.status, s.message)
    assert s.pinf <= tol and s.rel_gap <= tol
    assert abs(s.objective - ref_obj) <= tol * (1 + abs(ref_obj)), (s.objective, ref_obj)


@pytest.mark.parametrize("precision,runner", [(64, 0), (64, 1), (32, -1)])
def test_random_eq_solves_to_fixture(precision, runner, objectives):
    lp = lpgen.gen_random_eq(seed=0)
    ent = objectives["gen_random_eq(seed=0)"]
    s = igaos_amd.solve_lp(lp, tolerance=1e-6, time_limit=60, precision=precision,
                           small_runner=runner)
    _check_solution(s, ent["ref_simplex"])
    _check_solution(s, ent["scipy_highs"])
    # the reported pair satisfies the KKT bar in the oracle's own evaluation
    A = lp.to_scipy()
    mt = oracle.evaluate(lp.rmin, lp.rmax, lp.cl, lp.cu, lp.c, s.x, s.y,
                         A @ s.x, lp.c + A.T @ s.y)
    assert mt.pinf <= 1e-6 and mt.gap <= 1e-6


@pytest.mark.parametrize("name", ["afiro", "sc50a", "adlittle", "sc205"])
def test_netlib_solves_match_reference_highs(name, objectives):
    s = igaos_amd.solve(netlib_path(name), engine="pdhg", tolerance=1e-6,
                        time_limit=120)
    ref = objectives[f"netlib/{name}"]["highs_ref"]
    assert s.status == "near-optimal", s.message
    assert abs(s.objective - ref) <= 1e-6 * (1 + abs(ref)) + 5e-10 * abs(ref)
    # row_activity is A x summed in CSR order, as the reference simplex
    # fills it (scipy's CSR product sums in the same order)
    lp = igaos_amd.read_mps(netlib_path(name)).to_lp()
    assert np.array_equal(s.row_activity, lp.to_scipy() @ s.x)


ENGINE_SMOKE = [
    ("NAME T1\nROWS\n N obj\n G r1\nCOLUMNS\n x obj 1.0 r1 1.0\n"
     "RHS\n RHS1 r1 -5.0\nBOUNDS\n UP BND x -3.0\nENDATA\n", -5.0),
    ("NAME T2\nROWS\n N obj\n G r1\nCOLUMNS\n x obj 1.0 r1 1.0\n"
     "RHS\n RHS1 r1 -2.0\nBOUNDS\n MI BND x\nENDATA\n", -2.0),
    ("NAME T3\nROWS\n N obj\n G r1\nCOLUMNS\n x obj 1.0 r1 1.0\n"
     "RHS\n RHS1 obj -7.113 r1 2.0\nRANGES\n RNG r1 3.0\nENDATA\n", 2.0 + 7.113),
    ("NAME T4\nROWS\n N obj\n E r1\n G r2\nCOLUMNS\n x obj 1.0 r1 1.0\n"
     " x r2 0.0\n y obj 2.0 r1 1.0\n y r2 1.0\nRHS\n RHS1 r1 4.0 r2 1.0\n"
     "BOUNDS\n FR BND y\nENDATA\n", 5.0),
]


@pytest.mark.parametrize("idx", range(len(ENGINE_SMOKE)))
def test_reference_engine_smoke_lps(idx, tmp_path):
    text, obj = ENGINE_SMOKE[idx]
    p = tmp_path / "t.mps"
    p.write_text(text)
    s = igaos_amd.solve(str(p), engine="pdhg", tolerance=1e-8, time_limit=30)
    assert s.status == "near-optimal" and abs(s.objective - obj) < 1e-6


def test_iteration_cap_and_time_limit():
    lp = lpgen.gen_random_eq(seed=2, m=60, n=150)
    s = igaos_amd.solve_lp(lp, tolerance=0.0, time_limit=1e7, max_iterations=100)
    r = oracle.solve(lp, tolerance=0.0, time_limit=1e7, max_iterations=100)
    assert s.iterations == r.iterations == 100 and s.status == "iteration-limit"
    s = igaos_amd.solve_lp(lp, tolerance=0.0, time_limit=1e-9)
    assert s.status == "time-limit" and s.iterations == 64
    assert s.message == "wall-clock budget exhausted"


def test_run_to_run_bit_identical():
    # columns of at most m/4 rows: distinct rows are drawn quickly
    lp = lpgen.gen_zipf_badly_scaled(seed=1, m=2000, n=5000, kmax=500)
    kw = dict(tolerance=0.0, time_limit=1e7, max_iterations=3000)
    a = igaos_amd.solve_lp(lp, **kw)
    b = igaos_amd.solve_lp(lp, **kw)
    assert a.iterations == b.iterations == 3000
    assert a.info["restarts"] == b.info["restarts"] > 0
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


@pytest.mark.parametrize("precision", [64, 32])
def test_badly_scaled_and_multiperiod(precision, objectives):
    cases = [("gen_zipf_badly_scaled(seed=0,m=400,n=1000)",
              lpgen.gen_zipf_badly_scaled(seed=0, m=400, n=1000)),
             ("gen_multiperiod(seed=0,periods=2,rows_per=200,cols_per=320,inv=40)",
              lpgen.gen_multiperiod(seed=0, periods=2, rows_per=200,
                                    cols_per=320, inv=40))]
    for key, lp in cases:
        s = igaos_amd.solve_lp(lp, tolerance=1e-6, time_limit=120,
                               precision=precision)
        _check_solution(s, objectives[key]["scipy_highs"])


@pytest.mark.parametrize("precision", [64, 32])
def test_eight_period_blending_lp_against_highs(precision, objectives):
    # the block-structured multi-period family at 8 periods of 1,000 x
    # 1,600 (200 inventory columns each, linking consecutive periods):
    # 8,000 rows, 12,800 columns, 79,800 nonzeros; HiGHS's objective from
    # its interior-point method with crossover (tests/golden/make_fixtures.py)
    key = "gen_multiperiod(seed=0,periods=8,rows_per=1000,cols_per=1600,inv=200)"
    lp = lpgen.gen_multiperiod(seed=0, periods=8, rows_per=1000, cols_per=1600, inv=200)
    ent = objectives[key]
    assert (lp.m, lp.n, lp.nnz) == (ent["m"], ent["n"], ent["nnz"])
    s = igaos_amd.solve_lp(lp, tolerance=1e-6, time_limit=120, precision=precision)
    assert s.status == "near-optimal", (s.status, s.message)
    assert s.pinf <= 1e-6 and s.dinf <= 1e-6 and s.rel_gap <= 1e-6
    # the KKT bar bounds the gap, not the distance to the optimum: 2e-6
    ref = ent["scipy_highs"]
    assert abs(s.objective - ref) <= 2e-6 * (1 + abs(ref)), (s.objective, ref)


def test_graph_and_direct_launches_agree():
    lp = lpgen.gen_random_eq(seed=1, m=200, n=600)
    a = igaos_amd.solve_lp(lp, tolerance=1e-6, time_limit=60, graph_slots=16)
    b = igaos_amd.solve_lp(lp, tolerance=1e-6, time_limit=60, graph_slots=0)
    assert a.iterations == b.iterations and np.array_equal(a.x, b.x)


def test_cli_pdhg_json():
    out = subprocess.run([f"{ROOT}/igaos_amd/igaos", "solve", netlib_path("afiro"),
                          "--engine", "pdhg", "--tol", "1e-6"],
                         capture_output=True, text=True, check=True).stdout
    js = json.loads(out)
    assert js["status"] == "near-optimal" and js["engine"] == "pdhg"
    assert abs(js["objective"] + 464.7531429) < 1e-3
    assert len(js["x"]) == 32 and len(js["y"]) == 27


def test_column_locality_order_keeps_parity():
    # n >= 65536, m > 4096 and Zipf column lengths: the engine renumbers
    # the columns by length internally; the trajectory and the returned x
    # (in the model's own column order) still follow the oracle
    lp = lpgen.gen_zipf_badly_scaled(seed=3, m=40_000, n=100_000)
    lens = np.diff(lp.cp)
    top = np.sort(lens)[::-1][: lp.n // 50].sum()
    assert top * 4 >= lp.nnz  # the engine's criterion holds for this LP
    K = 64
    r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=K)
    s = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=K)
    assert s.iterations == r.iterations == K
    assert s.info["attempts"] == r.attempts
    assert _rel(s.x, r.x) <= 1e-9 and _rel(s.y, r.y) <= 1e-9
    s32 = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=K,
                             precision=32)
    # the FP32 engine's x comes back in the model's order too: the
    # column-specific upper bounds (U[1, 100]) hold
    assert s32.iterations == K
    assert np.all(s32.x >= lp.cl - 1e-6) and np.all(s32.x <= lp.cu * (1 + 1e-6) + 1e-6)
    # column slices of the renumbered copy: A's multi-row blocks are stored
    # sorted by column (gather order) and each product lands at its
    # row-major slot, so rows sum as in plain slices; with nontemporal
    # matrix loads
    for slices in (1, 8):
        g = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=K,
                               column_slices=slices, stream_loads=1)
        assert g.iterations == K and g.info["attempts"] == r.attempts
        assert _rel(g.x, r.x) <= 1e-9 and _rel(g.y, r.y) <= 1e-9
        if slices == 1:  # the fused kernels' blocks and row sums
            dx, dy = _rel(g.x, s.x), _rel(g.y, s.y)
            assert dx <= 1e-12 and dy <= 1e-12, (dx, dy)


@pytest.mark.parametrize("precision", [64, 32])
def test_nontemporal_stream_loads_change_nothing(precision):
    # the hint changes how the matrix streams are cached, not one value
    lp = lpgen.gen_zipf_badly_scaled(seed=4, m=6_000, n=20_000, kmax=600)
    for slices in (0, 3):
        kw = dict(time_limit=1e7, tolerance=0.0, max_iterations=192, precision=precision,
                  column_slices=slices)
        a = igaos_amd.solve_lp(lp, stream_loads=0, **kw)
        b = igaos_amd.solve_lp(lp, stream_loads=1, **kw)
        assert a.iterations == b.iterations == 192
        assert a.info["attempts"] == b.info["attempts"]
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


@pytest.mark.parametrize("slices", [1, 3, 8])
def test_column_sliced_products_keep_parity(slices):
    # the sliced A / A' products (per-slice partial row sums added in slice
    # order) follow the oracle's trajectory like the fused row-block ones;
    # one slice holds each row's nonzeros in the original blocks and order,
    # so it sums every row as the fused kernels do (the split primal / dual
    # step may contract its multiply-adds differently: ulp-level drift)
    lp = lpgen.gen_zipf_badly_scaled(seed=4, m=6_000, n=20_000, kmax=600)
    K = 64
    r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=K)
    base = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=K,
                              column_slices=0)
    s = igaos_amd.solve_lp(lp, time_limit=1e7, tolerance=0.0, max_iterations=K,
                           column_slices=slices)
    assert s.iterations == r.iterations == K
    assert s.info["attempts"] == r.attempts
    assert _rel(s.x, r.x) <= 1e-9 and _rel(s.y, r.y) <= 1e-9
    if slices == 1:
        dx, dy = _rel(s.x, base.x), _rel(s.y, base.y)
        assert dx <= 1e-12 and dy <= 1e-12, (dx, dy)
    eq = lpgen.gen_random_eq(seed=5, m=3000, n=8000)
    for prec in (64, 32):
        got = igaos_amd.solve_lp(eq, time_limit=120, tolerance=1e-6, precision=prec,
                                 column_slices=slices)
        ref = igaos_amd.solve_lp(eq, time_limit=120, tolerance=1e-6, precision=prec,
                                 column_slices=0)
        assert got.status == ref.status == "near-optimal", (prec, got.message)
        assert abs(got.objective - ref.objective) <= 1e-6 * (1 + abs(ref.objective))


@pytest.mark.parametrize("slices", [3, 8, 16])
def test_xcd_pinned_slices_sum_like_per_slice_launches(slices):
    # one launch with every XCD on its own slices (igaos_engine_config.
    # xcd_pinned) runs the same blocks with the same partial sums as one
    # launch per slice: the runs are identical bit for bit
    lp = lpgen.gen_zipf_badly_scaled(seed=4, m=6_000, n=20_000, kmax=600)
    kw = dict(time_limit=1e7, tolerance=0.0, max_iterations=128, column_slices=slices)
    for prec in (64, 32):
        a = igaos_amd.solve_lp(lp, precision=prec, xcd_pinned=False, **kw)
        b = igaos_amd.solve_lp(lp, precision=prec, xcd_pinned=True, **kw)
        assert a.info["attempts"] == b.info["attempts"]
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y), prec


def test_reference_recorded_pdhg_battery():
    # the reference's own PDHG runs at its default tolerance 1e-4 (recorded,
    # not re-run): same status, and an objective within the tolerance band
    # of the recorded one (trajectories differ below the tolerance)
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "reference_pdhg_battery.json")) as f:
        bat = json.load(f)
    for name, rec in bat["instances"].items():
        s = igaos_amd.solve(netlib_path(name), tolerance=bat["tolerance"],
                            time_limit=120, engine="pdhg")
        assert s.status == rec["status"], (name, s.status)
        ref = rec["objective"]
        assert abs(s.objective - ref) <= 3e-4 * (1 + abs(ref)), (name, s.objective, ref)


@pytest.mark.parametrize("key", ["gen_random_eq(seed=1,m=200,n=600)",
                                 "gen_batch_instance(seed=0,index=1)",
                                 "gen_batch_instance(seed=0,index=5)",
                                 "netlib/afiro", "netlib/sc50a"])
def test_single_lp_polish_returns_the_vertex(key, objectives):
    # active-set polishing of the engine's 1e-6 solution (igaos_pdhg_polish).
    # When the polished pair is kept it is primal and dual feasible to 1e-9
    # and its objective is the HiGHS optimum to 1e-9; an unkept pair leaves
    # the input unchanged.  The active set read off complementarity is square
    # at a nondegenerate vertex (the random LPs: kept); degenerate Netlib
    # vertices may be refused, as afiro is.
    if key.startswith("netlib/"):
        lp = igaos_amd.read_mps(netlib_path(key[7:])).to_lp()
        ref = objectives[key]["highs_ref"]
    else:
        import re
        kw = dict((k, int(v)) for k, v in re.findall(r"(\w+)=(\d+)", key))
        lp = getattr(lpgen, key.split("(")[0])(**kw)
        ref = objectives[key]["scipy_highs"]
    md = igaos_amd.model_from_lp(lp)
    for prec in (64, 32):
        s = igaos_amd.solve_lp(md, time_limit=120, tolerance=1e-6, precision=prec)
        assert s.status == "near-optimal"
        x, y, obj, info = igaos_amd.api.polish(md, s.x, s.y)
        if not key.startswith("netlib/"):
            assert info["accepted"] == 1, info
        if not info["accepted"]:
            assert np.array_equal(x, s.x) and np.array_equal(y, s.y)
            continue
        assert abs(obj - ref) <= 1e-9 * (1 + abs(ref)) + 5e-10 * abs(ref), (obj, ref)
        A = lp.to_scipy()
        mt = oracle.evaluate(lp.rmin, lp.rmax, lp.cl, lp.cu, lp.c, x, y, A @ x,
                             lp.c + A.T @ y)
        assert mt.pinf <= 1e-9 and mt.dinf <= 1e-9 and mt.gap <= 1e-9, (mt.pinf, mt.dinf, mt.gap)


def test_crossover_from_the_engine_warm_starts_the_reference_simplex():
    # the engine's FP32 + refinement solution -> igaos_crossover_basis ->
    # the reference's revised simplex (built from its sources) finishes at
    # the cold optimum in a few pivots instead of 1,388
    if not oracle.have_ref_simplex():
        pytest.skip("reference simplex not built (oracle/_ref)")
    lp = lpgen.gen_random_eq(seed=1, m=200, n=600)
    md = igaos_amd.model_from_lp(lp)
    s = igaos_amd.solve_lp(md, time_limit=120, tolerance=1e-6, precision=32)
    assert s.status == "near-optimal"
    basis, st, k = igaos_amd.api.crossover_basis(md, s.x, s.y, 1e-6)
    cold = oracle.ref_simplex(lp, time_limit=60)
    warm = oracle.ref_simplex(lp, time_limit=60, warm=(basis, st))
    assert cold[0] == warm[0] == "optimal"
    assert abs(warm[1] - cold[1]) <= 1e-9 * (1 + abs(cold[1]))
    assert warm[4] <= 0.1 * cold[4], (warm[4], cold[4])
