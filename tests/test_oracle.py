"""Pins of the oracle (oracle/pdhg_oracle.c) — CPU only.

Each oracle function is checked against something other than its own
formula: a library routine (scipy.sparse, numpy), a closed form on inputs
where the answer is known, an invariant of the mathematics, or the
reference's recorded optima (tests/golden).  The mutation cases plant a
plausible mistake in the oracle (a flipped sign, a dropped factor, the
wrong averaging weight) and assert that the same pins then fail.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from igaos_amd import lpgen
from tests.conftest import netlib_path, tiny_lp


@pytest.fixture(autouse=True)
def _faithful():
    oracle.set_mutation(0)
    yield
    oracle.set_mutation(0)


def _rand_csr(rng, m, n, density, empty_rows=()):
    A = sp.random(m, n, density=density, random_state=rng, format="csr")
    A.data = rng.standard_normal(A.data.shape[0])
    A = A.tolil()
    for i in empty_rows:
        A[i, :] = 0
    return A.tocsr()


# ---------------------------------------------------------------- SpMV --
def test_spmv_matches_scipy():
    rng = np.random.default_rng(0)
    A = _rand_csr(rng, 300, 200, 0.05, empty_rows=(0, 17, 299))
    x = rng.standard_normal(200)
    y = oracle.spmv(A.indptr, A.indices, A.data, x)
    np.testing.assert_allclose(y, A @ x, rtol=1e-13, atol=1e-13)
    assert y[0] == 0.0 and y[17] == 0.0


def test_spmv_integer_exact_and_transpose():
    rng = np.random.default_rng(1)
    A = _rand_csr(rng, 50, 80, 0.1)
    A.data = rng.integers(-5, 6, A.data.shape[0]).astype(float)
    x = rng.integers(-3, 4, 80).astype(float)
    assert np.array_equal(oracle.spmv(A.indptr, A.indices, A.data, x), A @ x)
    At = A.T.tocsr()
    yv = rng.integers(-3, 4, 50).astype(float)
    assert np.array_equal(oracle.spmv(At.indptr, At.indices, At.data, yv),
                          A.T @ yv)


def test_spmv_threads_do_not_change_results():
    # rows are split among threads, each row still sums in storage order
    rng = np.random.default_rng(2)
    A = _rand_csr(rng, 9000, 3000, 0.004, empty_rows=(0, 4500, 8999))
    x = rng.standard_normal(3000)
    y1 = oracle.spmv(A.indptr, A.indices, A.data, x)
    lp = lpgen.gen_random_eq(seed=5, m=5000, n=12000)
    r1 = oracle.solve(lp, tolerance=0.0, time_limit=1e7, max_iterations=128)
    oracle.set_threads(4)
    try:
        y4 = oracle.spmv(A.indptr, A.indices, A.data, x)
        r4 = oracle.solve(lp, tolerance=0.0, time_limit=1e7, max_iterations=128)
    finally:
        oracle.set_threads(1)
    assert np.array_equal(y1, y4)
    np.testing.assert_allclose(y1, A @ x, rtol=1e-12, atol=1e-12)
    assert r1.iterations == r4.iterations == 128
    assert np.array_equal(r1.x, r4.x) and np.array_equal(r1.y, r4.y)


# ---------------------------------------------------------------- Ruiz --
def test_ruiz_equilibrates_rows_and_columns():
    lp = lpgen.gen_zipf_badly_scaled(seed=3, m=300, n=700)
    dr, dc = oracle.ruiz(lp)
    A = lp.to_scipy()
    S = sp.diags(1 / dr) @ A @ sp.diags(dc)
    rmax = abs(S).max(axis=1).toarray().ravel()
    cmax = abs(S).max(axis=0).toarray().ravel()
    # values span 1e+-3 before scaling; after 10 sweeps every nonempty row
    # and column has infinity norm near 1
    assert np.all(rmax[np.diff(lp.ap) > 0] > 0.5) and np.all(rmax < 2.0)
    assert np.all(cmax[np.diff(lp.cp) > 0] > 0.5) and np.all(cmax < 2.0)


def test_ruiz_diagonal_closed_form():
    # on a diagonal matrix each sweep takes the fourth root of the scaled
    # entry (row step: square root; column step: square root again), so
    # ten sweeps leave a^(4^-10)
    d = np.array([4.0, 1e-3, 25.0, 1e6])
    lp = lpgen.LP.from_triplets(4, 4, range(4), range(4), d, np.ones(4),
                                np.zeros(4), np.ones(4), np.ones(4),
                                np.ones(4))
    dr, dc = oracle.ruiz(lp)
    np.testing.assert_allclose(d * dc / dr, d ** (4.0 ** -10), rtol=1e-12)


def test_ruiz_mutation_is_caught():
    d = np.array([4.0, 1e-3, 25.0, 1e6])
    lp = lpgen.LP.from_triplets(4, 4, range(4), range(4), d, np.ones(4),
                                np.zeros(4), np.ones(4), np.ones(4),
                                np.ones(4))
    oracle.set_mutation(4)
    dr, dc = oracle.ruiz(lp)
    assert not np.allclose(d * dc / dr, d ** (4.0 ** -10), rtol=1e-6)


def _pc_scaled(lp, dr, dc):
    return (sp.diags(1 / dr) @ lp.to_scipy() @ sp.diags(dc)).toarray()


def test_pock_chambolle_rank_one_has_unit_norm():
    # Pock & Chambolle (ICCV 2011, Lemma 2), alpha = 1: with row sums r and
    # column sums c of |A|, A_ij / sqrt(r_i c_j) has spectral norm <= 1; on
    # a positive rank-one A = u v' it is s t' with |s| = |t| = 1, norm 1
    rng = np.random.default_rng(4)
    u, v = rng.uniform(0.1, 10, 7), rng.uniform(0.01, 100, 11)
    A = np.outer(u, v)
    rows, cols = np.nonzero(A)
    lp = lpgen.LP.from_triplets(7, 11, rows, cols, A[rows, cols], np.ones(11), np.zeros(11),
                                np.ones(11), np.ones(7), np.ones(7))
    dr, dc = oracle.pock_chambolle(lp, np.ones(7), np.ones(11))
    assert abs(np.linalg.norm(_pc_scaled(lp, dr, dc), 2) - 1.0) < 1e-13
    oracle.set_mutation(7)  # dr *= r instead of sqrt(r)
    dr, dc = oracle.pock_chambolle(lp, np.ones(7), np.ones(11))
    assert abs(np.linalg.norm(_pc_scaled(lp, dr, dc), 2) - 1.0) > 1e-3


def test_pock_chambolle_diagonal_and_norm_bound():
    # diagonal: every scaled entry becomes +-1 exactly up to rounding
    d = np.array([4.0, -1e-3, 25.0, 1e6])
    lp = lpgen.LP.from_triplets(4, 4, range(4), range(4), d, np.ones(4), np.zeros(4),
                                np.ones(4), np.ones(4), np.ones(4))
    dr, dc = oracle.pock_chambolle(lp, np.ones(4), np.ones(4))
    np.testing.assert_allclose(np.diag(_pc_scaled(lp, dr, dc)), np.sign(d), rtol=1e-15)
    # after Ruiz (as the engine applies it), on badly scaled sparse LPs
    for seed in range(3):
        lp = lpgen.gen_zipf_badly_scaled(seed=seed, m=120, n=300)
        dr0, dc0 = oracle.ruiz(lp)
        dr, dc = oracle.pock_chambolle(lp, dr0, dc0)
        assert np.linalg.norm(_pc_scaled(lp, dr, dc), 2) <= 1.0 + 1e-12


def test_scale_is_diag_similarity():
    lp = lpgen.gen_random_eq(seed=5, m=40, n=90)
    dr, dc = oracle.ruiz(lp)
    s = oracle.scale(lp, dr, dc)
    S = (sp.diags(1 / dr) @ lp.to_scipy() @ sp.diags(dc)).tocsr()
    S.sort_indices()
    np.testing.assert_allclose(s["sax"], S.data, rtol=1e-15)
    St = sp.csr_matrix((s["sacx"], lp.ci, lp.cp), shape=(lp.n, lp.m))
    np.testing.assert_allclose((St - S.T).toarray(), 0, atol=1e-15)
    np.testing.assert_allclose(s["scl"] * dc, lp.cl)
    np.testing.assert_allclose(s["srmax"] * dr, lp.rmax)


# ------------------------------------------------------- power / omega --
def test_power_iteration_diagonal_closed_form():
    d = np.array([3.0, 2.0, 1.0, 0.5])
    A = sp.diags(d).tocsr()
    At = A.T.tocsr()
    lam2 = oracle.power(4, 4, A.indptr, A.indices, A.data, At.indptr,
                        At.indices, At.data)
    assert abs(lam2 - 9.0) < 1e-9


def test_power_iteration_matches_svd():
    rng = np.random.default_rng(2)
    A = _rand_csr(rng, 60, 90, 0.1)
    At = A.T.tocsr()
    lam2 = oracle.power(60, 90, A.indptr, A.indices, A.data, At.indptr,
                        At.indices, At.data)
    s = np.linalg.svd(A.toarray(), compute_uv=False)
    # 30 power steps: a lower bound that is tight up to the spectral gap
    assert lam2 <= s[0] ** 2 * (1 + 1e-12)
    assert lam2 >= s[0] ** 2 * (1 - 2 * (s[1] / s[0]) ** 120)


def test_omega0_closed_form():
    sc = np.array([3.0, 4.0])            # ||c|| = 5
    srmin = np.array([-np.inf, 1.0, 2.0])
    srmax = np.array([np.inf, 1.0, np.inf])  # q = (0, 1, 2): ||q|| = sqrt5
    assert math.isclose(oracle.omega0(sc, srmin, srmax), math.sqrt(5.0),
                        rel_tol=1e-15)
    assert oracle.omega0(np.array([100.0]), np.array([1e-3]),
                         np.array([1e-3])) == 3.0  # clipped to [0.3, 3]
    assert oracle.omega0(np.zeros(2), srmin, srmax) == 1.0


# ---------------------------------------------------------------- dots --
def test_dots_match_numpy_and_integer_exact():
    rng = np.random.default_rng(4)
    for n, m in ((10, 7), (3000, 1000), (70000, 300000)):
        dx, dy, adx = (rng.standard_normal(n), rng.standard_normal(m),
                       rng.standard_normal(m))
        d = oracle.dots(dx, dy, adx)
        np.testing.assert_allclose(d, [dx @ dx, dy @ dy, dy @ adx],
                                   rtol=1e-12)
        ix, iy, ia = (rng.integers(-9, 10, n).astype(float),
                      rng.integers(-9, 10, m).astype(float),
                      rng.integers(-9, 10, m).astype(float))
        assert list(oracle.dots(ix, iy, ia)) == [ix @ ix, iy @ iy, iy @ ia]


# ----------------------------------------------------------- averaging --
def test_average_weights_are_linear_ramp():
    # x_avg = sum (j+1) x_j / sum (j+1) within an epoch
    w = oracle.mix_weights(9)
    ramp = np.arange(1, 10) / np.arange(1, 10).sum()
    np.testing.assert_allclose(w, ramp, rtol=1e-14)
    oracle.set_mutation(5)
    assert not np.allclose(oracle.mix_weights(9), ramp, rtol=1e-6)


# ---------------------------------------------------------- line search --
def test_line_search_properties():
    # PDLP Algorithm 2: accept iff eta <= etabar (the local bound);
    # the next eta never grows more than (1 + k^-0.6); a rejection shrinks
    # eta below etabar; non-positive curvature always accepts; more than
    # 20 retries accept.
    omega = 1.3
    dots = np.array([2.0, 3.0, 0.5])
    etabar = (omega * 2.0 + 3.0 / omega) / (2 * 0.5)
    for k in (2, 10, 1000):
        acc, e, r = oracle.ls_decide(k, 0.9 * etabar, omega, dots)
        assert acc and r == 0 and e <= 0.9 * etabar * (1 + k ** -0.6) * (1 + 1e-15)
        acc, e, r = oracle.ls_decide(k, 1.1 * etabar, omega, dots)
        assert not acc and r == 1 and e < etabar
    assert oracle.ls_decide(5, 1e9, omega, [1.0, 1.0, -1.0])[0]
    assert oracle.ls_decide(5, 1e9, omega, [0.0, 0.0, 1.0])[0]
    acc, _, r = oracle.ls_decide(5, 10 * etabar, omega, dots, retries=20)
    assert acc and r == 21
    # k = 1: (1 - 1^-0.3) = 0 so an accepted step keeps eta
    acc, e, _ = oracle.ls_decide(1, 0.5, omega, dots)
    assert acc and e == 0.5


def test_line_search_exponents_closed_form():
    # PDLP's adaptive step (Applegate et al. 2021, Algorithm 2):
    # eta' = min((1 - k^-0.3) etabar, (1 + k^-0.6) eta).  At k = 1024 the
    # factors are exactly 1 - 2^-3 and 1 + 2^-6, at k = 2^20 exactly
    # 1 - 2^-6 and 1 + 2^-12; dots (1, 1, 1) with omega 1 give etabar = 1.
    dots = [1.0, 1.0, 1.0]
    for k, a, b in ((1024, 2.0 ** -3, 2.0 ** -6), (1 << 20, 2.0 ** -6, 2.0 ** -12)):
        acc, e, _ = oracle.ls_decide(k, 0.5, 1.0, dots)   # accepted, growth binds
        assert acc and e == 0.5 * (1 + b)
        acc, e, _ = oracle.ls_decide(k, 0.99, 1.0, dots)  # accepted, etabar binds
        assert acc and e == min(1 - a, 0.99 * (1 + b))
        acc, e, _ = oracle.ls_decide(k, 2.0, 1.0, dots)   # rejected
        assert not acc and e == 1 - a


# ------------------------------------------------------------- evaluate --
def _tiny_eval(x, y):
    lp = tiny_lp()
    A = lp.to_scipy()
    d = lp.c + A.T @ y
    return oracle.evaluate(lp.rmin, lp.rmax, lp.cl, lp.cu, lp.c, x, y, A @ x, d)


def test_evaluate_at_optimum_hand_computed():
    mt = _tiny_eval(np.array([0.0, 0.0, 1.0]), np.zeros(2))
    assert (mt.pinf, mt.dinf, mt.prow, mt.op, mt.od, mt.gap) == (0, 0, 0, 0, 0, 0)


def test_evaluate_off_optimum_hand_computed():
    # Ax = (0.9, -0.1): row 0 short by 0.1 -> pinf = 0.1/(1+1), prow 0.1/2;
    # d = (0.75, 1.25, -0.5) all interior -> dinf = 1.25/(1+2);
    # c'x = 0.8; bound = 10*(-0.5) + 0.5*1 - 0.25*0.5 = -4.625
    mt = _tiny_eval(np.array([0.2, 0.3, 0.4]), np.array([-0.5, 0.25]))
    assert math.isclose(mt.pinf, 0.05, rel_tol=1e-14)
    assert math.isclose(mt.prow, 0.05, rel_tol=1e-14)
    assert math.isclose(mt.dinf, 1.25 / 3, rel_tol=1e-14)
    assert math.isclose(mt.op, 0.8, rel_tol=1e-14)
    assert math.isclose(mt.od, -4.625, rel_tol=1e-14)
    assert math.isclose(mt.gap, 5.425 / 1.8, rel_tol=1e-14)


def test_evaluate_wrong_sign_dual_has_no_bound():
    # y < 0 on a <= row with rmin = -inf: the Lagrangian bound is -inf
    mt = _tiny_eval(np.array([0.0, 0.0, 1.0]), np.array([0.0, -0.25]))
    assert mt.od == -math.inf and mt.gap == math.inf


def test_evaluate_weak_duality_random_points():
    rng = np.random.default_rng(7)
    lp = lpgen.gen_random_eq(seed=7, m=30, n=70)
    A = lp.to_scipy()
    from scipy.optimize import linprog
    pstar = linprog(lp.c, A_eq=A, b_eq=lp.rmin, bounds=[(0, 10)] * lp.n,
                    method="highs").fun
    for _ in range(20):
        y = rng.standard_normal(lp.m)
        x = rng.random(lp.n)
        mt = oracle.evaluate(lp.rmin, lp.rmax, lp.cl, lp.cu, lp.c, x, y,
                             A @ x, lp.c + A.T @ y)
        assert mt.od <= pstar + 1e-9  # weak duality for any y


def test_evaluate_dinf_mutation_is_caught():
    oracle.set_mutation(3)
    mt = _tiny_eval(np.array([0.0, 0.0, 1.0]), np.zeros(2))
    assert mt.dinf > 0  # would break the hand-computed optimum pin


# ---------------------------------------------------------- full solves --
def _highs(lp):
    from tests.golden.make_fixtures import scipy_highs
    return scipy_highs(lp)


def test_solve_tiny_to_known_optimum():
    r = oracle.solve(tiny_lp(), tolerance=1e-8, time_limit=30)
    assert r.status == "near-optimal"
    assert abs(r.objective) < 1e-7
    np.testing.assert_allclose(r.x, [0, 0, 1], atol=1e-6)


@pytest.mark.parametrize("name", ["afiro", "sc50a", "adlittle", "sc205"])
def test_solve_netlib_matches_reference_highs(name, objectives):
    from igaos_amd import read_mps
    lp = read_mps(netlib_path(name)).to_lp()
    r = oracle.solve(lp, tolerance=1e-6, time_limit=120)
    ref = objectives[f"netlib/{name}"]["highs_ref"]
    assert r.status == "near-optimal"
    # the reference csv records 10 significant digits
    assert abs(r.objective - ref) <= 1e-6 * (1 + abs(ref)) + 5e-10 * abs(ref)


def test_solve_random_eq_matches_fixtures(objectives):
    lp = lpgen.gen_random_eq(seed=1, m=200, n=600)
    ent = objectives["gen_random_eq(seed=1,m=200,n=600)"]
    r = oracle.solve(lp, tolerance=1e-6, time_limit=120)
    assert r.status == "near-optimal"
    for key in ("scipy_highs", "ref_simplex"):
        if key in ent:
            assert abs(r.objective - ent[key]) <= 1e-6 * (1 + abs(ent[key]))


@pytest.mark.parametrize("code", [1, 2])
def test_solve_mutations_are_caught(code, objectives):
    lp = lpgen.gen_random_eq(seed=2, m=60, n=150)
    ref = objectives["gen_random_eq(seed=2,m=60,n=150)"]["scipy_highs"]
    oracle.set_mutation(code)
    r = oracle.solve(lp, tolerance=1e-6, time_limit=1e7, max_iterations=20000)
    ok = r.status == "near-optimal" and abs(r.objective - ref) <= 1e-6 * (1 + abs(ref))
    assert not ok


def test_fixed_iterations_and_time_limit_semantics():
    lp = lpgen.gen_random_eq(seed=2, m=60, n=150)
    # max_iterations applies only when time_limit_s >= 1e6 (pdhg.cu:669-671)
    r = oracle.solve(lp, tolerance=0.0, time_limit=1e7, max_iterations=100)
    assert r.iterations == 100 and r.status == "iteration-limit"
    r = oracle.solve(lp, tolerance=0.0, time_limit=1e-9, max_iterations=100)
    assert r.status == "time-limit" and r.iterations == 64
