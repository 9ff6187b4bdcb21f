"""GPU tests of the batched persistent PDHG solver (igaos_batch_* C-ABI).

Bars: every instance of a batch certifies (near-optimal) at tolerance 1e-6
with its objective within 1e-6 relative of the reference simplex / HiGHS
fixture; in FP64 mode an instance follows the oracle's trajectory to 1e-9
over one check period and ends with the same status as the single-LP
engine; an instance's result does not depend on the batch around it.
"""
import numpy as np
import pytest

import igaos_amd
import oracle
from igaos_amd import lpgen
from igaos_amd.batch import BatchSolver, solve_batch
from tests.conftest import netlib_path

pytestmark = pytest.mark.gpu


def _fixture_obj(objectives, key):
    ent = objectives[key]
    return ent.get("ref_simplex", ent["scipy_highs"])


def _c4_key(i):
    return f"gen_batch_instance(seed=0,index={i})"


MODES = ["persistent", "lockstep"]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("precision", [32, 64])
def test_batch_solves_match_fixtures(precision, mode, objectives):
    lps = [lpgen.gen_batch_instance(0, i) for i in range(16)]
    sol = solve_batch(lps, precision=precision, mode=mode, tolerance=1e-6,
                      time_limit=120)
    for i, lp in enumerate(lps):
        ref = _fixture_obj(objectives, _c4_key(i))
        assert sol.status[i] == "near-optimal", (i, sol.status[i])
        assert sol.pinf[i] <= 1e-6 and sol.rel_gap[i] <= 1e-6
        assert abs(sol.objective[i] - ref) <= 1e-6 * (1 + abs(ref)), (i, sol.objective[i], ref)
        # the returned pair passes the oracle's KKT evaluation too
        A = lp.to_scipy()
        mt = oracle.evaluate(lp.rmin, lp.rmax, lp.cl, lp.cu, lp.c, sol.x[i],
                             sol.y[i], A @ sol.x[i], lp.c + A.T @ sol.y[i])
        assert mt.pinf <= 1e-6 and mt.gap <= 1e-6


@pytest.mark.parametrize("mode", MODES)
def test_batch_fp64_trajectory_matches_oracle(mode):
    lps = [lpgen.gen_batch_instance(0, i, m_hi=600) for i in range(6)]
    sol = solve_batch(lps, precision=64, mode=mode, tolerance=0.0, time_limit=1e7,
                      max_iterations=64)
    for i, lp in enumerate(lps):
        r = oracle.solve(lp, tolerance=0.0, time_limit=1e7, max_iterations=64)
        assert sol.iterations[i] == 64 and sol.attempts[i] == r.attempts
        err = max(np.abs(sol.x[i] - r.x).max() / max(1, np.abs(r.x).max()),
                  np.abs(sol.y[i] - r.y).max() / max(1, np.abs(r.y).max()))
        assert err <= 1e-9, (i, err)


@pytest.mark.parametrize("mode", MODES)
def test_wide_instance_takes_the_32_bit_index_path(mode):
    # an instance with n > 65,536 keeps the whole batch on 32-bit column
    # indices in row order (no gather-ordered 16-bit copies): same
    # one-period trajectory as the oracle
    wide = lpgen.gen_random_eq(seed=3, m=20_000, n=70_000, per_col=3)
    lps = [lpgen.gen_batch_instance(0, 1, m_hi=600), wide]
    sol = solve_batch(lps, precision=64, mode=mode, tolerance=0.0, time_limit=1e7,
                      max_iterations=64)
    for i, lp in enumerate(lps):
        r = oracle.solve(lp, tolerance=0.0, time_limit=1e7, max_iterations=64)
        assert sol.iterations[i] == 64 and sol.attempts[i] == r.attempts
        err = max(np.abs(sol.x[i] - r.x).max() / max(1, np.abs(r.x).max()),
                  np.abs(sol.y[i] - r.y).max() / max(1, np.abs(r.y).max()))
        assert err <= 1e-9, (i, err)


def test_batch_fp64_matches_single_engine():
    lps = [lpgen.gen_batch_instance(3, i, m_hi=800) for i in range(4)]
    sol = solve_batch(lps, precision=64, tolerance=1e-6, time_limit=60)
    for i, lp in enumerate(lps):
        s = igaos_amd.solve_lp(lp, tolerance=1e-6, time_limit=60)
        assert sol.status[i] == s.status == "near-optimal"
        assert abs(sol.objective[i] - s.objective) <= 1e-6 * (1 + abs(s.objective))


@pytest.mark.parametrize("mode", MODES)
def test_instance_independent_of_batch(mode):
    lps = [lpgen.gen_batch_instance(1, i, m_hi=400) for i in range(10)]
    whole = solve_batch(lps, precision=32, mode=mode, tolerance=1e-6, time_limit=60)
    alone = solve_batch([lps[4]], precision=32, mode=mode, tolerance=1e-6,
                        time_limit=60)
    assert whole.iterations[4] == alone.iterations[0]
    assert np.array_equal(whole.x[4], alone.x[0])
    assert np.array_equal(whole.y[4], alone.y[0])


@pytest.mark.parametrize("mode", MODES)
def test_batch_general_rows_and_bounds(mode, objectives):
    # Netlib instances: inequality rows, free and fixed columns, RANGES
    names = ["afiro", "sc50a", "adlittle"]
    lps = [igaos_amd.read_mps(netlib_path(n)).to_lp() for n in names]
    sol = solve_batch(lps, precision=32, mode=mode, tolerance=1e-6, time_limit=120)
    for i, n in enumerate(names):
        ref = objectives[f"netlib/{n}"]["highs_ref"]
        assert sol.status[i] == "near-optimal", (n, sol.status[i])
        assert abs(sol.objective[i] - ref) <= 1e-6 * (1 + abs(ref)) + 5e-10 * abs(ref)


def test_batch_iteration_cap_and_schedule():
    lps = [lpgen.gen_batch_instance(2, i, m_hi=300) for i in range(5)]
    b = BatchSolver(lps, precision=32)
    tpr, grid, waves = b.schedule()
    assert set(tpr) <= {0, 1, 2, 4, 8} and 1 <= grid <= 5 and waves in (2, 4)
    s1 = b.run(tolerance=0.0, time_limit=1e7, max_iterations=100)
    assert all(s == "iteration-limit" for s in s1.status)
    assert list(s1.iterations) == [100] * 5
    # another SpMV schedule changes the summation order only
    b.set_schedule(np.full(5, 4))
    s2 = b.run(tolerance=1e-6, time_limit=60)
    b.set_schedule(np.full(5, 0))
    s3 = b.run(tolerance=1e-6, time_limit=60)
    for i in range(5):
        assert s2.status[i] == s3.status[i] == "near-optimal"
        assert abs(s2.objective[i] - s3.objective[i]) <= 1e-6 * (1 + abs(s3.objective[i]))
    b.close()


def test_autotuner_picks_a_valid_schedule(tmp_path):
    from igaos_amd import autotune
    lps = [lpgen.gen_batch_instance(4, i, m_hi=700) for i in range(24)]
    cache = str(tmp_path / "tune.json")
    sched = autotune.tune(lps, classes=2, per_class=6, iterations=64, cache=cache)
    assert set(sched.tpr) <= set(autotune.TPR) and sched.waves in autotune.WAVES
    again = autotune.tune(lps, classes=2, per_class=6, iterations=64, cache=cache)
    assert again.tpr == sched.tpr and again.waves == sched.waves
    b = BatchSolver(lps, precision=32)
    autotune.apply(b, sched)
    s = b.run(tolerance=1e-6, time_limit=60)
    b.close()
    assert s.solved == len(lps)
    # with the mode choice: both modes timed on the probe, the faster one set
    sm = autotune.tune(lps, classes=2, per_class=6, iterations=64, modes=True,
                       mode_probe=24)
    assert set(sm.mode_ms) == set(autotune.MODES)
    assert sm.mode == min(sm.mode_ms, key=sm.mode_ms.get)
    b = BatchSolver(lps, precision=32)
    autotune.apply(b, sm)
    assert b.mode == sm.mode
    s = b.run(tolerance=1e-6, time_limit=60)
    b.close()
    assert s.solved == len(lps)


def test_lockstep_iteration_cap_and_modes_agree():
    lps = [lpgen.gen_batch_instance(2, i, m_hi=300) for i in range(5)]
    b = BatchSolver(lps, precision=32, mode="lockstep")
    s1 = b.run(tolerance=0.0, time_limit=1e7, max_iterations=100)
    assert all(s == "iteration-limit" for s in s1.status)
    assert list(s1.iterations) == [100] * 5
    s2 = b.run(tolerance=1e-6, time_limit=60)
    b.set_mode("persistent")
    s3 = b.run(tolerance=1e-6, time_limit=60)
    b.close()
    for i in range(5):
        assert s2.status[i] == s3.status[i] == "near-optimal"
        assert abs(s2.objective[i] - s3.objective[i]) <= 1e-6 * (1 + abs(s3.objective[i]))


@pytest.mark.parametrize("mode", MODES)
def test_infeasible_and_unbounded_instances_do_not_disturb_the_batch(mode):
    # PDHG never certifies infeasibility or unboundedness (the reference has
    # no detector either): such instances run to the cap and report
    # iteration-limit, while the rest of the batch solves as usual
    inf = lpgen.LP.from_triplets(1, 2, [0, 0], [0, 1], [1.0, 1.0], [1.0, 1.0],
                                 [0.0, 0.0], [10.0, 10.0], [-1.0], [-1.0])
    unb = lpgen.LP.from_triplets(1, 2, [0, 0], [0, 1], [1.0, -1.0], [-1.0, 0.0],
                                 [0.0, 0.0], [np.inf, np.inf], [0.0], [0.0])
    good = [lpgen.gen_batch_instance(6, i, m_hi=300) for i in range(3)]
    lps = [good[0], inf, good[1], unb, good[2]]
    sol = solve_batch(lps, precision=32, mode=mode, tolerance=1e-6, time_limit=1e7,
                      max_iterations=100000)
    assert sol.status[1] == "iteration-limit" and sol.pinf[1] > 1e-3
    assert sol.status[3] == "iteration-limit"
    ref = solve_batch(good, precision=32, mode=mode, tolerance=1e-6, time_limit=1e7,
                      max_iterations=100000)
    for k, i in enumerate((0, 2, 4)):
        assert sol.status[i] == "near-optimal"
        assert np.array_equal(sol.x[i], ref.x[k])  # independent of the batch


@pytest.mark.parametrize("mode", MODES)
def test_degenerate_shapes_in_a_batch(mode):
    # an instance without rows (box-constrained: x_j at the bound c_j
    # points away from, objective in closed form), one whose matrix has an
    # empty row and an empty column, and ordinary ones around them
    n0 = 7
    rng = np.random.default_rng(5)
    c0 = rng.standard_normal(n0)
    cl0, cu0 = -rng.uniform(0, 2, n0), rng.uniform(0, 3, n0)
    norows = lpgen.LP.from_triplets(0, n0, [], [], [], c0, cl0, cu0, [], [])
    x0 = np.where(c0 > 0, cl0, cu0)
    # rows 0..2, row 1 empty; column 3 appears in no row (its bound decides)
    gaps = lpgen.LP.from_triplets(3, 4, [0, 0, 2, 2], [0, 1, 1, 2], [1.0, 2.0, -1.0, 1.0],
                                  [1.0, -1.0, 0.5, 2.0], [0.0] * 4, [5.0] * 4,
                                  [1.0, -np.inf, 0.0], [1.0, np.inf, 0.0])
    good = [lpgen.gen_batch_instance(8, i, m_hi=200) for i in range(2)]
    lps = [good[0], norows, gaps, good[1]]
    sol = solve_batch(lps, precision=32, mode=mode, tolerance=1e-6, time_limit=60)
    assert all(st == "near-optimal" for st in sol.status), sol.status
    assert np.allclose(sol.x[1], x0, rtol=0, atol=1e-6)
    assert abs(sol.objective[1] - c0 @ x0) <= 1e-6 * (1 + abs(c0 @ x0))
    # gaps: x3 = 0 (c3 > 0, free of rows); rows force x0 + 2 x1 = 1 and
    # x2 = x1; c.x = x0 - x1 + 0.5 x1 = 1 - 2.5 x1 -> x1 = 0.5, x0 = 0
    assert np.allclose(sol.x[2], [0.0, 0.5, 0.5, 0.0], rtol=0, atol=1e-5)
    assert abs(sol.objective[2] - (-0.25)) <= 1e-5


@pytest.mark.parametrize("precision,min_certified", [(32, 30), (64, 31)])
def test_netlib_sweep_as_one_lockstep_batch(precision, min_certified):
    # the 36 small Netlib LPs of the reference's sweep as one batch (60 s,
    # tolerance 1e-6): most certify, each certified objective lies within
    # 2e-6 relative of HiGHS's (the reference's recorded sweep values), and
    # the rest end at the time limit rather than in a wrong status
    import csv
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "netlib", "sweep_highs.csv")) as f:
        rows = list(csv.DictReader(f))
    lps = [igaos_amd.read_mps(netlib_path(r["instance"])).to_lp() for r in rows]
    sol = solve_batch(lps, precision=precision, mode="lockstep", tolerance=1e-6,
                      time_limit=60)
    assert sol.solved >= min_certified, sol.solved
    for r, st, ob in zip(rows, sol.status, sol.objective):
        ref = float(r["highs_objective"])
        if st == "near-optimal":
            assert abs(ob - ref) <= 2e-6 * (1 + abs(ref)), (r["instance"], ob, ref)
        else:
            assert st == "time-limit", (r["instance"], st)


def test_branch_and_bound_node_relaxations_in_one_batch():
    # the LP relaxations of 16 branch-and-bound nodes (the root with one
    # column's bound tightened to floor / ceil of its root value, as
    # src/milp/milp.cpp:399-421 forms a node's LP) plus a node whose bounds
    # cross, solved in one batched call: every feasible node certifies at
    # 1e-6 with HiGHS's objective, and HiGHS-infeasible nodes do not certify
    from igaos_amd.batch import solve_node_relaxations
    from tests.golden.make_fixtures import scipy_highs
    lp = lpgen.gen_random_eq(seed=2, m=60, n=150)
    root = igaos_amd.solve_lp(lp, time_limit=60, tolerance=1e-6)
    assert root.status == "near-optimal"
    frac = np.abs(root.x - np.round(root.x))
    cols = np.argsort(-frac)[:8]
    nodes = []
    for j in cols:
        for up in (False, True):
            cl, cu = lp.cl.copy(), lp.cu.copy()
            if up:
                cl[j] = np.ceil(root.x[j])
            else:
                cu[j] = np.floor(root.x[j])
            nodes.append((cl, cu))
    bad_cl = lp.cl.copy()
    bad_cl[0] = lp.cu[0] + 1.0
    nodes.append((bad_cl, lp.cu.copy()))
    res = solve_node_relaxations(lp, nodes, tolerance=1e-6, time_limit=20)
    assert res.status[-1] == "infeasible"
    import dataclasses
    for k, (cl, cu) in enumerate(nodes[:-1]):
        ref = scipy_highs(dataclasses.replace(lp, cl=cl, cu=cu))
        if ref is None:
            assert res.status[k] != "near-optimal", k
            continue
        assert res.status[k] == "near-optimal", (k, res.status[k])
        assert abs(res.objective[k] - ref) <= 2e-6 * (1 + abs(ref)), (k, res.objective[k], ref)
        assert res.objective[k] >= root.objective - 1e-6 * (1 + abs(root.objective))
