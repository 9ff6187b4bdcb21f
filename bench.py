"""Benchmark of the MI355X PDHG engines (one JSON line on stdout).

Default workload ("batch"): 8,192 independent random LPs (seed 0) with
m log-uniform in [32, 4096], n = round(m * U[1.5, 3]), 4-8 nonzeros per
column with values N(0,1), rhs from a feasible x0 ~ U[0,1], box bounds
[0, 10], c ~ N(0,1) (igaos_amd.lpgen.gen_batch).  A step is one batched
solve of every instance to relative KKT 1e-6 with FP32 iterations and FP64
refinement (per-instance budget --time-limit seconds, 60 by default), in
the execution mode the autotuner measures as faster on a probe of the
batch (lockstep rounds or one persistent launch; --mode fixes it);
``value`` is instances certified near-optimal per wall second of the timed
steps, with the batch already resident in HBM.  A step lasts tens of
seconds, so warm-up and timed steps run only while the whole run stays
within --budget-s seconds of wall clock (at least one step is timed):
``steps`` / ``warmup`` are the counts actually run, next to the requested
ones.  Beside it: the plain FP64 single-LP engine run instance by instance
on a fixed sample (every 512th instance, within --fp64-budget-s), timed
with its own HBM-resident setup excluded, extrapolated to LPs/s.

``roofline`` (bound: HBM) uses the algorithmic byte model per accepted
iteration of an instance, B = 2 nnz (w_a + 4) + 4 (m + n + 2) +
w (7n + 8m) with w_a = w = 4 (FP32).  Lockstep mode: the dual attempt
kernel's share of it over the active instances of sampled rounds, divided
by those launches' durations (HIP events on the batch stream); persistent
mode: summed over every instance's attempts and divided by the launch's
duration.  ``traffic`` comes from the committed rocprofv3 --pmc summary
when present.

``--workload c3`` solves the block-structured multi-period LP (96 periods,
m = 480,000, n = 768,000) once per step in FP32 + FP64 refinement and
times its attempt kernels.
``--workload c1`` solves the reference's CPU-runnable LP (m = 1,000,
n = 3,000) with the single-LP engine and reports iterations/s.
``--workload zipf`` instead measures the single-LP attempt kernels on the
badly scaled LP (n = 1,000,000, m = 400,000, clip(Zipf(1.5), 2, 2000)
nonzeros per column, 68.5M nonzeros) in FP64 and FP32, plus a bounded
FP32 solve (the LP needs well over 10^4 iterations to reach 1e-6).
``--workload zipf-solve`` solves that LP once, in ``--precision`` 64 or 32
(FP32 + FP64 refinement), within ``--time-limit``.

``cpu_baseline`` is the oracle (a single-threaded C host replay of the
reference engine) on a bounded sample of the same workload, named in
``sample`` (the batch's three CPU legs share --cpu-budget-s).  Beside it, ``cpu_baseline_all_cores`` runs the same replay on
up to 16 host threads (one instance per thread for the batch, the rows of
each sparse product split among the threads for the Zipf LP), and
``cpu_baseline_reference_simplex`` times the reference's own CPU path for
these LPs, its revised simplex built from its sources (``oracle/_ref``,
present when it was built), on the single-thread sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

T_START = time.perf_counter()

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("LPs solved/sec to 1e-6 rel KKT (batched FP32+refine vs plain "
          "FP64); HBM GB/s")
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_iteration(m, n, nnz, w_a, w):
    return 2 * nnz * (w_a + 4) + 4 * (m + n + 2) + w * (7 * n + 8 * m)


def load_traffic(name, alg_bytes=None, kernel=None):
    """HBM bytes per launch from the committed rocprofv3 counter summary:
    measured directly for a fixed launch, or, where the summary holds the
    ratio of measured to algorithmic bytes on full launches, that ratio
    times this launch's algorithmic bytes.  ``kernel`` picks a kernel's
    entry of a summary that holds several."""
    path = os.path.join(ROOT, "profiles", f"{name}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if kernel is not None and isinstance(d.get(kernel), dict):
        d = d[kernel]
    if "hbm_bytes_per_launch" in d:
        return d["hbm_bytes_per_launch"]
    if alg_bytes is not None and "traffic_over_algorithmic" in d:
        return int(round(d["traffic_over_algorithmic"] * alg_bytes))
    return None


# ---------------------------------------------------------------- batch --
def bench_batch(args):
    """The 8,192-LP batch.  A step is one full batched solve; steps are
    timed while the wall-clock budget allows (see ``--budget-s``), so the
    line reports how many it timed next to how many were asked for."""
    import numpy as np
    import igaos_amd
    from igaos_amd import lpgen
    from igaos_amd.batch import BatchSolver

    count = args.count
    t = time.perf_counter()
    lps = lpgen.gen_batch(0, count)
    nnz = np.array([lp.nnz for lp in lps])
    ms = np.array([lp.m for lp in lps])
    ns = np.array([lp.n for lp in lps])
    log(f"[bench] {count} LPs generated in {time.perf_counter() - t:.1f} s: "
        f"sum m={ms.sum()} n={ns.sum()} nnz={nnz.sum()}")
    sched = None
    mode = args.mode
    tune_s = 0.0
    if args.mode in ("persistent", "auto"):
        from igaos_amd import autotune
        t = time.perf_counter()
        sched = autotune.tune(lps, log=log, modes=args.mode == "auto")
        tune_s = time.perf_counter() - t
        mode = sched.mode if args.mode == "auto" else "persistent"
        log(f"[bench] autotuned in {tune_s:.1f} s: mode {mode}, tpr by class "
            f"{sched.tpr} (nnz edges {sched.edges}), waves/SIMD {sched.waves}")
    t_create = time.perf_counter()
    b = BatchSolver(lps, precision=32, mode=mode)
    t_create = time.perf_counter() - t_create
    if sched is not None:
        autotune.apply(b, sched)
    tpr, grid, waves = b.schedule()
    log(f"[bench] batch resident on the GPU, mode {mode}"
        + (f", persistent grid {grid}" if mode == "persistent" else ""))
    t0 = [0.0]

    def prog(done, total):
        log(f"[bench]   {done}/{total} instances done at "
            f"{time.perf_counter() - t0[0]:.1f} s")

    def one_solve():
        t0[0] = time.perf_counter()
        s = b.run(tolerance=args.tolerance, time_limit=args.time_limit,
                  want_xy=False, progress=prog, poll_s=10.0)
        return s, time.perf_counter() - t0[0]

    # the time the legs after the timed steps may take (each is bounded)
    reserve = args.fp64_budget_s + (0.0 if args.no_cpu_baseline else args.cpu_budget_s) + 15.0
    # a step cannot outlast the per-instance budget by much: every instance
    # starts with the launch (lockstep) and stops at its time limit
    est = args.time_limit + 5.0

    def room():
        return args.budget_s - (time.perf_counter() - T_START) - reserve

    # timed steps come first: a warm-up solve runs only where the budget
    # holds every requested timed step as well (the autotuner's probe runs
    # have already run every kernel at the batch's shapes)
    fit = int(room() // est)
    warm_done = 0
    for w in range(min(args.warmup, max(0, fit - args.steps))):
        s, el = one_solve()
        warm_done += 1
        log(f"[bench] warmup {w}: {s.solved}/{count} solved in {el:.1f} s")
    solved = 0
    elapsed = 0.0
    kernel_ms = 0.0
    sol = None
    step_s = []
    lk = None
    for k in range(args.steps):
        if k > 0 and room() < est:
            break
        sol, el = one_solve()
        step_s.append(el)
        elapsed += el
        kernel_ms += sol.kernel_ms
        solved += sol.solved
        est = max(step_s) + 2.0
        log(f"[bench] step {k}: {sol.solved}/{count} solved in {el:.1f} s, launch "
            f"{sol.kernel_ms:.0f} ms")
        if mode == "lockstep" and lk is None:
            lk = b.lockstep_timing()
    steps = len(step_s)
    b.close()
    value = solved / elapsed
    its = np.array(sol.iterations)
    att = np.array(sol.attempts)
    ok32 = np.array([st == "near-optimal" for st in sol.status])
    tms = np.array(sol.solve_ms)
    alg = float((att * bytes_per_iteration(ms, ns, nnz, 4, 4)).sum())
    launch_ms = kernel_ms / steps
    achieved_run = alg / (launch_ms * 1e-3) / 1e9
    dom = None
    if lk is not None and lk["samples"] > 0:
        # dominant kernel: the attempt kernel with the larger time per round
        dom = "primal" if lk["primal_ms"] >= lk["dual_ms"] else "dual"
        achieved = lk[f"{dom}_bytes"] / (lk[f"{dom}_ms"] * 1e-3) / 1e9
    else:
        achieved = achieved_run

    # plain FP64 single-LP engine, instance by instance, on a fixed sample,
    # within its own budget (each instance gets what is left, at most the
    # per-instance limit); --fp64-sample 0 skips it
    sample = (list(range(0, count, max(1, count // args.fp64_sample)))
              if args.fp64_sample > 0 else [])
    t64 = 0.0
    solved64 = 0
    run64 = []
    rel32 = []  # batched FP32+refine objective vs the plain FP64 engine's
    tb = time.perf_counter()
    for i in sample:
        left = args.fp64_budget_s - (time.perf_counter() - tb)
        if left < 1.0:
            break
        e = igaos_amd.Engine(lps[i], precision=64)
        s = e.run(time_limit=min(args.time_limit, left), tolerance=args.tolerance)
        e.close()
        run64.append(i)
        t64 += s.solve_time_ms * 1e-3
        solved64 += s.status == "near-optimal"
        if s.status == "near-optimal" and sol.status[i] == "near-optimal":
            rel32.append(abs(sol.objective[i] - s.objective) / (1 + abs(s.objective)))
    fp64_rate = solved64 / t64 if t64 > 0 else 0.0
    log(f"[bench] plain FP64 per instance: {solved64}/{len(run64)} solved "
        f"in {t64:.1f} s -> {fp64_rate:.2f} LPs/s")

    cpu = cpu_par = cpu_ref = None
    if not args.no_cpu_baseline:
        cpu, cpu_par, cpu_ref = batch_cpu_legs(args, lps, sample)

    unsolved = [i for i in range(count) if not ok32[i]]
    return {
        "metric": METRIC, "value": value, "unit": "LPs/s", "n_gpus": 1,
        "steps": steps, "warmup": warm_done,
        "steps_requested": args.steps, "warmup_requested": args.warmup,
        "ms_per_step": elapsed * 1000.0 / steps,
        "step_seconds": [round(x, 2) for x in step_s],
        "budget_s": args.budget_s,
        "budget_note": (f"a step is one full solve of the batch (per-instance limit "
                        f"{args.time_limit:g} s, fixed); steps are timed while the "
                        f"run stays within {args.budget_s:g} s of wall clock, so "
                        f"'steps' is the number timed; the autotuner's probe runs "
                        f"(every kernel at the batch's shapes) warm the GPU up"),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64refine", "data": "synthetic",
        "config": {"workload": (f"batch of {count} random LPs, m log-uniform "
                                f"[32,4096], n=m*U[1.5,3], 4-8 nnz/col, seed 0"),
                   "instances": count, "sum_m": int(ms.sum()),
                   "sum_n": int(ns.sum()), "sum_nnz": int(nnz.sum()),
                   "tolerance": args.tolerance,
                   "time_limit_per_instance_s": args.time_limit,
                   "mode": mode,
                   "mode_choice": ("autotuned" if args.mode == "auto" else "fixed"),
                   "schedule": None if sched is None else {
                       "persistent_grid": int(grid),
                       "tpr_by_class": sched.tpr,
                       "nnz_class_edges": sched.edges,
                       "waves_per_simd": sched.waves,
                       "mode_probe_ms": sched.mode_ms,
                       "tuning_s": round(tune_s, 1)}},
        "solved": solved, "instances": count * steps,
        "last_step": {"certified": int(ok32.sum()), "instances": count,
                      "makespan_s": round(step_s[-1], 2),
                      "certified_by_s": {str(q): int((tms[ok32] <= q * 1e3).sum())
                                         for q in (10, 20, 30, 45, 60)},
                      "uncertified": unsolved[:64],
                      "uncertified_iterations": [int(its[i]) for i in unsolved[:64]]},
        "iterations_per_s": float(its.sum()) / (elapsed / steps),
        "end_to_end": {
            "value": solved / (t_create + elapsed), "unit": "LPs/s",
            "setup_s": round(t_create, 2),
            "note": ("from host arrays to certified results: upload, row blocks "
                     "and gather-ordered copies, then the step(s); the autotuner's "
                     "probe runs are not included")},
        "iterations_median": float(np.median(its)),
        "iterations_p99": float(np.percentile(its, 99)),
        "launch_ms": launch_ms,
        "plain_fp64_per_instance": {
            "value": fp64_rate, "unit": "LPs/s", "sample": run64,
            "solved": solved64, "solve_seconds": round(t64, 3),
            "budget_s": args.fp64_budget_s},
        "fp32_refined_error": {
            "objective_vs_fp64_engine_rel_max": max(rel32) if rel32 else None,
            "compared": len(rel32),
            "certified_pinf_max": float(np.max(np.where(ok32, sol.pinf, 0.0))),
            "certified_dinf_max": float(np.max(np.where(ok32, sol.dinf, 0.0))),
            "certified_rel_gap_max": float(np.max(np.where(ok32, sol.rel_gap, 0.0))),
            "note": ("FP64 KKT of every certified batched result (original units), "
                     "and its objective against the plain FP64 engine on the sample")},
        "batched_vs_plain_fp64": (value / fp64_rate) if fp64_rate > 0 else None,
        "roofline": {"bound": "hbm",
                     "kernel": ("k_batch<float> (one persistent launch)"
                                if dom is None else
                                f"k_lk_{dom}<float> (HIP events around it in sampled "
                                f"lockstep rounds of the first timed step)"),
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": (None if dom is None else
                                 load_traffic("lockstep_batch8192", lk[f"{dom}_bytes"],
                                              kernel=dom)),
                     "bytes_per_launch": alg if dom is None else lk[f"{dom}_bytes"],
                     "ms_per_launch": round(launch_ms if dom is None else lk[f"{dom}_ms"], 4),
                     "lockstep_samples": lk,
                     "whole_run_gbs": round(achieved_run, 1)},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_par,
        "cpu_baseline_reference_simplex": cpu_ref,
    }


def batch_cpu_legs(args, lps, sample):
    """CPU baselines of the batch, each on a bounded sample: the oracle on
    one thread, the oracle on up to 16 threads (one instance each), and the
    reference's revised simplex (oracle/_ref) when it was built."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    count = len(lps)
    # 1 thread: instances of the sample in order while the leg's share of
    # the budget lasts
    share = args.cpu_budget_s / 3.0
    tc = 0.0
    sc = 0
    done = []
    wall = time.perf_counter()
    for i in sample:
        left = share - (time.perf_counter() - wall)
        if left < 1.0:
            break
        r = oracle.solve(lps[i], tolerance=args.tolerance,
                         time_limit=min(args.cpu_time_limit, left))
        tc += r.solve_time_ms * 1e-3
        sc += r.status == "near-optimal"
        done.append(i)
    wall = time.perf_counter() - wall
    cpu = {"value": sc / tc if tc > 0 else 0.0, "unit": "LPs/s",
           "cores": 1, "kind": "port",
           "sample": (f"oracle host replay (1 thread) on instances {done} of the "
                      f"batch, {sc}/{len(done)} certified in {tc:.1f} s of solve time "
                      f"({wall:.1f} s wall incl. setup), at most "
                      f"{args.cpu_time_limit:g} s each")}
    log(f"[bench] oracle 1 thread: {cpu['value']:.3f} LPs/s")
    # all cores: one instance per thread, one instance each
    cores = max(1, min(16, os.cpu_count() or 1))
    step = max(1, count // cores)
    cs = list(range(0, count, step))[:cores]
    tl = min(args.cpu_time_limit, share)
    wall = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        res = list(ex.map(lambda i: oracle.solve(lps[i], tolerance=args.tolerance,
                                                 time_limit=tl), cs))
    wall = time.perf_counter() - wall
    sc = sum(r.status == "near-optimal" for r in res)
    cpu_par = {"value": sc / wall, "unit": "LPs/s", "cores": cores, "kind": "port",
               "host_cpus": os.cpu_count(), "cpu_model": cpu_model(),
               "sample": (f"oracle host replay, {cores} threads, one instance each: "
                          f"instances {cs}, {sc} certified in {wall:.1f} s wall, "
                          f"at most {tl:g} s each")}
    log(f"[bench] oracle {cores} threads: {cpu_par['value']:.3f} LPs/s")
    cpu_ref = None
    if oracle.have_ref_simplex():
        # the reference's own CPU path for these LPs: its revised simplex
        # (built from its sources), 5 s each, while the share lasts
        wall = time.perf_counter()
        sc = 0
        done = []
        for i in sample:
            if share - (time.perf_counter() - wall) < 5.0:
                break
            st = oracle.ref_simplex(lps[i], time_limit=5.0)[0]
            sc += st == "optimal"
            done.append(i)
        wall = time.perf_counter() - wall
        cpu_ref = {"value": sc / wall if wall > 0 else 0.0, "unit": "LPs/s",
                   "cores": 1, "kind": "reference",
                   "sample": (f"reference revised simplex (src/simplex, 1 thread) "
                              f"on instances {done}, {sc}/{len(done)} optimal in "
                              f"{wall:.1f} s wall, budget 5 s each")}
    return cpu, cpu_par, cpu_ref


# ----------------------------------------------------------------- zipf --
# ------------------------------------------------------------------ diff --
def bench_diff(args):
    """Differentiable batch: 512 LPs (m log-uniform [256, 1024], n = 2m,
    6 nnz/col), FP32 c, b, A values as requires_grad tensors, loss
    sum(w.x*) + 0.5 |y*|^2; a step is forward (batched solve to the layer's
    tolerance + active-set polish) and backward (batched adjoint)."""
    import numpy as np
    import torch
    from igaos_amd import lpgen
    from igaos_amd.autograd import DiffLPBatch

    count = 512
    t = time.perf_counter()
    lps = []
    for i in range(count):
        lp = lpgen.gen_batch_instance(0, i, m_lo=256, m_hi=1024, ratio_lo=2.0,
                                      ratio_hi=2.0, k_lo=6, k_hi=6)
        lps.append(lpgen.perturb_nondegenerate(lp, seed=i, scale=1e-6))
    log(f"[bench] {count} LPs generated in {time.perf_counter() - t:.1f} s")
    mode = "lockstep" if args.mode == "auto" else args.mode
    layer = DiffLPBatch(lps, precision=32, mode=mode, time_limit=args.diff_time_limit)
    c, rmin, rmax, ax = layer.data(torch.float32)
    w = torch.from_numpy(np.random.default_rng(7).standard_normal(c.numel())).float()
    fwd = bwd = 0.0
    for k in range(args.warmup + args.steps):
        cc, bb, aa = (v.clone().requires_grad_(True) for v in (c, rmin, ax))
        t0 = time.perf_counter()
        x, y, obj = layer.solve_eq(cc, bb, aa)
        t1 = time.perf_counter()
        loss = (w * x).sum() + 0.5 * (y * y).sum()
        loss.backward()
        t2 = time.perf_counter()
        if k >= args.warmup:
            fwd += t1 - t0
            bwd += t2 - t1
        log(f"[bench] step {k}: forward {t1 - t0:.2f} s backward {t2 - t1:.2f} s")
    sol = layer.last_solution
    pinfo = layer.last_polish_info
    ainfo = layer.last_adjoint_info
    layer.close()
    # the same loss through the unrolled mode: K fixed-step iterations with
    # device checkpoints every 64, reverse mode through recomputed segments
    K = args.unroll_iterations
    ul = DiffLPBatch(lps, precision=32, differentiation="unrolled", unroll_iterations=K)
    cc, bb, aa = (v.clone().double().requires_grad_(True) for v in (c, rmin, ax))
    t0 = time.perf_counter()
    x, y, obj = ul.solve_eq(cc, bb, aa)
    t1 = time.perf_counter()
    ((w.double() * x).sum() + 0.5 * (y * y).sum()).backward()
    t2 = time.perf_counter()
    ul.close()
    log(f"[bench] unrolled K={K}: forward {t1 - t0:.2f} s backward {t2 - t1:.2f} s")
    unrolled = {"iterations": K, "forward_s": t1 - t0, "backward_s": t2 - t1,
                "checkpoint_every": 64,
                "instance_iterations_per_s": count * K / (t1 - t0)}
    el = fwd + bwd
    ok = np.array([st == "near-optimal" for st in sol.status]) & (pinfo["accepted"] == 1)
    return {
        "metric": "differentiable batch: LPs/s (forward solve + polish, backward adjoint)",
        "value": count * args.steps / el, "unit": "LPs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1000.0 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64refine",
        "data": "synthetic",
        "config": {"workload": "512 LPs, m log-uniform [256,1024], n=2m, 6 nnz/col, "
                               "nondegenerate perturbation 1e-6, seed 0",
                   "tolerance": layer.tolerance, "mode": mode,
                   "time_limit_per_instance_s": args.diff_time_limit},
        "forward_s": fwd / args.steps, "backward_s": bwd / args.steps,
        "certified": sol.solved, "polished": int(np.sum(pinfo["accepted"])),
        "adjoint_cgls_iterations_max": int(max(ainfo["iters_p"][ok].max(),
                                               ainfo["iters_q"][ok].max())),
        "adjoint_residual_max": float(max(ainfo["res_p"][ok].max(),
                                          ainfo["res_q"][ok].max())),
        "note": "adjoint statistics over the certified and polished instances",
        "unrolled_mode": unrolled,
    }


def bench_zipf(args):
    import numpy as np
    import igaos_amd
    from igaos_amd import lpgen

    t = time.perf_counter()
    lp = lpgen.gen_zipf_badly_scaled(seed=0)
    log(f"[bench] zipf LP nnz={lp.nnz} generated in "
        f"{time.perf_counter() - t:.1f} s")
    kern = {}
    for prec in (64, 32):
        e = igaos_amd.Engine(lp, precision=prec)
        tm = e.profile_attempts(args.attempts)
        sl = e.column_slices()
        w = 8 if prec == 64 else 4
        bp = lp.nnz * (w + 4) + 4 * (lp.n + 1) + 7 * lp.n * w
        bd = lp.nnz * (w + 4) + 4 * (lp.m + 1) + 8 * lp.m * w
        kern[prec] = dict(ms_primal=tm["ms_primal"], ms_dual=tm["ms_dual"],
                          gbs_primal=bp / tm["ms_primal"] / 1e6,
                          gbs_dual=bd / tm["ms_dual"] / 1e6,
                          bytes_primal=bp, bytes_dual=bd,
                          kernels_primal=step_kernels("primal", sl["primal"], prec),
                          kernels_dual=step_kernels("dual", sl["dual"], prec))
        log(f"[bench] FP{prec}: primal step {tm['ms_primal']:.3f} ms "
            f"({kern[prec]['gbs_primal']:.0f} GB/s, {kern[prec]['kernels_primal']}), "
            f"dual step {tm['ms_dual']:.3f} ms ({kern[prec]['gbs_dual']:.0f} GB/s, "
            f"{kern[prec]['kernels_dual']})")
        if prec == 32:
            torch_sync()
            t0 = time.perf_counter()
            s = e.run(time_limit=args.time_limit, tolerance=args.tolerance,
                      verbosity=1)
            el = time.perf_counter() - t0
            log(f"[bench] FP32 solve: {s.status} {s.iterations} it {el:.1f} s")
        e.close()
    k64 = kern[64]
    dom = "dual" if k64["ms_dual"] >= k64["ms_primal"] else "primal"
    ms_dom = k64[f"ms_{dom}"]
    b_dom = k64[f"bytes_{dom}"]
    achieved = b_dom / ms_dom / 1e6
    solved = 1 if s.status == "near-optimal" else 0
    cpu = cpu_par = None
    if not args.no_cpu_baseline:
        import oracle
        t0 = time.perf_counter()
        r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=64)
        wall = time.perf_counter() - t0
        ips = r.iterations / (r.solve_time_ms * 1e-3)
        cpu = {"value": ips, "unit": "iterations/s", "cores": 1,
               "kind": "port",
               "sample": (f"oracle host replay (1 thread): setup + 64 "
                          f"iterations of the same LP, {wall:.1f} s wall")}
        log(f"[bench] oracle 1 thread: {ips:.2f} it/s")
        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        try:
            t0 = time.perf_counter()
            r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=64)
            wall = time.perf_counter() - t0
        finally:
            oracle.set_threads(1)
        ips = r.iterations / (r.solve_time_ms * 1e-3)
        cpu_par = {"value": ips, "unit": "iterations/s", "cores": threads,
                   "kind": "port",
                   "sample": (f"oracle host replay, sparse products row-split "
                              f"over {threads} threads: setup + 64 iterations "
                              f"of the same LP, {wall:.1f} s wall"),
                   "cpu_model": cpu_model()}
        log(f"[bench] oracle {threads} threads: {ips:.2f} it/s")
    return {
        "metric": METRIC, "value": solved / el, "unit": "LPs/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": el * 1000.0,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "badly scaled Zipf LP, m=400000 n=1000000 "
                               "nnz=68528436 seed 0",
                   "tolerance": args.tolerance,
                   "time_limit_s": args.time_limit},
        "fp32_solve": {"status": s.status, "iterations": s.iterations,
                       "iterations_per_s": s.iterations / el,
                       "pinf": s.pinf, "dinf": s.dinf, "rel_gap": s.rel_gap},
        "attempt_kernels": kern,
        "roofline": {"bound": "hbm", "kernel": f"FP64 {dom} step: {k64[f'kernels_{dom}']}",
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": load_traffic(f"{dom}_step_zipf"),
                     "bytes_per_launch": b_dom,
                     "ms_per_launch": round(ms_dom, 4)},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_par,
    }


def bench_zipf_solve(args):
    """One solve of the 1M-column LP to the tolerance in one precision
    (``--precision``, FP64 or FP32 + FP64 refinement) within
    ``--time-limit``: status, iterations/s, the KKT measures and the
    objective (the FP32 result's error is read against an FP64 run of the
    same LP), and the roofline of the precision's dominant attempt step."""
    import igaos_amd
    from igaos_amd import lpgen

    prec = args.precision
    t = time.perf_counter()
    lp = lpgen.gen_zipf_badly_scaled(seed=0)
    log(f"[bench] zipf LP nnz={lp.nnz} generated in {time.perf_counter() - t:.1f} s")
    e = igaos_amd.Engine(lp, precision=prec, scaling=args.scaling)
    tm = e.profile_attempts(args.attempts)
    sl = e.column_slices()
    w = 8 if prec == 64 else 4
    bp = lp.nnz * (w + 4) + 4 * (lp.n + 1) + 7 * lp.n * w
    bd = lp.nnz * (w + 4) + 4 * (lp.m + 1) + 8 * lp.m * w
    dom = "dual" if tm["ms_dual"] >= tm["ms_primal"] else "primal"
    ms_dom = tm[f"ms_{dom}"]
    b_dom = bp if dom == "primal" else bd
    torch_sync()
    t0 = time.perf_counter()
    s = e.run(time_limit=args.time_limit, tolerance=args.tolerance, verbosity=1)
    el = time.perf_counter() - t0
    e.close()
    log(f"[bench] FP{prec} solve: {s.status} {s.iterations} it {el:.1f} s "
        f"obj {s.objective:.12g}")
    achieved = b_dom / ms_dom / 1e6
    solved = 1 if s.status == "near-optimal" else 0
    return {
        "metric": METRIC, "value": solved / el, "unit": "LPs/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": el * 1000.0,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == 64 else "f32+f64refine", "data": "synthetic",
        "config": {"workload": "badly scaled Zipf LP, m=400000 n=1000000 "
                               "nnz=68528436 seed 0, one solve",
                   "precision": prec, "scaling": args.scaling, "tolerance": args.tolerance,
                   "time_limit_s": args.time_limit},
        "solve": {"status": s.status, "iterations": s.iterations,
                  "iterations_per_s": s.iterations / el, "objective": s.objective,
                  "pinf": s.pinf, "dinf": s.dinf, "rel_gap": s.rel_gap,
                  "solve_time_ms": s.solve_time_ms},
        "roofline": {"bound": "hbm",
                     "kernel": f"FP{prec} {dom} step: {step_kernels(dom, sl[dom], prec)}",
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": None, "bytes_per_launch": b_dom,
                     "ms_per_launch": round(ms_dom, 4)},
        "cpu_baseline": None,
    }


def bench_c3(args):
    """The block-structured multi-period blending LP (96 periods of 5,000
    rows x 8,000 columns, 6 nonzeros per column, 1,000 inventory columns
    per period linked to the next period, bounds [0, 10]): one FP32 + FP64
    refinement solve to 1e-6 (a step), its iterations/s, and the attempt
    kernels' roofline from timed attempts in FP32 and FP64."""
    import numpy as np
    import igaos_amd
    from igaos_amd import lpgen

    t = time.perf_counter()
    lp = lpgen.gen_multiperiod(seed=0)
    log(f"[bench] multi-period LP m={lp.m} n={lp.n} nnz={lp.nnz} generated in "
        f"{time.perf_counter() - t:.1f} s")
    kern = {}
    for prec in (64, 32):
        e = igaos_amd.Engine(lp, precision=prec)
        tm = e.profile_attempts(args.attempts)
        e.close()
        w = 8 if prec == 64 else 4
        bp = lp.nnz * (w + 4) + 4 * (lp.n + 1) + 7 * lp.n * w
        bd = lp.nnz * (w + 4) + 4 * (lp.m + 1) + 8 * lp.m * w
        kern[prec] = dict(ms_primal=tm["ms_primal"], ms_dual=tm["ms_dual"],
                          gbs_primal=bp / tm["ms_primal"] / 1e6,
                          gbs_dual=bd / tm["ms_dual"] / 1e6,
                          bytes_primal=bp, bytes_dual=bd)
        log(f"[bench] FP{prec}: primal {tm['ms_primal']:.3f} ms ({kern[prec]['gbs_primal']:.0f} "
            f"GB/s), dual {tm['ms_dual']:.3f} ms ({kern[prec]['gbs_dual']:.0f} GB/s)")
    e = igaos_amd.Engine(lp, precision=32)
    setup_ms = e.setup_ms
    solves = []
    for k in range(args.warmup + args.steps):
        s = e.run(time_limit=args.time_limit, tolerance=args.tolerance)
        log(f"[bench] FP32+refine solve {k}: {s.status} {s.iterations} it "
            f"{s.solve_time_ms:.0f} ms obj {s.objective:.10g}")
        if k >= args.warmup:
            solves.append(s)
    e.close()
    ms_tot = sum(s.solve_time_ms for s in solves)
    solved = sum(s.status == "near-optimal" for s in solves)
    its = sum(s.iterations for s in solves)
    k32 = kern[32]
    dom = "primal" if k32["ms_primal"] >= k32["ms_dual"] else "dual"
    achieved = k32[f"bytes_{dom}"] / k32[f"ms_{dom}"] / 1e6
    cpu = cpu_par = None
    if not args.no_cpu_baseline:
        import oracle
        t0 = time.perf_counter()
        r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=64)
        wall = time.perf_counter() - t0
        cpu = {"value": r.iterations / (r.solve_time_ms * 1e-3), "unit": "iterations/s",
               "cores": 1, "kind": "port",
               "sample": (f"oracle host replay (1 thread): setup + 64 iterations of the "
                          f"same LP, {wall:.1f} s wall")}
        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        try:
            t0 = time.perf_counter()
            r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=64)
            wall = time.perf_counter() - t0
        finally:
            oracle.set_threads(1)
        cpu_par = {"value": r.iterations / (r.solve_time_ms * 1e-3), "unit": "iterations/s",
                   "cores": threads, "kind": "port", "cpu_model": cpu_model(),
                   "sample": (f"oracle host replay, sparse products row-split over "
                              f"{threads} threads: setup + 64 iterations, {wall:.1f} s wall")}
    last = solves[-1]
    return {
        "metric": METRIC, "value": solved / (ms_tot * 1e-3), "unit": "LPs/s", "n_gpus": 1,
        "steps": len(solves), "warmup": args.warmup, "ms_per_step": ms_tot / len(solves),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64refine", "data": "synthetic",
        "config": {"workload": (f"multi-period blending LP, 96 periods x (5000 x 8000), "
                                f"m={lp.m} n={lp.n} nnz={lp.nnz}, seed 0"),
                   "tolerance": args.tolerance, "time_limit_s": args.time_limit},
        "solve": {"status": last.status, "iterations": last.iterations,
                  "iterations_per_s": its / (ms_tot * 1e-3), "objective": last.objective,
                  "pinf": last.pinf, "dinf": last.dinf, "rel_gap": last.rel_gap,
                  "setup_ms": setup_ms,
                  "note": "solve_time_ms excludes setup, as the reference reports it"},
        "attempt_kernels": kern,
        "roofline": {"bound": "hbm", "kernel": f"FP32 {dom} step",
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                     "bytes_per_launch": k32[f"bytes_{dom}"],
                     "ms_per_launch": round(k32[f"ms_{dom}"], 4)},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_par,
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def step_kernels(step, slices, prec):
    """Kernels of one primal / dual step of the single-LP engine."""
    t = "double" if prec == 64 else "float"
    if slices:
        return f"k_{step}_slices<{t}> x{slices} + k_{step}_epi<{t}>"
    return f"k_{step}<{t}>"


# ------------------------------------------------------------------- c1 --
def bench_c1(args):
    """The reference's CPU-runnable case: one random equality LP (m = 1,000,
    n = 3,000, 5 nonzeros per column, seed 0) to 1e-6, plain FP64 and FP32 +
    refinement, each solved `steps` times from a prepared engine.  The LP
    runs on the one-workgroup runner (a check period per launch), whose one
    CU streams the matrices from L2: the figure is iterations/s, bound by
    that CU's vector-memory instruction rate, not an HBM fraction."""
    import numpy as np
    import igaos_amd
    from igaos_amd import lpgen

    lp = lpgen.gen_random_eq(seed=0)
    out = {}
    for prec in (64, 32):
        e = igaos_amd.Engine(lp, precision=prec)
        for _ in range(args.warmup):
            e.run(time_limit=args.time_limit, tolerance=args.tolerance)
        its, ms, st, ob = 0, 0.0, None, None
        for _ in range(args.steps):
            s = e.run(time_limit=args.time_limit, tolerance=args.tolerance)
            its += s.iterations
            ms += s.solve_time_ms
            st, ob = s.status, s.objective
        e.close()
        out[prec] = {"status": st, "objective": ob, "iterations": its // args.steps,
                     "solve_ms": ms / args.steps, "iterations_per_s": its / (ms * 1e-3),
                     "solves_per_s": args.steps / (ms * 1e-3)}
        log(f"[bench] C1 FP{prec}: {st} obj {ob:.10g}, {its // args.steps} it, "
            f"{ms / args.steps:.1f} ms per solve")
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        t0 = time.perf_counter()
        r = oracle.solve(lp, tolerance=args.tolerance, time_limit=args.cpu_time_limit)
        wall = time.perf_counter() - t0
        cpu = {"value": r.iterations / (r.solve_time_ms * 1e-3), "unit": "iterations/s",
               "cores": 1, "kind": "port",
               "sample": (f"oracle host replay (1 thread) of the same LP: {r.status} after "
                          f"{r.iterations} iterations, {wall:.1f} s wall, budget "
                          f"{args.cpu_time_limit} s")}
    k64 = out[64]
    return {
        "metric": "C1 single-LP solve: iterations/s (plain FP64 engine)",
        "value": k64["iterations_per_s"], "unit": "iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": k64["solve_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "random equality LP m=1000 n=3000, 5 nnz/col, seed 0",
                   "tolerance": args.tolerance},
        "engines": out, "roofline": None,
        "note": ("one 1,024-thread workgroup runs each check period (the matrices, "
                 "0.3 MB per product, stream from L2): bound by one CU's "
                 "vector-memory instruction rate, not by HBM"),
        "cpu_baseline": cpu,
    }


def torch_sync():
    import torch
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", default="batch",
                    choices=["batch", "zipf", "zipf-solve", "diff", "c1", "c3"])
    ap.add_argument("--precision", type=int, default=32, choices=[32, 64],
                    help="zipf-solve workload: FP64, or FP32 + FP64 refinement")
    ap.add_argument("--scaling", default="ruiz", choices=["ruiz", "ruiz+pock-chambolle"],
                    help="zipf-solve workload: the equilibration (Ruiz is the reference's)")
    ap.add_argument("--count", type=int, default=8192)
    ap.add_argument("--mode", default="auto", choices=["auto", "lockstep", "persistent"],
                    help="batched execution mode (igaos_batch_set_mode); auto: "
                         "the autotuner measures both on a probe and picks")
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--diff-time-limit", type=float, default=10.0,
                    help="diff workload: per-instance budget of the forward solve")
    ap.add_argument("--unroll-iterations", type=int, default=1024,
                    help="diff workload: steps of the unrolled-mode measurement")
    ap.add_argument("--time-limit", type=float, default=60.0)
    ap.add_argument("--attempts", type=int, default=20)
    ap.add_argument("--fp64-sample", type=int, default=16)
    ap.add_argument("--fp64-budget-s", type=float, default=30.0,
                    help="batch workload: wall-clock budget of the plain FP64 sample")
    ap.add_argument("--cpu-time-limit", type=float, default=10.0)
    ap.add_argument("--cpu-budget-s", type=float, default=45.0,
                    help="batch workload: wall-clock budget of the three CPU legs")
    ap.add_argument("--budget-s", type=float, default=450.0,
                    help="batch workload: wall-clock seconds from start within "
                         "which warm-up and timed steps are run (at least one "
                         "timed step always runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device to device (a known-size "
                         "kernel for scaling rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the engine runs on one GPU; sharding a batch "
                                   "by instance or one LP by rows across GPUs is "
                                   "not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import igaos_amd
    if not (torch.cuda.is_available() and igaos_amd.has_pdhg):
        raise SystemExit("bench.py needs a HIP GPU")
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    line = {"batch": bench_batch, "zipf": bench_zipf, "zipf-solve": bench_zipf_solve,
            "diff": bench_diff, "c1": bench_c1, "c3": bench_c3}[args.workload](args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
