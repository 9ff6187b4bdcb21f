"""Benchmark of the MI355X PDHG engine (one JSON line on stdout).

Default workload: the large badly scaled LP — n = 1,000,000 columns,
m = 400,000 rows (half =, half <=), column counts clip(Zipf(1.5), 2, 2000)
taken literally (68.5M nonzeros), values U[-1,1]*10^U[-3,3], bounds
[0, U[1,100]], seed 0 (igaos_amd.lpgen.gen_zipf_badly_scaled).

A step is one full solve to relative KKT 1e-6 (status near-optimal) with
the FP32 + FP64-refinement engine, starting from x = y = 0 with the model
already resident in HBM (upload, Ruiz scaling and the power iteration run
once, untimed, as in the reference's solve_time_ms, src/pdhg/pdhg.cu:692).
``value`` is LPs solved per second.  The plain FP64 engine's solve of the
same LP is timed outside the timed region and reported beside it.

``roofline`` is taken from the dominant kernel of an attempt, timed with
HIP events on the engine's stream (igaos_pdhg_profile_attempts).  Its
algorithmic bytes per launch follow the per-iteration byte model of
B = 2 nnz (w_a + 4) + 4 (m + n + 2) + w (7n + 8m), split per kernel:
  k_primal (A'y + primal step): nnz (w_a + 4) + 4 (n + 1) + 7 n w
  k_dual   (A xbar + dual step): nnz (w_a + 4) + 4 (m + 1) + 8 m w
with w_a, w the value / vector widths (4 for FP32, 8 for FP64).
``cpu_baseline`` is the oracle (a single-threaded C host replay of the
reference engine) timed on a bounded sample: setup plus 64 iterations of
the same LP, extrapolated to solves/s at the GPU solve's iteration count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("LPs solved/sec to 1e-6 rel KKT (batched FP32+refine vs plain "
          "FP64); HBM GB/s")
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name):
    from igaos_amd import lpgen
    if name == "zipf":
        return lpgen.gen_zipf_badly_scaled(seed=0), dict(
            workload="badly scaled Zipf LP, m=400000 n=1000000 seed 0",
            m=400000, n=1000000)
    if name == "random_eq":
        return lpgen.gen_random_eq(seed=0), dict(
            workload="random sparse equality LP, m=1000 n=3000 seed 0",
            m=1000, n=3000)
    if name == "multiperiod":
        return lpgen.gen_multiperiod(seed=0), dict(
            workload="multi-period blending LP, 96 periods of 5000x8000 seed 0",
            m=480000, n=768000)
    raise SystemExit(f"unknown workload {name}")


def kernel_bytes(lp, w_a, w):
    nnz = lp.nnz
    primal = nnz * (w_a + 4) + 4 * (lp.n + 1) + 7 * lp.n * w
    dual = nnz * (w_a + 4) + 4 * (lp.m + 1) + 8 * lp.m * w
    return primal, dual


def roofline_of(lp, eng, precision, attempts):
    t = eng.profile_attempts(attempts)
    w = 4 if precision == 32 else 8
    bp, bd = kernel_bytes(lp, w, w)
    gp = bp / (t["ms_primal"] * 1e-3) / 1e9
    gd = bd / (t["ms_dual"] * 1e-3) / 1e9
    if t["ms_dual"] >= t["ms_primal"]:
        name, ms, b, g = "k_dual", t["ms_dual"], bd, gd
    else:
        name, ms, b, g = "k_primal", t["ms_primal"], bp, gp
    return dict(kernel=name, ms=ms, bytes=b, gbs=g, primal_gbs=gp,
                dual_gbs=gd, ms_primal=t["ms_primal"], ms_dual=t["ms_dual"])


def load_traffic(kernel, workload):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(lp, gpu_iters):
    import oracle
    t0 = time.perf_counter()
    r = oracle.solve(lp, time_limit=1e7, tolerance=0.0, max_iterations=64)
    wall = time.perf_counter() - t0
    its_per_s = r.iterations / (r.solve_time_ms * 1e-3)
    return {"value": its_per_s / max(gpu_iters, 1), "unit": "LPs/s",
            "cores": 1, "kind": "port",
            "sample": (f"oracle host replay, 1 thread: setup + 64 iterations "
                       f"({wall:.1f} s wall, {its_per_s:.2f} it/s), "
                       f"extrapolated at the GPU solve's {gpu_iters} "
                       f"iterations")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", default="zipf",
                    choices=["zipf", "random_eq", "multiperiod"])
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--time-limit", type=float, default=300.0)
    ap.add_argument("--attempts", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU "
                          "sharding is a design only (see DESIGN.md)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import igaos_amd
    assert torch.cuda.is_available() and igaos_amd.has_pdhg
    torch.cuda.init()

    t = time.perf_counter()
    lp, cfg = make_workload(args.workload)
    log(f"[bench] workload {cfg['workload']}: nnz={lp.nnz} "
        f"(generated in {time.perf_counter() - t:.1f} s)")

    # warm the HIP runtime on a tiny LP (not a step)
    from igaos_amd import lpgen
    igaos_amd.solve_lp(lpgen.gen_random_eq(seed=2, m=60, n=150),
                       tolerance=1e-6, time_limit=10)

    eng = igaos_amd.Engine(lp, precision=32)
    log(f"[bench] FP32+refine engine ready, setup {eng.setup_ms:.0f} ms")
    for w in range(args.warmup):
        s = eng.run(time_limit=args.time_limit, tolerance=args.tolerance)
        log(f"[bench] warmup {w}: {s.status} {s.iterations} it "
            f"{s.solve_time_ms:.0f} ms")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sols = []
    for k in range(args.steps):
        sols.append(eng.run(time_limit=args.time_limit,
                            tolerance=args.tolerance))
        log(f"[bench] step {k}: {sols[-1].status} {sols[-1].iterations} it "
            f"obj {sols[-1].objective:.10g} {sols[-1].solve_time_ms:.0f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    solved = sum(s.status == "near-optimal" for s in sols)
    ms_per_step = elapsed * 1000.0 / args.steps
    value = solved / elapsed
    s32 = sols[-1]
    rf32 = roofline_of(lp, eng, 32, args.attempts)
    log(f"[bench] FP32 attempt: k_primal {rf32['ms_primal']:.3f} ms "
        f"({rf32['primal_gbs']:.0f} GB/s), k_dual {rf32['ms_dual']:.3f} ms "
        f"({rf32['dual_gbs']:.0f} GB/s)")
    eng.close()

    fp64 = None
    if not args.no_fp64:
        e64 = igaos_amd.Engine(lp, precision=64)
        t64 = time.perf_counter()
        s64 = e64.run(time_limit=args.time_limit, tolerance=args.tolerance)
        t64 = time.perf_counter() - t64
        rf64 = roofline_of(lp, e64, 64, args.attempts)
        e64.close()
        fp64 = dict(status=s64.status, iterations=s64.iterations,
                    seconds=round(t64, 3),
                    solves_per_s=(1.0 / t64) if s64.status == "near-optimal" else 0.0,
                    objective=s64.objective,
                    k_primal_gbs=round(rf64["primal_gbs"], 1),
                    k_dual_gbs=round(rf64["dual_gbs"], 1),
                    hbm_frac_dominant=round(rf64["gbs"] / PEAK_HBM_GBS, 4))
        log(f"[bench] FP64 plain: {s64.status} {s64.iterations} it "
            f"{t64:.2f} s; k_primal {rf64['primal_gbs']:.0f} GB/s, "
            f"k_dual {rf64['dual_gbs']:.0f} GB/s")

    cpu = None if args.no_cpu_baseline else cpu_baseline(lp, s32.iterations)
    traffic = load_traffic(rf32["kernel"], args.workload)
    line = {
        "metric": METRIC, "value": value, "unit": "LPs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64refine",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "m": lp.m, "n": lp.n,
                   "nnz": lp.nnz, "tolerance": args.tolerance,
                   "engine": "FP32 inner + FP64 refinement, device line search, hipGraph attempts"},
        "iterations": s32.iterations,
        "iterations_per_s": s32.iterations / (s32.solve_time_ms * 1e-3),
        "objective": s32.objective,
        "status": s32.status,
        "setup_ms": round(s32.info["setup_ms"], 1),
        "plain_fp64": fp64,
        "fp32_vs_fp64_speedup": (round(fp64["seconds"] / (ms_per_step / 1000), 3)
                                 if fp64 else None),
        "roofline": {"bound": "hbm", "kernel": rf32["kernel"],
                     "achieved": round(rf32["gbs"], 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(rf32["gbs"] / PEAK_HBM_GBS, 4),
                     "traffic": traffic,
                     "bytes_per_launch": rf32["bytes"],
                     "ms_per_launch": round(rf32["ms"], 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
