"""Autotuner of the batched solver's schedule (no reference counterpart).

Three choices are tuned:
  * the SpMV schedule of each size class of instances (classes are
    quantile bins of nonzeros): 0 = row-block streaming through LDS, or
    1, 2, 4 lanes per row;
  * the persistent kernel variant by resident waves per SIMD (2: natural
    register allocation; 4: 128-VGPR cap with some spills, twice the
    resident workgroups);
  * the execution mode: one persistent launch (a workgroup per instance,
    with the two choices above) or lockstep rounds (every active
    instance's row blocks in each launch).

Method: a probe sub-batch (up to ``per_class`` instances of every class) is
run for a fixed number of iterations with tolerance 0, so every instance
does exactly the same work under every candidate.  For lanes per row the
per-instance device time per attempt (measured in-kernel) is compared
class by class; for the variant the probe's total time is compared.  The
mode is compared on a larger probe (up to ``mode_probe`` instances, drawn
evenly from the classes; lockstep rounds fill the GPU only with many
instances in flight) by the wall time of the same fixed-iteration run in
each mode.  The winning schedule is returned as per-instance tpr, the
variant and the mode, and can be cached as JSON keyed by the class
boundaries.
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, asdict

import numpy as np

from .batch import BatchSolver

TPR = (0, 1, 2, 4)
WAVES = (2, 4)


MODES = ("persistent", "lockstep")


@dataclass
class Schedule:
    edges: list          # nnz class boundaries (len = classes + 1)
    tpr: list            # lanes per row per class
    waves: int
    probe_ms: dict       # measured ms per attempt per (waves, class, tpr)
    mode: str = "persistent"
    mode_ms: dict = None  # wall ms of the mode probe per mode

    def tpr_for(self, nnz):
        cls = np.clip(np.searchsorted(self.edges, nnz, side="right") - 1, 0,
                      len(self.tpr) - 1)
        return np.asarray(self.tpr, np.int32)[cls]


def _classes(nnz, k):
    edges = np.unique(np.quantile(nnz, np.linspace(0, 1, k + 1)))
    edges[-1] = edges[-1] + 1
    return edges


def tune(lps, classes=4, per_class=128, iterations=256, log=None,
         cache=None, modes=False, mode_probe=2048) -> Schedule:
    """Measure and return the best schedule for a batch of LPs (and, with
    ``modes``, the faster execution mode)."""
    nnz = np.array([lp.nnz for lp in lps])
    edges = _classes(nnz, classes)
    key = ",".join(str(int(e)) for e in edges) + (";modes" if modes else "")
    if cache and os.path.exists(cache):
        with open(cache) as f:
            got = json.load(f).get(key)
        if got:
            return Schedule(**got)
    cls = np.clip(np.searchsorted(edges, nnz, side="right") - 1, 0,
                  len(edges) - 2)
    rng = np.random.default_rng(0)
    probe = []
    for c in range(len(edges) - 1):
        idx = np.flatnonzero(cls == c)
        probe.extend(rng.choice(idx, min(per_class, idx.size), replace=False))
    probe = np.array(sorted(probe))
    pcls = cls[probe]
    b = BatchSolver([lps[i] for i in probe], precision=32)
    measured = {}
    best_total = None
    best = None
    try:
        for waves in WAVES:
            per = np.zeros((len(edges) - 1, len(TPR)))
            totals = []
            for t_i, t in enumerate(TPR):
                b.set_schedule(np.full(len(probe), t, np.int32), waves=waves)
                s = b.run(tolerance=0.0, time_limit=1e7,
                          max_iterations=iterations, want_xy=False)
                cost = np.asarray(s.solve_ms) / np.maximum(np.asarray(s.attempts), 1)
                for c in range(len(edges) - 1):
                    per[c, t_i] = cost[pcls == c].mean()
                totals.append(s.kernel_ms)
                measured[f"w{waves}_t{t}"] = dict(
                    kernel_ms=s.kernel_ms,
                    ms_per_attempt_by_class=[float(v) for v in per[:, t_i]])
                if log:
                    log(f"[autotune] waves={waves} tpr={t}: probe "
                        f"{s.kernel_ms:.1f} ms")
            pick = [TPR[int(np.argmin(per[c]))] for c in range(len(edges) - 1)]
            # time of the class-wise pick, estimated from the probe runs
            b.set_schedule(np.asarray([pick[c] for c in pcls], np.int32),
                           waves=waves)
            s = b.run(tolerance=0.0, time_limit=1e7, max_iterations=iterations,
                      want_xy=False)
            if log:
                log(f"[autotune] waves={waves} pick {pick}: probe "
                    f"{s.kernel_ms:.1f} ms")
            if best_total is None or s.kernel_ms < best_total:
                best_total = s.kernel_ms
                best = (pick, waves)
    finally:
        b.close()
    sched = Schedule(edges=[int(e) for e in edges], tpr=best[0],
                     waves=best[1], probe_ms=measured)
    if modes:
        _tune_mode(lps, cls, sched, mode_probe, iterations, log)
    if cache:
        d = {}
        if os.path.exists(cache):
            with open(cache) as f:
                d = json.load(f)
        d[key] = asdict(sched)
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(d, f)
    return sched


def _tune_mode(lps, cls, sched, mode_probe, iterations, log):
    """Time the same fixed-iteration run of a probe in both modes."""
    ncls = int(cls.max()) + 1
    rng = np.random.default_rng(1)
    per = max(1, mode_probe // ncls)
    probe = []
    for c in range(ncls):
        idx = np.flatnonzero(cls == c)
        probe.extend(rng.choice(idx, min(per, idx.size), replace=False))
    probe = sorted(int(i) for i in probe)
    sub = [lps[i] for i in probe]
    ms = {}
    for mode in MODES:
        b = BatchSolver(sub, precision=32, mode=mode)
        try:
            if mode == "persistent":
                b.set_schedule(sched.tpr_for(b.nnz), waves=sched.waves)
            b.run(tolerance=0.0, time_limit=1e7, max_iterations=16, want_xy=False)
            t = time.perf_counter()
            b.run(tolerance=0.0, time_limit=1e7, max_iterations=iterations,
                  want_xy=False)
            ms[mode] = (time.perf_counter() - t) * 1e3
        finally:
            b.close()
        if log:
            log(f"[autotune] mode {mode}: {len(sub)} instances x {iterations} "
                f"iterations in {ms[mode]:.1f} ms")
    sched.mode = min(ms, key=ms.get)
    sched.mode_ms = ms


def apply(solver: BatchSolver, sched: Schedule):
    """Set a tuned schedule (and mode, when tuned) on a BatchSolver."""
    if sched.mode_ms and solver.mode != sched.mode:
        solver.set_mode(sched.mode)
    solver.set_schedule(sched.tpr_for(solver.nnz), waves=sched.waves)
