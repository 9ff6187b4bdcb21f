"""Autotuner of the batched solver's schedule (no reference counterpart).

Two choices are tuned:
  * the SpMV schedule of each size class of instances (classes are
    quantile bins of nonzeros): 0 = row-block streaming through LDS, or
    1, 2, 4 lanes per row;
  * the persistent kernel variant by resident waves per SIMD (2: natural
    register allocation; 4: 128-VGPR cap with some spills, twice the
    resident workgroups).

Method: a probe sub-batch (up to ``per_class`` instances of every class) is
run for a fixed number of iterations with tolerance 0, so every instance
does exactly the same work under every candidate.  For lanes per row the
per-instance device time per attempt (measured in-kernel) is compared
class by class; for the variant the probe's total time is compared.  The
winning schedule is returned as per-instance tpr plus the variant and can
be cached as JSON keyed by the class boundaries.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, asdict

import numpy as np

from .batch import BatchSolver

TPR = (0, 1, 2, 4)
WAVES = (2, 4)


@dataclass
class Schedule:
    edges: list          # nnz class boundaries (len = classes + 1)
    tpr: list            # lanes per row per class
    waves: int
    probe_ms: dict       # measured ms per attempt per (waves, class, tpr)

    def tpr_for(self, nnz):
        cls = np.clip(np.searchsorted(self.edges, nnz, side="right") - 1, 0,
                      len(self.tpr) - 1)
        return np.asarray(self.tpr, np.int32)[cls]


def _classes(nnz, k):
    edges = np.unique(np.quantile(nnz, np.linspace(0, 1, k + 1)))
    edges[-1] = edges[-1] + 1
    return edges


def tune(lps, classes=4, per_class=128, iterations=256, log=None,
         cache=None) -> Schedule:
    """Measure and return the best schedule for a batch of LPs."""
    nnz = np.array([lp.nnz for lp in lps])
    edges = _classes(nnz, classes)
    key = ",".join(str(int(e)) for e in edges)
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


def apply(solver: BatchSolver, sched: Schedule):
    """Set a tuned schedule on a BatchSolver."""
    solver.set_schedule(sched.tpr_for(solver.nnz), waves=sched.waves)
