"""Writes tests/golden/adjoint_fd.json: finite-difference fixtures for the
gradients of the differentiable batched solve.

Workload: the 512-LP differentiable batch (igaos_amd.lpgen.gen_diff_batch,
seed 0: m log-uniform in [256, 1024], n = 2m, 6 nonzeros per column, box
[0, 10]), each instance made nondegenerate by
lpgen.perturb_nondegenerate(lp, seed=index, scale=1e-6), with c, b and A's
values then rounded to FP32.  The first 32
instances whose exact solution (scipy HiGHS, tests/highs_util.py) is
nondegenerate -- exactly m columns strictly inside their bounds by more
than 1e-7 and every other reduced cost above 1e-7 in magnitude -- are kept.

Loss: L = w.x* + 0.5 |y*|^2 with w ~ N(0, 1) from default_rng([7, index])
and y* in the PDHG sign convention.  For each of c, b (= rmin = rmax) and
A's values (CSR order) a unit-norm direction u ~ N(0, I) / |.| is drawn
from default_rng([11, index, group, attempt]) and
    fd = (L(theta + h u) - L(theta - h u)) / (2 h),  h = 1e-4,
is recorded, FP64, retrying with the next attempt when the basis at either
end differs from the base basis (the solution map is smooth only inside a
basis).  The objective's own gradients need no fixture: they have closed
forms (dobj/dc = x*, dobj/db = -y*, dobj/dA = y* x*').

Run:  python tests/golden/make_adjoint_fixtures.py   (about 6 minutes)
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from igaos_amd import lpgen  # noqa: E402
from tests.highs_util import basis, highs_xy  # noqa: E402

H = 1e-4
COUNT = 32
GROUPS = ("c", "b", "A")


def instance(index):
    """Instance `index` of gen_diff_batch(seed=0), made nondegenerate, with
    c, b and A's values rounded to FP32 (the differentiable layer takes
    FP32 tensors; the finite differences are taken in FP64 around exactly
    those values)."""
    lp = lpgen.gen_batch_instance(0, index, m_lo=256, m_hi=1024, ratio_lo=2.0,
                                  ratio_hi=2.0, k_lo=6, k_hi=6)
    lp = lpgen.perturb_nondegenerate(lp, seed=index, scale=1e-6)
    f = lambda v: np.asarray(v, np.float32).astype(np.float64)  # noqa
    b = f(lp.rmin)
    return lp.with_values(ax=f(lp.ax), c=f(lp.c), rmin=b, rmax=b.copy())


def weights(index, n):
    return np.random.default_rng([7, index]).standard_normal(n)


def direction(index, group, attempt, size):
    u = np.random.default_rng([11, index, GROUPS.index(group), attempt]).standard_normal(size)
    return u / np.linalg.norm(u)


def perturbed(lp, group, s, u):
    if group == "c":
        return lp.with_values(c=lp.c + s * u)
    if group == "b":
        return lp.with_values(rmin=lp.rmin + s * u, rmax=lp.rmax + s * u)
    return lp.with_values(ax=lp.ax + s * u)


def main():
    out = {"h": H, "loss": "w.x + 0.5|y|^2", "generator_version": lpgen.GENERATOR_VERSION,
           "instances": []}
    t0 = time.time()
    index = 0
    while len(out["instances"]) < COUNT:
        lp = instance(index)
        x, y, obj = highs_xy(lp)
        B = basis(lp, x)
        d = lp.c + lp.to_scipy().T @ y
        ok = B.sum() == lp.m and np.abs(d[~B]).min() > 1e-7
        if not ok:
            print(f"index {index}: degenerate, skipped", flush=True)
            index += 1
            continue
        w = weights(index, lp.n)
        ent = {"index": index, "m": lp.m, "n": lp.n, "nnz": lp.nnz,
               "objective": obj, "loss": float(w @ x + 0.5 * y @ y), "fd": {}}
        for g in GROUPS:
            size = {"c": lp.n, "b": lp.m, "A": lp.nnz}[g]
            for attempt in range(8):
                u = direction(index, g, attempt, size)
                vals = []
                same = True
                for s in (H, -H):
                    xp, yp, _ = highs_xy(perturbed(lp, g, s, u))
                    same &= bool((basis(lp, xp) == B).all())
                    vals.append(float(w @ xp + 0.5 * yp @ yp))
                if same:
                    ent["fd"][g] = {"attempt": attempt,
                                    "value": (vals[0] - vals[1]) / (2 * H)}
                    break
        if len(ent["fd"]) == len(GROUPS):
            out["instances"].append(ent)
        print(f"index {index}: m={lp.m} kept={len(ent['fd']) == 3} "
              f"({len(out['instances'])}/{COUNT}, {time.time() - t0:.0f} s)", flush=True)
        index += 1
    with open(os.path.join(HERE, "adjoint_fd.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
