"""Writes tests/golden/objectives.json: objective fixtures for the tests.

Sources, per entry:
  * ``highs_ref``   — the reference's own recorded HiGHS objectives for the
                      Netlib spike set (spike/pdhg-spike/highs_results.csv,
                      copied verbatim to tests/golden/netlib/);
  * ``ref_simplex`` — the reference's revised simplex
                      (src/simplex/simplex.cpp) built from its sources into
                      oracle/_ref/ by oracle/Makefile, run in this container;
  * ``scipy_highs`` — scipy.optimize.linprog(method="highs") as an
                      independent check.
Generated LPs are named by generator, parameters and seed (see
igaos_amd/lpgen.py, GENERATOR_VERSION).

Run:  python tests/golden/make_fixtures.py   (about 3 minutes)
      python tests/golden/make_fixtures.py --add   (the EXTRA and LARGE
      entries the existing file lacks, added to it)

LARGE entries use HiGHS's interior-point method with crossover
(method="highs-ipm"): its default simplex needs over 15 minutes on the
8-period LP.  On the same family at 2 periods of 1,000 x 1,600 both
methods return the same objective to the last bit (-399.9303203719868).
"""
from __future__ import annotations

import csv
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from igaos_amd import lpgen  # noqa: E402

import oracle  # noqa: E402


def scipy_highs(lp, method="highs"):
    import scipy.sparse as sp
    from scipy.optimize import linprog
    A = sp.csr_matrix((lp.ax, lp.ai, lp.ap), shape=(lp.m, lp.n))
    eq = lp.rmin == lp.rmax
    fin_hi = np.isfinite(lp.rmax) & ~eq
    fin_lo = np.isfinite(lp.rmin) & ~eq
    A_ub = sp.vstack([A[fin_hi], -A[fin_lo]]).tocsr()
    b_ub = np.concatenate([lp.rmax[fin_hi], -lp.rmin[fin_lo]])
    bounds = [(None if not np.isfinite(l) else l, None if not np.isfinite(u) else u)
              for l, u in zip(lp.cl, lp.cu)]
    kw = {}
    if A_ub.shape[0]:
        kw.update(A_ub=A_ub, b_ub=b_ub)
    if eq.any():
        kw.update(A_eq=A[eq], b_eq=lp.rmin[eq])
    r = linprog(lp.c, bounds=bounds, method=method, **kw)
    return (float(r.fun) + lp.obj_const) if r.status == 0 else None


# (generator, parameters): solved with HiGHS's interior-point method
LARGE = [("gen_multiperiod", dict(seed=0, periods=8, rows_per=1000, cols_per=1600, inv=200))]
# small cases added after the first generation (also part of a full one)
EXTRA = [("gen_ragged", dict(seed=5))]


def key_of(fn, kw):
    return fn + "(" + ",".join(f"{k}={v}" for k, v in kw.items()) + ")"


def small_entry(lp, have_ref):
    ent = {"m": lp.m, "n": lp.n, "nnz": lp.nnz, "scipy_highs": scipy_highs(lp)}
    if have_ref and lp.m * lp.n <= 3_000_000:
        t = time.time()
        st, obj, _, _, its, ms = oracle.ref_simplex(lp, time_limit=600)
        ent.update(ref_simplex=obj, ref_simplex_status=st, ref_simplex_iterations=its,
                   ref_simplex_seconds=time.time() - t)
    return ent


def add_extra(E, have_ref):
    for fn, kw in EXTRA:
        if key_of(fn, kw) not in E:
            E[key_of(fn, kw)] = small_entry(getattr(lpgen, fn)(**kw), have_ref)
            print(key_of(fn, kw), E[key_of(fn, kw)], flush=True)


def add_large(E):
    for fn, kw in LARGE:
        if key_of(fn, kw) in E:
            continue
        lp = getattr(lpgen, fn)(**kw)
        t = time.time()
        E[key_of(fn, kw)] = {"m": lp.m, "n": lp.n, "nnz": lp.nnz,
                             "scipy_highs": scipy_highs(lp, method="highs-ipm"),
                             "scipy_highs_method": "highs-ipm",
                             "scipy_highs_seconds": time.time() - t}
        print(key_of(fn, kw), E[key_of(fn, kw)], flush=True)


def main():
    if "--add" in sys.argv:
        path = os.path.join(HERE, "objectives.json")
        with open(path) as f:
            out = json.load(f)
        add_extra(out["entries"], oracle.have_ref_simplex())
        add_large(out["entries"])
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
        return
    out = {"generator_version": lpgen.GENERATOR_VERSION, "entries": {}}
    E = out["entries"]
    with open(os.path.join(HERE, "netlib", "highs_results.csv")) as f:
        for row in csv.reader(f):
            E[f"netlib/{row[0]}"] = {"highs_ref": float(row[2]),
                                     "source": "spike/pdhg-spike/highs_results.csv"}
    from igaos_amd.api import read_mps
    for name in list(E):
        lp = read_mps(os.path.join(HERE, name + ".mps")).to_lp()
        E[name]["scipy_highs"] = scipy_highs(lp)
        E[name].update(m=lp.m, n=lp.n, nnz=lp.nnz)
    have_ref = oracle.have_ref_simplex()
    cases = [("gen_random_eq", dict(seed=0)),
             ("gen_random_eq", dict(seed=1, m=200, n=600)),
             ("gen_random_eq", dict(seed=2, m=60, n=150))]
    cases += [("gen_batch_instance", dict(seed=0, index=i)) for i in range(16)]
    cases += [("gen_batch_instance", dict(seed=0, index=i, m_lo=256, m_hi=1024,
                                          ratio_lo=2.0, ratio_hi=2.0, k_lo=6,
                                          k_hi=6)) for i in range(8)]
    cases += [("gen_multiperiod", dict(seed=0, periods=2, rows_per=200,
                                       cols_per=320, inv=40))]
    cases += [("gen_zipf_badly_scaled", dict(seed=0, m=400, n=1000))]
    for fn, kw in cases:
        E[key_of(fn, kw)] = small_entry(getattr(lpgen, fn)(**kw), have_ref)
        print(key_of(fn, kw), E[key_of(fn, kw)], flush=True)
    add_extra(E, have_ref)
    add_large(E)
    with open(os.path.join(HERE, "objectives.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
