"""Seeded synthetic LP generators and the host-side LP container.

The container mirrors the reference's ``io::Model`` (src/io/model.hpp:16-37):
``min c'x + obj_const  s.t.  rmin <= A x <= rmax,  cl <= x <= cu`` with the
matrix stored twice, as CSR (``ap/ai/ax``) and CSC (``cp/ci/acx``), FP64
values, int32 indices and +-inf for missing bounds.  ``LP.from_triplets``
follows the reference's ``counting_sort`` (src/io/model.hpp:39-67): entries
keep their input order inside each row and inside each column.

Every generator draws from ``numpy.random.default_rng`` in a fixed, documented
order so that a (generator, seed) pair names one LP exactly.  Rows of the
sampled columns are distinct within a column and sorted ascending, and
triplets are emitted column by column, so CSR rows come out column-sorted.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

INF = float("inf")

GENERATOR_VERSION = 1


@dataclass
class LP:
    m: int
    n: int
    c: np.ndarray
    cl: np.ndarray
    cu: np.ndarray
    rmin: np.ndarray
    rmax: np.ndarray
    ap: np.ndarray
    ai: np.ndarray
    ax: np.ndarray
    cp: np.ndarray
    ci: np.ndarray
    acx: np.ndarray
    obj_const: float = 0.0
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.ax.shape[0])

    @staticmethod
    def from_triplets(m, n, rows, cols, vals, c, cl, cu, rmin, rmax,
                      obj_const=0.0, name=""):
        """Build CSR + CSC from triplets in their given order.

        Mirrors ``counting_sort`` (src/io/model.hpp:39-67): a stable bucket
        sort by row gives CSR and a stable bucket sort by column gives CSC.
        """
        rows = np.asarray(rows, dtype=np.int64)
        cols = np.asarray(cols, dtype=np.int64)
        vals = np.asarray(vals, dtype=np.float64)
        ap = np.zeros(m + 1, dtype=np.int64)
        ap[1:] = np.bincount(rows, minlength=m)
        ap = np.cumsum(ap)
        cp = np.zeros(n + 1, dtype=np.int64)
        cp[1:] = np.bincount(cols, minlength=n)
        cp = np.cumsum(cp)
        if ap[-1] >= 2**31 - 1:
            raise ValueError("nnz exceeds the int32 index range")
        r_order = np.argsort(rows, kind="stable")
        c_order = np.argsort(cols, kind="stable")
        return LP(
            m=int(m), n=int(n),
            c=np.ascontiguousarray(c, dtype=np.float64),
            cl=np.ascontiguousarray(cl, dtype=np.float64),
            cu=np.ascontiguousarray(cu, dtype=np.float64),
            rmin=np.ascontiguousarray(rmin, dtype=np.float64),
            rmax=np.ascontiguousarray(rmax, dtype=np.float64),
            ap=ap.astype(np.int32), ai=cols[r_order].astype(np.int32),
            ax=np.ascontiguousarray(vals[r_order]),
            cp=cp.astype(np.int32), ci=rows[c_order].astype(np.int32),
            acx=np.ascontiguousarray(vals[c_order]),
            obj_const=float(obj_const), name=name)

    def csc_perm(self):
        """perm with acx == ax[perm]: the CSR position of every CSC entry
        (the pattern holds each (row, column) pair once)."""
        rows = np.repeat(np.arange(self.m, dtype=np.int64), np.diff(self.ap))
        cols = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.cp))
        key_r = rows * self.n + self.ai
        key_c = self.ci.astype(np.int64) * self.n + cols
        o_r = np.argsort(key_r, kind="stable")
        o_c = np.argsort(key_c, kind="stable")
        perm = np.empty(self.nnz, dtype=np.int64)
        perm[o_c] = o_r
        return perm

    def with_values(self, ax=None, c=None, rmin=None, rmax=None):
        """Copy with new numeric data on the same pattern (ax in CSR
        order; the CSC values follow)."""
        import copy
        q = copy.copy(self)
        if ax is not None:
            q.ax = np.ascontiguousarray(ax, dtype=np.float64)
            q.acx = np.ascontiguousarray(q.ax[self.csc_perm()])
        if c is not None:
            q.c = np.ascontiguousarray(c, dtype=np.float64)
        if rmin is not None:
            q.rmin = np.ascontiguousarray(rmin, dtype=np.float64)
        if rmax is not None:
            q.rmax = np.ascontiguousarray(rmax, dtype=np.float64)
        return q

    def matvec(self, x):
        """Row-sequential A @ x in FP64 (host helper for generators)."""
        x = np.asarray(x, dtype=np.float64)
        prod = self.ax * x[self.ai]
        out = np.zeros(self.m)
        if self.nnz:
            # reduceat over the rows up to the last non-empty one (a start
            # index equal to nnz is out of its range); empty rows are 0
            last = int(np.searchsorted(self.ap[:-1], self.nnz, side="left"))
            out[:last] = np.add.reduceat(prod, self.ap[:last])
            empty = self.ap[1:] == self.ap[:-1]
            out[empty] = 0.0
        return out

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.ax, self.ai, self.ap), shape=(self.m, self.n))


def _distinct_rows(rng, m, counts):
    """Per-column distinct uniform row indices, ascending within a column.

    Draws ``sum(counts)`` uniform rows at once, then redraws only the
    duplicated slots until every column's rows are distinct.
    """
    n = counts.shape[0]
    if np.any(counts > m):
        raise ValueError("a column asks for more distinct rows than m")
    total = int(counts.sum())
    cols = np.repeat(np.arange(n, dtype=np.int64), counts)
    rows = rng.integers(0, m, size=total, dtype=np.int64)
    while True:
        key = cols * m + rows
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(total, dtype=bool)
        dup[order[1:][ks[1:] == ks[:-1]]] = True
        nd = int(dup.sum())
        if nd == 0:
            break
        rows[dup] = rng.integers(0, m, size=nd, dtype=np.int64)
    key = cols * m + rows
    key.sort()
    return key // m, key % m  # (cols, rows) column-major, rows ascending


def gen_random_eq(seed=0, m=1000, n=3000, per_col=5, box=10.0, name=None):
    """Random sparse equality LP (the m=1000, n=3000 workload by default).

    Draw order: rows (per column, distinct), values N(0,1), x0 ~ U[0,1],
    c ~ N(0,1).  Rows are equalities with b = A x0; bounds 0 <= x <= box.
    """
    rng = np.random.default_rng(seed)
    counts = np.full(n, per_col, dtype=np.int64)
    cols, rows = _distinct_rows(rng, m, counts)
    vals = rng.standard_normal(cols.shape[0])
    x0 = rng.random(n)
    c = rng.standard_normal(n)
    lp = LP.from_triplets(m, n, rows, cols, vals, c, np.zeros(n),
                          np.full(n, box), np.zeros(m), np.zeros(m),
                          name=name or f"rand_eq_m{m}_n{n}_s{seed}")
    b = lp.matvec(x0)
    lp.rmin[:] = b
    lp.rmax[:] = b
    return lp


def gen_ragged(seed=5, m=700, n=1500, name=None):
    """Small LP with the structural edge cases of sparse kernels: two empty
    rows and three empty columns, one row and one column far longer than the
    rest (every third column; every second row), ranged, one-sided and
    free-below rows, and free, one-sided and fixed columns.  Rows: b = A x0
    with x0 = clip(5 U[0,1], bounds); the costs of columns without an upper
    bound are |N(0,1)| and free columns cost 0, so the LP is bounded."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for j in range(n - 3):  # columns n-3 .. n-1 stay empty
        k = 1 + int(rng.integers(0, 7))
        rows.extend(rng.choice(m - 2, size=k, replace=False))  # rows m-2, m-1 stay empty
        cols.extend([j] * k)
    rows.extend(range(0, m - 2, 2)); cols.extend([n - 4] * len(range(0, m - 2, 2)))  # long column
    long_cols = [c for c in range(0, n - 4, 3) if c != n - 4]
    rows.extend([1] * len(long_cols)); cols.extend(long_cols)  # long row
    trip = {}
    for r, c in zip(rows, cols):
        trip[(int(r), int(c))] = rng.standard_normal()
    keys = sorted(trip)
    rr = np.array([k[0] for k in keys]); cc = np.array([k[1] for k in keys])
    vv = np.array([trip[k] for k in keys])
    cl = np.zeros(n); cu = np.full(n, 10.0)
    cl[::11] = -np.inf; cu[::11] = np.inf          # free columns
    cu[5::13] = np.inf                              # one-sided
    cl[7::17] = cu[7::17] = 2.0                     # fixed
    cl[n - 3:] = 0.0; cu[n - 3:] = 4.0              # the empty columns boxed
    c = rng.standard_normal(n)
    c[cu == np.inf] = np.abs(c[cu == np.inf])  # bounded below: no improving ray
    c[cl == -np.inf] = 0.0
    x0 = np.clip(rng.random(n) * 5, cl, cu)
    lp = LP.from_triplets(m, n, rr, cc, vv, c, cl, cu, np.zeros(m), np.zeros(m),
                          name=name or f"ragged_m{m}_n{n}_s{seed}")
    b = lp.matvec(x0)
    lp.rmin[:] = b; lp.rmax[:] = b
    lp.rmin[::5] = b[::5] - 1.0                     # ranged rows
    lp.rmax[3::7] = np.inf                          # one-sided rows
    lp.rmin[4::9] = -np.inf; lp.rmax[4::9] = b[4::9] + 0.5
    lp.rmin[m - 2:] = -1.0; lp.rmax[m - 2:] = 1.0   # the empty rows
    return lp


def gen_zipf_badly_scaled(seed=0, m=400_000, n=1_000_000, alpha=1.5,
                          kmin=2, kmax=2000, name=None):
    """Large badly scaled LP: Zipf column counts, values spanning 1e+-3.

    Column counts are ``clip(zipf(alpha), kmin, kmax)`` taken literally
    (mean about 68.5 at alpha=1.5, so about 68.5M nonzeros at n=1M).
    Draw order: counts, rows, value signs/magnitudes U[-1,1]*10^U[-3,3],
    x0 ~ U[0,1], upper bounds U[1,100], c ~ N(0,1), slacks U[0,1].
    The first m//2 rows are equalities b = A x0; the rest are
    ``A x <= A x0 + slack``.
    """
    rng = np.random.default_rng(seed)
    counts = np.clip(rng.zipf(alpha, n), kmin, min(kmax, m)).astype(np.int64)
    cols, rows = _distinct_rows(rng, m, counts)
    nnz = cols.shape[0]
    vals = rng.uniform(-1.0, 1.0, nnz) * 10.0 ** rng.uniform(-3.0, 3.0, nnz)
    x0 = rng.random(n)
    cu = rng.uniform(1.0, 100.0, n)
    c = rng.standard_normal(n)
    meq = m // 2
    slack = rng.random(m - meq)
    lp = LP.from_triplets(m, n, rows, cols, vals, c, np.zeros(n), cu,
                          np.zeros(m), np.zeros(m),
                          name=name or f"zipf_m{m}_n{n}_s{seed}")
    ax0 = lp.matvec(x0)
    lp.rmin[:meq] = ax0[:meq]
    lp.rmax[:meq] = ax0[:meq]
    lp.rmin[meq:] = -INF
    lp.rmax[meq:] = ax0[meq:] + slack
    return lp


def gen_multiperiod(seed=0, periods=96, rows_per=5000, cols_per=8000,
                    inv=1000, per_col=6, box=10.0, name=None):
    """Block-structured multi-period blending LP.

    Each period owns a ``rows_per x cols_per`` block with ``per_col``
    nonzeros per column, values U[0.5,1.5].  The last ``inv`` columns of a
    period are inventory variables: inventory ``k`` of period ``p`` takes
    -1 in row ``k`` of its own block and +1 in row ``k`` of period p+1
    (its random block entries avoid rows ``< inv``).  All rows are
    equalities with b = A x0, x0 ~ U[0,1]; bounds [0, box]; c ~ U[-1,1].
    Draw order: block rows, block values, x0, c.
    """
    rng = np.random.default_rng(seed)
    n = periods * cols_per
    m = periods * rows_per
    # regular columns draw from the whole block, inventory from rows >= inv
    reg = cols_per - inv
    creg, rreg = _distinct_rows(rng, rows_per,
                                np.full(periods * reg, per_col, np.int64))
    cinv, rinv = _distinct_rows(rng, rows_per - inv,
                                np.full(periods * inv, per_col, np.int64))
    rinv = rinv + inv
    p_reg = creg // reg
    j_reg = p_reg * cols_per + creg % reg
    i_reg = p_reg * rows_per + rreg
    p_inv = cinv // inv
    k_inv = cinv % inv
    j_inv = p_inv * cols_per + reg + k_inv
    i_inv = p_inv * rows_per + rinv
    vals_reg = rng.uniform(0.5, 1.5, creg.shape[0])
    vals_inv = rng.uniform(0.5, 1.5, cinv.shape[0])
    kk = np.arange(inv, dtype=np.int64)
    pp = np.arange(periods, dtype=np.int64)
    # -1 in own block row k, +1 in next block row k
    j_out = (pp[:, None] * cols_per + reg + kk[None, :]).ravel()
    i_out = (pp[:, None] * rows_per + kk[None, :]).ravel()
    pn = pp[:-1]
    j_in = (pn[:, None] * cols_per + reg + kk[None, :]).ravel()
    i_in = ((pn[:, None] + 1) * rows_per + kk[None, :]).ravel()
    rows = np.concatenate([i_reg, i_inv, i_out, i_in])
    cols = np.concatenate([j_reg, j_inv, j_out, j_in])
    vals = np.concatenate([vals_reg, vals_inv, -np.ones(i_out.shape[0]),
                           np.ones(i_in.shape[0])])
    order = np.lexsort((rows, cols))  # column-major, rows ascending
    rows, cols, vals = rows[order], cols[order], vals[order]
    x0 = rng.random(n)
    c = rng.uniform(-1.0, 1.0, n)
    lp = LP.from_triplets(m, n, rows, cols, vals, c, np.zeros(n),
                          np.full(n, box), np.zeros(m), np.zeros(m),
                          name=name or f"multiperiod_P{periods}_s{seed}")
    b = lp.matvec(x0)
    lp.rmin[:] = b
    lp.rmax[:] = b
    return lp


def gen_batch_instance(seed, index, m_lo=32, m_hi=4096, ratio_lo=1.5,
                       ratio_hi=3.0, k_lo=4, k_hi=8, box=10.0):
    """One member of a batch of independent random equality LPs.

    Instance ``index`` draws from ``default_rng([seed, index])``:
    m = round(exp(U[ln m_lo, ln m_hi])), n = round(m * U[ratio_lo,
    ratio_hi]), per-column counts uniform in [k_lo, k_hi] (capped at m),
    rows, values N(0,1), x0 ~ U[0,1], c ~ N(0,1); b = A x0; bounds [0, box].
    """
    rng = np.random.default_rng([seed, index])
    m = int(round(np.exp(rng.uniform(np.log(m_lo), np.log(m_hi)))))
    n = int(round(m * rng.uniform(ratio_lo, ratio_hi)))
    counts = np.minimum(rng.integers(k_lo, k_hi + 1, n), m).astype(np.int64)
    cols, rows = _distinct_rows(rng, m, counts)
    vals = rng.standard_normal(cols.shape[0])
    x0 = rng.random(n)
    c = rng.standard_normal(n)
    lp = LP.from_triplets(m, n, rows, cols, vals, c, np.zeros(n),
                          np.full(n, box), np.zeros(m), np.zeros(m),
                          name=f"batch_s{seed}_i{index}")
    b = lp.matvec(x0)
    lp.rmin[:] = b
    lp.rmax[:] = b
    return lp


def gen_batch(seed=0, count=8192, **kw):
    """A batch of independent random LPs (8,192 instances by default)."""
    return [gen_batch_instance(seed, i, **kw) for i in range(count)]


def gen_diff_batch(seed=0, count=512):
    """512 LPs, m log-uniform in [256,1024], n = 2m, 6 nonzeros per column."""
    return [gen_batch_instance(seed, i, m_lo=256, m_hi=1024, ratio_lo=2.0,
                               ratio_hi=2.0, k_lo=6, k_hi=6)
            for i in range(count)]


def perturb_nondegenerate(lp: LP, seed=0, scale=1e-6):
    """Tiny random perturbation of c and b (equality rows) that makes ties
    between vertices and zero reduced costs unlikely."""
    rng = np.random.default_rng(seed)
    lp.c = lp.c + scale * rng.standard_normal(lp.n)
    eq = lp.rmin == lp.rmax
    shift = scale * rng.standard_normal(lp.m)
    lp.rmin = np.where(eq, lp.rmin + shift, lp.rmin)
    lp.rmax = np.where(eq, lp.rmax + shift, lp.rmax)
    return lp
