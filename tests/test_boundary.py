"""The C-ABI boundary, the MPS reader and the host-side API — CPU only.

No compute is launched here: these tests check that libigaos_pdhg.so
loads and exports every entry point include/igaos_pdhg.h declares, that
the model reader keeps the reference's MPS semantics (the edge cases of
tests/engine_smoke.cpp in the reference), that the Python and CLI
surfaces keep the reference's names, defaults and error behaviour, that
malformed caller patterns are refused before device work, the layout
invariants the SpMV kernels rely on, and the host-side pieces of the check
(KKT evaluation, dual repair, row activity) against the oracle and closed
forms.
"""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import igaos_amd
from igaos_amd import _core, lpgen
from tests.conftest import ROOT, netlib_path

HEADER = os.path.join(ROOT, "include", "igaos_pdhg.h")
CLI = os.path.join(ROOT, "igaos_amd", "igaos")
INF = math.inf


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(igaos_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _core.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the ctypes table covers them all with signatures
    from igaos_amd import _batch_sigs
    table = set(_core.SIGNATURES) | set(_batch_sigs.SIGNATURES)
    assert set(syms) <= table, sorted(set(syms) - table)


def test_abi_and_status_names():
    L = _core.lib()
    assert L.igaos_pdhg_abi_version() == 4
    names = [L.igaos_status_name(i).decode() for i in range(8)]
    assert names == ["optimal", "near-optimal", "feasible", "infeasible",
                     "unbounded", "iteration-limit", "time-limit", "error"]
    assert [s.name for s in igaos_amd.Status] == [
        "Optimal", "NearOptimal", "Feasible", "Infeasible", "Unbounded",
        "IterationLimit", "TimeLimit", "Error"]


def test_option_defaults_match_reference():
    import ctypes as C
    o = _core.Options()
    _core.lib().igaos_options_default(C.byref(o))
    # include/igaos/options.h:5-13 of the reference
    assert (o.time_limit_s, o.tolerance, o.max_iterations, o.seed,
            o.presolve, o.restarts, o.verbosity) == (120.0, 1e-4, 500000, 0,
                                                     1, 0, 0)


def _mps(tmp_path, text):
    p = tmp_path / "t.mps"
    p.write_text(text)
    return igaos_amd.read_mps(str(p))


def test_mps_negative_up_frees_lower_bound(tmp_path):
    m = _mps(tmp_path, "NAME T1\nROWS\n N obj\n G r1\nCOLUMNS\n x obj 1.0 r1 1.0\n"
                       "RHS\n RHS1 r1 -5.0\nBOUNDS\n UP BND x -3.0\nENDATA\n")
    a = m.arrays()
    assert a["cl"][0] == -INF and a["cu"][0] == -3.0 and a["rmin"][0] == -5.0


def test_mps_mi_ranges_objconst_fr(tmp_path):
    m = _mps(tmp_path, "NAME T2\nROWS\n N obj\n G r1\nCOLUMNS\n x obj 1.0 r1 1.0\n"
                       "RHS\n RHS1 r1 -2.0\nBOUNDS\n MI BND x\nENDATA\n")
    a = m.arrays()
    assert a["cl"][0] == -INF and a["cu"][0] == INF
    m = _mps(tmp_path, "NAME T3\nROWS\n N obj\n G r1\nCOLUMNS\n x obj 1.0 r1 1.0\n"
                       "RHS\n RHS1 obj -7.113 r1 2.0\nRANGES\n RNG r1 3.0\nENDATA\n")
    a = m.arrays()
    assert m.n_ranges_parsed == 1 and abs(m.obj_const - 7.113) < 1e-12
    assert a["rmin"][0] == 2.0 and a["rmax"][0] == 5.0
    m = _mps(tmp_path, "NAME T4\nROWS\n N obj\n E r1\n G r2\nCOLUMNS\n x obj 1.0 r1 1.0\n"
                       " x r2 0.0\n y obj 2.0 r1 1.0\n y r2 1.0\nRHS\n RHS1 r1 4.0 r2 1.0\n"
                       "BOUNDS\n FR BND y\nENDATA\n")
    assert m.n_fr_parsed == 1 and m.nnz() == 4
    a = m.arrays()
    assert a["cl"][1] == -INF and a["cu"][1] == INF


def test_mps_ranges_sign_rules_and_markers(tmp_path):
    m = _mps(tmp_path,
             "NAME T\nROWS\n N obj\n E e1\n E e2\n L l1\n G g1\nCOLUMNS\n"
             " M1 'MARKER' 'INTORG'\n x obj 1.0 e1 1.0\n M2 'MARKER' 'INTEND'\n"
             " x e2 1.0 l1 1.0\n x g1 1.0\nRHS\n R e1 1.0 e2 1.0 l1 4.0 g1 -1.0\n"
             "RANGES\n RR e1 2.0 e2 -2.0\n RR l1 -3.0 g1 -3.0\nBOUNDS\n"
             " BV B x\nENDATA\n")
    a = m.arrays()
    assert list(a["rmin"]) == [1.0, -1.0, 1.0, -1.0]
    assert list(a["rmax"]) == [3.0, 1.0, 4.0, 2.0]
    assert m.n_int == 1 and a["cu"][0] == 1.0


def test_mps_tokens_numbers_and_layout(tmp_path):
    # the reference tokenises with istringstream (classic blanks) and
    # converts with std::stod: '+', hexadecimal and zero values parse as
    # stod parses them; CRLF lines, a missing final newline, a column
    # whose entries are not contiguous and a repeated row name read as
    # the reference reads them
    text = ("NAME T\r\nROWS\r\n N obj\r\n E r1\r\n L r2\r\n G r1\r\nCOLUMNS\r\n"
            " x obj +2.5 r1 0x1p3\r\n y\tobj -0.0 r2 1e2\r\n x r2 0\r\n"
            " y r1 -.5\r\nRHS\r\n B r1 +1 r2 4.\r\nBOUNDS\r\n UP BND y 1.5E1\r\nENDATA")
    m = _mps(tmp_path, text)
    assert (m.m, m.n) == (3, 2)
    a = m.arrays()
    # r1 names the G row (index 2); the E row keeps its sense, loses its name
    assert list(a["rmin"]) == [0.0, -INF, 1.0] and list(a["rmax"]) == [0.0, 4.0, INF]
    assert list(a["c"]) == [2.5, 0.0] and a["cu"][1] == 15.0
    lp = m.to_lp()
    A = lp.to_scipy().toarray()
    assert A.tolist() == [[0.0, 0.0], [0.0, 100.0], [8.0, -0.5]]
    assert lp.nnz == 4 and list(lp.ai) == [1, 0, 0, 1]   # explicit zero kept
    # stod's range error (a subnormal) and an unknown BOUNDS column are errors
    for bad in (text.replace("1e2", "1e-310"), text.replace("BND y", "BND z")):
        with pytest.raises(RuntimeError):
            _mps(tmp_path, bad)


def test_mps_afiro_stats_and_csc_is_transpose():
    m = igaos_amd.read_mps(netlib_path("afiro"))
    assert (m.m, m.n, m.nnz()) == (27, 32, 83)
    lp = m.to_lp()
    A = lp.to_scipy()
    import scipy.sparse as sp
    At = sp.csr_matrix((lp.acx, lp.ci, lp.cp), shape=(lp.n, lp.m))
    assert abs(A.T - At).max() == 0


def test_triplet_order_matches_counting_sort():
    import ctypes as C
    rng = np.random.default_rng(0)
    m, n, z = 7, 9, 30
    rows = rng.integers(0, m, z).astype(np.int32)
    cols = rng.integers(0, n, z).astype(np.int32)
    vals = rng.standard_normal(z)
    ref = lpgen.LP.from_triplets(m, n, rows, cols, vals, np.zeros(n),
                                 np.zeros(n), np.ones(n), np.zeros(m),
                                 np.ones(m))
    dp = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))  # noqa
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa
    h = _core.lib().igaos_model_from_triplets(
        m, n, z, ip(rows), ip(cols), dp(vals), dp(np.zeros(n)),
        dp(np.zeros(n)), dp(np.ones(n)), dp(np.zeros(m)), dp(np.ones(m)), 0.0)
    a = igaos_amd.Model(h).arrays()
    for k in ("ap", "ai", "ax", "cp", "ci", "acx"):
        assert np.array_equal(a[k], getattr(ref, k)), k


def test_invalid_engine_raises_and_other_engines_report_error():
    with pytest.raises(ValueError):
        igaos_amd.solve(netlib_path("afiro"), engine="bogus")
    s = igaos_amd.solve(netlib_path("afiro"), engine="simplex")
    assert s.status == "error" and "not part of this build" in s.message


@pytest.mark.skipif(_core.device_count() > 0, reason="host has a GPU")
def test_no_gpu_gives_honest_error():
    # the reference's binding test expects 'CUDA' in the message of a
    # build without the GPU engine (demo/test_bindings.py:84-86)
    s = igaos_amd.solve(netlib_path("afiro"), engine="pdhg")
    assert s.status == "error" and "CUDA" in s.message
    assert not igaos_amd.has_pdhg


def test_cli_info_and_non_pdhg_engine():
    out = subprocess.run([CLI, "info", netlib_path("afiro")], check=True,
                         capture_output=True, text=True).stdout
    assert "rows           27" in out and "nnz            83" in out
    assert "row types      E=8 L=19 G=0 ranged=0" in out
    r = subprocess.run([CLI, "solve", netlib_path("afiro"), "--engine", "qp"],
                       capture_output=True, text=True)
    import json
    js = json.loads(r.stdout)
    assert js["status"] == "error" and js["engine"] == "qp"
    assert subprocess.run([CLI, "solve", netlib_path("afiro"), "--engine",
                           "bogus"], capture_output=True).returncode == 2


def _rowblocks(ptr):
    import ctypes as C
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    L = _core.lib()
    k = L.igaos_rowblocks_build(len(ptr) - 1, ptr.ctypes.data_as(C.POINTER(C.c_int)), None, 0)
    out = np.zeros(k, np.int32)
    L.igaos_rowblocks_build(len(ptr) - 1, ptr.ctypes.data_as(C.POINTER(C.c_int)),
                            out.ctypes.data_as(C.POINTER(C.c_int)), k)
    return out


def test_generator_row_products_with_empty_rows():
    # lpgen's row-sequential A x (the generators' right-hand sides) when the
    # last rows are empty (a reduceat start at nnz) and when an empty row
    # sits inside; and the ragged LP holds the shapes its docstring names
    lp = lpgen.gen_ragged(seed=5)
    A = lp.to_scipy().tocsr()
    x = np.random.default_rng(0).random(lp.n)
    ax = A @ x
    y = lp.matvec(x)
    assert np.abs(y - ax).max() <= 1e-12 * (1 + np.abs(ax).max())
    rl, cl = np.diff(A.indptr), np.diff(A.tocsc().indptr)
    assert (rl[-2:] == 0).all() and (y[-2:] == 0).all() and (cl[-3:] == 0).all()
    assert rl.max() > 20 * np.median(rl) and cl.max() > 20 * np.median(cl)
    assert np.isinf(lp.cl).any() and np.isinf(lp.cu).any() and (lp.cl == lp.cu).any()
    assert np.isinf(lp.rmax).any() and np.isinf(lp.rmin).any()
    mid = lpgen.LP.from_triplets(3, 2, [0, 2, 2], [0, 0, 1], [1.0, 2.0, 3.0], np.zeros(2),
                                 np.zeros(2), np.ones(2), np.zeros(3), np.zeros(3))
    assert np.array_equal(mid.matvec(np.array([1.0, 10.0])), np.array([1.0, 0.0, 32.0]))


@pytest.mark.parametrize("gen", ["zipf", "eq", "empty"])
def test_rowblock_invariants(gen):
    if gen == "zipf":
        lp = lpgen.gen_zipf_badly_scaled(seed=1, m=3000, n=20000)
        ptrs = [lp.ap, lp.cp]
    elif gen == "eq":
        lp = lpgen.gen_random_eq(seed=1)
        ptrs = [lp.ap, lp.cp]
    else:
        ptrs = [np.zeros(5, np.int32), np.array([0, 0, 3000, 3000, 3001])]
    for ptr in ptrs:
        pairs = _rowblocks(ptr).reshape(-1, 2)
        rb = pairs[:, 0]
        assert np.array_equal(pairs[:, 1], np.asarray(ptr)[rb])
        nrows = len(ptr) - 1
        assert rb[0] == 0 and rb[-1] == nrows and np.all(np.diff(rb) > 0)
        for a, b in zip(rb[:-1], rb[1:]):
            nz = ptr[b] - ptr[a]
            assert b - a <= 256
            assert nz <= 2048 or b - a == 1


def _slices(ptr, idx, ncols, S):
    import ctypes as C
    L = _core.lib()
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    nrows, nnz = len(ptr) - 1, int(ptr[-1])
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa
    nb = L.igaos_slices_build(nrows, ncols, ip(ptr), ip(idx), S, None, None, None, None,
                              None, 0)
    assert nb >= 0
    sptr = np.zeros(S * (nrows + 1), np.int32)
    sidx, src = np.zeros(max(nnz, 1), np.int32), np.zeros(max(nnz, 1), np.int32)
    blocks, bslice = np.zeros(4 * max(nb, 1), np.int32), np.zeros(max(nb, 1), np.int32)
    assert L.igaos_slices_build(nrows, ncols, ip(ptr), ip(idx), S, ip(sptr), ip(sidx),
                                ip(src), ip(blocks), ip(bslice), nb) == nb
    return (sptr.reshape(S, nrows + 1), sidx[:nnz], src[:nnz], blocks[:4 * nb].reshape(-1, 4),
            bslice[:nb])


@pytest.mark.parametrize("chunk", [4, 8])
@pytest.mark.parametrize("which", ["ragged-rows", "ragged-cols", "eq-rows", "eq-cols"])
def test_small_ell_layout_places_every_entry_once(which, chunk):
    # the small-LP runner's sliced-ELL copy: slots by decreasing row length
    # (ties by row), slice widths from their longest row, every CSR position
    # at its chunk position in storage order, padding -1; the kernel sums a
    # slot's entries in that order, so this layout fixes its row sums
    import ctypes as C
    L = _core.lib()
    lp = lpgen.gen_ragged(seed=5) if which.startswith("ragged") else lpgen.gen_random_eq(seed=0)
    ptr = np.ascontiguousarray(lp.ap if which.endswith("rows") else lp.cp, np.int32)
    rows = len(ptr) - 1
    slices = (rows + 63) // 64
    perm, off = np.zeros(rows, np.int32), np.zeros(slices + 1, np.int32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa
    total = L.igaos_small_ell_layout(rows, ip(ptr), chunk, ip(perm), ip(off), None, 0)
    src = np.full(total, 7, np.int32)
    assert L.igaos_small_ell_layout(rows, ip(ptr), chunk, ip(perm), ip(off), ip(src),
                                    total) == total
    lens = np.diff(ptr)
    assert np.array_equal(np.sort(perm), np.arange(rows))
    key = [(-lens[r], r) for r in perm]
    assert key == sorted(key)
    nch = -(-lens[perm] // chunk)
    for q in range(slices):
        assert off[q + 1] - off[q] == nch[q * 64:(q + 1) * 64].max()
    assert total == max(off[-1], 1) * 64 * chunk
    want = np.full(total, -1, np.int64)
    for s_, r in enumerate(perm):
        e = np.arange(lens[r])
        pos = ((off[s_ // 64] + e // chunk) * 64 + s_ % 64) * chunk + e % chunk
        want[pos] = ptr[r] + e
    assert np.array_equal(src, want)
    assert np.array_equal(np.sort(src[src >= 0]), np.arange(ptr[-1]))
    bad = np.array([0, 5, 1], np.int32)
    assert L.igaos_small_ell_layout(2, ip(bad), chunk, ip(perm), ip(off), None, 0) < 0
    assert L.igaos_small_ell_layout(rows, ip(ptr), 0, ip(perm), ip(off), None, 0) < 0


@pytest.mark.parametrize("S", [1, 3, 8])
def test_gather_ordered_slices_restore_row_order(S):
    # the gather order of A's slices (renumbered heavy columns): within
    # every block of two or more rows the entries ascend by column, and
    # slot sends each entry back to its place in the plain layout, so the
    # kernels' row sums do not change; other blocks keep their order
    import ctypes as C
    L = _core.lib()
    lp = lpgen.gen_zipf_badly_scaled(seed=2, m=2000, n=9000, kmax=600)
    ptr = np.ascontiguousarray(lp.ap, np.int32)
    idx = np.ascontiguousarray(lp.ai, np.int32)
    nrows, nnz = lp.m, lp.nnz
    sptr, sidx, src, blocks, bslice = _slices(ptr, idx, lp.n, S)
    gidx, gsrc = np.zeros(nnz, np.int32), np.zeros(nnz, np.int32)
    slot = np.zeros(nnz, np.uint16)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa
    assert L.igaos_slices_gather_order(nrows, lp.n, ip(ptr), ip(idx), S, ip(gidx), ip(gsrc),
                                       slot.ctypes.data_as(C.c_void_p)) == nnz
    sorted_blocks = 0
    for r0, r1, z0, z1 in blocks:
        g, s, q = gidx[z0:z1], gsrc[z0:z1], slot[z0:z1].astype(np.int64)
        if r1 - r0 > 1:
            assert np.all(np.diff(g) >= 0)
            assert np.array_equal(np.sort(q), np.arange(z1 - z0))
            back_idx, back_src = np.empty_like(g), np.empty_like(s)
            back_idx[q], back_src[q] = g, s
            assert np.array_equal(back_idx, sidx[z0:z1]) and np.array_equal(back_src, src[z0:z1])
            sorted_blocks += not np.array_equal(g, sidx[z0:z1])
        else:
            assert np.array_equal(g, sidx[z0:z1]) and np.array_equal(s, src[z0:z1])
    assert sorted_blocks > 0
    bad = np.array([0, 5, 1], np.int32)
    assert L.igaos_slices_gather_order(2, 9, ip(bad), ip(idx), S, None, None, None) < 0


@pytest.mark.parametrize("S", [1, 3, 8])
def test_column_slice_invariants(S):
    # the sliced copy of the large-LP products (host builder): a
    # permutation of the nonzeros, each slice a contiguous column range with
    # ranges in order, each row's nonzeros in their own order, and row
    # blocks per slice that cover every row once within the block limits
    lp = lpgen.gen_zipf_badly_scaled(seed=2, m=2000, n=9000, kmax=600)
    for ptr, idx, ncols in ((lp.ap, lp.ai, lp.n), (lp.cp, lp.ci, lp.m)):
        ptr, idx = np.asarray(ptr), np.asarray(idx)
        nrows = len(ptr) - 1
        sptr, sidx, src, blocks, bslice = _slices(ptr, idx, ncols, S)
        assert np.array_equal(np.sort(src), np.arange(ptr[-1]))
        assert np.array_equal(sidx, idx[src])
        rows_of = np.repeat(np.arange(nrows), np.diff(ptr))
        last_max = -1
        for sc in range(S):
            p = sptr[sc]
            assert np.all(np.diff(p) >= 0)
            seg = slice(p[0], p[-1])
            if p[-1] > p[0]:
                assert sidx[seg].min() > last_max
                last_max = sidx[seg].max()
            # each row's entries in this slice: the row's own, in order
            r = np.repeat(np.arange(nrows), np.diff(p))
            assert np.array_equal(rows_of[src[seg]], r)
            for i in range(0, nrows, max(1, nrows // 50)):
                assert np.all(np.diff(src[p[i]:p[i + 1]]) > 0)
            b = blocks[bslice == sc]
            assert np.array_equal(b[:, 0], np.concatenate([[0], b[:-1, 1]]))
            assert b[-1, 1] == nrows and np.all(b[:, 1] > b[:, 0])
            assert np.array_equal(b[:, 2], p[b[:, 0]]) and np.array_equal(b[:, 3], p[b[:, 1]])
            assert np.all((b[:, 1] - b[:, 0] <= 256))
            assert np.all((b[:, 3] - b[:, 2] <= 2048) | (b[:, 1] - b[:, 0] == 1))
        assert np.all(np.diff(bslice) >= 0)
        # equal nonzero shares: no slice holds more than its share plus one column
        counts = np.array([sptr[sc][-1] - sptr[sc][0] for sc in range(S)])
        colmax = np.bincount(idx, minlength=ncols).max()
        assert counts.max() <= ptr[-1] / S + colmax


def test_exact_mode_config_is_checked_before_any_device_work():
    # the exact mode is FP64 only; the C-ABI refuses the pair as an invalid
    # argument (negative return, igaos_last_error), GPU or not
    import igaos_amd
    from igaos_amd import lpgen
    lp = lpgen.gen_random_eq(seed=3, m=20, n=50)
    with pytest.raises(ValueError, match="exact mode needs precision 64"):
        igaos_amd.solve_lp(lp, precision=32, exact=True, time_limit=1e7, max_iterations=10)


def test_bench_refuses_more_than_one_gpu():
    import json
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert json.loads(r.stdout.strip().splitlines()[-1])["n_gpus"] == 2


def test_gather_order_invariants():
    # the lockstep kernels' gather order (host builder): inside each
    # multi-row block a stable sort by column with slots back to row order;
    # single-row blocks unchanged
    import ctypes as C
    L = _core.lib()
    lp = lpgen.gen_batch_instance(0, 3)
    for ptr, idx, ncols in ((lp.ap, lp.ai, lp.n), (lp.cp, lp.ci, lp.m)):
        ptr = np.ascontiguousarray(ptr, dtype=np.int32)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        nnz = int(ptr[-1])
        sidx, slot = np.zeros(nnz, np.uint16), np.zeros(nnz, np.uint16)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa
        up = lambda a: a.ctypes.data_as(C.POINTER(C.c_ushort))  # noqa
        assert L.igaos_gather_order_build(len(ptr) - 1, ncols, ip(ptr), ip(idx), up(sidx),
                                          up(slot)) == 0
        pairs = _rowblocks(ptr).reshape(-1, 2)
        for (r0, p0), (r1, p1) in zip(pairs[:-1], pairs[1:]):
            sl = slot[p0:p1].astype(np.int64)
            assert np.array_equal(np.sort(sl), np.arange(p1 - p0))
            assert np.array_equal(sidx[p0:p1], idx[p0 + sl])
            if r1 - r0 > 1:
                assert np.all(np.diff(sidx[p0:p1].astype(np.int64)) >= 0)
                # stable: equal columns keep their row order
                eq = np.diff(sidx[p0:p1].astype(np.int64)) == 0
                assert np.all(np.diff(sl)[eq] > 0)
            else:
                assert np.array_equal(sl, np.arange(p1 - p0))
    assert L.igaos_gather_order_build(2, 70000, ip(np.array([0, 0, 0], np.int32)), None,
                                      up(np.zeros(1, np.uint16)),
                                      up(np.zeros(1, np.uint16))) != 0


def test_layer_and_engine_argument_errors():
    # argument errors surface before any device work
    from igaos_amd.autograd import DiffLPBatch
    lp = lpgen.gen_batch_instance(0, 0, m_hi=40)
    with pytest.raises(ValueError):
        DiffLPBatch([lp], differentiation="both")
    # the engine configuration rejects an out-of-range slice count
    import ctypes as C
    cfg = igaos_amd._core.EngineConfig()
    _core.lib().igaos_engine_config_default(C.byref(cfg))
    assert cfg.column_slices == -1 and cfg.precision == 64
    cfg.column_slices = 17
    m = igaos_amd.api.model_from_lp(lp)
    assert not _core.lib().igaos_pdhg_create(m._h, C.byref(cfg))
    assert "column_slices" in _core.last_error()


def test_malformed_patterns_are_rejected_before_device_work():
    # the kernels index vectors by the pattern arrays unchecked, so both
    # entry points that take caller arrays refuse a malformed pattern
    import dataclasses
    from igaos_amd.batch import BatchSolver
    lp = lpgen.gen_batch_instance(0, 0, m_hi=40)
    assert lp.nnz > 2

    def bad(**kw):
        return dataclasses.replace(lp, **{k: np.array(v, dtype=np.int32) for k, v in kw.items()})

    ai = lp.ai.copy()
    ai[1] = lp.n                       # column index out of range
    ap = lp.ap.copy()
    ap[1], ap[2] = ap[2], ap[1] - 1    # row pointer decreases
    ci = lp.ci.copy()
    ci[0] = (ci[0] + 1) % lp.m         # CSC entry moved to another row
    cases = [(bad(ai=ai), "outside"), (bad(ap=ap), "decreases"),
             (bad(ci=ci), "CSR and CSC differ|not the transpose")]
    for b, what in cases:
        with pytest.raises(ValueError, match=what):
            igaos_amd.api.model_from_lp(b)
        with pytest.raises(RuntimeError, match="instance 1: .*" + what):
            BatchSolver([lp, b])
    # the well-formed instance passes the checks (and then needs a device)
    m = igaos_amd.api.model_from_lp(lp)
    assert m.n == lp.n


def test_csc_must_be_the_transpose_of_csr():
    # same entry counts per row and per column, but not the same matrix:
    # an A' that is not the transpose of A would make the engine iterate on
    # a different operator, so both model entry points refuse it; a CSC
    # listing each column's rows in another order is the same matrix
    import dataclasses
    from igaos_amd.batch import BatchSolver
    lp = lpgen.gen_random_eq(seed=4, m=30, n=70)
    j1, j2 = 3, 11
    p1, p2 = lp.cp[j1], lp.cp[j2]
    assert lp.ci[p1] != lp.ci[p2]
    ci = lp.ci.copy()
    ci[p1], ci[p2] = lp.ci[p2], lp.ci[p1]   # row counts and column counts kept
    acx = lp.acx.copy()
    acx[p1] = acx[p1] + 1.0                   # same pattern, another value
    for b in (dataclasses.replace(lp, ci=ci), dataclasses.replace(lp, acx=acx)):
        with pytest.raises(ValueError, match="not the transpose"):
            igaos_amd.api.model_from_lp(b)
        with pytest.raises(RuntimeError, match="instance 1: .*not the transpose"):
            BatchSolver([lp, b])
    rev_ci, rev_acx = lp.ci.copy(), lp.acx.copy()
    for j in range(lp.n):
        a, b = lp.cp[j], lp.cp[j + 1]
        rev_ci[a:b] = lp.ci[a:b][::-1]
        rev_acx[a:b] = lp.acx[a:b][::-1]
    m = igaos_amd.api.model_from_lp(dataclasses.replace(lp, ci=rev_ci, acx=rev_acx))
    assert m.nnz() == lp.nnz


def test_model_from_csr_needs_a_device_and_a_well_formed_csr():
    lp = lpgen.gen_random_eq(seed=4, m=30, n=70)
    ai = lp.ai.copy()
    ai[0] = lp.n
    args = (lp.m, lp.n, lp.c, lp.cl, lp.cu, lp.rmin, lp.rmax, lp.ap)
    with pytest.raises(ValueError, match="outside"):
        igaos_amd.model_from_csr(*args, ai, lp.ax)
    if _core.device_count() == 0:
        with pytest.raises(ValueError, match="none is visible"):
            igaos_amd.model_from_csr(*args, lp.ai, lp.ax)


def test_layout_builders_refuse_malformed_patterns():
    # the host builders index their own arrays by the caller's pattern
    import ctypes as C
    L = _core.lib()
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa
    up = lambda a: a.ctypes.data_as(C.POINTER(C.c_ushort))  # noqa
    ptr = np.array([0, 2, 3], np.int32)
    good = np.array([0, 4, 1], np.int32)
    out = np.zeros(64, np.int32)
    nb = np.zeros(1, np.int32)
    assert L.igaos_rowblocks_build(2, ip(ptr), ip(out), 64) == 4
    assert L.igaos_rowblocks_build(2, ip(np.array([1, 2, 3], np.int32)), ip(out), 64) < 0
    assert "start at 0" in _core.last_error()
    assert L.igaos_rowblocks_build(2, ip(np.array([0, 3, 2], np.int32)), ip(out), 64) < 0
    assert "decreases" in _core.last_error()
    big = np.zeros(16, np.int32)
    args = lambda idx: (2, 5, ip(ptr), ip(idx), 2, ip(big), ip(big), ip(big), ip(big),  # noqa
                        ip(nb), 1)
    assert L.igaos_slices_build(*args(good)) > 0
    assert L.igaos_slices_build(*args(np.array([0, 5, 1], np.int32))) < 0
    assert "outside" in _core.last_error()
    s16 = np.zeros(3, np.uint16)
    assert L.igaos_gather_order_build(2, 5, ip(ptr), ip(good), up(s16), up(s16)) == 0
    assert L.igaos_gather_order_build(2, 5, ip(ptr), ip(np.array([0, -1, 1], np.int32)),
                                      up(s16), up(s16)) < 0
    assert "outside" in L.igaos_batch_last_error().decode()


def test_row_activity_is_csr_ordered_product():
    # igaos_model_row_activity sums each row from 0 in CSR order, as the
    # reference simplex fills Solution::row_activity; scipy's CSR product
    # sums in the same order, so the two agree bit for bit
    import ctypes as C
    m = igaos_amd.read_mps(netlib_path("sc205"))
    lp = m.to_lp()
    x = np.random.default_rng(0).standard_normal(lp.n)
    out = np.zeros(lp.m)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa
    assert _core.lib().igaos_model_row_activity(m._h, dp(x), dp(out)) == 0
    assert np.array_equal(out, lp.to_scipy() @ x)
    assert _core.lib().igaos_model_row_activity(m._h, None, dp(out)) != 0


def test_bench_reads_committed_traffic_summaries():
    # bench.py's roofline "traffic" comes from the committed counter
    # summaries; each must parse and give bytes near the algorithmic ones
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    t = bench.load_traffic("lockstep_batch8192", 1.0e9)
    assert t is not None and 1.0e9 <= t <= 1.3e9
    for dom in ("primal", "dual"):
        t = bench.load_traffic(f"{dom}_step_zipf")
        assert t is not None and 8.0e8 <= t <= 1.5e9
    assert bench.load_traffic("no_such_kernel") is None


@pytest.mark.parametrize("name", ["afiro", "sc50a", "kb2", "adlittle", "sc205"])
def test_dual_repair_matches_oracle_and_weak_duality(name, objectives):
    # the engine's host dual repair (pdhg.cu:288-448) against the oracle's
    # restatement on the same point, bit for bit, and against the
    # mathematics: the repaired dual is feasible to the repair's tolerance,
    # its Lagrangian bound is finite, lies below the known optimum (weak
    # duality) and close to it for a near-optimal x
    import ctypes as C
    import oracle
    m = igaos_amd.read_mps(netlib_path(name))
    lp = m.to_lp()
    r = oracle.solve(lp, tolerance=1e-4, time_limit=60)
    x, y = np.ascontiguousarray(r.x), np.ascontiguousarray(r.y)
    A = lp.to_scipy()
    Ax = A @ x
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa
    yp = np.zeros(lp.m)
    ok = _core.lib().igaos_dual_repair(m._h, dp(x), dp(Ax), dp(y), 1e-6, dp(yp))
    ok_o, yo = oracle.dual_repair(lp, x, Ax, y, 1e-6)
    assert ok == 1 and ok_o and np.array_equal(yp, yo)
    mt = oracle.evaluate(lp.rmin, lp.rmax, lp.cl, lp.cu, lp.c, x, yp, Ax, lp.c + A.T @ yp)
    ref = objectives[f"netlib/{name}"]["highs_ref"]
    bound = mt.od + lp.obj_const
    assert math.isfinite(mt.gap) and mt.dinf <= 1e-5
    assert bound <= ref + 1e-7 * (1 + abs(ref))
    assert ref - bound <= 1e-4 * (1 + abs(ref))


def test_dual_repair_closed_forms():
    # every column strictly inside its bounds: the repair must reach the
    # unique y with c + A'y = 0 from a far (and, for the G row, cone-
    # violating) iterate: y = -1 for min x1 + x2, x1 + x2 = 2 at x = (1, 1);
    # y = -1.5 for min 3x, 2x >= 2 at x = 1 (sign cone y <= 0)
    import ctypes as C
    import oracle
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa
    cases = [
        (lpgen.LP.from_triplets(1, 2, [0, 0], [0, 1], [1.0, 1.0], np.ones(2), np.zeros(2),
                                np.full(2, 10.0), np.array([2.0]), np.array([2.0])),
         np.array([1.0, 1.0]), np.array([0.5]), -1.0),
        (lpgen.LP.from_triplets(1, 1, [0], [0], [2.0], np.array([3.0]), np.zeros(1),
                                np.full(1, INF), np.array([2.0]), np.array([INF])),
         np.array([1.0]), np.array([0.7]), -1.5),
    ]
    for lp, x, y_it, want in cases:
        m = igaos_amd.api.model_from_lp(lp)
        Ax = lp.to_scipy() @ x
        y = np.zeros(lp.m)
        assert _core.lib().igaos_dual_repair(m._h, dp(x), dp(Ax), dp(y_it), 1e-6, dp(y)) == 1
        assert abs(y[0] - want) <= 1e-12
        ok, yo = oracle.dual_repair(lp, x, Ax, y_it, 1e-6)
        assert ok and abs(yo[0] - want) <= 1e-12


def test_dual_repair_closed_form_catches_a_flipped_projection():
    # the closed form fails when the oracle's affine projection step is
    # applied with the wrong sign (mutation 6)
    import oracle
    lp = lpgen.LP.from_triplets(1, 2, [0, 0], [0, 1], [1.0, 1.0], np.ones(2), np.zeros(2),
                                np.full(2, 10.0), np.array([2.0]), np.array([2.0]))
    x = np.array([1.0, 1.0])
    Ax = lp.to_scipy() @ x
    oracle.set_mutation(6)
    try:
        ok, y = oracle.dual_repair(lp, x, Ax, np.array([0.5]), 1e-6)
    finally:
        oracle.set_mutation(0)
    assert not (ok and abs(y[0] + 1.0) <= 1e-12)


@pytest.mark.parametrize("name", ["afiro", "kb2", "sc205", "israel"])
def test_host_kkt_evaluate_matches_oracle(name):
    # the host evaluation the repair uses (pdhg.cu:173-254) equals the
    # oracle's, bit for bit, at a near-optimal pair and at a perturbed one
    # (whose Lagrangian bound is -inf and gap inf)
    import ctypes as C
    import oracle
    m = igaos_amd.read_mps(netlib_path(name))
    lp = m.to_lp()
    r = oracle.solve(lp, tolerance=1e-4, time_limit=60)
    A = lp.to_scipy()
    rng = np.random.default_rng(0)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa
    pairs = ((r.x, r.y), (r.x + 0.01 * rng.standard_normal(lp.n),
                          r.y + 0.01 * rng.standard_normal(lp.m)))
    for k, (x, y) in enumerate(pairs):
        x, y = np.ascontiguousarray(x), np.ascontiguousarray(y)
        Ax, d = A @ x, lp.c + A.T @ y
        out = np.zeros(6)
        assert _core.lib().igaos_kkt_evaluate(m._h, dp(x), dp(y), dp(Ax), dp(d), dp(out)) == 0
        t = oracle.evaluate(lp.rmin, lp.rmax, lp.cl, lp.cu, lp.c, x, y, Ax, d)
        assert np.array_equal(out, [t.pinf, t.dinf, t.gap, t.op, t.od, t.prow])
        if k == 1:
            assert out[2] == INF and out[4] == -INF


def test_autotuner_classes_and_cached_schedule(tmp_path):
    # host logic of the autotuner: quantile size classes cover every
    # instance, each instance maps to its class's lanes per row, and a
    # cached schedule is returned without probing the device
    import json
    from dataclasses import asdict
    from igaos_amd import autotune
    lps = [lpgen.gen_batch_instance(0, i, m_hi=80) for i in range(40)]
    nnz = np.array([lp.nnz for lp in lps])
    edges = autotune._classes(nnz, 4)
    assert edges[0] == nnz.min() and edges[-1] == nnz.max() + 1
    assert np.all(np.diff(edges) > 0)
    k = len(edges) - 1
    s = autotune.Schedule(edges=[float(e) for e in edges], tpr=[0, 1, 2, 4][:k], waves=4,
                          probe_ms={}, mode="lockstep",
                          mode_ms={"persistent": 2.0, "lockstep": 1.0})
    cls = np.searchsorted(edges, nnz, side="right") - 1
    assert np.all((cls >= 0) & (cls < k))
    assert np.array_equal(s.tpr_for(nnz), np.array(s.tpr, np.int32)[cls])
    key = ",".join(str(int(e)) for e in edges) + ";modes"
    cache = tmp_path / "schedule.json"
    cache.write_text(json.dumps({key: asdict(s)}))
    assert autotune.tune(lps, cache=str(cache), modes=True) == s


def test_batch_solution_sizes_are_checked_on_the_host():
    # set_solution refuses arrays whose length is not the batch's sum of n
    # (or m) before the C-ABI reads them
    from igaos_amd.batch import BatchSolver
    b = BatchSolver.__new__(BatchSolver)
    b._h, b._coff, b._roff = None, np.array([0, 3, 5]), np.array([0, 2, 4])
    with pytest.raises(ValueError, match="5 entries"):
        b.set_solution(x=np.zeros(4))
    with pytest.raises(ValueError, match="4 entries"):
        b.set_solution(y=[np.zeros(2), np.zeros(1)])


@pytest.mark.parametrize("case", ["gen_random_eq(seed=1,m=200,n=600)", "netlib/sc50a",
                                  "netlib/blend", "netlib/grow7"])
def test_crossover_basis_warm_starts_the_reference_simplex(case):
    # PDHG -> simplex crossover (igaos_crossover_basis): from a 1e-6 PDHG
    # point (the oracle's, standing in for the engine's on the CPU) a basis
    # in the reference simplex's WarmStart layout; the reference's own
    # revised simplex (oracle/_ref, built from its sources) started there
    # reaches the cold solve's optimum in no more pivots, and in a few on
    # the random LP the cold solve needs 1,388 for
    import oracle
    from tests.conftest import netlib_path
    if not oracle.have_ref_simplex():
        pytest.skip("reference simplex not built (oracle/_ref)")
    if case.startswith("netlib/"):
        lp = igaos_amd.read_mps(netlib_path(case[7:])).to_lp()
    else:
        lp = lpgen.gen_random_eq(seed=1, m=200, n=600)
    r = oracle.solve(lp, tolerance=1e-6, time_limit=60)
    assert r.status == "near-optimal"
    md = igaos_amd.model_from_lp(lp)
    basis, st, k = igaos_amd.api.crossover_basis(md, r.x, r.y, 1e-6)
    assert k == lp.m
    placed = basis[basis >= 0]
    assert len(set(placed.tolist())) == len(placed) and placed.max() < lp.n + lp.m
    assert np.all(st[placed] == 2) and np.sum(st == 2) == len(placed)
    assert set(np.unique(st).tolist()) <= {0, 1, 2, 3}
    cold = oracle.ref_simplex(lp, time_limit=60)
    warm = oracle.ref_simplex(lp, time_limit=60, warm=(basis, st))
    assert cold[0] == warm[0] == "optimal"
    assert abs(warm[1] - cold[1]) <= 1e-9 * (1 + abs(cold[1]))
    assert warm[4] <= cold[4]
    if case.startswith("gen_random_eq"):
        assert warm[4] <= 0.1 * cold[4], (warm[4], cold[4])
