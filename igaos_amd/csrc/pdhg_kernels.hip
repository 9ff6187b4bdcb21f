// Single-LP PDHG kernels for gfx950: row-block SpMV, the fused primal and
// dual steps with the on-device line-search decision, KKT evaluation in
// original units, Ruiz equilibration and fixed-order reductions.
//
// Per attempt the engine launches three kernels:
//   k_primal : A'y (transposed CSR = CSC of A) fused with the primal step
//              x' = clamp(x - tau (g + A'y), lo, hi), xbar = 2x' - x and the
//              |dx|^2 partial (src/pdhg/pdhg.cu:32-45), plus the pending
//              averaging of x (pdhg.cu:127-141);
//   k_dual   : A xbar fused with the dual step in Moreau form
//              y' = t - sigma clamp((t + yoff)/sigma, rlo, rhi),
//              t = y + sigma A xbar (pdhg.cu:47-61), Adx = (A xbar - Ax)/2,
//              the |dy|^2 and dy'Adx partials, the pending averaging of y
//              and the A x commit;
//   k_decide : one workgroup reduces all partials in a fixed order and
//              applies the line-search rule (pdhg.cu:751-794) to the
//              device-resident LsState.
// A rejected attempt re-runs both kernels from the same (x, y); k_primal
// then reuses the stored A'y.  On large LPs either product may instead run
// column-sliced (k_primal_slices / k_dual_slices plus an epilogue kernel
// with the same step, below).

#include <algorithm>
#include <numeric>

#include "dev_common.hpp"
#include "pdhg_kernels.hpp"

namespace igaos::dev {

using pdhg::LsState;
using pdhg::Metrics;

std::vector<int> build_row_blocks(const int* ptr, int nrows) {
    std::vector<int> rb;
    rb.reserve(static_cast<size_t>(nrows) / 8 + 2);
    rb.push_back(0);
    int r = 0;
    while (r < nrows) {
        const int len0 = ptr[r + 1] - ptr[r];
        if (len0 > NNZ_MAX) {
            ++r;
            rb.push_back(r);
            continue;
        }
        long nz = 0;
        int cnt = 0;
        while (r < nrows && cnt < ROWS_MAX) {
            const int len = ptr[r + 1] - ptr[r];
            if (len > NNZ_MAX || nz + len > NNZ_MAX) break;
            nz += len;
            ++cnt;
            ++r;
        }
        rb.push_back(r);
    }
    std::vector<int> out(2 * rb.size());
    for (size_t b = 0; b < rb.size(); ++b) {
        out[2 * b] = rb[b];
        out[2 * b + 1] = ptr[rb[b]];
    }
    return out;
}

static inline int grid_of(int nblk) { return nblk > 0 ? nblk : 1; }

// ---------------------------------------------------------------- SpMV --

template <typename Acc, typename Val, typename G>
__global__ __launch_bounds__(BS) void k_spmv(RowBlocks rb,
                                             const int* __restrict__ ptr,
                                             const int* __restrict__ idx,
                                             const Val* __restrict__ val,
                                             G gx, Acc* __restrict__ y) {
    __shared__ Acc s_prod[NNZ_MAX];
    __shared__ Acc s_rows[ROWS_MAX];
    __shared__ Acc red[BS / WAVE];
    int r0, r1, p0, p1;
    block_rows(rb, blockIdx.x, r0, r1, p0, p1);
    rowblock_sums<Acc>(ptr, idx, val, gx, r0, r1, p0, p1, s_rows, s_prod, red);
    for (int t = threadIdx.x; t < r1 - r0; t += BS) y[r0 + t] = s_rows[t];
}

void spmv(const RowBlocks& rb, const int* ptr, const int* idx,
          const double* val, const double* x, double* y, hipStream_t s) {
    if (rb.nblk == 0) return;
    k_spmv<double, double><<<rb.nblk, BS, 0, s>>>(rb, ptr, idx, val,
                                                   GatherPlain<double>{x}, y);
}
void spmv(const RowBlocks& rb, const int* ptr, const int* idx,
          const float* val, const float* x, float* y, hipStream_t s) {
    if (rb.nblk == 0) return;
    k_spmv<float, float><<<rb.nblk, BS, 0, s>>>(rb, ptr, idx, val,
                                                 GatherPlain<float>{x}, y);
}
void spmv_ref(const RowBlocks& rb, const int* ptr, const int* idx,
              const double* val, const double* xref, const float* xd,
              double* y, hipStream_t s) {
    if (rb.nblk == 0) return;
    k_spmv<double, double><<<rb.nblk, BS, 0, s>>>(
        rb, ptr, idx, val, GatherRef<float>{xref, xd}, y);
}

// ----------------------------------------------------- primal / dual --

template <typename V, bool NTL>
__global__ __launch_bounds__(BS) void k_primal(IterBufs<V> b) {
    __shared__ V s_prod[NNZ_MAX];
    __shared__ V s_rows[ROWS_MAX];
    __shared__ V red_v[BS / WAVE];
    __shared__ double red[BS / WAVE];
    const LsState st = *b.st;
    if (st.frozen) return;
    const int cur = st.cur;
    const V tau = static_cast<V>(st.eta / st.omega);
    const V w = static_cast<V>(st.mix_w);
    const bool fresh = st.fresh != 0;
    const bool mix = fresh && st.mix_pending != 0;
    const V* __restrict__ x = b.x[cur];
    V* __restrict__ xn = b.x[cur ^ 1];
    const V* __restrict__ y = b.y[cur];
    int r0, r1, p0, p1;
    block_rows(b.rbT, blockIdx.x, r0, r1, p0, p1);
    if (fresh)
        rowblock_sums_t<NTL, V>(b.cp, b.ci, b.acx, GatherPlain<V>{y}, r0, r1, p0, p1, s_rows,
                                s_prod, red_v, nullptr);
    double acc = 0.0;
    for (int t = threadIdx.x; t < r1 - r0; t += BS) {
        const int j = r0 + t;
        V a;
        if (fresh) {
            a = s_rows[t];
            b.ata[j] = a;
        } else {
            a = b.ata[j];
        }
        const V xo = x[j];
        if (mix) {
            const V xa = b.xavg[j];
            b.xavg[j] = xa + w * (xo - xa);
        }
        const V g = b.gref[j] + a;
        const V v = fmin(fmax(xo - tau * g, b.lo[j]), b.hi[j]);
        xn[j] = v;
        b.xbar[j] = V(2) * v - xo;
        const V d = v - xo;
        acc += static_cast<double>(d) * static_cast<double>(d);
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) b.partT[blockIdx.x] = acc;
}

template <typename V, bool NTL>
__global__ __launch_bounds__(BS) void k_dual(IterBufs<V> b) {
    __shared__ V s_prod[NNZ_MAX];
    __shared__ V s_rows[ROWS_MAX];
    __shared__ V red_v[BS / WAVE];
    __shared__ double red[BS / WAVE];
    const LsState st = *b.st;
    if (st.frozen) return;
    const int cur = st.cur;
    const V sig = static_cast<V>(st.eta * st.omega);
    const V w = static_cast<V>(st.mix_w);
    const bool mix = st.fresh != 0 && st.mix_pending != 0;
    const V* __restrict__ y = b.y[cur];
    V* __restrict__ yn = b.y[cur ^ 1];
    int r0, r1, p0, p1;
    block_rows(b.rbA, blockIdx.x, r0, r1, p0, p1);
    rowblock_sums_t<NTL, V>(b.ap, b.ai, b.ax, GatherPlain<V>{b.xbar}, r0, r1, p0, p1, s_rows,
                            s_prod, red_v, nullptr);
    double a1 = 0.0, a2 = 0.0;
    for (int t = threadIdx.x; t < r1 - r0; t += BS) {
        const int i = r0 + t;
        const V tm = s_rows[t];
        const V yo = y[i];
        V axc = b.axcur[i];
        if (mix) {
            const V ya = b.yavg[i];
            b.yavg[i] = ya + w * (yo - ya);
            axc = axc + b.adx[i];
            b.axcur[i] = axc;
        }
        const V tt = yo + sig * tm;
        V u = b.yoff ? tt + b.yoff[i] : tt;
        u = u / sig;
        const V yv = tt - sig * fmin(fmax(u, b.rlo[i]), b.rhi[i]);
        yn[i] = yv;
        const V dy = yv - yo;
        const V ad = V(0.5) * (tm - axc);
        b.adx[i] = ad;
        a1 += static_cast<double>(dy) * static_cast<double>(dy);
        a2 += static_cast<double>(dy) * static_cast<double>(ad);
    }
    a1 = block_sum(a1, red);
    a2 = block_sum(a2, red);
    if (threadIdx.x == 0) {
        b.partA[2 * blockIdx.x] = a1;
        b.partA[2 * blockIdx.x + 1] = a2;
    }
}

// Line-search decision: fixed-order sums of the attempt's partials, then
// the accept/reject rule on the device-resident state.
template <typename V>
__global__ __launch_bounds__(FIN) void k_decide(IterBufs<V> b) {
    __shared__ double red[3 * FIN / WAVE];
    if (b.st->frozen) return;
    double v[3] = {fin_part_sum(b.partT, b.nT_part, 1), fin_part_sum(b.partA, b.nA_part, 2),
                   fin_part_sum(b.partA + 1, b.nA_part, 2)};
    block_reduce_k<3, FIN>(v, 0u, red);
    if (threadIdx.x == 0) {
        LsState s = *b.st;
        pdhg::ls_decide(s, v[0], v[1], v[2]);
        *b.st = s;
    }
}

template <typename V>
__global__ __launch_bounds__(BS) void k_flush(int n, int m, const V* x,
                                              V* xavg, const V* y, V* yavg,
                                              const V* adx, V* axcur, V w) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < n) xavg[i] += w * (x[i] - xavg[i]);
    if (i < m) {
        yavg[i] += w * (y[i] - yavg[i]);
        axcur[i] += adx[i];
    }
}


// ------------------------------------------------------- sliced path --
// Large LPs: the products of A' (gathers y) and A (gathers xbar) run over
// the column-sliced copies (one launch, slice-major blocks, partial row
// sums per slice), then an epilogue kernel per column / row adds the
// partials in slice order and applies the same primal / dual step as
// k_primal / k_dual, one thread per column / row.

template <bool NTL, typename V>
__device__ __forceinline__ void slice_rows(const Sliced<V>& sl, int blk, const V* x,
                                           V* s_prod, V* s_rows, V* red_v) {
    const int4 q = sl.quad[blk];
    const int sc = sl.bslice[blk];
    rowblock_sums_t<NTL, V>(sl.ptr + static_cast<long>(sc) * (sl.rows + 1), sl.idx, sl.val,
                            GatherPlain<V>{x}, q.x, q.y, q.z, q.w, s_rows, s_prod, red_v,
                            sl.slot);
    V* out = sl.part + static_cast<long>(sc) * sl.rows;
    for (int t = threadIdx.x; t < q.y - q.x; t += BS) out[q.x + t] = s_rows[t];
}

template <typename V, bool NTL>
__global__ __launch_bounds__(BS) void k_primal_slices(IterBufs<V> b, int b0) {
    __shared__ V s_prod[NNZ_MAX];
    __shared__ V s_rows[ROWS_MAX];
    __shared__ V red_v[BS / WAVE];
    const LsState st = *b.st;
    if (st.frozen || !st.fresh) return;  // a retry reuses the stored A'y
    slice_rows<NTL>(b.sT, b0 + static_cast<int>(blockIdx.x), b.y[st.cur], s_prod, s_rows, red_v);
}

template <typename V>
__global__ __launch_bounds__(BS) void k_primal_epi(IterBufs<V> b) {
    __shared__ double red[BS / WAVE];
    const LsState st = *b.st;
    if (st.frozen) return;
    const int cur = st.cur;
    const V tau = static_cast<V>(st.eta / st.omega);
    const V w = static_cast<V>(st.mix_w);
    const bool fresh = st.fresh != 0;
    const bool mix = fresh && st.mix_pending != 0;
    const int n = b.n, S = b.sT.S;
    const int j = blockIdx.x * BS + threadIdx.x;
    double acc = 0.0;
    if (j < n) {
        V a;
        if (fresh) {
            a = b.sT.part[j];
            for (int sc = 1; sc < S; ++sc) a += b.sT.part[static_cast<long>(sc) * n + j];
            b.ata[j] = a;
        } else {
            a = b.ata[j];
        }
        const V xo = b.x[cur][j];
        if (mix) {
            const V xa = b.xavg[j];
            b.xavg[j] = xa + w * (xo - xa);
        }
        const V g = b.gref[j] + a;
        const V v = fmin(fmax(xo - tau * g, b.lo[j]), b.hi[j]);
        b.x[cur ^ 1][j] = v;
        b.xbar[j] = V(2) * v - xo;
        const V d = v - xo;
        acc = static_cast<double>(d) * static_cast<double>(d);
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) b.partT[blockIdx.x] = acc;
}

template <typename V, bool NTL>
__global__ __launch_bounds__(BS) void k_dual_slices(IterBufs<V> b, int b0) {
    __shared__ V s_prod[NNZ_MAX];
    __shared__ V s_rows[ROWS_MAX];
    __shared__ V red_v[BS / WAVE];
    if (b.st->frozen) return;
    slice_rows<NTL>(b.sA, b0 + static_cast<int>(blockIdx.x), b.xbar, s_prod, s_rows, red_v);
}

template <typename V>
__global__ __launch_bounds__(BS) void k_dual_epi(IterBufs<V> b) {
    __shared__ double red[BS / WAVE];
    const LsState st = *b.st;
    if (st.frozen) return;
    const int cur = st.cur;
    const V sig = static_cast<V>(st.eta * st.omega);
    const V w = static_cast<V>(st.mix_w);
    const bool mix = st.fresh != 0 && st.mix_pending != 0;
    const int m = b.m, S = b.sA.S;
    const int i = blockIdx.x * BS + threadIdx.x;
    double a1 = 0.0, a2 = 0.0;
    if (i < m) {
        V tm = b.sA.part[i];
        for (int sc = 1; sc < S; ++sc) tm += b.sA.part[static_cast<long>(sc) * m + i];
        const V yo = b.y[cur][i];
        V axc = b.axcur[i];
        if (mix) {
            const V ya = b.yavg[i];
            b.yavg[i] = ya + w * (yo - ya);
            axc = axc + b.adx[i];
            b.axcur[i] = axc;
        }
        const V tt = yo + sig * tm;
        V u = b.yoff ? tt + b.yoff[i] : tt;
        u = u / sig;
        const V yv = tt - sig * fmin(fmax(u, b.rlo[i]), b.rhi[i]);
        b.y[cur ^ 1][i] = yv;
        const V dy = yv - yo;
        const V ad = V(0.5) * (tm - axc);
        b.adx[i] = ad;
        a1 = static_cast<double>(dy) * static_cast<double>(dy);
        a2 = static_cast<double>(dy) * static_cast<double>(ad);
    }
    a1 = block_sum(a1, red);
    a2 = block_sum(a2, red);
    if (threadIdx.x == 0) {
        b.partA[2 * blockIdx.x] = a1;
        b.partA[2 * blockIdx.x + 1] = a2;
    }
}

template <typename T>
__global__ void k_gather_values(const double* src_val, const int* src, T* dst, long nnz) {
    const long q = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (q < nnz) dst[q] = static_cast<T>(src_val[src[q]]);
}

void gather_values(const double* src_val, const int* src, double* dst, long nnz,
                   hipStream_t s) {
    if (nnz > 0)
        k_gather_values<double><<<(nnz + BS - 1) / BS, BS, 0, s>>>(src_val, src, dst, nnz);
}
void gather_values(const double* src_val, const int* src, float* dst, long nnz,
                   hipStream_t s) {
    if (nnz > 0)
        k_gather_values<float><<<(nnz + BS - 1) / BS, BS, 0, s>>>(src_val, src, dst, nnz);
}

HostSlices build_slices(int rows, int cols, const int* ptr, const int* idx, int S,
                        bool gather_order) {
    HostSlices h;
    h.S = S;
    const long nnz = ptr[rows];
    // column ranges with (about) equal nonzero counts
    std::vector<long> ccount(static_cast<size_t>(cols) + 1, 0);
    for (long p = 0; p < nnz; ++p) ccount[idx[p] + 1] += 1;
    for (int j = 0; j < cols; ++j) ccount[j + 1] += ccount[j];
    std::vector<int> bound(S + 1, cols);
    bound[0] = 0;
    for (int sc = 1; sc < S; ++sc) {
        const long target = nnz * sc / S;
        int j = bound[sc - 1];
        while (j < cols && ccount[j] < target) ++j;
        bound[sc] = j;
    }
    std::vector<int> slice_of(cols);
    for (int sc = 0; sc < S; ++sc)
        for (int j = bound[sc]; j < bound[sc + 1]; ++j) slice_of[j] = sc;
    h.ptr.assign(static_cast<size_t>(S) * (rows + 1), 0);
    std::vector<long> rowcnt(static_cast<size_t>(S) * rows, 0);
    for (int r = 0; r < rows; ++r)
        for (int p = ptr[r]; p < ptr[r + 1]; ++p)
            rowcnt[static_cast<size_t>(slice_of[idx[p]]) * rows + r] += 1;
    long off = 0;
    for (int sc = 0; sc < S; ++sc) {
        int* sp = h.ptr.data() + static_cast<size_t>(sc) * (rows + 1);
        for (int r = 0; r < rows; ++r) {
            sp[r] = static_cast<int>(off);
            off += rowcnt[static_cast<size_t>(sc) * rows + r];
        }
        sp[rows] = static_cast<int>(off);
    }
    h.idx.resize(nnz);
    h.src.resize(nnz);
    std::vector<int> fill(static_cast<size_t>(S) * rows);
    for (int sc = 0; sc < S; ++sc)
        for (int r = 0; r < rows; ++r)
            fill[static_cast<size_t>(sc) * rows + r] = h.ptr[static_cast<size_t>(sc) * (rows + 1) + r];
    for (int r = 0; r < rows; ++r)
        for (int p = ptr[r]; p < ptr[r + 1]; ++p) {
            const int sc = slice_of[idx[p]];
            const int q = fill[static_cast<size_t>(sc) * rows + r]++;
            h.idx[q] = idx[p];
            h.src[q] = p;
        }
    for (int sc = 0; sc < S; ++sc) {
        const std::vector<int> rb =
            build_row_blocks(h.ptr.data() + static_cast<size_t>(sc) * (rows + 1), rows);
        const int nb = static_cast<int>(rb.size() / 2) - 1;
        for (int k = 0; k < nb; ++k) {
            h.quad.insert(h.quad.end(), {rb[2 * k], rb[2 * k + 2], rb[2 * k + 1], rb[2 * k + 3]});
            h.bslice.push_back(sc);
        }
    }
    if (gather_order) {
        // sort each block of two or more rows by column: a wavefront's
        // gathers then fall on few cache lines wherever columns are hot
        // (the renumbered heavy columns); slot restores row-major order
        h.slot.assign(nnz, 0);
        std::vector<int> ord, ti, ts;
        for (size_t k = 0; k < h.bslice.size(); ++k) {
            const int r0 = h.quad[4 * k], r1 = h.quad[4 * k + 1];
            const int z0 = h.quad[4 * k + 2], z1 = h.quad[4 * k + 3];
            const int nz = z1 - z0;
            if (nz > NNZ_MAX) continue;  // a long single row: summed in storage order
            ord.resize(nz);
            std::iota(ord.begin(), ord.end(), 0);
            if (r1 - r0 > 1)
                std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
                    return h.idx[z0 + a] < h.idx[z0 + b];
                });
            ti.resize(nz);
            ts.resize(nz);
            for (int t = 0; t < nz; ++t) {
                ti[t] = h.idx[z0 + ord[t]];
                ts[t] = h.src[z0 + ord[t]];
                h.slot[z0 + t] = static_cast<unsigned short>(ord[t]);
            }
            std::copy(ti.begin(), ti.end(), h.idx.begin() + z0);
            std::copy(ts.begin(), ts.end(), h.src.begin() + z0);
        }
    }
    return h;
}

std::vector<int> build_xcd_map(const std::vector<int>& bslice, int S) {
    std::vector<std::vector<int>> of_slice(S);
    for (int k = 0; k < static_cast<int>(bslice.size()); ++k) of_slice[bslice[k]].push_back(k);
    std::vector<std::vector<int>> lists(XCDS);
    for (int sc = 0; sc < S; ++sc) {
        std::vector<int> xs;  // XCDs serving slice sc
        for (int x = 0; x < XCDS; ++x)
            if ((S <= XCDS ? x % S : sc % XCDS) == (S <= XCDS ? sc : x)) xs.push_back(x);
        const std::vector<int>& blk = of_slice[sc];
        const size_t g = xs.size(), nb = blk.size();
        for (size_t i = 0; i < g; ++i)
            for (size_t k = nb * i / g; k < nb * (i + 1) / g; ++k) lists[xs[i]].push_back(blk[k]);
    }
    size_t len = 0;
    for (const auto& l : lists) len = std::max(len, l.size());
    std::vector<int> map(len * XCDS, -1);
    for (int x = 0; x < XCDS; ++x)
        for (size_t k = 0; k < lists[x].size(); ++k) map[k * XCDS + x] = lists[x][k];
    return map;
}

template <typename V, bool NTL>
__global__ __launch_bounds__(BS) void k_primal_slices_x(IterBufs<V> b) {
    __shared__ V s_prod[NNZ_MAX];
    __shared__ V s_rows[ROWS_MAX];
    __shared__ V red_v[BS / WAVE];
    const LsState st = *b.st;
    if (st.frozen || !st.fresh) return;  // a retry reuses the stored A'y
    const int blk = b.sT.xmap[blockIdx.x];
    if (blk < 0) return;
    slice_rows<NTL>(b.sT, blk, b.y[st.cur], s_prod, s_rows, red_v);
}

template <typename V, bool NTL>
__global__ __launch_bounds__(BS) void k_dual_slices_x(IterBufs<V> b) {
    __shared__ V s_prod[NNZ_MAX];
    __shared__ V s_rows[ROWS_MAX];
    __shared__ V red_v[BS / WAVE];
    if (b.st->frozen) return;
    const int blk = b.sA.xmap[blockIdx.x];
    if (blk < 0) return;
    slice_rows<NTL>(b.sA, blk, b.xbar, s_prod, s_rows, red_v);
}

template <typename V, bool NTL>
static void primal_launch(const IterBufs<V>& b, hipStream_t s) {
    if (b.sT.S && b.sT.xmap) {
        if (b.sT.xgrid > 0) k_primal_slices_x<V, NTL><<<b.sT.xgrid, BS, 0, s>>>(b);
        if (b.nT_part > 0) k_primal_epi<V><<<b.nT_part, BS, 0, s>>>(b);
    } else if (b.sT.S) {
        for (int sc = 0; sc < b.sT.S; ++sc) {
            const int nb = b.sT.sb[sc + 1] - b.sT.sb[sc];
            if (nb > 0) k_primal_slices<V, NTL><<<nb, BS, 0, s>>>(b, b.sT.sb[sc]);
        }
        if (b.nT_part > 0) k_primal_epi<V><<<b.nT_part, BS, 0, s>>>(b);
    } else if (b.rbT.nblk > 0) {
        k_primal<V, NTL><<<b.rbT.nblk, BS, 0, s>>>(b);
    }
}
template <typename V>
static void primal_impl(const IterBufs<V>& b, hipStream_t s) {
    if (b.nt) primal_launch<V, true>(b, s);
    else primal_launch<V, false>(b, s);
}
template <typename V, bool NTL>
static void dual_launch(const IterBufs<V>& b, hipStream_t s) {
    if (b.sA.S && b.sA.xmap) {
        if (b.sA.xgrid > 0) k_dual_slices_x<V, NTL><<<b.sA.xgrid, BS, 0, s>>>(b);
        k_dual_epi<V><<<grid_of(b.nA_part), BS, 0, s>>>(b);
    } else if (b.sA.S) {
        for (int sc = 0; sc < b.sA.S; ++sc) {
            const int nb = b.sA.sb[sc + 1] - b.sA.sb[sc];
            if (nb > 0) k_dual_slices<V, NTL><<<nb, BS, 0, s>>>(b, b.sA.sb[sc]);
        }
        k_dual_epi<V><<<grid_of(b.nA_part), BS, 0, s>>>(b);
    } else {
        k_dual<V, NTL><<<grid_of(b.rbA.nblk), BS, 0, s>>>(b);
    }
}
template <typename V>
static void dual_impl(const IterBufs<V>& b, hipStream_t s) {
    if (b.nt) dual_launch<V, true>(b, s);
    else dual_launch<V, false>(b, s);
}
template <typename V>
static void attempt_impl(const IterBufs<V>& b, hipStream_t s) {
    primal_impl(b, s);
    dual_impl(b, s);
    k_decide<V><<<1, FIN, 0, s>>>(b);
}
void launch_attempt(const IterBufs<double>& b, hipStream_t s) { attempt_impl(b, s); }
void launch_primal(const IterBufs<double>& b, hipStream_t s) { primal_impl(b, s); }
void launch_primal(const IterBufs<float>& b, hipStream_t s) { primal_impl(b, s); }
void launch_dual(const IterBufs<double>& b, hipStream_t s) { dual_impl(b, s); }
void launch_dual(const IterBufs<float>& b, hipStream_t s) { dual_impl(b, s); }
void launch_decide(const IterBufs<double>& b, hipStream_t s) {
    k_decide<double><<<1, FIN, 0, s>>>(b);
}
void launch_decide(const IterBufs<float>& b, hipStream_t s) {
    k_decide<float><<<1, FIN, 0, s>>>(b);
}
void launch_attempt(const IterBufs<float>& b, hipStream_t s) { attempt_impl(b, s); }

template <typename V>
static void flush_impl(const IterBufs<V>& b, int cur, double w, hipStream_t s) {
    const long mx = b.n > b.m ? b.n : b.m;
    if (mx == 0) return;
    k_flush<V><<<(mx + BS - 1) / BS, BS, 0, s>>>(b.n, b.m, b.x[cur], b.xavg,
                                                  b.y[cur], b.yavg, b.adx,
                                                  b.axcur, static_cast<V>(w));
}
void launch_flush(const IterBufs<double>& b, int cur, double w, hipStream_t s) { flush_impl(b, cur, w, s); }
void launch_flush(const IterBufs<float>& b, int cur, double w, hipStream_t s) { flush_impl(b, cur, w, s); }

// The flush as the device state asks for it (a no-op without a pending
// step).  The flag is cleared by a later kernel (k_kkt_done with
// KktData::clear_mix): a block of the flush must not see it cleared by
// another.
template <typename V>
__global__ __launch_bounds__(BS) void k_flush_state(IterBufs<V> b) {
    const LsState* st = b.st;
    if (!st->mix_pending) return;
    const int cur = st->cur;
    const V w = static_cast<V>(st->mix_w);
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < b.n) b.xavg[i] += w * (b.x[cur][i] - b.xavg[i]);
    if (i < b.m) {
        b.yavg[i] += w * (b.y[cur][i] - b.yavg[i]);
        b.axcur[i] += b.adx[i];
    }
}
template <typename V>
static void flush_state_impl(const IterBufs<V>& b, hipStream_t s) {
    const long mx = b.n > b.m ? b.n : b.m;
    if (mx > 0) k_flush_state<V><<<(mx + BS - 1) / BS, BS, 0, s>>>(b);
}
void launch_flush_state(const IterBufs<double>& b, hipStream_t s) { flush_state_impl(b, s); }
void launch_flush_state(const IterBufs<float>& b, hipStream_t s) { flush_state_impl(b, s); }

// ------------------------------------------------------------------ KKT --
// Column pass: d = c + (A~'y~)/dc, dinf numerator, c'x and the column part
// of the Lagrangian bound (pdhg.cu:206-233).  Row pass: Ax = dr (A~x~),
// pinf / prow numerators and max|y| (pdhg.cu:180-201, 229).  Second row
// pass: the row part of the bound, which needs max|y| (pdhg.cu:234-244).

template <typename V>
__global__ __launch_bounds__(BS) void k_kkt_cols(KktData d, KktIter<V> it) {
    __shared__ double s_prod[NNZ_MAX];
    __shared__ double s_rows[ROWS_MAX];
    __shared__ double red[4 * BS / WAVE];
    const int q = blockIdx.y;
    int r0, r1, p0, p1;
    block_rows(d.rbT, blockIdx.x, r0, r1, p0, p1);
    rowblock_sums<double>(d.cp, d.ci, d.sacx, GatherRef<V>{it.yref[q], it.yd[q]},
                          r0, r1, p0, p1, s_rows, s_prod, red);
    const GatherRef<V> gx{it.xref[q], it.xd[q]};
    const double dz = 1e-8 * (1.0 + d.cmax);
    double dv = 0.0, op = 0.0, od = 0.0, unb = 0.0;
    for (int t = threadIdx.x; t < r1 - r0; t += BS) {
        const int j = r0 + t;
        const double ata = s_rows[t];
        if (it.ata_out[q]) it.ata_out[q][j] = ata;
        const double hx = gx(j) * d.dc[j];
        const double hd = d.c[j] + ata / d.dc[j];
        const double cl = d.cl[j], cu = d.cu[j];
        const double bnd = fabs(__builtin_isfinite(cu) ? cu : cl);
        const double tol = 1e-9 * (1.0 + bnd);
        double v;
        if (hx <= cl + tol && hd > 0) v = 0.0;
        else if (hx >= cu - tol && hd < 0) v = 0.0;
        else v = fabs(hd);
        dv = fmax(dv, v);
        op += d.c[j] * hx;
        if (hd >= 0) {
            if (__builtin_isfinite(cl)) od += hd * cl;
            else if (hd > dz) unb = 1.0;
        } else {
            if (__builtin_isfinite(cu)) od += hd * cu;
            else if (-hd > dz) unb = 1.0;
        }
    }
    double v[4] = {dv, op, od, unb};
    block_reduce_k<4, BS>(v, 0x9u, red);  // maxima: dv, unb
    if (threadIdx.x == 0) {
        double* p = d.partC + (static_cast<long>(q) * gridDim.x + blockIdx.x) * 4;
        for (int k = 0; k < 4; ++k) p[k] = v[k];
    }
}

template <typename V>
__global__ __launch_bounds__(BS) void k_kkt_rows(KktData d, KktIter<V> it) {
    __shared__ double s_prod[NNZ_MAX];
    __shared__ double s_rows[ROWS_MAX];
    __shared__ double red[4 * BS / WAVE];
    const int q = blockIdx.y;
    int r0, r1, p0, p1;
    block_rows(d.rbA, blockIdx.x, r0, r1, p0, p1);
    rowblock_sums<double>(d.ap, d.ai, d.sax, GatherRef<V>{it.xref[q], it.xd[q]},
                          r0, r1, p0, p1, s_rows, s_prod, red);
    const GatherRef<V> gy{it.yref[q], it.yd[q]};
    double pv = 0.0, pa = 0.0, prow = 0.0, ymax = 0.0;
    for (int t = threadIdx.x; t < r1 - r0; t += BS) {
        const int i = r0 + t;
        const double tm = s_rows[t];
        if (it.tm_out[q]) it.tm_out[q][i] = tm;
        const double ax = tm * d.dr[i];
        const double lo = d.rmin[i], hi = d.rmax[i];
        const double e = fmax(lo - ax, ax - hi);
        pv = fmax(pv, e);
        pa = fmax(pa, fabs(fmin(fmax(ax, lo), hi)));
        const double viol = fmax(0.0, e);
        double sc = fabs(ax);
        if (__builtin_isfinite(lo)) sc = fmax(sc, fabs(lo));
        if (__builtin_isfinite(hi)) sc = fmax(sc, fabs(hi));
        prow = fmax(prow, viol / (1.0 + sc));
        const double hy = gy(i) * (1.0 / d.dr[i]);
        ymax = fmax(ymax, fabs(hy));
    }
    double v[4] = {pv, pa, prow, ymax};
    block_reduce_k<4, BS>(v, 0xFu, red);
    if (threadIdx.x == 0) {
        double* p = d.partR + (static_cast<long>(q) * gridDim.x + blockIdx.x) * 4;
        for (int k = 0; k < 4; ++k) p[k] = v[k];
    }
}

// Row-pass maxima of both iterates -> rowres[q*4 + k].
__global__ __launch_bounds__(FIN) void k_kkt_rowfin(KktData d, int nb) {
    __shared__ double red[8 * FIN / WAVE];
    double v[8];
    for (int q = 0; q < 2; ++q)
        for (int k = 0; k < 4; ++k)
            v[q * 4 + k] = fin_part_max(d.partR + static_cast<long>(q) * nb * 4 + k, nb, 4);
    block_reduce_k<8, FIN>(v, 0xFFu, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; ++k) d.rowres[k] = v[k];
}

template <typename V>
__global__ __launch_bounds__(BS) void k_kkt_bound(KktData d, KktIter<V> it) {
    __shared__ double red[2 * BS / WAVE];
    const int q = blockIdx.y;
    const double yz = 1e-8 * (1.0 + d.rowres[q * 4 + 3]);
    const GatherRef<V> gy{it.yref[q], it.yd[q]};
    double od = 0.0, unb = 0.0;
    for (long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x; i < d.m;
         i += static_cast<long>(gridDim.x) * BS) {
        const double hy = gy(static_cast<int>(i)) * (1.0 / d.dr[i]);
        if (hy > yz) {
            if (__builtin_isfinite(d.rmax[i])) od -= hy * d.rmax[i];
            else unb = 1.0;
        } else if (hy < -yz) {
            if (__builtin_isfinite(d.rmin[i])) od -= hy * d.rmin[i];
            else unb = 1.0;
        }
    }
    double v[2] = {od, unb};
    block_reduce_k<2, BS>(v, 0x2u, red);
    if (threadIdx.x == 0) {
        double* p = d.partR2 + (static_cast<long>(q) * gridDim.x + blockIdx.x) * 2;
        p[0] = v[0];
        p[1] = v[1];
    }
}

// Final metrics of both iterates (pdhg.cu:195-253).
__global__ __launch_bounds__(FIN) void k_kkt_done(KktData d, int nbC, int nbB,
                                                  Metrics* out) {
    __shared__ double red[12 * FIN / WAVE];
    double v[12];
    for (int q = 0; q < 2; ++q) {
        const double* pc = d.partC + static_cast<long>(q) * nbC * 4;
        const double* pb = d.partR2 + static_cast<long>(q) * nbB * 2;
        v[6 * q + 0] = fin_part_max(pc + 0, nbC, 4);
        v[6 * q + 1] = fin_part_sum(pc + 1, nbC, 4);
        v[6 * q + 2] = fin_part_sum(pc + 2, nbC, 4);
        v[6 * q + 3] = fin_part_max(pc + 3, nbC, 4);
        v[6 * q + 4] = fin_part_sum(pb + 0, nbB, 2);
        v[6 * q + 5] = fin_part_max(pb + 1, nbB, 2);
    }
    block_reduce_k<12, FIN>(v, 0x29u | (0x29u << 6), red);  // maxima: dv, unbc, unbr
    for (int q = 0; q < 2; ++q) {
        const double dv = v[6 * q + 0], op = v[6 * q + 1], odc = v[6 * q + 2];
        const double unbc = v[6 * q + 3], odr = v[6 * q + 4], unbr = v[6 * q + 5];
        if (threadIdx.x == 0) {
            const double pv = d.rowres[q * 4 + 0];
            const double pa = d.rowres[q * 4 + 1];
            const double prow = d.rowres[q * 4 + 2];
            Metrics mt;
            mt.pinf = fmax(pv, 0.0) / (1.0 + pa);
            mt.prow = prow;
            mt.dinf = dv / (1.0 + d.cmax);
            mt.op = op;
            const bool bounded = unbc == 0.0 && unbr == 0.0;
            const double odv = odc + odr;
            mt.od = bounded ? odv : -__builtin_huge_val();
            mt.gap = bounded ? fmax(0.0, op - odv) / (1.0 + fabs(op))
                             : __builtin_huge_val();
            out[q] = mt;
        }
    }
    if (threadIdx.x == 0 && d.clear_mix) d.clear_mix->mix_pending = 0;
}

static int kkt_rows2_blocks(int m) {
    long b = (static_cast<long>(m) + BS * 4 - 1) / (BS * 4);
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return static_cast<int>(b);
}

template <typename V>
static void kkt_impl(const KktData& d, const KktIter<V>& it, Metrics* out,
                     hipStream_t s) {
    const int nbC = grid_of(d.rbT.nblk);
    const int nbR = grid_of(d.rbA.nblk);
    const int nbB = kkt_rows2_blocks(d.m);
    k_kkt_cols<V><<<dim3(nbC, 2), BS, 0, s>>>(d, it);
    k_kkt_rows<V><<<dim3(nbR, 2), BS, 0, s>>>(d, it);
    k_kkt_rowfin<<<1, FIN, 0, s>>>(d, nbR);
    k_kkt_bound<V><<<dim3(nbB, 2), BS, 0, s>>>(d, it);
    k_kkt_done<<<1, FIN, 0, s>>>(d, nbC, nbB, out);
}
void launch_kkt(const KktData& d, const KktIter<double>& it, Metrics* out,
                hipStream_t s) { kkt_impl(d, it, out, s); }
void launch_kkt(const KktData& d, const KktIter<float>& it, Metrics* out,
                hipStream_t s) { kkt_impl(d, it, out, s); }

// ------------------------------------------------------- Ruiz / scale --
// One wavefront per row; a row's maximum is exact in any order, so the
// factors match a sequential host loop bit for bit.

__global__ __launch_bounds__(BS) void k_ruiz_rows(int m, const int* ap,
                                                  const int* ai,
                                                  const double* ax,
                                                  const double* dc,
                                                  double* dr) {
    const long wid = (static_cast<long>(blockIdx.x) * BS + threadIdx.x) / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    if (wid >= m) return;
    const int i = static_cast<int>(wid);
    const double dri = dr[i];
    double mx = 0.0;
    for (int p = ap[i] + lane; p < ap[i + 1]; p += WAVE)
        mx = fmax(mx, fabs(ax[p]) * dc[ai[p]] / dri);
    mx = wave_max(mx);
    if (lane == 0 && mx > 0 && __builtin_isfinite(mx)) dr[i] = dri * sqrt(mx);
}

__global__ __launch_bounds__(BS) void k_ruiz_cols(int n, const int* cp,
                                                  const int* ci,
                                                  const double* acx,
                                                  const double* dr,
                                                  double* dc) {
    const long wid = (static_cast<long>(blockIdx.x) * BS + threadIdx.x) / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    if (wid >= n) return;
    const int j = static_cast<int>(wid);
    const double dcj = dc[j];
    double mx = 0.0;
    for (int p = cp[j] + lane; p < cp[j + 1]; p += WAVE)
        mx = fmax(mx, fabs(acx[p]) / dr[ci[p]] * dcj);
    mx = wave_max(mx);
    if (lane == 0 && mx > 0 && __builtin_isfinite(mx)) dc[j] = dcj / sqrt(mx);
}

__global__ void k_fix_ones(double* v, long len) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < len && !(v[i] > 0 && __builtin_isfinite(v[i]))) v[i] = 1.0;
}

static int wave_grid(long rows) {
    return static_cast<int>((rows * WAVE + BS - 1) / BS);
}
static int elem_grid(long len) { return static_cast<int>((len + BS - 1) / BS); }

void fill(double* v, long n, double val, hipStream_t s);

void ruiz(int m, int n, const int* ap, const int* ai, const double* ax,
          const int* cp, const int* ci, const double* acx, double* dr,
          double* dc, hipStream_t s) {
    fill(dr, m, 1.0, s);
    fill(dc, n, 1.0, s);
    for (int sweep = 0; sweep < 10; ++sweep) {
        if (m > 0) k_ruiz_rows<<<wave_grid(m), BS, 0, s>>>(m, ap, ai, ax, dc, dr);
        if (n > 0) k_ruiz_cols<<<wave_grid(n), BS, 0, s>>>(n, cp, ci, acx, dr, dc);
    }
    if (m > 0) k_fix_ones<<<elem_grid(m), BS, 0, s>>>(dr, m);
    if (n > 0) k_fix_ones<<<elem_grid(n), BS, 0, s>>>(dc, n);
}

// Pock-Chambolle (alpha = 1) on top of the Ruiz factors, opt-in: row and
// column absolute sums of the currently scaled matrix, one thread per row /
// column in storage order (the same sums as a sequential host loop), then
// dr_i *= sqrt(r_i), dc_j /= sqrt(c_j).  The reference implements Ruiz only
// (src/pdhg/pdhg.cu:469-492); its algorithm sheet pairs the two
// (docs/research/pdlp-algorithm-sheet.md:62).
__global__ __launch_bounds__(BS) void k_pc_sums(int rows, const int* ptr, const int* idx,
                                                const double* val, const double* dr,
                                                const double* dc, int by_col,
                                                double* out) {
    const long r = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (r >= rows) return;
    double s = 0.0;
    for (int p = ptr[r]; p < ptr[r + 1]; ++p) {
        const double t = by_col ? fabs(val[p]) * dc[r] / dr[idx[p]]
                                : fabs(val[p]) * dc[idx[p]] / dr[r];
        s = s + t;
    }
    out[r] = s;
}
__global__ void k_pc_apply(int m, int n, const double* rs, const double* cs, double* dr,
                           double* dc) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < m && rs[i] > 0 && __builtin_isfinite(rs[i])) dr[i] *= sqrt(rs[i]);
    if (i < n && cs[i] > 0 && __builtin_isfinite(cs[i])) dc[i] /= sqrt(cs[i]);
}
void pock_chambolle(int m, int n, const int* ap, const int* ai, const double* ax,
                    const int* cp, const int* ci, const double* acx, double* dr,
                    double* dc, double* rs, double* cs, hipStream_t s) {
    if (m > 0) k_pc_sums<<<elem_grid(m), BS, 0, s>>>(m, ap, ai, ax, dr, dc, 0, rs);
    if (n > 0) k_pc_sums<<<elem_grid(n), BS, 0, s>>>(n, cp, ci, acx, dr, dc, 1, cs);
    const long mx = m > n ? m : n;
    if (mx > 0) k_pc_apply<<<elem_grid(mx), BS, 0, s>>>(m, n, rs, cs, dr, dc);
}

// CSR: out = val * dc[col] / dr[row];  CSC: out = val * dc[col] / dr[row]
// with the CSC "row" being the column (pdhg.cu:503-512 operand order).
__global__ __launch_bounds__(BS) void k_scale_csr(int m, const int* ap,
                                                  const int* ai,
                                                  const double* ax,
                                                  const double* dc,
                                                  const double* dr,
                                                  double* sax) {
    const long wid = (static_cast<long>(blockIdx.x) * BS + threadIdx.x) / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    if (wid >= m) return;
    const int i = static_cast<int>(wid);
    const double d = dr[i];
    for (int p = ap[i] + lane; p < ap[i + 1]; p += WAVE)
        sax[p] = ax[p] * dc[ai[p]] / d;
}
__global__ __launch_bounds__(BS) void k_scale_csc(int n, const int* cp,
                                                  const int* ci,
                                                  const double* acx,
                                                  const double* dc,
                                                  const double* dr,
                                                  double* sacx) {
    const long wid = (static_cast<long>(blockIdx.x) * BS + threadIdx.x) / WAVE;
    const int lane = threadIdx.x & (WAVE - 1);
    if (wid >= n) return;
    const int j = static_cast<int>(wid);
    const double d = dc[j];
    for (int p = cp[j] + lane; p < cp[j + 1]; p += WAVE)
        sacx[p] = acx[p] * d / dr[ci[p]];
}
__global__ void k_scale_vecs(int m, int n, const double* dr, const double* dc,
                             const double* c, const double* cl,
                             const double* cu, const double* rmin,
                             const double* rmax, double* sc, double* scl,
                             double* scu, double* srmin, double* srmax) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < n) {
        sc[i] = c[i] * dc[i] / 1.0;
        scl[i] = cl[i] / dc[i];
        scu[i] = cu[i] / dc[i];
    }
    if (i < m) {
        srmin[i] = rmin[i] / dr[i];
        srmax[i] = rmax[i] / dr[i];
    }
}

void scale_data(int m, int n, const int* ap, const int* ai, const double* ax,
                const int* cp, const int* ci, const double* acx,
                const double* dr, const double* dc, const double* c,
                const double* cl, const double* cu, const double* rmin,
                const double* rmax, double* sax, double* sacx, double* sc,
                double* scl, double* scu, double* srmin, double* srmax,
                hipStream_t s) {
    if (m > 0) k_scale_csr<<<wave_grid(m), BS, 0, s>>>(m, ap, ai, ax, dc, dr, sax);
    if (n > 0) k_scale_csc<<<wave_grid(n), BS, 0, s>>>(n, cp, ci, acx, dc, dr, sacx);
    const long mx = m > n ? m : n;
    if (mx > 0)
        k_scale_vecs<<<elem_grid(mx), BS, 0, s>>>(m, n, dr, dc, c, cl, cu, rmin,
                                                  rmax, sc, scl, scu, srmin, srmax);
}

// ------------------------------------------------------- reductions --

int reduce_blocks(long n) {
    long b = (n + BS * 8 - 1) / (BS * 8);
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return static_cast<int>(b);
}

template <typename V>
__global__ __launch_bounds__(BS) void k_sumsq_diff(const V* a, const V* b,
                                                   long n, double* part) {
    __shared__ double red[BS / WAVE];
    double s = 0.0;
    for (long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * BS) {
        const double d = b ? static_cast<double>(a[i]) - static_cast<double>(b[i])
                           : static_cast<double>(a[i]);
        s += d * d;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(FIN) void k_fin(const double* part, int cnt,
                                             int is_max, double* out) {
    __shared__ double red[FIN / WAVE];
    const double t = is_max ? fin_max(part, cnt, 1, red) : fin_sum(part, cnt, 1, red);
    if (threadIdx.x == 0) *out = t;
}

void sumsq_diff(const double* a, const double* b, long n, double* part, double* out,
                hipStream_t s) {
    const int g = reduce_blocks(n);
    k_sumsq_diff<double><<<g, BS, 0, s>>>(a, b, n, part);
    k_fin<<<1, FIN, 0, s>>>(part, g, 0, out);
}
void sumsq_diff(const float* a, const float* b, long n, double* part, double* out,
                hipStream_t s) {
    const int g = reduce_blocks(n);
    k_sumsq_diff<float><<<g, BS, 0, s>>>(a, b, n, part);
    k_fin<<<1, FIN, 0, s>>>(part, g, 0, out);
}

__global__ __launch_bounds__(BS) void k_sumsq_q(const double* srmin,
                                                const double* srmax, long m,
                                                double* part) {
    __shared__ double red[BS / WAVE];
    double s = 0.0;
    for (long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x; i < m;
         i += static_cast<long>(gridDim.x) * BS) {
        const double q = __builtin_isfinite(srmax[i])
                             ? srmax[i]
                             : (__builtin_isfinite(srmin[i]) ? srmin[i] : 0.0);
        s += q * q;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
void sumsq_q(const double* srmin, const double* srmax, long m, double* part,
             double* out, hipStream_t s) {
    const int g = reduce_blocks(m);
    k_sumsq_q<<<g, BS, 0, s>>>(srmin, srmax, m, part);
    k_fin<<<1, FIN, 0, s>>>(part, g, 0, out);
}

__global__ __launch_bounds__(BS) void k_absmax(const double* v, long n,
                                               double* part) {
    __shared__ double red[BS / WAVE];
    double s = 0.0;
    for (long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * BS)
        s = fmax(s, fabs(v[i]));
    s = block_max(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
void absmax(const double* v, long n, double* part,
            double* out, hipStream_t s) {
    const int g = reduce_blocks(n);
    k_absmax<<<g, BS, 0, s>>>(v, n, part);
    k_fin<<<1, FIN, 0, s>>>(part, g, 1, out);
}

// ------------------------------------------------------ elementwise --

template <typename T>
__global__ void k_fill(T* v, long n, T val) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < n) v[i] = val;
}
void fill(double* v, long n, double val, hipStream_t s) {
    if (n > 0) k_fill<double><<<elem_grid(n), BS, 0, s>>>(v, n, val);
}
void fill(float* v, long n, float val, hipStream_t s) {
    if (n > 0) k_fill<float><<<elem_grid(n), BS, 0, s>>>(v, n, val);
}

__global__ void k_scale_into(double* out, const double* in, long n, double f) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < n) out[i] = in[i] * f;
}
void scale_into(double* out, const double* in, long n, double f, hipStream_t s) {
    if (n > 0) k_scale_into<<<elem_grid(n), BS, 0, s>>>(out, in, n, f);
}

__global__ void k_to_float(const double* in, float* out, long n) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < n) out[i] = static_cast<float>(in[i]);
}
void to_float(const double* in, float* out, long n, hipStream_t s) {
    if (n > 0) k_to_float<<<elem_grid(n), BS, 0, s>>>(in, out, n);
}

// ref + d, or the bound itself where d sits exactly on its clamp bound
__device__ __forceinline__ double snapped(const double* ref, const float* d, long i,
                                          BoundSnap b) {
    const float v = d[i];
    if (b.hi) {
        if (v == b.hi[i]) return b.scu[i];
        if (v == b.lo[i]) return b.scl[i];
    }
    const double dv = static_cast<double>(v);
    return ref ? ref[i] + dv : dv;
}

__global__ void k_combine(const double* ref, const float* d, double* out, long n,
                          BoundSnap b) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < n) out[i] = snapped(ref, d, i, b);
}
void combine(const double* ref, const float* d, double* out, long n, hipStream_t s,
             const BoundSnap* snap) {
    if (n > 0)
        k_combine<<<elem_grid(n), BS, 0, s>>>(ref, d, out, n, snap ? *snap : BoundSnap{});
}

__global__ void k_combine4(Combine4 c) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < c.n) {
        c.ox[0][i] = snapped(c.xref, c.xd[0], i, c.snap);
        c.ox[1][i] = snapped(c.xref, c.xd[1], i, c.snap);
    }
    if (i < c.m) {
        c.oy[0][i] = snapped(c.yref, c.yd[0], i, BoundSnap{});
        c.oy[1][i] = snapped(c.yref, c.yd[1], i, BoundSnap{});
    }
}
void combine4(const Combine4& c, hipStream_t s) {
    const long mx = c.n > c.m ? c.n : c.m;
    if (mx > 0) k_combine4<<<elem_grid(mx), BS, 0, s>>>(c);
}

__global__ void k_fold(double* ref, float* d, long n, BoundSnap b) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < n) {
        ref[i] = snapped(ref, d, i, b);
        d[i] = 0.0f;
    }
}
void fold(double* ref, float* d, long n, hipStream_t s, const BoundSnap* snap) {
    if (n > 0) k_fold<<<elem_grid(n), BS, 0, s>>>(ref, d, n, snap ? *snap : BoundSnap{});
}

__global__ void k_refine_data(int m, int n, const double* scl,
                              const double* scu, const double* sc,
                              const double* xref, const double* ata_ref,
                              const double* srmin, const double* srmax,
                              const double* aref, const double* yref,
                              float* lo, float* hi, float* gref, float* rlo,
                              float* rhi, float* yoff) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < n) {
        lo[i] = static_cast<float>(scl[i] - xref[i]);
        hi[i] = static_cast<float>(scu[i] - xref[i]);
        gref[i] = static_cast<float>(sc[i] + ata_ref[i]);
    }
    if (i < m) {
        rlo[i] = static_cast<float>(srmin[i] - aref[i]);
        rhi[i] = static_cast<float>(srmax[i] - aref[i]);
        yoff[i] = static_cast<float>(yref[i]);
    }
}
void refine_data(int m, int n, const double* scl, const double* scu,
                 const double* sc, const double* xref, const double* ata_ref,
                 const double* srmin, const double* srmax,
                 const double* aref, const double* yref, float* lo, float* hi,
                 float* gref, float* rlo, float* rhi, float* yoff,
                 hipStream_t s) {
    const long mx = m > n ? m : n;
    if (mx > 0)
        k_refine_data<<<elem_grid(mx), BS, 0, s>>>(m, n, scl, scu, sc, xref,
                                                   ata_ref, srmin, srmax, aref,
                                                   yref, lo, hi, gref, rlo,
                                                   rhi, yoff);
}

__global__ void k_unscale_out(int m, int n, const double* dc, const double* dr,
                              const double* xref, const double* xd64,
                              const float* xd32, const double* yref,
                              const double* yd64, const float* yd32,
                              double* x, double* y, BoundSnap b) {
    const long i = static_cast<long>(blockIdx.x) * BS + threadIdx.x;
    if (i < n) {
        double sx;
        if (xd64) {
            sx = xd64[i];
            if (xref) sx = xref[i] + sx;
        } else {
            sx = snapped(xref, xd32, i, b);
        }
        x[i] = sx * dc[i];
    }
    if (i < m) {
        double sy = yd64 ? yd64[i] : static_cast<double>(yd32[i]);
        if (yref) sy = yref[i] + sy;
        y[i] = 1.0 * sy / dr[i];
    }
}
void unscale_out(int m, int n, const double* dc, const double* dr,
                 const double* xref, const double* xd64, const float* xd32,
                 const double* yref, const double* yd64, const float* yd32,
                 double* x, double* y, hipStream_t s, const BoundSnap* snap) {
    const long mx = m > n ? m : n;
    if (mx > 0)
        k_unscale_out<<<elem_grid(mx), BS, 0, s>>>(m, n, dc, dr, xref, xd64, xd32, yref,
                                                   yd64, yd32, x, y,
                                                   snap ? *snap : BoundSnap{});
}

}  // namespace igaos::dev
