#pragma once
// Device-side building blocks shared by the single-LP and batched PDHG
// kernels: wavefront/workgroup reductions and the row-block SpMV core.
//
// Row-block SpMV (CSR-Adaptive, Greathouse & Daga, SC'14): rows are cut on
// the host into contiguous blocks holding at most NNZ_MAX nonzeros and at
// most ROWS_MAX rows; a row longer than NNZ_MAX gets a block of its own.
// A multi-row block streams its values and column indices with fully
// coalesced loads into LDS as products val*x[col], then 2^k lanes per row
// (k chosen from the block's row count) reduce each row with xor shuffles
// inside a 64-lane wavefront.  A single long row is reduced by the whole
// workgroup.  Results land in LDS so that the fused epilogue walks the
// block's rows with consecutive threads (coalesced vector reads/writes).
// The summation order depends only on the block layout, so results are
// run-to-run deterministic.

#include <hip/hip_runtime.h>

#include <cstdint>

namespace igaos::dev {

constexpr int WAVE = 64;
constexpr int BS = 256;         // threads per row-block workgroup
constexpr int NNZ_MAX = 2048;   // nonzeros staged per multi-row block
constexpr int ROWS_MAX = 256;   // rows per block (one epilogue row per thread)

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// Workgroup sum with a fixed combination order; every thread gets the
// result.  `red` must hold NT/WAVE entries.  Contains __syncthreads().
template <typename T, int NT = BS>
__device__ __forceinline__ T block_sum(T v, T* red) {
    v = wave_sum(v);
    const int w = threadIdx.x / WAVE;
    __syncthreads();
    if ((threadIdx.x & (WAVE - 1)) == 0) red[w] = v;
    __syncthreads();
    T s = red[0];
#pragma unroll
    for (int i = 1; i < NT / WAVE; ++i) s += red[i];
    return s;
}

template <typename T, int NT = BS>
__device__ __forceinline__ T block_max(T v, T* red) {
    v = wave_max(v);
    const int w = threadIdx.x / WAVE;
    __syncthreads();
    if ((threadIdx.x & (WAVE - 1)) == 0) red[w] = v;
    __syncthreads();
    T s = red[0];
#pragma unroll
    for (int i = 1; i < NT / WAVE; ++i) s = fmax(s, red[i]);
    return s;
}

// Reductions across workgroups: every workgroup of a kernel writes one
// partial per quantity, and a follow-up kernel of one FIN-thread workgroup
// combines them in a fixed order (thread t takes partials t, t+FIN, ...;
// then the workgroup tree).  No atomics and no cross-workgroup fences, so
// results are run-to-run deterministic.
constexpr int FIN = 1024;

// This thread's part of fin_sum / fin_max.
__device__ __forceinline__ double fin_part_sum(const double* p, long cnt, long stride) {
    double s0 = 0.0, s1 = 0.0;
    long i = threadIdx.x;
    for (; i + FIN < cnt; i += 2 * FIN) {
        s0 += p[i * stride];
        s1 += p[(i + FIN) * stride];
    }
    if (i < cnt) s0 += p[i * stride];
    return s0 + s1;
}
__device__ __forceinline__ double fin_part_max(const double* p, long cnt, long stride) {
    double s = 0.0;
    for (long i = threadIdx.x; i < cnt; i += FIN) s = fmax(s, p[i * stride]);
    return s;
}
__device__ __forceinline__ double fin_sum(const double* p, long cnt,
                                          long stride, double* red) {
    return block_sum<double, FIN>(fin_part_sum(p, cnt, stride), red);
}
__device__ __forceinline__ double fin_max(const double* p, long cnt,
                                          long stride, double* red) {
    return block_max<double, FIN>(fin_part_max(p, cnt, stride), red);
}

// K workgroup reductions at once: v[k] becomes, in every thread, the sum
// of v[k] over the workgroup (the maximum where bit k of max_mask is set),
// combined in block_sum's (block_max's) order, so to the same value; one
// pair of barriers for all K.  `red` must hold K * NT / WAVE entries.
template <int K, int NT>
__device__ __forceinline__ void block_reduce_k(double* v, unsigned max_mask, double* red) {
    constexpr int W = NT / WAVE;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = (max_mask >> k & 1u) ? wave_max(v[k]) : wave_sum(v[k]);
    const int w = threadIdx.x / WAVE;
    __syncthreads();
    if ((threadIdx.x & (WAVE - 1)) == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) red[k * W + w] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const bool mx = max_mask >> k & 1u;
        double s = red[k * W];
        for (int i = 1; i < W; ++i) s = mx ? fmax(s, red[k * W + i]) : s + red[k * W + i];
        v[k] = s;
    }
}

// Gather functors: x[j] for plain vectors, xref[j] + (double)xd[j] for an
// FP64 reference plus an FP32 correction.
template <typename V>
struct GatherPlain {
    const V* x;
    __device__ __forceinline__ V operator()(int j) const { return x[j]; }
};
template <typename V>
struct GatherRef {
    const double* ref;
    const V* d;
    __device__ __forceinline__ double operator()(int j) const {
        return ref ? ref[j] + static_cast<double>(d[j])
                   : static_cast<double>(d[j]);
    }
};

// Row-block layout: rb[2b] = first row of block b, rb[2b+1] = its first
// nonzero; the pair after the last block holds (nrows, nnz).
struct RowBlocks {
    const int* rb = nullptr;
    int nblk = 0;
};

__device__ __forceinline__ void block_rows(const RowBlocks& rb, int blk,
                                           int& r0, int& r1, int& p0,
                                           int& p1) {
    if (blk < rb.nblk) {
        const int2 u = *reinterpret_cast<const int2*>(rb.rb + 2 * blk);
        const int2 v = *reinterpret_cast<const int2*>(rb.rb + 2 * blk + 2);
        r0 = u.x;
        p0 = u.y;
        r1 = v.x;
        p1 = v.y;
    } else {
        r0 = r1 = p0 = p1 = 0;
    }
}

// A load of a matrix stream, with a nontemporal hint when NTL.
template <bool NTL, typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
    if constexpr (NTL) return __builtin_nontemporal_load(p);
    else return *p;
}

// Row sums of one row block: s_rows[r - r0] = sum_p val[p] * x(col[p]).
// `s_prod` holds NNZ_MAX entries of Acc, `s_rows` ROWS_MAX, `red`
// BS/WAVE.  Contains __syncthreads(); all threads must call it.
//
// Multi-row block: every thread issues its (at most NNZ_MAX/BS = 8) index
// and value loads first, then the 8 gathers, then the products into LDS,
// so a thread has 8 independent HBM requests in flight.  2^k threads per
// row then reduce: inside a wavefront with xor shuffles, across the 2 or 4
// wavefronts of a block through LDS when the block holds only 1-2 rows.
// Single row longer than NNZ_MAX: the workgroup strides over it four
// loads deep per thread.  Column indices are int, or 16-bit where every
// index fits (the batched solver's instances).
//
// `slot` (optional): the block's nonzeros are stored in another order
// (idx, val permuted alike) and nonzero q of the block belongs at position
// slot[q] of the block's row-major order; products are placed there, so
// every row sums exactly as without the permutation.  A block of a single
// row keeps its own order (slot[q] = q).
template <typename Acc, typename Val, typename G, typename I>
__device__ __forceinline__ void rowblock_sums(const int* __restrict__ ptr,
                                              const I* __restrict__ idx,
                                              const Val* __restrict__ val,
                                              const G& gx, int r0, int r1,
                                              int p0, int p1, Acc* s_rows,
                                              Acc* s_prod, Acc* red,
                                              const unsigned short* __restrict__ slot =
                                                  nullptr) {
    const int tid = threadIdx.x;
    const int nrows = r1 - r0;
    const int nz = p1 - p0;
    if (nrows == 1 && nz > NNZ_MAX) {
        Acc s0 = Acc(0), s1 = Acc(0), s2 = Acc(0), s3 = Acc(0);
        int p = p0 + tid;
        for (; p + 3 * BS < p1; p += 4 * BS) {
            const int c0 = idx[p], c1 = idx[p + BS], c2 = idx[p + 2 * BS],
                      c3 = idx[p + 3 * BS];
            const Val v0 = val[p], v1 = val[p + BS], v2 = val[p + 2 * BS],
                      v3 = val[p + 3 * BS];
            s0 += static_cast<Acc>(v0) * static_cast<Acc>(gx(c0));
            s1 += static_cast<Acc>(v1) * static_cast<Acc>(gx(c1));
            s2 += static_cast<Acc>(v2) * static_cast<Acc>(gx(c2));
            s3 += static_cast<Acc>(v3) * static_cast<Acc>(gx(c3));
        }
        for (; p < p1; p += BS) s0 += static_cast<Acc>(val[p]) * static_cast<Acc>(gx(idx[p]));
        Acc s = block_sum((s0 + s1) + (s2 + s3), red);
        if (tid == 0) s_rows[0] = s;
        __syncthreads();
        return;
    }
    int tpr = 1;
    while (tpr < BS && tpr * 2 * nrows <= BS) tpr *= 2;
    const int row = tid / tpr;
    const int lane = tid & (tpr - 1);
    int a = 0, b = 0;
    if (row < nrows) {  // row bounds, loaded early and independently
        a = ptr[r0 + row] - p0;
        b = ptr[r0 + row + 1] - p0;
    }
    if (nz > 0) {
        // Branch-free loads: slots past the block's end re-read its last
        // nonzero (same cache line), so all 8 requests issue back to back.
        constexpr int U = NNZ_MAX / BS;
        const int last = nz - 1;
        int ci[U];
        Val vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = min(tid + u * BS, last);
            ci[u] = idx[p0 + q];
            vv[u] = val[p0 + q];
        }
        if (slot) {
            int sq[U];
#pragma unroll
            for (int u = 0; u < U; ++u) sq[u] = slot[p0 + min(tid + u * BS, last)];
            Acc xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = static_cast<Acc>(gx(ci[u]));
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = tid + u * BS;
                if (q < nz) s_prod[sq[u]] = static_cast<Acc>(vv[u]) * xv[u];
            }
        } else {
            Acc xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = static_cast<Acc>(gx(ci[u]));
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = tid + u * BS;
                if (q < nz) s_prod[q] = static_cast<Acc>(vv[u]) * xv[u];
            }
        }
    }
    __syncthreads();
    Acc s = Acc(0);
    for (int q = a + lane; q < b; q += tpr) s += s_prod[q];
    const int wl = tpr < WAVE ? tpr : WAVE;
    for (int o = wl / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (tpr <= WAVE) {
        if (row < nrows && lane == 0) s_rows[row] = s;
    } else {
        if ((tid & (WAVE - 1)) == 0) red[tid / WAVE] = s;
        __syncthreads();
        if (lane == 0 && row < nrows) {
            const int wpr = tpr / WAVE;
            Acc t = red[row * wpr];
            for (int k = 1; k < wpr; ++k) t += red[row * wpr + k];
            s_rows[row] = t;
        }
    }
    __syncthreads();
}

// rowblock_sums with the matrix streams (indices, values, slots) loaded
// with nontemporal hints.  On gfx950 such loads stream from HBM at about
// 6 TB/s against about 4 TB/s for loads that allocate in the caches, and
// they leave L2 to the gathered vector; they only pay when the matrices
// are too large to stay in the 256 MB last-level cache between
// iterations.  Same sums as rowblock_sums.  A copy rather than a flag on
// rowblock_sums: the batch's persistent kernel (register-capped, with
// spills) inlines rowblock_sums, and a flag changed its code generation.
template <typename Acc, typename Val, typename G, typename I>
__device__ __forceinline__ void rowblock_sums_nt(const int* __restrict__ ptr,
                                                 const I* __restrict__ idx,
                                                 const Val* __restrict__ val,
                                                 const G& gx, int r0, int r1,
                                                 int p0, int p1, Acc* s_rows,
                                                 Acc* s_prod, Acc* red,
                                                 const unsigned short* __restrict__ slot) {
    const int tid = threadIdx.x;
    const int nrows = r1 - r0;
    const int nz = p1 - p0;
    if (nrows == 1 && nz > NNZ_MAX) {
        Acc s0 = Acc(0), s1 = Acc(0), s2 = Acc(0), s3 = Acc(0);
        int p = p0 + tid;
        for (; p + 3 * BS < p1; p += 4 * BS) {
            const int c0 = ld_stream<true>(idx + p), c1 = ld_stream<true>(idx + p + BS),
                      c2 = ld_stream<true>(idx + p + 2 * BS),
                      c3 = ld_stream<true>(idx + p + 3 * BS);
            const Val v0 = ld_stream<true>(val + p), v1 = ld_stream<true>(val + p + BS),
                      v2 = ld_stream<true>(val + p + 2 * BS),
                      v3 = ld_stream<true>(val + p + 3 * BS);
            s0 += static_cast<Acc>(v0) * static_cast<Acc>(gx(c0));
            s1 += static_cast<Acc>(v1) * static_cast<Acc>(gx(c1));
            s2 += static_cast<Acc>(v2) * static_cast<Acc>(gx(c2));
            s3 += static_cast<Acc>(v3) * static_cast<Acc>(gx(c3));
        }
        for (; p < p1; p += BS)
            s0 += static_cast<Acc>(ld_stream<true>(val + p)) *
                  static_cast<Acc>(gx(ld_stream<true>(idx + p)));
        Acc s = block_sum((s0 + s1) + (s2 + s3), red);
        if (tid == 0) s_rows[0] = s;
        __syncthreads();
        return;
    }
    int tpr = 1;
    while (tpr < BS && tpr * 2 * nrows <= BS) tpr *= 2;
    const int row = tid / tpr;
    const int lane = tid & (tpr - 1);
    int a = 0, b = 0;
    if (row < nrows) {  // row bounds, loaded early and independently
        a = ptr[r0 + row] - p0;
        b = ptr[r0 + row + 1] - p0;
    }
    if (nz > 0) {
        // Branch-free loads: slots past the block's end re-read its last
        // nonzero (same cache line), so all 8 requests issue back to back.
        constexpr int U = NNZ_MAX / BS;
        const int last = nz - 1;
        int ci[U];
        Val vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = min(tid + u * BS, last);
            ci[u] = ld_stream<true>(idx + p0 + q);
            vv[u] = ld_stream<true>(val + p0 + q);
        }
        if (slot) {
            int sq[U];
#pragma unroll
            for (int u = 0; u < U; ++u) sq[u] = ld_stream<true>(slot + p0 + min(tid + u * BS, last));
            Acc xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = static_cast<Acc>(gx(ci[u]));
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = tid + u * BS;
                if (q < nz) s_prod[sq[u]] = static_cast<Acc>(vv[u]) * xv[u];
            }
        } else {
            Acc xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = static_cast<Acc>(gx(ci[u]));
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = tid + u * BS;
                if (q < nz) s_prod[q] = static_cast<Acc>(vv[u]) * xv[u];
            }
        }
    }
    __syncthreads();
    Acc s = Acc(0);
    for (int q = a + lane; q < b; q += tpr) s += s_prod[q];
    const int wl = tpr < WAVE ? tpr : WAVE;
    for (int o = wl / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (tpr <= WAVE) {
        if (row < nrows && lane == 0) s_rows[row] = s;
    } else {
        if ((tid & (WAVE - 1)) == 0) red[tid / WAVE] = s;
        __syncthreads();
        if (lane == 0 && row < nrows) {
            const int wpr = tpr / WAVE;
            Acc t = red[row * wpr];
            for (int k = 1; k < wpr; ++k) t += red[row * wpr + k];
            s_rows[row] = t;
        }
    }
    __syncthreads();
}

// rowblock_sums (NTL false) or rowblock_sums_nt (NTL true).
template <bool NTL, typename Acc, typename Val, typename G, typename I>
__device__ __forceinline__ void rowblock_sums_t(const int* __restrict__ ptr,
                                                const I* __restrict__ idx,
                                                const Val* __restrict__ val,
                                                const G& gx, int r0, int r1,
                                                int p0, int p1, Acc* s_rows,
                                                Acc* s_prod, Acc* red,
                                                const unsigned short* __restrict__ slot) {
    if constexpr (NTL)
        rowblock_sums_nt<Acc>(ptr, idx, val, gx, r0, r1, p0, p1, s_rows, s_prod, red, slot);
    else
        rowblock_sums<Acc>(ptr, idx, val, gx, r0, r1, p0, p1, s_rows, s_prod, red, slot);
}

}  // namespace igaos::dev
