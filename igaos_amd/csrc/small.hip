// One-workgroup runner of the default mode for small LPs (gfx950).
//
// An LP with max(n, m) <= 3 * 1024 whose two matrices' products, x and y
// fit in LDS runs a whole check period of line-search attempts in one
// launch of one 1024-thread workgroup, with the attempt of the three-kernel
// path (pdhg_kernels.hip: k_primal, k_dual, k_decide) restated so that
// almost nothing but the matrices leaves the CU:
//   * thread t owns columns t, t + 1024, ... and rows t, t + 1024, ...; its
//     columns' x, average and A'y and its rows' y, average, A x commit and
//     A dx live in registers for the whole launch (the per-epoch constants
//     are read from L2 where they are used);
//   * xbar and y, the vectors the products gather, live in LDS, and so do
//     the staged products: each product streams its matrix (CSC for A'y,
//     CSR for A xbar) from L2 with coalesced loads into prod[], and every
//     owned row then sums its slice of prod[] in storage order;
//   * the three line-search dots are reduced once per attempt in a fixed
//     order and thread 0 applies the decision rule to the state, which
//     lives in LDS, so no global memory traffic is needed between
//     attempts.
// On exit the state and the vectors go back to the engine's buffers exactly
// as the three kernels leave them (current iterate in x[cur], y[cur]; the
// pending averaging flagged in the state), so the host's check code is the
// same for both paths.  Summation orders: every sparse row in storage
// order, dots per thread then wavefront trees then the 16 wavefronts in
// order; the results agree with the oracle to rounding like the
// three-kernel path's.

#include "dev_common.hpp"
#include "small.hpp"

namespace igaos::dev {

using pdhg::LsState;

namespace {

constexpr int ST = 1024;  // threads of the runner's workgroup
constexpr int SW = ST / WAVE;
constexpr int SMALL_PW = 65;  // line-search powers tabulated per launch

// Per-wavefront sums of three per-thread values into red[w], red[SW + w],
// red[2 SW + w]; after a barrier, fin3 adds the SW values of each in order.
__device__ __forceinline__ void part3(double a, double b, double c, double* red) {
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        red[w] = a;
        red[SW + w] = b;
        red[2 * SW + w] = c;
    }
}
__device__ __forceinline__ void fin3(const double* red, double& a, double& b, double& c) {
    a = red[0];
    b = red[SW];
    c = red[2 * SW];
    for (int k = 1; k < SW; ++k) {
        a += red[k];
        b += red[SW + k];
        c += red[2 * SW + k];
    }
}

// prod[p] = val[p] * x[idx[p]] for p < nnz, U index/value pairs in flight
// per thread (U = 8 in FP32, 4 in FP64 to stay within the registers of a
// 1024-thread workgroup); x is in LDS.
template <typename V>
__device__ __forceinline__ void stage_products(int nnz, const int* __restrict__ idx,
                                               const V* __restrict__ val, const V* x,
                                               V* prod) {
    constexpr int U = sizeof(V) == 4 ? 8 : 4;
    int p = threadIdx.x;
    for (; p + (U - 1) * ST < nnz; p += U * ST) {
        int c[U];
        V v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = idx[p + u * ST];
            v[u] = val[p + u * ST];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) prod[p + u * ST] = v[u] * x[c[u]];
    }
    for (; p < nnz; p += ST) prod[p] = val[p] * x[idx[p]];
}

template <typename V>
__device__ __forceinline__ V row_sum(const V* prod, int a, int b) {
    V s = V(0);
    for (int p = a; p < b; ++p) s += prod[p];
    return s;
}

template <typename V, int K>
__global__ __launch_bounds__(ST) void k_small_run(IterBufs<V> b, int cap) {
    extern __shared__ double smem_d[];
    double* red = smem_d;                                // 3 * SW
    double* pw = smem_d + 3 * SW;                        // 2 * SMALL_PW
    V* ys = reinterpret_cast<V*>(pw + 2 * SMALL_PW);     // m
    V* xbs = ys + b.m;                                   // n
    V* prod = xbs + b.n;                                 // nnz
    // the line-search state: one copy in LDS, decided by thread 0
    __shared__ LsState st;
    const int tid = threadIdx.x, n = b.n, m = b.m;
    if (tid == 0) st = *b.st;
    __syncthreads();
    const int cur0 = st.cur;
    // k^-0.3 and k^-0.6 of the iterations this launch can reach (a check
    // period is at most 64 accepted iterations), before the vectors are
    // loaded: pow's registers are then not live beside them
    const long long k0 = st.k;
    if (tid < SMALL_PW) {
        const double kk = static_cast<double>(k0 + tid);
        pw[tid] = pow(kk, -0.3);
        pw[SMALL_PW + tid] = pow(kk, -0.6);
    }
    // owned columns j = tid + k ST and rows i = tid + k ST: their evolving
    // state in registers (the shifted gradient, the bounds and the row
    // pointers are constant within a launch and are read from L2 at use)
    V xc[K], xa[K], at[K], xn[K];
    V yc[K], ya[K], ax[K], ad[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * ST;
        const bool on = j < n;
        xc[k] = on ? b.x[cur0][j] : V(0);
        xa[k] = on ? b.xavg[j] : V(0);
        at[k] = on ? b.ata[j] : V(0);
        xn[k] = xc[k];
        const int i = tid + k * ST;
        const bool ro = i < m;
        yc[k] = ro ? b.y[cur0][i] : V(0);
        ya[k] = ro ? b.yavg[i] : V(0);
        ax[k] = ro ? b.axcur[i] : V(0);
        ad[k] = ro ? b.adx[i] : V(0);
        if (ro) ys[i] = yc[k];
    }
    const int nnzT = b.cp[n], nnzA = b.ap[m];
    __syncthreads();
    for (int att = 0; att < cap && !st.frozen && st.k - k0 < SMALL_PW; ++att) {
        const V tau = static_cast<V>(st.eta / st.omega);
        const V sig = static_cast<V>(st.eta * st.omega);
        const V w = static_cast<V>(st.mix_w);
        const bool fresh = st.fresh != 0;
        const bool mix = fresh && st.mix_pending != 0;
        // A'y (a retry reuses the stored one) and the primal step (k_primal)
        if (fresh) {
            stage_products(nnzT, b.ci, b.acx, ys, prod);
            __syncthreads();
        }
        double d0 = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int j = tid + k * ST;
            if (j < n) {
                const V g0 = b.gref[j], l0 = b.lo[j], h0 = b.hi[j];
                if (fresh) at[k] = row_sum(prod, b.cp[j], b.cp[j + 1]);
                const V x0 = xc[k];
                if (mix) xa[k] = xa[k] + w * (x0 - xa[k]);
                const V g = g0 + at[k];
                const V v = fmin(fmax(x0 - tau * g, l0), h0);
                xn[k] = v;
                xbs[j] = V(2) * v - x0;
                const V d = v - x0;
                d0 += static_cast<double>(d) * static_cast<double>(d);
            }
        }
        __syncthreads();  // xbar complete; prod free
        // A xbar and the dual step (k_dual)
        stage_products(nnzA, b.ai, b.ax, xbs, prod);
        __syncthreads();
        double d1 = 0.0, d2 = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = tid + k * ST;
            if (i < m) {
                const V r0 = b.rlo[i], r1 = b.rhi[i];
                const V of = b.yoff ? b.yoff[i] : V(0);
                const V tm = row_sum(prod, b.ap[i], b.ap[i + 1]);
                const V y0 = yc[k];
                V axc = ax[k];
                if (mix) {
                    ya[k] = ya[k] + w * (y0 - ya[k]);
                    axc = axc + ad[k];
                    ax[k] = axc;
                }
                const V tt = y0 + sig * tm;
                V u = b.yoff ? tt + of : tt;
                u = u / sig;
                const V yv = tt - sig * fmin(fmax(u, r0), r1);
                // the attempt's y goes straight to LDS: only the A'y of a
                // fresh attempt reads ys, and a fresh attempt follows an
                // accepted one, whose y this then is
                ys[i] = yv;
                const V dy = yv - y0;
                const V a = V(0.5) * (tm - axc);
                ad[k] = a;
                d1 += static_cast<double>(dy) * static_cast<double>(dy);
                d2 += static_cast<double>(dy) * static_cast<double>(a);
            }
        }
        // the decision (k_decide): fixed-order sums, the rule on thread 0
        const int before = st.cur;
        part3(d0, d1, d2, red);
        __syncthreads();
        if (tid == 0) {
            double s0, s1, s2;
            fin3(red, s0, s1, s2);
            const int kq = static_cast<int>(st.k - k0);
            LsState t = st;
            pdhg::ls_decide_pw(t, s0, s1, s2, pw[kq], pw[SMALL_PW + kq]);
            st = t;
        }
        __syncthreads();  // the state is decided; red is free again
        if (st.cur != before) {  // accepted: the attempt becomes the iterate
#pragma unroll
            for (int k = 0; k < K; ++k) {
                xc[k] = xn[k];
                const int i = tid + k * ST;
                if (i < m) yc[k] = ys[i];  // written by this thread
            }
        }
    }
    // back to the engine's buffers, as the three kernels leave them
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * ST;
        if (j < n) {
            b.x[st.cur][j] = xc[k];
            b.x[st.cur ^ 1][j] = xn[k];
            b.xavg[j] = xa[k];
            b.ata[j] = at[k];
            b.xbar[j] = xbs[j];
        }
        const int i = tid + k * ST;
        if (i < m) {
            b.y[st.cur][i] = yc[k];
            b.y[st.cur ^ 1][i] = ys[i];
            b.yavg[i] = ya[k];
            b.axcur[i] = ax[k];
            b.adx[i] = ad[k];
        }
    }
    if (tid == 0) *b.st = st;
}

template <typename V>
size_t small_lds(int n, int m, long nnz) {
    return (3 * SW + 2 * SMALL_PW) * sizeof(double) +
           sizeof(V) * (static_cast<size_t>(n) + m + nnz);
}

template <typename V, int K>
hipError_t launch_k(const IterBufs<V>& b, int cap, size_t lds, hipStream_t s) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small_run<V, K>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(lds));
    if (e != hipSuccess) return e;
    k_small_run<V, K><<<1, ST, lds, s>>>(b, cap);
    return hipGetLastError();
}

template <typename V>
hipError_t launch_small(const IterBufs<V>& b, long nnz, int cap, hipStream_t s) {
    const int mx = b.n > b.m ? b.n : b.m;
    const size_t lds = small_lds<V>(b.n, b.m, nnz);
    if (mx <= ST) return launch_k<V, 1>(b, cap, lds, s);
    if (mx <= 2 * ST) return launch_k<V, 2>(b, cap, lds, s);
    return launch_k<V, 3>(b, cap, lds, s);
}

}  // namespace

bool small_fits(int n, int m, long nnz, bool f32) {
    int dev = 0, cap = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cap, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return false;
    const int mx = n > m ? n : m;
    const size_t lds = f32 ? small_lds<float>(n, m, nnz) : small_lds<double>(n, m, nnz);
    // the kernel's static LDS (the state) comes on top of the dynamic part
    static_assert(sizeof(LsState) <= 256, "state outgrew its LDS allowance");
    return n > 0 && m > 0 && mx <= SMALL_MAX_DIM && lds + 256 <= static_cast<size_t>(cap);
}

hipError_t launch_small_run(const IterBufs<double>& b, long nnz, int cap, hipStream_t s) {
    return launch_small(b, nnz, cap, s);
}
hipError_t launch_small_run(const IterBufs<float>& b, long nnz, int cap, hipStream_t s) {
    return launch_small(b, nnz, cap, s);
}

}  // namespace igaos::dev
