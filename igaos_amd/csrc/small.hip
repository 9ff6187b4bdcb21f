// One-workgroup runner of the default mode for small LPs (gfx950).
//
// An LP with max(n, m) <= 3 * 1024 whose x, y and bounds fit in LDS (and
// whose matrix is at most SMALL_MAX_BYTES per product) runs a whole check
// period of line-search attempts in one launch of one 1024-thread
// workgroup, with the attempt of the three-kernel path (pdhg_kernels.hip:
// k_primal, k_dual, k_decide) restated so that almost nothing but the
// matrices leaves the CU:
//   * columns and rows sit in slots in order of decreasing length, and
//     thread t owns the column slots t, t + 1024, ... (KC of them) and the
//     row slots likewise (KR); their x, average and A'y, and y, average,
//     A x commit and A dx, live in registers for the whole launch;
//   * xbar and y, the vectors the products gather, and the per-launch
//     constants (shifted gradient, bounds) live in LDS;
//   * each product streams a sliced-ELL copy of its matrix (A' for A'y, A
//     for A xbar; small.hpp) from L2: a slice's 64 slots hold their
//     entries in chunk-major order, so the wavefront that owns the slice
//     loads one contiguous run per chunk, and every thread sums its own
//     rows, in storage order, straight from the loads and LDS gathers (no
//     staged products, no barrier between a product and its step);
//   * the three line-search dots are reduced once per attempt in a fixed
//     order and thread 0 applies the decision rule to the state, which
//     lives in LDS, while every wavefront already sums the next attempt's
//     A'y (used if the attempt is accepted), so no global memory traffic
//     but the matrices is needed between attempts.
// On exit the state and the vectors go back to the engine's buffers as the
// three kernels leave them, with the current iterate always in x[0], y[0]
// (the state's cur set to 0; nothing reads the other slot before the next
// attempt writes it) and the pending averaging flagged in the state, so
// the host's check code is the same for both paths and a check's device
// work can be captured once with cur = 0.  Summation orders: every sparse
// row in storage order, dots per thread then wavefront butterflies then the
// 16 wavefronts in order; the results agree with the oracle to rounding
// like the three-kernel path's.

#include <algorithm>

#include "dev_common.hpp"
#include "small.hpp"

namespace igaos::dev {

using pdhg::LsState;

namespace {

constexpr int ST = 1024;  // threads of the runner's workgroup
constexpr int SW = ST / WAVE;
constexpr int SMALL_PW = 65;  // line-search powers tabulated per launch

// v from the lane that DPP control CTRL names, both halves moved
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(u), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// The wavefront's sum of v in lane 0 by a butterfly that pairs lanes 1,
// 2, 4, 8, 16 and 32 apart in that order (v + partner at each level):
// lanes 1 and 2 apart by DPP quad permutes; once groups of 4 (8) lanes hold
// equal values, the DPP mirror within 8 (16) lanes reads a lane of the
// partner group, as an xor by 4 (8) would; 16 apart by a swizzle, 32 apart
// by readlane.  Every lane of the wavefront must be active.
__device__ __forceinline__ double wave_sum_lane0(double v) {
    v += dpp_d<0xB1>(v);   // quad_perm 1,0,3,2
    v += dpp_d<0x4E>(v);   // quad_perm 2,3,0,1
    v += dpp_d<0x141>(v);  // row_half_mirror
    v += dpp_d<0x140>(v);  // row_mirror
    {
        const long long u = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_ds_swizzle(static_cast<int>(u), 0x401F);  // xor 16
        const int hi = __builtin_amdgcn_ds_swizzle(static_cast<int>(u >> 32), 0x401F);
        v += __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
    }
    const long long u = __double_as_longlong(v);
    const double v0 = __longlong_as_double(
        (static_cast<long long>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), 0)) << 32) |
        static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u), 0)));
    const double v32 = __longlong_as_double(
        (static_cast<long long>(__builtin_amdgcn_readlane(static_cast<int>(u >> 32), 32)) << 32) |
        static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(u), 32)));
    return v0 + v32;
}

// Per-wavefront sums of three per-thread values into red[w], red[SW + w],
// red[2 SW + w]; after a barrier, fin3 adds the SW values of each in order.
// Called by every thread of the workgroup.
__device__ __forceinline__ void part3(double a, double b, double c, double* red) {
    a = wave_sum_lane0(a);
    b = wave_sum_lane0(b);
    c = wave_sum_lane0(c);
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

// a lane's chunk: C values, C 16-bit indices
template <typename T, int C>
struct LaneVec {
    typedef T type __attribute__((ext_vector_type(C)));
};

// out[k] = sum over the entries of the row in owned slot tid + k ST of
// val * x[idx], in storage order, the product rounded before the add (a
// separate multiply and add, like the three-kernel path).  The row's
// entries come in chunks of small_chunk<V>() (one 32-byte value load and
// one 8- or 16-byte index load per chunk): cb[k] is the chunk row where
// the row's slice starts (the same for the whole wavefront), nch[k] the
// row's chunks.  A chunk's padding entries are 0 with index 0, and add +-0
// (x is finite), which leaves a sum that started at +0 unchanged bit for
// bit.  All K rows' chunks are loaded together; a lane past its row's last
// chunk re-loads it (or chunk 0, which always exists) and adds it times 0,
// so neither loads nor gathers branch.
template <typename V, int K>
__device__ __forceinline__ void ell_sums(const SmallEll<V>& E, const int* cb, const int* nch,
                                         const V* x, V* out) {
#pragma clang fp contract(off)
    constexpr int C = small_chunk<V>();
    using VC = typename LaneVec<V, C>::type;
    using IC = typename LaneVec<unsigned short, C>::type;
    constexpr int UC = K == 1 ? 2 : 1;  // chunks of each row in flight
    const int lane = threadIdx.x & (WAVE - 1);
    const VC* val = reinterpret_cast<const VC*>(E.val);
    const IC* idx = reinterpret_cast<const IC*>(E.idx);
    int most = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        most = max(most, nch[k]);
        out[k] = V(0);
    }
    for (int c = 0; c < most; c += UC) {
        VC v[K][UC];
        IC ix[K][UC];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int u = 0; u < UC; ++u) {
                const int q = nch[k] > 0 ? (cb[k] + min(c + u, nch[k] - 1)) * WAVE + lane : 0;
                v[k][u] = val[q];
                ix[k][u] = idx[q];
            }
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int u = 0; u < UC; ++u) {
                const V on = static_cast<V>(c + u < nch[k]);
#pragma unroll
                for (int g = 0; g < C; ++g) out[k] += (v[k][u][g] * on) * x[ix[k][u][g]];
            }
    }
}

// The ELL position of owned slot s = tid + k ST: the chunk row where its
// slice starts (wave-uniform: a scalar load) and the chunk count of its row
// (0 past the last slot).
template <typename V, int K>
__device__ __forceinline__ void ell_rows(const SmallEll<V>& E, const int* __restrict__ ptr,
                                         int rows, int* cb, int* nch) {
    const int slices = (rows + WAVE - 1) / WAVE;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int s = threadIdx.x + k * ST;
        const int q = __builtin_amdgcn_readfirstlane(s / WAVE);
        cb[k] = q < slices ? E.off[q] : 0;
        const int r = s < rows ? E.perm[s] : 0;
        constexpr int C = small_chunk<V>();
        nch[k] = s < rows ? (ptr[r + 1] - ptr[r] + C - 1) / C : 0;
    }
}

// KC, KR: columns and rows a thread owns (ceil(n / 1024), ceil(m / 1024))
template <typename V, int KC, int KR>
__global__ __launch_bounds__(ST) void k_small_run(IterBufs<V> b, SmallEll<V> eT, SmallEll<V> eA,
                                                  int cap) {
    extern __shared__ double smem_d[];
    double* red = smem_d;                                // 3 * SW
    double* pw = smem_d + 3 * SW;                        // 2 * SMALL_PW
    V* ys = reinterpret_cast<V*>(pw + 2 * SMALL_PW);     // m
    V* xbs = ys + b.m;                                   // n
    // the per-launch constants: shifted gradient and bounds of the columns,
    // bounds (and the correction offset) of the rows
    V* cg = xbs + b.n;                                   // n
    V* cl = cg + b.n;                                    // n
    V* ch = cl + b.n;                                    // n
    V* rl = ch + b.n;                                    // m
    V* rh = rl + b.m;                                    // m
    V* ro = rh + b.m;                                    // m, with yoff
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
    // owned column slots s = tid + k ST (column jc = eT.perm[s]) and row
    // slots (row ir = eA.perm[s]): their evolving state in registers
    int jc[KC], ir[KR];
    V xc[KC], xa[KC], at[KC], xn[KC];
    V yc[KR], ya[KR], ax[KR], ad[KR];
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const int s = tid + k * ST;
        const bool on = s < n;
        const int j = on ? eT.perm[s] : 0;
        jc[k] = j;
        xc[k] = on ? b.x[cur0][j] : V(0);
        xa[k] = on ? b.xavg[j] : V(0);
        at[k] = on ? b.ata[j] : V(0);
        xn[k] = xc[k];
    }
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        const int s = tid + k * ST;
        const bool on = s < m;
        const int i = on ? eA.perm[s] : 0;
        ir[k] = i;
        yc[k] = on ? b.y[cur0][i] : V(0);
        ya[k] = on ? b.yavg[i] : V(0);
        ax[k] = on ? b.axcur[i] : V(0);
        ad[k] = on ? b.adx[i] : V(0);
        if (on) ys[i] = yc[k];
    }
    // the constants in slot order
    for (int s = tid; s < n; s += ST) {
        const int j = eT.perm[s];
        cg[s] = b.gref[j];
        cl[s] = b.lo[j];
        ch[s] = b.hi[j];
    }
    for (int s = tid; s < m; s += ST) {
        const int i = eA.perm[s];
        rl[s] = b.rlo[i];
        rh[s] = b.rhi[i];
        if (b.yoff) ro[s] = b.yoff[i];
    }
    int cbT[KC], nT[KC], cbA[KR], nA[KR];
    ell_rows<V, KC>(eT, b.cp, n, cbT, nT);
    ell_rows<V, KR>(eA, b.ap, m, cbA, nA);
    __syncthreads();
    // A'y of the last attempt's y, computed while thread 0 decides: the
    // next attempt's A'y if the attempt is accepted (the same sums of the
    // same ys, so the same values)
    V atn[KC];
    bool spec = false;
    for (int att = 0; att < cap && !st.frozen && st.k - k0 < SMALL_PW; ++att) {
        const V tau = static_cast<V>(st.eta / st.omega);
        const V sig = static_cast<V>(st.eta * st.omega);
        const V w = static_cast<V>(st.mix_w);
        const bool fresh = st.fresh != 0;
        const bool mix = fresh && st.mix_pending != 0;
        // A'y (a retry reuses the stored one) and the primal step (k_primal)
        if (fresh) {
            if (spec) {
#pragma unroll
                for (int k = 0; k < KC; ++k) at[k] = atn[k];
            } else {
                ell_sums<V, KC>(eT, cbT, nT, ys, at);
            }
        }
        // the step's data for every owned column in one batch of reads (a
        // slot past n reads slot n - 1 and is not stepped)
        V g0[KC], l0[KC], h0[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int s = min(tid + k * ST, n - 1);
            g0[k] = cg[s];
            l0[k] = cl[s];
            h0[k] = ch[s];
        }
        double d0 = 0.0;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            if (tid + k * ST < n) {
                const V x0 = xc[k];
                if (mix) xa[k] = xa[k] + w * (x0 - xa[k]);
                const V g = g0[k] + at[k];
                const V v = fmin(fmax(x0 - tau * g, l0[k]), h0[k]);
                xn[k] = v;
                xbs[jc[k]] = V(2) * v - x0;
                const V d = v - x0;
                d0 += static_cast<double>(d) * static_cast<double>(d);
            }
        }
        __syncthreads();  // xbar complete
        // A xbar and the dual step (k_dual)
        V tms[KR];
        ell_sums<V, KR>(eA, cbA, nA, xbs, tms);
        V r0[KR], r1[KR], of[KR];
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int s = min(tid + k * ST, m - 1);
            r0[k] = rl[s];
            r1[k] = rh[s];
            of[k] = b.yoff ? ro[s] : V(0);
        }
        double d1 = 0.0, d2 = 0.0;
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            if (tid + k * ST < m) {
                const V tm = tms[k];
                const V y0 = yc[k];
                V axc = ax[k];
                if (mix) {
                    ya[k] = ya[k] + w * (y0 - ya[k]);
                    axc = axc + ad[k];
                    ax[k] = axc;
                }
                const V tt = y0 + sig * tm;
                V u = b.yoff ? tt + of[k] : tt;
                u = u / sig;
                const V yv = tt - sig * fmin(fmax(u, r0[k]), r1[k]);
                // the attempt's y goes straight to LDS: only the A'y of a
                // fresh attempt reads ys, and a fresh attempt follows an
                // accepted one, whose y this then is
                ys[ir[k]] = yv;
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
        ell_sums<V, KC>(eT, cbT, nT, ys, atn);
        spec = true;
        __syncthreads();  // the state is decided; red is free again
        if (st.cur != before) {  // accepted: the attempt becomes the iterate
#pragma unroll
            for (int k = 0; k < KC; ++k) xc[k] = xn[k];
#pragma unroll
            for (int k = 0; k < KR; ++k) {
                if (tid + k * ST < m) yc[k] = ys[ir[k]];  // written by this thread
            }
        }
    }
    // back to the engine's buffers, the current iterate in slot 0
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        if (tid + k * ST < n) {
            const int j = jc[k];
            b.x[0][j] = xc[k];
            b.x[1][j] = xn[k];
            b.xavg[j] = xa[k];
            b.ata[j] = at[k];
        }
    }
    for (int j = tid; j < n; j += ST) b.xbar[j] = xbs[j];
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        if (tid + k * ST < m) {
            const int i = ir[k];
            b.y[0][i] = yc[k];
            b.y[1][i] = ys[i];
            b.yavg[i] = ya[k];
            b.axcur[i] = ax[k];
            b.adx[i] = ad[k];
        }
    }
    if (tid == 0) {
        LsState t = st;
        t.cur = 0;
        *b.st = t;
    }
}

// x, y and the constants (4 n + 3 m values, and m offsets with yoff)
template <typename V>
size_t small_lds(int n, int m, bool yoff) {
    return (3 * SW + 2 * SMALL_PW) * sizeof(double) +
           sizeof(V) * (4 * static_cast<size_t>(n) + (yoff ? 4 : 3) * static_cast<size_t>(m));
}

template <typename V>
__global__ void k_ell_fill(const int* __restrict__ src, long total, const int* __restrict__ idx,
                           const V* __restrict__ val, unsigned short* __restrict__ eidx,
                           V* __restrict__ eval) {
    const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
    if (p >= total) return;
    const int q = src[p];
    eidx[p] = q >= 0 ? static_cast<unsigned short>(idx[q]) : 0;
    eval[p] = q >= 0 ? val[q] : V(0);
}

template <typename V>
hipError_t ell_fill(const int* src, long total, const int* idx, const V* val,
                    unsigned short* eidx, V* eval, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    const int tpb = 256;
    k_ell_fill<V><<<static_cast<unsigned>((total + tpb - 1) / tpb), tpb, 0, s>>>(
        src, total, idx, val, eidx, eval);
    return hipGetLastError();
}

template <typename V, int KC, int KR>
hipError_t launch_k(const IterBufs<V>& b, const SmallEll<V>& eT, const SmallEll<V>& eA, int cap,
                    size_t lds, hipStream_t s) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_small_run<V, KC, KR>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(lds));
    if (e != hipSuccess) return e;
    k_small_run<V, KC, KR><<<1, ST, lds, s>>>(b, eT, eA, cap);
    return hipGetLastError();
}

template <typename V, int KC>
hipError_t launch_kr(const IterBufs<V>& b, const SmallEll<V>& eT, const SmallEll<V>& eA, int cap,
                     size_t lds, hipStream_t s) {
    const int kr = (b.m + ST - 1) / ST;
    if (kr <= 1) return launch_k<V, KC, 1>(b, eT, eA, cap, lds, s);
    if (kr == 2) return launch_k<V, KC, 2>(b, eT, eA, cap, lds, s);
    return launch_k<V, KC, 3>(b, eT, eA, cap, lds, s);
}

template <typename V>
hipError_t launch_small(const IterBufs<V>& b, const SmallEll<V>& eT, const SmallEll<V>& eA,
                        int cap, hipStream_t s) {
    const size_t lds = small_lds<V>(b.n, b.m, b.yoff != nullptr);
    const int kc = (b.n + ST - 1) / ST;
    if (kc <= 1) return launch_kr<V, 1>(b, eT, eA, cap, lds, s);
    if (kc == 2) return launch_kr<V, 2>(b, eT, eA, cap, lds, s);
    return launch_kr<V, 3>(b, eT, eA, cap, lds, s);
}

}  // namespace

bool small_fits(int n, int m, long nnz, bool f32) {
    int dev = 0, cap = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cap, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return false;
    const int mx = n > m ? n : m;
    // FP32 engines carry the correction offsets, FP64 engines none
    const size_t lds = f32 ? small_lds<float>(n, m, true) : small_lds<double>(n, m, false);
    // the kernel's static LDS (the state) comes on top of the dynamic part
    static_assert(sizeof(LsState) <= 256, "state outgrew its LDS allowance");
    const long bytes = nnz * static_cast<long>((f32 ? sizeof(float) : sizeof(double)) + 2);
    return n > 0 && m > 0 && mx <= SMALL_MAX_DIM && bytes <= SMALL_MAX_BYTES &&
           lds + 256 <= static_cast<size_t>(cap);
}

long small_ell_map(int rows, const int* ptr, int chunk, std::vector<int>& perm,
                   std::vector<int>& off, std::vector<int>& src) {
    // slots in order of decreasing row length (ties by row), so that the 64
    // rows of a slice are alike in length
    perm.resize(rows);
    for (int r = 0; r < rows; ++r) perm[r] = r;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
        return ptr[a + 1] - ptr[a] > ptr[b + 1] - ptr[b];
    });
    const int slices = (rows + WAVE - 1) / WAVE;
    off.assign(slices + 1, 0);
    for (int q = 0; q < slices; ++q) {
        int width = 0;
        for (int s = q * WAVE; s < std::min(rows, (q + 1) * WAVE); ++s) {
            const int r = perm[s];
            width = std::max(width, (ptr[r + 1] - ptr[r] + chunk - 1) / chunk);
        }
        off[q + 1] = off[q] + width;
    }
    // at least one chunk row, so that chunk 0 exists
    const long total = static_cast<long>(std::max(off[slices], 1)) * WAVE * chunk;
    src.assign(static_cast<size_t>(total), -1);
    for (int s = 0; s < rows; ++s) {
        const int r = perm[s], q = s / WAVE, lane = s % WAVE;
        for (int e = 0; e < ptr[r + 1] - ptr[r]; ++e) {
            const long c = static_cast<long>(off[q] + e / chunk) * WAVE + lane;
            src[static_cast<size_t>(c * chunk + e % chunk)] = ptr[r] + e;
        }
    }
    return total;
}

hipError_t small_ell_fill(const int* src, long total, const int* idx, const double* val,
                          unsigned short* eidx, double* eval, hipStream_t s) {
    return ell_fill(src, total, idx, val, eidx, eval, s);
}
hipError_t small_ell_fill(const int* src, long total, const int* idx, const float* val,
                          unsigned short* eidx, float* eval, hipStream_t s) {
    return ell_fill(src, total, idx, val, eidx, eval, s);
}

hipError_t launch_small_run(const IterBufs<double>& b, const SmallEll<double>& eT,
                            const SmallEll<double>& eA, int cap, hipStream_t s) {
    return launch_small(b, eT, eA, cap, s);
}
hipError_t launch_small_run(const IterBufs<float>& b, const SmallEll<float>& eT,
                            const SmallEll<float>& eA, int cap, hipStream_t s) {
    return launch_small(b, eT, eA, cap, s);
}

}  // namespace igaos::dev
