// Exact FP64 mode of the single-LP engine: the reference iteration
// (src/pdhg/pdhg.cu:32-141, 605-663, 694-733) with the arithmetic fixed
// operation by operation, so that a run is bit-reproducible against a
// sequential FP64 host replay of the same steps.
//
//  * every sparse product sums each row in storage order, one thread per
//    row, product then add (no fused multiply-add anywhere in this file:
//    it is compiled with -ffp-contract=off and the pragma below);
//  * the line-search dots use the reference's two-stage layout
//    (kern_dots + kern_dotfin, pdhg.cu:63-125): gD = min(ceil(max(n,m)/256),
//    4096) workgroups of 256 threads accumulate grid-stride, each reduces
//    by a halving tree in LDS, and one workgroup of 256 threads sums the gD
//    partials strided and reduces them by the same tree;
//  * the step-size exponents k^-0.3 and k^-0.6 come from a table the host
//    fills with its own pow, as the reference's host controller computes
//    them (pdhg.cu:780-781);
//  * norms, c'x and the Lagrangian bound are sequential sums in index
//    order, as the reference's host loops compute them (pdhg.cu:173-254,
//    617-626, 647-663, 1034-1060).
// Maxima and elementwise steps need no ordering.  The fast kernels of
// pdhg_kernels.hip remain the engine's default; this mode trades their
// bandwidth for reproducibility and exists for parity checks.

#include "exact.hpp"

#pragma clang fp contract(off)

namespace igaos::dev {

using pdhg::LsState;
using pdhg::Metrics;

constexpr int XB = 256;  // the reference's dot-product block size

__device__ __forceinline__ double seq_row(const int* __restrict__ ptr,
                                          const int* __restrict__ idx,
                                          const double* __restrict__ val,
                                          const double* __restrict__ x, int r) {
    double s = 0.0;
    const int p1 = ptr[r + 1];
    for (int p = ptr[r]; p < p1; ++p) s = s + val[p] * x[idx[p]];
    return s;
}

__global__ __launch_bounds__(XB) void k_x_spmv(int rows, const int* ptr, const int* idx,
                                               const double* val, const double* x,
                                               double* y) {
    const long r = static_cast<long>(blockIdx.x) * XB + threadIdx.x;
    if (r < rows) y[r] = seq_row(ptr, idx, val, x, static_cast<int>(r));
}

void spmv_seq(int rows, const int* ptr, const int* idx, const double* val,
              const double* x, double* y, hipStream_t s) {
    if (rows > 0)
        k_x_spmv<<<(rows + XB - 1) / XB, XB, 0, s>>>(rows, ptr, idx, val, x, y);
}

// ----------------------------------------------------------- attempt --

// A'y of a fresh attempt (a retry keeps the stored one, pdhg.cu:748).
__global__ __launch_bounds__(XB) void k_x_ata(IterBufs<double> b) {
    const LsState* st = b.st;
    if (st->frozen || !st->fresh) return;
    const long j = static_cast<long>(blockIdx.x) * XB + threadIdx.x;
    if (j < b.n) b.ata[j] = seq_row(b.cp, b.ci, b.acx, b.y[st->cur], static_cast<int>(j));
}

// kern_primal (pdhg.cu:32-45), with the pending averaging of x
// (pdhg.cu:127-141) applied first.
__global__ __launch_bounds__(XB) void k_x_primal(IterBufs<double> b, ExactBufs e) {
    const LsState st = *b.st;
    if (st.frozen) return;
    const long j = static_cast<long>(blockIdx.x) * XB + threadIdx.x;
    if (j >= b.n) return;
    const int cur = st.cur;
    const double tau = st.eta / st.omega;
    const double xo = b.x[cur][j];
    if (st.fresh && st.mix_pending) {
        const double xa = b.xavg[j];
        b.xavg[j] = xa + st.mix_w * (xo - xa);
    }
    const double g = b.gref[j] + b.ata[j];
    const double v = fmin(fmax(xo - tau * g, b.lo[j]), b.hi[j]);
    b.x[cur ^ 1][j] = v;
    b.xbar[j] = 2.0 * v - xo;
    e.dx[j] = v - xo;
}

__global__ __launch_bounds__(XB) void k_x_axbar(IterBufs<double> b, ExactBufs e) {
    if (b.st->frozen) return;
    const long i = static_cast<long>(blockIdx.x) * XB + threadIdx.x;
    if (i < b.m) e.tm[i] = seq_row(b.ap, b.ai, b.ax, b.xbar, static_cast<int>(i));
}

// kern_dual (pdhg.cu:47-61), with the pending averaging of y and the A x
// commit (pdhg.cu:795-799) applied first.
__global__ __launch_bounds__(XB) void k_x_dual(IterBufs<double> b, ExactBufs e) {
    const LsState st = *b.st;
    if (st.frozen) return;
    const long i = static_cast<long>(blockIdx.x) * XB + threadIdx.x;
    if (i >= b.m) return;
    const int cur = st.cur;
    const double sig = st.eta * st.omega;
    const double yo = b.y[cur][i];
    double axc = b.axcur[i];
    if (st.fresh && st.mix_pending) {
        const double ya = b.yavg[i];
        b.yavg[i] = ya + st.mix_w * (yo - ya);
        axc = axc + b.adx[i];
        b.axcur[i] = axc;
    }
    const double tm = e.tm[i];
    const double t = yo + sig * tm;
    const double yv = t - sig * fmin(fmax(t / sig, b.rlo[i]), b.rhi[i]);
    b.y[cur ^ 1][i] = yv;
    e.dy[i] = yv - yo;
    b.adx[i] = 0.5 * (tm - axc);
}

// First stage of the dots (kern_dots, pdhg.cu:63-100).
__device__ __forceinline__ void dots_stage1(long n, long m, int gD, const double* dx,
                                            const double* dy, const double* adx,
                                            double* part) {
    __shared__ double sh[3][XB];
    const int tid = threadIdx.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (long i = static_cast<long>(blockIdx.x) * XB + tid; i < n || i < m;
         i += static_cast<long>(gD) * XB) {
        if (i < n) a0 = a0 + dx[i] * dx[i];
        if (i < m) {
            a1 = a1 + dy[i] * dy[i];
            a2 = a2 + dy[i] * adx[i];
        }
    }
    sh[0][tid] = a0;
    sh[1][tid] = a1;
    sh[2][tid] = a2;
    __syncthreads();
    for (int s = XB / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int q = 0; q < 3; ++q) sh[q][tid] = sh[q][tid] + sh[q][tid + s];
        __syncthreads();
    }
    if (tid == 0)
        for (int q = 0; q < 3; ++q) part[q * gD + blockIdx.x] = sh[q][0];
}

// Second stage (kern_dotfin, pdhg.cu:102-125): all threads of the single
// workgroup return the three sums.
__device__ __forceinline__ void dots_stage2(int gD, const double* part, double out[3]) {
    __shared__ double sh[3][XB];
    const int tid = threadIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int i = tid; i < gD; i += XB) {
        s0 = s0 + part[i];
        s1 = s1 + part[gD + i];
        s2 = s2 + part[2 * gD + i];
    }
    sh[0][tid] = s0;
    sh[1][tid] = s1;
    sh[2][tid] = s2;
    __syncthreads();
    for (int s = XB / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int q = 0; q < 3; ++q) sh[q][tid] = sh[q][tid] + sh[q][tid + s];
        __syncthreads();
    }
    for (int q = 0; q < 3; ++q) out[q] = sh[q][0];
}

int dots_blocks(long n, long m) {
    const long mx = n > m ? n : m;
    long g = (mx + XB - 1) / XB;
    if (g > DOTS_MAX_BLOCKS) g = DOTS_MAX_BLOCKS;
    if (g < 1) g = 1;
    return static_cast<int>(g);
}

__global__ __launch_bounds__(XB) void k_x_dots(IterBufs<double> b, ExactBufs e, int gD) {
    if (b.st->frozen) return;
    dots_stage1(b.n, b.m, gD, e.dx, e.dy, b.adx, e.part);
}

// Line-search decision on the device state (pdhg.cu:751-794), with the
// exponent table for k^-0.3, k^-0.6.
__global__ __launch_bounds__(XB) void k_x_decide(IterBufs<double> b, ExactBufs e, int gD) {
    if (b.st->frozen) return;
    double d[3];
    dots_stage2(gD, e.part, d);
    if (threadIdx.x == 0) {
        LsState s = *b.st;
        const long long k = s.k;
        double p3, p6;
        if (k >= 1 && k <= e.pow_cap) {
            p3 = e.powtab[2 * (k - 1)];
            p6 = e.powtab[2 * (k - 1) + 1];
        } else {  // beyond the table (the host keeps runs inside it)
            p3 = pow(static_cast<double>(k), -0.3);
            p6 = pow(static_cast<double>(k), -0.6);
        }
        pdhg::ls_decide_pw(s, d[0], d[1], d[2], p3, p6);
        *b.st = s;
    }
}

void launch_attempt_exact(const IterBufs<double>& b, const ExactBufs& e, hipStream_t s) {
    const int gn = (b.n + XB - 1) / XB, gm = (b.m + XB - 1) / XB;
    const int gD = dots_blocks(b.n, b.m);
    if (gn > 0) k_x_ata<<<gn, XB, 0, s>>>(b);
    if (gn > 0) k_x_primal<<<gn, XB, 0, s>>>(b, e);
    if (gm > 0) k_x_axbar<<<gm, XB, 0, s>>>(b, e);
    if (gm > 0) k_x_dual<<<gm, XB, 0, s>>>(b, e);
    k_x_dots<<<gD, XB, 0, s>>>(b, e, gD);
    k_x_decide<<<1, XB, 0, s>>>(b, e, gD);
}

// ------------------------------------------------- one-workgroup runner --
// Small LPs: one workgroup of RT threads runs line-search attempts until
// the state freezes (or `cap` attempts), with exactly the arithmetic of the
// six attempt kernels above: A'y and A xbar summed per row in storage
// order, the same steps, and the dots in the reference's two-stage layout.
// Element i of that layout is thread i % 256 of virtual workgroup i / 256;
// a round of the runner covers virtual workgroups 4q .. 4q + 3 (elements
// 1024 q + r, r = threadIdx.x) and each of its four 256-thread quarters
// reduces one of them by the same halving tree (the levels below 64 with
// shuffles inside the first wavefront: the same additions in the same
// order).  Element i is stepped and then read back by the same thread, so
// only the gathers of y and xbar cross threads, after a barrier.  One
// launch replaces up to a whole check period of six-kernel attempts.
constexpr int RT = 1024;

// Halving trees of kern_dots / kern_dotfin over the three quantities'
// 256 values of one quarter (tree[q][base .. base + 255]): every thread of
// the workgroup calls it; lane 0 of the quarter's first wavefront returns
// the three sums.
__device__ __forceinline__ void quarter_trees(double (*tree)[RT], int base, double out[3]) {
    const int t = threadIdx.x & (XB - 1);
    __syncthreads();
    if (t < 128)
        for (int q = 0; q < 3; ++q) tree[q][base + t] = tree[q][base + t] + tree[q][base + t + 128];
    __syncthreads();
    if (t < 64)
        for (int q = 0; q < 3; ++q) tree[q][base + t] = tree[q][base + t] + tree[q][base + t + 64];
    __syncthreads();
    if (t < 64)
        for (int q = 0; q < 3; ++q) {
            double a = tree[q][base + t];
#pragma unroll
            for (int s = 32; s > 0; s >>= 1) a = a + __shfl_down(a, s, 64);
            out[q] = a;
        }
}

// Products val[p] * x[idx[p]] of all nnz nonzeros into prod[p]: every
// thread issues its loads back to back (no dependence between them).
__device__ __forceinline__ void run_products(long nnz, const int* __restrict__ idx,
                                             const double* __restrict__ val,
                                             const double* __restrict__ x, double* prod) {
#pragma unroll 8
    for (long p = threadIdx.x; p < nnz; p += RT) prod[p] = val[p] * x[idx[p]];
}
// Row r of a product staged by run_products, summed in storage order.
__device__ __forceinline__ double run_row(const int* __restrict__ ptr, const double* prod,
                                          int r) {
    double s = 0.0;
    const int p1 = ptr[r + 1];
    for (int p = ptr[r]; p < p1; ++p) s = s + prod[p];
    return s;
}

__global__ __launch_bounds__(RT) void k_x_run(IterBufs<double> b, ExactBufs e, int gD,
                                              int cap) {
    __shared__ LsState sh;
    // the staged products of one SpMV; the dots' trees reuse the space
    __shared__ double prod[RUN_MAX_NNZ];
    __shared__ double part[3][RUN_MAX_GD];
    double (*tree)[RT] = reinterpret_cast<double (*)[RT]>(prod);
    const int tid = threadIdx.x;
    const int n = b.n, m = b.m;
    const int quarter = tid / XB, t = tid & (XB - 1);
    if (tid == 0) sh = *b.st;
    __syncthreads();
    for (int att = 0; att < cap && !sh.frozen; ++att) {
        const int cur = sh.cur;
        const double tau = sh.eta / sh.omega;
        const double sig = sh.eta * sh.omega;
        const double w = sh.mix_w;
        const bool fresh = sh.fresh != 0;
        const bool mix = fresh && sh.mix_pending != 0;
        const double* __restrict__ x = b.x[cur];
        const double* __restrict__ y = b.y[cur];
        // k_x_ata + k_x_primal
        if (fresh) run_products(b.cp[n], b.ci, b.acx, y, prod);
        __syncthreads();
        for (int j = tid; j < n; j += RT) {
            double a;
            if (fresh) {
                a = run_row(b.cp, prod, j);
                b.ata[j] = a;
            } else {
                a = b.ata[j];
            }
            const double xo = x[j];
            if (mix) {
                const double xa = b.xavg[j];
                b.xavg[j] = xa + w * (xo - xa);
            }
            const double g = b.gref[j] + a;
            const double v = fmin(fmax(xo - tau * g, b.lo[j]), b.hi[j]);
            b.x[cur ^ 1][j] = v;
            b.xbar[j] = 2.0 * v - xo;
            e.dx[j] = v - xo;
        }
        __syncthreads();  // xbar complete, prod free
        // k_x_axbar + k_x_dual
        run_products(b.ap[m], b.ai, b.ax, b.xbar, prod);
        __syncthreads();
        for (int i = tid; i < m; i += RT) {
            const double tm = run_row(b.ap, prod, i);
            e.tm[i] = tm;
            const double yo = y[i];
            double axc = b.axcur[i];
            if (mix) {
                const double ya = b.yavg[i];
                b.yavg[i] = ya + w * (yo - ya);
                axc = axc + b.adx[i];
                b.axcur[i] = axc;
            }
            const double tt = yo + sig * tm;
            const double yv = tt - sig * fmin(fmax(tt / sig, b.rlo[i]), b.rhi[i]);
            b.y[cur ^ 1][i] = yv;
            e.dy[i] = yv - yo;
            b.adx[i] = 0.5 * (tm - axc);
        }
        __syncthreads();  // prod free for the trees
        // k_x_dots: virtual workgroups 4q .. 4q + 3 per round
        for (int q = 0; 4 * q < gD; ++q) {
            const long i = static_cast<long>(q) * RT + tid;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            if (i < n) a0 = a0 + e.dx[i] * e.dx[i];
            if (i < m) {
                a1 = a1 + e.dy[i] * e.dy[i];
                a2 = a2 + e.dy[i] * b.adx[i];
            }
            tree[0][tid] = a0;
            tree[1][tid] = a1;
            tree[2][tid] = a2;
            const int vb = 4 * q + quarter;
            double r[3] = {0.0, 0.0, 0.0};
            quarter_trees(tree, quarter * XB, r);
            if (t == 0 && vb < gD) {
                part[0][vb] = r[0];
                part[1][vb] = r[1];
                part[2][vb] = r[2];
            }
            __syncthreads();  // the tree slices are reused by the next round
        }
        // k_x_decide: kern_dotfin on the first quarter, then the rule
        {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            if (quarter == 0)
                for (int k = t; k < gD; k += XB) {
                    s0 = s0 + part[0][k];
                    s1 = s1 + part[1][k];
                    s2 = s2 + part[2][k];
                }
            tree[0][tid] = s0;
            tree[1][tid] = s1;
            tree[2][tid] = s2;
            double d[3] = {0.0, 0.0, 0.0};
            quarter_trees(tree, quarter * XB, d);
            const double d0 = d[0], d1 = d[1], d2 = d[2];
            if (tid == 0) {
                LsState st = sh;
                const long long k = st.k;
                double p3, p6;
                if (e.powtab && k >= 1 && k <= e.pow_cap) {
                    p3 = e.powtab[2 * (k - 1)];
                    p6 = e.powtab[2 * (k - 1) + 1];
                } else {
                    p3 = pow(static_cast<double>(k), -0.3);
                    p6 = pow(static_cast<double>(k), -0.6);
                }
                pdhg::ls_decide_pw(st, d0, d1, d2, p3, p6);
                sh = st;
            }
            __syncthreads();
        }
    }
    if (tid == 0) *b.st = sh;
}

bool run_exact_fits(long n, long m, long nnz) {
    const long mx = n > m ? n : m;
    return n > 0 && m >= 0 && nnz <= RUN_MAX_NNZ && (mx + XB - 1) / XB <= RUN_MAX_GD;
}

void launch_run_exact(const IterBufs<double>& b, const ExactBufs& e, int cap, hipStream_t s) {
    k_x_run<<<1, RT, 0, s>>>(b, e, dots_blocks(b.n, b.m), cap);
}

__global__ __launch_bounds__(XB) void k_x_flush(int n, int m, const double* x, double* xavg,
                                                const double* y, double* yavg,
                                                const double* adx, double* axcur, double w) {
    const long i = static_cast<long>(blockIdx.x) * XB + threadIdx.x;
    if (i < n) xavg[i] = xavg[i] + w * (x[i] - xavg[i]);
    if (i < m) {
        yavg[i] = yavg[i] + w * (y[i] - yavg[i]);
        axcur[i] = axcur[i] + adx[i];
    }
}
void launch_flush_exact(const IterBufs<double>& b, int cur, double w, hipStream_t s) {
    const long mx = b.n > b.m ? b.n : b.m;
    if (mx > 0)
        k_x_flush<<<(mx + XB - 1) / XB, XB, 0, s>>>(b.n, b.m, b.x[cur], b.xavg, b.y[cur],
                                                    b.yavg, b.adx, b.axcur, w);
}

// ------------------------------------------------------------- dots --

__global__ __launch_bounds__(XB) void k_x_dots_free(int n, int m, int gD, const double* dx,
                                                    const double* dy, const double* adx,
                                                    double* part) {
    dots_stage1(n, m, gD, dx, dy, adx, part);
}
__global__ __launch_bounds__(XB) void k_x_dotfin(int gD, const double* part, double* out) {
    double d[3];
    dots_stage2(gD, part, d);
    if (threadIdx.x == 0) {
        out[0] = d[0];
        out[1] = d[1];
        out[2] = d[2];
    }
}
void ls_dots(int n, int m, const double* dx, const double* dy, const double* adx,
             double* part, double* out, hipStream_t s) {
    const int gD = dots_blocks(n, m);
    k_x_dots_free<<<gD, XB, 0, s>>>(n, m, gD, dx, dy, adx, part);
    k_x_dotfin<<<1, XB, 0, s>>>(gD, part, out);
}

// ----------------------------------------------------- sequential sums --

// out = sum_i (a_i - b_i)^2 in index order (b may be null), one thread.
__global__ void k_x_sumsq(const double* a, const double* b, long n, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (long i = 0; i < n; ++i) {
        const double d = b ? a[i] - b[i] : a[i];
        s = s + d * d;
    }
    *out = s;
}
void sumsq_seq(const double* a, const double* b, long n, double* out, hipStream_t s) {
    k_x_sumsq<<<1, 64, 0, s>>>(a, b, n, out);
}

// out = sum_i q_i^2, q = srmax if finite, else srmin if finite, else 0
// (pdhg.cu:652-657), in index order.
__global__ void k_x_sumsq_q(const double* srmin, const double* srmax, long m, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (long i = 0; i < m; ++i) {
        const double q = __builtin_isfinite(srmax[i])
                             ? srmax[i]
                             : (__builtin_isfinite(srmin[i]) ? srmin[i] : 0.0);
        s = s + q * q;
    }
    *out = s;
}
void sumsq_q_seq(const double* srmin, const double* srmax, long m, double* out,
                 hipStream_t s) {
    k_x_sumsq_q<<<1, 64, 0, s>>>(srmin, srmax, m, out);
}

// -------------------------------------------------------------- KKT --
// metrics_of + evaluate (pdhg.cu:173-254, 694-733) of the two iterates
// q = 0, 1 (one workgroup each): A~'y~ and A~x~ come from k_x_spmv; the
// elementwise terms and maxima are formed in parallel, then thread 0 adds
// c'x and the Lagrangian bound in index order.  A skipped term is stored
// as +0, which leaves a running sum that starts at +0 unchanged, and the
// row terms are stored negated (s + (-t) == s - t).

constexpr int KT = 1024;

__global__ __launch_bounds__(KT) void k_x_kkt(KktData d, ExactKkt k, Metrics* out) {
    __shared__ double red[KT / 64];
    const int q = blockIdx.x;
    const int n = d.n, m = d.m;
    const double* x = k.x[q];
    const double* y = k.y[q];
    const double* ata = k.ata[q];
    const double* tm = k.tm[q];
    double* top = k.term + static_cast<long>(q) * (2L * n + m);
    double* tod = top + n;
    double* trow = tod + n;
    const double dz = 1e-8 * (1.0 + d.cmax);
    double dv = 0.0, unb = 0.0, pv = 0.0, pa = 0.0, prow = 0.0, ymax = 0.0;
    for (int j = threadIdx.x; j < n; j += KT) {
        const double hx = x[j] * d.dc[j];
        const double hd = d.c[j] + 1.0 * ata[j] / d.dc[j];
        const double cl = d.cl[j], cu = d.cu[j];
        const double bnd = fabs(__builtin_isfinite(cu) ? cu : cl);
        const double tol = 1e-9 * (1.0 + bnd);
        double v;
        if (hx <= cl + tol && hd > 0) v = 0.0;
        else if (hx >= cu - tol && hd < 0) v = 0.0;
        else v = fabs(hd);
        dv = fmax(dv, v);
        top[j] = d.c[j] * hx;
        double t = 0.0;
        if (hd >= 0) {
            if (__builtin_isfinite(cl)) t = hd * cl;
            else if (hd > dz) unb = 1.0;
        } else {
            if (__builtin_isfinite(cu)) t = hd * cu;
            else if (-hd > dz) unb = 1.0;
        }
        tod[j] = t;
    }
    for (int i = threadIdx.x; i < m; i += KT) {
        const double ax = tm[i] * d.dr[i];
        const double lo = d.rmin[i], hi = d.rmax[i];
        const double e = fmax(lo - ax, ax - hi);
        pv = fmax(pv, e);
        pa = fmax(pa, fabs(fmin(fmax(ax, lo), hi)));
        double sc = fabs(ax);
        if (__builtin_isfinite(lo)) sc = fmax(sc, fabs(lo));
        if (__builtin_isfinite(hi)) sc = fmax(sc, fabs(hi));
        prow = fmax(prow, fmax(0.0, e) / (1.0 + sc));
        ymax = fmax(ymax, fabs(y[i] * (1.0 / d.dr[i])));
    }
    dv = block_max<double, KT>(dv, red);
    unb = block_max<double, KT>(unb, red);
    pv = block_max<double, KT>(pv, red);
    pa = block_max<double, KT>(pa, red);
    prow = block_max<double, KT>(prow, red);
    ymax = block_max<double, KT>(ymax, red);
    const double yz = 1e-8 * (1.0 + ymax);
    for (int i = threadIdx.x; i < m; i += KT) {
        const double hy = y[i] * (1.0 / d.dr[i]);
        double t = 0.0;
        if (hy > yz) {
            if (__builtin_isfinite(d.rmax[i])) t = -(hy * d.rmax[i]);
            else unb = 1.0;
        } else if (hy < -yz) {
            if (__builtin_isfinite(d.rmin[i])) t = -(hy * d.rmin[i]);
            else unb = 1.0;
        }
        trow[i] = t;
    }
    unb = block_max<double, KT>(unb, red);
    __syncthreads();
    if (threadIdx.x == 0) {
        double op = 0.0, od = 0.0;
        for (int j = 0; j < n; ++j) op = op + top[j];
        for (int j = 0; j < n; ++j) od = od + tod[j];
        for (int i = 0; i < m; ++i) od = od + trow[i];
        Metrics mt;
        mt.pinf = fmax(pv, 0.0) / (1.0 + pa);
        mt.prow = prow;
        mt.dinf = dv / (1.0 + d.cmax);
        mt.op = op;
        const bool bounded = unb == 0.0;
        mt.od = bounded ? od : -__builtin_huge_val();
        mt.gap = bounded ? fmax(0.0, op - od) / (1.0 + fabs(op)) : __builtin_huge_val();
        out[q] = mt;
    }
}

void launch_kkt_exact(const KktData& d, const ExactKkt& k, Metrics* out, hipStream_t s) {
    for (int q = 0; q < 2; ++q) {
        spmv_seq(d.n, d.cp, d.ci, d.sacx, k.y[q], k.ata[q], s);
        spmv_seq(d.m, d.ap, d.ai, d.sax, k.x[q], k.tm[q], s);
    }
    k_x_kkt<<<2, KT, 0, s>>>(d, k, out);
}

}  // namespace igaos::dev
