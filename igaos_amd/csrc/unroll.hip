// Unrolled differentiation kernels (see unroll.hpp): K fixed-step PDHG
// steps per instance with checkpoints every 64 steps, and reverse mode
// through them.  One 256-thread workgroup per instance, FP64, original
// units; the instance SpMVs are the batched solver's (row-block streaming
// or lanes per row, the instance's schedule) on the original values.

#include "batch_dev.hpp"
#include "unroll.hpp"

namespace igaos::batch {

namespace {

constexpr double RHO = 0.9;

// clamp with the layer's tie rule: the lower bound wins when v <= lo,
// then the upper when v >= hi; code 0 inside, 1 lower, 2 upper
__device__ __forceinline__ double clamp_code(double v, double lo, double hi,
                                             unsigned char& code) {
    if (v <= lo) {
        code = 1;
        return lo;
    }
    if (v >= hi) {
        code = 2;
        return hi;
    }
    code = 0;
    return v;
}

// One PDHG step (x, y) -> (xo, yo); xo may alias x and yo may alias y (the
// primal half gathers only y, the dual half only xb).  Codes are written
// when cx / cy are not null.
template <typename V>
__device__ void step(const LP<V>& L, const double* x, const double* y, double* xo,
                     double* yo, double* xb, double tau, double sigma, unsigned char* cx,
                     unsigned char* cy, Lds& S) {
    const double* c = L.c();
    const double* cl = L.cl();
    const double* cu = L.cu();
    spmv_T<double>(L, L.acx(), dev::GatherPlain<double>{y}, S, [&](int j, double s) {
        const double v = x[j] - tau * (c[j] + s);
        unsigned char k;
        const double xn = clamp_code(v, cl[j], cu[j], k);
        xb[j] = 2.0 * xn - x[j];
        xo[j] = xn;
        if (cx) cx[j] = k;
    });
    __syncthreads();
    const double* rmin = L.rmin();
    const double* rmax = L.rmax();
    spmv_A<double>(L, L.ax(), dev::GatherPlain<double>{xb}, S, [&](int i, double s) {
        const double t = y[i] + sigma * s;
        unsigned char k;
        const double p = clamp_code(t / sigma, rmin[i], rmax[i], k);
        yo[i] = t - sigma * p;
        if (cy) cy[i] = k;
    });
    __syncthreads();
}

struct Inst {
    long co, ro, zo;
};

__device__ __forceinline__ Inst inst_of(const LPDesc& D) {
    return Inst{D.coff, D.roff, D.zoff};
}

// Step sizes from the data: |A| by 30 power iterations of A'A from
// v = 1/sqrt(n) (pdhg.cu:605-631), the primal weight of pdhg.cu:647-663 in
// original units.
template <typename V>
__device__ void steps_of(const LP<V>& L, const UnrollArrays& U, const Inst& I, Lds& S,
                         double& tau, double& sigma) {
    const int m = L.m, n = L.n, tid = threadIdx.x;
    double* v = U.tmpn + I.co;
    double* w = U.gb + I.co;
    double* t = U.tmpm + I.ro;
    const double v0 = n > 0 ? 1.0 / sqrt(static_cast<double>(n)) : 0.0;
    for (int j = tid; j < n; j += NT) v[j] = v0;
    __syncthreads();
    for (int it = 0; it < 30; ++it) {
        spmv_A<double>(L, L.ax(), dev::GatherPlain<double>{v}, S,
                       [&](int i, double s) { t[i] = s; });
        __syncthreads();
        double q = 0.0;
        spmv_T<double>(L, L.acx(), dev::GatherPlain<double>{t}, S,
                       [&](int j, double s) { w[j] = s; q += s * s; });
        const double nw = sqrt(bsum(q, S.red));
        if (nw < 1e-300) break;
        const double inv = 1.0 / nw;
        for (int j = tid; j < n; j += NT) v[j] = w[j] * inv;
        __syncthreads();
    }
    double q = 0.0;
    spmv_A<double>(L, L.ax(), dev::GatherPlain<double>{v}, S,
                   [&](int i, double s) { q += s * s; });
    double lam2 = bsum(q, S.red);
    if (!(lam2 > 0) || !__builtin_isfinite(lam2)) lam2 = 1.0;
    // primal weight (pdhg.cu:647-663, original units)
    double cn = 0.0, qn = 0.0;
    for (int j = tid; j < n; j += NT) cn += L.c()[j] * L.c()[j];
    for (int i = tid; i < m; i += NT) {
        const double qq = __builtin_isfinite(L.rmax()[i])   ? L.rmax()[i]
                          : __builtin_isfinite(L.rmin()[i]) ? L.rmin()[i]
                                                            : 0.0;
        qn += qq * qq;
    }
    cn = bsum(cn, S.red);
    qn = bsum(qn, S.red);
    double om = 1.0;
    if (cn > 1e-24 && qn > 1e-24) om = fmin(fmax(sqrt(cn / qn), 0.3), 3.0);
    const double na = sqrt(lam2);
    tau = RHO / (om * na);
    sigma = RHO * om / na;
}

template <typename V>
__global__ __launch_bounds__(NT) void k_unroll_fwd(const Arrays<V>* A, const LPDesc* lps,
                                                   const UnrollArrays* Up) {
    __shared__ Lds S;
    const UnrollArrays& U = *Up;
    const LPDesc D = lps[blockIdx.x];
    const LP<V> L = view(A, D);
    const Inst I = inst_of(D);
    const int m = L.m, n = L.n, tid = threadIdx.x;
    double* x = U.xK + I.co;
    double* y = U.yK + I.ro;
    double* xb = U.xb + I.co;

    double tau, sigma;
    if (U.fixed) {
        tau = U.tau[blockIdx.x];
        sigma = U.sigma[blockIdx.x];
    } else {
        steps_of(L, U, I, S, tau, sigma);
    }

    for (int j = tid; j < n; j += NT) x[j] = 0.0;
    for (int i = tid; i < m; i += NT) y[i] = 0.0;
    __syncthreads();
    for (int k = 0; k < U.K; ++k) {
        if (k % UNROLL_SEG == 0) {
            const long s = k / UNROLL_SEG;
            for (int j = tid; j < n; j += NT) U.ckx[s * U.sum_n + I.co + j] = x[j];
            for (int i = tid; i < m; i += NT) U.cky[s * U.sum_m + I.ro + i] = y[i];
        }
        step(L, x, y, x, y, xb, tau, sigma, nullptr, nullptr, S);
    }
    double ob = 0.0;
    for (int j = tid; j < n; j += NT) ob += L.c()[j] * x[j];
    ob = bsum(ob, S.red);
    if (tid == 0) {
        U.tau[blockIdx.x] = tau;
        U.sigma[blockIdx.x] = sigma;
        U.obj[blockIdx.x] = ob + D.obj_const;
    }
}


template <typename V>
__global__ __launch_bounds__(NT) void k_unroll_bwd(const Arrays<V>* A, const LPDesc* lps,
                                                   const UnrollArrays* Up) {
    __shared__ Lds S;
    const UnrollArrays& U = *Up;
    const LPDesc D = lps[blockIdx.x];
    const LP<V> L = view(A, D);
    const Inst I = inst_of(D);
    const int m = L.m, n = L.n, tid = threadIdx.x;
    const double tau = U.tau[blockIdx.x], sigma = U.sigma[blockIdx.x];
    const double gobj = U.gobj[blockIdx.x];
    const double* c = L.c();
    double* ax = U.ax + I.co;
    double* ay = U.ay + I.ro;
    double* xb = U.xb + I.co;
    double* gb = U.gb + I.co;
    double* tb = U.tb + I.ro;
    double* gc = U.gc + I.co;
    double* gcl = U.gcl + I.co;
    double* gcu = U.gcu + I.co;
    double* grmin = U.grmin + I.ro;
    double* grmax = U.grmax + I.ro;
    double* gax = U.gax + I.zo;
    unsigned char* cx = U.cx + I.co;
    unsigned char* cy = U.cy + I.ro;
    for (int j = tid; j < n; j += NT) {
        ax[j] = U.gx[I.co + j] + gobj * c[j];
        gc[j] = gobj * U.xK[I.co + j];
        gcl[j] = gcu[j] = 0.0;
    }
    for (int i = tid; i < m; i += NT) {
        ay[i] = U.gy[I.ro + i];
        grmin[i] = grmax[i] = 0.0;
    }
    const int* ap = L.ap();
    const int* ai = L.ai();
    for (int i = tid; i < m; i += NT)
        for (int p = ap[i]; p < ap[i + 1]; ++p) gax[p] = 0.0;
    __syncthreads();

    const int nseg = (U.K + UNROLL_SEG - 1) / UNROLL_SEG;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int k0 = sg * UNROLL_SEG;
        const int len = min(UNROLL_SEG, U.K - k0);
        // the segment's iterates: slot t holds (x, y) before step k0 + t
        double* sx0 = U.segx + I.co;
        double* sy0 = U.segy + I.ro;
        for (int j = tid; j < n; j += NT) sx0[j] = U.ckx[sg * U.sum_n + I.co + j];
        for (int i = tid; i < m; i += NT) sy0[i] = U.cky[sg * U.sum_m + I.ro + i];
        __syncthreads();
        for (int t = 0; t + 1 < len; ++t)
            step(L, U.segx + t * U.sum_n + I.co, U.segy + t * U.sum_m + I.ro,
                 U.segx + (t + 1) * U.sum_n + I.co, U.segy + (t + 1) * U.sum_m + I.ro, xb,
                 tau, sigma, nullptr, nullptr, S);
        for (int t = len - 1; t >= 0; --t) {
            const double* xt = U.segx + t * U.sum_n + I.co;
            const double* yt = U.segy + t * U.sum_m + I.ro;
            // forward pieces of the step: xb, codes (outputs unused)
            double* xo = U.tmpn + I.co;
            double* yo = U.tmpm + I.ro;
            step(L, xt, yt, xo, yo, xb, tau, sigma, cx, cy, S);
            // dual half: tbar and the row-bound gradients
            for (int i = tid; i < m; i += NT) {
                const double g = ay[i];
                const unsigned char k = cy[i];
                // y' = t - sigma bound where t / sigma is clamped (dy'/dt = 1),
                // y' = 0 inside (dy'/dt = 0)
                tb[i] = k != 0 ? g : 0.0;
                if (k == 1) grmin[i] -= sigma * g;
                if (k == 2) grmax[i] -= sigma * g;
                ay[i] = k != 0 ? g : 0.0;
            }
            __syncthreads();
            // sigma A'tbar, extrapolation and primal half
            spmv_T<double>(L, L.acx(), dev::GatherPlain<double>{tb}, S, [&](int j, double s) {
                const double xbb = sigma * s;
                const double h = ax[j] + 2.0 * xbb;
                const unsigned char k = cx[j];
                const double vb = k == 0 ? h : 0.0;
                if (k == 1) gcl[j] += h;
                if (k == 2) gcu[j] += h;
                const double g = -tau * vb;
                gc[j] += g;
                gb[j] = g;
                ax[j] = vb - xbb;
            });
            __syncthreads();
            spmv_A<double>(L, L.ax(), dev::GatherPlain<double>{gb}, S,
                           [&](int i, double s) { ay[i] += s; });
            // A's values: sigma tbar_i xb_j + y_i gbar_j
            for (int i = tid; i < m; i += NT) {
                const double ti = sigma * tb[i], yi = yt[i];
                for (int p = ap[i]; p < ap[i + 1]; ++p) {
                    const int j = ai[p];
                    gax[p] += ti * xb[j] + yi * gb[j];
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace

void launch_unroll_forward(const Arrays<float>* a, const LPDesc* lps, int count,
                           const UnrollArrays* u, hipStream_t s) {
    k_unroll_fwd<float><<<count, NT, 0, s>>>(a, lps, u);
}
void launch_unroll_forward(const Arrays<double>* a, const LPDesc* lps, int count,
                           const UnrollArrays* u, hipStream_t s) {
    k_unroll_fwd<double><<<count, NT, 0, s>>>(a, lps, u);
}
void launch_unroll_backward(const Arrays<float>* a, const LPDesc* lps, int count,
                            const UnrollArrays* u, hipStream_t s) {
    k_unroll_bwd<float><<<count, NT, 0, s>>>(a, lps, u);
}
void launch_unroll_backward(const Arrays<double>* a, const LPDesc* lps, int count,
                            const UnrollArrays* u, hipStream_t s) {
    k_unroll_bwd<double><<<count, NT, 0, s>>>(a, lps, u);
}

}  // namespace igaos::batch
