// Device dual repair: the check-time certification fallback of
// src/pdhg/pdhg.cu:288-448 (schedule 854-931) for LPs of any size.
//
// The reference repairs the dual when the residuals are at tolerance but
// the duality gap is not certified: with x fixed, it alternates projections
// of y (POCS) onto {y : d_J = 0} (d = c + A'y, J the columns strictly
// inside their bounds) -- through a dense Cholesky factor of A_J A_J' +
// ridge, O(m^2) memory and O(m^3) time, its dense index overflowing int
// above m = 46,340 (pdhg.cu:327) -- onto the sign halfspaces of the other
// columns, one column after another, and onto the rows' sign cones, and
// keeps the y with the best finite gap.  Here every step runs on the GPU
// over the sparse matrix:
//  * the affine projection solves (A_J A_J' + ridge I) delta = -A_J d_J by
//    conjugate gradients (two sparse products per iteration, the scalars
//    kept in device memory so the host launches without synchronising);
//  * the halfspace steps are taken simultaneously (component averaging:
//    row i moves by the mean of the steps of the violated columns touching
//    it) instead of one column after another;
//  * the row cones are applied elementwise, and the gap of each iterate is
//    reduced on the device; the host reads three numbers per iteration.
// The ridge, the column tolerance tiers, the two anchors (the iterate's
// dual and zero), the 300-iteration cap and the early stop of pdhg.cu are
// kept.  All reductions have a fixed order: runs are deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "dev_common.hpp"
#include "repair_dev.hpp"

namespace igaos::dev {

namespace {

constexpr int RB = 256;
constexpr int NRED = 1024;  // partials per reduction (fixed grid)

#define RCK(expr)                                                               \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess)                                                   \
            throw std::runtime_error(std::string("HIP error in device repair: ") + \
                                     hipGetErrorString(e_) + " at " #expr);     \
    } while (0)

int eg(long k) { return static_cast<int>((k + RB - 1) / RB); }
int rg(long k) {
    long b = (k + RB - 1) / RB;
    return static_cast<int>(b < 1 ? 1 : (b > NRED ? NRED : b));
}

// column types (pdhg.cu:303-316): 0 lower bound only (d >= 0), 1 upper only
// (d <= 0), 2 free (d = 0), 3 boxed (any d); J = strictly inside by tol
__global__ void k_rep_cols(int n, const double* cl, const double* cu, const double* x,
                           double col_tol, const int* cp, const double* acx,
                           unsigned char* ct, unsigned char* jm, double* cn2) {
    const long j = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (j >= n) return;
    const bool up_inf = !__builtin_isfinite(cu[j]), lo_inf = !__builtin_isfinite(cl[j]);
    ct[j] = (up_inf && lo_inf) ? 2 : up_inf ? 0 : lo_inf ? 1 : 3;
    const double bl = col_tol * (1.0 + (lo_inf ? 0.0 : fabs(cl[j])));
    const double bu = col_tol * (1.0 + (up_inf ? 0.0 : fabs(cu[j])));
    jm[j] = (x[j] - cl[j] > bl && cu[j] - x[j] > bu) ? 1 : 0;
    double s = 0.0;
    for (int p = cp[j]; p < cp[j + 1]; ++p) s += acx[p] * acx[p];
    cn2[j] = s;
}

// row sign cones (pdhg.cu:292-301)
__global__ void k_rep_rows(int m, const double* rmin, const double* rmax, double* lo,
                           double* hi) {
    const long i = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (i >= m) return;
    const bool fl = __builtin_isfinite(rmin[i]), fu = __builtin_isfinite(rmax[i]);
    const double INF = __builtin_huge_val();
    lo[i] = (fl && fu) ? -INF : fu ? 0.0 : fl ? -INF : 0.0;
    hi[i] = (fl && fu) ? INF : fu ? INF : fl ? 0.0 : 0.0;
}

__global__ void k_clamp(int m, const double* src, const double* lo, const double* hi,
                        double* y) {
    const long i = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (i < m) y[i] = fmin(fmax(src ? src[i] : 0.0, lo[i]), hi[i]);
}

// d = c + t (t = A'y), optionally masked to J and negated: out_j = -d_j on J
__global__ void k_reduced(int n, const double* c, const double* t, const unsigned char* jm,
                          double* d, double* bj) {
    const long j = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (j >= n) return;
    const double v = c[j] + t[j];
    d[j] = v;
    if (bj) bj[j] = jm[j] ? -v : 0.0;
}

// partial sums of a.b (b may be a) into part[blockIdx.x]
__global__ __launch_bounds__(RB) void k_dot(long n, const double* a, const double* b,
                                            double* part) {
    __shared__ double red[RB / WAVE];
    double s = 0.0;
    for (long i = static_cast<long>(blockIdx.x) * RB + threadIdx.x; i < n;
         i += static_cast<long>(gridDim.x) * RB)
        s += a[i] * b[i];
    s = block_sum<double, RB>(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// CG scalars live in sc[]: 0 rr, 1 rr0, 2 alpha, 3 beta, 4 done
__global__ __launch_bounds__(FIN) void k_cg_init(const double* part, int cnt, double* sc) {
    __shared__ double red[FIN / WAVE];
    const double rr = fin_sum(part, cnt, 1, red);
    if (threadIdx.x == 0) {
        sc[0] = rr;
        sc[1] = rr;
        sc[4] = rr > 0.0 ? 0.0 : 1.0;
    }
}
__global__ __launch_bounds__(FIN) void k_cg_alpha(const double* part, int cnt, double* sc) {
    __shared__ double red[FIN / WAVE];
    const double pq = fin_sum(part, cnt, 1, red);
    if (threadIdx.x == 0) sc[2] = (sc[4] == 0.0 && pq > 0.0) ? sc[0] / pq : 0.0;
}
__global__ __launch_bounds__(FIN) void k_cg_beta(const double* part, int cnt, double tol2,
                                                 double* sc) {
    __shared__ double red[FIN / WAVE];
    const double rr = fin_sum(part, cnt, 1, red);
    if (threadIdx.x == 0 && sc[4] == 0.0) {
        sc[3] = sc[0] > 0.0 ? rr / sc[0] : 0.0;
        sc[0] = rr;
        if (!(rr > tol2 * sc[1])) sc[4] = 1.0;
    }
}
// q += ridge p (q = A_J A_J' p computed before)
__global__ void k_cg_ridge(int m, const double* p, double ridge, double* q) {
    const long i = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (i < m) q[i] = q[i] + ridge * p[i];
}
__global__ void k_mask(int n, const unsigned char* jm, double* t) {
    const long j = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (j < n && !jm[j]) t[j] = 0.0;
}
// delta += alpha p; r -= alpha q (no-ops once done)
__global__ void k_cg_step(int m, const double* sc, const double* p, const double* q,
                          double* delta, double* r) {
    const long i = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (i >= m || sc[4] != 0.0) return;
    const double a = sc[2];
    delta[i] += a * p[i];
    r[i] -= a * q[i];
}
// p = r + beta p (uses the beta of the step just taken)
__global__ void k_cg_dir(int m, const double* sc, const double* r, double* p) {
    const long i = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (i < m && sc[4] == 0.0) p[i] = r[i] + sc[3] * p[i];
}
__global__ void k_add(int m, const double* delta, double* y) {
    const long i = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (i < m) y[i] += delta[i];
}

// halfspace steps of the sign-constrained columns (pdhg.cu:372-392), taken
// together: w_j = v_j / |a_j|^2 for a violated column, then each row moves
// by the mean of its violated columns' steps and is clamped to its cone
__global__ void k_half_w(int n, const unsigned char* ct, const double* cn2, const double* d,
                         double* w) {
    const long j = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (j >= n) return;
    double v = 0.0;
    if (ct[j] != 3 && cn2[j] != 0.0) {
        const double dj = d[j];
        if (ct[j] == 0 && dj < 0.0) v = -dj;
        else if (ct[j] == 1 && dj > 0.0) v = -dj;
        else if (ct[j] == 2 && dj != 0.0) v = -dj;
    }
    w[j] = v != 0.0 ? v / cn2[j] : 0.0;
}
__global__ void k_half_apply(int m, const int* ap, const int* ai, const double* ax,
                             const double* w, const double* lo, const double* hi,
                             double* y) {
    const long i = static_cast<long>(blockIdx.x) * RB + threadIdx.x;
    if (i >= m) return;
    double s = 0.0;
    int cnt = 0;
    for (int p = ap[i]; p < ap[i + 1]; ++p) {
        const double wj = w[ai[p]];
        if (wj != 0.0) {
            s += wj * ax[p];
            ++cnt;
        }
    }
    const double v = cnt ? y[i] + s / cnt : y[i];
    y[i] = fmin(fmax(v, lo[i]), hi[i]);
}

// per-block partials of the Lagrangian bound (pdhg.cu:224-244), the dual
// residual numerator and the sign violation (pdhg.cu:396-414): part layout
// [od, unb, dv, viol, ymax] x gridDim.x
__global__ __launch_bounds__(RB) void k_rep_eval_cols(int n, const double* cl, const double* cu,
                                                      const double* x, const double* d,
                                                      const unsigned char* ct, double dz,
                                                      double* part) {
    __shared__ double red[RB / WAVE];
    double od = 0.0, unb = 0.0, dv = 0.0, vi = 0.0;
    for (long j = static_cast<long>(blockIdx.x) * RB + threadIdx.x; j < n;
         j += static_cast<long>(gridDim.x) * RB) {
        const double dj = d[j];
        if (dj >= 0) {
            if (__builtin_isfinite(cl[j])) od += dj * cl[j];
            else if (dj > dz) unb = 1.0;
        } else {
            if (__builtin_isfinite(cu[j])) od += dj * cu[j];
            else if (-dj > dz) unb = 1.0;
        }
        const double bnd = fabs(__builtin_isfinite(cu[j]) ? cu[j] : cl[j]);
        const double tol = 1e-9 * (1.0 + bnd);
        double v = fabs(dj);
        if (x[j] <= cl[j] + tol && dj > 0) v = 0.0;
        else if (x[j] >= cu[j] - tol && dj < 0) v = 0.0;
        dv = fmax(dv, v);
        double s = 0.0;
        if (ct[j] == 0 && dj < 0) s = -dj;
        else if (ct[j] == 1 && dj > 0) s = dj;
        else if (ct[j] == 2) s = fabs(dj);
        vi = fmax(vi, s);
    }
    od = block_sum<double, RB>(od, red);
    unb = block_max<double, RB>(unb, red);
    dv = block_max<double, RB>(dv, red);
    vi = block_max<double, RB>(vi, red);
    if (threadIdx.x == 0) {
        part[0 * NRED + blockIdx.x] = od;
        part[1 * NRED + blockIdx.x] = unb;
        part[2 * NRED + blockIdx.x] = dv;
        part[3 * NRED + blockIdx.x] = vi;
    }
}
__global__ __launch_bounds__(RB) void k_rep_ymax(int m, const double* y, double* part) {
    __shared__ double red[RB / WAVE];
    double s = 0.0;
    for (long i = static_cast<long>(blockIdx.x) * RB + threadIdx.x; i < m;
         i += static_cast<long>(gridDim.x) * RB)
        s = fmax(s, fabs(y[i]));
    s = block_max<double, RB>(s, red);
    if (threadIdx.x == 0) part[4 * NRED + blockIdx.x] = s;
}
__global__ __launch_bounds__(RB) void k_rep_eval_rows(int m, const double* rmin,
                                                      const double* rmax, const double* y,
                                                      const double* lo, const double* hi,
                                                      const double* ymx, double* part) {
    __shared__ double red[RB / WAVE];
    const double yz = 1e-8 * (1.0 + *ymx);
    double od = 0.0, unb = 0.0, vi = 0.0;
    for (long i = static_cast<long>(blockIdx.x) * RB + threadIdx.x; i < m;
         i += static_cast<long>(gridDim.x) * RB) {
        const double yi = y[i];
        if (yi > yz) {
            if (__builtin_isfinite(rmax[i])) od -= yi * rmax[i];
            else unb = 1.0;
        } else if (yi < -yz) {
            if (__builtin_isfinite(rmin[i])) od -= yi * rmin[i];
            else unb = 1.0;
        }
        double v = 0.0;
        if (lo[i] == 0.0 && hi[i] == 0.0) v = fabs(yi);
        else if (hi[i] == 0.0 && yi > 0) v = yi;
        else if (lo[i] == 0.0 && yi < 0) v = -yi;
        vi = fmax(vi, v);
    }
    od = block_sum<double, RB>(od, red);
    unb = block_max<double, RB>(unb, red);
    vi = block_max<double, RB>(vi, red);
    if (threadIdx.x == 0) {
        part[5 * NRED + blockIdx.x] = od;
        part[6 * NRED + blockIdx.x] = unb;
        part[7 * NRED + blockIdx.x] = vi;
    }
}
// out: 0 od (cols + rows), 1 unbounded flag, 2 dinf numerator, 3 sign
// violation, 4 max|y|
__global__ __launch_bounds__(FIN) void k_rep_fin(const double* part, int gc, int gr,
                                                 double* out) {
    __shared__ double red[FIN / WAVE];
    const double odc = fin_sum(part + 0 * NRED, gc, 1, red);
    const double unc = fin_max(part + 1 * NRED, gc, 1, red);
    const double dv = fin_max(part + 2 * NRED, gc, 1, red);
    const double vc = fin_max(part + 3 * NRED, gc, 1, red);
    const double odr = fin_sum(part + 5 * NRED, gr, 1, red);
    const double unr = fin_max(part + 6 * NRED, gr, 1, red);
    const double vr = fin_max(part + 7 * NRED, gr, 1, red);
    if (threadIdx.x == 0) {
        out[0] = odc + odr;
        out[1] = fmax(unc, unr);
        out[2] = dv;
        out[3] = fmax(vc, vr);
    }
}
__global__ __launch_bounds__(FIN) void k_rep_fin_ymax(const double* part, int gr, double* out) {
    __shared__ double red[FIN / WAVE];
    const double v = fin_max(part + 4 * NRED, gr, 1, red);
    if (threadIdx.x == 0) out[4] = v;
}

__global__ __launch_bounds__(RB) void k_trace(int n, const unsigned char* jm, const double* cn2,
                                              double* part) {
    __shared__ double red[RB / WAVE];
    double tr = 0.0, k = 0.0;
    for (long j = static_cast<long>(blockIdx.x) * RB + threadIdx.x; j < n;
         j += static_cast<long>(gridDim.x) * RB)
        if (jm[j]) {
            tr += cn2[j];
            k += 1.0;
        }
    tr = block_sum<double, RB>(tr, red);
    k = block_sum<double, RB>(k, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = tr;
        part[NRED + blockIdx.x] = k;
    }
}
__global__ __launch_bounds__(FIN) void k_trace_fin(const double* part, int cnt, double* out) {
    __shared__ double red[FIN / WAVE];
    const double tr = fin_sum(part, cnt, 1, red);
    const double k = fin_sum(part + NRED, cnt, 1, red);
    if (threadIdx.x == 0) {
        out[5] = tr;
        out[6] = k;
    }
}

}  // namespace

void RepairDev::alloc(int m_, int n_) {
    if (m_ == m && n_ == n && buf) return;
    release();
    m = m_;
    n = n_;
    const size_t dn = n > 0 ? n : 1, dm = m > 0 ? m : 1;
    const size_t doubles = 5 * dn + 10 * dm + 8 * NRED + 16;
    RCK(hipMalloc(&buf, doubles * sizeof(double) + 2 * dn));
    double* p = static_cast<double*>(buf);
    cn2 = p; p += dn;
    d = p; p += dn;
    t = p; p += dn;
    w = p; p += dn;
    bj = p; p += dn;
    lo = p; p += dm;
    hi = p; p += dm;
    y = p; p += dm;
    ybest = p; p += dm;
    yviol = p; p += dm;
    delta = p; p += dm;
    r = p; p += dm;
    pp = p; p += dm;
    q = p; p += dm;
    tmpm = p; p += dm;
    part = p; p += 8 * NRED;
    sc = p; p += 16;
    ct = reinterpret_cast<unsigned char*>(p);
    jm = ct + dn;
}

void RepairDev::release() {
    if (buf) (void)hipFree(buf);
    buf = nullptr;
}

// The repair's evaluation of a dual y (x fixed): out = od, unbounded flag,
// dinf numerator, sign violation, max|y| (on the host)
void RepairDev::evaluate(const RepairData& D, const double* x, const double* yv, double cmax,
                         hipStream_t s, double out[5]) {
    spmv(D.rbT, D.cp, D.ci, D.acx, yv, t, s);
    k_reduced<<<eg(n), RB, 0, s>>>(n, D.c, t, jm, d, nullptr);
    const int gc = rg(n), gr = rg(m);
    k_rep_ymax<<<gr, RB, 0, s>>>(m, yv, part);
    k_rep_fin_ymax<<<1, FIN, 0, s>>>(part, gr, sc + 8);
    k_rep_eval_cols<<<gc, RB, 0, s>>>(n, D.cl, D.cu, x, d, ct, 1e-8 * (1.0 + cmax), part);
    k_rep_eval_rows<<<gr, RB, 0, s>>>(m, D.rmin, D.rmax, yv, lo, hi, sc + 12, part);
    k_rep_fin<<<1, FIN, 0, s>>>(part, gc, gr, sc + 8);
    RCK(hipMemcpyAsync(out, sc + 8, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
    RCK(hipStreamSynchronize(s));
}

bool RepairDev::run(const RepairData& D, const double* x, const double* y_it, double col_tol,
                    double op, double cmax, hipStream_t s, double* y_out, RepairInfo* info) {
    const double INF = std::numeric_limits<double>::infinity();
    if (m <= 0 || n <= 0) return false;
    const int gn = rg(n), gm = rg(m);
    k_rep_cols<<<eg(n), RB, 0, s>>>(n, D.cl, D.cu, x, col_tol, D.cp, D.acx, ct, jm, cn2);
    k_rep_rows<<<eg(m), RB, 0, s>>>(m, D.rmin, D.rmax, lo, hi);
    k_trace<<<gn, RB, 0, s>>>(n, jm, cn2, part);
    k_trace_fin<<<1, FIN, 0, s>>>(part, gn, sc + 8);
    double tk[2];
    RCK(hipMemcpyAsync(tk, sc + 13, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    RCK(hipStreamSynchronize(s));
    const double tr = tk[0];
    const long kJ = static_cast<long>(tk[1]);
    const double ridge = tr > 0 ? 1e-10 * tr / (m > 0 ? m : 1) : 1e-14;
    double best_gap = INF, best_viol = INF;
    bool have = false, have_viol = false;
    long cg_total = 0, its_total = 0;
    for (int anch = 0; anch < 2; ++anch) {
        k_clamp<<<eg(m), RB, 0, s>>>(m, anch == 0 ? y_it : nullptr, lo, hi, y);
        double vref = -1.0;
        for (int it = 0; it < 300; ++it) {
            ++its_total;
            if (kJ > 0 && m > 0) {
                // affine projection onto d_J = 0: (A_J A_J' + ridge) delta = -A_J d_J
                spmv(D.rbT, D.cp, D.ci, D.acx, y, t, s);
                k_reduced<<<eg(n), RB, 0, s>>>(n, D.c, t, jm, d, bj);
                spmv(D.rbA, D.ap, D.ai, D.ax, bj, r, s);
                RCK(hipMemcpyAsync(pp, r, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
                RCK(hipMemsetAsync(delta, 0, sizeof(double) * m, s));
                k_dot<<<gm, RB, 0, s>>>(m, r, r, part);
                k_cg_init<<<1, FIN, 0, s>>>(part, gm, sc);
                for (int k = 0; k < max_cg; ++k) {
                    spmv(D.rbT, D.cp, D.ci, D.acx, pp, t, s);
                    k_mask<<<eg(n), RB, 0, s>>>(n, jm, t);
                    spmv(D.rbA, D.ap, D.ai, D.ax, t, q, s);
                    k_cg_ridge<<<eg(m), RB, 0, s>>>(m, pp, ridge, q);
                    k_dot<<<gm, RB, 0, s>>>(m, pp, q, part);
                    k_cg_alpha<<<1, FIN, 0, s>>>(part, gm, sc);
                    k_cg_step<<<eg(m), RB, 0, s>>>(m, sc, pp, q, delta, r);
                    k_dot<<<gm, RB, 0, s>>>(m, r, r, part);
                    k_cg_beta<<<1, FIN, 0, s>>>(part, gm, cg_tol * cg_tol, sc);
                    k_cg_dir<<<eg(m), RB, 0, s>>>(m, sc, r, pp);
                    ++cg_total;
                    if (k % 8 == 7) {  // stop launching once converged
                        double done = 0.0;
                        RCK(hipMemcpyAsync(&done, sc + 4, sizeof(double),
                                           hipMemcpyDeviceToHost, s));
                        RCK(hipStreamSynchronize(s));
                        if (done != 0.0) break;
                    }
                }
                k_add<<<eg(m), RB, 0, s>>>(m, delta, y);
            }
            // halfspace steps of the sign-constrained columns, then the cones
            spmv(D.rbT, D.cp, D.ci, D.acx, y, t, s);
            k_reduced<<<eg(n), RB, 0, s>>>(n, D.c, t, jm, d, nullptr);
            k_half_w<<<eg(n), RB, 0, s>>>(n, ct, cn2, d, w);
            k_half_apply<<<eg(m), RB, 0, s>>>(m, D.ap, D.ai, D.ax, w, lo, hi, y);
            double ev[5];
            evaluate(D, x, y, cmax, s, ev);
            const double gap = ev[1] != 0.0 ? INF : std::fmax(0.0, op - ev[0]) / (1.0 + std::fabs(op));
            if (it % 25 == 0) {
                const double viol = ev[3];
                if (it == 25) vref = viol;
                if (viol < best_viol) {
                    best_viol = viol;
                    RCK(hipMemcpyAsync(yviol, y, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
                    have_viol = true;
                }
                if (it == 100 && vref > 0 && viol > 0.25 * vref) break;
            }
            if (std::isfinite(gap) && gap < best_gap) {
                best_gap = gap;
                RCK(hipMemcpyAsync(ybest, y, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
                have = true;
            }
        }
    }
    const double* src = have ? ybest : (have_viol ? yviol : nullptr);
    if (src && m > 0)
        RCK(hipMemcpyAsync(y_out, src, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    if (info) {
        info->gap = best_gap;
        info->dinf = INF;
        if (have) {
            double ev[5];
            evaluate(D, x, ybest, cmax, s, ev);
            info->dinf = ev[2] / (1.0 + cmax);
            info->od = ev[0];
        }
        info->iterations = its_total;
        info->cg_iterations = cg_total;
        info->columns_in_J = kJ;
    }
    RCK(hipStreamSynchronize(s));
    return have;
}

namespace {
// device copies freed on every exit
struct UploadSet {
    std::vector<void*> ps;
    ~UploadSet() {
        for (void* p : ps) (void)hipFree(p);
    }
    template <typename T>
    T* up(const T* h, size_t k) {
        void* p = nullptr;
        RCK(hipMalloc(&p, sizeof(T) * (k > 0 ? k : 1)));
        ps.push_back(p);
        if (k && h) RCK(hipMemcpy(p, h, sizeof(T) * k, hipMemcpyHostToDevice));
        return static_cast<T*>(p);
    }
};
}  // namespace

// Host-array entry: uploads the model, runs the repair, downloads y.
bool dual_repair_device(const io::Model& md, const double* x, const double* y_it,
                        double col_tol, double* y_out, RepairInfo* info) {
    const int m = md.m, n = md.n;
    const long nnz = md.nnz();
    UploadSet M;
    const std::vector<int> rbA = build_row_blocks(md.ap.data(), m);
    const std::vector<int> rbT = build_row_blocks(md.cp.data(), n);
    RepairData D;
    D.ap = M.up(md.ap.data(), md.ap.size());
    D.ai = M.up(md.ai.data(), nnz);
    D.ax = M.up(md.ax.data(), nnz);
    D.cp = M.up(md.cp.data(), md.cp.size());
    D.ci = M.up(md.ci.data(), nnz);
    D.acx = M.up(md.acx.data(), nnz);
    D.rbA = {M.up(rbA.data(), rbA.size()), static_cast<int>(rbA.size() / 2) - 1};
    D.rbT = {M.up(rbT.data(), rbT.size()), static_cast<int>(rbT.size() / 2) - 1};
    D.c = M.up(md.c.data(), n);
    D.cl = M.up(md.cl.data(), n);
    D.cu = M.up(md.cu.data(), n);
    D.rmin = M.up(md.rmin.data(), m);
    D.rmax = M.up(md.rmax.data(), m);
    const double* dx = M.up(x, n);
    const double* dy = M.up(y_it, m);
    double* yo = M.up<double>(nullptr, m);
    double op = 0.0, cmax = 0.0;
    for (int j = 0; j < n; ++j) {
        op += md.c[j] * x[j];
        cmax = std::fmax(cmax, std::fabs(md.c[j]));
    }
    RepairDev R;
    R.alloc(m, n);
    const bool have = R.run(D, dx, dy, col_tol, op, cmax, nullptr, yo, info);
    if (m > 0) RCK(hipMemcpy(y_out, yo, sizeof(double) * m, hipMemcpyDeviceToHost));
    return have;
}

}  // namespace igaos::dev
