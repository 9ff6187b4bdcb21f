// Host-side pieces of the check that run only when the reference's dual
// repair schedule fires: KKT evaluation in original units
// (src/pdhg/pdhg.cu:173-254) and the POCS dual repair with a dense
// Cholesky of A_J A_J' + ridge (pdhg.cu:255-448).  They follow the
// reference's steps exactly so that status parity holds when only the
// duality-gap certificate lags.  The engine calls the repair only for
// m <= EngineConfig::max_repair_m (the dense m x m factor is O(m^3)).
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "pdhg.hpp"

namespace igaos::pdhg {

HostMetrics host_evaluate(const io::Model& lp, const double* x,
                          const double* y, const double* Ax, const double* d) {
    const int m = lp.m, n = lp.n;
    const double INF = std::numeric_limits<double>::infinity();
    HostMetrics mt{};
    double pv = 0.0, pa = 0.0, prow = 0.0;
    for (int i = 0; i < m; ++i) {
        const double lo = lp.rmin[i], hi = lp.rmax[i];
        const double e = std::max(lo - Ax[i], Ax[i] - hi);
        pv = std::max(pv, e);
        pa = std::max(pa, std::fabs(std::min(std::max(Ax[i], lo), hi)));
        double sc = std::fabs(Ax[i]);
        if (std::isfinite(lo)) sc = std::max(sc, std::fabs(lo));
        if (std::isfinite(hi)) sc = std::max(sc, std::fabs(hi));
        prow = std::max(prow, std::max(0.0, e) / (1.0 + sc));
    }
    mt.pinf = std::max(pv, 0.0) / (1.0 + pa);
    mt.prow = prow;
    double cmax = 0.0, ymax = 0.0, dv = 0.0;
    for (int j = 0; j < n; ++j) cmax = std::max(cmax, std::fabs(lp.c[j]));
    for (int i = 0; i < m; ++i) ymax = std::max(ymax, std::fabs(y[i]));
    for (int j = 0; j < n; ++j) {
        const double bnd = std::fabs(std::isfinite(lp.cu[j]) ? lp.cu[j] : lp.cl[j]);
        const double tol = 1e-9 * (1.0 + bnd);
        double v = std::fabs(d[j]);
        if (x[j] <= lp.cl[j] + tol && d[j] > 0) v = 0.0;
        else if (x[j] >= lp.cu[j] - tol && d[j] < 0) v = 0.0;
        dv = std::max(dv, v);
    }
    mt.dinf = dv / (1.0 + cmax);
    mt.op = 0.0;
    for (int j = 0; j < n; ++j) mt.op += lp.c[j] * x[j];
    const double dz = 1e-8 * (1.0 + cmax), yz = 1e-8 * (1.0 + ymax);
    double od = 0.0;
    bool bounded = true;
    for (int j = 0; j < n && bounded; ++j) {
        if (d[j] >= 0) {
            if (std::isfinite(lp.cl[j])) od += d[j] * lp.cl[j];
            else if (d[j] > dz) bounded = false;
        } else {
            if (std::isfinite(lp.cu[j])) od += d[j] * lp.cu[j];
            else if (-d[j] > dz) bounded = false;
        }
    }
    for (int i = 0; i < m && bounded; ++i) {
        if (y[i] > yz) {
            if (std::isfinite(lp.rmax[i])) od -= y[i] * lp.rmax[i];
            else bounded = false;
        } else if (y[i] < -yz) {
            if (std::isfinite(lp.rmin[i])) od -= y[i] * lp.rmin[i];
            else bounded = false;
        }
    }
    mt.od = bounded ? od : -INF;
    mt.gap = bounded ? std::max(0.0, mt.op - od) / (1.0 + std::fabs(mt.op)) : INF;
    return mt;
}

namespace {

// In-place lower Cholesky of a row-major nr x nr matrix; false if not PD.
bool cholesky(std::vector<double>& G, int nr) {
    for (int i = 0; i < nr; ++i) {
        double* gi = &G[static_cast<size_t>(i) * nr];
        for (int j = 0; j <= i; ++j) {
            const double* gj = &G[static_cast<size_t>(j) * nr];
            double s = gi[j];
            for (int t = 0; t < j; ++t) s -= gi[t] * gj[t];
            if (i == j) {
                if (!(s > 0)) return false;
                gi[i] = std::sqrt(s);
            } else {
                gi[j] = s / gj[j];
            }
        }
    }
    return true;
}

void chol_solve(const std::vector<double>& G, int nr, std::vector<double>& b) {
    for (int i = 0; i < nr; ++i) {
        const double* gi = &G[static_cast<size_t>(i) * nr];
        double s = b[i];
        for (int t = 0; t < i; ++t) s -= gi[t] * b[t];
        b[i] = s / gi[i];
    }
    for (int i = nr - 1; i >= 0; --i) {
        double s = b[i];
        for (int t = i + 1; t < nr; ++t) s -= G[static_cast<size_t>(t) * nr + i] * b[t];
        b[i] = s / G[static_cast<size_t>(i) * nr + i];
    }
}

double col_dot(const io::Model& lp, int j, const std::vector<double>& y) {
    double dj = lp.c[j];
    for (int p = lp.cp[j]; p < lp.cp[j + 1]; ++p) dj += lp.acx[p] * y[lp.ci[p]];
    return dj;
}

}  // namespace

bool host_dual_repair(const io::Model& lp, const double* x, const double* Ax,
                      const double* y_it, double col_tol, double* y_out) {
    const int m = lp.m, n = lp.n;
    const double INF = std::numeric_limits<double>::infinity();
    // Row dual cones: y >= 0 on <= rows, y <= 0 on >= rows, free on ranged
    // or equality rows, zero on free rows (d = c + A'y convention).
    std::vector<double> lo(m), hi(m);
    for (int i = 0; i < m; ++i) {
        const bool fl = std::isfinite(lp.rmin[i]), fu = std::isfinite(lp.rmax[i]);
        if (fl && fu) { lo[i] = -INF; hi[i] = INF; }
        else if (fu) { lo[i] = 0.0; hi[i] = INF; }
        else if (fl) { lo[i] = -INF; hi[i] = 0.0; }
        else { lo[i] = hi[i] = 0.0; }
    }
    // Column classes: 0 d>=0, 1 d<=0, 2 d==0, 3 unconstrained.
    std::vector<char> ct(n);
    std::vector<double> cn2(n, 0.0);
    std::vector<int> J;
    for (int j = 0; j < n; ++j) {
        const bool ui = !std::isfinite(lp.cu[j]), li = !std::isfinite(lp.cl[j]);
        ct[j] = (ui && li) ? 2 : ui ? 0 : li ? 1 : 3;
        for (int p = lp.cp[j]; p < lp.cp[j + 1]; ++p) cn2[j] += lp.acx[p] * lp.acx[p];
        const double bl = col_tol * (1.0 + (std::isfinite(lp.cl[j]) ? std::fabs(lp.cl[j]) : 0.0));
        const double bu = col_tol * (1.0 + (std::isfinite(lp.cu[j]) ? std::fabs(lp.cu[j]) : 0.0));
        if (x[j] - lp.cl[j] > bl && lp.cu[j] - x[j] > bu) J.push_back(j);
    }
    const int k = static_cast<int>(J.size());
    std::vector<double> G;
    if (k > 0) {
        G.assign(static_cast<size_t>(m) * m, 0.0);
        for (int j : J)
            for (int p = lp.cp[j]; p < lp.cp[j + 1]; ++p) {
                const double v1 = lp.acx[p];
                double* row = &G[static_cast<size_t>(lp.ci[p]) * m];
                for (int q = lp.cp[j]; q < lp.cp[j + 1]; ++q) row[lp.ci[q]] += v1 * lp.acx[q];
            }
        double tr = 0.0;
        for (int i = 0; i < m; ++i) tr += G[static_cast<size_t>(i) * m + i];
        double ridge = tr > 0 ? 1e-10 * tr / m : 1e-14;
        bool ok = false;
        for (int att = 0; att < 6 && !ok; ++att) {
            std::vector<double> Gc = G;
            for (int i = 0; i < m; ++i) Gc[static_cast<size_t>(i) * m + i] += ridge;
            ok = cholesky(Gc, m);
            if (ok) G.swap(Gc);
            else ridge = ridge > 0 ? ridge * 100.0 : 1e-14;
        }
        if (!ok) return false;
    }
    double best_gap = INF, best_viol = INF;
    bool have = false, have_viol = false;
    std::vector<double> d(n), rhs(m), yc(m), ybest, yviol;
    for (int anch = 0; anch < 2; ++anch) {
        for (int i = 0; i < m; ++i)
            yc[i] = std::min(std::max(anch == 0 ? y_it[i] : 0.0, lo[i]), hi[i]);
        double vref = -1.0;
        for (int it = 0; it < 300; ++it) {
            if (k > 0) {  // affine projection onto d_J = 0
                std::fill(rhs.begin(), rhs.end(), 0.0);
                double res2 = 0.0;
                for (int j : J) {
                    const double dj = col_dot(lp, j, yc);
                    res2 += dj * dj;
                    if (dj != 0.0)
                        for (int p = lp.cp[j]; p < lp.cp[j + 1]; ++p) rhs[lp.ci[p]] -= lp.acx[p] * dj;
                }
                if (res2 > 0.0) {
                    chol_solve(G, m, rhs);
                    for (int i = 0; i < m; ++i) yc[i] += rhs[i];
                }
            }
            for (int j = 0; j < n; ++j) {  // Hildreth half-space steps
                if (ct[j] == 3 || cn2[j] == 0.0) continue;
                const double dj = col_dot(lp, j, yc);
                double v = 0.0;
                if (ct[j] == 0 && dj < 0.0) v = -dj;
                else if (ct[j] == 1 && dj > 0.0) v = -dj;
                else if (ct[j] == 2 && dj != 0.0) v = -dj;
                if (v != 0.0) {
                    const double st = v / cn2[j];
                    for (int p = lp.cp[j]; p < lp.cp[j + 1]; ++p) yc[lp.ci[p]] += st * lp.acx[p];
                }
            }
            for (int i = 0; i < m; ++i) yc[i] = std::min(std::max(yc[i], lo[i]), hi[i]);
            for (int j = 0; j < n; ++j) d[j] = col_dot(lp, j, yc);
            if (it % 25 == 0) {
                double viol = 0.0;
                for (int j = 0; j < n; ++j) {
                    double v = 0.0;
                    if (ct[j] == 0 && d[j] < 0) v = -d[j];
                    else if (ct[j] == 1 && d[j] > 0) v = d[j];
                    else if (ct[j] == 2) v = std::fabs(d[j]);
                    viol = std::max(viol, v);
                }
                for (int i = 0; i < m; ++i) {
                    double v = 0.0;
                    if (lo[i] == 0.0 && hi[i] == 0.0) v = std::fabs(yc[i]);
                    else if (hi[i] == 0.0 && yc[i] > 0) v = yc[i];
                    else if (lo[i] == 0.0 && yc[i] < 0) v = -yc[i];
                    viol = std::max(viol, v);
                }
                if (it == 25) vref = viol;
                if (viol < best_viol) {
                    best_viol = viol;
                    yviol = yc;
                    have_viol = true;
                }
                if (it == 100 && vref > 0 && viol > 0.25 * vref) break;
            }
            const HostMetrics t = host_evaluate(lp, x, yc.data(), Ax, d.data());
            if (std::isfinite(t.gap) && t.gap < best_gap) {
                best_gap = t.gap;
                ybest = yc;
                have = true;
            }
        }
    }
    if (have) std::copy(ybest.begin(), ybest.end(), y_out);
    else if (have_viol) std::copy(yviol.begin(), yviol.end(), y_out);
    return have;
}

}  // namespace igaos::pdhg
