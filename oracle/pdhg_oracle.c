/* TEST INFRASTRUCTURE ONLY — the oracle.
 *
 * A plain-C FP64 host restatement of the reference restarted adaptive PDHG
 * LP engine, src/pdhg/pdhg.cu (igaos::pdhg::solve, pdhg.cu:457-1156).  Each
 * function names the reference lines it follows.  It is the checker for the
 * HIP engine: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it; the product never does.
 *
 * Numerics: compiled with -ffp-contract=off so every a*b+c is two roundings,
 * as written.  SpMV sums each row sequentially in storage order.  The
 * line-search dot products follow the reference's fixed two-stage layout
 * (pdhg.cu:63-125, gD = min(ceil(max(n,m)/256), 4096) blocks of 256
 * threads, grid-stride accumulation, shared-memory tree) so that the
 * accept/reject branch sees the same partial order the reference's GPU
 * code would.  The norm in the power iteration is a sequential sum of
 * squares (the reference calls cuBLAS Dnrm2, whose order is not published).
 *
 * Pinning (what each function is checked against) is stated in
 * tests/test_oracle.py (the dual repair: tests/test_boundary.py); a
 * mutation switch (orc_set_mutation) plants
 * plausible mistakes so those tests can show the pins catch them.
 */
#include "pdhg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static int g_mut = 0;
void orc_set_mutation(int code) { g_mut = code; }

/* Host threads of the SpMV (rows split among them; each row still sums in
 * storage order, so results do not depend on the count).  Default 1. */
static int g_threads = 1;
void orc_set_threads(int k) { g_threads = k > 0 ? k : 1; }

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }
/* CUDA fmin/fmax semantics (NaN-ignoring), as used in pdhg.cu:41,56 */
static double cfmin(double a, double b) { return fmin(a, b); }
static double cfmax(double a, double b) { return fmax(a, b); }
static double dclamp(double v, double lo, double hi) {
    /* std::clamp(v, lo, hi) as used in pdhg.cu:662,1057 */
    return v < lo ? lo : (hi < v ? hi : v);
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* CSR SpMV, beta = 0: the operation cusparseSpMV performs at
 * pdhg.cu:150-156.  Sequential per-row sum in storage order. */
void orc_spmv(int nrows, const int* ptr, const int* idx, const double* val,
              const double* x, double* y) {
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1 && nrows > 4096)
    for (int i = 0; i < nrows; ++i) {
        double s = 0.0;
        for (int p = ptr[i]; p < ptr[i + 1]; ++p) s += val[p] * x[idx[p]];
        y[i] = s;
    }
}

/* Ruiz equilibration, pdhg.cu:469-492: 10 sweeps, rows then columns,
 * dr *= sqrt(max), dc /= sqrt(max); non-finite / non-positive -> 1. */
void orc_ruiz(const orc_model* lp, double* dr, double* dc) {
    const int m = lp->m, n = lp->n;
    for (int i = 0; i < m; ++i) dr[i] = 1.0;
    for (int j = 0; j < n; ++j) dc[j] = 1.0;
    for (int sweep = 0; sweep < 10; ++sweep) {
        for (int i = 0; i < m; ++i) {
            double mx = 0.0;
            for (int p = lp->ap[i]; p < lp->ap[i + 1]; ++p)
                mx = dmax(mx, fabs(lp->ax[p]) * dc[lp->ai[p]] / dr[i]);
            if (!(mx > 0) || !isfinite(mx)) continue;
            dr[i] *= sqrt(mx);
        }
        for (int j = 0; j < n; ++j) {
            double mx = 0.0;
            for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
                mx = dmax(mx, fabs(lp->acx[p]) / dr[lp->ci[p]] * dc[j]);
            if (!(mx > 0) || !isfinite(mx)) continue;
            if (g_mut == 4) dc[j] *= sqrt(mx);
            else dc[j] /= sqrt(mx);
        }
    }
    for (int i = 0; i < m; ++i)
        if (!(dr[i] > 0) || !isfinite(dr[i])) dr[i] = 1.0;
    for (int j = 0; j < n; ++j)
        if (!(dc[j] > 0) || !isfinite(dc[j])) dc[j] = 1.0;
}

/* Pock-Chambolle diagonal preconditioning with alpha = 1 (Pock and
 * Chambolle, ICCV 2011, Lemma 2), applied after Ruiz as the reference's
 * algorithm sheet prescribes (docs/research/pdlp-algorithm-sheet.md:62,141)
 * but pdhg.cu does not implement (pdhg.cu:469-492 is Ruiz only): with the
 * current factors, r_i = sum_j |a_ij| dc_j / dr_i and c_j = sum_i |a_ij|
 * dc_j / dr_i, both from the same matrix, each summed in storage order;
 * then dr_i *= sqrt(r_i), dc_j /= sqrt(c_j) (a zero or non-finite sum
 * leaves the factor).  The scaled matrix then has spectral norm <= 1. */
void orc_pock_chambolle(const orc_model* lp, double* dr, double* dc) {
    const int m = lp->m, n = lp->n;
    double* r = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double* c = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int p = lp->ap[i]; p < lp->ap[i + 1]; ++p)
            s += fabs(lp->ax[p]) * dc[lp->ai[p]] / dr[i];
        r[i] = s;
    }
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
            s += fabs(lp->acx[p]) * dc[j] / dr[lp->ci[p]];
        c[j] = s;
    }
    for (int i = 0; i < m; ++i)
        if (r[i] > 0 && isfinite(r[i])) dr[i] *= g_mut == 7 ? r[i] : sqrt(r[i]);
    for (int j = 0; j < n; ++j)
        if (c[j] > 0 && isfinite(c[j])) dc[j] /= sqrt(c[j]);
    free(r);
    free(c);
}

/* Scaled data, pdhg.cu:499-513 (gamma = 1, pdhg.cu:498). */
void orc_scale(const orc_model* lp, const double* dr, const double* dc,
               double* sax, double* sacx, double* sc, double* scl,
               double* scu, double* srmin, double* srmax) {
    const double gamma = 1.0;
    for (int i = 0; i < lp->m; ++i) {
        srmin[i] = lp->rmin[i] / dr[i];
        srmax[i] = lp->rmax[i] / dr[i];
        for (int p = lp->ap[i]; p < lp->ap[i + 1]; ++p)
            sax[p] = lp->ax[p] * dc[lp->ai[p]] / dr[i];
    }
    for (int j = 0; j < lp->n; ++j) {
        sc[j] = lp->c[j] * dc[j] / gamma;
        scl[j] = lp->cl[j] / dc[j];
        scu[j] = lp->cu[j] / dc[j];
        for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
            sacx[p] = lp->acx[p] * dc[j] / dr[lp->ci[p]];
    }
}

static double nrm2(int k, const double* v) {
    double s = 0.0;
    for (int i = 0; i < k; ++i) s += v[i] * v[i];
    return sqrt(s);
}

/* Power iteration for ||A~||_2^2, pdhg.cu:605-631. */
double orc_power(int m, int n, const int* ap, const int* ai,
                 const double* sax, const int* cp, const int* ci,
                 const double* sacx) {
    double* v = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* w = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* t = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    for (int j = 0; j < n; ++j) v[j] = 1.0 / sqrt((double)n);
    for (int it = 0; it < 30; ++it) {
        orc_spmv(m, ap, ai, sax, v, t);
        orc_spmv(n, cp, ci, sacx, t, w);
        double nw = nrm2(n, w);
        if (nw < 1e-300) break;
        double inv = 1.0 / nw;
        for (int j = 0; j < n; ++j) v[j] = w[j] * inv;
    }
    orc_spmv(m, ap, ai, sax, v, t);
    double na = nrm2(m, t);
    double lam2 = na * na;
    if (!(lam2 > 0) || !isfinite(lam2)) lam2 = 1.0;
    free(v);
    free(w);
    free(t);
    return lam2;
}

/* Initial primal weight, pdhg.cu:647-663. */
double orc_omega0(int m, int n, const double* sc, const double* srmin,
                  const double* srmax) {
    double cn = 0.0, qn = 0.0, omega = 1.0;
    for (int j = 0; j < n; ++j) cn += sc[j] * sc[j];
    for (int i = 0; i < m; ++i) {
        double q = isfinite(srmax[i]) ? srmax[i]
                                      : (isfinite(srmin[i]) ? srmin[i] : 0.0);
        qn += q * q;
    }
    if (cn > 1e-24 && qn > 1e-24) omega = dclamp(sqrt(cn / qn), 0.3, 3.0);
    return omega;
}

/* Line-search dots with the reference's fixed partial layout,
 * pdhg.cu:63-125 (kern_dots + kern_dotfin): gD blocks x 256 threads,
 * grid-stride accumulation, then a halving tree per block; the gD partials
 * are then strided over 256 threads and tree-reduced again. */
void orc_dots(int n, int m, const double* dx, const double* dy,
              const double* adx, double out[3]) {
    const int B = 256;
    long mx = n > m ? n : m;
    int gD = (int)((mx + 255) / 256);
    if (gD > 4096) gD = 4096;
    if (gD < 1) gD = 1;
    double* part = (double*)malloc(sizeof(double) * 3 * (size_t)gD);
    double sh[3][256];
    for (int b = 0; b < gD; ++b) {
        for (int tid = 0; tid < B; ++tid) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            for (long i = (long)b * B + tid; i < n || i < m;
                 i += (long)gD * B) {
                if (i < n) a0 += dx[i] * dx[i];
                if (i < m) {
                    a1 += dy[i] * dy[i];
                    a2 += dy[i] * adx[i];
                }
            }
            sh[0][tid] = a0;
            sh[1][tid] = a1;
            sh[2][tid] = a2;
        }
        for (int s = B / 2; s > 0; s >>= 1)
            for (int tid = 0; tid < s; ++tid)
                for (int q = 0; q < 3; ++q) sh[q][tid] += sh[q][tid + s];
        for (int q = 0; q < 3; ++q) part[q * gD + b] = sh[q][0];
    }
    for (int tid = 0; tid < B; ++tid) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int i = tid; i < gD; i += B) {
            s0 += part[i];
            s1 += part[gD + i];
            s2 += part[2 * gD + i];
        }
        sh[0][tid] = s0;
        sh[1][tid] = s1;
        sh[2][tid] = s2;
    }
    for (int s = B / 2; s > 0; s >>= 1)
        for (int tid = 0; tid < s; ++tid)
            for (int q = 0; q < 3; ++q) sh[q][tid] += sh[q][tid + s];
    out[0] = sh[0][0];
    out[1] = sh[1][0];
    out[2] = sh[2][0];
    free(part);
}


/* Averaging weight of the hn-th accepted step of an epoch, pdhg.cu:795. */
double orc_avg_weight(long hn) {
    if (g_mut == 5) return 1.0 / ((double)hn + 1.0);
    return 2.0 / ((double)hn + 2.0);
}

/* kern_mix2's running average, pdhg.cu:127-141. */
static void mix_into(int len, double w, const double* v, double* avg) {
    for (int i = 0; i < len; ++i) avg[i] += w * (v[i] - avg[i]);
}

/* Averages `count` basis vectors e_0..e_{count-1} with the same recurrence
 * the solve uses; out[j] is then the weight of step j in the average. */
void orc_mix_weights(int count, double* out) {
    double* e = (double*)calloc((size_t)(count > 0 ? count : 1), sizeof(double));
    for (int j = 0; j < count; ++j) out[j] = 0.0;
    for (long h = 0; h < count; ++h) {
        for (int j = 0; j < count; ++j) e[j] = (j == h) ? 1.0 : 0.0;
        mix_into(count, orc_avg_weight(h), e, out);
    }
    free(e);
}

/* Line-search controller, pdhg.cu:770-793.  Returns 1 on accept; updates
 * eta and the retry counter exactly as the reference does. */
int orc_ls_decide(long k, double* eta, double omega, const double dots[3],
                  int* retries) {
    const double INF = INFINITY;
    const double num = omega * dots[0] + dots[1] / omega;
    const double cross = dots[2];
    const double etabar = (num > 0.0 && cross > 0.0) ? num / (2.0 * cross) : INF;
    const double kk = (double)k;
    const double eta_new = isfinite(etabar)
                               ? dmin((1.0 - pow(kk, -0.3)) * etabar,
                                      (1.0 + pow(kk, -0.6)) * *eta)
                               : INF;
    const int accept = !(*eta > etabar) || num <= 0.0 || cross <= 0.0 ||
                       ++*retries > 20;
    if (accept) {
        if (isfinite(eta_new) && eta_new > 1e-12) *eta = dmin(eta_new, 1e12);
        return 1;
    }
    *eta = dmax(eta_new, 1e-12);
    return 0;
}

/* KKT evaluation in original units, pdhg.cu:173-254. */
orc_metrics orc_evaluate(int m, int n, const double* rmin,
                         const double* rmax, const double* cl,
                         const double* cu, const double* c, const double* x,
                         const double* y, const double* Ax, const double* d) {
    const double INF = INFINITY;
    orc_metrics mt;
    double pv = 0.0, pa = 0.0, prow = 0.0;
    for (int i = 0; i < m; ++i) {
        pv = dmax(pv, dmax(rmin[i] - Ax[i], Ax[i] - rmax[i]));
        pa = dmax(pa, fabs(dmin(dmax(Ax[i], rmin[i]), rmax[i])));
        double viol = dmax(0.0, dmax(rmin[i] - Ax[i], Ax[i] - rmax[i]));
        double sc = fabs(Ax[i]);
        if (isfinite(rmin[i])) sc = dmax(sc, fabs(rmin[i]));
        if (isfinite(rmax[i])) sc = dmax(sc, fabs(rmax[i]));
        prow = dmax(prow, viol / (1.0 + sc));
    }
    mt.pinf = dmax(pv, 0.0) / (1.0 + pa);
    mt.prow = prow;
    double dv = 0.0;
    for (int j = 0; j < n; ++j) {
        double bnd = fabs(isfinite(cu[j]) ? cu[j] : cl[j]);
        double tol = 1e-9 * (1.0 + bnd);
        double v;
        if (g_mut == 3) v = fabs(d[j]);
        else if (x[j] <= cl[j] + tol && d[j] > 0) v = 0.0;
        else if (x[j] >= cu[j] - tol && d[j] < 0) v = 0.0;
        else v = fabs(d[j]);
        dv = dmax(dv, v);
    }
    double cmax = 0.0, ymax = 0.0;
    for (int j = 0; j < n; ++j) cmax = dmax(cmax, fabs(c[j]));
    for (int i = 0; i < m; ++i) ymax = dmax(ymax, fabs(y[i]));
    mt.dinf = dv / (1.0 + cmax);
    mt.op = 0.0;
    for (int j = 0; j < n; ++j) mt.op += c[j] * x[j];
    double od = 0.0;
    int bounded = 1;
    const double dz = 1e-8 * (1.0 + cmax);
    const double yz = 1e-8 * (1.0 + ymax);
    for (int j = 0; j < n; ++j) {
        if (d[j] >= 0) {
            if (isfinite(cl[j])) od += d[j] * cl[j];
            else if (d[j] > dz) { bounded = 0; break; }
        } else {
            if (isfinite(cu[j])) od += d[j] * cu[j];
            else if (-d[j] > dz) { bounded = 0; break; }
        }
    }
    if (bounded) {
        for (int i = 0; i < m; ++i) {
            if (y[i] > yz) {
                if (isfinite(rmax[i])) od -= y[i] * rmax[i];
                else { bounded = 0; break; }
            } else if (y[i] < -yz) {
                if (isfinite(rmin[i])) od -= y[i] * rmin[i];
                else { bounded = 0; break; }
            }
        }
    }
    mt.od = bounded ? od : -INF;
    mt.gap = bounded ? dmax(0.0, mt.op - od) / (1.0 + fabs(mt.op)) : INF;
    return mt;
}

static int finite_res(const orc_metrics* t) {
    return isfinite(t->pinf) && isfinite(t->dinf);
}
static double err_of(const orc_metrics* t) {
    double g = isfinite(t->gap) ? t->gap : 0.0;
    return dmax(t->pinf, dmax(t->dinf, g));
}

/* Dense Cholesky, pdhg.cu:255-271 (lower triangle, row-major). */
static int chol_factor(double* G, int nr) {
    for (int i = 0; i < nr; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = G[(size_t)i * nr + j];
            for (int t = 0; t < j; ++t)
                s -= G[(size_t)i * nr + t] * G[(size_t)j * nr + t];
            if (i == j) {
                if (!(s > 0)) return 0;
                G[(size_t)i * nr + i] = sqrt(s);
            } else {
                G[(size_t)i * nr + j] = s / G[(size_t)j * nr + j];
            }
        }
    }
    return 1;
}

/* pdhg.cu:273-286 */
static void chol_apply(const double* G, int nr, double* rhs) {
    for (int i = 0; i < nr; ++i) {
        double s = rhs[i];
        for (int t = 0; t < i; ++t) s -= G[(size_t)i * nr + t] * rhs[t];
        rhs[i] = s / G[(size_t)i * nr + i];
    }
    for (int i = nr - 1; i >= 0; --i) {
        double s = rhs[i];
        for (int t = i + 1; t < nr; ++t) s -= G[(size_t)t * nr + i] * rhs[t];
        rhs[i] = s / G[(size_t)i * nr + i];
    }
}

/* POCS dual repair, pdhg.cu:288-448.  The oracle declines m > 20000
 * (the reference's dense m*m int index overflows above 46,340). */
int orc_dual_repair(const orc_model* lp, const double* x, const double* Ax,
                    const double* y_it, double col_tol, double* y) {
    const int m = lp->m, n = lp->n;
    if (m > 20000) return 0;
    const double NL = -INFINITY, PL = INFINITY;
    double* lo = (double*)malloc(sizeof(double) * (size_t)m);
    double* hi = (double*)malloc(sizeof(double) * (size_t)m);
    for (int i = 0; i < m; ++i) {
        int fin_lo = isfinite(lp->rmin[i]), fin_hi = isfinite(lp->rmax[i]);
        if (fin_lo && fin_hi) { lo[i] = NL; hi[i] = PL; }
        else if (fin_hi) { lo[i] = 0.0; hi[i] = PL; }
        else if (fin_lo) { lo[i] = NL; hi[i] = 0.0; }
        else { lo[i] = hi[i] = 0.0; }
    }
    char* ct = (char*)malloc((size_t)n);
    double* cn2 = (double*)calloc((size_t)n, sizeof(double));
    int* ji = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int k = 0;
    for (int j = 0; j < n; ++j) {
        int up_inf = !isfinite(lp->cu[j]), lo_inf = !isfinite(lp->cl[j]);
        ct[j] = (up_inf && lo_inf) ? 2 : up_inf ? 0 : lo_inf ? 1 : 3;
        for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
            cn2[j] += lp->acx[p] * lp->acx[p];
        double bl = col_tol * (1.0 + (isfinite(lp->cl[j]) ? fabs(lp->cl[j]) : 0.0));
        double bu = col_tol * (1.0 + (isfinite(lp->cu[j]) ? fabs(lp->cu[j]) : 0.0));
        if (x[j] - lp->cl[j] > bl && lp->cu[j] - x[j] > bu) ji[k++] = j;
    }
    double* G = NULL;
    int ok_all = 1;
    if (k > 0) {
        G = (double*)calloc((size_t)m * m, sizeof(double));
        for (int jj = 0; jj < k; ++jj) {
            const int j = ji[jj];
            for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p) {
                const int i1 = lp->ci[p];
                const double v1 = lp->acx[p];
                for (int q = lp->cp[j]; q < lp->cp[j + 1]; ++q)
                    G[(size_t)i1 * m + lp->ci[q]] += v1 * lp->acx[q];
            }
        }
        double tr = 0.0;
        for (int i = 0; i < m; ++i) tr += G[(size_t)i * m + i];
        double ridge = tr > 0 ? 1e-10 * tr / m : 1e-14;
        int ok = 0;
        double* Gc = (double*)malloc(sizeof(double) * (size_t)m * m);
        for (int att = 0; att < 6 && !ok; ++att) {
            memcpy(Gc, G, sizeof(double) * (size_t)m * m);
            for (int i = 0; i < m; ++i) Gc[(size_t)i * m + i] += ridge;
            ok = chol_factor(Gc, m);
            if (ok) memcpy(G, Gc, sizeof(double) * (size_t)m * m);
            else ridge = ridge > 0 ? ridge * 100.0 : 1e-14;
        }
        free(Gc);
        if (!ok) ok_all = 0;
    }
    int have = 0;
    if (ok_all) {
        double best_gap = INFINITY, best_viol = INFINITY;
        int have_viol = 0;
        double* d = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        double* rhs = (double*)malloc(sizeof(double) * (size_t)m);
        double* yc = (double*)malloc(sizeof(double) * (size_t)m);
        double* ybest = (double*)malloc(sizeof(double) * (size_t)m);
        double* yviol = (double*)malloc(sizeof(double) * (size_t)m);
        for (int anch = 0; anch < 2; ++anch) {
            for (int i = 0; i < m; ++i) {
                yc[i] = anch == 0 ? y_it[i] : 0.0;
                yc[i] = dmin(dmax(yc[i], lo[i]), hi[i]);
            }
            double vref = -1.0;
            for (int it = 0; it < 300; ++it) {
                if (k > 0) {
                    for (int i = 0; i < m; ++i) rhs[i] = 0.0;
                    double res2 = 0.0;
                    for (int jj = 0; jj < k; ++jj) {
                        const int j = ji[jj];
                        double dj = lp->c[j];
                        for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
                            dj += lp->acx[p] * yc[lp->ci[p]];
                        res2 += dj * dj;
                        if (dj != 0.0)
                            for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
                                rhs[lp->ci[p]] -= lp->acx[p] * dj;
                    }
                    if (res2 > 0.0) {
                        chol_apply(G, m, rhs);
                        const double sg = g_mut == 6 ? -1.0 : 1.0;
                        for (int i = 0; i < m; ++i) yc[i] += sg * rhs[i];
                    }
                }
                for (int j = 0; j < n; ++j) {
                    if (ct[j] == 3 || cn2[j] == 0.0) continue;
                    double dj = lp->c[j];
                    for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
                        dj += lp->acx[p] * yc[lp->ci[p]];
                    double v = 0.0;
                    if (ct[j] == 0 && dj < 0.0) v = -dj;
                    else if (ct[j] == 1 && dj > 0.0) v = -dj;
                    else if (ct[j] == 2 && dj != 0.0) v = -dj;
                    if (v != 0.0) {
                        double st = v / cn2[j];
                        for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
                            yc[lp->ci[p]] += st * lp->acx[p];
                    }
                }
                for (int i = 0; i < m; ++i) yc[i] = dmin(dmax(yc[i], lo[i]), hi[i]);
                for (int j = 0; j < n; ++j) {
                    double dj = lp->c[j];
                    for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
                        dj += lp->acx[p] * yc[lp->ci[p]];
                    d[j] = dj;
                }
                if (it % 25 == 0) {
                    double viol = 0.0;
                    for (int j = 0; j < n; ++j) {
                        if (ct[j] == 3) continue;
                        double v = 0.0;
                        if (ct[j] == 0 && d[j] < 0) v = -d[j];
                        else if (ct[j] == 1 && d[j] > 0) v = d[j];
                        else if (ct[j] == 2) v = fabs(d[j]);
                        viol = dmax(viol, v);
                    }
                    for (int i = 0; i < m; ++i) {
                        double v = 0.0;
                        if (lo[i] == 0.0 && hi[i] == 0.0) v = fabs(yc[i]);
                        else if (hi[i] == 0.0 && yc[i] > 0) v = yc[i];
                        else if (lo[i] == 0.0 && yc[i] < 0) v = -yc[i];
                        viol = dmax(viol, v);
                    }
                    if (it == 25) vref = viol;
                    if (viol < best_viol) {
                        best_viol = viol;
                        memcpy(yviol, yc, sizeof(double) * (size_t)m);
                        have_viol = 1;
                    }
                    if (it == 100 && vref > 0 && viol > 0.25 * vref) break;
                }
                orc_metrics t = orc_evaluate(m, n, lp->rmin, lp->rmax, lp->cl,
                                             lp->cu, lp->c, x, yc, Ax, d);
                if (isfinite(t.gap) && t.gap < best_gap) {
                    best_gap = t.gap;
                    memcpy(ybest, yc, sizeof(double) * (size_t)m);
                    have = 1;
                }
            }
        }
        if (have) memcpy(y, ybest, sizeof(double) * (size_t)m);
        else if (have_viol) memcpy(y, yviol, sizeof(double) * (size_t)m);
        free(d); free(rhs); free(yc); free(ybest); free(yviol);
    }
    free(G); free(lo); free(hi); free(ct); free(cn2); free(ji);
    return have;
}

/* ---------------------------------------------------------------------- */
/* Full solve: pdhg.cu:457-1156 replayed on the host.                      */

typedef struct {
    const orc_model* lp;
    int m, n;
    const double *dr, *dc;
    const int *ap, *ai, *cp, *ci;
    const double *sax, *sacx;
    double *hx, *hy, *hax, *hd, *tm, *ata;
} orc_ctx;

/* metrics_of, pdhg.cu:694-733: unscale (x = dc x~, y = (gamma/dr) y~,
 * Ax = dr (A~x~), d = c + gamma (A~'y~)/dc) and evaluate. */
static orc_metrics metrics_of(orc_ctx* C, const double* sx, const double* sy) {
    const double gamma = 1.0;
    const int m = C->m, n = C->n;
    for (int j = 0; j < n; ++j) C->hx[j] = sx[j] * C->dc[j];
    for (int i = 0; i < m; ++i) C->hy[i] = sy[i] * (gamma / C->dr[i]);
    orc_spmv(m, C->ap, C->ai, C->sax, sx, C->tm);
    for (int i = 0; i < m; ++i) C->hax[i] = C->tm[i] * C->dr[i];
    orc_spmv(n, C->cp, C->ci, C->sacx, sy, C->ata);
    for (int j = 0; j < n; ++j)
        C->hd[j] = C->lp->c[j] + gamma * C->ata[j] / C->dc[j];
    return orc_evaluate(m, n, C->lp->rmin, C->lp->rmax, C->lp->cl,
                        C->lp->cu, C->lp->c, C->hx, C->hy, C->hax, C->hd);
}

#define ALLOC(T, k) ((T*)calloc((size_t)((k) > 0 ? (k) : 1), sizeof(T)))
#define COPY(dst, src, k) memcpy((dst), (src), sizeof(double) * (size_t)(k))

int orc_pdhg_solve(const orc_model* lp, const orc_options* opt,
                   double* x_out, double* y_out, orc_result* res,
                   double* trace_eta, long trace_cap) {
    const int m = lp->m, n = lp->n;
    const int nnz = lp->ap[m];
    const double INF = INFINITY;
    const double gamma = 1.0;
    memset(res, 0, sizeof(*res));
    res->status = 7;

    double *dr = ALLOC(double, m), *dc = ALLOC(double, n);
    orc_ruiz(lp, dr, dc);
    if (opt->scaling == 1) orc_pock_chambolle(lp, dr, dc);
    double *sax = ALLOC(double, nnz), *sacx = ALLOC(double, nnz);
    double *sc = ALLOC(double, n), *scl = ALLOC(double, n),
           *scu = ALLOC(double, n), *srmin = ALLOC(double, m),
           *srmax = ALLOC(double, m);
    orc_scale(lp, dr, dc, sax, sacx, sc, scl, scu, srmin, srmax);

    double *x = ALLOC(double, n), *xnew = ALLOC(double, n),
           *xbar = ALLOC(double, n), *xavg = ALLOC(double, n),
           *dx = ALLOC(double, n);
    double *y = ALLOC(double, m), *ynew = ALLOC(double, m),
           *yavg = ALLOC(double, m), *dy = ALLOC(double, m),
           *adx = ALLOC(double, m), *axcur = ALLOC(double, m);
    double *ata = ALLOC(double, n), *tm = ALLOC(double, m);
    double *xb = ALLOC(double, n), *yb = ALLOC(double, m);
    double *pesx = ALLOC(double, n), *pesy = ALLOC(double, m);
    orc_ctx C = {lp, m, n, dr, dc, lp->ap, lp->ai, lp->cp, lp->ci, sax, sacx,
                 ALLOC(double, n), ALLOC(double, m), ALLOC(double, m),
                 ALLOC(double, n), ALLOC(double, m), ALLOC(double, n)};

    double lam2 = orc_power(m, n, lp->ap, lp->ai, sax, lp->cp, lp->ci, sacx);
    const double eta0 = 1.2 / sqrt(lam2);
    double eta = eta0;
    double omega = orc_omega0(m, n, sc, srmin, srmax);
    res->lam2 = lam2;
    res->eta0 = eta0;
    res->omega0 = omega;

    const int CHECK = 64;
    long MAXIT = (opt->time_limit_s > 0 && opt->time_limit_s < 1e6)
                     ? 20000000L : opt->max_iterations;
    const double TOL = opt->tolerance;
    long hn = 0;
    double epoch_kkt = -1.0, prev_cand = INF;
    long epoch_start = 0;
    int have_pe = 0;
    long iters_run = 0;
    double t0 = now_s();

    orc_metrics last;
    memset(&last, 0, sizeof(last));
    int last_is_avg = 0;
    orc_metrics inc;
    memset(&inc, 0, sizeof(inc));
    int have_inc = 0;
    double inc_op = 0.0;
    long last_rollback = 0;
    long next_repair = 5000, repair_spacing = 2000;
    int repaired = 0;
    int time_limit = 0;
    long attempts = 0, restarts = 0;
    double hdots[3];

    for (long k = 1; k <= MAXIT; ++k) {
        iters_run = k;
        orc_spmv(n, lp->cp, lp->ci, sacx, y, ata);
        int retries = 0;
        for (;;) {
            const double tau = eta / omega, sigma = eta * omega;
            ++attempts;
            /* kern_primal, pdhg.cu:32-45 */
            for (int j = 0; j < n; ++j) {
                double g = sc[j] + ata[j];
                double step = g_mut == 1 ? x[j] + tau * g : x[j] - tau * g;
                double xn = cfmin(cfmax(step, scl[j]), scu[j]);
                xnew[j] = xn;
                xbar[j] = 2.0 * xn - x[j];
                dx[j] = xn - x[j];
            }
            orc_spmv(m, lp->ap, lp->ai, sax, xbar, tm);
            /* kern_dual, pdhg.cu:47-61 */
            for (int i = 0; i < m; ++i) {
                double t = y[i] + sigma * tm[i];
                double pr = cfmin(cfmax(t / sigma, srmin[i]), srmax[i]);
                double yn = g_mut == 2 ? t + sigma * pr : t - sigma * pr;
                ynew[i] = yn;
                dy[i] = yn - y[i];
                adx[i] = 0.5 * (tm[i] - axcur[i]);
            }
            orc_dots(n, m, dx, dy, adx, hdots);
            if (orc_ls_decide(k, &eta, omega, hdots, &retries)) {
                double* t1 = x; x = xnew; xnew = t1;
                double* t2 = y; y = ynew; ynew = t2;
                break;
            }
        }
        if (trace_eta && k - 1 < trace_cap) trace_eta[k - 1] = eta;
        /* kern_mix2, pdhg.cu:127-141,795-799 */
        double weta = orc_avg_weight(hn);
        ++hn;
        mix_into(n, weta, x, xavg);
        mix_into(m, weta, y, yavg);
        for (int i = 0; i < m; ++i) axcur[i] += adx[i];

        if (k % CHECK != 0) continue;
        /* check, pdhg.cu:801-1084 */
        orc_metrics mc = metrics_of(&C, x, y);
        orc_metrics ma = metrics_of(&C, xavg, yavg);
        double ec = finite_res(&mc) ? err_of(&mc) : INF;
        double ea = finite_res(&ma) ? err_of(&ma) : INF;
        last = (ec <= ea) ? mc : ma;
        last_is_avg = (ea < ec);
        if (opt->verbosity > 0 && k % (CHECK * 40) == 0)
            fprintf(stderr, "[pdhg] k=%ld pinf=%.3e dinf=%.3e gap=%.3e\n", k,
                    last.pinf, last.dinf, last.gap);
#define PASSES(t) ((t).pinf <= TOL && (t).dinf <= TOL && (t).prow <= TOL && \
                   isfinite((t).gap) && (t).gap <= TOL)
        if (PASSES(mc)) { last = mc; last_is_avg = 0; break; }
        if (PASSES(ma)) { last = ma; last_is_avg = 1; break; }

        /* dual repair schedule, pdhg.cu:854-931 */
        if (k >= 5000 && k >= next_repair) {
            next_repair = k + repair_spacing;
            repair_spacing = repair_spacing * 3 < 30000 ? repair_spacing * 3
                                                        : 30000;
            for (int cc = 1; cc <= 2; ++cc) {
                const orc_metrics* t = cc == 1 ? &ma : &inc;
                if (cc == 2 && !have_inc) break;
                if (!(t->pinf <= TOL && t->prow <= TOL &&
                      t->dinf <= 10.0 * TOL &&
                      !(isfinite(t->gap) && t->gap <= TOL)))
                    continue;
                double* rx = ALLOC(double, n);
                double* ry_it = ALLOC(double, m);
                double* ry = ALLOC(double, m);
                double* rAx = ALLOC(double, m);
                double* rd = ALLOC(double, n);
                const double* dxs = cc == 1 ? xavg : xb;
                double* dys = cc == 1 ? yavg : y;
                for (int j = 0; j < n; ++j) rx[j] = dxs[j] * dc[j];
                for (int i = 0; i < m; ++i) ry_it[i] = gamma * dys[i] / dr[i];
                for (int j = 0; j < n; ++j)
                    for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
                        rAx[lp->ci[p]] += lp->acx[p] * rx[j];
                static const double ctol[2] = {1e-6, 1e-1};
                int ok_rep = 0;
                orc_metrics mr;
                memset(&mr, 0, sizeof(mr));
                for (int tier = 0; tier < 2 && !ok_rep; ++tier) {
                    if (!orc_dual_repair(lp, rx, rAx, ry_it, ctol[tier], ry))
                        continue;
                    for (int j = 0; j < n; ++j) {
                        rd[j] = lp->c[j];
                        for (int p = lp->cp[j]; p < lp->cp[j + 1]; ++p)
                            rd[j] += lp->acx[p] * ry[lp->ci[p]];
                    }
                    orc_metrics t2 = orc_evaluate(m, n, lp->rmin, lp->rmax,
                                                  lp->cl, lp->cu, lp->c, rx,
                                                  ry, rAx, rd);
                    if (t2.pinf <= TOL && t2.prow <= TOL && isfinite(t2.gap) &&
                        t2.gap <= TOL) {
                        ok_rep = 1;
                        mr = t2;
                    }
                }
                if (ok_rep) {
                    for (int i = 0; i < m; ++i) dys[i] = ry[i] * dr[i] / gamma;
                    if (cc == 2) COPY(x, xb, n);
                    last = mr;
                    last_is_avg = (cc == 1);
                    repaired = 1;
                }
                free(rx); free(ry_it); free(ry); free(rAx); free(rd);
                if (ok_rep) break;
            }
            if (repaired) break;
        }

        /* incumbent, pdhg.cu:933-958 */
        const double* inc_x_src = NULL;
        const double* inc_y_src = NULL;
        for (int cc = 0; cc < 2; ++cc) {
            const orc_metrics* t = cc ? &ma : &mc;
            if (t->pinf <= 1e-4 && isfinite(t->op) &&
                (!have_inc || t->op < inc_op)) {
                have_inc = 1;
                inc_op = t->op;
                inc = *t;
                inc_x_src = cc ? xavg : x;
                inc_y_src = cc ? yavg : y;
            }
        }
        if (inc_x_src) {
            COPY(xb, inc_x_src, n);
            COPY(yb, inc_y_src, m);
        }
        /* rollback, pdhg.cu:965-988 */
        if (mc.pinf >= 0.5 && have_inc && k - last_rollback > 2000) {
            COPY(x, xb, n);
            COPY(y, yb, m);
            COPY(xavg, xb, n);
            COPY(yavg, yb, m);
            eta = dmin(eta, 0.5 * eta0);
            hn = 0;
            epoch_start = k;
            epoch_kkt = -1.0;
            prev_cand = INF;
            orc_spmv(m, lp->ap, lp->ai, sax, x, tm);
            COPY(axcur, tm, m);
            last_rollback = k;
            ++res->rollbacks;
        }
        /* restart + primal weight, pdhg.cu:989-1075 */
        double cand = dmin(ec, ea);
        if (epoch_kkt < 0.0) {
            epoch_kkt = cand;
            prev_cand = cand;
        }
        const long t_epoch = k - epoch_start;
        int restart_now =
            (isfinite(cand) && isfinite(epoch_kkt) && epoch_kkt > 0.0 &&
             (cand <= 0.2 * epoch_kkt ||
              (cand <= 0.8 * epoch_kkt && isfinite(prev_cand) &&
               cand > prev_cand))) ||
            t_epoch >= 0.36 * (double)k;
        prev_cand = cand;
        if (restart_now) {
            ++restarts;
            if (ea < ec) {
                COPY(x, xavg, n);
                COPY(y, yavg, m);
            }
            COPY(xavg, x, n);
            COPY(yavg, y, m);
            hn = 0;
            epoch_start = k;
            epoch_kkt = cand;
            prev_cand = INF;
            orc_spmv(m, lp->ap, lp->ai, sax, x, tm);
            COPY(axcur, tm, m);
            if (have_pe) {
                double dxn = 0.0, dyn = 0.0;
                for (int j = 0; j < n; ++j) {
                    double d = x[j] - pesx[j];
                    dxn += d * d;
                }
                for (int i = 0; i < m; ++i) {
                    double d = y[i] - pesy[i];
                    dyn += d * d;
                }
                if (dxn > 1e-24 && dyn > 1e-24) {
                    double om = sqrt(omega * sqrt(dyn / dxn));
                    if (isfinite(om) && om > 0.0)
                        omega = dclamp(om, 1e-6, 1e6);
                }
            }
            COPY(pesx, x, n);
            COPY(pesy, y, m);
            have_pe = 1;
        }
        if (now_s() - t0 > opt->time_limit_s) {
            res->status = 6;
            time_limit = 1;
            break;
        }
    }
    double secs = now_s() - t0;
#define CERTIFIED(t) ((t).pinf <= TOL && (t).prow <= TOL && \
                      isfinite((t).gap) && (t).gap <= TOL)
    if (!time_limit) res->status = CERTIFIED(last) ? 1 : 5;
    const double* fx = last_is_avg ? xavg : x;
    const double* fy = last_is_avg ? yavg : y;
    res->used_incumbent = 0;
    if (!CERTIFIED(last) && have_inc && inc.pinf <= TOL && isfinite(inc.op) &&
        (inc.op < last.op || last.pinf > TOL)) {
        last = inc;
        last_is_avg = 0;
        fx = xb;
        fy = yb;
        res->used_incumbent = 1;
    }
    res->objective = last.op + lp->obj_const;
    res->pinf = last.pinf;
    res->dinf = last.dinf;
    res->rel_gap = last.gap;
    res->iterations = iters_run;
    res->solve_time_ms = secs * 1000.0;
    res->attempts = attempts;
    res->restarts = restarts;
    res->eta = eta;
    res->omega = omega;
    res->last_is_avg = last_is_avg;
    res->repaired = repaired;
    for (int j = 0; j < n; ++j) x_out[j] = fx[j] * dc[j];
    for (int i = 0; i < m; ++i) y_out[i] = gamma * fy[i] / dr[i];

    free(dr); free(dc); free(sax); free(sacx); free(sc); free(scl);
    free(scu); free(srmin); free(srmax); free(x); free(xnew); free(xbar);
    free(xavg); free(dx); free(y); free(ynew); free(yavg); free(dy);
    free(adx); free(axcur); free(ata); free(tm); free(xb); free(yb);
    free(pesx); free(pesy); free(C.hx); free(C.hy); free(C.hax); free(C.hd);
    free(C.tm); free(C.ata);
    return res->status;
}
