// Batched active-set kernels (see adjoint.hpp): polishing of PDHG
// solutions on their active set, and the adjoint of the solution map.
//
// One 256-thread workgroup per instance.  Shared pieces:
//   * masks(): active sets from complementarity in scaled units -- column
//     j is in B when its distance to the nearer bound exceeds |d~_j|
//     (d~ = c~ + A~'y~); an inequality row is in R when |y~_i| exceeds its
//     distance to the nearer row bound; equality rows are always in R;
//   * cgls_pair(): CGLS for  A~_RB z1 = r1  and  A~_RB' z2 = r2  iterated
//     together (four masked instance SpMVs per iteration), each stopped
//     when its normal-equation residual falls below tol relative to its
//     start, or at maxit.
// k_polish solves the active-set equations for the corrections of x_B
// and y_R (nonbasic columns snapped to their bounds, inactive rows' duals
// to zero) and keeps the polished pair only if it is primal and dual
// feasible; k_adjoint solves the transposed system and forms the
// gradients.  All reductions use the workgroup tree in a fixed order, so
// results are run-to-run deterministic.

#include "adjoint.hpp"
#include "batch_dev.hpp"

namespace igaos::batch {

namespace {

struct GatherMul {  // v[j] * s[j]
    const double* v;
    const double* s;
    __device__ __forceinline__ double operator()(int j) const { return v[j] * s[j]; }
};
struct GatherDiv {  // v[j] / s[j]
    const double* v;
    const double* s;
    __device__ __forceinline__ double operator()(int j) const { return v[j] / s[j]; }
};

struct Work {
    double *z1, *s1, *p1, *r1, *t1;
    double *z2, *s2, *p2, *r2, *t2;
};

__device__ inline Work work_of(const AdjArrays& J, long co, long ro) {
    return Work{J.z1 + co, J.s1 + co, J.p1 + co, J.r1 + ro, J.t1 + ro,
                J.z2 + ro, J.s2 + ro, J.p2 + ro, J.r2 + co, J.t2 + co};
}

// Active sets of the pair (x, y) in original units; returns |B| and |R|.
// A column's score is its distance to the nearer bound over |d~_j| (both
// scaled); B holds the columns scoring above 1.  At a nondegenerate vertex
// |B| = |R|; when an approximate solution leaves them unequal, the
// lowest-scoring columns of B are moved out (or the highest-scoring ones
// outside moved in), at most 64, one workgroup argmin/argmax at a time.
// `score` is n doubles of scratch.
template <typename V>
__device__ void masks(const LP<V>& L, const double* x, const double* y,
                      unsigned char* mc, unsigned char* mr, double* score,
                      Lds& S, double& nb, double& nr) {
    const double* dr = L.dr();
    const double* dc = L.dc();
    const double INF = __builtin_huge_val();
    nr = 0.0;
    nb = 0.0;
    spmv_A<double>(L, L.sax(), GatherDiv{x, dc}, S, [&](int i, double a) {
        bool in = true;
        if (L.rmin()[i] != L.rmax()[i]) {
            const double sl = fmin(a - L.srmin()[i], L.srmax()[i] - a);
            in = fabs(y[i] * dr[i]) > sl;
        }
        mr[i] = in;
        nr += in;
    });
    __syncthreads();
    spmv_T<double>(L, L.sacx(), GatherMul{y, dr}, S, [&](int j, double a) {
        const double d = fabs(L.sc()[j] + a);
        const double xt = x[j] / dc[j];
        const double sl = fmin(xt - L.scl()[j], L.scu()[j] - xt);
        const double sc = d > 0.0 ? sl / d : (sl > 0.0 ? INF : 0.0);
        score[j] = sc;
        const bool in = sc > 1.0;
        mc[j] = in;
        nb += in;
    });
    nr = bsum(nr, S.red);
    nb = bsum(nb, S.red);
    for (int k = 0; k < 64 && nb != nr; ++k) {
        const bool drop = nb > nr;  // uniform
        // best candidate: lowest score in B (drop) or highest outside B
        double best = -INF;
        for (int j = threadIdx.x; j < L.n; j += NT)
            if ((mc[j] != 0) == drop) best = fmax(best, drop ? -score[j] : score[j]);
        best = bmax(best, S.red);
        if (!(best > -INF)) break;
        double idx = -INF;  // smallest index attaining it (as -index)
        for (int j = threadIdx.x; j < L.n; j += NT)
            if ((mc[j] != 0) == drop && (drop ? -score[j] : score[j]) == best)
                idx = fmax(idx, -static_cast<double>(j));
        idx = bmax(idx, S.red);
        if (threadIdx.x == 0) mc[static_cast<int>(-idx)] = drop ? 0 : 1;
        nb += drop ? -1.0 : 1.0;
        __syncthreads();
    }
}

// CGLS on both masked systems; w.r1 (m) and w.r2 (n) hold the right-hand
// sides (zero off R and B); w.z1 (n) and w.z2 (m) receive the solutions.
// Returns the final relative residuals through e1, e2.
template <typename V>
__device__ void cgls_pair(const LP<V>& L, const unsigned char* mc,
                          const unsigned char* mr, const Work& w, double tol,
                          int maxit, Lds& S, int& it1, int& it2, double& e1,
                          double& e2) {
    const int m = L.m, n = L.n, tid = threadIdx.x;
    double nf = 0.0, nh = 0.0;
    for (int i = tid; i < m; i += NT) {
        nf += w.r1[i] * w.r1[i];
        w.z2[i] = 0.0;
    }
    for (int j = tid; j < n; j += NT) {
        nh += w.r2[j] * w.r2[j];
        w.z1[j] = 0.0;
    }
    nf = bsum(nf, S.red);
    nh = bsum(nh, S.red);
    double g1 = 0.0, g2 = 0.0;
    spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{w.r1}, S, [&](int j, double a) {
        const double v = mc[j] ? a : 0.0;
        w.s1[j] = v;
        w.p1[j] = v;
        g1 += v * v;
    });
    spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{w.r2}, S, [&](int i, double a) {
        const double v = mr[i] ? a : 0.0;
        w.s2[i] = v;
        w.p2[i] = v;
        g2 += v * v;
    });
    g1 = bsum(g1, S.red);
    g2 = bsum(g2, S.red);
    const double tol2 = tol * tol;
    const double g10 = g1, g20 = g2;
    bool done1 = !(g1 > 0.0), done2 = !(g2 > 0.0);
    it1 = it2 = 0;
    for (int it = 0; it < maxit && !(done1 && done2); ++it) {
        double d1 = 0.0, d2 = 0.0;
        if (!done1)
            spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{w.p1}, S, [&](int i, double a) {
                const double v = mr[i] ? a : 0.0;
                w.t1[i] = v;
                d1 += v * v;
            });
        if (!done2)
            spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{w.p2}, S, [&](int j, double a) {
                const double v = mc[j] ? a : 0.0;
                w.t2[j] = v;
                d2 += v * v;
            });
        d1 = bsum(d1, S.red);
        d2 = bsum(d2, S.red);
        if (!done1 && !(d1 > 0.0)) done1 = true;
        if (!done2 && !(d2 > 0.0)) done2 = true;
        const double a1 = done1 ? 0.0 : g1 / d1;
        const double a2 = done2 ? 0.0 : g2 / d2;
        if (!done1) {
            for (int j = tid; j < n; j += NT) w.z1[j] += a1 * w.p1[j];
            for (int i = tid; i < m; i += NT) w.r1[i] -= a1 * w.t1[i];
        }
        if (!done2) {
            for (int i = tid; i < m; i += NT) w.z2[i] += a2 * w.p2[i];
            for (int j = tid; j < n; j += NT) w.r2[j] -= a2 * w.t2[j];
        }
        __syncthreads();
        double h1 = 0.0, h2 = 0.0;
        if (!done1)
            spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{w.r1}, S, [&](int j, double a) {
                const double v = mc[j] ? a : 0.0;
                w.s1[j] = v;
                h1 += v * v;
            });
        if (!done2)
            spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{w.r2}, S, [&](int i, double a) {
                const double v = mr[i] ? a : 0.0;
                w.s2[i] = v;
                h2 += v * v;
            });
        h1 = bsum(h1, S.red);
        h2 = bsum(h2, S.red);
        if (!done1) {
            const double b1 = h1 / g1;
            for (int j = tid; j < n; j += NT) w.p1[j] = w.s1[j] + b1 * w.p1[j];
            g1 = h1;
            ++it1;
            if (g1 <= tol2 * g10) done1 = true;
        }
        if (!done2) {
            const double b2 = h2 / g2;
            for (int i = tid; i < m; i += NT) w.p2[i] = w.s2[i] + b2 * w.p2[i];
            g2 = h2;
            ++it2;
            if (g2 <= tol2 * g20) done2 = true;
        }
        __syncthreads();
    }
    // residuals r1 = f - A~_RB z1 and r2 = h - A~_RB' z2 as the recursions
    // carry them (the polish verifies its result independently)
    double q1 = 0.0, q2 = 0.0;
    for (int i = tid; i < m; i += NT) q1 += w.r1[i] * w.r1[i];
    for (int j = tid; j < n; j += NT) q2 += w.r2[j] * w.r2[j];
    q1 = bsum(q1, S.red);
    q2 = bsum(q2, S.red);
    e1 = nf > 0.0 ? sqrt(q1 / nf) : sqrt(q1);
    e2 = nh > 0.0 ? sqrt(q2 / nh) : sqrt(q2);
}

template <typename V>
__global__ __launch_bounds__(NT) void k_polish(const Arrays<V>* A,
                                               const LPDesc* lps,
                                               const AdjArrays* Jp) {
    __shared__ Lds S;
    const AdjArrays& J = *Jp;
    const LPDesc D = lps[blockIdx.x];
    const LP<V> L = view(A, D);
    const int m = L.m, n = L.n, tid = threadIdx.x;
    const long co = D.coff, ro = D.roff;
    unsigned char* mc = J.mc + co;
    unsigned char* mr = J.mr + ro;
    double* x = L.xout();
    double* y = L.yout();
    const double* dr = L.dr();
    const double* dc = L.dc();
    const Work w = work_of(J, co, ro);
    double* xs = J.xs + co;  // scaled candidate x~
    double* ys = J.ys + ro;  // scaled candidate y~
    double nb, nr;
    masks(L, x, y, mc, mr, xs, S, nb, nr);
    __syncthreads();
    // candidate start: nonbasic columns on their nearer bound, inactive
    // rows' duals zero
    for (int j = tid; j < n; j += NT) {
        const double xt = x[j] / dc[j];
        double v = xt;
        if (!mc[j]) v = (xt - L.scl()[j] <= L.scu()[j] - xt) ? L.scl()[j] : L.scu()[j];
        xs[j] = v;
    }
    for (int i = tid; i < m; i += NT) ys[i] = mr[i] ? y[i] * dr[i] : 0.0;
    __syncthreads();
    // right-hand sides: r1 = b~_R - A~ x~ on R, r2 = -(c~ + A~'y~) on B
    spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{xs}, S, [&](int i, double a) {
        const double b = y[i] >= 0.0 ? L.srmax()[i] : L.srmin()[i];
        w.r1[i] = mr[i] ? b - a : 0.0;
    });
    spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{ys}, S, [&](int j, double a) {
        w.r2[j] = mc[j] ? -(L.sc()[j] + a) : 0.0;
    });
    __syncthreads();
    int it1, it2;
    double e1, e2;
    cgls_pair(L, mc, mr, w, J.tol, J.maxit, S, it1, it2, e1, e2);
    for (int j = tid; j < n; j += NT)
        if (mc[j]) xs[j] += w.z1[j];
    for (int i = tid; i < m; i += NT)
        if (mr[i]) ys[i] += w.z2[i];
    __syncthreads();
    // feasibility of the polished pair: largest relative violation of each
    // condition (column bounds, rows, active-row dual signs, d_B = 0,
    // reduced-cost signs on N)
    double cmax = 0.0;
    for (int j = tid; j < n; j += NT) cmax = fmax(cmax, fabs(L.c()[j]));
    cmax = bmax(cmax, S.red);
    double vb = 0.0, vr = 0.0, vs = 0.0, vd = 0.0, vn = 0.0;
    for (int j = tid; j < n; j += NT) {
        const double xv = xs[j] * dc[j];
        const double lo = L.cl()[j], hi = L.cu()[j];
        vb = fmax(vb, fmax(lo - xv, 0.0) / (1.0 + fabs(lo)));
        vb = fmax(vb, fmax(xv - hi, 0.0) / (1.0 + fabs(hi)));
    }
    spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{xs}, S, [&](int i, double a) {
        const double ax = a * dr[i];
        const double lo = L.rmin()[i], hi = L.rmax()[i];
        vr = fmax(vr, fmax(lo - ax, 0.0) / (1.0 + fabs(lo)));
        vr = fmax(vr, fmax(ax - hi, 0.0) / (1.0 + fabs(hi)));
        if (mr[i] && lo != hi) {  // an active inequality row's dual sign
            const double yv = ys[i] / dr[i];
            vs = fmax(vs, y[i] >= 0.0 ? fmax(-yv, 0.0) : fmax(yv, 0.0));
        }
    });
    spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{ys}, S, [&](int j, double a) {
        const double d = (L.c()[j] + a / dc[j]) / (1.0 + cmax);
        if (mc[j]) {
            vd = fmax(vd, fabs(d));  // an interior column has d_j = 0
        } else {
            const bool lower = xs[j] == L.scl()[j];
            vn = fmax(vn, lower ? fmax(-d, 0.0) : fmax(d, 0.0));
        }
    });
    vb = bmax(vb, S.red);
    vr = bmax(vr, S.red);
    vs = bmax(vs, S.red);
    vd = bmax(vd, S.red);
    vn = bmax(vn, S.red);
    const double TOLP = 1e-9;
    const int flags = (vb > TOLP) | (vr > TOLP) << 1 | (vs > TOLP) << 2 |
                      (vd > TOLP) << 3 | (vn > TOLP) << 4;
    const double viol = fmax(fmax(vb, vr), fmax(vs, fmax(vd, vn)));
    const bool ok = flags == 0;
    if (ok) {
        double op = 0.0;
        for (int j = tid; j < n; j += NT) {
            x[j] = xs[j] * dc[j];
            op += L.c()[j] * x[j];
        }
        for (int i = tid; i < m; i += NT) y[i] = ys[i] / dr[i];
        op = bsum(op, S.red);
        if (tid == 0) J.obj[blockIdx.x] = op + D.obj_const;
    }
    if (tid == 0) {
        AdjInfo inf{};
        inf.basic_cols = static_cast<int>(nb);
        inf.active_rows = static_cast<int>(nr);
        inf.iters_p = it1;
        inf.iters_q = it2;
        inf.accepted = ok;
        inf.flags = flags;
        inf.res_p = e1;
        inf.res_q = e2;
        inf.viol = viol;
        J.info[blockIdx.x] = inf;
    }
}

template <typename V>
__global__ __launch_bounds__(NT) void k_adjoint(const Arrays<V>* A,
                                                const LPDesc* lps,
                                                const AdjArrays* Jp) {
    __shared__ Lds S;
    const AdjArrays& J = *Jp;
    const LPDesc D = lps[blockIdx.x];
    const LP<V> L = view(A, D);
    const int m = L.m, n = L.n, tid = threadIdx.x;
    const long co = D.coff, ro = D.roff;
    unsigned char* mc = J.mc + co;
    unsigned char* mr = J.mr + ro;
    const double* x = L.xout();
    const double* y = L.yout();
    const double* dr = L.dr();
    const double* dc = L.dc();
    const double* c = L.c();
    const double* gx = J.gx + co;
    const double* gy = J.gy + ro;
    const double gobj = J.gobj[blockIdx.x];
    const Work w = work_of(J, co, ro);

    double nb, nr;
    masks(L, x, y, mc, mr, J.xs + co, S, nb, nr);
    for (int i = tid; i < m; i += NT) w.r1[i] = mr[i] ? gy[i] / dr[i] : 0.0;
    for (int j = tid; j < n; j += NT)
        w.r2[j] = mc[j] ? dc[j] * (gx[j] + gobj * c[j]) : 0.0;
    __syncthreads();
    int it1, it2;
    double e1, e2;
    cgls_pair(L, mc, mr, w, J.tol, J.maxit, S, it1, it2, e1, e2);
    const double* z1 = w.z1;  // p~ on B
    const double* z2 = w.z2;  // q~ on R

    // gradients of c and of the active column bounds: (A~'q~)_j / dc_j
    spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{z2}, S, [&](int j, double a) {
        const double wj = gx[j] + gobj * c[j];
        if (mc[j]) {
            J.gcl[co + j] = 0.0;
            J.gcu[co + j] = 0.0;
        } else {
            const double g = wj - a / dc[j];
            const double xt = x[j] / dc[j];
            const bool lower = (xt - L.scl()[j]) <= (L.scu()[j] - xt);
            J.gcl[co + j] = lower ? g : 0.0;
            J.gcu[co + j] = lower ? 0.0 : g;
        }
        J.gc[co + j] = gobj * x[j] - (mc[j] ? dc[j] * z1[j] : 0.0);
    });
    // row-bound gradients: q_i goes to the bound the row sits on, which the
    // dual's sign names (y_i > 0: rmax binding, y_i < 0: rmin binding, as
    // in the dual objective -sum y_i * (y_i > 0 ? rmax_i : rmin_i))
    for (int i = tid; i < m; i += NT) {
        const double q = mr[i] ? z2[i] / dr[i] : 0.0;
        const bool upper = y[i] >= 0.0;
        J.grmin[ro + i] = upper ? 0.0 : q;
        J.grmax[ro + i] = upper ? q : 0.0;
    }
    // A's values, CSR order: -[i in R] ([j in B] p_j y_i + q_i x_j)
    const int* ap = L.ap();
    const int* ai = L.ai();
    double* gax = J.gax + D.zoff;
    for (int i = tid; i < m; i += NT) {
        const double q = mr[i] ? z2[i] / dr[i] : 0.0;
        const double yi = mr[i] ? y[i] : 0.0;
        for (int p = ap[i]; p < ap[i + 1]; ++p) {
            const int j = ai[p];
            const double pj = mc[j] ? dc[j] * z1[j] : 0.0;
            gax[p] = -(pj * yi + q * x[j]);
        }
    }
    if (tid == 0) {
        AdjInfo inf{};
        inf.basic_cols = static_cast<int>(nb);
        inf.active_rows = static_cast<int>(nr);
        inf.iters_p = it1;
        inf.iters_q = it2;
        inf.accepted = 1;
        inf.res_p = e1;
        inf.res_q = e2;
        J.info[blockIdx.x] = inf;
    }
}

}  // namespace

void launch_adjoint(const Arrays<float>* a, const LPDesc* lps, int count,
                    const AdjArrays* adj, hipStream_t s) {
    k_adjoint<float><<<count, NT, 0, s>>>(a, lps, adj);
}
void launch_adjoint(const Arrays<double>* a, const LPDesc* lps, int count,
                    const AdjArrays* adj, hipStream_t s) {
    k_adjoint<double><<<count, NT, 0, s>>>(a, lps, adj);
}
void launch_polish(const Arrays<float>* a, const LPDesc* lps, int count,
                   const AdjArrays* adj, hipStream_t s) {
    k_polish<float><<<count, NT, 0, s>>>(a, lps, adj);
}
void launch_polish(const Arrays<double>* a, const LPDesc* lps, int count,
                   const AdjArrays* adj, hipStream_t s) {
    k_polish<double><<<count, NT, 0, s>>>(a, lps, adj);
}

}  // namespace igaos::batch
