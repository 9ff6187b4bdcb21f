// Batched adjoint kernel (see adjoint.hpp for the system it solves).
//
// One 256-thread workgroup per instance.  Phases:
//   1. active-set masks from complementarity in scaled units: column j is
//      in B when its distance to the nearer bound exceeds |d~_j|
//      (d~ = c~ + A~'y~); an inequality row is in R when |y~_i| exceeds its
//      distance to the nearer bound; equality rows are always in R;
//   2. CGLS for  A~_RB p~ = D_r^-1 gy  and  A~_RB' q~ = D_c w, iterated
//      together (four masked instance SpMVs per iteration), each stopped
//      when its normal-equation residual falls below tol relative to its
//      start, or at maxit;
//   3. true residuals, then the gradients of c, the active row bounds, the
//      active column bounds and A's values (CSR order) in original units.
// All reductions use the workgroup tree in a fixed order, so results are
// run-to-run deterministic.

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
    double *z1 = J.z1 + co, *s1 = J.s1 + co, *p1 = J.p1 + co;
    double *r1 = J.r1 + ro, *t1 = J.t1 + ro;
    double *z2 = J.z2 + ro, *s2 = J.s2 + ro, *p2 = J.p2 + ro;
    double *r2 = J.r2 + co, *t2 = J.t2 + co;

    // ---- 1. masks
    double nr = 0.0, nb = 0.0;
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
        const double d = L.sc()[j] + a;
        const double xt = x[j] / dc[j];
        const double sl = fmin(xt - L.scl()[j], L.scu()[j] - xt);
        const bool in = sl > fabs(d);
        mc[j] = in;
        nb += in;
    });
    nr = bsum(nr, S.red);
    nb = bsum(nb, S.red);

    // ---- 2. right-hand sides and CGLS start
    double nf = 0.0, nh = 0.0;
    for (int i = tid; i < m; i += NT) {
        const double f = mr[i] ? gy[i] / dr[i] : 0.0;
        r1[i] = f;
        z2[i] = 0.0;
        nf += f * f;
    }
    for (int j = tid; j < n; j += NT) {
        const double h = mc[j] ? dc[j] * (gx[j] + gobj * c[j]) : 0.0;
        r2[j] = h;
        z1[j] = 0.0;
        nh += h * h;
    }
    nf = bsum(nf, S.red);
    nh = bsum(nh, S.red);
    double g1 = 0.0, g2 = 0.0;
    spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{r1}, S, [&](int j, double a) {
        const double v = mc[j] ? a : 0.0;
        s1[j] = v;
        p1[j] = v;
        g1 += v * v;
    });
    spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{r2}, S, [&](int i, double a) {
        const double v = mr[i] ? a : 0.0;
        s2[i] = v;
        p2[i] = v;
        g2 += v * v;
    });
    g1 = bsum(g1, S.red);
    g2 = bsum(g2, S.red);
    const double tol2 = J.tol * J.tol;
    const double g10 = g1, g20 = g2;
    bool done1 = !(g1 > 0.0), done2 = !(g2 > 0.0);
    int it1 = 0, it2 = 0;
    for (int it = 0; it < J.maxit && !(done1 && done2); ++it) {
        double d1 = 0.0, d2 = 0.0;
        if (!done1)
            spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{p1}, S, [&](int i, double a) {
                const double v = mr[i] ? a : 0.0;
                t1[i] = v;
                d1 += v * v;
            });
        if (!done2)
            spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{p2}, S, [&](int j, double a) {
                const double v = mc[j] ? a : 0.0;
                t2[j] = v;
                d2 += v * v;
            });
        d1 = bsum(d1, S.red);
        d2 = bsum(d2, S.red);
        if (!done1 && !(d1 > 0.0)) done1 = true;
        if (!done2 && !(d2 > 0.0)) done2 = true;
        const double a1 = done1 ? 0.0 : g1 / d1;
        const double a2 = done2 ? 0.0 : g2 / d2;
        if (!done1) {
            for (int j = tid; j < n; j += NT) z1[j] += a1 * p1[j];
            for (int i = tid; i < m; i += NT) r1[i] -= a1 * t1[i];
        }
        if (!done2) {
            for (int i = tid; i < m; i += NT) z2[i] += a2 * p2[i];
            for (int j = tid; j < n; j += NT) r2[j] -= a2 * t2[j];
        }
        __syncthreads();
        double h1 = 0.0, h2 = 0.0;
        if (!done1)
            spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{r1}, S, [&](int j, double a) {
                const double v = mc[j] ? a : 0.0;
                s1[j] = v;
                h1 += v * v;
            });
        if (!done2)
            spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{r2}, S, [&](int i, double a) {
                const double v = mr[i] ? a : 0.0;
                s2[i] = v;
                h2 += v * v;
            });
        h1 = bsum(h1, S.red);
        h2 = bsum(h2, S.red);
        if (!done1) {
            const double b1 = h1 / g1;
            for (int j = tid; j < n; j += NT) p1[j] = s1[j] + b1 * p1[j];
            g1 = h1;
            ++it1;
            if (g1 <= tol2 * g10) done1 = true;
        }
        if (!done2) {
            const double b2 = h2 / g2;
            for (int i = tid; i < m; i += NT) p2[i] = s2[i] + b2 * p2[i];
            g2 = h2;
            ++it2;
            if (g2 <= tol2 * g20) done2 = true;
        }
        __syncthreads();
    }

    // ---- 3. true residuals: t1 = A~_RB z1 - f, t2 = A~_RB' z2 - h; and
    // (A~'q~)_j for the column-bound gradients (t2 before masking)
    double e1 = 0.0, e2 = 0.0;
    spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{z1}, S, [&](int i, double a) {
        const double f = mr[i] ? gy[i] / dr[i] : 0.0;
        const double v = mr[i] ? a - f : 0.0;
        e1 += v * v;
    });
    spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{z2}, S, [&](int j, double a) {
        const double w = gx[j] + gobj * c[j];
        if (mc[j]) {
            const double v = a - dc[j] * w;
            e2 += v * v;
            J.gcl[co + j] = 0.0;
            J.gcu[co + j] = 0.0;
        } else {
            const double g = w - a / dc[j];
            const double xt = x[j] / dc[j];
            const bool lower = (xt - L.scl()[j]) <= (L.scu()[j] - xt);
            J.gcl[co + j] = lower ? g : 0.0;
            J.gcu[co + j] = lower ? 0.0 : g;
        }
        J.gc[co + j] = gobj * x[j] - (mc[j] ? dc[j] * z1[j] : 0.0);
    });
    e1 = bsum(e1, S.red);
    e2 = bsum(e2, S.red);
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
        AdjInfo inf;
        inf.basic_cols = static_cast<int>(nb);
        inf.active_rows = static_cast<int>(nr);
        inf.iters_p = it1;
        inf.iters_q = it2;
        inf.res_p = nf > 0.0 ? sqrt(e1 / nf) : sqrt(e1);
        inf.res_q = nh > 0.0 ? sqrt(e2 / nh) : sqrt(e2);
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

}  // namespace igaos::batch
