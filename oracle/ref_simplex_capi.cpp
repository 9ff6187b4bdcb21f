// TEST INFRASTRUCTURE ONLY.  A C-ABI shim, compiled against the reference's
// own simplex sources where they lie (src/simplex/simplex.cpp and its
// headers) into oracle/_ref/, so that tests can compute objective fixtures
// with the reference's revised simplex (src/simplex/simplex.hpp:24-28) and
// bench.py can time it as a CPU baseline.  No reference source is copied.
#include <cstring>
#include <vector>

#include "model.hpp"
#include "simplex.hpp"

// warm_basis (m) / warm_state (n + m), when not null, are passed to the
// simplex as its WarmStart (src/simplex/simplex.hpp:10-15).
extern "C" int ref_simplex_solve_warm(int m, int n, const double* rmin,
                                      const double* rmax, const double* cl,
                                      const double* cu, const double* c,
                                      const int* ap, const int* ai,
                                      const double* ax, const int* cp,
                                      const int* ci, const double* acx,
                                      double obj_const, double time_limit_s,
                                      double* objective, double* x_out,
                                      double* y_out, long* iterations,
                                      double* solve_ms, const int* warm_basis,
                                      const unsigned char* warm_state);

extern "C" int ref_simplex_solve(int m, int n, const double* rmin,
                                 const double* rmax, const double* cl,
                                 const double* cu, const double* c,
                                 const int* ap, const int* ai,
                                 const double* ax, const int* cp,
                                 const int* ci, const double* acx,
                                 double obj_const, double time_limit_s,
                                 double* objective, double* x_out,
                                 double* y_out, long* iterations,
                                 double* solve_ms) {
    return ref_simplex_solve_warm(m, n, rmin, rmax, cl, cu, c, ap, ai, ax, cp, ci, acx,
                                  obj_const, time_limit_s, objective, x_out, y_out,
                                  iterations, solve_ms, nullptr, nullptr);
}

extern "C" int ref_simplex_solve_warm(int m, int n, const double* rmin,
                                      const double* rmax, const double* cl,
                                      const double* cu, const double* c,
                                      const int* ap, const int* ai,
                                      const double* ax, const int* cp,
                                      const int* ci, const double* acx,
                                      double obj_const, double time_limit_s,
                                      double* objective, double* x_out,
                                      double* y_out, long* iterations,
                                      double* solve_ms, const int* warm_basis,
                                      const unsigned char* warm_state) {
    igaos::io::Model md;
    md.m = m;
    md.n = n;
    md.rmin.assign(rmin, rmin + m);
    md.rmax.assign(rmax, rmax + m);
    md.cl.assign(cl, cl + n);
    md.cu.assign(cu, cu + n);
    md.c.assign(c, c + n);
    const int nnz = ap[m];
    md.ap.assign(ap, ap + m + 1);
    md.ai.assign(ai, ai + nnz);
    md.ax.assign(ax, ax + nnz);
    md.cp.assign(cp, cp + n + 1);
    md.ci.assign(ci, ci + nnz);
    md.acx.assign(acx, acx + nnz);
    md.integ.assign(n, 0);
    md.obj_const = obj_const;
    igaos::Options opt;
    opt.time_limit_s = time_limit_s;
    opt.tolerance = 1e-9;
    igaos::simplex::WarmStart ws;
    if (warm_basis && warm_state) {
        ws.basis.assign(warm_basis, warm_basis + m);
        ws.nb_state.assign(warm_state, warm_state + n + m);
    }
    igaos::Solution s = igaos::simplex::solve(md, opt, warm_basis ? &ws : nullptr);
    *objective = s.objective;
    *iterations = s.iterations;
    *solve_ms = s.solve_time_ms;
    if (x_out && (int)s.x.size() == n) std::memcpy(x_out, s.x.data(), sizeof(double) * n);
    if (y_out && (int)s.y.size() == m) std::memcpy(y_out, s.y.data(), sizeof(double) * m);
    return static_cast<int>(s.status);
}
