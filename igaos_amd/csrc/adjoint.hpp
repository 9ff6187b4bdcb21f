#pragma once
// Batched adjoint of the PDHG solution map (no reference counterpart).
//
// At a solution (x*, y*) of  min c'x  s.t. rmin <= Ax <= rmax, cl <= x <= cu
// (dual sign convention d = c + A'y), the projections of the PDHG step
// split the coordinates into
//   B: columns strictly inside their bounds (the primal projection is the
//      identity there, so the reduced cost d_j vanishes), and
//   R: rows whose dual is free (equality rows, and inequality rows sitting
//      on a bound; elsewhere y_i = 0 is pinned by the dual projection).
// With those active-set masks, the fixed point of the PDHG map linearised
// at (x*, y*) is the pair of equations
//   c_B + A_RB' y_R = 0,          A_RB x_B + A_RN x_N = b_R,
// and the transposed (adjoint) system for upstream gradients gx, gy, gobj is
//   A_RB p = gy_R,                A_RB' q = w_B,   w = gx + gobj * c,
// whose solution gives
//   dL/dc_j    = gobj x_j - [j in B] p_j
//   dL/db_i    = [i in R] q_i          (b_i = the bound row i sits on)
//   dL/dA_ij   = -[i in R] ([j in B] p_j y_i + q_i x_j)
//   dL/dbnd_j  = [j not in B] (w_j - (A_R' q)_j)   (the bound x_j sits on).
// Polishing uses the same masks: it solves the active-set equations for
// corrections of x_B and y_R (x_N on its bounds, y off R zero), and keeps
// the result only if it is primal and dual feasible to 1e-9 (relative).
// Both systems are solved per instance by CGLS in FP64 on the Ruiz-scaled
// matrix (the scaled systems are  A~_RB p~ = D_r^-1 gy,  A~_RB' q~ = D_c w
// with p = D_c p~, q = q~ / D_r), reusing the batched solver's instance
// SpMVs with the masks applied to their inputs and outputs.  CGLS started
// at zero returns the minimum-norm least-squares solution when the masks
// are not square (a degenerate solution).
#include <hip/hip_runtime.h>

#include "batch.hpp"

namespace igaos::batch {

struct AdjInfo {
    int basic_cols, active_rows, iters_p, iters_q;
    int accepted;         // polish: the polished pair passed the checks
    int flags;            // polish: failed checks (bits: column bounds, rows,
                          // active-row dual signs, d_B = 0, reduced-cost signs)
    double res_p, res_q;  // final relative residuals of the two systems
    double viol;          // polish: largest relative violation
};

struct AdjArrays {
    unsigned char *mc, *mr;                 // masks: columns in B, rows in R
    double *z1, *s1, *p1, *r1, *t1;         // A_RB p = gy   (z1, s1, p1: n; r1, t1: m)
    double *z2, *s2, *p2, *r2, *t2;         // A_RB' q = w   (z2, s2, p2: m; r2, t2: n)
    const double *gx, *gy, *gobj;           // upstream gradients
    double *gc, *gcl, *gcu, *grmin, *grmax, *gax;  // outputs (gax in CSR order)
    double *xs, *ys;                        // polish: scaled candidate pair
    double* obj;                            // polish: objective of accepted pairs
    AdjInfo* info;
    double tol;
    int maxit;
};

void launch_adjoint(const Arrays<float>* a, const LPDesc* lps, int count,
                    const AdjArrays* adj, hipStream_t s);
void launch_polish(const Arrays<float>* a, const LPDesc* lps, int count,
                   const AdjArrays* adj, hipStream_t s);
void launch_polish(const Arrays<double>* a, const LPDesc* lps, int count,
                   const AdjArrays* adj, hipStream_t s);
void launch_adjoint(const Arrays<double>* a, const LPDesc* lps, int count,
                    const AdjArrays* adj, hipStream_t s);

}  // namespace igaos::batch
