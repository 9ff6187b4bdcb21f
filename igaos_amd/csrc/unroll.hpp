#pragma once
// Unrolled differentiation of a batch (no reference counterpart).
//
// Forward: K steps of fixed-step PDHG on every instance, in original units
// and FP64, from x = y = 0 (the iteration of src/pdhg/pdhg.cu:32-61 with
// its dual sign convention d = c + A'y, without line search, restarts or
// averaging):
//   x' = clamp(x - tau (c + A'y), cl, cu),    xb = 2 x' - x,
//   t  = y + sigma A xb,                       y' = t - sigma clamp(t / sigma, rmin, rmax),
// with tau = rho / (omega |A|), sigma = rho omega / |A|, rho = 0.9 (so
// tau sigma |A|^2 = 0.81 < 1), |A| from 30 power iterations of A'A (as in
// pdhg.cu:605-631) and omega = clamp(|c| / |q|, 0.3, 3) (q_i = rmax_i if
// finite, else rmin_i if finite, else 0; pdhg.cu:647-663), or given step
// sizes (`fixed`).  The step sizes are constants of the map: gradients do
// not flow through them.  The iterate pair is checkpointed every 64 steps
// (k = 0, 64, 128, ...).
//
// Backward: reverse mode through the K steps, segment by segment from the
// last checkpoint: each segment's iterates are recomputed from its
// checkpoint into a 64-step buffer, then the steps are differentiated in
// reverse.  For one step with output adjoints (gx', gy'):
//   tbar = gy' where t/sigma is clamped (<= rmin or >= rmax), else 0 (inside,
//   y' = 0);  d rmin += -sigma gy' where t/sigma <= rmin;  d rmax += -sigma gy'
//   where t/sigma >= rmax (rmin first);
//   gy = tbar;  xbb = sigma A'tbar;  dA_ij += sigma tbar_i xb_j;
//   h = gx' + 2 xbb;  gx = -xbb;  vbar = h inside (cl, cu), else 0;
//   d cl += h where v <= cl;  d cu += h where v >= cu (cl first);
//   gx += vbar;  gbar = -tau vbar;  d c += gbar;  gy += A gbar;
//   dA_ij += y_i gbar_j.
// The objective output is c'x_K + obj_const (its gradient adds gobj c to
// the adjoint of x_K and gobj x_K to d c).
#include <hip/hip_runtime.h>

#include "batch.hpp"

namespace igaos::batch {

constexpr int UNROLL_SEG = 64;  // steps per checkpoint segment

struct UnrollArrays {
    int K;                  // steps
    int fixed;              // tau, sigma given (else computed from the data)
    long sum_n, sum_m;      // strides of the per-step / per-checkpoint slabs
    double *ckx, *cky;      // checkpoints: (K / 64 + 1) slabs of n, m
    double *segx, *segy;    // one segment's iterates: 64 slabs of n, m
    double *ax, *ay;        // adjoints of the current step's outputs (n, m)
    double *xb, *gb;        // recomputed xb and gbar (n)
    double *tb, *tmpm;      // tbar and A gbar (m)
    double *tmpn;           // sigma A'tbar (n)
    unsigned char *cx, *cy; // projection codes (0 inside, 1 lower, 2 upper)
    double *tau, *sigma;    // per instance step sizes
    double *xK, *yK, *obj;  // forward outputs
    const double *gx, *gy, *gobj;  // upstream gradients (n, m, count)
    double *gc, *gcl, *gcu, *grmin, *grmax, *gax;  // outputs (gax CSR order)
};

void launch_unroll_forward(const Arrays<float>* a, const LPDesc* lps, int count,
                           const UnrollArrays* u, hipStream_t s);
void launch_unroll_forward(const Arrays<double>* a, const LPDesc* lps, int count,
                           const UnrollArrays* u, hipStream_t s);
void launch_unroll_backward(const Arrays<float>* a, const LPDesc* lps, int count,
                            const UnrollArrays* u, hipStream_t s);
void launch_unroll_backward(const Arrays<double>* a, const LPDesc* lps, int count,
                            const UnrollArrays* u, hipStream_t s);

}  // namespace igaos::batch
