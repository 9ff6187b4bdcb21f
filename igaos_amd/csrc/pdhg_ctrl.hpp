#pragma once
// PDHG step-size controller and KKT bookkeeping shared by host and device.
//
// LsState is the iteration state the single-LP engine keeps in device
// memory so that the adaptive line search (src/pdhg/pdhg.cu:751-794) is
// decided on the GPU by the last workgroup of the dual kernel instead of a
// blocking 24-byte copy to the host per attempt (pdhg.cu:768).  ls_decide()
// is the reference's accept/reject rule verbatim in meaning:
//   etabar  = (w|dx|^2 + |dy|^2/w) / (2 dy'A dx)        (inf if num<=0 or cross<=0)
//   eta_new = min((1 - k^-0.3) etabar, (1 + k^-0.6) eta)
//   accept  = eta <= etabar || num <= 0 || cross <= 0 || ++retries > 20
// with eta clamped to [1e-12, 1e12] on the two branches.

#include <hip/hip_runtime.h>

#include <cmath>

namespace igaos::pdhg {

constexpr int CHECK_EVERY = 64;  // pdhg.cu:665

struct LsState {
    double eta;
    double omega;
    double mix_w;        // averaging weight of the pending accepted step
    long long k;         // iteration being computed (1-based)
    long long hn;        // averaging counter (epoch-local)
    long long maxit;     // last iteration allowed (MAXIT)
    long long attempts;  // primal/dual attempts including retries
    int retries;         // retries of iteration k so far
    int fresh;           // next primal kernel starts iteration k (computes A'y)
    int cur;             // buffer index holding the current x, y
    int frozen;          // iteration k-1 ended on a check boundary or MAXIT
    int mix_pending;     // the last accepted step still needs averaging
    int no_freeze;       // timing mode: never freeze at check boundaries
    double dots[3];      // dots of the last attempt (diagnostics)
};

// p3 = k^-0.3 and p6 = k^-0.6 of the iteration s.k being decided.  The
// rule is written without fused multiply-adds (the reference's controller
// runs on the host, pdhg.cu:770-793).
__host__ __device__ inline void ls_decide_pw(LsState& s, double d0, double d1,
                                             double d2, double p3, double p6) {
#pragma clang fp contract(off)
    const double INF = __builtin_huge_val();
    const double num = s.omega * d0 + d1 / s.omega;
    const double cross = d2;
    const double etabar = (num > 0.0 && cross > 0.0) ? num / (2.0 * cross) : INF;
    const double eta_new =
        std::isfinite(etabar) ? fmin((1.0 - p3) * etabar, (1.0 + p6) * s.eta) : INF;
    s.dots[0] = d0;
    s.dots[1] = d1;
    s.dots[2] = d2;
    s.attempts += 1;
    const bool accept =
        !(s.eta > etabar) || num <= 0.0 || cross <= 0.0 || ++s.retries > 20;
    if (accept) {
        if (std::isfinite(eta_new) && eta_new > 1e-12) s.eta = fmin(eta_new, 1e12);
        s.cur ^= 1;
        s.mix_w = 2.0 / (static_cast<double>(s.hn) + 2.0);
        s.hn += 1;
        s.mix_pending = 1;
        const long long done = s.k;
        s.k += 1;
        s.retries = 0;
        s.fresh = 1;
        if (!s.no_freeze && (done % CHECK_EVERY == 0 || done >= s.maxit)) s.frozen = 1;
    } else {
        s.eta = fmax(eta_new, 1e-12);
        s.fresh = 0;
        s.mix_pending = 0;
    }
}

__host__ __device__ inline void ls_decide(LsState& s, double d0, double d1,
                                          double d2) {
    const double kk = static_cast<double>(s.k);
    ls_decide_pw(s, d0, d1, d2, pow(kk, -0.3), pow(kk, -0.6));
}

// pdhg.cu:159-170
struct Metrics {
    double pinf, dinf, gap, op, od, prow;
    __host__ __device__ bool finite_res() const {
        return std::isfinite(pinf) && std::isfinite(dinf);
    }
    __host__ __device__ double err() const {
        const double g = std::isfinite(gap) ? gap : 0.0;
        return fmax(pinf, fmax(dinf, g));
    }
};

// Termination test on an iterate pair (pdhg.cu:835-840).
__host__ __device__ inline bool passes(const Metrics& t, double tol) {
    return t.pinf <= tol && t.dinf <= tol && t.prow <= tol &&
           std::isfinite(t.gap) && t.gap <= tol;
}
// Final certification (pdhg.cu:1095-1098).
__host__ __device__ inline bool certified(const Metrics& t, double tol) {
    return t.pinf <= tol && t.prow <= tol && std::isfinite(t.gap) &&
           t.gap <= tol;
}

// Restart rule (pdhg.cu:989-1012): returns whether to restart and updates
// the epoch bookkeeping the same way the reference does.
struct RestartState {
    double epoch_kkt;   // < 0 = unset
    double prev_cand;
    long long epoch_start;
};

__host__ __device__ inline bool restart_decide(RestartState& r, double cand,
                                               long long k) {
    if (r.epoch_kkt < 0.0) {
        r.epoch_kkt = cand;
        r.prev_cand = cand;
    }
    const long long t_epoch = k - r.epoch_start;
    const bool now =
        (std::isfinite(cand) && std::isfinite(r.epoch_kkt) && r.epoch_kkt > 0.0 &&
         (cand <= 0.2 * r.epoch_kkt ||
          (cand <= 0.8 * r.epoch_kkt && std::isfinite(r.prev_cand) &&
           cand > r.prev_cand))) ||
        static_cast<double>(t_epoch) >= 0.36 * static_cast<double>(k);
    r.prev_cand = cand;
    return now;
}

// Primal-weight update (pdhg.cu:1034-1060): omega <- sqrt(omega*sqrt(Dy/Dx)).
__host__ __device__ inline double primal_weight(double omega, double dxn,
                                                double dyn) {
    if (dxn > 1e-24 && dyn > 1e-24) {
        const double om = sqrt(omega * sqrt(dyn / dxn));
        if (std::isfinite(om) && om > 0.0) return fmin(fmax(om, 1e-6), 1e6);
    }
    return omega;
}

}  // namespace igaos::pdhg
