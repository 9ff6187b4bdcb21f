#pragma once
// Launchers of the exact FP64 mode (exact.hip): the single-LP iteration
// with every reduction in a fixed, reference-defined order and no fused
// multiply-add, so that a run is bit-reproducible against a sequential FP64
// host replay of src/pdhg/pdhg.cu.

#include "pdhg_kernels.hpp"

namespace igaos::dev {

// The reference's dot-product grid is capped at 4096 workgroups
// (pdhg.cu:571-572), so the partials never exceed 3 * 4096 doubles.
constexpr int DOTS_MAX_BLOCKS = 4096;
int dots_blocks(long n, long m);

// Per-attempt vectors of the exact mode and the exponent table.
struct ExactBufs {
    double* dx = nullptr;    // n: x' - x
    double* dy = nullptr;    // m: y' - y
    double* tm = nullptr;    // m: A~ xbar
    double* part = nullptr;  // 3 * dots_blocks(n, m) dot partials
    const double* powtab = nullptr;  // k^-0.3, k^-0.6 for k = 1..pow_cap
    long long pow_cap = 0;
};

// One line-search attempt (no-ops once the state is frozen).
void launch_attempt_exact(const IterBufs<double>& b, const ExactBufs& e, hipStream_t s);
// Small LPs (max(n, m) <= 256 RUN_MAX_GD, nnz <= RUN_MAX_NNZ): one workgroup
// runs attempts with the arithmetic of launch_attempt_exact until the state
// freezes or `cap` attempts have run, in one launch; each sparse product is
// staged in LDS (120 KB) and summed per row from there.  e.powtab may be
// null (device pow).
constexpr int RUN_MAX_GD = 64;
constexpr int RUN_MAX_NNZ = 15360;
bool run_exact_fits(long n, long m, long nnz);
void launch_run_exact(const IterBufs<double>& b, const ExactBufs& e, int cap, hipStream_t s);
// The pending averaging / A x commit of the last accepted step.
void launch_flush_exact(const IterBufs<double>& b, int cur, double w, hipStream_t s);
// y = A x, one thread per row, each row summed in storage order.
void spmv_seq(int rows, const int* ptr, const int* idx, const double* val,
              const double* x, double* y, hipStream_t s);
// The line-search dots |dx|^2, |dy|^2, dy'adx with the reference's layout
// (pdhg.cu:63-125) into out[0..2] (device); part holds 3 * dots_blocks(n, m)
// doubles.
void ls_dots(int n, int m, const double* dx, const double* dy, const double* adx,
             double* part, double* out, hipStream_t s);
// out[0] = sum_i (a_i - b_i)^2 (b may be null), in index order.
void sumsq_seq(const double* a, const double* b, long n, double* out, hipStream_t s);
// out[0] = sum_i q_i^2 (q as in pdhg.cu:652-657), in index order.
void sumsq_q_seq(const double* srmin, const double* srmax, long m, double* out,
                 hipStream_t s);

// KKT metrics of two scaled FP64 iterates in original units.
struct ExactKkt {
    const double* x[2] = {nullptr, nullptr};
    const double* y[2] = {nullptr, nullptr};
    double* ata[2] = {nullptr, nullptr};  // n each: A~'y~ (written)
    double* tm[2] = {nullptr, nullptr};   // m each: A~x~ (written)
    double* term = nullptr;               // 2 (2n + m) summands
};
void launch_kkt_exact(const KktData& d, const ExactKkt& k, pdhg::Metrics* out,
                      hipStream_t s);

}  // namespace igaos::dev
