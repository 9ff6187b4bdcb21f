#pragma once
// Host-callable launchers of the single-LP PDHG kernels (gfx950).
// All pointers are device pointers unless stated; every launcher enqueues
// on the given stream and returns without synchronising.

#include <hip/hip_runtime.h>

#include <vector>

#include "dev_common.hpp"
#include "pdhg_ctrl.hpp"

namespace igaos::dev {

// Host: cut rows [0, nrows) of a CSR row pointer into row blocks; returns
// the interleaved (first row, first nonzero) pairs of the RowBlocks layout
// (dev_common.hpp), nblk + 1 pairs.
std::vector<int> build_row_blocks(const int* ptr, int nrows);

// Column-sliced copy of a matrix for gather locality on large LPs.  The
// columns are cut into S contiguous ranges holding equal numbers of
// nonzeros; each row's nonzeros are split by range (keeping their order)
// and stored slice-major.  The slices' row blocks run as S launches, one
// per slice, so that while a slice runs only its range of the gathered
// vector is in use and stays resident in every XCD's L2.  Each block
// writes partial row sums part[s * rows + r]; an epilogue kernel adds the
// S partials of a row in slice order.
struct HostSlices {
    int S = 0;
    std::vector<int> ptr;     // S * (rows + 1) offsets into idx
    std::vector<int> idx;     // column indices, slice-major
    std::vector<int> src;     // position of each entry in the original CSR
    std::vector<unsigned short> slot;  // gather order only: see Sliced::slot
    std::vector<int> quad;    // per block: first row, end row, first nz, end nz
    std::vector<int> bslice;  // per block: its slice
};
// With `gather_order`, the entries of every block of two or more rows are
// stored sorted by column (stable) and `slot` holds their row-major places.
HostSlices build_slices(int rows, int cols, const int* ptr, const int* idx, int S,
                        bool gather_order = false);

constexpr int MAX_SLICES = 16;
template <typename V>
struct Sliced {
    int S = 0, rows = 0, nblk = 0;
    int sb[MAX_SLICES + 1] = {};  // blocks of slice s: [sb[s], sb[s + 1])
    const int* ptr = nullptr;
    const int* idx = nullptr;
    const V* val = nullptr;
    // gather order (optional): the nonzeros of every multi-row block are
    // stored sorted by column, slot[q] = the entry's place in row-major
    // order (the row sums do not change; see rowblock_sums)
    const unsigned short* slot = nullptr;
    const int4* quad = nullptr;
    const int* bslice = nullptr;
    V* part = nullptr;        // S * rows partial row sums
    // XCD-pinned schedule: all slices run in one launch of xgrid
    // workgroups; workgroup b (dispatched to XCD b % 8) takes block
    // xmap[b] (-1: none), and every XCD only takes blocks of the slices
    // assigned to it, so its L2 holds only those slices' range of the
    // gathered vector
    const int* xmap = nullptr;
    int xgrid = 0;
};
constexpr int XCDS = 8;  // MI355X: 8 XCDs, workgroups dispatched round robin
// Host: the XCD-pinned block map of a sliced matrix (see Sliced::xmap).
// With S <= 8 slices, XCD x serves slice x % S (S = 1, 2, 4, 8 split the
// XCDs evenly); with S > 8, XCD x serves slices x, x + 8, ...  A slice's
// blocks are cut into contiguous runs, one per XCD serving it.
std::vector<int> build_xcd_map(const std::vector<int>& bslice, int S);
// dst[q] = V(src_val[src[q]]) for q < nnz (the sliced copy of the values).
void gather_values(const double* src_val, const int* src, double* dst, long nnz,
                   hipStream_t s);
void gather_values(const double* src_val, const int* src, float* dst, long nnz,
                   hipStream_t s);

// y = A x  (plain SpMV).
void spmv(const RowBlocks& rb, const int* ptr, const int* idx,
          const double* val, const double* x, double* y, hipStream_t s);
void spmv(const RowBlocks& rb, const int* ptr, const int* idx,
          const float* val, const float* x, float* y, hipStream_t s);
// y = A (xref + xd) in FP64 (xref may be null).
void spmv_ref(const RowBlocks& rb, const int* ptr, const int* idx,
              const double* val, const double* xref, const float* xd,
              double* y, hipStream_t s);

// Buffers of one single-LP PDHG iteration.  x[c], y[c] hold the current
// iterate for c = LsState::cur; the other slot receives the attempt.
template <typename V>
struct IterBufs {
    int m = 0, n = 0;
    const int *ap = nullptr, *ai = nullptr, *cp = nullptr, *ci = nullptr;
    const V *ax = nullptr, *acx = nullptr;
    RowBlocks rbA, rbT;
    V* x[2] = {nullptr, nullptr};
    V* y[2] = {nullptr, nullptr};
    V *xbar = nullptr, *ata = nullptr, *xavg = nullptr, *yavg = nullptr;
    V *axcur = nullptr, *adx = nullptr;
    const V *gref = nullptr, *lo = nullptr, *hi = nullptr;   // primal data
    const V *rlo = nullptr, *rhi = nullptr, *yoff = nullptr; // dual data
    double* partT = nullptr;  // nT_part partials of |dx|^2
    double* partA = nullptr;  // nA_part pairs (|dy|^2, dy'Adx)
    int nT_part = 0, nA_part = 0;
    // column-sliced A' (gathers y) and A (gathers xbar); S = 0: unsliced,
    // the fused row-block kernels run instead
    Sliced<V> sT, sA;
    pdhg::LsState* st = nullptr;
    // the attempt kernels load the matrix streams with nontemporal hints
    // (for matrices too large to stay in the last-level cache)
    bool nt = false;
};

// One line-search attempt: fused A'y + primal step, fused A xbar + dual
// step + line-search decision (both no-ops once the state is frozen).
void launch_attempt(const IterBufs<double>& b, hipStream_t s);
void launch_attempt(const IterBufs<float>& b, hipStream_t s);
// The two kernels of an attempt, separately (for timing).
void launch_primal(const IterBufs<double>& b, hipStream_t s);
void launch_primal(const IterBufs<float>& b, hipStream_t s);
void launch_dual(const IterBufs<double>& b, hipStream_t s);
void launch_dual(const IterBufs<float>& b, hipStream_t s);
void launch_decide(const IterBufs<double>& b, hipStream_t s);
void launch_decide(const IterBufs<float>& b, hipStream_t s);
// Apply the averaging / A x commit of the last accepted step.
void launch_flush(const IterBufs<double>& b, int cur, double w, hipStream_t s);
void launch_flush(const IterBufs<float>& b, int cur, double w, hipStream_t s);
// The same, with cur, w and whether a step is pending read from the device
// state (b.st), for captured graphs; a later kernel must clear the state's
// pending flag (k_kkt_done does, given KktData::clear_mix).
void launch_flush_state(const IterBufs<double>& b, hipStream_t s);
void launch_flush_state(const IterBufs<float>& b, hipStream_t s);

// KKT evaluation in original units for two iterates at once
// (src/pdhg/pdhg.cu:173-254 via metrics_of, 694-733).
struct KktData {
    int m = 0, n = 0;
    const int *ap = nullptr, *ai = nullptr, *cp = nullptr, *ci = nullptr;
    const double *sax = nullptr, *sacx = nullptr;  // scaled FP64 A, A'
    RowBlocks rbA, rbT;
    const double *c = nullptr, *cl = nullptr, *cu = nullptr;
    const double *rmin = nullptr, *rmax = nullptr;  // original units
    const double *dr = nullptr, *dc = nullptr;
    double cmax = 0.0;
    double *partC = nullptr, *partR = nullptr, *partR2 = nullptr;
    double* rowres = nullptr;        // 2 x 4 row-pass results
    pdhg::LsState* clear_mix = nullptr;  // k_kkt_done clears this state's pending
                                         // flag (captured check graphs)
};
template <typename V>
struct KktIter {
    const double* xref[2] = {nullptr, nullptr};
    const V* xd[2] = {nullptr, nullptr};
    const double* yref[2] = {nullptr, nullptr};
    const V* yd[2] = {nullptr, nullptr};
    double* ata_out[2] = {nullptr, nullptr};  // optional A~'y~ (scaled)
    double* tm_out[2] = {nullptr, nullptr};   // optional A~x~ (scaled)
};
void launch_kkt(const KktData& d, const KktIter<double>& it,
                pdhg::Metrics* out, hipStream_t s);
void launch_kkt(const KktData& d, const KktIter<float>& it,
                pdhg::Metrics* out, hipStream_t s);

// Ruiz equilibration (10 sweeps, src/pdhg/pdhg.cu:469-492) on device.
void ruiz(int m, int n, const int* ap, const int* ai, const double* ax,
          const int* cp, const int* ci, const double* acx, double* dr,
          double* dc, hipStream_t s);
// Pock-Chambolle (alpha = 1) on top of the current factors (opt-in);
// rs (m), cs (n) are scratch.
void pock_chambolle(int m, int n, const int* ap, const int* ai, const double* ax,
                    const int* cp, const int* ci, const double* acx, double* dr,
                    double* dc, double* rs, double* cs, hipStream_t s);
// Scaled data (pdhg.cu:499-513).
void scale_data(int m, int n, const int* ap, const int* ai, const double* ax,
                const int* cp, const int* ci, const double* acx,
                const double* dr, const double* dc, const double* c,
                const double* cl, const double* cu, const double* rmin,
                const double* rmax, double* sax, double* sacx, double* sc,
                double* scl, double* scu, double* srmin, double* srmax,
                hipStream_t s);

// out[0] = sum (a - b)^2 (b may be null), fixed order (per-workgroup
// partials, then one workgroup); `part` holds at least reduce_blocks(n)
// doubles.
int reduce_blocks(long n);
void sumsq_diff(const double* a, const double* b, long n, double* part, double* out,
                hipStream_t s);
void sumsq_diff(const float* a, const float* b, long n, double* part, double* out,
                hipStream_t s);
// out[0] = sum_j q_j^2 with q = srmax if finite, else srmin if finite, else 0
void sumsq_q(const double* srmin, const double* srmax, long m, double* part,
             double* out, hipStream_t s);
// out[0] = max |v|
void absmax(const double* v, long n, double* part, double* out, hipStream_t s);

// Elementwise helpers.
void fill(double* v, long n, double val, hipStream_t s);
void fill(float* v, long n, float val, hipStream_t s);
void scale_into(double* out, const double* in, long n, double f, hipStream_t s);
void to_float(const double* in, float* out, long n, hipStream_t s);
// FP64 iterate of an FP32 correction: out = ref + d.  With `snap` (the
// primal side: the correction's clamp bounds lo, hi and the scaled column
// bounds scl, scu), a correction exactly at lo or hi stands for the column
// at its bound and gives scl or scu exactly: ref + (double)hi differs from
// scu by the FP32 rounding of scu - ref, up to 6e-8 |scu - ref|, which the
// KKT test's bound tolerance (1e-9 (1 + |bound|), pdhg.cu:206-215) does
// not absorb, so the column would count as dual infeasible.
struct BoundSnap {
    const float* lo = nullptr;
    const float* hi = nullptr;
    const double* scl = nullptr;
    const double* scu = nullptr;
};
void combine(const double* ref, const float* d, double* out, long n,
             hipStream_t s, const BoundSnap* snap = nullptr);
// Four combines in one launch, each element as combine computes it:
// ox[q] = xref + xd[q] (snapped), oy[q] = yref + yd[q].
struct Combine4 {
    const double *xref = nullptr, *yref = nullptr;
    const float* xd[2] = {nullptr, nullptr};
    const float* yd[2] = {nullptr, nullptr};
    double* ox[2] = {nullptr, nullptr};
    double* oy[2] = {nullptr, nullptr};
    long n = 0, m = 0;
    BoundSnap snap;
};
void combine4(const Combine4& c, hipStream_t s);
// ref = ref + d (snapped as combine does), in place on ref; d zeroed.
void fold(double* ref, float* d, long n, hipStream_t s, const BoundSnap* snap = nullptr);
// Shifted FP32 data of the correction problem (see engine.cpp).
void refine_data(int m, int n, const double* scl, const double* scu,
                 const double* sc, const double* xref, const double* ata_ref,
                 const double* srmin, const double* srmax,
                 const double* aref, const double* yref, float* lo, float* hi,
                 float* gref, float* rlo, float* rhi, float* yoff,
                 hipStream_t s);
// Original-unit output: x = dc * sx (sx = ref + d, snapped as combine
// does when `snap` is given), y = sy / dr.
void unscale_out(int m, int n, const double* dc, const double* dr,
                 const double* xref, const double* xd64, const float* xd32,
                 const double* yref, const double* yd64, const float* yd32,
                 double* x, double* y, hipStream_t s, const BoundSnap* snap = nullptr);

}  // namespace igaos::dev
