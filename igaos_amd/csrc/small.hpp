#pragma once
// One-workgroup runner of the default mode for small LPs (small.hip): a
// check period of line-search attempts per launch, the LP's vectors in
// registers and LDS, sliced-ELL copies of its matrices streamed from L2.

#include <hip/hip_runtime.h>

#include <vector>

#include "pdhg_kernels.hpp"

namespace igaos::dev {

constexpr int SMALL_MAX_DIM = 3 * 1024;  // max(n, m): up to 3 columns / rows a thread
// Matrix bytes per product (values and 16-bit indices) up to which one CU
// streams an attempt's two products faster than the whole GPU runs the
// three-kernel attempt (MI355X: 2.0x at 150 KB, 1.5x at 300 KB, even at
// 600 KB, 0.83x at 720 KB).
constexpr long SMALL_MAX_BYTES = 480L * 1024;

// Sliced-ELL copy of a scaled matrix for the runner (rows of A for A xbar,
// columns for A'y).  Rows sit in slots in order of decreasing length
// (perm[s] = the row in slot s), and slots go in slices of 64 (q = s / 64,
// lane s % 64); a row's entries, in storage order, go in chunks of
// small_chunk<V>() entries, chunk c of slot s at chunk index
// (off[q] + c) * 64 + lane, so that a wavefront loads one contiguous run
// per chunk; a slice is as many chunk rows wide as its longest row needs,
// and padding entries are 0 with index 0.  Values: val[chunk * C + g];
// 16-bit indices likewise.  C = 8 in FP32 (32-byte value and 16-byte index
// loads), 4 in FP64 (32-byte and 8-byte), within the register budget.
template <typename V>
constexpr int small_chunk() {
    return sizeof(V) == 4 ? 8 : 4;
}
template <typename V>
struct SmallEll {
    const int* perm = nullptr;  // rows: the row in each slot
    const int* off = nullptr;   // ceil(rows / 64) + 1 chunk-row offsets
    const unsigned short* idx = nullptr;
    const V* val = nullptr;
};

// Whether an LP (n columns, m rows, nnz nonzeros) fits the runner on the
// current device: x, y and the per-launch constants (bounds, gradient)
// within the device's LDS per workgroup, at most 3 columns and rows per
// thread, at most SMALL_MAX_BYTES of matrix per product.
bool small_fits(int n, int m, long nnz, bool f32);
// The slot order, chunk-row offsets of a CSR (CSC) pattern's ELL copy and,
// per entry position, the CSR (CSC) position it holds (-1: padding);
// returns the number of entry positions (at least one chunk row).
long small_ell_map(int rows, const int* ptr, int chunk, std::vector<int>& perm,
                   std::vector<int>& off, std::vector<int>& src);
// Fills an ELL copy's indices and values from the CSR (CSC) arrays.
hipError_t small_ell_fill(const int* src, long total, const int* idx, const double* val,
                          unsigned short* eidx, double* eval, hipStream_t s);
hipError_t small_ell_fill(const int* src, long total, const int* idx, const float* val,
                          unsigned short* eidx, float* eval, hipStream_t s);
// Attempts until the state freezes at a check boundary (or MAXIT), at most
// cap, in one launch; the buffers and the state are left as the
// three-kernel attempts leave them.  Requires small_fits, unsliced
// products and the ELL copies of A' (eT) and A (eA).  Returns the launch's
// error (the engine raises it).
hipError_t launch_small_run(const IterBufs<double>& b, const SmallEll<double>& eT,
                            const SmallEll<double>& eA, int cap, hipStream_t s);
hipError_t launch_small_run(const IterBufs<float>& b, const SmallEll<float>& eT,
                            const SmallEll<float>& eA, int cap, hipStream_t s);

}  // namespace igaos::dev
