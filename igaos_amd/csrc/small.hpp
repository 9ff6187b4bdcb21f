#pragma once
// One-workgroup runner of the default mode for small LPs (small.hip): a
// check period of line-search attempts per launch, the LP's vectors in
// registers and LDS, its matrices streamed from L2.

#include <hip/hip_runtime.h>

#include "pdhg_kernels.hpp"

namespace igaos::dev {

constexpr int SMALL_MAX_DIM = 3 * 1024;  // max(n, m): up to 3 columns / rows a thread

// Whether an LP (n columns, m rows, nnz nonzeros) fits the runner on the
// current device: its products, x and y within the device's LDS per
// workgroup (160 KB on gfx950) and at most 3 columns and rows per thread.
bool small_fits(int n, int m, long nnz, bool f32);
// Attempts until the state freezes at a check boundary (or MAXIT), at most
// cap, in one launch; the buffers and the state are left as the
// three-kernel attempts leave them.  Requires small_fits and unsliced
// products.  Returns the launch's error (the engine raises it).
hipError_t launch_small_run(const IterBufs<double>& b, long nnz, int cap, hipStream_t s);
hipError_t launch_small_run(const IterBufs<float>& b, long nnz, int cap, hipStream_t s);

}  // namespace igaos::dev
