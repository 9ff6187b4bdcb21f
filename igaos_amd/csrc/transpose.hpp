#pragma once
// CSR -> CSC on the device (transpose.hip).
#include <hip/hip_runtime.h>

namespace igaos::dev {

// cp (n + 1), ci and acx (nnz) receive the transpose of the CSR (ap, ai,
// ax), each column's entries in increasing row order; all device arrays.
// Allocates its scratch and synchronises the stream before returning.
// Throws std::runtime_error on a HIP error.
void csr_to_csc(int m, int n, long nnz, const int* ap, const int* ai, const double* ax,
                int* cp, int* ci, double* acx, hipStream_t s);

}  // namespace igaos::dev
