// CSR -> CSC on the device: the transpose a model needs for the A'y
// products, built from an uploaded CSR instead of by the host's bucket
// sort (counting_sort, src/io/model.hpp:39-67).
//
// A stable radix sort of the nonzeros' column indices (hipCUB, LSD radix
// sort, stable) with their CSR positions as values puts each column's
// entries in CSR order, i.e. in increasing row order -- the order
// counting_sort produces whenever the triplets arrive row by row.  Row
// indices come from expanding the row pointer (one thread per row), values
// are gathered by the sorted positions, and the column pointer is the
// exclusive scan of the column counts (integer atomics: exact in any
// order).  Every step is deterministic.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "transpose.hpp"

namespace igaos::dev {

namespace {

constexpr int TB = 256;

#define TCK(expr)                                                                \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess)                                                    \
            throw std::runtime_error(std::string("HIP error in CSR transpose: ") + \
                                     hipGetErrorString(e_) + " at " #expr);      \
    } while (0)

__global__ void k_iota_rows(int m, const int* ap, int* row_of, int* pos) {
    const long i = static_cast<long>(blockIdx.x) * TB + threadIdx.x;
    if (i >= m) return;
    for (int p = ap[i]; p < ap[i + 1]; ++p) {
        row_of[p] = static_cast<int>(i);
        pos[p] = p;
    }
}

__global__ void k_count_cols(long nnz, const int* ai, int* cnt) {
    const long p = static_cast<long>(blockIdx.x) * TB + threadIdx.x;
    if (p < nnz) atomicAdd(&cnt[ai[p]], 1);
}

__global__ void k_gather_csc(long nnz, const int* sorted_pos, const int* row_of,
                             const double* ax, int* ci, double* acx) {
    const long q = static_cast<long>(blockIdx.x) * TB + threadIdx.x;
    if (q >= nnz) return;
    const int p = sorted_pos[q];
    ci[q] = row_of[p];
    acx[q] = ax[p];
}

int grid(long k) { return static_cast<int>((k + TB - 1) / TB); }

struct Tmp {  // device scratch, freed on every exit
    void* p = nullptr;
    ~Tmp() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

void csr_to_csc(int m, int n, long nnz, const int* ap, const int* ai, const double* ax,
                int* cp, int* ci, double* acx, hipStream_t s) {
    if (m < 0 || n < 0 || nnz < 0) throw std::invalid_argument("negative size");
    // column counts -> exclusive scan into cp[0..n]
    TCK(hipMemsetAsync(cp, 0, sizeof(int) * (static_cast<size_t>(n) + 1), s));
    if (nnz > 0) k_count_cols<<<grid(nnz), TB, 0, s>>>(nnz, ai, cp);
    size_t scan_bytes = 0;
    TCK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, cp, cp, n + 1, s));
    if (nnz == 0) {
        Tmp t;
        TCK(hipMalloc(&t.p, scan_bytes > 0 ? scan_bytes : 1));
        TCK(hipcub::DeviceScan::ExclusiveSum(t.p, scan_bytes, cp, cp, n + 1, s));
        TCK(hipStreamSynchronize(s));
        return;
    }
    // stable sort of (column, CSR position) pairs by column
    int key_bits = 1;
    while (key_bits < 31 && (1L << key_bits) < n) ++key_bits;
    size_t sort_bytes = 0;
    TCK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, ai, ci,
                                           static_cast<const int*>(nullptr),
                                           static_cast<int*>(nullptr), static_cast<int>(nnz),
                                           0, key_bits, s));
    const size_t work = std::max(scan_bytes, sort_bytes);
    // scratch: row of each CSR entry, positions in, positions out, cub work
    Tmp t;
    const size_t ints = static_cast<size_t>(nnz);
    const size_t off = (3 * ints * sizeof(int) + 255) & ~static_cast<size_t>(255);
    TCK(hipMalloc(&t.p, off + work));
    int* row_of = static_cast<int*>(t.p);
    int* pos_in = row_of + ints;
    int* pos_out = pos_in + ints;
    void* cub = static_cast<char*>(t.p) + off;  // 256-byte aligned
    TCK(hipcub::DeviceScan::ExclusiveSum(cub, scan_bytes, cp, cp, n + 1, s));
    if (m > 0) k_iota_rows<<<grid(m), TB, 0, s>>>(m, ap, row_of, pos_in);
    // the sorted keys (columns) are not needed: they land in ci and are
    // overwritten by the row indices below
    TCK(hipcub::DeviceRadixSort::SortPairs(cub, sort_bytes, ai, ci, pos_in, pos_out,
                                           static_cast<int>(nnz), 0, key_bits, s));
    k_gather_csc<<<grid(nnz), TB, 0, s>>>(nnz, pos_out, row_of, ax, ci, acx);
    TCK(hipGetLastError());
    TCK(hipStreamSynchronize(s));  // the scratch is freed on return
}

}  // namespace igaos::dev
