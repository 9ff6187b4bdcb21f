#pragma once
// LP model container and MPS reader of the pdhg engine slot.
//
// io::Model has the same fields and meaning as the reference's
// src/io/model.hpp:16-37: min c'x + obj_const s.t. rmin <= Ax <= rmax,
// cl <= x <= cu, with A stored both as CSR (ap/ai/ax) and CSC (cp/ci/acx),
// FP64 values, int32 indices and +-inf for missing bounds.  The engine
// borrows it read-only.

#include <string>
#include <vector>

namespace igaos::io {

struct Model {
    int m = 0, n = 0;
    std::vector<double> rmin, rmax;
    std::vector<double> cl, cu;
    std::vector<double> c;
    std::vector<int> ap, ai;
    std::vector<double> ax;
    std::vector<int> cp, ci;
    std::vector<double> acx;
    std::vector<unsigned char> integ;
    double obj_const = 0.0;
    int n_fr_parsed = 0;
    int n_ranges_parsed = 0;
    std::vector<int> q_i, q_j;
    std::vector<double> q_v;

    bool has_quadratic() const { return !q_v.empty(); }
    int nnz() const { return static_cast<int>(ax.size()); }
};

struct Triplet {
    int row, col;
    double val;
};

// Triplets -> CSR + CSC by two stable bucket sorts: entries keep their
// input order inside each row and each column (src/io/model.hpp:39-67).
void counting_sort(const std::vector<Triplet>& trips, int rows, int cols,
                   std::vector<int>& ap, std::vector<int>& ai,
                   std::vector<double>& ax, std::vector<int>& cp,
                   std::vector<int>& ci, std::vector<double>& acx);

// Free-format MPS reader with the reference's section semantics
// (src/io/model.hpp:69-218).  Throws std::runtime_error on I/O errors and
// std::out_of_range for a BOUNDS entry naming an unknown column.
void parse_mps(const std::string& path, Model& model);

// Row activity A x of a solution, in FP64: each row summed from 0 in CSR
// order, the computation of the reference simplex (src/simplex/
// simplex.cpp:1727-1730).  out has m entries.
void row_activity(const Model& model, const double* x, double* out);

// Empty when (ptr, idx) is a compressed pattern of `rows` rows over `cols`
// columns (ptr[0] = 0, ptr non-decreasing, every index in [0, cols)), else
// what is wrong with it.  The kernels index vectors by these arrays
// unchecked, so every array that comes in from a caller passes here first.
std::string check_compressed(int rows, int cols, const int* ptr, const int* idx);

// check_compressed of the CSR (ap, ai) and of the CSC (cp, ci) of an m x n
// matrix, and that both hold the same number of entries in every row and
// in every column.
std::string check_csr_csc(int m, int n, const int* ap, const int* ai, const int* cp,
                          const int* ci);
// check_csr_csc, and the CSC holds exactly the CSR's entries: per column
// the same (row, value) pairs, in any order (O(nnz); a column listed in
// row order is compared directly).
std::string check_transpose(int m, int n, const int* ap, const int* ai, const double* ax,
                            const int* cp, const int* ci, const double* acx);

}  // namespace igaos::io
