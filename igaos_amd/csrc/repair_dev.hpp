#pragma once
// Device dual repair (repair_dev.hip): the certification fallback of
// src/pdhg/pdhg.cu:288-448 for LPs of any size, on the original (unscaled)
// matrix resident on the GPU.
#include <hip/hip_runtime.h>

#include "model.hpp"
#include "pdhg_kernels.hpp"

namespace igaos::dev {

// The model in original units on the device (CSR, CSC with their row
// blocks, bounds, costs).
struct RepairData {
    const int *ap = nullptr, *ai = nullptr, *cp = nullptr, *ci = nullptr;
    const double *ax = nullptr, *acx = nullptr;
    RowBlocks rbA, rbT;
    const double *c = nullptr, *cl = nullptr, *cu = nullptr;
    const double *rmin = nullptr, *rmax = nullptr;
};

struct RepairInfo {
    double gap = 0.0;           // best relative gap found (inf if none finite)
    double dinf = 0.0;          // relative dual residual of the returned y
    double od = 0.0;            // its Lagrangian dual bound
    long iterations = 0;        // POCS iterations over both anchors
    long cg_iterations = 0;     // conjugate-gradient iterations in all
    long columns_in_J = 0;      // columns strictly inside their bounds
};

// Scratch of the repair (allocated on first use, 5n + 10m doubles).
struct RepairDev {
    int m = -1, n = -1;
    int max_cg = 200;       // CG iterations per affine projection
    double cg_tol = 1e-10;  // relative residual at which CG stops
    void* buf = nullptr;
    double *cn2 = nullptr, *d = nullptr, *t = nullptr, *w = nullptr, *bj = nullptr;
    double *lo = nullptr, *hi = nullptr, *y = nullptr, *ybest = nullptr, *yviol = nullptr;
    double *delta = nullptr, *r = nullptr, *pp = nullptr, *q = nullptr, *tmpm = nullptr;
    double *part = nullptr, *sc = nullptr;
    unsigned char *ct = nullptr, *jm = nullptr;

    RepairDev() = default;
    RepairDev(const RepairDev&) = delete;
    RepairDev& operator=(const RepairDev&) = delete;
    ~RepairDev() { release(); }
    void alloc(int m, int n);
    void release();
    // From the point x (n, original units, device) and the iterate's dual
    // y_it (m, device): y_out (m, device) receives the dual with the best
    // finite gap (returns true) or the least sign-violating one (false).
    // op = c'x.  Synchronises the stream.  Throws std::runtime_error on a
    // HIP error.
    bool run(const RepairData& D, const double* x, const double* y_it, double col_tol,
             double op, double cmax, hipStream_t s, double* y_out, RepairInfo* info);

  private:
    void evaluate(const RepairData& D, const double* x, const double* yv, double cmax,
                  hipStream_t s, double out[5]);
};

// Uploads the model (original units) and repairs from host x (n), y_it
// (m); y_out (m, host) as in RepairDev::run.
bool dual_repair_device(const io::Model& md, const double* x, const double* y_it,
                        double col_tol, double* y_out, RepairInfo* info);

}  // namespace igaos::dev
