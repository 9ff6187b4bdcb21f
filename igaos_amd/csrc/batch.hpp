#pragma once
// Batched persistent PDHG: thousands of independent LPs of uneven sizes
// solved in one launch (no reference counterpart; the reference solves one
// LP per call).  Each workgroup pulls LPs from a queue (largest first) and
// runs a whole solve for each: Ruiz scaling, power iteration, the adaptive
// line search, averaging, FP64 KKT checks every 64 iterations, restarts
// with primal-weight updates, incumbent and rollback -- the control flow
// of the single-LP engine (src/pdhg/pdhg.cu:457-1156) with per-instance
// state and termination.  The dense host dual repair (pdhg.cu:288-448) is
// not run in batched mode.

#include <hip/hip_runtime.h>

#include "pdhg_ctrl.hpp"

namespace igaos::batch {

constexpr int NT = 256;  // threads per workgroup (one LP at a time)

struct LPDesc {
    int m, n, nnz;
    int tpr;     // SpMV schedule: 0 = row-block streaming, else lanes per row
    int nblkA, nblkT;  // row blocks of A and A' (row-block streaming)
    long rbao, rbto;   // offsets of the instance's block tables (ints)
    long roff;   // offset into m-sized arrays
    long coff;   // offset into n-sized arrays
    long zoff;   // offset into nnz-sized arrays
    long rpo;    // offset into the concatenated row pointers (m+1 per LP)
    long cpo;    // offset into the concatenated column pointers (n+1 per LP)
    double obj_const;
};

struct LPResult {
    int status;
    int last_is_avg;
    double objective, pinf, dinf, gap, prow;
    long long iterations, attempts, restarts;
    double solve_ms;
    double eta, omega, lam2;
};

// Concatenated arrays of the batch (device pointers).  V is the iteration
// precision; in FP64 mode the V* data arrays alias the FP64 ones and the
// refinement arrays are unused.
template <typename V>
struct Arrays {
    const int *ap, *ai, *cp, *ci;     // local indices, per-LP pointer arrays
    const unsigned short *ai16, *ci16;  // 16-bit copies of ai, ci when every
                                        // instance has m, n <= 65536 (else null)
    // Gather-ordered copies for the lockstep attempt kernels (with ai16):
    // inside every multi-row block the nonzeros sorted by column (A) / row
    // (A'), so that a wavefront's gathers fall into few cache lines;
    // aslot / cslot give each one's position in the block's row order, and
    // faxs / facxs hold the values in the same order (filled per run).
    const unsigned short *ais16, *cis16, *aslot, *cslot;
    V *faxs, *facxs;
    // FP32: the same three streams packed into one 8-byte record per
    // nonzero, {column or row | slot << 16, value bits} (filled per run),
    // so the attempt kernels issue one load per nonzero instead of three
    uint2 *apk, *cpk;
    const int *rbA, *rbT;             // per-LP row-block tables (pairs)
    const double *ax, *acx;           // original values
    const double *c, *cl, *cu, *rmin, *rmax;
    double *dr, *dc, *sax, *sacx, *sc, *scl, *scu, *srmin, *srmax;
    V *fax, *facx;
    V *x0, *x1, *y0, *y1, *xbar, *ata, *xavg, *yavg, *axcur, *adx;
    double *xref, *yref, *aref, *atref;
    V *lo, *hi, *gref, *rlo, *rhi, *yoff;
    double *xb, *yb, *pesx, *pesy;
    double *xout, *yout;
};

struct Params {
    const LPDesc* lps;
    const int* order;     // queue order (largest first)
    int count;
    int* queue;           // zeroed counter
    LPResult* res;
    double tol;
    long long maxit;
    double time_limit_s;
    double ticks_per_s;   // wall-clock counter rate
    unsigned int* progress;  // optional, pinned host memory: LPs finished per workgroup
};

// `a` points to a device copy of the pointer table; `waves` (2 or 4) picks
// the register-allocation variant of the persistent kernel.
void launch_batch(const Arrays<float>* a, const Params& p, int grid,
                  int waves, hipStream_t s);
void launch_batch(const Arrays<double>* a, const Params& p, int grid,
                  int waves, hipStream_t s);
int batch_occupancy(bool f32, int waves);

// Lockstep mode (batch.hip): per-instance control blocks live in device
// memory (lockstep_ctl_bytes() each); the host keeps tables of the active
// instances' row blocks (one LkEnt per block, carrying the block's rows and
// nonzero range so that a workgroup reads its whole work description in
// one load) and one LsRange per active instance.  lockstep_round records
// ev[0..2] (if not null) before the primal kernel, between the two attempt
// kernels and after the dual kernel.
struct alignas(32) LkEnt {
    int lp;        // instance
    int r0, r1;    // rows [r0, r1) of the instance's matrix (A or A')
    int p0, p1;    // their nonzeros [p0, p1) (instance-local offsets)
    int pad[3];
};
struct LsRange {
    int lp;        // instance
    int bT0, nT;   // its blocks in the A' table
    int bA0, nA;   // its blocks in the A table
};
size_t lockstep_ctl_bytes();
void lockstep_setup(const Arrays<float>* a, const Params& p, void* ctl, int* done,
                    hipStream_t s);
void lockstep_setup(const Arrays<double>* a, const Params& p, void* ctl, int* done,
                    hipStream_t s);
// idx16: read the 16-bit gather-ordered copies (Arrays::ais16, cis16 with
// their slots and values; lockstep_sort_values fills the values).
void lockstep_sort_values(const Arrays<float>* a, const LPDesc* lps, const LkEnt* tblT,
                          int nbT, const LkEnt* tblA, int nbA, hipStream_t s);
void lockstep_sort_values(const Arrays<double>* a, const LPDesc* lps, const LkEnt* tblT,
                          int nbT, const LkEnt* tblA, int nbA, hipStream_t s);
void lockstep_round(const Arrays<float>* a, const LPDesc* lps, const LkEnt* tblT,
                    int nbT, const LkEnt* tblA, int nbA, const LsRange* rng, int nact,
                    void* ctl, double* partT, double* partA, int* nfrozen,
                    hipEvent_t* ev, bool idx16, hipStream_t s);
void lockstep_round(const Arrays<double>* a, const LPDesc* lps, const LkEnt* tblT,
                    int nbT, const LkEnt* tblA, int nbA, const LsRange* rng, int nact,
                    void* ctl, double* partT, double* partA, int* nfrozen,
                    hipEvent_t* ev, bool idx16, hipStream_t s);
// The check of every active instance; workgroup b takes active instance
// ord[b] (largest first, so the longest checks start first) and writes
// done[ord[b]].
void lockstep_check(const Arrays<float>* a, const Params& p, const LsRange* rng,
                    const int* ord, int nact, void* ctl, int* done, hipStream_t s);
void lockstep_check(const Arrays<double>* a, const Params& p, const LsRange* rng,
                    const int* ord, int nact, void* ctl, int* done, hipStream_t s);

}  // namespace igaos::batch
