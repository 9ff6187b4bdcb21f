#pragma once
// The pdhg engine slot: igaos::pdhg::solve, same signature and semantics
// as the reference's src/pdhg/pdhg.hpp:9 (dispatched by
// solve_with_engine, src/api/engine.hpp:32-38), backed by HIP kernels on
// one MI355X.  solve_ex adds engine-private knobs without touching Options.

#include <memory>

#include "igaos/options.h"
#include "igaos/solution.h"
#include "model.hpp"

namespace igaos::pdhg {

struct EngineConfig {
    int precision = 64;      // 64: plain FP64 engine; 32: FP32 inner + FP64 refinement
    int graph_slots = 16;    // attempts captured per hipGraph replay (0 = no graph)
    int max_repair_m = 4096; // dual repair: host dense (pdhg.cu's) up to this many rows,
                             // device CG repair above
    int device = 0;
    int column_slices = -1;  // column slices of A / A' products: -1 auto, 0 off, 1..16 forced
    int xcd_pinned = 0;      // column slices: 1 = all in one launch, each XCD on its own slices
    int scaling = 0;         // 0: Ruiz (the reference); 1: Ruiz then Pock-Chambolle (alpha 1)
    int exact = 0;           // 1: exact FP64 mode (reference summation orders, no FMA;
                             // precision 64, iteration-bounded runs only)
    int small_runner = -1;   // small LPs: one workgroup runs a check period's attempts per
                             // launch (-1 or 1: wherever the LP fits, 0: never); exact mode:
                             // exact.hip's runner, default mode: small.hip's

    int stream_loads = -1;   // matrix streams of the attempt kernels: -1 nontemporal when
                             // both matrices exceed the last-level cache, 0 cached, 1 nontemporal
};

struct SolveInfo {
    double setup_ms = 0.0;     // upload, scaling, power iteration
    double solve_ms = 0.0;     // iteration loop (the reference's solve_time_ms)
    long attempts = 0;         // primal/dual attempts incl. line-search retries
    long restarts = 0;
    long checks = 0;
    long refinements = 0;      // FP64 re-anchors of the FP32 correction problem
    double eta = 0.0, omega = 0.0, lam2 = 0.0, eta0 = 0.0, omega0 = 0.0;
    int last_is_avg = 0;
    int used_incumbent = 0;
    int repaired = 0;
    double kernel_ms = 0.0;    // device time of the attempt kernels (events; with the
                               // small-LP runner's check graph, its check work too)
    long rollbacks = 0;        // incumbent rollbacks (pdhg.cu:965-988)
};

// The reference's entry point (src/pdhg/pdhg.hpp:9).  Besides the
// reference's fields it fills row_activity = A x as the reference simplex
// does (io::row_activity); solve_ex leaves that to its caller.
Solution solve(const io::Model& model, const Options& options);

// A prepared engine: construction uploads the model and runs the setup
// (Ruiz scaling, power iteration, graph capture) once; every run() is a
// fresh solve from x = y = 0 on the HBM-resident data.  The model must
// outlive the engine.
struct AttemptTiming;
class EngineBase {
  public:
    virtual ~EngineBase() = default;
    virtual Solution run(const Options& options, SolveInfo* info) = 0;
    virtual AttemptTiming profile(int attempts) = 0;
    virtual double setup_ms() const = 0;
    // column slices of the primal-step (A'y) and dual-step (A xbar) products
    virtual void column_slices(int& primal, int& dual) const = 0;
};
// Throws std::runtime_error when no HIP device is visible or a HIP call
// fails, std::invalid_argument for a malformed model.
std::unique_ptr<EngineBase> make_engine(const io::Model& model,
                                        const EngineConfig& cfg);
Solution solve_ex(const io::Model& model, const Options& options,
                  const EngineConfig& cfg, SolveInfo* info);

// Times `attempts` line-search attempts (k_primal, k_dual) with HIP events
// on the engine's stream after the usual setup; returns the mean device
// milliseconds per launch of each kernel.  Checks and restarts are off.
struct AttemptTiming {
    double ms_primal = 0.0, ms_dual = 0.0;
    int nblk_primal = 0, nblk_dual = 0;
};
AttemptTiming profile_attempts(const io::Model& model, const EngineConfig& cfg,
                               int attempts);

// Host-side KKT evaluation and dense dual repair in original units
// (src/pdhg/pdhg.cu:173-254, 288-448); used at check time only.
struct HostMetrics {
    double pinf, dinf, gap, op, od, prow;
};
HostMetrics host_evaluate(const io::Model& lp, const double* x,
                          const double* y, const double* Ax, const double* d);
bool host_dual_repair(const io::Model& lp, const double* x, const double* Ax,
                      const double* y_it, double col_tol, double* y);

// PDHG -> simplex crossover basis (crossover.cpp); see igaos_crossover_basis.
int crossover_basis(const io::Model& md, const double* x, const double* y, double tol,
                    int* basis, unsigned char* nb_state, bool fill);

}  // namespace igaos::pdhg
