/* C-ABI of the MI355X PDHG engine (libigaos_pdhg.so).
 *
 * This is the drop-in boundary for the reference's `pdhg` engine slot.
 * Every entry point takes plain pointers and sizes; none takes or returns
 * a C++ or torch type.  Functions return 0 on success and a negative code
 * on failure (igaos_last_error() then describes it) unless stated.
 * Pointers documented as "device" must be HIP device allocations on the
 * current device; "host" pointers are ordinary host memory.
 */
#ifndef IGAOS_PDHG_H
#define IGAOS_PDHG_H

#ifdef __cplusplus
extern "C" {
#endif

#define IGAOS_PDHG_ABI_VERSION 4

/* Status codes: the order of igaos::Status (include/igaos/status.h:5-14). */
enum {
    IGAOS_OPTIMAL = 0,
    IGAOS_NEAR_OPTIMAL = 1,
    IGAOS_FEASIBLE = 2,
    IGAOS_INFEASIBLE = 3,
    IGAOS_UNBOUNDED = 4,
    IGAOS_ITERATION_LIMIT = 5,
    IGAOS_TIME_LIMIT = 6,
    IGAOS_ERROR = 7
};

int igaos_pdhg_abi_version(void);
/* Number of visible HIP devices (0 when no GPU is present). */
int igaos_pdhg_device_count(void);
/* Text of the last error raised on this thread ("" if none). */
const char* igaos_last_error(void);
/* igaos::status_name (include/igaos/status.h:16-28). */
const char* igaos_status_name(int status);

/* ---- model: the reference's io::Model (src/io/model.hpp:16-37) ------- */
typedef struct igaos_model igaos_model;

/* Replaces igaos::io::parse_mps (src/io/model.hpp:69-218).  Returns NULL
 * on failure. */
igaos_model* igaos_model_read_mps(const char* path);
/* Builds a model from triplets in the given order; CSR and CSC keep that
 * order inside each row / column like counting_sort (model.hpp:39-67). */
igaos_model* igaos_model_from_triplets(int m, int n, int nnz,
                                       const int* rows, const int* cols,
                                       const double* vals, const double* c,
                                       const double* cl, const double* cu,
                                       const double* rmin, const double* rmax,
                                       double obj_const);
/* Builds a model from ready CSR + CSC arrays (copied).  Returns NULL when
 * a pointer array does not start at 0 or decreases, an index is out of
 * range, or the CSC is not the CSR's transpose: every column must hold the
 * same (row, value) pairs in both, in any order (checked in O(nnz)). */
igaos_model* igaos_model_from_csr_csc(int m, int n, const double* c,
                                      const double* cl, const double* cu,
                                      const double* rmin, const double* rmax,
                                      const int* ap, const int* ai,
                                      const double* ax, const int* cp,
                                      const int* ci, const double* acx,
                                      double obj_const);
/* Builds a model from CSR arrays alone (copied): the CSC is built on the
 * GPU (igaos_csr_to_csc_f64) instead of by the host's bucket sort
 * (counting_sort, src/io/model.hpp:39-67), with each column's entries in
 * row order -- counting_sort's order whenever triplets arrive row by row.
 * Needs a visible HIP device.  Returns NULL on a malformed CSR or a HIP
 * error (igaos_last_error). */
igaos_model* igaos_model_from_csr(int m, int n, const double* c, const double* cl,
                                  const double* cu, const double* rmin,
                                  const double* rmax, const int* ap, const int* ai,
                                  const double* ax, double obj_const);
void igaos_model_free(igaos_model* model);
/* Host: row activity A x (m entries) of a point x (n entries), each row
 * summed in FP64 from 0 in CSR order -- how the reference simplex fills
 * Solution::row_activity (src/simplex/simplex.cpp:1727-1730).  The
 * reference's pdhg leaves that field empty; igaos::pdhg::solve here fills
 * it this way.  0 on success. */
int igaos_model_row_activity(const igaos_model* model, const double* x, double* out);
/* Sizes and MPS counters (any output may be NULL). */
void igaos_model_dims(const igaos_model* model, int* m, int* n, int* nnz,
                      double* obj_const, int* n_fr_parsed,
                      int* n_ranges_parsed, int* n_int);
/* Copies the model's arrays out (any output may be NULL). */
void igaos_model_get(const igaos_model* model, double* rmin, double* rmax,
                     double* cl, double* cu, double* c, int* ap, int* ai,
                     double* ax, int* cp, int* ci, double* acx);

/* ---- solve: igaos::pdhg::solve (src/pdhg/pdhg.hpp:9) ------------------- */
/* Mirrors igaos::Options (include/igaos/options.h:5-13). */
typedef struct {
    double time_limit_s;
    double tolerance;
    long max_iterations;
    int seed;
    int presolve;
    int restarts;
    int verbosity;
} igaos_options;
void igaos_options_default(igaos_options* opt);

/* Engine-private knobs (no reference counterpart). */
typedef struct {
    int precision;    /* 64: plain FP64; 32: FP32 inner + FP64 refinement */
    int graph_slots;  /* attempts per hipGraph replay; 0 = direct launches */
    int max_repair_m; /* dual repair (pdhg.cu:288-448, schedule 854-931):
                         the reference's dense host repair for m <= this,
                         the device repair (conjugate gradients on
                         A_J A_J' + ridge, simultaneous halfspace steps)
                         above it */
    int device;
    int column_slices; /* column slices of the A / A' products (gather
                          locality on large LPs): -1 auto, 0 off, 1..16 */
    int xcd_pinned;    /* column slices (above): 1 = all slices in one
                          launch, each of the 8 XCDs taking only blocks of
                          its own slices (its L2 then holds only their range
                          of the gathered vector); 0 (default) = one launch
                          per slice over all XCDs, which measured faster on
                          the 1M-column LP at every slice count but 16 */
    int scaling;       /* 0 (default): Ruiz equilibration, 10 sweeps, as the
                          reference (pdhg.cu:469-492); 1: Ruiz, then
                          Pock-Chambolle diagonal preconditioning with
                          alpha = 1 (the reference's algorithm sheet,
                          docs/research/pdlp-algorithm-sheet.md:62) */
    int exact;         /* 1: exact FP64 mode -- every sparse row summed in
                          storage order, the line-search dots in the
                          reference's two-stage layout (pdhg.cu:63-125),
                          norms and the KKT sums in index order, no fused
                          multiply-add, k^-0.3 / k^-0.6 from the host's pow:
                          a run is bit-reproducible against a sequential
                          FP64 host replay of pdhg.cu.  Precision 64 only,
                          iteration-bounded runs only (time_limit_s >= 1e6
                          or <= 0, max_iterations <= 2^21); far slower than
                          the default kernels.  0: default. */
    int small_runner;  /* small LPs: one workgroup of 1024 threads runs the
                          attempts of a whole check period in one launch.
                          Exact mode (FP64 LPs with max(n, m) <= 16384 and
                          nnz <= 15360): the exact runner, bit-identical to
                          the exact mode's six kernels per attempt.  Default
                          mode (max(n, m) <= 3072, x, y and the bounds
                          within the LDS of a workgroup, nnz * (value + 2
                          bytes) <= 480 KiB): the register / LDS runner,
                          each column's and row's state in registers,
                          length-sorted sliced-ELL copies of the matrices
                          streamed from L2, rows summed in storage order
                          (2.0x the three kernels per attempt on the
                          1,000 x 3,000 LP in FP64, 2.3x in FP32).  -1 (default) and 1: wherever
                          the LP fits; 0: never. */
    int stream_loads;  /* how the attempt kernels load the matrix streams
                          (indices, values): -1 (default) nontemporal when
                          both matrices together exceed the 256 MB
                          last-level cache, else cached; 0: cached; 1:
                          nontemporal.  Results are identical either way. */
} igaos_engine_config;
void igaos_engine_config_default(igaos_engine_config* cfg);

/* Solution fields of igaos::Solution (include/igaos/solution.h:10-22)
 * plus engine diagnostics. */
typedef struct {
    int status;
    double objective, pinf, dinf, rel_gap;
    long iterations;
    double solve_time_ms;
    double setup_ms;
    long attempts, restarts, checks, refinements;
    double eta, omega, lam2, eta0, omega0;
    int last_is_avg, used_incumbent, repaired;
    double kernel_ms; /* device time of the attempts (events); with the
                         small-LP runner's check graph, of the check's
                         device work too */
    long rollbacks; /* incumbent rollbacks (pdhg.cu:965-988) */
} igaos_solve_info;

/* Replaces igaos::pdhg::solve (src/pdhg/pdhg.hpp:9, pdhg.cu:457-1156).
 * x (host, n) and y (host, m) receive the solution in original units with
 * the dual sign convention d = c + A'y; msg (host, msglen bytes) receives
 * Solution::message.  cfg may be NULL (plain FP64 defaults).  Like the
 * reference it never fails loudly for solver outcomes: the status field
 * carries error / iteration-limit / time-limit; the return value is
 * negative only for invalid arguments. */
int igaos_pdhg_solve(const igaos_model* model, const igaos_options* opt,
                     const igaos_engine_config* cfg, double* x, double* y,
                     igaos_solve_info* info, char* msg, int msglen);

/* Prepared engine (no reference counterpart; the reference re-uploads and
 * re-scales on every call).  igaos_pdhg_create uploads the model and runs
 * the setup of pdhg.cu:467-663 (Ruiz, scaling, power iteration) once;
 * every igaos_pdhg_run is then a fresh solve from x = y = 0 on the
 * HBM-resident data with the same results as igaos_pdhg_solve.  The model
 * must outlive the engine.  create returns NULL on failure. */
typedef struct igaos_engine igaos_engine;
igaos_engine* igaos_pdhg_create(const igaos_model* model,
                                const igaos_engine_config* cfg);
int igaos_pdhg_run(igaos_engine* engine, const igaos_options* opt, double* x,
                   double* y, igaos_solve_info* info, char* msg, int msglen);
/* Setup time of the engine in milliseconds. */
double igaos_pdhg_setup_ms(const igaos_engine* engine);
void igaos_pdhg_destroy(igaos_engine* engine);
/* Timing (no reference counterpart): runs `attempts` line-search attempts
 * without checks and reports the mean device time per launch of the two
 * attempt kernels, measured with HIP events on the engine's stream, and
 * their grid sizes.  A later igaos_pdhg_run starts afresh. */
int igaos_pdhg_profile_attempts(igaos_engine* engine, int attempts,
                                double* ms_primal, double* ms_dual,
                                int* blocks_primal, int* blocks_dual);
/* Column slices the engine uses for the A' y product of the primal step and
 * the A xbar product of the dual step (0: fused row-block kernels). */
int igaos_pdhg_column_slices(const igaos_engine* engine, int* slices_primal,
                             int* slices_dual);

/* ---- batched persistent solve (no reference counterpart) ----------------
 * Thousands of independent LPs solved in one persistent launch: each
 * workgroup takes instances from a queue (largest first) and runs the whole
 * PDHG solve of pdhg.cu:457-1156 for each, with per-instance line search,
 * restarts, termination and results (the dense dual repair is not run).
 * Host arrays are the instances' arrays concatenated in order: c, cl, cu
 * by n_i; rmin, rmax by m_i; ap (m_i + 1 entries per instance) and cp
 * (n_i + 1) hold instance-local offsets; ai, ci instance-local indices. */
typedef struct igaos_batch igaos_batch;
/* Per-instance result; status uses the IGAOS_* codes. */
typedef struct {
    int status;
    int last_is_avg;
    double objective, pinf, dinf, rel_gap, prow;
    long long iterations, attempts, restarts;
    double solve_ms;
    double eta, omega, lam2;
} igaos_batch_result;
/* precision: 32 = FP32 iterations + FP64 refinement, 64 = plain FP64.
 * Uploads the batch; returns NULL on failure. */
igaos_batch* igaos_batch_create(int count, const int* m, const int* n,
                                const int* nnz, const double* c,
                                const double* cl, const double* cu,
                                const double* rmin, const double* rmax,
                                const int* ap, const int* ai,
                                const double* ax, const int* cp,
                                const int* ci, const double* acx,
                                const double* obj_const, int precision);
/* Solves every instance from x = y = 0 with the options' tolerance,
 * max_iterations / time_limit_s (per instance, as in igaos_pdhg_solve).
 * x (host, sum n_i), y (host, sum m_i) and res (count) may be NULL. */
int igaos_batch_run(igaos_batch* batch, const igaos_options* opt, double* x,
                    double* y, igaos_batch_result* res);
/* The same as launch + wait: launch returns once the persistent kernel is
 * enqueued; progress returns how many instances have finished so far (read
 * from pinned host memory, no synchronisation); wait blocks and copies. */
int igaos_batch_launch(igaos_batch* batch, const igaos_options* opt);
int igaos_batch_progress(const igaos_batch* batch);
int igaos_batch_wait(igaos_batch* batch, double* x, double* y,
                     igaos_batch_result* res);
/* Schedule: SpMV mode of each instance (0 = row-block streaming through
 * LDS, the default; 1, 2, 4 or 8 = lanes per row; NULL keeps the current
 * ones), the persistent grid size (<= 0 keeps it) and the
 * kernel variant by resident waves per SIMD (2 or 4; 0 keeps it; setting
 * it resets the grid to the variant's full occupancy unless grid > 0
 * follows in a later call). */
int igaos_batch_set_schedule(igaos_batch* batch, const int* tpr, int grid,
                             int waves);
int igaos_batch_get_schedule(const igaos_batch* batch, int* tpr, int* grid,
                             int* waves);
/* Execution mode of later runs: 0 = one persistent launch, each workgroup
 * solving whole instances taken from a queue (the default); 1 = lockstep
 * rounds: every active instance advances one line-search attempt per
 * round, each round being three launches over the row blocks of all active
 * instances, with a per-instance check (one workgroup each) every 64
 * accepted iterations and finished instances dropped from the next
 * rounds.  Both modes run the same per-instance algorithm and controller;
 * their SpMV summation orders differ.  In mode 1 launch starts a host
 * thread that drives the rounds; progress and wait work as in mode 0. */
int igaos_batch_set_mode(igaos_batch* batch, int mode);
int igaos_batch_get_mode(const igaos_batch* batch);
/* Lockstep timing of the last run: rounds 0, 16, 32 and 48 of every check
 * period (all active instances work in them) are timed with HIP events
 * around the primal and the dual attempt kernel; returns the summed
 * milliseconds and algorithmic bytes (k_primal: nnz (w_a + 4) + 4 (n + 1)
 * + 7 n w, k_dual: nnz (w_a + 4) + 4 (m + 1) + 8 m w, summed over the
 * active instances) and the number of sampled rounds. */
int igaos_batch_lockstep_timing(const igaos_batch* batch, double* primal_ms,
                                double* primal_bytes, double* dual_ms,
                                double* dual_bytes, long* samples);
/* Device time of the last run (HIP events around all its launches), ms. */
double igaos_batch_kernel_ms(const igaos_batch* batch);
/* Replaces the batch's numeric data in place (same sparsity pattern and
 * sizes; host arrays in the layout of igaos_batch_create, NULL keeps an
 * array).  ax and acx must hold the same values in CSR and CSC order.
 * The next run re-scales from the new data. */
int igaos_batch_update(igaos_batch* batch, const double* c, const double* cl,
                       const double* cu, const double* rmin,
                       const double* rmax, const double* ax,
                       const double* acx);

/* ---- adjoint of the batched solve (no reference counterpart) -----------
 * Gradients of a scalar loss L(x*, y*, obj) through the solution map of
 * every instance, at the solution of the last finished run.  With the
 * active sets B (columns strictly inside their bounds) and R (equality
 * rows and inequality rows on a bound) read from complementarity, it
 * solves  A_RB p = gy_R  and  A_RB' q = (gx + gobj c)_B  per instance by
 * CGLS in FP64 (the transposed fixed-point system of the PDHG map
 * linearised at the solution) and returns, in original units:
 *   gc    (sum n): dL/dc = gobj x - [B] p
 *   grmin, grmax (sum m): dL/d(bound row i sits on) = q_i, on the side the
 *         sign of y_i names (y_i >= 0: rmax), 0 on the other side
 *   gcl, gcu (sum n): dL/d(bound column j sits on) for j not in B
 *   gax   (sum nnz, CSR order): dL/dA_ij = -[R](p_j y_i + q_i x_j)
 * gx (sum n), gy (sum m) and gobj (count) are the upstream gradients (NULL
 * = zero); every output may be NULL.  CGLS stops when the normal-equation
 * residual falls below `tolerance` relative to its start, or after
 * max_iterations iterations. */
typedef struct {
    int basic_cols, active_rows;  /* |B|, |R| */
    int iters_p, iters_q;         /* CGLS iterations of the two systems */
    int accepted;                 /* polish: the polished pair was kept */
    int flags;                    /* polish: failed checks, bit 0 column
                                     bounds, 1 rows, 2 active-row dual
                                     signs, 3 d_B = 0, 4 reduced-cost signs */
    double res_p, res_q;          /* final relative residuals */
    double viol;                  /* polish: largest relative violation */
} igaos_adjoint_info;
/* Active-set polishing of the last run's solutions (no reference
 * counterpart): with the same masks, x_N is put on its bounds, y off R to
 * zero, and the corrections of x_B and y_R solving A_RB x_B = b_R - A_RN x_N
 * and A_RB' y_R = -c_B are found by CGLS (FP64); the polished pair replaces
 * the instance's solution (also for a later igaos_batch_adjoint) only if
 * it is primal feasible, has d_B = 0 and correct reduced-cost and dual
 * signs, all to 1e-9 relative (info.accepted).  x, y, objective (count)
 * receive the resulting solutions; any may be NULL. */
/* Replaces the solutions of the last run (host x: sum n, y: sum m; NULL
 * keeps one) that igaos_batch_polish and igaos_batch_adjoint start from,
 * e.g. with solutions of the same data obtained elsewhere. */
int igaos_batch_set_solution(igaos_batch* batch, const double* x, const double* y);
int igaos_batch_polish(igaos_batch* batch, double tolerance, int max_iterations,
                       double* x, double* y, double* objective,
                       igaos_adjoint_info* info);
int igaos_batch_adjoint(igaos_batch* batch, const double* gx, const double* gy,
                        const double* gobj, double tolerance,
                        int max_iterations, double* gc, double* grmin,
                        double* grmax, double* gcl, double* gcu, double* gax,
                        igaos_adjoint_info* info);
/* Unrolled differentiation (no reference counterpart): `iterations` steps
 * of fixed-step PDHG on every instance of the current data, in FP64 and
 * original units, from x = y = 0 (the step of src/pdhg/pdhg.cu:32-61
 * without line search, restarts or averaging).  The step sizes are
 * `steps` (host, tau of every instance then sigma of every instance) or,
 * when it is NULL, tau = 0.9 / (omega |A|), sigma = 0.9 omega / |A| from
 * the data (|A| from 30 power iterations, omega the engine's primal weight
 * rule in original units); either way they are constants of the map (no
 * gradient flows through them).  The iterates are checkpointed every 64
 * steps on the device.  Outputs (host; any may be NULL): x (sum n), y
 * (sum m) after the last step, objective c'x + obj_const, and the tau and
 * sigma used (count each). */
int igaos_batch_unroll(igaos_batch* batch, int iterations, const double* steps,
                       double* x, double* y, double* objective, double* tau,
                       double* sigma);
/* Reverse mode through the last igaos_batch_unroll on the current data,
 * recomputing each 64-step segment from its checkpoint: gradients of
 * L(x_K, y_K, objective) for upstream gx (sum n), gy (sum m), gobj (count),
 * NULL = zero.  Outputs as for igaos_batch_adjoint (gax in CSR order); a
 * clamp at a tie passes the gradient to the bound (the lower one first). */
int igaos_batch_unroll_backward(igaos_batch* batch, const double* gx, const double* gy,
                                const double* gobj, double* gc, double* grmin,
                                double* grmax, double* gcl, double* gcu, double* gax);
const char* igaos_batch_last_error(void);
/* Host: the gather order the lockstep kernels read (no reference
 * counterpart): over the row blocks of a CSR matrix with at most 65,536
 * columns, within every block of more than one row the nonzeros sorted by
 * column (stable); sidx[q] is the column of the q-th nonzero in that order
 * (16-bit) and slot[q] its position within the block's row order; blocks
 * of one row keep their order.  sidx, slot: nnz each.  0 on success; -1
 * (igaos_batch_last_error) for a malformed pattern or more columns. */
int igaos_gather_order_build(int nrows, int ncols, const int* ptr, const int* idx,
                             unsigned short* sidx, unsigned short* slot);
void igaos_batch_destroy(igaos_batch* batch);

/* ---- kernel layer (device pointers, HIP stream or NULL) ----------------- */
/* Host: the KKT evaluation of src/pdhg/pdhg.cu:173-254 in original units
 * that the dual repair runs (exported so that host tests can pin it): for
 * x (n), y (m), Ax (m) and d = c + A'y (n), out receives pinf, dinf, gap,
 * primal objective, dual objective (without obj_const) and the largest
 * relative row violation.  0 on success. */
int igaos_kkt_evaluate(const igaos_model* model, const double* x, const double* y,
                       const double* Ax, const double* d, double* out);

/* Host: the check-time dual repair of src/pdhg/pdhg.cu:288-448 that the
 * engine runs when only the duality-gap certificate lags (exported so that
 * host tests can pin it): from the point x (n) with activity Ax (m) and the
 * iterate's dual y_it (m), POCS projections onto {d_J = 0 for columns
 * strictly inside their bounds (tolerance col_tol)} and the sign cones,
 * with a dense Cholesky of A_J A_J' + ridge.  y (m) receives the dual with
 * the best finite gap (returns 1), or the least-violating projection when
 * none has one (returns 0); -1 for invalid arguments. */
int igaos_dual_repair(const igaos_model* model, const double* x, const double* Ax,
                      const double* y_it, double col_tol, double* y);

/* Active-set polishing of a single LP's solution (no reference
 * counterpart; the batched engine's igaos_batch_polish on a batch of one):
 * from x (n), y (m) in original units it reads the active sets off
 * complementarity, puts x_N on its bounds and y off the active rows to
 * zero, and solves A_RB x_B = b_R - A_RN x_N and A_RB' y_R = -c_B by CGLS in
 * FP64 on the GPU.  x_out, y_out and objective receive the polished pair if
 * it is primal feasible, has d_B = 0 and correct reduced-cost and dual
 * signs to 1e-9 relative (info->accepted = 1), else the input pair.  A
 * 1e-6 PDHG solution then becomes the vertex it approximates, exact to
 * FP64 rounding.  Any output may be NULL.  0 on success. */
int igaos_pdhg_polish(const igaos_model* model, const double* x, const double* y,
                      double tolerance, int max_iterations, double* x_out, double* y_out,
                      double* objective, igaos_adjoint_info* info);

/* Host: PDHG -> simplex crossover.  From a near-optimal pair x (n), y (m;
 * this engine's sign, d = c + A'y) of the model, a starting basis for the
 * reference's revised simplex in its WarmStart layout
 * (src/simplex/simplex.hpp:10-15; variables: columns [0, n), row slacks
 * n + i valued (A x)_i in [rmin_i, rmax_i]): basis (m) receives the basic
 * variable of each slot or -1 (the simplex puts an artificial there),
 * nb_state (n + m) each variable's state (0 at lower, 1 at upper, 2 basic,
 * 3 free nonbasic).  A variable within tol of a bound (relative to 1 +
 * |bound|) is nonbasic there unless its reduced cost is zero to tol; the
 * most interior of the rest fill the basis.  Returns the number of basic
 * variables placed (at most m), or -1 (igaos_last_error). */
int igaos_crossover_basis(const igaos_model* model, const double* x, const double* y,
                          double tol, int fill, int* basis, unsigned char* nb_state);

/* The device dual repair the engine runs for m > max_repair_m (no
 * counterpart: the reference's repair, pdhg.cu:288-448, is dense and
 * stops at m = 46,340): the same POCS alternation with x fixed -- the
 * projection onto d_J = 0 by conjugate gradients on A_J A_J' + ridge (at
 * most 200 iterations to 1e-10 relative residual each), the halfspace
 * steps of the sign-constrained columns taken together (each row moves by
 * the mean of its violated columns' steps), the row cones -- from the two
 * anchors, at most 300 iterations each, keeping the y with the best finite
 * gap.  x (n), y_it (m) host, original units; y (m, host) receives the
 * dual; out (6 doubles, may be NULL) gap, dinf, dual bound, POCS and CG
 * iterations, |J|.  Needs a visible HIP device.  Returns 1 when a finite
 * gap was found, 0 otherwise, -1 on error. */
int igaos_dual_repair_device(const igaos_model* model, const double* x, const double* y_it,
                             double col_tol, double* y, double* out);

/* Host: row-block layout of a CSR row pointer (host) for the SpMV
 * kernels: pairs (first row, first nonzero) of each block followed by the
 * pair (nrows, nnz), i.e. 2 * (blocks + 1) ints.  A block holds at most
 * 2048 nonzeros and 256 rows, or a single longer row.  Writes at most cap
 * ints; returns the number of ints, or the required count if cap is too
 * small, or -1 (igaos_last_error) when ptr does not start at 0 or
 * decreases.  The device copy must be 8-byte aligned. */
int igaos_rowblocks_build(int nrows, const int* ptr, int* out, int cap);
/* Host: the column-sliced copy of a CSR matrix that the engine's large-LP
 * products use (no reference counterpart): the columns cut into `slices`
 * (1..16) ranges of equal nonzero count, each row's nonzeros split by
 * range in their order, slices stored one after the other.  Writes sptr
 * (slices * (nrows + 1) offsets), sidx and src (nnz each: column index,
 * position in the input), and per row block (slice-major) its (first row,
 * end row, first nonzero, end nonzero) into blocks and its slice into
 * bslice.  Returns the number of blocks; when cap (blocks' capacity in
 * blocks) is smaller, writes nothing and returns the count needed; -1
 * (igaos_last_error) for a malformed pattern. */
int igaos_slices_build(int nrows, int ncols, const int* ptr, const int* idx, int slices,
                       int* sptr, int* sidx, int* src, int* blocks, int* bslice, int cap);
/* Host: the gather order the engine gives the slices of A when its columns
 * are renumbered by length (no reference counterpart): the layout of
 * igaos_slices_build with the entries of every block of two or more rows
 * (at most 2048 nonzeros) sorted by column (stable), and slot[q] = the
 * place of entry q in its block's row-major order (0 in a block of one
 * row longer than 2048, which keeps its order).  Writes sidx, src and
 * slot (nnz each); returns nnz, or -1 (igaos_last_error) for a malformed
 * pattern. */
int igaos_slices_gather_order(int nrows, int ncols, const int* ptr, const int* idx,
                              int slices, int* sidx, int* src, unsigned short* slot);
/* Host: the sliced-ELL layout the small-LP runner streams (no reference
 * counterpart): rows go in slots in order of decreasing length (perm[s] =
 * the row in slot s, ties by row), slots in slices of 64, and a row's
 * entries, in storage order, in chunks of `chunk` (the runner uses 4 in
 * FP64, 8 in FP32): entry e of slot s sits at position
 * ((off[s / 64] + e / chunk) * 64 + s % 64) * chunk + e % chunk, and
 * src[p] is the CSR position held at position p (-1: padding).  Writes
 * perm (nrows) and off (ceil(nrows / 64) + 1), and src when src_cap holds
 * all positions; returns the number of positions (at least one chunk row),
 * or -1 (igaos_last_error) for a malformed pattern or chunk outside 1..16. */
long igaos_small_ell_layout(int nrows, const int* ptr, int chunk, int* perm, int* off, int* src,
                            long src_cap);
/* y = A x with A in CSR (device arrays); replaces the cusparseSpMV calls
 * of src/pdhg/pdhg.cu:150-156 for A and A' (CSR of A' = CSC of A). */
int igaos_spmv_csr_f64(int nblk, const int* rowblocks, const int* ptr,
                       const int* idx, const double* val, const double* x,
                       double* y, void* stream);
int igaos_spmv_csr_f32(int nblk, const int* rowblocks, const int* ptr,
                       const int* idx, const float* val, const float* x,
                       float* y, void* stream);
/* CSR -> CSC of device arrays (replaces the transpose half of
 * counting_sort, src/io/model.hpp:39-67): cp (n + 1), ci, acx (nnz)
 * receive the transpose, each column's entries in row order, by a stable
 * radix sort of the column indices.  Allocates its own scratch and
 * synchronises the stream before returning.  0 on success. */
int igaos_csr_to_csc_f64(int m, int n, int nnz, const int* ap, const int* ai,
                         const double* ax, int* cp, int* ci, double* acx, void* stream);
/* Ruiz equilibration, 10 sweeps (pdhg.cu:469-492), device arrays. */
int igaos_ruiz_f64(int m, int n, const int* ap, const int* ai,
                   const double* ax, const int* cp, const int* ci,
                   const double* acx, double* dr, double* dc, void* stream);
/* The three line-search dot products |dx|^2, |dy|^2 and dy'adx of an
 * attempt; replaces kern_dots + kern_dotfin (pdhg.cu:63-125, launched at
 * 764-767) with the same two-stage fixed order: gD = min(ceil(max(n,m) /
 * 256), 4096) workgroups of 256 threads accumulate grid-stride and reduce
 * by a halving tree, then one workgroup sums the gD partials strided and
 * reduces them by the same tree, so results are bit-reproducible.  dx (n),
 * dy and adx (m) and out (3 doubles) are device arrays; work is device
 * scratch of at least igaos_ls_dots_work(n, m) doubles (3 gD).  Enqueued on
 * the stream; nothing is allocated. */
long igaos_ls_dots_work(int n, int m);
int igaos_ls_dots_f64(int n, int m, const double* dx, const double* dy,
                      const double* adx, double* work, double* out, void* stream);
/* No counterpart in the reference: igaos_ls_dots_f64 for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops that forbid hipMalloc) and have no workspace to hand over.  The
 * partials go to a scratch array the library holds per GPU for the life of
 * the process (96 KiB: 3 x 4096 doubles, which the layout's 4096-workgroup
 * cap makes enough for every n, m); sizes that would need more are refused
 * with igaos_last_error().  Same arguments as igaos_ls_dots_f64 without
 * `work`; the stream must belong to the current device.  Calls that overlap
 * on one GPU share the scratch and race; each call rewrites every partial
 * it reads, so no clearing is needed between calls.  Allocates nothing. */
int igaos_ls_dots_f64_noalloc(int n, int m, const double* dx, const double* dy,
                              const double* adx, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
