/* TEST INFRASTRUCTURE ONLY.  The oracle is a plain-C FP64 host restatement
 * of the reference PDHG engine (src/pdhg/pdhg.cu) used as the checker by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Nothing
 * in the product (igaos_amd/, include/, the HIP library) may link, import or
 * call it. */
#ifndef IGAOS_PDHG_ORACLE_H
#define IGAOS_PDHG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int m, n;
    const double *rmin, *rmax, *cl, *cu, *c;
    const int *ap, *ai;
    const double *ax;
    const int *cp, *ci;
    const double *acx;
    double obj_const;
} orc_model;

typedef struct {
    double time_limit_s;
    double tolerance;
    long max_iterations;
    int verbosity;
    int scaling; /* 0: Ruiz (the reference); 1: Ruiz, then Pock-Chambolle */
} orc_options;

/* pdhg.cu:159-170 (Metrics) */
typedef struct {
    double pinf, dinf, gap, op, od, prow;
} orc_metrics;

typedef struct {
    int status; /* igaos::Status order: 0 Optimal .. 7 Error */
    double objective, pinf, dinf, rel_gap;
    long iterations;
    double solve_time_ms;
    long attempts;  /* primal/dual attempts including line-search retries */
    long restarts;
    double eta, omega, lam2, eta0, omega0;
    int last_is_avg;
    int used_incumbent;
    int repaired;
    long rollbacks; /* incumbent rollbacks (pdhg.cu:965-988) */
} orc_result;

/* mutation switch for pin tests: 0 = faithful.  See pdhg_oracle.c. */
void orc_set_mutation(int code);
void orc_set_threads(int k);

void orc_spmv(int nrows, const int* ptr, const int* idx, const double* val,
              const double* x, double* y);
void orc_ruiz(const orc_model* lp, double* dr, double* dc);
void orc_pock_chambolle(const orc_model* lp, double* dr, double* dc);
void orc_scale(const orc_model* lp, const double* dr, const double* dc,
               double* sax, double* sacx, double* sc, double* scl,
               double* scu, double* srmin, double* srmax);
double orc_power(int m, int n, const int* ap, const int* ai,
                 const double* sax, const int* cp, const int* ci,
                 const double* sacx);
double orc_omega0(int m, int n, const double* sc, const double* srmin,
                  const double* srmax);
void orc_dots(int n, int m, const double* dx, const double* dy,
              const double* adx, double out[3]);
orc_metrics orc_evaluate(int m, int n, const double* rmin,
                         const double* rmax, const double* cl,
                         const double* cu, const double* c, const double* x,
                         const double* y, const double* Ax, const double* d);
double orc_avg_weight(long hn);
void orc_mix_weights(int count, double* out);
int orc_ls_decide(long k, double* eta, double omega, const double dots[3],
                  int* retries);
int orc_dual_repair(const orc_model* lp, const double* x, const double* Ax,
                    const double* y_it, double col_tol, double* y);
/* Full host replay of pdhg::solve.  x_out (n) and y_out (m) receive the
 * reported solution in original units.  trace_eta (optional, length
 * trace_cap) receives eta after each accepted iteration. */
int orc_pdhg_solve(const orc_model* lp, const orc_options* opt,
                   double* x_out, double* y_out, orc_result* res,
                   double* trace_eta, long trace_cap);

#ifdef __cplusplus
}
#endif
#endif
