// extern "C" layer of libigaos_pdhg.so: see include/igaos_pdhg.h.
#include "igaos_pdhg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "igaos/status.h"
#include "model.hpp"
#include "pdhg.hpp"
#include "small.hpp"
#include "exact.hpp"
#include "repair_dev.hpp"
#include "transpose.hpp"
#include "pdhg_kernels.hpp"

struct igaos_model {
    igaos::io::Model m;
};

namespace {

thread_local std::string g_err;

int fail(const std::string& what) {
    g_err = what;
    return -1;
}

}  // namespace

namespace igaos {
// Sets the text igaos_last_error() returns on this thread (for entry
// points defined in other translation units); returns -1.
int set_last_error(const std::string& what) { return fail(what); }
}  // namespace igaos

extern "C" {

int igaos_pdhg_abi_version(void) { return IGAOS_PDHG_ABI_VERSION; }

int igaos_pdhg_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

const char* igaos_last_error(void) { return g_err.c_str(); }

const char* igaos_status_name(int status) {
    if (status < 0 || status > IGAOS_ERROR) return "unknown";
    return igaos::status_name(static_cast<igaos::Status>(status));
}

igaos_model* igaos_model_read_mps(const char* path) {
    try {
        auto* h = new igaos_model;
        try {
            igaos::io::parse_mps(path ? path : "", h->m);
        } catch (...) {
            delete h;
            throw;
        }
        return h;
    } catch (const std::exception& e) {
        fail(e.what());
        return nullptr;
    }
}

igaos_model* igaos_model_from_triplets(int m, int n, int nnz,
                                       const int* rows, const int* cols,
                                       const double* vals, const double* c,
                                       const double* cl, const double* cu,
                                       const double* rmin, const double* rmax,
                                       double obj_const) {
    if (m < 0 || n < 0 || nnz < 0) {
        fail("negative size");
        return nullptr;
    }
    if ((nnz > 0 && (!rows || !cols || !vals)) || (n > 0 && (!c || !cl || !cu)) ||
        (m > 0 && (!rmin || !rmax))) {
        fail("null data array");
        return nullptr;
    }
    try {
        std::vector<igaos::io::Triplet> t(nnz);
        for (int p = 0; p < nnz; ++p) {
            if (rows[p] < 0 || rows[p] >= m || cols[p] < 0 || cols[p] >= n) {
                fail("triplet index out of range");
                return nullptr;
            }
            t[p] = {rows[p], cols[p], vals[p]};
        }
        auto h = std::make_unique<igaos_model>();
        auto& md = h->m;
        md.m = m;
        md.n = n;
        md.c.assign(c, c + n);
        md.cl.assign(cl, cl + n);
        md.cu.assign(cu, cu + n);
        md.rmin.assign(rmin, rmin + m);
        md.rmax.assign(rmax, rmax + m);
        md.integ.assign(n, 0);
        md.obj_const = obj_const;
        igaos::io::counting_sort(t, m, n, md.ap, md.ai, md.ax, md.cp, md.ci, md.acx);
        return h.release();
    } catch (const std::exception& e) {
        fail(e.what());
        return nullptr;
    }
}

igaos_model* igaos_model_from_csr_csc(int m, int n, const double* c,
                                      const double* cl, const double* cu,
                                      const double* rmin, const double* rmax,
                                      const int* ap, const int* ai,
                                      const double* ax, const int* cp,
                                      const int* ci, const double* acx,
                                      double obj_const) {
    if (m < 0 || n < 0 || !ap || !cp) {
        fail("invalid sizes or pointers");
        return nullptr;
    }
    std::string bad = igaos::io::check_csr_csc(m, n, ap, ai, cp, ci);
    if (!bad.empty()) {
        fail(bad);
        return nullptr;
    }
    const int nnz = ap[m];
    if ((nnz > 0 && (!ax || !acx)) || (n > 0 && (!c || !cl || !cu)) ||
        (m > 0 && (!rmin || !rmax))) {
        fail("null data array");
        return nullptr;
    }
    bad = igaos::io::check_transpose(m, n, ap, ai, ax, cp, ci, acx);
    if (!bad.empty()) {
        fail(bad);
        return nullptr;
    }
    try {
        auto h = std::make_unique<igaos_model>();
        auto& md = h->m;
        md.m = m;
        md.n = n;
        md.c.assign(c, c + n);
        md.cl.assign(cl, cl + n);
        md.cu.assign(cu, cu + n);
        md.rmin.assign(rmin, rmin + m);
        md.rmax.assign(rmax, rmax + m);
        md.ap.assign(ap, ap + m + 1);
        md.ai.assign(ai, ai + nnz);
        md.ax.assign(ax, ax + nnz);
        md.cp.assign(cp, cp + n + 1);
        md.ci.assign(ci, ci + nnz);
        md.acx.assign(acx, acx + nnz);
        md.integ.assign(n, 0);
        md.obj_const = obj_const;
        return h.release();
    } catch (const std::exception& e) {
        fail(e.what());
        return nullptr;
    }
}

igaos_model* igaos_model_from_csr(int m, int n, const double* c, const double* cl,
                                  const double* cu, const double* rmin,
                                  const double* rmax, const int* ap, const int* ai,
                                  const double* ax, double obj_const) {
    const std::string bad = igaos::io::check_compressed(m, n, ap, ai);
    if (!bad.empty()) {
        fail("CSR: " + bad);
        return nullptr;
    }
    const int nnz = ap[m];
    if ((nnz > 0 && !ax) || (n > 0 && (!c || !cl || !cu)) || (m > 0 && (!rmin || !rmax))) {
        fail("null data array");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        fail("igaos_model_from_csr builds the CSC on a HIP device; none is visible");
        return nullptr;
    }
    struct Buf {  // device arrays, freed on every exit
        void* p = nullptr;
        ~Buf() {
            if (p) (void)hipFree(p);
        }
    } d_ap, d_ai, d_ax, d_cp, d_ci, d_acx;
    try {
        auto h = std::make_unique<igaos_model>();
        auto& md = h->m;
        md.m = m;
        md.n = n;
        md.c.assign(c, c + n);
        md.cl.assign(cl, cl + n);
        md.cu.assign(cu, cu + n);
        md.rmin.assign(rmin, rmin + m);
        md.rmax.assign(rmax, rmax + m);
        md.ap.assign(ap, ap + m + 1);
        md.ai.assign(ai, ai + nnz);
        md.ax.assign(ax, ax + nnz);
        md.cp.resize(static_cast<size_t>(n) + 1);
        md.ci.resize(nnz);
        md.acx.resize(nnz);
        md.integ.assign(n, 0);
        md.obj_const = obj_const;
        auto chk = [](hipError_t e) {
            if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
        };
        const size_t zi = sizeof(int) * (nnz > 0 ? nnz : 1);
        const size_t zd = sizeof(double) * (nnz > 0 ? nnz : 1);
        chk(hipMalloc(&d_ap.p, sizeof(int) * (static_cast<size_t>(m) + 1)));
        chk(hipMalloc(&d_ai.p, zi));
        chk(hipMalloc(&d_ax.p, zd));
        chk(hipMalloc(&d_cp.p, sizeof(int) * (static_cast<size_t>(n) + 1)));
        chk(hipMalloc(&d_ci.p, zi));
        chk(hipMalloc(&d_acx.p, zd));
        chk(hipMemcpy(d_ap.p, ap, sizeof(int) * (static_cast<size_t>(m) + 1), hipMemcpyHostToDevice));
        if (nnz > 0) {
            chk(hipMemcpy(d_ai.p, ai, zi, hipMemcpyHostToDevice));
            chk(hipMemcpy(d_ax.p, ax, zd, hipMemcpyHostToDevice));
        }
        igaos::dev::csr_to_csc(m, n, nnz, static_cast<int*>(d_ap.p), static_cast<int*>(d_ai.p),
                               static_cast<double*>(d_ax.p), static_cast<int*>(d_cp.p),
                               static_cast<int*>(d_ci.p), static_cast<double*>(d_acx.p),
                               nullptr);
        chk(hipMemcpy(md.cp.data(), d_cp.p, sizeof(int) * (static_cast<size_t>(n) + 1),
                      hipMemcpyDeviceToHost));
        if (nnz > 0) {
            chk(hipMemcpy(md.ci.data(), d_ci.p, zi, hipMemcpyDeviceToHost));
            chk(hipMemcpy(md.acx.data(), d_acx.p, zd, hipMemcpyDeviceToHost));
        }
        return h.release();
    } catch (const std::exception& e) {
        fail(e.what());
        return nullptr;
    }
}

void igaos_model_free(igaos_model* model) { delete model; }

int igaos_model_row_activity(const igaos_model* model, const double* x, double* out) {
    if (!model || (model->m.n > 0 && !x) || (model->m.m > 0 && !out))
        return fail("null model or array");
    igaos::io::row_activity(model->m, x, out);
    return 0;
}

void igaos_model_dims(const igaos_model* h, int* m, int* n, int* nnz,
                      double* obj_const, int* n_fr, int* n_ranges,
                      int* n_int) {
    if (!h) return;
    const auto& md = h->m;
    if (m) *m = md.m;
    if (n) *n = md.n;
    if (nnz) *nnz = md.nnz();
    if (obj_const) *obj_const = md.obj_const;
    if (n_fr) *n_fr = md.n_fr_parsed;
    if (n_ranges) *n_ranges = md.n_ranges_parsed;
    if (n_int) {
        int k = 0;
        for (unsigned char v : md.integ) k += v != 0;
        *n_int = k;
    }
}

void igaos_model_get(const igaos_model* h, double* rmin, double* rmax,
                     double* cl, double* cu, double* c, int* ap, int* ai,
                     double* ax, int* cp, int* ci, double* acx) {
    if (!h) return;
    const auto& md = h->m;
    auto cpy = [](auto* dst, const auto& src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(src[0]) * src.size());
    };
    cpy(rmin, md.rmin);
    cpy(rmax, md.rmax);
    cpy(cl, md.cl);
    cpy(cu, md.cu);
    cpy(c, md.c);
    cpy(ap, md.ap);
    cpy(ai, md.ai);
    cpy(ax, md.ax);
    cpy(cp, md.cp);
    cpy(ci, md.ci);
    cpy(acx, md.acx);
}

void igaos_options_default(igaos_options* opt) {
    if (!opt) return;
    igaos::Options o;
    opt->time_limit_s = o.time_limit_s;
    opt->tolerance = o.tolerance;
    opt->max_iterations = o.max_iterations;
    opt->seed = o.seed;
    opt->presolve = o.presolve ? 1 : 0;
    opt->restarts = o.restarts ? 1 : 0;
    opt->verbosity = o.verbosity;
}

void igaos_engine_config_default(igaos_engine_config* cfg) {
    if (!cfg) return;
    igaos::pdhg::EngineConfig e;
    cfg->precision = e.precision;
    cfg->graph_slots = e.graph_slots;
    cfg->max_repair_m = e.max_repair_m;
    cfg->device = e.device;
    cfg->column_slices = e.column_slices;
    cfg->xcd_pinned = e.xcd_pinned;
    cfg->scaling = e.scaling;
    cfg->exact = e.exact;
    cfg->small_runner = e.small_runner;
    cfg->stream_loads = e.stream_loads;
}

namespace {

bool to_options(const igaos_options* opt, igaos::Options& o) {
    if (!opt) return false;
    o.time_limit_s = opt->time_limit_s;
    o.tolerance = opt->tolerance;
    o.max_iterations = opt->max_iterations;
    o.seed = opt->seed;
    o.presolve = opt->presolve != 0;
    o.restarts = opt->restarts != 0;
    o.verbosity = opt->verbosity;
    return true;
}

// nullptr on success, else what is wrong with the configuration
const char* to_config(const igaos_engine_config* cfg, igaos::pdhg::EngineConfig& ec) {
    if (!cfg) return nullptr;
    if (cfg->precision != 32 && cfg->precision != 64) return "precision must be 32 or 64";
    if (cfg->column_slices < -1 || cfg->column_slices > 16)
        return "column_slices must be -1 (engine's choice), 0 (off) or 1..16";
    ec.precision = cfg->precision;
    ec.graph_slots = std::max(0, cfg->graph_slots);
    ec.max_repair_m = cfg->max_repair_m;
    ec.device = cfg->device;
    ec.column_slices = cfg->column_slices;
    if (cfg->xcd_pinned != 0 && cfg->xcd_pinned != 1) return "xcd_pinned must be 0 or 1";
    ec.xcd_pinned = cfg->xcd_pinned;
    if (cfg->scaling != 0 && cfg->scaling != 1) return "scaling must be 0 (Ruiz) or 1 (Ruiz + Pock-Chambolle)";
    ec.scaling = cfg->scaling;
    if (cfg->exact != 0 && cfg->exact != 1) return "exact must be 0 or 1";
    if (cfg->exact && cfg->precision != 64) return "exact mode needs precision 64";
    ec.exact = cfg->exact;
    if (cfg->small_runner < -1 || cfg->small_runner > 1) return "small_runner must be -1, 0 or 1";
    ec.small_runner = cfg->small_runner;
    if (cfg->stream_loads < -1 || cfg->stream_loads > 1) return "stream_loads must be -1, 0 or 1";
    ec.stream_loads = cfg->stream_loads;
    return nullptr;
}

void export_solution(const igaos::Solution& s, const igaos::pdhg::SolveInfo& si,
                     int m, int n, double* x, double* y, igaos_solve_info* info,
                     char* msg, int msglen) {
    if (x && s.x.size() == static_cast<size_t>(n))
        std::memcpy(x, s.x.data(), sizeof(double) * s.x.size());
    if (y && s.y.size() == static_cast<size_t>(m))
        std::memcpy(y, s.y.data(), sizeof(double) * s.y.size());
    if (msg && msglen > 0) {
        std::strncpy(msg, s.message.c_str(), msglen - 1);
        msg[msglen - 1] = '\0';
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->status = static_cast<int>(s.status);
        info->objective = s.objective;
        info->pinf = s.pinf;
        info->dinf = s.dinf;
        info->rel_gap = s.rel_gap;
        info->iterations = s.iterations;
        info->solve_time_ms = s.solve_time_ms;
        info->setup_ms = si.setup_ms;
        info->attempts = si.attempts;
        info->restarts = si.restarts;
        info->checks = si.checks;
        info->refinements = si.refinements;
        info->eta = si.eta;
        info->omega = si.omega;
        info->lam2 = si.lam2;
        info->eta0 = si.eta0;
        info->omega0 = si.omega0;
        info->last_is_avg = si.last_is_avg;
        info->used_incumbent = si.used_incumbent;
        info->repaired = si.repaired;
        info->kernel_ms = si.kernel_ms;
        info->rollbacks = si.rollbacks;
    }
    if (s.status == igaos::Status::Error) g_err = s.message;
}

}  // namespace

struct igaos_engine {
    std::unique_ptr<igaos::pdhg::EngineBase> e;
    int m = 0, n = 0;
};

int igaos_pdhg_solve(const igaos_model* model, const igaos_options* opt,
                     const igaos_engine_config* cfg, double* x, double* y,
                     igaos_solve_info* info, char* msg, int msglen) {
    igaos::Options o;
    igaos::pdhg::EngineConfig ec;
    if (!model || !to_options(opt, o)) return fail("null model or options");
    if (const char* err = to_config(cfg, ec)) return fail(err);
    igaos::pdhg::SolveInfo si;
    const igaos::Solution s = igaos::pdhg::solve_ex(model->m, o, ec, &si);
    export_solution(s, si, model->m.m, model->m.n, x, y, info, msg, msglen);
    return 0;
}

igaos_engine* igaos_pdhg_create(const igaos_model* model,
                                const igaos_engine_config* cfg) {
    igaos::pdhg::EngineConfig ec;
    if (!model) {
        fail("null model");
        return nullptr;
    }
    if (const char* err = to_config(cfg, ec)) {
        fail(err);
        return nullptr;
    }
    try {
        auto* h = new igaos_engine;
        h->m = model->m.m;
        h->n = model->m.n;
        try {
            h->e = igaos::pdhg::make_engine(model->m, ec);
        } catch (...) {
            delete h;
            throw;
        }
        return h;
    } catch (const std::exception& ex) {
        fail(ex.what());
        return nullptr;
    }
}

int igaos_pdhg_run(igaos_engine* engine, const igaos_options* opt, double* x,
                   double* y, igaos_solve_info* info, char* msg, int msglen) {
    igaos::Options o;
    if (!engine || !to_options(opt, o)) return fail("null engine or options");
    igaos::pdhg::SolveInfo si;
    igaos::Solution s;
    try {
        s = engine->e->run(o, &si);
    } catch (const std::exception& ex) {
        s = igaos::Solution{};
        s.status = igaos::Status::Error;
        s.message = ex.what();
    }
    export_solution(s, si, engine->m, engine->n, x, y, info, msg, msglen);
    return 0;
}

double igaos_pdhg_setup_ms(const igaos_engine* engine) {
    return engine ? engine->e->setup_ms() : 0.0;
}

void igaos_pdhg_destroy(igaos_engine* engine) { delete engine; }

int igaos_pdhg_column_slices(const igaos_engine* engine, int* slices_primal,
                             int* slices_dual) {
    if (!engine) return fail("null engine");
    int p = 0, d = 0;
    engine->e->column_slices(p, d);
    if (slices_primal) *slices_primal = p;
    if (slices_dual) *slices_dual = d;
    return 0;
}

int igaos_pdhg_profile_attempts(igaos_engine* engine, int attempts,
                                double* ms_primal, double* ms_dual,
                                int* blocks_primal, int* blocks_dual) {
    if (!engine) return fail("null engine");
    try {
        const auto t = engine->e->profile(attempts < 1 ? 1 : attempts);
        if (ms_primal) *ms_primal = t.ms_primal;
        if (ms_dual) *ms_dual = t.ms_dual;
        if (blocks_primal) *blocks_primal = t.nblk_primal;
        if (blocks_dual) *blocks_dual = t.nblk_dual;
    } catch (const std::exception& ex) {
        return fail(ex.what());
    }
    return 0;
}

int igaos_pdhg_polish(const igaos_model* model, const double* x, const double* y,
                      double tolerance, int max_iterations, double* x_out, double* y_out,
                      double* objective, igaos_adjoint_info* info) {
    if (!model) return fail("null model");
    const auto& md = model->m;
    if ((md.n > 0 && !x) || (md.m > 0 && !y)) return fail("null array");
    int m = md.m, n = md.n, nnz = md.nnz();
    const double oc = md.obj_const;
    struct Batch {  // destroyed on every exit
        igaos_batch* b = nullptr;
        ~Batch() {
            if (b) igaos_batch_destroy(b);
        }
    } B;
    B.b = igaos_batch_create(1, &m, &n, &nnz, md.c.data(), md.cl.data(), md.cu.data(),
                             md.rmin.data(), md.rmax.data(), md.ap.data(), md.ai.data(),
                             md.ax.data(), md.cp.data(), md.ci.data(), md.acx.data(), &oc, 64);
    if (!B.b) return fail(igaos_batch_last_error());
    // a zero-iteration run computes the scaling the active sets are read in
    igaos_options o;
    igaos_options_default(&o);
    o.time_limit_s = 1e7;
    o.max_iterations = 0;
    o.tolerance = 0.0;
    if (igaos_batch_run(B.b, &o, nullptr, nullptr, nullptr) != 0 ||
        igaos_batch_set_solution(B.b, x, y) != 0 ||
        igaos_batch_polish(B.b, tolerance, max_iterations, x_out, y_out, objective, info) != 0)
        return fail(igaos_batch_last_error());
    return 0;
}

int igaos_crossover_basis(const igaos_model* model, const double* x, const double* y,
                          double tol, int fill, int* basis, unsigned char* nb_state) {
    if (!model || !basis || !nb_state) return fail("null model or output");
    const auto& md = model->m;
    if ((md.n > 0 && !x) || (md.m > 0 && !y)) return fail("null array");
    if (!(tol >= 0)) return fail("tolerance must be >= 0");
    return igaos::pdhg::crossover_basis(md, x, y, tol, basis, nb_state, fill != 0);
}

int igaos_kkt_evaluate(const igaos_model* model, const double* x, const double* y,
                       const double* Ax, const double* d, double* out) {
    if (!model || !out) return fail("null model or output");
    const auto& md = model->m;
    if ((md.n > 0 && (!x || !d)) || (md.m > 0 && (!y || !Ax))) return fail("null array");
    const igaos::pdhg::HostMetrics t = igaos::pdhg::host_evaluate(md, x, y, Ax, d);
    out[0] = t.pinf;
    out[1] = t.dinf;
    out[2] = t.gap;
    out[3] = t.op;
    out[4] = t.od;
    out[5] = t.prow;
    return 0;
}

int igaos_dual_repair(const igaos_model* model, const double* x, const double* Ax,
                      const double* y_it, double col_tol, double* y) {
    if (!model) return fail("null model");
    const auto& md = model->m;
    if ((md.n > 0 && !x) || (md.m > 0 && (!Ax || !y_it || !y)))
        return fail("null array");
    try {
        return igaos::pdhg::host_dual_repair(md, x, Ax, y_it, col_tol, y) ? 1 : 0;
    } catch (const std::exception& e) {
        return fail(e.what());
    }
}

int igaos_dual_repair_device(const igaos_model* model, const double* x, const double* y_it,
                             double col_tol, double* y, double* out) {
    if (!model) return fail("null model");
    const auto& md = model->m;
    if ((md.n > 0 && !x) || (md.m > 0 && (!y_it || !y))) return fail("null array");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail("the device repair needs a HIP device; none is visible");
    }
    try {
        igaos::dev::RepairInfo info;
        const bool have = igaos::dev::dual_repair_device(md, x, y_it, col_tol, y, &info);
        if (out) {
            out[0] = info.gap;
            out[1] = info.dinf;
            out[2] = info.od;
            out[3] = static_cast<double>(info.iterations);
            out[4] = static_cast<double>(info.cg_iterations);
            out[5] = static_cast<double>(info.columns_in_J);
        }
        return have ? 1 : 0;
    } catch (const std::exception& e) {
        return fail(e.what());
    }
}

int igaos_rowblocks_build(int nrows, const int* ptr, int* out, int cap) {
    if (nrows < 0 || !ptr) return fail("invalid row pointer");
    if (ptr[0] != 0) return fail("pointer array does not start at 0");
    for (int i = 0; i < nrows; ++i)
        if (ptr[i + 1] < ptr[i]) return fail("pointer array decreases at " + std::to_string(i));
    const std::vector<int> rb = igaos::dev::build_row_blocks(ptr, nrows);
    const int k = static_cast<int>(rb.size());
    if (out && cap >= k) std::memcpy(out, rb.data(), sizeof(int) * k);
    return k;
}

int igaos_slices_build(int nrows, int ncols, const int* ptr, const int* idx, int slices,
                       int* sptr, int* sidx, int* src, int* blocks, int* bslice, int cap) {
    if (nrows < 0 || ncols < 0 || !ptr || (ptr[nrows] > 0 && !idx))
        return fail("invalid CSR arrays");
    if (slices < 1 || slices > igaos::dev::MAX_SLICES) return fail("slices must be 1..16");
    const std::string bad = igaos::io::check_compressed(nrows, ncols, ptr, idx);
    if (!bad.empty()) return fail(bad);
    try {
        const igaos::dev::HostSlices h =
            igaos::dev::build_slices(nrows, ncols, ptr, idx, slices);
        const int nb = static_cast<int>(h.bslice.size());
        if (cap < nb) return nb;
        if (sptr) std::memcpy(sptr, h.ptr.data(), sizeof(int) * h.ptr.size());
        if (sidx) std::memcpy(sidx, h.idx.data(), sizeof(int) * h.idx.size());
        if (src) std::memcpy(src, h.src.data(), sizeof(int) * h.src.size());
        if (blocks) std::memcpy(blocks, h.quad.data(), sizeof(int) * h.quad.size());
        if (bslice) std::memcpy(bslice, h.bslice.data(), sizeof(int) * nb);
        return nb;
    } catch (const std::exception& e) {
        return fail(e.what());
    }
}

int igaos_slices_gather_order(int nrows, int ncols, const int* ptr, const int* idx,
                              int slices, int* sidx, int* src, unsigned short* slot) {
    if (nrows < 0 || ncols < 0 || !ptr || (ptr[nrows] > 0 && !idx))
        return fail("invalid CSR arrays");
    if (slices < 1 || slices > igaos::dev::MAX_SLICES) return fail("slices must be 1..16");
    const std::string bad = igaos::io::check_compressed(nrows, ncols, ptr, idx);
    if (!bad.empty()) return fail(bad);
    try {
        const igaos::dev::HostSlices h =
            igaos::dev::build_slices(nrows, ncols, ptr, idx, slices, true);
        const size_t nnz = h.idx.size();
        if (sidx) std::memcpy(sidx, h.idx.data(), sizeof(int) * nnz);
        if (src) std::memcpy(src, h.src.data(), sizeof(int) * nnz);
        if (slot) std::memcpy(slot, h.slot.data(), sizeof(unsigned short) * nnz);
        return static_cast<int>(nnz);
    } catch (const std::exception& e) {
        return fail(e.what());
    }
}

long igaos_small_ell_layout(int nrows, const int* ptr, int chunk, int* perm, int* off, int* src,
                            long src_cap) {
    if (nrows < 0 || !ptr || !perm || !off) return fail("invalid arrays");
    if (chunk < 1 || chunk > 16) return fail("chunk must be 1..16");
    if (ptr[0] != 0) return fail("ptr[0] must be 0");
    for (int r = 0; r < nrows; ++r)
        if (ptr[r + 1] < ptr[r]) return fail("ptr must not decrease");
    try {
        std::vector<int> hperm, hoff, hsrc;
        const long total = igaos::dev::small_ell_map(nrows, ptr, chunk, hperm, hoff, hsrc);
        std::copy(hperm.begin(), hperm.end(), perm);
        std::copy(hoff.begin(), hoff.end(), off);
        if (src && src_cap >= total) std::copy(hsrc.begin(), hsrc.end(), src);
        return total;
    } catch (const std::exception& e) {
        return fail(e.what());
    }
}

int igaos_spmv_csr_f64(int nblk, const int* rowblocks, const int* ptr,
                       const int* idx, const double* val, const double* x,
                       double* y, void* stream) {
    if (nblk < 0) return fail("negative block count");
    igaos::dev::spmv({rowblocks, nblk}, ptr, idx, val, x, y,
                     static_cast<hipStream_t>(stream));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail(hipGetErrorString(e));
}

int igaos_spmv_csr_f32(int nblk, const int* rowblocks, const int* ptr,
                       const int* idx, const float* val, const float* x,
                       float* y, void* stream) {
    if (nblk < 0) return fail("negative block count");
    igaos::dev::spmv({rowblocks, nblk}, ptr, idx, val, x, y,
                     static_cast<hipStream_t>(stream));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail(hipGetErrorString(e));
}

int igaos_ruiz_f64(int m, int n, const int* ap, const int* ai,
                   const double* ax, const int* cp, const int* ci,
                   const double* acx, double* dr, double* dc, void* stream) {
    if (m < 0 || n < 0) return fail("negative size");
    igaos::dev::ruiz(m, n, ap, ai, ax, cp, ci, acx, dr, dc,
                     static_cast<hipStream_t>(stream));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail(hipGetErrorString(e));
}

int igaos_csr_to_csc_f64(int m, int n, int nnz, const int* ap, const int* ai,
                         const double* ax, int* cp, int* ci, double* acx, void* stream) {
    if (m < 0 || n < 0 || nnz < 0) return fail("negative size");
    if (!ap || !cp || (nnz > 0 && (!ai || !ax || !ci || !acx))) return fail("null device array");
    try {
        igaos::dev::csr_to_csc(m, n, nnz, ap, ai, ax, cp, ci, acx,
                               static_cast<hipStream_t>(stream));
    } catch (const std::exception& e) {
        return fail(e.what());
    }
    return 0;
}

long igaos_ls_dots_work(int n, int m) {
    if (n < 0 || m < 0) return fail("negative size");
    return 3L * igaos::dev::dots_blocks(n, m);
}

int igaos_ls_dots_f64(int n, int m, const double* dx, const double* dy,
                      const double* adx, double* work, double* out, void* stream) {
    if (n < 0 || m < 0) return fail("negative size");
    if ((n > 0 && !dx) || (m > 0 && (!dy || !adx)) || !work || !out)
        return fail("null device array");
    igaos::dev::ls_dots(n, m, dx, dy, adx, work, out, static_cast<hipStream_t>(stream));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail(hipGetErrorString(e));
}

}  // extern "C"
