// igaos::pdhg::solve on one MI355X: the reference's restarted adaptive PDHG
// (src/pdhg/pdhg.cu:457-1156) with every iterate, the line search and the
// KKT reductions resident in HBM.
//
// Structure of a solve:
//  * setup (untimed, as in the reference, pdhg.cu:692): upload the model,
//    Ruiz equilibration and scaling on the device, power iteration for
//    ||A~||_2^2, eta0 = 1.2/||A~||, omega0 = clamp(||c~||/||q~||, 0.3, 3);
//  * iterations: hipGraph replays of `graph_slots` attempt slots; each slot
//    is one line-search attempt (k_primal + k_dual).  The dual kernel's last
//    workgroup decides acceptance on the device and freezes the state when
//    an accepted iteration lands on a multiple of 64 (or MAXIT), so the host
//    only synchronises once per check period;
//  * check (host control, device reductions): flush the pending averaging,
//    evaluate both iterates' KKT metrics in FP64 on the device, then apply
//    the reference's termination, dual repair, incumbent, rollback,
//    restart and primal-weight rules (pdhg.cu:801-1084).
//
// Precision 32 runs the same iteration in FP32 on a correction problem:
// x = x_ref + dx, y = y_ref + dy with (x_ref, y_ref) in FP64.  The FP32
// kernels see the shifted bounds clamp(dx .. scl - x_ref, scu - x_ref), the
// FP64-computed reduced cost g_ref = c~ + A~'y_ref, the row bounds less
// A~x_ref, and y_ref as the dual clamp offset; A~ is held in FP32 as well.
// KKT checks rebuild x, y in FP64 and use the FP64 matrix.  Every restart
// folds the anchor's correction into (x_ref, y_ref) in FP64, recomputes the
// residual data with FP64 SpMVs and restarts the FP32 correction at zero:
// an outer iterative refinement whose inner solver is PDHG in FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "exact.hpp"
#include "small.hpp"
#include "repair_dev.hpp"
#include "pdhg.hpp"
#include "pdhg_kernels.hpp"

namespace igaos::pdhg {

namespace {

struct HipFailure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCK(expr)                                                          \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess)                                                \
            throw HipFailure(std::string("HIP error in PDHG engine: ") +     \
                             hipGetErrorString(e_) + " at " #expr);          \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t k) {
        n = k;
        HIPCK(hipMalloc(&p, sizeof(T) * (k > 0 ? k : 1)));
    }
    void upload(const T* h, size_t k, hipStream_t s) {
        alloc(k);
        if (k) HIPCK(hipMemcpyAsync(p, h, sizeof(T) * k, hipMemcpyHostToDevice, s));
    }
    void zero(hipStream_t s) {
        HIPCK(hipMemsetAsync(p, 0, sizeof(T) * (n > 0 ? n : 1), s));
    }
};

double now_s() {
    return std::chrono::duration<double>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Column order for gather locality.  A' row sums gather y and A row sums
// gather x at the columns of each row; when a few columns hold a large
// share of the nonzeros (heavy-tailed column lengths), numbering those
// columns first packs half of all x gathers into a few cache lines' worth
// of x.  The engine then works on a column-permuted copy: x, c and the
// bounds are permuted, every row keeps its nonzeros in their original
// order (so A x sums in the same order), and the solution is permuted
// back on output.  Applied when n >= 65536 and the top 2% of columns hold
// at least a quarter of the nonzeros (never in exact mode); both dual
// repairs work on the permuted copy like every other step.
std::unique_ptr<io::Model> locality_copy(const io::Model& md,
                                         const EngineConfig& cfg,
                                         std::vector<int>& perm) {
    const int n = md.n;
    const long nnz = md.nnz();
    if (cfg.exact || n < (1 << 16) || nnz == 0) return nullptr;
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        return md.cp[a + 1] - md.cp[a] > md.cp[b + 1] - md.cp[b];
    });
    long heavy = 0;
    for (int k = 0; k < n / 50; ++k) heavy += md.cp[ord[k] + 1] - md.cp[ord[k]];
    if (heavy * 4 < nnz) return nullptr;
    std::vector<int> inv(n);
    for (int k = 0; k < n; ++k) inv[ord[k]] = k;
    auto M = std::make_unique<io::Model>();
    M->m = md.m;
    M->n = n;
    M->rmin = md.rmin;
    M->rmax = md.rmax;
    M->ap = md.ap;
    M->ax = md.ax;
    M->ai.resize(nnz);
    for (long p = 0; p < nnz; ++p) M->ai[p] = inv[md.ai[p]];
    M->c.resize(n);
    M->cl.resize(n);
    M->cu.resize(n);
    M->cp.resize(n + 1);
    M->ci.resize(nnz);
    M->acx.resize(nnz);
    long q = 0;
    M->cp[0] = 0;
    for (int k = 0; k < n; ++k) {
        const int j = ord[k];
        M->c[k] = md.c[j];
        M->cl[k] = md.cl[j];
        M->cu[k] = md.cu[j];
        for (int p = md.cp[j]; p < md.cp[j + 1]; ++p, ++q) {
            M->ci[q] = md.ci[p];
            M->acx[q] = md.acx[p];
        }
        M->cp[k + 1] = static_cast<int>(q);
    }
    if (!md.integ.empty()) {
        M->integ.resize(n);
        for (int k = 0; k < n; ++k) M->integ[k] = md.integ[ord[k]];
    }
    M->obj_const = md.obj_const;
    M->n_fr_parsed = md.n_fr_parsed;
    M->n_ranges_parsed = md.n_ranges_parsed;
    perm.swap(ord);
    return M;
}

template <typename V>
class Engine final : public EngineBase {
  public:
    Engine(const io::Model& lp, const EngineConfig& cfg)
        : own_(locality_copy(lp, cfg, colperm_)), lp_(own_ ? *own_ : lp),
          cfg_(cfg), m_(lp.m), n_(lp.n) {
        const double t0 = now_s();
        setup();
        setup_ms_ = (now_s() - t0) * 1000.0;
    }
    ~Engine() override {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        if (cexec_) (void)hipGraphExecDestroy(cexec_);
        if (cgraph_) (void)hipGraphDestroy(cgraph_);
        if (hck_) (void)hipHostFree(hck_);
        if (ev0_) (void)hipEventDestroy(ev0_);
        if (ev1_) (void)hipEventDestroy(ev1_);
        if (s_) (void)hipStreamDestroy(s_);
    }
    Solution run(const Options& opt, SolveInfo* info) override;
    AttemptTiming profile(int attempts) override;
    double setup_ms() const override { return setup_ms_; }
    void column_slices(int& primal, int& dual) const override {
        primal = ib_.sT.S;
        dual = ib_.sA.S;
    }

  private:
    static constexpr bool F32 = std::is_same<V, float>::value;
    std::vector<int> colperm_;        // internal column k is model column colperm_[k]
    std::unique_ptr<io::Model> own_;  // column-permuted copy (locality_copy)
    const io::Model& lp_;
    EngineConfig cfg_;
    double setup_ms_ = 0.0;
    const int m_, n_;
    hipStream_t s_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;

    // original data (FP64) and scaled data
    DBuf<int> ap_, ai_, cp_, ci_, rbA_, rbT_;
    DBuf<double> ax_, acx_, c_, cl_, cu_, rmin_, rmax_;
    DBuf<double> dr_, dc_, sax_, sacx_, sc_, scl_, scu_, srmin_, srmax_;
    // iteration vectors of precision V
    DBuf<V> vax_, vacx_;  // FP32 copies of the scaled matrix
    DBuf<V> x0_, x1_, y0_, y1_, xbar_, ata_, xavg_, yavg_, axcur_, adx_;
    DBuf<V> lo_, hi_, gref_, rlo_, rhi_, yoff_;  // FP32 correction data
    // FP64 references of the FP32 engine and scratch
    DBuf<double> xref_, yref_, aref_, atref_;
    DBuf<double> kx_[2], ky_[2];  // FP64 iterates of a check (FP32 engine)
    DBuf<double> xb_, yb_, pesx_, pesy_, tmpm_, tmpn_;
    DBuf<double> partT_, partA_, partC_, partR_, partR2_, rowres_, red_, scal_;
    DBuf<LsState> st_;
    DBuf<Metrics> met_;
    // column-sliced copies of A (gathers x) and A' (gathers y), when used
    struct SliceBufs {
        DBuf<int> ptr, idx, quad, bslice, xmap;
        DBuf<unsigned short> slot;
        DBuf<V> val, part;
    } slA_, slT_;
    dev::IterBufs<V> ib_;
    dev::KktData kd_;
    // exact FP64 mode (exact.hip): per-attempt vectors, dot partials, the
    // k^-0.3 / k^-0.6 table (filled up to the largest MAXIT run so far)
    // and the KKT products and summands
    DBuf<double> xdx_, xdy_, xtm_, xpart_, powtab_, kata_[2], ktm_[2], kterm_;
    // device dual repair (m > max_repair_m): its scratch and the point
    dev::RepairDev repair_;
    DBuf<double> rpx_, rpy_, rpyit_;
    long repair_cg_ = 0, repair_its_ = 0;
    long long pow_filled_ = 0;
    dev::ExactBufs eb_;
    // small FP64 LPs in exact mode: one workgroup runs a check period's
    // attempts per launch (exact.hip, k_x_run), with the exact arithmetic
    bool runner_ = false;
    // small LPs in the default mode: the register / LDS runner (small.hip)
    // with sliced-ELL copies of A' and A
    bool small_ = false;
    DBuf<int> epermT_, epermA_, eoffT_, eoffA_;
    DBuf<unsigned short> eidxT_, eidxA_;
    DBuf<V> evalT_, evalA_;
    dev::SmallEll<V> eT_, eA_;
    // the runner's check period as one graph (graph_slots > 0): the runner,
    // the state's pending flush, the KKT metrics at cur = 0 (where the
    // runner leaves the current iterate; k_kkt_done clears the flush's
    // flag) and the state and metrics copied to pinned host memory, so a
    // check costs one launch and one wait
    struct CheckOut {
        LsState st;
        Metrics met[2];
    };
    hipGraph_t cgraph_ = nullptr;
    hipGraphExec_t cexec_ = nullptr;
    CheckOut* hck_ = nullptr;
    static constexpr long long EXACT_POW_CAP = 1LL << 21;
    // attempts per runner launch: a check period of 64 accepted
    // iterations takes at most 64 x 21 attempts (pdhg.cu:785: the 21st
    // retry is accepted)
    static constexpr int RUNNER_CAP = 64 * 22;
    dev::RowBlocks rA_, rT_;
    double lam2_ = 1.0, eta0_ = 1.0, omega0_ = 1.0;

    void setup();
    int slices_for(int cols, int rows, long nnz) const;
    void make_slices(dev::Sliced<V>& sl, SliceBufs& B, int rows, int cols,
                     const std::vector<int>& ptr, const std::vector<int>& idx,
                     const double* sval, int S, bool gather_order);
    void make_ell(dev::SmallEll<V>& E, DBuf<int>& perm, DBuf<int>& off, DBuf<unsigned short>& idx,
                  DBuf<V>& val, int rows, const std::vector<int>& ptr, const int* didx,
                  const V* dval);
    void reset();
    double power_iteration();
    double dsum(const double* a, const double* b, long n);
    void check_metrics(int cur, Metrics& mc, Metrics& ma);
    void kkt_launch(int cur);
    void recompute_refs();
    // y = A~ x (rows of A) in the engine's summation order
    void spmv_rows(const double* x, double* y) {
        if (cfg_.exact) dev::spmv_seq(m_, ap_.p, ai_.p, sax_.p, x, y, s_);
        else dev::spmv(rA_, ap_.p, ai_.p, sax_.p, x, y, s_);
    }
    void spmv_cols(const double* yv, double* out) {
        if (cfg_.exact) dev::spmv_seq(n_, cp_.p, ci_.p, sacx_.p, yv, out, s_);
        else dev::spmv(rT_, cp_.p, ci_.p, sacx_.p, yv, out, s_);
    }
    void attempt(hipStream_t s) {
        if constexpr (F32) dev::launch_attempt(ib_, s);
        else if (cfg_.exact) dev::launch_attempt_exact(ib_, eb_, s);
        else dev::launch_attempt(ib_, s);
    }
    void fill_powtab(long long maxit);
    bool device_repair(const Metrics& t, double tol, Metrics& mr, std::vector<double>& rx,
                       std::vector<double>& ry);
    void sync() { HIPCK(hipStreamSynchronize(s_)); }
    LsState read_state() {
        LsState h;
        HIPCK(hipMemcpyAsync(&h, st_.p, sizeof(LsState), hipMemcpyDeviceToHost, s_));
        sync();
        return h;
    }
    void write_state(const LsState& h) {
        HIPCK(hipMemcpyAsync(st_.p, &h, sizeof(LsState), hipMemcpyHostToDevice, s_));
    }
    V* xcur(int c) { return c ? x1_.p : x0_.p; }
    V* ycur(int c) { return c ? y1_.p : y0_.p; }
    void copy_v(V* dst, const V* src, long k) {
        if (k) HIPCK(hipMemcpyAsync(dst, src, sizeof(V) * k, hipMemcpyDeviceToDevice, s_));
    }
    void copy_d(double* dst, const double* src, long k) {
        if (k) HIPCK(hipMemcpyAsync(dst, src, sizeof(double) * k, hipMemcpyDeviceToDevice, s_));
    }
    void zero_v(V* p, long k) { dev::fill(p, k, V(0), s_); }
    // full FP64 scaled iterate (x or y) of a V vector, into dst
    void full(double* dst, const double* ref, const V* d, long k);
    // the FP32 correction's primal clamp bounds, for the FP64 iterate
    dev::BoundSnap snap_x() const {
        dev::BoundSnap b;
        if constexpr (F32) {
            b.lo = lo_.p;
            b.hi = hi_.p;
            b.scl = scl_.p;
            b.scu = scu_.p;
        }
        return b;
    }
    // full() of a primal iterate: corrections on their clamp bounds give
    // the bounds exactly
    void full_x(double* dst, const double* ref, const V* d) {
        if constexpr (F32) {
            const dev::BoundSnap b = snap_x();
            dev::combine(ref, d, dst, n_, s_, &b);
        } else {
            full(dst, ref, d, n_);
        }
    }
};

template <typename V>
void Engine<V>::full(double* dst, const double* ref, const V* d, long k) {
    if constexpr (F32) {
        dev::combine(ref, d, dst, k, s_);
    } else {
        (void)ref;
        copy_d(dst, d, k);
    }
}

template <typename V>
double Engine<V>::dsum(const double* a, const double* b, long n) {
    if (cfg_.exact) dev::sumsq_seq(a, b, n, scal_.p, s_);
    else dev::sumsq_diff(a, b, n, red_.p, scal_.p, s_);
    double h = 0.0;
    HIPCK(hipMemcpyAsync(&h, scal_.p, sizeof(double), hipMemcpyDeviceToHost, s_));
    sync();
    return h;
}

template <typename V>
double Engine<V>::power_iteration() {
    // pdhg.cu:605-631: v0 = 1/sqrt(n); 30 x { w = A'A v; v = w/||w|| };
    // lam2 = ||A v||^2 (non-finite or <= 0 -> 1).
    DBuf<double> v, w;
    v.alloc(n_);
    w.alloc(n_);
    dev::fill(v.p, n_, 1.0 / std::sqrt(static_cast<double>(n_)), s_);
    for (int it = 0; it < 30; ++it) {
        spmv_rows(v.p, tmpm_.p);
        spmv_cols(tmpm_.p, w.p);
        const double nw = std::sqrt(dsum(w.p, nullptr, n_));
        if (nw < 1e-300) break;
        dev::scale_into(v.p, w.p, n_, 1.0 / nw, s_);
    }
    spmv_rows(v.p, tmpm_.p);
    const double na = std::sqrt(dsum(tmpm_.p, nullptr, m_));
    double lam2 = na * na;
    if (!(lam2 > 0) || !std::isfinite(lam2)) lam2 = 1.0;
    return lam2;
}

template <typename V>
void Engine<V>::setup() {
    HIPCK(hipSetDevice(cfg_.device));
    HIPCK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    HIPCK(hipEventCreate(&ev0_));
    HIPCK(hipEventCreate(&ev1_));
    const long nnz = lp_.nnz();
    ap_.upload(lp_.ap.data(), m_ + 1, s_);
    ai_.upload(lp_.ai.data(), nnz, s_);
    ax_.upload(lp_.ax.data(), nnz, s_);
    cp_.upload(lp_.cp.data(), n_ + 1, s_);
    ci_.upload(lp_.ci.data(), nnz, s_);
    acx_.upload(lp_.acx.data(), nnz, s_);
    c_.upload(lp_.c.data(), n_, s_);
    cl_.upload(lp_.cl.data(), n_, s_);
    cu_.upload(lp_.cu.data(), n_, s_);
    rmin_.upload(lp_.rmin.data(), m_, s_);
    rmax_.upload(lp_.rmax.data(), m_, s_);
    const std::vector<int> hbA = dev::build_row_blocks(lp_.ap.data(), m_);
    const std::vector<int> hbT = dev::build_row_blocks(lp_.cp.data(), n_);
    rbA_.upload(hbA.data(), hbA.size(), s_);
    rbT_.upload(hbT.data(), hbT.size(), s_);
    rA_ = {rbA_.p, static_cast<int>(hbA.size() / 2) - 1};
    rT_ = {rbT_.p, static_cast<int>(hbT.size() / 2) - 1};

    dr_.alloc(m_);
    dc_.alloc(n_);
    dev::ruiz(m_, n_, ap_.p, ai_.p, ax_.p, cp_.p, ci_.p, acx_.p, dr_.p, dc_.p, s_);
    tmpm_.alloc(m_);
    tmpn_.alloc(n_);
    if (cfg_.scaling == 1)
        dev::pock_chambolle(m_, n_, ap_.p, ai_.p, ax_.p, cp_.p, ci_.p, acx_.p, dr_.p, dc_.p,
                            tmpm_.p, tmpn_.p, s_);
    sax_.alloc(nnz);
    sacx_.alloc(nnz);
    sc_.alloc(n_);
    scl_.alloc(n_);
    scu_.alloc(n_);
    srmin_.alloc(m_);
    srmax_.alloc(m_);
    dev::scale_data(m_, n_, ap_.p, ai_.p, ax_.p, cp_.p, ci_.p, acx_.p, dr_.p,
                    dc_.p, c_.p, cl_.p, cu_.p, rmin_.p, rmax_.p, sax_.p,
                    sacx_.p, sc_.p, scl_.p, scu_.p, srmin_.p, srmax_.p, s_);

    red_.alloc(1024);
    scal_.alloc(4);

    lam2_ = power_iteration();
    eta0_ = 1.2 / std::sqrt(lam2_);
    {  // pdhg.cu:647-663
        const double cn = dsum(sc_.p, nullptr, n_);
        if (cfg_.exact) dev::sumsq_q_seq(srmin_.p, srmax_.p, m_, scal_.p, s_);
        else dev::sumsq_q(srmin_.p, srmax_.p, m_, red_.p, scal_.p, s_);
        double qn = 0.0;
        HIPCK(hipMemcpyAsync(&qn, scal_.p, sizeof(double), hipMemcpyDeviceToHost, s_));
        sync();
        omega0_ = 1.0;
        if (cn > 1e-24 && qn > 1e-24)
            omega0_ = std::clamp(std::sqrt(cn / qn), 0.3, 3.0);
    }
    double cmax = 0.0;
    dev::absmax(c_.p, n_, red_.p, scal_.p, s_);
    HIPCK(hipMemcpyAsync(&cmax, scal_.p, sizeof(double), hipMemcpyDeviceToHost, s_));
    sync();

    // iteration buffers
    x0_.alloc(n_); x1_.alloc(n_); xbar_.alloc(n_); ata_.alloc(n_); xavg_.alloc(n_);
    y0_.alloc(m_); y1_.alloc(m_); yavg_.alloc(m_); axcur_.alloc(m_); adx_.alloc(m_);
    xb_.alloc(n_); yb_.alloc(m_); pesx_.alloc(n_); pesy_.alloc(m_);
    {
        const int sT = slices_for(m_, n_, nnz), sA = slices_for(n_, m_, nnz);
        // A's slices gather x: on a column-renumbered copy its first columns
        // are the heavy ones, and blocks sorted by column gather them from
        // few cache lines; y has no such order, so A' keeps row order
        if (sT) make_slices(ib_.sT, slT_, n_, m_, lp_.cp, lp_.ci, sacx_.p, sT, false);
        if (sA) make_slices(ib_.sA, slA_, m_, n_, lp_.ap, lp_.ai, sax_.p, sA, own_ != nullptr);
        ib_.nT_part = sT ? (n_ + dev::BS - 1) / dev::BS : rT_.nblk;
        ib_.nA_part = sA ? std::max((m_ + dev::BS - 1) / dev::BS, 1) : std::max(rA_.nblk, 1);
    }
    partT_.alloc(std::max(ib_.nT_part, 1));
    partA_.alloc(2 * ib_.nA_part);
    partC_.alloc(8 * std::max(rT_.nblk, 1));
    partR_.alloc(8 * std::max(rA_.nblk, 1));
    partR2_.alloc(2 * 2 * 1024);
    rowres_.alloc(8);
    st_.alloc(1);
    met_.alloc(2);

    ib_.m = m_;
    ib_.n = n_;
    ib_.ap = ap_.p; ib_.ai = ai_.p; ib_.cp = cp_.p; ib_.ci = ci_.p;
    ib_.rbA = rA_; ib_.rbT = rT_;
    ib_.x[0] = x0_.p; ib_.x[1] = x1_.p; ib_.y[0] = y0_.p; ib_.y[1] = y1_.p;
    ib_.xbar = xbar_.p; ib_.ata = ata_.p; ib_.xavg = xavg_.p; ib_.yavg = yavg_.p;
    ib_.axcur = axcur_.p; ib_.adx = adx_.p;
    ib_.partT = partT_.p; ib_.partA = partA_.p;
    ib_.st = st_.p;
    // matrices that cannot stay in the 256 MB last-level cache between
    // iterations stream with nontemporal loads (about 6 TB/s against 4 TB/s
    // for cached loads on gfx950, and the gathered vectors keep L2)
    ib_.nt = cfg_.stream_loads == 1 ||
             (cfg_.stream_loads < 0 &&
              2.0 * static_cast<double>(nnz) * (sizeof(V) + sizeof(int)) > 256.0e6);
    if constexpr (F32) {
        vax_.alloc(nnz);
        vacx_.alloc(nnz);
        dev::to_float(sax_.p, vax_.p, nnz, s_);
        dev::to_float(sacx_.p, vacx_.p, nnz, s_);
        ib_.ax = vax_.p;
        ib_.acx = vacx_.p;
        xref_.alloc(n_); yref_.alloc(m_); aref_.alloc(m_); atref_.alloc(n_);
        for (int q = 0; q < 2; ++q) {
            kx_[q].alloc(n_);
            ky_[q].alloc(m_);
        }
        lo_.alloc(n_); hi_.alloc(n_); gref_.alloc(n_);
        rlo_.alloc(m_); rhi_.alloc(m_); yoff_.alloc(m_);
        ib_.lo = lo_.p; ib_.hi = hi_.p; ib_.gref = gref_.p;
        ib_.rlo = rlo_.p; ib_.rhi = rhi_.p; ib_.yoff = yoff_.p;
    } else {
        ib_.ax = sax_.p;
        ib_.acx = sacx_.p;
        ib_.lo = scl_.p; ib_.hi = scu_.p; ib_.gref = sc_.p;
        ib_.rlo = srmin_.p; ib_.rhi = srmax_.p; ib_.yoff = nullptr;
    }

    kd_.m = m_;
    kd_.n = n_;
    kd_.ap = ap_.p; kd_.ai = ai_.p; kd_.cp = cp_.p; kd_.ci = ci_.p;
    kd_.sax = sax_.p; kd_.sacx = sacx_.p;
    kd_.rbA = rA_; kd_.rbT = rT_;
    kd_.c = c_.p; kd_.cl = cl_.p; kd_.cu = cu_.p;
    kd_.rmin = rmin_.p; kd_.rmax = rmax_.p;
    kd_.dr = dr_.p; kd_.dc = dc_.p;
    kd_.cmax = cmax;
    kd_.partC = partC_.p; kd_.partR = partR_.p; kd_.partR2 = partR2_.p;
    kd_.rowres = rowres_.p;

    {
        const int sr = cfg_.small_runner;
        const bool unsliced = !ib_.sT.S && !ib_.sA.S;
        if (cfg_.exact) {
            // exact mode: the exact runner (the reference's summation orders)
            if constexpr (!F32)
                runner_ = sr != 0 && dev::run_exact_fits(n_, m_, nnz) && unsliced;
        } else {
            // default mode: the register / LDS runner (small.hip)
            small_ = sr != 0 && dev::small_fits(n_, m_, nnz, F32) && unsliced;
            if (small_) {
                make_ell(eT_, epermT_, eoffT_, eidxT_, evalT_, n_, lp_.cp, ib_.ci, ib_.acx);
                make_ell(eA_, epermA_, eoffA_, eidxA_, evalA_, m_, lp_.ap, ib_.ai, ib_.ax);
            }
        }
    }
    if (cfg_.exact) {
        xdx_.alloc(n_);
        xdy_.alloc(m_);
        xtm_.alloc(m_);
        xpart_.alloc(3 * dev::DOTS_MAX_BLOCKS);
        powtab_.alloc(2 * EXACT_POW_CAP);
        for (int q = 0; q < 2; ++q) {
            kata_[q].alloc(n_);
            ktm_[q].alloc(m_);
        }
        kterm_.alloc(2 * (2 * static_cast<size_t>(n_) + m_));
        eb_.dx = xdx_.p;
        eb_.dy = xdy_.p;
        eb_.tm = xtm_.p;
        eb_.part = xpart_.p;
        eb_.powtab = powtab_.p;
        eb_.pow_cap = EXACT_POW_CAP;
    }

    if (small_ && cfg_.graph_slots > 0) {
        HIPCK(hipHostMalloc(reinterpret_cast<void**>(&hck_), sizeof(CheckOut)));
        HIPCK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        HIPCK(dev::launch_small_run(ib_, eT_, eA_, RUNNER_CAP, s_));
        dev::launch_flush_state(ib_, s_);
        kd_.clear_mix = st_.p;  // k_kkt_done clears the flag in the graph
        kkt_launch(0);
        kd_.clear_mix = nullptr;
        HIPCK(hipMemcpyAsync(&hck_->st, st_.p, sizeof(LsState), hipMemcpyDeviceToHost, s_));
        HIPCK(hipMemcpyAsync(hck_->met, met_.p, sizeof(hck_->met), hipMemcpyDeviceToHost, s_));
        HIPCK(hipStreamEndCapture(s_, &cgraph_));
        HIPCK(hipGraphInstantiate(&cexec_, cgraph_, nullptr, nullptr, 0));
    }
    const int G = (runner_ || small_) ? 0 : cfg_.graph_slots;
    if (G > 0) {  // G attempt slots, replayed until the state freezes
        HIPCK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        for (int g = 0; g < G; ++g) attempt(s_);
        HIPCK(hipStreamEndCapture(s_, &graph_));
        HIPCK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    }
    sync();
}

// Column slicing of the product that gathers a vector of `cols` entries
// over `rows` rows: used when that vector (cols * sizeof(V) bytes)
// outgrows what an XCD's 4 MB L2 keeps under the gathers (3 MB and up) and
// the matrix has enough nonzeros (8M) that every slice still fills the GPU.
// Every slice adds a pass over the rows (pointers and partial sums), so the
// slice size follows the row length: 0.5 MB of the vector per slice when
// rows average 128 nonzeros or more, 1.6 MB otherwise, 2 to 8 slices.  On
// the 1M-column Zipf LP (m = 400k, 68.5M nonzeros; A rows average 171,
// A' rows 68.5) this gives 8 slices of x, which took the A xbar step from
// 450 to 401 us (FP64) and 382 to 285 us (FP32; 4 slices: 326 us), and 2
// slices of the FP64 y, which took the A'y step from 467 to 416 us (4
// slices: 435 us, 6: 463 us).
template <typename V>
int Engine<V>::slices_for(int cols, int rows, long nnz) const {
    if (cfg_.exact) return 0;
    const int k = cfg_.column_slices;
    if (k >= 0) return std::min(k, dev::MAX_SLICES);
    const double bytes = static_cast<double>(cols) * sizeof(V);
    if (bytes < 3.0e6 || nnz < 8000000 || rows <= 0) return 0;
    const double per = static_cast<double>(nnz) / rows >= 128.0 ? 0.5e6 : 1.6e6;
    return static_cast<int>(std::clamp(std::lround(bytes / per), 2L, 8L));
}

// The runner's sliced-ELL copy of a CSR (CSC) pattern: the map is built on
// the host from the pointers, the indices and scaled values are gathered on
// the device.
template <typename V>
void Engine<V>::make_ell(dev::SmallEll<V>& E, DBuf<int>& perm, DBuf<int>& off,
                         DBuf<unsigned short>& idx, DBuf<V>& val, int rows,
                         const std::vector<int>& ptr, const int* didx, const V* dval) {
    std::vector<int> hperm, hoff, hsrc;
    const long total = dev::small_ell_map(rows, ptr.data(), dev::small_chunk<V>(), hperm, hoff, hsrc);
    DBuf<int> src;
    perm.upload(hperm.data(), hperm.size(), s_);
    off.upload(hoff.data(), hoff.size(), s_);
    src.upload(hsrc.data(), hsrc.size(), s_);
    idx.alloc(total);
    val.alloc(total);
    HIPCK(dev::small_ell_fill(src.p, total, didx, dval, idx.p, val.p, s_));
    sync();  // src and the host maps go out of scope
    E.perm = perm.p;
    E.off = off.p;
    E.idx = idx.p;
    E.val = val.p;
}

template <typename V>
void Engine<V>::make_slices(dev::Sliced<V>& sl, SliceBufs& B, int rows, int cols,
                            const std::vector<int>& ptr, const std::vector<int>& idx,
                            const double* sval, int S, bool gather_order) {
    const dev::HostSlices h =
        dev::build_slices(rows, cols, ptr.data(), idx.data(), S, gather_order);
    const long nnz = static_cast<long>(h.idx.size());
    B.ptr.upload(h.ptr.data(), h.ptr.size(), s_);
    B.idx.upload(h.idx.data(), h.idx.size(), s_);
    B.quad.upload(h.quad.data(), h.quad.size(), s_);
    B.bslice.upload(h.bslice.data(), h.bslice.size(), s_);
    if (gather_order) B.slot.upload(h.slot.data(), h.slot.size(), s_);
    B.val.alloc(nnz);
    B.part.alloc(static_cast<size_t>(S) * rows);
    {
        DBuf<int> src;
        src.upload(h.src.data(), h.src.size(), s_);
        dev::gather_values(sval, src.p, B.val.p, nnz, s_);
        sync();  // src is freed on scope exit
    }
    sl.S = S;
    sl.rows = rows;
    sl.nblk = static_cast<int>(h.bslice.size());
    sl.sb[0] = 0;
    for (int sc = 0, k = 0; sc < S; ++sc) {
        while (k < sl.nblk && h.bslice[k] == sc) ++k;
        sl.sb[sc + 1] = k;
    }
    if (cfg_.xcd_pinned) {
        const std::vector<int> xm = dev::build_xcd_map(h.bslice, S);
        B.xmap.upload(xm.data(), xm.size(), s_);
        sync();  // xm is freed on scope exit
        sl.xmap = B.xmap.p;
        sl.xgrid = static_cast<int>(xm.size());
    }
    sl.ptr = B.ptr.p;
    sl.idx = B.idx.p;
    sl.val = B.val.p;
    sl.slot = gather_order ? B.slot.p : nullptr;
    sl.quad = reinterpret_cast<const int4*>(B.quad.p);
    sl.bslice = B.bslice.p;
    sl.part = B.part.p;
}

// Fresh start of a solve on the resident data: x = y = 0 (pdhg.cu:576-580).
template <typename V>
void Engine<V>::reset() {
    for (DBuf<V>* b : {&x0_, &x1_, &xbar_, &ata_, &xavg_, &y0_, &y1_, &yavg_, &axcur_, &adx_})
        zero_v(b->p, b->n);
    if constexpr (F32) {
        dev::fill(xref_.p, n_, 0.0, s_);
        dev::fill(yref_.p, m_, 0.0, s_);
        recompute_refs();
    }
}

// FP32 engine: residual data of the correction problem at (x_ref, y_ref).
template <typename V>
void Engine<V>::recompute_refs() {
    if constexpr (F32) {
        dev::spmv(rA_, ap_.p, ai_.p, sax_.p, xref_.p, aref_.p, s_);
        dev::spmv(rT_, cp_.p, ci_.p, sacx_.p, yref_.p, atref_.p, s_);
        dev::refine_data(m_, n_, scl_.p, scu_.p, sc_.p, xref_.p, atref_.p,
                         srmin_.p, srmax_.p, aref_.p, yref_.p, lo_.p, hi_.p,
                         gref_.p, rlo_.p, rhi_.p, yoff_.p, s_);
    }
}

template <typename V>
void Engine<V>::check_metrics(int cur, Metrics& mc, Metrics& ma) {
    kkt_launch(cur);
    Metrics h[2];
    HIPCK(hipMemcpyAsync(h, met_.p, sizeof(h), hipMemcpyDeviceToHost, s_));
    sync();
    mc = h[0];
    ma = h[1];
}

// Both iterates' KKT metrics into met_ (device work only).
template <typename V>
void Engine<V>::kkt_launch(int cur) {
    if constexpr (F32) {
        // both FP64 iterates (reference + FP32 correction) formed once, so
        // the KKT SpMVs gather one double per nonzero instead of a double
        // and a float (same values)
        dev::Combine4 c4;
        c4.xref = xref_.p;
        c4.yref = yref_.p;
        c4.xd[0] = xcur(cur);
        c4.xd[1] = xavg_.p;
        c4.yd[0] = ycur(cur);
        c4.yd[1] = yavg_.p;
        for (int q = 0; q < 2; ++q) {
            c4.ox[q] = kx_[q].p;
            c4.oy[q] = ky_[q].p;
        }
        c4.n = n_;
        c4.m = m_;
        c4.snap = snap_x();
        dev::combine4(c4, s_);
        dev::KktIter<double> it;
        for (int q = 0; q < 2; ++q) {
            it.xd[q] = kx_[q].p;
            it.yd[q] = ky_[q].p;
        }
        dev::launch_kkt(kd_, it, met_.p, s_);
    } else if (cfg_.exact) {
        dev::ExactKkt k;
        k.x[0] = xcur(cur);
        k.y[0] = ycur(cur);
        k.x[1] = xavg_.p;
        k.y[1] = yavg_.p;
        for (int q = 0; q < 2; ++q) {
            k.ata[q] = kata_[q].p;
            k.tm[q] = ktm_[q].p;
        }
        k.term = kterm_.p;
        dev::launch_kkt_exact(kd_, k, met_.p, s_);
    } else {
        dev::KktIter<V> it;
        it.xd[0] = xcur(cur);
        it.yd[0] = ycur(cur);
        it.xd[1] = xavg_.p;
        it.yd[1] = yavg_.p;
        dev::launch_kkt(kd_, it, met_.p, s_);
    }
}

// Device dual repair of the scaled point in tmpn_ / tmpm_ (its metrics t):
// both column tolerance tiers of pdhg.cu:897-925, accepted when the
// repaired pair's gap is certified (x, hence pinf and prow, unchanged).
// On success rx, ry receive the pair in original units.
template <typename V>
bool Engine<V>::device_repair(const Metrics& t, double tol, Metrics& mr,
                              std::vector<double>& rx, std::vector<double>& ry) {
    if (!rpx_.p) {
        rpx_.alloc(n_);
        rpy_.alloc(m_);
        rpyit_.alloc(m_);
        repair_.alloc(m_, n_);
    }
    dev::unscale_out(m_, n_, dc_.p, dr_.p, nullptr, tmpn_.p, nullptr, nullptr, tmpm_.p, nullptr,
                     rpx_.p, rpyit_.p, s_);
    dev::RepairData D;
    D.ap = ap_.p; D.ai = ai_.p; D.ax = ax_.p;
    D.cp = cp_.p; D.ci = ci_.p; D.acx = acx_.p;
    D.rbA = rA_; D.rbT = rT_;
    D.c = c_.p; D.cl = cl_.p; D.cu = cu_.p; D.rmin = rmin_.p; D.rmax = rmax_.p;
    static const double ctol[2] = {1e-6, 1e-1};
    for (int tier = 0; tier < 2; ++tier) {
        dev::RepairInfo info;
        const bool have = repair_.run(D, rpx_.p, rpyit_.p, ctol[tier], t.op, kd_.cmax, s_,
                                      rpy_.p, &info);
        repair_cg_ += info.cg_iterations;
        repair_its_ += info.iterations;
        if (!have || !(t.pinf <= tol && t.prow <= tol && std::isfinite(info.gap) &&
                       info.gap <= tol))
            continue;
        mr = Metrics{t.pinf, info.dinf, info.gap, t.op, info.od, t.prow};
        rx.resize(n_);
        ry.resize(m_);
        HIPCK(hipMemcpyAsync(rx.data(), rpx_.p, sizeof(double) * n_, hipMemcpyDeviceToHost, s_));
        HIPCK(hipMemcpyAsync(ry.data(), rpy_.p, sizeof(double) * m_, hipMemcpyDeviceToHost, s_));
        sync();
        return true;
    }
    return false;
}

// Exact mode: k^-0.3 and k^-0.6 for k = 1..maxit with the host's pow, as
// the reference's host controller computes them (pdhg.cu:780-781).
template <typename V>
void Engine<V>::fill_powtab(long long maxit) {
    if (maxit > EXACT_POW_CAP)
        throw std::invalid_argument(
            "exact mode runs at most 2097152 iterations per solve (set "
            "time_limit_s >= 1e6 or <= 0 and max_iterations within that)");
    if (maxit <= pow_filled_) return;
    std::vector<double> h(2 * static_cast<size_t>(maxit - pow_filled_));
    for (long long k = pow_filled_ + 1; k <= maxit; ++k) {
        const double kk = static_cast<double>(k);
        h[2 * (k - pow_filled_ - 1)] = std::pow(kk, -0.3);
        h[2 * (k - pow_filled_ - 1) + 1] = std::pow(kk, -0.6);
    }
    HIPCK(hipMemcpyAsync(powtab_.p + 2 * pow_filled_, h.data(), sizeof(double) * h.size(),
                         hipMemcpyHostToDevice, s_));
    sync();
    pow_filled_ = maxit;
}

template <typename V>
Solution Engine<V>::run(const Options& opt_, SolveInfo* info) {
    Solution sol;
    const double INF = std::numeric_limits<double>::infinity();
    reset();

    const long MAXIT = (opt_.time_limit_s > 0 && opt_.time_limit_s < 1e6)
                           ? 20000000L
                           : opt_.max_iterations;
    const double TOL = opt_.tolerance;
    if (cfg_.exact) fill_powtab(MAXIT);

    LsState st{};
    st.eta = eta0_;
    st.omega = omega0_;
    st.k = 1;
    st.hn = 0;
    st.maxit = MAXIT;
    st.fresh = 1;
    st.cur = 0;
    st.frozen = MAXIT < 1 ? 1 : 0;
    write_state(st);
    sync();

    const int G = cfg_.graph_slots;
    auto launch_slots = [&](long slots) {
        if constexpr (!F32) {
            if (runner_) {  // attempts until the freeze, in one launch
                dev::launch_run_exact(ib_, eb_, RUNNER_CAP, s_);
                return;
            }
        }
        if (small_) {  // attempts until the freeze, in one launch
            HIPCK(dev::launch_small_run(ib_, eT_, eA_, RUNNER_CAP, s_));
            return;
        }
        if (G > 0) {
            const long reps = (slots + G - 1) / G;
            for (long r = 0; r < reps; ++r) HIPCK(hipGraphLaunch(exec_, s_));
        } else {
            for (long r = 0; r < slots; ++r) attempt(s_);
        }
    };

    RestartState rs{-1.0, INF, 0};
    bool have_pe = false;
    Metrics last{};
    bool last_is_avg = false;
    Metrics inc{};
    bool have_inc = false;
    double inc_op = 0.0;
    long last_rollback = 0;
    long next_repair = 5000, repair_spacing = 2000;
    bool repaired = false, time_limit = false;
    long iters_run = 0, restarts = 0, checks = 0, refinements = 0, rollbacks = 0;
    std::vector<double> rep_x, rep_y;  // repaired pair (original units)
    float kernel_ms_total = 0.0f;

    const double t0 = now_s();
    while (true) {
        bool have_met = false;  // the check graph brought the metrics
        Metrics gmet[2];
        if (!st.frozen) {
            // attempts until the device freezes at the next check boundary
            HIPCK(hipEventRecord(ev0_, s_));
            if (cexec_) {
                HIPCK(hipGraphLaunch(cexec_, s_));
                sync();
                st = hck_->st;  // flushed; cur = 0
                if (st.frozen) {
                    have_met = true;
                    gmet[0] = hck_->met[0];
                    gmet[1] = hck_->met[1];
                }
            }
            while (!st.frozen) {
                const long k = st.k;
                long next = ((k - 1) / CHECK_EVERY + 1) * CHECK_EVERY;
                if (next > MAXIT) next = MAXIT;
                launch_slots(std::max(1L, next - k + 1));
                st = read_state();
            }
            HIPCK(hipEventRecord(ev1_, s_));
            HIPCK(hipEventSynchronize(ev1_));
            float ms = 0.0f;
            HIPCK(hipEventElapsedTime(&ms, ev0_, ev1_));
            kernel_ms_total += ms;
        }
        const long kdone = st.k - 1;
        iters_run = kdone;
        if (st.mix_pending) {
            if constexpr (F32) dev::launch_flush(ib_, st.cur, st.mix_w, s_);
            else if (cfg_.exact) dev::launch_flush_exact(ib_, st.cur, st.mix_w, s_);
            else dev::launch_flush(ib_, st.cur, st.mix_w, s_);
            st.mix_pending = 0;
        }
        if (kdone <= 0 || kdone % CHECK_EVERY != 0) break;  // MAXIT reached
        const int cur = st.cur;
        ++checks;

        // ---- check (pdhg.cu:801-1084)
        Metrics mc, ma;
        if (have_met) {
            mc = gmet[0];
            ma = gmet[1];
        } else {
            check_metrics(cur, mc, ma);
        }
        const double ec = mc.finite_res() ? mc.err() : INF;
        const double ea = ma.finite_res() ? ma.err() : INF;
        last = (ec <= ea) ? mc : ma;
        last_is_avg = (ea < ec);
        if (opt_.verbosity > 0 && kdone % (CHECK_EVERY * 40) == 0)
            std::fprintf(stderr, "[pdhg] k=%ld pinf=%.3e dinf=%.3e gap=%.3e\n",
                         kdone, last.pinf, last.dinf, last.gap);
        if (passes(mc, TOL)) { last = mc; last_is_avg = false; break; }
        if (passes(ma, TOL)) { last = ma; last_is_avg = true; break; }

        if (kdone >= 5000 && kdone >= next_repair) {  // pdhg.cu:854-931
            next_repair = kdone + repair_spacing;
            repair_spacing = std::min(repair_spacing * 3, 30000L);
            for (int cc = 1; cc <= 2; ++cc) {
                const Metrics& t = cc == 1 ? ma : inc;
                if (cc == 2 && !have_inc) break;
                if (!(t.pinf <= TOL && t.prow <= TOL && t.dinf <= 10.0 * TOL &&
                      !(std::isfinite(t.gap) && t.gap <= TOL)))
                    continue;
                if (cc == 1) {
                    full_x(tmpn_.p, F32 ? xref_.p : nullptr, xavg_.p);
                    full(tmpm_.p, F32 ? yref_.p : nullptr, yavg_.p, m_);
                } else {
                    copy_d(tmpn_.p, xb_.p, n_);
                    full(tmpm_.p, F32 ? yref_.p : nullptr, ycur(cur), m_);
                }
                if (m_ > cfg_.max_repair_m) {
                    // device repair (any m): the same schedule and tiers
                    Metrics mr{};
                    if (!device_repair(t, TOL, mr, rep_x, rep_y)) continue;
                    if (opt_.verbosity > 1)
                        std::fprintf(stderr, "[pdhg] device dual repair c=%d @k=%ld gap=%.3e\n",
                                     cc, kdone, mr.gap);
                    last = mr;
                    last_is_avg = (cc == 1);
                    repaired = true;
                    break;
                }
                std::vector<double> sx(n_), sy(m_), rx(n_), ry_it(m_), ry(m_);
                HIPCK(hipMemcpyAsync(sx.data(), tmpn_.p, sizeof(double) * n_, hipMemcpyDeviceToHost, s_));
                HIPCK(hipMemcpyAsync(sy.data(), tmpm_.p, sizeof(double) * m_, hipMemcpyDeviceToHost, s_));
                std::vector<double> hdc(n_), hdr(m_);
                HIPCK(hipMemcpyAsync(hdc.data(), dc_.p, sizeof(double) * n_, hipMemcpyDeviceToHost, s_));
                HIPCK(hipMemcpyAsync(hdr.data(), dr_.p, sizeof(double) * m_, hipMemcpyDeviceToHost, s_));
                sync();
                for (int j = 0; j < n_; ++j) rx[j] = sx[j] * hdc[j];
                for (int i = 0; i < m_; ++i) ry_it[i] = 1.0 * sy[i] / hdr[i];
                std::vector<double> rAx(m_, 0.0), rd(n_);
                for (int j = 0; j < n_; ++j)
                    for (int p = lp_.cp[j]; p < lp_.cp[j + 1]; ++p)
                        rAx[lp_.ci[p]] += lp_.acx[p] * rx[j];
                static const double ctol[2] = {1e-6, 1e-1};
                bool ok_rep = false;
                Metrics mr{};
                for (int tier = 0; tier < 2 && !ok_rep; ++tier) {
                    if (!host_dual_repair(lp_, rx.data(), rAx.data(), ry_it.data(), ctol[tier], ry.data()))
                        continue;
                    for (int j = 0; j < n_; ++j) {
                        double dj = lp_.c[j];
                        for (int p = lp_.cp[j]; p < lp_.cp[j + 1]; ++p) dj += lp_.acx[p] * ry[lp_.ci[p]];
                        rd[j] = dj;
                    }
                    const HostMetrics t2 = host_evaluate(lp_, rx.data(), ry.data(), rAx.data(), rd.data());
                    if (t2.pinf <= TOL && t2.prow <= TOL && std::isfinite(t2.gap) && t2.gap <= TOL) {
                        ok_rep = true;
                        mr = Metrics{t2.pinf, t2.dinf, t2.gap, t2.op, t2.od, t2.prow};
                    }
                }
                if (!ok_rep) continue;
                rep_x = rx;
                rep_y.resize(m_);
                for (int i = 0; i < m_; ++i) rep_y[i] = 1.0 * (ry[i] * hdr[i] / 1.0) / hdr[i];
                if (opt_.verbosity > 1)
                    std::fprintf(stderr, "[pdhg] dual repair c=%d @k=%ld gap=%.3e\n", cc, kdone, mr.gap);
                last = mr;
                last_is_avg = (cc == 1);
                repaired = true;
                break;
            }
            if (repaired) break;
        }

        // incumbent (pdhg.cu:933-958)
        int inc_src = -1;
        for (int cc = 0; cc < 2; ++cc) {
            const Metrics& t = cc ? ma : mc;
            if (t.pinf <= 1e-4 && std::isfinite(t.op) && (!have_inc || t.op < inc_op)) {
                have_inc = true;
                inc_op = t.op;
                inc = t;
                inc_src = cc;
            }
        }
        if (inc_src >= 0) {
            const double* xr = F32 ? xref_.p : nullptr;
            const double* yr = F32 ? yref_.p : nullptr;
            full_x(xb_.p, xr, inc_src ? xavg_.p : xcur(cur));
            full(yb_.p, yr, inc_src ? yavg_.p : ycur(cur), m_);
        }
        // rollback (pdhg.cu:965-988)
        if (mc.pinf >= 0.5 && have_inc && kdone - last_rollback > 2000) {
            if constexpr (F32) {
                copy_d(xref_.p, xb_.p, n_);
                copy_d(yref_.p, yb_.p, m_);
                for (V* p : {x0_.p, x1_.p, xavg_.p}) zero_v(p, n_);
                for (V* p : {y0_.p, y1_.p, yavg_.p, axcur_.p}) zero_v(p, m_);
                recompute_refs();
                ++refinements;
            } else {
                copy_v(xcur(cur), xb_.p, n_);
                copy_v(ycur(cur), yb_.p, m_);
                copy_v(xavg_.p, xb_.p, n_);
                copy_v(yavg_.p, yb_.p, m_);
                spmv_rows(xcur(cur), axcur_.p);
            }
            st.eta = std::min(st.eta, 0.5 * eta0_);
            st.hn = 0;
            rs.epoch_start = kdone;
            rs.epoch_kkt = -1.0;
            rs.prev_cand = INF;
            last_rollback = kdone;
            ++rollbacks;
            if (opt_.verbosity > 1) std::fprintf(stderr, "[pdhg] rollback @k=%ld to inc\n", kdone);
        }
        // restart + primal weight (pdhg.cu:989-1075)
        const double cand = std::min(ec, ea);
        const long t_epoch = kdone - rs.epoch_start;
        if (restart_decide(rs, cand, kdone)) {
            ++restarts;
            if constexpr (F32) {
                // fold the anchor's correction into the FP64 reference
                const dev::BoundSnap b = snap_x();
                if (ea < ec) {
                    dev::fold(xref_.p, xavg_.p, n_, s_, &b);
                    dev::fold(yref_.p, yavg_.p, m_, s_);
                } else {
                    dev::fold(xref_.p, xcur(cur), n_, s_, &b);
                    dev::fold(yref_.p, ycur(cur), m_, s_);
                }
                for (V* p : {x0_.p, x1_.p, xavg_.p}) zero_v(p, n_);
                for (V* p : {y0_.p, y1_.p, yavg_.p, axcur_.p}) zero_v(p, m_);
                recompute_refs();
                ++refinements;
            } else {
                if (ea < ec) {
                    copy_v(xcur(cur), xavg_.p, n_);
                    copy_v(ycur(cur), yavg_.p, m_);
                }
                copy_v(xavg_.p, xcur(cur), n_);
                copy_v(yavg_.p, ycur(cur), m_);
                spmv_rows(xcur(cur), axcur_.p);
            }
            st.hn = 0;
            rs.epoch_start = kdone;
            rs.epoch_kkt = cand;
            rs.prev_cand = INF;
            const double* sxp;
            const double* syp;
            if constexpr (F32) {
                sxp = xref_.p;
                syp = yref_.p;
            } else {
                sxp = xcur(cur);
                syp = ycur(cur);
            }
            if (have_pe) {
                const double dxn = dsum(sxp, pesx_.p, n_);
                const double dyn = dsum(syp, pesy_.p, m_);
                st.omega = primal_weight(st.omega, dxn, dyn);
            }
            copy_d(pesx_.p, sxp, n_);
            copy_d(pesy_.p, syp, m_);
            have_pe = true;
            if (opt_.verbosity > 1)
                std::fprintf(stderr, "[pdhg] restart @k=%ld t=%ld r=%.3e eta=%.3e om=%.3e\n",
                             kdone, t_epoch, cand, st.eta, st.omega);
        }
        if (now_s() - t0 > opt_.time_limit_s) {
            sol.status = Status::TimeLimit;
            sol.message = "wall-clock budget exhausted";
            time_limit = true;
            break;
        }
        if (kdone >= MAXIT) break;
        st.frozen = 0;
        st.fresh = 1;
        st.retries = 0;
        write_state(st);
    }
    sync();
    const double secs = now_s() - t0;

    if (!time_limit) {
        if (certified(last, TOL)) {
            sol.status = Status::NearOptimal;
            sol.message = "converged: residuals + duality gap @tol";
        } else {
            sol.status = Status::IterationLimit;
            sol.message = last.pinf <= TOL && last.dinf <= TOL
                              ? "iteration cap: residuals @tol but gap not certified (dual bound -inf or > tol)"
                              : "iteration cap reached before tolerance";
        }
    }
    int src = last_is_avg ? 1 : 0;  // 0 current, 1 average, 2 incumbent
    bool used_inc = false;
    if (!certified(last, TOL) && have_inc && inc.pinf <= TOL && std::isfinite(inc.op) &&
        (inc.op < last.op || last.pinf > TOL)) {
        last = inc;
        last_is_avg = false;
        src = 2;
        used_inc = true;
    }
    sol.objective = last.op + lp_.obj_const;
    sol.pinf = last.pinf;
    sol.dinf = last.dinf;
    sol.rel_gap = last.gap;
    sol.iterations = iters_run;
    sol.solve_time_ms = secs * 1000.0;
    sol.x.assign(n_, 0.0);
    sol.y.assign(m_, 0.0);
    if (repaired && !used_inc) {
        sol.x = rep_x;
        sol.y = rep_y;
    } else {
        DBuf<double> ox, oy;
        ox.alloc(n_);
        oy.alloc(m_);
        const double* xr = F32 ? xref_.p : nullptr;
        const double* yr = F32 ? yref_.p : nullptr;
        if (src == 2) {
            dev::unscale_out(m_, n_, dc_.p, dr_.p, nullptr, xb_.p, nullptr, nullptr, yb_.p, nullptr, ox.p, oy.p, s_);
        } else {
            const V* fx = src == 1 ? xavg_.p : xcur(st.cur);
            const V* fy = src == 1 ? yavg_.p : ycur(st.cur);
            if constexpr (F32) {
                const dev::BoundSnap b = snap_x();
                dev::unscale_out(m_, n_, dc_.p, dr_.p, xr, nullptr, fx, yr, nullptr, fy, ox.p,
                                 oy.p, s_, &b);
            }
            else
                dev::unscale_out(m_, n_, dc_.p, dr_.p, nullptr, fx, nullptr, nullptr, fy, nullptr, ox.p, oy.p, s_);
        }
        if (n_) HIPCK(hipMemcpyAsync(sol.x.data(), ox.p, sizeof(double) * n_, hipMemcpyDeviceToHost, s_));
        if (m_) HIPCK(hipMemcpyAsync(sol.y.data(), oy.p, sizeof(double) * m_, hipMemcpyDeviceToHost, s_));
        sync();
    }
    if (!colperm_.empty()) {  // back to the model's column order
        std::vector<double> xo(n_);
        for (int k = 0; k < n_; ++k) xo[colperm_[k]] = sol.x[k];
        sol.x.swap(xo);
    }
    if (info) {
        info->setup_ms = setup_ms_;
        info->solve_ms = secs * 1000.0;
        info->attempts = static_cast<long>(st.attempts);
        info->restarts = restarts;
        info->checks = checks;
        info->refinements = refinements;
        info->eta = st.eta;
        info->omega = st.omega;
        info->lam2 = lam2_;
        info->eta0 = eta0_;
        info->omega0 = omega0_;
        info->last_is_avg = last_is_avg ? 1 : 0;
        info->used_incumbent = used_inc ? 1 : 0;
        info->repaired = repaired ? 1 : 0;
        info->kernel_ms = kernel_ms_total;
        info->rollbacks = rollbacks;
    }
    return sol;
}

template <typename V>
AttemptTiming Engine<V>::profile(int attempts) {
    if (cfg_.exact) throw std::invalid_argument("attempt timing is not offered in exact mode");
    reset();
    LsState st{};
    st.eta = eta0_;
    st.omega = omega0_;
    st.k = 1;
    st.maxit = 1LL << 60;
    st.fresh = 1;
    st.no_freeze = 1;
    write_state(st);
    sync();
    struct Events {  // destroyed on every exit, thrown errors included
        hipEvent_t ev[3] = {};
        ~Events() {
            for (hipEvent_t x : ev)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    hipEvent_t* e = evs.ev;
    for (int k = 0; k < 3; ++k) HIPCK(hipEventCreate(&e[k]));
    AttemptTiming t;
    t.nblk_primal = rT_.nblk;
    t.nblk_dual = rA_.nblk;
    double a = 0.0, b = 0.0;
    const int warm = 3;
    for (int it = 0; it < warm + attempts; ++it) {
        HIPCK(hipEventRecord(e[0], s_));
        dev::launch_primal(ib_, s_);
        HIPCK(hipEventRecord(e[1], s_));
        dev::launch_dual(ib_, s_);
        HIPCK(hipEventRecord(e[2], s_));
        dev::launch_decide(ib_, s_);
        HIPCK(hipEventSynchronize(e[2]));
        float ms1 = 0.0f, ms2 = 0.0f;
        HIPCK(hipEventElapsedTime(&ms1, e[0], e[1]));
        HIPCK(hipEventElapsedTime(&ms2, e[1], e[2]));
        if (it >= warm) {
            a += ms1;
            b += ms2;
        }
    }
    t.ms_primal = a / attempts;
    t.ms_dual = b / attempts;
    return t;
}

}  // namespace

std::unique_ptr<EngineBase> make_engine(const io::Model& model,
                                        const EngineConfig& cfg) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg.device) {
        (void)hipGetLastError();
        throw std::runtime_error(
            "PDHG engine unavailable: no HIP GPU device visible "
            "(this build runs the engine on MI355X via HIP, not CUDA)");
    }
    if (model.m < 0 || model.n < 0 || static_cast<int>(model.ap.size()) != model.m + 1 ||
        static_cast<int>(model.cp.size()) != model.n + 1)
        throw std::invalid_argument("malformed model: CSR/CSC pointer sizes do not match m, n");
    // the primal step's launch covers the rows of A' (one per column)
    if (model.n == 0) throw std::invalid_argument("the LP has no columns (n = 0)");
    if (cfg.exact && cfg.precision != 64)
        throw std::invalid_argument("exact mode is an FP64 mode (precision 64)");
    if (cfg.precision == 32) return std::make_unique<Engine<float>>(model, cfg);
    return std::make_unique<Engine<double>>(model, cfg);
}

AttemptTiming profile_attempts(const io::Model& model, const EngineConfig& cfg,
                               int attempts) {
    return make_engine(model, cfg)->profile(attempts < 1 ? 1 : attempts);
}

Solution solve_ex(const io::Model& model, const Options& options,
                  const EngineConfig& cfg, SolveInfo* info) {
    try {
        auto e = make_engine(model, cfg);
        return e->run(options, info);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[pdhg] %s\n", e.what());
        Solution sol;
        sol.status = Status::Error;
        sol.message = e.what();
        return sol;
    }
}

Solution solve(const io::Model& model, const Options& options) {
    Solution sol = solve_ex(model, options, EngineConfig{}, nullptr);
    if (sol.status != Status::Error && sol.x.size() == static_cast<size_t>(model.n)) {
        sol.row_activity.resize(model.m);
        io::row_activity(model, sol.x.data(), sol.row_activity.data());
    }
    return sol;
}

}  // namespace igaos::pdhg
