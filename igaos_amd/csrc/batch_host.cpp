// Host side of the batched persistent PDHG solver: packs a batch of LPs
// into concatenated device arrays, picks a per-LP SpMV schedule, launches
// the persistent kernel once per run and returns per-instance results.
// C entry points: igaos_batch_* in include/igaos_pdhg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "adjoint.hpp"
#include "batch.hpp"
#include "unroll.hpp"
#include "igaos_pdhg.h"
#include "model.hpp"
#include "pdhg_kernels.hpp"

namespace {

#define BCK(expr)                                                            \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess)                                                \
            throw std::runtime_error(std::string("HIP error in batched PDHG: ") + \
                                     hipGetErrorString(e_) + " at " #expr);  \
    } while (0)

struct Mem {
    std::vector<void*> ptrs;
    ~Mem() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T* alloc(size_t k) {
        void* p = nullptr;
        BCK(hipMalloc(&p, sizeof(T) * (k > 0 ? k : 1)));
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    template <typename T>
    T* upload(const T* h, size_t k) {
        T* d = alloc<T>(k);
        if (k) BCK(hipMemcpy(d, h, sizeof(T) * k, hipMemcpyHostToDevice));
        return d;
    }
};

using igaos::batch::LPDesc;
using igaos::batch::LPResult;

}  // namespace

struct igaos_batch {
    int count = 0;
    int precision = 32;
    long sum_m = 0, sum_n = 0, sum_z = 0;
    std::vector<LPDesc> desc;
    std::vector<int> order;
    std::vector<int> rbA, rbT;  // concatenated per-instance row-block tables
    Mem mem;
    LPDesc* d_desc = nullptr;
    int* d_order = nullptr;
    int* d_queue = nullptr;
    LPResult* d_res = nullptr;
    igaos::batch::Arrays<float> a32{};
    igaos::batch::Arrays<double> a64{};
    igaos::batch::Arrays<float>* d_a32 = nullptr;   // device copies
    igaos::batch::Arrays<double>* d_a64 = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    unsigned int* h_prog = nullptr;  // pinned, coherent: per-workgroup progress
    unsigned int* d_prog = nullptr;
    bool running = false;
    double last_ms = 0.0;
    int grid = 0;
    int waves = 2;
    int cus = 1;
    bool solved = false;  // a run finished and the data are unchanged since
    int device = 0;
    int mode = 0;         // 0: persistent kernel, 1: lockstep rounds
    bool idx16 = false;   // lockstep kernels read 16-bit column indices
    std::thread worker;   // lockstep: drives the rounds of one run
    std::string worker_err;
    std::atomic<int> lk_done{0};
    Mem lkmem;            // lockstep buffers, allocated on first use
    void* d_ctl = nullptr;
    igaos::batch::LkEnt* d_tblT = nullptr;
    igaos::batch::LkEnt* d_tblA = nullptr;
    igaos::batch::LsRange* d_rng = nullptr;
    double* d_partT = nullptr;
    double* d_partA = nullptr;
    int* d_done = nullptr;
    int* d_ord = nullptr;  // check launch order of the active instances
    // pinned host staging of the per-period tables (asynchronous uploads)
    igaos::batch::LkEnt* h_tblT = nullptr;
    igaos::batch::LkEnt* h_tblA = nullptr;
    igaos::batch::LsRange* h_rng = nullptr;
    int* h_ord = nullptr;
    int* d_nfrozen = nullptr;
    // lockstep timing samples: rounds 0, 16, 32, 48 of each check period
    // (every active instance works in those), HIP events around the two
    // attempt kernels, with their algorithmic bytes
    hipEvent_t lk_ev[4][3] = {};
    long lk_samples = 0;
    double lk_primal_ms = 0.0, lk_dual_ms = 0.0;
    double lk_primal_bytes = 0.0, lk_dual_bytes = 0.0;
    Mem umem;             // unrolled-mode workspace, allocated on first use
    std::unique_ptr<Mem> ckmem;  // its checkpoints (sized for un_cap steps)
    igaos::batch::UnrollArrays un{};
    igaos::batch::UnrollArrays* d_un = nullptr;
    int un_cap = -1;      // steps the checkpoint slabs hold
    int un_K = -1;        // steps of the last forward on the current data (-1: none)
    Mem amem;             // adjoint workspace, allocated on first use
    igaos::batch::AdjArrays adj{};
    igaos::batch::AdjArrays* d_adj = nullptr;
    ~igaos_batch() {
        if (worker.joinable()) worker.join();
        for (auto& t : lk_ev)
            for (hipEvent_t e : t)
                if (e) (void)hipEventDestroy(e);
        if (running && e1) (void)hipEventSynchronize(e1);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (h_prog) (void)hipHostFree(h_prog);
        for (void* p : {static_cast<void*>(h_tblT), static_cast<void*>(h_tblA),
                        static_cast<void*>(h_rng), static_cast<void*>(h_ord)})
            if (p) (void)hipHostFree(p);
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

thread_local std::string g_berr;

int bfail(const std::string& w) {
    g_berr = w;
    return -1;
}

// Gather order of one matrix (A: rows ptr = ap, indices ai, blocks rbA;
// or A' alike): inside every row block of more than one row the nonzeros
// sorted by index (stable), as 16-bit indices plus each one's position in
// the block's row order.  Instances are spread over host threads.
// One matrix's row blocks rb (nb blocks, pairs of first row, first
// nonzero): within every block of more than one row the nonzeros' order
// sorted by index (stable), as 16-bit indices and slots.
void gather_order_blocks(const int* rb, int nb, const int* ix, unsigned short* si,
                         unsigned short* sl, std::vector<int>& ord) {
    for (int k = 0; k < nb; ++k) {
        const int r0 = rb[2 * k], p0 = rb[2 * k + 1];
        const int r1 = rb[2 * k + 2], p1 = rb[2 * k + 3];
        ord.resize(p1 - p0);
        std::iota(ord.begin(), ord.end(), 0);
        if (r1 - r0 > 1)
            std::stable_sort(ord.begin(), ord.end(),
                             [&](int u, int v) { return ix[p0 + u] < ix[p0 + v]; });
        for (int q = 0; q < p1 - p0; ++q) {
            si[p0 + q] = static_cast<unsigned short>(ix[p0 + ord[q]]);
            sl[p0 + q] = static_cast<unsigned short>(ord[q]);
        }
    }
}

void gather_order(const std::vector<LPDesc>& desc, const std::vector<int>& rb, bool trans,
                  const int* idx, long z, std::vector<unsigned short>& sidx,
                  std::vector<unsigned short>& slot) {
    sidx.resize(z);
    slot.resize(z);
    const int nt = static_cast<int>(
        std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::atomic<int> next{0};
    auto work = [&]() {
        std::vector<int> ord;
        for (int lp = next++; lp < static_cast<int>(desc.size()); lp = next++) {
            const LPDesc& d = desc[lp];
            gather_order_blocks(rb.data() + (trans ? d.rbto : d.rbao),
                                trans ? d.nblkT : d.nblkA, idx + d.zoff,
                                sidx.data() + d.zoff, slot.data() + d.zoff, ord);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

template <typename V>
void fill_arrays(igaos_batch& B, igaos::batch::Arrays<V>& A, const int* ap,
                 const int* ai, const double* ax, const int* cp,
                 const int* ci, const double* acx, const double* c,
                 const double* cl, const double* cu, const double* rmin,
                 const double* rmax) {
    Mem& M = B.mem;
    const long m = B.sum_m, n = B.sum_n, z = B.sum_z;
    A.rbA = M.upload(B.rbA.data(), B.rbA.size());
    A.rbT = M.upload(B.rbT.data(), B.rbT.size());
    A.ap = M.upload(ap, m + B.count);
    A.ai = M.upload(ai, z);
    A.ax = M.upload(ax, z);
    A.cp = M.upload(cp, n + B.count);
    A.ci = M.upload(ci, z);
    A.acx = M.upload(acx, z);
    A.ai16 = A.ci16 = A.ais16 = A.cis16 = A.aslot = A.cslot = nullptr;
    A.faxs = A.facxs = nullptr;
    A.apk = A.cpk = nullptr;
    bool fits = true;
    for (const LPDesc& d : B.desc) fits = fits && d.m <= 65536 && d.n <= 65536;
    if (fits && z > 0) {  // 16-bit column indices for the lockstep attempt kernels
        std::vector<unsigned short> h(z), hs;
        for (long p = 0; p < z; ++p) h[p] = static_cast<unsigned short>(ai[p]);
        A.ai16 = M.upload(h.data(), z);
        for (long p = 0; p < z; ++p) h[p] = static_cast<unsigned short>(ci[p]);
        A.ci16 = M.upload(h.data(), z);
        // block sizes are at most 2048 nonzeros except single long rows,
        // which keep their order: every slot fits 16 bits
        gather_order(B.desc, B.rbA, false, ai, z, h, hs);
        A.ais16 = M.upload(h.data(), z);
        A.aslot = M.upload(hs.data(), z);
        gather_order(B.desc, B.rbT, true, ci, z, h, hs);
        A.cis16 = M.upload(h.data(), z);
        A.cslot = M.upload(hs.data(), z);
        if constexpr (sizeof(V) == 4) {
            A.apk = M.alloc<uint2>(z);
            A.cpk = M.alloc<uint2>(z);
        } else {
            A.faxs = M.alloc<V>(z);
            A.facxs = M.alloc<V>(z);
        }
    }
    A.c = M.upload(c, n);
    A.cl = M.upload(cl, n);
    A.cu = M.upload(cu, n);
    A.rmin = M.upload(rmin, m);
    A.rmax = M.upload(rmax, m);
    A.dr = M.alloc<double>(m);
    A.dc = M.alloc<double>(n);
    A.sax = M.alloc<double>(z);
    A.sacx = M.alloc<double>(z);
    A.sc = M.alloc<double>(n);
    A.scl = M.alloc<double>(n);
    A.scu = M.alloc<double>(n);
    A.srmin = M.alloc<double>(m);
    A.srmax = M.alloc<double>(m);
    constexpr bool F32 = sizeof(V) == 4;
    if constexpr (F32) {
        A.fax = M.alloc<V>(z);
        A.facx = M.alloc<V>(z);
        A.xref = M.alloc<double>(n);
        A.yref = M.alloc<double>(m);
        A.aref = M.alloc<double>(m);
        A.atref = M.alloc<double>(n);
        A.lo = M.alloc<V>(n);
        A.hi = M.alloc<V>(n);
        A.gref = M.alloc<V>(n);
        A.rlo = M.alloc<V>(m);
        A.rhi = M.alloc<V>(m);
        A.yoff = M.alloc<V>(m);
    } else {
        A.fax = A.sax;
        A.facx = A.sacx;
        A.xref = A.yref = A.aref = A.atref = nullptr;
        A.lo = A.scl;
        A.hi = A.scu;
        A.gref = A.sc;
        A.rlo = A.srmin;
        A.rhi = A.srmax;
        A.yoff = nullptr;
    }
    for (V** p : {&A.x0, &A.x1, &A.xbar, &A.ata, &A.xavg}) *p = M.alloc<V>(n);
    for (V** p : {&A.y0, &A.y1, &A.yavg, &A.axcur, &A.adx}) *p = M.alloc<V>(m);
    A.xb = M.alloc<double>(n);
    A.yb = M.alloc<double>(m);
    A.pesx = M.alloc<double>(n);
    A.pesy = M.alloc<double>(m);
    A.xout = M.alloc<double>(n);
    A.yout = M.alloc<double>(m);
}

template <typename V>
void lockstep_loop(igaos_batch* B, const igaos::batch::Arrays<V>* A,
                   const igaos::batch::Params& P) {
    using igaos::batch::LsRange;
    hipStream_t s = B->s;
    const int count = B->count;
    BCK(hipEventRecord(B->e0, s));
    igaos::batch::lockstep_setup(A, P, B->d_ctl, B->d_done, s);
    BCK(hipGetLastError());
    std::vector<int> flags(count);
    BCK(hipMemcpyAsync(flags.data(), B->d_done, sizeof(int) * count,
                       hipMemcpyDeviceToHost, s));
    BCK(hipStreamSynchronize(s));
    std::vector<int> active;
    for (int i = 0; i < count; ++i)
        if (!flags[i]) active.push_back(i);
    long periods_done = 0;
    B->lk_done = count - static_cast<int>(active.size());
    using igaos::batch::LkEnt;
    // one entry per row block b of an instance's table rb (pairs of first
    // row, first nonzero)
    auto entry = [](int lp, const int* rb, int b) {
        LkEnt e{};
        e.lp = lp;
        e.r0 = rb[2 * b];
        e.p0 = rb[2 * b + 1];
        e.r1 = rb[2 * b + 2];
        e.p1 = rb[2 * b + 3];
        return e;
    };
    // the checks run the largest instances first (their workgroups take
    // longest): one order of all instances, filtered each period
    std::vector<int> lpt(count), pos(count, -1);
    std::iota(lpt.begin(), lpt.end(), 0);
    std::stable_sort(lpt.begin(), lpt.end(), [&](int u, int v) {
        const LPDesc& du = B->desc[u];
        const LPDesc& dv = B->desc[v];
        return du.nnz + du.m + du.n > dv.nnz + dv.m + dv.n;
    });
    while (!active.empty()) {
        // tables of the active instances' row blocks, built in pinned host
        // memory (the previous period's uploads have completed: the host
        // synchronised on the check)
        const int nact = static_cast<int>(active.size());
        int nbT = 0, nbA = 0;
        for (int a = 0; a < nact; ++a) {
            const int lp = active[a];
            const LPDesc& d = B->desc[lp];
            B->h_rng[a] = LsRange{lp, nbT, d.nblkT, nbA, d.nblkA};
            const int* rbT = B->rbT.data() + d.rbto;
            const int* rbA = B->rbA.data() + d.rbao;
            for (int b = 0; b < d.nblkT; ++b) B->h_tblT[nbT++] = entry(lp, rbT, b);
            for (int b = 0; b < d.nblkA; ++b) B->h_tblA[nbA++] = entry(lp, rbA, b);
            pos[lp] = a;
        }
        // algorithmic bytes of one round's attempt kernels over the active set
        const double wa = sizeof(V), wv = sizeof(V);
        double bp = 0.0, bd = 0.0;
        for (int lp : active) {
            const LPDesc& d = B->desc[lp];
            bp += d.nnz * (wa + 4) + 4.0 * (d.n + 1) + 7.0 * d.n * wv;
            bd += d.nnz * (wa + 4) + 4.0 * (d.m + 1) + 8.0 * d.m * wv;
        }
        int no = 0;
        for (int lp : lpt)
            if (pos[lp] >= 0) B->h_ord[no++] = pos[lp];
        for (int lp : active) pos[lp] = -1;
        if (nbT) BCK(hipMemcpyAsync(B->d_tblT, B->h_tblT, sizeof(LkEnt) * nbT,
                                    hipMemcpyHostToDevice, s));
        if (nbA) BCK(hipMemcpyAsync(B->d_tblA, B->h_tblA, sizeof(LkEnt) * nbA,
                                    hipMemcpyHostToDevice, s));
        BCK(hipMemcpyAsync(B->d_rng, B->h_rng, sizeof(LsRange) * nact, hipMemcpyHostToDevice, s));
        BCK(hipMemcpyAsync(B->d_ord, B->h_ord, sizeof(int) * nact, hipMemcpyHostToDevice, s));
        BCK(hipMemsetAsync(B->d_nfrozen, 0, sizeof(int), s));
        if (B->idx16 && periods_done == 0)  // values in gather order, once per run
            igaos::batch::lockstep_sort_values(A, B->d_desc, B->d_tblT, nbT, B->d_tblA,
                                               nbA, s);
        // attempt rounds until every active instance froze on its check
        // boundary (64 accepted iterations, plus retries)
        int chunk = igaos::pdhg::CHECK_EVERY;
        bool first = true;
        for (;;) {
            int ns = 0;
            for (int r = 0; r < chunk; ++r) {
                hipEvent_t* ev = (first && r % 16 == 0 && ns < 4) ? B->lk_ev[ns++] : nullptr;
                igaos::batch::lockstep_round(A, B->d_desc, B->d_tblT, nbT, B->d_tblA, nbA,
                                             B->d_rng, nact, B->d_ctl, B->d_partT,
                                             B->d_partA, B->d_nfrozen, ev, B->idx16, s);
            }
            BCK(hipGetLastError());
            int nf = 0;
            BCK(hipMemcpyAsync(&nf, B->d_nfrozen, sizeof(int), hipMemcpyDeviceToHost, s));
            BCK(hipStreamSynchronize(s));
            for (int k = 0; k < ns; ++k) {
                float a = 0.0f, b = 0.0f;
                BCK(hipEventElapsedTime(&a, B->lk_ev[k][0], B->lk_ev[k][1]));
                BCK(hipEventElapsedTime(&b, B->lk_ev[k][1], B->lk_ev[k][2]));
                B->lk_primal_ms += a;
                B->lk_dual_ms += b;
                B->lk_primal_bytes += bp;
                B->lk_dual_bytes += bd;
                B->lk_samples += 1;
            }
            first = false;
            if (nf >= nact) break;
            chunk = 4;
        }
        ++periods_done;
        igaos::batch::lockstep_check(A, P, B->d_rng, B->d_ord, nact, B->d_ctl, B->d_done, s);
        BCK(hipGetLastError());
        BCK(hipMemcpyAsync(flags.data(), B->d_done, sizeof(int) * nact,
                           hipMemcpyDeviceToHost, s));
        BCK(hipStreamSynchronize(s));
        std::vector<int> next;
        next.reserve(active.size());
        for (int a = 0; a < nact; ++a)
            if (!flags[a]) next.push_back(active[a]);
        B->lk_done += nact - static_cast<int>(next.size());
        active.swap(next);
    }
    BCK(hipEventRecord(B->e1, s));
}

void lockstep_alloc(igaos_batch* B) {
    // d_ctl is set last: after a failed allocation it stays null and the
    // next run tries again, keeping what was already allocated
    if (B->d_ctl) return;
    for (auto& t : B->lk_ev)
        for (hipEvent_t& e : t)
            if (!e) BCK(hipEventCreate(&e));
    Mem& M = B->lkmem;
    long bt = 0, ba = 0;
    for (const LPDesc& d : B->desc) {
        bt += d.nblkT;
        ba += d.nblkA;
    }
    unsigned char* ctl = M.alloc<unsigned char>(igaos::batch::lockstep_ctl_bytes() * B->count);
    B->d_tblT = M.alloc<igaos::batch::LkEnt>(bt);
    B->d_tblA = M.alloc<igaos::batch::LkEnt>(ba);
    B->d_rng = M.alloc<igaos::batch::LsRange>(B->count);
    B->d_partT = M.alloc<double>(bt);
    B->d_partA = M.alloc<double>(2 * ba);
    B->d_done = M.alloc<int>(B->count);
    B->d_ord = M.alloc<int>(B->count);
    if (!B->h_tblT)
        BCK(hipHostMalloc(reinterpret_cast<void**>(&B->h_tblT),
                          sizeof(igaos::batch::LkEnt) * std::max(bt, 1L)));
    if (!B->h_tblA)
        BCK(hipHostMalloc(reinterpret_cast<void**>(&B->h_tblA),
                          sizeof(igaos::batch::LkEnt) * std::max(ba, 1L)));
    if (!B->h_rng)
        BCK(hipHostMalloc(reinterpret_cast<void**>(&B->h_rng),
                          sizeof(igaos::batch::LsRange) * std::max(B->count, 1)));
    if (!B->h_ord)
        BCK(hipHostMalloc(reinterpret_cast<void**>(&B->h_ord),
                          sizeof(int) * std::max(B->count, 1)));
    B->d_nfrozen = M.alloc<int>(1);
    B->d_ctl = ctl;
}

}  // namespace

extern "C" {

const char* igaos_batch_last_error(void) { return g_berr.c_str(); }

int igaos_gather_order_build(int nrows, int ncols, const int* ptr, const int* idx,
                             unsigned short* sidx, unsigned short* slot) {
    if (nrows < 0 || ncols < 0 || ncols > 65536 || !ptr || !sidx || !slot ||
        (ptr[nrows] > 0 && !idx))
        return bfail("invalid CSR arrays (at most 65,536 columns)");
    const std::string bad = igaos::io::check_compressed(nrows, ncols, ptr, idx);
    if (!bad.empty()) return bfail(bad);
    try {
        const std::vector<int> rb = igaos::dev::build_row_blocks(ptr, nrows);
        std::vector<int> ord;
        gather_order_blocks(rb.data(), static_cast<int>(rb.size() / 2) - 1, idx, sidx, slot,
                            ord);
    } catch (const std::exception& e) {
        return bfail(e.what());
    }
    return 0;
}

int igaos_batch_set_mode(igaos_batch* B, int mode) {
    if (!B) return bfail("null batch");
    if (B->running) return bfail("batch is running");
    if (mode != 0 && mode != 1) return bfail("mode must be 0 (persistent) or 1 (lockstep)");
    B->mode = mode;
    return 0;
}

int igaos_batch_get_mode(const igaos_batch* B) { return B ? B->mode : -1; }

int igaos_batch_lockstep_timing(const igaos_batch* B, double* primal_ms,
                                double* primal_bytes, double* dual_ms,
                                double* dual_bytes, long* samples) {
    if (!B) return bfail("null batch");
    if (primal_ms) *primal_ms = B->lk_primal_ms;
    if (primal_bytes) *primal_bytes = B->lk_primal_bytes;
    if (dual_ms) *dual_ms = B->lk_dual_ms;
    if (dual_bytes) *dual_bytes = B->lk_dual_bytes;
    if (samples) *samples = B->lk_samples;
    return 0;
}

igaos_batch* igaos_batch_create(int count, const int* m, const int* n,
                                const int* nnz, const double* c,
                                const double* cl, const double* cu,
                                const double* rmin, const double* rmax,
                                const int* ap, const int* ai,
                                const double* ax, const int* cp,
                                const int* ci, const double* acx,
                                const double* obj_const, int precision) {
    if (count <= 0 || !m || !n || !nnz) {
        bfail("empty batch");
        return nullptr;
    }
    if (precision != 32 && precision != 64) {
        bfail("precision must be 32 or 64");
        return nullptr;
    }
    if (!c || !cl || !cu || !rmin || !rmax || !ap || !ai || !ax || !cp || !ci || !acx) {
        bfail("null data array");
        return nullptr;
    }
    try {
        // every instance's pattern is checked before anything is uploaded:
        // the kernels index the vectors by these arrays unchecked
        std::vector<long> rpo(count), cpo(count), zo(count);
        long r = 0, q = 0, z = 0;
        for (int i = 0; i < count; ++i) {
            if (m[i] < 0 || n[i] < 0 || nnz[i] < 0) throw std::invalid_argument("negative size");
            rpo[i] = r;
            cpo[i] = q;
            zo[i] = z;
            if (ap[r + m[i]] != nnz[i] || cp[q + n[i]] != nnz[i])
                throw std::invalid_argument("pointer arrays do not match nnz of instance " +
                                            std::to_string(i));
            r += m[i] + 1;
            q += n[i] + 1;
            z += nnz[i];
        }
        std::vector<std::string> bad(count);
        const int nt = static_cast<int>(
            std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        std::atomic<int> next{0};
        auto work = [&]() {
            for (int i = next++; i < count; i = next++) {
                try {
                    bad[i] = igaos::io::check_transpose(m[i], n[i], ap + rpo[i], ai + zo[i],
                                                        ax + zo[i], cp + cpo[i], ci + zo[i],
                                                        acx + zo[i]);
                } catch (const std::exception& e) {
                    bad[i] = e.what();
                }
            }
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
        for (int i = 0; i < count; ++i)
            if (!bad[i].empty())
                throw std::invalid_argument("instance " + std::to_string(i) + ": " + bad[i]);
    } catch (const std::exception& e) {
        bfail(e.what());
        return nullptr;
    }
    try {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
            (void)hipGetLastError();
            throw std::runtime_error("batched PDHG unavailable: no HIP GPU device visible");
        }
        auto B = std::make_unique<igaos_batch>();
        B->count = count;
        B->precision = precision;
        B->desc.resize(count);
        long ro = 0, co = 0, zo = 0, rpo = 0, cpo = 0;
        for (int i = 0; i < count; ++i) {
            LPDesc& d = B->desc[i];
            d.m = m[i];
            d.n = n[i];
            d.nnz = nnz[i];
            d.tpr = 0;  // row-block streaming; the autotuner may change it
            d.roff = ro;
            d.coff = co;
            d.zoff = zo;
            d.rpo = rpo;
            d.cpo = cpo;
            d.obj_const = obj_const ? obj_const[i] : 0.0;
            const std::vector<int> ba = igaos::dev::build_row_blocks(ap + rpo, m[i]);
            const std::vector<int> bt = igaos::dev::build_row_blocks(cp + cpo, n[i]);
            d.rbao = static_cast<long>(B->rbA.size());
            d.rbto = static_cast<long>(B->rbT.size());
            d.nblkA = static_cast<int>(ba.size() / 2) - 1;
            d.nblkT = static_cast<int>(bt.size() / 2) - 1;
            B->rbA.insert(B->rbA.end(), ba.begin(), ba.end());
            B->rbT.insert(B->rbT.end(), bt.begin(), bt.end());
            ro += m[i];
            co += n[i];
            zo += nnz[i];
            rpo += m[i] + 1;
            cpo += n[i] + 1;
        }
        B->sum_m = ro;
        B->sum_n = co;
        B->sum_z = zo;
        // largest first, so the long solves start early (LPT order)
        B->order.resize(count);
        std::iota(B->order.begin(), B->order.end(), 0);
        std::stable_sort(B->order.begin(), B->order.end(), [&](int a, int b) {
            return B->desc[a].nnz + B->desc[a].m + B->desc[a].n >
                   B->desc[b].nnz + B->desc[b].m + B->desc[b].n;
        });
        BCK(hipStreamCreateWithFlags(&B->s, hipStreamNonBlocking));
        if (precision == 32) {
            fill_arrays(*B, B->a32, ap, ai, ax, cp, ci, acx, c, cl, cu, rmin, rmax);
            B->idx16 = B->a32.ai16 != nullptr;
            B->d_a32 = B->mem.upload(&B->a32, 1);
        } else {
            fill_arrays(*B, B->a64, ap, ai, ax, cp, ci, acx, c, cl, cu, rmin, rmax);
            B->idx16 = B->a64.ai16 != nullptr;
            B->d_a64 = B->mem.upload(&B->a64, 1);
        }
        B->d_desc = B->mem.upload(B->desc.data(), count);
        B->d_order = B->mem.upload(B->order.data(), count);
        B->d_queue = B->mem.alloc<int>(1);
        B->d_res = B->mem.alloc<LPResult>(count);
        BCK(hipGetDevice(&B->device));
        hipDeviceProp_t prop;
        BCK(hipGetDeviceProperties(&prop, B->device));
        B->cus = prop.multiProcessorCount;
        const int occ = igaos::batch::batch_occupancy(precision == 32, B->waves);
        B->grid = std::min(count, B->cus * occ);
        BCK(hipEventCreate(&B->e0));
        BCK(hipEventCreate(&B->e1));
        BCK(hipHostMalloc(reinterpret_cast<void**>(&B->h_prog),
                          sizeof(unsigned int) * count,
                          hipHostMallocMapped | hipHostMallocCoherent));
        BCK(hipHostGetDevicePointer(reinterpret_cast<void**>(&B->d_prog), B->h_prog, 0));
        return B.release();
    } catch (const std::exception& e) {
        bfail(e.what());
        return nullptr;
    }
}

int igaos_batch_set_schedule(igaos_batch* B, const int* tpr, int grid,
                             int waves) {
    if (!B) return bfail("null batch");
    if (B->running) return bfail("batch is running");
    try {
        if (waves == 2 || waves == 4) {
            B->waves = waves;
            const int occ = igaos::batch::batch_occupancy(B->precision == 32, waves);
            B->grid = std::min(B->count, B->cus * occ);
        } else if (waves != 0) {
            return bfail("waves must be 0, 2 or 4");
        }
        if (tpr) {
            for (int i = 0; i < B->count; ++i) {
                const int t = tpr[i];
                if (t != 0 && t != 1 && t != 2 && t != 4 && t != 8)
                    return bfail("schedule must be 0 (row blocks) or 1, 2, 4, 8 lanes per row");
                B->desc[i].tpr = t;
            }
            BCK(hipMemcpy(B->d_desc, B->desc.data(), sizeof(LPDesc) * B->count,
                          hipMemcpyHostToDevice));
        }
        if (grid > 0) B->grid = std::min(grid, B->count);
    } catch (const std::exception& e) {
        return bfail(e.what());
    }
    return 0;
}

int igaos_batch_get_schedule(const igaos_batch* B, int* tpr, int* grid,
                             int* waves) {
    if (!B) return bfail("null batch");
    if (tpr)
        for (int i = 0; i < B->count; ++i) tpr[i] = B->desc[i].tpr;
    if (grid) *grid = B->grid;
    if (waves) *waves = B->waves;
    return 0;
}

int igaos_batch_launch(igaos_batch* B, const igaos_options* opt) {
    if (!B || !opt) return bfail("null batch or options");
    if (B->running) return bfail("batch already running");
    try {
        igaos::batch::Params P{};
        P.lps = B->d_desc;
        P.order = B->d_order;
        P.count = B->count;
        P.queue = B->d_queue;
        P.res = B->d_res;
        P.tol = opt->tolerance;
        // pdhg.cu:669-671: the iteration cap applies only without a budget
        P.maxit = (opt->time_limit_s > 0 && opt->time_limit_s < 1e6) ? 20000000LL
                                                                     : opt->max_iterations;
        P.time_limit_s = opt->time_limit_s;
        int khz = 0;
        BCK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0));
        P.ticks_per_s = static_cast<double>(khz) * 1e3;
        P.progress = B->d_prog;
        if (B->mode == 1) {
            lockstep_alloc(B);
            P.progress = nullptr;
            B->lk_done = 0;
            B->lk_samples = 0;
            B->lk_primal_ms = B->lk_dual_ms = 0.0;
            B->lk_primal_bytes = B->lk_dual_bytes = 0.0;
            B->worker_err.clear();
            B->running = true;
            B->worker = std::thread([B, P]() {
                try {
                    BCK(hipSetDevice(B->device));
                    if (B->precision == 32) lockstep_loop(B, B->d_a32, P);
                    else lockstep_loop(B, B->d_a64, P);
                } catch (const std::exception& e) {
                    B->worker_err = e.what();
                }
            });
            return 0;
        }
        std::memset(B->h_prog, 0, sizeof(unsigned int) * B->count);
        BCK(hipMemsetAsync(B->d_queue, 0, sizeof(int), B->s));
        BCK(hipEventRecord(B->e0, B->s));
        if (B->precision == 32) igaos::batch::launch_batch(B->d_a32, P, B->grid, B->waves, B->s);
        else igaos::batch::launch_batch(B->d_a64, P, B->grid, B->waves, B->s);
        BCK(hipGetLastError());
        BCK(hipEventRecord(B->e1, B->s));
        B->running = true;
    } catch (const std::exception& e) {
        return bfail(e.what());
    }
    return 0;
}

int igaos_batch_progress(const igaos_batch* B) {
    if (!B) return bfail("null batch");
    if (B->mode == 1) return B->lk_done.load();
    long done = 0;
    const volatile unsigned int* p = B->h_prog;
    for (int g = 0; g < B->grid; ++g) done += p[g];
    return static_cast<int>(done);
}

int igaos_batch_wait(igaos_batch* B, double* x, double* y,
                     igaos_batch_result* res) {
    if (!B) return bfail("null batch");
    if (!B->running) return bfail("batch not launched");
    static_assert(sizeof(igaos_batch_result) == sizeof(LPResult), "result layout");
    try {
        B->running = false;
        B->solved = false;
        if (B->worker.joinable()) {
            B->worker.join();
            if (!B->worker_err.empty()) throw std::runtime_error(B->worker_err);
        }
        BCK(hipEventSynchronize(B->e1));
        float ms = 0.0f;
        BCK(hipEventElapsedTime(&ms, B->e0, B->e1));
        B->last_ms = ms;
        const double* xo = B->precision == 32 ? B->a32.xout : B->a64.xout;
        const double* yo = B->precision == 32 ? B->a32.yout : B->a64.yout;
        if (x) BCK(hipMemcpy(x, xo, sizeof(double) * B->sum_n, hipMemcpyDeviceToHost));
        if (y) BCK(hipMemcpy(y, yo, sizeof(double) * B->sum_m, hipMemcpyDeviceToHost));
        if (res)
            BCK(hipMemcpy(res, B->d_res, sizeof(LPResult) * B->count, hipMemcpyDeviceToHost));
        B->solved = true;
    } catch (const std::exception& e) {
        return bfail(e.what());
    }
    return 0;
}

int igaos_batch_run(igaos_batch* B, const igaos_options* opt, double* x,
                    double* y, igaos_batch_result* res) {
    const int rc = igaos_batch_launch(B, opt);
    if (rc != 0) return rc;
    return igaos_batch_wait(B, x, y, res);
}

double igaos_batch_kernel_ms(const igaos_batch* B) { return B ? B->last_ms : 0.0; }

int igaos_batch_update(igaos_batch* B, const double* c, const double* cl,
                       const double* cu, const double* rmin,
                       const double* rmax, const double* ax,
                       const double* acx) {
    if (!B) return bfail("null batch");
    if (B->running) return bfail("batch is running");
    try {
        auto put = [&](const double* src, const double* dst, long k) {
            if (src && k)
                BCK(hipMemcpy(const_cast<double*>(dst), src, sizeof(double) * k,
                              hipMemcpyHostToDevice));
        };
        auto go = [&](const auto& A) {
            put(c, A.c, B->sum_n);
            put(cl, A.cl, B->sum_n);
            put(cu, A.cu, B->sum_n);
            put(rmin, A.rmin, B->sum_m);
            put(rmax, A.rmax, B->sum_m);
            put(ax, A.ax, B->sum_z);
            put(acx, A.acx, B->sum_z);
        };
        if (B->precision == 32) go(B->a32);
        else go(B->a64);
        B->solved = false;
        B->un_K = -1;
    } catch (const std::exception& e) {
        return bfail(e.what());
    }
    return 0;
}

namespace {

void unroll_alloc(igaos_batch* B, int K) {
    using igaos::batch::UNROLL_SEG;
    const long m = B->sum_m, n = B->sum_n, z = B->sum_z;
    igaos::batch::UnrollArrays& U = B->un;
    if (!B->d_un) {
        Mem& M = B->umem;
        U.segx = M.alloc<double>(static_cast<size_t>(UNROLL_SEG) * n);
        U.segy = M.alloc<double>(static_cast<size_t>(UNROLL_SEG) * m);
        for (double** p : {&U.ax, &U.xb, &U.gb, &U.tmpn, &U.xK, &U.gc, &U.gcl, &U.gcu})
            *p = M.alloc<double>(n);
        for (double** p : {&U.ay, &U.tb, &U.tmpm, &U.yK, &U.grmin, &U.grmax})
            *p = M.alloc<double>(m);
        U.cx = M.alloc<unsigned char>(n);
        U.cy = M.alloc<unsigned char>(m);
        for (double** p : {&U.tau, &U.sigma, &U.obj}) *p = M.alloc<double>(B->count);
        U.gx = M.alloc<double>(n);
        U.gy = M.alloc<double>(m);
        U.gobj = M.alloc<double>(B->count);
        U.gax = M.alloc<double>(z);
        U.sum_n = n;
        U.sum_m = m;
        B->d_un = M.alloc<igaos::batch::UnrollArrays>(1);
    }
    if (K > B->un_cap) {  // checkpoints at steps 0, 64, ... < K
        const long nck = std::max(1L, (static_cast<long>(K) + UNROLL_SEG - 1) / UNROLL_SEG);
        B->ckmem.reset();
        B->ckmem = std::make_unique<Mem>();
        U.ckx = B->ckmem->alloc<double>(nck * n);
        U.cky = B->ckmem->alloc<double>(nck * m);
        B->un_cap = K;
    }
    U.K = K;
}

void adjoint_alloc(igaos_batch* B) {
    if (B->d_adj) return;
    const long m = B->sum_m, n = B->sum_n, z = B->sum_z;
    igaos::batch::AdjArrays& J = B->adj;
    Mem& M = B->amem;
    J.mc = M.alloc<unsigned char>(n);
    J.mr = M.alloc<unsigned char>(m);
    for (double** p : {&J.z1, &J.s1, &J.p1, &J.r2, &J.t2, &J.xs}) *p = M.alloc<double>(n);
    for (double** p : {&J.z2, &J.s2, &J.p2, &J.r1, &J.t1, &J.ys}) *p = M.alloc<double>(m);
    J.gx = M.alloc<double>(n);
    J.gy = M.alloc<double>(m);
    J.gobj = M.alloc<double>(B->count);
    J.obj = M.alloc<double>(B->count);
    for (double** p : {&J.gc, &J.gcl, &J.gcu}) *p = M.alloc<double>(n);
    J.grmin = M.alloc<double>(m);
    J.grmax = M.alloc<double>(m);
    J.gax = M.alloc<double>(z);
    J.info = M.alloc<igaos::batch::AdjInfo>(B->count);
    B->d_adj = M.alloc<igaos::batch::AdjArrays>(1);
}

int check_active_set_call(igaos_batch* B, double tolerance, int max_iterations) {
    if (!B) return bfail("null batch");
    if (B->running) return bfail("batch is running");
    if (!B->solved) return bfail("needs a finished run on the current data");
    if (!(tolerance >= 0.0) || max_iterations < 0)
        return bfail("tolerance must be >= 0 and max_iterations >= 0");
    return 0;
}

}  // namespace

int igaos_batch_set_solution(igaos_batch* B, const double* x, const double* y) {
    if (!B) return bfail("null batch");
    if (B->running) return bfail("batch is running");
    if (!B->solved) return bfail("needs a finished run on the current data");
    try {
        double* xo = B->precision == 32 ? B->a32.xout : B->a64.xout;
        double* yo = B->precision == 32 ? B->a32.yout : B->a64.yout;
        if (x) BCK(hipMemcpy(xo, x, sizeof(double) * B->sum_n, hipMemcpyHostToDevice));
        if (y) BCK(hipMemcpy(yo, y, sizeof(double) * B->sum_m, hipMemcpyHostToDevice));
    } catch (const std::exception& e) {
        return bfail(e.what());
    }
    return 0;
}

int igaos_batch_polish(igaos_batch* B, double tolerance, int max_iterations,
                       double* x, double* y, double* objective,
                       igaos_adjoint_info* info) {
    if (check_active_set_call(B, tolerance, max_iterations) != 0) return -1;
    try {
        adjoint_alloc(B);
        igaos::batch::AdjArrays& J = B->adj;
        J.tol = tolerance;
        J.maxit = max_iterations;
        BCK(hipMemcpyAsync(B->d_adj, &J, sizeof(J), hipMemcpyHostToDevice, B->s));
        // objectives of rejected pairs stay as the run reported them
        std::vector<LPResult> res(B->count);
        BCK(hipMemcpy(res.data(), B->d_res, sizeof(LPResult) * B->count,
                      hipMemcpyDeviceToHost));
        std::vector<double> obj(B->count);
        for (int i = 0; i < B->count; ++i) obj[i] = res[i].objective;
        BCK(hipMemcpy(J.obj, obj.data(), sizeof(double) * B->count, hipMemcpyHostToDevice));
        if (B->precision == 32)
            igaos::batch::launch_polish(B->d_a32, B->d_desc, B->count, B->d_adj, B->s);
        else
            igaos::batch::launch_polish(B->d_a64, B->d_desc, B->count, B->d_adj, B->s);
        BCK(hipGetLastError());
        BCK(hipStreamSynchronize(B->s));
        const double* xo = B->precision == 32 ? B->a32.xout : B->a64.xout;
        const double* yo = B->precision == 32 ? B->a32.yout : B->a64.yout;
        if (x) BCK(hipMemcpy(x, xo, sizeof(double) * B->sum_n, hipMemcpyDeviceToHost));
        if (y) BCK(hipMemcpy(y, yo, sizeof(double) * B->sum_m, hipMemcpyDeviceToHost));
        if (objective)
            BCK(hipMemcpy(objective, J.obj, sizeof(double) * B->count, hipMemcpyDeviceToHost));
        if (info)
            BCK(hipMemcpy(info, J.info, sizeof(igaos::batch::AdjInfo) * B->count,
                          hipMemcpyDeviceToHost));
    } catch (const std::exception& e) {
        return bfail(e.what());
    }
    return 0;
}

int igaos_batch_unroll(igaos_batch* B, int iterations, const double* steps, double* x,
                       double* y, double* objective, double* tau, double* sigma) {
    if (!B) return bfail("null batch");
    if (B->running) return bfail("batch is running");
    if (iterations < 0) return bfail("iterations must be >= 0");
    if (steps)
        for (int i = 0; i < 2 * B->count; ++i)
            if (!(steps[i] > 0.0) || !std::isfinite(steps[i]))
                return bfail("step sizes must be positive and finite");
    try {
        B->un_K = -1;
        unroll_alloc(B, iterations);
        igaos::batch::UnrollArrays& U = B->un;
        U.fixed = steps != nullptr;
        if (steps) {
            BCK(hipMemcpyAsync(U.tau, steps, sizeof(double) * B->count,
                               hipMemcpyHostToDevice, B->s));
            BCK(hipMemcpyAsync(U.sigma, steps + B->count, sizeof(double) * B->count,
                               hipMemcpyHostToDevice, B->s));
        }
        BCK(hipMemcpyAsync(B->d_un, &U, sizeof(U), hipMemcpyHostToDevice, B->s));
        if (B->precision == 32)
            igaos::batch::launch_unroll_forward(B->d_a32, B->d_desc, B->count, B->d_un, B->s);
        else
            igaos::batch::launch_unroll_forward(B->d_a64, B->d_desc, B->count, B->d_un, B->s);
        BCK(hipGetLastError());
        BCK(hipStreamSynchronize(B->s));
        auto out = [&](double* dst, const double* src, long k) {
            if (dst) BCK(hipMemcpy(dst, src, sizeof(double) * k, hipMemcpyDeviceToHost));
        };
        out(x, U.xK, B->sum_n);
        out(y, U.yK, B->sum_m);
        out(objective, U.obj, B->count);
        out(tau, U.tau, B->count);
        out(sigma, U.sigma, B->count);
        B->un_K = iterations;
    } catch (const std::exception& e) {
        return bfail(e.what());
    }
    return 0;
}

int igaos_batch_unroll_backward(igaos_batch* B, const double* gx, const double* gy,
                                const double* gobj, double* gc, double* grmin,
                                double* grmax, double* gcl, double* gcu, double* gax) {
    if (!B) return bfail("null batch");
    if (B->running) return bfail("batch is running");
    if (B->un_K < 0) return bfail("needs igaos_batch_unroll on the current data first");
    try {
        igaos::batch::UnrollArrays& U = B->un;
        auto in = [&](const double* src, const double* dst, long k) {
            double* d = const_cast<double*>(dst);
            if (src) BCK(hipMemcpyAsync(d, src, sizeof(double) * k, hipMemcpyHostToDevice, B->s));
            else BCK(hipMemsetAsync(d, 0, sizeof(double) * k, B->s));
        };
        in(gx, U.gx, B->sum_n);
        in(gy, U.gy, B->sum_m);
        in(gobj, U.gobj, B->count);
        U.K = B->un_K;
        BCK(hipMemcpyAsync(B->d_un, &U, sizeof(U), hipMemcpyHostToDevice, B->s));
        if (B->precision == 32)
            igaos::batch::launch_unroll_backward(B->d_a32, B->d_desc, B->count, B->d_un, B->s);
        else
            igaos::batch::launch_unroll_backward(B->d_a64, B->d_desc, B->count, B->d_un, B->s);
        BCK(hipGetLastError());
        BCK(hipStreamSynchronize(B->s));
        auto out = [&](double* dst, const double* src, long k) {
            if (dst) BCK(hipMemcpy(dst, src, sizeof(double) * k, hipMemcpyDeviceToHost));
        };
        out(gc, U.gc, B->sum_n);
        out(gcl, U.gcl, B->sum_n);
        out(gcu, U.gcu, B->sum_n);
        out(grmin, U.grmin, B->sum_m);
        out(grmax, U.grmax, B->sum_m);
        out(gax, U.gax, B->sum_z);
    } catch (const std::exception& e) {
        return bfail(e.what());
    }
    return 0;
}

int igaos_batch_adjoint(igaos_batch* B, const double* gx, const double* gy,
                        const double* gobj, double tolerance,
                        int max_iterations, double* gc, double* grmin,
                        double* grmax, double* gcl, double* gcu, double* gax,
                        igaos_adjoint_info* info) {
    if (check_active_set_call(B, tolerance, max_iterations) != 0) return -1;
    static_assert(sizeof(igaos_adjoint_info) == sizeof(igaos::batch::AdjInfo),
                  "adjoint info layout");
    try {
        const long m = B->sum_m, n = B->sum_n, z = B->sum_z;
        adjoint_alloc(B);
        igaos::batch::AdjArrays& J = B->adj;
        auto in = [&](const double* src, const double* dst, long k) {
            double* d = const_cast<double*>(dst);
            if (src) BCK(hipMemcpyAsync(d, src, sizeof(double) * k, hipMemcpyHostToDevice, B->s));
            else BCK(hipMemsetAsync(d, 0, sizeof(double) * k, B->s));
        };
        in(gx, J.gx, n);
        in(gy, J.gy, m);
        in(gobj, J.gobj, B->count);
        J.tol = tolerance;
        J.maxit = max_iterations;
        BCK(hipMemcpyAsync(B->d_adj, &J, sizeof(J), hipMemcpyHostToDevice, B->s));
        if (B->precision == 32)
            igaos::batch::launch_adjoint(B->d_a32, B->d_desc, B->count, B->d_adj, B->s);
        else
            igaos::batch::launch_adjoint(B->d_a64, B->d_desc, B->count, B->d_adj, B->s);
        BCK(hipGetLastError());
        BCK(hipStreamSynchronize(B->s));
        auto out = [&](double* dst, const double* src, long k) {
            if (dst) BCK(hipMemcpy(dst, src, sizeof(double) * k, hipMemcpyDeviceToHost));
        };
        out(gc, J.gc, n);
        out(gcl, J.gcl, n);
        out(gcu, J.gcu, n);
        out(grmin, J.grmin, m);
        out(grmax, J.grmax, m);
        out(gax, J.gax, z);
        if (info)
            BCK(hipMemcpy(info, J.info, sizeof(igaos::batch::AdjInfo) * B->count,
                          hipMemcpyDeviceToHost));
    } catch (const std::exception& e) {
        return bfail(e.what());
    }
    return 0;
}

void igaos_batch_destroy(igaos_batch* B) { delete B; }

}  // extern "C"
