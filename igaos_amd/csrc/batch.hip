// Persistent batched PDHG kernel (see batch.hpp).
//
// One 256-thread workgroup owns one LP at a time.  Per attempt it runs the
// two SpMV phases with 2^k threads per row (k from the LP's schedule; rows
// strided over the workgroup, lanes of a row reduced with xor shuffles),
// fused with the primal / dual updates exactly as the single-LP kernels do,
// then reduces |dx|^2, |dy|^2 and dy'Adx with the workgroup tree and
// applies the line-search rule (pdhg_ctrl.hpp) on thread 0.  Every 64
// accepted iterations it evaluates both iterates' KKT metrics in FP64 in
// original units and runs the check logic of src/pdhg/pdhg.cu:801-1084
// (termination, incumbent, rollback, restart, primal weight, time limit)
// with the instance's own state.  In FP32 mode a restart folds the anchor's
// correction into the FP64 reference point and rebuilds the correction
// problem's data with FP64 SpMVs (iterative refinement).
//
// Every phase uses a fixed thread-to-data mapping, so an LP's result does
// not depend on which workgroup solves it or on the rest of the batch.

#include "batch_dev.hpp"

namespace igaos::batch {

using pdhg::LsState;
using pdhg::Metrics;

// Shared per-workgroup control block (thread 0 writes, all read).
struct Ctl {
    LsState st;
    pdhg::RestartState rs;
    Metrics mc, ma, last, inc;
    double eta0, omega0, inc_op, cmax, lam2;
    long long last_rollback, restarts;
    int lp, have_inc, have_pe, last_is_avg, stop, time_limit;
    int inc_src, rollback, restart, anchor_avg, accepted, done;
    unsigned long long t0;
};

// x = xref + xd in FP64 (xref may be null)
template <typename V>
__device__ __forceinline__ double full(const double* ref, const V* d, int j) {
    return ref ? ref[j] + static_cast<double>(d[j]) : static_cast<double>(d[j]);
}

// The FP64 primal iterate of column j.  In FP32 a correction exactly on
// its clamp bound (lo, hi) stands for the column at its bound and gives
// that bound exactly: x_ref + hi misses it by the FP32 rounding of
// scu - x_ref, which the KKT test's bound tolerance (1e-9 (1 + |bound|),
// pdhg.cu:206-215) does not absorb.
template <typename V>
__device__ __forceinline__ double xfull(const LP<V>& L, const V* d, int j) {
    if constexpr (sizeof(V) == 4) {
        const V v = d[j];
        if (v == L.hi()[j]) return L.scu()[j];
        if (v == L.lo()[j]) return L.scl()[j];
    }
    return full(L.xref(), d, j);
}

template <typename V>
__device__ void refresh_refs(const LP<V>& L, Lds& S) {
    // FP32: aref = A~ xref, atref = A~' yref (FP64), then the shifted data
    if (!L.xref()) return;
    spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{L.xref()}, S,
                       [&](int i, double s) { L.aref()[i] = s; });
    spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{L.yref()}, S,
                       [&](int j, double s) { L.atref()[j] = s; });
    __syncthreads();
    for (int j = threadIdx.x; j < L.n; j += NT) {
        L.lo()[j] = static_cast<V>(L.scl()[j] - L.xref()[j]);
        L.hi()[j] = static_cast<V>(L.scu()[j] - L.xref()[j]);
        L.gref()[j] = static_cast<V>(L.sc()[j] + L.atref()[j]);
    }
    for (int i = threadIdx.x; i < L.m; i += NT) {
        L.rlo()[i] = static_cast<V>(L.srmin()[i] - L.aref()[i]);
        L.rhi()[i] = static_cast<V>(L.srmax()[i] - L.aref()[i]);
        L.yoff()[i] = static_cast<V>(L.yref()[i]);
    }
    __syncthreads();
}

// KKT metrics in original units of the iterate (xref + xd, yref + yd)
// (src/pdhg/pdhg.cu:173-254 through metrics_of, 694-733).
template <typename V>
__device__ Metrics metrics(const LP<V>& L, const V* xd, const V* yd,
                           double cmax, Lds& S) {
    // the FP64 iterate gathered by the two SpMVs: an FP32 iterate is
    // combined with its reference point once, into the output arrays (free
    // until finish_lp writes them), so that each nonzero gathers one double
    // instead of a double and a float (same values: x_ref + double(dx))
    const double* xg;
    const double* yg;
    if constexpr (sizeof(V) == 4) {
        const double* yr = L.yref();
        double* xf = L.xout();
        double* yf = L.yout();
        for (int j = threadIdx.x; j < L.n; j += NT) xf[j] = xfull(L, xd, j);
        for (int i = threadIdx.x; i < L.m; i += NT) yf[i] = full(yr, yd, i);
        __syncthreads();
        xg = xf;
        yg = yf;
    } else {
        xg = xd;
        yg = yd;
    }
    const double dz = 1e-8 * (1.0 + cmax);
    double dv = 0.0, op = 0.0, odc = 0.0, unbc = 0.0;
    double pv = 0.0, pa = 0.0, prow = 0.0, ymax = 0.0;
    {  // column pass: d = c + A'y / dc, dinf, c'x and the column bound
        const double* dc = L.dc();
        const double* c = L.c();
        const double* clp = L.cl();
        const double* cup = L.cu();
        spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{yg}, S, [&](int j, double ata) {
            const double hx = xg[j] * dc[j];
            const double hd = c[j] + ata / dc[j];
            const double cl = clp[j], cu = cup[j];
            const double bnd = fabs(__builtin_isfinite(cu) ? cu : cl);
            const double tol = 1e-9 * (1.0 + bnd);
            double v;
            if (hx <= cl + tol && hd > 0) v = 0.0;
            else if (hx >= cu - tol && hd < 0) v = 0.0;
            else v = fabs(hd);
            dv = fmax(dv, v);
            op += c[j] * hx;
            if (hd >= 0) {
                if (__builtin_isfinite(cl)) odc += hd * cl;
                else if (hd > dz) unbc = 1.0;
            } else {
                if (__builtin_isfinite(cu)) odc += hd * cu;
                else if (-hd > dz) unbc = 1.0;
            }
        });
    }
    {  // row pass: Ax = dr (A~x~), pinf / prow numerators, max|y|
        const double* dr = L.dr();
        const double* rminp = L.rmin();
        const double* rmaxp = L.rmax();
        spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{xg}, S, [&](int i, double tm) {
            const double ax = tm * dr[i];
            const double lo = rminp[i], hi = rmaxp[i];
            const double e = fmax(lo - ax, ax - hi);
            pv = fmax(pv, e);
            pa = fmax(pa, fabs(fmin(fmax(ax, lo), hi)));
            double sc = fabs(ax);
            if (__builtin_isfinite(lo)) sc = fmax(sc, fabs(lo));
            if (__builtin_isfinite(hi)) sc = fmax(sc, fabs(hi));
            prow = fmax(prow, fmax(0.0, e) / (1.0 + sc));
            ymax = fmax(ymax, fabs(yg[i] * (1.0 / dr[i])));
        });
    }
    dv = bmax(dv, S.red);
    op = bsum(op, S.red);
    odc = bsum(odc, S.red);
    unbc = bmax(unbc, S.red);
    pv = bmax(pv, S.red);
    pa = bmax(pa, S.red);
    prow = bmax(prow, S.red);
    ymax = bmax(ymax, S.red);
    const double yz = 1e-8 * (1.0 + ymax);
    double odr = 0.0, unbr = 0.0;
    for (int i = threadIdx.x; i < L.m; i += NT) {
        const double hy = yg[i] * (1.0 / L.dr()[i]);
        if (hy > yz) {
            if (__builtin_isfinite(L.rmax()[i])) odr -= hy * L.rmax()[i];
            else unbr = 1.0;
        } else if (hy < -yz) {
            if (__builtin_isfinite(L.rmin()[i])) odr -= hy * L.rmin()[i];
            else unbr = 1.0;
        }
    }
    odr = bsum(odr, S.red);
    unbr = bmax(unbr, S.red);
    Metrics mt;
    mt.pinf = fmax(pv, 0.0) / (1.0 + pa);
    mt.prow = prow;
    mt.dinf = dv / (1.0 + cmax);
    mt.op = op;
    const bool bounded = unbc == 0.0 && unbr == 0.0;
    mt.od = bounded ? odc + odr : -__builtin_huge_val();
    mt.gap = bounded ? fmax(0.0, op - (odc + odr)) / (1.0 + fabs(op))
                     : __builtin_huge_val();
    return mt;
}

template <typename V>
__device__ void setup(const LP<V>& L, Ctl& C, Lds& S) {
    const int m = L.m, n = L.n, tid = threadIdx.x;
    // Ruiz, 10 sweeps (pdhg.cu:469-492); one thread per row / column, so
    // each factor is exactly the host loop's value.
    for (int i = tid; i < m; i += NT) L.dr()[i] = 1.0;
    for (int j = tid; j < n; j += NT) L.dc()[j] = 1.0;
    __syncthreads();
    for (int sweep = 0; sweep < 10; ++sweep) {
        for (int i = tid; i < m; i += NT) {
            double mx = 0.0;
            const double d = L.dr()[i];
            for (int p = L.ap()[i]; p < L.ap()[i + 1]; ++p)
                mx = fmax(mx, fabs(L.ax()[p]) * L.dc()[L.ai()[p]] / d);
            if (mx > 0 && __builtin_isfinite(mx)) L.dr()[i] = d * sqrt(mx);
        }
        __syncthreads();
        for (int j = tid; j < n; j += NT) {
            double mx = 0.0;
            const double d = L.dc()[j];
            for (int p = L.cp()[j]; p < L.cp()[j + 1]; ++p)
                mx = fmax(mx, fabs(L.acx()[p]) / L.dr()[L.ci()[p]] * d);
            if (mx > 0 && __builtin_isfinite(mx)) L.dc()[j] = d / sqrt(mx);
        }
        __syncthreads();
    }
    for (int i = tid; i < m; i += NT) {
        double d = L.dr()[i];
        if (!(d > 0 && __builtin_isfinite(d))) L.dr()[i] = d = 1.0;
        L.srmin()[i] = L.rmin()[i] / d;
        L.srmax()[i] = L.rmax()[i] / d;
    }
    for (int j = tid; j < n; j += NT) {
        double d = L.dc()[j];
        if (!(d > 0 && __builtin_isfinite(d))) L.dc()[j] = d = 1.0;
        L.sc()[j] = L.c()[j] * d / 1.0;
        L.scl()[j] = L.cl()[j] / d;
        L.scu()[j] = L.cu()[j] / d;
    }
    __syncthreads();
    for (int i = tid; i < m; i += NT) {
        const double d = L.dr()[i];
        for (int p = L.ap()[i]; p < L.ap()[i + 1]; ++p) {
            const double v = L.ax()[p] * L.dc()[L.ai()[p]] / d;
            L.sax()[p] = v;
            L.fax()[p] = static_cast<V>(v);
        }
    }
    for (int j = tid; j < n; j += NT) {
        const double d = L.dc()[j];
        for (int p = L.cp()[j]; p < L.cp()[j + 1]; ++p) {
            const double v = L.acx()[p] * d / L.dr()[L.ci()[p]];
            L.sacx()[p] = v;
            L.facx()[p] = static_cast<V>(v);
        }
    }
    __syncthreads();
    // power iteration (pdhg.cu:605-631); xb, pesx hold v, w and yb A v
    double* v = L.xb();
    double* w = L.pesx();
    double* t = L.yb();
    const double v0 = 1.0 / sqrt(static_cast<double>(n));
    for (int j = tid; j < n; j += NT) v[j] = v0;
    __syncthreads();
    for (int it = 0; it < 30; ++it) {
        spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{v}, S,
                           [&](int i, double s) { t[i] = s; });
        __syncthreads();
        double q = 0.0;
        spmv_T<double>(L, L.sacx(), dev::GatherPlain<double>{t}, S,
                           [&](int j, double s) { w[j] = s; q += s * s; });
        const double nw = sqrt(bsum(q, S.red));
        if (nw < 1e-300) break;
        const double inv = 1.0 / nw;
        for (int j = tid; j < n; j += NT) v[j] = w[j] * inv;
        __syncthreads();
    }
    double q = 0.0;
    spmv_A<double>(L, L.sax(), dev::GatherPlain<double>{v}, S,
                       [&](int i, double s) { q += s * s; });
    double lam2 = bsum(q, S.red);
    if (!(lam2 > 0) || !__builtin_isfinite(lam2)) lam2 = 1.0;
    // omega0 (pdhg.cu:647-663), max|c|
    double cn = 0.0, qn = 0.0, cm = 0.0;
    for (int j = tid; j < n; j += NT) {
        cn += L.sc()[j] * L.sc()[j];
        cm = fmax(cm, fabs(L.c()[j]));
    }
    for (int i = tid; i < m; i += NT) {
        const double qq = __builtin_isfinite(L.srmax()[i]) ? L.srmax()[i]
                          : (__builtin_isfinite(L.srmin()[i]) ? L.srmin()[i] : 0.0);
        qn += qq * qq;
    }
    cn = bsum(cn, S.red);
    qn = bsum(qn, S.red);
    cm = bmax(cm, S.red);
    if (tid == 0) {
        C.lam2 = lam2;
        C.eta0 = 1.2 / sqrt(lam2);
        double om = 1.0;
        if (cn > 1e-24 && qn > 1e-24) om = fmin(fmax(sqrt(cn / qn), 0.3), 3.0);
        C.omega0 = om;
        C.cmax = cm;
    }
    // fresh iterates: x = y = 0
    for (int j = tid; j < n; j += NT) {
        L.x(0)[j] = L.x(1)[j] = L.xbar()[j] = L.ata()[j] = L.xavg()[j] = V(0);
        if (L.xref()) L.xref()[j] = 0.0;
    }
    for (int i = tid; i < m; i += NT) {
        L.y(0)[i] = L.y(1)[i] = L.yavg()[i] = L.axcur()[i] = L.adx()[i] = V(0);
        if (L.yref()) L.yref()[i] = 0.0;
    }
    __syncthreads();
    refresh_refs(L, S);
}

// Control state of a fresh solve (thread 0).  no_freeze = 1 lets the
// persistent kernel run the check itself every 64 accepted iterations; the
// lockstep mode freezes the state there instead (pdhg_ctrl.hpp).
__device__ inline void init_ctl(Ctl& C, const Params& P, int no_freeze) {
    const double INF = __builtin_huge_val();
    LsState& s = C.st;
    s = LsState{};
    s.eta = C.eta0;
    s.omega = C.omega0;
    s.k = 1;
    s.maxit = P.maxit;
    s.fresh = 1;
    s.no_freeze = no_freeze;
    C.rs = pdhg::RestartState{-1.0, INF, 0};
    C.last = Metrics{0, 0, 0, 0, 0, 0};
    C.inc = Metrics{0, 0, 0, 0, 0, 0};
    C.last_is_avg = 0;
    C.have_inc = 0;
    C.have_pe = 0;
    C.inc_op = 0.0;
    C.last_rollback = 0;
    C.restarts = 0;
    C.stop = P.maxit < 1;
    C.time_limit = 0;
    C.done = 0;
    C.t0 = wall_clock64();
}

// Averaging and A x commit of the last accepted step (pdhg.cu:795-799).
template <typename V>
__device__ inline void apply_mix(const LP<V>& L, const Ctl& C) {
    const int cur = C.st.cur;
    const V w = static_cast<V>(C.st.mix_w);
    for (int j = threadIdx.x; j < L.n; j += NT) L.xavg()[j] += w * (L.x(cur)[j] - L.xavg()[j]);
    for (int i = threadIdx.x; i < L.m; i += NT) {
        L.yavg()[i] += w * (L.y(cur)[i] - L.yavg()[i]);
        L.axcur()[i] += L.adx()[i];
    }
}

// The check after `kdone` accepted iterations, a multiple of 64
// (src/pdhg/pdhg.cu:801-1084): metrics of both iterates, termination,
// incumbent, rollback, restart (FP32: refinement of the reference point),
// primal weight, time limit.  Returns true when the solve stops.
template <typename V>
__device__ bool check_lp(const LP<V>& L, const Params& P, Ctl& C, Lds& S,
                         long long kdone) {
    const int m = L.m, n = L.n, tid = threadIdx.x;
    constexpr bool F32 = sizeof(V) == 4;
    const double INF = __builtin_huge_val();
    const double TOL = P.tol;
    const int cur = C.st.cur;
    const Metrics mc = metrics(L, L.x(cur), L.y(cur), C.cmax, S);
    const Metrics ma = metrics(L, L.xavg(), L.yavg(), C.cmax, S);
    if (tid == 0) {
        const double ec = mc.finite_res() ? mc.err() : INF;
        const double ea = ma.finite_res() ? ma.err() : INF;
        C.last = (ec <= ea) ? mc : ma;
        C.last_is_avg = (ea < ec);
        C.inc_src = -1;
        C.rollback = 0;
        C.restart = 0;
        if (pdhg::passes(mc, TOL)) {
            C.last = mc;
            C.last_is_avg = 0;
            C.stop = 1;
        } else if (pdhg::passes(ma, TOL)) {
            C.last = ma;
            C.last_is_avg = 1;
            C.stop = 1;
        } else {
            for (int cc = 0; cc < 2; ++cc) {
                const Metrics& t = cc ? ma : mc;
                if (t.pinf <= 1e-4 && __builtin_isfinite(t.op) &&
                    (!C.have_inc || t.op < C.inc_op)) {
                    C.have_inc = 1;
                    C.inc_op = t.op;
                    C.inc = t;
                    C.inc_src = cc;
                }
            }
            if (mc.pinf >= 0.5 && C.have_inc && kdone - C.last_rollback > 2000) {
                C.rollback = 1;
                C.st.eta = fmin(C.st.eta, 0.5 * C.eta0);
                C.st.hn = 0;
                C.rs.epoch_start = kdone;
                C.rs.epoch_kkt = -1.0;
                C.rs.prev_cand = INF;
                C.last_rollback = kdone;
            }
            const double cand = fmin(ec, ea);
            if (pdhg::restart_decide(C.rs, cand, kdone)) {
                C.restart = 1;
                C.anchor_avg = ea < ec;
                C.restarts += 1;
                C.st.hn = 0;
                C.rs.epoch_start = kdone;
                C.rs.epoch_kkt = cand;
                C.rs.prev_cand = INF;
            }
            const double el = static_cast<double>(wall_clock64() - C.t0) / P.ticks_per_s;
            if (el > P.time_limit_s) C.time_limit = 1;
        }
    }
    __syncthreads();
    if (C.stop) return true;  // a pair passed the termination test
    if (C.inc_src >= 0) {
        const V* sx = C.inc_src ? L.xavg() : L.x(cur);
        const V* sy = C.inc_src ? L.yavg() : L.y(cur);
        for (int j = tid; j < n; j += NT) L.xb()[j] = xfull(L, sx, j);
        for (int i = tid; i < m; i += NT) L.yb()[i] = full(L.yref(), sy, i);
    }
    __syncthreads();
    if (C.rollback) {
        if constexpr (F32) {
            for (int j = tid; j < n; j += NT) {
                L.xref()[j] = L.xb()[j];
                L.x(0)[j] = L.x(1)[j] = L.xavg()[j] = V(0);
            }
            for (int i = tid; i < m; i += NT) {
                L.yref()[i] = L.yb()[i];
                L.y(0)[i] = L.y(1)[i] = L.yavg()[i] = L.axcur()[i] = V(0);
            }
            __syncthreads();
            refresh_refs(L, S);
        } else {
            for (int j = tid; j < n; j += NT) L.x(cur)[j] = L.xavg()[j] = static_cast<V>(L.xb()[j]);
            for (int i = tid; i < m; i += NT) L.y(cur)[i] = L.yavg()[i] = static_cast<V>(L.yb()[i]);
            __syncthreads();
            spmv_A<V>(L, L.fax(), dev::GatherPlain<V>{L.x(cur)}, S,
                          [&](int i, V s) { L.axcur()[i] = s; });
        }
        __syncthreads();
    }
    if (C.restart) {
        if constexpr (F32) {
            const V* ax = C.anchor_avg ? L.xavg() : L.x(cur);
            const V* ay = C.anchor_avg ? L.yavg() : L.y(cur);
            for (int j = tid; j < n; j += NT) L.xref()[j] = xfull(L, ax, j);
            for (int i = tid; i < m; i += NT) L.yref()[i] += static_cast<double>(ay[i]);
            __syncthreads();
            for (int j = tid; j < n; j += NT) L.x(0)[j] = L.x(1)[j] = L.xavg()[j] = V(0);
            for (int i = tid; i < m; i += NT) L.y(0)[i] = L.y(1)[i] = L.yavg()[i] = L.axcur()[i] = V(0);
            __syncthreads();
            refresh_refs(L, S);
        } else {
            if (C.anchor_avg) {
                for (int j = tid; j < n; j += NT) L.x(cur)[j] = L.xavg()[j];
                for (int i = tid; i < m; i += NT) L.y(cur)[i] = L.yavg()[i];
            } else {
                for (int j = tid; j < n; j += NT) L.xavg()[j] = L.x(cur)[j];
                for (int i = tid; i < m; i += NT) L.yavg()[i] = L.y(cur)[i];
            }
            __syncthreads();
            spmv_A<V>(L, L.fax(), dev::GatherPlain<V>{L.x(cur)}, S,
                          [&](int i, V s) { L.axcur()[i] = s; });
            __syncthreads();
        }
        // primal weight on the epoch-start points (scaled units)
        double dxn = 0.0, dyn = 0.0;
        for (int j = tid; j < n; j += NT) {
            const double v = F32 ? L.xref()[j] : static_cast<double>(L.x(cur)[j]);
            const double d = v - L.pesx()[j];
            dxn += d * d;
            L.pesx()[j] = v;
        }
        for (int i = tid; i < m; i += NT) {
            const double v = F32 ? L.yref()[i] : static_cast<double>(L.y(cur)[i]);
            const double d = v - L.pesy()[i];
            dyn += d * d;
            L.pesy()[i] = v;
        }
        dxn = bsum(dxn, S.red);
        dyn = bsum(dyn, S.red);
        if (tid == 0) {
            if (C.have_pe) C.st.omega = pdhg::primal_weight(C.st.omega, dxn, dyn);
            C.have_pe = 1;
        }
        __syncthreads();
    }
    if (C.time_limit) return true;  // checked last, as in pdhg.cu:1077-1083
    return kdone >= P.maxit;
}

// Exit (src/pdhg/pdhg.cu:1090-1141): status, incumbent fallback, result
// record and the solution in original units.
template <typename V>
__device__ void finish_lp(const LP<V>& L, const Params& P, const LPDesc& D,
                          Ctl& C, Lds& S, LPResult& R, long long kdone) {
    const int m = L.m, n = L.n, tid = threadIdx.x;
    const double TOL = P.tol;
    __syncthreads();
    if (tid == 0) {
        Metrics last = C.last;
        int st;
        if (C.time_limit) st = 6;
        else st = pdhg::certified(last, TOL) ? 1 : 5;
        int src = C.last_is_avg ? 1 : 0;
        if (!pdhg::certified(last, TOL) && C.have_inc && C.inc.pinf <= TOL &&
            __builtin_isfinite(C.inc.op) && (C.inc.op < last.op || last.pinf > TOL)) {
            last = C.inc;
            src = 2;
        }
        R.status = st;
        R.last_is_avg = src == 1;
        R.objective = last.op + D.obj_const;
        R.pinf = last.pinf;
        R.dinf = last.dinf;
        R.gap = last.gap;
        R.prow = last.prow;
        R.iterations = kdone;
        R.attempts = C.st.attempts;
        R.restarts = C.restarts;
        R.solve_ms = static_cast<double>(wall_clock64() - C.t0) / P.ticks_per_s * 1e3;
        R.eta = C.st.eta;
        R.omega = C.st.omega;
        R.lam2 = C.lam2;
        C.inc_src = src;
    }
    __syncthreads();
    const int src = C.inc_src;
    const int cur = C.st.cur;
    for (int j = tid; j < n; j += NT) {
        const double sx = src == 2 ? L.xb()[j]
                          : xfull(L, src == 1 ? L.xavg() : L.x(cur), j);
        L.xout()[j] = sx * L.dc()[j];
    }
    for (int i = tid; i < m; i += NT) {
        const double sy = src == 2 ? L.yb()[i]
                          : full(L.yref(), src == 1 ? L.yavg() : L.y(cur), i);
        L.yout()[i] = 1.0 * sy / L.dr()[i];
    }
    __syncthreads();
}

template <typename V>
__device__ void solve_one(const LP<V>& L, const Params& P, const LPDesc& D,
                          Ctl& C, Lds& S, LPResult& R) {
    const int m = L.m, n = L.n, tid = threadIdx.x;
    setup(L, C, S);
    if (tid == 0) init_ctl(C, P, 1);
    __syncthreads();
    long long kdone = 0;
    while (!C.stop) {
        // ---- one accepted iteration (pdhg.cu:748-799)
        bool fresh = true;
        for (;;) {
            const int cur = C.st.cur;
            const V tau = static_cast<V>(C.st.eta / C.st.omega);
            const V sig = static_cast<V>(C.st.eta * C.st.omega);
            const V* x = L.x(cur);
            V* xn = L.x(cur ^ 1);
            const V* y = L.y(cur);
            V* yn = L.y(cur ^ 1);
            double pacc = 0.0;
            auto primal = [&](int j, V a) {
                const V xo = x[j];
                const V g = L.gref()[j] + a;
                const V v = fmin(fmax(xo - tau * g, L.lo()[j]), L.hi()[j]);
                xn[j] = v;
                L.xbar()[j] = V(2) * v - xo;
                const double d = static_cast<double>(v - xo);
                pacc += d * d;
            };
            if (fresh) {
                spmv_T<V>(L, L.facx(), dev::GatherPlain<V>{y}, S,
                              [&](int j, V a) { L.ata()[j] = a; primal(j, a); });
            } else {
                for (int j = tid; j < n; j += NT) primal(j, L.ata()[j]);
            }
            __syncthreads();
            double a1 = 0.0, a2 = 0.0;
            spmv_A<V>(L, L.fax(), dev::GatherPlain<V>{L.xbar()}, S,
                          [&](int i, V tm) {
                const V yo = y[i];
                const V tt = yo + sig * tm;
                V u = L.yoff() ? tt + L.yoff()[i] : tt;
                u = u / sig;
                const V yv = tt - sig * fmin(fmax(u, L.rlo()[i]), L.rhi()[i]);
                yn[i] = yv;
                const V dy = yv - yo;
                const V ad = V(0.5) * (tm - L.axcur()[i]);
                L.adx()[i] = ad;
                a1 += static_cast<double>(dy) * static_cast<double>(dy);
                a2 += static_cast<double>(dy) * static_cast<double>(ad);
            });
            pacc = bsum(pacc, S.red);
            a1 = bsum(a1, S.red);
            a2 = bsum(a2, S.red);
            if (tid == 0) {
                pdhg::ls_decide(C.st, pacc, a1, a2);
                C.accepted = C.st.mix_pending;
            }
            __syncthreads();
            if (C.accepted) break;
            fresh = false;
        }
        apply_mix(L, C);
        kdone = C.st.k - 1;
        __syncthreads();
        if (kdone % pdhg::CHECK_EVERY != 0) {
            if (kdone >= P.maxit) break;
            continue;
        }
        if (check_lp(L, P, C, S, kdone)) break;
    }
    finish_lp(L, P, D, C, S, R, kdone);
}

// W = waves per SIMD the register allocation must allow: 2 leaves the
// compiler its natural allocation; 4 caps it at 128 VGPRs (some spills,
// twice the resident workgroups).  The autotuner picks per batch.
template <typename V, int W>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(W, 8)))
void k_batch(const Arrays<V>* A, Params P) {
    __shared__ Ctl C;
    __shared__ Lds S;
    unsigned int done = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            const int q = atomicAdd(P.queue, 1);
            C.lp = q < P.count ? P.order[q] : -1;
        }
        __syncthreads();
        const int id = C.lp;
        if (id < 0) break;
        const LPDesc D = P.lps[id];
        const LP<V> L = view(A, D);
        solve_one(L, P, D, C, S, P.res[id]);
        __syncthreads();
        if (threadIdx.x == 0 && P.progress)  // this workgroup's own slot
            __hip_atomic_store(P.progress + blockIdx.x, ++done, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

void launch_batch(const Arrays<float>* a, const Params& p, int grid, int waves,
                  hipStream_t s) {
    if (waves >= 4) k_batch<float, 4><<<grid, NT, 0, s>>>(a, p);
    else k_batch<float, 2><<<grid, NT, 0, s>>>(a, p);
}
void launch_batch(const Arrays<double>* a, const Params& p, int grid,
                  int waves, hipStream_t s) {
    if (waves >= 4) k_batch<double, 4><<<grid, NT, 0, s>>>(a, p);
    else k_batch<double, 2><<<grid, NT, 0, s>>>(a, p);
}

int batch_occupancy(bool f32, int waves) {
    int nb = 0;
    if (f32) {
        if (waves >= 4) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<float, 4>, NT, 0);
        else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<float, 2>, NT, 0);
    } else {
        if (waves >= 4) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<double, 4>, NT, 0);
        else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<double, 2>, NT, 0);
    }
    return nb > 0 ? nb : 1;
}

// ---------------------------------------------------------------------------
// Lockstep mode.  Every active instance advances by one line-search attempt
// per round; a round is three launches over the row blocks of all active
// instances (one table entry per block: instance, rows, nonzeros), so thousands of
// small SpMVs run as a few wide, low-register kernels that keep HBM busy,
// instead of one workgroup walking its instance's blocks one after another.
// The attempt kernels follow the single-LP engine's k_primal / k_dual
// (pending averaging folded into the next fresh attempt) and each
// instance's LsState freezes on its 64-iteration check boundary; once every
// active instance is frozen, one workgroup per instance runs the check with
// the persistent kernel's own code (check_lp / finish_lp).  Per-instance
// reductions run in a fixed order over the instance's own blocks, so every
// result is independent of the rest of the batch.

// XCD-aware work mapping.  Workgroups are dispatched to the 8 XCDs round
// robin (workgroup b runs on XCD b % 8), and each XCD has its own L2.  The
// tables list an instance's blocks consecutively, so the identity mapping
// would spread every instance over all XCDs and each XCD would fetch the
// instance's gathered vector on its own.  This bijection of [0, n) gives
// XCD x the contiguous range of table entries [x q + min(x, r), ...) of
// length q + (x < r), with q = n / 8, r = n % 8.
constexpr int XCDS = 8;
__device__ __forceinline__ int xcd_entry(int b, int n) {
    const int q = n / XCDS, r = n % XCDS;
    const int x = b % XCDS, slot = b / XCDS;
    return x * q + min(x, r) + slot;
}

// Resident waves per SIMD the attempt kernels are compiled for: 8 (the
// LDS limit of 8 workgroups per CU) in FP32, which caps them at 64 VGPRs;
// 6 in FP64, whose wider values would spill under that cap.
template <typename V>
constexpr int LK_WAVES = sizeof(V) == 4 ? 8 : 6;

// Index type of the attempt kernels: int (the CSR / CSC arrays as they
// are), or 16-bit on the gather-ordered copies (Arrays::ais16 ...).
template <typename I>
constexpr bool GATHER_ORDERED = sizeof(I) == 2;

// Values of the gather-ordered copies: per table entry (one row block),
// dst[p] = src[p0 + slot[p]] over the block's nonzeros.
// FP32: the packed records {index | slot << 16, value bits} instead.
template <typename V>
__global__ __launch_bounds__(NT) void k_lk_sortvals(const Arrays<V>* A, const LPDesc* lps,
                                                    const LkEnt* tbl, bool transposed) {
    const LkEnt e = tbl[blockIdx.x];
    const LP<V> L = view(A, lps[e.lp]);
    const V* src = transposed ? L.facx() : L.fax();
    const unsigned short* slot = transposed ? L.cslot() : L.aslot();
    if constexpr (sizeof(V) == 4) {
        const unsigned short* idx = transposed ? L.cis16() : L.ais16();
        uint2* dst = transposed ? L.cpk() : L.apk();
        for (int p = e.p0 + threadIdx.x; p < e.p1; p += NT) {
            const unsigned int sq = slot[p];
            dst[p] = make_uint2(static_cast<unsigned int>(idx[p]) | (sq << 16),
                                __float_as_uint(src[e.p0 + sq]));
        }
    } else {
        V* dst = transposed ? L.facxs() : L.faxs();
        for (int p = e.p0 + threadIdx.x; p < e.p1; p += NT) dst[p] = src[e.p0 + slot[p]];
    }
}

// Row sums of one row block (dev::rowblock_sums with a slot permutation)
// from packed FP32 records: one 8-byte load per nonzero carries its
// index, its slot in the block's row order and its value.  Same products
// at the same LDS positions, so the sums are bit-identical to the three
// separate streams.
__device__ __forceinline__ void rowblock_sums_packed(const int* __restrict__ ptr,
                                                     const uint2* __restrict__ rec,
                                                     const float* __restrict__ x, int r0,
                                                     int r1, int p0, int p1, float* s_rows,
                                                     float* s_prod, float* red) {
    using dev::BS;
    using dev::NNZ_MAX;
    const int tid = threadIdx.x;
    const int nrows = r1 - r0;
    const int nz = p1 - p0;
    if (nrows == 1 && nz > NNZ_MAX) {  // a single long row keeps its order
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        int p = p0 + tid;
        for (; p + 3 * BS < p1; p += 4 * BS) {
            const uint2 a = rec[p], b = rec[p + BS], c = rec[p + 2 * BS], d = rec[p + 3 * BS];
            s0 += __uint_as_float(a.y) * x[a.x & 0xffffu];
            s1 += __uint_as_float(b.y) * x[b.x & 0xffffu];
            s2 += __uint_as_float(c.y) * x[c.x & 0xffffu];
            s3 += __uint_as_float(d.y) * x[d.x & 0xffffu];
        }
        for (; p < p1; p += BS) {
            const uint2 a = rec[p];
            s0 += __uint_as_float(a.y) * x[a.x & 0xffffu];
        }
        const float s = dev::block_sum((s0 + s1) + (s2 + s3), red);
        if (tid == 0) s_rows[0] = s;
        __syncthreads();
        return;
    }
    int tpr = 1;
    while (tpr < BS && tpr * 2 * nrows <= BS) tpr *= 2;
    const int row = tid / tpr;
    const int lane = tid & (tpr - 1);
    int a = 0, b = 0;
    if (row < nrows) {
        a = ptr[r0 + row] - p0;
        b = ptr[r0 + row + 1] - p0;
    }
    if (nz > 0) {
        constexpr int U = NNZ_MAX / BS;
        const int last = nz - 1;
        uint2 r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) r[u] = rec[p0 + min(tid + u * BS, last)];
        float xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[r[u].x & 0xffffu];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (tid + u * BS < nz) s_prod[r[u].x >> 16] = __uint_as_float(r[u].y) * xv[u];
    }
    __syncthreads();
    float s = 0.0f;
    for (int q = a + lane; q < b; q += tpr) s += s_prod[q];
    const int wl = tpr < dev::WAVE ? tpr : dev::WAVE;
    for (int o = wl / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (tpr <= dev::WAVE) {
        if (row < nrows && lane == 0) s_rows[row] = s;
    } else {
        if ((tid & (dev::WAVE - 1)) == 0) red[tid / dev::WAVE] = s;
        __syncthreads();
        if (lane == 0 && row < nrows) {
            const int wpr = tpr / dev::WAVE;
            float t = red[row * wpr];
            for (int k = 1; k < wpr; ++k) t += red[row * wpr + k];
            s_rows[row] = t;
        }
    }
    __syncthreads();
}

template <typename V>
__global__ __launch_bounds__(NT) void k_lk_setup(const Arrays<V>* A, Params P,
                                                 Ctl* ctl, int* done) {
    __shared__ Ctl C;
    __shared__ Lds S;
    const int id = blockIdx.x;
    const LPDesc D = P.lps[id];
    const LP<V> L = view(A, D);
    setup(L, C, S);
    if (threadIdx.x == 0) init_ctl(C, P, 0);
    __syncthreads();
    if (C.stop) finish_lp(L, P, D, C, S, P.res[id], 0);
    __syncthreads();
    if (threadIdx.x == 0) {
        C.done = C.stop;
        C.lp = id;
        ctl[id] = C;
        done[id] = C.done;
    }
}

template <typename V, typename I>
__global__ __launch_bounds__(NT, LK_WAVES<V>) void k_lk_primal(const Arrays<V>* A,
                                                  const LPDesc* lps,
                                                  const LkEnt* tbl, Ctl* ctl,
                                                  double* part) {
    __shared__ Lds S;
    const int te = xcd_entry(blockIdx.x, gridDim.x);
    const LkEnt e = tbl[te];
    // both loads depend only on the entry: issue them together
    const LPDesc D = lps[e.lp];
    const LsState st = ctl[e.lp].st;
    if (st.frozen) {
        if (threadIdx.x == 0) part[te] = 0.0;
        return;
    }
    const LP<V> L = view(A, D);
    const int cur = st.cur;
    const V tau = static_cast<V>(st.eta / st.omega);
    const V w = static_cast<V>(st.mix_w);
    const bool fresh = st.fresh != 0;
    const bool mix = fresh && st.mix_pending != 0;
    const V* __restrict__ x = L.x(cur);
    V* __restrict__ xn = L.x(cur ^ 1);
    const V* __restrict__ y = L.y(cur);
    const int r0 = e.r0, p0 = e.p0, r1 = e.r1, p1 = e.p1;
    V* rows = reinterpret_cast<V*>(S.rows);
    // a block holds at most NT rows: thread t steps row r0 + t.  Its
    // operands are loaded before the product so that they arrive while
    // the matrix streams (they depend on the table entry only)
    static_assert(dev::ROWS_MAX <= NT, "one epilogue row per thread");
    const int t = threadIdx.x;
    const bool own = t < r1 - r0;
    const int j = r0 + (own ? t : 0);
    V xo = V(0), xa = V(0), g0 = V(0), lo = V(0), hi = V(0);
    if (own) {
        xo = x[j];
        if (mix) xa = L.xavg()[j];
        g0 = L.gref()[j];
        lo = L.lo()[j];
        hi = L.hi()[j];
    }
    // A'y is recomputed on a retry too (same y, same sums) rather than
    // stored for it: retries are rare, the store would be paid every round
    if constexpr (GATHER_ORDERED<I> && sizeof(V) == 4)
        rowblock_sums_packed(L.cp(), L.cpk(), y, r0, r1, p0, p1, rows,
                             reinterpret_cast<V*>(S.prod), reinterpret_cast<V*>(S.red));
    else if constexpr (GATHER_ORDERED<I>)
        dev::rowblock_sums<V>(L.cp(), L.cis16(), L.facxs(), dev::GatherPlain<V>{y}, r0, r1,
                              p0, p1, rows, reinterpret_cast<V*>(S.prod),
                              reinterpret_cast<V*>(S.red), L.cslot());
    else
        dev::rowblock_sums<V>(L.cp(), L.ci(), L.facx(), dev::GatherPlain<V>{y}, r0, r1, p0,
                              p1, rows, reinterpret_cast<V*>(S.prod),
                              reinterpret_cast<V*>(S.red));
    double acc = 0.0;
    if (own) {
        const V a = rows[t];
        if (mix) L.xavg()[j] = xa + w * (xo - xa);
        const V g = g0 + a;
        const V v = fmin(fmax(xo - tau * g, lo), hi);
        xn[j] = v;
        L.xbar()[j] = V(2) * v - xo;
        const V d = v - xo;
        acc = static_cast<double>(d) * static_cast<double>(d);
    }
    acc = bsum(acc, S.red);
    if (threadIdx.x == 0) part[te] = acc;
}

template <typename V, typename I>
__global__ __launch_bounds__(NT, LK_WAVES<V>) void k_lk_dual(const Arrays<V>* A,
                                                const LPDesc* lps,
                                                const LkEnt* tbl, Ctl* ctl,
                                                double* part) {
    __shared__ Lds S;
    const int te = xcd_entry(blockIdx.x, gridDim.x);
    const LkEnt e = tbl[te];
    const LPDesc D = lps[e.lp];
    const LsState st = ctl[e.lp].st;
    if (st.frozen) {
        if (threadIdx.x == 0) part[2 * te] = part[2 * te + 1] = 0.0;
        return;
    }
    const LP<V> L = view(A, D);
    const int cur = st.cur;
    const V sig = static_cast<V>(st.eta * st.omega);
    const V w = static_cast<V>(st.mix_w);
    const bool mix = st.fresh != 0 && st.mix_pending != 0;
    const V* __restrict__ y = L.y(cur);
    V* __restrict__ yn = L.y(cur ^ 1);
    const int r0 = e.r0, p0 = e.p0, r1 = e.r1, p1 = e.p1;
    V* rows = reinterpret_cast<V*>(S.rows);
    // the row's operands, loaded before the product (see k_lk_primal)
    const V* yoff = L.yoff();
    const int t = threadIdx.x;
    const bool own = t < r1 - r0;
    const int i = r0 + (own ? t : 0);
    V yo = V(0), axc = V(0), ya = V(0), adx = V(0), yof = V(0), rlo = V(0), rhi = V(0);
    if (own) {
        yo = y[i];
        axc = L.axcur()[i];
        if (mix) {
            ya = L.yavg()[i];
            adx = L.adx()[i];
        }
        if (yoff) yof = yoff[i];
        rlo = L.rlo()[i];
        rhi = L.rhi()[i];
    }
    if constexpr (GATHER_ORDERED<I> && sizeof(V) == 4)
        rowblock_sums_packed(L.ap(), L.apk(), L.xbar(), r0, r1, p0, p1, rows,
                             reinterpret_cast<V*>(S.prod), reinterpret_cast<V*>(S.red));
    else if constexpr (GATHER_ORDERED<I>)
        dev::rowblock_sums<V>(L.ap(), L.ais16(), L.faxs(), dev::GatherPlain<V>{L.xbar()},
                              r0, r1, p0, p1, rows, reinterpret_cast<V*>(S.prod),
                              reinterpret_cast<V*>(S.red), L.aslot());
    else
        dev::rowblock_sums<V>(L.ap(), L.ai(), L.fax(), dev::GatherPlain<V>{L.xbar()}, r0,
                              r1, p0, p1, rows, reinterpret_cast<V*>(S.prod),
                              reinterpret_cast<V*>(S.red));
    double a1 = 0.0, a2 = 0.0;
    if (own) {
        const V tm = rows[t];
        if (mix) {
            L.yavg()[i] = ya + w * (yo - ya);
            axc = axc + adx;
            L.axcur()[i] = axc;
        }
        const V tt = yo + sig * tm;
        V u = yoff ? tt + yof : tt;
        u = u / sig;
        const V yv = tt - sig * fmin(fmax(u, rlo), rhi);
        yn[i] = yv;
        const V dy = yv - yo;
        const V ad = V(0.5) * (tm - axc);
        L.adx()[i] = ad;
        a1 = static_cast<double>(dy) * static_cast<double>(dy);
        a2 = static_cast<double>(dy) * static_cast<double>(ad);
    }
    a1 = bsum(a1, S.red);
    a2 = bsum(a2, S.red);
    if (threadIdx.x == 0) {
        part[2 * te] = a1;
        part[2 * te + 1] = a2;
    }
}

// One wavefront per active instance: fixed-order sums of its block
// partials, then the line-search rule on its state.
__global__ __launch_bounds__(NT) void k_lk_decide(const LsRange* rng, int nact,
                                                  Ctl* ctl, const double* partT,
                                                  const double* partA,
                                                  int* nfrozen) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int a = blockIdx.x * (NT / WAVE) + threadIdx.x / WAVE;
    if (a >= nact) return;
    const LsRange r = rng[a];
    LsState* sp = &ctl[r.lp].st;
    if (sp->frozen) return;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    for (int b = lane; b < r.nT; b += WAVE) d0 += partT[r.bT0 + b];
    for (int b = lane; b < r.nA; b += WAVE) {
        d1 += partA[2 * (r.bA0 + b)];
        d2 += partA[2 * (r.bA0 + b) + 1];
    }
    d0 = dev::wave_sum(d0);
    d1 = dev::wave_sum(d1);
    d2 = dev::wave_sum(d2);
    if (lane == 0) {
        LsState s = *sp;
        pdhg::ls_decide(s, d0, d1, d2);
        *sp = s;
        if (s.frozen) atomicAdd(nfrozen, 1);
    }
}

// The check of every active instance (all frozen): one workgroup each.
template <typename V>
__global__ __launch_bounds__(NT) void k_lk_check(const Arrays<V>* A, Params P,
                                                 const LsRange* rng, const int* ord,
                                                 Ctl* ctl, int* done) {
    __shared__ Ctl C;
    __shared__ Lds S;
    const int a = ord[blockIdx.x];
    const int id = rng[a].lp;
    if (threadIdx.x == 0) C = ctl[id];
    __syncthreads();
    const LPDesc D = P.lps[id];
    const LP<V> L = view(A, D);
    if (C.st.mix_pending) {
        apply_mix(L, C);
        __syncthreads();
    }
    const long long kdone = C.st.k - 1;
    bool stop;
    if (kdone % pdhg::CHECK_EVERY != 0) stop = kdone >= P.maxit;  // MAXIT off a boundary
    else stop = check_lp(L, P, C, S, kdone);
    if (stop) finish_lp(L, P, D, C, S, P.res[id], kdone);
    __syncthreads();
    if (threadIdx.x == 0) {
        C.st.mix_pending = 0;
        if (stop) C.done = 1;
        else C.st.frozen = 0;
        ctl[id] = C;
        done[a] = C.done;
    }
}

size_t lockstep_ctl_bytes() { return sizeof(Ctl); }

template <typename V>
static void lk_setup_impl(const Arrays<V>* a, const Params& p, void* ctl,
                          int* done, hipStream_t s) {
    k_lk_setup<V><<<p.count, NT, 0, s>>>(a, p, static_cast<Ctl*>(ctl), done);
}
template <typename V>
static void lk_round_impl(const Arrays<V>* a, const LPDesc* lps, const LkEnt* tblT,
                          int nbT, const LkEnt* tblA, int nbA, const LsRange* rng,
                          int nact, void* ctl, double* partT, double* partA,
                          int* nfrozen, hipEvent_t* ev, bool idx16, hipStream_t s) {
    Ctl* c = static_cast<Ctl*>(ctl);
    if (ev) (void)hipEventRecord(ev[0], s);
    if (nbT > 0) {
        if (idx16) k_lk_primal<V, unsigned short><<<nbT, NT, 0, s>>>(a, lps, tblT, c, partT);
        else k_lk_primal<V, int><<<nbT, NT, 0, s>>>(a, lps, tblT, c, partT);
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    if (nbA > 0) {
        if (idx16) k_lk_dual<V, unsigned short><<<nbA, NT, 0, s>>>(a, lps, tblA, c, partA);
        else k_lk_dual<V, int><<<nbA, NT, 0, s>>>(a, lps, tblA, c, partA);
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    const int per = NT / WAVE;
    k_lk_decide<<<(nact + per - 1) / per, NT, 0, s>>>(rng, nact, c, partT, partA, nfrozen);
}
template <typename V>
static void lk_sort_impl(const Arrays<V>* a, const LPDesc* lps, const LkEnt* tblT, int nbT,
                         const LkEnt* tblA, int nbA, hipStream_t s) {
    if (nbT > 0) k_lk_sortvals<V><<<nbT, NT, 0, s>>>(a, lps, tblT, true);
    if (nbA > 0) k_lk_sortvals<V><<<nbA, NT, 0, s>>>(a, lps, tblA, false);
}
void lockstep_sort_values(const Arrays<float>* a, const LPDesc* lps, const LkEnt* tblT,
                          int nbT, const LkEnt* tblA, int nbA, hipStream_t s) {
    lk_sort_impl(a, lps, tblT, nbT, tblA, nbA, s);
}
void lockstep_sort_values(const Arrays<double>* a, const LPDesc* lps, const LkEnt* tblT,
                          int nbT, const LkEnt* tblA, int nbA, hipStream_t s) {
    lk_sort_impl(a, lps, tblT, nbT, tblA, nbA, s);
}

template <typename V>
static void lk_check_impl(const Arrays<V>* a, const Params& p, const LsRange* rng,
                          const int* ord, int nact, void* ctl, int* done, hipStream_t s) {
    k_lk_check<V><<<nact, NT, 0, s>>>(a, p, rng, ord, static_cast<Ctl*>(ctl), done);
}

void lockstep_setup(const Arrays<float>* a, const Params& p, void* ctl, int* done,
                    hipStream_t s) { lk_setup_impl(a, p, ctl, done, s); }
void lockstep_setup(const Arrays<double>* a, const Params& p, void* ctl, int* done,
                    hipStream_t s) { lk_setup_impl(a, p, ctl, done, s); }
void lockstep_round(const Arrays<float>* a, const LPDesc* lps, const LkEnt* tblT,
                    int nbT, const LkEnt* tblA, int nbA, const LsRange* rng, int nact,
                    void* ctl, double* partT, double* partA, int* nfrozen,
                    hipEvent_t* ev, bool idx16, hipStream_t s) {
    lk_round_impl(a, lps, tblT, nbT, tblA, nbA, rng, nact, ctl, partT, partA, nfrozen, ev,
                  idx16, s);
}
void lockstep_round(const Arrays<double>* a, const LPDesc* lps, const LkEnt* tblT,
                    int nbT, const LkEnt* tblA, int nbA, const LsRange* rng, int nact,
                    void* ctl, double* partT, double* partA, int* nfrozen,
                    hipEvent_t* ev, bool idx16, hipStream_t s) {
    lk_round_impl(a, lps, tblT, nbT, tblA, nbA, rng, nact, ctl, partT, partA, nfrozen, ev,
                  idx16, s);
}
void lockstep_check(const Arrays<float>* a, const Params& p, const LsRange* rng,
                    const int* ord, int nact, void* ctl, int* done, hipStream_t s) {
    lk_check_impl(a, p, rng, ord, nact, ctl, done, s);
}
void lockstep_check(const Arrays<double>* a, const Params& p, const LsRange* rng,
                    const int* ord, int nact, void* ctl, int* done, hipStream_t s) {
    lk_check_impl(a, p, rng, ord, nact, ctl, done, s);
}

}  // namespace igaos::batch
