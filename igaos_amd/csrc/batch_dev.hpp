#pragma once
// Device helpers shared by the batched solver and the batched adjoint: the
// per-instance view of the concatenated batch arrays and the instance SpMVs
// (row-block streaming through LDS, or 2^k lanes per row).

#include "batch.hpp"
#include "dev_common.hpp"

namespace igaos::batch {

using dev::WAVE;

__device__ __forceinline__ double bsum(double v, double* red) {
    return dev::block_sum<double, NT>(v, red);
}
__device__ __forceinline__ double bmax(double v, double* red) {
    return dev::block_max<double, NT>(v, red);
}

// LDS scratch of a workgroup: row-block products and row sums (as FP64 or
// reinterpreted as FP32) and the reduction slots.
struct Lds {
    double prod[dev::NNZ_MAX];
    double rows[dev::ROWS_MAX];
    double red[NT / WAVE];
};

// Rows [0, nrows) of a CSR block: tpr lanes per row, f(row, sum) on the
// row's first lane.  The outer loop trip count is uniform, so every lane
// reaches the shuffles.
template <typename Acc, typename Val, typename G, typename F>
__device__ __forceinline__ void rows_apply(int nrows, const int* __restrict__ ptr,
                                           const int* __restrict__ idx,
                                           const Val* __restrict__ val,
                                           const G& gx, int tpr, F&& f) {
    const int groups = NT / tpr;
    const int g = threadIdx.x / tpr;
    const int lane = threadIdx.x & (tpr - 1);
    for (int base = 0; base < nrows; base += groups) {
        const int r = base + g;
        Acc s = Acc(0);
        if (r < nrows) {
            const int p1 = ptr[r + 1];
            for (int p = ptr[r] + lane; p < p1; p += tpr)
                s += static_cast<Acc>(val[p]) * static_cast<Acc>(gx(idx[p]));
        }
        for (int o = tpr / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0 && r < nrows) f(r, s);
    }
}

// Per-instance view: array pointers are formed on demand from the batch's
// pointer table (device memory, read through scalar loads) and the
// instance's offsets, which keeps register pressure low.
template <typename V>
struct LP {
    const Arrays<V>* A;
    int m, n, tpr, nblkA, nblkT;
    long ro, co, zo, rpo, cpo, rbao, rbto;
    static constexpr bool F32 = sizeof(V) == 4;
    __device__ const int* ap() const { return A->ap + rpo; }
    __device__ const int* ai() const { return A->ai + zo; }
    __device__ const int* cp() const { return A->cp + cpo; }
    __device__ const int* ci() const { return A->ci + zo; }
    __device__ const unsigned short* ai16() const { return A->ai16 + zo; }
    __device__ const unsigned short* ci16() const { return A->ci16 + zo; }
    __device__ const unsigned short* ais16() const { return A->ais16 + zo; }
    __device__ const unsigned short* cis16() const { return A->cis16 + zo; }
    __device__ const unsigned short* aslot() const { return A->aslot + zo; }
    __device__ const unsigned short* cslot() const { return A->cslot + zo; }
    __device__ V* faxs() const { return A->faxs + zo; }
    __device__ V* facxs() const { return A->facxs + zo; }
    __device__ uint2* apk() const { return A->apk + zo; }
    __device__ uint2* cpk() const { return A->cpk + zo; }
    __device__ const int* rbA() const { return A->rbA + rbao; }
    __device__ const int* rbT() const { return A->rbT + rbto; }
    __device__ const double* ax() const { return A->ax + zo; }
    __device__ const double* acx() const { return A->acx + zo; }
    __device__ const double* c() const { return A->c + co; }
    __device__ const double* cl() const { return A->cl + co; }
    __device__ const double* cu() const { return A->cu + co; }
    __device__ const double* rmin() const { return A->rmin + ro; }
    __device__ const double* rmax() const { return A->rmax + ro; }
    __device__ double* dr() const { return A->dr + ro; }
    __device__ double* dc() const { return A->dc + co; }
    __device__ double* sax() const { return A->sax + zo; }
    __device__ double* sacx() const { return A->sacx + zo; }
    __device__ double* sc() const { return A->sc + co; }
    __device__ double* scl() const { return A->scl + co; }
    __device__ double* scu() const { return A->scu + co; }
    __device__ double* srmin() const { return A->srmin + ro; }
    __device__ double* srmax() const { return A->srmax + ro; }
    __device__ V* fax() const { return A->fax + zo; }
    __device__ V* facx() const { return A->facx + zo; }
    __device__ V* x(int k) const { return (k ? A->x1 : A->x0) + co; }
    __device__ V* y(int k) const { return (k ? A->y1 : A->y0) + ro; }
    __device__ V* xbar() const { return A->xbar + co; }
    __device__ V* ata() const { return A->ata + co; }
    __device__ V* xavg() const { return A->xavg + co; }
    __device__ V* yavg() const { return A->yavg + ro; }
    __device__ V* axcur() const { return A->axcur + ro; }
    __device__ V* adx() const { return A->adx + ro; }
    __device__ double* xref() const { return F32 ? A->xref + co : nullptr; }
    __device__ double* yref() const { return F32 ? A->yref + ro : nullptr; }
    __device__ double* aref() const { return F32 ? A->aref + ro : nullptr; }
    __device__ double* atref() const { return F32 ? A->atref + co : nullptr; }
    __device__ V* lo() const { return A->lo + co; }
    __device__ V* hi() const { return A->hi + co; }
    __device__ V* gref() const { return A->gref + co; }
    __device__ V* rlo() const { return A->rlo + ro; }
    __device__ V* rhi() const { return A->rhi + ro; }
    __device__ V* yoff() const { return F32 ? A->yoff + ro : nullptr; }
    __device__ double* xb() const { return A->xb + co; }
    __device__ double* yb() const { return A->yb + ro; }
    __device__ double* pesx() const { return A->pesx + co; }
    __device__ double* pesy() const { return A->pesy + ro; }
    __device__ double* xout() const { return A->xout + co; }
    __device__ double* yout() const { return A->yout + ro; }
};

template <typename V>
__device__ __forceinline__ LP<V> view(const Arrays<V>* A, const LPDesc& d) {
    LP<V> L;
    L.A = A;
    L.m = d.m;
    L.n = d.n;
    L.tpr = d.tpr;
    L.nblkA = d.nblkA;
    L.nblkT = d.nblkT;
    L.rbao = d.rbao;
    L.rbto = d.rbto;
    L.ro = d.roff;
    L.co = d.coff;
    L.zo = d.zoff;
    L.rpo = d.rpo;
    L.cpo = d.cpo;
    return L;
}

// Row-block streaming over one instance matrix: each block of <= 2048
// nonzeros is staged through LDS with 8 independent loads per thread
// (dev::rowblock_sums), then f(row, sum) runs on consecutive threads.
template <typename Acc, typename Val, typename G, typename F>
__device__ __forceinline__ void blocks_apply(const int* rb, int nblk,
                                             const int* ptr, const int* idx,
                                             const Val* val, const G& gx,
                                             Lds& S, F&& f) {
    Acc* prod = reinterpret_cast<Acc*>(S.prod);
    Acc* rows = reinterpret_cast<Acc*>(S.rows);
    Acc* red = reinterpret_cast<Acc*>(S.red);
    for (int b = 0; b < nblk; ++b) {
        const int r0 = rb[2 * b], p0 = rb[2 * b + 1];
        const int r1 = rb[2 * b + 2], p1 = rb[2 * b + 3];
        dev::rowblock_sums<Acc>(ptr, idx, val, gx, r0, r1, p0, p1, rows, prod, red);
        for (int t = threadIdx.x; t < r1 - r0; t += NT) f(r0 + t, rows[t]);
        __syncthreads();
    }
}

// y = A x (rows of A) and A'y (rows of A') with the instance's schedule.
template <typename Acc, typename V, typename Val, typename G, typename F>
__device__ __forceinline__ void spmv_A(const LP<V>& L, const Val* val,
                                       const G& gx, Lds& S, F&& f) {
    if (L.tpr == 0) blocks_apply<Acc>(L.rbA(), L.nblkA, L.ap(), L.ai(), val, gx, S, f);
    else rows_apply<Acc>(L.m, L.ap(), L.ai(), val, gx, L.tpr, f);
}
template <typename Acc, typename V, typename Val, typename G, typename F>
__device__ __forceinline__ void spmv_T(const LP<V>& L, const Val* val,
                                       const G& gx, Lds& S, F&& f) {
    if (L.tpr == 0) blocks_apply<Acc>(L.rbT(), L.nblkT, L.cp(), L.ci(), val, gx, S, f);
    else rows_apply<Acc>(L.n, L.cp(), L.ci(), val, gx, L.tpr, f);
}

}  // namespace igaos::batch
