// Residual and restriction of a V-cycle level in one launch (the V-cycle's
// r = f - A u followed by f_c = R r, par_cycle.c:443 and :704; HYPRE runs a
// matvec and then a transposed matvec).  The plan holds, for a level's A and
// R = P^T:
//  * R in CSR with ascending fine rows (each coarse row's sum order);
//  * per tile of A's block SpMV kernel (rpb block rows) the coarse rows its
//    fine rows touch, and per coarse row the number of such tiles (an empty
//    coarse row is given tile 0, whose sum over no entries writes its zero);
//  * one arrival counter per coarse row, zero between calls (the launch puts
//    each back to zero), so a call allocates nothing and can be captured.
// The kernel is spmv.hip's block kernel with the EpResidRestrict epilogue.
#include <new>

#include "internal.hpp"

namespace amgt {

namespace {

constexpr int kThreads = 256;
inline unsigned grid_of(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

// tiles of 4 rpb fine rows touched by coarse row c (its fine rows ascend)
__global__ __launch_bounds__(kThreads) void k_nspan(int32_t nc, int32_t rows_per_tile,
                                                    const int32_t *__restrict__ ri,
                                                    const int32_t *__restrict__ rj,
                                                    int32_t *__restrict__ nspan) {
    const int c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= nc) return;
    int k = 0, last = -1;
    for (int p = ri[c]; p < ri[c + 1]; ++p) {
        const int t = rj[p] / rows_per_tile;
        if (t != last) {
            ++k;
            last = t;
        }
    }
    nspan[c] = k > 0 ? k : 1;
}

__global__ __launch_bounds__(kThreads) void k_span_tiles(int32_t nc, int32_t rows_per_tile,
                                                         const int32_t *__restrict__ ri,
                                                         const int32_t *__restrict__ rj,
                                                         const int32_t *__restrict__ pptr,
                                                         int32_t *__restrict__ ptile) {
    const int c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= nc) return;
    int o = pptr[c], last = -1;
    for (int p = ri[c]; p < ri[c + 1]; ++p) {
        const int t = rj[p] / rows_per_tile;
        if (t != last) {
            ptile[o++] = t;
            last = t;
        }
    }
    if (o == pptr[c]) ptile[o] = 0;  // empty coarse row: tile 0 writes its zero
}

}  // namespace

struct RestrictPlan {
    int32_t nfine = 0, nc = 0, ntiles = 0, rpb = 0;
    int64_t nnzR = 0;
    int32_t *ri = nullptr, *rj = nullptr;
    double *rv = nullptr;
    int32_t *tile_ptr = nullptr, *tile_c = nullptr, *nspan = nullptr, *cnt = nullptr;
};

static void plan_release(RestrictPlan *p, hipStream_t s) {
    dev_free(p->ri, s);
    dev_free(p->rj, s);
    dev_free(p->rv, s);
    dev_free(p->tile_ptr, s);
    dev_free(p->tile_c, s);
    dev_free(p->nspan, s);
    dev_free(p->cnt, s);
}

static int plan_build(const Mbsr &A, const Mbsr &R, hipStream_t s, RestrictPlan *p) {
    const int lanes = A.spmv_lanes ? A.spmv_lanes : spmv_default_lanes(A);
    p->nfine = A.nrows;
    p->nc = R.nrows;
    p->rpb = blk_tile_rows(lanes);
    p->ntiles = int32_t((int64_t(A.brow) + p->rpb - 1) / p->rpb);
    AMGT_TRY(mbsr_to_csr(R, s, &p->nnzR, &p->ri, &p->rj, &p->rv));
    const int32_t nc = p->nc, rows_per_tile = 4 * p->rpb;
    int32_t *pptr = nullptr, *ptile = nullptr;
    p->nspan = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(nc) + 1), s));
    p->cnt = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(nc) + 1), s));
    p->tile_ptr = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(p->ntiles) + 1), s));
    if (!p->nspan || !p->cnt || !p->tile_ptr) {
        set_error("restriction plan: allocation failed");
        return AMGT_ERR_ALLOC;
    }
    AMGT_TRY(tmp_alloc_t(&pptr, size_t(nc) + 1, s));
    int st = AMGT_OK;
    if (nc > 0) {
        hipLaunchKernelGGL(k_nspan, dim3(grid_of(nc)), dim3(kThreads), 0, s, nc, rows_per_tile,
                           p->ri, p->rj, p->nspan);
        if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
    }
    if (st == AMGT_OK) st = exclusive_scan_i32(p->nspan, pptr, nc, s);
    int32_t npairs = 0;
    if (st == AMGT_OK) {
        const Readback rb[1] = {{&npairs, pptr + nc, sizeof(int32_t)}};
        st = readback(rb, 1, s);
    }
    if (st == AMGT_OK) st = tmp_alloc_t(&ptile, size_t(npairs), s);
    if (st == AMGT_OK) {
        p->tile_c = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * size_t(npairs > 0 ? npairs : 1), s));
        if (!p->tile_c) {
            set_error("restriction plan: allocation failed");
            st = AMGT_ERR_ALLOC;
        }
    }
    if (st == AMGT_OK && nc > 0) {
        hipLaunchKernelGGL(k_span_tiles, dim3(grid_of(nc)), dim3(kThreads), 0, s, nc,
                           rows_per_tile, p->ri, p->rj, pptr, ptile);
        if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
    }
    // per tile its coarse rows: the (coarse row -> tiles) lists transposed
    if (st == AMGT_OK)
        st = csr_transpose_device(nc, p->ntiles, pptr, ptile, nullptr, npairs, p->tile_ptr,
                                  p->tile_c, nullptr, s);
    if (st == AMGT_OK && hipMemsetAsync(p->cnt, 0, sizeof(int32_t) * (size_t(nc) + 1), s) !=
                             hipSuccess)
        st = AMGT_ERR_HIP;
    if (st == AMGT_OK && hipStreamSynchronize(s) != hipSuccess) st = AMGT_ERR_HIP;
    if (st == AMGT_ERR_HIP) set_error("restriction plan: a HIP call failed");
    tmp_free(pptr, s);
    tmp_free(ptile, s);
    return st;
}

}  // namespace amgt

struct amgt_restrict_plan {
    amgt::RestrictPlan p;
};

using namespace amgt;

extern "C" {

int amgt_restrict_plan_create(amgt_mbsr_t A, amgt_mbsr_t R, void *stream,
                              amgt_restrict_plan_t *plan) {
    AMGT_CHECK_ARG(A && R && plan, "NULL argument");
    *plan = nullptr;
    const Mbsr &a = A->m, &r = R->m;
    AMGT_CHECK_ARG(a.nrows == a.ncols, "A must be square");
    AMGT_CHECK_ARG(r.ncols == a.nrows, "R's columns must be A's rows");
    AMGT_CHECK_ARG(!r.pattern_only, "pattern-only R has no values");
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto *h = new (std::nothrow) amgt_restrict_plan;
    if (!h) {
        set_error("restriction plan: host allocation failed");
        return AMGT_ERR_ALLOC;
    }
    const int st = plan_build(a, r, s, &h->p);
    if (st != AMGT_OK) {
        (void)hipStreamSynchronize(s);
        plan_release(&h->p, s);
        delete h;
        return st;
    }
    *plan = h;
    return AMGT_OK;
}

int amgt_restrict_plan_destroy(amgt_restrict_plan_t plan) {
    if (!plan) return AMGT_OK;
    hipStream_t s = nullptr;
    AMGT_HIP(hipDeviceSynchronize());  // no launch of this plan may still run
    plan_release(&plan->p, s);
    delete plan;
    return AMGT_OK;
}

int amgt_mbsr_residual_restrict(amgt_mbsr_t A, amgt_restrict_plan_t plan, const double *d_f,
                                const double *d_u, double *d_r, double *d_fc, void *stream) {
    AMGT_CHECK_ARG(A && plan, "NULL handle or plan");
    Mbsr &a = A->m;
    const RestrictPlan &p = plan->p;
    AMGT_CHECK_ARG(a.nrows == p.nfine && a.ncols == p.nfine, "A does not match the plan");
    AMGT_CHECK_ARG(int32_t((int64_t(a.brow) + p.rpb - 1) / p.rpb) == p.ntiles,
                   "A's tiling does not match the plan");
    AMGT_CHECK_ARG(a.nrows == 0 || (d_f && d_u && d_r), "NULL vector");
    AMGT_CHECK_ARG(p.nc == 0 || d_fc, "NULL coarse vector");
    AMGT_CHECK_ARG(d_u != d_r && d_f != d_r, "r must not alias f or u");
    AMGT_CHECK_ARG((reinterpret_cast<uintptr_t>(d_u) & 15u) == 0, "u must be 16-byte aligned");
    AMGT_CHECK_ARG(!a.pattern_only, "pattern-only matrix has no values");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int prec = effective_precision(a);
    AMGT_TRY(mbsr_ensure_lowp(&a, &prec, s));
    const RestrictTables t{p.rpb, p.tile_ptr, p.tile_c, p.nspan, p.cnt, p.ri, p.rj, p.rv};
    return residual_restrict(a, prec, t, d_f, d_u, d_r, d_fc, s);
}

}  // extern "C"
