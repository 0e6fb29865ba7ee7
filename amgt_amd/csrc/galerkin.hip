// The Galerkin product R A P (R = P^T) in the summation order and row layout
// of HYPRE's host setup, hypre_BoomerAMGBuildCoarseOperatorKT
// (par_rap.c:1525-1930) on one process:
//  * R = P^T with every row listing its fine rows in ascending order
//    (hypre_CSRMatrixTranspose);
//  * row ic of R A: R's row in order, each A row in its stored order; a new
//    column takes the next slot, a seen one accumulates (:1759-1788);
//  * row ic of (R A) P: the diagonal first (value 0, :1607-1613), then the
//    R A row's entries in their slot order times P's rows in their stored
//    order, first touch taking the next slot (:1855-1890).
// Products and sums are rounded one at a time (no fused multiply-add), so the
// coarse operator, and every level coarsened from it, equals HYPRE's host
// hierarchy bit for bit.  This is the order a bit-exact hierarchy needs; the
// mBSR SpGEMM (spgemm.hip) is the fast product, whose sums run in another
// order.
//
// One thread per row keeps each row's order (the rows are independent): a
// per-row open-addressing table maps a column to its slot, and the row is
// written in slot order into a temporary sized by the row's product count,
// then compacted.
#include "internal.hpp"

#include <hipcub/hipcub.hpp>

#include <vector>

namespace amgt {

namespace {

constexpr int kThreads = 256;
inline unsigned grid_of(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

struct Tmp {
    hipStream_t s;
    std::vector<void *> ptrs;
    explicit Tmp(hipStream_t st) : s(st) {}
    ~Tmp() {
        for (void *p : ptrs) tmp_free(p, s);
    }
    template <typename T>
    int get(T **p, size_t n) {
        const int st = tmp_alloc_t(p, n, s);
        if (st == AMGT_OK) ptrs.push_back(*p);
        return st;
    }
};

__global__ __launch_bounds__(kThreads) void k_row_of(int32_t n, const int32_t *__restrict__ I,
                                                     int32_t *__restrict__ row) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    for (int p = I[i]; p < I[i + 1]; ++p) row[p] = i;
}

__global__ __launch_bounds__(kThreads) void k_iota(int64_t n, int32_t *__restrict__ v) {
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) v[i] = int32_t(i);
}

// row pointer of R from P's column indices sorted ascending: first position
// of each column (lower bound), R_i[nc] = nnz
__global__ __launch_bounds__(kThreads) void k_rowptr_sorted(int32_t nc, int32_t nnz,
                                                            const int32_t *__restrict__ keys,
                                                            int32_t *__restrict__ ri) {
    const int c = blockIdx.x * kThreads + threadIdx.x;
    if (c > nc) return;
    int a = 0, b = nnz;
    while (a < b) {
        const int mid = (a + b) >> 1;
        if (keys[mid] < c) a = mid + 1; else b = mid;
    }
    ri[c] = a;
}

// the sort is stable, so each R row keeps P's row order: ascending fine rows
__global__ __launch_bounds__(kThreads) void k_gather_t(int32_t nnz, const int32_t *__restrict__ perm,
                                                       const int32_t *__restrict__ row_of,
                                                       const double *__restrict__ pa,
                                                       int32_t *__restrict__ rj,
                                                       double *__restrict__ ra) {
    const int q = blockIdx.x * kThreads + threadIdx.x;
    if (q >= nnz) return;
    const int p = perm[q];
    rj[q] = row_of[p];
    ra[q] = pa[p];
}

// products of row i of X*Y (+1 for a diagonal slot)
__global__ __launch_bounds__(kThreads) void k_prod_ub(int32_t m, const int32_t *__restrict__ XI,
                                                      const int32_t *__restrict__ XJ,
                                                      const int32_t *__restrict__ YI, int diag,
                                                      int64_t *__restrict__ ub) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= m) return;
    int64_t s = diag;
    for (int p = XI[i]; p < XI[i + 1]; ++p) s += YI[XJ[p] + 1] - YI[XJ[p]];
    ub[i] = s;
}

// table slots of row i: the power of two at or above 2 ub (0 for an empty row)
__global__ __launch_bounds__(kThreads) void k_tab_size(int32_t m, const int64_t *__restrict__ ub,
                                                       int64_t *__restrict__ ts) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= m) return;
    int64_t t = 0;
    if (ub[i] > 0) {
        t = 8;
        while (t < 2 * ub[i]) t <<= 1;
    }
    ts[i] = t;
}

__device__ __forceinline__ uint32_t tab_slot(int k, uint32_t mask) {
    return (uint32_t(k) * 2654435761u) & mask;
}

// Row i of X*Y in first-touch order: columns into tj, sums into tv (both at
// off[i]), the number of columns into cnt[i].  With diag, column i takes slot
// 0 with the value 0 before any product (the Galerkin row layout).
#pragma clang fp contract(off)
__global__ __launch_bounds__(kThreads) void k_first_touch(
    int32_t m, const int32_t *__restrict__ XI, const int32_t *__restrict__ XJ,
    const double *__restrict__ XV, const int32_t *__restrict__ YI, const int32_t *__restrict__ YJ,
    const double *__restrict__ YV, int diag, const int64_t *__restrict__ off,
    const int64_t *__restrict__ toff, int2 *__restrict__ kv, int32_t *__restrict__ tj,
    double *__restrict__ tv, int32_t *__restrict__ cnt) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= m) return;
    const int64_t tsz = toff[i + 1] - toff[i];
    if (tsz == 0) {
        cnt[i] = 0;
        return;
    }
    int2 *tab = kv + toff[i];
    const uint32_t mask = uint32_t(tsz) - 1u;
    for (int64_t h = 0; h < tsz; ++h) tab[h] = make_int2(-1, 0);
    int32_t *oj = tj + off[i];
    double *ov = tv + off[i];
    int c = 0;
    if (diag) {
        tab[tab_slot(i, mask)] = make_int2(i, 0);
        oj[0] = i;
        ov[0] = 0.0;
        c = 1;
    }
    for (int p = XI[i]; p < XI[i + 1]; ++p) {
        const int k = XJ[p];
        const double x = XV[p];
        for (int q = YI[k]; q < YI[k + 1]; ++q) {
            const int col = YJ[q];
            const double v = x * YV[q];
            for (uint32_t h = tab_slot(col, mask);; h = (h + 1) & mask) {
                const int2 e = tab[h];
                if (e.x == col) {
                    ov[e.y] += v;
                    break;
                }
                if (e.x == -1) {
                    tab[h] = make_int2(col, c);
                    oj[c] = col;
                    ov[c] = v;
                    ++c;
                    break;
                }
            }
        }
    }
    cnt[i] = c;
}
#pragma clang fp contract(fast)

__global__ __launch_bounds__(kThreads) void k_compact(int32_t m, const int64_t *__restrict__ off,
                                                      const int32_t *__restrict__ cnt,
                                                      const int32_t *__restrict__ oi,
                                                      const int32_t *__restrict__ tj,
                                                      const double *__restrict__ tv,
                                                      int32_t *__restrict__ cj,
                                                      double *__restrict__ cv) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= m) return;
    const int64_t s = off[i];
    const int o = oi[i];
    for (int q = 0; q < cnt[i]; ++q) {
        cj[o + q] = tj[s + q];
        cv[o + q] = tv[s + q];
    }
}

__global__ __launch_bounds__(kThreads) void k_gather_pattern(int32_t nnz,
                                                             const int32_t *__restrict__ perm,
                                                             const int32_t *__restrict__ row_of,
                                                             int32_t *__restrict__ tj) {
    const int q = blockIdx.x * kThreads + threadIdx.x;
    if (q >= nnz) return;
    tj[q] = row_of[perm[q]];
}

int d2h_sync(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes <= kReadbackBytes) {  // small: through the pinned staging area
        const Readback rb[1] = {{dst, src, bytes}};
        return readback(rb, 1, s);
    }
    AMGT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    AMGT_HIP(hipStreamSynchronize(s));
    return AMGT_OK;
}

// C = X * Y (m rows) in first-touch order; C's arrays come from `alloc` when
// `owned` (the library's allocator for results), else from tmp.
int first_touch_product(int32_t m, const int32_t *XI, const int32_t *XJ, const double *XV,
                        const int32_t *YI, const int32_t *YJ, const double *YV, int diag,
                        bool owned, Tmp &tmp, hipStream_t s, int32_t **CI, int32_t **CJ,
                        double **CV, int64_t *nnz) {
    int64_t *ub = nullptr, *off = nullptr, *ts = nullptr, *toff = nullptr;
    int32_t *cnt = nullptr;
    AMGT_TRY(tmp.get(&ub, size_t(m) + 1));
    AMGT_TRY(tmp.get(&off, size_t(m) + 1));
    AMGT_TRY(tmp.get(&ts, size_t(m) + 1));
    AMGT_TRY(tmp.get(&toff, size_t(m) + 1));
    AMGT_TRY(tmp.get(&cnt, size_t(m) + 1));
    int32_t *ci = owned ? static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(m) + 1), s))
                        : nullptr;
    if (!owned) AMGT_TRY(tmp.get(&ci, size_t(m) + 1));
    if (!ci) {
        set_error("Galerkin product: row pointer allocation failed");
        return AMGT_ERR_ALLOC;
    }
    *CI = ci;
    if (m > 0) {
        hipLaunchKernelGGL(k_prod_ub, dim3(grid_of(m)), dim3(kThreads), 0, s, m, XI, XJ, YI, diag,
                           ub);
        hipLaunchKernelGGL(k_tab_size, dim3(grid_of(m)), dim3(kThreads), 0, s, m, ub, ts);
        AMGT_HIP(hipGetLastError());
    }
    AMGT_TRY(exclusive_scan_i64(ub, off, m, s));
    AMGT_TRY(exclusive_scan_i64(ts, toff, m, s));
    int64_t tot[2] = {0, 0};
    {
        const Readback rb[2] = {{&tot[0], off + m, sizeof(int64_t)},
                                {&tot[1], toff + m, sizeof(int64_t)}};
        AMGT_TRY(readback(rb, 2, s));
    }
    int2 *kv = nullptr;
    int32_t *tj = nullptr;
    double *tv = nullptr;
    AMGT_TRY(tmp.get(&kv, size_t(tot[1])));
    AMGT_TRY(tmp.get(&tj, size_t(tot[0])));
    AMGT_TRY(tmp.get(&tv, size_t(tot[0])));
    if (m > 0) {
        hipLaunchKernelGGL(k_first_touch, dim3(grid_of(m)), dim3(kThreads), 0, s, m, XI, XJ, XV, YI,
                           YJ, YV, diag, off, toff, kv, tj, tv, cnt);
        AMGT_HIP(hipGetLastError());
    }
    AMGT_TRY(exclusive_scan_i32(cnt, ci, m, s));
    int32_t total = 0;
    AMGT_TRY(d2h_sync(&total, ci + m, sizeof(int32_t), s));
    if (total < 0) {
        set_error("Galerkin product: 2^31 or more entries");
        return AMGT_ERR_UNSUPPORTED;
    }
    int32_t *cj = nullptr;
    double *cv = nullptr;
    if (owned) {
        cj = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * size_t(total > 0 ? total : 1), s));
        cv = static_cast<double *>(dev_alloc(sizeof(double) * size_t(total > 0 ? total : 1), s));
        if (!cj || !cv) {
            dev_free(cj, s);
            dev_free(cv, s);
            set_error("Galerkin product: result allocation failed");
            return AMGT_ERR_ALLOC;
        }
    } else {
        AMGT_TRY(tmp.get(&cj, size_t(total)));
        AMGT_TRY(tmp.get(&cv, size_t(total)));
    }
    *CJ = cj;
    *CV = cv;
    *nnz = total;
    if (m > 0) {
        hipLaunchKernelGGL(k_compact, dim3(grid_of(m)), dim3(kThreads), 0, s, m, off, cnt, ci, tj,
                           tv, cj, cv);
        AMGT_HIP(hipGetLastError());
    }
    return AMGT_OK;
}

}  // namespace

int csr_transpose_device(int32_t nrows, int32_t ncols, const int32_t *I, const int32_t *J,
                         const double *V, int32_t nnz, int32_t *TI, int32_t *TJ, double *TV,
                         hipStream_t s) {
    Tmp tmp(s);
    int32_t *row_of = nullptr, *idx = nullptr, *perm = nullptr, *keys = nullptr;
    AMGT_TRY(tmp.get(&row_of, size_t(nnz)));
    AMGT_TRY(tmp.get(&idx, size_t(nnz)));
    AMGT_TRY(tmp.get(&perm, size_t(nnz)));
    AMGT_TRY(tmp.get(&keys, size_t(nnz)));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_row_of, dim3(grid_of(nrows)), dim3(kThreads), 0, s, nrows, I, row_of);
        hipLaunchKernelGGL(k_iota, dim3(grid_of(nnz)), dim3(kThreads), 0, s, int64_t(nnz), idx);
        AMGT_HIP(hipGetLastError());
        int bits = 1;
        while (bits < 31 && (int64_t(1) << bits) < int64_t(ncols)) ++bits;
        size_t bytes = 0;
        AMGT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, J, keys, idx, perm, nnz, 0,
                                                    bits, s));
        char *sort_tmp = nullptr;
        AMGT_TRY(tmp.get(&sort_tmp, bytes));
        AMGT_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp, bytes, J, keys, idx, perm, nnz, 0,
                                                    bits, s));
        if (V) {
            hipLaunchKernelGGL(k_gather_t, dim3(grid_of(nnz)), dim3(kThreads), 0, s, nnz, perm,
                               row_of, V, TJ, TV);
        } else {
            hipLaunchKernelGGL(k_gather_pattern, dim3(grid_of(nnz)), dim3(kThreads), 0, s, nnz,
                               perm, row_of, TJ);
        }
        AMGT_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_rowptr_sorted, dim3(grid_of(int64_t(ncols) + 1)), dim3(kThreads), 0, s,
                       ncols, nnz, keys, TI);
    AMGT_HIP(hipGetLastError());
    // the temporaries go back to the pool stream-ordered: no wait needed
    return AMGT_OK;
}

}  // namespace amgt

using namespace amgt;

extern "C" {

int amgt_amg_rap_host_order(int32_t n, int32_t nc, const int32_t *A_i, const int32_t *A_j,
                            const double *A_data, const int32_t *P_i, const int32_t *P_j,
                            const double *P_data, int64_t *nnz, int32_t **RAP_i,
                            int32_t **RAP_j, double **RAP_data, void *stream) {
    AMGT_CHECK_ARG(nnz && RAP_i && RAP_j && RAP_data, "NULL output argument");
    AMGT_CHECK_ARG(n >= 0 && nc >= 0, "negative size");
    AMGT_CHECK_ARG(n == 0 || (A_i && P_i), "NULL row pointer");
    *nnz = 0;
    *RAP_i = nullptr;
    *RAP_j = nullptr;
    *RAP_data = nullptr;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Tmp tmp(s);
    int32_t nnzP = 0, nnzA = 0;
    if (n > 0) {
        const Readback rb[2] = {{&nnzP, P_i + n, sizeof(int32_t)},
                                {&nnzA, A_i + n, sizeof(int32_t)}};
        AMGT_TRY(readback(rb, 2, s));
    }
    AMGT_CHECK_ARG(nnzP == 0 || (P_j && P_data), "NULL P column or value array");
    AMGT_CHECK_ARG(nnzA == 0 || (A_j && A_data), "NULL A column or value array");
    // R = P^T, rows ascending in the fine index
    int32_t *ri = nullptr, *rj = nullptr;
    double *ra = nullptr;
    AMGT_TRY(tmp.get(&ri, size_t(nc) + 1));
    AMGT_TRY(tmp.get(&rj, size_t(nnzP)));
    AMGT_TRY(tmp.get(&ra, size_t(nnzP)));
    AMGT_TRY(csr_transpose_device(n, nc, P_i, P_j, P_data, nnzP, ri, rj, ra, s));
    // R A, then (R A) P with the diagonal first
    int32_t *rai = nullptr, *raj = nullptr;
    double *rav = nullptr;
    int64_t nnzRA = 0;
    AMGT_TRY(first_touch_product(nc, ri, rj, ra, A_i, A_j, A_data, 0, false, tmp, s, &rai, &raj,
                                 &rav, &nnzRA));
    int32_t *ci = nullptr, *cj = nullptr;
    double *cv = nullptr;
    int64_t nnzC = 0;
    const int st = first_touch_product(nc, rai, raj, rav, P_i, P_j, P_data, 1, true, tmp, s, &ci,
                                       &cj, &cv, &nnzC);
    if (st != AMGT_OK) {
        dev_free(ci, s);
        dev_free(cj, s);
        dev_free(cv, s);
        return st;
    }
    AMGT_HIP(hipStreamSynchronize(s));  // the temporaries go back to the pool
    *nnz = nnzC;
    *RAP_i = ci;
    *RAP_j = cj;
    *RAP_data = cv;
    return AMGT_OK;
}

}  // extern "C"
