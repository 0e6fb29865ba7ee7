// mBSR transpose: builds A^T as an mBSR so that trans = 1 SpMV runs the same
// row kernel (the reference ignores trans and relies on a kept transpose,
// par_csr_matvec_device.c:445-447).
//
// Block (I, J) with bitmap m becomes block (J, I) with the transposed bitmap.
// A stable radix sort of block indices by block column keeps the source block
// rows ascending inside every transposed row, so the result is deterministic.
#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace amgt {

namespace {

__global__ void k_block_rows(int32_t brow, const int32_t *__restrict__ blk_ptr,
                             int32_t *__restrict__ row_of) {
    const int br = blockIdx.x * blockDim.x + threadIdx.x;
    if (br >= brow) return;
    for (int j = blk_ptr[br]; j < blk_ptr[br + 1]; ++j) row_of[j] = br;
}

__global__ void k_iota(int64_t n, int32_t *__restrict__ p) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = int32_t(i);
}

__global__ void k_zero(int64_t n, int32_t *__restrict__ p) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_col_counts(int64_t nblk, const int32_t *__restrict__ blk_col,
                             int32_t *__restrict__ cnt) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nblk) atomicAdd(&cnt[blk_col[i]], 1);
}

__global__ void k_popc_u16(int64_t n, const uint16_t *__restrict__ m, int32_t *__restrict__ out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = __popc(uint32_t(m[i]));
}

// dest block t takes source block perm[t]
__global__ void k_transpose_blocks(int64_t nblk, const int32_t *__restrict__ perm,
                                   const int32_t *__restrict__ row_of,
                                   const uint16_t *__restrict__ map,
                                   const int32_t *__restrict__ src_voff,
                                   const int32_t *__restrict__ dst_voff,
                                   const double *__restrict__ val, int32_t *__restrict__ tcol,
                                   uint16_t *__restrict__ tmap, double *__restrict__ tval) {
    int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nblk) return;
    const int j = perm[t];
    const uint32_t m = map[j];
    const uint32_t mt = bitmap_transpose(m);
    tcol[t] = row_of[j];
    tmap[t] = uint16_t(mt);
    if (!val) return;
    const int so = src_voff[j];
    int d = dst_voff[t];
    uint32_t bits = mt;
    while (bits) {
        const int b = __builtin_ctz(bits);  // dest bit i'*4 + j'
        bits &= bits - 1u;
        const int sb = ((b & 3) << 2) | (b >> 2);  // source bit j'*4 + i'
        tval[d++] = val[so + __popc(m & ((1u << sb) - 1u))];
    }
}

__global__ void k_gather_popc(int64_t n, const int32_t *__restrict__ perm,
                              const uint16_t *__restrict__ m, int32_t *__restrict__ out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = __popc(uint32_t(m[perm[i]]));
}

__global__ void k_tval_ptr(int32_t tbrow, const int32_t *__restrict__ tptr,
                           const int32_t *__restrict__ dst_voff, int64_t nblk, int64_t nnz,
                           int32_t *__restrict__ tval_ptr) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > tbrow) return;
    const int p = tptr[c];
    tval_ptr[c] = (p < nblk) ? dst_voff[p] : int32_t(nnz);
}

}  // namespace

int mbsr_transpose(const Mbsr &A, hipStream_t s, Mbsr *out) {
    Mbsr &T = *out;
    T = Mbsr();
    T.nrows = A.ncols;
    T.ncols = A.nrows;
    T.brow = A.bcol;
    T.bcol = A.brow;
    T.nblk = A.nblk;
    T.nnz = A.nnz;
    T.pattern_only = A.pattern_only;
    const int64_t nb = A.nblk;
    T.blk_ptr = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(T.brow) + 1), s));
    T.val_ptr = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(T.brow) + 1), s));
    T.blk_col = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * size_t(nb ? nb : 1), s));
    T.blk_map = static_cast<uint16_t *>(dev_alloc(sizeof(uint16_t) * size_t(nb ? nb : 1), s));
    if (!A.pattern_only)
        T.val = static_cast<double *>(dev_alloc(sizeof(double) * size_t(A.nnz ? A.nnz : 1), s));
    if (!T.blk_ptr || !T.val_ptr || !T.blk_col || !T.blk_map || (!A.pattern_only && !T.val)) {
        set_error("transpose: allocation failed");
        return AMGT_ERR_ALLOC;
    }
    int32_t *row_of = nullptr, *idx = nullptr, *perm = nullptr, *keys = nullptr, *cnt = nullptr;
    int32_t *pc = nullptr, *src_voff = nullptr, *dst_voff = nullptr;
    RowStats *rstats = nullptr;
    AMGT_TRY(tmp_alloc_t(&row_of, size_t(nb), s));
    AMGT_TRY(tmp_alloc_t(&idx, size_t(nb), s));
    AMGT_TRY(tmp_alloc_t(&perm, size_t(nb), s));
    AMGT_TRY(tmp_alloc_t(&keys, size_t(nb), s));
    AMGT_TRY(tmp_alloc_t(&cnt, size_t(T.brow) + 1, s));
    AMGT_TRY(tmp_alloc_t(&pc, size_t(nb) + 1, s));
    AMGT_TRY(tmp_alloc_t(&src_voff, size_t(nb) + 1, s));
    AMGT_TRY(tmp_alloc_t(&dst_voff, size_t(nb) + 1, s));
    AMGT_TRY(tmp_alloc_t(&rstats, 1, s));
    const unsigned gb = unsigned((nb + 255) / 256);
    if (A.brow > 0)
        hipLaunchKernelGGL(k_block_rows, dim3((A.brow + 255) / 256), dim3(256), 0, s, A.brow,
                           A.blk_ptr, row_of);
    hipLaunchKernelGGL(k_zero, dim3((T.brow + 256) / 256), dim3(256), 0, s, int64_t(T.brow) + 1,
                       cnt);
    if (nb > 0) {
        hipLaunchKernelGGL(k_iota, dim3(gb), dim3(256), 0, s, nb, idx);
        hipLaunchKernelGGL(k_col_counts, dim3(gb), dim3(256), 0, s, nb, A.blk_col, cnt);
        int end_bit = 1;
        while ((int64_t(1) << end_bit) <= int64_t(A.bcol) && end_bit < 31) ++end_bit;
        size_t bytes = 0;
        AMGT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, A.blk_col, keys, idx, perm,
                                                    int(nb), 0, end_bit, s));
        void *tmp = nullptr;
        AMGT_TRY(tmp_alloc(&tmp, bytes, s));
        AMGT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, A.blk_col, keys, idx, perm,
                                                    int(nb), 0, end_bit, s));
        tmp_free(tmp, s);
        hipLaunchKernelGGL(k_popc_u16, dim3(gb), dim3(256), 0, s, nb, A.blk_map, pc);
    }
    AMGT_TRY(exclusive_scan_i32(cnt, T.blk_ptr, T.brow, s));
    AMGT_TRY(exclusive_scan_i32(pc, src_voff, nb, s));
    if (nb > 0) {
        // popcount is invariant under transposition: dest block t has the
        // popcount of source block perm[t]
        hipLaunchKernelGGL(k_gather_popc, dim3(gb), dim3(256), 0, s, nb, perm, A.blk_map, keys);
    }
    AMGT_TRY(exclusive_scan_i32(keys, dst_voff, nb, s));
    if (nb > 0)
        hipLaunchKernelGGL(k_transpose_blocks, dim3(gb), dim3(256), 0, s, nb, perm, row_of,
                           A.blk_map, src_voff, dst_voff, A.pattern_only ? nullptr : A.val,
                           T.blk_col, T.blk_map, T.val);
    hipLaunchKernelGGL(k_tval_ptr, dim3((T.brow + 256) / 256), dim3(256), 0, s, T.brow, T.blk_ptr,
                       dst_voff, nb, A.nnz, T.val_ptr);
    AMGT_HIP(hipGetLastError());
    AMGT_TRY(row_stats_launch(T.blk_ptr, T.brow, rstats, s));
    RowStats hrs{};
    {
        const Readback rb[1] = {{&hrs, rstats, sizeof(RowStats)}};
        AMGT_TRY(readback(rb, 1, s));
    }
    tmp_free(row_of, s);
    tmp_free(idx, s);
    tmp_free(perm, s);
    tmp_free(keys, s);
    tmp_free(cnt, s);
    tmp_free(pc, s);
    tmp_free(src_voff, s);
    tmp_free(dst_voff, s);
    tmp_free(rstats, s);
    T.max_row_blocks = hrs.max_blocks;
    T.avg_nnz = nb ? double(A.nnz) / double(nb) : 0.0;
    T.stand = stand_from(hrs, nb, T.brow);
    T.spmv_kernel = default_spmv_kernel(T);
    return AMGT_OK;
}

}  // namespace amgt
