// mBSR -> CSR export (counterpart of BSR2CSR_step1-3 with bsr2csr_get_ptr and
// bsr2csr_get_idx_val, csr_spgemm_device.c:1344-1440, 1501-1525).
//
// Columns come out ascending within a row: blocks are stored by ascending block
// column and the bits of a row nibble by ascending column.  One wave per block
// row, 16 block slots x 4 rows per step; each lane's output offset is an
// exclusive scan (over slots) of its nibble popcount plus a per-row carry.
#include "internal.hpp"

namespace amgt {

namespace {

__global__ void k_row_counts(int32_t nrows, const int32_t *__restrict__ blk_ptr,
                             const uint16_t *__restrict__ blk_map, int32_t *__restrict__ cnt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int br = r >> 2;
    const int sh = (r & 3) * 4;
    int c = 0;
    for (int j = blk_ptr[br]; j < blk_ptr[br + 1]; ++j) c += __popc((blk_map[j] >> sh) & 0xFu);
    cnt[r] = c;
}

__global__ __launch_bounds__(256) void k_fill_csr(
    int32_t nrows, int32_t brow, const int32_t *__restrict__ blk_ptr,
    const int32_t *__restrict__ blk_col, const uint16_t *__restrict__ blk_map,
    const int32_t *__restrict__ val_ptr, const double *__restrict__ val,
    const int32_t *__restrict__ I, int32_t *__restrict__ J, double *__restrict__ A) {
    const int lane = threadIdx.x & 63;
    const int br = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (br >= brow) return;
    const int r = lane & 3;
    const int s = lane >> 2;  // 16 slots
    const int row = 4 * br + r;
    const int p0 = blk_ptr[br], p1 = blk_ptr[br + 1];
    int voff = val_ptr[br];
    int woff = (row < nrows) ? I[row] : 0;
    const uint32_t low = (1u << (4 * r)) - 1u;
    for (int jb = p0; jb < p1; jb += 16) {
        const int j = jb + s;
        uint32_t m = 0;
        int col = 0;
        if (j < p1) {
            m = blk_map[j];
            col = blk_col[j];
        }
        const uint32_t nib = (m >> (4 * r)) & 0xFu;
        const int cb = __popc(m);
        const int cn = __popc(nib);
        int ib = cb, in = cn;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            int tb = __shfl_up(ib, 4 * d);
            int tn = __shfl_up(in, 4 * d);
            if (s >= d) {
                ib += tb;
                in += tn;
            }
        }
        const int totb = __shfl(ib, 60 + r);
        const int totn = __shfl(in, 60 + r);
        int o = voff + (ib - cb) + __popc(m & low);
        int w = woff + (in - cn);
        if (row < nrows) {
            uint32_t t = nib;
            while (t) {
                const int c = __builtin_ctz(t);
                t &= t - 1u;
                J[w] = 4 * col + c;
                if (A) A[w] = val[o];
                ++w;
                ++o;
            }
        }
        voff += totb;
        woff += totn;
    }
}

}  // namespace

int mbsr_to_csr(const Mbsr &M, hipStream_t s, int64_t *nnz, int32_t **rowptr, int32_t **colidx,
                double **val) {
    *nnz = 0;
    *rowptr = nullptr;
    *colidx = nullptr;
    *val = nullptr;
    int32_t *I = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(M.nrows) + 1), s));
    if (!I) {
        set_error("CSR export: row pointer allocation failed");
        return AMGT_ERR_ALLOC;
    }
    int32_t *cnt = nullptr;
    int st = tmp_alloc_t(&cnt, size_t(M.nrows) + 1, s);
    if (st != AMGT_OK) {
        dev_free(I, s);
        return st;
    }
    if (M.nrows > 0)
        hipLaunchKernelGGL(k_row_counts, dim3((M.nrows + 255) / 256), dim3(256), 0, s, M.nrows,
                           M.blk_ptr, M.blk_map, cnt);
    st = exclusive_scan_i32(cnt, I, M.nrows, s);
    tmp_free(cnt, s);
    if (st != AMGT_OK) {
        dev_free(I, s);
        return st;
    }
    // nnz of the export equals the handle's nnz (sum of popcounts)
    const int64_t n = M.nnz;
    int32_t *J = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * size_t(n ? n : 1), s));
    double *A = nullptr;
    if (!M.pattern_only) A = static_cast<double *>(dev_alloc(sizeof(double) * size_t(n ? n : 1), s));
    if (!J || (!M.pattern_only && !A)) {
        dev_free(I, s);
        dev_free(J, s);
        dev_free(A, s);
        set_error("CSR export: allocation failed");
        return AMGT_ERR_ALLOC;
    }
    if (M.brow > 0)
        hipLaunchKernelGGL(k_fill_csr, dim3((M.brow + 3) / 4), dim3(256), 0, s, M.nrows, M.brow,
                           M.blk_ptr, M.blk_col, M.blk_map, M.val_ptr, M.val, I, J, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        dev_free(I, s);
        dev_free(J, s);
        dev_free(A, s);
        set_error(std::string("CSR export: ") + hipGetErrorString(e));
        return AMGT_ERR_HIP;
    }
    *nnz = n;
    *rowptr = I;
    *colidx = J;
    *val = A;
    return AMGT_OK;
}

}  // namespace amgt
