// Smoother support for the V-cycle: the matrix diagonal and the weighted
// Jacobi sweep fused into the SpMV epilogue (spmv.hip, EpJacobi).
//
// HYPRE's relax type 18 on the GPU (hypre_BoomerAMGRelax18WeightedL1Jacobi ->
// hypre_BoomerAMGRelax7Jacobi, par_relax.c:338-355, 1178-1250) is three
// kernels: Vtemp = f, Vtemp = w f - w A u (matvec), u += Vtemp ./ l1_norms.
// Here one SpMV launch writes u_new = u + w (f - A u) ./ l1 directly.  The l1
// norms of relax 18 (hypre_ParCSRComputeL1Norms option 4, ams.c:659-690) are
// |a_ii| plus half the l1 norm of the off-processor part; on one GPU there is
// no off-processor part, so l1_i = |a_ii|.
#include "internal.hpp"

namespace amgt {

namespace {

// one thread per row: find the diagonal block (block column == block row) and
// the entry (r, r) in it
__global__ __launch_bounds__(256) void k_diagonal(int32_t n, int kind,
                                                  const int32_t *__restrict__ blk_ptr,
                                                  const int32_t *__restrict__ blk_col,
                                                  const uint16_t *__restrict__ blk_map,
                                                  const int32_t *__restrict__ val_ptr,
                                                  const double *__restrict__ val,
                                                  double *__restrict__ out) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const int br = row >> 2, rr = row & 3;
    const int p0 = blk_ptr[br], p1 = blk_ptr[br + 1];
    double d = 0.0;
    int lo = p0, hi = p1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (blk_col[mid] < br) lo = mid + 1; else hi = mid;
    }
    if (lo < p1 && blk_col[lo] == br) {
        const uint32_t m = blk_map[lo];
        const int bit = 5 * rr;  // entry (rr, rr) of the 4x4 block
        if ((m >> bit) & 1u) {
            int off = val_ptr[br];
            for (int q = p0; q < lo; ++q) off += __popc(uint32_t(blk_map[q]));
            d = val[off + __popc(m & ((1u << bit) - 1u))];
        }
    }
    if (kind == AMGT_DIAG_INV_ABS) d = (d != 0.0) ? 1.0 / fabs(d) : 0.0;
    out[row] = d;
}

}  // namespace

int mbsr_diagonal(const Mbsr &A, int kind, double *out, hipStream_t s) {
    const int32_t n = A.nrows < A.ncols ? A.nrows : A.ncols;
    if (n == 0) return AMGT_OK;
    hipLaunchKernelGGL(k_diagonal, dim3((n + 255) / 256), dim3(256), 0, s, n, kind, A.blk_ptr,
                       A.blk_col, A.blk_map, A.val_ptr, A.val, out);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

}  // namespace amgt
