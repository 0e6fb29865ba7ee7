// Smoother support for the V-cycle: the matrix diagonal, the l1 norms of the
// rows, and the weighted Jacobi sweep fused into the SpMV epilogue
// (spmv.hip, EpJacobi).
//
// HYPRE's relax types 7 and 18 on one process (hypre_BoomerAMGRelax7Jacobi,
// par_relax.c:1178-1250; relax 18 calls it too, :338-355) are three steps:
// Vtemp = f, Vtemp = w f - w A u (matvec), u += Vtemp ./ l1.  Here one SpMV
// launch writes u_new = u + w (f - A u) .* (1 ./ l1) directly.  The l1 vector
// comes from hypre_ParCSRComputeL1Norms (ams.c:527-800) as par_amg_setup.c
// selects it:
//  * relax 18: option 1 (par_amg_setup.c:3393-3406), the l1 norm of the whole
//    row, sum_j |a_ij| (ams.c:593-602), negated where a_ii < 0 (:786-799);
//  * relax 7: option 5 (par_amg_setup.c:3521-3536), the diagonal a_ii, with 1
//    where it is 0 (ams.c:700-725).
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
    double l1 = 0.0;
    if (kind == AMGT_DIAG_INV_L1) {
        // sum of |a_ij| over row `row`: the row's nibble of every block
        int off = val_ptr[br];
        for (int q = p0; q < p1; ++q) {
            const uint32_t m = blk_map[q];
            const uint32_t nib = (m >> (4 * rr)) & 0xFu;
            int v = off + __popc(m & ((1u << (4 * rr)) - 1u));
            for (int c = 0; c < 4; ++c)
                if ((nib >> c) & 1u) l1 += fabs(val[v++]);
            off += __popc(m);
        }
    }
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
    if (kind == AMGT_DIAG_INV_ABS) {
        d = (d != 0.0) ? 1.0 / fabs(d) : 0.0;
    } else if (kind == AMGT_DIAG_INV_L1) {
        if (d < 0.0) l1 = -l1;
        d = (l1 != 0.0) ? 1.0 / l1 : 0.0;
    } else if (kind == AMGT_DIAG_INV) {
        d = (d != 0.0) ? 1.0 / d : 1.0;
    }
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
