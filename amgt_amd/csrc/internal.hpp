// Internal declarations shared by the HIP translation units of libamgt_mbsr.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/amgt_mbsr.h"
#include "../../include/amgt_policy.h"

namespace amgt {

constexpr int kWave = 64;

// ---- error plumbing -------------------------------------------------------
void set_error(const std::string &msg);

struct Status {
    int code = AMGT_OK;
};

#define AMGT_HIP(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::amgt::set_error(std::string(#call) + ": " + hipGetErrorString(e_));          \
            return AMGT_ERR_HIP;                                                           \
        }                                                                                  \
    } while (0)

#define AMGT_TRY(expr)                 \
    do {                               \
        int s_ = (expr);               \
        if (s_ != AMGT_OK) return s_;  \
    } while (0)

#define AMGT_CHECK_ARG(cond, msg)                 \
    do {                                          \
        if (!(cond)) {                            \
            ::amgt::set_error(msg);               \
            return AMGT_ERR_ARG;                  \
        }                                         \
    } while (0)

// ---- wave primitives ----------------------------------------------------------
// Word 3 of a raw buffer resource on gfx9 parts (32-bit data format; no
// swizzle): loads past num_records return zeros instead of faulting.
constexpr int kBufferDword3 = 0x00020000;

// Inclusive sum / max over the 64 lanes of a wave from DPP moves: row shifts
// by 1, 2, 4, 8 within each row of 16 lanes, then the last lane of a row
// broadcast into the next row (row_bcast:15) and lane 31 into rows 2 and 3
// (row_bcast:31).  Lanes a move does not reach add 0 (the `old` operand).
// One VALU operation per step, no LDS traffic.
__device__ __forceinline__ int wave_incl_sum(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
    return v;
}
// (v >= 0)
__device__ __forceinline__ int wave_incl_max(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));
    return v;
}

// ---- allocation -------------------------------------------------------------
// Long-lived arrays (handle arrays, CSR exports) go through the user hooks
// when set, else the library's device pool, stream-ordered on s.
void *dev_alloc(size_t bytes, hipStream_t s);
void dev_free(void *p, hipStream_t s);
// Stream-ordered temporaries from the library's device pool.
int tmp_alloc(void **p, size_t bytes, hipStream_t s);
void tmp_free(void *p, hipStream_t s);

template <typename T>
int tmp_alloc_t(T **p, size_t n, hipStream_t s) {
    return tmp_alloc(reinterpret_cast<void **>(p), sizeof(T) * (n ? n : 1), s);
}

// Device->host readbacks of sizes and counters, all at once: each item is
// copied into a pinned per-thread staging area, one stream synchronisation
// covers them, then they are copied to their destinations.  (A copy into
// pageable memory waits for the stream and stages through a driver buffer
// on its own: about 12 us each on MI355X, against one sync for a batch.)
// At most kReadbackBytes in total; every caller synchronises before it
// returns, so the area is free again.
struct Readback {
    void *dst;
    const void *src;
    size_t bytes;
};
constexpr size_t kReadbackBytes = 4096;
int readback(const Readback *items, int n, hipStream_t s);

// ---- handle -----------------------------------------------------------------
struct Mbsr {
    int32_t nrows = 0, ncols = 0, brow = 0, bcol = 0;
    int64_t nblk = 0, nnz = 0;
    int32_t *blk_ptr = nullptr;
    int32_t *blk_col = nullptr;
    uint16_t *blk_map = nullptr;
    int32_t *val_ptr = nullptr;
    double *val = nullptr;
    float *val32 = nullptr;   // lazily made FP32 copy
    __half *val16 = nullptr;  // lazily made FP16 copy
    int8_t half_fits = -1;    // values finite, FP16 copy possible (1), not (0), not checked yet (-1)
    double half_scale = 1.0;  // val16 = half_scale * val (a power of two: max |val16| near 2^14)
    bool pattern_only = false;
    double stand = 0.0, avg_nnz = 0.0;
    int32_t max_row_blocks = 0;
    int32_t level = 0;
    int32_t precision = -1;   // -1: follow the level policy
    int32_t spmv_lanes = 0;   // 0: heuristic
    int32_t spmv_kernel = 2;  // kSpmv* kind (amgt_mbsr_set_spmv_kernel); block-lane by default
    Mbsr *trans = nullptr;    // cached A^T
    // tile tables of the balanced SpMV (built on first use, for one tile size)
    int32_t bal_capb = 0, bal_ntiles = 0;
    int32_t *bal_lo = nullptr, *bal_head = nullptr, *bal_vb = nullptr;
    double *bal_carry = nullptr;  // per-tile partial sums of rows cut by tile edges
};

}  // namespace amgt

struct amgt_mbsr {
    amgt::Mbsr m;
};

namespace amgt {

bool mixed_precision_enabled();
// precision the SpMV of A runs in: its explicit one, else the level policy
// when mixed precision is on, else FP64 (FP32 for values without an FP16 copy)
int effective_precision(const Mbsr &A);

// SpMV kernel kinds (amgt_mbsr_set_spmv_kernel)
enum : int { kSpmvDirect = 0, kSpmvStaged = 1, kSpmvBlock = 2, kSpmvBalanced = 3 };

// Spread of blocks per block row above which the balanced SpMV is the default
// for matrices too large for the direct walk (the reference's switch,
// csr_matvec_device.c:1336)
constexpr double kBalancedStand = 12.0;
// SpMV kernel kind chosen from the matrix statistics (spmv.hip): the direct
// walk for small operators and short-row ones, the balanced kernel for a large
// spread of blocks per row, else the block-lane kernel.
int default_spmv_kernel(const Mbsr &A);
// applies the statistics choice to a handle (kind; lanes left to the rule)
inline void apply_spmv_defaults(Mbsr &A) {
    A.spmv_kernel = default_spmv_kernel(A);
    A.spmv_lanes = 0;
}

// ---- component entry points (each in its own translation unit) -------------
int csr_to_mbsr(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *rowptr,
                const int32_t *colidx, const double *val, hipStream_t s, Mbsr *out);
int mbsr_refresh_values(Mbsr *A, const int32_t *rowptr, const int32_t *colidx,
                        const double *val, hipStream_t s);
int mbsr_transpose(const Mbsr &A, hipStream_t s, Mbsr *out);
// Makes the value copy that *precision needs.  The FP16 copy holds the values
// times a power of two (half_scale) that puts max |a| in [2^14, 2^15), so that
// small values stay out of FP16's subnormal range and large ones fit (max |a|
// is measured once per value set); every consumer divides by it exactly.  A
// matrix with non-finite values gets no FP16 copy: *precision becomes FP32.
int mbsr_ensure_lowp(Mbsr *A, int *precision, hipStream_t s);
// builds what the handle's SpMV kernel kind needs for this lanes setting (the
// balanced kernel's tile tables); no-op for the other kinds
int spmv_prepare(Mbsr *A, int lanes, hipStream_t s);
int balanced_bpt(int lanes);
void mbsr_release_balanced(Mbsr *A, hipStream_t s);
int spmv(const Mbsr &A, int precision, int lanes, double alpha, const double *x, double beta,
         const double *b, double *y, hipStream_t s);
// y = alpha*A*x + beta*b (FP64 values) with the balanced kernel over a caller's
// workspace of ws_bytes instead of the handle's tile tables: nothing is
// allocated, the tables are rebuilt on every call.  AMGT_ERR_UNSUPPORTED when
// the workspace is smaller than spmv_balanced_ws_bytes(A, lanes).  x must be
// 16-byte aligned.
size_t spmv_balanced_ws_bytes(const Mbsr &A, int lanes);
int spmv_balanced_ws(const Mbsr &A, int lanes, double alpha, const double *x, double beta,
                     const double *b, double *y, void *ws, size_t ws_bytes, hipStream_t s);
int spmm(const Mbsr &A, int precision, int lanes, int k, double alpha, const double *x,
         int64_t ldx, double beta, const double *b, int64_t ldb, double *y, int64_t ldy,
         hipStream_t s);
int spmv_default_lanes(const Mbsr &A);
// k right-hand sides interleaved (entry (row i, vector v) at x[i * k + v])
int spmm_interleaved(const Mbsr &A, int precision, int k, double alpha, const double *x,
                     double beta, const double *b, double *y, hipStream_t s);
// Residual with its restriction in one launch (the block SpMV kernel with
// tiles of rpb block rows): r = f - A u, and fc = R r, each coarse row summed
// over R's row c in order by the last tile of A that touches it.  The tables
// come from amgt_restrict_plan_create: per tile its coarse rows, per coarse
// row the number of tiles touching it and an arrival counter (zero between
// calls), R in CSR with ascending fine rows.  u 16-byte aligned.
struct RestrictTables {
    int32_t rpb;
    const int32_t *tile_ptr, *tile_c, *nspan;
    int32_t *cnt;
    const int32_t *ri, *rj;
    const double *rv;
};
int residual_restrict(const Mbsr &A, int precision, const RestrictTables &t, const double *f,
                      const double *u, double *r, double *fc, hipStream_t s);
// block rows per tile of the block SpMV kernel for a lanes setting
int blk_tile_rows(int lanes);
// u_new = u + omega * dinv .* (f - A u) in one SpMV launch (square A)
int relax_jacobi(const Mbsr &A, int precision, int lanes, double omega, const double *dinv,
                 const double *f, const double *u, double *u_new, hipStream_t s);
// diagonal of A (kind AMGT_DIAG_VALUE) or 1/|a_ii| (AMGT_DIAG_INV_ABS, 0 where a_ii = 0)
int mbsr_diagonal(const Mbsr &A, int kind, double *out, hipStream_t s);
// y = beta*b over n entries (beta == 0 writes zeros without reading b)
int scale_vector(int32_t n, double beta, const double *b, double *y, hipStream_t s);
// C = A*B with operand values in `precision` (AMGT_DOUBLE / FLOAT / HALF; the
// FP32/FP16 copies of A and B must exist, mbsr_ensure_lowp)
int spgemm(const Mbsr &A, const Mbsr &B, int precision, hipStream_t s, Mbsr *C);
int mbsr_to_csr(const Mbsr &A, hipStream_t s, int64_t *nnz, int32_t **rowptr, int32_t **colidx,
                double **val);
void mbsr_release(Mbsr *A, hipStream_t s);

// Totals of a row pass (rowpass.hpp): rows per bin, the largest value and the
// sum of squares.  For blocks per block row (row_stats_launch) max_blocks and
// sumsq give max_row_blocks and, through stand_from, the standard deviation
// of blocks per block row.
struct RowStats {
    int32_t cnt[3];
    int32_t max_blocks;
    unsigned long long sumsq;
};
int row_stats_launch(const int32_t *blk_ptr, int32_t brow, RowStats *d_stats, hipStream_t s);
double stand_from(const RowStats &h, int64_t nblk, int32_t brow);

// exclusive scan of n int32 counts into out[0..n] (out[n] = total); int64 total
// is returned through *total if non-null (synchronises the stream)
int exclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t s);
int exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s);
// CSR transpose on the device (nrows x ncols -> ncols x nrows): every row of
// the transpose lists its columns (the input rows) in ascending order, values
// (when V and TV are given) moved with them; a stable radix sort by column.
// TI[ncols+1], TJ[nnz], TV[nnz] are the caller's.
int csr_transpose_device(int32_t nrows, int32_t ncols, const int32_t *I, const int32_t *J,
                         const double *V, int32_t nnz, int32_t *TI, int32_t *TJ, double *TV,
                         hipStream_t s);

// ---- device helpers ----------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t bitmap_mul(uint32_t a, uint32_t b) {
    // bitMatrixMul, csr_spgemm_device.c:189-205: OR_k (col_k(a) x row_k(b))
    uint32_t res = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t col = (a >> k) & 0x1111u;
        uint32_t row = (b >> (4 * k)) & 0xFu;
        res |= (col * 0xFu) & (row * 0x1111u);
    }
    return res;
}

// transpose of a 4x4 occupancy bitmap: bit r*4+c -> bit c*4+r
__device__ __forceinline__ uint32_t bitmap_transpose(uint32_t m) {
    uint32_t t = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
            t |= ((m >> (r * 4 + c)) & 1u) << (c * 4 + r);
    return t;
}

}  // namespace amgt
