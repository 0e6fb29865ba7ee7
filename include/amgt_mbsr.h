/*
 * amgt_mbsr.h -- C-ABI of the MI355X (gfx950) mBSR library.
 *
 * The library replaces AmgT's mBSR engine inside HYPRE's sequential CSR layer
 * (reference: AmgT_HYPRE/src/seq_mv/csr_matvec_device.c and
 * csr_spgemm_device.c).  Plain C types only: device pointers, sizes, a HIP
 * stream passed as `void *` (hipStream_t), int status codes.  Every entry point
 * is stream-ordered on the stream it is given.  Global mutable state: the
 * allocator hooks, the last-error string (per thread) and, without hooks, the
 * library's cache of device blocks (freed blocks are kept for reuse until the
 * process ends).
 *
 * mBSR layout (device memory, owned by the handle):
 *   blk_ptr[brow+1]  int32   first block of each 4-row block row
 *   blk_col[nblk]    int32   block column, strictly increasing within a row
 *   blk_map[nblk]    uint16  occupancy bitmap, bit r*4+c = entry (r, c)
 *   val_ptr[brow+1]  int32   first value of each block row (= CSR I[4*br])
 *   val[nnz]         double  nonzeros only, in (block, bit) order
 * Semantics of blk_ptr/blk_col/blk_map follow bsrMAT, seq_mv.h:105-129; the
 * reference stores 16 dense values per block (csr_matvec_device.c:936-945),
 * this library stores only the nonzeros.
 */
#ifndef AMGT_MBSR_H
#define AMGT_MBSR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
enum {
    AMGT_OK = 0,
    AMGT_ERR_ARG = 1,          /* invalid argument (null pointer, negative size, x == y ...) */
    AMGT_ERR_DIM = 2,          /* incompatible dimensions (csr_matop_device.c:257-263) */
    AMGT_ERR_ALLOC = 3,        /* device allocation failed */
    AMGT_ERR_HIP = 4,          /* a HIP runtime call failed */
    AMGT_ERR_INPUT = 5,        /* CSR input invalid: duplicate entry or column out of range */
    AMGT_ERR_UNSUPPORTED = 6   /* valid input outside the sizes this build handles */
};

/* AMGT_PRECISION, seq_mv.h:131-136 */
enum { AMGT_DOUBLE = 0, AMGT_FLOAT = 1, AMGT_HALF = 2 };

typedef struct amgt_mbsr *amgt_mbsr_t;
typedef struct amgt_restrict_plan *amgt_restrict_plan_t;

typedef struct {
    int32_t nrows, ncols, brow, bcol;
    int64_t nblk, nnz;
    double stand;       /* std-dev of blocks per block row (getStand, csr_matvec_device.c:638-671) */
    double avg_nnz;     /* nnz / nblk (csr_matvec_device.c:1225) */
    double fill;        /* nnz / (16 * nblk) */
    int32_t max_row_blocks;
    int32_t level;      /* coarsening level tag (hypre_Level, seq_mv.h:199) */
    int32_t precision;  /* value precision used by SpMV (AMGT_DOUBLE/FLOAT/HALF) */
    int32_t spmv_lanes; /* lanes per block row chosen by the tuner (4..64) */
    int32_t spmv_kernel; /* AMGT_SPMV_DIRECT / _STAGED / _BLOCK / _BALANCED */
} amgt_mbsr_info_t;

typedef struct {
    const int32_t *blk_ptr, *blk_col, *val_ptr;
    const uint16_t *blk_map;
    const double *val;
} amgt_mbsr_arrays_t;

/* ---- library ---- */
const char *amgt_status_string(int status);
/* message of the last failing call on the calling thread */
const char *amgt_last_error(void);
/* Device allocator hooks.  HYPRE passes wrappers of hypre_TAlloc/hypre_TFree
 * (HYPRE_MEMORY_DEVICE) so that CSR arrays returned by amgt_csr_spgemm are
 * owned and freed by hypre_CSRMatrixDestroy.  NULL restores hipMalloc/hipFree. */
typedef void *(*amgt_alloc_fn)(size_t bytes, void *ctx);
typedef void (*amgt_free_fn)(void *ptr, void *ctx);
int amgt_set_allocator(amgt_alloc_fn alloc_fn, amgt_free_fn free_fn, void *ctx);
/* level -> precision policy, get_calculationPrecison csr_spgemm_device.c:1847-1862 */
int amgt_precision_for_level(int level);
/* 1: SpMV/SpGEMM follow the level policy (reference MIXED_PRESION build,
 * seq_mv.h:75); 0 (default): FP64 everywhere (reference AmgT_FP64 build). */
int amgt_set_mixed_precision(int enable);

/* ---- mBSR handle ---- */
/* CSR2BSR_GPU, csr_matvec_device.c:1201-1267.  Rows may be in any column
 * order (HYPRE keeps the diagonal first).  d_val may be NULL (pattern only). */
int amgt_mbsr_create_from_csr(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *d_rowptr,
                              const int32_t *d_colidx, const double *d_val, void *stream,
                              amgt_mbsr_t *out);
/* frees the handle and every array it owns (csr_matrix.c:94-103) */
int amgt_mbsr_destroy(amgt_mbsr_t A);
/* Re-reads values from a CSR with the same pattern (the reference never
 * invalidates its cache; this is the refresh hook). */
int amgt_mbsr_update_values(amgt_mbsr_t A, const int32_t *d_rowptr, const int32_t *d_colidx,
                            const double *d_val, void *stream);
int amgt_mbsr_get_info(amgt_mbsr_t A, amgt_mbsr_info_t *info);
int amgt_mbsr_get_arrays(amgt_mbsr_t A, amgt_mbsr_arrays_t *arrays);
int amgt_mbsr_set_level(amgt_mbsr_t A, int level);
/* explicit value precision of SpMV with A and of SpGEMM with A as left
 * operand; overrides the level policy for this handle (-1 restores it) */
int amgt_mbsr_set_precision(amgt_mbsr_t A, int precision);
/* lanes per block row for SpMV: 4, 8, 16, 32 or 64; 0 = choose from the
 * matrix statistics */
int amgt_mbsr_set_spmv_lanes(amgt_mbsr_t A, int lanes);
/* SpMV kernel kind:
 *   AMGT_SPMV_DIRECT    lane per (block slot, row) walks bitmaps, columns and values in
 *                       global memory
 *   AMGT_SPMV_STAGED    a workgroup stages its tile of block rows in LDS, lanes walk it
 *   AMGT_SPMV_BLOCK     lane per 4x4 block over an LDS-staged tile of block rows, whole
 *                       x segments, per-row partial sums reduced in LDS
 *   AMGT_SPMV_BALANCED  the block kernel over equal tiles of blocks whatever rows they
 *                       belong to; rows cut by tile edges are finished by a second pass
 *                       (bsr_spmv_balanced_tc_fp64 + BSR_BALANCED_PREPROCESS_GPU,
 *                       csr_matvec_device.c:556-613, 1269-1300)
 * Default: BALANCED when the spread of blocks per block row (stand) is >= 12, as in
 * the reference (csr_matvec_device.c:1336), else BLOCK.  The staged, block and
 * balanced kernels need a 16-byte aligned x; otherwise the direct kernel runs. */
enum { AMGT_SPMV_DIRECT = 0, AMGT_SPMV_STAGED = 1, AMGT_SPMV_BLOCK = 2, AMGT_SPMV_BALANCED = 3 };
int amgt_mbsr_set_spmv_kernel(amgt_mbsr_t A, int kind);
/* times every lanes-per-block-row variant on this matrix and keeps the fastest */
int amgt_mbsr_autotune(amgt_mbsr_t A, void *stream);
/* The kernel kind and lanes chosen from the matrix statistics (what a new handle gets):
 * the direct walk up to 150k blocks, up to 1M blocks with <= 5 blocks per block row or
 * up to 350k with <= 8, with lanes per block row so that about 2^18 lanes share the
 * walk; else the balanced kernel when the spread of blocks per block row is >= 12
 * (tiles of 256 blocks), else the block-lane kernel with the largest tile whose average
 * load fits 90 % of its 512 blocks.  Fitted to amgt_mbsr_spmv_sweep over every
 * operator of the AMG hierarchies of the benchmark configurations.  Resets settings made
 * with amgt_mbsr_set_spmv_kernel / _lanes or amgt_mbsr_autotune. */
int amgt_mbsr_tune_from_stats(amgt_mbsr_t A);
/* The timing sweep behind amgt_mbsr_autotune without changing the handle:
 * times_ms[k * 5 + c] = mean time of y = A x with kernel kind k (AMGT_SPMV_BLOCK,
 * _BALANCED, _STAGED, _DIRECT in that order) and lanes {4, 8, 16, 32, 64}[c] (best of
 * two sweeps of 10 launches; 1e30 where the pair does not exist).  Synchronises. */
int amgt_mbsr_spmv_sweep(amgt_mbsr_t A, void *stream, float *times_ms);

/* ---- SpMV: y = alpha*op(A)*x + beta*b ----
 * spmv_amgT_fp64 csr_matvec_device.c:1302-1411 (+ fp32 2205-2309, fp16
 * 2098-2203).  op = A (trans 0) or A^T (trans 1, through a cached transposed
 * mBSR).  b may equal y; x must not alias y.  beta == 0 never reads b.
 * Values in the handle's precision; FP32 accumulation below FP64; x, b, y are
 * FP64. */
int amgt_mbsr_spmv(amgt_mbsr_t A, int trans, double alpha, const double *d_x, double beta,
                   const double *d_b, double *d_y, void *stream);
/* ---- PCG vector kernels (hypre_PCGSolve, krylov/pcg.c:320-900, two-norm form) ----
 * For the solve driver: every scalar stays in device memory (d_*: one double each),
 * so one iteration is a fixed sequence of launches that a graph can capture.  Dots are
 * deterministic (per-workgroup partials over fixed ranges, summed in order by one
 * workgroup).  d_work: amgt_pcg_work_bytes() bytes of device memory, shared by the calls
 * of one stream.
 *   amgt_dot            d_out = x . y
 *   amgt_pcg_update_xr  alpha = rz / pAp; x += alpha p; r -= alpha Ap; d_rr = r . r
 *   amgt_pcg_update_p   beta = rz_new / rz; p = z + beta p; then rz = rz_new */
size_t amgt_pcg_work_bytes(void);
int amgt_dot(int64_t n, const double *d_x, const double *d_y, double *d_out, void *d_work,
             void *stream);
int amgt_pcg_update_xr(int64_t n, const double *d_rz, const double *d_pap, const double *d_p,
                       const double *d_ap, double *d_x, double *d_r, double *d_rr, void *d_work,
                       void *stream);
int amgt_pcg_update_p(int64_t n, const double *d_rz_new, double *d_rz, const double *d_z,
                      double *d_p, void *stream);

/* ---- residual and restriction in one launch ----
 * The V-cycle's r = f - A u and f_c = R r (par_cycle.c:443 and :704: a matvec, then
 * the restriction matvec with R = P^T) as one launch of the block SpMV kernel: each
 * tile stores its rows of r, then arrives at the coarse rows its fine rows touch; the
 * last tile to arrive at coarse row c forms f_c[c] over R's row c in order (the same
 * sum whichever tile is last, so results are deterministic).  The plan (per level:
 * R in CSR, the tiles' coarse rows, arrival counters kept at zero between calls) is
 * built once for A's lanes setting; the call allocates nothing.  A square, R's
 * columns = A's rows; values in A's precision, R in FP64; u 16-byte aligned; r must
 * not alias f or u.  Calls with one plan must not overlap (shared counters). */
int amgt_restrict_plan_create(amgt_mbsr_t A, amgt_mbsr_t R, void *stream,
                              amgt_restrict_plan_t *plan);
int amgt_restrict_plan_destroy(amgt_restrict_plan_t plan);
int amgt_mbsr_residual_restrict(amgt_mbsr_t A, amgt_restrict_plan_t plan, const double *d_f,
                                const double *d_u, double *d_r, double *d_fc, void *stream);

/* ---- SpMV that allocates nothing ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  y = alpha*A*x + beta*b over the handle's FP64 values with
 * the balanced kernel (AMGT_SPMV_BALANCED at the handle's lanes setting; the
 * same results as amgt_mbsr_spmv with that kernel and FP64).  Its tile tables
 * and carry sums live in one library-owned device array of
 * AMGT_NOALLOC_SCRATCH_BYTES per GPU (76 bytes per tile of 256/512/1024
 * blocks: up to about 28M/56M/112M blocks), rebuilt on every call by three
 * small launches; larger matrices return AMGT_ERR_UNSUPPORTED.  Calls that
 * overlap on one GPU (two streams or host threads) share the array and race:
 * serialise them.  x must be 16-byte aligned; b may equal y; x must not alias
 * y; beta == 0 never reads b. */
#define AMGT_NOALLOC_SCRATCH_BYTES (8u * 1024u * 1024u)
int amgt_mbsr_spmv_noalloc(amgt_mbsr_t A, double alpha, const double *d_x, double beta,
                           const double *d_b, double *d_y, void *stream);
/* k right-hand sides, column-major with leading dimensions ldx/ldb/ldy
 * (HYPRE multivector storage method 0, vector.c) */
int amgt_mbsr_spmm(amgt_mbsr_t A, int trans, int k, double alpha, const double *d_x, int64_t ldx,
                   double beta, const double *d_b, int64_t ldb, double *d_y, int64_t ldy,
                   void *stream);

/* New handle holding A^T (hypre_CSRMatrixTransposeDevice, csr_matop_device.c:2471, as
 * used for R = P^T): blocks sorted by block column with a stable radix sort,
 * bitmaps and values transposed; block columns ascending within a row. */
int amgt_mbsr_transpose(amgt_mbsr_t A, void *stream, amgt_mbsr_t *AT);

/* ---- smoother support (V-cycle) ---- */
enum { AMGT_DIAG_VALUE = 0, AMGT_DIAG_INV_ABS = 1, AMGT_DIAG_INV_L1 = 2, AMGT_DIAG_INV = 3 };
/* For i < min(nrows, ncols):
 *   AMGT_DIAG_VALUE    d[i] = a_ii
 *   AMGT_DIAG_INV_ABS  d[i] = 1/|a_ii| (0 where a_ii = 0)
 *   AMGT_DIAG_INV_L1   d[i] = 1/l1_i with l1_i = sum_j |a_ij|, negated where a_ii < 0 (0
 *                      where the row is empty): the inverse l1 norms of relax type 18
 *                      (hypre_ParCSRComputeL1Norms option 1, ams.c:593-602, 786-799, as
 *                      par_amg_setup.c:3393-3406 selects it)
 *   AMGT_DIAG_INV      d[i] = 1/a_ii (1 where a_ii = 0): the diagonal of relax type 7
 *                      (option 5, ams.c:700-725, par_amg_setup.c:3521-3536) */
int amgt_mbsr_diagonal(amgt_mbsr_t A, int kind, double *d_diag, void *stream);
/* One weighted l1-Jacobi sweep, u_new = u + omega * dinv .* (f - A u), fused into a
 * single SpMV launch (relax type 18: hypre_BoomerAMGRelax18WeightedL1Jacobi ->
 * hypre_BoomerAMGRelax7Jacobi, par_relax.c:338-355, 1178-1250, which takes a copy, a
 * matvec and an element-wise update).  A square; u_new must not alias u; values in
 * the handle's precision as for SpMV. */
int amgt_mbsr_relax_jacobi(amgt_mbsr_t A, double omega, const double *d_dinv, const double *d_f,
                           const double *d_u, double *d_u_new, void *stream);

/* k right-hand sides interleaved (HYPRE multivector storage method 1): entry (row i,
 * vector v) at x[i*k + v], likewise b and y.  MFMA kernel: one v_mfma_f64_4x4x4f64
 * advances 16 right-hand sides of a 4x4 block; amgt_mbsr_spmm with k >= 8 transposes
 * to this layout and back. */
int amgt_mbsr_spmm_interleaved(amgt_mbsr_t A, int trans, int k, double alpha, const double *d_x,
                               double beta, const double *d_b, double *d_y, void *stream);

/* ---- SpGEMM: C = A*B on mBSR (spgemm_amgT_fp64 csr_spgemm_device.c:1527-1846) ----
 * Operand precision follows A (level policy in mixed mode, or A's explicit
 * precision): FP64, or FP32 / FP16 copies of both operands (spgemm_amgT_fp32
 * 2224-2578, spgemm_amgT_fp16 1863-2223); products of the rounded operands
 * accumulate in FP64 and, below FP64, every C entry is rounded to FP32.  C's
 * values are returned as FP64. */
int amgt_mbsr_spgemm(amgt_mbsr_t A, amgt_mbsr_t B, void *stream, amgt_mbsr_t *C);
/* mBSR -> CSR (BSR2CSR_step1-3, csr_spgemm_device.c:1501-1525), columns
 * ascending within a row; arrays come from the allocator hooks */
int amgt_mbsr_to_csr(amgt_mbsr_t A, void *stream, int64_t *nnz, int32_t **d_rowptr,
                     int32_t **d_colidx, double **d_val);
/* frees a pointer obtained from this library's allocator */
int amgt_free(void *ptr);

/* ---- host matrix readers (host CSR out, ascending columns; free with amgt_host_free) ----
 * IJ files <prefix>.00000 .. <prefix>.<nfiles-1> (hypre_IJMatrixRead, IJ_mv/IJMatrix.c:112-242;
 * hypre_CSRMatrixRead / Print* in seq_mv.h:524-530): header "ilower iupper jlower jupper",
 * lines "i j value" (0-based); entries in the file's own rows are set (the last one wins),
 * others are added to the owner's value.  All files are merged into one matrix. */
int amgt_read_ij(const char *prefix, int nfiles, int32_t *nrows, int32_t *ncols, int64_t *nnz,
                 int32_t **I, int32_t **J, double **V);
/* Matrix Market coordinate real/integer, general or symmetric (HYPRE_IJMatrixReadMM) */
int amgt_read_mm(const char *path, int32_t *nrows, int32_t *ncols, int64_t *nnz, int32_t **I,
                 int32_t **J, double **V);
void amgt_host_free(void *p);

/* ---- HYPRE-shaped entry points (plain pointers; the mBSR cache lives in a
 *      caller-owned slot, the `bsr_mat` field of hypre_CSRMatrix,
 *      seq_mv.h:156-163) ---- */
/* hypreDevice_CSRSpGemm, csr_spgemm_device.c:2579-2608 (seq_mv.h:727).
 * A_data or B_data NULL: pattern only, C_data returned NULL.  nnzA or
 * nnzB == 0: C has a zeroed row pointer (csr_spgemm_device.c:2633-2638). */
int amgt_csr_spgemm(int32_t m, int32_t k, int32_t n,
                    int64_t nnzA, const int32_t *A_i, const int32_t *A_j, const double *A_data,
                    amgt_mbsr_t *A_cache,
                    int64_t nnzB, const int32_t *B_i, const int32_t *B_j, const double *B_data,
                    amgt_mbsr_t *B_cache,
                    int64_t *nnzC, int32_t **C_i, int32_t **C_j, double **C_data,
                    amgt_mbsr_t *C_cache, void *stream);
/* hypre_CSRMatrixMatvecDevice, csr_matvec_device.c:128-191 (seq_mv.h:562):
 * y = alpha*op(A)*x + beta*b, num_vectors == 1, offset == 0. */
int amgt_csr_matvec(int trans, double alpha, int32_t nrows, int32_t ncols, int64_t nnz,
                    const int32_t *A_i, const int32_t *A_j, const double *A_data,
                    amgt_mbsr_t *A_cache, const double *d_x, double beta, const double *d_b,
                    double *d_y, void *stream);

/* ---- BoomerAMG setup on the device, one process (CSR in, device pointers;
 *      results equal HYPRE's host algorithms: par_strength.c:75,
 *      par_coarsen.c:2101, par_lr_interp.c:1025, par_csr_matrix.c:2866).
 *      The diagonal of a row is found by column.  Outputs allocated by the
 *      library are freed with amgt_free. ---- */
/* hypre_BoomerAMGCreateSDevice, par_strength_device.c:34 (num_functions 1):
 * S_i[n+1], S_j[nnzS] strong columns in row order, no diagonal. */
int amgt_amg_strength(int32_t n, const int32_t *A_i, const int32_t *A_j, const double *A_data,
                      double theta, double max_row_sum, int64_t *nnzS, int32_t **S_i,
                      int32_t **S_j, void *stream);
/* hypre_BoomerAMGCoarsenPMISDevice, par_coarsen_device.c:30 (CF_init 0 or 2,
 * coarsen types 8/9; CF_init 1 below): cf_marker[n] (caller's device array) = 1 C,
 * -1 F, -3 no strong dependence; measures |S^T row| + hypre_Rand() from seed 2747. */
int amgt_amg_pmis(int32_t n, const int32_t *S_i, const int32_t *S_j, int cf_init,
                  int32_t *cf_marker, int32_t *n_coarse, void *stream);
/* PMIS with CF_init 1 (hypre_BoomerAMGCoarsenPMISHost, par_coarsen.c:2353-2386, 2494)
 * reads cf_marker as the markers of amgt_amg_ruge_first_pass (1 C, -1 F, -2 Z,
 * -3 no strong dependence) and overwrites it with the splitting. */
/* First pass of Ruge-Stueben coarsening as HMIS runs it, on HOST arrays
 * (hypre_BoomerAMGCoarsenRuge, par_coarsen.c:911-1398, coarsen_type 10, measure_type 0,
 * one process): cf_marker[n] = 1 C, -1 F, -2 Z (influences no point), -3 no strong
 * dependence.  The undecided point of largest measure is taken first, the oldest among
 * equals, as HYPRE's list of lists does (utilities/amg_linklist.c). */
int amgt_amg_ruge_first_pass(int32_t n, const int32_t *S_i, const int32_t *S_j,
                             int32_t *cf_marker);
/* hypre_BoomerAMGCoarsenHMIS, par_coarsen.c:2845-2866 (coarsen type 10; HYPRE's default):
 * the first pass on the host (S is copied to pinned host memory and the markers back),
 * then amgt_amg_pmis with CF_init 1 on the device.  Device S and cf_marker[n]; result
 * markers as amgt_amg_pmis.  Synchronises the stream. */
int amgt_amg_hmis(int32_t n, const int32_t *S_i, const int32_t *S_j, int32_t *cf_marker,
                  int32_t *n_coarse, void *stream);
/* hypre_BoomerAMGBuildExtPIInterpDevice, par_lr_interp_device.c:991, with
 * hypre_BoomerAMGInterpTruncationDevice (max_elmts > 0, trunc_factor 0):
 * P (n x n_coarse) in CSR, row entries in HYPRE's host order. */
int amgt_amg_extpi_interp(int32_t n, const int32_t *A_i, const int32_t *A_j,
                          const double *A_data, const int32_t *S_i, const int32_t *S_j,
                          const int32_t *cf_marker, int max_elmts, int32_t *n_coarse,
                          int64_t *nnzP, int32_t **P_i, int32_t **P_j, double **P_data,
                          void *stream);

/* The Galerkin product R A P (R = P^T) in HYPRE's host order,
 * hypre_BoomerAMGBuildCoarseOperatorKT par_rap.c:1525-1930 on one process: R A with
 * first-touch column slots over R's rows (fine rows ascending) and A's rows in their
 * stored order, then (R A) P with the diagonal first and first-touch slots; products and
 * sums rounded one at a time.  The coarse operator (n_coarse x n_coarse, device CSR, rows
 * in HYPRE's layout) equals HYPRE's host one bit for bit, so a hierarchy built with it
 * equals HYPRE's CPU hierarchy level by level.  Device inputs (A n x n, P n x n_coarse,
 * any row order); outputs freed with amgt_free.  Synchronises the stream. */
int amgt_amg_rap_host_order(int32_t n, int32_t n_coarse, const int32_t *A_i, const int32_t *A_j,
                            const double *A_data, const int32_t *P_i, const int32_t *P_j,
                            const double *P_data, int64_t *nnz, int32_t **RAP_i,
                            int32_t **RAP_j, double **RAP_data, void *stream);

/* ---- geometric prolongation (the V-cycle driver's structured hierarchy; no
 *      counterpart in the reference, whose BoomerAMG builds P algebraically,
 *      amgt_amg_extpi_interp).  Outputs freed with amgt_free. ---- */
/* Block-diagonal trilinear P of nprob independent grids (grids: HOST array of
 * nprob (nx, ny, nz) triples, x fastest): coarse points at even indices,
 * weights 1 / (1/2, 1/2) per direction, zero Dirichlet beyond the last coarse
 * point; rows in (z, y, x)-major order with ascending columns, at most 8
 * entries.  Synchronises the stream. */
int amgt_interp_trilinear(int32_t nprob, const int32_t *grids, int32_t *n_fine,
                          int32_t *n_coarse, int64_t *nnzP, int32_t **P_i, int32_t **P_j,
                          double **P_data, void *stream);

#ifdef __cplusplus
}
#endif

#endif
