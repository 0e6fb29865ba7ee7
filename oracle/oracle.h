/*
 * CPU oracle for the mBSR data path (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the checker for the HIP product library, never part of it:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  Every function restates the reference algorithm it names (file:line
 * relative to the reference tree, `src/` = AmgT_HYPRE/src/) in plain C.
 *
 * Pinning of each function (how it is checked before it is trusted) is listed
 * in tests/test_oracle.py; functions without a pin say "parity unpinned".
 *
 * Index type is int32 (HYPRE_Int); nnz counts and offsets that can exceed 2^31
 * in symbolic counting use int64.
 */
#ifndef AMGT_ORACLE_H
#define AMGT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Park-Miller "minimal standard" RNG: src/utilities/random.c:53-101 ---- */
int32_t orc_pm_seed(int64_t seed);                 /* hypre_SeedRand clamp, random.c:53-66 */
int32_t orc_pm_next(int32_t *state);               /* hypre_RandI, random.c:75-91 */
/* hypre_SeqVectorSetRandomValues, src/seq_mv/vector.c:395-418: v = 2*Rand-1 */
void orc_rand_uniform(int64_t seed, int64_t n, double *out);
/* N(0,1) by Box-Muller on the same stream (project-defined transform) */
void orc_rand_normal(int64_t seed, int64_t n, double *out);

/* ---- stencil generators: src/parcsr_ls/par_laplace.c:15 (7pt, 5pt = nz 1),
 *      par_laplace_9pt.c, par_laplace_27pt.c; single process ----
 * kind: 7 (7pt, includes 2D 5pt when nz==1), 9 (2D 9pt), 27 (3D 27pt).
 * Row layout matches HYPRE: diagonal first, then neighbours by ascending
 * column.  cx,cy,cz: 7pt couplings (ij.c:9565-9580); 9pt/27pt use diag
 * 8/26 and off-diagonal -1. */
int64_t orc_stencil_nnz(int kind, int nx, int ny, int nz);
int orc_stencil(int kind, int nx, int ny, int nz, double cx, double cy, double cz,
                int *I, int *J, double *A);

/* ---- 4x4 occupancy bitmap product: csr_spgemm_device.c:189-205 with
 *      Bitmap_col_x_row seq_mv.h:149 ---- */
uint16_t orc_bitmap_mul(uint16_t a, uint16_t b);

/* ---- CSR -> mBSR: csr_matvec_device.c:683-1267 (structure), compressed
 *      values in (block, bit) order ----
 * count: blk_ptr[brow+1] (exclusive prefix), returns nblk, or -1 on a
 * duplicate (row, col) entry, -2 on a column index out of range. */
int64_t orc_csr2mbsr_count(int nrows, int ncols, const int *I, const int *J, int *blk_ptr);
int orc_csr2mbsr_fill(int nrows, int ncols, const int *I, const int *J, const double *A,
                      const int *blk_ptr, int *blk_col, uint16_t *blk_map, int *val_ptr,
                      double *val);
/* the reference's dense 16-slot block values (csr_matvec_device.c:936-945) */
void orc_mbsr_dense_values(int brow, const int *blk_ptr, const uint16_t *blk_map,
                           const int *val_ptr, const double *val, double *dense16);
/* stand / avg_nnz: getStand csr_matvec_device.c:638-671 + 1240-1249 */
void orc_mbsr_stats(int brow, const int *blk_ptr, int64_t nnz, double *stand, double *avg_nnz);

/* ---- mBSR -> CSR: bsr2csr_get_ptr / _get_idx_val csr_spgemm_device.c:1344-1440 ---- */
int64_t orc_mbsr2csr_count(int nrows, int brow, const int *blk_ptr, const uint16_t *blk_map, int *I);
void orc_mbsr2csr_fill(int nrows, int brow, const int *blk_ptr, const int *blk_col,
                       const uint16_t *blk_map, const int *val_ptr, const double *val,
                       const int *I, int *J, double *A);

/* ---- SpMV on mBSR: y = alpha*A*x + beta*b; bsr_spmv_cc_fp64
 *      csr_matvec_device.c:509-555 (bit-gated, block order) ----
 * beta == 0 ignores b. */
void orc_mbsr_spmv(int nrows, int brow, const int *blk_ptr, const int *blk_col,
                   const uint16_t *blk_map, const int *val_ptr, const double *val,
                   double alpha, const double *x, double beta, const double *b, double *y);
/* y = alpha*A^T*x + beta*b on mBSR (the transposed product the reference
 * leaves to a kept transpose, par_csr_matvec_device.c:445-447). */
void orc_mbsr_spmv_t(int nrows, int ncols, int brow, const int *blk_ptr, const int *blk_col,
                     const uint16_t *blk_map, const int *val_ptr, const double *val,
                     double alpha, const double *x, double beta, const double *b, double *y);

/* ---- HYPRE host CSR SpMV: hypre_CSRMatrixMatvecOutOfPlaceHost csr_matvec.c:22
 *      (single vector, no rownnz) and MatvecTHost csr_matvec.c:914 ---- */
void orc_csr_matvec(int nrows, const int *I, const int *J, const double *A, double alpha,
                    const double *x, double beta, const double *b, double *y);
/* orc_csr_matvec with the rows split over nthreads OpenMP threads (same results) */
void orc_csr_matvec_omp(int nthreads, int nrows, const int *I, const int *J, const double *A,
                        double alpha, const double *x, double beta, const double *b, double *y);
void orc_csr_matvec_t(int nrows, int ncols, const int *I, const int *J, const double *A,
                      double alpha, const double *x, double beta, const double *b, double *y);

/* ---- HYPRE host SpGEMM: hypre_CSRMatrixMultiplyHost csr_matop.c:665-900
 *      (one thread; square C gets its diagonal first, then first-touch order) ---- */
int64_t orc_csr_spgemm_count(int m, int n, const int *AI, const int *AJ, const int *BI,
                             const int *BJ, int diag_first, int *CI);
void orc_csr_spgemm_fill(int m, int n, const int *AI, const int *AJ, const double *AA,
                         const int *BI, const int *BJ, const double *BA, int diag_first,
                         const int *CI, int *CJ, double *CA);
/* symbolic row-count only (for the big A^2 nnz fixtures); returns total */
int64_t orc_csr_spgemm_nnz(int m, int n, const int *AI, const int *AJ, const int *BI,
                           const int *BJ, int diag_first);

/* ---- mBSR SpGEMM: symbolic csr_spgemm_device.c:206-752 (blocks whose bitmap
 *      product is non-zero, sorted block columns), numeric 1142-1342
 *      (A-block order, mapC |= product) ---- */
int64_t orc_mbsr_spgemm_count(int brow_a, const int *ap, const int *ac, const uint16_t *am,
                              int bcol_b, const int *bp, const int *bc, const uint16_t *bm,
                              int *cp);
void orc_mbsr_spgemm_fill(int brow_a, const int *ap, const int *ac, const uint16_t *am,
                          const int *avp, const double *av,
                          int bcol_b, const int *bp, const int *bc, const uint16_t *bm,
                          const int *bvp, const double *bv,
                          const int *cp, int *cc, uint16_t *cm, int *cvp, double *cv);

/* ---- BoomerAMG setup on one process (CSR, any column order; the diagonal is
 *      found by column, HYPRE keeps it first) ---- */
/* strength, hypre_BoomerAMGCreateSHost par_strength.c:75-528 (one function):
 * SI[n+1], SJ[nnz(A)] (strong columns in row order, no diagonal); returns nnz(S) */
int64_t orc_amg_strength(int n, const int *AI, const int *AJ, const double *AA, double theta,
                         double max_row_sum, int *SI, int *SJ);
/* PMIS, hypre_BoomerAMGCoarsenPMISHost par_coarsen.c:2101-2813 with measures
 * from hypre_BoomerAMGIndepSetInit par_indepset.c (seed 2747): cf = 1 (C),
 * -1 (F), -3 (no strong dependence); returns the number of C points.
 * cf_init 1 (HMIS) reads cf as the markers of orc_amg_ruge_first_pass. */
int orc_amg_pmis(int n, const int *SI, const int *SJ, int cf_init, int *cf);
/* Ruge-Stueben first pass as HMIS runs it, hypre_BoomerAMGCoarsenRuge
 * par_coarsen.c:911-1398 (coarsen_type 10, measure_type 0, one process):
 * cf = 1 (C), -1 (F), -2 (Z: influences no point), -3 (no strong dependence) */
void orc_amg_ruge_first_pass(int n, const int *SI, const int *SJ, int *cf);
/* HMIS, hypre_BoomerAMGCoarsenHMIS par_coarsen.c:2845-2866: the first pass,
 * then PMIS with CF_init 1; returns the number of C points */
int orc_amg_hmis(int n, const int *SI, const int *SJ, int *cf);
/* measures |S^T row| + hypre_Rand() (seed 2747), for checking */
void orc_amg_pmis_measures(int n, const int *SI, const int *SJ, double *measure);
/* extended+i interpolation, hypre_BoomerAMGBuildExtPIInterpHost
 * par_lr_interp.c:1025-1928 (one process): pass 1 fills PI[n+1] and returns
 * nnz, pass 2 fills PJ/PA (row entries in discovery order) */
int64_t orc_amg_extpi_count(int n, const int *AI, const int *AJ, const int *SI, const int *SJ,
                            const int *cf, int *PI);
void orc_amg_extpi_fill(int n, const int *AI, const int *AJ, const double *AA, const int *SI,
                        const int *SJ, const int *cf, const int *PI, int *PJ, double *PA);
/* Galerkin product P^T A P in the order of HYPRE's host setup,
 * hypre_BoomerAMGBuildCoarseOperatorKT par_rap.c:1525-1930 (one process):
 * (R*A)*P row by row with first-touch slots, the diagonal first.  Pass 1
 * fills RAPI[nc+1] and returns nnz; pass 2 fills RAPJ/RAPA. */
int64_t orc_amg_rap_kt_count(int n, int nc, const int *AI, const int *AJ, const int *PI,
                             const int *PJ, int *RAPI);
void orc_amg_rap_kt_fill(int n, int nc, const int *AI, const int *AJ, const double *AA,
                         const int *PI, const int *PJ, const double *PA, const int *RAPI,
                         int *RAPJ, double *RAPA);
/* keep the max_elmts largest |p_ij| per row and rescale to the row sum,
 * hypre_ParCSRMatrixTruncate par_csr_matrix.c:2866-3260 (tol 0, rescale, the
 * row's order from hypre_qsort2_abs qsort.c:395-417); in place, returns nnz */
int64_t orc_amg_truncate(int n, int *PI, int *PJ, double *PA, int max_elmts);

/* ---- per-level policy network (no counterpart in the reference: it replaces
 *      the fixed rule of get_calculationPrecison, csr_spgemm_device.c:1847-1862).
 *      12 -> 64 tanh -> 64 tanh -> 4, parameters W1[64][12] b1 W2[64][64] b2
 *      W3[4][64] b3 (5252 values), in double. ---- */
/* out[n][4] */
void orc_policy_forward(const double *params, const double *x, int64_t n, double *out);
/* returns the mean over the samples of -log softmax(o[0..2])[cls] +
 * lambda (o[3] - target)^2 and writes its gradient to grad[5252] */
double orc_policy_loss_grad(const double *params, const double *x, const int *cls,
                            const double *target, int64_t n, double lambda, double *grad);
/* one bias-corrected Adam step (step >= 1), in place */
void orc_policy_adam(double *params, const double *grad, double *m, double *v, int64_t np,
                     int64_t step, double lr, double b1, double b2, double eps);

#ifdef __cplusplus
}
#endif

#endif
