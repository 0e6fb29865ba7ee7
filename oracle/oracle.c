/*
 * CPU oracle for the mBSR data path -- TEST INFRASTRUCTURE ONLY.
 *
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, as the checker.  The product library (amgt_amd/libamgt_mbsr.so) never
 * links or calls it.  Each function restates the reference algorithm whose
 * file:line it cites (relative to the reference tree; `src/` = AmgT_HYPRE/src/).
 * How each function is pinned is listed in tests/test_oracle.py.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Park-Miller minimal standard generator (Schrage decomposition).
 * src/utilities/random.c:43-101. */
#define PM_A 16807
#define PM_M 2147483647
#define PM_Q 127773
#define PM_R 2836

int32_t orc_pm_seed(int64_t seed)
{
    /* random.c:56-63: seed < 1 -> 1, seed >= m -> m-1 */
    if (seed < 1) return 1;
    if (seed >= PM_M) return PM_M - 1;
    return (int32_t)seed;
}

int32_t orc_pm_next(int32_t *state)
{
    int32_t s = *state;
    int32_t high = s / PM_Q;
    int32_t low = s % PM_Q;
    int32_t test = PM_A * low - PM_R * high;
    *state = (test > 0) ? test : test + PM_M;
    return *state;
}

static double pm_rand(int32_t *state) /* hypre_Rand, random.c:99-101: (0, 1] */
{
    return (double)orc_pm_next(state) / (double)PM_M;
}

void orc_rand_uniform(int64_t seed, int64_t n, double *out)
{
    /* vector.c:406-413 */
    int32_t st = orc_pm_seed(seed);
    for (int64_t i = 0; i < n; i++) out[i] = 2.0 * pm_rand(&st) - 1.0;
}

void orc_rand_normal(int64_t seed, int64_t n, double *out)
{
    /* Box-Muller on consecutive pairs of the same stream. u1 in (0,1] so the
     * log is finite. */
    int32_t st = orc_pm_seed(seed);
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < n; i += 2) {
        double u1 = pm_rand(&st);
        double u2 = pm_rand(&st);
        double rad = sqrt(-2.0 * log(u1));
        out[i] = rad * cos(two_pi * u2);
        if (i + 1 < n) out[i + 1] = rad * sin(two_pi * u2);
    }
}

/* ------------------------------------------------------------------------ */
/* Stencils.  Single-process restatement of GenerateLaplacian
 * (src/parcsr_ls/par_laplace.c:202-300: diagonal first, then -z,-y,-x,+x,+y,+z),
 * GenerateLaplacian9pt (par_laplace_9pt.c:243-360) and GenerateLaplacian27pt
 * (par_laplace_27pt.c:778-1300): diagonal first, neighbours in ascending
 * (dz, dy, dx) order, which is ascending column order. */
static int stencil_offsets(int kind, int *dx, int *dy, int *dz)
{
    int k = 0;
    if (kind == 7) {
        const int o[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int t = 0; t < 6; t++) { dx[k] = o[t][0]; dy[k] = o[t][1]; dz[k] = o[t][2]; k++; }
    } else if (kind == 9) {
        for (int j = -1; j <= 1; j++)
            for (int i = -1; i <= 1; i++)
                if (i || j) { dx[k] = i; dy[k] = j; dz[k] = 0; k++; }
    } else if (kind == 27) {
        for (int l = -1; l <= 1; l++)
            for (int j = -1; j <= 1; j++)
                for (int i = -1; i <= 1; i++)
                    if (i || j || l) { dx[k] = i; dy[k] = j; dz[k] = l; k++; }
    } else {
        return -1;
    }
    return k;
}

int64_t orc_stencil_nnz(int kind, int nx, int ny, int nz)
{
    int dx[26], dy[26], dz[26];
    int k = stencil_offsets(kind, dx, dy, dz);
    if (k < 0) return -1;
    int64_t nnz = 0;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                nnz++;
                for (int t = 0; t < k; t++) {
                    int xx = x + dx[t], yy = y + dy[t], zz = z + dz[t];
                    if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz) nnz++;
                }
            }
    return nnz;
}

int orc_stencil(int kind, int nx, int ny, int nz, double cx, double cy, double cz,
                int *I, int *J, double *A)
{
    int dx[26], dy[26], dz[26];
    int k = stencil_offsets(kind, dx, dy, dz);
    if (k < 0) return -1;
    double diag, off[26];
    if (kind == 7) {
        /* ij.c:9565-9580: diagonal = sum of 2c over dimensions longer than 1 */
        diag = 0.0;
        if (nx > 1) diag += 2.0 * cx;
        if (ny > 1) diag += 2.0 * cy;
        if (nz > 1) diag += 2.0 * cz;
        for (int t = 0; t < k; t++) off[t] = dx[t] ? -cx : (dy[t] ? -cy : -cz);
    } else {
        diag = (kind == 9) ? 8.0 : 26.0;
        for (int t = 0; t < k; t++) off[t] = -1.0;
    }
    int64_t cnt = 0;
    int row = 0;
    const int64_t nxy = (int64_t)nx * ny;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                I[row] = (int)cnt;
                J[cnt] = row;
                A[cnt++] = diag;
                for (int t = 0; t < k; t++) {
                    int xx = x + dx[t], yy = y + dy[t], zz = z + dz[t];
                    if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz) {
                        J[cnt] = (int)(zz * nxy + (int64_t)yy * nx + xx);
                        A[cnt++] = off[t];
                    }
                }
                row++;
            }
    I[row] = (int)cnt;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* bitMatrixMul, csr_spgemm_device.c:189-205: OR over k of
 * Bitmap_col_x_row(col_k(A), row_k(B)) = outer product of column k of A
 * (bits 4r+k) with row k of B (bits 4k+c). */
uint16_t orc_bitmap_mul(uint16_t a, uint16_t b)
{
    uint32_t res = 0;
    uint32_t ca = a, rb = b;
    for (int k = 0; k < 4; k++) {
        uint32_t col = ca & 0x1111u;      /* A[r][k] at bit 4r */
        uint32_t row = rb & 0xFu;         /* B[k][c] at bit c  */
        res |= (col * 0xFu) & (row * 0x1111u);
        ca >>= 1;
        rb >>= 4;
    }
    return (uint16_t)res;
}

/* ------------------------------------------------------------------------ */
/* CSR -> mBSR.  Structure per csr_matvec_device.c:683-1092: a block (br,bc)
 * exists iff an entry (r,c) of rows 4br..4br+3 has c/4 == bc; block columns
 * ascending; bit (r%4)*4 + c%4.  Values are stored compressed: the nonzeros of
 * a block row in (block column, bit) order, starting at I[4br]. */
typedef struct { uint32_t bc; uint32_t bit; double v; } ent_t;

static int ent_cmp(const void *p, const void *q)
{
    const ent_t *a = (const ent_t *)p, *b = (const ent_t *)q;
    if (a->bc != b->bc) return a->bc < b->bc ? -1 : 1;
    if (a->bit != b->bit) return a->bit < b->bit ? -1 : 1;
    return 0;
}

/* gathers and sorts the entries of block row br; returns the entry count or
 * -1 (duplicate) / -2 (column out of range) */
static int64_t gather_block_row(int nrows, int ncols, const int *I, const int *J,
                                const double *A, int br, ent_t **buf, int64_t *cap)
{
    int r0 = 4 * br;
    int r1 = r0 + 4 < nrows ? r0 + 4 : nrows;
    int64_t n = (int64_t)I[r1] - I[r0];
    if (n > *cap) {
        *cap = n * 2 + 16;
        *buf = (ent_t *)realloc(*buf, sizeof(ent_t) * (size_t)*cap);
    }
    int64_t t = 0;
    for (int r = r0; r < r1; r++)
        for (int p = I[r]; p < I[r + 1]; p++) {
            int c = J[p];
            if (c < 0 || c >= ncols) return -2;
            (*buf)[t].bc = (uint32_t)(c >> 2);
            (*buf)[t].bit = (uint32_t)(((r & 3) << 2) | (c & 3));
            (*buf)[t].v = A ? A[p] : 0.0;
            t++;
        }
    qsort(*buf, (size_t)n, sizeof(ent_t), ent_cmp);
    for (int64_t q = 1; q < n; q++)
        if ((*buf)[q].bc == (*buf)[q - 1].bc && (*buf)[q].bit == (*buf)[q - 1].bit) return -1;
    return n;
}

int64_t orc_csr2mbsr_count(int nrows, int ncols, const int *I, const int *J, int *blk_ptr)
{
    int brow = (nrows + 3) / 4;
    ent_t *buf = NULL;
    int64_t cap = 0, total = 0;
    blk_ptr[0] = 0;
    for (int br = 0; br < brow; br++) {
        int64_t n = gather_block_row(nrows, ncols, I, J, NULL, br, &buf, &cap);
        if (n < 0) { free(buf); return n; }
        int64_t cnt = 0;
        for (int64_t q = 0; q < n; q++)
            if (q == 0 || buf[q].bc != buf[q - 1].bc) cnt++;
        total += cnt;
        blk_ptr[br + 1] = (int)total;
    }
    free(buf);
    return total;
}

int orc_csr2mbsr_fill(int nrows, int ncols, const int *I, const int *J, const double *A,
                      const int *blk_ptr, int *blk_col, uint16_t *blk_map, int *val_ptr,
                      double *val)
{
    int brow = (nrows + 3) / 4;
    ent_t *buf = NULL;
    int64_t cap = 0;
    for (int br = 0; br < brow; br++) {
        int64_t n = gather_block_row(nrows, ncols, I, J, A, br, &buf, &cap);
        if (n < 0) { free(buf); return (int)n; }
        int64_t base = I[4 * br];
        val_ptr[br] = (int)base;
        int b = blk_ptr[br] - 1;
        for (int64_t q = 0; q < n; q++) {
            if (q == 0 || buf[q].bc != buf[q - 1].bc) {
                b++;
                blk_col[b] = (int)buf[q].bc;
                blk_map[b] = 0;
            }
            blk_map[b] |= (uint16_t)(1u << buf[q].bit);
            val[base + q] = buf[q].v;
        }
    }
    val_ptr[brow] = I[nrows];
    free(buf);
    return 0;
}

void orc_mbsr_dense_values(int brow, const int *blk_ptr, const uint16_t *blk_map,
                           const int *val_ptr, const double *val, double *dense16)
{
    for (int br = 0; br < brow; br++) {
        int64_t o = val_ptr[br];
        for (int b = blk_ptr[br]; b < blk_ptr[br + 1]; b++) {
            for (int s = 0; s < 16; s++)
                dense16[(int64_t)b * 16 + s] = ((blk_map[b] >> s) & 1) ? val[o++] : 0.0;
        }
    }
}

void orc_mbsr_stats(int brow, const int *blk_ptr, int64_t nnz, double *stand, double *avg_nnz)
{
    /* csr_matvec_device.c:1239-1249: avg_len = nblk/brow, stand = sqrt of the
     * mean squared deviation of blocks per block row; avg_nnz = nnz/nblk. */
    int64_t nblk = blk_ptr[brow];
    double avg = brow ? (double)nblk / (double)brow : 0.0;
    double s = 0.0;
    for (int br = 0; br < brow; br++) {
        double d = (double)(blk_ptr[br + 1] - blk_ptr[br]) - avg;
        s += d * d;
    }
    *stand = brow ? sqrt(s / brow) : 0.0;
    *avg_nnz = nblk ? (double)nnz / (double)nblk : 0.0;
}

/* ------------------------------------------------------------------------ */
/* mBSR -> CSR, columns ascending within a row (bsr2csr_get_ptr and
 * bsr2csr_get_idx_val, csr_spgemm_device.c:1344-1440). */
int64_t orc_mbsr2csr_count(int nrows, int brow, const int *blk_ptr, const uint16_t *blk_map, int *I)
{
    int64_t total = 0;
    I[0] = 0;
    for (int r = 0; r < nrows; r++) {
        int br = r >> 2, sh = (r & 3) * 4;
        for (int b = blk_ptr[br]; b < blk_ptr[br + 1]; b++)
            total += __builtin_popcount((blk_map[b] >> sh) & 0xF);
        I[r + 1] = (int)total;
    }
    (void)brow;
    return total;
}

void orc_mbsr2csr_fill(int nrows, int brow, const int *blk_ptr, const int *blk_col,
                       const uint16_t *blk_map, const int *val_ptr, const double *val,
                       const int *I, int *J, double *A)
{
    for (int r = 0; r < nrows; r++) {
        int br = r >> 2, rr = r & 3;
        int64_t o = val_ptr[br];
        int64_t w = I[r];
        for (int b = blk_ptr[br]; b < blk_ptr[br + 1]; b++) {
            uint32_t m = blk_map[b];
            int64_t ob = o + __builtin_popcount(m & ((1u << (4 * rr)) - 1u));
            uint32_t nib = (m >> (4 * rr)) & 0xF;
            for (int c = 0; c < 4; c++)
                if ((nib >> c) & 1) {
                    J[w] = blk_col[b] * 4 + c;
                    A[w] = val[ob++];
                    w++;
                }
            o += __builtin_popcount(m);
        }
    }
    (void)brow;
}

/* ------------------------------------------------------------------------ */
/* mBSR SpMV, bsr_spmv_cc_fp64 semantics (csr_matvec_device.c:509-555) with
 * beta applied as y = alpha*acc + beta*b (beta_vecY 673-679 fused). */
void orc_mbsr_spmv(int nrows, int brow, const int *blk_ptr, const int *blk_col,
                   const uint16_t *blk_map, const int *val_ptr, const double *val,
                   double alpha, const double *x, double beta, const double *b, double *y)
{
    for (int br = 0; br < brow; br++) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        int64_t o = val_ptr[br];
        for (int k = blk_ptr[br]; k < blk_ptr[br + 1]; k++) {
            uint32_t m = blk_map[k];
            const double *xs = x + (int64_t)blk_col[k] * 4;
            for (int s = 0; s < 16; s++)
                if ((m >> s) & 1) acc[s >> 2] += val[o++] * xs[s & 3];
        }
        for (int r = 0; r < 4; r++) {
            int row = 4 * br + r;
            if (row >= nrows) break;
            y[row] = (beta == 0.0) ? alpha * acc[r] : alpha * acc[r] + beta * b[row];
        }
    }
}

void orc_mbsr_spmv_t(int nrows, int ncols, int brow, const int *blk_ptr, const int *blk_col,
                     const uint16_t *blk_map, const int *val_ptr, const double *val,
                     double alpha, const double *x, double beta, const double *b, double *y)
{
    double *acc = (double *)calloc((size_t)ncols + 4, sizeof(double));
    for (int br = 0; br < brow; br++) {
        int64_t o = val_ptr[br];
        for (int k = blk_ptr[br]; k < blk_ptr[br + 1]; k++) {
            uint32_t m = blk_map[k];
            int64_t c0 = (int64_t)blk_col[k] * 4;
            for (int s = 0; s < 16; s++)
                if ((m >> s) & 1) acc[c0 + (s & 3)] += val[o++] * x[4 * br + (s >> 2)];
        }
    }
    for (int c = 0; c < ncols; c++) y[c] = (beta == 0.0) ? alpha * acc[c] : alpha * acc[c] + beta * b[c];
    free(acc);
    (void)nrows;
}

/* ------------------------------------------------------------------------ */
/* hypre_CSRMatrixMatvecOutOfPlaceHost, csr_matvec.c:22-700, single vector
 * without rownnz: temp = beta/alpha; alpha == 0 -> y = beta*b; temp == 0 ->
 * y = alpha*(A x); otherwise y = beta*b + alpha*(A x) with the row sum in CSR
 * order starting from 0. */
void orc_csr_matvec(int nrows, const int *I, const int *J, const double *A, double alpha,
                    const double *x, double beta, const double *b, double *y)
{
    if (alpha == 0.0) {
        for (int i = 0; i < nrows; i++) y[i] = beta * b[i];
        return;
    }
    double temp = beta / alpha;
    for (int i = 0; i < nrows; i++) {
        double t = 0.0;
        for (int p = I[i]; p < I[i + 1]; p++) t += A[p] * x[J[p]];
        if (temp == 0.0) {
            y[i] = (alpha == 1.0) ? t : ((alpha == -1.0) ? -t : alpha * t);
        } else if (alpha == 1.0) {
            y[i] = beta * b[i] + t;
        } else if (alpha == -1.0) {
            y[i] = -temp * b[i] - t;
        } else {
            y[i] = beta * b[i] + alpha * t;
        }
    }
}

/* The same product with the rows split over nthreads OpenMP threads in
 * contiguous static chunks, as HYPRE's OpenMP build of
 * hypre_CSRMatrixMatvecOutOfPlaceHost splits them (csr_matvec.c:22-700 under
 * HYPRE_USING_OPENMP); every row is summed as in orc_csr_matvec, so the result
 * is the same bit for bit.  The CPU baseline of the SpMV timing. */
void orc_csr_matvec_omp(int nthreads, int nrows, const int *I, const int *J, const double *A,
                        double alpha, const double *x, double beta, const double *b, double *y)
{
    const double temp = alpha == 0.0 ? 0.0 : beta / alpha;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int i = 0; i < nrows; i++) {
        if (alpha == 0.0) {
            y[i] = beta * b[i];
            continue;
        }
        double t = 0.0;
        for (int p = I[i]; p < I[i + 1]; p++) t += A[p] * x[J[p]];
        if (temp == 0.0) {
            y[i] = (alpha == 1.0) ? t : ((alpha == -1.0) ? -t : alpha * t);
        } else if (alpha == 1.0) {
            y[i] = beta * b[i] + t;
        } else if (alpha == -1.0) {
            y[i] = -temp * b[i] - t;
        } else {
            y[i] = beta * b[i] + alpha * t;
        }
    }
}

/* hypre_CSRMatrixMatvecTHost, csr_matvec.c:914: y = alpha*A^T*x + beta*b,
 * scattered in row order. */
void orc_csr_matvec_t(int nrows, int ncols, const int *I, const int *J, const double *A,
                      double alpha, const double *x, double beta, const double *b, double *y)
{
    double *acc = (double *)calloc((size_t)ncols + 1, sizeof(double));
    for (int i = 0; i < nrows; i++)
        for (int p = I[i]; p < I[i + 1]; p++) acc[J[p]] += A[p] * x[i];
    for (int c = 0; c < ncols; c++) y[c] = (beta == 0.0) ? alpha * acc[c] : alpha * acc[c] + beta * b[c];
    free(acc);
}

/* ------------------------------------------------------------------------ */
/* hypre_CSRMatrixMultiplyHost, csr_matop.c:665-900 with one thread and no
 * rownnz: first pass counts with a marker array (the diagonal is counted first
 * when C is square), second pass fills in first-touch order and accumulates
 * a_ik*b_kj in A then B order. */
int64_t orc_csr_spgemm_count(int m, int n, const int *AI, const int *AJ, const int *BI,
                             const int *BJ, int diag_first, int *CI)
{
    int *marker = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; j++) marker[j] = -1;
    int64_t nnz = 0;
    for (int i = 0; i < m; i++) {
        CI[i] = (int)nnz;
        if (diag_first && i < n) { marker[i] = i; nnz++; }
        for (int p = AI[i]; p < AI[i + 1]; p++) {
            int k = AJ[p];
            for (int q = BI[k]; q < BI[k + 1]; q++) {
                int j = BJ[q];
                if (marker[j] != i) { marker[j] = i; nnz++; }
            }
        }
    }
    CI[m] = (int)nnz;
    free(marker);
    return nnz;
}

int64_t orc_csr_spgemm_nnz(int m, int n, const int *AI, const int *AJ, const int *BI,
                           const int *BJ, int diag_first)
{
    int *marker = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; j++) marker[j] = -1;
    int64_t nnz = 0;
    for (int i = 0; i < m; i++) {
        if (diag_first && i < n) { marker[i] = i; nnz++; }
        for (int p = AI[i]; p < AI[i + 1]; p++) {
            int k = AJ[p];
            for (int q = BI[k]; q < BI[k + 1]; q++) {
                int j = BJ[q];
                if (marker[j] != i) { marker[j] = i; nnz++; }
            }
        }
    }
    free(marker);
    return nnz;
}

void orc_csr_spgemm_fill(int m, int n, const int *AI, const int *AJ, const double *AA,
                         const int *BI, const int *BJ, const double *BA, int diag_first,
                         const int *CI, int *CJ, double *CA)
{
    int *marker = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; j++) marker[j] = -1;
    int64_t counter = 0;
    for (int i = 0; i < m; i++) {
        int64_t row_start = counter;
        if (diag_first && i < n) {
            marker[i] = (int)counter;
            CJ[counter] = i;
            CA[counter] = 0.0;
            counter++;
        }
        for (int p = AI[i]; p < AI[i + 1]; p++) {
            int k = AJ[p];
            double a = AA[p];
            for (int q = BI[k]; q < BI[k + 1]; q++) {
                int j = BJ[q];
                double bv = BA[q];
                if (marker[j] < row_start) {
                    marker[j] = (int)counter;
                    CJ[counter] = j;
                    CA[counter] = a * bv;
                    counter++;
                } else {
                    CA[marker[j]] += a * bv;
                }
            }
        }
        (void)CI;
    }
    free(marker);
}

/* ------------------------------------------------------------------------ */
/* mBSR SpGEMM.  Symbolic (csr_spgemm_device.c:206-752): C block (i,j) exists
 * iff some A block (i,k) and B block (k,j) have a non-zero bitmap product;
 * block columns ascending; mapC = OR of the products.  Numeric (1142-1342):
 * per A block in order, per B block in order with a non-zero product,
 * C[r][c] += sum over t ascending of A[r][t]*B[t][c] (the 4-term partial sum is
 * formed first, then added, as the scalar path's temp_valC does). */
static int cmp_int(const void *p, const void *q)
{
    int a = *(const int *)p, b = *(const int *)q;
    return (a > b) - (a < b);
}

int64_t orc_mbsr_spgemm_count(int brow_a, const int *ap, const int *ac, const uint16_t *am,
                              int bcol_b, const int *bp, const int *bc, const uint16_t *bm,
                              int *cp)
{
    int *marker = (int *)malloc(sizeof(int) * (size_t)(bcol_b > 0 ? bcol_b : 1));
    for (int j = 0; j < bcol_b; j++) marker[j] = -1;
    int64_t total = 0;
    cp[0] = 0;
    for (int i = 0; i < brow_a; i++) {
        for (int p = ap[i]; p < ap[i + 1]; p++) {
            int k = ac[p];
            for (int q = bp[k]; q < bp[k + 1]; q++) {
                if (!orc_bitmap_mul(am[p], bm[q])) continue;
                int j = bc[q];
                if (marker[j] != i) { marker[j] = i; total++; }
            }
        }
        cp[i + 1] = (int)total;
    }
    free(marker);
    return total;
}

void orc_mbsr_spgemm_fill(int brow_a, const int *ap, const int *ac, const uint16_t *am,
                          const int *avp, const double *av,
                          int bcol_b, const int *bp, const int *bc, const uint16_t *bm,
                          const int *bvp, const double *bv,
                          const int *cp, int *cc, uint16_t *cm, int *cvp, double *cv)
{
    int *slot = (int *)malloc(sizeof(int) * (size_t)(bcol_b > 0 ? bcol_b : 1));
    for (int j = 0; j < bcol_b; j++) slot[j] = -1;
    /* per-block value offsets of A and B (prefix of popcount within rows) */
    int64_t nblk_a = ap[brow_a];
    int64_t *aoff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nblk_a + 1));
    for (int i = 0; i < brow_a; i++) {
        int64_t o = avp[i];
        for (int p = ap[i]; p < ap[i + 1]; p++) { aoff[p] = o; o += __builtin_popcount(am[p]); }
    }
    /* B's block-row count is A's block-column count; derive it from the largest
     * block column A references */
    int brow_b = 0;
    for (int64_t p = 0; p < nblk_a; p++) if (ac[p] + 1 > brow_b) brow_b = ac[p] + 1;
    int64_t nblk_b = brow_b ? bp[brow_b] : 0;
    int64_t *boff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nblk_b + 1));
    for (int k = 0; k < brow_b; k++) {
        int64_t o = bvp[k];
        for (int q = bp[k]; q < bp[k + 1]; q++) { boff[q] = o; o += __builtin_popcount(bm[q]); }
    }
    int64_t vo = 0;
    double *dense = NULL;
    int64_t dense_cap = 0;
    for (int i = 0; i < brow_a; i++) {
        int64_t c0 = cp[i], c1 = cp[i + 1];
        int64_t cnt = 0;
        /* symbolic: collect block columns, then sort */
        for (int p = ap[i]; p < ap[i + 1]; p++) {
            int k = ac[p];
            for (int q = bp[k]; q < bp[k + 1]; q++) {
                if (!orc_bitmap_mul(am[p], bm[q])) continue;
                int j = bc[q];
                if (slot[j] < 0) { slot[j] = 1; cc[c0 + cnt++] = j; }
            }
        }
        qsort(cc + c0, (size_t)cnt, sizeof(int), cmp_int);
        for (int64_t t = 0; t < cnt; t++) { slot[cc[c0 + t]] = (int)t; cm[c0 + t] = 0; }
        if (cnt > dense_cap) {
            dense_cap = cnt * 2;
            dense = (double *)realloc(dense, sizeof(double) * 16 * (size_t)dense_cap);
        }
        memset(dense, 0, sizeof(double) * 16 * (size_t)cnt);
        /* numeric */
        for (int p = ap[i]; p < ap[i + 1]; p++) {
            int k = ac[p];
            uint32_t ma = am[p];
            double a[16];
            int64_t o = aoff[p];
            for (int s = 0; s < 16; s++) a[s] = ((ma >> s) & 1) ? av[o++] : 0.0;
            for (int q = bp[k]; q < bp[k + 1]; q++) {
                uint16_t mc = orc_bitmap_mul((uint16_t)ma, bm[q]);
                if (!mc) continue;
                uint32_t mb = bm[q];
                double b[16];
                int64_t ob = boff[q];
                for (int s = 0; s < 16; s++) b[s] = ((mb >> s) & 1) ? bv[ob++] : 0.0;
                int t = slot[bc[q]];
                cm[c0 + t] |= mc;
                double *d = dense + 16 * (int64_t)t;
                for (int r = 0; r < 4; r++)
                    for (int c = 0; c < 4; c++) {
                        if (!((mc >> (4 * r + c)) & 1)) continue;
                        double tmp = 0.0;
                        for (int kk = 0; kk < 4; kk++)
                            if (((ma >> (4 * r + kk)) & 1) && ((mb >> (4 * kk + c)) & 1))
                                tmp += a[4 * r + kk] * b[4 * kk + c];
                        d[4 * r + c] += tmp;
                    }
            }
        }
        /* compress */
        cvp[i] = (int)vo;
        for (int64_t t = 0; t < cnt; t++) {
            uint32_t m = cm[c0 + t];
            for (int s = 0; s < 16; s++)
                if ((m >> s) & 1) cv[vo++] = dense[16 * t + s];
        }
        for (int64_t t = 0; t < cnt; t++) slot[cc[c0 + t]] = -1;
        (void)c1;
    }
    cvp[brow_a] = (int)vo;
    free(dense);
    free(slot);
    free(aoff);
    free(boff);
}

/* ------------------------------------------------------------------------ */
/* BoomerAMG setup on one process.  HYPRE keeps the diagonal first in every
 * row of a ParCSR diag block and reads it as A_diag_i[i]; here it is found by
 * column, so rows in any column order work.  Off-diagonal entries are visited
 * in row order, as HYPRE visits A_diag_i[i]+1 .. A_diag_i[i+1]-1. */
static int64_t amg_diag_pos(const int *AI, const int *AJ, int i)
{
    for (int64_t p = AI[i]; p < AI[i + 1]; p++)
        if (AJ[p] == i) return p;
    return -1;
}

/* par_strength.c:75-528, num_functions = 1, one process */
int64_t orc_amg_strength(int n, const int *AI, const int *AJ, const double *AA, double theta,
                         double max_row_sum, int *SI, int *SJ)
{
    int64_t cnt = 0;
    for (int i = 0; i < n; i++) {
        SI[i] = (int)cnt;
        const int64_t d = amg_diag_pos(AI, AJ, i);
        const double diag = d >= 0 ? AA[d] : 0.0;
        double row_scale = 0.0, row_sum = diag;
        for (int64_t p = AI[i]; p < AI[i + 1]; p++) {
            if (p == d) continue;
            const double a = AA[p];
            if (diag < 0) row_scale = a > row_scale ? a : row_scale;
            else row_scale = a < row_scale ? a : row_scale;
            row_sum += a;
        }
        /* par_strength.c:325-339: a row that is too diagonally dominant has
         * only weak dependencies */
        if (fabs(row_sum) > fabs(diag) * max_row_sum && max_row_sum < 1.0) continue;
        for (int64_t p = AI[i]; p < AI[i + 1]; p++) {
            if (p == d) continue;
            const double a = AA[p];
            const int weak = diag < 0 ? (a <= theta * row_scale) : (a >= theta * row_scale);
            if (!weak) SJ[cnt++] = AJ[p];
        }
    }
    SI[n] = (int)cnt;
    return cnt;
}

/* par_indepset.c (hypre_BoomerAMGIndepSetInit): seed 2747 (one process, and
 * the sequential stream skips FirstRowIndex = 0 draws), one draw per row */
void orc_amg_pmis_measures(int n, const int *SI, const int *SJ, double *measure)
{
    for (int i = 0; i < n; i++) measure[i] = 0.0;
    for (int64_t p = 0; p < SI[n]; p++) measure[SJ[p]] += 1.0;
    int32_t st = orc_pm_seed(2747);
    for (int i = 0; i < n; i++) measure[i] += pm_rand(&st);
}

/* par_coarsen.c:2101-2813 on one process */
int orc_amg_pmis(int n, const int *SI, const int *SJ, int cf_init, int *cf)
{
    double *measure = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int *graph = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *graph2 = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    orc_amg_pmis_measures(n, SI, SJ, measure);
    int gsize = 0;
    if (cf_init == 1) {
        /* :2353-2386, markers of the Ruge first pass (HMIS): F points are
         * undecided again, C points stay C, a Z point (no influence) goes back
         * to the graph if its measure reaches 1 or it depends on some point,
         * else it is F; special F points stay out of the graph */
        for (int i = 0; i < n; i++) {
            if (cf[i] != -3) {
                if (cf[i] == -1) cf[i] = 0;
                if (cf[i] == -2) {
                    if (measure[i] >= 1.0 || SI[i + 1] - SI[i] > 0) {
                        cf[i] = 0;
                        graph[gsize++] = i;
                    } else {
                        cf[i] = -1;
                    }
                } else {
                    graph[gsize++] = i;
                }
            } else {
                measure[i] = 0.0;
            }
        }
    } else {
        for (int i = 0; i < n; i++) {
            cf[i] = 0;
            if (SI[i + 1] - SI[i] == 0) { /* :2385-2396, no strong dependence */
                cf[i] = (cf_init == 3 || cf_init == 4) ? 1 : -3;
                measure[i] = 0.0;
            } else {
                graph[gsize++] = i;
            }
        }
    }
    int iter = 0;
    while (gsize > 0) {
        if (!cf_init || iter) { /* :2488-2573 independent set */
            for (int g = 0; g < gsize; g++) {
                const int i = graph[g];
                if (measure[i] > 1) cf[i] = 1;
            }
            for (int g = 0; g < gsize; g++) {
                const int i = graph[g];
                if (!(measure[i] > 1)) continue;
                for (int p = SI[i]; p < SI[i + 1]; p++) {
                    const int j = SJ[p];
                    if (measure[j] > 1) {
                        if (measure[i] > measure[j]) cf[j] = 0;
                        else if (measure[j] > measure[i]) cf[i] = 0;
                    }
                }
            }
        }
        iter++;
        for (int g = 0; g < gsize; g++) { /* :2583-2666 C and F points */
            const int i = graph[g];
            if (measure[i] < 1) cf[i] = -1;
            if (cf[i] > 0) {
                cf[i] = 1;
            } else {
                for (int p = SI[i]; p < SI[i + 1]; p++)
                    if (cf[SJ[p]] > 0) cf[i] = -1;
            }
        }
        int ng = 0; /* :2690-2760 subgraph update */
        for (int g = 0; g < gsize; g++) {
            const int i = graph[g];
            if (cf[i] != 0) measure[i] = 0.0;
            else graph2[ng++] = i;
        }
        int *t = graph;
        graph = graph2;
        graph2 = t;
        gsize = ng;
    }
    int nc = 0;
    for (int i = 0; i < n; i++) nc += cf[i] >= 0;
    free(measure);
    free(graph);
    free(graph2);
    return nc;
}

/* Points kept in lists by measure, with the behaviour of hypre_LinkList
 * (utilities/amg_linklist.c:40-260): the lists are chained in descending
 * measure, a point entering a list goes to its tail, a list that loses its
 * last point is dropped.  next[p] / prev[p] link the points of one list
 * (-2 ends a list, -1 starts it). */
typedef struct amg_list {
    int measure, head, tail;
    struct amg_list *up, *down; /* larger / smaller measure */
} amg_list;

typedef struct {
    amg_list *top, *bottom;
    int *next, *prev;
} amg_lists;

static void amg_lists_remove(amg_lists *L, int measure, int p)
{
    for (amg_list *l = L->top; l; l = l->down) {
        if (l->measure != measure) continue;
        if (l->head == p && l->tail == p) {
            if (l->up) l->up->down = l->down; else L->top = l->down;
            if (l->down) l->down->up = l->up; else L->bottom = l->up;
            free(l);
        } else if (l->head == p) {
            l->head = L->next[p];
            L->prev[L->next[p]] = -1;
        } else if (l->tail == p) {
            l->tail = L->prev[p];
            L->next[L->prev[p]] = -2;
        } else {
            L->next[L->prev[p]] = L->next[p];
            L->prev[L->next[p]] = L->prev[p];
        }
        return;
    }
}

static amg_list *amg_list_new(int measure, int p, amg_lists *L)
{
    amg_list *l = (amg_list *)calloc(1, sizeof(amg_list));
    l->measure = measure;
    l->head = l->tail = p;
    L->next[p] = -2;
    L->prev[p] = -1;
    return l;
}

static void amg_lists_enter(amg_lists *L, int measure, int p)
{
    for (amg_list *l = L->top; l; l = l->down) {
        if (measure > l->measure) { /* a new list above l */
            amg_list *e = amg_list_new(measure, p, L);
            e->up = l->up;
            e->down = l;
            if (l->up) l->up->down = e; else L->top = e;
            l->up = e;
            return;
        }
        if (measure == l->measure) { /* to the tail of l */
            L->next[l->tail] = p;
            L->prev[p] = l->tail;
            L->next[p] = -2;
            l->tail = p;
            return;
        }
    }
    amg_list *e = amg_list_new(measure, p, L); /* below every list */
    e->up = L->bottom;
    if (L->bottom) L->bottom->down = e; else L->top = e;
    L->bottom = e;
}

/* The first pass of Ruge-Stueben coarsening as HMIS runs it:
 * hypre_BoomerAMGCoarsenRuge par_coarsen.c:911-1398 with coarsen_type 10
 * (:1119-1123: points that influence nobody become Z points, -2, and the
 * routine returns after the first pass, :1391-1398), measure_type 0,
 * cut_factor 0, one process, a fresh marker array.  cf[i] = 1 (C), -1 (F),
 * -2 (Z), -3 (no strong dependence). */
void orc_amg_ruge_first_pass(int n, const int *SI, const int *SJ, int *cf)
{
    const size_t nn = (size_t)(n > 0 ? n : 1);
    int *STI = (int *)calloc(nn + 1, sizeof(int));
    int *STJ = (int *)malloc(sizeof(int) * (size_t)(SI[n] > 0 ? SI[n] : 1));
    int *meas = (int *)malloc(sizeof(int) * nn);
    amg_lists L = {NULL, NULL, (int *)malloc(sizeof(int) * nn), (int *)malloc(sizeof(int) * nn)};
    /* S^T (:1069-1094): rows of S^T in ascending order of the rows of S */
    for (int64_t p = 0; p < SI[n]; p++) STI[SJ[p] + 1]++;
    for (int i = 0; i < n; i++) STI[i + 1] += STI[i];
    {
        int *pos = (int *)malloc(sizeof(int) * nn);
        for (int i = 0; i < n; i++) pos[i] = STI[i];
        for (int i = 0; i < n; i++)
            for (int p = SI[i]; p < SI[i + 1]; p++) STJ[pos[SJ[p]]++] = i;
        free(pos);
    }
    for (int i = 0; i < n; i++) meas[i] = STI[i + 1] - STI[i]; /* :1108-1111 */
    int num_left = 0;                                            /* :1176-1201 */
    for (int j = 0; j < n; j++) {
        if (SI[j + 1] - SI[j] == 0) {
            cf[j] = -3;
            meas[j] = 0;
        } else {
            cf[j] = 0;
            num_left++;
        }
    }
    for (int j = 0; j < n; j++) { /* :1222-1266 */
        if (cf[j] == -3) continue;
        if (meas[j] > 0) {
            amg_lists_enter(&L, meas[j], j);
            continue;
        }
        cf[j] = -2;
        for (int k = SI[j]; k < SI[j + 1]; k++) {
            const int nb = SJ[k];
            if (cf[nb] == -3) continue;
            if (nb < j) {
                if (meas[nb] > 0) amg_lists_remove(&L, meas[nb], nb);
                amg_lists_enter(&L, ++meas[nb], nb);
            } else {
                ++meas[nb];
            }
        }
        --num_left;
    }
    while (num_left > 0 && L.top) { /* :1288-1366 */
        const int index = L.top->head;
        cf[index] = 1;
        const int m = meas[index];
        meas[index] = 0;
        --num_left;
        amg_lists_remove(&L, m, index);
        for (int p = STI[index]; p < STI[index + 1]; p++) {
            const int nb = STJ[p];
            if (cf[nb] != 0) continue;
            cf[nb] = -1;
            amg_lists_remove(&L, meas[nb], nb);
            --num_left;
            for (int k = SI[nb]; k < SI[nb + 1]; k++) {
                const int n2 = SJ[k];
                if (cf[n2] != 0) continue;
                amg_lists_remove(&L, meas[n2], n2);
                amg_lists_enter(&L, ++meas[n2], n2);
            }
        }
        for (int p = SI[index]; p < SI[index + 1]; p++) {
            const int nb = SJ[p];
            if (cf[nb] != 0) continue;
            amg_lists_remove(&L, meas[nb], nb);
            const int mm = --meas[nb];
            if (mm > 0) {
                amg_lists_enter(&L, mm, nb);
            } else {
                cf[nb] = -1;
                --num_left;
                for (int k = SI[nb]; k < SI[nb + 1]; k++) {
                    const int n2 = SJ[k];
                    if (cf[n2] != 0) continue;
                    amg_lists_remove(&L, meas[n2], n2);
                    amg_lists_enter(&L, ++meas[n2], n2);
                }
            }
        }
    }
    while (L.top) {
        amg_list *d = L.top->down;
        free(L.top);
        L.top = d;
    }
    free(L.next);
    free(L.prev);
    free(meas);
    free(STI);
    free(STJ);
}

/* hypre_BoomerAMGCoarsenHMIS par_coarsen.c:2845-2866: the Ruge first pass,
 * then PMIS with CF_init 1 on its markers; returns the number of C points */
int orc_amg_hmis(int n, const int *SI, const int *SJ, int *cf)
{
    orc_amg_ruge_first_pass(n, SI, SJ, cf);
    return orc_amg_pmis(n, SI, SJ, 1, cf);
}

/* The Galerkin product as HYPRE's host setup forms it,
 * hypre_BoomerAMGBuildCoarseOperatorKT par_rap.c:1525-1930 on one process
 * (square RAP, R = P^T by hypre_CSRMatrixTranspose, so a row of R lists its
 * fine rows in ascending order): row ic starts with its diagonal (zero,
 * :1607-1613); row ic of R*A is accumulated over R's row in order and each A
 * row in its stored order, a new column taking the next slot (:1759-1788);
 * then the RA row's entries, in that order, are multiplied into P's rows in
 * their stored order, a new column taking the next slot of the RAP row
 * (:1855-1890).  Products r*a and (r*a)*p are rounded separately (no fused
 * multiply-add).  Pass 1 (fill = 0) counts RAP_i[nc+1] and returns nnz; pass 2
 * fills RAP_j/RAP_a. */
#pragma GCC push_options
#pragma GCC optimize("fp-contract=off")
static int64_t amg_rap_kt(int n, int nc, const int *AI, const int *AJ, const double *AA,
                          const int *PI, const int *PJ, const double *PA, int *RAPI, int *RAPJ,
                          double *RAPA, int fill)
{
    /* R = P^T */
    int *RI = (int *)calloc((size_t)nc + 1, sizeof(int));
    const int64_t nnzP = PI[n];
    int *RJ = (int *)malloc(sizeof(int) * (size_t)(nnzP > 0 ? nnzP : 1));
    double *RA = (double *)malloc(sizeof(double) * (size_t)(nnzP > 0 ? nnzP : 1));
    for (int64_t p = 0; p < nnzP; p++) RI[PJ[p] + 1]++;
    for (int c = 0; c < nc; c++) RI[c + 1] += RI[c];
    int *pos = (int *)malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
    for (int c = 0; c < nc; c++) pos[c] = RI[c];
    for (int i = 0; i < n; i++)
        for (int p = PI[i]; p < PI[i + 1]; p++) {
            RJ[pos[PJ[p]]] = i;
            RA[pos[PJ[p]]++] = fill ? PA[p] : 0.0;
        }
    free(pos);
    int64_t *amark = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t *pmark = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nc > 0 ? nc : 1));
    int *raj = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    double *rav = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) amark[i] = -1;
    for (int c = 0; c < nc; c++) pmark[c] = -1;
    int64_t cnt = 0;
    for (int ic = 0; ic < nc; ic++) {
        const int64_t row_begin = cnt;
        if (!fill) RAPI[ic] = (int)cnt;
        else cnt = RAPI[ic];
        const int64_t rb = fill ? RAPI[ic] : row_begin;
        pmark[ic] = cnt;
        if (fill) {
            RAPJ[cnt] = ic;
            RAPA[cnt] = 0.0;
        }
        cnt++;
        /* row ic of R*A */
        int nra = 0;
        for (int p1 = RI[ic]; p1 < RI[ic + 1]; p1++) {
            const int i1 = RJ[p1];
            const double r = RA[p1];
            for (int p2 = AI[i1]; p2 < AI[i1 + 1]; p2++) {
                const int i2 = AJ[p2];
                const double ra = fill ? r * AA[p2] : 0.0;
                if (amark[i2] < 0) {
                    amark[i2] = nra;
                    raj[nra] = i2;
                    rav[nra++] = ra;
                } else {
                    rav[amark[i2]] += ra;
                }
            }
        }
        /* (R*A)*P */
        for (int q = 0; q < nra; q++) {
            const int i1 = raj[q];
            const double ra = rav[q];
            amark[i1] = -1;
            for (int p2 = PI[i1]; p2 < PI[i1 + 1]; p2++) {
                const int i2 = PJ[p2];
                const double v = fill ? ra * PA[p2] : 0.0;
                if (pmark[i2] < rb) {
                    pmark[i2] = cnt;
                    if (fill) {
                        RAPJ[cnt] = i2;
                        RAPA[cnt] = v;
                    }
                    cnt++;
                } else if (fill) {
                    RAPA[pmark[i2]] += v;
                }
            }
        }
    }
    if (!fill) RAPI[nc] = (int)cnt;
    free(RI);
    free(RJ);
    free(RA);
    free(amark);
    free(pmark);
    free(raj);
    free(rav);
    return cnt;
}
#pragma GCC pop_options

int64_t orc_amg_rap_kt_count(int n, int nc, const int *AI, const int *AJ, const int *PI,
                             const int *PJ, int *RAPI)
{
    return amg_rap_kt(n, nc, AI, AJ, NULL, PI, PJ, NULL, RAPI, NULL, NULL, 0);
}

void orc_amg_rap_kt_fill(int n, int nc, const int *AI, const int *AJ, const double *AA,
                         const int *PI, const int *PJ, const double *PA, const int *RAPI,
                         int *RAPJ, double *RAPA)
{
    amg_rap_kt(n, nc, AI, AJ, AA, PI, PJ, PA, (int *)RAPI, RAPJ, RAPA, 1);
}

/* par_lr_interp.c:1025-1928 on one process: the structure (C points of S row
 * i, then the strong C points of each strong F neighbour, first-touch order)
 * and the weights of extended+i interpolation. */
static int amg_pass(int n, const int *AI, const int *AJ, const double *AA, const int *SI,
                    const int *SJ, const int *cf, const int *f2c, int64_t *pm, const int *PI,
                    int *PJ, double *PA, int fill)
{
    int64_t counter = 0;
    int64_t strong_f_marker = -2;
    for (int i = 0; i < n; i++) pm[i] = -1;
    for (int i = 0; i < n; i++) {
        const int64_t begin = fill ? PI[i] : counter;
        if (fill) counter = begin;
        if (cf[i] >= 0) { /* :1488-1495 C point: identity */
            if (fill) {
                PJ[counter] = f2c[i];
                PA[counter] = 1.0;
            }
            counter++;
            continue;
        }
        if (cf[i] == -3) continue;
        strong_f_marker--;
        for (int p = SI[i]; p < SI[i + 1]; p++) { /* :1501-1560 */
            const int i1 = SJ[p];
            if (cf[i1] >= 0) {
                if (pm[i1] < begin) {
                    pm[i1] = counter;
                    if (fill) {
                        PJ[counter] = f2c[i1];
                        PA[counter] = 0.0;
                    }
                    counter++;
                }
            } else if (cf[i1] != -3) {
                pm[i1] = strong_f_marker;
                for (int q = SI[i1]; q < SI[i1 + 1]; q++) {
                    const int k1 = SJ[q];
                    if (cf[k1] >= 0 && pm[k1] < begin) {
                        pm[k1] = counter;
                        if (fill) {
                            PJ[counter] = f2c[k1];
                            PA[counter] = 0.0;
                        }
                        counter++;
                    }
                }
            }
        }
        if (!fill) continue;
        /* :1630-1708 weights */
        const int64_t di = amg_diag_pos(AI, AJ, i);
        double diagonal = di >= 0 ? AA[di] : 0.0;
        for (int64_t jj = AI[i]; jj < AI[i + 1]; jj++) {
            if (jj == di) continue;
            const int i1 = AJ[jj];
            if (pm[i1] >= begin) {
                PA[pm[i1]] += AA[jj];
            } else if (pm[i1] == strong_f_marker) {
                const int64_t d1 = amg_diag_pos(AI, AJ, i1);
                const int sgn = (d1 >= 0 && AA[d1] < 0) ? -1 : 1;
                double sum = 0.0;
                for (int64_t q = AI[i1]; q < AI[i1 + 1]; q++) {
                    if (q == d1) continue;
                    const int i2 = AJ[q];
                    if ((pm[i2] >= begin || i2 == i) && (sgn * AA[q]) < 0) sum += AA[q];
                }
                if (sum != 0) {
                    const double distribute = AA[jj] / sum;
                    for (int64_t q = AI[i1]; q < AI[i1 + 1]; q++) {
                        if (q == d1) continue;
                        const int i2 = AJ[q];
                        if (pm[i2] >= begin && (sgn * AA[q]) < 0) PA[pm[i2]] += distribute * AA[q];
                        if (i2 == i && (sgn * AA[q]) < 0) diagonal += distribute * AA[q];
                    }
                } else {
                    diagonal += AA[jj];
                }
            } else if (cf[i1] != -3) {
                diagonal += AA[jj];
            }
        }
        if (diagonal != 0) /* :1790-1800 */
            for (int64_t jj = begin; jj < counter; jj++) PA[jj] /= -diagonal;
    }
    return (int)counter;
}

static int *amg_fine_to_coarse(int n, const int *cf)
{
    int *f2c = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int c = 0;
    for (int i = 0; i < n; i++) f2c[i] = cf[i] >= 0 ? c++ : -1;
    return f2c;
}

int64_t orc_amg_extpi_count(int n, const int *AI, const int *AJ, const int *SI, const int *SJ,
                            const int *cf, int *PI)
{
    int *f2c = amg_fine_to_coarse(n, cf);
    int64_t *pm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    /* count row by row: a pass that records each row's start */
    int64_t total = 0;
    for (int i = 0; i < n; i++) pm[i] = -1;
    int64_t strong_f_marker = -2;
    for (int i = 0; i < n; i++) {
        PI[i] = (int)total;
        const int64_t begin = total;
        if (cf[i] >= 0) { total++; continue; }
        if (cf[i] == -3) continue;
        strong_f_marker--;
        for (int p = SI[i]; p < SI[i + 1]; p++) {
            const int i1 = SJ[p];
            if (cf[i1] >= 0) {
                if (pm[i1] < begin) { pm[i1] = total; total++; }
            } else if (cf[i1] != -3) {
                pm[i1] = strong_f_marker;
                for (int q = SI[i1]; q < SI[i1 + 1]; q++) {
                    const int k1 = SJ[q];
                    if (cf[k1] >= 0 && pm[k1] < begin) { pm[k1] = total; total++; }
                }
            }
        }
    }
    PI[n] = (int)total;
    free(pm);
    free(f2c);
    return total;
}

void orc_amg_extpi_fill(int n, const int *AI, const int *AJ, const double *AA, const int *SI,
                        const int *SJ, const int *cf, const int *PI, int *PJ, double *PA)
{
    int *f2c = amg_fine_to_coarse(n, cf);
    int64_t *pm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    amg_pass(n, AI, AJ, AA, SI, SJ, cf, f2c, pm, PI, PJ, PA, 1);
    free(pm);
    free(f2c);
}

/* qsort.c:395-417, descending |w| */
static void amg_swap2(int *v, double *w, int i, int j)
{
    int t = v[i];
    v[i] = v[j];
    v[j] = t;
    double t2 = w[i];
    w[i] = w[j];
    w[j] = t2;
}

static void amg_qsort2_abs(int *v, double *w, int left, int right)
{
    if (left >= right) return;
    amg_swap2(v, w, left, (left + right) / 2);
    int last = left;
    for (int i = left + 1; i <= right; i++)
        if (fabs(w[i]) > fabs(w[left])) amg_swap2(v, w, ++last, i);
    amg_swap2(v, w, left, last);
    amg_qsort2_abs(v, w, left, last - 1);
    amg_qsort2_abs(v, w, last + 1, right);
}

/* par_csr_matrix.c:3085-3230 (tol = 0, rescale = 1, one process) */
int64_t orc_amg_truncate(int n, int *PI, int *PJ, double *PA, int max_elmts)
{
    if (max_elmts <= 0) return PI[n];
    int64_t out = 0;
    for (int i = 0; i < n; i++) {
        const int64_t b = PI[i], e = PI[i + 1];
        const int cnt = (int)(e - b);
        PI[i] = (int)out;
        if (cnt > max_elmts) {
            double row_sum = 0.0;
            for (int64_t p = b; p < e; p++) row_sum += PA[p];
            /* sort a copy in place: entries b.. are not read again */
            int *vj = PJ + b;
            double *vw = PA + b;
            amg_qsort2_abs(vj, vw, 0, cnt - 1);
            double scale = 0.0;
            for (int k = 0; k < max_elmts; k++) scale += vw[k];
            for (int k = 0; k < max_elmts; k++) {
                PJ[out + k] = vj[k];
                PA[out + k] = vw[k];
            }
            if (scale != 0. && scale != row_sum) {
                scale = row_sum / scale;
                for (int k = 0; k < max_elmts; k++) PA[out + k] *= scale;
            }
            out += max_elmts;
        } else {
            for (int64_t p = b; p < e; p++) {
                PJ[out] = PJ[p];
                PA[out] = PA[p];
                out++;
            }
        }
    }
    PI[n] = (int)out;
    return out;
}

/* ------------------------------------------------------------------------ */
/* Per-level policy network: plain loops over the layers, double precision.
 * Layer l maps h_(l-1) to z_l = W_l h_(l-1) + b_l, h_l = tanh(z_l) for the two
 * hidden layers and o = z_3 for the output.  Backward by the chain rule:
 * do = d loss / d o, dW_l = dz_l h_(l-1)^T, db_l = dz_l, and
 * dz_(l-1) = (W_l^T dz_l) * (1 - h_(l-1)^2). */
#define POL_IN 12
#define POL_HID 64
#define POL_OUT 4
#define POL_W1 0
#define POL_B1 (POL_W1 + POL_HID * POL_IN)
#define POL_W2 (POL_B1 + POL_HID)
#define POL_B2 (POL_W2 + POL_HID * POL_HID)
#define POL_W3 (POL_B2 + POL_HID)
#define POL_B3 (POL_W3 + POL_OUT * POL_HID)

static void pol_layer(const double *W, const double *b, const double *in, int nin, int nout,
                      int act, double *out)
{
    for (int o = 0; o < nout; ++o) {
        double acc = b[o];
        for (int i = 0; i < nin; ++i) acc += W[o * nin + i] * in[i];
        out[o] = act ? tanh(acc) : acc;
    }
}

static void pol_sample(const double *p, const double *x, double *h1, double *h2, double *o)
{
    pol_layer(p + POL_W1, p + POL_B1, x, POL_IN, POL_HID, 1, h1);
    pol_layer(p + POL_W2, p + POL_B2, h1, POL_HID, POL_HID, 1, h2);
    pol_layer(p + POL_W3, p + POL_B3, h2, POL_HID, POL_OUT, 0, o);
}

void orc_policy_forward(const double *params, const double *x, int64_t n, double *out)
{
    double h1[POL_HID], h2[POL_HID];
    for (int64_t s = 0; s < n; ++s) pol_sample(params, x + s * POL_IN, h1, h2, out + s * POL_OUT);
}

double orc_policy_loss_grad(const double *params, const double *x, const int *cls,
                            const double *target, int64_t n, double lambda, double *grad)
{
    double h1[POL_HID], h2[POL_HID], o[POL_OUT], dout[POL_OUT], dz2[POL_HID], dz1[POL_HID];
    double loss = 0.0;
    memset(grad, 0, sizeof(double) * (POL_B3 + POL_OUT));
    for (int64_t s = 0; s < n; ++s) {
        const double *xs = x + s * POL_IN;
        pol_sample(params, xs, h1, h2, o);
        /* softmax cross-entropy over the three precision logits */
        double m = o[0];
        for (int c = 1; c < 3; ++c) m = o[c] > m ? o[c] : m;
        double sum = 0.0, e[3];
        for (int c = 0; c < 3; ++c) { e[c] = exp(o[c] - m); sum += e[c]; }
        const int y = cls[s];
        const double r = o[3] - target[s];
        loss += log(sum) + m - o[y] + lambda * r * r;
        for (int c = 0; c < 3; ++c) dout[c] = (e[c] / sum - (c == y)) / (double)n;
        dout[3] = 2.0 * lambda * r / (double)n;
        /* output layer */
        for (int k = 0; k < POL_OUT; ++k) {
            for (int i = 0; i < POL_HID; ++i) grad[POL_W3 + k * POL_HID + i] += dout[k] * h2[i];
            grad[POL_B3 + k] += dout[k];
        }
        for (int i = 0; i < POL_HID; ++i) {
            double a = 0.0;
            for (int k = 0; k < POL_OUT; ++k) a += params[POL_W3 + k * POL_HID + i] * dout[k];
            dz2[i] = a * (1.0 - h2[i] * h2[i]);
        }
        /* second hidden layer */
        for (int k = 0; k < POL_HID; ++k) {
            for (int i = 0; i < POL_HID; ++i) grad[POL_W2 + k * POL_HID + i] += dz2[k] * h1[i];
            grad[POL_B2 + k] += dz2[k];
        }
        for (int i = 0; i < POL_HID; ++i) {
            double a = 0.0;
            for (int k = 0; k < POL_HID; ++k) a += params[POL_W2 + k * POL_HID + i] * dz2[k];
            dz1[i] = a * (1.0 - h1[i] * h1[i]);
        }
        /* first hidden layer */
        for (int k = 0; k < POL_HID; ++k) {
            for (int i = 0; i < POL_IN; ++i) grad[POL_W1 + k * POL_IN + i] += dz1[k] * xs[i];
            grad[POL_B1 + k] += dz1[k];
        }
    }
    return loss / (double)n;
}

void orc_policy_adam(double *params, const double *grad, double *m, double *v, int64_t np,
                     int64_t step, double lr, double b1, double b2, double eps)
{
    const double bc1 = 1.0 - pow(b1, (double)step), bc2 = 1.0 - pow(b2, (double)step);
    for (int64_t i = 0; i < np; ++i) {
        m[i] = b1 * m[i] + (1.0 - b1) * grad[i];
        v[i] = b2 * v[i] + (1.0 - b2) * grad[i] * grad[i];
        params[i] -= lr * (m[i] / bc1) / (sqrt(v[i] / bc2) + eps);
    }
}
