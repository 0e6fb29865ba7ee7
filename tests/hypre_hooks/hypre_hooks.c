/*
 * The HYPRE side of the drop-in boundary, compiled and run.
 *
 * HYPRE 2.31 reaches the mBSR path through two device entry points that AmgT
 * hooks: hypreDevice_CSRSpGemm (src/seq_mv/csr_spgemm_device.c:2579-2608,
 * declared seq_mv.h:727) and hypre_CSRMatrixMatvecDevice2
 * (src/seq_mv/csr_matvec_device.c:56-124, reached from
 * hypre_CSRMatrixMatvecDevice seq_mv.h:562).  The blocks between the
 * "hook:" markers below are the code a maintainer puts into those functions;
 * INTEGRATION.md quotes them from this file.  Everything else here stands in
 * for the few parts of HYPRE the hooks touch (the hypre_CSRMatrix and
 * hypre_Vector fields, hypre_TAlloc/hypre_TFree on device memory,
 * hypre_error_w_msg, the compute stream), so that the hooks compile as HYPRE
 * would compile them and can be run against the oracle.
 *
 * main(nx, ny, nz, outdir): a non-symmetric 7-point operator on nx*ny*nz
 * (rows as HYPRE stores them: the diagonal first), then
 *   C = A*A through the SpGEMM hook (and once more with A pattern-only),
 *   y = 2.5 A x - 0.5 y and z = A^T x through the SpMV hook,
 * all arrays written to outdir as raw little-endian files.
 */
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "amgt_mbsr.h"

/* ---- stand-ins for the parts of HYPRE the hooks use ---------------------------------- */
typedef int HYPRE_Int;
typedef double HYPRE_Complex;
typedef enum { HYPRE_MEMORY_HOST = 0, HYPRE_MEMORY_DEVICE = 1 } HYPRE_MemoryLocation;
#define HYPRE_ERROR_GENERIC 1
static HYPRE_Int hypre_error_flag = 0;
static void hypre_error_w_msg(HYPRE_Int code, const char *msg)
{
   hypre_error_flag |= code;
   fprintf(stderr, "hypre error: %s\n", msg);
}

/* hypre_CSRMatrix (seq_mv.h) with the two fields the integration adds */
typedef struct {
   HYPRE_Int *i, *j;
   HYPRE_Complex *data;
   HYPRE_Int num_rows, num_cols, num_nonzeros;
   HYPRE_Int owns_data, pattern_only;
   HYPRE_MemoryLocation memory_location;
   amgt_mbsr_t mbsr;          /* lazily built mBSR of (i, j, data); NULL = not built */
   HYPRE_Int mbsr_level;      /* coarsening level tag (AmgT's coasen_level) */
} hypre_CSRMatrix;
#define hypre_CSRMatrixI(m) ((m)->i)
#define hypre_CSRMatrixJ(m) ((m)->j)
#define hypre_CSRMatrixData(m) ((m)->data)
#define hypre_CSRMatrixNumRows(m) ((m)->num_rows)
#define hypre_CSRMatrixNumCols(m) ((m)->num_cols)
#define hypre_CSRMatrixNumNonzeros(m) ((m)->num_nonzeros)
#define hypre_CSRMatrixPatternOnly(m) ((m)->pattern_only)
#define hypre_CSRMatrixMemoryLocation(m) ((m)->memory_location)
#define hypre_CSRMatrixMbsr(m) ((m)->mbsr)
#define hypre_CSRMatrixMbsrLevel(m) ((m)->mbsr_level)

typedef struct {
   HYPRE_Complex *data;
   HYPRE_Int size, num_vectors;
} hypre_Vector;
#define hypre_VectorData(v) ((v)->data)
#define hypre_VectorNumVectors(v) ((v)->num_vectors)

static hipStream_t compute_stream;
#define hypre_handle() NULL
#define hypre_HandleComputeStream(h) ((void *) compute_stream)

static void *device_alloc(size_t bytes)
{
   void *p = NULL;
   return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess ? p : NULL;
}
#define hypre_TAlloc(type, count, location) ((type *) device_alloc(sizeof(type) * (size_t)(count)))
#define hypre_TFree(ptr, location) ((void) hipFree(ptr))

static hypre_CSRMatrix *hypre_CSRMatrixCreate(HYPRE_Int rows, HYPRE_Int cols, HYPRE_Int nnz)
{
   hypre_CSRMatrix *m = (hypre_CSRMatrix *) calloc(1, sizeof(hypre_CSRMatrix));
   m->num_rows = rows;
   m->num_cols = cols;
   m->num_nonzeros = nnz;
   m->owns_data = 1;
   return m;
}

/* hypre_CSRMatrixDestroy (csr_matrix.c) with the cache release the integration adds */
static void hypre_CSRMatrixDestroy(hypre_CSRMatrix *A)
{
   amgt_mbsr_destroy(hypre_CSRMatrixMbsr(A));
   hypre_CSRMatrixMbsr(A) = NULL;
   if (A->owns_data) {
      hypre_TFree(A->i, HYPRE_MEMORY_DEVICE);
      hypre_TFree(A->j, HYPRE_MEMORY_DEVICE);
      hypre_TFree(A->data, HYPRE_MEMORY_DEVICE);
   }
   free(A);
}

/* ---- hook: allocator (once, after HYPRE_Init) ---------------------------------------- */
static void *amgt_hypre_alloc(size_t n, void *ctx) { return hypre_TAlloc(char, n, HYPRE_MEMORY_DEVICE); }
static void  amgt_hypre_free(void *p, void *ctx)   { hypre_TFree(p, HYPRE_MEMORY_DEVICE); }
/* ---- end hook ---- */

/* ---- hook: SpGEMM (csr_spgemm_device.c, hypreDevice_CSRSpGemm) ----------------------- */
HYPRE_Int
hypreDevice_CSRSpGemm(hypre_CSRMatrix *A, hypre_CSRMatrix *B, hypre_CSRMatrix **C_ptr)
{
   HYPRE_Int m = hypre_CSRMatrixNumRows(A), k = hypre_CSRMatrixNumCols(A);
   HYPRE_Int n = hypre_CSRMatrixNumCols(B);
   int64_t nnzC; int32_t *ci, *cj; double *cd; amgt_mbsr_t cache = NULL;
   hypre_CSRMatrix *C;
   if (hypre_CSRMatrixMbsr(A)) amgt_mbsr_set_level(hypre_CSRMatrixMbsr(A), hypre_CSRMatrixMbsrLevel(A));
   int st = amgt_csr_spgemm(m, k, n,
              hypre_CSRMatrixNumNonzeros(A), hypre_CSRMatrixI(A), hypre_CSRMatrixJ(A),
              hypre_CSRMatrixPatternOnly(A) ? NULL : hypre_CSRMatrixData(A), &A->mbsr,
              hypre_CSRMatrixNumNonzeros(B), hypre_CSRMatrixI(B), hypre_CSRMatrixJ(B),
              hypre_CSRMatrixPatternOnly(B) ? NULL : hypre_CSRMatrixData(B), &B->mbsr,
              &nnzC, &ci, &cj, &cd, &cache, hypre_HandleComputeStream(hypre_handle()));
   if (st != AMGT_OK) { hypre_error_w_msg(HYPRE_ERROR_GENERIC, amgt_last_error()); return hypre_error_flag; }
   *C_ptr = C = hypre_CSRMatrixCreate(m, n, (HYPRE_Int) nnzC);
   hypre_CSRMatrixI(C) = ci; hypre_CSRMatrixJ(C) = cj; hypre_CSRMatrixData(C) = cd;
   hypre_CSRMatrixMemoryLocation(C) = HYPRE_MEMORY_DEVICE;
   hypre_CSRMatrixMbsr(C) = cache; hypre_CSRMatrixMbsrLevel(C) = hypre_CSRMatrixMbsrLevel(A);
   return hypre_error_flag;
}
/* ---- end hook ---- */

/* hypre_CSRMatrixMatvecDevice2 (csr_matvec_device.c:56): y = alpha op(A) x + beta y */
HYPRE_Int
hypre_CSRMatrixMatvecDevice2(HYPRE_Int trans, HYPRE_Complex alpha, hypre_CSRMatrix *A,
                             hypre_Vector *x, HYPRE_Complex beta, hypre_Vector *y,
                             HYPRE_Int offset)
{
   HYPRE_Int num_vectors = hypre_VectorNumVectors(x);
   if (offset != 0) { hypre_error_w_msg(HYPRE_ERROR_GENERIC, "offset != 0"); return hypre_error_flag; }
/* ---- hook: SpMV (csr_matvec_device.c, hypre_CSRMatrixMatvecDevice2) ----------------- */
   if (num_vectors == 1) {
      if (hypre_CSRMatrixMbsr(A)) amgt_mbsr_set_level(hypre_CSRMatrixMbsr(A), hypre_CSRMatrixMbsrLevel(A));
      int st = amgt_csr_matvec(trans, alpha, hypre_CSRMatrixNumRows(A), hypre_CSRMatrixNumCols(A),
                 hypre_CSRMatrixNumNonzeros(A), hypre_CSRMatrixI(A), hypre_CSRMatrixJ(A),
                 hypre_CSRMatrixData(A), &A->mbsr, hypre_VectorData(x), beta,
                 hypre_VectorData(y) /* b == y here */, hypre_VectorData(y),
                 hypre_HandleComputeStream(hypre_handle()));
      if (st != AMGT_OK) hypre_error_w_msg(HYPRE_ERROR_GENERIC, amgt_last_error());
      return hypre_error_flag;
   }
/* ---- end hook ---- */
   hypre_error_w_msg(HYPRE_ERROR_GENERIC, "num_vectors > 1 is not driven here");
   return hypre_error_flag;
}

/* ---- driver ------------------------------------------------------------------------- */
#define CHECK(call)                                                         \
   do {                                                                     \
      hipError_t e_ = (call);                                               \
      if (e_ != hipSuccess) {                                               \
         fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #call,       \
                 hipGetErrorString(e_));                                    \
         exit(2);                                                           \
      }                                                                     \
   } while (0)

static void *to_device(const void *h, size_t bytes)
{
   void *d = device_alloc(bytes);
   if (!d) { fprintf(stderr, "device allocation failed\n"); exit(2); }
   CHECK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
   return d;
}

static void write_device(const char *dir, const char *name, const void *d, size_t bytes)
{
   char path[4096];
   snprintf(path, sizeof(path), "%s/%s", dir, name);
   void *h = malloc(bytes ? bytes : 1);
   if (bytes) CHECK(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
   FILE *f = fopen(path, "wb");
   if (!f || fwrite(h, 1, bytes, f) != bytes) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
   fclose(f);
   free(h);
}

int main(int argc, char **argv)
{
   if (argc != 5) { fprintf(stderr, "usage: %s nx ny nz outdir\n", argv[0]); return 2; }
   const int nx = atoi(argv[1]), ny = atoi(argv[2]), nz = atoi(argv[3]);
   const char *out = argv[4];
   const int n = nx * ny * nz;
   int *I = (int *) malloc(sizeof(int) * (size_t)(n + 1));
   int *J = (int *) malloc(sizeof(int) * (size_t) n * 7);
   double *V = (double *) malloc(sizeof(double) * (size_t) n * 7);
   int nnz = 0;
   I[0] = 0;
   for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
         for (int x = 0; x < nx; x++) {
            const int r = (z * ny + y) * nx + x;
            const int nb[6] = {z > 0 ? r - nx * ny : -1, y > 0 ? r - nx : -1, x > 0 ? r - 1 : -1,
                               x < nx - 1 ? r + 1 : -1, y < ny - 1 ? r + nx : -1,
                               z < nz - 1 ? r + nx * ny : -1};
            J[nnz] = r;
            V[nnz++] = 6.0 + 0.01 * (r % 5);
            for (int q = 0; q < 6; q++)
               if (nb[q] >= 0) {
                  J[nnz] = nb[q];
                  V[nnz++] = -(1.0 + 0.001 * ((r + 2 * nb[q]) % 11));
               }
            I[r + 1] = nnz;
         }
   CHECK(hipStreamCreate(&compute_stream));
   amgt_set_allocator(amgt_hypre_alloc, amgt_hypre_free, NULL);

   hypre_CSRMatrix *A = hypre_CSRMatrixCreate(n, n, nnz);
   A->i = (int *) to_device(I, sizeof(int) * (size_t)(n + 1));
   A->j = (int *) to_device(J, sizeof(int) * (size_t) nnz);
   A->data = (double *) to_device(V, sizeof(double) * (size_t) nnz);
   A->memory_location = HYPRE_MEMORY_DEVICE;

   hypre_CSRMatrix *C = NULL;
   if (hypreDevice_CSRSpGemm(A, A, &C)) return 1;
   CHECK(hipStreamSynchronize(compute_stream));
   write_device(out, "C_i.i32", C->i, sizeof(int) * (size_t)(n + 1));
   write_device(out, "C_j.i32", C->j, sizeof(int) * (size_t) C->num_nonzeros);
   write_device(out, "C_data.f64", C->data, sizeof(double) * (size_t) C->num_nonzeros);

   /* the same product with the left operand pattern-only: C's data is NULL */
   hypre_CSRMatrix *Ap = hypre_CSRMatrixCreate(n, n, nnz);
   Ap->i = A->i; Ap->j = A->j; Ap->data = A->data; Ap->owns_data = 0; Ap->pattern_only = 1;
   hypre_CSRMatrix *Cp = NULL;
   if (hypreDevice_CSRSpGemm(Ap, A, &Cp)) return 1;
   CHECK(hipStreamSynchronize(compute_stream));
   if (Cp->data != NULL || Cp->num_nonzeros != C->num_nonzeros) {
      fprintf(stderr, "pattern-only product: data %p nnz %d (expected NULL, %d)\n",
              (void *) Cp->data, Cp->num_nonzeros, C->num_nonzeros);
      return 1;
   }
   write_device(out, "Cp_j.i32", Cp->j, sizeof(int) * (size_t) Cp->num_nonzeros);

   double *hx = (double *) malloc(sizeof(double) * (size_t) n);
   double *hy = (double *) malloc(sizeof(double) * (size_t) n);
   for (int r = 0; r < n; r++) {
      hx[r] = 1.0 / (1.0 + r % 13);
      hy[r] = 0.25 * (r % 7) - 0.5;
   }
   hypre_Vector x = {(double *) to_device(hx, sizeof(double) * (size_t) n), n, 1};
   hypre_Vector y = {(double *) to_device(hy, sizeof(double) * (size_t) n), n, 1};
   hypre_Vector z = {(double *) to_device(hy, sizeof(double) * (size_t) n), n, 1};
   if (hypre_CSRMatrixMatvecDevice2(0, 2.5, A, &x, -0.5, &y, 0)) return 1;
   if (hypre_CSRMatrixMatvecDevice2(1, 1.0, A, &x, 0.0, &z, 0)) return 1;
   CHECK(hipStreamSynchronize(compute_stream));
   write_device(out, "y.f64", y.data, sizeof(double) * (size_t) n);
   write_device(out, "z.f64", z.data, sizeof(double) * (size_t) n);

   const int nnzC = C->num_nonzeros;
   hypre_CSRMatrixDestroy(Cp);
   hypre_CSRMatrixDestroy(C);
   hypre_CSRMatrixDestroy(Ap);
   hypre_CSRMatrixDestroy(A);
   hypre_TFree(x.data, HYPRE_MEMORY_DEVICE);
   hypre_TFree(y.data, HYPRE_MEMORY_DEVICE);
   hypre_TFree(z.data, HYPRE_MEMORY_DEVICE);
   CHECK(hipStreamDestroy(compute_stream));
   printf("hooks ok: n=%d nnz(A)=%d nnz(C)=%d\n", n, nnz, nnzC);
   free(I); free(J); free(V); free(hx); free(hy);
   return 0;
}
