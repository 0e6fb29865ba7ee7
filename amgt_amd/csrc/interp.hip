// Trilinear prolongation of structured grids, for a batch of independent
// grids at once (block-diagonal P: problem k's fine rows and coarse columns
// follow problem k-1's).  Used by the geometric driver of the V-cycle; the
// algebraic counterpart, as BoomerAMG builds it, is amgt_amg_extpi_interp.
//
// Per grid and direction the coarse points are the even fine indices
// (nc = (n + 1) / 2).  An even fine index takes weight 1 from its coarse
// point; an odd one 1/2 from each even neighbour, a neighbour beyond the last
// coarse point being a zero Dirichlet value (no entry).  A row holds the
// tensor products of its three directions' entries, at most 8, in
// (z, y, x)-major order, which is ascending column order.
#include "internal.hpp"

#include <vector>

namespace amgt {

namespace {

constexpr int kThreads = 256;

struct Grid {          // one problem of the batch
    int64_t row0;      // its first fine row
    int64_t col0;      // its first coarse column
    int32_t nx, ny, nz;
    int32_t pad;
};

// problem of fine row r: the last grid whose first row is <= r
__device__ __forceinline__ int find_grid(const Grid *__restrict__ g, int nprob, int64_t r) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g[mid].row0 <= r) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// entries of fine index f along a direction of n points: 1 (even f), 2 (odd f
// with an upper coarse neighbour), 1 (odd f at the upper end)
__device__ __forceinline__ int axis_count(int f, int n) {
    if ((f & 1) == 0) return 1;
    return (f / 2 + 1) < (n + 1) / 2 ? 2 : 1;
}

__global__ __launch_bounds__(kThreads) void k_tri_count(int64_t n, const Grid *__restrict__ g,
                                                       int nprob, int32_t *__restrict__ cnt) {
    const int64_t r = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (r >= n) return;
    const Grid G = g[find_grid(g, nprob, r)];
    const int64_t loc = r - G.row0;
    const int x = int(loc % G.nx);
    const int y = int((loc / G.nx) % G.ny);
    const int z = int(loc / (int64_t(G.nx) * G.ny));
    cnt[r] = axis_count(x, G.nx) * axis_count(y, G.ny) * axis_count(z, G.nz);
}

__global__ __launch_bounds__(kThreads) void k_tri_fill(int64_t n, const Grid *__restrict__ g,
                                                      int nprob, const int32_t *__restrict__ pi,
                                                      int32_t *__restrict__ pj,
                                                      double *__restrict__ pa) {
    const int64_t r = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (r >= n) return;
    const Grid G = g[find_grid(g, nprob, r)];
    const int64_t loc = r - G.row0;
    const int f[3] = {int(loc % G.nx), int((loc / G.nx) % G.ny),
                      int(loc / (int64_t(G.nx) * G.ny))};
    const int nd[3] = {G.nx, G.ny, G.nz};
    int c[3][2], k[3];
    double w[3][2];
    for (int d = 0; d < 3; ++d) {
        const int nc = (nd[d] + 1) / 2;
        c[d][0] = f[d] / 2;
        if ((f[d] & 1) == 0) {
            w[d][0] = 1.0;
            k[d] = 1;
        } else {
            w[d][0] = 0.5;
            c[d][1] = f[d] / 2 + 1;
            w[d][1] = 0.5;
            k[d] = c[d][1] < nc ? 2 : 1;
        }
    }
    const int ncx = (G.nx + 1) / 2, ncy = (G.ny + 1) / 2;
    int64_t p = pi[r];
    for (int a = 0; a < k[2]; ++a)
        for (int b = 0; b < k[1]; ++b)
            for (int e = 0; e < k[0]; ++e) {
                pj[p] = int32_t(G.col0 + (int64_t(c[2][a]) * ncy + c[1][b]) * ncx + c[0][e]);
                pa[p] = w[0][e] * w[1][b] * w[2][a];
                ++p;
            }
}

}  // namespace

}  // namespace amgt

using namespace amgt;

extern "C" {

int amgt_interp_trilinear(int32_t nprob, const int32_t *grids, int32_t *n_fine,
                          int32_t *n_coarse, int64_t *nnzP, int32_t **P_i, int32_t **P_j,
                          double **P_data, void *stream) {
    AMGT_CHECK_ARG(nprob > 0 && grids, "no grids");
    AMGT_CHECK_ARG(n_fine && n_coarse && nnzP && P_i && P_j && P_data, "NULL output argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<Grid> h(static_cast<size_t>(nprob));
    int64_t rows = 0, cols = 0;
    for (int32_t q = 0; q < nprob; ++q) {
        const int32_t nx = grids[3 * q], ny = grids[3 * q + 1], nz = grids[3 * q + 2];
        AMGT_CHECK_ARG(nx > 0 && ny > 0 && nz > 0, "grid sides must be positive");
        h[size_t(q)] = Grid{rows, cols, nx, ny, nz, 0};
        rows += int64_t(nx) * ny * nz;
        cols += int64_t((nx + 1) / 2) * ((ny + 1) / 2) * ((nz + 1) / 2);
    }
    // 8 entries per row at most: nnz < 2^31 keeps int32 row pointers
    if (rows * 8 >= (int64_t(1) << 31)) {
        set_error("trilinear P: more than 2^28 fine rows");
        return AMGT_ERR_UNSUPPORTED;
    }
    Grid *dg = nullptr;
    int32_t *cnt = nullptr;
    AMGT_TRY(tmp_alloc_t(&dg, size_t(nprob), s));
    int st = tmp_alloc_t(&cnt, size_t(rows) + 1, s);
    int32_t *pi = nullptr, *pj = nullptr;
    double *pa = nullptr;
    int32_t nnz = 0;
    if (st == AMGT_OK) {
        pi = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(rows) + 1), s));
        if (!pi) {
            set_error("allocation of P failed");
            st = AMGT_ERR_ALLOC;
        }
    }
    const unsigned grid = unsigned((rows + kThreads - 1) / kThreads);
    if (st == AMGT_OK) {
        // the table is read by the kernels only after the synchronisation below
        if (hipMemcpyAsync(dg, h.data(), sizeof(Grid) * size_t(nprob), hipMemcpyHostToDevice,
                           s) != hipSuccess)
            st = AMGT_ERR_HIP;
    }
    if (st == AMGT_OK) {
        hipLaunchKernelGGL(k_tri_count, dim3(grid), dim3(kThreads), 0, s, rows, dg, nprob, cnt);
        if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
    }
    if (st == AMGT_OK) st = exclusive_scan_i32(cnt, pi, rows, s);
    if (st == AMGT_OK) {
        const Readback rb[1] = {{&nnz, pi + rows, sizeof(int32_t)}};
        st = readback(rb, 1, s);
    }
    if (st == AMGT_OK) {
        pj = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(nnz) + 1), s));
        pa = static_cast<double *>(dev_alloc(sizeof(double) * (size_t(nnz) + 1), s));
        if (!pj || !pa) {
            set_error("allocation of P failed");
            st = AMGT_ERR_ALLOC;
        }
    }
    if (st == AMGT_OK) {
        hipLaunchKernelGGL(k_tri_fill, dim3(grid), dim3(kThreads), 0, s, rows, dg, nprob, pi, pj,
                           pa);
        if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
    }
    if (st != AMGT_OK) (void)hipStreamSynchronize(s);  // the host table may still be read
    if (cnt) tmp_free(cnt, s);
    tmp_free(dg, s);
    if (st != AMGT_OK) {
        if (st == AMGT_ERR_HIP) set_error("trilinear P: kernel launch or copy failed");
        if (pi) dev_free(pi, s);
        if (pj) dev_free(pj, s);
        if (pa) dev_free(pa, s);
        return st;
    }
    *n_fine = int32_t(rows);
    *n_coarse = int32_t(cols);
    *nnzP = nnz;
    *P_i = pi;
    *P_j = pj;
    *P_data = pa;
    return AMGT_OK;
}

}  // extern "C"
