// Vector kernels of preconditioned CG (hypre_PCGSolve, krylov/pcg.c:320-900,
// two-norm form) for the solve driver: deterministic dot products and the
// fused updates, with every scalar (alpha, beta, the inner products) kept in
// device memory so that an iteration is a fixed sequence of launches that can
// be captured in a graph.  Dots: one partial per workgroup over a fixed
// grid-stride range, then one workgroup sums the partials in order: the same
// bits on every call, whatever the timing.
#include "internal.hpp"

namespace amgt {

namespace {

constexpr int kThreads = 256;
constexpr int kParts = 1024;  // workgroups of the first stage (partials)

__device__ __forceinline__ double block_sum(double v, double *s_w) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) s_w[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kThreads / 64; ++k) t += s_w[k];
    return t;  // valid on thread 0
}

// partial[wg] = sum over this workgroup's grid-stride range of x[i] * y[i]
__global__ __launch_bounds__(kThreads) void k_dot_part(int64_t n, const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       double *__restrict__ part) {
    __shared__ double s_w[kThreads / 64];
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kThreads)
        v = fma(x[i], y[i], v);
    const double t = block_sum(v, s_w);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out[0] = sum of the partials in order (one workgroup)
__global__ __launch_bounds__(kThreads) void k_dot_final(int nparts, const double *__restrict__ part,
                                                        double *__restrict__ out) {
    __shared__ double s_w[kThreads / 64];
    double v = 0.0;
    for (int k = threadIdx.x; k < nparts; k += kThreads) v += part[k];
    const double t = block_sum(v, s_w);
    if (threadIdx.x == 0) out[0] = t;
}

// alpha = rz / pAp; x += alpha p; r -= alpha Ap; partials of r.r
__global__ __launch_bounds__(kThreads) void k_update_xr(int64_t n, const double *__restrict__ rz,
                                                        const double *__restrict__ pap,
                                                        const double *__restrict__ p,
                                                        const double *__restrict__ ap,
                                                        double *__restrict__ x,
                                                        double *__restrict__ r,
                                                        double *__restrict__ part) {
    __shared__ double s_w[kThreads / 64];
    const double alpha = rz[0] / pap[0];
    double v = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kThreads) {
        x[i] = fma(alpha, p[i], x[i]);
        const double ri = fma(-alpha, ap[i], r[i]);
        r[i] = ri;
        v = fma(ri, ri, v);
    }
    const double t = block_sum(v, s_w);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// beta = rz_new / rz; p = z + beta p.  rz = rz_new follows as its own launch
// (k_copy_scalar): every workgroup reads rz here.
__global__ __launch_bounds__(kThreads) void k_update_p(int64_t n, const double *__restrict__ rz_new,
                                                       const double *__restrict__ rz,
                                                       const double *__restrict__ z,
                                                       double *__restrict__ p) {
    const double beta = rz_new[0] / rz[0];
    for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * kThreads)
        p[i] = fma(beta, p[i], z[i]);
}

__global__ void k_copy_scalar(const double *__restrict__ src, double *__restrict__ dst) {
    if (threadIdx.x == 0) dst[0] = src[0];
}

unsigned grid_for(int64_t n) {
    const int64_t g = (n + kThreads - 1) / kThreads;
    return unsigned(g < kParts ? (g > 0 ? g : 1) : kParts);
}

int dot(int64_t n, const double *x, const double *y, double *out, double *part, hipStream_t s) {
    const unsigned g = grid_for(n);
    hipLaunchKernelGGL(k_dot_part, dim3(g), dim3(kThreads), 0, s, n, x, y, part);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kThreads), 0, s, int(g), part, out);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

}  // namespace

}  // namespace amgt

using namespace amgt;

extern "C" {

size_t amgt_pcg_work_bytes(void) { return sizeof(double) * kParts; }

int amgt_dot(int64_t n, const double *d_x, const double *d_y, double *d_out, void *d_work,
             void *stream) {
    AMGT_CHECK_ARG(n >= 0 && d_out && d_work, "NULL argument or negative size");
    AMGT_CHECK_ARG(n == 0 || (d_x && d_y), "NULL vector");
    return dot(n, d_x, d_y, d_out, static_cast<double *>(d_work), static_cast<hipStream_t>(stream));
}

int amgt_pcg_update_xr(int64_t n, const double *d_rz, const double *d_pap, const double *d_p,
                       const double *d_ap, double *d_x, double *d_r, double *d_rr, void *d_work,
                       void *stream) {
    AMGT_CHECK_ARG(n >= 0 && d_rz && d_pap && d_rr && d_work, "NULL argument or negative size");
    AMGT_CHECK_ARG(n == 0 || (d_p && d_ap && d_x && d_r), "NULL vector");
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *part = static_cast<double *>(d_work);
    const unsigned g = grid_for(n);
    hipLaunchKernelGGL(k_update_xr, dim3(g), dim3(kThreads), 0, s, n, d_rz, d_pap, d_p, d_ap, d_x,
                       d_r, part);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kThreads), 0, s, int(g), part, d_rr);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

int amgt_pcg_update_p(int64_t n, const double *d_rz_new, double *d_rz, const double *d_z,
                      double *d_p, void *stream) {
    AMGT_CHECK_ARG(n >= 0 && d_rz_new && d_rz, "NULL argument or negative size");
    AMGT_CHECK_ARG(n == 0 || (d_z && d_p), "NULL vector");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_update_p, dim3(grid_for(n)), dim3(kThreads), 0, s, n, d_rz_new, d_rz,
                       d_z, d_p);
    hipLaunchKernelGGL(k_copy_scalar, dim3(1), dim3(64), 0, s, d_rz_new, d_rz);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

}  // extern "C"
