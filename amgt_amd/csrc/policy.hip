// Per-level precision / strength-threshold policy network: a 12-64-64-4 MLP
// (tanh hidden layers; outputs = three logits for FP64 / FP32 / FP16 and one
// strength threshold), its forward pass, the gradient of its training loss and
// an Adam step, all on the device in FP32.
//
// The reference picks a level's precision with a fixed rule
// (get_calculationPrecison, csr_spgemm_device.c:1847-1862: level 0 FP64, 1
// FP32, >= 2 FP16) and uses one strength threshold for every level; the
// network replaces that rule with per-level predictions from the level's
// matrix statistics.
//
// Layout of the parameter vector (FP32, kParams entries, row-major weights
// W[out][in]): W1[64][12] b1[64] W2[64][64] b2[64] W3[4][64] b3[4].
// Loss per sample s (mean over the batch):
//   -log softmax(o[0..2])[cls_s] + lambda * (o[3] - target_s)^2.
// Gradients over the batch are summed in fixed chunks of samples with FP64
// partials and the partials added in chunk order: the result does not depend
// on scheduling (no atomics).
#include "internal.hpp"

namespace amgt {

namespace {

constexpr int kIn = AMGT_POLICY_FEATURES;
constexpr int kHid = AMGT_POLICY_HIDDEN;
constexpr int kOut = AMGT_POLICY_OUTPUTS;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kIn;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kOut * kHid;
constexpr int kParams = kOffB3 + kOut;
static_assert(kParams == AMGT_POLICY_PARAMS, "parameter layout");

constexpr int kThreads = 256;
constexpr int kChunk = 512;                          // samples per gradient partial
constexpr int kMaxLayerParams = kHid * (kHid + 1);   // largest W + b of one layer

inline unsigned blocks_of(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }
inline int64_t chunks_of(int64_t n) { return (n + kChunk - 1) / kChunk; }

// work buffer layout, in floats then doubles (8-byte aligned)
struct Work {
    float *h1, *h2, *o, *dout, *dz2, *dz1, *lossp;
    double *part;
};

inline size_t floats_of(int64_t n) { return size_t(n) * (2 * kHid + 2 * kOut + 2 * kHid + 1); }

inline size_t work_bytes(int64_t n) {
    const size_t f = (floats_of(n) * sizeof(float) + 7) / 8 * 8;
    return f + size_t(chunks_of(n)) * kMaxLayerParams * sizeof(double);
}

inline Work carve(void *work, int64_t n) {
    Work w;
    float *p = static_cast<float *>(work);
    w.h1 = p;
    w.h2 = w.h1 + n * kHid;
    w.o = w.h2 + n * kHid;
    w.dout = w.o + n * kOut;
    w.dz2 = w.dout + n * kOut;
    w.dz1 = w.dz2 + n * kHid;
    w.lossp = w.dz1 + n * kHid;
    const size_t f = (floats_of(n) * sizeof(float) + 7) / 8 * 8;
    w.part = reinterpret_cast<double *>(static_cast<char *>(work) + f);
    return w;
}

// y[s][o] = act(b[o] + sum_i x[s][i] w[o][i]), one thread per (s, o)
template <int kI, int kO, bool kTanh>
__global__ __launch_bounds__(kThreads) void k_dense_fwd(const float *__restrict__ x,
                                                       const float *__restrict__ w,
                                                       const float *__restrict__ b, int64_t n,
                                                       float *__restrict__ y) {
    const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (t >= n * kO) return;
    const int64_t s = t / kO;
    const int o = int(t - s * kO);
    const float *xs = x + s * kI;
    const float *wo = w + o * kI;
    float acc = b[o];
#pragma unroll 4
    for (int i = 0; i < kI; ++i) acc += xs[i] * wo[i];
    y[t] = kTanh ? tanhf(acc) : acc;
}

// softmax cross-entropy on o[0..2] and lambda-weighted squared error on o[3];
// dout = d(mean loss)/d o, lossp[s] = the sample's loss (NaN for a label
// outside 0..2, whose gradient is left zero)
__global__ __launch_bounds__(kThreads) void k_out_grad(const float *__restrict__ o,
                                                      const int32_t *__restrict__ cls,
                                                      const float *__restrict__ target, int64_t n,
                                                      float lambda, float inv_n,
                                                      float *__restrict__ dout,
                                                      float *__restrict__ lossp) {
    const int64_t s = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (s >= n) return;
    const float *os = o + s * kOut;
    float *ds = dout + s * kOut;
    const int y = cls[s];
    if (y < 0 || y > 2) {
        for (int c = 0; c < kOut; ++c) ds[c] = 0.0f;
        lossp[s] = __builtin_nanf("");
        return;
    }
    const float m = fmaxf(os[0], fmaxf(os[1], os[2]));
    float e[3], sum = 0.0f;
    for (int c = 0; c < 3; ++c) {
        e[c] = expf(os[c] - m);
        sum += e[c];
    }
    for (int c = 0; c < 3; ++c) ds[c] = (e[c] / sum - (c == y ? 1.0f : 0.0f)) * inv_n;
    const float r = os[3] - target[s];
    ds[3] = 2.0f * lambda * r * inv_n;
    lossp[s] = (logf(sum) + m - os[y]) + lambda * r * r;
}

// dprev[s][i] = (sum_o dz[s][o] w[o][i]) * (1 - h[s][i]^2): the gradient at the
// pre-activation of the previous tanh layer
template <int kI, int kO>
__global__ __launch_bounds__(kThreads) void k_dense_bwd_x(const float *__restrict__ dz,
                                                         const float *__restrict__ w,
                                                         const float *__restrict__ h, int64_t n,
                                                         float *__restrict__ dprev) {
    const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (t >= n * kI) return;
    const int64_t s = t / kI;
    const int i = int(t - s * kI);
    const float *dzs = dz + s * kO;
    float acc = 0.0f;
#pragma unroll 4
    for (int o = 0; o < kO; ++o) acc += dzs[o] * w[o * kI + i];
    const float hv = h[t];
    dprev[t] = acc * (1.0f - hv * hv);
}

// partial[c][o*(kI+1)+i] = sum over the samples of chunk c of dz[s][o] x[s][i]
// (i < kI) or of dz[s][o] (i == kI, the bias); grid (param blocks, chunks)
template <int kI, int kO>
__global__ __launch_bounds__(kThreads) void k_dense_bwd_w(const float *__restrict__ dz,
                                                         const float *__restrict__ x, int64_t n,
                                                         double *__restrict__ part) {
    constexpr int kP = kO * (kI + 1);
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= kP) return;
    const int o = p / (kI + 1);
    const int i = p - o * (kI + 1);
    const int64_t c = blockIdx.y;
    const int64_t s0 = c * kChunk;
    const int64_t s1 = s0 + kChunk < n ? s0 + kChunk : n;
    double acc = 0.0;
    if (i < kI) {
        for (int64_t s = s0; s < s1; ++s) acc += double(dz[s * kO + o]) * double(x[s * kI + i]);
    } else {
        for (int64_t s = s0; s < s1; ++s) acc += double(dz[s * kO + o]);
    }
    part[c * kP + p] = acc;
}

// grad[W] / grad[b] of one layer = the chunk partials added in chunk order
template <int kI, int kO>
__global__ __launch_bounds__(kThreads) void k_sum_parts(const double *__restrict__ part,
                                                       int64_t nchunks, float *__restrict__ gw,
                                                       float *__restrict__ gb) {
    constexpr int kP = kO * (kI + 1);
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= kP) return;
    double acc = 0.0;
    for (int64_t c = 0; c < nchunks; ++c) acc += part[c * kP + p];
    const int o = p / (kI + 1);
    const int i = p - o * (kI + 1);
    if (i < kI)
        gw[o * kI + i] = float(acc);
    else
        gb[o] = float(acc);
}

// *out = sum of v[0..n) / n, one workgroup, fixed order (thread-strided FP64
// sums, then a tree in LDS)
__global__ __launch_bounds__(kThreads) void k_mean(const float *__restrict__ v, int64_t n,
                                                  float *__restrict__ out) {
    __shared__ double s_acc[kThreads];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kThreads) acc += double(v[i]);
    s_acc[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s_acc[threadIdx.x] += s_acc[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = float(s_acc[0] / double(n));
}

__global__ __launch_bounds__(kThreads) void k_adam(float *__restrict__ p,
                                                  const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v,
                                                  int64_t np, float lr, float b1, float b2,
                                                  float eps, float bc1, float bc2) {
    const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (t >= np) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / bc1) / (sqrtf(vt / bc2) + eps);
}

// ---- level statistics (the matrix features the policy reads) -----------------
// per row: |a_ii| (diagonal found by column), s = sum_(j != i) |a_ij|, the row's
// count of positive off-diagonals and max |a_ij|; per workgroup the sums / min /
// max of the eight statistics of amgt_level_stats, reduced in LDS in a fixed
// order, then one workgroup adds the partials in workgroup order.
constexpr int kStats = 8;

__global__ __launch_bounds__(kThreads) void k_level_stats(int32_t n, const int32_t *__restrict__ I,
                                                         const int32_t *__restrict__ J,
                                                         const double *__restrict__ A,
                                                         double *__restrict__ part) {
    __shared__ double s_v[kStats][kThreads];
    const int i = blockIdx.x * kThreads + threadIdx.x;
    double v[kStats] = {0.0, 0.0, 0.0, 0.0, INFINITY, 0.0, 0.0, 0.0};
    if (i < n) {
        double d = 0.0, off = 0.0, amax = 0.0, npos = 0.0;
        for (int64_t p = I[i]; p < I[i + 1]; ++p) {
            const double a = A[p];
            const double aa = fabs(a);
            amax = fmax(amax, aa);
            if (J[p] == i) {
                d += aa;
            } else {
                off += aa;
                npos += a > 0.0 ? 1.0 : 0.0;
            }
        }
        const double dom = off > 0.0 ? fmin(d / off, 10.0) : 10.0;
        v[0] = dom;                              // sum of dominance ratios
        v[1] = dom < 1.0 - 1e-12 ? 1.0 : 0.0;    // rows not diagonally dominant
        v[2] = npos;                             // positive off-diagonals
        v[3] = amax;                             // max |a|
        v[4] = d > 0.0 ? d : INFINITY;           // min positive |a_ii|
        const double l = d > 0.0 ? log10(d) : 0.0;
        v[5] = l;                                // sum log10 |a_ii| (rows with a_ii != 0)
        v[6] = l * l;                            // sum of squares
        v[7] = d > 0.0 ? 1.0 : 0.0;              // rows with a_ii != 0
    }
    for (int k = 0; k < kStats; ++k) s_v[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            for (int k = 0; k < kStats; ++k) {
                const double a = s_v[k][threadIdx.x], b = s_v[k][threadIdx.x + w];
                s_v[k][threadIdx.x] = k == 3 ? fmax(a, b) : (k == 4 ? fmin(a, b) : a + b);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < kStats) part[size_t(blockIdx.x) * kStats + threadIdx.x] = s_v[threadIdx.x][0];
}

// one workgroup: thread t combines partials t, t + kThreads, ... in order,
// then the same fixed tree as k_level_stats (deterministic)
__global__ __launch_bounds__(kThreads) void k_level_stats_final(int64_t nparts,
                                                               const double *__restrict__ part,
                                                               double *__restrict__ out) {
    __shared__ double s_v[kStats][kThreads];
    double v[kStats] = {0.0, 0.0, 0.0, 0.0, INFINITY, 0.0, 0.0, 0.0};
    for (int64_t b = threadIdx.x; b < nparts; b += kThreads) {
        for (int k = 0; k < kStats; ++k) {
            const double x = part[b * kStats + k];
            v[k] = k == 3 ? fmax(v[k], x) : (k == 4 ? fmin(v[k], x) : v[k] + x);
        }
    }
    for (int k = 0; k < kStats; ++k) s_v[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            for (int k = 0; k < kStats; ++k) {
                const double a = s_v[k][threadIdx.x], b = s_v[k][threadIdx.x + w];
                s_v[k][threadIdx.x] = k == 3 ? fmax(a, b) : (k == 4 ? fmin(a, b) : a + b);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < kStats) out[threadIdx.x] = s_v[threadIdx.x][0];
}

int forward_into(const float *params, const float *x, int64_t n, const Work &w, float *out,
                 hipStream_t s) {
    k_dense_fwd<kIn, kHid, true><<<blocks_of(n * kHid), kThreads, 0, s>>>(
        x, params + kOffW1, params + kOffB1, n, w.h1);
    k_dense_fwd<kHid, kHid, true><<<blocks_of(n * kHid), kThreads, 0, s>>>(
        w.h1, params + kOffW2, params + kOffB2, n, w.h2);
    k_dense_fwd<kHid, kOut, false><<<blocks_of(n * kOut), kThreads, 0, s>>>(
        w.h2, params + kOffW3, params + kOffB3, n, out);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

template <int kI, int kO>
int layer_grad(const float *dz, const float *x, int64_t n, double *part, float *gw, float *gb,
               hipStream_t s) {
    constexpr int kP = kO * (kI + 1);
    const int64_t nc = chunks_of(n);
    k_dense_bwd_w<kI, kO><<<dim3(blocks_of(kP), unsigned(nc)), kThreads, 0, s>>>(dz, x, n, part);
    k_sum_parts<kI, kO><<<blocks_of(kP), kThreads, 0, s>>>(part, nc, gw, gb);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

}  // namespace

}  // namespace amgt

using namespace amgt;

extern "C" {

size_t amgt_policy_work_bytes(int64_t nsamples) {
    return nsamples > 0 ? work_bytes(nsamples) : 0;
}

int amgt_policy_forward(const float *params, const float *features, int64_t n, float *out,
                        void *work, size_t work_size, void *stream) {
    AMGT_CHECK_ARG(n >= 0, "negative sample count");
    if (n == 0) return AMGT_OK;
    AMGT_CHECK_ARG(params && features && out && work, "NULL argument");
    AMGT_CHECK_ARG(n <= (int64_t(1) << 24), "more than 2^24 samples");
    AMGT_CHECK_ARG(work_size >= work_bytes(n), "work buffer smaller than amgt_policy_work_bytes");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Work w = carve(work, n);
    return forward_into(params, features, n, w, out, s);
}

int amgt_policy_loss_grad(const float *params, const float *features, const int32_t *cls,
                          const float *target, int64_t n, float lambda, float *grad, float *loss,
                          void *work, size_t work_size, void *stream) {
    AMGT_CHECK_ARG(n > 0, "empty batch");
    AMGT_CHECK_ARG(params && features && cls && target && grad && loss && work, "NULL argument");
    AMGT_CHECK_ARG(n <= (int64_t(1) << 24), "more than 2^24 samples");
    AMGT_CHECK_ARG(work_size >= work_bytes(n), "work buffer smaller than amgt_policy_work_bytes");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Work w = carve(work, n);
    AMGT_TRY(forward_into(params, features, n, w, w.o, s));
    k_out_grad<<<blocks_of(n), kThreads, 0, s>>>(w.o, cls, target, n, lambda, 1.0f / float(n),
                                                 w.dout, w.lossp);
    k_mean<<<1, kThreads, 0, s>>>(w.lossp, n, loss);
    AMGT_HIP(hipGetLastError());
    // output layer, then back through the two tanh layers
    AMGT_TRY((layer_grad<kHid, kOut>(w.dout, w.h2, n, w.part, grad + kOffW3, grad + kOffB3, s)));
    k_dense_bwd_x<kHid, kOut><<<blocks_of(n * kHid), kThreads, 0, s>>>(w.dout, params + kOffW3,
                                                                      w.h2, n, w.dz2);
    AMGT_TRY((layer_grad<kHid, kHid>(w.dz2, w.h1, n, w.part, grad + kOffW2, grad + kOffB2, s)));
    k_dense_bwd_x<kHid, kHid><<<blocks_of(n * kHid), kThreads, 0, s>>>(w.dz2, params + kOffW2,
                                                                      w.h1, n, w.dz1);
    AMGT_TRY((layer_grad<kIn, kHid>(w.dz1, features, n, w.part, grad + kOffW1, grad + kOffB1, s)));
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

size_t amgt_level_stats_work_bytes(int32_t nrows) {
    return nrows > 0 ? size_t(blocks_of(nrows)) * kStats * sizeof(double) + kStats * sizeof(double)
                     : 0;
}

int amgt_level_stats(int32_t n, const int32_t *A_i, const int32_t *A_j, const double *A_data,
                     double *stats, void *work, size_t work_size, void *stream) {
    AMGT_CHECK_ARG(n > 0, "empty matrix");
    AMGT_CHECK_ARG(A_i && A_j && A_data && stats && work, "NULL argument");
    AMGT_CHECK_ARG(work_size >= amgt_level_stats_work_bytes(n),
                   "work buffer smaller than amgt_level_stats_work_bytes");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned nb = blocks_of(n);
    double *part = static_cast<double *>(work);
    double *dev_out = part + size_t(nb) * kStats;
    k_level_stats<<<nb, kThreads, 0, s>>>(n, A_i, A_j, A_data, part);
    k_level_stats_final<<<1, kThreads, 0, s>>>(nb, part, dev_out);
    AMGT_HIP(hipGetLastError());
    const Readback rb[1] = {{stats, dev_out, kStats * sizeof(double)}};
    return readback(rb, 1, s);
}

int amgt_policy_adam(float *params, const float *grad, float *m, float *v, int64_t nparams,
                     int64_t step, float lr, float beta1, float beta2, float eps, void *stream) {
    AMGT_CHECK_ARG(nparams >= 0 && step >= 1, "negative size or step < 1");
    if (nparams == 0) return AMGT_OK;
    AMGT_CHECK_ARG(params && grad && m && v, "NULL argument");
    AMGT_CHECK_ARG(beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f && eps > 0.0f,
                   "Adam moments must lie in [0, 1) and eps > 0");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const float bc1 = float(1.0 - std::pow(double(beta1), double(step)));
    const float bc2 = float(1.0 - std::pow(double(beta2), double(step)));
    k_adam<<<blocks_of(nparams), kThreads, 0, s>>>(params, grad, m, v, nparams, lr, beta1, beta2,
                                                   eps, bc1, bc2);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

}  // extern "C"
