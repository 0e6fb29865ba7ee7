// Row passes without same-address atomics.
//
// A pass looks at one integer per row (entries of a block row, the SpGEMM
// upper bound of a C row, blocks of an mBSR row ...).  Every workgroup of 256
// rows writes one partial: how many of its rows fall in each of three bins
// (v <= t0, v <= t1, larger), the largest value and the sum of squares.
// k_pass_finish (one workgroup) turns the partials into per-workgroup list
// offsets, bin by bin, and totals; k_pass_scatter then writes the row lists,
// rows ascending within a bin, so lists are deterministic.
//
// Counters updated with one atomic per wave on a single address serialise at
// roughly 12 ns per atomic on this part (measured: 8192 wave atomics took
// ~95 us), which is why no pass here uses them.
#pragma once

#include "internal.hpp"

#include <climits>

namespace amgt {
namespace rowpass {

struct Part {
    int32_t cnt[3];  // rows per bin; after k_pass_finish: this workgroup's list offsets
    int32_t mx;
    unsigned long long sumsq;
};

using Totals = RowStats;  // cnt[3], max_blocks (the largest value), sumsq

constexpr int kThreads = 256;

__device__ __forceinline__ int bin3(int64_t v, int64_t t0, int64_t t1) {
    return v <= t0 ? 0 : (v <= t1 ? 1 : 2);
}

// Every thread of the (256-thread) workgroup calls this once with its row:
// vb picks the bin, v feeds the largest value and the sum of squares.
__device__ __forceinline__ void write_part(bool valid, int64_t vb, int64_t v, int64_t t0,
                                           int64_t t1, Part *parts) {
    __shared__ int s_cnt[4][3];
    __shared__ int s_mx[4];
    __shared__ unsigned long long s_sq[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b = valid ? bin3(vb, t0, t1) : -1;
    int c[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) c[q] = __popcll(__ballot(b == q));
    int mx = valid ? int(v > INT_MAX ? INT_MAX : (v < 0 ? 0 : v)) : 0;
    unsigned long long sq =
        valid ? static_cast<unsigned long long>(v) * static_cast<unsigned long long>(v) : 0ull;
    for (int off = 32; off > 0; off >>= 1) {
        mx = max(mx, __shfl_xor(mx, off));
        sq += __shfl_xor(sq, off);
    }
    if (lane == 0) {
        s_cnt[w][0] = c[0];
        s_cnt[w][1] = c[1];
        s_cnt[w][2] = c[2];
        s_mx[w] = mx;
        s_sq[w] = sq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Part p;
#pragma unroll
        for (int q = 0; q < 3; ++q) p.cnt[q] = s_cnt[0][q] + s_cnt[1][q] + s_cnt[2][q] + s_cnt[3][q];
        p.mx = max(max(s_mx[0], s_mx[1]), max(s_mx[2], s_mx[3]));
        p.sumsq = s_sq[0] + s_sq[1] + s_sq[2] + s_sq[3];
        parts[blockIdx.x] = p;
    }
}

// fb(i) picks the bin, fs(i) feeds the largest value and the sum of squares
template <typename FB, typename FS>
__global__ __launch_bounds__(256) void k_pass_count(int32_t n, FB fb, FS fs, int64_t t0,
                                                    int64_t t1, Part *parts) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    const bool valid = i < n;
    write_part(valid, valid ? fb(i) : 0, valid ? fs(i) : 0, t0, t1, parts);
}

// One workgroup of kFinishThreads: partial counts -> list offsets per
// workgroup and bin; totals.  Each pass scans kFinishThreads partials (DPP
// scans within the waves, then the wave totals from LDS).
constexpr int kFinishThreads = 1024;
static __global__ __launch_bounds__(kFinishThreads) void k_pass_finish(int32_t ng, Part *parts,
                                                                       Totals *tot) {
    constexpr int NW = kFinishThreads / 64;
    __shared__ int s_w[NW][3];
    __shared__ int s_mx[NW];
    __shared__ unsigned long long s_sq[NW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int run[3] = {0, 0, 0};
    int mx = 0;
    unsigned long long sq = 0ull;
    for (int base = 0; base < ng; base += kFinishThreads) {
        const int g = base + threadIdx.x;
        Part p{};
        if (g < ng) p = parts[g];
        int incl[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) incl[q] = wave_incl_sum(p.cnt[q]);
        if (lane == 63) {
#pragma unroll
            for (int q = 0; q < 3; ++q) s_w[w][q] = incl[q];
        }
        __syncthreads();
        int off[3], all[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            int before = 0, sum = 0;
            for (int k = 0; k < NW; ++k) {
                const int c = s_w[k][q];
                before += k < w ? c : 0;
                sum += c;
            }
            off[q] = run[q] + before + incl[q] - p.cnt[q];
            all[q] = sum;
        }
        __syncthreads();  // s_w is rewritten by the next pass
        if (g < ng) {
            parts[g].cnt[0] = off[0];
            parts[g].cnt[1] = off[1];
            parts[g].cnt[2] = off[2];
            mx = max(mx, p.mx);
            sq += p.sumsq;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) run[q] += all[q];
    }
    for (int o = 32; o > 0; o >>= 1) {
        mx = max(mx, __shfl_xor(mx, o));
        sq += __shfl_xor(sq, o);
    }
    if (lane == 0) {
        s_mx[w] = mx;
        s_sq[w] = sq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = 0;
        unsigned long long t = 0ull;
        for (int k = 0; k < NW; ++k) {
            m = max(m, s_mx[k]);
            t += s_sq[k];
        }
        tot->cnt[0] = run[0];
        tot->cnt[1] = run[1];
        tot->cnt[2] = run[2];
        tot->max_blocks = m;
        tot->sumsq = t;
    }
}

// Writes list[b * stride + k] = row for the k-th row (ascending) of bin b, for
// the bins whose bit is set in bin_mask.
template <typename F>
__global__ __launch_bounds__(256) void k_pass_scatter(int32_t n, F f, int64_t t0, int64_t t1,
                                                      const Part *__restrict__ parts,
                                                      int32_t *__restrict__ list, int64_t stride,
                                                      unsigned bin_mask) {
    __shared__ int s_wc[4][3];
    const int i = blockIdx.x * kThreads + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b = (i < n) ? bin3(f(i), t0, t1) : -1;
    uint64_t m[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) m[q] = __ballot(b == q);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) s_wc[w][q] = __popcll(m[q]);
    }
    __syncthreads();
    if (b >= 0 && ((bin_mask >> b) & 1u)) {
        int pos = parts[blockIdx.x].cnt[b];
        for (int q = 0; q < w; ++q) pos += s_wc[q][b];
        const uint64_t lt = (1ull << lane) - 1ull;
        pos += __popcll(m[b] & lt);
        list[int64_t(b) * stride + pos] = i;
    }
}

// ---- value functors ----
struct BlockDiff {  // ptr[i+1] - ptr[i]
    const int32_t *ptr;
    __device__ int64_t operator()(int i) const { return ptr[i + 1] - ptr[i]; }
};
struct ArrayI32 {
    const int32_t *p;
    __device__ int64_t operator()(int i) const { return p[i]; }
};
struct ArrayI64 {
    const int64_t *p;
    __device__ int64_t operator()(int i) const { return p[i]; }
};

inline int32_t groups(int32_t n) { return (n + kThreads - 1) / kThreads; }

// count + finish; parts needs groups(n) entries, *tot is written on the device
template <typename FB, typename FS>
int count(int32_t n, FB fb, FS fs, int64_t t0, int64_t t1, Part *parts, Totals *tot,
          hipStream_t s) {
    const int32_t ng = groups(n);
    if (ng > 0)
        hipLaunchKernelGGL((k_pass_count<FB, FS>), dim3(unsigned(ng)), dim3(kThreads), 0, s, n, fb,
                           fs, t0, t1, parts);
    hipLaunchKernelGGL(k_pass_finish, dim3(1), dim3(kFinishThreads), 0, s, ng, parts, tot);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

template <typename F>
int count(int32_t n, F f, int64_t t0, int64_t t1, Part *parts, Totals *tot, hipStream_t s) {
    return count(n, f, f, t0, t1, parts, tot, s);
}

template <typename F>
int scatter(int32_t n, F f, int64_t t0, int64_t t1, const Part *parts, int32_t *list,
            int64_t stride, unsigned bin_mask, hipStream_t s) {
    const int32_t ng = groups(n);
    if (ng > 0)
        hipLaunchKernelGGL((k_pass_scatter<F>), dim3(unsigned(ng)), dim3(kThreads), 0, s, n, f,
                           t0, t1, parts, list, stride, bin_mask);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

}  // namespace rowpass
}  // namespace amgt
