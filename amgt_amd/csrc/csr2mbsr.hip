// CSR -> mBSR conversion (counterpart of CSR2BSR_GPU, csr_matvec_device.c:1201-1267).
//
// Design: a block row's entries (the up-to-4 CSR rows 4br..4br+3) are sorted by
// the 32-bit key (block column << 4 | r*4 + c).  Sorted order IS the compressed
// value order of the mBSR layout (blocks ascending, bits ascending), so values
// land at val[I[4br] + rank] directly, block columns come from the key heads
// and bitmaps from OR-ing the low nibble of keys that share a head.  Sorting
// makes the conversion independent of the CSR column order (HYPRE rows start
// with the diagonal), unlike the reference's windowed bitmask that assumes
// sorted columns once ncols > 32768 (csr_matvec_device.c:701-722).
//
// Bins by entries per block row: <= 64 (one wave, register bitonic sort over
// lanes), <= 1024 (256-thread workgroup, LDS bitonic), <= 8192 (1024-thread
// workgroup, LDS bitonic).  Larger block rows return AMGT_ERR_UNSUPPORTED.
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace amgt {

namespace {

constexpr uint32_t kPadKey = 0xFFFFFFFFu;

enum : int { kErrDup = 1, kErrCol = 2 };

// Per-conversion device scratch counters.
struct ConvCounters {
    int32_t bin_count[4];
    int32_t max_entries;
    int32_t err;
    int32_t max_blocks;
    int32_t pad;
    double sq_dev;
};

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_zero_counters(ConvCounters *c) {
    if (threadIdx.x == 0) {
        for (int i = 0; i < 4; ++i) c->bin_count[i] = 0;
        c->max_entries = 0;
        c->err = 0;
        c->max_blocks = 0;
        c->pad = 0;
        c->sq_dev = 0.0;
    }
}

__device__ __forceinline__ int bin_of(int n) {
    return n <= 64 ? 0 : (n <= 1024 ? 1 : (n <= 8192 ? 2 : 3));
}

// Bins block rows by entry count; list[bin*brow + t] = block row.
// One atomic per wave and bin (wave-aggregated): a single hot counter would
// otherwise serialise every block row.
__global__ __launch_bounds__(256) void k_bin_rows(int32_t nrows, int32_t brow,
                                                  const int32_t *__restrict__ I,
                                                  int32_t *__restrict__ list, ConvCounters *cnt) {
    const int br = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = br < brow;
    int n = 0, b = -1;
    if (valid) {
        const int r0 = 4 * br;
        const int r1 = min(r0 + 4, nrows);
        n = I[r1] - I[r0];
        b = bin_of(n);
    }
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint64_t mask = __ballot(b == q);
        if (mask == 0ull) continue;
        const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&cnt->bin_count[q], __popcll(mask));
        base = __shfl(base, leader);
        if (b == q) list[int64_t(q) * brow + base + __popcll(mask & lt)] = br;
    }
    int mx = n;
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
    if (lane == 0 && mx > 0) atomicMax(&cnt->max_entries, mx);
}

// Key of CSR entry p in block row starting at row r0.
__device__ __forceinline__ uint32_t entry_key(int p, int r0, int nrows, int ncols,
                                              const int32_t *__restrict__ I,
                                              const int32_t *__restrict__ J, int *err) {
    int c = J[p];
    if (c < 0 || c >= ncols) {
        *err |= kErrCol;
        c = 0;
    }
    int r = (p >= I[min(r0 + 1, nrows)]) + (p >= I[min(r0 + 2, nrows)]) +
            (p >= I[min(r0 + 3, nrows)]);
    return (uint32_t(c >> 2) << 4) | uint32_t(r << 2) | uint32_t(c & 3);
}

// ---- bin 0: one wave per block row, <= 64 entries, bitonic sort across lanes.
// tmp_col/tmp_map receive the row's blocks at offset I[4br] (upper-bound slot).
__global__ __launch_bounds__(256) void k_sort_wave(
    int32_t nrows, int32_t ncols, const int32_t *__restrict__ I, const int32_t *__restrict__ J,
    const double *__restrict__ A, const int32_t *__restrict__ list, int32_t count,
    double *__restrict__ val, int32_t *__restrict__ tmp_col, uint32_t *__restrict__ tmp_map,
    int32_t *__restrict__ blk_cnt, ConvCounters *cnt, int values_only) {
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= count) return;  // whole wave exits together
    const int br = list[w];
    const int r0 = 4 * br;
    const int s = I[r0];
    const int e = I[min(r0 + 4, nrows)];
    const int n = e - s;
    int err = 0;
    uint32_t key = kPadKey;
    int src = 0;
    if (lane < n) {
        src = s + lane;
        key = entry_key(src, r0, nrows, ncols, I, J, &err);
    }
    // bitonic sort, ascending by key, across the 64 lanes
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            uint32_t pk = __shfl_xor(key, j);
            int ps = __shfl_xor(src, j);
            bool up = (lane & k) == 0;
            bool lower = (lane & j) == 0;
            bool take = (lower == up) ? (pk < key) : (pk > key);
            if (take) {
                key = pk;
                src = ps;
            }
        }
    }
    uint32_t prev = __shfl_up(key, 1);
    const bool valid = lane < n;
    if (valid && lane > 0 && prev == key) err |= kErrDup;
    if (A && valid) val[s + lane] = A[src];
    if (!values_only) {
        const bool head = valid && (lane == 0 || (prev >> 4) != (key >> 4));
        const uint64_t heads = __ballot(head);
        // block rank of this lane's entry
        const uint64_t le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
        const int rank = __popcll(heads & le) - 1;
        // OR of the bits of the block: walk forward from each head (<= 16 entries)
        uint32_t map = 0;
#pragma unroll
        for (int d = 0; d < 16; ++d) {
            uint32_t kk = __shfl(key, min(lane + d, 63));
            bool same = (lane + d < n) && ((kk >> 4) == (key >> 4));
            // stop at the first entry of another block (keys are sorted)
            if (same) map |= 1u << (kk & 15u);
        }
        if (head) {
            tmp_col[s + rank] = int32_t(key >> 4);
            tmp_map[s + rank] = map;
        }
        if (lane == 0) {
            int nb = __popcll(heads);
            blk_cnt[br] = nb;
            atomicMax(&cnt->max_blocks, nb);
        }
    }
    if (err) atomicOr(&cnt->err, err);
}

// ---- bins 1/2: one workgroup per block row, LDS bitonic sort of CAP keys.
template <int THREADS, int CAP>
__global__ __launch_bounds__(THREADS) void k_sort_block(
    int32_t nrows, int32_t ncols, const int32_t *__restrict__ I, const int32_t *__restrict__ J,
    const double *__restrict__ A, const int32_t *__restrict__ list, int32_t count,
    double *__restrict__ val, int32_t *__restrict__ tmp_col, uint32_t *__restrict__ tmp_map,
    int32_t *__restrict__ blk_cnt, ConvCounters *cnt, int values_only) {
    __shared__ uint32_t s_key[CAP];
    __shared__ int32_t s_aux[CAP];  // source index, then block rank
    __shared__ int32_t s_wave_tot[THREADS / 64];
    __shared__ int s_err;
    const int t = threadIdx.x;
    const int w = blockIdx.x;
    if (w >= count) return;
    const int br = list[w];
    const int r0 = 4 * br;
    const int s = I[r0];
    const int e = I[min(r0 + 4, nrows)];
    const int n = e - s;
    int np = 64;
    while (np < n) np <<= 1;
    if (t == 0) s_err = 0;
    int err = 0;
    for (int i = t; i < np; i += THREADS) {
        if (i < n) {
            s_key[i] = entry_key(s + i, r0, nrows, ncols, I, J, &err);
            s_aux[i] = s + i;
        } else {
            s_key[i] = kPadKey;
            s_aux[i] = 0;
        }
    }
    __syncthreads();
    for (int k = 2; k <= np; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < np; i += THREADS) {
                int p = i ^ j;
                if (p > i) {
                    bool up = (i & k) == 0;
                    uint32_t a = s_key[i], b = s_key[p];
                    if ((a > b) == up) {
                        s_key[i] = b;
                        s_key[p] = a;
                        int x = s_aux[i];
                        s_aux[i] = s_aux[p];
                        s_aux[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = t; i < n; i += THREADS) {
        if (A) val[s + i] = A[s_aux[i]];
        if (i > 0 && s_key[i] == s_key[i - 1]) err |= kErrDup;
    }
    if (err) atomicOr(&s_err, err);
    __syncthreads();
    if (!values_only) {
        // block ranks: inclusive scan of head flags over the sorted keys,
        // processed in chunks of THREADS with a running carry
        int carry = 0;
        const int lane = t & 63, wv = t >> 6;
        for (int base = 0; base < n; base += THREADS) {
            int i = base + t;
            int head = 0;
            if (i < n) head = (i == 0) || ((s_key[i] >> 4) != (s_key[i - 1] >> 4));
            uint64_t bal = __ballot(head);
            uint64_t le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
            int incl = __popcll(bal & le);
            if (lane == 63) s_wave_tot[wv] = __popcll(bal);
            __syncthreads();
            int before = 0, total = 0;
            for (int q = 0; q < THREADS / 64; ++q) {
                int v = s_wave_tot[q];
                if (q < wv) before += v;
                total += v;
            }
            if (i < n) s_aux[i] = carry + before + incl - 1;  // block rank
            carry += total;
            __syncthreads();
        }
        // heads write block column and bitmap (<= 16 entries per block)
        for (int i = t; i < n; i += THREADS) {
            uint32_t key = s_key[i];
            bool head = (i == 0) || ((s_key[i - 1] >> 4) != (key >> 4));
            if (head) {
                uint32_t map = 0;
                for (int d = 0; d < 16 && i + d < n; ++d) {
                    uint32_t kk = s_key[i + d];
                    if ((kk >> 4) != (key >> 4)) break;
                    map |= 1u << (kk & 15u);
                }
                int rank = s_aux[i];
                tmp_col[s + rank] = int32_t(key >> 4);
                tmp_map[s + rank] = map;
            }
        }
        if (t == 0) {
            blk_cnt[br] = carry;
            atomicMax(&cnt->max_blocks, carry);
        }
    }
    if (t == 0 && s_err) atomicOr(&cnt->err, s_err);
}

// Copies each block row's blocks from the upper-bound slot to the compact
// arrays; one wave per block row.
__global__ __launch_bounds__(256) void k_compact(int32_t nrows, int32_t brow,
                                                 const int32_t *__restrict__ I,
                                                 const int32_t *__restrict__ blk_ptr,
                                                 const int32_t *__restrict__ tmp_col,
                                                 const uint32_t *__restrict__ tmp_map,
                                                 int32_t *__restrict__ blk_col,
                                                 uint16_t *__restrict__ blk_map,
                                                 int32_t *__restrict__ val_ptr) {
    const int lane = threadIdx.x & 63;
    const int br = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (br > brow) return;
    if (br == brow) {
        if (lane == 0) val_ptr[brow] = I[nrows];
        return;
    }
    const int s = I[4 * br];
    const int p0 = blk_ptr[br], p1 = blk_ptr[br + 1];
    for (int t = lane; t < p1 - p0; t += 64) {
        blk_col[p0 + t] = tmp_col[s + t];
        blk_map[p0 + t] = uint16_t(tmp_map[s + t]);
    }
    if (lane == 0) val_ptr[br] = s;
}

// Sum of squared deviations of blocks per block row (getStand,
// csr_matvec_device.c:638-671).
__global__ __launch_bounds__(256) void k_stand(int32_t brow, const int32_t *__restrict__ blk_cnt,
                                               double avg, ConvCounters *cnt) {
    __shared__ double s_part[4];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0.0;
    if (i < brow) {
        double v = double(blk_cnt[i]) - avg;
        d = v * v;
    }
    for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&cnt->sq_dev, s_part[0] + s_part[1] + s_part[2] + s_part[3]);
}

int launch_sorts(int32_t nrows, int32_t ncols, const int32_t *I, const int32_t *J, const double *A,
                 const int32_t *list, int32_t brow, const int32_t bin_count[4], double *val,
                 int32_t *tmp_col, uint32_t *tmp_map, int32_t *blk_cnt, ConvCounters *cnt,
                 int values_only, hipStream_t s) {
    if (bin_count[3] > 0) {
        set_error("CSR->mBSR: a block row has more than 8192 entries (4 rows with > 2048 "
                  "nonzeros on average); this build does not convert such rows");
        return AMGT_ERR_UNSUPPORTED;
    }
    if (bin_count[0] > 0) {
        int grid = (bin_count[0] + 3) / 4;
        hipLaunchKernelGGL(k_sort_wave, dim3(grid), dim3(256), 0, s, nrows, ncols, I, J, A,
                           list, bin_count[0], val, tmp_col, tmp_map, blk_cnt, cnt, values_only);
        AMGT_HIP(hipGetLastError());
    }
    if (bin_count[1] > 0) {
        hipLaunchKernelGGL((k_sort_block<256, 1024>), dim3(bin_count[1]), dim3(256), 0, s, nrows,
                           ncols, I, J, A, list + brow, bin_count[1], val, tmp_col, tmp_map,
                           blk_cnt, cnt, values_only);
        AMGT_HIP(hipGetLastError());
    }
    if (bin_count[2] > 0) {
        hipLaunchKernelGGL((k_sort_block<1024, 8192>), dim3(bin_count[2]), dim3(1024), 0, s,
                           nrows, ncols, I, J, A, list + 2 * int64_t(brow), bin_count[2], val,
                           tmp_col, tmp_map, blk_cnt, cnt, values_only);
        AMGT_HIP(hipGetLastError());
    }
    return AMGT_OK;
}

struct HostHeader {
    int32_t i0, in;
};

}  // namespace

static int bin_block_rows(int32_t nrows, int32_t brow, const int32_t *I, int32_t **list,
                          ConvCounters *cnt, ConvCounters *h_cnt, hipStream_t s) {
    AMGT_TRY(tmp_alloc_t(list, size_t(brow) * 4, s));
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, cnt);
    if (brow > 0)
        hipLaunchKernelGGL(k_bin_rows, dim3((brow + 255) / 256), dim3(256), 0, s, nrows, brow, I,
                           *list, cnt);
    AMGT_HIP(hipGetLastError());
    AMGT_HIP(hipMemcpyAsync(h_cnt, cnt, sizeof(ConvCounters), hipMemcpyDeviceToHost, s));
    AMGT_HIP(hipStreamSynchronize(s));
    return AMGT_OK;
}

int csr_to_mbsr(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *rowptr,
                const int32_t *colidx, const double *A, hipStream_t s, Mbsr *out) {
    Mbsr &m = *out;
    m = Mbsr();
    m.nrows = nrows;
    m.ncols = ncols;
    m.brow = (nrows + 3) / 4;
    m.bcol = (ncols + 3) / 4;
    m.nnz = nnz;
    // a matrix without nonzeros has (empty) values even when no value pointer is given
    m.pattern_only = (A == nullptr && nnz > 0);
    if (nnz >= (int64_t(1) << 31)) {
        set_error("CSR->mBSR: nnz exceeds int32 indexing");
        return AMGT_ERR_UNSUPPORTED;
    }
    if (ncols > (1 << 30) - 8) {
        // block column << 4 must stay below the padding key 0xFFFFFFFF
        set_error("CSR->mBSR: ncols > 2^30 - 8 does not fit the 32-bit sort key");
        return AMGT_ERR_UNSUPPORTED;
    }
    // verify I[0] == 0 and I[nrows] == nnz (one small D2H, the stream syncs below anyway)
    HostHeader hh{0, 0};
    AMGT_HIP(hipMemcpyAsync(&hh.i0, rowptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    AMGT_HIP(hipMemcpyAsync(&hh.in, rowptr + nrows, sizeof(int32_t), hipMemcpyDeviceToHost, s));

    m.blk_ptr = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(m.brow) + 1)));
    m.val_ptr = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(m.brow) + 1)));
    if (!m.blk_ptr || !m.val_ptr) {
        set_error("CSR->mBSR: allocation failed");
        return AMGT_ERR_ALLOC;
    }
    if (!m.pattern_only) {
        m.val = static_cast<double *>(dev_alloc(sizeof(double) * size_t(nnz ? nnz : 1)));
        if (!m.val) {
            set_error("CSR->mBSR: value allocation failed");
            return AMGT_ERR_ALLOC;
        }
    }
    ConvCounters *cnt = nullptr;
    AMGT_TRY(tmp_alloc_t(&cnt, 1, s));
    ConvCounters hc;
    int32_t *list = nullptr;
    int st = bin_block_rows(nrows, m.brow, rowptr, &list, cnt, &hc, s);
    if (st != AMGT_OK) {
        tmp_free(cnt, s);
        return st;
    }
    if (hh.i0 != 0 || int64_t(hh.in) != nnz) {
        tmp_free(list, s);
        tmp_free(cnt, s);
        set_error("CSR->mBSR: rowptr[0] must be 0 and rowptr[nrows] must equal nnz");
        return AMGT_ERR_INPUT;
    }
    int32_t *tmp_col = nullptr, *blk_cnt = nullptr;
    uint32_t *tmp_map = nullptr;
    st = tmp_alloc_t(&tmp_col, size_t(nnz), s);
    if (st == AMGT_OK) st = tmp_alloc_t(&tmp_map, size_t(nnz), s);
    if (st == AMGT_OK) st = tmp_alloc_t(&blk_cnt, size_t(m.brow) + 1, s);
    if (st == AMGT_OK && m.brow > 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3((m.brow + 255) / 256), dim3(256), 0, s, blk_cnt, 0,
                           int64_t(m.brow));
        st = launch_sorts(nrows, ncols, rowptr, colidx, A, list, m.brow, hc.bin_count, m.val,
                          tmp_col, tmp_map, blk_cnt, cnt, 0, s);
    }
    if (st == AMGT_OK) st = exclusive_scan_i32(blk_cnt, m.blk_ptr, m.brow, s);
    int32_t nblk = 0;
    if (st == AMGT_OK) {
        hipError_t e = hipMemcpyAsync(&nblk, m.blk_ptr + m.brow, sizeof(int32_t),
                                      hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            set_error(std::string("CSR->mBSR: ") + hipGetErrorString(e));
            st = AMGT_ERR_HIP;
        }
    }
    if (st == AMGT_OK) {
        m.nblk = nblk;
        m.blk_col = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * size_t(nblk ? nblk : 1)));
        m.blk_map = static_cast<uint16_t *>(dev_alloc(sizeof(uint16_t) * size_t(nblk ? nblk : 1)));
        if (!m.blk_col || !m.blk_map) {
            set_error("CSR->mBSR: block allocation failed");
            st = AMGT_ERR_ALLOC;
        }
    }
    if (st == AMGT_OK) {
        hipLaunchKernelGGL(k_compact, dim3((m.brow + 1 + 3) / 4), dim3(256), 0, s, nrows, m.brow,
                           rowptr, m.blk_ptr, tmp_col, tmp_map, m.blk_col, m.blk_map, m.val_ptr);
        double avg = m.brow ? double(nblk) / double(m.brow) : 0.0;
        if (m.brow > 0)
            hipLaunchKernelGGL(k_stand, dim3((m.brow + 255) / 256), dim3(256), 0, s, m.brow,
                               blk_cnt, avg, cnt);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(&hc, cnt, sizeof(ConvCounters), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            set_error(std::string("CSR->mBSR: ") + hipGetErrorString(e));
            st = AMGT_ERR_HIP;
        } else {
            // reference computes sqrtf on the double sum (csr_matvec_device.c:1249);
            // here the square root is taken in double
            m.stand = m.brow ? std::sqrt(hc.sq_dev / double(m.brow)) : 0.0;
            m.avg_nnz = nblk ? double(nnz) / double(nblk) : 0.0;
            m.max_row_blocks = hc.max_blocks;
            if (hc.err & kErrCol) {
                set_error("CSR->mBSR: column index out of range");
                st = AMGT_ERR_INPUT;
            } else if (hc.err & kErrDup) {
                set_error("CSR->mBSR: duplicate (row, col) entry");
                st = AMGT_ERR_INPUT;
            }
        }
    }
    tmp_free(tmp_col, s);
    tmp_free(tmp_map, s);
    tmp_free(blk_cnt, s);
    tmp_free(list, s);
    tmp_free(cnt, s);
    return st;
}

int mbsr_refresh_values(Mbsr *Ap, const int32_t *rowptr, const int32_t *colidx,
                        const double *A, hipStream_t s) {
    Mbsr &m = *Ap;
    if (m.pattern_only || !m.val) {
        m.val = static_cast<double *>(dev_alloc(sizeof(double) * size_t(m.nnz ? m.nnz : 1)));
        if (!m.val) {
            set_error("value allocation failed");
            return AMGT_ERR_ALLOC;
        }
        m.pattern_only = false;
    }
    ConvCounters *cnt = nullptr;
    AMGT_TRY(tmp_alloc_t(&cnt, 1, s));
    ConvCounters hc;
    int32_t *list = nullptr;
    int st = bin_block_rows(m.nrows, m.brow, rowptr, &list, cnt, &hc, s);
    if (st == AMGT_OK && m.brow > 0)
        st = launch_sorts(m.nrows, m.ncols, rowptr, colidx, A, list, m.brow, hc.bin_count, m.val,
                          nullptr, nullptr, nullptr, cnt, 1, s);
    tmp_free(list, s);
    tmp_free(cnt, s);
    // low-precision copies and the transpose are stale now
    dev_free(m.val32);
    m.val32 = nullptr;
    dev_free(m.val16);
    m.val16 = nullptr;
    if (m.trans) {
        mbsr_release(m.trans);
        delete m.trans;
        m.trans = nullptr;
    }
    return st;
}

}  // namespace amgt
