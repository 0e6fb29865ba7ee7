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
// Bins by entries per block row: <= 64 (one wave, rank sort over an LDS copy
// of the keys), <= 1024 (256-thread workgroup, LDS bitonic), <= 8192
// (1024-thread workgroup, LDS bitonic).  Larger block rows return
// AMGT_ERR_UNSUPPORTED.
#include "internal.hpp"
#include "rowpass.hpp"

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
    int32_t bin_count[4];  // host side only: rows per bin, [3] = 1 if a row is too large
    int32_t err;
    int32_t pad;
};

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_zero_counters(ConvCounters *c) {
    if (threadIdx.x == 0) {
        for (int i = 0; i < 4; ++i) c->bin_count[i] = 0;
        c->err = 0;
        c->pad = 0;
    }
}

__device__ __forceinline__ int bin_of(int n) {
    return n <= 64 ? 0 : (n <= 1024 ? 1 : (n <= 8192 ? 2 : 3));
}

// Entries of block row i (CSR rows 4i .. 4i+3), for the row passes.
struct RowEntries {
    const int32_t *I;
    int32_t nrows;
    __device__ int64_t operator()(int i) const { return I[min(4 * i + 4, nrows)] - I[4 * i]; }
};

__device__ __forceinline__ void wave_sync_conv() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
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

// ---- bin 0: one wave per block row, <= 64 entries, rank sort in LDS.
// tmp_col/tmp_map receive the row's blocks at offset I[4br] (upper-bound slot).
__global__ __launch_bounds__(256) void k_sort_wave(
    int32_t nrows, int32_t ncols, const int32_t *__restrict__ I, const int32_t *__restrict__ J,
    const double *__restrict__ A, int32_t brow, double *__restrict__ val,
    int32_t *__restrict__ tmp_col, uint32_t *__restrict__ tmp_map,
    int32_t *__restrict__ blk_cnt, ConvCounters *cnt, int values_only, int packed) {
    __shared__ __align__(16) uint32_t s_key[4][64];
    __shared__ uint32_t s_sorted[4][64];
    const int lane = threadIdx.x & 63;
    const int br = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (br >= brow) return;  // whole wave exits together
    const int r0 = 4 * br;
    const int s = I[r0];
    const int e = I[min(r0 + 4, nrows)];
    const int n = e - s;
    if (n > 64) return;  // listed for the workgroup sorts
    int err = 0;
    uint32_t key = kPadKey;
    int src = 0;
    if (lane < n) {
        src = s + lane;
        key = entry_key(src, r0, nrows, ncols, I, J, &err);
    }
    // Rank sort: every lane counts the keys below its own over an LDS copy of
    // the row's keys (four per read; ties, i.e. duplicate entries, by lane), so
    // the work is a few VALU operations per key instead of a 21-stage shuffle
    // network (this kernel is bound by VALU issue).  With fewer than 2^24 - 4
    // columns the key and the lane fit one 32-bit integer, all distinct, and
    // the rank is one comparison per key.
    const int w = threadIdx.x >> 6;
    const bool valid = lane < n;
    int rank = 0;
    if (packed) {
        const uint32_t pk = valid ? (key << 6) | uint32_t(lane) : kPadKey;
        s_key[w][lane] = pk;
        wave_sync_conv();
        for (int j = 0; j < n; j += 4) {
            const uint4 kk = *reinterpret_cast<const uint4 *>(&s_key[w][j]);
            rank += kk.x < pk ? 1 : 0;
            rank += kk.y < pk ? 1 : 0;
            rank += kk.z < pk ? 1 : 0;
            rank += kk.w < pk ? 1 : 0;
        }
    } else {
        s_key[w][lane] = key;
        wave_sync_conv();
        for (int j = 0; j < n; j += 4) {
            const uint4 kk = *reinterpret_cast<const uint4 *>(&s_key[w][j]);
            rank += (kk.x < key || (kk.x == key && j < lane)) ? 1 : 0;
            rank += (kk.y < key || (kk.y == key && j + 1 < lane)) ? 1 : 0;
            rank += (kk.z < key || (kk.z == key && j + 2 < lane)) ? 1 : 0;
            rank += (kk.w < key || (kk.w == key && j + 3 < lane)) ? 1 : 0;
        }
    }
    if (valid) {
        s_sorted[w][rank] = key;
        if (A) val[s + rank] = A[src];
    }
    wave_sync_conv();
    // lane p now holds the p-th smallest key
    key = valid ? s_sorted[w][lane] : kPadKey;
    const uint32_t prev = __shfl_up(key, 1);
    if (valid && lane > 0 && prev == key) err |= kErrDup;
    if (!values_only) {
        const bool head = valid && (lane == 0 || (prev >> 4) != (key >> 4));
        const uint64_t heads = __ballot(head);
        // OR of the bits of each block: segmented inclusive scan over equal
        // block columns; the block's last entry holds the whole bitmap
        uint32_t map = valid ? (1u << (key & 15u)) : 0u;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const uint32_t k2 = __shfl_up(key, d);
            const uint32_t m2 = __shfl_up(map, d);
            if (lane >= d && (k2 >> 4) == (key >> 4)) map |= m2;
        }
        const uint32_t nxt = __shfl_down(key, 1);
        const bool tail = valid && (lane == n - 1 || (nxt >> 4) != (key >> 4));
        if (tail) {
            const uint64_t le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
            const int brank = __popcll(heads & le) - 1;
            tmp_col[s + brank] = int32_t(key >> 4);
            tmp_map[s + brank] = map;
        }
        if (lane == 0) blk_cnt[br] = __popcll(heads);
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
        if (t == 0) blk_cnt[br] = carry;
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

}  // namespace

int row_stats_launch(const int32_t *blk_ptr, int32_t brow, RowStats *d_stats, hipStream_t s) {
    rowpass::Part *parts = nullptr;
    AMGT_TRY(tmp_alloc_t(&parts, size_t(rowpass::groups(brow)) + 1, s));
    const int st = rowpass::count(brow, rowpass::BlockDiff{blk_ptr}, INT64_MAX, INT64_MAX, parts,
                                  d_stats, s);
    tmp_free(parts, s);
    return st;
}

double stand_from(const RowStats &h, int64_t nblk, int32_t brow) {
    // sqrt(sum (c - avg)^2 / brow) with sum (c - avg)^2 = (brow*sum c^2 - nblk^2) / brow,
    // the numerator exact in 128-bit integers (the reference sums doubles with
    // atomics and takes sqrtf, csr_matvec_device.c:1245-1249)
    if (brow <= 0) return 0.0;
    const __int128 num = static_cast<__int128>(h.sumsq) * brow - static_cast<__int128>(nblk) * nblk;
    const double var = double(num) / (double(brow) * double(brow));
    return var > 0.0 ? std::sqrt(var) : 0.0;
}

namespace {

int launch_sorts(int32_t nrows, int32_t ncols, const int32_t *I, const int32_t *J, const double *A,
                 const int32_t *list, int32_t brow, const int32_t bin_count[4], double *val,
                 int32_t *tmp_col, uint32_t *tmp_map, int32_t *blk_cnt, ConvCounters *cnt,
                 int values_only, hipStream_t s) {
    if (bin_count[3] > 0) {
        set_error("CSR->mBSR: a block row has more than 8192 entries (4 rows with > 2048 "
                  "nonzeros on average); this build does not convert such rows");
        return AMGT_ERR_UNSUPPORTED;
    }
    if (brow > 0) {
        int grid = (brow + 3) / 4;
        // columns <= 2^24 - 5: keys below 2^26 - 16, packed keys below kPadKey
        const int packed = ncols < (1 << 24) - 4 ? 1 : 0;
        hipLaunchKernelGGL(k_sort_wave, dim3(grid), dim3(256), 0, s, nrows, ncols, I, J, A, brow,
                           val, tmp_col, tmp_map, blk_cnt, cnt, values_only, packed);
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

// Bins block rows by entries: <= 64 (wave sort, not listed: it walks every
// block row), <= 1024 and larger (listed at list + brow and list + 2*brow for
// the workgroup sorts).  bin_count[3] flags a block row over 8192 entries.
// hh (optional): I[0] and I[nrows], read back with the bins.
static int bin_block_rows(int32_t nrows, int32_t brow, const int32_t *I, int32_t **list,
                          ConvCounters *cnt, ConvCounters *h_cnt, HostHeader *hh,
                          hipStream_t s) {
    AMGT_TRY(tmp_alloc_t(list, size_t(brow) * 3 + 1, s));
    rowpass::Part *parts = nullptr;
    RowStats *tot = nullptr;
    AMGT_TRY(tmp_alloc_t(&parts, size_t(rowpass::groups(brow)) + 1, s));
    AMGT_TRY(tmp_alloc_t(&tot, 1, s));
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, cnt);
    AMGT_HIP(hipGetLastError());
    const RowEntries f{I, nrows};
    int st = rowpass::count(brow, f, 64, 1024, parts, tot, s);
    RowStats ht{};
    if (st == AMGT_OK) {
        const Readback rb[3] = {{&ht, tot, sizeof(RowStats)},
                                {hh ? &hh->i0 : nullptr, I, sizeof(int32_t)},
                                {hh ? &hh->in : nullptr, I + nrows, sizeof(int32_t)}};
        st = readback(rb, hh ? 3 : 1, s);
    }
    *h_cnt = ConvCounters{};
    h_cnt->bin_count[0] = ht.cnt[0];
    h_cnt->bin_count[1] = ht.cnt[1];
    h_cnt->bin_count[2] = ht.cnt[2];
    h_cnt->bin_count[3] = ht.max_blocks > 8192 ? 1 : 0;
    if (st == AMGT_OK && h_cnt->bin_count[3] == 0 && ht.cnt[1] + ht.cnt[2] > 0)
        st = rowpass::scatter(brow, f, 64, 1024, parts, *list, int64_t(brow), 6u, s);
    tmp_free(parts, s);
    tmp_free(tot, s);
    return st;
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
    // I[0] == 0 and I[nrows] == nnz are verified with the bins' readback below
    HostHeader hh{0, 0};

    m.blk_ptr = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(m.brow) + 1), s));
    m.val_ptr = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(m.brow) + 1), s));
    if (!m.blk_ptr || !m.val_ptr) {
        set_error("CSR->mBSR: allocation failed");
        return AMGT_ERR_ALLOC;
    }
    if (!m.pattern_only) {
        m.val = static_cast<double *>(dev_alloc(sizeof(double) * size_t(nnz ? nnz : 1), s));
        if (!m.val) {
            set_error("CSR->mBSR: value allocation failed");
            return AMGT_ERR_ALLOC;
        }
    }
    ConvCounters *cnt = nullptr;
    AMGT_TRY(tmp_alloc_t(&cnt, 1, s));
    ConvCounters hc;
    int32_t *list = nullptr;
    int st = bin_block_rows(nrows, m.brow, rowptr, &list, cnt, &hc, &hh, s);
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
    RowStats *rstats = nullptr;
    RowStats hrs{};
    st = tmp_alloc_t(&rstats, 1, s);
    if (st == AMGT_OK) st = tmp_alloc_t(&tmp_col, size_t(nnz), s);
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
        const Readback rb[1] = {{&nblk, m.blk_ptr + m.brow, sizeof(int32_t)}};
        st = readback(rb, 1, s);
    }
    if (st == AMGT_OK) {
        m.nblk = nblk;
        m.blk_col = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * size_t(nblk ? nblk : 1), s));
        m.blk_map = static_cast<uint16_t *>(dev_alloc(sizeof(uint16_t) * size_t(nblk ? nblk : 1), s));
        if (!m.blk_col || !m.blk_map) {
            set_error("CSR->mBSR: block allocation failed");
            st = AMGT_ERR_ALLOC;
        }
    }
    if (st == AMGT_OK) {
        hipLaunchKernelGGL(k_compact, dim3((m.brow + 1 + 3) / 4), dim3(256), 0, s, nrows, m.brow,
                           rowptr, m.blk_ptr, tmp_col, tmp_map, m.blk_col, m.blk_map, m.val_ptr);
        st = row_stats_launch(m.blk_ptr, m.brow, rstats, s);
        hipError_t e = hipGetLastError();
        if (st == AMGT_OK && e != hipSuccess) {
            set_error(std::string("CSR->mBSR: ") + hipGetErrorString(e));
            st = AMGT_ERR_HIP;
        }
        if (st == AMGT_OK) {
            const Readback rb[2] = {{&hc, cnt, sizeof(ConvCounters)},
                                    {&hrs, rstats, sizeof(RowStats)}};
            st = readback(rb, 2, s);
        }
        if (st == AMGT_OK) {
            m.stand = stand_from(hrs, nblk, m.brow);
            m.avg_nnz = nblk ? double(nnz) / double(nblk) : 0.0;
            m.max_row_blocks = hrs.max_blocks;
            m.spmv_kernel = default_spmv_kernel(m);
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
    tmp_free(rstats, s);
    tmp_free(list, s);
    tmp_free(cnt, s);
    return st;
}

int mbsr_refresh_values(Mbsr *Ap, const int32_t *rowptr, const int32_t *colidx,
                        const double *A, hipStream_t s) {
    Mbsr &m = *Ap;
    if (m.pattern_only || !m.val) {
        m.val = static_cast<double *>(dev_alloc(sizeof(double) * size_t(m.nnz ? m.nnz : 1), s));
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
    int st = bin_block_rows(m.nrows, m.brow, rowptr, &list, cnt, &hc, nullptr, s);
    if (st == AMGT_OK && m.brow > 0)
        st = launch_sorts(m.nrows, m.ncols, rowptr, colidx, A, list, m.brow, hc.bin_count, m.val,
                          nullptr, nullptr, nullptr, cnt, 1, s);
    tmp_free(list, s);
    tmp_free(cnt, s);
    // low-precision copies and the transpose are stale now
    dev_free(m.val32, s);
    m.val32 = nullptr;
    dev_free(m.val16, s);
    m.val16 = nullptr;
    m.half_fits = -1;
    if (m.trans) {
        mbsr_release(m.trans, s);
        delete m.trans;
        m.trans = nullptr;
    }
    return st;
}

}  // namespace amgt
