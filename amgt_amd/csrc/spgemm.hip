// mBSR SpGEMM C = A*B (counterpart of spgemm_amgT_fp64, csr_spgemm_device.c:1527-1846).
//
// 1. upper bound  ub_i = sum over A blocks (i,k) of |B row k|  (compute_Cub_bin,
//    csr_spgemm_device.c:82-134), scanned into per-row slots of a temporary
//    buffer; rows are binned by ub.
// 2. symbolic: every product (A block, B block) whose 4x4 bitmap product
//    (bitMatrixMul, 189-205) is non-zero inserts its block column into a hash
//    table (LDS for ub <= 2048, a global slab otherwise) and ORs the product
//    bitmap into the entry; the distinct block columns are then sorted with
//    their bitmaps and written to the row's slot.  One pass replaces the
//    reference's count + fill passes (206-752): the slot layout is sized by ub.
//    Products are flattened over the lanes of a group (A blocks x B blocks),
//    not "one warp per A block".
// 3. scans of block counts and value counts give C.blk_ptr / C.val_ptr; one
//    device->host copy sizes C.
// 4. numeric (1142-1342): one wave per C block row keeps the row's 4x4
//    accumulators in LDS (global scratch for rows with more than 64 blocks);
//    for each A block in order, four B blocks at a time are multiplied with one
//    v_mfma_f64_4x4x4f64 (four independent 4x4x4 products: the A block
//    replicated, B block g in block slot g) and added to the C slots of their
//    block columns.  A B row never repeats a block column, so the four slots
//    are distinct: no atomics, deterministic results.
// 5. accumulators are compressed to the mBSR value layout.
#include "internal.hpp"
#include "rowpass.hpp"

#include <algorithm>
#include <string>
#include <type_traits>

namespace amgt {

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kNoEntryOff = 0xFFFFFFF0u;  // a buffer offset past every resource: reads 0
constexpr int kBin1Ub = 2048;   // workgroup per row, LDS table of 4096
constexpr int kNumCap = 64;     // C blocks per row kept in LDS by the numeric kernel
constexpr int kBigCap = 256;    // ... and by the long-row kernel before its global slab
// long rows whose C values (compressed, as the symbolic phase fixed them) fit
// kCmpVal accumulators and whose blocks fit kCmpBlk: LDS per wave ~16 KB
constexpr int kCmpBlk = 512;
constexpr int kCmpVal = 1024;
constexpr int kCmpValShort = 256;  // ... and rows of at most kNumCap blocks
constexpr int kCmpMidBlk = 256, kCmpMidVal = 512;  // the long bin's smaller launch

// a C block row of at most BLK blocks and VAL values (row pointers of C)
template <int BLK, int VAL>
__device__ __forceinline__ bool fits_cmp(const int32_t *cptr, const int32_t *cvp, int i) {
    return cptr[i + 1] - cptr[i] <= BLK && cvp[i + 1] - cvp[i] <= VAL;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int NT>
__device__ __forceinline__ void grp_sync() {
    if constexpr (NT == 64) {
        wave_sync();
    } else {
        __syncthreads();
    }
}

// inclusive scan over the NT threads of a group; `tot` gets the group total.
// sc: NT/64 ints of LDS scratch for NT > 64.
template <int NT>
__device__ __forceinline__ int grp_scan(int v, int *sc, int *tot) {
    const int lane = threadIdx.x & 63;
    v = wave_incl_sum(v);  // every lane of the group calls this
    if constexpr (NT == 64) {
        *tot = __builtin_amdgcn_readlane(v, 63);
        return v;
    } else {
        const int w = (threadIdx.x % NT) >> 6;
        __syncthreads();
        if (lane == 63) sc[w] = v;
        __syncthreads();
        int before = 0, all = 0;
#pragma unroll
        for (int q = 0; q < NT / 64; ++q) {
            int s = sc[q];
            if (q < w) before += s;
            all += s;
        }
        *tot = all;
        return v + before;
    }
}

__device__ __forceinline__ uint32_t hash_slot(uint32_t key, int hs) {
    return (key * 2654435761u) & uint32_t(hs - 1);
}

// Symbolic work for one C block row by a group of NT threads (tid in [0, NT)).
// hkey/hmap: hash table of hs slots (LDS or global), initialised here.
// s_excl/s_q0/s_ma: NT-entry LDS staging for the product enumeration.
// kTry (a row whose upper bound exceeds the table): insertion gives up once
// more than hs/2 distinct keys arrived or a probe sequence passed the whole
// table; the row then writes nothing but retry[i] = 1 (retry[i] = 0 when it
// completes), for the global-table kernel.  s_flag: two LDS ints.
template <int NT, bool kTry = false>
__device__ void symbolic_row(int i, int tid, uint32_t *hkey, uint32_t *hmap, int hs,
                             int *s_excl, int *s_q0, int *s_ma, int *s_sc,
                             const int32_t *__restrict__ ap, const int32_t *__restrict__ ac,
                             const uint16_t *__restrict__ am, const int32_t *__restrict__ bp,
                             const int32_t *__restrict__ bc, const uint16_t *__restrict__ bm,
                             int64_t slot, int32_t *__restrict__ tcol, uint32_t *__restrict__ tmap,
                             int32_t *__restrict__ cnt_out, int32_t *__restrict__ nnz_out,
                             int *s_flag = nullptr, int32_t *__restrict__ retry = nullptr) {
    for (int s = tid; s < hs; s += NT) {
        hkey[s] = kEmpty;
        hmap[s] = 0u;
    }
    if constexpr (kTry) {
        if (tid == 0) {
            s_flag[0] = 0;  // distinct keys inserted
            s_flag[1] = 0;  // given up
        }
    }
    grp_sync<NT>();
    const int a_beg = ap[i], a_end = ap[i + 1];
    for (int a0 = a_beg; a0 < a_end; a0 += NT) {
        const int a = a0 + tid;
        int len = 0, q0 = 0, ma = 0;
        if (a < a_end) {
            const int k = ac[a];
            ma = am[a];
            q0 = bp[k];
            len = bp[k + 1] - q0;
        }
        int total;
        const int incl = grp_scan<NT>(len, s_sc, &total);
        s_excl[tid] = incl - len;
        s_q0[tid] = q0;
        s_ma[tid] = ma;
        grp_sync<NT>();
        const int na = min(NT, a_end - a0);
        for (int t = tid; t < total; t += NT) {
            // last staged A block with excl <= t
            int lo = 0, hi = na - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_excl[mid] <= t) lo = mid; else hi = mid - 1;
            }
            const int q = s_q0[lo] + (t - s_excl[lo]);
            const uint32_t mc = bitmap_mul(uint32_t(s_ma[lo]), bm[q]);
            if (mc) {
                const uint32_t key = uint32_t(bc[q]);
                uint32_t h = hash_slot(key, hs);
                if constexpr (kTry) {
                    if (__atomic_load_n(&s_flag[1], __ATOMIC_RELAXED)) continue;
                    int probe = 0;
                    for (; probe < hs; ++probe) {
                        const uint32_t old = atomicCAS(&hkey[h], kEmpty, key);
                        if (old == kEmpty || old == key) {
                            atomicOr(&hmap[h], mc);
                            if (old == kEmpty && atomicAdd(&s_flag[0], 1) >= hs / 2)
                                atomicExch(&s_flag[1], 1);
                            break;
                        }
                        h = (h + 1) & uint32_t(hs - 1);
                    }
                    if (probe == hs) atomicExch(&s_flag[1], 1);
                } else {
                    while (true) {
                        const uint32_t old = atomicCAS(&hkey[h], kEmpty, key);
                        if (old == kEmpty || old == key) {
                            atomicOr(&hmap[h], mc);
                            break;
                        }
                        h = (h + 1) & uint32_t(hs - 1);
                    }
                }
            }
        }
        grp_sync<NT>();
        if constexpr (kTry) {
            if (__atomic_load_n(&s_flag[1], __ATOMIC_RELAXED)) break;  // uniform after the sync
        }
    }
    if constexpr (kTry) {
        const int gave_up = __atomic_load_n(&s_flag[1], __ATOMIC_RELAXED);
        grp_sync<NT>();
        if (tid == 0) retry[i] = gave_up;
        if (gave_up) return;  // uniform: every thread read the flag after the last sync
    }
    // compact occupied slots to the front (in place: destinations never pass sources)
    int base = 0;
    for (int s0 = 0; s0 < hs; s0 += NT) {
        const int s = s0 + tid;
        uint32_t k = kEmpty, m = 0;
        if (s < hs) {
            k = hkey[s];
            m = hmap[s];
        }
        const int occ = (k != kEmpty);
        int tot, incl;
        if constexpr (NT == 64) {  // a ballot counts the occupied slots
            const uint64_t bal = __ballot(occ);
            incl = __popcll(bal & ((2ull << (tid & 63)) - 1ull));
            tot = __popcll(bal);
        } else {
            incl = grp_scan<NT>(occ, s_sc, &tot);
        }
        grp_sync<NT>();
        if (occ) {
            hkey[base + incl - 1] = k;
            hmap[base + incl - 1] = m;
        }
        base += tot;
        grp_sync<NT>();
    }
    const int cnt = base;
    if constexpr (NT == 64) {
        if (cnt <= 64) {
            // at most one key per lane and all keys distinct: each lane's rank
            // among the row's keys is its sorted position
            const uint32_t key = tid < cnt ? hkey[tid] : kEmpty;
            const uint32_t m = tid < cnt ? hmap[tid] : 0u;
            int rank = 0;
            for (int j = 0; j < cnt; ++j) rank += hkey[j] < key ? 1 : 0;
            if (tid < cnt) {
                tcol[slot + rank] = int32_t(key);
                tmap[slot + rank] = m;
            }
            const int nz_tot = __builtin_amdgcn_readlane(wave_incl_sum(__popc(m)), 63);
            if (tid == 0) {
                cnt_out[i] = cnt;
                nnz_out[i] = nz_tot;
            }
            grp_sync<NT>();
            return;
        }
    }
    int np = 1;
    while (np < cnt) np <<= 1;
    for (int s = cnt + tid; s < np; s += NT) hkey[s] = kEmpty;
    grp_sync<NT>();
    // bitonic sort of (key, map) by key
    for (int k = 2; k <= np; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int x = tid; x < np; x += NT) {
                const int p = x ^ j;
                if (p > x) {
                    const bool up = (x & k) == 0;
                    const uint32_t kx = hkey[x], kp = hkey[p];
                    if ((kx > kp) == up) {
                        hkey[x] = kp;
                        hkey[p] = kx;
                        const uint32_t mx = hmap[x];
                        hmap[x] = hmap[p];
                        hmap[p] = mx;
                    }
                }
            }
            grp_sync<NT>();
        }
    }
    int nz = 0;
    for (int t = tid; t < cnt; t += NT) {
        const uint32_t m = hmap[t];
        tcol[slot + t] = int32_t(hkey[t]);
        tmap[slot + t] = m;
        nz += __popc(m);
    }
    int nz_tot;
    grp_scan<NT>(nz, s_sc, &nz_tot);
    if (tid == 0) {
        cnt_out[i] = cnt;
        nnz_out[i] = nz_tot;
    }
    grp_sync<NT>();
}

struct SpgemmArgs {
    const int32_t *ap, *ac, *avp;
    const uint16_t *am;
    const void *av;   // A values: double, float or __half (the numeric kernels' VT)
    const int32_t *bp, *bc, *bvp;
    const uint16_t *bm;
    const void *bv;   // B values, same type as av
    int64_t bnnz;     // values of B
    int64_t annz;     // values of A
    int fp64_rs;      // FP64 values of A and B both under 2^32 - 16 bytes: whole-array
                      // buffer resources with 32-bit offsets
};

// value of type VT at index idx, widened to double
template <typename VT>
__device__ __forceinline__ double ldv(const void *p, int64_t idx) {
    return double(static_cast<const VT *>(p)[idx]);
}
template <>
__device__ __forceinline__ double ldv<__half>(const void *p, int64_t idx) {
    return double(__half2float(static_cast<const __half *>(p)[idx]));
}

// C entry as stored: the FP64 product for FP64 operands; below FP64 the
// entry is an FP32 number (the reference keeps C in the operands' precision,
// csr_spgemm_device.c:938-1140 TF32 and 753-937 FP16 with FP16 accumulation;
// here FP32 / FP16 operands run on the 16-block MFMAs with FP32 accumulation,
// and the entry is stored as FP32)
template <typename VT>
__device__ __forceinline__ double round_out(double x) {
    return double(float(x));
}
template <>
__device__ __forceinline__ double round_out<double>(double x) {
    return x;
}

// ---- bin 0: rows with at most 64 products and 64 A blocks, RPW rows per wave.
// Lane j holds A block j, lane t product t (its owner A block: every A block
// with products marks its first one, a max-scan spreads the marks).  The
// block columns of the products with a non-zero bitmap product are
// rank-sorted (each lane counts the keys below its own over an LDS copy of
// the 64 keys, ties by lane) and placed in sorted order; the first lane of
// every run of equal columns opens a C block, and every product ORs its
// bitmap product into its C block's bitmap in LDS (OR does not depend on the
// order).  No hashing, deterministic.  The kernel waits on chains of
// dependent loads (row pointer, A blocks, B row pointers, B blocks), so a
// wave takes RPW rows and issues each stage's loads for all of them before
// it waits.
//
// rows / ub_off NULL: the bounded form, every block row 0 .. count-1 with its
// slot of the temporaries at row * stride, where the host's bound (largest A
// row times largest B row, at most 64) holds every row's products.  A row
// that breaks the bound writes no structure and a block count of -1, which
// the size pass reports as an error.
constexpr int kSymRegRows = 2;  // rows per wave of k_symbolic_reg
template <int RPW>
__global__ __launch_bounds__(256) void k_symbolic_reg(int count, const int32_t *__restrict__ rows,
                                                      SpgemmArgs g,
                                                      const int64_t *__restrict__ ub_off,
                                                      int32_t stride, int packed,
                                                      int32_t *__restrict__ tcol,
                                                      uint32_t *__restrict__ tmap,
                                                      int32_t *__restrict__ cnt_out,
                                                      int32_t *__restrict__ nnz_out) {
    __shared__ __align__(16) int32_t s_key[4][RPW][64];
    __shared__ int32_t s_sk[4][RPW][64];  // owner marks, then the sorted keys
    __shared__ uint32_t s_sm[4][RPW][64], s_cm[4][RPW][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * 4 + w) * RPW;
    if (r0 >= count) return;  // whole wave; the waves share no LDS
    int i[RPW], na[RPW], a_beg[RPW];
    bool on[RPW];
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        on[u] = r0 + u < count;
        i[u] = on[u] ? (rows ? rows[r0 + u] : r0 + u) : 0;
    }
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        a_beg[u] = g.ap[i[u]];
        na[u] = on[u] ? g.ap[i[u] + 1] - a_beg[u] : 0;
    }
    int kk[RPW], ma[RPW], bq0[RPW], blen[RPW];
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        kk[u] = 0;
        ma[u] = 0;
        if (lane < na[u]) {
            kk[u] = g.ac[a_beg[u] + lane];
            ma[u] = g.am[a_beg[u] + lane];
        }
    }
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        bq0[u] = 0;
        blen[u] = 0;
        if (lane < na[u]) {
            bq0[u] = g.bp[kk[u]];
            blen[u] = g.bp[kk[u] + 1] - bq0[u];
        }
    }
    int pst[RPW], ptot[RPW];
    bool ok[RPW];
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        const int linc = wave_incl_sum(blen[u]);
        pst[u] = linc - blen[u];
        ptot[u] = __builtin_amdgcn_readlane(linc, 63);
        ok[u] = on[u] && na[u] <= 64 && ptot[u] <= (ub_off ? 64 : stride);
        if (on[u] && !ok[u] && lane == 0) {  // outside the row's bin or bound
            cnt_out[i[u]] = -1;
            nnz_out[i[u]] = 0;
        }
        s_sk[w][u][lane] = 0;
        s_cm[w][u][lane] = 0u;
    }
    wave_sync();
#pragma unroll
    for (int u = 0; u < RPW; ++u)
        if (ok[u] && blen[u] > 0) s_sk[w][u][pst[u]] = lane;  // pst < ptot <= 64, distinct
    wave_sync();
    int q[RPW];
    uint32_t pma[RPW];
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        const int own = wave_incl_max(s_sk[w][u][lane]);
        q[u] = __shfl(bq0[u], own) + (lane - __shfl(pst[u], own));
        pma[u] = uint32_t(__shfl(ma[u], own));
    }
    int key[RPW];
    uint32_t mb[RPW];
#pragma unroll
    for (int u = 0; u < RPW; ++u) {  // the products' B blocks, all rows' loads together
        key[u] = INT_MAX;
        mb[u] = 0u;
        if (ok[u] && lane < ptot[u]) {
            mb[u] = uint32_t(g.bm[q[u]]);
            key[u] = g.bc[q[u]];
        }
    }
    uint32_t mc[RPW];
    int nvalid[RPW], rank[RPW];
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        mc[u] = bitmap_mul(pma[u], mb[u]);
        if (!mc[u]) key[u] = INT_MAX;
        nvalid[u] = __popcll(__ballot(key[u] != INT_MAX));
        rank[u] = 0;
    }
    if (packed) {
        // block columns below 2^25 - 1: the key and the lane in one integer,
        // all distinct, so the rank is one comparison per key
        int pk[RPW];
#pragma unroll
        for (int u = 0; u < RPW; ++u) {
            pk[u] = key[u] != INT_MAX ? key[u] * 64 + lane : INT_MAX;
            s_key[w][u][lane] = pk[u];
        }
        wave_sync();
#pragma unroll
        for (int u = 0; u < RPW; ++u) {
            if (!ok[u]) continue;  // wave-uniform
            for (int j = 0; j < ptot[u]; j += 4) {
                const int4 k4 = *reinterpret_cast<const int4 *>(&s_key[w][u][j]);
                rank[u] += k4.x < pk[u] ? 1 : 0;
                rank[u] += k4.y < pk[u] ? 1 : 0;
                rank[u] += k4.z < pk[u] ? 1 : 0;
                rank[u] += k4.w < pk[u] ? 1 : 0;
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < RPW; ++u) s_key[w][u][lane] = key[u];
        wave_sync();
#pragma unroll
        for (int u = 0; u < RPW; ++u) {
            if (!ok[u]) continue;  // wave-uniform
            const int ky = key[u];
            for (int j = 0; j < ptot[u]; j += 4) {
                const int4 k4 = *reinterpret_cast<const int4 *>(&s_key[w][u][j]);
                rank[u] += (k4.x < ky || (k4.x == ky && j < lane)) ? 1 : 0;
                rank[u] += (k4.y < ky || (k4.y == ky && j + 1 < lane)) ? 1 : 0;
                rank[u] += (k4.z < ky || (k4.z == ky && j + 2 < lane)) ? 1 : 0;
                rank[u] += (k4.w < ky || (k4.w == ky && j + 3 < lane)) ? 1 : 0;
            }
        }
    }
    // lanes at or past ptot hold INT_MAX and rank past every valid key
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        if (ok[u] && key[u] != INT_MAX) {
            s_sk[w][u][rank[u]] = key[u];
            s_sm[w][u][rank[u]] = mc[u];
        }
    }
    wave_sync();
    uint64_t heads[RPW];
    int sk[RPW];
    bool in[RPW];
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        in[u] = ok[u] && lane < nvalid[u];
        sk[u] = in[u] ? s_sk[w][u][lane] : INT_MAX;
        const bool head = in[u] && (lane == 0 || s_sk[w][u][lane > 0 ? lane - 1 : 0] != sk[u]);
        heads[u] = __ballot(head);
        const int cslot = __popcll(heads[u] & ((2ull << lane) - 1ull)) - 1;  // sorted lane's block
        if (in[u]) atomicOr(&s_cm[w][u][cslot], s_sm[w][u][lane]);
    }
    wave_sync();
#pragma unroll
    for (int u = 0; u < RPW; ++u) {
        if (!ok[u]) continue;  // wave-uniform
        const int ncb = __popcll(heads[u]);
        const int64_t slot = ub_off ? ub_off[i[u]] : int64_t(i[u]) * stride;
        const bool head = ((heads[u] >> lane) & 1ull) != 0ull;
        const int cslot = __popcll(heads[u] & ((2ull << lane) - 1ull)) - 1;
        if (head) tcol[slot + cslot] = sk[u];
        const uint32_t cm = lane < ncb ? s_cm[w][u][lane] : 0u;
        if (lane < ncb) tmap[slot + lane] = cm;
        const int nz = __builtin_amdgcn_readlane(wave_incl_sum(__popc(cm)), 63);
        if (lane == 0) {
            cnt_out[i[u]] = ncb;
            nnz_out[i[u]] = nz;
        }
    }
}

// Symbolic bin of a C block row: 0 = k_symbolic_reg (<= 64 products and A
// blocks), else by the upper bound: <= kBin1Ub workgroup hash in LDS, larger
// global hash.
struct SymbolicBin {
    const int64_t *ub;
    const int32_t *ap;
    __device__ int64_t operator()(int i) const {
        const int64_t u = ub[i];
        if (u <= 64 && ap[i + 1] - ap[i] <= 64) return 0;
        return u > 0 ? u : 1;
    }
};

// hash slots for one row: a power of two >= twice the row's upper bound (the
// distinct keys never fill more than half of it), at least 256, at most cap
__device__ __forceinline__ int row_table(const int64_t *__restrict__ ub_off, int i, int cap) {
    const int64_t want = 2 * (ub_off[i + 1] - ub_off[i]);
    int hs = 256;
    while (hs < want && hs < cap) hs <<= 1;
    return hs < cap ? hs : cap;
}

// ---- bin 1, rows with ub <= kWaveUb: one wave per row, each with its own LDS
// table of up to 2 kWaveUb slots (four rows per workgroup)
constexpr int kWaveUb = 512;
__global__ __launch_bounds__(256) void k_symbolic_wave(int count, const int32_t *__restrict__ rows,
                                                       SpgemmArgs g,
                                                       const int64_t *__restrict__ ub_off,
                                                       int32_t *tcol, uint32_t *tmap,
                                                       int32_t *cnt_out, int32_t *nnz_out) {
    constexpr int HS = 2 * kWaveUb;
    __shared__ uint32_t s_key[4][HS];
    __shared__ uint32_t s_map[4][HS];
    __shared__ int s_excl[4][64], s_q0[4][64], s_ma[4][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + w;
    if (r >= count) return;  // whole wave
    const int i = rows[r];
    if (ub_off[i + 1] - ub_off[i] > kWaveUb) return;  // k_symbolic_block's row
    symbolic_row<64>(i, lane, s_key[w], s_map[w], row_table(ub_off, i, HS), s_excl[w], s_q0[w],
                     s_ma[w], nullptr, g.ap, g.ac, g.am, g.bp, g.bc, g.bm, ub_off[i], tcol, tmap,
                     cnt_out, nnz_out);
}

// ---- bin 1: one workgroup per row, LDS table of up to 4096
__global__ __launch_bounds__(256) void k_symbolic_block(int count, const int32_t *__restrict__ rows,
                                                        SpgemmArgs g,
                                                        const int64_t *__restrict__ ub_off,
                                                        int32_t *tcol, uint32_t *tmap,
                                                        int32_t *cnt_out, int32_t *nnz_out) {
    constexpr int HS = 2 * kBin1Ub;
    __shared__ uint32_t s_key[HS];
    __shared__ uint32_t s_map[HS];
    __shared__ int s_excl[256], s_q0[256], s_ma[256], s_sc[4];
    if (int(blockIdx.x) >= count) return;
    const int i = rows[blockIdx.x];
    if (ub_off[i + 1] - ub_off[i] <= kWaveUb) return;  // k_symbolic_wave's row
    symbolic_row<256>(i, threadIdx.x, s_key, s_map, row_table(ub_off, i, HS), s_excl, s_q0, s_ma,
                      s_sc, g.ap, g.ac, g.am,
                      g.bp, g.bc, g.bm, ub_off[i], tcol, tmap, cnt_out, nnz_out);
}

// ---- bin 2, first try: one workgroup per row, the LDS table of bin 1 (the
// rows' distinct C blocks are usually far fewer than their upper bound)
__global__ __launch_bounds__(256) void k_symbolic_try(int count, const int32_t *__restrict__ rows,
                                                      SpgemmArgs g,
                                                      const int64_t *__restrict__ ub_off,
                                                      int32_t *tcol, uint32_t *tmap,
                                                      int32_t *cnt_out, int32_t *nnz_out,
                                                      int32_t *retry) {
    constexpr int HS = 2 * kBin1Ub;
    __shared__ uint32_t s_key[HS];
    __shared__ uint32_t s_map[HS];
    __shared__ int s_excl[256], s_q0[256], s_ma[256], s_sc[4], s_flag[2];
    if (int(blockIdx.x) >= count) return;
    const int i = rows[blockIdx.x];
    symbolic_row<256, true>(i, threadIdx.x, s_key, s_map, HS, s_excl, s_q0, s_ma, s_sc, g.ap,
                            g.ac, g.am, g.bp, g.bc, g.bm, ub_off[i], tcol, tmap, cnt_out,
                            nnz_out, s_flag, retry);
}

// ---- bin 2: workgroups loop over rows, each with a global slab of up to hs slots
__global__ __launch_bounds__(256) void k_symbolic_global(int count, const int32_t *__restrict__ rows,
                                                         SpgemmArgs g,
                                                         const int64_t *__restrict__ ub_off,
                                                         int32_t *tcol, uint32_t *tmap,
                                                         int32_t *cnt_out, int32_t *nnz_out,
                                                         uint32_t *slab, int hs,
                                                         const int32_t *__restrict__ retry) {
    __shared__ int s_excl[256], s_q0[256], s_ma[256], s_sc[4];
    uint32_t *hkey = slab + int64_t(blockIdx.x) * 2 * hs;
    uint32_t *hmap = hkey + hs;
    for (int r = blockIdx.x; r < count; r += gridDim.x) {
        const int i = rows[r];
        if (!retry[i]) continue;  // completed in LDS by k_symbolic_try
        symbolic_row<256>(i, threadIdx.x, hkey, hmap, row_table(ub_off, i, hs), s_excl, s_q0,
                          s_ma, s_sc, g.ap, g.ac,
                          g.am, g.bp, g.bc, g.bm, ub_off[i], tcol, tmap, cnt_out, nnz_out);
    }
}

// ub_i = sum over A blocks (i,k) of |B row k|; the workgroup's partial bins
// rows (SymbolicBin: register kernel, <= kBin1Ub, larger) and keeps the
// largest ub.
__global__ __launch_bounds__(256) void k_upper_bound(int32_t brow, const int32_t *__restrict__ ap,
                                                     const int32_t *__restrict__ ac,
                                                     const int32_t *__restrict__ bp,
                                                     int64_t *__restrict__ ub,
                                                     rowpass::Part *parts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t v = 0, vb = 0;
    if (i < brow) {
        for (int p = ap[i]; p < ap[i + 1]; ++p) {
            const int k = ac[p];
            v += bp[k + 1] - bp[k];
        }
        ub[i] = v;
        vb = (v <= 64 && ap[i + 1] - ap[i] <= 64) ? 0 : (v > 0 ? v : 1);  // as SymbolicBin
    }
    rowpass::write_part(i < brow, vb, v, 0, kBin1Ub, parts);
}

// (rows: the block rows to copy, or every row 0 .. count-1 when NULL; ub_off
// NULL: the row's slot is at row * stride)
__global__ __launch_bounds__(256) void k_compact_c(int32_t count, const int32_t *__restrict__ rows,
                                                   const int64_t *__restrict__ ub_off,
                                                   int32_t stride,
                                                   const int32_t *__restrict__ cptr,
                                                   const int32_t *__restrict__ tcol,
                                                   const uint32_t *__restrict__ tmap,
                                                   int32_t *__restrict__ ccol,
                                                   uint16_t *__restrict__ cmap) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= count) return;
    const int i = rows ? rows[r] : r;
    const int64_t s = ub_off ? ub_off[i] : int64_t(i) * stride;
    const int p0 = cptr[i], n = cptr[i + 1] - p0;
    for (int t = lane; t < n; t += 64) {
        ccol[p0 + t] = tcol[s + t];
        cmap[p0 + t] = uint16_t(tmap[s + t]);
    }
}

// ---- numeric -------------------------------------------------------------------------
// Operand layout of v_mfma_f64_4x4x4f64 on gfx950 (measured with an in-bounds
// probe: unit B entries against distinct A entries), block slot b of the four:
//   A_b[i][k] in lane 16k + 4b + i,  B_b[k][n] in lane 16k + 4b + n,
//   D_b[i][n] in lane 16i + 4b + n.
// So in lane L: slot b = (L >> 2) & 3; the A operand is bit 4*(L&3) + (L>>4)
// of the A bitmap; the B operand and the result are bit 4*(L>>4) + (L&3).
__device__ __forceinline__ double block_product(double a, double b) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, 0.0, 0, 0, 0);
}

// FP32 / FP16 operands: the 16-block MFMAs, FP32 accumulation.  Operand layout
// (measured on the box): lane 4b + r holds row r of block b's A (a[k], k =
// 0..3) and column r of block b's B (b[k]); the result entry (v, r) of block b
// is register v of lane 4b + r.  FP16: one v_mfma_f32_4x4x4_16b_f16 per 16
// block products; FP32: four v_mfma_f32_4x4x1_16b_f32 (one per k), chained.
typedef _Float16 mf_half4 __attribute__((ext_vector_type(4)));
typedef float mf_float4 __attribute__((ext_vector_type(4)));

// operand element type of the 16-block MFMA for value type VT, and the
// value's bits read as that type (no conversion: the copies hold FP16 / FP32
// numbers)
template <typename VT> struct Mf16;
template <> struct Mf16<__half> {
    using T = _Float16;
    using V = mf_half4;
    __device__ static T load(const void *p, int64_t idx) {
        return static_cast<const _Float16 *>(p)[idx];
    }
};
template <> struct Mf16<float> {
    using T = float;
    using V = mf_float4;
    __device__ static T load(const void *p, int64_t idx) {
        return static_cast<const float *>(p)[idx];
    }
};

template <typename VT>
__device__ __forceinline__ mf_float4 mfma16(typename Mf16<VT>::V a, typename Mf16<VT>::V b,
                                            mf_float4 c);
template <>
__device__ __forceinline__ mf_float4 mfma16<__half>(mf_half4 a, mf_half4 b, mf_float4 c) {
    return __builtin_amdgcn_mfma_f32_4x4x4f16(a, b, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ mf_float4 mfma16<float>(mf_float4 a, mf_float4 b, mf_float4 c) {
#pragma unroll
    for (int k = 0; k < 4; ++k) c = __builtin_amdgcn_mfma_f32_4x4x1f32(a[k], b[k], c, 0, 0, 0);
    return c;
}

// this lane's operands of a product (A block map xa at value offset oa, B
// block map xb at ob) for the 16-block MFMAs: row r of A, column r of B
template <typename VT>
__device__ __forceinline__ void gather16(const SpgemmArgs &g, uint32_t xa, int oa, uint32_t xb,
                                         int ob, int r, typename Mf16<VT>::V &a,
                                         typename Mf16<VT>::V &b) {
    using T = typename Mf16<VT>::T;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int ea = 4 * r + k, eb = 4 * k + r;
        a[k] = ((xa >> ea) & 1u) ? Mf16<VT>::load(g.av, oa + __popc(xa & ((1u << ea) - 1u))) : T(0);
        b[k] = ((xb >> eb) & 1u) ? Mf16<VT>::load(g.bv, ob + __popc(xb & ((1u << eb) - 1u))) : T(0);
    }
}

__device__ __forceinline__ int lower_bound_key(const int32_t *keys, int n, int key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// LDS staging of one wave in the numeric pass: the current chunk of A blocks
// and the compacted list of products whose bitmap product is non-zero.
struct NumStage {
    int32_t a_bq0[64], a_pst[64], a_ma[64], a_aoff[64];
    int32_t p_ma[64], p_aoff[64], p_mb[64], p_boff[64], p_slot[64];
};

// FP32 / FP16 rows of the LDS and global-slab bins (FP64 rows:
// numeric_row_outer).  One wave per C block row; acc points to LDS (cnt <=
// kNumCap) or a global slab, keys to the row's block columns (LDS copy or
// global).  The row's products (A block, B block) are flattened over the 64
// lanes, so the gathers of a chunk of 64 products are all in flight together;
// products with an empty bitmap product are dropped; the rest run 16 at a
// time on one 16-block MFMA (block slot b16 = product g0 + b16).  Products
// may share a C block, so their results are added one block slot after the
// other (fixed order: deterministic, no atomics).  boff[q] is the value
// offset of B block q (values are stored in block order).
// kCmp: acc holds only the row's C values, in their final (block, bit) order:
// entry e of block t at scvo[t] + popc(scm[t] below e); an entry outside the C
// bitmap only ever receives +0.0 products and is not stored.  The additions
// into stored entries are the same, in the same order, as with dense blocks.
template <typename VT, bool kCmp = false>
__device__ void numeric_row(int i, const SpgemmArgs &g, const int32_t *__restrict__ boff,
                            const int32_t *__restrict__ cptr, const int32_t *__restrict__ ccol,
                            const uint16_t *__restrict__ cmap, const int32_t *__restrict__ cvp,
                            double *__restrict__ cval, int32_t *skey, double *acc,
                            NumStage &st, int32_t *scvo = nullptr, uint32_t *scm = nullptr) {
    static_assert(!std::is_same<VT, double>::value, "FP64 rows: numeric_row_outer");
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    const int c0 = cptr[i], cnt = cptr[i + 1] - c0;
    const int32_t *keys = ccol + c0;
    if (skey) {
        for (int t = lane; t < cnt; t += 64) skey[t] = keys[t];
        keys = skey;
    }
    int nacc = cnt * 16;
    if constexpr (kCmp) {
        // value offset of every C block of the row: a scan of the bitmap popcounts
        int run = 0;
        for (int t0 = 0; t0 < cnt; t0 += 64) {
            const int t = t0 + lane;
            const uint32_t m = t < cnt ? uint32_t(cmap[c0 + t]) : 0u;
            const int pc = __popc(m);
            int inc = pc;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int v = __shfl_up(inc, d);
                if (lane >= d) inc += v;
            }
            if (t < cnt) {
                scm[t] = m;
                scvo[t] = run + inc - pc;
            }
            run += __shfl(inc, 63);
        }
        nacc = run;
    }
    for (int x = lane; x < nacc; x += 64) acc[x] = 0.0;
    wave_sync();
    int aoff_run = g.avp[i];
    const int a_beg = g.ap[i], a_end = g.ap[i + 1];
    for (int a0 = a_beg; a0 < a_end; a0 += 64) {
        const int p = a0 + lane;
        int ma = 0, bq0 = 0, blen = 0;
        if (p < a_end) {
            const int k = g.ac[p];
            ma = g.am[p];
            bq0 = g.bp[k];
            blen = g.bp[k + 1] - bq0;
        }
        const int pc = __popc(uint32_t(ma));
        int pinc = pc, linc = blen;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t1 = __shfl_up(pinc, d), t2 = __shfl_up(linc, d);
            if (lane >= d) {
                pinc += t1;
                linc += t2;
            }
        }
        const int ptot = __shfl(linc, 63), atot = __shfl(pinc, 63);
        wave_sync();  // the previous chunk's staging is consumed
        st.a_bq0[lane] = bq0;
        st.a_pst[lane] = linc - blen;
        st.a_ma[lane] = ma;
        st.a_aoff[lane] = aoff_run + pinc - pc;
        aoff_run += atot;
        const int na = min(64, a_end - a0);
        wave_sync();
        for (int t0 = 0; t0 < ptot; t0 += 64) {
            const int t = t0 + lane;
            bool valid = false;
            int pma = 0, pmb = 0, paoff = 0, pboff = 0, slot = 0;
            if (t < ptot) {
                // owner: the last A block of the chunk whose first product is <= t
                int lo = 0, hi = na - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (st.a_pst[mid] <= t) lo = mid; else hi = mid - 1;
                }
                const int q = st.a_bq0[lo] + (t - st.a_pst[lo]);
                pma = st.a_ma[lo];
                pmb = g.bm[q];
                if (bitmap_mul(uint32_t(pma), uint32_t(pmb))) {
                    valid = true;
                    paoff = st.a_aoff[lo];
                    pboff = boff[q];
                    slot = lower_bound_key(keys, cnt, g.bc[q]);
                }
            }
            const uint64_t vm = __ballot(valid);
            const int np = __popcll(vm);
            wave_sync();  // the previous product chunk is consumed
            if (valid) {
                const int r = __popcll(vm & lt);
                st.p_ma[r] = pma;
                st.p_mb[r] = pmb;
                st.p_aoff[r] = paoff;
                st.p_boff[r] = pboff;
                st.p_slot[r] = slot;
            }
            wave_sync();
            // FP32 / FP16: 16 products per 16-block MFMA, slot b16 = lane / 4
            // takes product g0 + b16; their results are added one slot
            // after the other (fixed order: deterministic, no atomics)
            const int b16 = lane >> 2, r4 = lane & 3;
            for (int g0 = 0; g0 < np; g0 += 16) {
                const int idx = g0 + b16;
                const bool on16 = idx < np;
                typename Mf16<VT>::V fa = {}, fb = {};
                int sl = 0;
                if (on16) {
                    sl = st.p_slot[idx];
                    gather16<VT>(g, uint32_t(st.p_ma[idx]), st.p_aoff[idx],
                                 uint32_t(st.p_mb[idx]), st.p_boff[idx], r4, fa, fb);
                }
                const mf_float4 d = mfma16<VT>(fa, fb, mf_float4{0.f, 0.f, 0.f, 0.f});
                // slot b2's 16 entries go to lanes 0..15 (entry e = lane, from
                // register e / 4 of lane 4 b2 + e % 4), which add them
                const int e16 = lane & 15, v16 = e16 >> 2;
                const int nslot = min(16, np - g0);  // wave-uniform
                for (int b2 = 0; b2 < nslot; ++b2) {
                    const int from = 4 * b2 + (e16 & 3);
                    const float t0 = __shfl(d[0], from), t1 = __shfl(d[1], from),
                                t2 = __shfl(d[2], from), t3 = __shfl(d[3], from);
                    const float val = v16 == 0 ? t0 : (v16 == 1 ? t1 : (v16 == 2 ? t2 : t3));
                    const int sl2 = __shfl(sl, 4 * b2);
                    if (lane < 16) {
                        if constexpr (kCmp) {
                            const uint32_t m = scm[sl2];
                            if ((m >> e16) & 1u)
                                acc[scvo[sl2] + __popc(m & ((1u << e16) - 1u))] += double(val);
                        } else {
                            acc[sl2 * 16 + e16] += double(val);
                        }
                    }
                    wave_sync();
                }
            }
        }
    }
    wave_sync();
    if constexpr (kCmp) {  // already in the value order of the row
        const int v0 = cvp[i];
        for (int x = lane; x < nacc; x += 64) cval[v0 + x] = round_out<VT>(acc[x]);
        return;
    }
    // compress: four C blocks per step, slot gs takes block t0 + gs
    const int gs = lane >> 4, e = lane & 15;
    const uint32_t below = (1u << e) - 1u;
    int voff = cvp[i];
    for (int t0 = 0; t0 < cnt; t0 += 4) {
        const int t = t0 + gs;
        const uint32_t m = (t < cnt) ? uint32_t(cmap[c0 + t]) : 0u;
        const int pc = __popc(m);
        const int p0c = __shfl(pc, 0), p1c = __shfl(pc, 16), p2c = __shfl(pc, 32),
                  p3c = __shfl(pc, 48);
        const int excl = (gs > 0 ? p0c : 0) + (gs > 1 ? p1c : 0) + (gs > 2 ? p2c : 0);
        if ((m >> e) & 1u) cval[voff + excl + __popc(m & below)] = round_out<VT>(acc[t * 16 + e]);
        voff += p0c + p1c + p2c + p3c;
    }
}

// FP64 rows of the LDS and global-slab bins (compressed or dense
// accumulators), four products per MFMA that all come from one A block: each A block's products are padded to a multiple
// of four positions, so the four block slots of an instruction hold products
// of one B row, whose block columns are distinct, and the four results are
// added to their C blocks' compressed entries in one step (numeric_row adds
// the four slots one after the other, since its flattened products may share
// a C block).  A chunk of 64 padded positions is staged at a time: lane t
// finds its position's A block (binary search over the padded starts), the
// product's bitmap, B value offset and C slot; then 16 MFMAs read their
// operands' descriptions from LDS.  Every C entry sums its products in A-block
// order: deterministic, no atomics.
struct OuterStage {
    int32_t a_bq0[64], a_pps[64], a_ma[64], a_aoff[64], a_blen[64];
    int32_t p_info[64], p_boff[64], p_ma[64], p_aoff[64];
};

// kCmp: compressed accumulators (scvo / scm as in numeric_row); else dense
// 4x4 blocks.  skey: LDS copy of the row's block columns, or NULL to search
// them in C's column array; acc: LDS or the wave's global slab.
template <bool kCmp>
__device__ void numeric_row_outer(int i, const SpgemmArgs &g, const int32_t *__restrict__ boff,
                                  const int32_t *__restrict__ cptr,
                                  const int32_t *__restrict__ ccol,
                                  const uint16_t *__restrict__ cmap,
                                  const int32_t *__restrict__ cvp, double *__restrict__ cval,
                                  int32_t *skey, double *acc, OuterStage &st,
                                  int32_t *scvo = nullptr, uint32_t *scm = nullptr) {
    const int lane = threadIdx.x & 63;
    const int bs = (lane >> 2) & 3;
    const int eA = 4 * (lane & 3) + (lane >> 4);
    const int eB = 4 * (lane >> 4) + (lane & 3);
    const uint32_t belowA = (1u << eA) - 1u, belowB = (1u << eB) - 1u;
    const int c0 = cptr[i], cnt = cptr[i + 1] - c0;
    const int32_t *keys = ccol + c0;
    if (skey) {
        for (int t = lane; t < cnt; t += 64) skey[t] = keys[t];
        keys = skey;
    }
    int nacc = cnt * 16;
    if constexpr (kCmp) {
        // value offset of every C block of the row: a scan of the bitmap popcounts
        int run = 0;
        for (int t0 = 0; t0 < cnt; t0 += 64) {
            const int t = t0 + lane;
            const uint32_t m = t < cnt ? uint32_t(cmap[c0 + t]) : 0u;
            const int pc = __popc(m);
            const int inc = wave_incl_sum(pc);
            if (t < cnt) {
                scm[t] = m;
                scvo[t] = run + inc - pc;
            }
            run += __builtin_amdgcn_readlane(inc, 63);
        }
        nacc = run;
    }
    for (int x = lane; x < nacc; x += 64) acc[x] = 0.0;
    wave_sync();
    // whole-array resources over A's and B's values (used when g.fp64_rs)
    const __amdgpu_buffer_rsrc_t a_rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(g.av), short(0),
                                          g.fp64_rs ? int(uint32_t(g.annz * 8)) : 0, kBufferDword3);
    const __amdgpu_buffer_rsrc_t b_rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(g.bv), short(0),
                                          g.fp64_rs ? int(uint32_t(g.bnnz * 8)) : 0, kBufferDword3);
    int aoff_run = g.avp[i];
    const int a_beg = g.ap[i], a_end = g.ap[i + 1];
    for (int a0 = a_beg; a0 < a_end; a0 += 64) {
        const int p = a0 + lane;
        int ma = 0, bq0 = 0, blen = 0;
        if (p < a_end) {
            const int k = g.ac[p];
            ma = g.am[p];
            bq0 = g.bp[k];
            blen = g.bp[k + 1] - bq0;
        }
        const int pc = __popc(uint32_t(ma));
        const int pinc = wave_incl_sum(pc);
        const int pad = (blen + 3) & ~3;
        const int ppinc = wave_incl_sum(pad);
        const int ptot = __builtin_amdgcn_readlane(ppinc, 63);
        const int na = min(64, a_end - a0);
        wave_sync();  // the previous chunk's staging is consumed
        st.a_bq0[lane] = bq0;
        st.a_pps[lane] = ppinc - pad;
        st.a_ma[lane] = ma;
        st.a_aoff[lane] = aoff_run + pinc - pc;
        st.a_blen[lane] = blen;
        aoff_run += __builtin_amdgcn_readlane(pinc, 63);
        wave_sync();
        for (int t0 = 0; t0 < ptot; t0 += 64) {
            const int t = t0 + lane;
            int info = -1, pboff = 0, pma = 0, paoff = 0;
            if (t < ptot) {
                // owner: the last A block of the chunk whose padded start is <= t
                int lo = 0, hi = na - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (st.a_pps[mid] <= t) lo = mid; else hi = mid - 1;
                }
                const int r = t - st.a_pps[lo];
                pma = st.a_ma[lo];
                paoff = st.a_aoff[lo];
                if (r < st.a_blen[lo]) {
                    const int q = st.a_bq0[lo] + r;
                    const uint32_t mb = g.bm[q];
                    if (bitmap_mul(uint32_t(pma), mb)) {
                        info = int(mb) | (lower_bound_key(keys, cnt, g.bc[q]) << 16);
                        pboff = boff[q];
                    }
                }
            }
            wave_sync();  // the previous positions are consumed
            st.p_info[lane] = info;
            st.p_boff[lane] = pboff;
            st.p_ma[lane] = pma;
            st.p_aoff[lane] = paoff;
            wave_sync();
            const int ng = min(64, ptot - t0) >> 2;  // whole groups of four (padded)
            for (int g0 = 0; g0 < ng; g0 += 4) {
                double av[4], bv[4];
                int inf[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int idx = 4 * (g0 + u) + bs;  // < 64
                    inf[u] = g0 + u < ng ? st.p_info[idx] : -1;
                    const uint32_t xa = uint32_t(st.p_ma[idx]);
                    const uint32_t xb = uint32_t(inf[u]) & 0xFFFFu;
                    const bool ina = (inf[u] >= 0) & (((xa >> eA) & 1u) != 0u);
                    const bool inb = (inf[u] >= 0) & (((xb >> eB) & 1u) != 0u);
                    const int ia = st.p_aoff[idx] + __popc(xa & belowA);
                    const int ib = st.p_boff[idx] + __popc(xb & belowB);
                    if (g.fp64_rs) {
                        // raw buffer loads: an offset past the resource reads 0.0
                        const auto a = __builtin_amdgcn_raw_buffer_load_b64(
                            a_rs, ina ? uint32_t(ia) * 8u : kNoEntryOff, 0, 0);
                        const auto b = __builtin_amdgcn_raw_buffer_load_b64(
                            b_rs, inb ? uint32_t(ib) * 8u : kNoEntryOff, 0, 0);
                        __builtin_memcpy(&av[u], &a, 8);
                        __builtin_memcpy(&bv[u], &b, 8);
                    } else {
                        av[u] = ina ? ldv<double>(g.av, ia) : 0.0;
                        bv[u] = inb ? ldv<double>(g.bv, ib) : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (g0 + u >= ng) break;  // wave-uniform: the MFMA runs with full EXEC
                    const double d = block_product(av[u], bv[u]);
                    if (inf[u] >= 0) {
                        const int sl = inf[u] >> 16;
                        if constexpr (kCmp) {
                            const uint32_t m = scm[sl];
                            if ((m >> eB) & 1u) acc[scvo[sl] + __popc(m & belowB)] += d;
                        } else {
                            acc[sl * 16 + eB] += d;
                        }
                    }
                    wave_sync();  // a C entry's next product may come from another lane
                }
            }
        }
    }
    wave_sync();
    if constexpr (kCmp) {  // already in the value order of the row
        const int v0 = cvp[i];
        for (int x = lane; x < nacc; x += 64) cval[v0 + x] = acc[x];
        return;
    }
    // compress: four C blocks per step, slot gs takes block t0 + gs
    const int gs = lane >> 4, e = lane & 15;
    const uint32_t below = (1u << e) - 1u;
    int voff = cvp[i];
    for (int t0 = 0; t0 < cnt; t0 += 4) {
        const int t = t0 + gs;
        const uint32_t m = (t < cnt) ? uint32_t(cmap[c0 + t]) : 0u;
        const int pc = __popc(m);
        const int p0c = __shfl(pc, 0), p1c = __shfl(pc, 16), p2c = __shfl(pc, 32),
                  p3c = __shfl(pc, 48);
        const int excl = (gs > 0 ? p0c : 0) + (gs > 1 ? p1c : 0) + (gs > 2 ? p2c : 0);
        if ((m >> e) & 1u) cval[voff + excl + __popc(m & below)] = acc[t * 16 + e];
        voff += p0c + p1c + p2c + p3c;
    }
}

template <typename VT>
__global__ __launch_bounds__(256) void k_numeric_lds(int count, const int32_t *__restrict__ rows,
                                                     SpgemmArgs g,
                                                     const int32_t *__restrict__ boff,
                                                     const int32_t *__restrict__ cptr,
                                                     const int32_t *__restrict__ ccol,
                                                     const uint16_t *__restrict__ cmap,
                                                     const int32_t *__restrict__ cvp,
                                                     double *__restrict__ cval) {
    __shared__ double s_acc[4][kNumCap * 16];
    __shared__ int32_t s_key[4][kNumCap];
    using Stage = typename std::conditional<std::is_same<VT, double>::value, OuterStage,
                                            NumStage>::type;
    __shared__ Stage s_st[4];
    const int w = threadIdx.x >> 6;
    const int r = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + w);
    if (r >= count) return;
    const int i = __builtin_amdgcn_readfirstlane(rows[r]);
    if (fits_cmp<kNumCap, kCmpValShort>(cptr, cvp, i)) return;  // k_numeric_cmp's row
    if constexpr (std::is_same<VT, double>::value)
        numeric_row_outer<false>(i, g, boff, cptr, ccol, cmap, cvp, cval, s_key[w], s_acc[w],
                                 s_st[w]);
    else
        numeric_row<VT>(i, g, boff, cptr, ccol, cmap, cvp, cval, s_key[w], s_acc[w], s_st[w]);
}

// rows that fit compressed accumulators of BLK blocks and VAL values (and not
// those of a smaller launch, SBLK / SVAL): one wave per row.  For the long bin
// <kCmpMidBlk, kCmpMidVal> (about 9 KB of LDS per wave, four workgroups per
// CU), then <kCmpBlk, kCmpVal> (about 16 KB per wave, two workgroups per CU);
// <kNumCap, kCmpValShort> for the bin of at most 64 blocks (about 5 KB per
// wave, against 10 KB for dense blocks)
template <typename VT, int BLK, int VAL, int SBLK = 0, int SVAL = 0>
__global__ __launch_bounds__(256) void k_numeric_cmp(int count, const int32_t *__restrict__ rows,
                                                     SpgemmArgs g,
                                                     const int32_t *__restrict__ boff,
                                                     const int32_t *__restrict__ cptr,
                                                     const int32_t *__restrict__ ccol,
                                                     const uint16_t *__restrict__ cmap,
                                                     const int32_t *__restrict__ cvp,
                                                     double *__restrict__ cval) {
    __shared__ double s_acc[4][VAL];
    __shared__ int32_t s_key[4][BLK];
    __shared__ int32_t s_cvo[4][BLK];
    __shared__ uint32_t s_cm[4][BLK];
    using Stage = typename std::conditional<std::is_same<VT, double>::value, OuterStage,
                                            NumStage>::type;
    __shared__ Stage s_st[4];
    const int w = threadIdx.x >> 6;
    const int r = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + w);
    if (r >= count) return;  // whole wave
    const int i = __builtin_amdgcn_readfirstlane(rows[r]);
    if (!fits_cmp<BLK, VAL>(cptr, cvp, i)) return;  // the dense kernel's row
    if constexpr (SBLK > 0)
        if (fits_cmp<SBLK, SVAL>(cptr, cvp, i)) return;  // the smaller launch's row
    if constexpr (std::is_same<VT, double>::value)
        numeric_row_outer<true>(i, g, boff, cptr, ccol, cmap, cvp, cval, s_key[w], s_acc[w],
                                s_st[w], s_cvo[w], s_cm[w]);
    else
        numeric_row<VT, true>(i, g, boff, cptr, ccol, cmap, cvp, cval, s_key[w], s_acc[w],
                              s_st[w], s_cvo[w], s_cm[w]);
}

template <typename VT>
__global__ __launch_bounds__(256) void k_numeric_global(int count, const int32_t *__restrict__ rows,
                                                        SpgemmArgs g,
                                                        const int32_t *__restrict__ boff,
                                                        const int32_t *__restrict__ cptr,
                                                        const int32_t *__restrict__ ccol,
                                                        const uint16_t *__restrict__ cmap,
                                                        const int32_t *__restrict__ cvp,
                                                        double *__restrict__ cval, double *slab,
                                                        int64_t slab_elems) {
    // rows up to kBigCap blocks accumulate in LDS, longer ones in the wave's slab
    __shared__ double s_acc[4][kBigCap * 16];
    __shared__ int32_t s_key[4][kBigCap];
    using Stage = typename std::conditional<std::is_same<VT, double>::value, OuterStage,
                                            NumStage>::type;
    __shared__ Stage s_st[4];
    const int w = threadIdx.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + w);
    const int nwaves = gridDim.x * 4;
    double *acc = slab + int64_t(wave) * slab_elems;
    for (int r = wave; r < count; r += nwaves) {
        const int i = __builtin_amdgcn_readfirstlane(rows[r]);
        if (fits_cmp<kCmpBlk, kCmpVal>(cptr, cvp, i)) continue;  // k_numeric_cmp's row
        const bool in_lds = cptr[i + 1] - cptr[i] <= kBigCap;
        if constexpr (std::is_same<VT, double>::value) {
            if (in_lds)
                numeric_row_outer<false>(i, g, boff, cptr, ccol, cmap, cvp, cval, s_key[w],
                                         s_acc[w], s_st[w]);
            else
                numeric_row_outer<false>(i, g, boff, cptr, ccol, cmap, cvp, cval, nullptr, acc,
                                         s_st[w]);
        } else {
            if (in_lds)
                numeric_row<VT>(i, g, boff, cptr, ccol, cmap, cvp, cval, s_key[w], s_acc[w],
                                s_st[w]);
            else
                numeric_row<VT>(i, g, boff, cptr, ccol, cmap, cvp, cval, nullptr, acc, s_st[w]);
        }
    }
}

// The register bin in FP64, outer-product order: for each A block of the row
// (in stored order) its products with the blocks of B's row, four per
// v_mfma_f64_4x4x4f64: the A block in every block slot, four consecutive
// blocks of the B row in the slots.  A B row never repeats a block column, so
// the four results belong to four distinct C blocks; each is added to its C
// block's dense 4x4 accumulator in LDS.  One wave owns a row and its LDS
// accumulators, and its LDS accesses execute in program order, so every C
// entry sums its products in A-block order: deterministic, no atomics.
// (Grouping the products by C block instead, one C block per block slot with
// the products chained through the MFMA accumulator, pads every group of four
// slots to its longest one and needs per-slot ballots and four shuffles per
// product: 0.85 against 0.70 ms for the numeric pass of A^2 on 7-pt 128^3.)
// Lane t of the prologue holds product t (owner A block and C slot found
// with branchless shuffle searches).  CAPB: C blocks per row the wave's
// accumulators hold (at least the bin's longest C row).
template <int CAPB>
__global__ __launch_bounds__(256) void k_numeric_outer(int count, const int32_t *__restrict__ rows,
                                                       SpgemmArgs g,
                                                       const int32_t *__restrict__ boff,
                                                       const int32_t *__restrict__ cptr,
                                                       const int64_t *__restrict__ ub_off,
                                                       int32_t stride,
                                                       const int32_t *__restrict__ tcol,
                                                       const uint32_t *__restrict__ tmap,
                                                       int32_t *__restrict__ ccol,
                                                       uint16_t *__restrict__ cmap,
                                                       const int32_t *__restrict__ cvp,
                                                       double *__restrict__ cval) {
    __shared__ double s_acc[4][CAPB * 16];
    __shared__ int32_t s_own[4][64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int r = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + w);
    if (r >= count) return;  // whole wave; the waves share no LDS
    double *acc = s_acc[w];
    int32_t *own_of = s_own[w];
    // the row's scalars first, all in flight together (wave-uniform: scalar loads)
    const int i = __builtin_amdgcn_readfirstlane(rows ? rows[r] : r);
    const int c0 = cptr[i], cnt = cptr[i + 1] - c0;  // cnt <= min(64, CAPB): the bin's rule
    const int64_t so = ub_off ? ub_off[i] : int64_t(i) * stride;
    const int a_beg = g.ap[i], na = g.ap[i + 1] - a_beg;
    const int avp0 = g.avp[i];
    const int vbase = cvp[i];
    const int bs = (lane >> 2) & 3;
    const int eA = 4 * (lane & 3) + (lane >> 4);
    const int eB = 4 * (lane >> 4) + (lane & 3);
    const uint32_t belowA = (1u << eA) - 1u, belowB = (1u << eB) - 1u;
    // C's block columns and bitmaps as the symbolic pass left them in the
    // row's slot of the temporaries, copied to C here
    const int key = lane < cnt ? tcol[so + lane] : INT_MAX;
    const uint32_t cmy = lane < cnt ? tmap[so + lane] : 0u;
    int ma = 0, bq0 = 0, blen = 0;
    if (lane < na) {
        const int k = g.ac[a_beg + lane];
        ma = g.am[a_beg + lane];
        bq0 = g.bp[k];
        blen = g.bp[k + 1] - bq0;
    }
    if (lane < cnt) {
        ccol[c0 + lane] = key;
        cmap[c0 + lane] = uint16_t(cmy);
    }
    for (int e = lane; e < cnt * 16; e += 64) acc[e] = 0.0;
    own_of[lane] = 0;
    const int pc = __popc(uint32_t(ma));
    const int aoff = avp0 + wave_incl_sum(pc) - pc;
    const int linc = wave_incl_sum(blen);
    const int pst = linc - blen;
    const int ptot = __builtin_amdgcn_readlane(linc, 63);
    // product t = lane: owner = last A block j < na with pst[j] <= t: every
    // A block with products marks its first product, a max-scan spreads it
    wave_sync();
    if (blen > 0) own_of[pst] = lane;  // pst < ptot <= 64, distinct
    wave_sync();
    const int t = lane;
    const int own = wave_incl_max(own_of[lane]);
    const int q = __shfl(bq0, own) + (t - __shfl(pst, own));
    const int pma = __shfl(ma, own);
    int pmb = 0, pboff = 0, bcol = 0;
    if (t < ptot) {
        pmb = g.bm[q];
        bcol = g.bc[q];
        pboff = boff[q];
    }
    const bool valid = (t < ptot) && bitmap_mul(uint32_t(pma), uint32_t(pmb)) != 0u;
    int slot = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
        const int cand = slot + step;
        const int kv = __shfl(key, cand - 1);
        if (cand <= cnt && kv < bcol) slot = cand;
    }
    // product record: B bitmap and C slot, or -1 for a product with an empty
    // bitmap product (it adds nothing)
    const int pinfo = valid ? (pmb | (slot << 16)) : -1;
    // Operand loads are raw buffer loads whose range check returns 0.0 for
    // an offset past the resource, so an entry outside the bitmap (offset
    // kNoEntry) needs no branch.  A values: the row's own, from avp0.
    constexpr uint32_t kNoEntry = 0xFFFFFFF0u;
    const double *av_row = static_cast<const double *>(g.av) + avp0;
    const __amdgpu_buffer_rsrc_t a_rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<double *>(av_row), short(0), (g.avp[i + 1] - avp0) * 8, kBufferDword3);
    // lane's operands of step s of A block j: the A block's entry eA, and
    // entry eB of B block s * 4 + bs of the A block's B row (bq: B's value
    // offset the resource b_rs starts at).  The product's lane is addressed
    // directly (ds_bpermute uses the lane bits only); a lane past the A
    // block's products discards what it fetched.
    auto gather_b = [&](int st, int bl, int s4, int bq, __amdgpu_buffer_rsrc_t b_rs,
                        int &info) -> double {
        const bool in = s4 + bs < bl;
        const int addr = (st + s4 + bs) << 2;
        const int inf = __builtin_amdgcn_ds_bpermute(addr, pinfo);
        const int ob = __builtin_amdgcn_ds_bpermute(addr, pboff);
        info = in ? inf : -1;
        const uint32_t xb = uint32_t(info) & 0xFFFFu;
        const uint32_t idx = uint32_t(ob - bq + __popc(xb & belowB)) * 8u;  // no branch
        const uint32_t off = ((info >= 0) & (((xb >> eB) & 1u) != 0u)) ? idx : kNoEntry;
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(b_rs, off, 0, 0);
        double d;
        __builtin_memcpy(&d, &v, 8);
        return d;
    };
    auto add_step = [&](double av, double bv, int info) {
        const double d = __builtin_amdgcn_mfma_f64_4x4x4f64(av, bv, 0.0, 0, 0, 0);
        // lane (slot bs, entry eB) holds entry eB of its product's C block
        if (info >= 0) acc[(info >> 16) * 16 + eB] += d;
    };
    wave_sync();  // the zeroed accumulators, before other lanes add to them
    // four A blocks at a time: the operands of their first eight products
    // each are gathered together, then multiplied in A-block order
    for (int j0 = 0; j0 < na; j0 += 4) {  // na is wave-uniform
        double av[4], bv[4][2];
        int info[4][2], st[4], bl[4], bq[4];
        __amdgpu_buffer_rsrc_t b_rs[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + u < na ? j0 + u : 0;
            const uint32_t xa = uint32_t(__builtin_amdgcn_readlane(ma, j));
            const int oa = __builtin_amdgcn_readlane(aoff, j);
            st[u] = __builtin_amdgcn_readlane(pst, j);
            bl[u] = j0 + u < na ? __builtin_amdgcn_readlane(blen, j) : 0;
            // the B row's first value: B's value offset of the block's first product
            bq[u] = __builtin_amdgcn_readlane(pboff, st[u] < 64 ? st[u] : 63);
            const int64_t brest = g.bnnz - bq[u];
            b_rs[u] = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<double *>(static_cast<const double *>(g.bv) + bq[u]), short(0),
                int(brest < 0x1FFFFFFE ? brest * 8 : 0xFFFFFFF0), kBufferDword3);
            const uint32_t aidx = uint32_t(oa - avp0 + __popc(xa & belowA)) * 8u;
            const uint32_t aoff_e = ((bl[u] > 0) & (((xa >> eA) & 1u) != 0u)) ? aidx : kNoEntry;
            const auto a = __builtin_amdgcn_raw_buffer_load_b64(a_rs, aoff_e, 0, 0);
            __builtin_memcpy(&av[u], &a, 8);
#pragma unroll
            for (int v = 0; v < 2; ++v)
                bv[u][v] = gather_b(st[u], bl[u], 4 * v, bq[u], b_rs[u], info[u][v]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int v = 0; v < 2; ++v)
                if (4 * v < bl[u]) add_step(av[u], bv[u][v], info[u][v]);  // uniform
            // B rows of more than eight blocks: the rest of this A block's products
            for (int s4 = 8; s4 < bl[u]; s4 += 4) {
                int inf;
                const double b = gather_b(st[u], bl[u], s4, bq[u], b_rs[u], inf);
                add_step(av[u], b, inf);
            }
            wave_sync();  // a C entry's next product may come from another lane
        }
    }
    // compress by the C bitmaps: block t's values start at the exclusive scan
    // of the row's popcounts; four lanes per block, four bits each
    const int cpc = __popc(cmy);
    const int cexcl = wave_incl_sum(cpc) - cpc;
    const int qd = lane & 3;
    for (int t0 = 0; t0 < cnt; t0 += 16) {
        const int tb = t0 + (lane >> 2);
        const int tc = tb < 64 ? tb : 63;
        const uint32_t cm = uint32_t(__shfl(int(cmy), tc));
        const int ex = __shfl(cexcl, tc);
        if (tb < cnt) {
            int o = vbase + ex + __popc(cm & ((1u << (4 * qd)) - 1u));
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if ((cm >> (4 * qd + b)) & 1u) cval[o++] = acc[tb * 16 + 4 * qd + b];
        }
    }
}

// The register bin for FP32 / FP16 operands: the product and slot assignment
// of k_numeric_outer's prologue; the 16-block MFMAs (FP32 accumulation)
// take 16 C blocks at a time: lane 4b + r works for C block g0 + b, gathers
// row r of A and column r of B of that block's next product and chains the
// block's products (in product order) through the MFMA accumulator.  Lane
// 4b + r then holds entries (v, r), v = 0..3, of the finished block.
template <typename VT>
__global__ __launch_bounds__(256) void k_numeric_reg16(int count, const int32_t *__restrict__ rows,
                                                       SpgemmArgs g,
                                                       const int32_t *__restrict__ boff,
                                                       const int32_t *__restrict__ cptr,
                                                       const int32_t *__restrict__ ccol,
                                                       const uint16_t *__restrict__ cmap,
                                                       const int32_t *__restrict__ cvp,
                                                       double *__restrict__ cval) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= count) return;  // whole wave
    const int i = rows[r];
    const int b16 = lane >> 2, r4 = lane & 3;
    const int c0 = cptr[i], cnt = cptr[i + 1] - c0;
    const int key = lane < cnt ? ccol[c0 + lane] : INT_MAX;
    const int a_beg = g.ap[i], na = g.ap[i + 1] - a_beg;
    int ma = 0, bq0 = 0, blen = 0;
    if (lane < na) {
        const int k = g.ac[a_beg + lane];
        ma = g.am[a_beg + lane];
        bq0 = g.bp[k];
        blen = g.bp[k + 1] - bq0;
    }
    const int pc = __popc(uint32_t(ma));
    int pinc = pc, linc = blen;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t1 = __shfl_up(pinc, d), t2 = __shfl_up(linc, d);
        if (lane >= d) {
            pinc += t1;
            linc += t2;
        }
    }
    const int aoff = g.avp[i] + pinc - pc;
    const int pst = linc - blen;
    const int ptot = __shfl(linc, 63);
    // product t = lane: owner = last A block j < na with pst[j] <= t
    const int t = lane;
    int own = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
        const int cand = own + step;
        const int v = __shfl(pst, cand < 64 ? cand : 63);
        if (cand < na && v <= t) own = cand;
    }
    const int q = __shfl(bq0, own) + (t - __shfl(pst, own));
    const int pma = __shfl(ma, own);
    const int paoff = __shfl(aoff, own);
    int pmb = 0, pboff = 0, bcol = 0;
    if (t < ptot) {
        pmb = g.bm[q];
        bcol = g.bc[q];
        pboff = boff[q];
    }
    const bool valid = (t < ptot) && bitmap_mul(uint32_t(pma), uint32_t(pmb)) != 0u;
    // C slot: number of C block columns below bcol
    int slot = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
        const int cand = slot + step;
        const int kv = __shfl(key, cand - 1);
        if (cand <= cnt && kv < bcol) slot = cand;
    }
    int voff = cvp[i];
    for (int g0 = 0; g0 < cnt; g0 += 16) {
        uint64_t mine = 0ull;
        int iters = 0;
#pragma unroll
        for (int sb = 0; sb < 16; ++sb) {
            const uint64_t m = __ballot(valid && slot == g0 + sb);
            if (sb == b16) mine = m;
            iters = max(iters, __popcll(m));
        }
        mf_float4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int it = 0; it < iters; it += 2) {
            // two products per step: both gathers before the MFMAs
            typename Mf16<VT>::V fa[2] = {}, fb[2] = {};
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const bool on = mine != 0ull;
                const int src = on ? __builtin_ctzll(mine) : 0;
                if (on) mine &= mine - 1ull;
                const uint32_t xa = uint32_t(__shfl(pma, src)), xb = uint32_t(__shfl(pmb, src));
                const int oa = __shfl(paoff, src), ob = __shfl(pboff, src);
                if (on) gather16<VT>(g, xa, oa, xb, ob, r4, fa[u], fb[u]);
            }
            acc = mfma16<VT>(fa[0], fb[0], acc);
            if (it + 1 < iters) acc = mfma16<VT>(fa[1], fb[1], acc);
        }
        // lane (b16, r4): entries (v, r4) of C block g0 + b16
        const int tb = g0 + b16;
        const uint32_t cm = tb < cnt ? uint32_t(cmap[c0 + tb]) : 0u;
        const int cpc = __popc(cm);
        // exclusive prefix of the 16 blocks' value counts (lanes 4b carry them)
        const int x = r4 == 0 ? cpc : 0;
        int inc = x;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(inc, d);
            if (lane >= d) inc += v;
        }
        const int excl = __shfl(inc, 4 * b16) - cpc;
        const int tot = __shfl(inc, 63);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int e = 4 * v + r4;
            if ((cm >> e) & 1u)
                cval[voff + excl + __popc(cm & ((1u << e) - 1u))] = round_out<VT>(double(acc[v]));
        }
        voff += tot;
    }
}

// Numeric bin of a C block row: 0 = register kernel (<= 64 products and A
// blocks), 1 = LDS accumulators (<= kNumCap C blocks), 2 = global slab.
struct NumericBin {
    const int64_t *ub;
    const int32_t *ap;
    const int32_t *ccnt;
    __device__ int64_t operator()(int i) const {
        if (ub[i] <= 64 && ap[i + 1] - ap[i] <= 64) return 0;
        return ccnt[i] <= kNumCap ? 1 : 2;
    }
};

// every row in the register bin (the bounded form of the symbolic pass)
struct RegBin {
    __device__ int64_t operator()(int) const { return 0; }
};

// ---- C's sizes in three kernels: every workgroup of 256 block rows sums its
// rows' block and value counts and bins its rows (k_csize_part), one
// workgroup scans the partials (k_csize_finish), then every workgroup writes
// its rows' block and value offsets and numeric list entries (k_csize_write).
// Together they replace two device scans and a row pass.
struct CSizePart {
    int32_t f[5];  // blocks, values, rows per numeric bin; after k_csize_finish: offsets
    int32_t mx;    // largest block count
    int32_t bad;   // a row that broke the register bound (block count -1)
    int32_t big[3];  // largest values of a bin-1 row, blocks and values of a bin-2 row
    unsigned long long sumsq;
};

struct CSizeTotals {
    RowStats rs;  // rows per numeric bin, largest block count, sum of squares
    int32_t nblk, nnz, bad;
    int32_t big[3];  // as CSizePart: which numeric launches have rows
};

template <typename FB>
__global__ __launch_bounds__(256) void k_csize_part(int32_t n, FB fb,
                                                    const int32_t *__restrict__ ccnt,
                                                    const int32_t *__restrict__ cnnz,
                                                    CSizePart *__restrict__ parts) {
    __shared__ int s_f[4][10];
    __shared__ unsigned long long s_sq[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool valid = i < n;
    const int c = valid ? ccnt[i] : 0, v = valid ? cnnz[i] : 0;
    const int b = valid ? int(fb(i)) : -1;
    const int cs = __builtin_amdgcn_readlane(wave_incl_sum(c), 63);
    const int vs = __builtin_amdgcn_readlane(wave_incl_sum(v), 63);
    const int mx = __builtin_amdgcn_readlane(wave_incl_max(c), 63);
    const int v1 = __builtin_amdgcn_readlane(wave_incl_max(b == 1 ? v : 0), 63);
    const int b2 = __builtin_amdgcn_readlane(wave_incl_max(b == 2 ? c : 0), 63);
    const int v2 = __builtin_amdgcn_readlane(wave_incl_max(b == 2 ? v : 0), 63);
    const bool bad = __ballot(c < 0) != 0ull;
    int nb[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) nb[q] = __popcll(__ballot(b == q));
    unsigned long long sq = c > 0 ? static_cast<unsigned long long>(c) * unsigned(c) : 0ull;
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
    if (lane == 0) {
        s_f[w][0] = cs;
        s_f[w][1] = vs;
        s_f[w][2] = nb[0];
        s_f[w][3] = nb[1];
        s_f[w][4] = nb[2];
        s_f[w][5] = mx;
        s_f[w][6] = bad ? 1 : 0;
        s_f[w][7] = v1;
        s_f[w][8] = b2;
        s_f[w][9] = v2;
        s_sq[w] = sq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        CSizePart p{};
#pragma unroll
        for (int f = 0; f < 5; ++f) p.f[f] = s_f[0][f] + s_f[1][f] + s_f[2][f] + s_f[3][f];
        p.mx = max(max(s_f[0][5], s_f[1][5]), max(s_f[2][5], s_f[3][5]));
        p.bad = s_f[0][6] | s_f[1][6] | s_f[2][6] | s_f[3][6];
#pragma unroll
        for (int q = 0; q < 3; ++q)
            p.big[q] = max(max(s_f[0][7 + q], s_f[1][7 + q]), max(s_f[2][7 + q], s_f[3][7 + q]));
        p.sumsq = s_sq[0] + s_sq[1] + s_sq[2] + s_sq[3];
        parts[blockIdx.x] = p;
    }
}

// one workgroup: partials -> per-workgroup offsets (five exclusive scans),
// totals, and the last entries of C's block and value offsets
__global__ __launch_bounds__(rowpass::kFinishThreads) void k_csize_finish(
    int32_t ng, CSizePart *__restrict__ parts, CSizeTotals *__restrict__ tot,
    int32_t *__restrict__ cptr_end, int32_t *__restrict__ cvp_end) {
    constexpr int NT = rowpass::kFinishThreads, NW = NT / 64;
    __shared__ int s_w[NW][5];
    __shared__ int s_mx[NW][4], s_bad[NW];
    __shared__ unsigned long long s_sq[NW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int run[5] = {0, 0, 0, 0, 0};
    int mx[4] = {0, 0, 0, 0}, bad = 0;
    unsigned long long sq = 0ull;
    for (int base = 0; base < ng; base += NT) {
        const int gi = base + threadIdx.x;
        CSizePart p{};
        if (gi < ng) p = parts[gi];
        int incl[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) incl[q] = wave_incl_sum(p.f[q]);
        if (lane == 63) {
#pragma unroll
            for (int q = 0; q < 5; ++q) s_w[w][q] = incl[q];
        }
        __syncthreads();
        int off[5], all[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            int before = 0, sum = 0;
            for (int k = 0; k < NW; ++k) {
                const int c = s_w[k][q];
                before += k < w ? c : 0;
                sum += c;
            }
            off[q] = run[q] + before + incl[q] - p.f[q];
            all[q] = sum;
        }
        __syncthreads();  // s_w is rewritten by the next pass
        if (gi < ng) {
#pragma unroll
            for (int q = 0; q < 5; ++q) parts[gi].f[q] = off[q];
            mx[0] = max(mx[0], p.mx);
#pragma unroll
            for (int q = 0; q < 3; ++q) mx[1 + q] = max(mx[1 + q], p.big[q]);
            bad |= p.bad;
            sq += p.sumsq;
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) run[q] += all[q];
    }
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) mx[q] = max(mx[q], __shfl_xor(mx[q], o));
        bad |= __shfl_xor(bad, o);
        sq += __shfl_xor(sq, o);
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) s_mx[w][q] = mx[q];
        s_bad[w] = bad;
        s_sq[w] = sq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int m[4] = {0, 0, 0, 0}, b = 0;
        unsigned long long t = 0ull;
        for (int k = 0; k < NW; ++k) {
#pragma unroll
            for (int q = 0; q < 4; ++q) m[q] = max(m[q], s_mx[k][q]);
            b |= s_bad[k];
            t += s_sq[k];
        }
        tot->rs.cnt[0] = run[2];
        tot->rs.cnt[1] = run[3];
        tot->rs.cnt[2] = run[4];
        tot->rs.max_blocks = m[0];
        tot->big[0] = m[1];
        tot->big[1] = m[2];
        tot->big[2] = m[3];
        tot->rs.sumsq = t;
        tot->nblk = run[0];
        tot->nnz = run[1];
        tot->bad = b;
        *cptr_end = run[0];
        *cvp_end = run[1];
    }
}

// C's block and value offsets of every row, and list[b * stride + k] = row for
// the k-th row (ascending) of numeric bin b
template <typename FB>
__global__ __launch_bounds__(256) void k_csize_write(int32_t n, FB fb,
                                                     const int32_t *__restrict__ ccnt,
                                                     const int32_t *__restrict__ cnnz,
                                                     const CSizePart *__restrict__ parts,
                                                     int32_t *__restrict__ cptr,
                                                     int32_t *__restrict__ cvp,
                                                     int32_t *__restrict__ list, int64_t stride) {
    __shared__ int s_w[4][5];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool valid = i < n;
    const int c = valid ? ccnt[i] : 0, v = valid ? cnnz[i] : 0;
    const int b = valid ? int(fb(i)) : -1;
    const int ci = wave_incl_sum(c), vi = wave_incl_sum(v);
    const uint64_t m0 = __ballot(b == 0), m1 = __ballot(b == 1), m2 = __ballot(b == 2);
    if (lane == 63) {
        s_w[w][0] = ci;
        s_w[w][1] = vi;
    }
    if (lane == 0) {
        s_w[w][2] = __popcll(m0);
        s_w[w][3] = __popcll(m1);
        s_w[w][4] = __popcll(m2);
    }
    __syncthreads();
    if (!valid) return;
    const CSizePart &p = parts[blockIdx.x];
    int ob = p.f[0], ov = p.f[1];
    for (int q = 0; q < w; ++q) {
        ob += s_w[q][0];
        ov += s_w[q][1];
    }
    cptr[i] = ob + ci - c;
    cvp[i] = ov + vi - v;
    if (b >= 0) {
        const uint64_t mb = b == 0 ? m0 : (b == 1 ? m1 : m2);
        int pos = p.f[2 + b];
        for (int q = 0; q < w; ++q) pos += s_w[q][2 + b];
        pos += __popcll(mb & ((1ull << lane) - 1ull));
        list[int64_t(b) * stride + pos] = i;
    }
}

template <typename FB>
int c_sizes(int32_t brow, FB fb, const int32_t *ccnt, const int32_t *cnnz, CSizePart *parts,
            CSizeTotals *tot, int32_t *cptr, int32_t *cvp, int32_t *list, hipStream_t s) {
    const int32_t ng = rowpass::groups(brow);
    if (ng > 0)
        hipLaunchKernelGGL((k_csize_part<FB>), dim3(unsigned(ng)), dim3(256), 0, s, brow, fb, ccnt,
                           cnnz, parts);
    hipLaunchKernelGGL(k_csize_finish, dim3(1), dim3(rowpass::kFinishThreads), 0, s, ng, parts,
                       tot, cptr + brow, cvp + brow);
    if (ng > 0)
        hipLaunchKernelGGL((k_csize_write<FB>), dim3(unsigned(ng)), dim3(256), 0, s, brow, fb,
                           ccnt, cnnz, parts, cptr, cvp, list, int64_t(brow));
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

__global__ __launch_bounds__(256) void k_widen(int32_t n, const int32_t *__restrict__ in,
                                               int64_t *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// value offset of every block of B: its block row's value offset plus the
// popcounts of the row's blocks before it (one thread per block row; rows of
// at most kVoffRowMax blocks, else k_block_popc and a scan)
constexpr int kVoffRowMax = 256;
__global__ __launch_bounds__(256) void k_block_voff(int32_t brow, const int32_t *__restrict__ bp,
                                                    const int32_t *__restrict__ vp,
                                                    const uint16_t *__restrict__ map,
                                                    int32_t *__restrict__ off) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= brow) return;
    int o = vp[r];
    for (int q = bp[r]; q < bp[r + 1]; ++q) {
        off[q] = o;
        o += __popc(uint32_t(map[q]));
    }
}

// popcount of every block bitmap (values are stored in block order, so the
// exclusive scan of these is each block's value offset)
__global__ __launch_bounds__(256) void k_block_popc(int64_t nblk, const uint16_t *__restrict__ map,
                                                    int32_t *__restrict__ out) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < nblk) out[q] = __popc(uint32_t(map[q]));
}

__global__ __launch_bounds__(256) void k_scale_values(int64_t n, double f, double *__restrict__ v) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] *= f;
}

// numeric kernels of one value precision; *nslab receives the global slab of
// the rows over kNumCap blocks (freed by the caller)
// The FP64 register-bin kernel also copies its rows' block columns and
// bitmaps from the symbolic temporaries (ub_off, tcol, tmap) to C; the other
// kernels read them from C.
template <typename VT>
int launch_numeric(const SpgemmArgs &g, const int32_t *boff, const Mbsr &C, int32_t brow,
                   const int32_t *list, int32_t reg_rows, int32_t lds_rows, int32_t big_rows,
                   const int64_t *ub_off, int32_t stride, const int32_t *tcol,
                   const uint32_t *tmap, const int32_t *big, double **nslab, hipStream_t s) {
    if (reg_rows > 0) {
        if constexpr (std::is_same<VT, double>::value) {
            // accumulators for the bin's longest C row (at most 64 blocks)
            const dim3 grid((reg_rows + 3) / 4);
            if (C.max_row_blocks <= 16)
                hipLaunchKernelGGL((k_numeric_outer<16>), grid, dim3(256), 0, s, reg_rows, list,
                                   g, boff, C.blk_ptr, ub_off, stride, tcol, tmap, C.blk_col,
                                   C.blk_map, C.val_ptr, C.val);
            else if (C.max_row_blocks <= 32)
                hipLaunchKernelGGL((k_numeric_outer<32>), grid, dim3(256), 0, s, reg_rows, list,
                                   g, boff, C.blk_ptr, ub_off, stride, tcol, tmap, C.blk_col,
                                   C.blk_map, C.val_ptr, C.val);
            else
                hipLaunchKernelGGL((k_numeric_outer<64>), grid, dim3(256), 0, s, reg_rows, list,
                                   g, boff, C.blk_ptr, ub_off, stride, tcol, tmap, C.blk_col,
                                   C.blk_map, C.val_ptr, C.val);
        } else
            hipLaunchKernelGGL((k_numeric_reg16<VT>), dim3((reg_rows + 3) / 4), dim3(256), 0, s,
                               reg_rows, list, g, boff, C.blk_ptr, C.blk_col, C.blk_map,
                               C.val_ptr, C.val);
    }
    if (lds_rows > 0) {
        hipLaunchKernelGGL((k_numeric_cmp<VT, kNumCap, kCmpValShort>), dim3((lds_rows + 3) / 4),
                           dim3(256), 0, s, lds_rows, list + brow, g, boff, C.blk_ptr, C.blk_col,
                           C.blk_map, C.val_ptr, C.val);
        if (big[0] > kCmpValShort)  // a row the compressed accumulators do not hold
            hipLaunchKernelGGL((k_numeric_lds<VT>), dim3((lds_rows + 3) / 4), dim3(256), 0, s,
                               lds_rows, list + brow, g, boff, C.blk_ptr, C.blk_col, C.blk_map,
                               C.val_ptr, C.val);
    }
    if (big_rows > 0) {
        hipLaunchKernelGGL((k_numeric_cmp<VT, kCmpMidBlk, kCmpMidVal>), dim3((big_rows + 3) / 4),
                           dim3(256), 0, s, big_rows, list + 2 * int64_t(brow), g, boff,
                           C.blk_ptr, C.blk_col, C.blk_map, C.val_ptr, C.val);
        if (big[1] > kCmpMidBlk || big[2] > kCmpMidVal)
            hipLaunchKernelGGL((k_numeric_cmp<VT, kCmpBlk, kCmpVal, kCmpMidBlk, kCmpMidVal>),
                               dim3((big_rows + 3) / 4), dim3(256), 0, s, big_rows,
                               list + 2 * int64_t(brow), g, boff, C.blk_ptr, C.blk_col,
                               C.blk_map, C.val_ptr, C.val);
    }
    // rows over the larger compressed launch's limits: dense LDS or a global slab
    if (big_rows > 0 && (big[1] > kCmpBlk || big[2] > kCmpVal)) {
        const int64_t elems = int64_t(C.max_row_blocks) * 16;
        int nwaves = std::min(big_rows, 1024);
        while (nwaves > 4 && elems * nwaves * 8 > (int64_t(1) << 31)) nwaves >>= 1;
        nwaves = (nwaves + 3) / 4 * 4;  // every wave of the grid owns a slab
        AMGT_TRY(tmp_alloc_t(nslab, size_t(elems * nwaves), s));
        hipLaunchKernelGGL((k_numeric_global<VT>), dim3((nwaves + 3) / 4), dim3(256), 0, s,
                           big_rows, list + 2 * int64_t(brow), g, boff, C.blk_ptr, C.blk_col,
                           C.blk_map, C.val_ptr, C.val, *nslab, elems);
    }
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

}  // namespace

int spgemm(const Mbsr &A, const Mbsr &B, int precision, hipStream_t s, Mbsr *Cp) {
    Mbsr &C = *Cp;
    C = Mbsr();
    C.nrows = A.nrows;
    C.ncols = B.ncols;
    C.brow = A.brow;
    C.bcol = B.bcol;
    C.pattern_only = A.pattern_only || B.pattern_only;
    if (A.bcol != B.brow) {
        set_error("SpGEMM: block dimensions of A and B differ");
        return AMGT_ERR_DIM;
    }
    const int32_t brow = A.brow;
    C.blk_ptr = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(brow) + 1), s));
    C.val_ptr = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(brow) + 1), s));
    if (!C.blk_ptr || !C.val_ptr) {
        set_error("SpGEMM: allocation failed");
        return AMGT_ERR_ALLOC;
    }
    // operand values in the product's precision (FP32/FP16 copies made by the caller)
    const void *avals = A.val, *bvals = B.val;
    if (precision == AMGT_FLOAT) {
        avals = A.val32;
        bvals = B.val32;
    } else if (precision == AMGT_HALF) {
        avals = A.val16;
        bvals = B.val16;
    }
    if (!C.pattern_only && (!avals || !bvals)) {
        set_error("SpGEMM: operand values missing for the requested precision");
        return AMGT_ERR_ARG;
    }
    const int64_t rs_max = int64_t(kNoEntryOff) / 8;  // values a 32-bit byte offset reaches
    const int fp64_rs = precision != AMGT_FLOAT && precision != AMGT_HALF && A.nnz <= rs_max &&
                        B.nnz <= rs_max;
    SpgemmArgs g{A.blk_ptr, A.blk_col, A.val_ptr, A.blk_map, avals,
                 B.blk_ptr, B.blk_col, B.val_ptr, B.blk_map, bvals, B.nnz, A.nnz, fp64_rs};

    // Bounded form: when the largest A row times the largest B row is at most
    // 64 products, every block row fits the register kernels, and the
    // symbolic temporaries take a fixed slot per row, so the upper-bound pass,
    // its scan and its readback are not needed.
    const bool bounded = brow > 0 && A.max_row_blocks >= 1 && A.max_row_blocks <= 64 &&
                         B.max_row_blocks >= 1 &&
                         int64_t(A.max_row_blocks) * B.max_row_blocks <= 64;
    const int32_t stride = bounded ? A.max_row_blocks * B.max_row_blocks : 0;
    // the register symbolic kernel ranks (block column, lane) pairs packed in one int
    const int packed = B.bcol < (1 << 25) - 1 ? 1 : 0;

    rowpass::Part *parts = nullptr;
    RowStats *tots = nullptr;  // the upper-bound pass
    RowStats ht = {};
    CSizePart *cparts = nullptr;
    CSizeTotals *ctot = nullptr;
    int64_t *ub = nullptr, *ub_off = nullptr;
    int32_t *list = nullptr, *ccnt = nullptr, *cnnz = nullptr;
    int32_t *tcol = nullptr;
    uint32_t *tmap = nullptr, *slab = nullptr;
    int32_t *retry = nullptr;
    double *nslab = nullptr;
    int32_t *bpop = nullptr, *boff = nullptr;
    int st = AMGT_OK;
    auto cleanup = [&]() {
        tmp_free(parts, s);
        tmp_free(tots, s);
        tmp_free(cparts, s);
        tmp_free(ctot, s);
        tmp_free(ub, s);
        tmp_free(ub_off, s);
        tmp_free(list, s);
        tmp_free(ccnt, s);
        tmp_free(cnnz, s);
        tmp_free(tcol, s);
        tmp_free(tmap, s);
        tmp_free(slab, s);
        tmp_free(retry, s);
        tmp_free(nslab, s);
        tmp_free(bpop, s);
        tmp_free(boff, s);
    };
#define SG_TRY(expr)               \
    do {                           \
        st = (expr);               \
        if (st != AMGT_OK) {       \
            cleanup();             \
            return st;             \
        }                          \
    } while (0)
#define SG_HIP(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            set_error(std::string("SpGEMM: ") + #expr + ": " + hipGetErrorString(e_)); \
            cleanup();                                                            \
            return AMGT_ERR_HIP;                                                  \
        }                                                                         \
    } while (0)

    SG_TRY(tmp_alloc_t(&cparts, size_t(rowpass::groups(brow)) + 1, s));
    SG_TRY(tmp_alloc_t(&ctot, 1, s));
    SG_TRY(tmp_alloc_t(&list, size_t(brow) * 3 + 1, s));
    SG_TRY(tmp_alloc_t(&ccnt, size_t(brow) + 1, s));
    SG_TRY(tmp_alloc_t(&cnnz, size_t(brow) + 1, s));
    int64_t total_ub = 0;
    const int32_t *bin = ht.cnt;
    if (bounded) {
        total_ub = int64_t(brow) * stride;
    } else {
        SG_TRY(tmp_alloc_t(&parts, size_t(rowpass::groups(brow)) + 1, s));
        SG_TRY(tmp_alloc_t(&tots, 1, s));
        SG_TRY(tmp_alloc_t(&ub, size_t(brow) + 1, s));
        SG_TRY(tmp_alloc_t(&ub_off, size_t(brow) + 1, s));
        const unsigned gr = unsigned(rowpass::groups(brow));
        if (brow > 0)
            hipLaunchKernelGGL(k_upper_bound, dim3(gr), dim3(256), 0, s, brow, A.blk_ptr,
                               A.blk_col, B.blk_ptr, ub, parts);
        hipLaunchKernelGGL(rowpass::k_pass_finish, dim3(1), dim3(rowpass::kFinishThreads), 0, s,
                           int32_t(gr), parts, tots);
        SG_HIP(hipGetLastError());
        SG_TRY(rowpass::scatter(brow, SymbolicBin{ub, A.blk_ptr}, 0, kBin1Ub, parts, list,
                                int64_t(brow), 7u, s));
        SG_TRY(exclusive_scan_i64(ub, ub_off, brow, s));
        // sync 1: sizes of the temporaries and bins
        const Readback rb[2] = {{&total_ub, ub_off + brow, sizeof(int64_t)},
                                {&ht, tots, sizeof(RowStats)}};
        SG_TRY(readback(rb, 2, s));
    }
    SG_TRY(tmp_alloc_t(&tcol, size_t(total_ub), s));
    SG_TRY(tmp_alloc_t(&tmap, size_t(total_ub), s));
    if (!C.pattern_only && brow > 0) {
        // value offset of every B block, for the numeric pass; enqueued here
        // so that it does not follow the readback of C's sizes
        SG_TRY(tmp_alloc_t(&boff, size_t(B.nblk) + 1, s));
        if (B.max_row_blocks <= kVoffRowMax) {
            if (B.brow > 0)
                hipLaunchKernelGGL(k_block_voff, dim3(unsigned((B.brow + 255) / 256)), dim3(256),
                                   0, s, B.brow, B.blk_ptr, B.val_ptr, B.blk_map, boff);
        } else {
            SG_TRY(tmp_alloc_t(&bpop, size_t(B.nblk) + 1, s));
            if (B.nblk > 0)
                hipLaunchKernelGGL(k_block_popc, dim3(unsigned((B.nblk + 255) / 256)), dim3(256),
                                   0, s, B.nblk, B.blk_map, bpop);
            SG_HIP(hipGetLastError());
            SG_TRY(exclusive_scan_i32(bpop, boff, B.nblk, s));
        }
        SG_HIP(hipGetLastError());
    }
    constexpr int kSymRows = 4 * kSymRegRows;  // rows per workgroup of k_symbolic_reg
    if (bounded)
        hipLaunchKernelGGL((k_symbolic_reg<kSymRegRows>), dim3((brow + kSymRows - 1) / kSymRows),
                           dim3(256), 0, s, brow, static_cast<const int32_t *>(nullptr), g,
                           static_cast<const int64_t *>(nullptr), stride, packed, tcol, tmap,
                           ccnt, cnnz);
    if (!bounded && bin[0] > 0)
        hipLaunchKernelGGL((k_symbolic_reg<kSymRegRows>),
                           dim3((bin[0] + kSymRows - 1) / kSymRows), dim3(256), 0, s, bin[0],
                           list, g, ub_off, 0, packed, tcol, tmap, ccnt, cnnz);
    if (!bounded && bin[1] > 0) {
        hipLaunchKernelGGL(k_symbolic_wave, dim3((bin[1] + 3) / 4), dim3(256), 0, s, bin[1],
                           list + brow, g, ub_off, tcol, tmap, ccnt, cnnz);
        if (ht.max_blocks > kWaveUb)  // a row over the wave kernel's table
            hipLaunchKernelGGL(k_symbolic_block, dim3(bin[1]), dim3(256), 0, s, bin[1],
                               list + brow, g, ub_off, tcol, tmap, ccnt, cnnz);
    }
    if (!bounded && bin[2] > 0) {
        int hs = 1;
        const int64_t want = std::min<int64_t>(2 * int64_t(ht.max_blocks), 2 * int64_t(B.bcol) + 2);
        while (hs < want) hs <<= 1;
        const int64_t per_wg = 2 * int64_t(hs);
        int nwg = std::min(bin[2], 512);
        const int64_t budget = int64_t(1) << 29;  // 2 GiB of uint32 slots
        while (nwg > 1 && per_wg * nwg > budget) nwg >>= 1;
        if (per_wg > budget) {
            set_error("SpGEMM: a block row's product is too large for the global hash table");
            cleanup();
            return AMGT_ERR_UNSUPPORTED;
        }
        SG_TRY(tmp_alloc_t(&slab, size_t(per_wg * nwg), s));
        SG_TRY(tmp_alloc_t(&retry, size_t(brow) + 1, s));
        hipLaunchKernelGGL(k_symbolic_try, dim3(bin[2]), dim3(256), 0, s, bin[2],
                           list + 2 * int64_t(brow), g, ub_off, tcol, tmap, ccnt, cnnz, retry);
        hipLaunchKernelGGL(k_symbolic_global, dim3(nwg), dim3(256), 0, s, bin[2],
                           list + 2 * int64_t(brow), g, ub_off, tcol, tmap, ccnt, cnnz, slab, hs,
                           retry);
    }
    SG_HIP(hipGetLastError());
    // C's block and value offsets are int32: nblk <= total_ub and nnz <= 16 nblk,
    // so only a product whose bound reaches 2^31 values needs its exact totals,
    // summed in 64 bits before anything of C is allocated
    if (16 * total_ub >= (int64_t(1) << 31)) {
        int64_t *wide = nullptr, *wscan = nullptr;  // wscan: one scan per count array
        SG_TRY(tmp_alloc_t(&wide, size_t(brow) + 1, s));
        SG_TRY(tmp_alloc_t(&wscan, 2 * (size_t(brow) + 1), s));
        int64_t tot[2] = {0, 0};
        const int32_t *cnts[2] = {ccnt, cnnz};
        for (int w = 0; w < 2; ++w) {
            if (brow > 0)
                hipLaunchKernelGGL(k_widen, dim3((brow + 255) / 256), dim3(256), 0, s, brow,
                                   cnts[w], wide);
            SG_HIP(hipGetLastError());
            SG_TRY(exclusive_scan_i64(wide, wscan + w * (int64_t(brow) + 1), brow, s));
        }
        {
            const Readback rb[2] = {{&tot[0], wscan + brow, sizeof(int64_t)},
                                    {&tot[1], wscan + 2 * int64_t(brow) + 1, sizeof(int64_t)}};
            SG_TRY(readback(rb, 2, s));
        }
        tmp_free(wide, s);
        tmp_free(wscan, s);
        if (tot[0] >= (int64_t(1) << 31) || tot[1] >= (int64_t(1) << 31)) {
            set_error("SpGEMM: the product has 2^31 or more blocks or values (int32 offsets)");
            cleanup();
            return AMGT_ERR_UNSUPPORTED;
        }
    }
    // C's block and value offsets, the numeric bins (rows over kNumCap blocks go
    // to the global-memory numeric kernel), the largest block count and the
    // sum of squares (max_row_blocks and stand)
    if (bounded)
        SG_TRY(c_sizes(brow, RegBin{}, ccnt, cnnz, cparts, ctot, C.blk_ptr, C.val_ptr, list, s));
    else
        SG_TRY(c_sizes(brow, NumericBin{ub, A.blk_ptr, ccnt}, ccnt, cnnz, cparts, ctot, C.blk_ptr,
                       C.val_ptr, list, s));
    CSizeTotals ct{};
    {  // sync 2: sizes of C
        const Readback rb[1] = {{&ct, ctot, sizeof(CSizeTotals)}};
        SG_TRY(readback(rb, 1, s));
    }
    if (ct.bad) {
        set_error("SpGEMM: a block row exceeds the bound of its operands' row lengths");
        cleanup();
        return AMGT_ERR_INPUT;
    }
    const int32_t nblk = ct.nblk, nnz = ct.nnz;
    const int32_t reg_rows = ct.rs.cnt[0], lds_rows = ct.rs.cnt[1], big_rows = ct.rs.cnt[2];
    C.nblk = nblk;
    C.nnz = nnz;
    C.max_row_blocks = ct.rs.max_blocks;
    C.avg_nnz = nblk ? double(nnz) / double(nblk) : 0.0;
    C.stand = stand_from(ct.rs, nblk, brow);
    C.spmv_kernel = default_spmv_kernel(C);
    C.blk_col = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * size_t(nblk ? nblk : 1), s));
    C.blk_map = static_cast<uint16_t *>(dev_alloc(sizeof(uint16_t) * size_t(nblk ? nblk : 1), s));
    if (!C.pattern_only)
        C.val = static_cast<double *>(dev_alloc(sizeof(double) * size_t(nnz ? nnz : 1), s));
    if (!C.blk_col || !C.blk_map || (!C.pattern_only && !C.val)) {
        set_error("SpGEMM: allocation of C failed");
        cleanup();
        return AMGT_ERR_ALLOC;
    }
    // C's block columns and bitmaps out of the symbolic temporaries; in FP64
    // the register-bin numeric kernel copies its own rows
    const bool fp64_numeric = !C.pattern_only && precision != AMGT_FLOAT && precision != AMGT_HALF;
    if (brow > 0 && !fp64_numeric)
        hipLaunchKernelGGL(k_compact_c, dim3((brow + 3) / 4), dim3(256), 0, s, brow,
                           static_cast<const int32_t *>(nullptr), ub_off, stride, C.blk_ptr, tcol,
                           tmap, C.blk_col, C.blk_map);
    if (fp64_numeric && lds_rows > 0)
        hipLaunchKernelGGL(k_compact_c, dim3((lds_rows + 3) / 4), dim3(256), 0, s, lds_rows,
                           list + brow, ub_off, stride, C.blk_ptr, tcol, tmap, C.blk_col,
                           C.blk_map);
    if (fp64_numeric && big_rows > 0)
        hipLaunchKernelGGL(k_compact_c, dim3((big_rows + 3) / 4), dim3(256), 0, s, big_rows,
                           list + 2 * int64_t(brow), ub_off, stride, C.blk_ptr, tcol, tmap,
                           C.blk_col, C.blk_map);
    SG_HIP(hipGetLastError());
    if (!C.pattern_only && brow > 0) {
        switch (precision) {
            case AMGT_FLOAT:
                st = launch_numeric<float>(g, boff, C, brow, list, reg_rows, lds_rows, big_rows, ub_off, stride, tcol, tmap, ct.big, &nslab, s);
                break;
            case AMGT_HALF:
                st = launch_numeric<__half>(g, boff, C, brow, list, reg_rows, lds_rows, big_rows, ub_off, stride, tcol, tmap, ct.big, &nslab, s);
                // the FP16 copies hold A and B times half_scale: undo both (powers of two, exact)
                if (st == AMGT_OK && C.nnz > 0)
                    hipLaunchKernelGGL(k_scale_values, dim3(unsigned((C.nnz + 255) / 256)), dim3(256),
                                       0, s, C.nnz, 1.0 / (A.half_scale * B.half_scale), C.val);
                break;
            default:
                st = launch_numeric<double>(g, boff, C, brow, list, reg_rows, lds_rows, big_rows, ub_off, stride, tcol, tmap, ct.big, &nslab, s);
        }
        if (st != AMGT_OK) {
            cleanup();
            return st;
        }
        SG_HIP(hipGetLastError());
    }
    cleanup();
#undef SG_TRY
#undef SG_HIP
    return AMGT_OK;
}

}  // namespace amgt
