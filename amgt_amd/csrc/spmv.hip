// mBSR SpMV / SpMM (counterparts of spmv_amgT_fp64/fp32/fp16,
// csr_matvec_device.c:1302-1411, 2098-2309).
//
// Kernel shape: G lanes per 4-row block row (G = 4, 8, 16, 32, 64 chosen from
// the matrix statistics or by the tuner), 64/G block rows per wave, 4 waves per
// workgroup.  Within a block row, lane (slot s, row r) walks blocks s, s+G/4,
// ... and, for each, the set bits of row r's nibble of the 16-bit bitmap.  Its
// values sit at  val_ptr[br] + (popcounts of the earlier blocks) +
// popc(map & ((1 << 4r) - 1)),  so the popcount walk needs no per-nonzero
// index.  Row sums are reduced across slots with shuffles; beta*b is fused in
// the store (no separate beta pass, no atomics: results are deterministic).
// Values may be FP64, FP32 or FP16; below FP64 the products accumulate in FP32.
// x, b, y are FP64.
#include "internal.hpp"

#include <string>
#include <type_traits>

namespace amgt {

namespace {

template <typename VT>
__device__ __forceinline__ float to_f(VT v);
template <>
__device__ __forceinline__ float to_f<float>(float v) { return v; }
template <>
__device__ __forceinline__ float to_f<__half>(__half v) { return __half2float(v); }

template <typename VT, typename AT>
__device__ __forceinline__ AT mul_acc(AT acc, VT a, double x);

template <>
__device__ __forceinline__ double mul_acc<double, double>(double acc, double a, double x) {
    return fma(a, x, acc);
}
template <>
__device__ __forceinline__ float mul_acc<float, float>(float acc, float a, double x) {
    return fmaf(a, float(x), acc);
}
template <>
__device__ __forceinline__ float mul_acc<__half, float>(float acc, __half a, double x) {
    return fmaf(__half2float(a), float(x), acc);
}

// XCD-aware workgroup remap: hardware hands out workgroups round-robin over
// the 8 XCDs; give each XCD a contiguous range of block rows so that x reuse
// between neighbouring block rows stays within one L2.
__device__ __forceinline__ int remap_wg(int wg, int nwg) {
    const int xcd = wg & 7;
    const int idx = wg >> 3;
    const int per = nwg >> 3;
    const int rem = nwg & 7;
    return xcd * per + min(xcd, rem) + idx;
}

// Row epilogues of the SpMV kernels: y[row] = ep(row, (A x)[row]).
struct EpAxpby {  // alpha (A x) + beta b; beta == 0 never reads b
    double alpha, beta;
    const double *b;
    __device__ __forceinline__ double operator()(int row, double ax) const {
        double v = alpha * ax;
        if (beta != 0.0) v = fma(beta, b[row], v);
        return v;
    }
};
template <typename EP>
struct EpScaled {  // the inner epilogue of sc (A x): A's values were stored times 1/sc
    EP inner;
    double sc;
    __device__ __forceinline__ double operator()(int row, double ax) const {
        return inner(row, ax * sc);
    }
};
struct EpJacobi {  // one weighted Jacobi sweep: u + omega dinv (f - A u), with x = u
    double omega;
    const double *dinv, *f, *u;
    __device__ __forceinline__ double operator()(int row, double ax) const {
        return fma(omega * dinv[row], f[row] - ax, u[row]);
    }
};
// The residual r = f - A u of a V-cycle level (x = u, y = r) whose tiles also
// restrict it, fc = R r: once a tile has stored its rows of r it arrives at
// every coarse row its fine rows touch (tile_c[tile_ptr[t] .. tile_ptr[t+1])),
// and the last tile to arrive at coarse row c (nspan[c] arrivals) forms
// fc[c] over R's row c in R's order -- the same sum whichever tile is last --
// and puts the counter back to zero for the next call.  No tile waits for
// another, so nothing can hang.  r is stored and read by the last tile at
// device scope (written through to memory, read past this XCD's L2), and each
// wave waits for its stores before the arrivals: agent-scope fences instead
// write back and invalidate the whole L2 of the XCD per wave, which made the
// cycle 7x slower.
struct EpResidRestrict {
    const double *f;
    const int32_t *tile_ptr, *tile_c, *nspan;
    int32_t *cnt;
    const int32_t *ri, *rj;
    const double *rv;
    const double *r;
    double *fc;
    __device__ __forceinline__ double operator()(int row, double ax) const { return f[row] - ax; }
};

// Work of a whole tile after its rows are stored (every thread of the
// workgroup calls it): nothing for the plain epilogues.
template <typename EP>
__device__ __forceinline__ void tile_post(const EP &, int) {}
template <typename EP>
__device__ __forceinline__ void tile_post(const EpScaled<EP> &ep, int t) {
    tile_post(ep.inner, t);
}
__device__ __forceinline__ void tile_post(const EpResidRestrict &ep, int t) {
    // the tile's rows of r were stored at device scope (written through to
    // memory, store_y): once every wave's stores have completed, arrive
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int k = ep.tile_ptr[t] + int(threadIdx.x); k < ep.tile_ptr[t + 1]; k += int(blockDim.x)) {
        const int c = ep.tile_c[k];
        if (atomicAdd(&ep.cnt[c], 1) == ep.nspan[c] - 1) {
            // last arrival: every tile's rows of r are in memory; read them at
            // device scope (not from a stale line of this XCD's L2)
            double acc = 0.0;
            for (int p = ep.ri[c]; p < ep.ri[c + 1]; ++p) {
                const double rj = __hip_atomic_load(ep.r + ep.rj[p], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
                acc = fma(ep.rv[p], rj, acc);
            }
            ep.fc[c] = acc;
            ep.cnt[c] = 0;
        }
    }
}

// y[row] = v: a plain store, or for the residual of the fused restriction a
// device-scope store that other XCDs' last arrivals can read
template <typename EP>
__device__ __forceinline__ void store_y(const EP &, double *y, int row, double v) {
    y[row] = v;
}
template <typename EP>
__device__ __forceinline__ void store_y(const EpScaled<EP> &ep, double *y, int row, double v) {
    store_y(ep.inner, y, row, v);
}
__device__ __forceinline__ void store_y(const EpResidRestrict &, double *y, int row, double v) {
    __hip_atomic_store(y + row, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int G, typename VT, typename AT, typename EP>
__global__ __launch_bounds__(256) void k_spmv(
    int32_t brow, int32_t nrows, const int32_t *__restrict__ blk_ptr,
    const int32_t *__restrict__ blk_col, const uint16_t *__restrict__ blk_map,
    const int32_t *__restrict__ val_ptr, const VT *__restrict__ val, const double *__restrict__ x,
    EP ep, double *y) {
    constexpr int S = G / 4;
    constexpr int RPW = 64 / G;
    const int lane = threadIdx.x & 63;
    const int wg = remap_wg(blockIdx.x, gridDim.x);
    const int wave = wg * 4 + (threadIdx.x >> 6);
    const int br = wave * RPW + lane / G;
    const int sub = lane % G;
    const int r = sub & 3;
    const int s = sub >> 2;
    if (br >= brow) return;  // whole G-lane group leaves together
    const int p0 = blk_ptr[br];
    const int p1 = blk_ptr[br + 1];
    int voff = val_ptr[br];
    const uint32_t low = (1u << (4 * r)) - 1u;
    AT acc = AT(0);
    for (int jb = p0; jb < p1; jb += S) {
        const int j = jb + s;
        uint32_t m = 0;
        int col = 0;
        if (j < p1) {
            m = blk_map[j];
            col = blk_col[j];
        }
        const int cnt = __popc(m);
        int incl = cnt;
        int tot = cnt;
        if constexpr (S > 1) {
#pragma unroll
            for (int d = 1; d < S; d <<= 1) {
                int t = __shfl_up(incl, 4 * d, G);
                if (s >= d) incl += t;
            }
            tot = __shfl(incl, 4 * (S - 1) + r, G);
        }
        int o = voff + (incl - cnt) + __popc(m & low);
        uint32_t nib = (m >> (4 * r)) & 0xFu;
        const double *xs = x + 4 * int64_t(col);
        while (nib) {
            const int c = __builtin_ctz(nib);
            nib &= nib - 1u;
            acc = mul_acc<VT, AT>(acc, val[o], xs[c]);
            ++o;
        }
        voff += tot;
    }
    if constexpr (S > 1) {
#pragma unroll
        for (int d = S / 2; d >= 1; d >>= 1) acc += __shfl_xor(acc, 4 * d, G);
    }
    if (s == 0) {
        const int row = 4 * br + r;
        if (row < nrows) {
            y[row] = ep(row, double(acc));
        }
    }
}

// Staged variant: a workgroup owns 256/G consecutive block rows.  Their block
// columns, bitmaps and values (contiguous ranges) are first copied to LDS with
// coalesced loads that depend only on the two pointer pairs, so every matrix
// byte is requested at once; the popcount walk then reads LDS and only the x
// gathers wait on loaded indices.  Tiles larger than the LDS capacity read
// global memory directly (same arithmetic).
template <typename T>
__device__ __forceinline__ T stream_load(const T *p) {
    return *p;  // plain load (non-temporal loads measured slower here)
}
template <>
__device__ __forceinline__ __half stream_load<__half>(const __half *p) {
    return *p;
}

template <int G, typename VT, typename AT>
__device__ __forceinline__ AT row_walk(const int32_t *__restrict__ col, const uint16_t *__restrict__ map,
                                       const VT *__restrict__ val, int p0, int p1, int voff,
                                       int s, int r, const double *__restrict__ x) {
    constexpr int S = G / 4;
    const uint32_t low = (1u << (4 * r)) - 1u;
    AT acc = AT(0);
    for (int jb = p0; jb < p1; jb += S) {
        const int j = jb + s;
        uint32_t m = 0;
        int c4 = 0;
        if (j < p1) {
            m = map[j];
            c4 = col[j];
        }
        const int cnt = __popc(m);
        int incl = cnt;
        int tot = cnt;
        if constexpr (S > 1) {
#pragma unroll
            for (int d = 1; d < S; d <<= 1) {
                int t = __shfl_up(incl, 4 * d, G);
                if (s >= d) incl += t;
            }
            tot = __shfl(incl, 4 * (S - 1) + r, G);
        }
        int o = voff + (incl - cnt) + __popc(m & low);
        uint32_t nib = (m >> (4 * r)) & 0xFu;
        const double *xs = x + 4 * int64_t(c4);
        while (nib) {
            const int c = __builtin_ctz(nib);
            nib &= nib - 1u;
            acc = mul_acc<VT, AT>(acc, val[o], xs[c]);
            ++o;
        }
        voff += tot;
    }
    if constexpr (S > 1) {
#pragma unroll
        for (int d = S / 2; d >= 1; d >>= 1) acc += __shfl_xor(acc, 4 * d, G);
    }
    return acc;
}

// Walk over LDS-resident blocks that first gathers the x segments (4 doubles,
// two 16-byte loads) of up to MAXB blocks per lane, then consumes them with
// straight-line predicated FMAs, so the x loads of a lane are all in flight
// together.  x must be 16-byte aligned (checked by the launcher); the last
// block column is read element-wise when ncols is not a multiple of 4.
template <int G, int MAXB, bool XSEG, typename VT, typename AT>
__device__ __forceinline__ AT row_walk_prefetch(const int32_t *s_col, const uint16_t *s_map,
                                                const VT *s_val, int p0, int p1, int voff, int s,
                                                int r, const double *__restrict__ x, int ncols) {
    constexpr int S = G / 4;
    const uint32_t low = (1u << (4 * r)) - 1u;
    AT acc = AT(0);
    for (int j0 = p0; j0 < p1; j0 += S * MAXB) {
        double xs[MAXB][4];
        uint32_t m[MAXB];
#pragma unroll
        for (int u = 0; u < MAXB; ++u) {
            const int j = j0 + s + S * u;
            m[u] = 0u;
            xs[u][0] = xs[u][1] = xs[u][2] = xs[u][3] = 0.0;
            if (j < p1) {
                m[u] = s_map[j];
                const int c4 = 4 * s_col[j];
                const double *xp = x + c4;
                if constexpr (!XSEG) {
                    // gather only the columns of this lane's row nibble
                    const uint32_t nib = (m[u] >> (4 * r)) & 0xFu;
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if ((nib >> c) & 1u) xs[u][c] = xp[c];
                } else if (c4 + 4 <= ncols) {
                    const double2 lo2 = *reinterpret_cast<const double2 *>(xp);
                    const double2 hi2 = *reinterpret_cast<const double2 *>(xp + 2);
                    xs[u][0] = lo2.x;
                    xs[u][1] = lo2.y;
                    xs[u][2] = hi2.x;
                    xs[u][3] = hi2.y;
                } else {
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (c4 + c < ncols) xs[u][c] = xp[c];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < MAXB; ++u) {
            const int cnt = __popc(m[u]);
            int incl = cnt;
            int tot = cnt;
            if constexpr (S > 1) {
#pragma unroll
                for (int d = 1; d < S; d <<= 1) {
                    int t = __shfl_up(incl, 4 * d, G);
                    if (s >= d) incl += t;
                }
                tot = __shfl(incl, 4 * (S - 1) + r, G);
            }
            int o = voff + (incl - cnt) + __popc(m[u] & low);
            const uint32_t nib = (m[u] >> (4 * r)) & 0xFu;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if ((nib >> c) & 1u) {
                    acc = mul_acc<VT, AT>(acc, s_val[o], xs[u][c]);
                    ++o;
                }
            }
            voff += tot;
        }
    }
    if constexpr (S > 1) {
#pragma unroll
        for (int d = S / 2; d >= 1; d >>= 1) acc += __shfl_xor(acc, 4 * d, G);
    }
    return acc;
}

// Block-lane variant: one lane per 4x4 block.  A workgroup owns RPB block rows
// (at most 512 blocks, 2048 values).  Thread t takes tile blocks 2t and 2t+1;
// the tile's values are staged in LDS with coalesced loads; each lane loads the
// whole 32-byte x segment of its block once (two 16-byte loads), walks the 16
// bitmap bits against its LDS values into 4 row partials, and after a barrier
// the partials replace the values in LDS.  Lane (block row, r) then sums row
// r's partials over its blocks in ascending block order, so results are
// deterministic.  Block value offsets come from a workgroup scan of the
// bitmap popcounts (blocks and values of a tile are both contiguous).  Tiles
// over the capacity fall back to the walk on global memory.  x must be
// 16-byte aligned.
template <int RPB, int BPT, typename VT, typename AT, typename EP>
__global__ __launch_bounds__(256) void k_spmv_blk(
    int32_t brow, int32_t nrows, int32_t ncols, const int32_t *__restrict__ blk_ptr,
    const int32_t *__restrict__ blk_col, const uint16_t *__restrict__ blk_map,
    const int32_t *__restrict__ val_ptr, const VT *__restrict__ val, const double *__restrict__ x,
    EP ep, double *y) {
    constexpr int CAPB = 256 * BPT, CAPV = 4 * CAPB, VPT = CAPV / 256;
    static_assert(RPB < 256, "one thread per row offset");
    __shared__ AT s_buf[CAPV];
    __shared__ int32_t s_bp[RPB + 1], s_vp[RPB + 1];
    __shared__ int32_t s_sc[4];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wg = remap_wg(blockIdx.x, gridDim.x);
    const int br0 = wg * RPB;
    const int nr = min(RPB, brow - br0);
    // the tile's ends: wave-uniform addresses, so scalar loads; the value and
    // block loads wait for nothing else
    const int pb = blk_ptr[br0], pe = blk_ptr[br0 + nr];
    const int vb = val_ptr[br0], ve = val_ptr[br0 + nr];
    const int nb = pe - pb, nv = ve - vb;
    // this thread's row offset for the row sums at the end (clamped index: an
    // unconditional load; stored to LDS once the tile's loads are in flight)
    const int bp_t = blk_ptr[br0 + min(tid, nr)];
    if (nb > CAPB || nv > CAPV) {
        for (int t = tid; t <= nr; t += 256) {
            s_bp[t] = blk_ptr[br0 + t];
            s_vp[t] = val_ptr[br0 + t];
        }
        __syncthreads();
        // oversized tile: 256/RPB lanes per block row walk global memory
        constexpr int G = 256 / RPB >= 4 ? (256 / RPB > 64 ? 64 : 256 / RPB) : 4;
        constexpr int RPW = 64 / G;
        for (int base = 0; base < nr; base += 4 * RPW) {  // uniform trip count
            const int lr = base + (tid >> 6) * RPW + lane / G;
            if (lr >= nr) continue;  // whole G-lane group
            const int sub = lane % G, r = sub & 3, s = sub >> 2;
            const AT acc = row_walk<G, VT, AT>(blk_col, blk_map, val, s_bp[lr], s_bp[lr + 1],
                                               s_vp[lr], s, r, x);
            if (s == 0) {
                const int row = 4 * (br0 + lr) + r;
                if (row < nrows) {
                    store_y(ep, y, row, ep(row, double(acc)));
                }
            }
        }
        tile_post(ep, wg);
        return;
    }
    // Loads are issued in dependency order and consumed late, so a tile costs
    // two memory round trips after its row pointers, not one per load: block
    // maps and columns, then the tile's values, then the x segments (which
    // need the columns) while the values are still in flight; the values go
    // to LDS last.
    // (1) this thread's BPT consecutive blocks; lanes past the tile's end load
    // its last block (clamped index: an unconditional load nothing waits for)
    if (nb == 0) {  // every row of the tile is empty
        for (int rr = tid; rr < 4 * nr; rr += 256) {
            const int row = 4 * (br0 + (rr >> 2)) + (rr & 3);
            if (row < nrows) store_y(ep, y, row, ep(row, 0.0));
        }
        tile_post(ep, wg);
        return;
    }
    uint32_t mr[BPT];
    int cr[BPT];
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        const int q = min(BPT * tid + u, nb - 1);
        mr[u] = blk_map[pb + q];
        cr[u] = blk_col[pb + q];
    }
    // (2) the tile's values
    constexpr bool kPair = std::is_same<VT, double>::value && std::is_same<AT, double>::value;
    const int head = vb & 1;
    const int npair = (nv - head) >> 1;
    double2 tv[kPair ? VPT / 2 : 1];
    double edge = 0.0;
    // FP32 / FP16 values: E per 16-byte load from the first index vb + head
    // that is a multiple of E, the head and tail elements singly
    constexpr int E = kPair ? 1 : int(16 / sizeof(VT));
    constexpr int NVPT = kPair ? 1 : (VPT / E > 0 ? VPT / E : 1);
    uint4 tq[NVPT];
    VT tsingle = VT(0.0f);
    int vhead = 0, nvec = 0;
    if constexpr (kPair) {
        // 16-byte loads from the first even value index on, the odd head and
        // tail element singly
        // (val is 16-byte aligned: the launcher checks it).  Raw buffer loads
        // over exactly the tile's pairs: a lane past the last pair gets zeros
        // from the range check, so the loads need no branch.
        const double *dval = reinterpret_cast<const double *>(val);
        const __amdgpu_buffer_rsrc_t vrs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<double *>(dval + vb + head), short(0), npair * 16, kBufferDword3);
#pragma unroll
        for (int u = 0; u < VPT / 2; ++u) {
            const auto q = __builtin_amdgcn_raw_buffer_load_b128(vrs, (u * 256 + tid) * 16, 0, 0);
            __builtin_memcpy(&tv[u], &q, 16);
        }
        // the first (even threads) or last (odd threads) value of the tile,
        // for threads 0 and 1 (a buffer load: zero if the tile had no values)
        const __amdgpu_buffer_rsrc_t ers = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<double *>(dval + vb), short(0), nv * 8, kBufferDword3);
        const auto e = __builtin_amdgcn_raw_buffer_load_b64(ers, ((tid & 1) ? nv - 1 : 0) * 8, 0, 0);
        __builtin_memcpy(&edge, &e, 8);
    } else {
        // (val is 16-byte aligned: handle arrays are).  Raw buffer loads over
        // exactly the tile's vectors: lanes past the last one read zeros.
        vhead = (E - (vb % E)) % E;
        if (vhead > nv) vhead = nv;
        nvec = (nv - vhead) / E;
        const __amdgpu_buffer_rsrc_t vrs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<VT *>(val + vb + vhead), short(0), nvec * 16, kBufferDword3);
#pragma unroll
        for (int u = 0; u < NVPT; ++u) {
            const auto q = __builtin_amdgcn_raw_buffer_load_b128(vrs, (u * 256 + tid) * 16, 0, 0);
            __builtin_memcpy(&tq[u], &q, 16);
        }
        // head (threads 0 .. vhead-1) and tail (threads 8 .. 8 + tail - 1) elements
        const int ntail = nv - vhead - nvec * E;
        if (tid < vhead) tsingle = val[vb + tid];
        else if (tid >= 8 && tid < 8 + ntail) tsingle = val[vb + vhead + nvec * E + (tid - 8)];
    }
    // (3) x segments of the blocks
    uint32_t m[BPT];
    int c4[BPT];
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        const bool ok = BPT * tid + u < nb;
        m[u] = ok ? mr[u] : 0u;
        c4[u] = ok ? 4 * cr[u] : 0;
    }
    // Unconditional 16-byte loads (no branch for the compiler to wait at): a
    // block reaching past the last column, and an empty lane, read x[0..3]
    // instead; an empty lane's segment is never used (no bit is set), a
    // partial block is reloaded element-wise below.
    double xs[BPT][4];
    const bool xvec = ncols >= 4;
    int xb[BPT];
#pragma unroll
    for (int u = 0; u < BPT; ++u) xb[u] = (c4[u] + 4 <= ncols) ? c4[u] : 0;
    if (xvec) {
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const double2 lo2 = *reinterpret_cast<const double2 *>(x + xb[u]);
            const double2 hi2 = *reinterpret_cast<const double2 *>(x + xb[u] + 2);
            xs[u][0] = lo2.x;
            xs[u][1] = lo2.y;
            xs[u][2] = hi2.x;
            xs[u][3] = hi2.y;
        }
    }
    // (4) the values and row offsets into LDS
    if (tid <= nr) s_bp[tid] = bp_t;
    if constexpr (kPair) {
        if (tid == 0 && head) s_buf[0] = edge;
        if (tid == 1 && head + 2 * npair < nv) s_buf[nv - 1] = edge;
#pragma unroll
        for (int u = 0; u < VPT / 2; ++u) {
            const int p = u * 256 + tid;
            if (p < npair) {
                s_buf[head + 2 * p] = tv[u].x;
                s_buf[head + 2 * p + 1] = tv[u].y;
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < NVPT; ++u) {
            const int p = u * 256 + tid;
            if (p < nvec) {
                VT e[E];
                __builtin_memcpy(e, &tq[u], 16);
#pragma unroll
                for (int k = 0; k < E; ++k) s_buf[vhead + E * p + k] = AT(e[k]);
            }
        }
        const int ntail = nv - vhead - nvec * E;
        if (tid < vhead) s_buf[tid] = AT(tsingle);
        else if (tid >= 8 && tid < 8 + ntail) s_buf[vhead + nvec * E + (tid - 8)] = AT(tsingle);
    }
    // (5) blocks reaching past the last column: x element-wise
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        if (m[u] && (!xvec || c4[u] + 4 > ncols)) {
#pragma unroll
            for (int c = 0; c < 4; ++c) xs[u][c] = (c4[u] + c < ncols) ? x[c4[u] + c] : 0.0;
        }
    }
    // exclusive scan of the popcounts over the workgroup -> value offsets
    int pcs[BPT];
    int mine = 0;
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        pcs[u] = __popc(m[u]);
        mine += pcs[u];
    }
    int incl = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    if (lane == 63) s_sc[tid >> 6] = incl;
    __syncthreads();  // also publishes the staged values
    int o = incl - mine;
    for (int q = 0; q < (tid >> 6); ++q) o += s_sc[q];
    AT part[BPT][4];
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        part[u][0] = part[u][1] = part[u][2] = part[u][3] = AT(0);
#pragma unroll
        for (int bit = 0; bit < 16; ++bit) {
            if ((m[u] >> bit) & 1u) {
                part[u][bit >> 2] = mul_acc<AT, AT>(part[u][bit >> 2], s_buf[o], xs[u][bit & 3]);
                ++o;
            }
        }
    }
    __syncthreads();  // every value read: the buffer now takes 4 partials per block
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        const int q = BPT * tid + u;
        if (q < nb) {
#pragma unroll
            for (int r = 0; r < 4; ++r) s_buf[4 * q + r] = part[u][r];
        }
    }
    __syncthreads();
    // row sums: LPR lanes per row when the tile has fewer than 256 rows (long
    // block rows: lane `sub` adds blocks sub, sub + LPR, ... in ascending
    // order, then a fixed xor tree combines the lanes: deterministic)
    constexpr int LPR = 4 * RPB >= 256 ? 1 : 256 / (4 * RPB);
    if constexpr (LPR == 1) {
        for (int rr = tid; rr < 4 * nr; rr += 256) {
            const int lr = rr >> 2, r = rr & 3;
            AT acc = AT(0);
            for (int j = s_bp[lr] - pb; j < s_bp[lr + 1] - pb; ++j) acc += s_buf[4 * j + r];
            const int row = 4 * (br0 + lr) + r;
            if (row < nrows) store_y(ep, y, row, ep(row, double(acc)));
        }
    } else {
        const int rr = tid / LPR, sub = tid % LPR;  // rr < 4 RPB: every thread has a row slot
        AT acc = AT(0);
        if (rr < 4 * nr) {
            const int lr = rr >> 2, r = rr & 3;
            for (int j = s_bp[lr] - pb + sub; j < s_bp[lr + 1] - pb; j += LPR)
                acc += s_buf[4 * j + r];
        }
#pragma unroll
        for (int d = LPR / 2; d >= 1; d >>= 1) acc += __shfl_xor(acc, d);
        if (sub == 0 && rr < 4 * nr) {
            const int row = 4 * (br0 + (rr >> 2)) + (rr & 3);
            if (row < nrows) store_y(ep, y, row, ep(row, double(acc)));
        }
    }
    tile_post(ep, wg);
}

// Balanced variant (counterpart of bsr_spmv_balanced_tc_fp64,
// csr_matvec_device.c:556-613, with the split of BSR_BALANCED_PREPROCESS_GPU
// :1269-1300): workgroup t owns blocks [t*CAPB, (t+1)*CAPB) whatever rows they
// belong to, so a block row of 10^5 blocks is spread over many workgroups and
// no workgroup waits on a long row.  The products are the block-lane ones
// (lane per block, whole x segment, LDS row reduction).  A row that lies
// entirely in the tile is stored to y; a row cut by a tile edge leaves its
// partial sums in carry[t] (slots 0-3: the row the tile starts inside, slots
// 4-7: the row the tile ends inside) and k_spmv_bal_fix adds them in tile
// order.  No atomics: results are deterministic.
//   tile_lo[t]   first block row whose first block lies in tile t
//                (tile_lo[ntiles] = brow, so trailing empty rows belong to the
//                last tile)
//   tile_head[t] block row that starts before the tile and continues into it,
//                or -1
//   tile_vb[t]   value offset of block t*CAPB (tile_vb[ntiles] = nnz)
// Values are staged in LDS when the tile holds at most 8 per block on
// average; denser tiles read them from global memory.  x must be 16-byte
// aligned.
template <int BPT, typename VT, typename AT, typename EP>
__global__ __launch_bounds__(256) void k_spmv_bal(
    int32_t ntiles, int32_t nblk, int32_t nrows, int32_t ncols,
    const int32_t *__restrict__ blk_ptr, const int32_t *__restrict__ blk_col,
    const uint16_t *__restrict__ blk_map, const VT *__restrict__ val,
    const int32_t *__restrict__ tile_lo, const int32_t *__restrict__ tile_head,
    const int32_t *__restrict__ tile_vb, double *__restrict__ carry, const double *__restrict__ x,
    EP ep, double *y) {
    constexpr int CAPB = 256 * BPT, CAPV = 8 * CAPB, VPT = CAPV / 256;
    __shared__ AT s_buf[CAPV];
    __shared__ int32_t s_sc[4];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int t = remap_wg(blockIdx.x, gridDim.x);
    const int q0 = t * CAPB;
    const int q1 = int(min(int64_t(q0) + CAPB, int64_t(nblk)));
    const int nb = q1 - q0;
    const int vb = tile_vb[t];
    const int nv = tile_vb[t + 1] - vb;
    const bool staged = nv <= CAPV;
    uint32_t m[BPT];
    int c4[BPT];
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        const int q = BPT * tid + u;
        m[u] = 0u;
        c4[u] = 0;
        if (q < nb) {
            m[u] = blk_map[q0 + q];
            c4[u] = 4 * blk_col[q0 + q];
        }
    }
    if (staged) {
        VT tv[VPT];
#pragma unroll
        for (int u = 0; u < VPT; ++u) {
            const int q = u * 256 + tid;
            if (q < nv) tv[u] = val[vb + q];
        }
#pragma unroll
        for (int u = 0; u < VPT; ++u) {
            const int q = u * 256 + tid;
            if (q < nv) s_buf[q] = AT(tv[u]);
        }
    }
    double xs[BPT][4];
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        xs[u][0] = xs[u][1] = xs[u][2] = xs[u][3] = 0.0;
        if (m[u]) {
            const double *xp = x + c4[u];
            if (c4[u] + 4 <= ncols) {
                const double2 lo2 = *reinterpret_cast<const double2 *>(xp);
                const double2 hi2 = *reinterpret_cast<const double2 *>(xp + 2);
                xs[u][0] = lo2.x;
                xs[u][1] = lo2.y;
                xs[u][2] = hi2.x;
                xs[u][3] = hi2.y;
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    if (c4[u] + c < ncols) xs[u][c] = xp[c];
            }
        }
    }
    int mine = 0;
#pragma unroll
    for (int u = 0; u < BPT; ++u) mine += __popc(m[u]);
    int incl = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d);
        if (lane >= d) incl += v;
    }
    if (lane == 63) s_sc[tid >> 6] = incl;
    __syncthreads();  // also publishes the staged values
    int o = incl - mine;
    for (int q = 0; q < (tid >> 6); ++q) o += s_sc[q];
    AT part[BPT][4];
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        part[u][0] = part[u][1] = part[u][2] = part[u][3] = AT(0);
#pragma unroll
        for (int bit = 0; bit < 16; ++bit) {
            if ((m[u] >> bit) & 1u) {
                const AT v = staged ? s_buf[o] : AT(val[vb + o]);
                part[u][bit >> 2] = mul_acc<AT, AT>(part[u][bit >> 2], v, xs[u][bit & 3]);
                ++o;
            }
        }
    }
    __syncthreads();  // every staged value read: the buffer takes 4 partials per block
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        const int q = BPT * tid + u;
        if (q < nb) {
#pragma unroll
            for (int r = 0; r < 4; ++r) s_buf[4 * q + r] = part[u][r];
        }
    }
    __syncthreads();
    const int lo = tile_lo[t], hi = tile_lo[t + 1];
    const int head = tile_head[t];
    const int nh = head >= 0 ? 1 : 0;
    const int nrow_t = nh + (hi - lo);
    // lanes per row: as many as 256 threads allow for the tile's rows (a tile
    // of a few long block rows), a power of two; lane `sub` adds the row's
    // blocks sub, sub + lpr, ... in ascending order, a fixed xor tree combines
    // the lanes (deterministic).  lpr is the same for the whole workgroup.
    int lpr = 1;
    while (lpr < 64 && 8 * nrow_t * lpr <= 256) lpr <<= 1;
    for (int rr0 = 0; rr0 < 4 * nrow_t * lpr; rr0 += 256) {  // uniform trip count
        const int rr = (rr0 + tid) / lpr, sub = (rr0 + tid) % lpr;
        const bool live = rr < 4 * nrow_t;
        int R = 0, rb = 0, re = 0, r = 0;
        AT acc = AT(0);
        if (live) {
            const int idx = rr >> 2;
            r = rr & 3;
            R = (idx < nh) ? head : lo + idx - nh;
            rb = blk_ptr[R];
            re = blk_ptr[R + 1];
            const int j0 = max(rb, q0) - q0, j1 = min(re, q1) - q0;
            for (int j = j0 + sub; j < j1; j += lpr) acc += s_buf[4 * j + r];
        }
        for (int d = lpr / 2; d >= 1; d >>= 1) acc += __shfl_xor(acc, d);
        if (live && sub == 0) {
            if (rb < q0) {
                carry[8 * int64_t(t) + r] = double(acc);  // row started in an earlier tile
            } else if (re > q1) {
                carry[8 * int64_t(t) + 4 + r] = double(acc);  // row continues in later tiles
            } else {
                const int row = 4 * R + r;
                if (row < nrows) {
                    y[row] = ep(row, double(acc));
                }
            }
        }
    }
}

// Finishes the block rows cut by tile edges: thread (t, r) owns the row that
// starts in tile t and ends in a later one, and adds its partials in tile
// order (tile t's tail, then the heads of t+1, t+2, ...).
template <typename EP>
__global__ __launch_bounds__(256) void k_spmv_bal_fix(
    int32_t ntiles, int32_t capb, int32_t nblk, int32_t nrows,
    const int32_t *__restrict__ blk_ptr, const int32_t *__restrict__ tile_lo,
    const int32_t *__restrict__ tile_head, const double *__restrict__ carry, EP ep, double *y) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = gid >> 2, r = gid & 3;
    if (t >= ntiles) return;
    const int lo = tile_lo[t], hi = tile_lo[t + 1];
    if (hi == lo) return;  // no row starts in this tile
    const int R = hi - 1;
    const int q1 = int(min(int64_t(t + 1) * capb, int64_t(nblk)));
    if (blk_ptr[R + 1] <= q1) return;  // the tile's last row ends inside it
    double acc = carry[8 * int64_t(t) + 4 + r];
    for (int t2 = t + 1; t2 < ntiles && tile_head[t2] == R; ++t2) acc += carry[8 * int64_t(t2) + r];
    const int row = 4 * R + r;
    if (row < nrows) {
        y[row] = ep(row, acc);
    }
}

// Tile tables of the balanced kernel, one thread per tile (tile_vb comes from
// a scan of k_tile_popc's per-tile value counts).
__global__ __launch_bounds__(256) void k_tile_rows(int32_t ntiles, int32_t capb, int32_t brow,
                                                   const int32_t *__restrict__ blk_ptr,
                                                   int32_t *__restrict__ tile_lo,
                                                   int32_t *__restrict__ tile_head) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    if (t == ntiles) {
        tile_lo[t] = brow;
        return;
    }
    const int q0 = t * capb;
    // lo = first block row R with blk_ptr[R] >= q0 (blk_ptr[brow] = nblk > q0)
    int a = 0, c = brow;
    while (a < c) {
        const int mid = (a + c) >> 1;
        if (blk_ptr[mid] >= q0) c = mid;
        else a = mid + 1;
    }
    const int lo = (t == 0) ? 0 : a;
    tile_lo[t] = lo;
    // the block row holding block q0 starts before it: it is this tile's head
    tile_head[t] = (t > 0 && blk_ptr[lo] > q0) ? lo - 1 : -1;
}

__global__ __launch_bounds__(256) void k_tile_popc(int32_t nblk, int32_t capb,
                                                   const uint16_t *__restrict__ blk_map,
                                                   int32_t *__restrict__ tile_cnt) {
    __shared__ int s_part[4];
    const int t = blockIdx.x;
    const int q0 = t * capb, q1 = int(min(int64_t(q0) + capb, int64_t(nblk)));
    int n = 0;
    for (int q = q0 + int(threadIdx.x); q < q1; q += 256) n += __popc(uint32_t(blk_map[q]));
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[t] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
}

// Exclusive scan of v[0..n-1] in place into v[0..n] (v[n] = total) by one
// workgroup of 1024 threads, each owning a contiguous chunk: the balanced
// kernel's value offsets per tile when no scan temporaries may be allocated.
__global__ __launch_bounds__(1024) void k_scan_excl_1wg(int32_t n, int32_t *__restrict__ v) {
    __shared__ int32_t s_w[16];
    const int tid = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int i0 = min(n, tid * per), i1 = min(n, i0 + per);
    int sum = 0;
    for (int i = i0; i < i1; ++i) sum += v[i];
    int incl = sum;
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d);
        if ((tid & 63) >= d) incl += t;
    }
    if ((tid & 63) == 63) s_w[tid >> 6] = incl;
    __syncthreads();
    int off = incl - sum;
    for (int w = 0; w < (tid >> 6); ++w) off += s_w[w];
    for (int i = i0; i < i1; ++i) {
        const int c = v[i];
        v[i] = off;
        off += c;
    }
    if (tid == 1023) v[n] = off;  // the last chunk ends at the total
}

template <int G, typename VT, typename AT, int CAPV, int CAPB, int MAXB, bool XSEG, typename EP>
__global__ __launch_bounds__(256) void k_spmv_staged(
    int32_t brow, int32_t nrows, int32_t ncols, const int32_t *__restrict__ blk_ptr,
    const int32_t *__restrict__ blk_col, const uint16_t *__restrict__ blk_map,
    const int32_t *__restrict__ val_ptr, const VT *__restrict__ val, const double *__restrict__ x,
    EP ep, double *y) {
    constexpr int RPW = 64 / G;
    constexpr int RPB = 256 / G;
    __shared__ VT s_val[CAPV];
    __shared__ int32_t s_col[CAPB];
    __shared__ uint16_t s_map[CAPB];
    __shared__ int32_t s_bp[RPB + 1], s_vp[RPB + 1];
    const int wg = remap_wg(blockIdx.x, gridDim.x);
    const int br0 = wg * RPB;
    const int nr = min(RPB, brow - br0);
    for (int t = threadIdx.x; t <= nr; t += 256) {
        s_bp[t] = blk_ptr[br0 + t];
        s_vp[t] = val_ptr[br0 + t];
    }
    __syncthreads();
    const int pb = s_bp[0], pe = s_bp[nr];
    const int vb = s_vp[0], ve = s_vp[nr];
    const bool staged = (pe - pb <= CAPB) && (ve - vb <= CAPV);
    if (staged) {
        // all loads of a thread are issued before any LDS store (memory-level parallelism)
        constexpr int VPT = CAPV / 256;
        constexpr int BPT = CAPB / 256;
        const int nv = ve - vb, nb = pe - pb;
        VT tv[VPT];
        int32_t tc[BPT];
        uint16_t tm[BPT];
#pragma unroll
        for (int u = 0; u < VPT; ++u) {
            const int t = u * 256 + threadIdx.x;
            if (t < nv) tv[u] = stream_load(val + vb + t);
        }
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const int t = u * 256 + threadIdx.x;
            if (t < nb) {
                tc[u] = stream_load(blk_col + pb + t);
                tm[u] = stream_load(blk_map + pb + t);
            }
        }
#pragma unroll
        for (int u = 0; u < VPT; ++u) {
            const int t = u * 256 + threadIdx.x;
            if (t < nv) s_val[t] = tv[u];
        }
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const int t = u * 256 + threadIdx.x;
            if (t < nb) {
                s_col[t] = tc[u];
                s_map[t] = tm[u];
            }
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int lr = (threadIdx.x >> 6) * RPW + lane / G;  // block row within the tile
    if (lr >= nr) return;
    const int sub = lane % G;
    const int r = sub & 3;
    const int s = sub >> 2;
    AT acc;
    if (staged) {
        acc = row_walk_prefetch<G, MAXB, XSEG, VT, AT>(s_col, s_map, s_val, s_bp[lr] - pb,
                                              s_bp[lr + 1] - pb, s_vp[lr] - vb, s, r, x, ncols);
    } else {
        acc = row_walk<G, VT, AT>(blk_col, blk_map, val, s_bp[lr], s_bp[lr + 1], s_vp[lr], s, r, x);
    }
    if (s == 0) {
        const int row = 4 * (br0 + lr) + r;
        if (row < nrows) {
            y[row] = ep(row, double(acc));
        }
    }
}

// k right-hand sides (column-major, leading dimensions), KT at a time.
template <int G, int KT, typename VT, typename AT>
__global__ __launch_bounds__(256) void k_spmm(
    int32_t brow, int32_t nrows, const int32_t *__restrict__ blk_ptr,
    const int32_t *__restrict__ blk_col, const uint16_t *__restrict__ blk_map,
    const int32_t *__restrict__ val_ptr, const VT *__restrict__ val, int k0, int kn,
    const double *__restrict__ x, int64_t ldx, double alpha, double beta, const double *b,
    int64_t ldb, double *y, int64_t ldy) {
    constexpr int S = G / 4;
    constexpr int RPW = 64 / G;
    const int lane = threadIdx.x & 63;
    const int wg = remap_wg(blockIdx.x, gridDim.x);
    const int wave = wg * 4 + (threadIdx.x >> 6);
    const int br = wave * RPW + lane / G;
    const int sub = lane % G;
    const int r = sub & 3;
    const int s = sub >> 2;
    if (br >= brow) return;
    const int p0 = blk_ptr[br];
    const int p1 = blk_ptr[br + 1];
    int voff = val_ptr[br];
    const uint32_t low = (1u << (4 * r)) - 1u;
    AT acc[KT];
#pragma unroll
    for (int v = 0; v < KT; ++v) acc[v] = AT(0);
    const double *xk = x + int64_t(k0) * ldx;
    for (int jb = p0; jb < p1; jb += S) {
        const int j = jb + s;
        uint32_t m = 0;
        int col = 0;
        if (j < p1) {
            m = blk_map[j];
            col = blk_col[j];
        }
        const int cnt = __popc(m);
        int incl = cnt;
        int tot = cnt;
        if constexpr (S > 1) {
#pragma unroll
            for (int d = 1; d < S; d <<= 1) {
                int t = __shfl_up(incl, 4 * d, G);
                if (s >= d) incl += t;
            }
            tot = __shfl(incl, 4 * (S - 1) + r, G);
        }
        int o = voff + (incl - cnt) + __popc(m & low);
        uint32_t nib = (m >> (4 * r)) & 0xFu;
        while (nib) {
            const int c = __builtin_ctz(nib);
            nib &= nib - 1u;
            const VT a = val[o];
            const double *xs = xk + 4 * int64_t(col) + c;
#pragma unroll
            for (int v = 0; v < KT; ++v)
                if (v < kn) acc[v] = mul_acc<VT, AT>(acc[v], a, xs[v * ldx]);
            ++o;
        }
        voff += tot;
    }
#pragma unroll
    for (int v = 0; v < KT; ++v) {
        if constexpr (S > 1) {
#pragma unroll
            for (int d = S / 2; d >= 1; d >>= 1) acc[v] += __shfl_xor(acc[v], 4 * d, G);
        }
    }
    if (s == 0) {
        const int row = 4 * br + r;
        if (row < nrows) {
#pragma unroll
            for (int v = 0; v < KT; ++v) {
                if (v < kn) {
                    double val_out = alpha * double(acc[v]);
                    if (beta != 0.0) val_out = fma(beta, b[int64_t(k0 + v) * ldb + row], val_out);
                    y[int64_t(k0 + v) * ldy + row] = val_out;
                }
            }
        }
    }
}

// MFMA SpMM for k >= 8 right-hand sides: one wave per block row.  For each
// block j of the row, v_mfma_f64_4x4x4f64 multiplies the 4x4 block (the same
// A operand in all four block slots) by four 4x4 slices of the X segment
// (block slot b takes right-hand sides 16g + 4b .. 16g + 4b + 3), so one
// instruction advances 16 right-hand sides and the accumulators stay in
// registers across the row's blocks; NG groups of 16 share each block's A
// operand.  The row's block bitmaps, columns and value offsets are loaded
// across the lanes first, and the operands of U = 4 blocks are gathered before
// their MFMAs.  Operand layout (measured on gfx950, block slot b): A_b[i][k] in
// lane 16k+4b+i, B_b[k][n] in lane 16k+4b+n, D_b[i][n] in lane 16i+4b+n.
// Multivectors are interleaved (HYPRE multivector storage
// method 1): entry (row i, vector v) at x[i * kx + v], so the B operand of one
// instruction is four rows x 16 consecutive right-hand sides (four 128-byte
// lines) instead of 16 separate 32-byte pieces of column-major vectors.
// Writes y[row * ky + v] = alpha (A x)[row][v] + beta b[row * kb + v].
template <int NG, typename VT>
__global__ __launch_bounds__(256) void k_spmm_mfma_il(
    int32_t brow, int32_t nrows, int32_t ncols, const int32_t *__restrict__ blk_ptr,
    const int32_t *__restrict__ blk_col, const uint16_t *__restrict__ blk_map,
    const int32_t *__restrict__ val_ptr, const VT *__restrict__ val, int k0, int kn,
    const double *__restrict__ x, int64_t kx, double alpha, double beta, const double *b,
    int64_t kb, double *y, int64_t ky) {
    constexpr int U = 4;
    const int lane = threadIdx.x & 63;
    const int br = remap_wg(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (br >= brow) return;  // whole wave
    const int kk = lane >> 4, bs = (lane >> 2) & 3, nn = lane & 3;
    const int eA = 4 * nn + kk;
    const uint32_t belowA = (1u << eA) - 1u;
    const int p0 = blk_ptr[br], p1 = blk_ptr[br + 1];
    int vrel[NG];
    bool xon[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int rhs = 16 * g + 4 * bs + nn;
        xon[g] = rhs < kn;
        vrel[g] = k0 + (xon[g] ? rhs : 0);
    }
    double acc[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) acc[g] = 0.0;
    int voff = val_ptr[br];
    for (int c0 = p0; c0 < p1; c0 += 64) {
        const int nb = min(64, p1 - c0);
        uint32_t mj = 0u;
        int cj = 0;
        if (lane < nb) {
            mj = blk_map[c0 + lane];
            cj = blk_col[c0 + lane];
        }
        int inc = __popc(mj);
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(inc, d);
            if (lane >= d) inc += t;
        }
        const int oj = voff + inc - __popc(mj);
        voff += __shfl(inc, 63);
        for (int j0 = 0; j0 < nb; j0 += U) {
            double a[U], bx[U][NG];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = j0 + u < nb ? j0 + u : nb - 1;
                const uint32_t m = uint32_t(__shfl(int(mj), j));
                const int xrow = 4 * __shfl(cj, j) + kk;
                const int o = __shfl(oj, j);
                const bool on = j0 + u < nb;
                a[u] = (on && ((m >> eA) & 1u)) ? double(val[o + __popc(m & belowA)]) : 0.0;
#pragma unroll
                for (int g = 0; g < NG; ++g)
                    bx[u][g] = (on && xon[g] && xrow < ncols) ? x[int64_t(xrow) * kx + vrel[g]] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int g = 0; g < NG; ++g)
                    acc[g] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[u], bx[u][g], acc[g], 0, 0, 0);
        }
    }
    const int row = 4 * br + kk;
    if (row < nrows) {
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            if (!xon[g]) continue;
            double out = alpha * acc[g];
            if (beta != 0.0) out = fma(beta, b[int64_t(row) * kb + vrel[g]], out);
            y[int64_t(row) * ky + vrel[g]] = out;
        }
    }
}

// Tiled transposes between the two multivector layouts (32 x 32 tiles in LDS,
// both sides coalesced).  cm -> il: xi[i * k + v] = x[v * ldx + i].
__global__ __launch_bounds__(256) void k_cm_to_il(int32_t n, int32_t k,
                                                  const double *__restrict__ x, int64_t ldx,
                                                  double *__restrict__ xi) {
    __shared__ double tile[32][33];
    const int i0 = blockIdx.x * 32, v0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int v = v0 + r, i = i0 + tx;
        tile[r][tx] = (v < k && i < n) ? x[int64_t(v) * ldx + i] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, v = v0 + tx;
        if (i < n && v < k) xi[int64_t(i) * k + v] = tile[tx][r];
    }
}

// il -> cm with the SpMM epilogue: y[v * ldy + i] = alpha yi[i * k + v] + beta b[v * ldb + i]
__global__ __launch_bounds__(256) void k_il_to_cm(int32_t n, int32_t k,
                                                  const double *__restrict__ yi, double alpha,
                                                  double beta, const double *b, int64_t ldb,
                                                  double *y, int64_t ldy) {
    __shared__ double tile[32][33];
    const int i0 = blockIdx.x * 32, v0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, v = v0 + tx;
        tile[r][tx] = (i < n && v < k) ? yi[int64_t(i) * k + v] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int v = v0 + r, i = i0 + tx;
        if (v < k && i < n) {
            double out = alpha * tile[tx][r];
            if (beta != 0.0) out = fma(beta, b[int64_t(v) * ldb + i], out);
            y[int64_t(v) * ldy + i] = out;
        }
    }
}

// y = beta*b (beta == 0 -> y = 0): used when the matrix has no blocks
__global__ void k_scale(int32_t n, double beta, const double *b, double *y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (beta == 0.0) ? 0.0 : beta * b[i];
}

__global__ void k_cast_f32(int64_t n, const double *__restrict__ in, float *__restrict__ out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = float(in[i]);
}

// per-workgroup maximum of |in[i]| over a grid-stride range (NaN propagates
// as a failed "fits" test by the caller)
__global__ __launch_bounds__(256) void k_absmax_part(int64_t n, const double *__restrict__ in,
                                                     double *__restrict__ out) {
    __shared__ double s_m[4];
    double m = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const double v = fabs(in[i]);
        m = (v > m || v != v) ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off);
        m = (o > m || o != o) ? o : m;
    }
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = s_m[0];
        for (int w = 1; w < 4; ++w) r = (s_m[w] > r || s_m[w] != s_m[w]) ? s_m[w] : r;
        out[blockIdx.x] = r;
    }
}

__global__ void k_cast_f16(int64_t n, const double *__restrict__ in, double scale,
                           __half *__restrict__ out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = __float2half(float(in[i] * scale));
}

template <int G, typename VT, typename AT, typename EP>
int launch_spmv_g(const Mbsr &A, const VT *val, const double *x, EP ep, double *y, hipStream_t s) {
    constexpr int RPW = 64 / G;
    const int64_t waves = (int64_t(A.brow) + RPW - 1) / RPW;
    const int64_t grid = (waves + 3) / 4;
    if (grid > 0x7fffffff) {
        set_error("SpMV grid too large");
        return AMGT_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL((k_spmv<G, VT, AT, EP>), dim3(unsigned(grid)), dim3(256), 0, s, A.brow,
                       A.nrows, A.blk_ptr, A.blk_col, A.blk_map, A.val_ptr, val, x, ep, y);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

constexpr int kStageVals = 2048;
constexpr int kStageBlocks = 1024;

template <int G, typename VT, typename AT, typename EP>
int launch_spmv_staged_g(const Mbsr &A, const VT *val, const double *x, EP ep, double *y,
                         hipStream_t s) {
    constexpr int RPB = 256 / G;
    const int64_t grid = (int64_t(A.brow) + RPB - 1) / RPB;
    hipLaunchKernelGGL((k_spmv_staged<G, VT, AT, kStageVals, kStageBlocks, 4, true, EP>),
                       dim3(unsigned(grid)), dim3(256), 0, s, A.brow, A.nrows, A.ncols,
                       A.blk_ptr, A.blk_col, A.blk_map, A.val_ptr, val, x, ep, y);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

template <int RPB, int BPT, typename VT, typename AT, typename EP>
int launch_spmv_blk_g(const Mbsr &A, const VT *val, const double *x, EP ep, double *y,
                      hipStream_t s) {
    const int64_t grid = (int64_t(A.brow) + RPB - 1) / RPB;
    hipLaunchKernelGGL((k_spmv_blk<RPB, BPT, VT, AT, EP>), dim3(unsigned(grid)), dim3(256), 0, s,
                       A.brow, A.nrows, A.ncols, A.blk_ptr, A.blk_col, A.blk_map, A.val_ptr, val,
                       x, ep, y);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

// Tile tables and carry buffer of the balanced kernel: the handle's own
// (spmv_prepare) or a caller's workspace (spmv_balanced_ws).
struct BalTables {
    int32_t ntiles, capb;
    const int32_t *lo, *head, *vb;
    double *carry;
};

inline BalTables handle_tables(const Mbsr &A) {
    return BalTables{A.bal_ntiles, A.bal_capb, A.bal_lo, A.bal_head, A.bal_vb, A.bal_carry};
}

template <int BPT, typename VT, typename AT, typename EP>
int launch_spmv_bal_g(const Mbsr &A, const BalTables &tb, const VT *val, const double *x, EP ep,
                      double *y, hipStream_t s) {
    const int nt = tb.ntiles;
    hipLaunchKernelGGL((k_spmv_bal<BPT, VT, AT, EP>), dim3(unsigned(nt)), dim3(256), 0, s, nt,
                       int32_t(A.nblk), A.nrows, A.ncols, A.blk_ptr, A.blk_col, A.blk_map, val,
                       tb.lo, tb.head, tb.vb, tb.carry, x, ep, y);
    AMGT_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_spmv_bal_fix<EP>), dim3(unsigned((4 * int64_t(nt) + 255) / 256)),
                       dim3(256), 0, s, nt, tb.capb, int32_t(A.nblk), A.nrows, A.blk_ptr,
                       tb.lo, tb.head, tb.carry, ep, y);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

// y[row] = ep(row, (A x)[row]) with the handle's kernel kind and lanes setting
template <typename VT, typename AT, typename EP>
int launch_spmv_t(const Mbsr &A, const VT *val, int lanes, const double *x, EP ep, double *y,
                  hipStream_t s) {
    if (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64) {
        set_error("lanes per block row must be 4, 8, 16, 32 or 64");
        return AMGT_ERR_ARG;
    }
    // the staged, block and balanced kernels read x in 16-byte pieces;
    // otherwise the direct kernel
    const bool x_aligned = (reinterpret_cast<uintptr_t>(x) & 15u) == 0;
    if (A.spmv_kernel == kSpmvBalanced && x_aligned && A.bal_ntiles > 0 &&
        A.bal_capb == 256 * balanced_bpt(lanes)) {
        const BalTables tb = handle_tables(A);
        switch (balanced_bpt(lanes)) {
            case 1: return launch_spmv_bal_g<1, VT, AT>(A, tb, val, x, ep, y, s);
            case 4: return launch_spmv_bal_g<4, VT, AT>(A, tb, val, x, ep, y, s);
            default: return launch_spmv_bal_g<2, VT, AT>(A, tb, val, x, ep, y, s);
        }
    }
    // values are staged in 16-byte pieces (handle arrays always are aligned)
    const bool val_aligned = (reinterpret_cast<uintptr_t>(val) & 15u) == 0;
    if ((A.spmv_kernel == kSpmvBlock || A.spmv_kernel == kSpmvBalanced) && x_aligned &&
        val_aligned) {
        // lanes selects the tile: 4 -> 64 block rows (2 blocks per thread),
        // 8 -> 128 (4 per thread), 16 -> 32, 32 -> 16, 64 -> 8 block rows
        switch (lanes) {
            case 4: return launch_spmv_blk_g<64, 2, VT, AT>(A, val, x, ep, y, s);
            case 8: return launch_spmv_blk_g<128, 4, VT, AT>(A, val, x, ep, y, s);
            case 16: return launch_spmv_blk_g<32, 2, VT, AT>(A, val, x, ep, y, s);
            case 32: return launch_spmv_blk_g<16, 2, VT, AT>(A, val, x, ep, y, s);
            default: return launch_spmv_blk_g<8, 2, VT, AT>(A, val, x, ep, y, s);
        }
    }
    if (A.spmv_kernel == kSpmvStaged && x_aligned) {
        switch (lanes) {
            case 4: return launch_spmv_staged_g<4, VT, AT>(A, val, x, ep, y, s);
            case 8: return launch_spmv_staged_g<8, VT, AT>(A, val, x, ep, y, s);
            case 16: return launch_spmv_staged_g<16, VT, AT>(A, val, x, ep, y, s);
            case 32: return launch_spmv_staged_g<32, VT, AT>(A, val, x, ep, y, s);
            default: return launch_spmv_staged_g<64, VT, AT>(A, val, x, ep, y, s);
        }
    }
    switch (lanes) {
        case 4: return launch_spmv_g<4, VT, AT>(A, val, x, ep, y, s);
        case 8: return launch_spmv_g<8, VT, AT>(A, val, x, ep, y, s);
        case 16: return launch_spmv_g<16, VT, AT>(A, val, x, ep, y, s);
        case 32: return launch_spmv_g<32, VT, AT>(A, val, x, ep, y, s);
        default: return launch_spmv_g<64, VT, AT>(A, val, x, ep, y, s);
    }
}

// the handle's values in `precision`, with the kernel's accumulation type
template <typename EP>
int launch_spmv_prec(const Mbsr &A, int precision, int lanes, const double *x, EP ep, double *y,
                     hipStream_t s) {
    switch (precision) {
        case AMGT_DOUBLE: return launch_spmv_t<double, double>(A, A.val, lanes, x, ep, y, s);
        case AMGT_FLOAT: return launch_spmv_t<float, float>(A, A.val32, lanes, x, ep, y, s);
        case AMGT_HALF:
            return launch_spmv_t<__half, float>(A, A.val16, lanes, x,
                                                EpScaled<EP>{ep, 1.0 / A.half_scale}, y, s);
        default:
            set_error("unknown precision");
            return AMGT_ERR_ARG;
    }
}

template <int G, int KT, typename VT, typename AT>
int launch_spmm_g(const Mbsr &A, const VT *val, int k, double alpha, const double *x,
                  int64_t ldx, double beta, const double *b, int64_t ldb, double *y, int64_t ldy,
                  hipStream_t s) {
    constexpr int RPW = 64 / G;
    const int64_t waves = (int64_t(A.brow) + RPW - 1) / RPW;
    const int64_t grid = (waves + 3) / 4;
    for (int k0 = 0; k0 < k; k0 += KT) {
        int kn = k - k0 < KT ? k - k0 : KT;
        hipLaunchKernelGGL((k_spmm<G, KT, VT, AT>), dim3(unsigned(grid)), dim3(256), 0, s, A.brow,
                           A.nrows, A.blk_ptr, A.blk_col, A.blk_map, A.val_ptr, val, k0, kn, x,
                           ldx, alpha, beta, b, ldb, y, ldy);
        AMGT_HIP(hipGetLastError());
    }
    return AMGT_OK;
}

// interleaved multivectors: 32 right-hand sides per launch
template <typename VT>
int launch_spmm_il(const Mbsr &A, const VT *val, int k, double alpha, const double *x,
                   int64_t kx, double beta, const double *b, int64_t kb, double *y, int64_t ky,
                   hipStream_t s) {
    const int64_t grid = (int64_t(A.brow) + 3) / 4;
    for (int k0 = 0; k0 < k; k0 += 32) {
        const int kn = k - k0 < 32 ? k - k0 : 32;
        if (kn > 16)
            hipLaunchKernelGGL((k_spmm_mfma_il<2, VT>), dim3(unsigned(grid)), dim3(256), 0, s,
                               A.brow, A.nrows, A.ncols, A.blk_ptr, A.blk_col, A.blk_map,
                               A.val_ptr, val, k0, kn, x, kx, alpha, beta, b, kb, y, ky);
        else
            hipLaunchKernelGGL((k_spmm_mfma_il<1, VT>), dim3(unsigned(grid)), dim3(256), 0, s,
                               A.brow, A.nrows, A.ncols, A.blk_ptr, A.blk_col, A.blk_map,
                               A.val_ptr, val, k0, kn, x, kx, alpha, beta, b, kb, y, ky);
        AMGT_HIP(hipGetLastError());
    }
    return AMGT_OK;
}

// column-major multivectors with k >= 8: transpose X to the interleaved
// layout, run the interleaved MFMA kernel, transpose back with the epilogue
template <typename VT>
int launch_spmm_cm_via_il(const Mbsr &A, const VT *val, int k, double alpha, const double *x,
                          int64_t ldx, double beta, const double *b, int64_t ldb, double *y,
                          int64_t ldy, hipStream_t s) {
    double *xi = nullptr, *yi = nullptr;
    AMGT_TRY(tmp_alloc_t(&xi, size_t(A.ncols) * size_t(k), s));
    int st = tmp_alloc_t(&yi, size_t(A.nrows) * size_t(k), s);
    if (st == AMGT_OK) {
        const dim3 gx(unsigned((A.ncols + 31) / 32), unsigned((k + 31) / 32));
        hipLaunchKernelGGL(k_cm_to_il, gx, dim3(256), 0, s, A.ncols, k, x, ldx, xi);
        st = launch_spmm_il<VT>(A, val, k, 1.0, xi, k, 0.0, nullptr, k, yi, k, s);
    }
    if (st == AMGT_OK) {
        const dim3 gy(unsigned((A.nrows + 31) / 32), unsigned((k + 31) / 32));
        hipLaunchKernelGGL(k_il_to_cm, gy, dim3(256), 0, s, A.nrows, k, yi, alpha, beta, b, ldb,
                           y, ldy);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error(std::string("SpMM transpose: ") + hipGetErrorString(e));
            st = AMGT_ERR_HIP;
        }
    }
    tmp_free(xi, s);
    tmp_free(yi, s);
    return st;
}

template <typename VT, typename AT>
int launch_spmm_t(const Mbsr &A, const VT *val, int lanes, int k, double alpha, const double *x,
                  int64_t ldx, double beta, const double *b, int64_t ldb, double *y, int64_t ldy,
                  hipStream_t s) {
    if (k >= 8)
        return launch_spmm_cm_via_il<VT>(A, val, k, alpha, x, ldx, beta, b, ldb, y, ldy, s);
    switch (lanes) {
        case 4: return launch_spmm_g<4, 8, VT, AT>(A, val, k, alpha, x, ldx, beta, b, ldb, y, ldy, s);
        case 8: return launch_spmm_g<8, 8, VT, AT>(A, val, k, alpha, x, ldx, beta, b, ldb, y, ldy, s);
        case 16: return launch_spmm_g<16, 8, VT, AT>(A, val, k, alpha, x, ldx, beta, b, ldb, y, ldy, s);
        case 32: return launch_spmm_g<32, 8, VT, AT>(A, val, k, alpha, x, ldx, beta, b, ldb, y, ldy, s);
        default: return launch_spmm_g<64, 8, VT, AT>(A, val, k, alpha, x, ldx, beta, b, ldb, y, ldy, s);
    }
}

}  // namespace

// The rule below was fitted to the timing sweep (amgt_mbsr_spmv_sweep) of
// every A, P and R of the geometric and algebraic 7-pt 128^3 hierarchies, the
// algebraic anisotropic 160^3 and 1024^2 GRF ones, the geometric 27-pt 64^3
// one and the 48-problem batch (126 matrices, 3 to 4.1M rows): summed over
// them its choices take 1.3 % longer than each matrix's fastest setting (the
// block kernel everywhere took 12 % longer).
//  * up to 150k blocks, up to 1M blocks with at most 5 blocks per block row
//    (interpolation and restriction operators) or up to 350k with at most 8:
//    the direct walk.  These launches are latency bound; the block kernel's
//    LDS staging, scans and barriers cost more than the coalescing gains
//    (about 3 against 4.5-5 us at the small end), whatever the spread;
//  * a spread of blocks per block row >= 12 above that: the balanced kernel;
//  * else the block-lane kernel.
int default_spmv_kernel(const Mbsr &A) {
    const double avg = A.brow ? double(A.nblk) / double(A.brow) : 0.0;
    if (A.nblk <= 150000 || (avg <= 5.0 && A.nblk <= 1000000) || (avg <= 8.0 && A.nblk <= 350000))
        return kSpmvDirect;
    if (A.stand >= kBalancedStand) return kSpmvBalanced;
    return kSpmvBlock;
}

int spmv_default_lanes(const Mbsr &A) {
    const double avg = A.brow ? double(A.nblk) / double(A.brow) : 0.0;
    if (A.spmv_kernel == kSpmvDirect) {
        // lanes per block row so that about 2^18 lanes share the walk: the
        // power of two at or above 0.9 * 2^18 / brow, at least 8 from 64k block
        // rows down, within 4..64
        const double want = 0.9 * 262144.0 / double(A.brow > 0 ? A.brow : 1);
        int g = 4;
        while (g < want && g < 64) g <<= 1;
        if (A.brow <= 65536 && g < 8) g = 8;
        return g;
    }
    if (A.spmv_kernel == kSpmvBalanced) return 16;  // tiles of 256 blocks
    if (A.spmv_kernel == kSpmvBlock) {
        // block-lane tiles of 512 blocks: the most block rows per tile (64, 32,
        // 16 or 8, lanes 4 / 16 / 32 / 64) whose average load stays within 90 %
        // of the capacity, so that few tiles overflow to the global-memory walk
        if (avg * 64.0 <= 460.0) return 4;
        if (avg * 32.0 <= 460.0) return 16;
        if (avg * 16.0 <= 460.0) return 32;
        return 64;
    }
    // enough slots that a typical block row takes about two block steps
    if (avg <= 8.0) return 4;
    if (avg <= 16.0) return 8;
    if (avg <= 32.0) return 16;
    if (avg <= 64.0) return 32;
    return 64;
}

int balanced_bpt(int lanes) { return lanes == 16 ? 1 : (lanes == 8 ? 4 : 2); }

int spmv_prepare(Mbsr *A, int lanes, hipStream_t s) {
    if (A->spmv_kernel != kSpmvBalanced || A->nblk == 0) return AMGT_OK;
    const int capb = 256 * balanced_bpt(lanes);
    if (A->bal_capb == capb && A->bal_ntiles > 0) return AMGT_OK;
    mbsr_release_balanced(A, s);
    const int64_t nt = (A->nblk + capb - 1) / capb;
    A->bal_lo = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * size_t(nt + 1), s));
    A->bal_head = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * size_t(nt), s));
    A->bal_vb = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * size_t(nt + 1), s));
    A->bal_carry = static_cast<double *>(dev_alloc(sizeof(double) * size_t(8 * nt), s));
    if (!A->bal_lo || !A->bal_head || !A->bal_vb || !A->bal_carry) {
        mbsr_release_balanced(A, s);
        set_error("balanced SpMV: tile table allocation failed");
        return AMGT_ERR_ALLOC;
    }
    int32_t *cnt = nullptr;
    AMGT_TRY(tmp_alloc_t(&cnt, size_t(nt), s));
    hipLaunchKernelGGL(k_tile_rows, dim3(unsigned((nt + 1 + 255) / 256)), dim3(256), 0, s,
                       int32_t(nt), capb, A->brow, A->blk_ptr, A->bal_lo, A->bal_head);
    hipLaunchKernelGGL(k_tile_popc, dim3(unsigned(nt)), dim3(256), 0, s, int32_t(A->nblk), capb,
                       A->blk_map, cnt);
    hipError_t e = hipGetLastError();
    int st = AMGT_OK;
    if (e != hipSuccess) {
        set_error(std::string("balanced SpMV setup: ") + hipGetErrorString(e));
        st = AMGT_ERR_HIP;
    }
    if (st == AMGT_OK) st = exclusive_scan_i32(cnt, A->bal_vb, nt, s);
    tmp_free(cnt, s);
    if (st != AMGT_OK) {
        mbsr_release_balanced(A, s);
        return st;
    }
    A->bal_capb = capb;
    A->bal_ntiles = int32_t(nt);
    return AMGT_OK;
}

// Workspace of the balanced kernel for nt tiles: tile_lo[nt+1], tile_head[nt],
// tile_vb[nt+1] (16-byte aligned pieces), then carry[8 nt] doubles.
struct BalWsLayout {
    size_t head, vb, carry, bytes;
};

static size_t up16(size_t b) { return (b + 15) & ~size_t(15); }

static BalWsLayout bal_ws_layout(int64_t nt) {
    BalWsLayout w;
    w.head = up16(4 * size_t(nt + 1));
    w.vb = w.head + up16(4 * size_t(nt));
    w.carry = w.vb + up16(4 * size_t(nt + 1));
    w.bytes = w.carry + 64 * size_t(nt);
    return w;
}

size_t spmv_balanced_ws_bytes(const Mbsr &A, int lanes) {
    if (A.nblk == 0) return 0;
    const int64_t capb = 256 * balanced_bpt(lanes);
    return bal_ws_layout((A.nblk + capb - 1) / capb).bytes;
}

int spmv_balanced_ws(const Mbsr &A, int lanes, double alpha, const double *x, double beta,
                     const double *b, double *y, void *ws, size_t ws_bytes, hipStream_t s) {
    if (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64) {
        set_error("lanes per block row must be 4, 8, 16, 32 or 64");
        return AMGT_ERR_ARG;
    }
    if (A.nrows == 0) return AMGT_OK;
    if (A.nblk == 0 || alpha == 0.0) return scale_vector(A.nrows, beta, b, y, s);
    const int capb = 256 * balanced_bpt(lanes);
    const int64_t nt = (A.nblk + capb - 1) / capb;
    const BalWsLayout L = bal_ws_layout(nt);
    if (L.bytes > ws_bytes) {
        set_error("balanced SpMV workspace: " + std::to_string(L.bytes) + " bytes needed for " +
                  std::to_string(nt) + " tiles, " + std::to_string(ws_bytes) + " available");
        return AMGT_ERR_UNSUPPORTED;
    }
    char *base = static_cast<char *>(ws);
    int32_t *lo = reinterpret_cast<int32_t *>(base);
    int32_t *head = reinterpret_cast<int32_t *>(base + L.head);
    int32_t *vb = reinterpret_cast<int32_t *>(base + L.vb);
    double *carry = reinterpret_cast<double *>(base + L.carry);
    hipLaunchKernelGGL(k_tile_rows, dim3(unsigned((nt + 1 + 255) / 256)), dim3(256), 0, s,
                       int32_t(nt), capb, A.brow, A.blk_ptr, lo, head);
    hipLaunchKernelGGL(k_tile_popc, dim3(unsigned(nt)), dim3(256), 0, s, int32_t(A.nblk), capb,
                       A.blk_map, vb);
    hipLaunchKernelGGL(k_scan_excl_1wg, dim3(1), dim3(1024), 0, s, int32_t(nt), vb);
    AMGT_HIP(hipGetLastError());
    const BalTables tb{int32_t(nt), capb, lo, head, vb, carry};
    const EpAxpby ep{alpha, beta, b};
    switch (balanced_bpt(lanes)) {
        case 1: return launch_spmv_bal_g<1, double, double>(A, tb, A.val, x, ep, y, s);
        case 4: return launch_spmv_bal_g<4, double, double>(A, tb, A.val, x, ep, y, s);
        default: return launch_spmv_bal_g<2, double, double>(A, tb, A.val, x, ep, y, s);
    }
}

int mbsr_ensure_lowp(Mbsr *A, int *prec, hipStream_t s) {
    if (*prec == AMGT_DOUBLE || A->nnz == 0 || A->pattern_only) return AMGT_OK;
    const int64_t n = A->nnz;
    const unsigned grid = unsigned((n + 255) / 256);
    if (*prec == AMGT_HALF && A->half_fits < 0) {
        // max |value| in two stages (per-workgroup partials, one workgroup)
        const int parts = int(std::min<int64_t>((n + 255) / 256, 1024));
        double *part = nullptr, h_max = 0.0;
        AMGT_TRY(tmp_alloc_t(&part, size_t(parts) + 1, s));
        hipLaunchKernelGGL(k_absmax_part, dim3(unsigned(parts)), dim3(256), 0, s, n, A->val, part);
        hipLaunchKernelGGL(k_absmax_part, dim3(1), dim3(256), 0, s, int64_t(parts), part,
                           part + parts);
        hipError_t e = hipGetLastError();
        int st = AMGT_OK;
        if (e == hipSuccess) {
            const Readback rb[1] = {{&h_max, part + parts, sizeof(double)}};
            st = readback(rb, 1, s);
        }
        tmp_free(part, s);
        if (e != hipSuccess) {
            set_error(std::string("FP16 range check: ") + hipGetErrorString(e));
            return AMGT_ERR_HIP;
        }
        AMGT_TRY(st);
        A->half_fits = std::isfinite(h_max) ? 1 : 0;
        // 2^e with max |a| * 2^e in [2^14, 2^15): ilogb gives max |a| = f 2^k, f in [1, 2)
        const int ex = h_max > 0.0 && A->half_fits
                           ? std::max(-64, std::min(64, 14 - std::ilogb(h_max)))
                           : 0;
        A->half_scale = std::ldexp(1.0, ex);
    }
    if (*prec == AMGT_HALF && !A->half_fits) *prec = AMGT_FLOAT;
    const int precision = *prec;
    if (precision == AMGT_FLOAT && !A->val32) {
        A->val32 = static_cast<float *>(dev_alloc(sizeof(float) * size_t(n), s));
        if (!A->val32) {
            set_error("FP32 value copy allocation failed");
            return AMGT_ERR_ALLOC;
        }
        hipLaunchKernelGGL(k_cast_f32, dim3(grid), dim3(256), 0, s, n, A->val, A->val32);
        AMGT_HIP(hipGetLastError());
    }
    if (precision == AMGT_HALF && !A->val16) {
        A->val16 = static_cast<__half *>(dev_alloc(sizeof(__half) * size_t(n), s));
        if (!A->val16) {
            set_error("FP16 value copy allocation failed");
            return AMGT_ERR_ALLOC;
        }
        hipLaunchKernelGGL(k_cast_f16, dim3(grid), dim3(256), 0, s, n, A->val, A->half_scale,
                           A->val16);
        AMGT_HIP(hipGetLastError());
    }
    return AMGT_OK;
}

int spmv(const Mbsr &A, int precision, int lanes, double alpha, const double *x, double beta,
         const double *b, double *y, hipStream_t s) {
    if (A.nrows == 0) return AMGT_OK;
    if (A.nblk == 0 || alpha == 0.0) {
        hipLaunchKernelGGL(k_scale, dim3((A.nrows + 255) / 256), dim3(256), 0, s, A.nrows, beta,
                           b, y);
        AMGT_HIP(hipGetLastError());
        return AMGT_OK;
    }
    return launch_spmv_prec(A, precision, lanes, x, EpAxpby{alpha, beta, b}, y, s);
}

int residual_restrict(const Mbsr &A, int precision, const RestrictTables &t, const double *f,
                      const double *u, double *r, double *fc, hipStream_t s) {
    if (A.nrows == 0) return AMGT_OK;
    const EpResidRestrict ep{f, t.tile_ptr, t.tile_c, t.nspan, t.cnt, t.ri, t.rj, t.rv, r, fc};
    auto go = [&](auto vt_tag, auto at_tag, const auto *val, auto epx) -> int {
        using VT = decltype(vt_tag);
        using AT = decltype(at_tag);
        switch (t.rpb) {
            case 64: return launch_spmv_blk_g<64, 2, VT, AT>(A, val, u, epx, r, s);
            case 128: return launch_spmv_blk_g<128, 4, VT, AT>(A, val, u, epx, r, s);
            case 32: return launch_spmv_blk_g<32, 2, VT, AT>(A, val, u, epx, r, s);
            case 16: return launch_spmv_blk_g<16, 2, VT, AT>(A, val, u, epx, r, s);
            default: return launch_spmv_blk_g<8, 2, VT, AT>(A, val, u, epx, r, s);
        }
    };
    switch (precision) {
        case AMGT_DOUBLE: return go(double(), double(), A.val, ep);
        case AMGT_FLOAT: return go(float(), float(), A.val32, ep);
        case AMGT_HALF:
            return go(__half(), float(), A.val16, EpScaled<EpResidRestrict>{ep, 1.0 / A.half_scale});
        default:
            set_error("unknown precision");
            return AMGT_ERR_ARG;
    }
}

int blk_tile_rows(int lanes) {
    switch (lanes) {
        case 4: return 64;
        case 8: return 128;
        case 16: return 32;
        case 32: return 16;
        default: return 8;
    }
}

int relax_jacobi(const Mbsr &A, int precision, int lanes, double omega, const double *dinv,
                 const double *f, const double *u, double *u_new, hipStream_t s) {
    if (A.nrows == 0) return AMGT_OK;
    return launch_spmv_prec(A, precision, lanes, u, EpJacobi{omega, dinv, f, u}, u_new, s);
}

int spmm(const Mbsr &A, int precision, int lanes, int k, double alpha, const double *x,
         int64_t ldx, double beta, const double *b, int64_t ldb, double *y, int64_t ldy,
         hipStream_t s) {
    if (A.nrows == 0) return AMGT_OK;
    if (k == 1) return spmv(A, precision, lanes, alpha, x, beta, b, y, s);
    if (A.nblk == 0 || alpha == 0.0) {
        for (int v = 0; v < k; ++v) {
            hipLaunchKernelGGL(k_scale, dim3((A.nrows + 255) / 256), dim3(256), 0, s, A.nrows,
                               beta, b ? b + int64_t(v) * ldb : nullptr, y + int64_t(v) * ldy);
        }
        AMGT_HIP(hipGetLastError());
        return AMGT_OK;
    }
    switch (precision) {
        case AMGT_DOUBLE:
            return launch_spmm_t<double, double>(A, A.val, lanes, k, alpha, x, ldx, beta, b, ldb,
                                                 y, ldy, s);
        case AMGT_FLOAT:
            return launch_spmm_t<float, float>(A, A.val32, lanes, k, alpha, x, ldx, beta, b, ldb,
                                               y, ldy, s);
        case AMGT_HALF:
            return launch_spmm_t<__half, float>(A, A.val16, lanes, k, alpha / A.half_scale, x,
                                                ldx, beta, b, ldb, y, ldy, s);
        default:
            set_error("unknown precision");
            return AMGT_ERR_ARG;
    }
}

int spmm_interleaved(const Mbsr &A, int precision, int k, double alpha, const double *x,
                     double beta, const double *b, double *y, hipStream_t s) {
    if (A.nrows == 0 || k == 0) return AMGT_OK;
    switch (precision) {
        case AMGT_DOUBLE:
            return launch_spmm_il<double>(A, A.val, k, alpha, x, k, beta, b, k, y, k, s);
        case AMGT_FLOAT:
            return launch_spmm_il<float>(A, A.val32, k, alpha, x, k, beta, b, k, y, k, s);
        case AMGT_HALF:
            return launch_spmm_il<__half>(A, A.val16, k, alpha / A.half_scale, x, k, beta, b, k,
                                          y, k, s);
        default:
            set_error("unknown precision");
            return AMGT_ERR_ARG;
    }
}

int scale_vector(int32_t n, double beta, const double *b, double *y, hipStream_t s) {
    if (n == 0) return AMGT_OK;
    hipLaunchKernelGGL(k_scale, dim3((n + 255) / 256), dim3(256), 0, s, n, beta, b, y);
    AMGT_HIP(hipGetLastError());
    return AMGT_OK;
}

}  // namespace amgt
