// BoomerAMG setup on the device, one process: strength of connection, PMIS
// coarsening and extended+i interpolation (with max-elements truncation) on
// CSR operators.
//
// Counterparts: hypre_BoomerAMGCreateSDevice (par_strength_device.c:34),
// hypre_BoomerAMGCoarsenPMISDevice (par_coarsen_device.c:30),
// hypre_BoomerAMGBuildExtPIInterpDevice (par_lr_interp_device.c:991) and
// hypre_BoomerAMGInterpTruncationDevice (par_interp_trunc_device.c:731).  The
// results follow HYPRE's host algorithms on one process (par_strength.c:75,
// par_coarsen.c:2101, par_lr_interp.c:1025, par_csr_matrix.c:2866 with
// qsort.c:395), which the reference's own regression outputs pin, so the
// device hierarchy equals the CPU one entry for entry:
//  * S: strong columns in row order, no diagonal;
//  * PMIS measures |S^T row| + hypre_Rand() from seed 2747 (the Park-Miller
//    stream jumped to each row's draw), the independent set of each pass taken
//    exactly as the host loop leaves it;
//  * P rows in the host's discovery order, weights accumulated in the host's
//    order with contraction off (no fused multiply-add), truncated rows
//    ordered by the host's quicksort.
// The diagonal of a row is found by column (HYPRE keeps it first; exports of
// the mBSR handles have ascending columns).
#include "internal.hpp"

#include <vector>

namespace amgt {

namespace {

constexpr int kThreads = 256;
// levels with few rows (<= kWaveMaxRows) that hold at least kWaveMinRowLen
// entries per row on average use the wave kernel of extended+i for their F
// rows (one wave per row holds ~38 KB of LDS: worth it only when the thread
// kernel would leave the chip idle behind a few long serial rows)
constexpr int kWaveMinRowLen = 12;
constexpr int kWaveMaxRows = 1 << 18;       // the kernel of rows up to 256 entries
constexpr int kWaveMaxRowsSmall = 1 << 21;  // the kernel of rows up to 64 entries

inline unsigned grid_of(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

// position of a_ii in row i, or -1: the first entry (HYPRE's diagonal-first
// rows), else a binary search (rows with ascending columns, the exports of
// mBSR handles), else a scan.  A search can only return an entry of column i,
// and a CSR row holds at most one, so any row order gives the same answer.
__device__ __forceinline__ int64_t diag_pos(const int32_t *__restrict__ I,
                                            const int32_t *__restrict__ J, int i) {
    const int64_t b = I[i], e = I[i + 1];
    if (b == e) return -1;
    if (J[b] == i) return b;
    int64_t lo = b, hi = e - 1;
    while (lo <= hi) {
        const int64_t mid = (lo + hi) >> 1;
        const int c = J[mid];
        if (c == i) return mid;
        if (c < i) lo = mid + 1; else hi = mid - 1;
    }
    for (int64_t p = b; p < e; ++p)
        if (J[p] == i) return p;
    return -1;
}

// ---- strength (par_strength.c:75-528, one function) ---------------------------
#pragma clang fp contract(off)
template <bool kFill>
__global__ __launch_bounds__(kThreads) void k_strength(int32_t n, const int32_t *__restrict__ I,
                                                      const int32_t *__restrict__ J,
                                                      const double *__restrict__ A, double theta,
                                                      double max_row_sum, int32_t *__restrict__ cnt,
                                                      const int32_t *__restrict__ SI,
                                                      int32_t *__restrict__ SJ) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int64_t d = diag_pos(I, J, i);
    const double diag = d >= 0 ? A[d] : 0.0;
    double row_scale = 0.0, row_sum = diag;
    for (int64_t p = I[i]; p < I[i + 1]; ++p) {
        if (p == d) continue;
        const double a = A[p];
        if (diag < 0) row_scale = a > row_scale ? a : row_scale;
        else row_scale = a < row_scale ? a : row_scale;
        row_sum += a;
    }
    int c = 0;
    int64_t out = kFill ? SI[i] : 0;
    // a row too dominant by the row-sum test keeps only weak dependencies
    if (!(fabs(row_sum) > fabs(diag) * max_row_sum && max_row_sum < 1.0)) {
        const double thr = theta * row_scale;
        for (int64_t p = I[i]; p < I[i + 1]; ++p) {
            if (p == d) continue;
            const double a = A[p];
            const bool weak = diag < 0 ? (a <= thr) : (a >= thr);
            if (!weak) {
                if (kFill) SJ[out++] = J[p];
                ++c;
            }
        }
    }
    if (!kFill) cnt[i] = c;
}

// ---- PMIS (par_coarsen.c:2101-2813 on one process) ------------------------------
__global__ __launch_bounds__(kThreads) void k_col_count(int32_t n, const int32_t *__restrict__ SI,
                                                       const int32_t *__restrict__ SJ,
                                                       int32_t *__restrict__ cnt) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    for (int p = SI[i]; p < SI[i + 1]; ++p) atomicAdd(&cnt[SJ[p]], 1);
}

// x_{k} = a^k x_0 mod m of the Park-Miller generator (random.c:43-101); all
// products stay below 2^62
__device__ __forceinline__ uint64_t pm_jump(uint64_t x0, uint64_t k) {
    constexpr uint64_t kA = 16807u, kM = 2147483647u;
    uint64_t r = 1u, b = kA;
    while (k) {
        if (k & 1u) r = (r * b) % kM;
        b = (b * b) % kM;
        k >>= 1;
    }
    return (r * x0) % kM;
}

// measures and the initial graph (par_coarsen.c:2230-2400, IndepSetInit with
// seed 2747); a row with no strong dependence is marked -3 and left out
__global__ __launch_bounds__(kThreads) void k_pmis_init(int32_t n, const int32_t *__restrict__ SI,
                                                       const int32_t *__restrict__ cnt,
                                                       double *__restrict__ measure,
                                                       int32_t *__restrict__ cf,
                                                       int32_t *__restrict__ any) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const double r = double(pm_jump(2747u, uint64_t(i) + 1u)) / 2147483647.0;
    const double m = double(cnt[i]) + r;
    if (SI[i + 1] == SI[i]) {
        cf[i] = -3;
        measure[i] = 0.0;
    } else {
        cf[i] = 0;
        measure[i] = m;
        *any = 1;
    }
}

// CF_init 1 (HMIS, par_coarsen.c:2353-2386): cf holds the markers of the Ruge
// first pass.  C points stay C and join the graph; F points are undecided
// again; a Z point (it influences nobody) joins the graph as undecided if its
// measure reaches 1 or it depends on some point, else it is F; points without
// strong dependence stay out.  Graph membership is a positive measure, so the
// points left out get measure 0 (the host leaves a Z point's measure below 1,
// which no comparison with > 1 can see).
__global__ __launch_bounds__(kThreads) void k_pmis_init_hmis(int32_t n,
                                                            const int32_t *__restrict__ SI,
                                                            const int32_t *__restrict__ cnt,
                                                            double *__restrict__ measure,
                                                            int32_t *__restrict__ cf,
                                                            int32_t *__restrict__ any) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const double r = double(pm_jump(2747u, uint64_t(i) + 1u)) / 2147483647.0;
    const double m = double(cnt[i]) + r;
    int c = cf[i];
    if (c == -3) {
        measure[i] = 0.0;
        return;
    }
    if (c == -1) c = 0;
    if (c == -2) {
        if (!(m >= 1.0) && SI[i + 1] == SI[i]) {
            cf[i] = -1;
            measure[i] = 0.0;
            return;
        }
        c = 0;
    }
    cf[i] = c;
    measure[i] = m;
    *any = 1;
}

// independent set, first half (par_coarsen.c:2510-2556): a node of the graph
// with measure > 1 loses against any S neighbour (either direction) with a
// larger measure > 1; the host's clearing stores are idempotent, so plain
// stores of 1 reproduce them
__global__ __launch_bounds__(kThreads) void k_pmis_clear(int32_t n, const int32_t *__restrict__ SI,
                                                        const int32_t *__restrict__ SJ,
                                                        const double *__restrict__ measure,
                                                        int32_t *__restrict__ lost) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const double mi = measure[i];
    if (!(mi > 1)) return;
    for (int p = SI[i]; p < SI[i + 1]; ++p) {
        const int j = SJ[p];
        const double mj = measure[j];
        if (mj > 1) {
            if (mi > mj) lost[j] = 1;
            else if (mj > mi) lost[i] = 1;
        }
    }
}

// graph nodes are the ones with a positive measure (decided and isolated nodes
// have measure 0; a random draw is > 0)
__global__ __launch_bounds__(kThreads) void k_pmis_set(int32_t n, const double *__restrict__ measure,
                                                      const int32_t *__restrict__ lost,
                                                      int32_t *__restrict__ cf) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    if (measure[i] > 1 && !lost[i]) cf[i] = 1;
}

// C and F points (par_coarsen.c:2583-2666): a graph node of measure < 1 is F;
// a node outside the set that depends on a C point is F
__global__ __launch_bounds__(kThreads) void k_pmis_cf(int32_t n, const int32_t *__restrict__ SI,
                                                     const int32_t *__restrict__ SJ,
                                                     const double *__restrict__ measure,
                                                     int32_t *__restrict__ cf) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const double mi = measure[i];
    if (!(mi > 0)) return;
    if (mi < 1) cf[i] = -1;
    if (cf[i] > 0) return;
    for (int p = SI[i]; p < SI[i + 1]; ++p)
        if (cf[SJ[p]] > 0) {
            cf[i] = -1;
            break;
        }
}

// subgraph update (par_coarsen.c:2690-2760)
__global__ __launch_bounds__(kThreads) void k_pmis_update(int32_t n, double *__restrict__ measure,
                                                         const int32_t *__restrict__ cf,
                                                         int32_t *__restrict__ any) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    if (!(measure[i] > 0)) return;
    if (cf[i] != 0) measure[i] = 0.0;
    else *any = 1;
}

__global__ __launch_bounds__(kThreads) void k_is_coarse(int32_t n, const int32_t *__restrict__ cf,
                                                       int32_t *__restrict__ flag) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) flag[i] = cf[i] >= 0 ? 1 : 0;
}

// ---- extended+i interpolation (par_lr_interp.c:1025-1928, one process) -----------
// upper bound of a row's entries: 1 for C, 0 for isolated, else the strong C
// neighbours plus the strong dependencies of the strong F neighbours
__global__ __launch_bounds__(kThreads) void k_extpi_ub(int32_t n, const int32_t *__restrict__ SI,
                                                      const int32_t *__restrict__ SJ,
                                                      const int32_t *__restrict__ cf,
                                                      int64_t *__restrict__ ub) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    int64_t u = 0;
    if (cf[i] >= 0) {
        u = 1;
    } else if (cf[i] != -3) {
        for (int p = SI[i]; p < SI[i + 1]; ++p) {
            const int i1 = SJ[p];
            if (cf[i1] >= 0) u += 1;
            else if (cf[i1] != -3) u += SI[i1 + 1] - SI[i1];
        }
    }
    ub[i] = u;
}

// Per-row open-addressing table (linear probing, power-of-two size, at least
// twice the entries it can receive): fine column -> position in the row (>= 0)
// or kStrongF for a strong F neighbour; keys start at -1.  It stands for the
// host's P_marker array, so every membership test is O(1).
constexpr int32_t kStrongF = -2;

__device__ __forceinline__ uint32_t h_slot(int k, uint32_t mask) {
    return (uint32_t(k) * 2654435761u) & mask;
}

// slots are (key, value) pairs: one 8-byte load per probe
__device__ __forceinline__ int h_find(const int2 *__restrict__ kv, uint32_t mask, int k) {
    for (uint32_t h = h_slot(k, mask);; h = (h + 1) & mask) {
        const int2 e = kv[h];
        if (e.x == k) return e.y;
        if (e.x == -1) return -1;
    }
}

__device__ __forceinline__ void h_put(int2 *__restrict__ kv, uint32_t mask, int k, int v) {
    for (uint32_t h = h_slot(k, mask);; h = (h + 1) & mask) {
        const int32_t kk = kv[h].x;
        if (kk == -1 || kk == k) {
            kv[h] = make_int2(k, v);
            return;
        }
    }
}

// table size of a row: 0 for C and isolated rows, else the power of two at or
// above 2 * (upper bound of its entries + its strong dependencies)
__global__ __launch_bounds__(kThreads) void k_extpi_hsize(int32_t n, const int32_t *__restrict__ SI,
                                                         const int32_t *__restrict__ cf,
                                                         const int64_t *__restrict__ ub,
                                                         int64_t *__restrict__ hs) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    int64_t z = 0;
    if (cf[i] < 0 && cf[i] != -3) {
        const int64_t need = 2 * (ub[i] + (SI[i + 1] - SI[i]));
        z = 1;
        while (z < need) z <<= 1;
    }
    hs[i] = z;
}

// qsort.c:395-417 (descending |w|) with an explicit stack in place of the
// recursion.  Each partition step is the host's; the two subranges it leaves
// are disjoint and sorted independently, so the order in which they are
// sorted does not change the result: the smaller one goes first, which keeps
// the stack within log2(length) entries.
__device__ void qsort2_abs(int32_t *v, double *w, int left, int right) {
    int stack[2 * 40];
    int top = 0;
    stack[top++] = left;
    stack[top++] = right;
    while (top > 0) {
        const int r = stack[--top];
        const int l = stack[--top];
        if (l >= r) continue;
        const int mid = (l + r) / 2;
        {
            const int32_t tv = v[l];
            v[l] = v[mid];
            v[mid] = tv;
            const double tw = w[l];
            w[l] = w[mid];
            w[mid] = tw;
        }
        int last = l;
        for (int i = l + 1; i <= r; ++i) {
            if (fabs(w[i]) > fabs(w[l])) {
                ++last;
                const int32_t tv = v[last];
                v[last] = v[i];
                v[i] = tv;
                const double tw = w[last];
                w[last] = w[i];
                w[i] = tw;
            }
        }
        {
            const int32_t tv = v[l];
            v[l] = v[last];
            v[last] = tv;
            const double tw = w[l];
            w[l] = w[last];
            w[last] = tw;
        }
        // larger subrange pushed first, smaller one popped next
        if (last - 1 - l > r - (last + 1)) {
            stack[top++] = l;
            stack[top++] = last - 1;
            stack[top++] = last + 1;
            stack[top++] = r;
        } else {
            stack[top++] = last + 1;
            stack[top++] = r;
            stack[top++] = l;
            stack[top++] = last - 1;
        }
    }
}

// One thread per row: the row's columns (fine indices) and weights in the
// host's discovery order into scratch at off[i]; truncated in place; cnt[i] =
// the row's final length.
__global__ __launch_bounds__(kThreads) void k_extpi_row(
    int32_t n, const int32_t *__restrict__ I, const int32_t *__restrict__ J,
    const double *__restrict__ A, const int32_t *__restrict__ SI, const int32_t *__restrict__ SJ,
    const int32_t *__restrict__ cf, const int64_t *__restrict__ off, int max_elmts,
    const int64_t *__restrict__ hoff, int2 *__restrict__ hkv,
    const int32_t *__restrict__ skip, int32_t *__restrict__ sj, double *__restrict__ sw,
    int32_t *__restrict__ cnt) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    if (skip && skip[i]) return;  // a row of the wave kernel
    int32_t *cols = sj + off[i];
    double *w = sw + off[i];
    int2 *kv = hkv + hoff[i];
    const uint32_t mask = uint32_t(hoff[i + 1] - hoff[i]) - 1u;
    if (cf[i] >= 0) {  // C point: identity (par_lr_interp.c:1488-1495)
        cols[0] = i;
        w[0] = 1.0;
        cnt[i] = 1;
        return;
    }
    if (cf[i] == -3) {
        cnt[i] = 0;
        return;
    }
    // structure (:1501-1560)
    int c = 0;
    for (int p = SI[i]; p < SI[i + 1]; ++p) {
        const int i1 = SJ[p];
        if (cf[i1] >= 0) {
            if (h_find(kv, mask, i1) < 0) {
                h_put(kv, mask, i1, c);
                cols[c++] = i1;
            }
        } else if (cf[i1] != -3) {
            h_put(kv, mask, i1, kStrongF);
            const int q1 = SI[i1 + 1];
            for (int q0 = SI[i1]; q0 < q1; q0 += 8) {  // eight strong neighbours at a time
                int k8[8], f8[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) k8[u] = q0 + u < q1 ? SJ[q0 + u] : -1;
#pragma unroll
                for (int u = 0; u < 8; ++u) f8[u] = k8[u] >= 0 ? cf[k8[u]] : -1;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int k1 = k8[u];
                    if (f8[u] >= 0 && h_find(kv, mask, k1) < 0) {
                        h_put(kv, mask, k1, c);
                        cols[c++] = k1;
                    }
                }
            }
        }
    }
    for (int q = 0; q < c; ++q) w[q] = 0.0;
    // weights (:1630-1708)
    const int64_t di = diag_pos(I, J, i);
    double diagonal = di >= 0 ? A[di] : 0.0;
    for (int64_t jj = I[i]; jj < I[i + 1]; ++jj) {
        if (jj == di) continue;
        const int i1 = J[jj];
        const int pos = h_find(kv, mask, i1);
        if (pos >= 0) {
            w[pos] += A[jj];
            continue;
        }
        if (pos == kStrongF) {
            const int64_t d1 = diag_pos(I, J, i1);
            const int sgn = (d1 >= 0 && A[d1] < 0) ? -1 : 1;
            // both passes over row i1 in the host's order, its entries loaded
            // eight at a time so that the loads overlap
            const int64_t b1 = I[i1], e1 = I[i1 + 1];
            double sum = 0.0;
            for (int64_t q0 = b1; q0 < e1; q0 += 8) {
                int j8[8];
                double a8[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    j8[u] = q0 + u < e1 ? J[q0 + u] : -1;
                    a8[u] = q0 + u < e1 ? A[q0 + u] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (q0 + u >= e1 || q0 + u == d1) continue;
                    const int i2 = j8[u];
                    if ((i2 == i || h_find(kv, mask, i2) >= 0) && (sgn * a8[u]) < 0)
                        sum += a8[u];
                }
            }
            if (sum != 0) {
                const double distribute = A[jj] / sum;
                for (int64_t q0 = b1; q0 < e1; q0 += 8) {
                    int j8[8];
                    double a8[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        j8[u] = q0 + u < e1 ? J[q0 + u] : -1;
                        a8[u] = q0 + u < e1 ? A[q0 + u] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (q0 + u >= e1 || q0 + u == d1) continue;
                        const int i2 = j8[u];
                        if ((sgn * a8[u]) < 0) {
                            const int p2 = h_find(kv, mask, i2);
                            if (p2 >= 0) w[p2] += distribute * a8[u];
                            if (i2 == i) diagonal += distribute * a8[u];
                        }
                    }
                }
            } else {
                diagonal += A[jj];
            }
        } else if (cf[i1] != -3) {
            diagonal += A[jj];
        }
    }
    if (diagonal != 0)
        for (int q = 0; q < c; ++q) w[q] /= -diagonal;
    // truncation (par_csr_matrix.c:3130-3215: tol 0, rescale)
    if (max_elmts > 0 && c > max_elmts) {
        double row_sum = 0.0;
        for (int q = 0; q < c; ++q) row_sum += w[q];
        qsort2_abs(cols, w, 0, c - 1);
        double scale = 0.0;
        for (int q = 0; q < max_elmts; ++q) scale += w[q];
        if (scale != 0. && scale != row_sum) {
            scale = row_sum / scale;
            for (int q = 0; q < max_elmts; ++q) w[q] *= scale;
        }
        c = max_elmts;
    }
    cnt[i] = c;
}

// ---- extended+i, one wave per row (long rows of coarse levels) ------------------
// The same results as k_extpi_row for rows whose columns are ascending (every
// level exported from an mBSR handle): the host's order of discovery and of
// accumulation is kept, the independent parts run across the lanes.
//  * discovery: the candidate columns in the host's order (lane per strong
//    dependency, offsets by a scan), first occurrences found with an LDS
//    table (atomicMin of the candidate index), positions by a prefix count;
//  * per entry jj of row i (lane per entry): its kind, and for a strong F
//    neighbour i1 the sum over row i1 in row order, the distribution factor
//    and its term for the diagonal;
//  * the weights, entry by entry jj in row order: a C neighbour's a_ij to its
//    column (one lane), a strong F neighbour's distribution over the entries
//    of row i1 that are P columns (a lane per entry of row i1, its column
//    looked up in the LDS table), so every weight receives its terms in the
//    host's order;
//  * the diagonal, the truncation sort and rescale on lane 0, in order.
// capacities: WA entries of row i (A and S), WT distinct P columns kept in
// LDS, WH table slots: at most WT + 64 candidate columns (the insertions stop
// once WT is passed) plus WA strong F neighbours ever enter the table, which
// stays at most about half full.  Two sets: rows of up to 64 entries (about
// 15 KB of LDS, many rows per CU) and of up to 256.
template <int WA, int WT, int WH>
struct WaveRow {
    static_assert(2 * (WT + 64 + WA) <= WH + 128 && (WH & (WH - 1)) == 0, "table size");
    int32_t acol[WA];
    double aval[WA];
    int32_t scol[WA];
    int32_t hkey[WH];
    int32_t hval[WH];
    int32_t pcol[WT];
    double pw[WT];
    double jdist[WA];
    double jdiag[WA];
    int32_t jinfo[WA];  // >= 0: P position; -1 skip/diag; -2 strong F (distribute);
                         // -3 add a_ij to the diagonal
    int32_t jsgn[WA];
    int32_t jb1[WA];    // strong F neighbour: its row's range in J / A
    int32_t je1[WA];
    int32_t sc[3];
};

template <int WH>
__device__ __forceinline__ int lds_find(const int32_t *hkey, const int32_t *hval, int k) {
    for (uint32_t h = h_slot(k, WH - 1);; h = (h + 1) & (WH - 1)) {
        const int32_t kk = hkey[h];
        if (kk == k) return hval[h];
        if (kk == -1) return -1;
    }
}

// slot of key k (inserted if absent) in the wave's table
template <int WH>
__device__ __forceinline__ uint32_t lds_slot(int32_t *hkey, int k) {
    for (uint32_t h = h_slot(k, WH - 1);; h = (h + 1) & (WH - 1)) {
        const int32_t prev = atomicCAS(&hkey[h], -1, k);
        if (prev == -1 || prev == k) return h;
    }
}

// exclusive prefix sum over the wave's 64 lanes; *tot = total
__device__ __forceinline__ int wave_excl(int v, int *tot) {
    const int lane = threadIdx.x & 63;
    int incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    *tot = __shfl(incl, 63);
    return incl - v;
}

template <int WA, int WT, int WH>
__global__ __launch_bounds__(64) void k_extpi_wave(
    int32_t count, const int32_t *__restrict__ rows, const int32_t *__restrict__ I,
    const int32_t *__restrict__ J, const double *__restrict__ A, const int32_t *__restrict__ SI,
    const int32_t *__restrict__ SJ, const int32_t *__restrict__ cf,
    const int64_t *__restrict__ off, int max_elmts, int32_t *__restrict__ sj,
    double *__restrict__ sw, int32_t *__restrict__ cnt, int32_t *__restrict__ use,
    int overflow_code) {
    __shared__ WaveRow<WA, WT, WH> s;
    if (blockIdx.x >= unsigned(count)) return;
    const int lane = threadIdx.x;
    const int i = rows[blockIdx.x];
    int32_t *cand = sj + off[i];  // the row's candidates, then (by the fallback) its entries
    const int a0 = I[i], la = I[i + 1] - a0;
    const int s0 = SI[i], ls = SI[i + 1] - s0;
    for (int t = lane; t < la; t += 64) {
        s.acol[t] = J[a0 + t];
        s.aval[t] = A[a0 + t];
    }
    for (int t = lane; t < ls; t += 64) s.scol[t] = SJ[s0 + t];
    for (int t = lane; t < WH; t += 64) {
        s.hkey[t] = -1;
        s.hval[t] = INT_MAX;
    }
    if (lane == 0) s.sc[2] = 0;  // distinct candidate columns
    __syncthreads();
    // candidates in the host's order (par_lr_interp.c:1501-1560)
    int T = 0;
    for (int b0 = 0; b0 < ls; b0 += 64) {
        const int p = b0 + lane;
        int i1 = 0, k = 0;
        int cf1 = 0, q0 = 0, q1 = 0;
        if (p < ls) {
            i1 = s.scol[p];
            cf1 = cf[i1];
            if (cf1 >= 0) {
                k = 1;
            } else if (cf1 != -3) {
                q0 = SI[i1];
                q1 = SI[i1 + 1];
                // the strong neighbours of i1 and their markers, eight at a
                // time so that the loads overlap
                for (int q = q0; q < q1; q += 8) {
                    int c8[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) c8[u] = q + u < q1 ? SJ[q + u] : -1;
#pragma unroll
                    for (int u = 0; u < 8; ++u) k += (c8[u] >= 0 && cf[c8[u]] >= 0) ? 1 : 0;
                }
            }
        }
        int tot;
        int o = T + wave_excl(k, &tot);
        if (p < ls && k > 0) {
            if (cf1 >= 0) {
                cand[o] = i1;
            } else {
                for (int q = q0; q < q1; q += 8) {  // in order, as counted
                    int c8[8], f8[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) c8[u] = q + u < q1 ? SJ[q + u] : -1;
#pragma unroll
                    for (int u = 0; u < 8; ++u) f8[u] = c8[u] >= 0 ? cf[c8[u]] : -1;
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        if (f8[u] >= 0) cand[o++] = c8[u];
                }
            }
        }
        T += tot;
    }
    __threadfence_block();
    __syncthreads();
    // first occurrences; a row with more distinct columns than WT is left to
    // the thread kernel (the insertions stop early enough that the table never
    // fills: at most WT + 64 keys before every lane sees the count)
    for (int t = lane; t < T; t += 64) {
        if (s.sc[2] > WT) break;
        const int col = cand[t];
        uint32_t h = h_slot(col, WH - 1);
        for (;; h = (h + 1) & (WH - 1)) {
            const int32_t prev = atomicCAS(&s.hkey[h], -1, col);
            if (prev == -1) {
                atomicAdd(&s.sc[2], 1);
                break;
            }
            if (prev == col) break;
        }
        atomicMin(&s.hval[h], t);
    }
    __syncthreads();
    if (s.sc[2] > WT) {  // to the next capacity (overflow_code), or the thread kernel (0)
        if (lane == 0) use[i] = overflow_code;
        return;
    }
    int c = 0;
    for (int b0 = 0; b0 < T; b0 += 64) {
        const int t = b0 + lane;
        bool first = false;
        int col = 0;
        if (t < T) {
            col = cand[t];
            first = lds_find<WH>(s.hkey, s.hval, col) == t;
        }
        int tot;
        const int pos = c + wave_excl(first ? 1 : 0, &tot);
        if (first) s.pcol[pos] = col;
        c += tot;
    }
    __syncthreads();
    // table: column -> position, strong F neighbours -> kStrongF
    for (int p = lane; p < c; p += 64) s.hval[lds_slot<WH>(s.hkey, s.pcol[p])] = p;
    for (int p = lane; p < ls; p += 64) {
        const int i1 = s.scol[p];
        if (cf[i1] < 0 && cf[i1] != -3) s.hval[lds_slot<WH>(s.hkey, i1)] = kStrongF;
    }
    __syncthreads();
    // per entry of row i
    for (int jj = lane; jj < la; jj += 64) {
        const int i1 = s.acol[jj];
        int info = -1, sgn = 1;
        double dist = 0.0, dterm = 0.0;
        if (i1 != i) {
            const int v = lds_find<WH>(s.hkey, s.hval, i1);
            if (v >= 0) {
                info = v;
            } else if (v == kStrongF) {
                const int64_t b1 = I[i1], e1 = I[i1 + 1];
                s.jb1[jj] = int32_t(b1);
                s.je1[jj] = int32_t(e1);
                const int64_t d1 = diag_pos(I, J, i1);
                sgn = (d1 >= 0 && A[d1] < 0) ? -1 : 1;
                // the host's sum in the host's order; the row's entries are
                // loaded eight at a time so that their loads overlap, and the
                // entry of column i is noted on the way (the row's columns are
                // distinct, so there is at most one)
                double sum = 0.0, aqi = 0.0;
                int64_t q = b1, qi = -1;
                for (; q + 8 <= e1; q += 8) {
                    int jq[8];
                    double aq[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        jq[u] = J[q + u];
                        aq[u] = A[q + u];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (q + u == d1) continue;
                        const int i2 = jq[u];
                        if (i2 == i) {
                            qi = q + u;
                            aqi = aq[u];
                        }
                        if ((i2 == i || lds_find<WH>(s.hkey, s.hval, i2) >= 0) &&
                            (sgn * aq[u]) < 0)
                            sum += aq[u];
                    }
                }
                for (; q < e1; ++q) {
                    if (q == d1) continue;
                    const int i2 = J[q];
                    const double aq1 = A[q];
                    if (i2 == i) {
                        qi = q;
                        aqi = aq1;
                    }
                    if ((i2 == i || lds_find<WH>(s.hkey, s.hval, i2) >= 0) && (sgn * aq1) < 0)
                        sum += aq1;
                }
                if (sum != 0) {
                    info = -2;
                    dist = s.aval[jj] / sum;
                    if (qi >= 0 && (sgn * aqi) < 0) dterm = dist * aqi;
                } else {
                    info = -3;
                }
            } else if (cf[i1] != -3) {
                info = -3;
            }
        }
        s.jinfo[jj] = info;
        s.jsgn[jj] = sgn;
        s.jdist[jj] = dist;
        s.jdiag[jj] = dterm;
    }
    __syncthreads();
    // weights, entry by entry in row order (the host's order of accumulation
    // for every column): a C neighbour adds a_ij to its column's weight; a
    // strong F neighbour i1 distributes over the entries of row i1 that are P
    // columns, a lane per entry (a row's columns are distinct, so no two lanes
    // add to one weight)
    for (int p = lane; p < c; p += 64) s.pw[p] = 0.0;
    __syncthreads();
    for (int jj = 0; jj < la; ++jj) {
        const int info = s.jinfo[jj];  // wave-uniform
        if (info >= 0) {
            if (lane == 0) s.pw[info] += s.aval[jj];
        } else if (info == -2) {
            const double dj = s.jdist[jj];
            const int sg = s.jsgn[jj];
            for (int q = s.jb1[jj] + lane; q < s.je1[jj]; q += 64) {
                const double aq = A[q];
                const int pos = lds_find<WH>(s.hkey, s.hval, J[q]);  // -1 / kStrongF: no P column
                if (pos >= 0 && (sg * aq) < 0) s.pw[pos] += dj * aq;
            }
        }
        __syncthreads();  // the next entry's terms may go to the same weights
    }
    if (lane == 0) {
        double diagonal = 0.0;
        for (int jj = 0; jj < la; ++jj)
            if (s.acol[jj] == i) {
                diagonal = s.aval[jj];
                break;
            }
        for (int jj = 0; jj < la; ++jj) {
            const int info = s.jinfo[jj];
            if (info == -2) {
                diagonal += s.jdiag[jj];
            } else if (info == -3) {
                diagonal += s.aval[jj];
            }
        }
        s.sc[0] = 0;
        if (diagonal != 0)
            for (int q = 0; q < c; ++q) s.pw[q] /= -diagonal;
        int cc = c;
        if (max_elmts > 0 && c > max_elmts) {
            double row_sum = 0.0;
            for (int q = 0; q < c; ++q) row_sum += s.pw[q];
            qsort2_abs(s.pcol, s.pw, 0, c - 1);
            double scale = 0.0;
            for (int q = 0; q < max_elmts; ++q) scale += s.pw[q];
            if (scale != 0. && scale != row_sum) {
                scale = row_sum / scale;
                for (int q = 0; q < max_elmts; ++q) s.pw[q] *= scale;
            }
            cc = max_elmts;
        }
        s.sc[1] = cc;
        cnt[i] = cc;
    }
    __syncthreads();
    const int cc = s.sc[1];
    for (int q = lane; q < cc; q += 64) {
        sj[off[i] + q] = s.pcol[q];
        sw[off[i] + q] = s.pw[q];
    }
}
#pragma clang fp contract(fast)

// rows the wave kernels cannot search: a row must be ascending, or hold its
// diagonal first and the rest ascending (HYPRE's layout); diag_pos relies on
// exactly that
__global__ __launch_bounds__(kThreads) void k_rows_unsorted(int32_t n, const int32_t *__restrict__ I,
                                                           const int32_t *__restrict__ J,
                                                           int32_t *__restrict__ flag) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    int64_t p = I[i];
    const int64_t e = I[i + 1];
    if (p < e && J[p] == i) ++p;
    int prev = -1;
    for (; p < e; ++p) {
        const int c = J[p];
        if (c <= prev) {
            *flag = 1;
            return;
        }
        prev = c;
    }
}

// rows for the wave kernels (0 = the thread kernel): kTier64 = rows of up to 64
// entries, kTier128 / kTier256 = up to 128 / 256 (when allowed).  Each tier
// first runs with a small table of P columns; a row with more distinct columns
// moves on (use[i] = the next code) to the tier's large table, and from there
// to the thread kernel.
constexpr int kTier64 = 1, kTier64Big = 2, kTier128 = 3, kTier128Big = 4, kTier256 = 5,
              kTier256Big = 6, kTier512 = 7;
__global__ __launch_bounds__(kThreads) void k_wave_rows(int32_t n, const int32_t *__restrict__ I,
                                                       const int32_t *__restrict__ SI,
                                                       const int32_t *__restrict__ cf,
                                                       int allow_big, int32_t *__restrict__ use) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int la = I[i + 1] - I[i], ls = SI[i + 1] - SI[i];
    int u = 0;
    if (cf[i] < 0 && cf[i] != -3) {
        if (la <= 64 && ls <= 64) u = kTier64;
        else if (allow_big && la <= 128 && ls <= 128) u = kTier128;
        else if (allow_big && la <= 256 && ls <= 256) u = kTier256;
        else if (allow_big && la <= 512 && ls <= 512) u = kTier512;
    }
    use[i] = u;
}

__global__ __launch_bounds__(kThreads) void k_flag_eq(int32_t n, const int32_t *__restrict__ use,
                                                     int which, int32_t *__restrict__ flag) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) flag[i] = use[i] == which ? 1 : 0;
}

__global__ __launch_bounds__(kThreads) void k_list_rows(int32_t n, const int32_t *__restrict__ flag,
                                                       const int32_t *__restrict__ pos,
                                                       int32_t *__restrict__ list) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n && flag[i]) list[pos[i]] = i;
}

__global__ __launch_bounds__(kThreads) void k_extpi_out(
    int32_t n, const int64_t *__restrict__ off, const int32_t *__restrict__ PI,
    const int32_t *__restrict__ sj, const double *__restrict__ sw, const int32_t *__restrict__ f2c,
    int32_t *__restrict__ PJ, double *__restrict__ PA) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int64_t s = off[i];
    const int len = PI[i + 1] - PI[i];
    for (int q = 0; q < len; ++q) {
        PJ[PI[i] + q] = f2c[sj[s + q]];
        PA[PI[i] + q] = sw[s + q];
    }
}

struct Tmp {
    hipStream_t s;
    std::vector<void *> ptrs;
    explicit Tmp(hipStream_t st) : s(st) {}
    ~Tmp() {
        for (void *p : ptrs) tmp_free(p, s);
    }
    template <typename T>
    int get(T **p, size_t n) {
        const int st = tmp_alloc_t(p, n, s);
        if (st == AMGT_OK) ptrs.push_back(*p);
        return st;
    }
};

}  // namespace

}  // namespace amgt

using namespace amgt;

extern "C" {

int amgt_amg_strength(int32_t n, const int32_t *A_i, const int32_t *A_j, const double *A_data,
                      double theta, double max_row_sum, int64_t *nnzS, int32_t **S_i,
                      int32_t **S_j, void *stream) {
    AMGT_CHECK_ARG(nnzS && S_i && S_j, "NULL output argument");
    AMGT_CHECK_ARG(n >= 0, "negative size");
    // column and value arrays may be NULL when the matrix has no entries
    AMGT_CHECK_ARG(n == 0 || A_i, "NULL row pointer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    *nnzS = 0;
    *S_i = nullptr;
    *S_j = nullptr;
    Tmp tmp(s);
    int32_t *cnt = nullptr;
    AMGT_TRY(tmp.get(&cnt, size_t(n) + 1));
    int32_t *si = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(n) + 1), s));
    if (!si) {
        set_error("allocation of S row pointer failed");
        return AMGT_ERR_ALLOC;
    }
    if (n > 0) {
        hipLaunchKernelGGL((k_strength<false>), dim3(grid_of(n)), dim3(kThreads), 0, s, n, A_i, A_j,
                           A_data, theta, max_row_sum, cnt, nullptr, nullptr);
        AMGT_HIP(hipGetLastError());
    }
    int st = exclusive_scan_i32(cnt, si, n, s);
    if (st != AMGT_OK) {
        dev_free(si, s);
        return st;
    }
    int32_t total = 0;
    {
        const Readback rb[1] = {{&total, si + n, sizeof(int32_t)}};
        AMGT_TRY(readback(rb, 1, s));
    }
    int32_t *sjj = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(total) + 1), s));
    if (!sjj) {
        dev_free(si, s);
        set_error("allocation of S columns failed");
        return AMGT_ERR_ALLOC;
    }
    if (n > 0) {
        hipLaunchKernelGGL((k_strength<true>), dim3(grid_of(n)), dim3(kThreads), 0, s, n, A_i, A_j,
                           A_data, theta, max_row_sum, nullptr, si, sjj);
        AMGT_HIP(hipGetLastError());
    }
    *nnzS = total;
    *S_i = si;
    *S_j = sjj;
    return AMGT_OK;
}

int amgt_amg_pmis(int32_t n, const int32_t *S_i, const int32_t *S_j, int cf_init,
                  int32_t *cf_marker, int32_t *n_coarse, void *stream) {
    AMGT_CHECK_ARG(n_coarse, "NULL output argument");
    AMGT_CHECK_ARG(n >= 0, "negative size");
    AMGT_CHECK_ARG(n == 0 || (S_i && cf_marker), "NULL row pointer or marker array");
    if (cf_init != 0 && cf_init != 1 && cf_init != 2) {
        set_error("PMIS: only CF_init 0, 1 (HMIS) and 2 (coarsen types 8, 10 and 9) are supported");
        return AMGT_ERR_UNSUPPORTED;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    *n_coarse = 0;
    if (n == 0) return AMGT_OK;
    Tmp tmp(s);
    int32_t *cnt = nullptr, *lost = nullptr, *any = nullptr, *flag = nullptr, *scan = nullptr;
    double *measure = nullptr;
    AMGT_TRY(tmp.get(&cnt, n));
    AMGT_TRY(tmp.get(&lost, n));
    AMGT_TRY(tmp.get(&any, 1));
    AMGT_TRY(tmp.get(&measure, n));
    AMGT_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * size_t(n), s));
    AMGT_HIP(hipMemsetAsync(any, 0, sizeof(int32_t), s));
    const dim3 g(grid_of(n)), b(kThreads);
    hipLaunchKernelGGL(k_col_count, g, b, 0, s, n, S_i, S_j, cnt);
    if (cf_init == 1)
        hipLaunchKernelGGL(k_pmis_init_hmis, g, b, 0, s, n, S_i, cnt, measure, cf_marker, any);
    else
        hipLaunchKernelGGL(k_pmis_init, g, b, 0, s, n, S_i, cnt, measure, cf_marker, any);
    AMGT_HIP(hipGetLastError());
    int32_t h_any = 0;
    const Readback rb_any[1] = {{&h_any, any, sizeof(int32_t)}};
    AMGT_TRY(readback(rb_any, 1, s));
    // every pass decides at least the nodes of largest measure, so the loop
    // ends within n passes
    for (int iter = 0; h_any && iter <= n; ++iter) {
        if (!cf_init || iter) {
            AMGT_HIP(hipMemsetAsync(lost, 0, sizeof(int32_t) * size_t(n), s));
            hipLaunchKernelGGL(k_pmis_clear, g, b, 0, s, n, S_i, S_j, measure, lost);
            hipLaunchKernelGGL(k_pmis_set, g, b, 0, s, n, measure, lost, cf_marker);
        }
        hipLaunchKernelGGL(k_pmis_cf, g, b, 0, s, n, S_i, S_j, measure, cf_marker);
        AMGT_HIP(hipMemsetAsync(any, 0, sizeof(int32_t), s));
        hipLaunchKernelGGL(k_pmis_update, g, b, 0, s, n, measure, cf_marker, any);
        AMGT_HIP(hipGetLastError());
        AMGT_TRY(readback(rb_any, 1, s));
    }
    if (h_any) {
        set_error("PMIS did not converge");
        return AMGT_ERR_INPUT;
    }
    AMGT_TRY(tmp.get(&flag, n));
    AMGT_TRY(tmp.get(&scan, size_t(n) + 1));
    hipLaunchKernelGGL(k_is_coarse, g, b, 0, s, n, cf_marker, flag);
    AMGT_HIP(hipGetLastError());
    AMGT_TRY(exclusive_scan_i32(flag, scan, n, s));
    const Readback rb[1] = {{n_coarse, scan + n, sizeof(int32_t)}};
    return readback(rb, 1, s);
}

int amgt_amg_extpi_interp(int32_t n, const int32_t *A_i, const int32_t *A_j,
                          const double *A_data, const int32_t *S_i, const int32_t *S_j,
                          const int32_t *cf_marker, int max_elmts, int32_t *n_coarse,
                          int64_t *nnzP, int32_t **P_i, int32_t **P_j, double **P_data,
                          void *stream) {
    AMGT_CHECK_ARG(n_coarse && nnzP && P_i && P_j && P_data, "NULL output argument");
    AMGT_CHECK_ARG(n >= 0 && max_elmts >= 0, "negative size");
    AMGT_CHECK_ARG(n == 0 || (A_i && S_i && cf_marker), "NULL row pointer or marker array");
    hipStream_t s = static_cast<hipStream_t>(stream);
    *n_coarse = 0;
    *nnzP = 0;
    *P_i = nullptr;
    *P_j = nullptr;
    *P_data = nullptr;
    Tmp tmp(s);
    int64_t *ub = nullptr, *off = nullptr;
    int32_t *flag = nullptr, *f2c = nullptr, *cnt = nullptr, *sj = nullptr;
    double *sw = nullptr;
    const dim3 g(grid_of(n)), b(kThreads);
    AMGT_TRY(tmp.get(&ub, size_t(n) + 1));
    AMGT_TRY(tmp.get(&off, size_t(n) + 1));
    AMGT_TRY(tmp.get(&flag, size_t(n) + 1));
    AMGT_TRY(tmp.get(&f2c, size_t(n) + 1));
    AMGT_TRY(tmp.get(&cnt, size_t(n) + 1));
    int32_t *pi = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(n) + 1), s));
    if (!pi) {
        set_error("allocation of P row pointer failed");
        return AMGT_ERR_ALLOC;
    }
    int st = AMGT_OK;
    int64_t total_ub = 0;
    int32_t nc = 0, nnz = 0;
    if (n > 0) {
        hipLaunchKernelGGL(k_is_coarse, g, b, 0, s, n, cf_marker, flag);
        hipLaunchKernelGGL(k_extpi_ub, g, b, 0, s, n, S_i, S_j, cf_marker, ub);
        if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
    }
    if (st == AMGT_OK) st = exclusive_scan_i32(flag, f2c, n, s);
    if (st == AMGT_OK) st = exclusive_scan_i64(ub, off, n, s);
    if (st == AMGT_OK) {
        const Readback rb[2] = {{&total_ub, off + n, sizeof(int64_t)},
                                {&nc, f2c + n, sizeof(int32_t)}};
        st = readback(rb, 2, s);
    }
    if (st == AMGT_OK) st = tmp.get(&sj, size_t(total_ub));
    if (st == AMGT_OK) st = tmp.get(&sw, size_t(total_ub));
    // per-row tables
    int64_t *hs = nullptr, *hoff = nullptr;
    int2 *hkv = nullptr;
    int64_t total_h = 0;
    if (st == AMGT_OK) st = tmp.get(&hs, size_t(n) + 1);
    if (st == AMGT_OK) st = tmp.get(&hoff, size_t(n) + 1);
    if (st == AMGT_OK && n > 0) {
        hipLaunchKernelGGL(k_extpi_hsize, g, b, 0, s, n, S_i, cf_marker, ub, hs);
        if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
    }
    if (st == AMGT_OK) st = exclusive_scan_i64(hs, hoff, n, s);
    if (st == AMGT_OK) {
        const Readback rb[1] = {{&total_h, hoff + n, sizeof(int64_t)}};
        st = readback(rb, 1, s);
    }
    if (st == AMGT_OK) st = tmp.get(&hkv, size_t(total_h));
    if (st == AMGT_OK && total_h > 0 &&
        hipMemsetAsync(hkv, 0xff, sizeof(int2) * size_t(total_h), s) != hipSuccess)
        st = AMGT_ERR_HIP;
    // long rows with ascending columns (coarse levels): a wave per row
    int32_t *use = nullptr, *upos = nullptr;
    int32_t nwave = 0;
    if (st == AMGT_OK && n > 0) {
        int64_t nnzA = 0;
        int32_t h_unsorted = 0;
        int32_t *unsorted = nullptr;
        int32_t lastI = 0;
        {
            const Readback rb[1] = {{&lastI, A_i + n, sizeof(int32_t)}};
            st = readback(rb, 1, s);
        }
        nnzA = lastI;
        if (st == AMGT_OK && n <= kWaveMaxRowsSmall && nnzA >= int64_t(kWaveMinRowLen) * n) {
            st = tmp.get(&unsorted, 1);
            if (st == AMGT_OK &&
                (hipMemsetAsync(unsorted, 0, sizeof(int32_t), s) != hipSuccess))
                st = AMGT_ERR_HIP;
            if (st == AMGT_OK) {
                hipLaunchKernelGGL(k_rows_unsorted, g, b, 0, s, n, A_i, A_j, unsorted);
                const Readback rb[1] = {{&h_unsorted, unsorted, sizeof(int32_t)}};
                st = hipGetLastError() != hipSuccess ? AMGT_ERR_HIP : readback(rb, 1, s);
            }
            if (st == AMGT_OK && !h_unsorted) {
                int32_t *flag2 = nullptr;
                st = tmp.get(&use, size_t(n) + 1);
                if (st == AMGT_OK) st = tmp.get(&upos, size_t(n) + 1);
                if (st == AMGT_OK) st = tmp.get(&flag2, size_t(n) + 1);
                if (st == AMGT_OK) {
                    hipLaunchKernelGGL(k_wave_rows, g, b, 0, s, n, A_i, S_i, cf_marker,
                                       n <= kWaveMaxRows ? 1 : 0, use);
                    if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
                }
                // (rows with this code, kernel, code for the rows that overflow it)
                const int stages[7][2] = {{kTier64, kTier64Big}, {kTier64Big, 0},
                                          {kTier128, kTier128Big}, {kTier128Big, 0},
                                          {kTier256, kTier256Big}, {kTier256Big, 0},
                                          {kTier512, 0}};
                for (int sg = 0; sg < 7 && st == AMGT_OK; ++sg) {
                    const int which = stages[sg][0], next = stages[sg][1];
                    int32_t cntw = 0;
                    hipLaunchKernelGGL(k_flag_eq, g, b, 0, s, n, use, which, flag2);
                    if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
                    if (st == AMGT_OK) st = exclusive_scan_i32(flag2, upos, n, s);
                    if (st == AMGT_OK) {
                        const Readback rb[1] = {{&cntw, upos + n, sizeof(int32_t)}};
                        st = readback(rb, 1, s);
                    }
                    if (st != AMGT_OK || cntw == 0) continue;
                    int32_t *wl = nullptr;
                    st = tmp.get(&wl, size_t(cntw));
                    if (st != AMGT_OK) break;
                    hipLaunchKernelGGL(k_list_rows, g, b, 0, s, n, flag2, upos, wl);
#define AMGT_EXTPI_WAVE(WA_, WT_, WH_)                                                     \
    hipLaunchKernelGGL((k_extpi_wave<WA_, WT_, WH_>), dim3(unsigned(cntw)), dim3(64), 0, s, cntw, \
                       wl, A_i, A_j, A_data, S_i, S_j, cf_marker, off, max_elmts, sj, sw, cnt, \
                       use, next)
                    switch (which) {
                        case kTier64: AMGT_EXTPI_WAVE(64, 128, 512); break;      // ~8.4 KB
                        case kTier64Big: AMGT_EXTPI_WAVE(64, 384, 1024); break;  // ~15.6 KB
                        case kTier128: AMGT_EXTPI_WAVE(128, 256, 1024); break;   // ~16 KB
                        case kTier256Big: AMGT_EXTPI_WAVE(256, 1536, 4096); break;  // ~60 KB
                        case kTier512: AMGT_EXTPI_WAVE(512, 1536, 4096); break;   // ~70 KB
                        default: AMGT_EXTPI_WAVE(256, 768, 2048); break;         // ~35 KB
                    }
#undef AMGT_EXTPI_WAVE
                    if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
                    nwave += cntw;
                }
            }
        }
    }
    if (st == AMGT_OK && n > 0) {
        hipLaunchKernelGGL(k_extpi_row, g, b, 0, s, n, A_i, A_j, A_data, S_i, S_j, cf_marker, off,
                           max_elmts, hoff, hkv, nwave > 0 ? use : nullptr, sj, sw, cnt);
        if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
    }
    if (st == AMGT_OK) st = exclusive_scan_i32(cnt, pi, n, s);
    if (st == AMGT_OK) {
        const Readback rb[1] = {{&nnz, pi + n, sizeof(int32_t)}};
        st = readback(rb, 1, s);
    }
    int32_t *pj = nullptr;
    double *pa = nullptr;
    if (st == AMGT_OK) {
        pj = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(nnz) + 1), s));
        pa = static_cast<double *>(dev_alloc(sizeof(double) * (size_t(nnz) + 1), s));
        if (!pj || !pa) {
            set_error("allocation of P failed");
            st = AMGT_ERR_ALLOC;
        }
    }
    if (st == AMGT_OK && n > 0) {
        hipLaunchKernelGGL(k_extpi_out, g, b, 0, s, n, off, pi, sj, sw, f2c, pj, pa);
        if (hipGetLastError() != hipSuccess) st = AMGT_ERR_HIP;
    }
    if (st != AMGT_OK) {
        if (st == AMGT_ERR_HIP) set_error("extended+i interpolation: kernel launch or copy failed");
        dev_free(pi, s);
        if (pj) dev_free(pj, s);
        if (pa) dev_free(pa, s);
        return st;
    }
    *n_coarse = nc;
    *nnzP = nnz;
    *P_i = pi;
    *P_j = pj;
    *P_data = pa;
    return AMGT_OK;
}

}  // extern "C"
