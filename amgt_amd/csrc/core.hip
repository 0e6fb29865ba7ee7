// C-ABI surface of libamgt_mbsr: handles, allocation hooks, errors, scans, and
// the HYPRE-shaped entry points (hypreDevice_CSRSpGemm and
// hypre_CSRMatrixMatvecDevice counterparts).
#include <hipcub/hipcub.hpp>

#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>

#include "internal.hpp"

namespace amgt {

namespace {
thread_local std::string g_last_error;
std::mutex g_alloc_mu;
amgt_alloc_fn g_alloc = nullptr;
amgt_free_fn g_free = nullptr;
void *g_alloc_ctx = nullptr;
std::atomic<int> g_mixed{0};
}  // namespace

void set_error(const std::string &msg) { g_last_error = msg; }

bool mixed_precision_enabled() { return g_mixed.load() != 0; }

// Caching device allocator of the library.  hipMalloc/hipFree cost 0.1-1 ms
// for large blocks and the stream-ordered pool of this ROCm returns memory to
// the driver inside hipFreeAsync (measured: 0.6-3 ms per call even with an
// unlimited release threshold), so blocks are kept here instead: sizes round
// up to a class (four steps per power of two), a freed block is reused at once
// by the stream that freed it (later work on that stream runs after the work
// that used it), and by another stream only after an event recorded at the
// free has completed.  Blocks stay cached until the process ends.
namespace {
constexpr int kMaxDevices = 64;

struct CachedBlock {
    void *p;
    hipStream_t stream;
    hipEvent_t ev;
};

struct DeviceCache {
    std::multimap<size_t, CachedBlock> free_blocks;  // by class size
    std::unordered_map<void *, std::pair<size_t, hipEvent_t>> live;  // ptr -> (class size, event)
};

std::mutex g_cache_mu;
DeviceCache g_cache[kMaxDevices];

size_t size_class(size_t n) {
    if (n <= 4096) return 4096;
    size_t p = 4096;
    while (p * 2 < n) p *= 2;  // p < n <= 2p
    const size_t step = p / 4;
    return (n + step - 1) / step * step;
}
}  // namespace

static void *pool_alloc(size_t bytes, hipStream_t s) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    const size_t sz = size_class(bytes);
    std::lock_guard<std::mutex> lk(g_cache_mu);
    DeviceCache &dc = g_cache[dev];
    // a block of this class, or up to twice as large: same stream first, then
    // any block whose last use has completed
    auto lo = dc.free_blocks.lower_bound(sz), hi = dc.free_blocks.upper_bound(2 * sz);
    auto pick = dc.free_blocks.end();
    for (auto it = lo; it != hi; ++it)
        if (it->second.stream == s) {
            pick = it;
            break;
        }
    if (pick == dc.free_blocks.end())
        for (auto it = lo; it != hi; ++it)
            if (hipEventQuery(it->second.ev) == hipSuccess) {
                pick = it;
                break;
            }
    if (pick != dc.free_blocks.end()) {
        void *p = pick->second.p;
        dc.live[p] = {pick->first, pick->second.ev};
        dc.free_blocks.erase(pick);
        return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        // give back the cached blocks that are idle, then try once more
        for (auto it = dc.free_blocks.begin(); it != dc.free_blocks.end();) {
            if (hipEventQuery(it->second.ev) == hipSuccess) {
                (void)hipFree(it->second.p);
                (void)hipEventDestroy(it->second.ev);
                it = dc.free_blocks.erase(it);
            } else {
                ++it;
            }
        }
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipFree(p);
        return nullptr;
    }
    dc.live[p] = {sz, ev};
    return p;
}

static void pool_free(void *p, hipStream_t s) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (auto &dc : g_cache) {
        auto it = dc.live.find(p);
        if (it == dc.live.end()) continue;
        const size_t sz = it->second.first;
        hipEvent_t ev = it->second.second;
        dc.live.erase(it);
        (void)hipEventRecord(ev, s);
        dc.free_blocks.emplace(sz, CachedBlock{p, s, ev});
        return;
    }
}

void *dev_alloc(size_t bytes, hipStream_t s) {
    if (bytes == 0) bytes = 1;
    amgt_alloc_fn fn;
    void *ctx;
    {
        std::lock_guard<std::mutex> lk(g_alloc_mu);
        fn = g_alloc;
        ctx = g_alloc_ctx;
    }
    if (fn) return fn(bytes, ctx);
    return pool_alloc(bytes, s);
}

void dev_free(void *p, hipStream_t s) {
    if (!p) return;
    amgt_free_fn fn;
    void *ctx;
    {
        std::lock_guard<std::mutex> lk(g_alloc_mu);
        fn = g_free;
        ctx = g_alloc_ctx;
    }
    if (fn) {
        fn(p, ctx);
        return;
    }
    pool_free(p, s);
}

int tmp_alloc(void **p, size_t bytes, hipStream_t s) {
    *p = nullptr;
    if (bytes == 0) bytes = 1;
    *p = pool_alloc(bytes, s);
    if (!*p) {
        set_error("device allocation of " + std::to_string(bytes) + " bytes failed");
        return AMGT_ERR_ALLOC;
    }
    return AMGT_OK;
}

void tmp_free(void *p, hipStream_t s) { pool_free(p, s); }

int readback(const Readback *items, int n, hipStream_t s) {
    // one pinned area per host thread, allocated on first use, kept for the process
    thread_local unsigned char *area = nullptr;
    if (!area) {
        void *p = nullptr;
        AMGT_HIP(hipHostMalloc(&p, kReadbackBytes, hipHostMallocDefault));
        area = static_cast<unsigned char *>(p);
    }
    size_t off = 0;
    for (int k = 0; k < n; ++k) {
        const size_t b = items[k].bytes;
        if (off + b > kReadbackBytes) {
            set_error("readback: more than kReadbackBytes requested");
            return AMGT_ERR_ARG;
        }
        AMGT_HIP(hipMemcpyAsync(area + off, items[k].src, b, hipMemcpyDeviceToHost, s));
        off = (off + b + 15) & ~size_t(15);
    }
    AMGT_HIP(hipStreamSynchronize(s));
    off = 0;
    for (int k = 0; k < n; ++k) {
        std::memcpy(items[k].dst, area + off, items[k].bytes);
        off = (off + items[k].bytes + 15) & ~size_t(15);
    }
    return AMGT_OK;
}

template <typename T>
__global__ void k_set_zero(T *p) {
    if (threadIdx.x == 0) p[0] = T(0);
}

template <typename T>
static int inclusive_into_exclusive(const T *in, T *out, int64_t n, hipStream_t s) {
    // out[0] = 0; out[1..n] = inclusive sums of in[0..n-1]
    hipLaunchKernelGGL(k_set_zero<T>, dim3(1), dim3(64), 0, s, out);
    AMGT_HIP(hipGetLastError());
    if (n == 0) return AMGT_OK;
    size_t bytes = 0;
    AMGT_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out + 1, (int)n, s));
    void *tmp = nullptr;
    AMGT_TRY(tmp_alloc(&tmp, bytes, s));
    hipError_t e = hipcub::DeviceScan::InclusiveSum(tmp, bytes, in, out + 1, (int)n, s);
    tmp_free(tmp, s);
    if (e != hipSuccess) {
        set_error(std::string("hipcub scan: ") + hipGetErrorString(e));
        return AMGT_ERR_HIP;
    }
    return AMGT_OK;
}

int exclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t s) {
    if (n >= (int64_t(1) << 31)) {
        set_error("scan length exceeds int32");
        return AMGT_ERR_UNSUPPORTED;
    }
    return inclusive_into_exclusive<int32_t>(in, out, n, s);
}

int exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s) {
    if (n >= (int64_t(1) << 31)) {
        set_error("scan length exceeds int32");
        return AMGT_ERR_UNSUPPORTED;
    }
    return inclusive_into_exclusive<int64_t>(in, out, n, s);
}

void mbsr_release_balanced(Mbsr *A, hipStream_t s) {
    dev_free(A->bal_lo, s);
    dev_free(A->bal_head, s);
    dev_free(A->bal_vb, s);
    dev_free(A->bal_carry, s);
    A->bal_lo = A->bal_head = A->bal_vb = nullptr;
    A->bal_carry = nullptr;
    A->bal_capb = A->bal_ntiles = 0;
}

void mbsr_release(Mbsr *A, hipStream_t s) {
    if (!A) return;
    mbsr_release_balanced(A, s);
    if (A->trans) {
        mbsr_release(A->trans, s);
        delete A->trans;
        A->trans = nullptr;
    }
    dev_free(A->blk_ptr, s);
    dev_free(A->blk_col, s);
    dev_free(A->blk_map, s);
    dev_free(A->val_ptr, s);
    dev_free(A->val, s);
    dev_free(A->val32, s);
    dev_free(A->val16, s);
    *A = Mbsr();
}

int effective_precision(const Mbsr &A) {
    int p = AMGT_DOUBLE;
    if (A.precision >= 0) p = A.precision;
    else if (mixed_precision_enabled()) p = amgt_precision_for_level(A.level);
    // values known to have no FP16 copy (non-finite) run in FP32 (mbsr_ensure_lowp)
    if (p == AMGT_HALF && A.half_fits == 0) p = AMGT_FLOAT;
    return p;
}

}  // namespace amgt

using namespace amgt;

static hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

extern "C" {

const char *amgt_status_string(int status) {
    switch (status) {
        case AMGT_OK: return "ok";
        case AMGT_ERR_ARG: return "invalid argument";
        case AMGT_ERR_DIM: return "incompatible dimensions";
        case AMGT_ERR_ALLOC: return "device allocation failed";
        case AMGT_ERR_HIP: return "HIP runtime error";
        case AMGT_ERR_INPUT: return "invalid CSR input";
        case AMGT_ERR_UNSUPPORTED: return "unsupported size";
        default: return "unknown status";
    }
}

const char *amgt_last_error(void) { return g_last_error.c_str(); }

int amgt_set_allocator(amgt_alloc_fn alloc_fn, amgt_free_fn free_fn, void *ctx) {
    if ((alloc_fn == nullptr) != (free_fn == nullptr)) {
        set_error("amgt_set_allocator: give both hooks or neither");
        return AMGT_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    g_alloc = alloc_fn;
    g_free = free_fn;
    g_alloc_ctx = ctx;
    return AMGT_OK;
}

int amgt_precision_for_level(int level) {
    // get_calculationPrecison, csr_spgemm_device.c:1847-1862
    if (level <= 0) return AMGT_DOUBLE;
    if (level <= 1) return AMGT_FLOAT;
    return AMGT_HALF;
}

int amgt_set_mixed_precision(int enable) {
    g_mixed.store(enable ? 1 : 0);
    return AMGT_OK;
}

int amgt_mbsr_create_from_csr(int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *d_rowptr,
                              const int32_t *d_colidx, const double *d_val, void *stream,
                              amgt_mbsr_t *out) {
    AMGT_CHECK_ARG(out != nullptr, "out is NULL");
    *out = nullptr;
    AMGT_CHECK_ARG(nrows >= 0 && ncols >= 0 && nnz >= 0, "negative size");
    AMGT_CHECK_ARG(d_rowptr != nullptr, "rowptr is NULL");
    AMGT_CHECK_ARG(nnz == 0 || d_colidx != nullptr, "colidx is NULL");
    hipStream_t s = as_stream(stream);
    auto *h = new amgt_mbsr();
    int st = csr_to_mbsr(nrows, ncols, nnz, d_rowptr, d_colidx, d_val, s, &h->m);
    if (st != AMGT_OK) {
        mbsr_release(&h->m, s);
        delete h;
        return st;
    }
    *out = h;
    return AMGT_OK;
}

int amgt_mbsr_destroy(amgt_mbsr_t A) {
    if (!A) return AMGT_OK;
    // no stream is given: wait for all work that may still read the arrays
    (void)hipDeviceSynchronize();
    mbsr_release(&A->m, nullptr);
    delete A;
    return AMGT_OK;
}

int amgt_mbsr_update_values(amgt_mbsr_t A, const int32_t *d_rowptr, const int32_t *d_colidx,
                            const double *d_val, void *stream) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    AMGT_CHECK_ARG(d_val != nullptr, "values are NULL");
    hipStream_t s = as_stream(stream);
    AMGT_TRY(mbsr_refresh_values(&A->m, d_rowptr, d_colidx, d_val, s));
    return AMGT_OK;
}

int amgt_mbsr_get_info(amgt_mbsr_t A, amgt_mbsr_info_t *info) {
    AMGT_CHECK_ARG(A != nullptr && info != nullptr, "NULL argument");
    const Mbsr &m = A->m;
    info->nrows = m.nrows;
    info->ncols = m.ncols;
    info->brow = m.brow;
    info->bcol = m.bcol;
    info->nblk = m.nblk;
    info->nnz = m.nnz;
    info->stand = m.stand;
    info->avg_nnz = m.avg_nnz;
    info->fill = m.nblk ? double(m.nnz) / (16.0 * double(m.nblk)) : 0.0;
    info->max_row_blocks = m.max_row_blocks;
    info->level = m.level;
    info->precision = effective_precision(m);
    info->spmv_lanes = m.spmv_lanes ? m.spmv_lanes : spmv_default_lanes(m);
    info->spmv_kernel = m.spmv_kernel;
    return AMGT_OK;
}

int amgt_mbsr_set_spmv_kernel(amgt_mbsr_t A, int kind) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    AMGT_CHECK_ARG(kind >= AMGT_SPMV_DIRECT && kind <= AMGT_SPMV_BALANCED, "unknown SpMV kernel kind");
    A->m.spmv_kernel = kind;
    return AMGT_OK;
}

int amgt_mbsr_get_arrays(amgt_mbsr_t A, amgt_mbsr_arrays_t *arrays) {
    AMGT_CHECK_ARG(A != nullptr && arrays != nullptr, "NULL argument");
    arrays->blk_ptr = A->m.blk_ptr;
    arrays->blk_col = A->m.blk_col;
    arrays->blk_map = A->m.blk_map;
    arrays->val_ptr = A->m.val_ptr;
    arrays->val = A->m.val;
    return AMGT_OK;
}

int amgt_mbsr_set_level(amgt_mbsr_t A, int level) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    A->m.level = level;
    return AMGT_OK;
}

int amgt_mbsr_set_precision(amgt_mbsr_t A, int precision) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    AMGT_CHECK_ARG(precision >= -1 && precision <= AMGT_HALF, "precision out of range");
    A->m.precision = precision;
    return AMGT_OK;
}

int amgt_mbsr_set_spmv_lanes(amgt_mbsr_t A, int lanes) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    AMGT_CHECK_ARG(lanes == 0 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 ||
                       lanes == 64,
                   "lanes must be 0, 4, 8, 16, 32 or 64");
    A->m.spmv_lanes = lanes;
    return AMGT_OK;
}

// Times y = A x with every timed kernel kind x lanes setting (two sweeps,
// clocks warmed, best of each): t[k * 5 + c] for kind kinds[k] and lanes
// cand[c], 1e30 where a pair does not exist.  Leaves the handle's kernel and
// lanes as they were.
static const int kTuneLanes[5] = {4, 8, 16, 32, 64};
static const int kTuneKinds[4] = {kSpmvBlock, kSpmvBalanced, kSpmvStaged, kSpmvDirect};

static int spmv_sweep(Mbsr &m, hipStream_t s, float *t_min) {
    for (int q = 0; q < 20; ++q) t_min[q] = 1e30f;
    if (m.nblk == 0 || m.pattern_only) return AMGT_OK;
    const int kind0 = m.spmv_kernel, lanes0 = m.spmv_lanes;
    int prec = effective_precision(m);
    AMGT_TRY(mbsr_ensure_lowp(&m, &prec, s));
    double *x = nullptr, *y = nullptr;
    AMGT_TRY(tmp_alloc_t(&x, size_t(m.ncols) + 4, s));
    AMGT_TRY(tmp_alloc_t(&y, size_t(m.nrows) + 4, s));
    AMGT_HIP(hipMemsetAsync(x, 0, sizeof(double) * (size_t(m.ncols) + 4), s));
    hipEvent_t e0, e1;
    AMGT_HIP(hipEventCreate(&e0));
    AMGT_HIP(hipEventCreate(&e1));
    // bring the clocks up before the first measurement
    AMGT_TRY(spmv_prepare(&m, spmv_default_lanes(m), s));
    for (int r = 0; r < 30; ++r) AMGT_TRY(spmv(m, prec, spmv_default_lanes(m), 1.0, x, 0.0, nullptr, y, s));
    for (int sweep = 0; sweep < 2; ++sweep) {
        for (int kk = 0; kk < 4; ++kk) {
            m.spmv_kernel = kTuneKinds[kk];
            for (int c = 0; c < 5; ++c) {
                // the balanced kernel has three tile sizes (lanes 4, 8, 16)
                if (kTuneKinds[kk] == kSpmvBalanced && kTuneLanes[c] > 16) continue;
                int st = spmv_prepare(&m, kTuneLanes[c], s);
                for (int r = 0; r < 3 && st == AMGT_OK; ++r)  // warm-up
                    st = spmv(m, prec, kTuneLanes[c], 1.0, x, 0.0, nullptr, y, s);
                if (st != AMGT_OK) continue;
                AMGT_HIP(hipEventRecord(e0, s));
                const int reps = 10;
                for (int r = 0; r < reps; ++r)
                    spmv(m, prec, kTuneLanes[c], 1.0, x, 0.0, nullptr, y, s);
                AMGT_HIP(hipEventRecord(e1, s));
                AMGT_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                AMGT_HIP(hipEventElapsedTime(&ms, e0, e1));
                ms /= float(reps);
                if (ms < t_min[kk * 5 + c]) t_min[kk * 5 + c] = ms;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tmp_free(x, s);
    tmp_free(y, s);
    m.spmv_kernel = kind0;
    m.spmv_lanes = lanes0;
    return spmv_prepare(&m, m.spmv_lanes ? m.spmv_lanes : spmv_default_lanes(m), s);
}

int amgt_mbsr_spmv_sweep(amgt_mbsr_t A, void *stream, float *times_ms) {
    AMGT_CHECK_ARG(A != nullptr && times_ms != nullptr, "NULL argument");
    return spmv_sweep(A->m, as_stream(stream), times_ms);
}

int amgt_mbsr_tune_from_stats(amgt_mbsr_t A) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    apply_spmv_defaults(A->m);
    return AMGT_OK;
}

int amgt_mbsr_autotune(amgt_mbsr_t A, void *stream) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    Mbsr &m = A->m;
    if (m.nblk == 0 || m.pattern_only) return AMGT_OK;
    hipStream_t s = as_stream(stream);
    float t[20];
    AMGT_TRY(spmv_sweep(m, s, t));
    float best = 1e30f;
    int best_lanes = spmv_default_lanes(m), best_kind = kTuneKinds[0];
    for (int kk = 0; kk < 4; ++kk)
        for (int c = 0; c < 5; ++c)
            if (t[kk * 5 + c] < best) {
                best = t[kk * 5 + c];
                best_lanes = kTuneLanes[c];
                best_kind = kTuneKinds[kk];
            }
    m.spmv_lanes = best_lanes;
    m.spmv_kernel = best_kind;
    return spmv_prepare(&m, best_lanes, s);
}

static int lanes_of(const Mbsr &m) { return m.spmv_lanes ? m.spmv_lanes : spmv_default_lanes(m); }

static int get_transpose(Mbsr &m, hipStream_t s, Mbsr **t) {
    if (!m.trans) {
        auto *tt = new Mbsr();
        int st = mbsr_transpose(m, s, tt);
        if (st != AMGT_OK) {
            mbsr_release(tt, s);
            delete tt;
            return st;
        }
        tt->level = m.level;
        tt->precision = m.precision;
        m.trans = tt;
    }
    *t = m.trans;
    return AMGT_OK;
}

int amgt_mbsr_spmv(amgt_mbsr_t A, int trans, double alpha, const double *d_x, double beta,
                   const double *d_b, double *d_y, void *stream) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    Mbsr &m = A->m;
    hipStream_t s = as_stream(stream);
    const int ny = trans ? m.ncols : m.nrows;
    const int nx = trans ? m.nrows : m.ncols;
    AMGT_CHECK_ARG(d_y != nullptr || ny == 0, "y is NULL");
    AMGT_CHECK_ARG(d_x != nullptr || nx == 0, "x is NULL");
    // hypre_CSRMatrixMatvecDevice2 csr_matvec_device.c:65-69
    AMGT_CHECK_ARG(d_x != d_y || ny == 0, "x and y are the same pointer");
    AMGT_CHECK_ARG(beta == 0.0 || d_b != nullptr, "b is NULL with beta != 0");
    AMGT_CHECK_ARG(!m.pattern_only, "pattern-only matrix has no values");
    Mbsr *op = &m;
    if (trans) AMGT_TRY(get_transpose(m, s, &op));
    int prec = effective_precision(m);
    AMGT_TRY(mbsr_ensure_lowp(op, &prec, s));
    AMGT_TRY(spmv_prepare(op, lanes_of(*op), s));
    return spmv(*op, prec, lanes_of(*op), alpha, d_x, beta, d_b, d_y, s);
}

int amgt_mbsr_spmm(amgt_mbsr_t A, int trans, int k, double alpha, const double *d_x, int64_t ldx,
                   double beta, const double *d_b, int64_t ldb, double *d_y, int64_t ldy,
                   void *stream) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    AMGT_CHECK_ARG(k >= 1, "k must be >= 1");
    Mbsr &m = A->m;
    hipStream_t s = as_stream(stream);
    const int ny = trans ? m.ncols : m.nrows;
    const int nx = trans ? m.nrows : m.ncols;
    AMGT_CHECK_ARG(ldx >= nx && ldy >= ny, "leading dimension too small");
    AMGT_CHECK_ARG(beta == 0.0 || (d_b != nullptr && ldb >= ny), "b is NULL or ldb too small");
    AMGT_CHECK_ARG(d_x != d_y || ny == 0, "x and y are the same pointer");
    AMGT_CHECK_ARG(!m.pattern_only, "pattern-only matrix has no values");
    Mbsr *op = &m;
    if (trans) AMGT_TRY(get_transpose(m, s, &op));
    int prec = effective_precision(m);
    AMGT_TRY(mbsr_ensure_lowp(op, &prec, s));
    AMGT_TRY(spmv_prepare(op, lanes_of(*op), s));  // k == 1 runs the SpMV kernels
    return spmm(*op, prec, lanes_of(*op), k, alpha, d_x, ldx, beta, d_b, ldb, d_y, ldy, s);
}

int amgt_mbsr_spmm_interleaved(amgt_mbsr_t A, int trans, int k, double alpha, const double *d_x,
                               double beta, const double *d_b, double *d_y, void *stream) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    AMGT_CHECK_ARG(k >= 1, "k must be >= 1");
    Mbsr &m = A->m;
    hipStream_t s = as_stream(stream);
    const int ny = trans ? m.ncols : m.nrows;
    AMGT_CHECK_ARG(d_y != nullptr || ny == 0, "y is NULL");
    AMGT_CHECK_ARG(beta == 0.0 || d_b != nullptr, "b is NULL with beta != 0");
    AMGT_CHECK_ARG(d_x != d_y || ny == 0, "x and y are the same pointer");
    AMGT_CHECK_ARG(!m.pattern_only, "pattern-only matrix has no values");
    Mbsr *op = &m;
    if (trans) AMGT_TRY(get_transpose(m, s, &op));
    int prec = effective_precision(m);
    AMGT_TRY(mbsr_ensure_lowp(op, &prec, s));
    if (op->nblk == 0 || alpha == 0.0) {
        // y = beta b over all k*ny entries
        return scale_vector(int32_t(int64_t(ny) * k), beta, d_b, d_y, s);
    }
    return spmm_interleaved(*op, prec, k, alpha, d_x, beta, d_b, d_y, s);
}

int amgt_mbsr_transpose(amgt_mbsr_t A, void *stream, amgt_mbsr_t *AT) {
    AMGT_CHECK_ARG(A != nullptr && AT != nullptr, "NULL argument");
    *AT = nullptr;
    hipStream_t s = as_stream(stream);
    auto *h = new amgt_mbsr();
    int st = mbsr_transpose(A->m, s, &h->m);
    if (st != AMGT_OK) {
        mbsr_release(&h->m, s);
        delete h;
        return st;
    }
    h->m.level = A->m.level;
    *AT = h;
    return AMGT_OK;
}

int amgt_mbsr_diagonal(amgt_mbsr_t A, int kind, double *d_diag, void *stream) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    AMGT_CHECK_ARG(kind == AMGT_DIAG_VALUE || kind == AMGT_DIAG_INV_ABS ||
                       kind == AMGT_DIAG_INV_L1 || kind == AMGT_DIAG_INV,
                   "unknown diagonal kind");
    AMGT_CHECK_ARG(!A->m.pattern_only, "pattern-only matrix has no values");
    const int32_t n = A->m.nrows < A->m.ncols ? A->m.nrows : A->m.ncols;
    AMGT_CHECK_ARG(d_diag != nullptr || n == 0, "output is NULL");
    return mbsr_diagonal(A->m, kind, d_diag, as_stream(stream));
}

int amgt_mbsr_relax_jacobi(amgt_mbsr_t A, double omega, const double *d_dinv, const double *d_f,
                           const double *d_u, double *d_u_new, void *stream) {
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    Mbsr &m = A->m;
    AMGT_CHECK_ARG(m.nrows == m.ncols, "Jacobi sweep needs a square matrix");
    AMGT_CHECK_ARG(!m.pattern_only, "pattern-only matrix has no values");
    AMGT_CHECK_ARG(m.nrows == 0 || (d_dinv && d_f && d_u && d_u_new), "NULL vector");
    AMGT_CHECK_ARG(d_u != d_u_new || m.nrows == 0, "u_new must not alias u");
    hipStream_t s = as_stream(stream);
    int prec = effective_precision(m);
    AMGT_TRY(mbsr_ensure_lowp(&m, &prec, s));
    AMGT_TRY(spmv_prepare(&m, lanes_of(m), s));
    return relax_jacobi(m, prec, lanes_of(m), omega, d_dinv, d_f, d_u, d_u_new, s);
}

int amgt_mbsr_spgemm(amgt_mbsr_t A, amgt_mbsr_t B, void *stream, amgt_mbsr_t *C) {
    AMGT_CHECK_ARG(A != nullptr && B != nullptr && C != nullptr, "NULL argument");
    *C = nullptr;
    if (A->m.ncols != B->m.nrows) {
        set_error("amgt_mbsr_spgemm: A.ncols != B.nrows");
        return AMGT_ERR_DIM;
    }
    hipStream_t s = as_stream(stream);
    // operand precision of A's level (get_calculationPrecison(hypre_Level(A)),
    // csr_spgemm_device.c:2592) or A's explicit precision
    int prec = effective_precision(A->m);
    if (!A->m.pattern_only && !B->m.pattern_only) {
        int pa = prec, pb = prec;
        AMGT_TRY(mbsr_ensure_lowp(&A->m, &pa, s));
        AMGT_TRY(mbsr_ensure_lowp(&B->m, &pb, s));
        if (pa != prec || pb != prec) {  // an operand without an FP16 copy: both in FP32
            prec = AMGT_FLOAT;
            AMGT_TRY(mbsr_ensure_lowp(&A->m, &prec, s));
            AMGT_TRY(mbsr_ensure_lowp(&B->m, &prec, s));
        }
    }
    auto *h = new amgt_mbsr();
    int st = spgemm(A->m, B->m, prec, s, &h->m);
    if (st != AMGT_OK) {
        mbsr_release(&h->m, s);
        delete h;
        return st;
    }
    h->m.level = A->m.level;
    *C = h;
    return AMGT_OK;
}

int amgt_mbsr_to_csr(amgt_mbsr_t A, void *stream, int64_t *nnz, int32_t **d_rowptr,
                     int32_t **d_colidx, double **d_val) {
    AMGT_CHECK_ARG(A && nnz && d_rowptr && d_colidx && d_val, "NULL argument");
    return mbsr_to_csr(A->m, as_stream(stream), nnz, d_rowptr, d_colidx, d_val);
}

int amgt_free(void *ptr) {
    if (!ptr) return AMGT_OK;
    // no stream is given: wait for all work that may still use the array
    (void)hipDeviceSynchronize();
    dev_free(ptr, nullptr);
    return AMGT_OK;
}

int amgt_csr_spgemm(int32_t m, int32_t k, int32_t n, int64_t nnzA, const int32_t *A_i,
                    const int32_t *A_j, const double *A_data, amgt_mbsr_t *A_cache, int64_t nnzB,
                    const int32_t *B_i, const int32_t *B_j, const double *B_data,
                    amgt_mbsr_t *B_cache, int64_t *nnzC, int32_t **C_i, int32_t **C_j,
                    double **C_data, amgt_mbsr_t *C_cache, void *stream) {
    AMGT_CHECK_ARG(nnzC && C_i && C_j && C_data, "NULL output argument");
    AMGT_CHECK_ARG(m >= 0 && k >= 0 && n >= 0 && nnzA >= 0 && nnzB >= 0, "negative size");
    hipStream_t s = as_stream(stream);
    *nnzC = 0;
    *C_i = nullptr;
    *C_j = nullptr;
    *C_data = nullptr;
    if (C_cache) *C_cache = nullptr;
    if (nnzA == 0 || nnzB == 0) {
        // trivial case, csr_spgemm_device.c:2633-2638
        int32_t *ci = static_cast<int32_t *>(dev_alloc(sizeof(int32_t) * (size_t(m) + 1), s));
        if (!ci) {
            set_error("allocation of C row pointer failed");
            return AMGT_ERR_ALLOC;
        }
        AMGT_HIP(hipMemsetAsync(ci, 0, sizeof(int32_t) * (size_t(m) + 1), s));
        *C_i = ci;
        return AMGT_OK;
    }
    amgt_mbsr_t a = A_cache ? *A_cache : nullptr;
    amgt_mbsr_t b = B_cache ? *B_cache : nullptr;
    bool own_a = false, own_b = false;
    if (!a) {
        AMGT_TRY(amgt_mbsr_create_from_csr(m, k, nnzA, A_i, A_j, A_data, stream, &a));
        if (A_cache) *A_cache = a; else own_a = true;
    }
    if (!b) {
        int st = amgt_mbsr_create_from_csr(k, n, nnzB, B_i, B_j, B_data, stream, &b);
        if (st != AMGT_OK) {
            if (own_a) amgt_mbsr_destroy(a);
            return st;
        }
        if (B_cache) *B_cache = b; else own_b = true;
    }
    amgt_mbsr_t c = nullptr;
    int st = amgt_mbsr_spgemm(a, b, stream, &c);
    if (st == AMGT_OK) {
        st = amgt_mbsr_to_csr(c, stream, nnzC, C_i, C_j, C_data);
        if (st == AMGT_OK && C_cache) {
            *C_cache = c;
            c = nullptr;
        }
    }
    if (c) amgt_mbsr_destroy(c);
    if (own_a) amgt_mbsr_destroy(a);
    if (own_b) amgt_mbsr_destroy(b);
    return st;
}

int amgt_csr_matvec(int trans, double alpha, int32_t nrows, int32_t ncols, int64_t nnz,
                    const int32_t *A_i, const int32_t *A_j, const double *A_data,
                    amgt_mbsr_t *A_cache, const double *d_x, double beta, const double *d_b,
                    double *d_y, void *stream) {
    hipStream_t s = as_stream(stream);
    const int ny = trans ? ncols : nrows;
    AMGT_CHECK_ARG(d_x != d_y || ny == 0, "x and y are the same pointer");
    if (nnz <= 0 || alpha == 0.0) {
        // csr_matvec_device.c:175-178: y = beta*b
        AMGT_CHECK_ARG(d_y != nullptr || ny == 0, "y is NULL");
        AMGT_CHECK_ARG(beta == 0.0 || d_b != nullptr || ny == 0, "b is NULL with beta != 0");
        return scale_vector(ny, beta, d_b, d_y, s);
    }
    amgt_mbsr_t a = A_cache ? *A_cache : nullptr;
    bool own = false;
    if (!a) {
        AMGT_TRY(amgt_mbsr_create_from_csr(nrows, ncols, nnz, A_i, A_j, A_data, stream, &a));
        if (A_cache) *A_cache = a; else own = true;
    }
    int st = amgt_mbsr_spmv(a, trans, alpha, d_x, beta, d_b, d_y, stream);
    if (own) {
        (void)hipStreamSynchronize(s);
        amgt_mbsr_destroy(a);
    }
    return st;
}

}  // extern "C"
