// SpMV entry point that allocates nothing, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
// forbid device allocation).  It runs the balanced kernel of amgt_mbsr_spmv
// (tile tables + carry sums of rows cut by tile edges) with those tables in a
// file-scope device array instead of per-handle allocations.
//
// The array is one per GPU, holds the last call's contents until the process
// ends, and is shared by every call on that GPU: calls that overlap on one GPU
// (two streams, two host threads) race on it and must be serialised by the
// caller.  Nothing in it has to start at zero: every call rebuilds the tables
// and the balanced kernel writes each carry slot before the fix-up reads it.
#include <mutex>

#include "internal.hpp"

namespace {

constexpr size_t kScratchBytes = AMGT_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// Host code may not use this name as a device address: hipGetSymbolAddress
// gives the address of each GPU's copy.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Address of the current GPU's copy of `scratch`, looked up once per device
// (with that device current) on its first call.
int scratch_address(void **out) {
    int dev = -1;
    AMGT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        amgt::set_error("no-alloc SpMV: device index out of range");
        return AMGT_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            amgt::set_error(std::string("no-alloc SpMV: hipGetSymbolAddress: ") +
                            hipGetErrorString(e));
            return AMGT_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return AMGT_OK;
}

}  // namespace

extern "C" {

int amgt_mbsr_spmv_noalloc(amgt_mbsr_t A, double alpha, const double *d_x, double beta,
                           const double *d_b, double *d_y, void *stream) {
    using namespace amgt;
    AMGT_CHECK_ARG(A != nullptr, "handle is NULL");
    const Mbsr &m = A->m;
    AMGT_CHECK_ARG(d_y != nullptr || m.nrows == 0, "y is NULL");
    AMGT_CHECK_ARG(d_x != nullptr || m.ncols == 0, "x is NULL");
    AMGT_CHECK_ARG(d_x != d_y || m.nrows == 0, "x and y are the same pointer");
    AMGT_CHECK_ARG(beta == 0.0 || d_b != nullptr, "b is NULL with beta != 0");
    AMGT_CHECK_ARG(!m.pattern_only, "pattern-only matrix has no values");
    AMGT_CHECK_ARG((reinterpret_cast<uintptr_t>(d_x) & 15u) == 0,
                   "x must be 16-byte aligned for the balanced kernel");
    // the handle's lanes setting, else the balanced kernel's default (tiles of
    // 256 blocks), as amgt_mbsr_spmv runs it with AMGT_SPMV_BALANCED
    const int lanes = m.spmv_lanes ? m.spmv_lanes : 16;
    const size_t need = spmv_balanced_ws_bytes(m, lanes);
    if (need > kScratchBytes) {
        set_error("no-alloc SpMV: the matrix needs " + std::to_string(need) +
                  " bytes of tile tables, the limit is " + std::to_string(kScratchBytes) +
                  " (AMGT_NOALLOC_SCRATCH_BYTES)");
        return AMGT_ERR_UNSUPPORTED;
    }
    void *ws = nullptr;
    AMGT_TRY(scratch_address(&ws));
    return spmv_balanced_ws(m, lanes, alpha, d_x, beta, d_b, d_y, ws, kScratchBytes,
                            reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
