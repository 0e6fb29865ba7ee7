// HMIS coarsening (hypre_BoomerAMGCoarsenHMIS, par_coarsen.c:2845-2866): the
// first pass of Ruge-Stueben coarsening on the host, then PMIS with CF_init 1
// on the device (amgt_amg_pmis).
//
// The first pass (hypre_BoomerAMGCoarsenRuge par_coarsen.c:911-1398 with
// coarsen_type 10, measure_type 0, one process) is a greedy sweep: the
// undecided point of largest measure (number of points it influences plus
// the F points that depend on it) becomes C, the points it influences become
// F, and measures are updated around them.  Which point is "largest" among
// equals decides the whole splitting, so the order is HYPRE's: points wait
// in one first-in-first-out list per measure, a point whose measure changes
// goes to the end of its new list, and the head of the largest non-empty list
// is taken (the behaviour of HYPRE's list of lists, utilities/amg_linklist.c).
// HYPRE keeps the lists in a linked list searched from the largest measure;
// here they are an array indexed by measure, so every step is O(1).  Points
// that influence nobody are marked Z (-2) and left for PMIS, which re-examines
// them, keeps the C points and decides the F points again
// (par_coarsen.c:2353-2386, first pass of its loop at :2494).
//
// The sweep is sequential by nature, as in HYPRE (the GPU build of HYPRE has
// no HMIS).  S and its transpose (built on the device) go to pinned host
// memory, the markers come back, and the rest runs on the device.
#include "internal.hpp"

#include <mutex>
#include <vector>

namespace amgt {

namespace {

constexpr int32_t kC = 1, kF = -1, kZ = -2, kSF = -3, kUndecided = 0;

// first-in-first-out lists of points, one per measure (measures > 0)
class MeasureLists {
  public:
    explicit MeasureLists(int32_t n) : next_(static_cast<size_t>(n)), prev_(static_cast<size_t>(n)) {}

    void enter(int32_t m, int32_t p) {
        if (size_t(m) >= head_.size()) {
            head_.resize(size_t(m) * 2 + 16, -1);
            tail_.resize(head_.size(), -1);
        }
        next_[p] = -1;
        if (tail_[m] < 0) {
            head_[m] = p;
            prev_[p] = -1;
        } else {
            next_[tail_[m]] = p;
            prev_[p] = tail_[m];
        }
        tail_[m] = p;
        if (m > top_) top_ = m;
    }

    void remove(int32_t m, int32_t p) {
        if (prev_[p] >= 0) next_[prev_[p]] = next_[p]; else head_[m] = next_[p];
        if (next_[p] >= 0) prev_[next_[p]] = prev_[p]; else tail_[m] = prev_[p];
    }

    // oldest point of the largest measure, or -1 when every list is empty
    int32_t first() {
        while (top_ > 0 && head_[top_] < 0) --top_;
        return top_ > 0 ? head_[top_] : -1;
    }

  private:
    std::vector<int32_t> head_, tail_, next_, prev_;
    int32_t top_ = 0;
};

// S^T on the host: the points each point influences, in ascending order
void transpose_host(int32_t n, const int32_t *SI, const int32_t *SJ, std::vector<int32_t> &ti,
                    std::vector<int32_t> &tj) {
    ti.assign(static_cast<size_t>(n) + 1, 0);
    tj.resize(static_cast<size_t>(SI[n]));
    for (int64_t p = 0; p < SI[n]; ++p) ++ti[size_t(SJ[p]) + 1];
    for (int32_t i = 0; i < n; ++i) ti[i + 1] += ti[i];
    std::vector<int32_t> pos(ti.begin(), ti.end() - 1);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t p = SI[i]; p < SI[i + 1]; ++p) tj[size_t(pos[SJ[p]]++)] = i;
}

// the sweep, given S and S^T (rows of S^T ascending)
void ruge_first_pass(int32_t n, const int32_t *SI, const int32_t *SJ, const int32_t *ti,
                     const int32_t *tj, int32_t *cf) {
    std::vector<int32_t> meas(static_cast<size_t>(n));
    int64_t left = 0;
    for (int32_t i = 0; i < n; ++i) {
        meas[i] = ti[i + 1] - ti[i];
        if (SI[i + 1] == SI[i]) {
            cf[i] = kSF;
            meas[i] = 0;
        } else {
            cf[i] = kUndecided;
            ++left;
        }
    }
    MeasureLists L(n);
    for (int32_t j = 0; j < n; ++j) {
        if (cf[j] == kSF) continue;
        if (meas[j] > 0) {
            L.enter(meas[j], j);
            continue;
        }
        cf[j] = kZ;
        for (int32_t k = SI[j]; k < SI[j + 1]; ++k) {
            const int32_t nb = SJ[k];
            if (cf[nb] == kSF) continue;
            if (nb < j) {
                if (meas[nb] > 0) L.remove(meas[nb], nb);
                L.enter(++meas[nb], nb);
            } else {
                ++meas[nb];
            }
        }
        --left;
    }
    // an undecided point gains one for each point that depends on it and turned F
    auto bump_dependencies = [&](int32_t f) {
        for (int32_t k = SI[f]; k < SI[f + 1]; ++k) {
            const int32_t q = SJ[k];
            if (cf[q] != kUndecided) continue;
            L.remove(meas[q], q);
            L.enter(++meas[q], q);
        }
    };
    while (left > 0) {
        const int32_t c = L.first();
        if (c < 0) break;
        cf[c] = kC;
        L.remove(meas[c], c);
        meas[c] = 0;
        --left;
        for (int32_t p = ti[c]; p < ti[c + 1]; ++p) {
            const int32_t f = tj[p];
            if (cf[f] != kUndecided) continue;
            cf[f] = kF;
            L.remove(meas[f], f);
            --left;
            bump_dependencies(f);
        }
        for (int32_t p = SI[c]; p < SI[c + 1]; ++p) {
            const int32_t q = SJ[p];
            if (cf[q] != kUndecided) continue;
            L.remove(meas[q], q);
            if (--meas[q] > 0) {
                L.enter(meas[q], q);
            } else {
                cf[q] = kF;
                --left;
                bump_dependencies(q);
            }
        }
    }
}

// flags an S whose row pointer descends or whose columns leave [0, n): the
// host sweep indexes with them (plain stores of 1, idempotent)
__global__ __launch_bounds__(256) void k_check_s(int32_t n, const int32_t *__restrict__ SI,
                                                 const int32_t *__restrict__ SJ,
                                                 int32_t *__restrict__ bad) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int p0 = SI[i], p1 = SI[i + 1];
    bool ok = p1 >= p0;
    for (int p = p0; ok && p < p1; ++p) ok = uint32_t(SJ[p]) < uint32_t(n);
    if (!ok) *bad = 1;
}

// pinned staging for S and the markers, grown as needed, kept for the process
struct Staging {
    std::mutex mu;
    void *buf = nullptr;
    size_t bytes = 0;
    int get(size_t need, void **out) {
        if (need > bytes) {
            if (buf) (void)hipHostFree(buf);
            buf = nullptr;
            bytes = 0;
            const size_t sz = need + need / 4;
            AMGT_HIP(hipHostMalloc(&buf, sz, hipHostMallocDefault));
            bytes = sz;
        }
        *out = buf;
        return AMGT_OK;
    }
};

Staging &staging() {
    static Staging s;
    return s;
}

}  // namespace

}  // namespace amgt

using namespace amgt;

extern "C" {

int amgt_amg_ruge_first_pass(int32_t n, const int32_t *S_i, const int32_t *S_j,
                             int32_t *cf_marker) {
    AMGT_CHECK_ARG(n >= 0, "negative size");
    AMGT_CHECK_ARG(n == 0 || (S_i && cf_marker), "NULL row pointer or marker array");
    if (n == 0) return AMGT_OK;
    AMGT_CHECK_ARG(S_i[n] == 0 || S_j, "NULL column array");
    for (int32_t i = 0; i < n; ++i)
        if (S_i[i + 1] < S_i[i]) {
            set_error("Ruge first pass: row pointer not ascending");
            return AMGT_ERR_INPUT;
        }
    for (int64_t p = 0; p < S_i[n]; ++p)
        if (S_j[p] < 0 || S_j[p] >= n) {
            set_error("Ruge first pass: column out of range");
            return AMGT_ERR_INPUT;
        }
    std::vector<int32_t> ti, tj;
    transpose_host(n, S_i, S_j, ti, tj);
    ruge_first_pass(n, S_i, S_j, ti.data(), tj.data(), cf_marker);
    return AMGT_OK;
}

int amgt_amg_hmis(int32_t n, const int32_t *S_i, const int32_t *S_j, int32_t *cf_marker,
                  int32_t *n_coarse, void *stream) {
    AMGT_CHECK_ARG(n_coarse, "NULL output argument");
    AMGT_CHECK_ARG(n >= 0, "negative size");
    AMGT_CHECK_ARG(n == 0 || (S_i && cf_marker), "NULL row pointer or marker array");
    *n_coarse = 0;
    if (n == 0) return AMGT_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Staging &st = staging();
    std::lock_guard<std::mutex> lock(st.mu);
    int32_t nnz = 0;
    {
        const Readback rb[1] = {{&nnz, S_i + n, sizeof(int32_t)}};
        AMGT_TRY(readback(rb, 1, s));
    }
    AMGT_CHECK_ARG(nnz == 0 || S_j, "NULL column array");
    AMGT_CHECK_ARG(nnz >= 0, "negative nonzero count in S");
    // S^T on the device (rows ascending, as the host sweep reads them)
    int32_t *dti = nullptr, *dtj = nullptr, *dbad = nullptr, hbad = 0;
    AMGT_TRY(tmp_alloc_t(&dti, size_t(n) + 1, s));
    int status = tmp_alloc_t(&dtj, size_t(nnz), s);
    if (status == AMGT_OK) status = tmp_alloc_t(&dbad, 1, s);
    if (status == AMGT_OK) {
        if (hipMemsetAsync(dbad, 0, sizeof(int32_t), s) != hipSuccess) status = AMGT_ERR_HIP;
        hipLaunchKernelGGL(k_check_s, dim3((n + 255) / 256), dim3(256), 0, s, n, S_i, S_j, dbad);
        if (hipGetLastError() != hipSuccess) status = AMGT_ERR_HIP;
    }
    if (status == AMGT_OK)
        status = csr_transpose_device(n, n, S_i, S_j, nullptr, nnz, dti, dtj, nullptr, s);
    void *buf = nullptr;
    if (status == AMGT_OK)
        status = st.get(sizeof(int32_t) * (3 * size_t(n) + 3 + 2 * size_t(nnz)), &buf);
    int32_t *hi = static_cast<int32_t *>(buf);
    int32_t *hj = hi + n + 1;
    int32_t *hti = hj + nnz;
    int32_t *htj = hti + n + 1;
    int32_t *hcf = htj + nnz;
    int32_t *hbad_p = hcf + n;  // the device check's flag, in the pinned area too
    if (status == AMGT_OK) {
        hipError_t e = hipMemcpyAsync(hi, S_i, sizeof(int32_t) * (size_t(n) + 1),
                                      hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(hti, dti, sizeof(int32_t) * (size_t(n) + 1), hipMemcpyDeviceToHost,
                               s);
        if (e == hipSuccess && nnz)
            e = hipMemcpyAsync(hj, S_j, sizeof(int32_t) * size_t(nnz), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && nnz)
            e = hipMemcpyAsync(htj, dtj, sizeof(int32_t) * size_t(nnz), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(hbad_p, dbad, sizeof(int32_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) hbad = *hbad_p;
        if (e != hipSuccess) {
            set_error(std::string("HMIS: copy of S to the host: ") + hipGetErrorString(e));
            status = AMGT_ERR_HIP;
        }
    }
    tmp_free(dti, s);
    tmp_free(dtj, s);
    tmp_free(dbad, s);
    AMGT_TRY(status);
    if (hbad) {
        set_error("HMIS: S row pointer not ascending or column out of range");
        return AMGT_ERR_INPUT;
    }
    ruge_first_pass(n, hi, hj, hti, htj, hcf);
    AMGT_HIP(hipMemcpyAsync(cf_marker, hcf, sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice, s));
    // amgt_amg_pmis synchronises the stream before it returns successfully; on
    // an error the copy out of the staging buffer is waited for here, so the
    // buffer is free again when the lock is released
    status = amgt_amg_pmis(n, S_i, S_j, 1, cf_marker, n_coarse, stream);
    if (status != AMGT_OK) (void)hipStreamSynchronize(s);
    return status;
}

}  // extern "C"
