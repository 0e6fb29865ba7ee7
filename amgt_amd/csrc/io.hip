// Host readers for the on-disk matrix formats HYPRE's drivers use, so test
// problems and per-level operators can be loaded without HYPRE:
//  * IJ files (hypre_IJMatrixRead, IJ_mv/IJMatrix.c:112-242): one file per
//    rank, <prefix>.%05d, header "ilower iupper jlower jupper", then lines
//    "i j value" with 0-based global indices.  An entry inside the file's row
//    range is set (a later one for the same (i, j) replaces it,
//    HYPRE_IJMatrixSetValues); an entry outside it is added to the owner's
//    value (HYPRE_IJMatrixAddToValues).  All files of the set are merged.
//  * Matrix Market (the same function with is_mm = 1): coordinate real or
//    integer, general or symmetric (off-diagonal entries mirrored), 1-based.
// The result is a host CSR with ascending columns per row (malloc'd; free
// with amgt_host_free).  Host code only.
#include "internal.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace amgt {

namespace {

struct Triplet {
    int64_t i, j;
    double v;
    int add;  // 1: add to the owner's value, 0: set
};

bool read_file(const std::string &path, std::string *out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out->resize(n > 0 ? size_t(n) : 0);
    const size_t got = n > 0 ? std::fread(&(*out)[0], 1, size_t(n), f) : 0;
    std::fclose(f);
    return got == out->size();
}

// whitespace-separated tokens of a text buffer
struct Scanner {
    const char *p, *end;
    void skip_ws() {
        while (p < end && std::isspace(static_cast<unsigned char>(*p))) ++p;
    }
    bool integer(int64_t *v) {
        skip_ws();
        if (p >= end) return false;
        char *e = nullptr;
        const long long x = std::strtoll(p, &e, 10);
        if (e == p) return false;
        p = e;
        *v = x;
        return true;
    }
    bool real(double *v) {
        skip_ws();
        if (p >= end) return false;
        char *e = nullptr;
        const double x = std::strtod(p, &e);
        if (e == p) return false;
        p = e;
        *v = x;
        return true;
    }
    bool at_end() {
        skip_ws();
        return p >= end;
    }
};

// Builds the CSR from triplets: per (i, j) the last set value plus the sum of
// the added ones; rows ascending, columns ascending.
int assemble(int64_t nrows, int64_t ncols, std::vector<Triplet> &t, int32_t *out_nrows,
             int32_t *out_ncols, int64_t *out_nnz, int32_t **I, int32_t **J, double **V) {
    if (nrows > INT32_MAX || ncols > INT32_MAX) {
        set_error("matrix reader: dimensions exceed int32");
        return AMGT_ERR_UNSUPPORTED;
    }
    for (const Triplet &e : t) {
        if (e.i < 0 || e.i >= nrows || e.j < 0 || e.j >= ncols) {
            set_error("matrix reader: index out of range (" + std::to_string(e.i) + ", " +
                      std::to_string(e.j) + ")");
            return AMGT_ERR_INPUT;
        }
    }
    // stable: file order is kept among equal (i, j), so "last set wins"
    std::stable_sort(t.begin(), t.end(), [](const Triplet &a, const Triplet &b) {
        return a.i != b.i ? a.i < b.i : a.j < b.j;
    });
    std::vector<int64_t> ri;
    std::vector<int32_t> cj;
    std::vector<double> va;
    for (size_t a = 0; a < t.size();) {
        size_t b = a;
        double set = 0.0, add = 0.0;
        while (b < t.size() && t[b].i == t[a].i && t[b].j == t[a].j) {
            if (t[b].add) add += t[b].v;
            else set = t[b].v;
            ++b;
        }
        ri.push_back(t[a].i);
        cj.push_back(int32_t(t[a].j));
        va.push_back(set + add);
        a = b;
    }
    const int64_t nnz = int64_t(cj.size());
    if (nnz >= (int64_t(1) << 31)) {
        set_error("matrix reader: nnz exceeds int32 indexing");
        return AMGT_ERR_UNSUPPORTED;
    }
    int32_t *hI = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * size_t(nrows + 1)));
    int32_t *hJ = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * size_t(nnz ? nnz : 1)));
    double *hV = static_cast<double *>(std::malloc(sizeof(double) * size_t(nnz ? nnz : 1)));
    if (!hI || !hJ || !hV) {
        std::free(hI);
        std::free(hJ);
        std::free(hV);
        set_error("matrix reader: host allocation failed");
        return AMGT_ERR_ALLOC;
    }
    std::fill(hI, hI + nrows + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) ++hI[ri[size_t(k)] + 1];
    for (int64_t r = 0; r < nrows; ++r) hI[r + 1] += hI[r];
    std::copy(cj.begin(), cj.end(), hJ);
    std::copy(va.begin(), va.end(), hV);
    *out_nrows = int32_t(nrows);
    *out_ncols = int32_t(ncols);
    *out_nnz = nnz;
    *I = hI;
    *J = hJ;
    *V = hV;
    return AMGT_OK;
}

}  // namespace

int read_ij(const char *prefix, int nfiles, int32_t *nrows, int32_t *ncols, int64_t *nnz,
            int32_t **I, int32_t **J, double **V) {
    std::vector<Triplet> t;
    int64_t maxrow = -1, maxcol = -1;
    std::string buf;
    for (int f = 0; f < nfiles; ++f) {
        char suffix[16];
        std::snprintf(suffix, sizeof(suffix), ".%05d", f);
        const std::string path = std::string(prefix) + suffix;
        if (!read_file(path, &buf)) {
            set_error("IJ reader: cannot read " + path);
            return AMGT_ERR_ARG;
        }
        Scanner sc{buf.data(), buf.data() + buf.size()};
        int64_t ilo, ihi, jlo, jhi;
        if (!sc.integer(&ilo) || !sc.integer(&ihi) || !sc.integer(&jlo) || !sc.integer(&jhi)) {
            set_error("IJ reader: bad header in " + path);
            return AMGT_ERR_INPUT;
        }
        maxrow = std::max(maxrow, ihi);
        maxcol = std::max(maxcol, jhi);
        while (!sc.at_end()) {
            int64_t i, j;
            double v;
            if (!sc.integer(&i) || !sc.integer(&j) || !sc.real(&v)) {
                set_error("IJ reader: error in IJ matrix input file " + path);
                return AMGT_ERR_INPUT;
            }
            t.push_back({i, j, v, (i < ilo || i > ihi) ? 1 : 0});
        }
    }
    return assemble(maxrow + 1, maxcol + 1, t, nrows, ncols, nnz, I, J, V);
}

int read_mm(const char *path, int32_t *nrows, int32_t *ncols, int64_t *nnz, int32_t **I,
            int32_t **J, double **V) {
    std::string buf;
    if (!read_file(path, &buf)) {
        set_error(std::string("MM reader: cannot read ") + path);
        return AMGT_ERR_ARG;
    }
    // banner: %%MatrixMarket matrix coordinate <real|integer> <general|symmetric>
    const size_t eol = buf.find('\n');
    std::string banner = buf.substr(0, eol);
    for (char &c : banner) c = char(std::tolower(static_cast<unsigned char>(c)));
    if (banner.rfind("%%matrixmarket", 0) != 0 || banner.find("matrix") == std::string::npos ||
        banner.find("coordinate") == std::string::npos) {
        set_error("MM reader: not a Matrix Market coordinate file");
        return AMGT_ERR_INPUT;
    }
    if (banner.find("real") == std::string::npos && banner.find("integer") == std::string::npos) {
        set_error("MM reader: only sparse real-valued/integer coordinate matrices are supported");
        return AMGT_ERR_UNSUPPORTED;
    }
    const bool sym = banner.find("symmetric") != std::string::npos;
    if (!sym && banner.find("general") == std::string::npos) {
        set_error("MM reader: only general or symmetric matrices are supported");
        return AMGT_ERR_UNSUPPORTED;
    }
    // skip comment lines
    size_t pos = eol == std::string::npos ? buf.size() : eol + 1;
    while (pos < buf.size() && buf[pos] == '%') {
        const size_t e = buf.find('\n', pos);
        pos = e == std::string::npos ? buf.size() : e + 1;
    }
    Scanner sc{buf.data() + pos, buf.data() + buf.size()};
    int64_t m, n, nz;
    if (!sc.integer(&m) || !sc.integer(&n) || !sc.integer(&nz)) {
        set_error("MM read size error");
        return AMGT_ERR_INPUT;
    }
    std::vector<Triplet> t;
    t.reserve(size_t(sym ? 2 * nz : nz));
    for (int64_t k = 0; k < nz; ++k) {
        int64_t i, j;
        double v;
        if (!sc.integer(&i) || !sc.integer(&j) || !sc.real(&v)) {
            set_error("MM reader: error in entry " + std::to_string(k));
            return AMGT_ERR_INPUT;
        }
        // every entry of a single MM file is set (ilower..iupper covers all rows)
        t.push_back({i - 1, j - 1, v, 0});
        if (sym && i != j) t.push_back({j - 1, i - 1, v, 0});
    }
    return assemble(m, n, t, nrows, ncols, nnz, I, J, V);
}

}  // namespace amgt

using namespace amgt;

extern "C" {

int amgt_read_ij(const char *prefix, int nfiles, int32_t *nrows, int32_t *ncols, int64_t *nnz,
                 int32_t **I, int32_t **J, double **V) {
    AMGT_CHECK_ARG(prefix && nrows && ncols && nnz && I && J && V, "NULL argument");
    AMGT_CHECK_ARG(nfiles >= 1, "nfiles must be >= 1");
    *I = nullptr;
    *J = nullptr;
    *V = nullptr;
    return read_ij(prefix, nfiles, nrows, ncols, nnz, I, J, V);
}

int amgt_read_mm(const char *path, int32_t *nrows, int32_t *ncols, int64_t *nnz, int32_t **I,
                 int32_t **J, double **V) {
    AMGT_CHECK_ARG(path && nrows && ncols && nnz && I && J && V, "NULL argument");
    *I = nullptr;
    *J = nullptr;
    *V = nullptr;
    return read_mm(path, nrows, ncols, nnz, I, J, V);
}

void amgt_host_free(void *p) { std::free(p); }

}  // extern "C"
