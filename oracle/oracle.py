"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module, and only as the checker.  The product library never calls
it.  The C restatements live in oracle.c; each cites the reference file:line
it follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32 = ctypes.c_int
_i64 = ctypes.c_int64
_f64 = ctypes.c_double


def build() -> str:
    """Compiles oracle.c with gcc (make) into oracle/liboracle.so."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_pm_seed": (ctypes.c_int32, [_i64]),
            "orc_pm_next": (ctypes.c_int32, [ctypes.POINTER(ctypes.c_int32)]),
            "orc_rand_uniform": (None, [_i64, _i64, _f64p]),
            "orc_rand_normal": (None, [_i64, _i64, _f64p]),
            "orc_stencil_nnz": (_i64, [_i32, _i32, _i32, _i32]),
            "orc_stencil": (_i32, [_i32, _i32, _i32, _i32, _f64, _f64, _f64, _i32p, _i32p, _f64p]),
            "orc_bitmap_mul": (ctypes.c_uint16, [ctypes.c_uint16, ctypes.c_uint16]),
            "orc_csr2mbsr_count": (_i64, [_i32, _i32, _i32p, _i32p, _i32p]),
            "orc_csr2mbsr_fill": (_i32, [_i32, _i32, _i32p, _i32p, _f64p, _i32p, _i32p, _u16p,
                                         _i32p, _f64p]),
            "orc_mbsr_dense_values": (None, [_i32, _i32p, _u16p, _i32p, _f64p, _f64p]),
            "orc_mbsr_stats": (None, [_i32, _i32p, _i64, ctypes.POINTER(_f64),
                                      ctypes.POINTER(_f64)]),
            "orc_mbsr2csr_count": (_i64, [_i32, _i32, _i32p, _u16p, _i32p]),
            "orc_mbsr2csr_fill": (None, [_i32, _i32, _i32p, _i32p, _u16p, _i32p, _f64p, _i32p,
                                         _i32p, _f64p]),
            "orc_mbsr_spmv": (None, [_i32, _i32, _i32p, _i32p, _u16p, _i32p, _f64p, _f64, _f64p,
                                     _f64, _f64p, _f64p]),
            "orc_mbsr_spmv_t": (None, [_i32, _i32, _i32, _i32p, _i32p, _u16p, _i32p, _f64p, _f64,
                                       _f64p, _f64, _f64p, _f64p]),
            "orc_csr_matvec": (None, [_i32, _i32p, _i32p, _f64p, _f64, _f64p, _f64, _f64p, _f64p]),
            "orc_csr_matvec_omp": (None, [_i32, _i32, _i32p, _i32p, _f64p, _f64, _f64p, _f64,
                                          _f64p, _f64p]),
            "orc_csr_matvec_t": (None, [_i32, _i32, _i32p, _i32p, _f64p, _f64, _f64p, _f64, _f64p,
                                        _f64p]),
            "orc_csr_spgemm_count": (_i64, [_i32, _i32, _i32p, _i32p, _i32p, _i32p, _i32, _i32p]),
            "orc_csr_spgemm_fill": (None, [_i32, _i32, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p,
                                           _i32, _i32p, _i32p, _f64p]),
            "orc_csr_spgemm_nnz": (_i64, [_i32, _i32, _i32p, _i32p, _i32p, _i32p, _i32]),
            "orc_mbsr_spgemm_count": (_i64, [_i32, _i32p, _i32p, _u16p, _i32, _i32p, _i32p, _u16p,
                                             _i32p]),
            "orc_mbsr_spgemm_fill": (None, [_i32, _i32p, _i32p, _u16p, _i32p, _f64p,
                                            _i32, _i32p, _i32p, _u16p, _i32p, _f64p,
                                            _i32p, _i32p, _u16p, _i32p, _f64p]),
            "orc_amg_strength": (_i64, [_i32, _i32p, _i32p, _f64p, _f64, _f64, _i32p, _i32p]),
            "orc_amg_pmis": (_i32, [_i32, _i32p, _i32p, _i32, _i32p]),
            "orc_amg_ruge_first_pass": (None, [_i32, _i32p, _i32p, _i32p]),
            "orc_amg_hmis": (_i32, [_i32, _i32p, _i32p, _i32p]),
            "orc_amg_pmis_measures": (None, [_i32, _i32p, _i32p, _f64p]),
            "orc_amg_extpi_count": (_i64, [_i32, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]),
            "orc_amg_extpi_fill": (None, [_i32, _i32p, _i32p, _f64p, _i32p, _i32p, _i32p, _i32p,
                                          _i32p, _f64p]),
            "orc_amg_truncate": (_i64, [_i32, _i32p, _i32p, _f64p, _i32]),
            "orc_amg_rap_kt_count": (_i64, [_i32, _i32, _i32p, _i32p, _i32p, _i32p, _i32p]),
            "orc_amg_rap_kt_fill": (None, [_i32, _i32, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p,
                                           _i32p, _i32p, _f64p]),
            "orc_policy_forward": (None, [_f64p, _f64p, _i64, _f64p]),
            "orc_policy_loss_grad": (_f64, [_f64p, _f64p, _i32p, _f64p, _i64, _f64, _f64p]),
            "orc_policy_adam": (None, [_f64p, _f64p, _f64p, _f64p, _i64, _i64, _f64, _f64, _f64,
                                       _f64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _cf(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# RNG (src/utilities/random.c:43-101, src/seq_mv/vector.c:395-418)

def pm_sequence(seed: int, n: int) -> np.ndarray:
    L = lib()
    st = ctypes.c_int32(L.orc_pm_seed(seed))
    out = np.empty(n, dtype=np.int64)
    for i in range(n):
        out[i] = L.orc_pm_next(ctypes.byref(st))
    return out


def rand_uniform(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().orc_rand_uniform(seed, n, out)
    return out


def rand_normal(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().orc_rand_normal(seed, n, out)
    return out


# --------------------------------------------------------------------------
# stencils (src/parcsr_ls/par_laplace*.c)

def stencil(kind: int, nx: int, ny: int, nz: int = 1, c=(1.0, 1.0, 1.0)):
    """CSR (I, J, A) of a HYPRE stencil matrix; rows are diagonal-first."""
    L = lib()
    nnz = L.orc_stencil_nnz(kind, nx, ny, nz)
    if nnz < 0:
        raise ValueError(f"unknown stencil kind {kind}")
    n = nx * ny * nz
    I = np.empty(n + 1, dtype=np.int32)
    J = np.empty(nnz, dtype=np.int32)
    A = np.empty(nnz, dtype=np.float64)
    rc = L.orc_stencil(kind, nx, ny, nz, c[0], c[1], c[2], I, J, A)
    if rc != 0:
        raise RuntimeError("orc_stencil failed")
    return I, J, A


def stencil_nnz(kind: int, nx: int, ny: int, nz: int = 1) -> int:
    return int(lib().orc_stencil_nnz(kind, nx, ny, nz))


# --------------------------------------------------------------------------
# mBSR

def bitmap_mul(a: int, b: int) -> int:
    return int(lib().orc_bitmap_mul(a, b))


def csr2mbsr(nrows: int, ncols: int, I, J, A):
    """Compressed mBSR of a CSR matrix: dict(blk_ptr, blk_col, blk_map, val_ptr, val)."""
    L = lib()
    I, J, A = _c32(I), _c32(J), _cf(A)
    brow = (nrows + 3) // 4
    blk_ptr = np.empty(brow + 1, dtype=np.int32)
    nblk = L.orc_csr2mbsr_count(nrows, ncols, I, J, blk_ptr)
    if nblk == -1:
        raise ValueError("duplicate (row, col) entry")
    if nblk == -2:
        raise ValueError("column index out of range")
    blk_col = np.empty(max(nblk, 1), dtype=np.int32)
    blk_map = np.empty(max(nblk, 1), dtype=np.uint16)
    val_ptr = np.empty(brow + 1, dtype=np.int32)
    val = np.empty(max(len(J), 1), dtype=np.float64)
    rc = L.orc_csr2mbsr_fill(nrows, ncols, I, J, A, blk_ptr, blk_col, blk_map, val_ptr, val)
    if rc != 0:
        raise RuntimeError(f"orc_csr2mbsr_fill rc={rc}")
    return dict(nrows=nrows, ncols=ncols, brow=brow, bcol=(ncols + 3) // 4, nblk=int(nblk),
                nnz=len(J), blk_ptr=blk_ptr, blk_col=blk_col[:nblk], blk_map=blk_map[:nblk],
                val_ptr=val_ptr, val=val[:len(J)])


def mbsr_dense_values(M) -> np.ndarray:
    out = np.zeros(max(M["nblk"], 1) * 16, dtype=np.float64)
    lib().orc_mbsr_dense_values(M["brow"], M["blk_ptr"], _u16(M["blk_map"]), M["val_ptr"],
                                _cf(M["val"]) if M["nnz"] else np.zeros(1), out)
    return out[: M["nblk"] * 16]


def _u16(a):
    a = np.ascontiguousarray(a, dtype=np.uint16)
    return a if a.size else np.zeros(1, dtype=np.uint16)


def _nz32(a):
    a = _c32(a)
    return a if a.size else np.zeros(1, dtype=np.int32)


def _nzf(a):
    a = _cf(a)
    return a if a.size else np.zeros(1, dtype=np.float64)


def mbsr_stats(M):
    s = ctypes.c_double()
    a = ctypes.c_double()
    lib().orc_mbsr_stats(M["brow"], M["blk_ptr"], M["nnz"], ctypes.byref(s), ctypes.byref(a))
    return s.value, a.value


def mbsr2csr(M):
    L = lib()
    I = np.empty(M["nrows"] + 1, dtype=np.int32)
    nnz = L.orc_mbsr2csr_count(M["nrows"], M["brow"], M["blk_ptr"], _u16(M["blk_map"]), I)
    J = np.empty(max(nnz, 1), dtype=np.int32)
    A = np.empty(max(nnz, 1), dtype=np.float64)
    L.orc_mbsr2csr_fill(M["nrows"], M["brow"], M["blk_ptr"], _nz32(M["blk_col"]),
                        _u16(M["blk_map"]), M["val_ptr"], _nzf(M["val"]), I, J, A)
    return I, J[:nnz], A[:nnz]


def mbsr_spmv(M, alpha, x, beta=0.0, b=None):
    y = np.empty(M["nrows"], dtype=np.float64)
    bb = _cf(b) if b is not None else np.zeros(max(M["nrows"], 1))
    lib().orc_mbsr_spmv(M["nrows"], M["brow"], M["blk_ptr"], _nz32(M["blk_col"]),
                        _u16(M["blk_map"]), M["val_ptr"], _nzf(M["val"]), alpha,
                        _nzf(np.asarray(x, dtype=np.float64)), beta, bb, y)
    return y


def mbsr_spmv_t(M, alpha, x, beta=0.0, b=None):
    y = np.empty(M["ncols"], dtype=np.float64)
    bb = _cf(b) if b is not None else np.zeros(max(M["ncols"], 1))
    lib().orc_mbsr_spmv_t(M["nrows"], M["ncols"], M["brow"], M["blk_ptr"], _nz32(M["blk_col"]),
                          _u16(M["blk_map"]), M["val_ptr"], _nzf(M["val"]), alpha,
                          _nzf(np.asarray(x, dtype=np.float64)), beta, bb, y)
    return y


# --------------------------------------------------------------------------
# HYPRE host CSR kernels

def csr_matvec(nrows, I, J, A, alpha, x, beta=0.0, b=None):
    y = np.empty(nrows, dtype=np.float64)
    bb = _cf(b) if b is not None else np.zeros(max(nrows, 1))
    lib().orc_csr_matvec(nrows, _c32(I), _nz32(J), _nzf(A), alpha, _nzf(x), beta, bb, y)
    return y


def csr_matvec_omp(nthreads, nrows, I, J, A, alpha, x, beta=0.0, b=None, out=None):
    """csr_matvec with the rows split over nthreads OpenMP threads."""
    y = np.empty(nrows, np.float64) if out is None else out
    bb = _cf(b) if b is not None else np.zeros(max(nrows, 1))
    lib().orc_csr_matvec_omp(int(nthreads), nrows, _c32(I), _c32(J), _cf(A), float(alpha),
                             _cf(x), float(beta), bb, y)
    return y


def csr_matvec_t(nrows, ncols, I, J, A, alpha, x, beta=0.0, b=None):
    y = np.empty(ncols, dtype=np.float64)
    bb = _cf(b) if b is not None else np.zeros(max(ncols, 1))
    lib().orc_csr_matvec_t(nrows, ncols, _c32(I), _nz32(J), _nzf(A), alpha, _nzf(x), beta, bb, y)
    return y


def csr_spgemm(m, n, AI, AJ, AA, BI, BJ, BA, diag_first=None):
    """HYPRE host SpGEMM; diag_first defaults to HYPRE's rule (C square)."""
    if diag_first is None:
        diag_first = int(m == n)
    L = lib()
    CI = np.empty(m + 1, dtype=np.int32)
    nnz = L.orc_csr_spgemm_count(m, n, _c32(AI), _nz32(AJ), _c32(BI), _nz32(BJ), diag_first, CI)
    CJ = np.empty(max(nnz, 1), dtype=np.int32)
    CA = np.empty(max(nnz, 1), dtype=np.float64)
    L.orc_csr_spgemm_fill(m, n, _c32(AI), _nz32(AJ), _nzf(AA), _c32(BI), _nz32(BJ), _nzf(BA),
                          diag_first, CI, CJ, CA)
    return CI, CJ[:nnz], CA[:nnz]


def csr_spgemm_nnz(m, n, AI, AJ, BI, BJ, diag_first=None) -> int:
    if diag_first is None:
        diag_first = int(m == n)
    return int(lib().orc_csr_spgemm_nnz(m, n, _c32(AI), _nz32(AJ), _c32(BI), _nz32(BJ),
                                        diag_first))


def mbsr_spgemm(MA, MB):
    """Block SpGEMM on compressed mBSR operands; returns the C mBSR dict."""
    L = lib()
    brow = MA["brow"]
    cp = np.empty(brow + 1, dtype=np.int32)
    nblk = L.orc_mbsr_spgemm_count(brow, MA["blk_ptr"], _nz32(MA["blk_col"]), _u16(MA["blk_map"]),
                                   MB["bcol"], MB["blk_ptr"], _nz32(MB["blk_col"]),
                                   _u16(MB["blk_map"]), cp)
    cc = np.empty(max(nblk, 1), dtype=np.int32)
    cm = np.empty(max(nblk, 1), dtype=np.uint16)
    cvp = np.empty(brow + 1, dtype=np.int32)
    cv = np.empty(max(16 * nblk, 1), dtype=np.float64)
    L.orc_mbsr_spgemm_fill(brow, MA["blk_ptr"], _nz32(MA["blk_col"]), _u16(MA["blk_map"]),
                           MA["val_ptr"], _nzf(MA["val"]),
                           MB["bcol"], MB["blk_ptr"], _nz32(MB["blk_col"]), _u16(MB["blk_map"]),
                           MB["val_ptr"], _nzf(MB["val"]), cp, cc, cm, cvp, cv)
    nnz = int(cvp[brow])
    return dict(nrows=MA["nrows"], ncols=MB["ncols"], brow=brow, bcol=MB["bcol"], nblk=int(nblk),
                nnz=nnz, blk_ptr=cp, blk_col=cc[:nblk], blk_map=cm[:nblk], val_ptr=cvp,
                val=cv[:nnz])


# --------------------------------------------------------------------------
# BoomerAMG setup on one process (par_strength.c, par_coarsen.c,
# par_lr_interp.c, par_csr_matrix.c truncation)

def amg_strength(n, I, J, A, theta=0.25, max_row_sum=1.0):
    """S (SI, SJ): strong columns per row in row order, no diagonal."""
    I, J, A = _c32(I), _c32(J), _cf(A)
    SI = np.zeros(n + 1, np.int32)
    SJ = np.zeros(max(len(J), 1), np.int32)
    nnz = lib().orc_amg_strength(n, I, J, A, float(theta), float(max_row_sum), SI, SJ)
    return SI, SJ[:nnz].copy()


def amg_pmis_measures(n, SI, SJ):
    m = np.zeros(max(n, 1), np.float64)
    lib().orc_amg_pmis_measures(n, _c32(SI), _c32(SJ), m)
    return m[:n]


def amg_pmis(n, SI, SJ, cf_init=0, cf_in=None):
    """CF marker: 1 C, -1 F, -3 no strong dependence; and the C count.
    cf_init 1 starts from cf_in, the markers of amg_ruge_first_pass."""
    cf = np.zeros(max(n, 1), np.int32)
    if cf_init == 1:
        cf[:n] = np.asarray(cf_in, np.int32)
    nc = lib().orc_amg_pmis(n, _c32(SI), _c32(SJ), int(cf_init), cf)
    return cf[:n], int(nc)


def amg_ruge_first_pass(n, SI, SJ):
    """Markers of the Ruge-Stueben first pass (coarsen type 10): 1 C, -1 F,
    -2 Z (influences no point), -3 no strong dependence."""
    cf = np.zeros(max(n, 1), np.int32)
    lib().orc_amg_ruge_first_pass(n, _c32(SI), _c32(SJ), cf)
    return cf[:n]


def amg_hmis(n, SI, SJ):
    """HMIS: the Ruge first pass, then PMIS with CF_init 1."""
    cf = np.zeros(max(n, 1), np.int32)
    nc = lib().orc_amg_hmis(n, _c32(SI), _c32(SJ), cf)
    return cf[:n], int(nc)


def amg_extpi(n, I, J, A, SI, SJ, cf, max_elmts=0):
    """Extended+i P (PI, PJ, PA, ncoarse), row entries in HYPRE's order,
    truncated to max_elmts per row when max_elmts > 0."""
    I, J, A, SI, SJ, cf = _c32(I), _c32(J), _cf(A), _c32(SI), _c32(SJ), _c32(cf)
    PI = np.zeros(n + 1, np.int32)
    nnz = lib().orc_amg_extpi_count(n, I, J, SI, SJ, cf, PI)
    PJ = np.zeros(max(nnz, 1), np.int32)
    PA = np.zeros(max(nnz, 1), np.float64)
    lib().orc_amg_extpi_fill(n, I, J, A, SI, SJ, cf, PI, PJ, PA)
    if max_elmts > 0:
        nnz = lib().orc_amg_truncate(n, PI, PJ, PA, int(max_elmts))
    return PI, PJ[:nnz].copy(), PA[:nnz].copy(), int((cf >= 0).sum())


def csr_transpose(m, n, I, J, A):
    """Rows of the transpose in ascending row order, entries in column order
    of the input (a counting sort)."""
    I, J, A = np.asarray(I), np.asarray(J), np.asarray(A)
    cnt = np.bincount(J, minlength=n)
    TI = np.zeros(n + 1, np.int64)
    np.cumsum(cnt, out=TI[1:])
    order = np.argsort(J, kind="stable")
    rows = np.repeat(np.arange(m), np.diff(I))
    return TI.astype(np.int32), rows[order].astype(np.int32), A[order].copy()


def amg_rap_kt(n, nc, I, J, A, PI, PJ, PA):
    """P^T A P in HYPRE's host order ((R A) P, first-touch slots, diagonal
    first): (RI, RJ, RA)."""
    I, J, A, PI, PJ, PA = _c32(I), _c32(J), _cf(A), _c32(PI), _c32(PJ), _cf(PA)
    RI = np.zeros(nc + 1, np.int32)
    nnz = lib().orc_amg_rap_kt_count(n, nc, I, J, PI, PJ, RI)
    RJ = np.zeros(max(nnz, 1), np.int32)
    RA = np.zeros(max(nnz, 1), np.float64)
    lib().orc_amg_rap_kt_fill(n, nc, I, J, A, PI, PJ, PA, RI, RJ, RA)
    return RI, RJ[:nnz].copy(), RA[:nnz].copy()


def amg_hierarchy(n, I, J, A, theta=0.25, max_row_sum=1.0, cf_init=0, max_elmts=0,
                  max_coarse_size=9, max_levels=25, coarsen="pmis", rap="two_step"):
    """The one-process BoomerAMG setup loop (par_amg_setup.c:994-3182) with
    strength -> PMIS (coarsen types 8/9, CF_init 0/2) or HMIS (coarsen type
    10, coarsen="hmis") -> extended+i -> R = P^T -> R (A P) (host SpGEMM, the
    diagonal first).  Returns the levels as dicts (n, I, J, A, cf, P...)."""
    levels = [{"n": n, "I": _c32(I), "J": _c32(J), "A": _cf(A)}]
    while True:
        L = levels[-1]
        if L["n"] <= max_coarse_size or len(levels) >= max_levels:
            break
        SI, SJ = amg_strength(L["n"], L["I"], L["J"], L["A"], theta, max_row_sum)
        if coarsen == "hmis":
            cf, nc = amg_hmis(L["n"], SI, SJ)
        else:
            cf, nc = amg_pmis(L["n"], SI, SJ, cf_init)
        if nc == 0 or nc == L["n"]:
            break
        PI, PJ, PA, nc = amg_extpi(L["n"], L["I"], L["J"], L["A"], SI, SJ, cf, max_elmts)
        if rap == "kt":
            CI, CJ, CA = amg_rap_kt(L["n"], nc, L["I"], L["J"], L["A"], PI, PJ, PA)
        else:
            RI, RJ, RA = csr_transpose(L["n"], nc, PI, PJ, PA)
            API, APJ, APA = csr_spgemm(L["n"], nc, L["I"], L["J"], L["A"], PI, PJ, PA, False)
            CI, CJ, CA = csr_spgemm(nc, nc, RI, RJ, RA, API, APJ, APA, True)
        L.update(S=(SI, SJ), cf=cf, P=(PI, PJ, PA), nc=nc)
        levels.append({"n": nc, "I": CI, "J": CJ, "A": CA})
    return levels


# ---- per-level policy network ---------------------------------------------------------------
POLICY_PARAMS = 5252


def policy_forward(params, X):
    X = _cf(X).reshape(-1, 12)
    out = np.empty((X.shape[0], 4))
    lib().orc_policy_forward(_cf(params), X, X.shape[0], out)
    return out


def policy_loss_grad(params, X, cls, target, lam):
    X = _cf(X).reshape(-1, 12)
    grad = np.empty(POLICY_PARAMS)
    loss = lib().orc_policy_loss_grad(_cf(params), X, _c32(cls), _cf(target), X.shape[0],
                                      float(lam), grad)
    return loss, grad


def policy_adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """In place on params, m, v (float64 arrays)."""
    lib().orc_policy_adam(params, _cf(grad), m, v, params.size, int(step), lr, b1, b2, eps)
