"""HYPRE-shaped host interface of the mBSR path.

Mirrors the two seq_mv entry points AmgT takes over and the AmgT functions
behind them, with HYPRE's argument meaning:

* ``hypreDevice_CSRSpGemm(A, B) -> C``      csr_spgemm_device.c:2579-2608 (seq_mv.h:727)
* ``hypre_CSRMatrixMatvecDevice(trans, alpha, A, x, beta, b, y, offset)``
                                            csr_matvec_device.c:128-191 (seq_mv.h:562)
* ``CSR2BSR_GPU(A)``                         csr_matvec_device.c:1201-1267
* ``spmv_amgT_fp64/fp32/fp16(...)``          csr_matvec_device.c:1302-1411, 2205-2309, 2098-2203
* ``spgemm_amgT_fp64(A, B) -> C``            csr_spgemm_device.c:1527-1846
* ``get_calculationPrecison(level)``         csr_spgemm_device.c:1847-1862

``CSRMatrix`` carries the fields of hypre_CSRMatrix the path reads
(i, j, data, num_rows, num_cols, pattern_only) plus AmgT's cache fields
(bsr_mat, coasen_level; seq_mv.h:156-163).  Errors: HYPRE returns its error
flag and keeps going; here a non-zero library status raises AmgtError, and the
functions return HYPRE's flag value (0) on success.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib as L
from .mbsr import MBSRMatrix, _need_cuda, _ptr, _stream_ptr, copy_from_device_ptr

AMGT_DOUBLE = L.AMGT_DOUBLE
AMGT_FLOAT = L.AMGT_FLOAT
AMGT_HALF = L.AMGT_HALF


class CSRMatrix:
    """Device CSR matrix with AmgT's lazily built mBSR cache."""

    def __init__(self, i: torch.Tensor, j: torch.Tensor, data: torch.Tensor | None,
                 num_rows: int, num_cols: int):
        _need_cuda(i, "i", torch.int32)
        _need_cuda(j, "j", torch.int32)
        _need_cuda(data, "data", torch.float64)
        self.i = i
        self.j = j
        self.data = data
        self.num_rows = int(num_rows)
        self.num_cols = int(num_cols)
        self.coasen_level = 0          # hypre_Level
        self.bsr_mat = ctypes.c_void_p()  # hypre_BSR cache slot (amgt_mbsr_t)

    @property
    def num_nonzeros(self) -> int:
        return int(self.j.numel())

    @property
    def pattern_only(self) -> bool:
        return self.data is None

    def bsr(self) -> MBSRMatrix | None:
        """A non-owning view of the cached mBSR (None before the first use)."""
        if not self.bsr_mat:
            return None
        m = MBSRMatrix.__new__(MBSRMatrix)
        m._h = self.bsr_mat
        m.device = self.i.device
        m.close = lambda: None  # the CSRMatrix owns the handle
        return m

    def invalidate(self):
        """Drops the mBSR cache (call after changing i/j/data)."""
        if self.bsr_mat:
            torch.cuda.synchronize()
            L.load().amgt_mbsr_destroy(self.bsr_mat)
            self.bsr_mat = ctypes.c_void_p()

    def __del__(self):
        try:
            self.invalidate()
        except Exception:
            pass


def get_calculationPrecison(level: int) -> int:  # noqa: N802 (reference spelling)
    return int(L.load().amgt_precision_for_level(level))


def CSR2BSR_GPU(A: CSRMatrix, stream=None) -> MBSRMatrix:  # noqa: N802
    """Builds (once) and returns the cached mBSR of A."""
    if not A.bsr_mat:
        h = ctypes.c_void_p()
        st = L.load().amgt_mbsr_create_from_csr(A.num_rows, A.num_cols, A.num_nonzeros,
                                                _ptr(A.i), _ptr(A.j), _ptr(A.data),
                                                _stream_ptr(stream), ctypes.byref(h))
        L.check(st, "CSR2BSR_GPU")
        A.bsr_mat = h
        L.check(L.load().amgt_mbsr_set_level(h, A.coasen_level), "amgt_mbsr_set_level")
    return A.bsr()


def hypreDevice_CSRSpGemm(A: CSRMatrix, B: CSRMatrix, stream=None) -> CSRMatrix:  # noqa: N802
    """C = A*B.  C owns new device arrays; C's mBSR is cached on C."""
    if A.num_cols != B.num_rows:
        raise L.AmgtError(L.AMGT_ERR_DIM, "hypreDevice_CSRSpGemm", "incompatible matrix dimensions")
    nnzC = ctypes.c_int64()
    ci, cj, cd = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    cache = ctypes.c_void_p()
    for M in (A, B):
        if M.bsr_mat:
            L.load().amgt_mbsr_set_level(M.bsr_mat, M.coasen_level)
    st = L.load().amgt_csr_spgemm(
        A.num_rows, A.num_cols, B.num_cols,
        A.num_nonzeros, _ptr(A.i), _ptr(A.j), _ptr(A.data), ctypes.byref(A.bsr_mat),
        B.num_nonzeros, _ptr(B.i), _ptr(B.j), _ptr(B.data), ctypes.byref(B.bsr_mat),
        ctypes.byref(nnzC), ctypes.byref(ci), ctypes.byref(cj), ctypes.byref(cd),
        ctypes.byref(cache), _stream_ptr(stream))
    L.check(st, "hypreDevice_CSRSpGemm")
    dev = A.i.device
    I = copy_from_device_ptr(ci.value, A.num_rows + 1, torch.int32, dev, stream)
    J = copy_from_device_ptr(cj.value, nnzC.value, torch.int32, dev, stream)
    D = copy_from_device_ptr(cd.value, nnzC.value, torch.float64, dev, stream) if cd.value else None
    C = CSRMatrix(I, J, D, A.num_rows, B.num_cols)
    C.bsr_mat = cache
    C.coasen_level = A.coasen_level
    return C


def spgemm_amgT_fp64(A: CSRMatrix, B: CSRMatrix, stream=None) -> CSRMatrix:  # noqa: N802
    return hypreDevice_CSRSpGemm(A, B, stream)


def hypre_CSRMatrixMatvecDevice(trans: int, alpha: float, A: CSRMatrix, x: torch.Tensor,  # noqa: N802
                                beta: float, b: torch.Tensor, y: torch.Tensor, offset: int = 0,
                                stream=None) -> int:
    """y = alpha*op(A)*x + beta*b.  offset must be 0 (csr_matvec_device.c:142)."""
    if offset != 0:
        raise L.AmgtError(L.AMGT_ERR_ARG, "hypre_CSRMatrixMatvecDevice", "offset must be 0")
    nx = A.num_rows if trans else A.num_cols
    ny = A.num_cols if trans else A.num_rows
    # csr_matvec_device.c:151-155: vectors may be longer than needed
    if x.numel() < nx or y.numel() < ny or (b is not None and b.numel() < ny):
        raise L.AmgtError(L.AMGT_ERR_DIM, "hypre_CSRMatrixMatvecDevice", "vector too short")
    if A.bsr_mat:
        L.load().amgt_mbsr_set_level(A.bsr_mat, A.coasen_level)
    st = L.load().amgt_csr_matvec(int(bool(trans)), float(alpha), A.num_rows, A.num_cols,
                                  A.num_nonzeros, _ptr(A.i), _ptr(A.j), _ptr(A.data),
                                  ctypes.byref(A.bsr_mat), _ptr(x), float(beta), _ptr(b), _ptr(y),
                                  _stream_ptr(stream))
    L.check(st, "hypre_CSRMatrixMatvecDevice")
    if A.bsr_mat:
        L.load().amgt_mbsr_set_level(A.bsr_mat, A.coasen_level)
    return 0


def _spmv_prec(precision, trans, alpha, A, x, beta, y, stream):
    M = CSR2BSR_GPU(A, stream)
    M.set_precision(precision)
    M.spmv(x, y, alpha=alpha, beta=beta, b=y, trans=bool(trans), stream=stream)
    M.set_precision(-1)
    return 0


def spmv_amgT_fp64(trans, alpha, A, x, beta, y, offset=0, stream=None):  # noqa: N802
    """y = alpha*A*x + beta*y in FP64 (csr_matvec_device.c:1302-1411)."""
    return _spmv_prec(AMGT_DOUBLE, trans, alpha, A, x, beta, y, stream)


def spmv_amgT_fp32(trans, alpha, A, x, beta, y, offset=0, stream=None):  # noqa: N802
    """FP32 values, FP32 accumulation; beta*y kept (the reference drops it, quirk Q2)."""
    return _spmv_prec(AMGT_FLOAT, trans, alpha, A, x, beta, y, stream)


def spmv_amgT_fp16(trans, alpha, A, x, beta, y, offset=0, stream=None):  # noqa: N802
    """FP16 values, FP32 accumulation (the reference accumulates in FP16, quirk Q3)."""
    return _spmv_prec(AMGT_HALF, trans, alpha, A, x, beta, y, stream)
