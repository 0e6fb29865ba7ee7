"""Python host side of the mBSR path over the C-ABI (torch supplies device memory
and the stream; all compute runs in libamgt_mbsr.so).

``MBSRMatrix`` owns one library handle (the bsrMAT of seq_mv.h:105-129).
Operator names mirroring AmgT's functions (CSR2BSR_GPU, spmv_amgT_fp64, ...)
live in ``amgt_amd.hypre`` next to the HYPRE-shaped boundary.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib as L


def _stream_ptr(stream=None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _need_rows(t, name):
    """(k, n) float64 device tensor whose rows are contiguous (any row stride)."""
    if t is None:
        return
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
        raise TypeError(f"{name} must be a float64 device tensor")
    if t.dim() != 2 or (t.shape[1] > 1 and t.stride(1) != 1) or t.stride(0) < t.shape[1]:
        raise ValueError(f"{name} must be (k, n) with contiguous rows")


def _need_cuda(t, name, dtype=None):
    if t is None:
        return
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a device tensor")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def _dev_key(d: torch.device):
    if d.type == "cuda" and d.index is None:
        return d.type, torch.cuda.current_device()
    return d.type, d.index


def _need_len(t, name, n, device):
    """t (already type-checked) is on `device` and holds at least n entries
    (the kernels index n of them): a shorter or foreign tensor would be read or
    written out of bounds, so it is refused before any launch."""
    if t is None:
        return
    if _dev_key(t.device) != _dev_key(device):
        raise ValueError(f"{name} is on {t.device}, the matrix on {device}")
    if t.numel() < n:
        raise ValueError(f"{name} holds {t.numel()} entries, {n} needed")


def _need_rows_len(t, name, k, n, device):
    if t is None:
        return
    if _dev_key(t.device) != _dev_key(device):
        raise ValueError(f"{name} is on {t.device}, the matrix on {device}")
    if t.shape[0] != k or (k > 0 and t.shape[1] < n):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, ({k}, >= {n}) needed")


def copy_from_device_ptr(ptr: int, n: int, dtype, device, stream=None, free=True):
    """Copies n elements at a library-owned device pointer into a new tensor and
    (optionally) frees the pointer through the library allocator."""
    out = torch.empty(n, dtype=dtype, device=device)
    sp = _stream_ptr(stream)
    if n and ptr:
        nbytes = n * out.element_size()
        rc = L.hip().hipMemcpyAsync(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ptr), nbytes,
                                    L.HIP_MEMCPY_D2D, sp)
        if rc != 0:
            raise RuntimeError(f"hipMemcpyAsync failed ({rc})")
    if free and ptr:
        torch.cuda.current_stream().synchronize() if stream is None else stream.synchronize()
        L.check(L.load().amgt_free(ctypes.c_void_p(ptr)), "amgt_free")
    return out


class MBSRMatrix:
    """A device mBSR matrix (4x4 blocks, 16-bit occupancy bitmaps, compressed values)."""

    def __init__(self, handle: ctypes.c_void_p, device: torch.device):
        self._h = handle
        self.device = device

    # -- construction -----------------------------------------------------------------
    @classmethod
    def from_csr(cls, rowptr: torch.Tensor, colidx: torch.Tensor, values, nrows: int, ncols: int,
                 stream=None) -> "MBSRMatrix":
        """CSR2BSR_GPU (csr_matvec_device.c:1201-1267).  values may be None (pattern only)."""
        _need_cuda(rowptr, "rowptr", torch.int32)
        _need_cuda(colidx, "colidx", torch.int32)
        _need_cuda(values, "values", torch.float64)
        h = ctypes.c_void_p()
        st = L.load().amgt_mbsr_create_from_csr(nrows, ncols, colidx.numel(), _ptr(rowptr),
                                                _ptr(colidx), _ptr(values), _stream_ptr(stream),
                                                ctypes.byref(h))
        L.check(st, "amgt_mbsr_create_from_csr")
        return cls(h, rowptr.device)

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise ValueError("matrix already destroyed")
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            L.load().amgt_mbsr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- properties -----------------------------------------------------------------------
    def info(self) -> dict:
        inf = L.MbsrInfo()
        L.check(L.load().amgt_mbsr_get_info(self.handle, ctypes.byref(inf)), "amgt_mbsr_get_info")
        return {f: getattr(inf, f) for f, _ in L.MbsrInfo._fields_}

    @property
    def shape(self):
        i = self.info()
        return (i["nrows"], i["ncols"])

    def arrays(self, stream=None) -> dict:
        """Device copies of blk_ptr, blk_col, blk_map, val_ptr, val."""
        inf = self.info()
        arr = L.MbsrArrays()
        L.check(L.load().amgt_mbsr_get_arrays(self.handle, ctypes.byref(arr)), "amgt_mbsr_get_arrays")
        dev = self.device
        out = {
            "blk_ptr": copy_from_device_ptr(arr.blk_ptr, inf["brow"] + 1, torch.int32, dev, stream, False),
            "blk_col": copy_from_device_ptr(arr.blk_col, inf["nblk"], torch.int32, dev, stream, False),
            "blk_map": copy_from_device_ptr(arr.blk_map, inf["nblk"], torch.int16, dev, stream, False),
            "val_ptr": copy_from_device_ptr(arr.val_ptr, inf["brow"] + 1, torch.int32, dev, stream, False),
            "val": copy_from_device_ptr(arr.val, inf["nnz"] if arr.val else 0, torch.float64, dev,
                                        stream, False),
        }
        return out

    def set_level(self, level: int):
        L.check(L.load().amgt_mbsr_set_level(self.handle, level), "amgt_mbsr_set_level")

    def set_precision(self, precision: int):
        L.check(L.load().amgt_mbsr_set_precision(self.handle, precision), "amgt_mbsr_set_precision")

    def set_spmv_lanes(self, lanes: int):
        L.check(L.load().amgt_mbsr_set_spmv_lanes(self.handle, lanes), "amgt_mbsr_set_spmv_lanes")

    def set_spmv_kernel(self, kind: int):
        """AMGT_SPMV_DIRECT / _STAGED / _BLOCK / _BALANCED."""
        L.check(L.load().amgt_mbsr_set_spmv_kernel(self.handle, int(kind)),
                "amgt_mbsr_set_spmv_kernel")

    def autotune(self, stream=None):
        L.check(L.load().amgt_mbsr_autotune(self.handle, _stream_ptr(stream)), "amgt_mbsr_autotune")

    def tune_from_stats(self):
        """SpMV kernel kind and lanes from the matrix statistics
        (amgt_mbsr_tune_from_stats; what a new handle starts with)."""
        L.check(L.load().amgt_mbsr_tune_from_stats(self.handle), "amgt_mbsr_tune_from_stats")

    def spmv_config(self) -> tuple:
        """(kernel kind name, lanes) the SpMV runs with."""
        inf = self.info()
        return L.SPMV_KERNEL_NAMES[inf["spmv_kernel"]], inf["spmv_lanes"]

    def spmv_sweep(self, stream=None) -> dict:
        """Measured SpMV time (ms) of every (kernel kind, lanes) pair, the
        handle unchanged: {(kind name, lanes): ms}."""
        t = (ctypes.c_float * 20)()
        L.check(L.load().amgt_mbsr_spmv_sweep(self.handle, _stream_ptr(stream), t),
                "amgt_mbsr_spmv_sweep")
        kinds = (L.AMGT_SPMV_BLOCK, L.AMGT_SPMV_BALANCED, L.AMGT_SPMV_STAGED, L.AMGT_SPMV_DIRECT)
        out = {}
        for k, kind in enumerate(kinds):
            for c, lanes in enumerate((4, 8, 16, 32, 64)):
                if t[k * 5 + c] < 1e29:
                    out[(L.SPMV_KERNEL_NAMES[kind], lanes)] = float(t[k * 5 + c])
        return out

    def update_values(self, rowptr, colidx, values, stream=None):
        L.check(L.load().amgt_mbsr_update_values(self.handle, _ptr(rowptr), _ptr(colidx),
                                                 _ptr(values), _stream_ptr(stream)),
                "amgt_mbsr_update_values")

    # -- products ----------------------------------------------------------------------------
    def spmv(self, x: torch.Tensor, y: torch.Tensor | None = None, alpha: float = 1.0,
             beta: float = 0.0, b: torch.Tensor | None = None, trans: bool = False,
             stream=None) -> torch.Tensor:
        """y = alpha*op(A)*x + beta*b (b defaults to y)."""
        nrows, ncols = self.shape
        ny = ncols if trans else nrows
        nx = nrows if trans else ncols
        _need_cuda(x, "x", torch.float64)
        if y is None:
            y = torch.empty(ny, dtype=torch.float64, device=self.device)
        _need_cuda(y, "y", torch.float64)
        if b is None and beta != 0.0:
            b = y
        _need_cuda(b, "b", torch.float64)
        _need_len(x, "x", nx, self.device)
        _need_len(y, "y", ny, self.device)
        if beta != 0.0:
            _need_len(b, "b", ny, self.device)
        st = L.load().amgt_mbsr_spmv(self.handle, int(bool(trans)), float(alpha), _ptr(x),
                                     float(beta), _ptr(b), _ptr(y), _stream_ptr(stream))
        L.check(st, "amgt_mbsr_spmv")
        return y

    def diagonal(self, inverse_abs: bool = False, out: torch.Tensor | None = None,
                 stream=None, kind: int | None = None) -> torch.Tensor:
        """a_ii; 1/|a_ii| with inverse_abs; or the given amgt_mbsr_diagonal kind
        (L.AMGT_DIAG_INV_L1: inverse l1 norms of relax type 18, L.AMGT_DIAG_INV:
        inverse diagonal of relax type 7)."""
        n = min(self.shape)
        if out is None:
            out = torch.empty(n, dtype=torch.float64, device=self.device)
        _need_cuda(out, "out", torch.float64)
        _need_len(out, "out", n, self.device)
        if kind is None:
            kind = L.AMGT_DIAG_INV_ABS if inverse_abs else L.AMGT_DIAG_VALUE
        L.check(L.load().amgt_mbsr_diagonal(self.handle, kind, _ptr(out), _stream_ptr(stream)),
                "amgt_mbsr_diagonal")
        return out

    def relax_jacobi(self, f: torch.Tensor, u: torch.Tensor, dinv: torch.Tensor,
                     omega: float = 1.0, out: torch.Tensor | None = None,
                     stream=None) -> torch.Tensor:
        """One weighted Jacobi sweep u + omega * dinv * (f - A u) (relax type 18), one launch."""
        n = self.shape[0]
        for t, name in ((f, "f"), (u, "u"), (dinv, "dinv")):
            _need_cuda(t, name, torch.float64)
        if out is None:
            out = torch.empty(n, dtype=torch.float64, device=self.device)
        _need_cuda(out, "out", torch.float64)
        for t, name in ((f, "f"), (u, "u"), (dinv, "dinv"), (out, "out")):
            _need_len(t, name, n, self.device)
        L.check(L.load().amgt_mbsr_relax_jacobi(self.handle, float(omega), _ptr(dinv), _ptr(f),
                                                 _ptr(u), _ptr(out), _stream_ptr(stream)),
                "amgt_mbsr_relax_jacobi")
        return out

    def spmm(self, X: torch.Tensor, Y: torch.Tensor | None = None, alpha: float = 1.0,
             beta: float = 0.0, B: torch.Tensor | None = None, trans: bool = False,
             stream=None) -> torch.Tensor:
        """Y = alpha*op(A)*X + beta*B for k columns.  X is (k, n) with each
        right-hand side contiguous and any row stride (HYPRE multivector
        storage method 0, vecstride = the row stride)."""
        nrows, ncols = self.shape
        ny = ncols if trans else nrows
        _need_rows(X, "X")
        k = X.shape[0]
        if Y is None:
            Y = torch.empty((k, ny), dtype=torch.float64, device=self.device)
        _need_rows(Y, "Y")
        if B is None and beta != 0.0:
            B = Y
        _need_rows(B, "B")
        _need_rows_len(X, "X", k, nrows if trans else ncols, self.device)
        _need_rows_len(Y, "Y", k, ny, self.device)
        if beta != 0.0:
            _need_rows_len(B, "B", k, ny, self.device)
        st = L.load().amgt_mbsr_spmm(self.handle, int(bool(trans)), k, float(alpha), _ptr(X),
                                     X.stride(0), float(beta), _ptr(B),
                                     B.stride(0) if B is not None else 0, _ptr(Y), Y.stride(0),
                                     _stream_ptr(stream))
        L.check(st, "amgt_mbsr_spmm")
        return Y

    def spmm_interleaved(self, X: torch.Tensor, Y: torch.Tensor | None = None,
                         alpha: float = 1.0, beta: float = 0.0, B: torch.Tensor | None = None,
                         trans: bool = False, stream=None) -> torch.Tensor:
        """Y = alpha*op(A)*X + beta*B with X of shape (n, k), contiguous: the k
        values of a row together (HYPRE multivector storage method 1)."""
        nrows, ncols = self.shape
        ny = ncols if trans else nrows
        _need_cuda(X, "X", torch.float64)
        if X.dim() != 2:
            raise ValueError("X must be (n, k)")
        k = X.shape[1]
        if Y is None:
            Y = torch.empty((ny, k), dtype=torch.float64, device=self.device)
        _need_cuda(Y, "Y", torch.float64)
        if B is None and beta != 0.0:
            B = Y
        _need_cuda(B, "B", torch.float64)
        nx = nrows if trans else ncols
        for t, name, rows in ((X, "X", nx), (Y, "Y", ny), (B if beta != 0.0 else None, "B", ny)):
            if t is None:
                continue
            if t.dim() != 2 or t.shape[1] != k or t.shape[0] < rows:
                raise ValueError(f"{name} has shape {tuple(t.shape)}, (>= {rows}, {k}) needed")
            _need_len(t, name, rows * k, self.device)
        st = L.load().amgt_mbsr_spmm_interleaved(self.handle, int(bool(trans)), k, float(alpha),
                                                 _ptr(X), float(beta), _ptr(B), _ptr(Y),
                                                 _stream_ptr(stream))
        L.check(st, "amgt_mbsr_spmm_interleaved")
        return Y

    def spgemm(self, other: "MBSRMatrix", stream=None) -> "MBSRMatrix":
        h = ctypes.c_void_p()
        st = L.load().amgt_mbsr_spgemm(self.handle, other.handle, _stream_ptr(stream), ctypes.byref(h))
        L.check(st, "amgt_mbsr_spgemm")
        return MBSRMatrix(h, self.device)

    def residual_restrict(self, plan: "RestrictPlan", f: torch.Tensor, u: torch.Tensor,
                          r: torch.Tensor, fc: torch.Tensor, stream=None) -> torch.Tensor:
        """r = f - A u and fc = R r in one launch (amgt_mbsr_residual_restrict);
        `plan` comes from RestrictPlan(A, R).  Returns fc."""
        n = self.shape[0]
        for t, name in ((f, "f"), (u, "u"), (r, "r"), (fc, "fc")):
            _need_cuda(t, name, torch.float64)
        for t, name in ((f, "f"), (u, "u"), (r, "r")):
            _need_len(t, name, n, self.device)
        _need_len(fc, "fc", plan.n_coarse, self.device)
        st = L.load().amgt_mbsr_residual_restrict(self.handle, plan.handle, _ptr(f), _ptr(u),
                                                  _ptr(r), _ptr(fc), _stream_ptr(stream))
        L.check(st, "amgt_mbsr_residual_restrict")
        return fc

    def transpose(self, stream=None) -> "MBSRMatrix":
        """A new handle holding A^T (amgt_mbsr_transpose)."""
        h = ctypes.c_void_p()
        st = L.load().amgt_mbsr_transpose(self.handle, _stream_ptr(stream), ctypes.byref(h))
        L.check(st, "amgt_mbsr_transpose")
        return MBSRMatrix(h, self.device)

    def __matmul__(self, other):
        if isinstance(other, MBSRMatrix):
            return self.spgemm(other)
        return self.spmv(other)

    def to_csr(self, stream=None):
        """(rowptr, colidx, values) with ascending columns per row; values None if
        the matrix is pattern-only."""
        nnz = ctypes.c_int64()
        pi, pj, pv = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        st = L.load().amgt_mbsr_to_csr(self.handle, _stream_ptr(stream), ctypes.byref(nnz),
                                       ctypes.byref(pi), ctypes.byref(pj), ctypes.byref(pv))
        L.check(st, "amgt_mbsr_to_csr")
        nrows = self.info()["nrows"]
        I = copy_from_device_ptr(pi.value, nrows + 1, torch.int32, self.device, stream)
        J = copy_from_device_ptr(pj.value, nnz.value, torch.int32, self.device, stream)
        V = (copy_from_device_ptr(pv.value, nnz.value, torch.float64, self.device, stream)
             if pv.value else None)
        return I, J, V


def set_mixed_precision(enable: bool) -> None:
    L.check(L.load().amgt_set_mixed_precision(int(bool(enable))), "amgt_set_mixed_precision")


def precision_for_level(level: int) -> int:
    return int(L.load().amgt_precision_for_level(level))


class RestrictPlan:
    """Tables of the one-launch residual + restriction of a V-cycle level
    (amgt_restrict_plan_create): A (square) and R = P^T as mBSR handles; built
    once for A's lanes setting."""

    def __init__(self, A: MBSRMatrix, R: MBSRMatrix, stream=None):
        h = ctypes.c_void_p()
        L.check(L.load().amgt_restrict_plan_create(A.handle, R.handle, _stream_ptr(stream),
                                                   ctypes.byref(h)), "amgt_restrict_plan_create")
        self._h = h
        self.n_coarse = R.shape[0]

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise ValueError("plan already destroyed")
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            L.load().amgt_restrict_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
