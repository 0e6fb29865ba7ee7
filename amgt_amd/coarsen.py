"""BoomerAMG setup on the device: strength, PMIS and HMIS coarsening and
extended+i interpolation (libamgt_mbsr.so's amgt_amg_* entry points).

Counterparts of hypre_BoomerAMGCreateS (par_strength.c:531, device version
par_strength_device.c:34), hypre_BoomerAMGCoarsenPMIS (par_coarsen.c:2816,
device par_coarsen_device.c:30) and hypre_BoomerAMGBuildExtPIInterp
(par_lr_interp.c:5553, device par_lr_interp_device.c:991), with HYPRE's
argument meaning (strong threshold, max row sum, CF_init, P max elements).
Every array is a CSR on the device (int32 indices, float64 values); there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib as L
from .mbsr import _stream_ptr, copy_from_device_ptr


def _p(t: torch.Tensor | None) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr() if t is not None and t.numel() else 0)


def _check_csr(n: int, I: torch.Tensor, J: torch.Tensor, A: torch.Tensor | None = None):
    if I.dtype != torch.int32 or J.dtype != torch.int32:
        raise TypeError("CSR indices must be int32")
    if A is not None and A.dtype != torch.float64:
        raise TypeError("CSR values must be float64")
    if I.numel() != n + 1:
        raise ValueError(f"row pointer has {I.numel()} entries, expected {n + 1}")
    for t in (I, J) + ((A,) if A is not None else ()):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("CSR arrays must be contiguous device tensors")


def strength(n: int, I: torch.Tensor, J: torch.Tensor, A: torch.Tensor, theta: float = 0.25,
             max_row_sum: float = 0.9, stream=None):
    """S (S_i, S_j): strong dependencies of every row (par_strength.c:75)."""
    _check_csr(n, I, J, A)
    lib = L.load()
    nnz = ctypes.c_int64()
    si, sj = ctypes.c_void_p(), ctypes.c_void_p()
    L.check(lib.amgt_amg_strength(n, _p(I), _p(J), _p(A), float(theta), float(max_row_sum),
                                  ctypes.byref(nnz), ctypes.byref(si), ctypes.byref(sj),
                                  _stream_ptr(stream)), "amgt_amg_strength")
    SI = copy_from_device_ptr(si.value, n + 1, torch.int32, I.device, stream)
    SJ = copy_from_device_ptr(sj.value, nnz.value, torch.int32, I.device, stream)
    return SI, SJ


def pmis(n: int, SI: torch.Tensor, SJ: torch.Tensor, cf_init: int = 0, stream=None):
    """C/F splitting (1 C, -1 F, -3 no strong dependence) and the C count
    (par_coarsen.c:2101; CF_init 0 or 2)."""
    if cf_init not in (0, 2):
        raise ValueError("pmis: CF_init 0 or 2 (use hmis for coarsen type 10)")
    _check_csr(n, SI, SJ)
    cf = torch.empty(max(n, 1), dtype=torch.int32, device=SI.device)
    nc = ctypes.c_int32()
    L.check(L.load().amgt_amg_pmis(n, _p(SI), _p(SJ), int(cf_init), _p(cf), ctypes.byref(nc),
                                   _stream_ptr(stream)), "amgt_amg_pmis")
    return cf[:n], int(nc.value)


def ruge_first_pass(n: int, SI, SJ):
    """Markers of the Ruge-Stueben first pass as HMIS runs it (coarsen type 10,
    par_coarsen.c:911): 1 C, -1 F, -2 Z, -3 no strong dependence.  Host
    arrays (numpy int32) in and out."""
    import numpy as np
    SI = np.ascontiguousarray(SI, dtype=np.int32)
    SJ = np.ascontiguousarray(SJ, dtype=np.int32)
    if SI.size != n + 1:
        raise ValueError(f"row pointer has {SI.size} entries, expected {n + 1}")
    if SJ.size < int(SI[n]):
        raise ValueError("column array shorter than the row pointer says")
    cf = np.zeros(max(n, 1), np.int32)
    vp = lambda a: ctypes.c_void_p(a.ctypes.data if a.size else 0)  # noqa: E731
    L.check(L.load().amgt_amg_ruge_first_pass(n, vp(SI), vp(SJ), vp(cf)),
            "amgt_amg_ruge_first_pass")
    return cf[:n]


def hmis(n: int, SI: torch.Tensor, SJ: torch.Tensor, stream=None):
    """HMIS C/F splitting (coarsen type 10, par_coarsen.c:2845): the Ruge first
    pass on the host, PMIS with CF_init 1 on the device.  (cf, C count)."""
    _check_csr(n, SI, SJ)
    cf = torch.empty(max(n, 1), dtype=torch.int32, device=SI.device)
    nc = ctypes.c_int32()
    L.check(L.load().amgt_amg_hmis(n, _p(SI), _p(SJ), _p(cf), ctypes.byref(nc),
                                   _stream_ptr(stream)), "amgt_amg_hmis")
    return cf[:n], int(nc.value)


def extpi_interp(n: int, I: torch.Tensor, J: torch.Tensor, A: torch.Tensor, SI: torch.Tensor,
                 SJ: torch.Tensor, cf: torch.Tensor, max_elmts: int = 0, stream=None):
    """Extended+i P (P_i, P_j, P_a, n_coarse), truncated to max_elmts per row
    when positive (par_lr_interp.c:1025, par_csr_matrix.c:2866)."""
    _check_csr(n, I, J, A)
    _check_csr(n, SI, SJ)
    if cf.dtype != torch.int32 or cf.numel() < n:
        raise ValueError("cf must hold n int32 markers")
    lib = L.load()
    nc = ctypes.c_int32()
    nnz = ctypes.c_int64()
    pi, pj, pa = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    L.check(lib.amgt_amg_extpi_interp(n, _p(I), _p(J), _p(A), _p(SI), _p(SJ), _p(cf),
                                      int(max_elmts), ctypes.byref(nc), ctypes.byref(nnz),
                                      ctypes.byref(pi), ctypes.byref(pj), ctypes.byref(pa),
                                      _stream_ptr(stream)), "amgt_amg_extpi_interp")
    PI = copy_from_device_ptr(pi.value, n + 1, torch.int32, I.device, stream)
    PJ = copy_from_device_ptr(pj.value, nnz.value, torch.int32, I.device, stream)
    PA = copy_from_device_ptr(pa.value, nnz.value, torch.float64, I.device, stream)
    return PI, PJ, PA, int(nc.value)


def rap_host_order(n: int, nc: int, I: torch.Tensor, J: torch.Tensor, A: torch.Tensor,
                   PI: torch.Tensor, PJ: torch.Tensor, PA: torch.Tensor, stream=None):
    """R A P with R = P^T in HYPRE's host order (par_rap.c:1525-1930): the
    coarse operator (RAP_i, RAP_j, RAP_a) bit for bit as HYPRE's CPU setup
    forms it, rows diagonal-first in HYPRE's layout."""
    _check_csr(n, I, J, A)
    _check_csr(n, PI, PJ, PA)
    lib = L.load()
    nnz = ctypes.c_int64()
    ci, cj, ca = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    L.check(lib.amgt_amg_rap_host_order(n, nc, _p(I), _p(J), _p(A), _p(PI), _p(PJ), _p(PA),
                                        ctypes.byref(nnz), ctypes.byref(ci), ctypes.byref(cj),
                                        ctypes.byref(ca), _stream_ptr(stream)),
            "amgt_amg_rap_host_order")
    CI = copy_from_device_ptr(ci.value, nc + 1, torch.int32, I.device, stream)
    CJ = copy_from_device_ptr(cj.value, nnz.value, torch.int32, I.device, stream)
    CA = copy_from_device_ptr(ca.value, nnz.value, torch.float64, I.device, stream)
    return CI, CJ, CA
