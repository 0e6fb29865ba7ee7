"""Matrix files HYPRE's drivers read (IJ per-rank files, Matrix Market),
through the library's host readers (amgt_read_ij / amgt_read_mm).  Returns
host CSR arrays (rows ascending, columns ascending within a row)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L


def _take(n, ptr, ctype, dtype):
    if n == 0 or not ptr:
        return np.zeros(n, dtype=dtype)
    arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(n,)).copy()
    return arr.astype(dtype, copy=False)


def _read(fn, *args):
    nr, nc = ctypes.c_int32(), ctypes.c_int32()
    nnz = ctypes.c_int64()
    pi, pj, pv = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    st = fn(*args, ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(nnz), ctypes.byref(pi),
            ctypes.byref(pj), ctypes.byref(pv))
    lib = L.load()
    try:
        L.check(st, fn.__name__)
        I = _take(nr.value + 1, pi.value, ctypes.c_int32, np.int32)
        J = _take(nnz.value, pj.value, ctypes.c_int32, np.int32)
        V = _take(nnz.value, pv.value, ctypes.c_double, np.float64)
    finally:
        for p in (pi, pj, pv):
            if p.value:
                lib.amgt_host_free(p)
    return nr.value, nc.value, I, J, V


def read_ij(prefix: str, nfiles: int = 1):
    """HYPRE IJ matrix files <prefix>.00000 ... (hypre_IJMatrixRead semantics)."""
    return _read(L.load().amgt_read_ij, prefix.encode(), int(nfiles))


def read_mm(path: str):
    """Matrix Market coordinate real/integer, general or symmetric."""
    return _read(L.load().amgt_read_mm, path.encode())


def write_ij(prefix: str, nrows: int, ncols: int, I, J, V, nfiles: int = 1):
    """Writes a CSR as nfiles IJ files (rows split evenly), the format of
    HYPRE_IJMatrixPrint / hypre_ParCSRMatrixPrintIJ with base 0."""
    bounds = np.linspace(0, nrows, nfiles + 1).astype(np.int64)
    for f in range(nfiles):
        r0, r1 = int(bounds[f]), int(bounds[f + 1])
        with open(f"{prefix}.{f:05d}", "w") as fh:
            fh.write(f"{r0} {r1 - 1} 0 {ncols - 1}\n")
            for r in range(r0, r1):
                for p in range(I[r], I[r + 1]):
                    fh.write(f"{r} {J[p]} {V[p]:.17e}\n")
