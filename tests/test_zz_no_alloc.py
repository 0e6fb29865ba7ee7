"""amgt_mbsr_spmv_noalloc, the SpMV entry point that allocates nothing,
against amgt_mbsr_spmv with the balanced kernel and FP64 values, both through
the C-ABI: equal bit for bit on two calls in a row (the second on a smaller
matrix, so the shared device array still holds the first call's larger
tables), for every tile size (lanes 16 / 4 / 8: tiles of 256 / 512 / 1024
blocks), alpha/beta forms and b aliasing y; plus its argument and size-limit
errors, which return before any launch."""
import ctypes
import functools

import numpy as np
import pytest
import torch

from amgt_amd import _lib as L
from amgt_amd.mbsr import MBSRMatrix
from oracle import oracle as O
from tests.util_matrices import skewed_csr

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def vp(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@functools.lru_cache(maxsize=1)
def big_and_small():
    """A skewed matrix (long block rows cut by tile edges) and a smaller 7-pt
    stencil."""
    m, n = 11999, 20000
    I, J, A = skewed_csr(m, n, (2, 500, 998, 2000, 2900), 2000, 5, 11,
                         empty_rows=(0, 7, 2001, 11998))
    Is, Js, As = O.stencil(7, 17, 13, 11)
    ns = 17 * 13 * 11
    return [((m, n), I, J, A), ((ns, ns), Is, Js, As)]


def existing(M, alpha, x, beta, b, y):
    st = L.load().amgt_mbsr_spmv(M.handle, 0, alpha, vp(x), beta, vp(b), vp(y), stream())
    L.check(st, "amgt_mbsr_spmv")


def noalloc(M, alpha, x, beta, b, y):
    return L.load().amgt_mbsr_spmv_noalloc(M.handle, alpha, vp(x), beta, vp(b), vp(y), stream())


@pytest.mark.parametrize("lanes", [16, 4, 8])
@pytest.mark.parametrize("form", ["axpby", "beta0", "b_is_y"])
def test_noalloc_matches_balanced_spmv(dev, lanes, form):
    print(f"[noalloc] lanes={lanes} form={form}", flush=True)
    mats = []
    for (m, n), I, J, A in big_and_small():
        M = MBSRMatrix.from_csr(T(I, dev), T(J, dev), T(A, dev), m, n)
        M.set_spmv_lanes(lanes)
        M.set_spmv_kernel(L.AMGT_SPMV_BALANCED)
        M.set_precision(L.AMGT_DOUBLE)
        rng = np.random.default_rng(lanes)
        x = T(rng.standard_normal(n), dev)
        b = T(rng.standard_normal(m), dev)
        mats.append((M, m, x, b))
    alpha, beta = (-1.5, 0.75) if form != "beta0" else (2.0, 0.0)
    refs = []
    for M, m, x, b in mats:
        y = torch.full((m,), 7.0, dtype=torch.float64, device=dev)
        if form == "b_is_y":
            y.copy_(b)
            existing(M, alpha, x, beta, y, y)
        else:
            existing(M, alpha, x, beta, b if form == "axpby" else None, y)
        refs.append(y)
    # two calls in a row, the second on the smaller matrix
    outs = []
    for M, m, x, b in mats:
        y = torch.full((m,), 7.0, dtype=torch.float64, device=dev)
        if form == "b_is_y":
            y.copy_(b)
            st = noalloc(M, alpha, x, beta, y, y)
        else:
            st = noalloc(M, alpha, x, beta, b if form == "axpby" else None, y)
        L.check(st, "amgt_mbsr_spmv_noalloc")
        outs.append(y)
    torch.cuda.synchronize()
    for k, (y, r) in enumerate(zip(outs, refs)):
        assert torch.equal(y, r), (k, (y - r).abs().max().item())
    # the balanced path really ran: a long block row spans several tiles
    assert mats[0][0].info()["max_row_blocks"] > 256 * 4


def test_noalloc_errors_return_before_launch(dev):
    (m, n), I, J, A = big_and_small()[1]
    M = MBSRMatrix.from_csr(T(I, dev), T(J, dev), T(A, dev), m, n)
    x = torch.zeros(n + 2, dtype=torch.float64, device=dev)
    y = torch.zeros(m, dtype=torch.float64, device=dev)
    lib = L.load()
    # x not 16-byte aligned
    st = lib.amgt_mbsr_spmv_noalloc(M.handle, 1.0, ctypes.c_void_p(x.data_ptr() + 8), 0.0, None,
                                    vp(y), stream())
    assert st == L.AMGT_ERR_ARG and b"aligned" in lib.amgt_last_error()
    # beta != 0 without b
    assert noalloc(M, 1.0, x, 1.0, None, y) == L.AMGT_ERR_ARG
    # x aliasing y
    assert noalloc(M, 1.0, y, 0.0, None, y) == L.AMGT_ERR_ARG


def test_noalloc_rejects_matrices_beyond_the_array(dev):
    """8 blocks (one entry each) per block row, 3.6M block rows: 28.8M blocks,
    112,500 tiles of 256 blocks at lanes 16 -- more than the 8 MiB array holds;
    at lanes 8 (tiles of 1024 blocks) the same matrix fits and matches."""
    brow = 3_600_000
    per = 8
    n = 4 * brow
    I = np.zeros(n + 1, np.int64)
    I[1::4] = per
    I = np.cumsum(I).astype(np.int32)
    bcols = (np.arange(brow, dtype=np.int64)[:, None] * 7 + np.arange(per) * (brow // per)) % brow
    bcols.sort(axis=1)
    J = (4 * bcols).astype(np.int32).ravel()
    A = np.linspace(-1.0, 1.0, len(J))
    M = MBSRMatrix.from_csr(T(I, dev), T(J, dev), T(A, dev), n, n)
    assert M.info()["nblk"] == brow * per
    M.set_spmv_kernel(L.AMGT_SPMV_BALANCED)
    M.set_precision(L.AMGT_DOUBLE)
    x = torch.linspace(-2.0, 3.0, n, dtype=torch.float64, device=dev)
    y = torch.full((n,), 5.0, dtype=torch.float64, device=dev)
    lib = L.load()
    M.set_spmv_lanes(16)
    st = noalloc(M, 1.0, x, 0.0, None, y)
    assert st == L.AMGT_ERR_UNSUPPORTED
    assert b"AMGT_NOALLOC_SCRATCH_BYTES" in lib.amgt_last_error()
    torch.cuda.synchronize()
    assert bool((y == 5.0).all())
    M.set_spmv_lanes(8)
    L.check(noalloc(M, 1.0, x, 0.0, None, y), "amgt_mbsr_spmv_noalloc")
    r = torch.empty_like(y)
    existing(M, 1.0, x, 0.0, None, r)
    torch.cuda.synchronize()
    assert torch.equal(y, r)
