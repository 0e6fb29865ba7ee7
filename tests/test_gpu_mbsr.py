"""GPU parity of CSR->mBSR conversion and mBSR SpMV/SpMM against the CPU oracle,
through the C-ABI (amgt_amd.mbsr / amgt_amd.hypre are thin ctypes layers).

Bars: integer/structure arrays and values moved by the conversion bit-exact;
FP64 products within 1e-12 relative to sum(|a_ij x_j|) + |beta b_i| per row;
FP32 / FP16 value storage within 2e-6 / 2e-3 of the same scale (FP32
accumulation; FP16 values round at 2^-11 relative).
"""
import numpy as np
import scipy.sparse as sp
import pytest
import torch

from amgt_amd import _lib as L
from amgt_amd import hypre as H
from amgt_amd.mbsr import MBSRMatrix
from oracle import oracle as O
from tests.util_matrices import block_count_csr, dense_rows_csr, rand_csr, skewed_csr

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def make(dev, m, n, I, J, A):
    return MBSRMatrix.from_csr(T(I, dev), T(J, dev), None if A is None else T(A, dev), m, n)


CASES = {
    "c1_5pt_128": lambda: ((16384, 16384), *O.stencil(7, 128, 128, 1)),
    "7pt_16": lambda: ((4096, 4096), *O.stencil(7, 16, 16, 16)),
    "7pt_aniso": lambda: ((6 * 7 * 9, 6 * 7 * 9), *O.stencil(7, 6, 7, 9, c=(1, 1, 0.01))),
    "27pt_10": lambda: ((1000, 1000), *O.stencil(27, 10, 10, 10)),
    "rand_rect": lambda: ((37, 53), *rand_csr(37, 53, 0.15, 1)),
    "rand_tall_empty": lambda: ((101, 9), *rand_csr(101, 9, 0.3, 2, empty_rows=(0, 5, 99, 100))),
    "rand_wide": lambda: ((13, 5000), *rand_csr(13, 5000, 0.05, 3)),
    "bin1_rows": lambda: ((45, 700), *dense_rows_csr(45, 700, 150, 4)),
    "bin2_rows": lambda: ((9, 9000), *dense_rows_csr(9, 9000, 1900, 5)),
    "one_row": lambda: ((1, 3), np.array([0, 2], np.int32), np.array([2, 0], np.int32),
                        np.array([1.5, -2.0])),
    # block rows 2, 500 and 998 hold ~8000 blocks each, the rest ~12: stand >> 12
    "skewed": lambda: ((3995, 40000), *skewed_csr(3995, 40000, (2, 500, 998), 2000, 3, 6,
                                                  empty_rows=(0, 7, 2001, 3994))),
    # block totals landing on multiples of 256 / 512 / 1024 (balanced tile edges)
    "tile_edges": lambda: _tile_edges(),
    # long runs of empty rows: whole SpMV tiles without a block
    "mostly_empty": lambda: ((3001, 2999), np.array([0] * 4 + [2] * 1500 + [3] * 1498, np.int32),
                             np.array([7, 2998, 1], np.int32), np.array([1.25, -3.0, 0.5])),
}


def _tile_edges():
    counts = [256, 0, 512, 0, 0, 100, 156, 1024, 1, 255, 0, 256, 2000, 48, 0, 0]
    m, I, J, A = block_count_csr(counts, 8192, 7)
    return (m, 8192), I, J, A


def check_structure(M, ref):
    a = M.arrays()
    assert torch.equal(a["blk_ptr"].cpu(), torch.from_numpy(ref["blk_ptr"]))
    assert torch.equal(a["blk_col"].cpu(), torch.from_numpy(ref["blk_col"]))
    assert np.array_equal(a["blk_map"].cpu().numpy().view(np.uint16), ref["blk_map"])
    assert torch.equal(a["val_ptr"].cpu(), torch.from_numpy(ref["val_ptr"]))
    if ref["nnz"]:
        assert np.array_equal(a["val"].cpu().numpy(), ref["val"])


@pytest.mark.parametrize("name", list(CASES))
def test_conversion_bit_exact(dev, name):
    print(f"[conversion] {name}", flush=True)
    (m, n), I, J, A = CASES[name]()
    ref = O.csr2mbsr(m, n, I, J, A)
    M = make(dev, m, n, I, J, A)
    inf = M.info()
    assert (inf["nrows"], inf["ncols"], inf["nblk"], inf["nnz"]) == (m, n, ref["nblk"], len(J))
    stand, avg = O.mbsr_stats(ref)
    assert inf["stand"] == pytest.approx(stand, rel=1e-12, abs=1e-12)
    assert inf["avg_nnz"] == pytest.approx(avg, rel=1e-15)
    assert inf["max_row_blocks"] == int(np.diff(ref["blk_ptr"]).max())
    check_structure(M, ref)


def test_conversion_pattern_only_and_empty(dev):
    (m, n), I, J, A = CASES["rand_rect"]()
    M = make(dev, m, n, I, J, None)
    ref = O.csr2mbsr(m, n, I, J, A)
    a = M.arrays()
    assert torch.equal(a["blk_col"].cpu(), torch.from_numpy(ref["blk_col"]))
    assert a["val"].numel() == 0
    # no nonzeros at all
    I0 = np.zeros(8, np.int32)
    M0 = make(dev, 7, 5, I0, np.zeros(0, np.int32), np.zeros(0))
    assert M0.info()["nblk"] == 0
    y = M0.spmv(T(np.ones(5), dev), alpha=2.0, beta=3.0, b=T(np.arange(7.0), dev))
    assert np.array_equal(y.cpu().numpy(), 3.0 * np.arange(7.0))


def test_conversion_errors(dev):
    I = np.array([0, 3], np.int32)
    with pytest.raises(L.AmgtError) as e:
        make(dev, 1, 8, I, np.array([1, 5, 1], np.int32), np.ones(3))
    assert e.value.status == L.AMGT_ERR_INPUT and "duplicate" in str(e.value)
    with pytest.raises(L.AmgtError) as e:
        make(dev, 1, 8, I, np.array([1, 5, 8], np.int32), np.ones(3))
    assert e.value.status == L.AMGT_ERR_INPUT and "out of range" in str(e.value)
    with pytest.raises(L.AmgtError) as e:
        I2, J2, A2 = dense_rows_csr(4, 9000, 2100, 6)  # 8400 entries in one block row
        make(dev, 4, 9000, I2, J2, A2)
    assert e.value.status == L.AMGT_ERR_UNSUPPORTED
    with pytest.raises(L.AmgtError) as e:
        make(dev, 1, 8, np.array([0, 5], np.int32), np.array([1, 2, 3], np.int32), np.ones(3))
    assert e.value.status == L.AMGT_ERR_INPUT


def row_scale(m, n, I, J, A, x, beta, b):
    import scipy.sparse as sp
    S = sp.csr_matrix((np.abs(A), J, I), shape=(m, n))
    return S @ np.abs(x) + abs(beta) * (np.abs(b) if b is not None else 0.0) + 1e-300


@pytest.mark.parametrize("name", ["c1_5pt_128", "7pt_aniso", "27pt_10", "rand_rect",
                                  "rand_tall_empty", "rand_wide", "bin1_rows", "bin2_rows",
                                  "skewed", "tile_edges", "one_row", "mostly_empty"])
@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_spmv_fp64(dev, name, lanes, kind):
    print(f"[spmv] {name} lanes={lanes} kernel={kind}", flush=True)
    (m, n), I, J, A = CASES[name]()
    ref = O.csr2mbsr(m, n, I, J, A)
    M = make(dev, m, n, I, J, A)
    M.set_spmv_lanes(lanes)
    M.set_spmv_kernel(kind)
    rng = np.random.default_rng(lanes)
    x, b = rng.standard_normal(n), rng.standard_normal(m)
    for alpha, beta in [(1.0, 0.0), (-1.0, 1.0), (0.5, -2.0)]:
        exp = O.mbsr_spmv(ref, alpha, x, beta, b)
        y = M.spmv(T(x, dev), alpha=alpha, beta=beta, b=T(b, dev)).cpu().numpy()
        scale = abs(alpha) * row_scale(m, n, I, J, A, x, 0.0, None) + abs(beta) * np.abs(b) + 1e-300
        assert np.all(np.abs(y - exp) <= 1e-12 * scale), (alpha, beta)
    # b aliasing y (residual form, hypre_CSRMatrixMatvecDevice b == y)
    yb = T(b, dev)
    M.spmv(T(x, dev), yb, alpha=-1.0, beta=1.0, b=yb)
    exp = O.mbsr_spmv(ref, -1.0, x, 1.0, b)
    assert np.allclose(yb.cpu().numpy(), exp, rtol=1e-12, atol=1e-12)
    # beta == 0 never reads b
    nan_b = T(np.full(m, np.nan), dev)
    y0 = M.spmv(T(x, dev), alpha=1.0, beta=0.0, b=nan_b).cpu().numpy()
    assert np.all(np.isfinite(y0))


@pytest.mark.parametrize("name,kind", [("27pt_10", 2), ("skewed", 3), ("tile_edges", 3)])
def test_spmv_deterministic(dev, name, kind):
    (m, n), I, J, A = CASES[name]()
    M = make(dev, m, n, I, J, A)
    M.set_spmv_kernel(kind)
    x = T(np.random.default_rng(0).standard_normal(n), dev)
    y1 = M.spmv(x).clone()
    y2 = M.spmv(x).clone()
    assert torch.equal(y1, y2)


def _big_skewed():
    """200,000 rows in a band of 8 entries each, plus 4 block rows of 2,000
    entries per row: 50,000 block rows, about 600k blocks (about 12 per block
    row), a spread of blocks per block row far above 12."""
    rng = np.random.default_rng(17)
    m = 200000
    band = np.array([0, 1, 7, 301, 4099, 9973, 30011, 77777])
    rows = []
    for r in range(m):
        if r // 4 in (3, 12000, 30000, 49999):
            rows.append(np.sort(rng.permutation(m)[:2000]))
        else:
            rows.append(np.sort((r + band) % m))
    I = np.zeros(m + 1, np.int32)
    I[1:] = np.cumsum([len(c) for c in rows])
    J = np.concatenate(rows).astype(np.int32)
    return (m, m), I, J, rng.standard_normal(len(J))


def test_default_kernel_from_statistics(dev):
    """The statistics choice of a new handle (amgt_mbsr_tune_from_stats): the
    direct walk up to 150k blocks (whatever the spread), the balanced kernel
    for a spread >= 12 above that (csr_matvec_device.c:1336's switch), else the
    block-lane kernel; lanes: direct ~2^18 lanes in all, balanced 16."""
    I7, J7, A7 = O.stencil(7, 64, 64, 64)
    cases = [("skewed", CASES["skewed"](), L.AMGT_SPMV_DIRECT),
             ("27pt_10", CASES["27pt_10"](), L.AMGT_SPMV_DIRECT),
             ("c1_5pt_128", CASES["c1_5pt_128"](), L.AMGT_SPMV_DIRECT),
             ("big_skewed", _big_skewed(), L.AMGT_SPMV_BALANCED),
             ("7pt_64", ((64 ** 3, 64 ** 3), I7, J7, A7), L.AMGT_SPMV_BLOCK)]
    for name, ((m, n), I, J, A), want in cases:
        ref = O.csr2mbsr(m, n, I, J, A)
        stand, _ = O.mbsr_stats(ref)
        M = make(dev, m, n, I, J, A)
        inf = M.info()
        nblk, brow = inf["nblk"], inf["brow"]
        assert inf["spmv_kernel"] == want, (name, nblk, stand)
        if want == L.AMGT_SPMV_BALANCED:
            assert stand >= 12.0 and nblk > 150000 and inf["spmv_lanes"] == 16
        if want == L.AMGT_SPMV_DIRECT:
            g = inf["spmv_lanes"]
            assert g in (4, 8, 16, 32, 64)
            assert g == 64 or g * brow >= 0.9 * 262144
        # overrides are undone by tune_from_stats
        M.set_spmv_kernel(L.AMGT_SPMV_STAGED)
        M.set_spmv_lanes(8)
        M.tune_from_stats()
        assert M.info()["spmv_kernel"] == want


@pytest.mark.parametrize("prec,tol", [(L.AMGT_FLOAT, 2e-6), (L.AMGT_HALF, 2e-3)])
def test_spmv_balanced_low_precision(dev, prec, tol):
    (m, n), I, J, A = CASES["skewed"]()
    ref = O.csr2mbsr(m, n, I, J, A)
    M = make(dev, m, n, I, J, A)
    M.set_precision(prec)
    M.set_spmv_kernel(L.AMGT_SPMV_BALANCED)
    rng = np.random.default_rng(3)
    x, b = rng.standard_normal(n), rng.standard_normal(m)
    exp = O.mbsr_spmv(ref, -1.0, x, 1.0, b)
    y = M.spmv(T(x, dev), alpha=-1.0, beta=1.0, b=T(b, dev)).cpu().numpy()
    scale = row_scale(m, n, I, J, A, x, 1.0, b)
    assert np.all(np.abs(y - exp) <= tol * scale)


@pytest.mark.parametrize("name", ["rand_rect", "rand_tall_empty", "7pt_aniso", "bin1_rows"])
def test_spmv_transposed(dev, name):
    (m, n), I, J, A = CASES[name]()
    ref = O.csr2mbsr(m, n, I, J, A)
    M = make(dev, m, n, I, J, A)
    rng = np.random.default_rng(3)
    x, b = rng.standard_normal(m), rng.standard_normal(n)
    exp = O.mbsr_spmv_t(ref, 1.5, x, -0.5, b)
    y = M.spmv(T(x, dev), alpha=1.5, beta=-0.5, b=T(b, dev), trans=True).cpu().numpy()
    At = np.abs(A)
    import scipy.sparse as sp
    scale = 1.5 * (sp.csr_matrix((At, J, I), shape=(m, n)).T @ np.abs(x)) + 0.5 * np.abs(b)
    assert np.all(np.abs(y - exp) <= 1e-12 * scale + 1e-300)


@pytest.mark.parametrize("prec,tol", [(L.AMGT_FLOAT, 2e-6), (L.AMGT_HALF, 2e-3)])
def test_spmv_low_precision(dev, prec, tol):
    (m, n), I, J, A = CASES["7pt_aniso"]()
    ref = O.csr2mbsr(m, n, I, J, A)
    M = make(dev, m, n, I, J, A)
    M.set_precision(prec)
    rng = np.random.default_rng(11)
    x, b = rng.standard_normal(n), rng.standard_normal(m)
    exp = O.mbsr_spmv(ref, -1.0, x, 1.0, b)
    y = M.spmv(T(x, dev), alpha=-1.0, beta=1.0, b=T(b, dev)).cpu().numpy()
    scale = row_scale(m, n, I, J, A, x, 1.0, b)
    err = np.abs(y - exp) / scale
    assert err.max() <= tol
    assert err.max() > 0  # really ran in reduced precision


def half_round(v):
    """The library's FP16 copy read back: values times 2^e (max |v| 2^e in
    [2^14, 2^15)), rounded through FP32 to FP16, divided by 2^e."""
    m = float(np.max(np.abs(v))) if v.size else 0.0
    if m == 0.0:
        return v.astype(np.float32).astype(np.float16).astype(np.float64)
    sc = 2.0 ** (15 - np.frexp(m)[1])
    return (v * sc).astype(np.float32).astype(np.float16).astype(np.float64) / sc


@pytest.mark.parametrize("factor", [3.0e5, 1.0e-9, 1.0])
def test_half_scaled_copy(dev, factor):
    """FP16 copies are scaled by a power of two: values far beyond FP16's range
    (x 3e5, max 1.9e6) or below its smallest subnormal (x 1e-9) run in FP16
    with FP16's relative precision; the result equals the FP64 product of the
    scaled-and-rounded values (FP32 accumulation) within 2e-6 of the row
    scale; the SpGEMM of two scaled copies is rescaled exactly."""
    (m, n), I, J, A = CASES["7pt_aniso"]()
    V = A * factor
    M = make(dev, m, n, I, J, V)
    M.set_precision(L.AMGT_HALF)
    x = np.random.default_rng(12).standard_normal(n)
    y = M.spmv(T(x, dev)).cpu().numpy()
    assert M.info()["precision"] == L.AMGT_HALF
    rs = row_scale(m, n, I, J, V, x, 0.0, None)
    exp16 = O.mbsr_spmv(O.csr2mbsr(m, n, I, J, half_round(V)), 1.0, x)
    assert np.all(np.abs(y - exp16) <= 2e-6 * rs)
    exp = O.mbsr_spmv(O.csr2mbsr(m, n, I, J, V), 1.0, x)
    assert np.all(np.abs(y - exp) <= 2e-3 * rs)
    # Jacobi sweep through the same scaled copy
    d = np.abs(V[np.asarray(J) == np.repeat(np.arange(m), np.diff(I))])
    dinv = 1.0 / d
    f, u = np.random.default_rng(13).standard_normal(m), np.random.default_rng(14).standard_normal(n)
    un = M.relax_jacobi(T(f, dev), T(u, dev), T(dinv, dev), 0.8).cpu().numpy()
    ref = u + 0.8 * dinv * (f - O.mbsr_spmv(O.csr2mbsr(m, n, I, J, half_round(V)), 1.0, u))
    assert np.all(np.abs(un - ref) <= 2e-6 * 0.8 * dinv * (np.abs(f) + row_scale(m, n, I, J, V, u, 0.0, None)) + 1e-300)
    # SpGEMM with both operands' scaled FP16 copies
    C = M.spgemm(M)
    CI, CJ, CV = C.to_csr()
    ref = O.mbsr_spgemm(O.csr2mbsr(m, n, I, J, half_round(V)), O.csr2mbsr(m, n, I, J, half_round(V)))
    RI, RJ, RV = O.mbsr2csr(ref)
    cv = CV.cpu().numpy()
    assert np.array_equal(CI.cpu().numpy(), RI) and np.array_equal(CJ.cpu().numpy(), RJ)
    assert np.all(np.abs(cv - RV) <= 2.0 ** -22 * np.abs(RV) + 1e-300 + 1e-7 * np.abs(RV).max())


def test_half_nonfinite_values_use_fp32(dev):
    """A matrix with an infinite value gets no FP16 copy: FP32 instead (the
    reported precision follows); new finite values make FP16 possible again."""
    (m, n), I, J, A = CASES["7pt_aniso"]()
    bad = A.copy()
    bad[5] = np.inf
    M = make(dev, m, n, I, J, bad)
    M.set_precision(L.AMGT_HALF)
    x = np.random.default_rng(12).standard_normal(n)
    M.spmv(T(x, dev))
    assert M.info()["precision"] == L.AMGT_FLOAT
    M.update_values(T(I, dev), T(J, dev), T(A, dev))
    y2 = M.spmv(T(x, dev)).cpu().numpy()
    assert M.info()["precision"] == L.AMGT_HALF
    err2 = np.abs(y2 - O.mbsr_spmv(O.csr2mbsr(m, n, I, J, A), 1.0, x)) / row_scale(m, n, I, J, A, x, 0.0, None)
    assert err2.max() <= 2e-3 and err2.max() > 2e-6


@pytest.mark.parametrize("name", ["rand_rect", "27pt_10", "skewed", "rand_tall_empty"])
@pytest.mark.parametrize("k", [1, 3, 8, 9, 16, 17, 32, 40])
def test_spmm(dev, name, k):
    """k right-hand sides (each contiguous, leading dimension padded past n):
    the scalar kernel below 8, the MFMA kernel (16 right-hand sides per
    instruction, 32 per launch) from 8; every column against the oracle."""
    print(f"[spmm] {name} k={k}", flush=True)
    (m, n), I, J, A = CASES[name]()
    ref = O.csr2mbsr(m, n, I, J, A)
    M = make(dev, m, n, I, J, A)
    rng = np.random.default_rng(k)
    X, B = rng.standard_normal((k, n + 3)), rng.standard_normal((k, m + 5))
    Y = T(np.full((k, m + 7), np.nan), dev)
    M.spmm(T(X, dev)[:, :n], Y[:, :m], alpha=0.5, beta=2.0, B=T(B, dev)[:, :m])
    Yh = Y.cpu().numpy()
    assert np.all(np.isnan(Yh[:, m:]))  # padding untouched
    for v in range(k):
        exp = O.mbsr_spmv(ref, 0.5, X[v, :n], 2.0, B[v, :m])
        scale = 0.5 * row_scale(m, n, I, J, A, X[v, :n], 0.0, None) + 2.0 * np.abs(B[v, :m])
        assert np.all(np.abs(Yh[v, :m] - exp) <= 1e-12 * scale), v


@pytest.mark.parametrize("name", ["rand_rect", "27pt_10", "skewed"])
@pytest.mark.parametrize("k", [1, 5, 16, 17, 40])
@pytest.mark.parametrize("trans", [False, True])
def test_spmm_interleaved(dev, name, k, trans):
    """Interleaved multivectors (entry (i, v) at x[i*k + v]) through the MFMA kernel;
    b aliasing y."""
    print(f"[spmm-il] {name} k={k} trans={trans}", flush=True)
    (m, n), I, J, A = CASES[name]()
    ref = O.csr2mbsr(m, n, I, J, A)
    M = make(dev, m, n, I, J, A)
    nx, ny = (m, n) if trans else (n, m)
    rng = np.random.default_rng(k + 100)
    X, B = rng.standard_normal((nx, k)), rng.standard_normal((ny, k))
    Yb = T(B, dev)
    M.spmm_interleaved(T(X, dev), Yb, alpha=-1.5, beta=0.5, B=Yb, trans=trans)
    Y = Yb.cpu().numpy()
    S = sp.csr_matrix((np.abs(A), J, I), shape=(m, n))
    for v in range(k):
        if trans:
            exp = O.mbsr_spmv_t(ref, -1.5, X[:, v], 0.5, B[:, v])
            scale = 1.5 * (S.T @ np.abs(X[:, v])) + 0.5 * np.abs(B[:, v])
        else:
            exp = O.mbsr_spmv(ref, -1.5, X[:, v], 0.5, B[:, v])
            scale = 1.5 * (S @ np.abs(X[:, v])) + 0.5 * np.abs(B[:, v])
        assert np.all(np.abs(Y[:, v] - exp) <= 1e-12 * scale + 1e-300), v


@pytest.mark.parametrize("prec,tol", [(L.AMGT_FLOAT, 2e-6), (L.AMGT_HALF, 2e-3)])
def test_spmm_low_precision_and_transposed(dev, prec, tol):
    (m, n), I, J, A = CASES["27pt_10"]()
    ref = O.csr2mbsr(m, n, I, J, A)
    M = make(dev, m, n, I, J, A)
    k = 24
    X = np.random.default_rng(9).standard_normal((k, n))
    Yt = M.spmm(T(X, dev), trans=True).cpu().numpy()
    for v in range(k):
        exp = O.mbsr_spmv_t(ref, 1.0, X[v], 0.0, None)
        assert np.allclose(Yt[v], exp, rtol=1e-12, atol=1e-12 * np.abs(exp).max())
    M.set_precision(prec)
    Y = M.spmm(T(X, dev)).cpu().numpy()
    for v in range(k):
        exp = O.mbsr_spmv(ref, 1.0, X[v], 0.0, None)
        scale = row_scale(m, n, I, J, A, X[v], 0.0, None)
        assert np.all(np.abs(Y[v] - exp) <= tol * scale)


def test_full_size_c2_spmv(dev):
    """3D 7-point 128^3 (2,097,152 rows): bit-exact structure and FP64 SpMV vs the oracle."""
    from amgt_amd import workloads as W
    I, J, A = W.stencil(7, 128, 128, 128)
    n = len(I) - 1
    print(f"[full C2] n={n} nnz={len(J)}", flush=True)
    ref = O.csr2mbsr(n, n, I, J, A)
    M = make(dev, n, n, I, J, A)
    check_structure(M, ref)
    x = W.rand_normal(1, n)
    exp = O.mbsr_spmv(ref, 1.0, x)
    y = M.spmv(T(x, dev)).cpu().numpy()
    scale = row_scale(n, n, I, J, A, x, 0.0, None)
    assert np.all(np.abs(y - exp) <= 1e-12 * scale)
    # closed form: A*1 = number of missing neighbours per row
    y1 = M.spmv(T(np.ones(n), dev)).cpu().numpy()
    exp1 = O.csr_matvec(n, I, J, A, 1.0, np.ones(n))
    assert np.array_equal(y1, exp1)


def test_hypre_matvec_boundary(dev):
    (m, n), I, J, A = CASES["7pt_aniso"]()
    C = H.CSRMatrix(T(I, dev), T(J, dev), T(A, dev), m, n)
    rng = np.random.default_rng(2)
    x, b = rng.standard_normal(n), rng.standard_normal(m)
    y = torch.empty(m, dtype=torch.float64, device=dev)
    assert H.hypre_CSRMatrixMatvecDevice(0, -1.0, C, T(x, dev), 1.0, T(b, dev), y) == 0
    exp = O.csr_matvec(m, I, J, A, -1.0, x, 1.0, b)
    assert np.allclose(y.cpu().numpy(), exp, rtol=1e-12, atol=1e-12)
    assert C.bsr_mat  # cached after the first call
    # transposed (MatvecT): y = A^T x
    yt = torch.empty(n, dtype=torch.float64, device=dev)
    H.hypre_CSRMatrixMatvecDevice(1, 1.0, C, T(b, dev), 0.0, yt, yt)
    expt = O.csr_matvec_t(m, n, I, J, A, 1.0, b, 0.0, np.zeros(n))
    assert np.allclose(yt.cpu().numpy(), expt, rtol=1e-12, atol=1e-12)
    # alpha == 0: y = beta*b without touching A
    H.hypre_CSRMatrixMatvecDevice(0, 0.0, C, T(x, dev), 2.0, T(b, dev), y)
    assert np.array_equal(y.cpu().numpy(), 2.0 * b)
    # x == y is rejected (csr_matvec_device.c:65-69)
    xx = T(x, dev)
    with pytest.raises(L.AmgtError):
        H.hypre_CSRMatrixMatvecDevice(0, 1.0, C, xx, 0.0, xx, xx)
    # precision follows the level tag only in mixed mode
    C.coasen_level = 2
    from amgt_amd.mbsr import set_mixed_precision
    set_mixed_precision(True)
    try:
        H.hypre_CSRMatrixMatvecDevice(0, 1.0, C, T(x, dev), 0.0, y, y)
        assert C.bsr().info()["precision"] == L.AMGT_HALF
        e16 = np.abs(y.cpu().numpy() - O.csr_matvec(m, I, J, A, 1.0, x, 0.0, b))
        e16 = (e16 / row_scale(m, n, I, J, A, x, 0.0, None)).max()
        assert 0 < e16 < 2e-3
    finally:
        set_mixed_precision(False)
    assert C.bsr().info()["precision"] == L.AMGT_DOUBLE
    C.invalidate()


def test_autotune_keeps_results(dev):
    (m, n), I, J, A = CASES["27pt_10"]()
    M = make(dev, m, n, I, J, A)
    x = T(np.random.default_rng(1).standard_normal(n), dev)
    y0 = M.spmv(x).cpu().numpy()
    M.autotune()
    assert M.info()["spmv_lanes"] in (4, 8, 16, 32, 64)
    y1 = M.spmv(x).cpu().numpy()
    assert np.allclose(y0, y1, rtol=1e-13, atol=1e-13)


def test_update_values(dev):
    (m, n), I, J, A = CASES["rand_rect"]()
    M = make(dev, m, n, I, J, A)
    A2 = A * 3.0 - 1.0
    M.update_values(T(I, dev), T(J, dev), T(A2, dev))
    ref = O.csr2mbsr(m, n, I, J, A2)
    assert np.array_equal(M.arrays()["val"].cpu().numpy(), ref["val"])
