"""GPU parity of the mBSR SpGEMM and the CSR export against the CPU oracle,
through the C-ABI.

Bars: C's block structure (blk_ptr, blk_col, blk_map, val_ptr) and the CSR
pattern bit-exact; FP64 values within 1e-12 relative to (|A| |B|) entrywise.
Patterns are also checked against HYPRE's host product (orc_csr_spgemm, without
its forced diagonal) and, at full size, against HYPRE's own A^2 nnz fixture.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from amgt_amd import _lib as L
from amgt_amd import hypre as H
from amgt_amd import workloads as W
from amgt_amd.mbsr import MBSRMatrix
from oracle import oracle as O
from tests.util_matrices import dense_rows_csr, rand_csr

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "hypre_fixtures.json")))


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def make(dev, m, n, I, J, A):
    return MBSRMatrix.from_csr(T(I, dev), T(J, dev), None if A is None else T(A, dev), m, n)


def abs_product_scale(m, k, n, AI, AJ, AA, BI, BJ, BA):
    SA = sp.csr_matrix((np.abs(AA), AJ, AI), shape=(m, k))
    SB = sp.csr_matrix((np.abs(BA), BJ, BI), shape=(k, n))
    return (SA @ SB).tocsr()


def check_product(dev, m, k, n, AI, AJ, AA, BI, BJ, BA, tag=""):
    print(f"[spgemm] {tag} {m}x{k}x{n}", flush=True)
    MA, MB = O.csr2mbsr(m, k, AI, AJ, AA), O.csr2mbsr(k, n, BI, BJ, BA)
    ref = O.mbsr_spgemm(MA, MB)
    GA, GB = make(dev, m, k, AI, AJ, AA), make(dev, k, n, BI, BJ, BA)
    GC = GA.spgemm(GB)
    inf = GC.info()
    assert (inf["nblk"], inf["nnz"]) == (ref["nblk"], ref["nnz"])
    a = GC.arrays()
    assert np.array_equal(a["blk_ptr"].cpu().numpy(), ref["blk_ptr"])
    assert np.array_equal(a["blk_col"].cpu().numpy(), ref["blk_col"])
    assert np.array_equal(a["blk_map"].cpu().numpy().view(np.uint16), ref["blk_map"])
    assert np.array_equal(a["val_ptr"].cpu().numpy(), ref["val_ptr"])
    # values through the CSR export (sorted columns)
    CI, CJ, CV = GC.to_csr()
    RI, RJ, RV = O.mbsr2csr(ref)
    assert np.array_equal(CI.cpu().numpy(), RI) and np.array_equal(CJ.cpu().numpy(), RJ)
    # |A||B| has exactly the structural pattern of A*B (no cancellation), in the
    # same sorted CSR order as the export
    scale = abs_product_scale(m, k, n, AI, AJ, AA, BI, BJ, BA)
    scale.sort_indices()
    assert np.array_equal(scale.indptr, RI) and np.array_equal(scale.indices, RJ)
    sc = scale.data
    err = np.abs(CV.cpu().numpy() - RV)
    assert np.all(err <= 1e-12 * sc + 1e-300)
    return GC, ref


@pytest.mark.parametrize("shape,dens,seed", [((37, 53, 29), (0.15, 0.2), 1),
                                             ((16, 16, 16), (0.3, 0.3), 2),
                                             ((9, 70, 13), (0.2, 0.25), 3),
                                             ((101, 9, 77), (0.3, 0.1), 4)])
def test_spgemm_random(dev, shape, dens, seed):
    m, k, n = shape
    AI, AJ, AA = rand_csr(m, k, dens[0], seed)
    BI, BJ, BA = rand_csr(k, n, dens[1], seed + 50)
    check_product(dev, m, k, n, AI, AJ, AA, BI, BJ, BA, "random")


def blocks_per_row_csr(nb_rows, nb_cols, kb, seed, empty_row=None):
    """CSR matrix of 4 nb_rows x 4 nb_cols whose every block row has exactly kb
    occupied 4x4 blocks (1-3 entries each, random positions), except block row
    `empty_row`, which has none."""
    rng = np.random.default_rng(seed)
    rows = [dict() for _ in range(4 * nb_rows)]
    for bi in range(nb_rows):
        if bi == empty_row:
            continue
        for bc in rng.choice(nb_cols, size=kb, replace=False):
            for _ in range(int(rng.integers(1, 4))):
                r, c = 4 * bi + int(rng.integers(4)), 4 * int(bc) + int(rng.integers(4))
                rows[r][c] = float(rng.uniform(-1.0, 1.0))
    I = np.zeros(4 * nb_rows + 1, dtype=np.int32)
    J, A = [], []
    for r, d in enumerate(rows):
        for c in sorted(d):
            J.append(c)
            A.append(d[c])
        I[r + 1] = len(J)
    return I, np.array(J, dtype=np.int32), np.array(A, dtype=np.float64)


@pytest.mark.parametrize("kb_b", [8, 9])
def test_spgemm_register_bound(dev, kb_b):
    # largest A block row (8 blocks) times largest B block row: 64 products is
    # the bounded register form (fixed temporaries per row, no upper-bound
    # pass), 72 the general form with bins; both against the oracle, with an
    # empty block row in A
    AI, AJ, AA = blocks_per_row_csr(40, 30, 8, 11, empty_row=7)
    BI, BJ, BA = blocks_per_row_csr(30, 25, kb_b, 12)
    GC, ref = check_product(dev, 160, 120, 100, AI, AJ, AA, BI, BJ, BA, f"bound kb_b={kb_b}")
    cnt = np.diff(ref["blk_ptr"])
    assert cnt[7] == 0 and cnt.max() >= 20  # the empty row, and rows near the bound
    assert GC.info()["max_row_blocks"] == int(cnt.max())


def test_spgemm_stencils(dev):
    for kind, dims in [(7, (10, 9, 8)), (27, (7, 6, 5)), (7, (33, 31, 1)), (9, (20, 17, 1))]:
        I, J, A = W.stencil(kind, *dims)
        n = len(I) - 1
        check_product(dev, n, n, n, I, J, A, I, J, A, f"stencil{kind}")


def test_spgemm_bins_and_big_rows(dev):
    # block rows with ub in (128, 2048] (workgroup LDS table) and > 2048 (global table),
    # and C block rows with more than 64 blocks (global numeric accumulators)
    AI, AJ, AA = dense_rows_csr(24, 600, 10, 7)      # ub ~ 350 per block row: LDS table bin
    BI, BJ, BA = rand_csr(600, 900, 0.01, 8)
    check_product(dev, 24, 600, 900, AI, AJ, AA, BI, BJ, BA, "bin1")
    AI, AJ, AA = dense_rows_csr(24, 600, 40, 7)      # ub up to ~13500: global table bin
    BI, BJ, BA = rand_csr(600, 900, 0.05, 8)
    check_product(dev, 24, 600, 900, AI, AJ, AA, BI, BJ, BA, "bin2")
    AI, AJ, AA = dense_rows_csr(8, 3000, 400, 9)
    BI, BJ, BA = rand_csr(3000, 5000, 0.004, 10)
    check_product(dev, 8, 3000, 5000, AI, AJ, AA, BI, BJ, BA, "bin2")
    # over 2048 distinct C blocks per block row: the LDS try gives up, the
    # global table redoes the row
    BI, BJ, BA = rand_csr(3000, 40000, 0.004, 15)
    C = sp.csr_matrix((AA, AJ, AI), shape=(8, 3000)) @ sp.csr_matrix((BA, BJ, BI), shape=(3000, 40000))
    assert len({c // 4 for r, c in zip(*C.nonzero()) if r < 4}) > 2048
    check_product(dev, 8, 3000, 40000, AI, AJ, AA, BI, BJ, BA, "bin2_retry")
    # C block rows of 65..128 blocks (LDS accumulators of the long-row kernel)
    # next to rows over 128 (its global slab)
    AI, AJ, AA = dense_rows_csr(16, 600, 60, 11)
    BI, BJ, BA = rand_csr(600, 480, 0.05, 12)
    C = sp.csr_matrix((AA, AJ, AI), shape=(16, 600)) @ sp.csr_matrix((BA, BJ, BI), shape=(600, 480))
    blocks = {(r // 4, c // 4) for r, c in zip(*C.nonzero())}
    per_row = np.bincount([b[0] for b in blocks], minlength=4)
    assert per_row.max() > 64
    check_product(dev, 16, 600, 480, AI, AJ, AA, BI, BJ, BA, "bin2_lds")
    BI, BJ, BA = rand_csr(600, 2000, 0.05, 13)
    check_product(dev, 16, 600, 2000, AI, AJ, AA, BI, BJ, BA, "bin2_slab")


def test_spgemm_long_sparse_rows(dev):
    """C block rows over 64 blocks with sparse blocks (coarse Galerkin rows):
    rows of at most 512 blocks and 1024 values accumulate compressed in LDS,
    the others (more blocks, or more values) in the long-row kernel, in the
    same launch; both against the oracle."""
    rng = np.random.default_rng(21)
    k = n = 4000
    rows = []
    for r in range(16):
        br = r // 4
        if br == 0:      # ~320 blocks, ~380 values: compressed accumulators
            cols = rng.choice(k, 80, replace=False)
        elif br == 1:    # ~540 blocks: long-row kernel (slab)
            cols = rng.choice(k, 160, replace=False)
        elif br == 2:    # short rows
            cols = rng.choice(k, 12, replace=False)
        else:            # ~300 blocks but ~1200 values: long-row kernel
            cols = rng.choice(1200, 300, replace=False)
        rows.append(np.sort(cols))
    AI = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.int32)
    AJ = np.concatenate(rows).astype(np.int32)
    AA = rng.standard_normal(AJ.size)
    Bm = (sp.identity(k, format="csr") * 2.0 + sp.random(k, n, density=0.00005, random_state=22,
                                                           format="csr")).tocsr()
    Bm.sort_indices()
    BI, BJ, BA = Bm.indptr.astype(np.int32), Bm.indices.astype(np.int32), Bm.data.copy()
    C = (sp.csr_matrix((AA, AJ, AI), shape=(16, k)) @ Bm).tocoo()
    nblk = [len({c // 4 for r, c in zip(C.row, C.col) if r // 4 == b}) for b in range(4)]
    nval = [int(np.sum(C.row // 4 == b)) for b in range(4)]
    assert 64 < nblk[0] <= 512 and nval[0] <= 1024
    assert nblk[1] > 512 and nval[3] > 1024
    check_product(dev, 16, k, n, AI, AJ, AA, BI, BJ, BA, "long_sparse")


def test_spgemm_rap(dev):
    nx, ny, nz = 17, 14, 11
    AI, AJ, AA = W.stencil(7, nx, ny, nz)
    n = len(AI) - 1
    PI, PJ, PA, nc = W.trilinear_interpolation(nx, ny, nz)
    RI, RJ, RA = W.csr_transpose(n, nc, PI, PJ, PA)
    GAP, refAP = check_product(dev, n, n, nc, AI, AJ, AA, PI, PJ, PA, "A*P")
    APi, APj, APv = O.mbsr2csr(refAP)
    check_product(dev, nc, n, nc, RI, RJ, RA, APi, APj, APv, "R*(AP)")


def test_spgemm_deterministic(dev):
    AI, AJ, AA = rand_csr(60, 70, 0.2, 11)
    BI, BJ, BA = rand_csr(70, 80, 0.2, 12)
    GA, GB = make(dev, 60, 70, AI, AJ, AA), make(dev, 70, 80, BI, BJ, BA)
    v1 = GA.spgemm(GB).arrays()["val"].cpu()
    v2 = GA.spgemm(GB).arrays()["val"].cpu()
    assert torch.equal(v1, v2)


def test_spgemm_full_size_c2_a_squared(dev):
    """HYPRE benchmark job 1 (ij_mm -n 128 128 128 -7pt): A^2 has 51,742,208 nonzeros."""
    I, J, A = W.stencil(7, 128, 128, 128)
    n = len(I) - 1
    G = make(dev, n, n, I, J, A)
    C = G.spgemm(G)
    nnz = C.info()["nnz"]
    assert nnz == GOLD["A_squared_nnz"]["7pt_128_128_128"]["nnz"]
    # values: row sums of A^2 equal A*(A*1) exactly (small integers)
    ones = torch.ones(n, dtype=torch.float64, device=dev)
    y = C.spmv(ones)
    z = G.spmv(G.spmv(ones))
    assert torch.equal(y, z)


def test_hypre_spgemm_boundary(dev):
    AI, AJ, AA = rand_csr(37, 53, 0.15, 21)
    BI, BJ, BA = rand_csr(53, 29, 0.2, 22)
    A = H.CSRMatrix(T(AI, dev), T(AJ, dev), T(AA, dev), 37, 53)
    B = H.CSRMatrix(T(BI, dev), T(BJ, dev), T(BA, dev), 53, 29)
    C = H.hypreDevice_CSRSpGemm(A, B)
    CI, CJ, CA = O.csr_spgemm(37, 29, AI, AJ, AA, BI, BJ, BA, diag_first=0)
    Sref = sp.csr_matrix((CA, CJ, CI), shape=(37, 29))
    Sref.sort_indices()
    Sg = sp.csr_matrix((C.data.cpu().numpy(), C.j.cpu().numpy(), C.i.cpu().numpy()), shape=(37, 29))
    assert np.array_equal(Sg.indptr, Sref.indptr) and np.array_equal(Sg.indices, Sref.indices)
    assert abs(Sg - Sref).max() < 1e-12
    assert A.bsr_mat and B.bsr_mat and C.bsr_mat  # caches kept
    # trivial case: an empty operand gives a zero row pointer
    E = H.CSRMatrix(T(np.zeros(54, np.int32), dev), T(np.zeros(0, np.int32), dev),
                    T(np.zeros(0), dev), 53, 29)
    Z = H.hypreDevice_CSRSpGemm(A, E)
    assert Z.num_nonzeros == 0 and torch.count_nonzero(Z.i).item() == 0
    # dimension mismatch
    with pytest.raises(L.AmgtError):
        H.hypreDevice_CSRSpGemm(B, B)
    # pattern-only operand: pattern computed, no values
    Ap = H.CSRMatrix(T(AI, dev), T(AJ, dev), None, 37, 53)
    Cp = H.hypreDevice_CSRSpGemm(Ap, B)
    assert Cp.data is None
    assert np.array_equal(Cp.j.cpu().numpy(), Sg.indices)


@pytest.mark.parametrize("prec,dt", [(L.AMGT_FLOAT, np.float32), (L.AMGT_HALF, np.float16)])
@pytest.mark.parametrize("case", ["random", "stencil_27", "big_rows"])
def test_spgemm_reduced_precision(dev, prec, dt, case):
    """spgemm_amgT_fp32 / _fp16 (csr_spgemm_device.c:2224-2578, 1863-2223):
    operands rounded to FP32 / FP16, block products on the 16-block FP32 /
    FP16 MFMAs with FP32 accumulation, C entries stored as FP32 numbers.
    Oracle: the FP64 product of the rounded operands.  Bar: the FP32 summation
    bound, (2 c + 2) 2^-24 of the (|A| |B|) scale of the rounded operands for
    an entry with c products (each product and each addition rounded once,
    plus the final FP32 rounding); and within the input rounding (2^-22 for
    FP32, 2^-9 for FP16) plus that bound of the FP64 product of the unrounded
    operands."""
    if case == "random":
        m, k, n = 61, 47, 53
        AI, AJ, AA = rand_csr(m, k, 0.2, 11)
        BI, BJ, BA = rand_csr(k, n, 0.25, 12)
    elif case == "stencil_27":
        AI, AJ, AA = O.stencil(27, 6, 5, 7)
        m = k = n = 6 * 5 * 7
        BI, BJ, BA = AI, AJ, AA
    else:
        m, k, n = 24, 3000, 40
        AI, AJ, AA = dense_rows_csr(m, k, 400, 13)
        BI, BJ, BA = rand_csr(k, n, 0.05, 14)
    print(f"[spgemm-lowp] {case} prec={prec}", flush=True)
    ra = AA.astype(dt).astype(np.float64)
    rb = BA.astype(dt).astype(np.float64)
    ref = O.mbsr_spgemm(O.csr2mbsr(m, k, AI, AJ, ra), O.csr2mbsr(k, n, BI, BJ, rb))
    RI, RJ, RV = O.mbsr2csr(ref)
    GA, GB = make(dev, m, k, AI, AJ, AA), make(dev, k, n, BI, BJ, BA)
    GA.set_precision(prec)
    GC = GA.spgemm(GB)
    CI, CJ, CV = GC.to_csr()
    assert np.array_equal(CI.cpu().numpy(), RI) and np.array_equal(CJ.cpu().numpy(), RJ)
    cv = CV.cpu().numpy()
    # every entry is an FP32 number
    assert np.array_equal(cv, cv.astype(np.float32).astype(np.float64))
    scale = abs_product_scale(m, k, n, AI, AJ, np.abs(ra), BI, BJ, np.abs(rb))
    scale.sort_indices()
    # products per C entry: the product of the patterns
    cnt = abs_product_scale(m, k, n, AI, AJ, np.ones_like(ra), BI, BJ, np.ones_like(rb))
    cnt.sort_indices()
    bound = (2.0 * cnt.data + 2.0) * 2.0 ** -24 * scale.data
    assert np.all(np.abs(cv - RV) <= bound + 1e-300)
    # and the FP64 product stays within FP32 / FP16 input rounding of the exact one
    GA.set_precision(-1)
    exact = GA.spgemm(GB).to_csr()[2].cpu().numpy()
    tol = 2.0 ** -22 if dt == np.float32 else 2.0 ** -9
    full = abs_product_scale(m, k, n, AI, AJ, AA, BI, BJ, BA).sorted_indices().data
    assert np.all(np.abs(cv - exact) <= tol * full + bound + 1e-300)
