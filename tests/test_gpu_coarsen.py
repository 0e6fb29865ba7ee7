"""GPU tests of the device BoomerAMG setup (amgt_amg_strength / _pmis /
_extpi_interp and the amg.BoomerAMG driver) against the oracle's restatement
of HYPRE's host algorithms.

* strength, PMIS and extended+i (with and without truncation) on the same CSR
  input as the oracle: S and the C/F split exactly equal; P's structure
  exactly equal (row entries in the same order) and its values equal bit for
  bit (the device accumulates in the host's order with contraction off);
* the hierarchy of HYPRE's TEST_ij default job (10^3 Laplacian, PMIS type 9,
  extended+i, P max 0): grid and operator complexity 1.407 / 3.252344 and the
  V(1,1) Jacobi convergence factor 0.678738 of default.saved;
* every level of a truncated (P max 4) hierarchy against the oracle fed with
  the device's own A_l, and R (A P) against the oracle's host SpGEMM;
* PCG with the algebraic V-cycle reaches 1e-8.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from amgt_amd import amg as G
from amgt_amd import coarsen as C
from oracle import oracle as O
from tests.util_matrices import rand_csr

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "hypre_fixtures.json")))


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _sorted_csr(n, I, J, A):
    M = sp.csr_matrix((A, J, I), shape=(n, n))
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64)


def _rand_mmatrix(n, density, seed):
    """Random rows with negative off-diagonals (and a few positive ones) and a
    dominant positive diagonal, columns in random order per row."""
    I, J, A = rand_csr(n, n, density, seed)
    rng = np.random.default_rng(seed)
    M = sp.csr_matrix((-np.abs(A), J, I), shape=(n, n)).tolil()
    M.setdiag(0.0)
    M = M.tocsr()
    M.eliminate_zeros()
    flip = rng.random(M.nnz) < 0.1
    M.data[flip] *= -0.3
    d = np.asarray(abs(M).sum(axis=1)).ravel() + 0.5
    M = (M + sp.diags(d)).tocsr()
    Ii, Jj, Aa = M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.copy()
    for i in range(n):  # shuffle each row
        s, e = Ii[i], Ii[i + 1]
        perm = rng.permutation(e - s)
        Jj[s:e], Aa[s:e] = Jj[s:e][perm], Aa[s:e][perm]
    return Ii, Jj, Aa


CASES = {
    "7pt_10": lambda: (1000,) + tuple(O.stencil(7, 10, 10, 10)),
    "7pt_aniso": lambda: (6 * 7 * 9,) + tuple(O.stencil(7, 6, 7, 9, c=(1, 1, 0.01))),
    "27pt_8": lambda: (512,) + tuple(O.stencil(27, 8, 8, 8)),
    "5pt_33": lambda: (33 * 33,) + tuple(O.stencil(7, 33, 33, 1)),
    "rand_m": lambda: (700,) + _rand_mmatrix(700, 0.01, 3),
    "single": lambda: (1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.0])),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("theta,mrs", [(0.25, 1.0), (0.5, 0.9)])
def test_strength_pmis_extpi_match_oracle(dev, name, theta, mrs):
    print(f"[setup] {name} theta={theta} max_row_sum={mrs}", flush=True)
    n, I, J, A = CASES[name]()
    dI, dJ, dA = T(I, dev), T(J, dev), T(A, dev)
    SI, SJ = C.strength(n, dI, dJ, dA, theta, mrs)
    rSI, rSJ = O.amg_strength(n, I, J, A, theta, mrs)
    assert np.array_equal(SI.cpu().numpy(), rSI) and np.array_equal(SJ.cpu().numpy(), rSJ)
    for cf_init in (0, 2):
        cf, nc = C.pmis(n, SI, SJ, cf_init)
        rcf, rnc = O.amg_pmis(n, rSI, rSJ, cf_init)
        assert nc == rnc and np.array_equal(cf.cpu().numpy(), rcf), cf_init
    for pmx in (0, 4, 2):
        PI, PJ, PA, pnc = C.extpi_interp(n, dI, dJ, dA, SI, SJ, cf, pmx)
        rPI, rPJ, rPA, rnc2 = O.amg_extpi(n, I, J, A, rSI, rSJ, rcf, pmx)
        assert pnc == rnc2
        assert np.array_equal(PI.cpu().numpy(), rPI), pmx
        assert np.array_equal(PJ.cpu().numpy(), rPJ), pmx
        assert np.array_equal(PA.cpu().numpy(), rPA), pmx


def _rand_long(n, per_row, seed):
    """Random M-matrix rows of about per_row negative off-diagonals (sorted
    columns, dominant diagonal): long rows with many distance-two C points."""
    rng = np.random.default_rng(seed)
    c = rng.integers(0, n - 1, size=(n, per_row + 8))
    rows, cols = [], []
    for i in range(n):
        cc = np.unique(c[i])[:per_row]
        cc = np.where(cc >= i, cc + 1, cc)
        rows.append(np.full(len(cc), i))
        cols.append(cc)
    r = np.concatenate(rows)
    cl = np.concatenate(cols)
    M = sp.csr_matrix((-rng.uniform(0.5, 1.5, r.size), (r, cl)), shape=(n, n))
    d = np.asarray(abs(M).sum(axis=1)).ravel() + 0.5
    M = (M + sp.diags(d)).tocsr()
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.copy()


@pytest.mark.parametrize("n,per_row", [(6000, 40), (20000, 100)])
def test_extpi_wave_tiers_match_oracle(dev, n, per_row):
    """Rows whose distinct P columns overflow the wave kernels' small tables
    (6000 x 40: a few rows over 128 columns of the 64-entry tier; 20000 x 100:
    every row over the 256 of the 128-entry tier) are redone with the large
    tables: S, the C/F split and P (P max 4) equal the oracle's exactly."""
    print(f"[extpi tiers] n={n} per_row={per_row}", flush=True)
    I, J, A = _rand_long(n, per_row, 7)
    dI, dJ, dA = T(I, dev), T(J, dev), T(A, dev)
    SI, SJ = C.strength(n, dI, dJ, dA, 0.25, 0.9)
    rSI, rSJ = O.amg_strength(n, I, J, A, 0.25, 0.9)
    assert np.array_equal(SI.cpu().numpy(), rSI) and np.array_equal(SJ.cpu().numpy(), rSJ)
    cf, nc = C.pmis(n, SI, SJ, 0)
    rcf, rnc = O.amg_pmis(n, rSI, rSJ, 0)
    assert nc == rnc and np.array_equal(cf.cpu().numpy(), rcf)
    PI, PJ, PA, _ = C.extpi_interp(n, dI, dJ, dA, SI, SJ, cf, 4)
    rPI, rPJ, rPA, _ = O.amg_extpi(n, I, J, A, rSI, rSJ, rcf, 4)
    assert np.array_equal(PI.cpu().numpy(), rPI)
    assert np.array_equal(PJ.cpu().numpy(), rPJ)
    assert np.array_equal(PA.cpu().numpy(), rPA)


def test_pmis_cf_init_rejected(dev):
    n, I, J, A = CASES["7pt_10"]()
    SI, SJ = C.strength(n, T(I, dev), T(J, dev), T(A, dev))
    with pytest.raises(RuntimeError):
        C.pmis(n, SI, SJ, cf_init=1)


def test_hierarchy_matches_ij_default(dev):
    g = GOLD["ij_default"]
    n, I, J, A = CASES["7pt_10"]()
    amg = G.BoomerAMG(T(I, dev), T(J, dev), T(A, dev), n, dev, theta=0.25, max_row_sum=1.0,
                      cf_init=2, max_elmts=0, omega=1.0, coarse_size=9)
    gc, oc = amg.complexities()
    assert [L_.n for L_ in amg.levels] == [1000, 354, 47, 6]
    assert round(gc, 6) == g["grid_complexity"] and round(oc, 6) == g["operator_complexity"]
    # stationary AMG iteration x += V(b - A x), as hypre_BoomerAMGSolve
    A0 = amg.levels[0].A
    b = A0.spmv(torch.ones(n, dtype=torch.float64, device=dev))
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    r = b.clone()
    z = torch.empty_like(b)
    r0 = bn = torch.linalg.vector_norm(b).item()
    it, rn = 0, r0
    while rn / bn >= 1e-8 and it < 100:
        amg.precondition(r, z)
        x += z
        A0.spmv(x, r, alpha=-1.0, beta=1.0, b=b)
        rn = torch.linalg.vector_norm(r).item()
        it += 1
    assert round((rn / r0) ** (1.0 / it), 6) == g["average_convergence_factor"]


def test_truncated_hierarchy_levels_match_oracle(dev):
    nx = 24
    I, J, A = O.stencil(7, nx, nx, nx)
    n = nx ** 3
    amg = G.BoomerAMG(T(I, dev), T(J, dev), T(A, dev), n, dev, theta=0.25, max_row_sum=0.9,
                      cf_init=0, max_elmts=4, coarse_size=20)
    assert len(amg.levels) >= 4
    for lvl in range(len(amg.levels) - 1):
        Lf, Lc = amg.levels[lvl], amg.levels[lvl + 1]
        dI, dJ, dA = Lf.A.to_csr()
        hI, hJ, hA = dI.cpu().numpy(), dJ.cpu().numpy(), dA.cpu().numpy()
        rSI, rSJ = O.amg_strength(Lf.n, hI, hJ, hA, 0.25, 0.9)
        rcf, rnc = O.amg_pmis(Lf.n, rSI, rSJ, 0)
        assert np.array_equal(amg.cf[lvl].cpu().numpy(), rcf), lvl
        rPI, rPJ, rPA, _ = O.amg_extpi(Lf.n, hI, hJ, hA, rSI, rSJ, rcf, 4)
        P = sp.csr_matrix((rPA, rPJ, rPI), shape=(Lf.n, rnc))
        P.sort_indices()
        gI, gJ, gA = Lf.P.to_csr()
        assert np.array_equal(gI.cpu().numpy(), P.indptr) and np.array_equal(gJ.cpu().numpy(), P.indices)
        assert np.array_equal(gA.cpu().numpy(), P.data), lvl
        # R (A P) against the oracle's host SpGEMM on the same operands
        RI, RJ, RA = O.csr_transpose(Lf.n, rnc, P.indptr, P.indices, P.data)
        AP = O.csr_spgemm(Lf.n, rnc, hI, hJ, hA, P.indptr, P.indices, P.data, False)
        CI, CJ, CA = O.csr_spgemm(rnc, rnc, RI, RJ, RA, *AP, True)
        ref = sp.csr_matrix((CA, CJ, CI), shape=(rnc, rnc))
        ref.sort_indices()
        cI, cJ, cA = Lc.A.to_csr()
        assert np.array_equal(cI.cpu().numpy(), ref.indptr)
        assert np.array_equal(cJ.cpu().numpy(), ref.indices)
        scale = (abs(sp.csr_matrix((RA, RJ, RI), shape=(rnc, Lf.n))) @
                 abs(sp.csr_matrix((hA, hJ, hI), shape=(Lf.n, Lf.n))) @ abs(P)).tocsr()
        scale.sort_indices()
        assert np.all(np.abs(cA.cpu().numpy() - ref.data) <= 1e-12 * scale.data + 1e-300)


def test_pcg_with_algebraic_vcycle(dev):
    nx = 32
    I, J, A = O.stencil(7, nx, nx, nx)
    n = nx ** 3
    amg = G.BoomerAMG(T(I, dev), T(J, dev), T(A, dev), n, dev, max_elmts=4, coarse_size=64)
    amg.capture()
    xs = O.rand_normal(1, n)
    b = T(sp.csr_matrix((A, J, I), shape=(n, n)) @ xs, dev)
    x, iters, hist = amg.pcg(b, tol=1e-8)
    # PMIS + extended+i (P max 4) + Jacobi weight 1: 27 iterations measured
    assert hist[-1] < 1e-8 and iters <= 35, (iters, hist[-1])
    assert np.max(np.abs(x.cpu().numpy() - xs)) < 1e-5 * np.max(np.abs(xs))
