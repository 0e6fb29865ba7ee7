"""GPU parity of the smoother support: the diagonal and the fused weighted
Jacobi sweep (relax type 18 on one process) against the CPU oracle's CSR
matvec, through the C-ABI.

Bars: the diagonal bit-exact (it is a copy); the sweep within 1e-12 relative
to |u| + omega |dinv| (|f| + |A| |u|) per row in FP64, 2e-6 / 2e-3 of that
scale with FP32 / FP16 values.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from amgt_amd import _lib as L
from amgt_amd.mbsr import MBSRMatrix
from oracle import oracle as O
from tests.util_matrices import rand_csr, skewed_csr

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def diag_of(n, I, J, A):
    d = np.zeros(n)
    for r in range(n):
        for p in range(I[r], I[r + 1]):
            if J[p] == r:
                d[r] = A[p]
    return d


CASES = {
    "7pt": lambda: ((6 * 7 * 5), *O.stencil(7, 6, 7, 5)),
    "27pt": lambda: ((5 * 5 * 6), *O.stencil(27, 5, 5, 6)),
    "rand_missing_diag": lambda: (97, *rand_csr(97, 97, 0.1, 21)),
    "skewed": lambda: (2003, *skewed_csr(2003, 2003, (3, 200), 800, 3, 22, empty_rows=(5, 2002))),
}


@pytest.mark.parametrize("name", list(CASES))
def test_diagonal(dev, name):
    n, I, J, A = CASES[name]()
    M = MBSRMatrix.from_csr(T(I, dev), T(J, dev), T(A, dev), n, n)
    d = diag_of(n, I, J, A)
    assert np.array_equal(M.diagonal().cpu().numpy(), d)
    inv = np.where(d != 0, 1.0 / np.where(d != 0, np.abs(d), 1.0), 0.0)
    assert np.array_equal(M.diagonal(inverse_abs=True).cpu().numpy(), inv)
    # relax 7: 1/a_ii, 1 where a_ii = 0 (hypre_ParCSRComputeL1Norms option 5)
    inv7 = 1.0 / np.where(d != 0, d, 1.0)
    assert np.array_equal(M.diagonal(kind=L.AMGT_DIAG_INV).cpu().numpy(), inv7)
    # relax 18: 1/l1 with l1 = sum |a_ij| over the row, negated where a_ii < 0
    # (option 1); the sum's order differs from the row's, so within 1 ulp-ish
    l1 = np.asarray(abs(sp.csr_matrix((A, J, I), shape=(n, n))).sum(axis=1)).ravel()
    l1 = np.where(d < 0, -l1, l1)
    inv18 = np.where(l1 != 0, 1.0 / np.where(l1 != 0, l1, 1.0), 0.0)
    got = M.diagonal(kind=L.AMGT_DIAG_INV_L1).cpu().numpy()
    assert np.all((got == 0) == (inv18 == 0))
    assert np.all(np.abs(got - inv18) <= 1e-13 * np.abs(inv18))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("kind", [L.AMGT_SPMV_DIRECT, L.AMGT_SPMV_BLOCK, L.AMGT_SPMV_BALANCED,
                                  L.AMGT_SPMV_STAGED])
@pytest.mark.parametrize("omega", [1.0, 2.0 / 3.0])
def test_relax_jacobi(dev, name, kind, omega):
    n, I, J, A = CASES[name]()
    print(f"[relax] {name} kind={kind} omega={omega}", flush=True)
    M = MBSRMatrix.from_csr(T(I, dev), T(J, dev), T(A, dev), n, n)
    M.set_spmv_kernel(kind)
    rng = np.random.default_rng(5)
    f, u = rng.standard_normal(n), rng.standard_normal(n)
    dinv = M.diagonal(inverse_abs=True)
    di = dinv.cpu().numpy()
    au = O.csr_matvec(n, I, J, A, 1.0, u)
    exp = u + omega * di * (f - au)
    got = M.relax_jacobi(T(f, dev), T(u, dev), dinv, omega).cpu().numpy()
    S = sp.csr_matrix((np.abs(A), J, I), shape=(n, n))
    scale = np.abs(u) + omega * np.abs(di) * (np.abs(f) + S @ np.abs(u)) + 1e-300
    assert np.all(np.abs(got - exp) <= 1e-12 * scale)


@pytest.mark.parametrize("prec,tol", [(L.AMGT_FLOAT, 2e-6), (L.AMGT_HALF, 2e-3)])
def test_relax_jacobi_low_precision(dev, prec, tol):
    n, I, J, A = CASES["27pt"]()
    M = MBSRMatrix.from_csr(T(I, dev), T(J, dev), T(A, dev), n, n)
    M.set_precision(prec)
    rng = np.random.default_rng(6)
    f, u = rng.standard_normal(n), rng.standard_normal(n)
    dinv = M.diagonal(inverse_abs=True)
    di = dinv.cpu().numpy()
    au = O.csr_matvec(n, I, J, A, 1.0, u)
    exp = u + di * (f - au)
    got = M.relax_jacobi(T(f, dev), T(u, dev), dinv).cpu().numpy()
    S = sp.csr_matrix((np.abs(A), J, I), shape=(n, n))
    scale = np.abs(u) + np.abs(di) * (np.abs(f) + S @ np.abs(u))
    assert np.all(np.abs(got - exp) <= tol * scale)


def test_relax_rejects_alias_and_rectangular(dev):
    n, I, J, A = CASES["7pt"]()
    M = MBSRMatrix.from_csr(T(I, dev), T(J, dev), T(A, dev), n, n)
    u = T(np.zeros(n), dev)
    d = M.diagonal(inverse_abs=True)
    with pytest.raises(L.AmgtError) as e:
        M.relax_jacobi(u, u, d, out=u)
    assert e.value.status == L.AMGT_ERR_ARG
    I2, J2, A2 = rand_csr(8, 12, 0.3, 3)
    R = MBSRMatrix.from_csr(T(I2, dev), T(J2, dev), T(A2, dev), 8, 12)
    with pytest.raises(L.AmgtError):
        R.relax_jacobi(T(np.zeros(8), dev), T(np.zeros(12), dev), T(np.ones(8), dev))
