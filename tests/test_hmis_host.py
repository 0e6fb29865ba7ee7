"""The library's host first pass of Ruge-Stueben coarsening (HMIS,
amgt_amg_ruge_first_pass) against the oracle, CPU only.

The library keeps one first-in-first-out list per measure in an array; the
oracle restates HYPRE's linked list of lists (utilities/amg_linklist.c) step
for step.  The markers must agree bit for bit: which point is taken among
equal measures decides the whole splitting."""
import numpy as np
import pytest
import scipy.sparse as sp

from amgt_amd import _lib as L
from amgt_amd import coarsen as C
from oracle import oracle as O
from tests.util_matrices import rand_csr, to_scipy


def _strength(M, theta=0.25, mrs=1.0):
    M = sp.csr_matrix(M)
    return O.amg_strength(M.shape[0], M.indptr, M.indices, M.data, theta, mrs)


def _cases():
    yield "7pt_10", O.stencil(7, 10, 10, 10), 1000
    yield "5pt_64", O.stencil(7, 64, 64, 1), 64 * 64
    yield "27pt_12", O.stencil(27, 12, 11, 10), 12 * 11 * 10
    yield "aniso_16", O.stencil(7, 16, 16, 16, c=(1.0, 1.0, 0.01)), 16 ** 3
    for seed in range(6):
        n = 300 + 97 * seed
        I, J, A = rand_csr(n, n, 0.02 + 0.01 * seed, seed)
        M = to_scipy(n, n, I, J, A)
        # an M-matrix-like pattern: negative off-diagonals, dominant diagonal,
        # some rows left without strong dependence
        M = -abs(M)
        M = (M + sp.diags(np.asarray(abs(M).sum(axis=1)).ravel() + 1.0)).tocsr()
        yield f"rand_{seed}", (M.indptr, M.indices, M.data), n


@pytest.mark.parametrize("name,mat,n", list(_cases()), ids=lambda v: v if isinstance(v, str) else "")
def test_library_first_pass_equals_oracle(name, mat, n):
    I, J, A = mat
    M = to_scipy(n, n, I, J, A)
    for theta in (0.25, 0.5):
        SI, SJ = _strength(M, theta)
        ref = O.amg_ruge_first_pass(n, SI, SJ)
        got = C.ruge_first_pass(n, SI, SJ)
        assert np.array_equal(got, ref), (name, theta, np.nonzero(got != ref)[0][:10])


def test_library_first_pass_on_coarse_levels():
    # the HMIS hierarchy of a 7-pt 16^3 problem: every level's S
    I, J, A = O.stencil(7, 16, 16, 16)
    levels = O.amg_hierarchy(16 ** 3, I, J, A, theta=0.25, max_row_sum=1.0, max_elmts=4,
                             max_coarse_size=9, coarsen="hmis", rap="kt")
    assert len(levels) >= 4
    for L_ in levels[:-1]:
        SI, SJ = O.amg_strength(L_["n"], L_["I"], L_["J"], L_["A"], 0.25, 1.0)
        assert np.array_equal(C.ruge_first_pass(L_["n"], SI, SJ),
                              O.amg_ruge_first_pass(L_["n"], SI, SJ))


def test_library_first_pass_rejects_bad_input():
    SI = np.array([0, 1, 2], np.int32)
    with pytest.raises(L.AmgtError):
        C.ruge_first_pass(2, SI, np.array([0, 5], np.int32))
    with pytest.raises(ValueError):
        C.ruge_first_pass(3, SI, np.array([1, 0], np.int32))
    assert C.ruge_first_pass(0, np.zeros(1, np.int32), np.zeros(0, np.int32)).size == 0
