"""The numpy workload generators used by bench.py match the pinned oracle."""
import numpy as np
import pytest

from amgt_amd import workloads as W
from oracle import oracle as O


@pytest.mark.parametrize("kind,dims,c", [(7, (9, 7, 5), (1, 1, 1)), (7, (12, 9, 1), (1, 1, 1)),
                                         (7, (6, 5, 7), (1.0, 1.0, 0.01)), (9, (8, 5, 1), (1, 1, 1)),
                                         (27, (5, 6, 4), (1, 1, 1))])
def test_stencil_matches_oracle(kind, dims, c):
    a = W.stencil(kind, *dims, c=c)
    b = O.stencil(kind, *dims, c=c)
    for x, y in zip(a, b):
        assert x.dtype == y.dtype and np.array_equal(x, y)


def test_rng_matches_oracle():
    assert np.array_equal(W.pm_stream(1, 10000), O.pm_sequence(1, 10000))
    assert np.array_equal(W.pm_stream(0, 100), O.pm_sequence(0, 100))
    assert np.array_equal(W.rand_uniform(0, 3001), O.rand_uniform(0, 3001))
    # numpy's and libm's log/cos/sin may differ in the last ulp
    assert np.allclose(W.rand_normal(2, 3001), O.rand_normal(2, 3001), rtol=1e-14, atol=1e-15)


def test_fv_diffusion_matches_face_loop():
    """Vectorised assembly against a loop over the faces of every cell."""
    k = np.exp(np.random.default_rng(0).normal(size=(5, 7)))
    I, J, A = W.fv_diffusion(k)
    ny, nx = k.shape
    n = nx * ny
    ref = np.zeros((n, n))
    for j in range(ny):
        for i in range(nx):
            p = i + nx * j
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < nx and 0 <= jj < ny:
                    t = 2 * k[j, i] * k[jj, ii] / (k[j, i] + k[jj, ii])
                    ref[p, p] += t
                    ref[p, ii + nx * jj] -= t
                else:
                    ref[p, p] += 2 * k[j, i]
    dense = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(I))
    dense[rows, J] = A
    assert np.allclose(dense, ref, rtol=1e-14, atol=0)
    assert np.all(J[I[:-1]] == np.arange(n))   # diagonal first
    for r in range(n):
        assert np.all(np.diff(J[I[r] + 1:I[r + 1]]) > 0)


def test_fv_constant_k_is_5pt_laplacian():
    """k = 1: interior rows 4 / -1; a boundary face adds 2 to the diagonal."""
    I, J, A = W.fv_diffusion(np.ones((4, 4)))
    d = A[I[:-1]]
    assert d[5] == 4.0 and d[0] == 6.0 and d[1] == 5.0
    assert np.all(A[np.setdiff1d(np.arange(I[-1]), I[:-1])] == -1.0)


def test_grf_covariance():
    """Ensemble mean of the periodic sample autocovariance equals
    sigma^2 exp(-r^2 / (2 ell^2)) (32 fields on 64^2 from the training split's
    seeds; statistical tolerance).  Seeds that were shifted copies of one stream
    fail this test (the fields repeat)."""
    side, sigma, ell = 64, 1.7, 0.08
    acc = np.zeros((side, side))
    nrep = 32
    for s in range(nrep):
        g = W.grf(side, sigma, ell, W.rand_normal(W.problem_seed("train", s, side), side * side))
        G = np.fft.fft2(g)
        acc += np.fft.ifft2(G * np.conj(G)).real / (side * side)
    acc /= nrep
    d = np.minimum(np.arange(side), side - np.arange(side)) / side
    c = sigma ** 2 * np.exp(-(d[:, None] ** 2 + d[None, :] ** 2) / (2 * ell ** 2))
    for lag in (0, 2, 5, 10):
        assert abs(acc[0, lag] - c[0, lag]) < 0.12 * sigma ** 2, (lag, acc[0, lag], c[0, lag])
        assert abs(acc[lag, 0] - c[lag, 0]) < 0.12 * sigma ** 2, (lag, acc[lag, 0], c[lag, 0])


def test_grf_problem_parameters():
    """sigma and ell come from the problem's first two draws, in their ranges."""
    for i in range(20):
        I, J, A, b, sigma, ell = W.grf_problem(16, W.problem_seed("train", i, 16))
        assert 0.5 <= sigma <= 3.0 and 0.01 <= ell <= 0.2
        assert I.size == 257 and b.size == 256 and np.all(A[I[:-1]] > 0)


def test_problem_seeds_disjoint():
    """Problem streams are disjoint stretches of the cycle: problem i + 1's
    seed is problem i's draw number `stride` and none of the earlier ones."""
    s0 = W.problem_seed("train", 0, 256)
    s1 = W.problem_seed("train", 1, 256)
    assert W.pm_stream(s0, 1 << 18)[-1] == s1
    assert s1 not in set(W.pm_stream(s0, (1 << 18) - 1).tolist())
    with pytest.raises(ValueError):
        W.problem_seed("train", 0, 1024)
    with pytest.raises(ValueError):
        W.problem_seed("test", 50, 1024)


def test_grf_problem_regression():
    """The committed policy labels and weights were produced from these
    problems: the generator's outputs for two of them stay as recorded
    (sigma and ell exactly; sums of the matrix and right-hand side to 1e-12)."""
    I, J, A, b, sigma, ell = W.grf_problem(64, W.problem_seed("train", 0, 64))
    assert (sigma, ell) == (0.5000195659231486, 0.03499217974720159)
    assert int(I[-1]) == 20224
    assert np.isclose(A.sum(), 604.7348820216386, rtol=1e-12)
    assert np.isclose(np.abs(A).sum(), 36982.36730485499, rtol=1e-12)
    assert np.isclose(b.sum(), -0.0029478719881590324, rtol=1e-10, atol=1e-15)
    I, J, A, b, sigma, ell = W.grf_problem(256, W.problem_seed("test", 3, 256))
    assert (sigma, ell) == (1.975880250323508, 0.019071906227186276)
    assert np.isclose(np.abs(A).sum(), 3511498.009229469, rtol=1e-12)
    assert np.isclose(b.sum(), -0.001801997472194001, rtol=1e-10, atol=1e-15)
