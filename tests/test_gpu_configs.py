"""GPU parity cases at the full sizes of the benchmark configurations that are
not the bench line itself.

* C3, mixed precision: 3D anisotropic 7-pt diffusion on 160^3 (couplings 1,
  1, 0.01; 4,096,000 rows), b ~ N(0,1) (Park-Miller seed 2), the algebraic
  hierarchy (HMIS, extended+i, P max 4, l1-Jacobi) with FP64 on the finest level, FP32 on levels 1-2 and FP16 below
  (FP32 accumulation), FP64 Krylov vectors: the true relative residual
  ||b - A x|| / ||b||, recomputed on the host in FP64, is below 1e-10 for the
  FP64 and the mixed preconditioner, and the two solutions agree to 1e-6
  (relative, max norm; both carry the 1e-10 residual times the anisotropic
  operator's conditioning).
* C4a, multi-RHS SpMM: 3D 27-pt Poisson 64^3 (262,144 rows, 6,859,000 nnz) with
  k = 1, 8 and 32 right-hand sides ~ N(0,1) (seed 3), column-major and
  interleaved layouts, against scipy's CSR product: within 1e-12 of the
  |A| |X| row scale.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from amgt_amd import amg as G
from amgt_amd import workloads as W
from amgt_amd.mbsr import MBSRMatrix

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def test_c3_aniso_mixed_precision(dev):
    n1 = 160
    I, J, V = W.stencil(7, n1, n1, n1, c=(1.0, 1.0, 0.01))
    n = n1 ** 3
    b = W.rand_normal(2, n)
    A = sp.csr_matrix((V, J, I), shape=(n, n))
    sols = {}
    for name, pol_v, pol_g in (("fp64", G.fp64_policy, G.fp64_policy),
                               ("mixed", G.mixed_spmv_policy, G.mixed_spgemm_policy)):
        print(f"[c3] {name}", flush=True)
        amg = G.BoomerAMG(T(I, dev), T(J, dev), T(V, dev), n, dev, spmv_policy=pol_v,
                          spgemm_policy=pol_g)
        if name == "mixed":
            precs = [L_.A.info()["precision"] for L_ in amg.levels]
            assert precs[0] == 0 and all(p == 1 for p in precs[1:3]) and all(p == 2 for p in precs[3:])
        amg.capture()
        x, iters, hist = amg.pcg(T(b, dev), tol=1e-10, max_iter=200)
        xh = x.cpu().numpy()
        true = np.linalg.norm(b - A @ xh) / np.linalg.norm(b)
        print(f"[c3] {name}: {iters} iterations, true relative residual {true:.3e}", flush=True)
        assert true < 1e-10, (name, true)
        sols[name] = xh
        del amg
    diff = np.max(np.abs(sols["mixed"] - sols["fp64"])) / np.max(np.abs(sols["fp64"]))
    print(f"[c3] mixed vs fp64 solution: {diff:.3e}", flush=True)
    assert diff < 1e-6


@pytest.mark.parametrize("k", [1, 8, 32])
def test_c4a_spmm_full_size(dev, k):
    n1 = 64
    I, J, V = W.stencil(27, n1, n1, n1)
    n = n1 ** 3
    A = sp.csr_matrix((V, J, I), shape=(n, n))
    M = MBSRMatrix.from_csr(T(I, dev), T(J, dev), T(V, dev), n, n)
    X = W.rand_normal(3, n * k).reshape(k, n)
    ref = (A @ X.T).T
    scale = (abs(A) @ np.abs(X).T).T + 1e-300
    Y = M.spmm(T(X, dev)).cpu().numpy()
    assert np.all(np.abs(Y - ref) <= 1e-12 * scale)
    Yi = M.spmm_interleaved(T(X.T.copy(), dev)).cpu().numpy()
    assert np.all(np.abs(Yi.T - ref) <= 1e-12 * scale)
