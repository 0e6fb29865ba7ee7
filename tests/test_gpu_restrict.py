"""GPU parity of the residual and restriction in one launch
(amgt_mbsr_residual_restrict, the V-cycle's r = f - A u and f_c = R r with
R = P^T) against the CPU oracle's HYPRE matvecs (csr_matvec.c:22 for A,
csr_matvec.c:1158's transposed form for P^T), through the C-ABI.

Bars: r within 1e-12 relative to |f| + |A| |u| per row; f_c within 1e-12
relative to |P|^T (|f| + |A| |u|) per coarse row (FP64); every lanes setting
(tiles of 8 to 128 block rows, so coarse rows span from one to many tiles);
the same bits on repeated calls (each coarse row is summed in R's order by
whichever tile arrives last); empty coarse rows give 0; and whole solves with
the fused launch take the two-launch path's iterations, with solutions that
agree to 1e-10.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from amgt_amd import _lib as L
from amgt_amd import amg as G
from amgt_amd.mbsr import MBSRMatrix, RestrictPlan
from oracle import oracle as O
from tests.util_matrices import skewed_csr

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _rand_p(n, nc, seed, empty_cols=()):
    """P-like CSR (n x nc): 1-4 entries per fine row, random order; no entry
    in the columns listed (empty coarse rows of R)."""
    rng = np.random.default_rng(seed)
    allowed = np.setdiff1d(np.arange(nc), np.asarray(empty_cols, dtype=np.int64))
    rows = [rng.permutation(allowed)[:rng.integers(1, 5)] for _ in range(n)]
    PI = np.zeros(n + 1, np.int32)
    PI[1:] = np.cumsum([len(r) for r in rows])
    return PI, np.concatenate(rows).astype(np.int32), rng.standard_normal(int(PI[-1]))


def _cases():
    I7, J7, A7 = O.stencil(7, 20, 18, 17)
    n7 = 20 * 18 * 17
    Is, Js, As = skewed_csr(4003, 4003, (3, 500), 1500, 4, 31, empty_rows=(6, 4002))
    return {
        # a geometric level: trilinear P of the 7-pt stencil (every coarse row
        # spans several tiles: its fine support is +-1 plane away)
        "7pt_trilinear": (n7, I7, J7, A7, None),
        # long rows, empty rows, a random P with two empty coarse rows
        "skewed_randp": (4003, Is, Js, As, _rand_p(4003, 997, 7, empty_cols=(0, 500)) + (997,)),
    }


def _trilinear_p(dev, I, J, A, grid):
    amg = G.GeometricAMG(T(I, dev), T(J, dev), T(A, dev), grid, dev, coarse_size=64,
                         max_levels=2, fused_restrict=False)
    PI, PJ, PA = (t.cpu().numpy() for t in amg.levels[0].P.to_csr())
    return PI, PJ, PA, amg.levels[1].n


@pytest.mark.parametrize("name", ["7pt_trilinear", "skewed_randp"])
@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
def test_residual_restrict_matches_oracle(dev, name, lanes):
    print(f"[restrict] {name} lanes={lanes}", flush=True)
    n, I, J, A, Pc = _cases()[name]
    if Pc is None:
        Pc = _trilinear_p(dev, I, J, A, (20, 18, 17))
    PI, PJ, PA, nc = Pc
    M = MBSRMatrix.from_csr(T(I, dev), T(J, dev), T(A, dev), n, n)
    M.set_spmv_lanes(lanes)
    P = MBSRMatrix.from_csr(T(PI, dev), T(PJ, dev), T(PA, dev), n, nc)
    R = P.transpose()
    plan = RestrictPlan(M, R)
    rng = np.random.default_rng(lanes)
    f, u = rng.standard_normal(n), rng.standard_normal(n)
    df, du = T(f, dev), T(u, dev)
    r = torch.empty(n, dtype=torch.float64, device=dev)
    fc = torch.full((nc,), 9.0, dtype=torch.float64, device=dev)
    M.residual_restrict(plan, df, du, r, fc)
    first = fc.clone()
    # oracle: r = f - A u, f_c = P^T r
    ref_r = O.csr_matvec(n, I, J, A, -1.0, u, 1.0, f)
    ref_fc = O.csr_matvec_t(n, nc, PI, PJ, PA, 1.0, ref_r)
    Aabs = sp.csr_matrix((np.abs(A), J, I), shape=(n, n))
    Pabs = sp.csr_matrix((np.abs(PA), PJ, PI), shape=(n, nc))
    scale_r = np.abs(f) + Aabs @ np.abs(u)
    scale_c = Pabs.T @ scale_r
    assert np.all(np.abs(r.cpu().numpy() - ref_r) <= 1e-12 * scale_r + 1e-300)
    assert np.all(np.abs(fc.cpu().numpy() - ref_fc) <= 1e-12 * scale_c + 1e-300)
    if name == "skewed_randp":
        assert fc[0].item() == 0.0 and fc[500].item() == 0.0
    # repeated calls: the counters are back at zero, the sums the same bits
    for _ in range(3):
        fc.fill_(-1.0)
        M.residual_restrict(plan, df, du, r, fc)
        assert torch.equal(fc, first)


def test_residual_restrict_rejects_bad_arguments(dev):
    n, I, J, A, _ = _cases()["7pt_trilinear"]
    PI, PJ, PA, nc = _trilinear_p(dev, I, J, A, (20, 18, 17))
    M = MBSRMatrix.from_csr(T(I, dev), T(J, dev), T(A, dev), n, n)
    P = MBSRMatrix.from_csr(T(PI, dev), T(PJ, dev), T(PA, dev), n, nc)
    plan = RestrictPlan(M, P.transpose())
    v = torch.zeros(n + 2, dtype=torch.float64, device=dev)
    fc = torch.zeros(nc, dtype=torch.float64, device=dev)
    lib = L.load()
    import ctypes
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # u misaligned, r aliasing u: refused before any launch
    st = lib.amgt_mbsr_residual_restrict(M.handle, plan.handle, ctypes.c_void_p(v.data_ptr()),
                                         ctypes.c_void_p(v.data_ptr() + 8),
                                         ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(
                                             fc.data_ptr()), s)
    assert st == L.AMGT_ERR_ARG
    # a matrix of another size than the plan's
    I2, J2, A2 = O.stencil(7, 5, 5, 5)
    M2 = MBSRMatrix.from_csr(T(I2, dev), T(J2, dev), T(A2, dev), 125, 125)
    w = torch.zeros(125, dtype=torch.float64, device=dev)
    st = lib.amgt_mbsr_residual_restrict(M2.handle, plan.handle, ctypes.c_void_p(w.data_ptr()),
                                         ctypes.c_void_p(w.data_ptr()),
                                         ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(
                                             fc.data_ptr()), s)
    assert st == L.AMGT_ERR_ARG and b"plan" in lib.amgt_last_error()


@pytest.mark.parametrize("hier", ["geometric", "boomer"])
def test_solves_with_fused_restriction_match_two_launches(dev, hier):
    """PCG with the V-cycle (hipGraph-captured) using the one-launch residual +
    restriction against the two-launch cycle: the same iterations, solutions
    within 1e-10; the fused cycle's graph replays equal its eager cycle bit
    for bit."""
    nx = 24
    I, J, A = O.stencil(7, nx, nx, nx)
    n = nx ** 3
    xs = O.rand_normal(3, n)
    b = T(sp.csr_matrix((A, J, I), shape=(n, n)) @ xs, dev)
    out = {}
    for fused in (False, True):
        if hier == "geometric":
            amg = G.GeometricAMG(T(I, dev), T(J, dev), T(A, dev), (nx, nx, nx), dev,
                                 coarse_size=64, fused_restrict=fused)
        else:
            amg = G.BoomerAMG(T(I, dev), T(J, dev), T(A, dev), n, dev, fused_restrict=fused)
        assert all((L_.plan is not None) == fused for L_ in amg.levels[:-1])
        z_eager = torch.empty_like(b)
        amg.precondition(b, z_eager)
        amg.capture()
        z_graph = torch.empty_like(b)
        for _ in range(3):
            amg.precondition(b, z_graph)
            torch.cuda.synchronize()
            assert torch.equal(z_graph, z_eager)
        x, it, hist = amg.pcg(b, tol=1e-8)
        out[fused] = (x.cpu().numpy(), it, hist[-1])
    assert out[True][1] == out[False][1]
    assert np.max(np.abs(out[True][0] - out[False][0])) <= 1e-10 * np.max(np.abs(xs))
    assert out[True][2] < 1e-8
