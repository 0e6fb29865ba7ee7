"""GPU tests of the V-cycle / PCG driver (amgt_amd.amg) built on the library.

* coarse operators: each level's A is the Galerkin product R (A P) of the level
  above, checked against the oracle's block SpGEMM (structure bit-exact, values
  within 1e-12 of the |R| |A| |P| scale);
* one V(1,1) cycle against the same steps written with scipy in FP64 (1e-12
  relative to the vector's max norm);
* the captured hipGraph: several replays, with eager work in between, each
  equal to the eager cycle bit for bit;
* PCG: FP64 reaches 1e-8 in the same number of iterations with and without the
  graph; the mixed-precision preconditioner also reaches 1e-8 (FP64 Krylov
  vectors), with its solution within 1e-6 of the FP64 one.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from amgt_amd import _lib as L
from amgt_amd import amg as G
from amgt_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def to_sp(M):
    I, J, V = M.to_csr()
    n, m = M.shape
    return sp.csr_matrix((V.cpu().numpy(), J.cpu().numpy(), I.cpu().numpy()), shape=(n, m))


def build(dev, grid, **kw):
    I, J, V = W.stencil(7, *grid)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    return G.GeometricAMG(t(I), t(J), t(V), grid, dev, **kw), (I, J, V)


def test_galerkin_hierarchy(dev):
    amg, (I, J, V) = build(dev, (12, 10, 9), coarse_size=20)
    assert len(amg.levels) >= 3
    for lvl in range(len(amg.levels) - 1):
        Lf, Lc = amg.levels[lvl], amg.levels[lvl + 1]
        A, P, R = to_sp(Lf.A), to_sp(Lf.P), to_sp(Lf.R)
        for M in (A, P, R):
            M.sort_indices()
        mA = O.csr2mbsr(A.shape[0], A.shape[1], A.indptr, A.indices, A.data)
        mP = O.csr2mbsr(P.shape[0], P.shape[1], P.indptr, P.indices, P.data)
        mR = O.csr2mbsr(R.shape[0], R.shape[1], R.indptr, R.indices, R.data)
        ref = O.mbsr_spgemm(mR, O.mbsr_spgemm(mA, mP))
        RI, RJ, RV = O.mbsr2csr(ref)
        C = to_sp(Lc.A)
        C.sort_indices()
        assert np.array_equal(C.indptr, RI) and np.array_equal(C.indices, RJ)
        scale = (abs(R) @ abs(A) @ abs(P)).tocsr()
        scale.sort_indices()
        # |R||A||P| may hold entries the bitmap product drops; compare on C's pattern
        sc = np.asarray(scale[np.repeat(np.arange(C.shape[0]), np.diff(C.indptr)), C.indices]).ravel()
        assert np.all(np.abs(C.data - RV) <= 1e-12 * sc + 1e-300)
        # and R is P^T
        assert abs(R - P.T).max() == 0.0


def vcycle_ref(amg, f):
    """The same V(1,1) steps with scipy in FP64."""
    mats = [(to_sp(L_.A), to_sp(L_.P) if L_.P else None, to_sp(L_.R) if L_.R else None,
             L_.dinv.cpu().numpy()) for L_ in amg.levels]
    Ainv = to_sp(amg.coarse_inv).toarray()
    w = amg.omega

    def cyc(l, f):
        A, P, R, d = mats[l]
        if l == len(mats) - 1:
            return Ainv @ f
        u = w * d * f
        r = f - A @ u
        u = u + P @ cyc(l + 1, R @ r)
        return u + w * d * (f - A @ u)
    return cyc(0, f)


@pytest.mark.parametrize("omega", [1.0, 0.8])
def test_vcycle_matches_scipy(dev, omega):
    amg, _ = build(dev, (16, 12, 10), coarse_size=64, omega=omega)
    f = np.random.default_rng(1).standard_normal(amg.levels[0].n)
    z = torch.empty(amg.levels[0].n, dtype=torch.float64, device=dev)
    got = amg.precondition(torch.from_numpy(f).to(dev), z).cpu().numpy()
    exp = vcycle_ref(amg, f)
    assert np.max(np.abs(got - exp)) <= 1e-12 * np.max(np.abs(exp))


def test_graph_replays_match_eager(dev):
    amg, _ = build(dev, (16, 16, 16), coarse_size=64)
    n = amg.levels[0].n
    rng = np.random.default_rng(2)
    fs = [torch.from_numpy(rng.standard_normal(n)).to(dev) for _ in range(3)]
    z = torch.empty(n, dtype=torch.float64, device=dev)
    eager = [amg.precondition(f, z).clone() for f in fs]
    amg.capture()
    for rep in range(3):
        for f, e in zip(fs, eager):
            got = amg.precondition(f, z).clone()
            torch.cuda.synchronize()
            assert torch.equal(got, e), rep
            # unrelated eager work between replays
            amg.levels[0].A.spmv(f, torch.empty_like(f))
            torch.cuda.synchronize()


def test_pcg_fp64_and_mixed(dev):
    nx, ny, nz = 24, 20, 18
    I, J, V = W.stencil(7, nx, ny, nz)
    n = nx * ny * nz
    xs = W.rand_normal(1, n)
    b = sp.csr_matrix((V, J, I), shape=(n, n)) @ xs
    r64 = G.solve_poisson(nx, ny, nz, b, dev, mixed=False, capture=False, coarse_size=64)
    r64g = G.solve_poisson(nx, ny, nz, b, dev, mixed=False, capture=True, coarse_size=64)
    assert r64["hist"][-1] < 1e-8 and r64g["hist"][-1] < 1e-8
    assert r64["iters"] == r64g["iters"]
    assert r64["iters"] <= 15
    rmx = G.solve_poisson(nx, ny, nz, b, dev, mixed=True, capture=True, coarse_size=64)
    assert rmx["hist"][-1] < 1e-8
    assert rmx["iters"] <= r64["iters"] + 2
    x64, xmx = r64["x"].cpu().numpy(), rmx["x"].cpu().numpy()
    assert np.max(np.abs(xmx - x64)) <= 1e-6 * np.max(np.abs(x64))
    # levels below the finest run in FP32 / FP16
    pol = [d["spmv_precision"] for d in rmx["amg"].describe()]
    assert pol[0] == L.AMGT_DOUBLE and L.AMGT_FLOAT in pol


def test_batched_interpolation_is_block_diagonal(dev):
    """amgt_interp_trilinear on a batch equals the block diagonal of the
    per-grid numpy construction (row pointers, columns and values exactly)."""
    import scipy.sparse as sp
    grids = [(5, 4, 3), (2, 3, 1), (7, 1, 1), (1, 1, 1), (9, 8, 2)]
    I, J, A, nc, cg = G.interpolation_batched(grids, dev)
    blocks = []
    for g in grids:
        PI, PJ, PA, pnc = W.trilinear_interpolation(*g)
        blocks.append(sp.csr_matrix((PA, PJ, PI), shape=(int(np.prod(g)), pnc)))
    ref = sp.block_diag(blocks, format="csr")
    ref.sort_indices()
    assert nc == ref.shape[1] and cg == [tuple((x + 1) // 2 for x in g) for g in grids]
    assert np.array_equal(I.cpu().numpy(), ref.indptr)
    assert np.array_equal(J.cpu().numpy(), ref.indices)
    assert np.array_equal(A.cpu().numpy(), ref.data)


@pytest.mark.parametrize("grid", [(5, 4, 3), (12, 10, 9), (33, 17, 5), (2, 3, 1)])
def test_device_interpolation_and_transpose(dev, grid):
    """The device-built trilinear P equals workloads' numpy construction; the
    mBSR transpose handle equals P^T."""
    PI, PJ, PA, nc = W.trilinear_interpolation(*grid)
    DI, DJ, DA, dnc = G.trilinear_interpolation_device(grid, dev)
    assert dnc == nc
    assert np.array_equal(DI.cpu().numpy(), PI) and np.array_equal(DJ.cpu().numpy(), PJ)
    assert np.array_equal(DA.cpu().numpy(), PA)
    from amgt_amd.mbsr import MBSRMatrix
    n = int(np.prod(grid))
    P = MBSRMatrix.from_csr(DI, DJ, DA, n, nc)
    R = to_sp(P.transpose())
    ref = sp.csr_matrix((PA, PJ, PI), shape=(n, nc)).T.tocsr()
    ref.sort_indices()
    assert np.array_equal(R.indptr, ref.indptr) and np.array_equal(R.indices, ref.indices)
    assert np.array_equal(R.data, ref.data)


def test_batch_blockdiag(dev):
    """Block-diagonal batch of 2D problems: each level is the block-diagonal of
    the problems' own Galerkin operators; one V-cycle equals the scipy steps;
    PCG brings every problem's own residual below 1e-8."""
    sizes = [7, 16, 3, 12, 2, 9]
    r = G.solve_batch(sizes, 4, dev, capture=True, coarse_size=8)
    amg = r["amg"]
    assert len(amg.levels) >= 3
    # level 1 = block-diagonal of the problems' own Galerkin operators, each
    # from the oracle's host SpGEMM on the host restatement of the trilinear P
    blocks = []
    for n in sizes:
        I, J, V = W.stencil(7, n, n, 1)
        PI, PJ, PA, nc = W.trilinear_interpolation(n, n, 1)
        RI, RJ, RA = O.csr_transpose(n * n, nc, PI, PJ, PA)
        AP = O.csr_spgemm(n * n, nc, I, J, V, PI, PJ, PA, False)
        CI, CJ, CA = O.csr_spgemm(nc, nc, RI, RJ, RA, *AP, True)
        blocks.append(sp.csr_matrix((CA, CJ, CI), shape=(nc, nc)))
    ref = sp.block_diag(blocks, format="csr")
    ref.sort_indices()
    got = to_sp(amg.levels[1].A)
    got.sort_indices()
    assert np.array_equal(got.indptr, ref.indptr) and np.array_equal(got.indices, ref.indices)
    assert np.max(np.abs(got.data - ref.data)) <= 1e-12 * np.max(np.abs(ref.data))
    # V-cycle against scipy
    f = np.random.default_rng(3).standard_normal(amg.levels[0].n)
    amg._graph = None
    z = torch.empty(amg.levels[0].n, dtype=torch.float64, device=dev)
    got_z = amg.precondition(torch.from_numpy(f).to(dev), z).cpu().numpy()
    exp_z = vcycle_ref(amg, f)
    assert np.max(np.abs(got_z - exp_z)) <= 1e-12 * np.max(np.abs(exp_z))
    # every problem converged
    I, J, V, b, xs, grids, seg = G.batch_poisson2d(sizes, 4, dev)
    A = sp.csr_matrix((V, J, I), shape=(len(seg), len(seg)))
    x = r["x"].cpu().numpy()
    res = b - A @ x
    for k in range(len(sizes)):
        m = seg == k
        assert np.linalg.norm(res[m]) <= 1e-8 * np.linalg.norm(b[m]), k


def test_batch_c4b_full_size(dev):
    """C4(b) at its configured size: the 48 problems with sides drawn from
    U{32..512} (Park-Miller seed 4), one block-diagonal hierarchy, one PCG.
    For the smallest, the largest and three random problems, the problem's
    block of the level-1 operator equals the oracle's R (A P) (HYPRE's host
    SpGEMM) on that problem's own A and P (the host restatement of the
    trilinear P): structure exact, values within 1e-12 of |R||A||P|.  Every
    problem's true relative residual, recomputed on the host, is below 1e-8."""
    sizes = G.batch_sizes(48, 4)
    assert len(sizes) == 48 and min(sizes) >= 32 and max(sizes) <= 512
    print(f"[c4b] {len(sizes)} problems, {sum(n * n for n in sizes)} rows", flush=True)
    r = G.solve_batch(sizes, 4, dev, capture=True)
    amg = r["amg"]
    L1 = to_sp(amg.levels[1].A).tocsr()
    L1.sort_indices()
    coarse = [((n + 1) // 2) ** 2 for n in sizes]
    offs = np.concatenate([[0], np.cumsum(coarse)])
    rng = np.random.default_rng(4)
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes))} |
                  {int(k) for k in rng.choice(len(sizes), 3, replace=False)})
    for k in pick:
        n = sizes[k]
        I, J, V = W.stencil(7, n, n, 1)
        PI, PJ, PA, nc = W.trilinear_interpolation(n, n, 1)
        assert nc == coarse[k]
        RI, RJ, RA = O.csr_transpose(n * n, nc, PI, PJ, PA)
        AP = O.csr_spgemm(n * n, nc, I, J, V, PI, PJ, PA, False)
        CI, CJ, CA = O.csr_spgemm(nc, nc, RI, RJ, RA, *AP, True)
        ref = sp.csr_matrix((CA, CJ, CI), shape=(nc, nc))
        ref.sort_indices()
        got = L1[offs[k]:offs[k + 1], offs[k]:offs[k + 1]].tocsr()
        got.sort_indices()
        # nothing of problem k's rows lies outside its diagonal block
        assert L1[offs[k]:offs[k + 1]].nnz == got.nnz, k
        assert np.array_equal(got.indptr, ref.indptr) and np.array_equal(got.indices, ref.indices), k
        A_ = sp.csr_matrix((V, J, I), shape=(n * n, n * n))
        P_ = sp.csr_matrix((PA, PJ, PI), shape=(n * n, nc))
        scale = (abs(P_.T) @ abs(A_) @ abs(P_)).tocsr()
        scale.sort_indices()
        sc = np.asarray(scale[np.repeat(np.arange(nc), np.diff(ref.indptr)), ref.indices]).ravel()
        assert np.all(np.abs(got.data - ref.data) <= 1e-12 * sc + 1e-300), k
    I, J, V, b, xs, grids, seg = G.batch_poisson2d(sizes, 4, dev)
    A = sp.csr_matrix((V, J, I), shape=(len(seg), len(seg)))
    x = r["x"].cpu().numpy()
    res = b - A @ x
    worst = 0.0
    for k in range(len(sizes)):
        m = seg == k
        rel = np.linalg.norm(res[m]) / np.linalg.norm(b[m])
        worst = max(worst, rel)
        assert rel < 1e-8, (k, sizes[k], rel)
    print(f"[c4b] {r['iters']} iterations, worst problem residual {worst:.2e}", flush=True)


@pytest.mark.parametrize("hier", ["geometric", "boomer", "batch"])
def test_graph_pcg_matches_eager_pcg(dev, hier):
    """PCG on the library's dot and update kernels (scalars on the device)
    against the torch-operation PCG around the same captured V-cycle: the same
    iterations, relative residual histories within 1e-6 of each other,
    solutions within 1e-10; a second solve repeats the first bit for bit
    (deterministic dots)."""
    import scipy.sparse as sp2
    segments = None
    if hier == "batch":
        sizes = [9, 17, 12]
        I, J, V, bh, xs, grids, seg = G.batch_poisson2d(sizes, 4, dev)
        amg = G.GeometricAMG(T(I, dev), T(J, dev), T(V, dev), grids, dev, coarse_size=16)
        segments = T(seg, dev)
        b = T(bh, dev)
    else:
        nx = 20
        I, J, V = O.stencil(7, nx, nx, nx)
        n = nx ** 3
        xs = O.rand_normal(5, n)
        b = T(sp2.csr_matrix((V, J, I), shape=(n, n)) @ xs, dev)
        if hier == "geometric":
            amg = G.GeometricAMG(T(I, dev), T(J, dev), T(V, dev), (nx, nx, nx), dev,
                                 coarse_size=64)
        else:
            amg = G.BoomerAMG(T(I, dev), T(J, dev), T(V, dev), n, dev)
    amg.capture()
    x1, it1, h1 = amg.pcg(b, tol=1e-8, segments=segments)
    x2, it2, h2 = amg.pcg_eager(b, tol=1e-8, segments=segments)
    x3, it3, h3 = amg.pcg(b, tol=1e-8, segments=segments)
    assert it1 == it2 and h1[-1] < 1e-8
    assert np.allclose(h1, h2, rtol=1e-6, atol=0)
    scale = np.abs(x2.cpu().numpy()).max()
    assert np.abs(x1.cpu().numpy() - x2.cpu().numpy()).max() <= 1e-10 * scale
    assert it3 == it1 and torch.equal(x3, x1) and h3 == h1
