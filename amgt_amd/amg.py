"""V-cycle and PCG on the mBSR path: the callers of the SpMV and SpGEMM.

This is the solve side that drives the library the way BoomerAMG does
(hypre_BoomerAMGCycle, par_cycle.c:443,704,751; relax type 18,
par_relax.c:338-355; PCG, krylov/pcg.c), so that setup+solve can be timed in
FP64 and in mixed precision on one MI355X.  The hierarchy is geometric, not
BoomerAMG's: P is trilinear interpolation to the grid of even-index points
(coarsening by 2 per direction), built on the device, R = P^T as the library's
mBSR transpose, and every coarse operator is the
Galerkin product R (A P) computed by the library's mBSR SpGEMM, as
hypre_BoomerAMGBuildCoarseOperatorKT / RAPKT does on the GPU
(par_csr_triplemat_device.c:844).  BoomerAMG below builds the algebraic
hierarchy instead (strength, HMIS/PMIS, extended+i on the device).

Per level: A, P, R as mBSR handles tagged with the level, the inverse of the
smoother's l1 vector, and fixed work vectors.  V(1,1):
    u  = w D^-1 f                      (first sweep from u = 0: no matvec, as HYPRE)
    r  = f - A u,  f' = R r            (two SpMVs; one launch of
                                        amgt_mbsr_residual_restrict with
                                        fused_restrict=True, correct but slower
                                        on MI355X: its cross-tile arrivals cost
                                        more than the second launch)
    coarse V-cycle -> u'
    u += P u'                          (SpMV with b aliasing y)
    u2 = u + w D^-1 (f - A u)          (one fused SpMV launch, amgt_mbsr_relax_jacobi)
The coarsest level (<= coarse_size rows) applies its dense inverse, held as
an mBSR matrix with full blocks, with one SpMV (HYPRE: Gaussian elimination,
relax type 9).  Every step is a library kernel on the current stream, so one
V-cycle can be captured in a hipGraph (torch.cuda.CUDAGraph) and replayed.

Smoother: relax type 18 (l1-Jacobi, HYPRE's GPU default) or 7 (Jacobi), weight
w = 1 by default as in HYPRE.  D is what hypre_ParCSRComputeL1Norms gives
(ams.c:527-800): for relax 18 the l1 norm of the whole row, sum_j |a_ij|
(option 1, par_amg_setup.c:3393-3406), for relax 7 the diagonal (option 5,
:3521-3536).

Precision (mixed mode): per-level policies for SpMV/smoothing and for the
Galerkin SpGEMM; the defaults follow the mixed configuration of the workload
definition: FP64 on the finest level, FP32 on levels 1-2, FP16 below (FP32
accumulation), FP64 SpGEMM on the two finest levels and the coarsest solve in
FP64.  PCG itself (level 0 vectors, dot products, updates) is FP64.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from . import workloads as W
from .mbsr import MBSRMatrix, RestrictPlan


def _excl_cumsum(v: torch.Tensor) -> torch.Tensor:
    out = torch.zeros_like(v)
    out[1:] = torch.cumsum(v[:-1], 0)
    return out


def interpolation_batched(grids, device):
    """CSR (I, J, A) of the block-diagonal trilinear prolongation of a batch of
    independent grids (problem k's rows and coarse columns follow problem
    k-1's), built on the device by amgt_interp_trilinear in one pass over all
    rows.

    Per grid and direction the coarse points are the even fine indices
    (nc = (n + 1) // 2); a fine index takes weight 1 from its even point, or
    1/2 from each even neighbour, a neighbour beyond the last coarse point
    being a zero Dirichlet value.  Up to 8 entries per row, ascending columns,
    no stored zeros.  Returns (I, J, A, ncoarse, coarse grids)."""
    import ctypes
    from .mbsr import _stream_ptr, copy_from_device_ptr
    flat = (ctypes.c_int32 * (3 * len(grids)))(*[int(v) for g in grids for v in g])
    nf, nc, nnz = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
    pi, pj, pa = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    L.check(L.load().amgt_interp_trilinear(len(grids), flat, ctypes.byref(nf), ctypes.byref(nc),
                                           ctypes.byref(nnz), ctypes.byref(pi), ctypes.byref(pj),
                                           ctypes.byref(pa), _stream_ptr(None)),
            "amgt_interp_trilinear")
    I = copy_from_device_ptr(pi.value, nf.value + 1, torch.int32, device)
    J = copy_from_device_ptr(pj.value, nnz.value, torch.int32, device)
    A = copy_from_device_ptr(pa.value, nnz.value, torch.float64, device)
    cgrids = [tuple((int(x) + 1) // 2 for x in gr) for gr in grids]
    return I, J, A, int(nc.value), cgrids


def trilinear_interpolation_device(grid, device):
    """CSR (I, J, A) of the trilinear prolongation from the even-index grid to
    the full nx x ny x nz grid, built on the device (same arrays as
    workloads.trilinear_interpolation).  Returns (I, J, A, ncoarse)."""
    I, J, A, nc, _ = interpolation_batched([tuple(grid)], device)
    return I, J, A, nc


def interpolation_blockdiag(grids, device):
    """Block-diagonal trilinear P for a batch of independent grids; see
    interpolation_batched.  Returns (I, J, A, ncoarse, coarse grids)."""
    return interpolation_batched(grids, device)


# smoother diagonal per relax type (amgt_mbsr_diagonal kinds)
RELAX_DIAG = {18: L.AMGT_DIAG_INV_L1, 7: L.AMGT_DIAG_INV}


def fp64_policy(level: int) -> int:
    return L.AMGT_DOUBLE


def mixed_spmv_policy(level: int) -> int:
    """FP64 finest, FP32 levels 1-2, FP16 below."""
    if level == 0:
        return L.AMGT_DOUBLE
    return L.AMGT_FLOAT if level <= 2 else L.AMGT_HALF


def mixed_spgemm_policy(level: int) -> int:
    """Galerkin products in FP64 on the two finest levels, FP32 below."""
    return L.AMGT_DOUBLE if level <= 1 else L.AMGT_FLOAT


@dataclass
class Level:
    n: int
    grid: list  # the grid of each problem on this level
    A: MBSRMatrix
    dinv: torch.Tensor
    P: MBSRMatrix | None = None
    R: MBSRMatrix | None = None
    f: torch.Tensor | None = None
    u: torch.Tensor | None = None
    u2: torch.Tensor | None = None
    r: torch.Tensor | None = None
    plan: object = None  # RestrictPlan of the fused residual + restriction


@dataclass
class SetupTimes:
    host_s: float = 0.0          # coarsest dense inverse (numpy)
    device_s: float = 0.0        # P, R = P^T, conversions, Galerkin SpGEMMs, diagonals
    total_s: float = 0.0
    spgemm_s: list = field(default_factory=list)
    coarsen_s: list = field(default_factory=list)  # strength + PMIS + interpolation per level


def stalled(hist: list, stall) -> bool:
    """stall = (window w, rate q): the last w iterations reduced the relative
    residual by less than q^w (mean rate per iteration above q)."""
    if stall is None:
        return False
    w, q = stall
    return len(hist) > 2 * w and hist[-1] > hist[-1 - w] * q ** w


class SegmentNorms:
    """Two-norms of the row segments of vectors: problem k of a batch owns the
    rows with segments == k, which are contiguous and ascending in k.

    Rows are gathered once per call into chunks that never straddle two
    problems (each problem padded to a whole number of chunks with a zero), the
    chunks are summed densely and the few chunk sums are added per problem, so
    no reduction funnels millions of atomics into a handful of addresses."""

    def __init__(self, segments: torch.Tensor, chunk: int = 2048):
        seg = segments.to(torch.int64)
        dev = seg.device
        n = seg.numel()
        if n and bool((seg[1:] < seg[:-1]).any().item()):
            raise ValueError("segments must be ascending (contiguous problems)")
        self.nseg = int(seg.max().item()) + 1 if n else 0
        lengths = torch.bincount(seg, minlength=self.nseg)
        nch = (lengths + chunk - 1) // chunk
        total = int(nch.sum().item())
        self.chunk_seg = torch.repeat_interleave(torch.arange(self.nseg, device=dev), nch,
                                                 output_size=total)
        slot = torch.arange(total * chunk, device=dev)
        c = slot // chunk
        k = self.chunk_seg[c]
        local = (c - _excl_cumsum(nch)[k]) * chunk + slot % chunk
        self.gidx = torch.where(local < lengths[k], _excl_cumsum(lengths)[k] + local,
                                torch.full_like(local, n))
        self.chunk = chunk
        self.n = n

    def __call__(self, v: torch.Tensor) -> torch.Tensor:
        sq = torch.zeros(self.n + 1, dtype=v.dtype, device=v.device)
        torch.mul(v, v, out=sq[:self.n])
        part = sq[self.gidx].view(-1, self.chunk).sum(1)
        return torch.zeros(self.nseg, dtype=v.dtype, device=v.device).index_add_(
            0, self.chunk_seg, part).sqrt_()


def _check_relax(relax_type: int) -> int:
    if relax_type not in RELAX_DIAG:
        raise ValueError(f"relax type {relax_type}: 7 (Jacobi) or 18 (l1-Jacobi)")
    return int(relax_type)


class GeometricAMG:
    """Galerkin hierarchy on a structured nx x ny x nz grid, or on a batch of
    independent grids held as one block-diagonal system (a list of grids), and
    its V(1,1) cycle.  In a batch every grid coarsens by 2 per level, so each
    level is one block-diagonal operator: one launch per step for all the
    problems."""

    def __init__(self, I, J, V, grid, device, omega: float = 1.0, coarse_size: int = 512,
                 spmv_policy=fp64_policy, spgemm_policy=fp64_policy, max_levels: int = 25,
                 relax_type: int = 18, fused_restrict: bool = False):
        self.device = device
        self.fused_restrict = bool(fused_restrict)
        self.omega = float(omega)
        self.relax_type = _check_relax(relax_type)
        self.coarse_size = int(coarse_size)
        self.spmv_policy = spmv_policy
        self.spgemm_policy = spgemm_policy
        self.max_levels = max_levels
        self.levels: list[Level] = []
        self.coarse_inv: MBSRMatrix | None = None
        self.times = SetupTimes()
        self._graph = None
        grids = [tuple(g) for g in grid] if isinstance(grid, (list, tuple)) and \
            len(grid) > 0 and isinstance(grid[0], (list, tuple)) else [tuple(grid)]
        self.batch = len(grids) > 1
        self._setup(I, J, V, grids)

    # ---- setup -------------------------------------------------------------------------
    def _t(self, a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.device)

    def _setup(self, I, J, V, grids):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        grid = grids
        n = int(sum(np.prod(g) for g in grids))
        A = MBSRMatrix.from_csr(I, J, V, n, n)
        lvl = 0
        while True:
            A.set_level(lvl)
            A.set_precision(self.level_precision(lvl))
            dinv = A.diagonal(kind=RELAX_DIAG[self.relax_type])
            L_ = Level(n=n, grid=grid, A=A, dinv=dinv)
            self.levels.append(L_)
            if n <= self.coarse_size or lvl + 1 >= self.max_levels or \
                    all(max(g) <= 1 for g in grid):
                break
            PI, PJ, PA, nc, cgrid = interpolation_batched(grid, self.device)
            P = MBSRMatrix.from_csr(PI, PJ, PA, n, nc)
            R = P.transpose()
            del PI, PJ, PA
            Ac = self._galerkin(lvl, A, P, R)
            L_.P, L_.R = P, R
            A, n = Ac, nc
            grid = cgrid
            lvl += 1
        self._finish(t0)

    def level_precision(self, lvl: int) -> int:
        """SpMV / smoother precision of level lvl."""
        return self.spmv_policy(lvl)

    def _galerkin(self, lvl, A, P, R):
        """R (A P) in the level's SpGEMM precision (it follows the left
        operand); the operands return to the level's SpMV precision."""
        for M in (P, R):
            M.set_level(lvl)
            M.set_precision(self.level_precision(lvl))
        torch.cuda.synchronize()
        tg = time.perf_counter()
        A.set_precision(self.spgemm_policy(lvl))
        AP = A.spgemm(P)
        R.set_precision(self.spgemm_policy(lvl))
        Ac = R.spgemm(AP)
        torch.cuda.synchronize()
        self.times.spgemm_s.append(time.perf_counter() - tg)
        A.set_precision(self.level_precision(lvl))
        R.set_precision(self.level_precision(lvl))
        return Ac

    def _finish(self, t0):
        """Coarsest level's dense inverse, work vectors, setup times."""
        # coarsest: dense inverse as an mBSR matrix with full blocks, applied in FP64
        torch.cuda.synchronize()
        th = time.perf_counter()
        Lc = self.levels[-1]
        CI, CJ, CV = Lc.A.to_csr()
        dense = np.zeros((Lc.n, Lc.n))
        ci, cj, cv = CI.cpu().numpy(), CJ.cpu().numpy(), CV.cpu().numpy()
        rows = np.repeat(np.arange(Lc.n), np.diff(ci))
        dense[rows, cj] = cv
        inv = np.linalg.inv(dense)
        II = (np.arange(Lc.n + 1) * Lc.n).astype(np.int32)
        JJ = np.tile(np.arange(Lc.n, dtype=np.int32), Lc.n)
        self.coarse_inv = MBSRMatrix.from_csr(self._t(II), self._t(JJ), self._t(inv.ravel()),
                                              Lc.n, Lc.n)
        self.coarse_inv.set_precision(L.AMGT_DOUBLE)
        torch.cuda.synchronize()
        host = time.perf_counter() - th
        # tables of the one-launch residual + restriction of every level with a P
        if getattr(self, "fused_restrict", False):
            for L_ in self.levels[:-1]:
                L_.plan = RestrictPlan(L_.A, L_.R)
        # work vectors (fixed addresses, so a captured V-cycle stays valid)
        for L_ in self.levels:
            z = lambda: torch.zeros(L_.n, dtype=torch.float64, device=self.device)  # noqa: E731
            L_.f, L_.u, L_.u2, L_.r = z(), z(), z(), z()
        torch.cuda.synchronize()
        self.times.total_s = time.perf_counter() - t0
        self.times.host_s = host
        self.times.device_s = self.times.total_s - host

    # ---- cycle ------------------------------------------------------------------------
    def _vcycle(self, lvl: int) -> torch.Tensor:
        """Solves approximately A_lvl u = f_lvl; returns the tensor holding u."""
        L_ = self.levels[lvl]
        if lvl == len(self.levels) - 1:
            return self.coarse_inv.spmv(L_.f, L_.u)
        # pre-smoothing from u = 0: u = w D^-1 f
        torch.mul(L_.dinv, L_.f, out=L_.u)
        if self.omega != 1.0:
            L_.u.mul_(self.omega)
        # residual and restriction (one launch with the level's plan)
        Lc = self.levels[lvl + 1]
        if L_.plan is not None:
            L_.A.residual_restrict(L_.plan, L_.f, L_.u, L_.r, Lc.f)
        else:
            L_.A.spmv(L_.u, L_.r, alpha=-1.0, beta=1.0, b=L_.f)
            L_.R.spmv(L_.r, Lc.f)
        uc = self._vcycle(lvl + 1)
        # coarse-grid correction u += P uc, then post-smoothing into u2
        L_.P.spmv(uc, L_.u, alpha=1.0, beta=1.0, b=L_.u)
        L_.A.relax_jacobi(L_.f, L_.u, L_.dinv, self.omega, out=L_.u2)
        return L_.u2

    def precondition(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        """z = M^-1 r with one V(1,1) cycle (captured graph if built)."""
        L0 = self.levels[0]
        L0.f.copy_(r)
        if self._graph is not None:
            self._graph.replay()
            out = self._graph_out
        else:
            out = self._vcycle(0)
        z.copy_(out)
        return z

    def capture(self, warmup: int = 1):
        """Captures one V-cycle (on the fixed work vectors) in a hipGraph.  The
        warm-up cycles run first, eagerly, so that everything a first call
        allocates (low-precision copies, SpMV tile tables) exists before the
        capture.  The capture is begun and ended directly on a side stream:
        torch.cuda.graph's context manager also runs gc.collect() and
        empty_cache() (10-40 ms here, against 0.9 ms for capture and
        instantiation of the cycle), and the cycle allocates nothing."""
        import gc
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self._vcycle(0)
            # no collection during the capture: a handle freed then would
            # synchronise the device inside it
            gc_was_on = gc.isenabled()
            gc.disable()
            try:
                g.capture_begin()
                try:
                    out = self._vcycle(0)
                finally:
                    g.capture_end()
            finally:
                if gc_was_on:
                    gc.enable()
        torch.cuda.current_stream().wait_stream(s)
        self._graph, self._graph_out = g, out
        return g

    # ---- Krylov ------------------------------------------------------------------------
    def pcg(self, b: torch.Tensor, tol: float = 1e-8, max_iter: int = 200,
            segments: torch.Tensor | None = None, stall=None):
        """Preconditioned CG (hypre_PCGSolve, krylov/pcg.c:320-900: two-norm
        relative residual, as ij -tol with PCG).  With `segments` (problem index
        of every row of a batch), the test is every problem's own relative
        residual.  With a captured V-cycle (capture()) an iteration is the
        library's kernels with every scalar in device memory: Ap = A p, p.Ap,
        one launch for x += alpha p, r -= alpha Ap and the partials of r.r,
        (the host reads r.r for the stopping test), the V-cycle replay z = M r,
        r.z and one launch for p = z + beta p.  r and z are the V-cycle's own
        input and output vectors (no copies).  Without a captured cycle:
        pcg_eager.  `stall` = (w, q) stops early, unconverged, when the last w
        iterations reduced the residual by less than q^w (`stalled`)."""
        if self._graph is None:
            return self.pcg_eager(b, tol, max_iter, segments, stall)
        L0 = self.levels[0]
        n = L0.n
        st = self._pcg_state()
        lib = L.load()
        # z: the cycle's output vector (u2 of the finest level, u of a
        # one-level hierarchy)
        x, r, z, p, ap = st["x"], L0.f, self._graph_out, st["p"], st["ap"]
        sc, work = st["sc"], st["work"]
        ptr = lambda t, k=0: ctypes.c_void_p(t.data_ptr() + 8 * k)  # noqa: E731
        sp_ = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        # sc: 0 rz, 1 pAp, 2 rr, 3 rz_new, 4 bb
        x.zero_()
        r.copy_(b)
        L.check(lib.amgt_dot(n, ptr(b), ptr(b), ptr(sc, 4), ptr(work), sp_), "amgt_dot")
        if segments is not None:
            norms = SegmentNorms(segments)
            bseg = norms(b)
            bseg = torch.where(bseg > 0, bseg, torch.ones_like(bseg))
        bb = sc[4].item()
        if bb == 0.0:
            return x.clone(), 0, [0.0]
        self._graph.replay()
        p.copy_(z)
        L.check(lib.amgt_dot(n, ptr(r), ptr(z), ptr(sc, 0), ptr(work), sp_), "amgt_dot")
        hist = [1.0]
        it = 0
        while it < max_iter:
            L0.A.spmv(p, ap)
            L.check(lib.amgt_dot(n, ptr(p), ptr(ap), ptr(sc, 1), ptr(work), sp_), "amgt_dot")
            L.check(lib.amgt_pcg_update_xr(n, ptr(sc, 0), ptr(sc, 1), ptr(p), ptr(ap), ptr(x),
                                           ptr(r), ptr(sc, 2), ptr(work), sp_),
                    "amgt_pcg_update_xr")
            it += 1
            if segments is not None:
                rel = (norms(r) / bseg).max().item()
            else:
                rel = (max(sc[2].item(), 0.0) / bb) ** 0.5
            hist.append(rel)
            if rel < tol or stalled(hist, stall):
                break
            self._graph.replay()
            L.check(lib.amgt_dot(n, ptr(r), ptr(z), ptr(sc, 3), ptr(work), sp_), "amgt_dot")
            L.check(lib.amgt_pcg_update_p(n, ptr(sc, 3), ptr(sc, 0), ptr(z), ptr(p), sp_),
                    "amgt_pcg_update_p")
        return x.clone(), it, hist

    def _pcg_state(self) -> dict:
        """PCG vectors and device scalars of this hierarchy (allocated once)."""
        if getattr(self, "_pcg_st", None) is None:
            n = self.levels[0].n
            dev = self.device
            z = lambda: torch.zeros(n, dtype=torch.float64, device=dev)  # noqa: E731
            nw = int(L.load().amgt_pcg_work_bytes()) // 8
            self._pcg_st = {"x": z(), "p": z(), "ap": z(),
                            "sc": torch.zeros(8, dtype=torch.float64, device=dev),
                            "work": torch.zeros(nw, dtype=torch.float64, device=dev)}
        return self._pcg_st

    def pcg_eager(self, b: torch.Tensor, tol: float = 1e-8, max_iter: int = 200,
                  segments: torch.Tensor | None = None, stall=None):
        """PCG with torch vector operations around precondition() (the form
        used before capture(); the reference for the graph-replayed pcg)."""
        A = self.levels[0].A
        x = torch.zeros_like(b)
        r = b.clone()
        z = torch.empty_like(b)
        Ap = torch.empty_like(b)
        if segments is not None:
            norms = SegmentNorms(segments)
            bseg = norms(b)
            bseg = torch.where(bseg > 0, bseg, torch.ones_like(bseg))

            def relres():
                return (norms(r) / bseg).max().item()
        else:
            bn = torch.linalg.vector_norm(b).item()

            def relres():
                return torch.linalg.vector_norm(r).item() / bn
            if bn == 0.0:
                return x, 0, [0.0]
        self.precondition(r, z)
        p = z.clone()
        rz = torch.dot(r, z)
        hist = [1.0]
        it = 0
        while it < max_iter:
            A.spmv(p, Ap)
            alpha = rz / torch.dot(p, Ap)
            x.add_(alpha * p)
            r.sub_(alpha * Ap)
            it += 1
            rel = relres()
            hist.append(rel)
            if rel < tol or stalled(hist, stall):
                break
            self.precondition(r, z)
            rz_new = torch.dot(r, z)
            p.mul_(rz_new / rz).add_(z)
            rz = rz_new
        return x, it, hist

    def operator_complexity(self) -> float:
        nnz = [L_.A.info()["nnz"] for L_ in self.levels]
        return sum(nnz) / nnz[0]

    def describe(self) -> list:
        """Per level: rows, nnz, precisions and the SpMV kernel/lanes of A, P
        and R (chosen from each operator's statistics when it was built,
        amgt_mbsr_tune_from_stats)."""
        return [{"level": i, "rows": L_.n, "nnz": L_.A.info()["nnz"],
                 "spmv_precision": self.level_precision(i),
                 "spgemm_precision": self.spgemm_policy(i) if L_.P is not None else None,
                 "spmv_config": {tag: "%s_%d" % M.spmv_config()
                                 for tag, M in (("A", L_.A), ("P", L_.P), ("R", L_.R))
                                 if M is not None}}
                for i, L_ in enumerate(self.levels)]


class BoomerAMG(GeometricAMG):
    """Algebraic Galerkin hierarchy built on the device the way BoomerAMG builds
    it on one process (par_amg_setup.c:994-3182): per level the strength matrix
    (theta, max row sum), HMIS coarsening (coarsen type 10, HYPRE's default:
    the Ruge first pass on the host, PMIS with CF_init 1 on the device) or PMIS
    (coarsen="pmis"; CF_init 0: coarsen type 8, 2: type 9), extended+i
    interpolation truncated to max_elmts entries per row, R = P^T (mBSR
    transpose) and R (A P) on the mBSR SpGEMM.  Coarsening stops at n <=
    coarse_size, at max_levels, or when a level would not shrink (no C point or
    all C points); the coarsest level is solved with its dense inverse.  The
    defaults are ij's (test/ij.c:244-290, 356): theta 0.25, max row sum 1.0,
    P max 4, coarse size 9, relax type 18 with weight 1.  The V-cycle and PCG
    are the same as for the geometric hierarchy.

    `level_hook(lvl, n_prev, A, CI, CJ, CV) -> (theta, precision)`, when given,
    is called for every level before it is coarsened and sets that level's
    strength threshold and SpMV/smoother precision (replacing `theta` and
    `spmv_policy`); the Galerkin products keep `spgemm_policy`.

    `galerkin` picks the coarse-operator product: "mbsr" (default) is the
    mBSR SpGEMM R (A P), the product hypre_CSRMatrixMultiplyDevice runs; its
    sums run in another order than HYPRE's host RAP, so coarse levels agree
    with HYPRE's CPU hierarchy to round-off, and where round-off decides a tie
    (strength threshold, truncation order) the coarsening can differ.
    "host_order" forms R A P in HYPRE's host order in FP64
    (amgt_amg_rap_host_order) and keeps every level's CSR in HYPRE's layout
    (`csr_levels`), so the whole hierarchy equals HYPRE's CPU one bit for
    bit; `spgemm_policy` does not apply to it."""

    def __init__(self, I, J, V, n, device, theta: float = 0.25, max_row_sum: float = 1.0,
                 cf_init: int = 0, max_elmts: int = 4, omega: float = 1.0, coarse_size: int = 9,
                 spmv_policy=fp64_policy, spgemm_policy=fp64_policy, max_levels: int = 25,
                 level_hook=None, coarsen: str = "hmis", relax_type: int = 18,
                 galerkin: str = "mbsr", fused_restrict: bool = False):
        if coarsen not in ("hmis", "pmis"):
            raise ValueError(f"coarsen {coarsen!r}: 'hmis' or 'pmis'")
        if galerkin not in ("mbsr", "host_order"):
            raise ValueError(f"galerkin {galerkin!r}: 'mbsr' or 'host_order'")
        self.galerkin = galerkin
        self.fused_restrict = bool(fused_restrict)
        self.csr_levels: list = []
        self.p_csr: list = []
        self.device = device
        self.omega = float(omega)
        self.relax_type = _check_relax(relax_type)
        self.coarsen = coarsen
        self.coarse_size = int(coarse_size)
        self.level_hook = level_hook
        self.level_theta: list[float] = []
        self.level_prec: list[int] = []
        self.spmv_policy = spmv_policy
        self.spgemm_policy = spgemm_policy
        self.max_levels = max_levels
        self.theta, self.max_row_sum = float(theta), float(max_row_sum)
        self.cf_init, self.max_elmts = int(cf_init), int(max_elmts)
        self.levels: list[Level] = []
        self.coarse_inv: MBSRMatrix | None = None
        self.times = SetupTimes()
        self._graph = None
        self.batch = False
        self.cf: list[torch.Tensor] = []
        self._setup_algebraic(I, J, V, int(n))

    def _setup_algebraic(self, I, J, V, n):
        from . import coarsen as C
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        A = MBSRMatrix.from_csr(I, J, V, n, n)
        CI, CJ, CV = I, J, V
        lvl = 0
        n_prev = n
        while True:
            A.set_level(lvl)
            if self.level_hook is not None:
                theta, prec = self.level_hook(lvl, n_prev, A, CI, CJ, CV)
                self.level_theta.append(float(theta))
                self.level_prec.append(int(prec))
            else:
                theta = self.theta
                self.level_theta.append(theta)
            A.set_precision(self.level_precision(lvl))
            dinv = A.diagonal(kind=RELAX_DIAG[self.relax_type])
            L_ = Level(n=n, grid=[], A=A, dinv=dinv)
            self.levels.append(L_)
            if n <= self.coarse_size or lvl + 1 >= self.max_levels:
                break
            torch.cuda.synchronize()
            tc = time.perf_counter()
            SI, SJ = C.strength(n, CI, CJ, CV, theta, self.max_row_sum)
            if self.coarsen == "hmis":
                cf, nc = C.hmis(n, SI, SJ)
            else:
                cf, nc = C.pmis(n, SI, SJ, self.cf_init)
            if nc == 0 or nc == n:
                break
            PI, PJ, PA, nc = C.extpi_interp(n, CI, CJ, CV, SI, SJ, cf, self.max_elmts)
            del SI, SJ
            torch.cuda.synchronize()
            self.times.coarsen_s.append(time.perf_counter() - tc)
            self.cf.append(cf)
            P = MBSRMatrix.from_csr(PI, PJ, PA, n, nc)
            R = P.transpose()
            if self.galerkin == "host_order":
                self.csr_levels.append((CI, CJ, CV))
                self.p_csr.append((PI, PJ, PA))
                for M in (P, R):
                    M.set_level(lvl)
                    M.set_precision(self.level_precision(lvl))
                torch.cuda.synchronize()
                tg = time.perf_counter()
                RI_, RJ_, RA_ = C.rap_host_order(n, nc, CI, CJ, CV, PI, PJ, PA)
                Ac = MBSRMatrix.from_csr(RI_, RJ_, RA_, nc, nc)
                torch.cuda.synchronize()
                self.times.spgemm_s.append(time.perf_counter() - tg)
                L_.P, L_.R = P, R
                A, n_prev, n = Ac, n, nc
                CI, CJ, CV = RI_, RJ_, RA_
                lvl += 1
                continue
            del PI, PJ, PA
            Ac = self._galerkin(lvl, A, P, R)
            L_.P, L_.R = P, R
            A, n_prev, n = Ac, n, nc
            CI, CJ, CV = A.to_csr()
            lvl += 1
        if self.galerkin == "host_order":
            self.csr_levels.append((CI, CJ, CV))
        self._finish(t0)

    def level_precision(self, lvl: int) -> int:
        if self.level_hook is not None:
            return self.level_prec[lvl]
        return self.spmv_policy(lvl)

    def complexities(self):
        """(grid complexity, operator complexity) as BoomerAMG reports them."""
        ns = [L_.n for L_ in self.levels]
        nnz = [L_.A.info()["nnz"] for L_ in self.levels]
        return sum(ns) / ns[0], sum(nnz) / nnz[0]


def solve_poisson_boomer(nx: int, ny: int, nz: int, b: np.ndarray, device, mixed: bool = False,
                         tol: float = 1e-8, capture: bool = True, max_elmts: int = 4,
                         coarse_size: int = 9, theta: float = 0.25, max_row_sum: float = 1.0,
                         omega: float = 1.0, coarsen: str = "hmis", relax_type: int = 18,
                         galerkin: str = "mbsr"):
    """7-pt Poisson with the algebraic hierarchy (device strength, HMIS or
    PMIS, extended+i; ij's defaults): setup + PCG/V(1,1) solve; timings as
    solve_poisson.  `galerkin` as for BoomerAMG."""
    I, J, V = W.stencil(7, nx, ny, nz)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    dI, dJ, dV = t(I), t(J), t(V)
    db = t(b)
    torch.cuda.synchronize()
    amg = BoomerAMG(dI, dJ, dV, nx * ny * nz, device, theta=theta, max_row_sum=max_row_sum,
                    cf_init=0, max_elmts=max_elmts, omega=omega, coarse_size=coarse_size,
                    spmv_policy=mixed_spmv_policy if mixed else fp64_policy,
                    spgemm_policy=mixed_spgemm_policy if mixed else fp64_policy,
                    coarsen=coarsen, relax_type=relax_type, galerkin=galerkin)
    tc = time.perf_counter()
    if capture:
        amg.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, iters, hist = amg.pcg(db, tol=tol)
    torch.cuda.synchronize()
    return {"amg": amg, "x": x, "iters": iters, "hist": hist, "setup": amg.times,
            "capture_s": t0 - tc, "solve_s": time.perf_counter() - t0}


def solve_poisson(nx: int, ny: int, nz: int, b: np.ndarray, device, mixed: bool = False,
                  tol: float = 1e-8, capture: bool = True, omega: float = 1.0,
                  coarse_size: int = 64):
    """7-pt Poisson on nx x ny x nz: setup + PCG/V(1,1) solve; returns a dict of
    timings and the solution (the timed solve includes the residual checks)."""
    I, J, V = W.stencil(7, nx, ny, nz)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    dI, dJ, dV = t(I), t(J), t(V)
    db = t(b)
    torch.cuda.synchronize()
    amg = GeometricAMG(dI, dJ, dV, (nx, ny, nz), device, omega=omega, coarse_size=coarse_size,
                       spmv_policy=mixed_spmv_policy if mixed else fp64_policy,
                       spgemm_policy=mixed_spgemm_policy if mixed else fp64_policy)
    tc = time.perf_counter()
    if capture:
        amg.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, iters, hist = amg.pcg(db, tol=tol)
    torch.cuda.synchronize()
    solve_s = time.perf_counter() - t0
    return {"amg": amg, "x": x, "iters": iters, "hist": hist, "setup": amg.times,
            "capture_s": t0 - tc, "solve_s": solve_s}


def batch_poisson2d(sizes, seed: int, device):
    """Block-diagonal batch of 2D 5-pt Poisson problems (side n_i each, diag 4,
    off-diagonal -1) with right-hand sides b_i = A_i x*_i, x* ~ N(0,1) from one
    Park-Miller stream (seed).  Returns (I, J, V, b, x*, grids, segments) on the
    host; row blocks follow the order of `sizes`."""
    Is, Js, Vs, segs = [np.zeros(1, np.int64)], [], [], []
    nnz = row = 0
    for k, n in enumerate(sizes):
        I, J, V = W.stencil(7, n, n, 1)
        Is.append(I[1:].astype(np.int64) + nnz)
        Js.append(J.astype(np.int64) + row)
        Vs.append(V)
        segs.append(np.full(n * n, k, np.int64))
        nnz += len(J)
        row += n * n
    I = np.concatenate(Is).astype(np.int32)
    J = np.concatenate(Js).astype(np.int32)
    V = np.concatenate(Vs)
    xs = W.rand_normal(seed, row)
    import scipy.sparse as sp
    b = sp.csr_matrix((V, J, I), shape=(row, row)) @ xs
    return I, J, V, b, xs, [(n, n, 1) for n in sizes], np.concatenate(segs)


def solve_batch(sizes, seed, device, tol: float = 1e-8, capture: bool = True,
                coarse_size: int = 512, omega: float = 1.0):
    """All problems of a batch in one block-diagonal setup and PCG/V(1,1) solve."""
    I, J, V, b, xs, grids, seg = batch_poisson2d(sizes, seed, device)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    dI, dJ, dV, db, dseg = t(I), t(J), t(V), t(b), t(seg)
    amg = GeometricAMG(dI, dJ, dV, grids, device, omega=omega, coarse_size=coarse_size)
    tc = time.perf_counter()
    if capture:
        amg.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, iters, hist = amg.pcg(db, tol=tol, segments=dseg)
    torch.cuda.synchronize()
    return {"amg": amg, "x": x, "xs": xs, "iters": iters, "hist": hist, "setup": amg.times,
            "capture_s": t0 - tc, "solve_s": time.perf_counter() - t0, "segments": seg}


def batch_sizes(count: int = 48, seed: int = 4, lo: int = 32, hi: int = 512):
    """Sides n_i ~ U{lo..hi} from the Park-Miller stream (seed)."""
    u = W.pm_stream(seed, count).astype(np.float64) / 2147483647.0
    return [int(min(hi, lo + int(v * (hi - lo + 1)))) for v in u]


def solve_one_by_one(sizes, seed, device, tol: float = 1e-8, coarse_size: int = 512,
                     omega: float = 1.0):
    """The same problems solved one after the other (one hierarchy each)."""
    I, J, V, b, xs, grids, seg = batch_poisson2d(sizes, seed, device)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    # inputs resident on the device before the clock starts, as in solve_batch
    inputs = []
    row = 0
    for n in sizes:
        m = n * n
        Ii, Ji, Vi = W.stencil(7, n, n, 1)
        inputs.append((t(Ii), t(Ji), t(Vi), t(b[row:row + m]), n))
        row += m
    out = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for dI, dJ, dV, db, n in inputs:
        amg = GeometricAMG(dI, dJ, dV, (n, n, 1), device, omega=omega, coarse_size=coarse_size)
        amg.capture()
        x, iters, hist = amg.pcg(db, tol=tol)
        out.append((iters, hist[-1]))
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out
