"""Synthetic inputs for the configurations the path is measured on.

Vectorised numpy generators (independent of the CPU oracle, which checks them):

* HYPRE stencil matrices, rows diagonal-first then ascending columns, as
  GenerateLaplacian / 9pt / 27pt build them (src/parcsr_ls/par_laplace*.c);
  7pt couplings c = (cx, cy, cz) give diagonal sum(2c) over dimensions > 1
  (src/test/ij.c:9565-9580).
* HYPRE's Park-Miller stream (src/utilities/random.c:43-101) for right-hand
  sides: uniform 2U-1 (src/seq_mv/vector.c:395-418) and N(0,1) by Box-Muller
  on consecutive pairs.
"""
from __future__ import annotations

import numpy as np

PM_A, PM_M = 16807, 2147483647


def _offsets(kind: int):
    if kind == 7:
        return [(0, 0, -1), (0, -1, 0), (-1, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    if kind == 9:
        return [(i, j, 0) for j in (-1, 0, 1) for i in (-1, 0, 1) if i or j]
    if kind == 27:
        return [(i, j, l) for l in (-1, 0, 1) for j in (-1, 0, 1) for i in (-1, 0, 1)
                if i or j or l]
    raise ValueError(f"unknown stencil kind {kind}")


def stencil(kind: int, nx: int, ny: int, nz: int = 1, c=(1.0, 1.0, 1.0)):
    """CSR (I int32, J int32, A float64) of a HYPRE stencil matrix."""
    offs = _offsets(kind)
    n = nx * ny * nz
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    row = np.arange(n, dtype=np.int64)
    if kind == 7:
        diag = (2.0 * c[0] if nx > 1 else 0.0) + (2.0 * c[1] if ny > 1 else 0.0) + \
               (2.0 * c[2] if nz > 1 else 0.0)
        offv = [-(c[0] if dx else (c[1] if dy else c[2])) for dx, dy, dz in offs]
    else:
        diag = 8.0 if kind == 9 else 26.0
        offv = [-1.0] * len(offs)
    # per-row slot masks: slot 0 diagonal, then neighbours in ascending column order
    cols = [row]
    vals = [np.full(n, diag)]
    masks = [np.ones(n, dtype=bool)]
    for (dx, dy, dz), v in zip(offs, offv):
        xx, yy, zz = x + dx, y + dy, z + dz
        m = (xx >= 0) & (xx < nx) & (yy >= 0) & (yy < ny) & (zz >= 0) & (zz < nz)
        cols.append(row + dz * nx * ny + dy * nx + dx)
        vals.append(np.full(n, v))
        masks.append(m)
    C = np.stack(cols, axis=1)
    V = np.stack(vals, axis=1)
    Mk = np.stack(masks, axis=1)
    counts = Mk.sum(axis=1)
    I = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=I[1:])
    J = C[Mk].astype(np.int32)
    A = V[Mk].astype(np.float64)
    return I.astype(np.int32), J, A


def pm_stream(seed: int, n: int) -> np.ndarray:
    """n consecutive hypre_RandI values after hypre_SeedRand(seed), as int64.

    Computed by jumping: x_k = a^k * x_0 mod m, in blocks with exact int64
    arithmetic (a^k mod m < 2^31, so products fit in 62 bits)."""
    s = 1 if seed < 1 else (PM_M - 1 if seed >= PM_M else int(seed))
    out = np.empty(n, dtype=np.int64)
    blk = 4096
    # powers a^1..a^blk mod m
    pw = np.empty(blk, dtype=np.int64)
    p = 1
    for k in range(blk):
        p = (p * PM_A) % PM_M
        pw[k] = p
    a_blk = int(pw[-1])
    cur = s
    for start in range(0, n, blk):
        m = min(blk, n - start)
        out[start:start + m] = (pw[:m] * cur) % PM_M
        cur = (cur * a_blk) % PM_M
    return out


def rand_uniform(seed: int, n: int) -> np.ndarray:
    """hypre_SeqVectorSetRandomValues: 2*Rand - 1 with Rand = RandI/m."""
    return 2.0 * (pm_stream(seed, n).astype(np.float64) / float(PM_M)) - 1.0


def rand_normal(seed: int, n: int) -> np.ndarray:
    """N(0,1): Box-Muller on consecutive (u1, u2) pairs of the Park-Miller stream."""
    m = n + (n & 1)
    u = pm_stream(seed, m).astype(np.float64) / float(PM_M)
    u1, u2 = u[0::2], u[1::2]
    rad = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(m, dtype=np.float64)
    out[0::2] = rad * np.cos(2.0 * np.pi * u2)
    out[1::2] = rad * np.sin(2.0 * np.pi * u2)
    return out[:n]


# Configurations (sizes the path is measured on)
CONFIGS = {
    "c1_5pt_128": dict(kind=7, n=(128, 128, 1), c=(1.0, 1.0, 1.0), rhs=("uniform", 0)),
    "c2_7pt_128": dict(kind=7, n=(128, 128, 128), c=(1.0, 1.0, 1.0), rhs=("normal", 1)),
    "c3_7pt_aniso_160": dict(kind=7, n=(160, 160, 160), c=(1.0, 1.0, 0.01), rhs=("normal", 2)),
    "c4_27pt_64": dict(kind=27, n=(64, 64, 64), c=(1.0, 1.0, 1.0), rhs=("normal", 3)),
}


def config_matrix(name: str):
    cfg = CONFIGS[name]
    nx, ny, nz = cfg["n"]
    return stencil(cfg["kind"], nx, ny, nz, cfg["c"])


def _interp_1d(n: int):
    """Fine index -> list of (coarse index, weight) for coarsening by 2 (coarse
    points at even fine indices, linear interpolation, zero Dirichlet outside)."""
    nc = (n + 1) // 2
    f = np.arange(n)
    even = (f % 2) == 0
    c_lo = f // 2
    c_hi = c_lo + 1
    w_lo = np.where(even, 1.0, 0.5)
    w_hi = np.where(even, 0.0, 0.5)
    has_hi = (~even) & (c_hi < nc)
    return nc, c_lo, w_lo, c_hi, np.where(has_hi, w_hi, 0.0), has_hi


def trilinear_interpolation(nx: int, ny: int, nz: int = 1):
    """CSR (I, J, A) of the geometric prolongation from the grid of even-index
    points to the full nx x ny x nz grid (tensor product of 1D linear weights,
    up to 8 entries per row, ascending columns).  Returns (I, J, A, ncoarse)."""
    parts = []
    for n in (nx, ny, nz):
        nc, c_lo, w_lo, c_hi, w_hi, has_hi = _interp_1d(n)
        cols = np.stack([c_lo, c_hi], axis=1)
        wts = np.stack([w_lo, w_hi], axis=1)
        ok = np.stack([np.ones(n, bool), has_hi], axis=1)
        parts.append((nc, cols, wts, ok))
    (ncx, cx, wx, okx), (ncy, cy, wy, oky), (ncz, cz, wz, okz) = parts
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    cols, vals, mask = [], [], []
    for kz in range(2):
        for ky in range(2):
            for kx in range(2):
                m = okx[x, kx] & oky[y, ky] & okz[z, kz]
                c = (cz[z, kz] * ncy + cy[y, ky]) * ncx + cx[x, kx]
                v = wx[x, kx] * wy[y, ky] * wz[z, kz]
                cols.append(c)
                vals.append(v)
                mask.append(m & (v != 0.0))
    # slots in (kz, ky, kx) order already give ascending columns: the high
    # neighbour of each direction is the low one plus 1
    C = np.stack(cols, axis=1)
    V = np.stack(vals, axis=1)
    M = np.stack(mask, axis=1)
    counts = M.sum(axis=1)
    I = np.zeros(len(x) + 1, dtype=np.int64)
    np.cumsum(counts, out=I[1:])
    return I.astype(np.int32), C[M].astype(np.int32), V[M].astype(np.float64), ncx * ncy * ncz


def csr_transpose(nrows: int, ncols: int, I, J, A):
    """CSR of the transpose (rows of the result in ascending column order)."""
    nnz = len(J)
    rows = np.repeat(np.arange(nrows, dtype=np.int64), np.diff(I.astype(np.int64)))
    order = np.lexsort((rows, J))
    TJ = rows[order].astype(np.int32)
    TA = np.asarray(A)[order]
    counts = np.bincount(J, minlength=ncols)
    TI = np.zeros(ncols + 1, dtype=np.int64)
    np.cumsum(counts, out=TI[1:])
    assert TI[-1] == nnz
    return TI.astype(np.int32), TJ, TA


# ---- Gaussian-random-field diffusion (learned-policy problems) ----------------------------
# Park-Miller x_k = a^k x_0 mod m is one cycle and a seed is only a position on it, so
# seeds drawn from one stream would give shifted copies of one stream.  Problem i of a
# split starts at its own position a^p (p = base + i * stride) and uses at most `stride`
# draws: training problems at [0, 1500 * 2^18), held-out problems at 2^30 + j * 2^22.
GRF_SPLITS = {"train": (0, 1 << 18, 1500), "test": (1 << 30, 1 << 22, 50)}


def problem_seed(split: str, i: int, side: int) -> int:
    """Seed of problem i of a split ("train" or "test"); the side^2 problem
    must fit its stride (2 side^2 + 2 draws)."""
    base, stride, count = GRF_SPLITS[split]
    if not 0 <= i < count:
        raise ValueError(f"{split} problem index {i} outside [0, {count})")
    if 2 * side * side + 2 > stride:
        raise ValueError(f"a {side}^2 problem needs more than the {split} stride of draws")
    return pow(PM_A, base + i * stride, PM_M)


def grf(side: int, sigma: float, ell: float, noise: np.ndarray) -> np.ndarray:
    """Stationary Gaussian random field on the side x side cell centres of the
    unit square (periodic distance), covariance sigma^2 exp(-|r|^2 / (2 ell^2)),
    by spectral synthesis: g = ifft2(fft2(w) * sqrt(Lambda)) with w ~ N(0,1)
    white noise and Lambda = fft2(c) the eigenvalues of the circulant
    covariance (negative round-off eigenvalues clipped to 0), so that
    E[g_x g_(x+r)] = c(r) exactly."""
    d = np.minimum(np.arange(side), side - np.arange(side)) / side
    r2 = d[:, None] ** 2 + d[None, :] ** 2
    c = sigma * sigma * np.exp(-r2 / (2.0 * ell * ell))
    lam = np.maximum(np.fft.rfft2(c).real, 0.0)
    w = noise.reshape(side, side)
    return np.fft.irfft2(np.fft.rfft2(w) * np.sqrt(lam), s=(side, side))


def fv_diffusion(k: np.ndarray):
    """5-point finite-volume matrix of -div(k grad u) on the unit square with
    u = 0 on the boundary, one cell per entry of k (side x side, cell h = 1/side,
    row index i + side*j for cell (i, j), i along x).  Interior face
    transmissibility = harmonic mean 2 k1 k2 / (k1 + k2), boundary face 2 k
    (centre-to-wall distance h/2); the face length h and the 1/h of the gradient
    cancel.  Rows diagonal first, then ascending columns."""
    ny, nx = k.shape
    n = nx * ny
    idx = np.arange(n).reshape(ny, nx)
    tx = 2.0 * k[:, :-1] * k[:, 1:] / (k[:, :-1] + k[:, 1:])   # face (i, j)-(i+1, j)
    ty = 2.0 * k[:-1, :] * k[1:, :] / (k[:-1, :] + k[1:, :])   # face (i, j)-(i, j+1)
    diag = np.zeros((ny, nx))
    diag[:, :-1] += tx
    diag[:, 1:] += tx
    diag[:-1, :] += ty
    diag[1:, :] += ty
    diag[:, 0] += 2.0 * k[:, 0]
    diag[:, -1] += 2.0 * k[:, -1]
    diag[0, :] += 2.0 * k[0, :]
    diag[-1, :] += 2.0 * k[-1, :]
    # (row, col, value) of the off-diagonals, then one stable sort per row
    r = np.concatenate([idx[:, :-1].ravel(), idx[:, 1:].ravel(), idx[:-1, :].ravel(), idx[1:, :].ravel()])
    c = np.concatenate([idx[:, 1:].ravel(), idx[:, :-1].ravel(), idx[1:, :].ravel(), idx[:-1, :].ravel()])
    v = -np.concatenate([tx.ravel(), tx.ravel(), ty.ravel(), ty.ravel()])
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    counts = np.bincount(r, minlength=n) + 1
    I = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=I[1:])
    J = np.empty(I[-1], dtype=np.int32)
    A = np.empty(I[-1], dtype=np.float64)
    J[I[:-1]] = np.arange(n)
    A[I[:-1]] = diag.reshape(-1)
    rest = np.ones(I[-1], dtype=bool)
    rest[I[:-1]] = False
    J[rest] = c
    A[rest] = v
    return I.astype(np.int32), J, A


def grf_problem(side: int, seed: int):
    """One learned-policy problem: -div(k grad u) = f on side^2 cells with
    k = exp(g), g a GRF with sigma ~ U[0.5, 3] and correlation length
    ell ~ U[0.01, 0.2] of the unit domain, f ~ N(0,1) (cell integrals f h^2).
    From the problem's seed: the first two Park-Miller draws give sigma and ell,
    the stream continued gives the field's white noise and then f.  Returns
    (I, J, A, b, sigma, ell)."""
    u = pm_stream(seed, 2)
    sigma = 0.5 + 2.5 * float(u[0]) / PM_M
    ell = 0.01 + 0.19 * float(u[1]) / PM_M
    z = rand_normal(int(u[1]), 2 * side * side)
    g = grf(side, sigma, ell, z[:side * side])
    I, J, A = fv_diffusion(np.exp(g))
    b = z[side * side:] / float(side * side)
    return I, J, A, b, sigma, ell
