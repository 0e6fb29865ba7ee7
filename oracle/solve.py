"""BoomerAMG's one-process solve phase restated with scipy -- TEST
INFRASTRUCTURE ONLY (the checker of hierarchies against HYPRE's own outputs,
and bench.py's CPU baseline of setup + solve).

* l1_norms: the smoother's diagonal, hypre_ParCSRComputeL1Norms
  (parcsr_ls/ams.c:527-800) as par_amg_setup.c selects it: relax type 7 uses
  option 5 (the diagonal, 0 replaced by 1, par_amg_setup.c:3521-3536), relax
  type 18 option 1 (the l1 norm of the whole row, negated where a_ii < 0,
  par_amg_setup.c:3393-3406, ams.c:593-602, 786-799).
* VCycle: hypre_BoomerAMGCycle V(1,1) (par_cycle.c) with the Jacobi sweep of
  hypre_BoomerAMGRelax7Jacobi (par_relax.c:1178-1250): Vtemp = w f - w A u,
  u += Vtemp ./ l1 (the matvec skipped from a zero u), residual, restriction
  with P^T, coarse-grid correction u += P u_c, post-sweep; Gaussian
  elimination (relax type 9) on the coarsest level.
* pcg: hypre_PCGSolve (krylov/pcg.c:320-900) with two_norm = 1, rtol only:
  stop when <r,r> / <b,b> < tol^2; the final relative residual norm is
  sqrt(<r,r> / <b,b>).
"""
import numpy as np
import scipy.sparse as sp


def l1_norms(M, relax):
    d = M.diagonal()
    if relax == 7:
        out = d.copy()
        out[out == 0.0] = 1.0
        return out
    if relax == 18:
        out = np.asarray(abs(M).sum(axis=1)).ravel()
        out[d < 0] *= -1.0
        return out
    raise ValueError(f"relax type {relax}")


def level_mats(levels):
    """(A, P) scipy pairs from the oracle's amg_hierarchy dicts."""
    mats = []
    for L in levels:
        A = sp.csr_matrix((L["A"], L["J"], L["I"]), shape=(L["n"], L["n"]))
        P = None
        if "P" in L:
            PI, PJ, PA = L["P"]
            P = sp.csr_matrix((PA, PJ, PI), shape=(L["n"], L["nc"]))
        mats.append((A, P))
    return mats


class VCycle:
    """z = M^-1 r for one V(1,1) cycle from a zero initial guess."""

    def __init__(self, mats, relax=18, weight=1.0):
        self.mats = mats
        self.w = float(weight)
        self.l1 = [l1_norms(A, relax) for A, _ in mats]
        self.coarse = mats[-1][0].toarray()

    def _cycle(self, lvl, f):
        A, P = self.mats[lvl]
        if lvl == len(self.mats) - 1:
            return np.linalg.solve(self.coarse, f)
        w, l1 = self.w, self.l1[lvl]
        u = (w * f) / l1
        r = f - A @ u
        uc = self._cycle(lvl + 1, P.T @ r)
        u = u + P @ uc
        u = u + (w * f - w * (A @ u)) / l1
        return u

    def __call__(self, r):
        return self._cycle(0, r)


def pcg(A, b, precond, tol=1e-8, max_iter=1000):
    """Returns (x, iterations, final relative residual norm)."""
    x = np.zeros_like(b)
    r = b.copy()
    p = precond(r)
    gamma = r @ p
    bi = b @ b
    eps = tol * tol
    i = 0
    iprod = r @ r
    while i + 1 <= max_iter:
        i += 1
        s = A @ p
        alpha = gamma / (s @ p)
        x += alpha * p
        r -= alpha * s
        iprod = r @ r
        if iprod / bi < eps:
            break
        s = precond(r)
        gamma_old = gamma
        gamma = r @ s
        p = s + (gamma / gamma_old) * p
    return x, i, float(np.sqrt(iprod / bi))


def complexities(levels_n, levels_nnz):
    return sum(levels_n) / levels_n[0], sum(levels_nnz) / levels_nnz[0]


def setup_and_solve(n, I, J, A, b, relax=18, tol=1e-8, rap="kt"):
    """HYPRE's CPU BoomerAMG-PCG on one process as the oracle restates it (ij's
    defaults: HMIS, theta 0.25, max row sum 1, extended+i P max 4, coarse size
    9, the host RAP order), timed: (x, iterations, final relative residual,
    levels, setup seconds, solve seconds)."""
    import time

    from . import oracle as O
    t0 = time.perf_counter()
    levels = O.amg_hierarchy(n, I, J, A, theta=0.25, max_row_sum=1.0, max_elmts=4,
                             max_coarse_size=9, coarsen="hmis", rap=rap)
    mats = level_mats(levels)
    M = VCycle(mats, relax=relax)
    t1 = time.perf_counter()
    x, it, rel = pcg(mats[0][0], b, M, tol=tol)
    t2 = time.perf_counter()
    return x, it, rel, [L["n"] for L in levels], t1 - t0, t2 - t1
