"""Learned per-level policy for the algebraic hierarchy: a 12-64-64-4 network
(libamgt_mbsr.so's amgt_policy_* kernels) that reads each level's matrix
statistics during setup and picks the level's SpMV/smoother precision (FP64 /
FP32 / FP16) and its strength threshold.

It replaces the reference's fixed rule (get_calculationPrecison,
csr_spgemm_device.c:1847-1862: level 0 FP64, 1 FP32, deeper FP16) and
BoomerAMG's single strong_threshold.  The finest level always stays FP64, and
the Galerkin products are FP64 on every level, as in the sweeps that made the
labels (the FP32/FP16 products are also slower than the FP64 ones on this
part: A^2 of 7-pt 128^3 1.68 / 1.85 against 1.06 ms).

Workload: -div(k grad u) = f, 5-point finite volumes, k = exp(GRF)
(workloads.grf_problem); training problems at 256^2, held-out ones at 1024^2.

Labels come from the solver itself (`sweep_labels`):
* threshold: every problem is set up and solved (capture + PCG to 1e-8) with
  each threshold of THETAS on every level, twice; the threshold with the
  smallest measured setup + solve time is every level's target, except that
  the default 0.25 is kept when it is within LABEL_TIE (5%) of the fastest
  (the per-threshold times are stored with the labels);
* precision: on that hierarchy, each level l >= 1 that has a coarser level is
  switched alone to FP16, then FP32 (the others FP64); the label is the lowest
  precision whose V-cycle contraction (`cycle_contraction`: three cycles of
  the stationary iteration from one random error) stays within 10% of the
  all-FP64 one (FP64 otherwise).  The finest level is labelled FP64.
At solve time (`solve_learned`) the check runs inside the solve: PCG on the
predicted hierarchy stops early when its residual stalls (mean reduction per
iteration over the last STALL[0] iterations above STALL[1]) or when it has not
reached the tolerance after ITER_CAP iterations; the problem is then solved
again with the plain FP64 settings.  (A separate check before the solve,
three cycles of the predicted hierarchy against three of the same hierarchy
in FP64, cost 2.3 ms per problem at 1024^2 and sent problems whose PCG would
have converged to the FP64 re-solve.)
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time

import numpy as np
import torch

from . import _lib as L
from . import amg as G
from . import workloads as W
from .mbsr import _stream_ptr

NF, NOUT, NPARAMS = L.POLICY_FEATURES, L.POLICY_OUTPUTS, L.POLICY_PARAMS
PRECISIONS = (L.AMGT_DOUBLE, L.AMGT_FLOAT, L.AMGT_HALF)
THETAS = (0.1, 0.25, 0.4, 0.55, 0.7)
THETA_RANGE = (THETAS[0], THETAS[-1])   # predictions are clipped to the swept range
FACTOR_SLACK = 1.10          # allowed convergence-factor ratio to FP64 (precision labels)
DEFAULT_THETA = 0.25         # BoomerAMG's strong threshold (ij's default)
LABEL_TIE = 0.05             # threshold labels: relative time difference taken as a tie
STALL = (3, 0.7)             # in-solve check: window, largest mean reduction per iteration
ITER_CAP = 40                # in-solve check: PCG iterations before the FP64 fallback
FEATURE_NAMES = ("log10_rows_rel", "log10_nnz_rel", "nnz_per_row", "block_fill", "block_row_std",
                 "level", "coarsening_ratio", "mean_diag_dominance", "frac_rows_not_dominant",
                 "frac_positive_offdiag", "log10_value_range", "std_log10_diag")
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "policy_weights.json")
AMG_ARGS = dict(max_row_sum=0.9, cf_init=0, max_elmts=4, omega=6.0 / 7.0, coarse_size=64,
                coarsen="pmis", relax_type=7)


# ---- features ------------------------------------------------------------------------------
_STATS_WORK: dict = {}


def level_stats(n: int, CI: torch.Tensor, CJ: torch.Tensor, CV: torch.Tensor) -> np.ndarray:
    """amgt_level_stats: the eight row statistics of a device CSR (one pass,
    one transfer)."""
    lib = L.load()
    need = int(lib.amgt_level_stats_work_bytes(n))
    w = _STATS_WORK.get(CV.device)
    if w is None or w.numel() < need:
        w = torch.empty(max(need, 1 << 16), dtype=torch.uint8, device=CV.device)
        _STATS_WORK[CV.device] = w
    out = (ctypes.c_double * 8)()
    L.check(lib.amgt_level_stats(n, ctypes.c_void_p(CI.data_ptr()), ctypes.c_void_p(CJ.data_ptr()),
                                 ctypes.c_void_p(CV.data_ptr()), out, ctypes.c_void_p(w.data_ptr()),
                                 w.numel(), _stream_ptr(None)), "amgt_level_stats")
    return np.array(out[:], dtype=np.float64)


def level_features(lvl: int, n_prev: int, A, CI: torch.Tensor, CJ: torch.Tensor,
                   CV: torch.Tensor, finest=None) -> np.ndarray:
    """The 12 statistics of one level's operator (FEATURE_NAMES): rows and nnz
    relative to the finest level (log10; `finest` = (rows, nnz) of level 0, so
    that the features do not depend on the problem size the network was
    trained at), row length, mBSR block fill and block-row spread, level index,
    coarsening ratio n_l / n_(l-1), diagonal dominance |a_ii| / sum_(j != i)
    |a_ij| (mean, and the fraction of rows below 1), the fraction of positive
    off-diagonals, the value range max |a| / min |a_ii| and the standard
    deviation of log10 |a_ii|."""
    info = A.info()
    n = int(CI.numel() - 1)
    nnz = int(CV.numel())
    n0, nnz0 = finest if finest is not None else (n, nnz)
    st = level_stats(n, CI, CJ, CV)
    cnt = st[7]                      # rows with a_ii != 0
    vr = st[3] / st[4] if cnt > 0 else 1.0
    mean_l = st[5] / cnt if cnt > 0 else 0.0
    var_l = max(st[6] - cnt * mean_l * mean_l, 0.0) / max(cnt - 1.0, 1.0) if cnt > 0 else 0.0
    return np.array([
        math.log10(max(n, 1) / max(n0, 1)), math.log10(max(nnz, 1) / max(nnz0, 1)),
        nnz / max(n, 1), info["fill"], info["stand"], float(lvl), n / max(n_prev, 1),
        st[0] / n, st[1] / n, st[2] / max(nnz - cnt, 1.0), math.log10(max(vr, 1.0)),
        math.sqrt(var_l)], dtype=np.float64)


class _FeatureHook:
    """level_hook of BoomerAMG that records every level's features (relative to
    the finest level) and asks `decide(lvl, features)` for (theta, precision)."""

    def __init__(self, decide):
        self.decide = decide
        self.features = []
        self.finest = None

    def __call__(self, lvl, n_prev, A, CI, CJ, CV):
        if lvl == 0:
            self.finest = (int(CI.numel() - 1), int(CV.numel()))
        f = level_features(lvl, n_prev, A, CI, CJ, CV, self.finest)
        self.features.append(f)
        return self.decide(lvl, f)


# ---- network -------------------------------------------------------------------------------
class PolicyNet:
    """Parameters (FP32, device) plus the feature standardisation and the
    threshold scaling of the targets; forward/gradient/Adam on the device."""

    def __init__(self, device, seed: int = 0):
        self.device = device
        rng = np.random.default_rng(seed)
        p = np.zeros(NPARAMS, dtype=np.float32)
        o = 0
        for fan_in, fan_out in ((NF, 64), (64, 64), (64, NOUT)):
            lim = math.sqrt(6.0 / (fan_in + fan_out))
            p[o:o + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
            o += fan_in * fan_out + fan_out
        self.params = torch.from_numpy(p).to(device)
        self.mean = np.zeros(NF)
        self.std = np.ones(NF)
        self.theta_mid, self.theta_half = 0.4, 0.3
        self.meta: dict = {}
        self._work = None

    # standardisation -------------------------------------------------------------------
    def fit_scaling(self, X: np.ndarray):
        self.mean = X.mean(axis=0)
        self.std = np.where(X.std(axis=0) > 1e-12, X.std(axis=0), 1.0)

    def _x(self, X: np.ndarray) -> torch.Tensor:
        Z = ((np.atleast_2d(X) - self.mean) / self.std).astype(np.float32)
        return torch.from_numpy(np.ascontiguousarray(Z)).to(self.device)

    def _workspace(self, n: int) -> torch.Tensor:
        need = int(L.load().amgt_policy_work_bytes(n))
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.uint8, device=self.device)
        return self._work

    # kernels ---------------------------------------------------------------------------
    def _check(self, t: torch.Tensor, shape: tuple, dtype, name: str):
        """The kernels index these arrays by the sizes given: anything else is
        rejected before a launch."""
        if tuple(t.shape) != shape or t.dtype != dtype or not t.is_contiguous() or \
                t.device != self.params.device:
            raise ValueError(f"{name}: expected a contiguous {dtype} tensor of shape {shape} on "
                             f"{self.params.device}, got {tuple(t.shape)} {t.dtype} on {t.device}")

    def forward_raw(self, Xd: torch.Tensor) -> torch.Tensor:
        n = Xd.shape[0]
        self._check(Xd, (n, NF), torch.float32, "features")
        self._check(self.params, (NPARAMS,), torch.float32, "params")
        out = torch.empty((n, NOUT), dtype=torch.float32, device=self.device)
        w = self._workspace(n)
        L.check(L.load().amgt_policy_forward(
            ctypes.c_void_p(self.params.data_ptr()), ctypes.c_void_p(Xd.data_ptr()), n,
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(w.data_ptr()), w.numel(),
            _stream_ptr(None)), "amgt_policy_forward")
        return out

    def loss_grad_raw(self, Xd, cls_d, tgt_d, lam: float):
        n = Xd.shape[0]
        self._check(Xd, (n, NF), torch.float32, "features")
        self._check(cls_d, (n,), torch.int32, "labels")
        self._check(tgt_d, (n,), torch.float32, "targets")
        self._check(self.params, (NPARAMS,), torch.float32, "params")
        grad = torch.empty(NPARAMS, dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        w = self._workspace(n)
        L.check(L.load().amgt_policy_loss_grad(
            ctypes.c_void_p(self.params.data_ptr()), ctypes.c_void_p(Xd.data_ptr()),
            ctypes.c_void_p(cls_d.data_ptr()), ctypes.c_void_p(tgt_d.data_ptr()), n,
            float(lam), ctypes.c_void_p(grad.data_ptr()), ctypes.c_void_p(loss.data_ptr()),
            ctypes.c_void_p(w.data_ptr()), w.numel(), _stream_ptr(None)), "amgt_policy_loss_grad")
        return loss, grad

    def adam_raw(self, grad, m, v, step: int, lr: float, b1=0.9, b2=0.999, eps=1e-8):
        for t, name in ((self.params, "params"), (grad, "grad"), (m, "m"), (v, "v")):
            self._check(t, (NPARAMS,), torch.float32, name)
        L.check(L.load().amgt_policy_adam(
            ctypes.c_void_p(self.params.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
            ctypes.c_void_p(m.data_ptr()), ctypes.c_void_p(v.data_ptr()), NPARAMS, int(step),
            float(lr), float(b1), float(b2), float(eps), _stream_ptr(None)), "amgt_policy_adam")

    # training / prediction -------------------------------------------------------------
    def fit(self, X: np.ndarray, cls: np.ndarray, theta: np.ndarray, epochs: int = 1500,
            lr: float = 3e-3, lam: float = 1.0) -> list:
        """Full-batch Adam on the device; returns the loss every 100 epochs."""
        if np.any((cls < 0) | (cls > 2)):
            raise ValueError("precision labels must be 0, 1 or 2")
        self.fit_scaling(X)
        Xd = self._x(X)
        cd = torch.from_numpy(cls.astype(np.int32)).to(self.device)
        td = torch.from_numpy(((theta - self.theta_mid) / self.theta_half).astype(np.float32)).to(
            self.device)
        m = torch.zeros(NPARAMS, dtype=torch.float32, device=self.device)
        v = torch.zeros_like(m)
        hist = []
        for ep in range(1, epochs + 1):
            loss, grad = self.loss_grad_raw(Xd, cd, td, lam)
            self.adam_raw(grad, m, v, ep, lr)
            if ep % 100 == 0 or ep == 1:
                hist.append(float(loss.item()))
        return hist

    def predict(self, X: np.ndarray):
        """(precision class per row, threshold per row)."""
        out = self.forward_raw(self._x(X)).cpu().numpy().astype(np.float64)
        cls = out[:, :3].argmax(axis=1)
        theta = np.clip(self.theta_mid + self.theta_half * out[:, 3], *THETA_RANGE)
        return cls, theta

    def save(self, path: str = WEIGHTS_PATH, meta: dict | None = None):
        d = {"params": self.params.cpu().numpy().astype(float).tolist(),
             "mean": self.mean.tolist(), "std": self.std.tolist(),
             "theta_mid": self.theta_mid, "theta_half": self.theta_half,
             "features": list(FEATURE_NAMES), "meta": meta or {}}
        with open(path, "w") as f:
            json.dump(d, f)

    @classmethod
    def load(cls, device, path: str = WEIGHTS_PATH) -> "PolicyNet":
        with open(path) as f:
            d = json.load(f)
        net = cls(device)
        net.params = torch.tensor(d["params"], dtype=torch.float32, device=device)
        net.mean, net.std = np.array(d["mean"]), np.array(d["std"])
        net.theta_mid, net.theta_half = d["theta_mid"], d["theta_half"]
        net.meta = d.get("meta", {})
        return net


# ---- solver runs ---------------------------------------------------------------------------
def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def set_level_precision(amg, lvl: int, prec: int):
    """SpMV/smoother precision of one level (its A, P and R)."""
    L_ = amg.levels[lvl]
    for M in (L_.A, L_.P, L_.R):
        if M is not None:
            M.set_precision(prec)


def cycle_contraction(amg, cycles: int = 3, seed: int = 7) -> float:
    """(||e_k|| / ||e_0||)^(1/k) of the stationary iteration e <- e - M^-1 A e
    from a N(0,1) error drawn on the device (torch generator, fixed seed): one
    application of the V-cycle (the captured graph when there is one) each."""
    A0 = amg.levels[0].A
    n = amg.levels[0].n
    g = torch.Generator(device=amg.device).manual_seed(seed)
    e = torch.randn(n, dtype=torch.float64, device=amg.device, generator=g)
    r = torch.empty_like(e)
    z = torch.empty_like(e)
    n0 = torch.linalg.vector_norm(e)
    for _ in range(cycles):
        A0.spmv(e, r)
        amg.precondition(r, z)
        e.sub_(z)
    return (torch.linalg.vector_norm(e) / n0).item() ** (1.0 / cycles)


def _timed_solve(amg, db, tol, t_setup0, capture=True, max_iter=200):
    tc = time.perf_counter()
    if capture:
        amg.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, iters, hist = amg.pcg(db, tol=tol, max_iter=max_iter)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return x, iters, hist, {"setup_s": tc - t_setup0, "capture_s": t0 - tc, "solve_s": t1 - t0,
                            "total_s": t1 - t_setup0}


def solve_fixed(dI, dJ, dV, n, db, device, theta=0.25, mixed=False, tol=1e-8, capture=True):
    """Plain settings: one threshold, FP64 or the fixed mixed schedule (FP64
    finest, FP32 levels 1-2, FP16 deeper)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    amg = G.BoomerAMG(dI, dJ, dV, n, device, theta=theta,
                      spmv_policy=G.mixed_spmv_policy if mixed else G.fp64_policy,
                      spgemm_policy=G.mixed_spgemm_policy if mixed else G.fp64_policy, **AMG_ARGS)
    x, iters, hist, t = _timed_solve(amg, db, tol, t0, capture)
    return amg, x, iters, hist, t


def solve_learned(net: PolicyNet, dI, dJ, dV, n, db, device, tol=1e-8, capture=True,
                  stall=STALL, iter_cap: int = ITER_CAP):
    """Setup with the network's per-level threshold and precision, then PCG
    with the in-solve check (module docstring): a stalled solve, or one that
    misses the tolerance within `iter_cap` iterations, falls back to the plain
    FP64 settings.  t["total_s"] includes a failed attempt."""
    def decide(lvl, f):
        cls, theta = net.predict(f[None, :])
        return float(theta[0]), (L.AMGT_DOUBLE if lvl == 0 else PRECISIONS[int(cls[0])])

    hook = _FeatureHook(decide)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    amg = G.BoomerAMG(dI, dJ, dV, n, device, level_hook=hook, spgemm_policy=G.fp64_policy,
                      **AMG_ARGS)
    torch.cuda.synchronize()
    tc = time.perf_counter()
    if capture:
        amg.capture()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    x, iters, hist = amg.pcg(db, tol=tol, max_iter=iter_cap, stall=stall)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    info = {"thetas": list(amg.level_theta), "precisions": list(amg.level_prec),
            "fallback": None, "features": [f.tolist() for f in hook.features],
            "learned_iters": iters}
    if hist[-1] < tol:
        t = {"setup_s": tc - t0, "capture_s": t1 - tc, "solve_s": t2 - t1, "total_s": t2 - t0}
        return amg, x, iters, hist, t, info
    info["fallback"] = (f"iteration cap {iter_cap}" if iters >= iter_cap else
                        f"residual stalled after {iters} iterations")
    del amg
    amg, x, iters, hist, t2 = solve_fixed(dI, dJ, dV, n, db, device, capture=capture, tol=tol)
    t = {k: v for k, v in t2.items()}
    t["total_s"] = time.perf_counter() - t0  # includes the failed attempt
    return amg, x, iters, hist, t, info


# ---- labels --------------------------------------------------------------------------------
def sweep_labels(I, J, A, b, n, device, tol=1e-8):
    """Per-level (features, precision class, threshold) of one problem from
    the solver's own sweeps (module docstring)."""
    dI, dJ, dV, db = _t(I, device), _t(J, device), _t(A, device), _t(b, device)
    best, best_t = None, float("inf")
    times = {}
    for theta in THETAS:
        tt = []
        for _ in range(2):
            amg, x, iters, hist, t = solve_fixed(dI, dJ, dV, n, db, device, theta=theta, tol=tol)
            ok = hist[-1] < tol
            tt.append(t["total_s"] if ok else float("inf"))
            del amg, x
        times[theta] = min(tt)
        if times[theta] < best_t:
            best, best_t = theta, times[theta]
    if best is None:
        return None
    # within LABEL_TIE of the fastest, BoomerAMG's default threshold wins: at
    # the training size the differences are within the timing noise, while a
    # lower threshold's larger strength graph costs setup time on larger grids
    if times.get(DEFAULT_THETA, float("inf")) <= best_t * (1.0 + LABEL_TIE):
        best = DEFAULT_THETA
    hook = _FeatureHook(lambda lvl, f: (best, L.AMGT_DOUBLE))
    amg = G.BoomerAMG(dI, dJ, dV, n, device, level_hook=hook, spgemm_policy=G.fp64_policy,
                      **AMG_ARGS)
    feats = hook.features
    rho64 = cycle_contraction(amg)
    cls = [0] * len(amg.levels)
    for lvl in range(1, len(amg.levels) - 1):
        for c in (2, 1):
            set_level_precision(amg, lvl, PRECISIONS[c])
            rho = cycle_contraction(amg)
            set_level_precision(amg, lvl, L.AMGT_DOUBLE)
            if rho <= FACTOR_SLACK * rho64:
                cls[lvl] = c
                break
    nlev = len(amg.levels)
    del amg
    return {"features": np.array(feats), "cls": np.array(cls[:nlev - 1] + [0]),
            "theta": best, "times": times, "rho_fp64": rho64, "levels": nlev}


def build_training_set(count: int, side: int, device, log=print):
    """Labels of the first `count` training problems (side^2 cells)."""
    X, C, T, S = [], [], [], []
    for i in range(count):
        I, J, A, b, sigma, ell = W.grf_problem(side, W.problem_seed("train", i, side))
        r = sweep_labels(I, J, A, b, side * side, device)
        if r is None:
            continue
        S.append([r["times"][th] for th in THETAS])
        # the coarsest level is solved directly: no precision or threshold to learn
        nl = r["levels"] - 1
        X.append(r["features"][:nl])
        C.append(r["cls"][:nl])
        T.append(np.full(nl, r["theta"]))
        if log and (i % 25 == 0 or i == count - 1):
            log(f"[labels] {i + 1}/{count} sigma {sigma:.2f} ell {ell:.3f} theta* {r['theta']} "
                f"cls {r['cls'].tolist()}")
    return np.concatenate(X), np.concatenate(C), np.concatenate(T), np.array(S)


# ---- evaluation ----------------------------------------------------------------------------
def evaluate(net: PolicyNet, count: int, side: int, device, tol=1e-8, log=print, split="test"):
    """Held-out problems solved three ways (plain FP64 at theta 0.25, the fixed
    mixed schedule, the learned policy with its in-solve check and fallback), each set up and
    solved twice with the second run timed.  Before the first problem every
    way runs once untimed, so that first-use costs (code-object loads,
    allocator growth, first graph instantiations) do not fall on the way that
    comes first.  Per problem: times, iterations, the true FP64 relative
    residual and the max-norm difference of each solution from the FP64 one."""
    import scipy.sparse as sp

    def run(name, dI, dJ, dV, n, db):
        if name == "learned":
            return solve_learned(net, dI, dJ, dV, n, db, device, tol=tol)
        return solve_fixed(dI, dJ, dV, n, db, device, mixed=name == "mixed", tol=tol) + (None,)

    rows = []
    for j in range(count):
        I, J, A, b, sigma, ell = W.grf_problem(side, W.problem_seed(split, j, side))
        n = side * side
        M = sp.csr_matrix((A, J, I), shape=(n, n))
        bn = np.linalg.norm(b)
        dI, dJ, dV, db = (_t(a, device) for a in (I, J, A, b))
        res = {"problem": j, "sigma": sigma, "ell": ell}
        sols = {}
        if j == 0:
            for name in ("fp64", "mixed", "learned"):
                run(name, dI, dJ, dV, n, db)
        for name in ("fp64", "mixed", "learned"):
            for rep in range(2):
                amg, x, iters, hist, t, info = run(name, dI, dJ, dV, n, db)
                if rep == 0:
                    del amg, x
            xh = x.cpu().numpy()
            sols[name] = xh
            r = {"iters": iters, "total_s": t["total_s"], "setup_s": t["setup_s"],
                 "solve_s": t["solve_s"], "true_rel_res": float(np.linalg.norm(b - M @ xh) / bn),
                 "levels": len(amg.levels), "op_complexity": amg.complexities()[1]}
            if info is not None:
                r.update({"fallback": info["fallback"], "thetas": info["thetas"],
                          "precisions": info["precisions"],
                          "learned_iters": info["learned_iters"]})
            res[name] = r
            del amg, x
        scale = np.abs(sols["fp64"]).max()
        for name in ("mixed", "learned"):
            res[name]["diff_vs_fp64"] = float(np.abs(sols[name] - sols["fp64"]).max() / scale)
        rows.append(res)
        if log:
            log(f"[eval] {j + 1}/{count} sigma {sigma:.2f} ell {ell:.3f} total ms fp64 "
                f"{1e3 * res['fp64']['total_s']:.1f} mixed {1e3 * res['mixed']['total_s']:.1f} "
                f"learned {1e3 * res['learned']['total_s']:.1f} its {res['fp64']['iters']}/"
                f"{res['mixed']['iters']}/{res['learned']['iters']} fallback "
                f"{res['learned']['fallback']} thetas {np.round(res['learned']['thetas'], 2).tolist()}")
    summary = {}
    for name in ("fp64", "mixed", "learned"):
        summary[name] = {k: float(np.mean([r[name][k] for r in rows]))
                         for k in ("total_s", "setup_s", "solve_s", "iters", "op_complexity")}
        summary[name]["max_true_rel_res"] = float(max(r[name]["true_rel_res"] for r in rows))
    for name in ("mixed", "learned"):
        summary[name]["max_diff_vs_fp64"] = float(max(r[name]["diff_vs_fp64"] for r in rows))
        summary[name]["speedup_vs_fp64"] = summary["fp64"]["total_s"] / summary[name]["total_s"]
    summary["learned"]["fallbacks"] = int(sum(r["learned"]["fallback"] is not None for r in rows))
    return rows, summary


def _main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="labels, training and evaluation of the policy")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("labels")
    a.add_argument("--count", type=int, default=1500)
    a.add_argument("--side", type=int, default=256)
    a.add_argument("--out", required=True)
    a = sub.add_parser("train")
    a.add_argument("--labels", required=True)
    a.add_argument("--epochs", type=int, default=3000)
    a.add_argument("--out", default=WEIGHTS_PATH)
    a = sub.add_parser("evaluate")
    a.add_argument("--count", type=int, default=50)
    a.add_argument("--side", type=int, default=1024)
    a.add_argument("--weights", default=WEIGHTS_PATH)
    a.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    if args.cmd == "labels":
        t0 = time.perf_counter()
        X, C, T, S = build_training_set(args.count, args.side, dev)
        np.savez_compressed(args.out, features=X.astype(np.float32), cls=C.astype(np.int8),
                            theta_times=S.astype(np.float32),
                            theta=T.astype(np.float32), side=args.side, count=args.count)
        print(f"[labels] {X.shape[0]} levels from {args.count} problems in "
              f"{time.perf_counter() - t0:.1f} s; class counts {np.bincount(C, minlength=3).tolist()}"
              f"; theta* counts {dict(zip(*np.unique(T, return_counts=True)))}", flush=True)
    elif args.cmd == "train":
        d = np.load(args.labels)
        X, C, T = d["features"].astype(np.float64), d["cls"].astype(np.int32), d["theta"].astype(np.float64)
        net = PolicyNet(dev, seed=0)
        t0 = time.perf_counter()
        hist = net.fit(X, C, T, epochs=args.epochs)
        torch.cuda.synchronize()
        cls, th = net.predict(X)
        meta = {"samples": int(X.shape[0]), "epochs": args.epochs, "loss": hist,
                "train_accuracy": float((cls == C).mean()),
                "train_theta_mae": float(np.abs(th - T).mean()),
                "train_s": time.perf_counter() - t0, "labels": os.path.basename(args.labels)}
        net.save(args.out, meta)
        print(json.dumps(meta), flush=True)
    else:
        net = PolicyNet.load(dev, args.weights)
        rows, summary = evaluate(net, args.count, args.side, dev)
        with open(args.out, "w") as f:
            json.dump({"summary": summary, "problems": rows}, f)
        print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    _main()
