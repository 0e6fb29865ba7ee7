"""GPU tests of the per-level policy network (amgt_policy_* through the C-ABI)
and of the learned setup on GRF diffusion problems.

* forward, loss and gradient against the oracle's double-precision
  restatement (itself pinned by finite differences and closed forms): FP32
  kernels within 2e-5 relative (outputs, loss) and 1e-4 of the largest
  gradient entry, for batches of 1, 300 and 1300 samples (several gradient
  chunks);
* a label outside 0..2 gives a NaN loss;
* Adam steps against the oracle's, and full-batch training drives the loss
  of a separable toy set down;
* the learned setup on 256^2 GRF problems: the hook's thresholds and
  precisions are the network's, the finest level is FP64, the solve reaches
  1e-8 (true FP64 residual) with or without the fallback, and a failed
  convergence check falls back to the plain FP64 settings.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from amgt_amd import _lib as L
from amgt_amd import policy as P
from amgt_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 12)).astype(np.float32)
    cls = rng.integers(0, 3, n).astype(np.int32)
    tgt = rng.normal(size=n).astype(np.float32)
    return X, cls, tgt


@pytest.mark.parametrize("n", [1, 300, 1300])
def test_forward_and_grad_match_oracle(dev, n):
    print(f"[policy] forward/grad n={n}", flush=True)
    net = P.PolicyNet(dev, seed=n)
    p32 = net.params.cpu().numpy()
    X, cls, tgt = _data(n, n + 1)
    Xd = torch.from_numpy(X).to(dev)
    out = net.forward_raw(Xd).cpu().numpy()
    ref = O.policy_forward(p32.astype(np.float64), X.astype(np.float64))
    assert np.allclose(out, ref, rtol=2e-5, atol=2e-5 * np.abs(ref).max())
    loss, grad = net.loss_grad_raw(Xd, torch.from_numpy(cls).to(dev), torch.from_numpy(tgt).to(dev),
                                   0.5)
    rl, rg = O.policy_loss_grad(p32.astype(np.float64), X.astype(np.float64), cls,
                                tgt.astype(np.float64), 0.5)
    assert abs(loss.item() - rl) <= 2e-5 * abs(rl)
    g = grad.cpu().numpy()
    assert np.all(np.abs(g - rg) <= 1e-4 * np.abs(rg).max()), np.abs(g - rg).max()


def test_bad_label_gives_nan_loss(dev):
    net = P.PolicyNet(dev)
    X, cls, tgt = _data(8, 9)
    cls[3] = 5
    loss, grad = net.loss_grad_raw(torch.from_numpy(X).to(dev), torch.from_numpy(cls).to(dev),
                                   torch.from_numpy(tgt).to(dev), 1.0)
    assert np.isnan(loss.item())
    with pytest.raises(ValueError):
        net.fit(X, cls, tgt, epochs=1)


def test_adam_matches_oracle_and_training_converges(dev):
    net = P.PolicyNet(dev, seed=3)
    X, _, _ = _data(600, 4)
    # separable toy labels: class by the sign pattern of two features, target linear
    cls = (X[:, 0] > 0).astype(np.int32) + (X[:, 1] > 0.5).astype(np.int32)
    tgt = (0.3 * X[:, 2] - 0.2 * X[:, 3]).astype(np.float32)
    Xd, cd, td = (torch.from_numpy(a).to(dev) for a in (X, cls, tgt))
    p = net.params.cpu().numpy().astype(np.float64)
    m = torch.zeros(P.NPARAMS, device=dev)
    v = torch.zeros_like(m)
    mo, vo = np.zeros(P.NPARAMS), np.zeros(P.NPARAMS)
    for step in range(1, 4):
        _, grad = net.loss_grad_raw(Xd, cd, td, 1.0)
        g = grad.cpu().numpy().astype(np.float64)
        net.adam_raw(grad, m, v, step, 1e-2)
        O.policy_adam(p, g, mo, vo, step, 1e-2)
        assert np.allclose(net.params.cpu().numpy(), p, rtol=1e-5, atol=1e-6), step
    net2 = P.PolicyNet(dev, seed=5)
    hist = net2.fit(X.astype(np.float64), cls, (0.4 + 0.3 * tgt).astype(np.float64), epochs=400,
                    lr=1e-2)
    assert hist[-1] < 0.25 * hist[0], hist
    pc, pt = net2.predict(X.astype(np.float64))
    assert (pc == cls).mean() > 0.9


def _grf(side, i, dev):
    I, J, A, b, _, _ = W.grf_problem(side, W.problem_seed("train", i, side))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    return (I, J, A, b), (t(I), t(J), t(A), t(b))


def _true_res(h, x):
    I, J, A, b = h
    n = b.size
    M = sp.csr_matrix((A, J, I), shape=(n, n))
    return np.linalg.norm(b - M @ x) / np.linalg.norm(b)


def test_learned_setup_uses_network(dev):
    print("[policy] learned setup 256^2", flush=True)
    net = P.PolicyNet(dev, seed=11)
    h, (dI, dJ, dV, db) = _grf(256, 3, dev)
    amg, x, iters, hist, t, info = P.solve_learned(net, dI, dJ, dV, 256 * 256, db, dev)
    assert _true_res(h, x.cpu().numpy()) < 1e-8
    if info["fallback"] is None:
        feats = np.array(info["features"])
        cls, theta = net.predict(feats)
        assert np.allclose(info["thetas"], theta)
        assert info["precisions"][0] == L.AMGT_DOUBLE
        assert info["precisions"][1:] == [P.PRECISIONS[c] for c in cls[1:]]
        # an FP16 request runs in FP32 where the level's values exceed FP16's range
        for want, lv in zip(info["precisions"], amg.levels):
            got = lv.A.info()["precision"]
            assert got == want or (want == L.AMGT_HALF and got == L.AMGT_FLOAT)


def test_fallback_to_fp64(dev):
    """With no iterations allowed every learned solve fails the in-solve
    check: the problem is solved again with the plain FP64 settings
    (threshold 0.25, FP64 on every level) and meets the tolerance.  Without
    the stall rule and with a large cap the network's hierarchy is kept (here
    threshold 0.9 and FP16 on every level but the finest)."""
    net = P.PolicyNet(dev, seed=12)
    p = net.params.cpu().numpy()
    p[4992:5248] = 0.0
    p[5248:5252] = [-10.0, -10.0, 10.0, 10.0]
    net.params = torch.from_numpy(p).to(dev)
    h, (dI, dJ, dV, db) = _grf(256, 4, dev)
    amg, x, iters, hist, t, info = P.solve_learned(net, dI, dJ, dV, 256 * 256, db, dev,
                                                   iter_cap=0)
    assert info["fallback"] is not None and info["learned_iters"] == 0
    assert all(th == P.THETA_RANGE[1] for th in info["thetas"])
    assert _true_res(h, x.cpu().numpy()) < 1e-8
    assert amg.level_theta == [0.25] * len(amg.levels)
    assert all(lv.A.info()["precision"] == L.AMGT_DOUBLE for lv in amg.levels)
    amg, x, iters, hist, t, info = P.solve_learned(net, dI, dJ, dV, 256 * 256, db, dev,
                                                   stall=None, iter_cap=200)
    assert info["fallback"] is None
    assert amg.level_theta == [P.THETA_RANGE[1]] * len(amg.levels)
    assert info["precisions"] == [L.AMGT_DOUBLE] + [L.AMGT_HALF] * (len(amg.levels) - 1)
    assert _true_res(h, x.cpu().numpy()) < 1e-8


@pytest.mark.parametrize("case", ["grf", "rand"])
def test_level_stats_match_numpy(dev, case):
    """amgt_level_stats against the same row statistics computed with numpy
    (a zero diagonal row and a row without off-diagonals included)."""
    if case == "grf":
        I, J, A, _, _, _ = W.grf_problem(64, W.problem_seed("train", 7, 64))
    else:
        rng = np.random.default_rng(5)
        M = sp.random(300, 300, density=0.03, random_state=6, format="lil")
        M.setdiag(rng.normal(size=300) * 5)
        M[10, 10] = 0.0
        M[20, :] = 0.0
        M[20, 20] = 3.0
        M = M.tocsr()
        I, J, A = M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.copy()
    n = I.size - 1
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    st = P.level_stats(n, t(I), t(J), t(A))
    rows = np.repeat(np.arange(n), np.diff(I))
    dmask = J == rows
    d = np.zeros(n)
    np.add.at(d, rows[dmask], np.abs(A[dmask]))
    off = np.zeros(n)
    np.add.at(off, rows[~dmask], np.abs(A[~dmask]))
    dom = np.where(off > 0, np.minimum(d / np.where(off > 0, off, 1), 10.0), 10.0)
    pos = d > 0
    ref = [dom.sum(), (dom < 1 - 1e-12).sum(), ((A > 0) & ~dmask).sum(), np.abs(A).max(),
           d[pos].min(), np.log10(d[pos]).sum(), (np.log10(d[pos]) ** 2).sum(), pos.sum()]
    assert np.allclose(st, ref, rtol=1e-12, atol=1e-12), (st, ref)


def test_trained_policy_on_heldout(dev):
    """The committed network on three held-out 1024^2 problems against plain
    FP64 and the fixed mixed schedule: every result reaches a true FP64
    relative residual below 1e-8, the learned solution is within 1e-6 of the
    FP64 one (max norm, relative), it needs at most one PCG iteration more per
    problem on average, and the times are printed (the network's gain or
    loss: on these problems it trades the 2-4 ms of its convergence check for
    about 1 ms of FP16 coarse levels)."""
    import json
    net = P.PolicyNet.load(dev)
    assert net.meta["samples"] > 5000 and net.meta["train_accuracy"] > 0.95
    rows, summary = P.evaluate(net, 3, 1024, dev, log=None)
    print("[policy] held-out 1024^2:", json.dumps(summary), flush=True)
    for name in ("fp64", "mixed", "learned"):
        assert summary[name]["max_true_rel_res"] < 1e-8, name
    assert summary["learned"]["max_diff_vs_fp64"] < 1e-6
    assert summary["learned"]["iters"] <= summary["fp64"]["iters"] + 1
    # the network's thresholds stay in the swept range and the finest level is FP64
    for r in rows:
        lr = r["learned"]
        assert all(P.THETA_RANGE[0] <= t <= P.THETA_RANGE[1] for t in lr["thetas"])
        assert lr["precisions"][0] == L.AMGT_DOUBLE


def test_wrappers_reject_bad_arrays(dev):
    """Arrays the kernels would index out of bounds never reach a launch."""
    net = P.PolicyNet(dev)
    with pytest.raises(ValueError):
        net.forward_raw(torch.zeros((4, 11), dtype=torch.float32, device=dev))
    with pytest.raises(ValueError):
        net.forward_raw(torch.zeros((4, 12), dtype=torch.float64, device=dev))
    X = torch.zeros((4, 12), dtype=torch.float32, device=dev)
    with pytest.raises(ValueError):
        net.loss_grad_raw(X, torch.zeros(3, dtype=torch.int32, device=dev),
                          torch.zeros(4, dtype=torch.float32, device=dev), 1.0)
    g = torch.zeros(P.NPARAMS, dtype=torch.float32, device=dev)
    with pytest.raises(ValueError):
        net.adam_raw(g, g, torch.zeros(10, dtype=torch.float32, device=dev), 1, 1e-3)
