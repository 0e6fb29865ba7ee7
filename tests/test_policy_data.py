"""The committed policy artifacts (CPU only): the solver-generated labels are
well formed, and the trained weights, evaluated by the oracle's double-precision
restatement of the network on the standardised features, reproduce the labels
(the accuracy and threshold error recorded at training time)."""
import json
import os

import numpy as np

from amgt_amd import policy as P
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_labels_well_formed():
    d = np.load(os.path.join(ROOT, "amgt_amd", "policy_labels.npz"))
    X, C, T = d["features"], d["cls"], d["theta"]
    assert X.shape[1] == P.NF and X.shape[0] == C.size == T.size > 5000
    assert set(np.unique(C)) <= {0, 1, 2}
    assert np.all(np.min(np.abs(T.astype(np.float64)[:, None] - np.array(P.THETAS)[None, :]), axis=1) < 1e-6)
    # finest levels (level index 0) are FP64 and have relative sizes 0
    fin = X[:, 5] == 0
    assert np.all(C[fin] == 0) and np.allclose(X[fin, :2], 0.0)
    assert int(d["count"]) == 1500 and int(d["side"]) == 256
    # per-problem times of every threshold; the label rule picks from them
    S = d["theta_times"].astype(np.float64)
    assert S.shape[1] == len(P.THETAS) and S.shape[0] <= 1500 and np.all(S > 0)
    best = S.min(axis=1)
    dflt = S[:, P.THETAS.index(P.DEFAULT_THETA)]
    want = np.where(dflt <= best * (1 + P.LABEL_TIE), P.DEFAULT_THETA,
                    np.array(P.THETAS)[S.argmin(axis=1)])
    # every level of a problem carries that problem's threshold; problems are
    # consecutive runs of level samples that start at a finest level
    starts = np.nonzero(X[:, 5] == 0)[0]
    assert starts.size == S.shape[0]
    assert np.allclose(T[starts].astype(np.float64), want, atol=1e-6)


def test_weights_reproduce_labels():
    w = json.load(open(P.WEIGHTS_PATH))
    assert len(w["params"]) == O.POLICY_PARAMS and tuple(w["features"]) == P.FEATURE_NAMES
    d = np.load(os.path.join(ROOT, "amgt_amd", "policy_labels.npz"))
    X, C, T = d["features"].astype(np.float64), d["cls"], d["theta"].astype(np.float64)
    Z = ((X - np.array(w["mean"])) / np.array(w["std"])).astype(np.float32).astype(np.float64)
    out = O.policy_forward(np.array(w["params"], dtype=np.float64), Z)
    acc = float((out[:, :3].argmax(axis=1) == C).mean())
    theta = np.clip(w["theta_mid"] + w["theta_half"] * out[:, 3], *P.THETA_RANGE)
    mae = float(np.abs(theta - T).mean())
    assert abs(acc - w["meta"]["train_accuracy"]) < 2e-3, (acc, w["meta"]["train_accuracy"])
    assert acc > 0.99 and mae < 0.05, (acc, mae)


def test_stall_rule():
    """The in-solve check of the learned policy: a solve stops when the last w
    iterations reduced the residual by less than q^w, never before 2w+1
    residuals exist, and not at all without a rule."""
    from amgt_amd.amg import stalled
    fast = [0.2 ** k for k in range(12)]
    assert not any(stalled(fast[:k], (3, 0.7)) for k in range(1, 13))
    slow = [0.9 ** k for k in range(12)]
    assert [stalled(slow[:k], (3, 0.7)) for k in range(1, 9)] == [False] * 6 + [True, True]
    assert not stalled(slow, None)
    # a rate exactly at the bound is not a stall; just above it is
    edge = [0.7 ** k for k in range(8)]
    assert not stalled(edge, (3, 0.7 + 1e-12)) and stalled(edge, (3, 0.7 - 1e-3))
