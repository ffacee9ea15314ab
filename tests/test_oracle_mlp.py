"""The oracle's restatement of the learned centre proposals, pinned on the CPU:
its FP64 gradient against central finite differences of its own loss (with a
planted mistake that the check must catch), the bf16 rounding points against
the FP64 network, and the features against values worked out by hand."""
import numpy as np
import pytest

import components_amd as ca
import oracle

P = oracle.MLP_PARAMS
GROUPS = {"W1": (0, 384), "b1": (384, 448), "W2": (448, 4544), "b2": (4544, 4608), "w3": (4608, 4672),
          "b3": (4672, 4673)}


def rand_params(rng, scale=0.3):
    w = rng.normal(0, scale, P)
    w[GROUPS["b1"][0]:GROUPS["b1"][1]] = rng.normal(0.1, 0.1, 64)  # keep most first-layer units active
    return w


def rand_data(rng, n):
    X = np.column_stack([rng.uniform(0, 8, n), rng.uniform(0, 8, n), rng.uniform(0, 10, n), rng.uniform(0, 14, n),
                         rng.uniform(0, 1, n), rng.integers(0, 4, n)]).astype(np.float32)
    y = rng.uniform(0, 6, n).astype(np.float32)
    return X, y


def loss_of(w, X, y):
    s = oracle.mlp_forward(w, X, bf16=False)
    return 0.5 * np.mean((s - y.astype(np.float64)) ** 2)


def fd_check(w, X, y, grad, idxs, h=1e-6):
    errs = []
    for i in idxs:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (loss_of(wp, X, y) - loss_of(wm, X, y)) / (2 * h)
        errs.append(abs(fd - grad[i]) / max(1e-8, abs(fd) + abs(grad[i])))
    return max(errs)


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = rand_params(rng)
    X, y = rand_data(rng, 40)
    g, loss = oracle.mlp_grad(w, X, y, bf16=False)
    assert loss == pytest.approx(loss_of(w, X, y), rel=1e-12)
    idxs = []
    for lo, hi in GROUPS.values():
        idxs += list(rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False))
    assert fd_check(w, X, y, g, idxs) < 1e-5


def test_finite_difference_check_catches_a_planted_mistake():
    rng = np.random.default_rng(4)
    w = rand_params(rng)
    X, y = rand_data(rng, 40)
    g, _ = oracle.mlp_grad(w, X, y, bf16=False)
    for name in ("W1", "b2", "w3"):
        lo, hi = GROUPS[name]
        bad = g.copy()
        bad[lo:hi] *= -1.0  # a sign error in one layer's gradient
        idxs = list(rng.choice(np.arange(lo, hi), size=min(8, hi - lo), replace=False))
        assert fd_check(w, X, y, bad, idxs) > 0.5, name


def test_bf16_points_stay_close_to_fp64():
    rng = np.random.default_rng(5)
    w = rand_params(rng)
    X, y = rand_data(rng, 200)
    s64 = oracle.mlp_forward(w, X, bf16=False)
    s16 = oracle.mlp_forward(w, X, bf16=True)
    assert np.linalg.norm(s16 - s64) / np.linalg.norm(s64) < 2e-2
    g64, _ = oracle.mlp_grad(w, X, y, bf16=False)
    g16, _ = oracle.mlp_grad(w, X, y, bf16=True)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 5e-2
    # weights and inputs already bf16, hidden layer exactly representable: identical forward
    w_b = np.round(w * 8) / 8
    X_b = np.round(X * 4) / 4
    w_b[GROUPS["W1"][0]:GROUPS["W2"][1]] = 0
    assert np.allclose(oracle.mlp_forward(w_b, X_b, bf16=True), oracle.mlp_forward(w_b, X_b, bf16=False), rtol=0,
                       atol=1e-12)


def test_features_by_hand():
    # vertex 0: star centre of 1..4 plus triangle 0-5-6 (neighbours sorted: 1,2,3,4,5,6)
    edges = [(0, 1), (0, 2), (0, 3), (0, 4), (0, 5), (0, 6), (5, 6), (1, 2)]
    g = ca.Graph.from_edges(7, edges)
    X = oracle.level_features(g.offsets, g.targets, round_index=3)
    deg = np.diff(g.offsets.astype(np.int64))
    assert list(deg) == [6, 2, 2, 1, 1, 2, 2]
    # vertex 0: samples at positions 0, 6//4=1, 12//4=3, 18//4=4 -> neighbours 1, 2, 4, 5
    sampled = [2, 2, 1, 2]
    assert X[0, 0] == pytest.approx(np.log2(7))
    assert X[0, 1] == pytest.approx(np.log2(1 + np.mean(sampled)))
    assert X[0, 2] == pytest.approx(np.log2(3))
    assert X[0, 3] == pytest.approx(np.log2(1 + 6 * np.mean(sampled)))
    # pairs (1,2): 2 in row(1) -> hit; (4,5): 5 not in row(4) = [0] -> miss
    assert X[0, 4] == pytest.approx(0.5)
    assert np.all(X[:, 5] == 3)
    # vertex 3: one neighbour (0, degree 6), no pair
    assert X[3, 1] == pytest.approx(np.log2(7)) and X[3, 2] == pytest.approx(np.log2(7)) and X[3, 4] == 0
    # vertex 5: neighbours 0 and 6, adjacent (6 is in row(0)'s first 8 entries)
    assert X[5, 4] == 1.0
