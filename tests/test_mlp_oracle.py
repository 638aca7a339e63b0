"""The numpy oracle of the learned limiter's network, pinned independently:
its float64 backward pass against central finite differences of its loss,
its BF16 rounding against hand-computed values, its features against closed
forms."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def _problem(seed, m=9):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.2, M.NPARAM)
    X = rng.normal(0, 1, (m, M.IN))
    X[:, 125:] = 0
    T = rng.normal(0, 1, (m, 16))
    return p, X, T


def test_backward_matches_finite_differences():
    p, X, T = _problem(1)
    loss, g = M.loss_grad(p, X, T, bf=False)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(M.NPARAM, 300, replace=False), np.arange(M.NPARAM - 130, M.NPARAM)])
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = M.loss_grad(q, X, T, bf=False)
        q[i] -= 2 * h
        lm, _ = M.loss_grad(q, X, T, bf=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # the check has teeth: some hidden units are inactive, so a backward pass
    # without the ReLU masks would differ
    w = M.unpack(p)
    assert np.any(np.maximum(X @ w["W1"].T + w["b1"], 0) @ w["W2"].T + w["b2"] < 0)


def test_bf16_rounding():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7 + 2 ** -9, -2.5, 0.0, 3.0e38], dtype=np.float32)
    got = M.bf16(x)
    exp = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 0.0], dtype=np.float32)
    assert np.array_equal(got[:6], exp)
    assert abs(got[6] - 3.0e38) / 3.0e38 < 2 ** -8


def test_features_closed_forms():
    nx, ny = 9, 7
    i = np.tile(np.arange(nx), ny).astype(float)
    rho = 2.0 + i          # linear in x
    e = 5.0 + 0 * i        # constant
    ux = np.zeros((ny + 1) * (nx + 1))
    uy = np.repeat(np.arange(ny + 1), nx + 1).astype(float)  # cell-mean v = j + 1/2
    X, S = M.features(nx, ny, rho, e, ux, uy)
    c = 3 * nx + 4  # interior cell (4, 3)
    di = np.tile(np.arange(-2, 3), 5)
    dj = np.repeat(np.arange(-2, 3), 5)
    assert np.allclose(X[c, 0:25], di / (2 + 1e-6 * 6), atol=1e-6)
    assert np.all(X[c, 25:50] == 0)
    assert np.allclose(X[c, 75:100], dj / (2 + 1e-6 * 3.5), atol=1e-6)
    assert np.all(X[:, 125:] == 0)
    # edge cells repeat the boundary: at i = 0 the three left columns equal the centre
    assert np.all(X[0, 0:25].reshape(5, 5)[:, :3] == 0)
    assert np.isclose(S[c, 0], 2 + 6e-6)
