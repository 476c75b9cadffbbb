"""Pins the network restatement (oracle/mlp.py): its float64 backward pass
against central finite differences of its loss, parameter by parameter on
a random subset of every block; the BF16 mode differs from float64 by the
BF16 rounding only; the padded input columns get zero gradient."""
import numpy as np
import pytest

from oracle import mlp


def _data(m=40, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (m, 8))
    t = rng.normal(0, 1, m)
    return X, t


def test_backward_matches_finite_differences():
    P = mlp.init(3).astype(np.float64)
    X, t = _data()
    g, loss = mlp.grad(P, X, t, bf=False)
    rng = np.random.default_rng(0)
    blocks = [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2),
              (mlp.OFF_W2, mlp.OFF_B2), (mlp.OFF_B2, mlp.OFF_W3),
              (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.NPARAM)]
    for a, b in blocks:
        idx = rng.choice(np.arange(a, b), size=min(12, b - a), replace=False)
        if a == mlp.OFF_W1:  # only the live input columns move the loss
            idx = [i for i in idx if (i - a) % mlp.KIN < mlp.NIN] or [a]
        for i in idx:
            h = 1e-6
            Pp, Pm = P.copy(), P.copy()
            Pp[i] += h
            Pm[i] -= h
            fd = (mlp.grad(Pp, X, t, bf=False)[1] -
                  mlp.grad(Pm, X, t, bf=False)[1]) / (2 * h)
            assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), (a, i)
    # padded input columns
    W1g = g[mlp.OFF_W1:mlp.OFF_B1].reshape(mlp.NH, mlp.KIN)
    assert np.all(W1g[:, mlp.NIN:] == 0)


def test_bf16_mode_is_a_rounding_of_float64():
    P = mlp.init(4)
    X, t = _data(200, 2)
    y64, _ = mlp.forward(P, X, bf=False)
    y16, _ = mlp.forward(P, X, bf=True)
    assert not np.array_equal(y64, y16)
    assert np.abs(y16 - y64).max() <= 0.05 * np.abs(y64).max()
    g64, _ = mlp.grad(P, X, t, bf=False)
    g16, _ = mlp.grad(P, X, t, bf=True)
    assert np.abs(g16 - g64).max() <= 0.05 * np.abs(g64).max()
    # BF16 rounding: nearest even, 8 significant bits
    v = np.array([1.0, 1.00390625, 1.005859375, 3.0e-3], np.float32)
    r = mlp.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] - 3e-3) <= 3e-3 * 2 ** -8


def test_adam_reduces_the_loss():
    P = mlp.init(5)
    X, t = _data(256, 3)
    t = np.tanh(X[:, 0]) + 0.5 * X[:, 1]  # a learnable target
    m = np.zeros_like(P)
    v = np.zeros_like(P)
    l0 = mlp.grad(P, X, t)[1]
    for k in range(1, 101):
        g, _ = mlp.grad(P, X, t)
        P, m, v = mlp.adam(P, g.astype(np.float32), m, v, k, 3e-3)
    assert mlp.grad(P, X, t)[1] < 0.3 * l0
