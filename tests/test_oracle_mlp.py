"""Pins of the warm-start network oracle: the analytic gradient against central finite differences,
and the loss at aligned / orthogonal outputs."""
import numpy as np

from oracle import mlp


def _params(K0, H, NO, rng):
    return rng.standard_normal(H * K0 + H + H * H + H + NO * H + NO) * 0.3


def test_gradient_matches_finite_differences():
    K0, H, NO, B = 16, 8, 8, 5
    rng = np.random.default_rng(1)
    p = _params(K0, H, NO, rng)
    X = rng.standard_normal((B, K0))
    V = rng.standard_normal((B, NO))
    g, _ = mlp.backward(p, X, V, K0, H, NO, round16=False)
    f = lambda q: mlp.loss(mlp.forward(q, X, K0, H, NO, round16=False)[0], V).mean()
    idx = rng.choice(p.size, 60, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (f(p + e) - f(p - e)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_loss_closed_form():
    V = np.array([[1.0, 2.0, 3.0]])
    assert abs(mlp.loss(-2.5 * V, V)[0]) < 1e-15             # aligned up to sign and scale
    assert abs(mlp.loss(np.array([[3.0, 0.0, -1.0]]), V)[0] - 1.0) < 1e-15   # orthogonal


def test_features_layout():
    S = np.array([[4.0, 2.0, 1.0], [10.0, 5.0, 0.0]])
    f = mlp.features(S, 3.5, 16)
    np.testing.assert_allclose(f[:7], [1.0, 0.5, 0.25, 1.0, 0.5, 0.0, 3.5])
    assert not f[7:].any()
