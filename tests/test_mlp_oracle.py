"""The MLP oracle itself: analytic gradients against central finite
differences in float64, and masked softmax invariants."""
import numpy as np

from oracle import mlp


def test_param_count():
    assert mlp.NPARAMS == 7728


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    flat = mlp.init_params(3).astype(np.float64)
    x = rng.normal(size=(16, mlp.IN))
    y = rng.integers(0, mlp.OUT, 16)
    mask = np.ones((16, mlp.OUT), dtype=np.uint8)
    mask[:, 40:] = 0
    y = np.minimum(y, 39)
    _, g = mlp.loss_and_grad(flat, x, y, mask)
    idx = rng.choice(mlp.NPARAMS, 60, replace=False)
    eps = 1e-6
    for i in idx:
        d = np.zeros_like(flat)
        d[i] = eps
        lp, _ = mlp.loss_and_grad(flat + d, x, y, mask)
        lm, _ = mlp.loss_and_grad(flat - d, x, y, mask)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_masked_softmax():
    rng = np.random.default_rng(2)
    flat = mlp.init_params(1)
    x = rng.normal(size=(8, mlp.IN))
    mask = (rng.random((8, mlp.OUT)) < 0.5).astype(np.uint8)
    mask[:, 0] = 1
    prob, _ = mlp.forward(flat, x, mask)
    assert np.allclose(prob.sum(1), 1)
    assert (prob[mask == 0] == 0).all()


def test_adam_closed_form_first_step():
    # first step moves each parameter by -lr * sign(g) (up to eps)
    flat = np.zeros(10)
    g = np.linspace(-1, 1, 10)
    g[5] = 0.3
    new, m, v = mlp.adam(flat, g, np.zeros(10), np.zeros(10), 1, 0.01)
    assert np.allclose(new, -0.01 * np.sign(g), atol=1e-7)
