"""Pins the CPU MLP oracle: backprop vs central finite differences (FP64),
bf16 rounding, Adam against its closed form on a quadratic."""
import numpy as np

from helpers import relerr


def test_bf16_rounding(oracle):
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0, 65504.0])
    b = oracle.bf16(x)
    assert b[0] == 1.0 and b[1] == 1.0          # tie to even
    assert b[2] == 1.0078125                    # rounds up
    assert abs(b[3] + 3.140625) < 1e-12 and b[4] == 0.0
    assert np.all(np.abs(b - x) <= np.abs(x) * 2.0 ** -8)


def test_backprop_matches_finite_differences(oracle):
    rng = np.random.default_rng(0)
    from fluxvortex_amd.mlp import init_params
    p = init_params(n_in=31, n_out=12, seed=1).astype(np.float64)
    p[oracle.MLP_NPARAMS - 16:] = rng.normal(0, 0.1, 16)      # non-zero b4
    X = np.zeros((7, oracle.MLP_IN))
    X[:, :31] = rng.normal(0, 1, (7, 31))
    T = rng.normal(0, 1, (7, 16))
    w = np.zeros(16)
    w[:12] = rng.uniform(0.5, 2, 12)
    g = oracle.mlp_gradients(p, X, T, w, emulate_bf16=False)
    idx = rng.choice(oracle.MLP_NPARAMS, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (oracle.mlp_loss(p + e, X, T, w) - oracle.mlp_loss(p - e, X, T, w)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_adam_quadratic(oracle):
    # minimise 0.5 |x - c|^2: Adam's first step moves every coordinate by ~lr
    c = np.array([3.0, -2.0, 0.5])
    x = np.zeros(3)
    m = np.zeros(3)
    v = np.zeros(3)
    x1, m, v = oracle.adam_step(x, x - c, m, v, 1, 0.1)
    assert relerr(x1, 0.1 * np.sign(c)) < 1e-6
    for k in range(2, 400):
        x1, m, v = oracle.adam_step(x1, x1 - c, m, v, k, 0.1)
    assert np.max(np.abs(x1 - c)) < 1e-2
