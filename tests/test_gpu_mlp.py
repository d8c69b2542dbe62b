"""GPU closure MLP (BF16 MFMA forward / backward / Adam) against the CPU
oracle, which emulates the same bf16 rounding points in FP64.  Bounds: forward
<= 5e-3 and gradients <= 2e-2 normwise relative per tensor (FP32 MFMA sums and
FP32 SiLU can flip a bf16 rounding); Adam <= 1e-5."""
import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mlpmod(ctx):
    from fluxvortex_amd import mlp
    return mlp


def batch(n, seed=0):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 32), np.float32)
    X[:, :31] = rng.normal(0, 1, (n, 31))
    T = np.zeros((n, 16), np.float32)
    T[:, :12] = rng.normal(0, 1, (n, 12))
    w = np.zeros(16, np.float32)
    w[:12] = 1.0
    return X, T, w


def test_forward_matches_oracle(oracle, mlpmod):
    p = mlpmod.init_params(n_in=31, n_out=12, seed=3)
    net = mlpmod.ClosureMLP(p)
    for n in (1, 16, 37, 1000):
        X, _, _ = batch(n, seed=n)
        y = net.forward(X)
        ref, _ = oracle.mlp_forward(p.astype(np.float64), X.astype(np.float64))
        assert relerr(y[:, :12], ref[:, :12]) < 5e-3, n
        assert np.all(y[:, 12:] == 0.0)


def test_gradients_match_oracle(oracle, mlpmod):
    p = mlpmod.init_params(n_in=31, n_out=12, seed=4)
    net = mlpmod.ClosureMLP(p)
    X, T, w = batch(300, seed=5)
    g = net.gradients(X, T, w)
    ref = oracle.mlp_gradients(p.astype(np.float64), X.astype(np.float64),
                               T.astype(np.float64), w.astype(np.float64))
    gd, rd = oracle.mlp_unpack(g.astype(np.float64)), oracle.mlp_unpack(ref)
    for name, _ in oracle.MLP_SHAPES:
        if name == "b4":
            assert relerr(gd[name][:12], rd[name][:12]) < 2e-2
            continue
        assert relerr(gd[name], rd[name]) < 2e-2, name
    assert np.all(gd["W4"][12:] == 0) and np.all(gd["W1"][:, 31:] == 0)


def test_adam_matches_oracle(oracle, mlpmod):
    p = mlpmod.init_params(n_in=31, n_out=12, seed=6)
    net = mlpmod.ClosureMLP(p)
    rng = np.random.default_rng(1)
    m = np.zeros(p.size)
    v = np.zeros(p.size)
    x = p.astype(np.float64)
    for k in range(1, 4):
        g = rng.normal(0, 1e-2, p.size).astype(np.float32)
        net.adam(g, 1e-3)
        x, m, v = oracle.adam_step(x, g.astype(np.float64), m, v, k, 1e-3)
    assert relerr(net.get_params(), x) < 1e-5


def test_training_reduces_loss(oracle, mlpmod):
    rng = np.random.default_rng(2)
    n = 2048
    X = np.zeros((n, 32), np.float32)
    X[:, :3] = rng.uniform(-1, 1, (n, 3))
    T = np.zeros((n, 16), np.float32)
    T[:, 0] = np.sin(2 * X[:, 0]) * X[:, 1]
    T[:, 1] = X[:, 2] ** 2
    w = np.zeros(16, np.float32)
    w[:2] = 1.0
    net = mlpmod.ClosureMLP(mlpmod.init_params(n_in=3, n_out=2, seed=7))
    y0 = net.forward(X)
    l0 = float(np.mean(np.sum((y0[:, :2] - T[:, :2]) ** 2, axis=1)))
    for _ in range(300):
        net.adam(net.gradients(X, T, w), 3e-3)
    y1 = net.forward(X)
    l1 = float(np.mean(np.sum((y1[:, :2] - T[:, :2]) ** 2, axis=1)))
    assert l1 < 0.05 * l0, (l0, l1)
