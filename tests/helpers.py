"""Shared test helpers (normwise errors, seeded inputs)."""
import numpy as np


def relerr(a, b):
    """||a - b||_inf / ||b||_inf (absolute when b == 0)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if a.size else 0.0
    return float(num / den) if den > 0 else float(num)


def particle_cloud(n, seed=7, sigma=0.1, box=1.0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-box, box, (n, 3))
    alpha = rng.normal(0.0, 0.01, (n, 3))
    sig = np.full(n, sigma)
    return pos, alpha, sig


def sheet_rings(nx, ny, seed=42, lx=20.0, ly=10.0):
    """M = nx*ny rings on a perturbed z=0 sheet (the raw-kernel sweep recipe):
    lattice over x in [0, lx), y in [-ly/2, ly/2), vertices perturbed by
    N(0, (0.05 min(dx, dy))^2), Gamma ~ N(0, 1), rc0 = 0.03, age ~ U(0, 1)."""
    rng = np.random.default_rng(seed)
    dx, dy = lx / nx, ly / ny
    xs = np.arange(nx + 1) * dx
    ys = -ly / 2 + np.arange(ny + 1) * dy
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    V = np.stack([X, Y, np.zeros_like(X)], axis=-1)
    V = V + rng.normal(0.0, 0.05 * min(dx, dy), V.shape)
    fl = V[:-1, :-1].reshape(-1, 3)
    fr = V[:-1, 1:].reshape(-1, 3)
    bl = V[1:, :-1].reshape(-1, 3)
    br = V[1:, 1:].reshape(-1, 3)
    m = fl.shape[0]
    gamma = rng.normal(0.0, 1.0, m)
    rc0 = np.full(m, 0.03)
    age = rng.uniform(0.0, 1.0, m)
    return br.copy(), fr.copy(), fl.copy(), bl.copy(), gamma, rc0, age


def shear_freestream(X):
    """A non-uniform freestream (N, 3) at points X (N, 3) for advect_rk3's
    callable U_inf_func (the reference-written vectors and the tests use it)."""
    X = np.asarray(X, dtype=np.float64)
    return np.stack([10.0 + 0.3 * X[:, 2], 0.1 * X[:, 0], 0.5 - 0.2 * X[:, 1]], axis=1)


def swirl_bound(X):
    """A smooth stand-in for the bound vortices' velocity (N, 3) at points X,
    for advect_rk3's bound_velocity_func."""
    X = np.asarray(X, dtype=np.float64)
    w = 0.05 / (1.0 + X[:, 0] ** 2 + X[:, 1] ** 2)
    return np.stack([-w * X[:, 1], w * X[:, 0], np.zeros(X.shape[0])], axis=1)
