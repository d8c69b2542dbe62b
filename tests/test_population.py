"""Pins the oracle's particle population control and circulation ledger
(newton_pc/adapters/flap.py:466-578) with hand-solved cases and invariants."""
import numpy as np
import pytest


def _cloud(n, seed, aligned=0.02):
    rng = np.random.default_rng(seed)
    pos = rng.uniform([0, -3, -0.5], [30, 3, 0.5], (n, 3))
    al = rng.normal(0, 0.01, (n, 3))
    al[:, 1] += aligned
    sg = rng.uniform(0.05, 0.15, n)
    age = np.arange(n, dtype=float)
    return pos, al, sg, age


REF = np.array([[0.5 + 0.2 * k, -1.0 + 0.25 * k, 0.0] for k in range(9)])


def test_drop_keeps_the_newest(oracle):
    pos, al, sg, age = _cloud(100, 0)
    p, a, s, g = oracle.pop_drop(pos, al, sg, age, 40)
    assert len(s) == 40
    assert np.array_equal(p, pos[60:]) and np.array_equal(a, al[60:])
    assert np.array_equal(s, sg[60:]) and np.array_equal(g, age[60:])
    p2, *_ = oracle.pop_drop(pos, al, sg, age, 200)
    assert np.array_equal(p2, pos)


def test_merge_two_aligned_particles_closed_form(oracle):
    """Two aligned particles in one cell far from the reference points (so the
    error test passes) plus `protect` newest ones: alpha sum, |alpha|-weighted
    centroid, max sigma; the survivor keeps the first slot and its age."""
    pos = np.array([[20.0, 0.01, 0.0], [20.02, 0.0, 0.01]] + [[1.0, 0.0, 0.0]] * 2)
    al = np.array([[0.0, 1.0, 0.0], [0.0, 3.0, 0.1]] + [[0.0, 1.0, 0.0]] * 2)
    sg = np.array([0.1, 0.2, 0.1, 0.1])
    age = np.array([5.0, 4.0, 1.0, 0.0])
    p, a, s, g, nm = oracle.pop_merge(pos, al, sg, age, REF, eps=1.0, vinf_norm=10.0, protect=2)
    assert nm == 1 and len(s) == 3
    w1, w2 = 1.0, np.sqrt(9.01)
    assert np.allclose(a[0], al[0] + al[1], rtol=0, atol=1e-15)
    assert np.allclose(p[0], (w1 * pos[0] + w2 * pos[1]) / (w1 + w2), rtol=0, atol=1e-14)
    assert s[0] == 0.2 and g[0] == 5.0
    assert np.array_equal(p[1:], pos[2:])


def test_merge_rules(oracle):
    """Misaligned pairs stay; the greedy scan takes consecutive pairs of one
    cell in particle order (three in a cell: the first two); protected
    particles never merge; a strict error threshold rejects everything."""
    base = np.array([20.0, 0.0, 0.0])
    pos = np.array([base, base + 0.01, base + 0.02, [2.0, 0, 0], [2.0, 0.0, 0.001]])
    al = np.array([[0, 1.0, 0], [0, 1.0, 0], [0, 1.0, 0], [0, 1.0, 0], [0, 1.0, 0]])
    sg = np.full(5, 0.1)
    age = np.zeros(5)
    _, a, _, _, nm = oracle.pop_merge(pos, al, sg, age, REF, eps=1.0, vinf_norm=10.0, protect=2)
    assert nm == 1 and np.allclose(a[0], [0, 2.0, 0]) and np.allclose(a[1], [0, 1.0, 0])
    al2 = al.copy()
    al2[1] = [0, -1.0, 0]   # 0-1 misaligned, 1-2 misaligned -> nothing in that cell
    _, a2, _, _, nm2 = oracle.pop_merge(pos, al2, sg, age, REF, eps=1.0, vinf_norm=10.0,
                                        protect=2)
    assert nm2 == 0 and len(a2) == 5
    _, _, _, _, nm3 = oracle.pop_merge(pos, al, sg, age, REF, eps=0.0, vinf_norm=10.0, protect=2)
    assert nm3 == 0
    # fewer than 2 * protect particles: no merge at all
    _, _, _, _, nm4 = oracle.pop_merge(pos, al, sg, age, REF, eps=1.0, vinf_norm=10.0, protect=3)
    assert nm4 == 0


def _vel_rm(R, X, A, S):
    """Velocity of the merge criterion's kernel g = rho^3 / (rho^2 + 1)^1.5."""
    r = R[:, None, :] - X[None, :, :]
    d = np.linalg.norm(r, axis=-1) + 1e-30
    rho = d / S[None, :]
    g = rho ** 3 / (rho ** 2 + 1.0) ** 1.5
    return ((g / (4 * np.pi * d ** 3))[..., None] * np.cross(A[None], r)).sum(axis=1)


@pytest.mark.parametrize("eps,protect_dist", [(1e-2, 0.0), (1e-4, 0.0), (1e-3, 8.0)])
def test_merge_invariants(oracle, eps, protect_dist):
    """On a random cloud: sum(alpha) is kept to round-off; every accepted pair
    moves the reference-point velocity by less than eps |V| per component, so
    the whole pass moves it by less than n_merged eps |V|; the protected
    particles are untouched."""
    pos, al, sg, age = _cloud(4000, 3)
    V = 10.0
    p, a, s, g, nm = oracle.pop_merge(pos, al, sg, age, REF, eps=eps, vinf_norm=V, protect=64,
                                      protect_dist=protect_dist)
    assert nm > 100 and len(s) == len(sg) - nm
    assert np.abs(a.sum(0) - al.sum(0)).max() < 1e-12 * np.abs(al).sum()
    dv = np.abs(_vel_rm(REF, p, a, s) - _vel_rm(REF, pos, al, sg)).max()
    assert dv < nm * eps * V
    c = REF.mean(axis=0)
    if protect_dist > 0:
        near = np.linalg.norm(pos - c, axis=1) <= protect_dist
        kept = {tuple(x) for x in p}
        assert all(tuple(x) in kept for x in pos[near])
    else:
        assert np.array_equal(p[-64:], pos[-64:]) and np.array_equal(a[-64:], al[-64:])
    # age follows the surviving slot: ages stay increasing
    assert np.all(np.diff(g) > 0)


def test_ledger(oracle):
    """Particles: sum of alpha; closed rings: G (b - a) over the four legs
    sums to zero."""
    pos, al, sg, _ = _cloud(500, 5)
    assert np.allclose(oracle.ledger(al), al.sum(0), rtol=1e-14, atol=1e-15)
    rng = np.random.default_rng(7)
    br, fr, fl, bl = (rng.normal(size=(20, 3)) for _ in range(4))
    g = rng.normal(size=20)
    out = oracle.ledger(al, br, fr, fl, bl, g)
    assert np.abs(out - al.sum(0)).max() < 1e-13
