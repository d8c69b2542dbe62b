"""Pins the mathematics of the closure's local (target-cell) expansion: the
field of a cell's monopole A + dipole M, u_i = -c eps_ijk (A_j dk phi -
M_jl dl dk phi) with phi = 1/R, its gradient J and Hessian H, checked by
finite differences, and the cross-product forms the device kernel evaluates
(fv_closure.hip, k_local_part) against the tensor forms."""
import itertools

import numpy as np

C = 1.0 / (4.0 * np.pi)
EPS = np.zeros((3, 3, 3))
for _i, _j, _k in itertools.permutations(range(3)):
    EPS[_i, _j, _k] = np.linalg.det(np.eye(3)[[_i, _j, _k]])


def _tensors(d):
    R = np.linalg.norm(d)
    I = np.eye(3)
    T1 = -d / R ** 3
    T2 = 3 * np.einsum("k,m->km", d, d) / R ** 5 - I / R ** 3
    T3 = (-15 * np.einsum("k,m,n->kmn", d, d, d) / R ** 7
          + 3 * (np.einsum("km,n->kmn", I, d) + np.einsum("kn,m->kmn", I, d)
                 + np.einsum("mn,k->kmn", I, d)) / R ** 5)
    dd = np.einsum("k,m->km", d, d)
    s2 = (np.einsum("km,nl->kmnl", I, dd) + np.einsum("kn,ml->kmnl", I, dd)
          + np.einsum("kl,mn->kmnl", I, dd) + np.einsum("mn,kl->kmnl", I, dd)
          + np.einsum("ml,kn->kmnl", I, dd) + np.einsum("nl,km->kmnl", I, dd))
    s0 = (np.einsum("km,nl->kmnl", I, I) + np.einsum("kn,ml->kmnl", I, I)
          + np.einsum("kl,mn->kmnl", I, I))
    T4 = 105 * np.einsum("k,m,n,l->kmnl", d, d, d, d) / R ** 9 - 15 * s2 / R ** 7 + 3 * s0 / R ** 5
    return T1, T2, T3, T4


def _tensor_form(d, A, M):
    T1, T2, T3, T4 = _tensors(d)
    u = -C * (np.einsum("ijk,j,k->i", EPS, A, T1) - np.einsum("ijk,jl,lk->i", EPS, M, T2))
    J = -C * (np.einsum("ijk,j,km->im", EPS, A, T2) - np.einsum("ijk,jl,lkm->im", EPS, M, T3))
    H = -C * (np.einsum("ijk,j,kmn->imn", EPS, A, T3) - np.einsum("ijk,jl,lkmn->imn", EPS, M, T4))
    return u, J, H


def _cross_form(d, A, M):
    """The device's evaluation order: v = M d, (j1, k1) = (i+1, i+2)."""
    R = np.sqrt(d @ d)
    iR3, iR5, iR7, iR9 = R ** -3, R ** -5, R ** -7, R ** -9
    v = M @ d
    Ad, vd = np.cross(A, d), np.cross(v, d)
    em = np.array([M[1, 2] - M[2, 1], M[2, 0] - M[0, 2], M[0, 1] - M[1, 0]])
    u = -C * (-iR3 * Ad - (3 * iR5 * vd - iR3 * em))
    J, H = np.zeros((3, 3)), np.zeros((3, 3, 3))
    for i in range(3):
        j1, k1 = (i + 1) % 3, (i + 2) % 3
        Y = M[j1] * d[k1] - M[k1] * d[j1]
        Am = np.array([(A[j1] if k1 == m else 0) - (A[k1] if j1 == m else 0) for m in range(3)])
        Vm = np.array([(v[j1] if k1 == m else 0) - (v[k1] if j1 == m else 0) for m in range(3)])
        for m in range(3):
            t2 = 3 * iR5 * Ad[i] * d[m] - iR3 * Am[m]
            s3 = -15 * iR7 * vd[i] * d[m] + 3 * iR5 * (em[i] * d[m] + Y[m] + Vm[m])
            J[i, m] = -C * (t2 - s3)
            for n in range(3):
                dmn = 1.0 if m == n else 0.0
                t3 = -15 * iR7 * Ad[i] * d[m] * d[n] + 3 * iR5 * (Am[m] * d[n] + Am[n] * d[m] + dmn * Ad[i])
                z = ((M[j1, m] if k1 == n else 0) - (M[k1, m] if j1 == n else 0)
                     + (M[j1, n] if k1 == m else 0) - (M[k1, n] if j1 == m else 0))
                s4 = (105 * iR9 * vd[i] * d[m] * d[n]
                      - 15 * iR7 * (em[i] * d[m] * d[n] + Y[m] * d[n] + Y[n] * d[m]
                                    + Vm[m] * d[n] + Vm[n] * d[m] + dmn * vd[i])
                      + 3 * iR5 * (em[i] * dmn + z))
                H[i, m, n] = -C * (t3 - s4)
    return u, J, H


def test_monopole_dipole_field_is_the_particle_sum_limit():
    """A cluster of particles far away: the first-order field of its moments
    approaches the exact point-vortex sum as (a / R)^2."""
    rng = np.random.default_rng(0)
    s = rng.uniform(-0.5, 0.5, (50, 3))
    al = rng.normal(size=(50, 3))
    A = al.sum(0)
    M = np.einsum("pj,pl->jl", al, s)
    errs = []
    for R in (20.0, 40.0):
        x = np.array([R, 0.3, -0.2])
        exact = -C * np.cross(x - s, al).T @ (1 / np.linalg.norm(x - s, axis=1) ** 3)
        u, _, _ = _tensor_form(x, A, M)
        errs.append(np.abs(u - exact).max() / np.abs(exact).max())
    assert 3.5 < errs[0] / errs[1] < 9.0


def test_gradient_and_hessian_by_finite_differences():
    rng = np.random.default_rng(1)
    A, M = rng.normal(size=3), rng.normal(size=(3, 3))
    x = np.array([12.0, -3.0, 2.0])
    u, J, H = _tensor_form(x, A, M)
    e = 1e-4
    Jfd = np.stack([(_tensor_form(x + e * np.eye(3)[m], A, M)[0]
                     - _tensor_form(x - e * np.eye(3)[m], A, M)[0]) / (2 * e) for m in range(3)], 1)
    Hfd = np.stack([(_tensor_form(x + e * np.eye(3)[n], A, M)[1]
                     - _tensor_form(x - e * np.eye(3)[n], A, M)[1]) / (2 * e) for n in range(3)], 2)
    assert np.abs(J - Jfd).max() < 1e-8 * np.abs(J).max()
    assert np.abs(H - Hfd).max() < 1e-7 * np.abs(H).max()
    assert np.allclose(H, H.transpose(0, 2, 1), rtol=0, atol=1e-15 * np.abs(H).max())


def test_cross_product_form_equals_tensor_form():
    rng = np.random.default_rng(2)
    for _ in range(20):
        d = rng.normal(size=3) * rng.uniform(1, 50)
        A, M = rng.normal(size=3), rng.normal(size=(3, 3))
        for a, b in zip(_tensor_form(d, A, M), _cross_form(d, A, M)):
            assert np.abs(a - b).max() < 1e-13 * np.abs(a).max()
