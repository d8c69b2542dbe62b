"""Pins the CPU oracle (oracle/fv_oracle.c) with closed forms, invariants,
an independent NumPy restatement and brute-force checks.  CPU only."""
import math

import numpy as np
import pytest

from helpers import particle_cloud, relerr


def seg(o, p, s, e, gamma=1.0, rc0=0.0, age=0.0, nu=0.0):
    """Velocity at points p of one straight segment s->e (via a 1-leg horseshoe)."""
    s = np.asarray(s, float)[None]
    e = np.asarray(e, float)[None]
    z = np.zeros((1, 3))
    # leg 0 of the ring mask runs Br -> Fr: put the segment there, mask chordwise
    # would add Fl->Bl; use the ring mask with the other 3 legs degenerate.
    return o.collapsed(np.atleast_2d(p), e * 0 + s, e, e, e, [gamma], [rc0], [age], nu,
                       "ring") if False else _one_leg(o, p, s, e, gamma, rc0, age, nu)


def _one_leg(o, p, s, e, gamma, rc0, age, nu):
    # corners (Br, Fr, Fl, Bl) = (s, e, e, s): legs Br->Fr = s->e, Fr->Fl zero
    # length, Fl->Bl = e->s (the reverse), Bl->Br zero length.  Chordwise mask
    # (legs 0 and 2) therefore cancels; the ring mask as well.  Use a horseshoe
    # (legs 0,1,2) with Fl == Fr so only leg 0 survives: Br=s, Fr=e, Fl=e, Bl=e.
    p = np.atleast_2d(np.asarray(p, float))
    return o.collapsed(p, s, e, e, e, np.array([gamma], float), np.array([rc0], float),
                       np.array([age], float), nu, "horseshoe")


def test_finite_segment_closed_form(oracle):
    L, h, G = 2.0, 0.7, 1.3
    v = _one_leg(oracle, [0.0, 0.0, h], np.array([[-L / 2, 0, 0]]), np.array([[L / 2, 0, 0]]),
                 G, 0.0, 0.0, 0.0)[0]
    expect = G / (4 * math.pi * h) * L / math.sqrt(h * h + L * L / 4)
    assert abs(np.linalg.norm(v) - expect) < 1e-14 * expect
    # direction: r1 x r2 with r1 = s - p, r2 = e - p
    r1 = np.array([-L / 2, 0, -h])
    r2 = np.array([L / 2, 0, -h])
    d = np.cross(r1, r2)
    assert np.allclose(v / np.linalg.norm(v), d / np.linalg.norm(d), atol=1e-14)


def test_square_ring_centre_closed_form(oracle):
    a, G = 0.8, 2.0
    h = a / 2
    Br, Fr, Fl, Bl = ([h, h, 0], [-h, h, 0], [-h, -h, 0], [h, -h, 0])
    v = oracle.collapsed(np.zeros((1, 3)), [Br], [Fr], [Fl], [Bl], [G], [0.0])[0]
    assert abs(np.linalg.norm(v) - 2 * math.sqrt(2) * G / (math.pi * a)) < 1e-14 * abs(v[2])
    assert v[0] == 0.0 and v[1] == 0.0
    # traversal Br->Fr->Fl->Bl is counter-clockwise about +z: positive G induces +z
    assert v[2] > 0


def test_infinite_line_limit(oracle):
    L, h, G = 1e5, 1.0, 1.0
    v = _one_leg(oracle, [0, h, 0], np.array([[-L / 2, 0, 0]]), np.array([[L / 2, 0, 0]]), G,
                 0.0, 0.0, 0.0)[0]
    assert abs(np.linalg.norm(v) - G / (2 * math.pi * h)) < 1e-9


def test_zero_length_and_on_line_give_zero(oracle):
    s = np.array([[0.0, 0, 0]])
    v = _one_leg(oracle, [[1.0, 2.0, 3.0]], s, s, 1.0, 0.0, 0.0, 0.0)
    assert np.all(v == 0.0)
    e = np.array([[1.0, 0, 0]])
    pts = [[0.5, 0, 0], [2.0, 0, 0], [-3.0, 0, 0], [0.0, 0, 0], [1.0, 0, 0]]
    v = _one_leg(oracle, pts, s, e, 1.0, 0.03, 0.5, 1e-5)
    assert np.all(v == 0.0)


def test_core_growth_law(oracle):
    # rc^2 = rc0^2 + 5.02572 (nu + 1e-4 |G|) age; on the bisector at distance
    # h the velocity is G/(4 pi) * (2l)(l^2 - (h^2 - L^2/4)) / (l^2 (h^2 L^2 + L^2 rc^2))
    # * h L with l = |r1| = |r2| = sqrt(h^2 + L^2/4).
    L, h, G, rc0, age, nu = 1.0, 0.05, -2.0, 0.03, 0.7, 1.5e-5
    v = _one_leg(oracle, [0, 0, h], np.array([[-L / 2, 0, 0]]), np.array([[L / 2, 0, 0]]), G,
                 rc0, age, nu)[0]
    rc2 = rc0 ** 2 + 5.02572 * (nu + 1e-4 * abs(G)) * age
    l = math.sqrt(h * h + L * L / 4)
    c3 = h * h - L * L / 4
    mag = abs(G) / (4 * math.pi) * (2 * l) * (l * l - c3) / (l * l * (h * h * L * L + L * L * rc2)) * h * L
    assert abs(np.linalg.norm(v) - mag) < 1e-13 * mag


def test_numpy_restatement_matches_c(oracle):
    pts, br, fr, fl, bl, g, rc, sing, ages, nu = oracle.seeded_rings()
    for legs in ("ring", "chordwise", "horseshoe"):
        a = ages if legs != "horseshoe" else None
        c = oracle.collapsed(pts, br, fr, fl, bl, g, rc, a, nu, legs)
        n = oracle.np_rings(pts, br, fr, fl, bl, g, rc, a, nu, legs)
        assert relerr(c, n) < 1e-13, legs


def test_collapsed_is_sum_of_expanded_and_legs(oracle):
    pts, br, fr, fl, bl, g, rc, sing, ages, nu = oracle.seeded_rings(N=50, M=40)
    c = oracle.collapsed(pts, br, fr, fl, bl, g, rc, ages, nu)
    e = oracle.expanded(pts, br, fr, fl, bl, g, rc, ages, nu)
    assert relerr(e.sum(axis=1), c) < 1e-13
    # ring = chordwise (legs 0, 2) + the other two legs; horseshoe = legs 0, 1, 2
    chord = oracle.collapsed(pts, br, fr, fl, bl, g, rc, ages, nu, "chordwise")
    hs = oracle.collapsed(pts, br, fr, fl, bl, g, rc, ages, nu, "horseshoe")
    # leg 3 alone = ring - horseshoe; leg 1 alone = horseshoe - chordwise
    leg3 = c - hs
    leg1 = hs - chord
    assert relerr(chord + leg1 + leg3, c) < 1e-13
    # leg 1 (Fr -> Fl) alone through a degenerate ring: corners (Fr, Fr, Fl, Fl)
    only1 = oracle.collapsed(pts, fr, fr, fl, fl, g, rc, ages, nu, "horseshoe")
    assert relerr(only1, leg1) < 1e-12


def test_linearity_in_strength(oracle):
    pts, br, fr, fl, bl, g, rc, sing, ages, nu = oracle.seeded_rings(N=30, M=20)
    a = oracle.collapsed(pts, br, fr, fl, bl, g, rc, None, nu)
    b = oracle.collapsed(pts, br, fr, fl, bl, 3.0 * g, rc, None, nu)
    assert relerr(b, 3.0 * a) < 1e-14


def test_divergence_free(oracle):
    pts, br, fr, fl, bl, g, rc, sing, ages, nu = oracle.seeded_rings(N=20, M=10)
    h = 1e-5
    div = np.zeros(pts.shape[0])
    scale = np.zeros(pts.shape[0])
    for k in range(3):
        d = np.zeros(3)
        d[k] = h
        vp = oracle.collapsed(pts + d, br, fr, fl, bl, g, rc, ages, nu)
        vm = oracle.collapsed(pts - d, br, fr, fl, bl, g, rc, ages, nu)
        div += (vp[:, k] - vm[:, k]) / (2 * h)
        scale += np.abs(vp - vm).sum(axis=1) / (2 * h)
    assert np.all(np.abs(div) <= 1e-6 * scale + 1e-9)


def test_aic_equals_collapsed_projection(oracle):
    rng = np.random.default_rng(3)
    P = 24
    br, fr, fl, bl = (rng.standard_normal((P, 3)) for _ in range(4))
    col = rng.standard_normal((P, 3))
    nrm = rng.standard_normal((P, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    rc = np.full(P, 0.02)
    A = oracle.aic(col, nrm, br, fr, fl, bl, rc, 1e-5)
    gam = rng.standard_normal(P)
    # AIC . Gamma == (collapsed with strengths Gamma, zero age) . n
    v = oracle.collapsed(col, br, fr, fl, bl, gam, rc, None, 1e-5)
    # the core uses |Gamma| only through the age term, which is zero here
    assert relerr(A @ gam, np.sum(v * nrm, axis=1)) < 1e-12


def test_lu_solve_matches_numpy(oracle):
    rng = np.random.default_rng(5)
    for n in (1, 7, 126):
        A = rng.standard_normal((n, n)) + n * np.eye(n) * 0.1
        b = rng.standard_normal(n)
        x = oracle.lu_solve(A, b)
        assert relerr(x, np.linalg.solve(A, b)) < 1e-11
    with pytest.raises(np.linalg.LinAlgError):
        oracle.lu_solve(np.zeros((3, 3)), np.ones(3))


# ------------------------------------------------------------------ particles
def test_winckelmans_g_half(oracle):
    # u_y at (0.5, 0, 0) from alpha = z_hat, sigma = 1 at the origin = g(0.5)/pi
    u = oracle.particle_vel([[0.5, 0, 0]], [[0, 0, 0]], [[0, 0, 1.0]], [1.0], "winckelmans")[0]
    g = 0.5 ** 3 * (0.25 + 2.5) / 1.25 ** 2.5
    assert abs(g - 0.19677) < 5e-6
    assert abs(u[1] - g / math.pi) < 1e-15 and u[0] == 0.0 and u[2] == 0.0


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_point_vortex_far_field(oracle, kernel):
    a = np.array([[0.3, -0.2, 0.5]])
    x = np.array([[40.0, 10.0, -25.0]])
    u = oracle.particle_vel(x, [[0, 0, 0]], a, [0.05], kernel)[0]
    r = np.linalg.norm(x[0])
    ref = -np.cross(x[0], a[0]) / (4 * math.pi * r ** 3)
    assert relerr(u, ref) < 1e-9


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_particle_numpy_restatement(oracle, kernel):
    pos, al, sg = particle_cloud(150, seed=11, sigma=0.2)
    tgt = np.concatenate([pos[:40], np.random.default_rng(1).uniform(-1, 1, (30, 3))])
    assert relerr(oracle.particle_vel(tgt, pos, al, sg, kernel),
                  oracle.np_particle_vel(tgt, pos, al, sg, kernel)) < 1e-12
    assert relerr(oracle.particle_jac(tgt, pos, al, sg, kernel),
                  oracle.np_particle_jac(tgt, pos, al, sg, kernel)) < 1e-12


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_jacobian_is_gradient_and_traceless(oracle, kernel):
    pos, al, sg = particle_cloud(60, seed=2, sigma=0.3)
    tgt = np.random.default_rng(9).uniform(-1, 1, (25, 3))
    J = oracle.particle_jac(tgt, pos, al, sg, kernel)
    h = 1e-5
    for k in range(3):
        d = np.zeros(3)
        d[k] = h
        fd = (oracle.particle_vel(tgt + d, pos, al, sg, kernel)
              - oracle.particle_vel(tgt - d, pos, al, sg, kernel)) / (2 * h)
        assert relerr(fd, J[:, :, k]) < 1e-7
    tr = np.trace(J, axis1=1, axis2=2)
    assert np.max(np.abs(tr)) < 1e-10 * np.max(np.abs(J))


def test_self_pair_excluded(oracle):
    pos, al, sg = particle_cloud(10, seed=4)
    J_all = oracle.particle_jac(pos, pos, al, sg, "winckelmans")
    J_others = np.stack([oracle.particle_jac(pos[i:i + 1], np.delete(pos, i, 0),
                                             np.delete(al, i, 0), np.delete(sg, i), "winckelmans")[0]
                         for i in range(10)])
    assert relerr(J_all, J_others) < 1e-14


# ------------------------------------------------------------------ VPM step
def test_rk3_constant_field_moves_by_U_dt(oracle):
    pos, al, sg = particle_cloud(20, seed=8)
    U = np.array([3.0, -1.0, 0.5])
    p2, g2, s2, a2 = oracle.vpm_advect_rk3(pos, al, sg, np.zeros(20), 0.01, U, False, False,
                                           "winckelmans", rlxf=0.0)
    assert np.allclose(p2, pos + U * 0.01, rtol=0, atol=1e-15)
    assert np.array_equal(g2, al) and np.array_equal(s2, sg)
    assert np.allclose(a2, 0.01)


def test_rk3_velocity_clamp(oracle):
    pos, al, sg = particle_cloud(5, seed=8)
    U = np.array([100.0, 0, 0])
    p2, *_ = oracle.vpm_advect_rk3(pos, al, sg, np.zeros(5), 0.01, U, False, False,
                                   "winckelmans", rlxf=0.0)
    assert np.allclose(p2 - pos, [0.5, 0, 0], atol=1e-14)


def test_rk3_gamma_and_sigma_clamps(oracle):
    pos, al, sg = particle_cloud(30, seed=3, sigma=0.3)
    al = al * 1e5          # |Gamma|^2 >> 1e4
    sg = sg.copy()
    sg[0] = 50.0           # above the sigma ceiling
    p2, g2, s2, _ = oracle.vpm_advect_rk3(pos, al, sg, np.zeros(30), 1e-3, np.zeros(3), True,
                                          True, "winckelmans", rlxf=0.0)
    assert np.all(np.sum(g2 * g2, axis=1) <= 1e4 * (1 + 1e-12))
    assert s2[0] <= 5.0 and np.all(s2 >= 1e-6)   # without the clip it stays near 50


def test_rk3_nan_protection(oracle):
    pos, al, sg = particle_cloud(6, seed=3)
    pos[2, 1] = np.nan
    _, g2, _, _ = oracle.vpm_advect_rk3(pos, al, sg, np.zeros(6), 1e-3, np.ones(3), False,
                                        False, "gaussianerf", rlxf=0.0)
    assert np.all(g2[2] == 0.0) and np.all(g2[[0, 1, 3, 4, 5]] == al[[0, 1, 3, 4, 5]])


def test_relaxation_full_alignment(oracle):
    pos, al, sg = particle_cloud(40, seed=21, sigma=0.4)
    _, g2, _, _ = oracle.vpm_advect_rk3(pos, al, sg, np.zeros(40), 1e-9, np.zeros(3), False,
                                        False, "winckelmans", rlxf=1.0)
    J = oracle.particle_jac(pos, pos, al, sg, "gaussianerf")
    w = np.stack([J[:, 2, 1] - J[:, 1, 2], J[:, 0, 2] - J[:, 2, 0], J[:, 1, 0] - J[:, 0, 1]], 1)
    wh = w / np.linalg.norm(w, axis=1, keepdims=True)
    gn = np.linalg.norm(al, axis=1, keepdims=True)
    assert relerr(g2, gn * wh) < 1e-6   # positions moved by 1e-9 * 0 only


def test_single_particle_relaxation_noop(oracle):
    _, g2, _, _ = oracle.vpm_advect_rk3([[0, 0, 0]], [[0.1, 0.2, 0.3]], [0.1], [0.0], 0.01,
                                        np.zeros(3), True, True, "gaussianerf", rlxf=0.3)
    assert np.array_equal(g2, [[0.1, 0.2, 0.3]])


# ------------------------------------------------------------------ conversion
def test_conversion_invariants(oracle):
    rng = np.random.default_rng(4)
    nr = 12
    fl = rng.standard_normal((nr, 3))
    fr = fl + [0, 0.3, 0]
    bl = fl + [0.2, 0, 0.01]
    br = fr + [0.2, 0, -0.01]
    g = rng.standard_normal(nr)
    g[3] = 0.0
    g[5] = 1e-16
    pos, al, sg, gout = oracle.convert_rings(br, fr, fl, bl, g, 0.05)
    assert pos.shape[0] == 4 * (nr - 2)
    assert np.all(gout[[i for i in range(nr) if i not in (3, 5)]] == 0.0)
    assert gout[3] == 0.0 and gout[5] == 1e-16
    k = 0
    for r in range(nr):
        if r in (3, 5):
            continue
        blk = al[k:k + 4]
        assert np.max(np.abs(blk.sum(axis=0))) < 1e-15 * np.max(np.abs(blk)) * 8
        assert np.allclose(pos[k], 0.5 * (fr[r] + fl[r]), atol=0)
        assert np.allclose(pos[k + 2], 0.5 * (bl[r] + br[r]), atol=0)
        k += 4
    assert np.all(sg == 0.05)


def test_conversion_far_field_matches_ring(oracle):
    a = 0.1
    fl, fr = np.array([[0, 0, 0.0]]), np.array([[0, a, 0.0]])
    bl, br = np.array([[a, 0, 0.0]]), np.array([[a, a, 0.0]])
    pos, al, sg, _ = oracle.convert_rings(br, fr, fl, bl, [1.5], 1e-4)
    x = np.array([[2.0, 1.0, 1.5]])   # |x| ~ 27 a
    up = oracle.particle_vel(x, pos, al, sg, "winckelmans")
    ur = oracle.collapsed(x, br, fr, fl, bl, [1.5], [0.0])
    assert relerr(up, ur) < 2e-2
