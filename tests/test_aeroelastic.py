"""Pins the aeroelastic host logic (beam element matrices, envelope growth,
flutter-speed interpolation) and the oracle's coupled UVLM-beam run, against
closed forms, the reference's published Goland numbers and its stored Goland
histories (tests/golden/goland_flutter_ref.npz, from figures/flutter_data.npz)."""
import math
from pathlib import Path

import numpy as np
import pytest

from fluxvortex_amd import aeroelastic as ae

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "goland_flutter_ref.npz")


def test_goland_beam_frequencies_match_published():
    """8 elements with the CG offset: 7.70 Hz bending, 21.14 Hz torsion
    (reference README.md:137)."""
    f = ae.goland_beam().natural_frequencies(2)
    assert abs(f[0] - 7.70) < 0.005, f
    assert abs(f[1] - 21.14) < 0.005, f


def test_uncoupled_beam_converges_to_closed_forms():
    """Without the CG offset the bending and torsion families separate; with
    32 elements the first cantilever frequencies reach Euler-Bernoulli's
    (1.8751^2 / L^2) sqrt(EI / m) and the shaft's (pi / 2L) sqrt(GJ / Ip)
    (the README's analytic 7.88 / 18.27 Hz) within 0.05 %."""
    L, EI, GJ, m, c = 6.096, 9.773e6, 0.988e6, 35.72, 1.8288
    Ip = m * c * c / 24.0
    b = ae.Beam(length=L, n_elements=32, EI=EI, GJ=GJ, m=m, Ip=Ip, x_ea_cg=0.0)
    f = b.natural_frequencies()
    fb = 1.875104068711961 ** 2 / L ** 2 * math.sqrt(EI / m) / (2 * math.pi)
    ft = math.pi / (2 * L) * math.sqrt(GJ / Ip) / (2 * math.pi)
    assert abs(fb - 7.877) < 0.001 and abs(ft - 18.27) < 0.01
    assert np.min(np.abs(f - fb)) < 5e-4 * fb
    assert np.min(np.abs(f - ft)) < 5e-4 * ft


def test_newmark_matrices_and_initial_state():
    """K_eff = K + gamma/(beta dt) C + M/(beta dt^2) on the free DOFs; the
    initial state is a tip displacement at rest with a = M^-1 (-K d)."""
    b = ae.goland_beam()
    dt = 0.003
    M, Cm, Kinv = b.newmark_matrices(dt)
    K, M2, C2 = b.reduced()
    Keff = K + 0.5 / (0.25 * dt) * C2 + M2 / (0.25 * dt * dt)
    assert np.allclose(np.linalg.inv(Kinv), Keff, rtol=1e-10, atol=1e-6 * np.abs(Keff).max())
    d, v, a = b.initial_state(0.05, math.radians(2.0))
    assert d[-3] == 0.05 and d[-1] == math.radians(2.0) and not v.any()
    assert np.allclose(M @ a, -K @ d, atol=1e-9 * np.abs(K @ d).max())
    # Rayleigh damping C = (2 zeta / omega_1) K
    w1 = 2 * math.pi * b.natural_frequencies(1)[0]
    assert np.allclose(Cm, 2 * 0.005 / w1 * K, rtol=1e-12, atol=0)


@pytest.mark.parametrize("sigma", [-0.9, -0.05, 0.3, 1.4])
def test_envelope_growth_recovers_exponential_rate(sigma):
    """A sampled A e^(sigma t) cos(omega t) at the Goland bending frequency:
    the peak fit returns sigma within 2e-3 1/s."""
    dt = 0.003
    t = np.arange(400) * dt
    s = 0.05 * np.exp(sigma * t) * np.cos(2 * math.pi * 7.7 * t + 0.3)
    assert abs(ae.envelope_growth(s, dt) - sigma) < 2e-3


def test_envelope_growth_short_signals_and_flutter_speed():
    assert ae.envelope_growth(np.ones(9), 0.01) == 0.0
    assert ae.envelope_growth(np.linspace(0, 1, 50), 0.01) == 0.0   # no peaks
    V = np.linspace(80, 180, 11)
    assert abs(ae.flutter_speed(V, 0.02 * (V - 137.3)) - 137.3) < 1e-9
    assert ae.flutter_speed(V, -np.ones_like(V)) is None
    s = np.array([-1.0, 0.5, -0.2, 0.7])   # first crossing wins
    assert abs(ae.flutter_speed([0, 1, 2, 3], s) - 1.0 / 1.5) < 1e-12


def _stations(F):
    """Station values from nodal loads on nodes 1..8: each station sits
    midway between two nodes and splits half and half; node 0 is clamped."""
    L = np.zeros_like(F)
    L[..., -1] = 2 * F[..., -1]
    for k in range(F.shape[-1] - 2, -1, -1):
        L[..., k] = 2 * F[..., k] - L[..., k + 1]
    return L


def _loads_from_histories(w, th, dt=0.003):
    """Nodal loads implied by stored nodal w / theta histories: each Newmark
    step K_eff d_new = F + M(...) + C(...) solved for F (zero on the slope
    DOFs) and the slopes."""
    b = ae.goland_beam()
    M, Cm, Kinv = b.newmark_matrices(dt)
    Keff = np.linalg.inv(Kinv)
    x, v, a = b.initial_state(0.05, math.radians(2.0))
    n = len(x)
    wi, si, ti = np.arange(0, n, 3), np.arange(1, n, 3), np.arange(2, n, 3)
    kn = np.concatenate([wi, ti])
    out = []
    for k in range(len(w)):
        S = M @ (x / (0.25 * dt * dt) + v / (0.25 * dt) + a) + Cm @ (2 * x / dt + v)
        dn = np.zeros(n)
        dn[wi], dn[ti] = w[k][1:], th[k][1:]
        dn[si] = np.linalg.solve(Keff[np.ix_(si, si)], S[si] - Keff[np.ix_(si, kn)] @ dn[kn])
        F = Keff @ dn - S
        an = (dn - x) / (0.25 * dt * dt) - v / (0.25 * dt) - a
        v = v + dt * 0.5 * (a + an)
        x, a = dn, an
        out.append((F[wi], F[ti]))
    return np.array(out)


@pytest.mark.parametrize("key,V,w_tol,th_tol", [("stable", 100.0, 0.02, 0.03),
                                                 ("flutter", 160.0, 0.004, 0.007)])
def test_oracle_reproduces_reference_goland_histories(oracle, key, V, w_tol, th_tol):
    """The reference's stored Goland runs (PteraSoftware ring wake, 60 chords):
    the oracle's tip heave / twist histories stay within w_tol / th_tol of
    the stored ones relative to their peaks, and the station lift / moment
    the stored histories imply (inverted through the beam) agree with the
    oracle's at the last step (steady) within 1e-4 -- which pins the beam,
    Newmark, the load distribution, the -F_z station lift, the moment arm
    and the half-and-half share of the lines between rows."""
    ref_w, ref_t = GOLD["w_" + key], GOLD["theta_" + key]
    inst = ae.goland_instance(V, num_chords=60, n_keep=100000)
    r = oracle.aeroelastic_run(inst)
    ok = ~np.isnan(r["tip_w"])
    w, th = r["tip_w"][ok], r["tip_theta"][ok]
    assert len(w) == len(ref_w)
    assert np.abs(w - ref_w[:, -1]).max() < w_tol * np.abs(ref_w[:, -1]).max()
    assert np.abs(th - ref_t[:, -1]).max() < th_tol * np.abs(ref_t[:, -1]).max()
    loads = _loads_from_histories(ref_w, ref_t)
    lift, mom = _stations(loads[-1, 0]), _stations(loads[-1, 1])
    steps = np.nonzero(ok)[0]
    my_lift, my_mom = r["station_lift"][steps[-1]], r["station_moment"][steps[-1]]
    assert np.abs(my_lift - lift).max() < 1e-4 * np.abs(lift).max()
    assert np.abs(my_mom - mom).max() < 1e-4 * np.abs(mom).max()


def test_oracle_ring_wake_sweep_reproduces_published_flutter_speed(oracle):
    """benchmark_goland.py's sweep (ring wake, 100 chords) as published in
    README.md:121-132: sigma_w per speed within 0.13 1/s and the interpolated
    flutter speed within 0.5 % of 140.2 m/s."""
    Vs = [80, 100, 120, 130, 140, 144, 160, 180]
    pub = [-0.628, -0.899, -0.249, -0.101, -0.037, 0.564, 0.794, 1.413]
    sig = []
    for V in Vs:
        inst = ae.goland_instance(float(V), n_keep=100000)
        r = oracle.aeroelastic_run(inst)
        sig.append(ae.envelope_growth(r["tip_w"][~np.isnan(r["tip_w"])], inst.dt))
    assert np.abs(np.array(sig) - pub).max() < 0.13, sig
    Vf = ae.flutter_speed(Vs, sig)
    assert Vf is not None and abs(Vf - 140.2) < 0.005 * 140.2, Vf


def test_oracle_feedback_switch(oracle):
    """With feedback the deformation reaches the aerodynamics (loads change);
    without it the loads are those of the rigid lattice, so the CL history
    equals the uncoupled run's exactly."""
    base = dict(num_chords=8)
    rigid = oracle.aeroelastic_run(ae.goland_instance(140.0, couple=False, **base))
    nofb = oracle.aeroelastic_run(ae.goland_instance(140.0, **base))
    fb = oracle.aeroelastic_run(ae.goland_instance(140.0, feedback=True, **base))
    assert np.array_equal(rigid["cl"], nofb["cl"])
    assert np.abs(fb["cl"] - rigid["cl"]).max() > 1e-4


def test_oracle_sweep_free_ring_wake_differs_from_prescribed(oracle):
    """The oracle's aeroelastic run with a free ring wake: finite, the same
    first step as the prescribed wake (no wake yet) and different loads once
    the wake has rolled."""
    a = oracle.aeroelastic_run(ae.goland_instance(120.0, num_chords=5, ring_free_wake=True))
    b = oracle.aeroelastic_run(ae.goland_instance(120.0, num_chords=5))
    assert np.all(np.isfinite(a["cl"])) and np.all(np.isfinite(a["gamma"]))
    assert a["cl"][0] == b["cl"][0]
    assert np.abs(a["cl"] - b["cl"]).max() > 1e-6


def test_oracle_sweep_wake_truncation(oracle):
    """Truncating the ring wake at x = 9 m changes the loads once the wake has
    passed it, and leaves the earlier steps unchanged."""
    a = oracle.aeroelastic_run(ae.goland_instance(140.0, num_chords=10, n_keep=100000,
                                                  truncate_x=9.0))
    b = oracle.aeroelastic_run(ae.goland_instance(140.0, num_chords=10, n_keep=100000))
    # the wake starts at the trailing edge (x ~ 1.9 m) and moves V dt per step:
    # no ring passes x = 9 m in the first k steps
    k = int((9.0 - 2.5) / (140.0 * 0.003))
    assert np.array_equal(a["cl"][:k], b["cl"][:k])
    assert np.abs(a["cl"] - b["cl"]).max() > 1e-8
