"""The oracle reproduces the committed golden fixtures (CPU)."""
from pathlib import Path

import numpy as np

from helpers import relerr

G = Path(__file__).parent / "golden"


def test_seeded_rings_fixture(oracle):
    z = np.load(G / "seeded_rings.npz")
    a = (z["points"], z["br"], z["fr"], z["fl"], z["bl"], z["strengths"], z["rc0s"])
    nu = float(z["nu"])
    assert relerr(oracle.collapsed(*a, z["ages"], nu, "ring"), z["collapsed_ring"]) < 1e-14
    assert relerr(oracle.collapsed(*a, z["ages"], nu, "chordwise"),
                  z["collapsed_chordwise"]) < 1e-14
    assert relerr(oracle.collapsed(*a, None, nu, "horseshoe"), z["collapsed_horseshoe"]) < 1e-14
    e = (z["points"][:50],) + a[1:]
    assert relerr(oracle.expanded(*e, z["ages"], nu, "ring"), z["expanded_ring"]) < 1e-14
    # the recipe itself is the reference's (tests/test_correctness.py:19-31)
    p2 = oracle.seeded_rings()
    assert np.array_equal(p2[0], z["points"]) and np.array_equal(p2[5], z["strengths"])


def test_particles_fixture(oracle):
    z = np.load(G / "particles.npz")
    pos, al, sg = z["pos"], z["alpha"], z["sigma"]
    for k in ("gaussianerf", "winckelmans"):
        assert relerr(oracle.particle_vel(pos, pos, al, sg, k), z[f"vel_{k}"]) < 1e-14
        assert relerr(oracle.particle_jac(pos, pos, al, sg, k), z[f"jac_{k}"]) < 1e-14
        p2, g2, s2, _ = oracle.vpm_advect_rk3(pos, al, sg, np.zeros(len(sg)), 0.01,
                                              np.array([10.0, 0, 0.5]), True, True, k)
        assert relerr(p2, z[f"rk3_{k}_pos"]) < 1e-14
        assert relerr(g2, z[f"rk3_{k}_gamma"]) < 1e-13
