"""Host-side geometry and case builders (CPU)."""
import math

import numpy as np

from fluxvortex_amd import geometry as G
from fluxvortex_amd import hybrid


def test_spacing():
    assert np.allclose(G.spacing(4, "uniform"), [0, 0.25, 0.5, 0.75, 1])
    c = G.spacing(6, "cosine")
    assert c[0] == 0 and abs(c[-1] - 1) < 1e-15 and abs(c[3] - 0.5) < 1e-15
    assert c[1] < 1 / 6


def test_naca_camber():
    assert np.all(G.naca4_camber(np.linspace(0, 1, 11), "0012") == 0)
    z = G.naca4_camber(np.array([0.0, 0.4, 1.0]), "2412")
    assert abs(z[1] - 0.02) < 1e-15 and z[0] == 0 and abs(z[2]) < 1e-15


def test_symmetric_lattice_mirrors():
    s = G.Surface(1.0, 1.0, (0.0, 5.0, 0.0), nc=4, ns=3)
    (g, side), = s.lattices()
    assert g.shape == (5, 7, 3)
    assert np.allclose(g[:, ::-1, 1], -g[:, :, 1])
    assert np.allclose(g[0, :, 0], 0) and np.allclose(g[-1, :, 0], 1)


def test_off_plane_root_gives_two_lattices():
    s = G.Surface(1.75, 1.5, (0.75, 6.0, 1.0), nc=6, ns=8, root_le=(0, 0.5, 0))
    lat = s.lattices()
    assert len(lat) == 2 and lat[0][0].shape == (7, 9, 3)
    assert np.allclose(lat[1][0][..., 1], -lat[0][0][:, ::-1, 1])


def test_motion_heave_and_flap():
    s = G.Surface(1.0, 1.0, (0.0, 5.0, 0.0), nc=2, ns=2, flap=True)
    ap = G.Airplane([s], G.Motion(heave_amp=0.1, heave_period=2.0))
    g = ap.grids(3, 0.5)
    assert np.allclose(g[1, :, 2] - g[0, :, 2], 0.1)
    ap = G.Airplane([s], G.Motion(flap_amp_deg=15, flap_period=1.0))
    g = ap.grids(2, 0.25)
    r0 = np.linalg.norm(g[0][:, 1:], axis=1)
    r1 = np.linalg.norm(g[1][:, 1:], axis=1)
    assert np.allclose(r0, r1)          # rotation about x
    tip = np.argmax(g[0][:, 1])
    assert abs(math.degrees(math.atan2(g[1][tip, 2], g[1][tip, 1])) - 15) < 1e-9


def test_config_builders():
    c2 = hybrid.heaving_wing_case()
    assert c2.panels == 120 and c2.num_steps == 150 and c2.n_keep == 10
    assert abs(c2.sigma - 0.5 * 10 * c2.dt) < 1e-15
    assert abs(c2.meta["period"] - 2 * math.pi / 10) < 1e-15
    c1 = hybrid.static_wing_case()
    assert c1.panels == 252 and c1.num_steps == 42
    c1h = hybrid.static_wing_case(symmetric=False)
    assert c1h.panels == 126
    c5 = hybrid.flapping_case()
    assert c5.panels == 192 and len(c5.nc) == 3 and c5.num_steps == 129
