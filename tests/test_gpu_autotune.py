"""The pair-sum tile-shape autotuner: tuned sums equal the default shape's to
round-off, each kind is tuned once and reported."""
import numpy as np
import pytest

from fluxvortex_amd import _lib
from fluxvortex_amd import kernel as K

pytestmark = pytest.mark.gpu


def test_autotuned_sums_match_default(ctx):
    rng = np.random.default_rng(5)
    n = 6000
    pos = rng.uniform(-5, 5, (n, 3))
    al = rng.normal(0, 0.01, (n, 3))
    sg = np.full(n, 0.2)
    ref_u = K.velocity_from_particles(pos, pos, al, sg, kernel="winckelmans")
    ref_J = K.jacobian_from_particles(pos, al, pos, al, sg, kernel="winckelmans")
    ctx.set_autotune(True)
    try:
        u = K.velocity_from_particles(pos, pos, al, sg, kernel="winckelmans")
        J = K.jacobian_from_particles(pos, al, pos, al, sg, kernel="winckelmans")
        rep = ctx.autotune_report()
        u2 = K.velocity_from_particles(pos, pos, al, sg, kernel="winckelmans")
        assert len(ctx.autotune_report()) == len(rep)          # cached: no new kinds
    finally:
        ctx.set_autotune(False)
    print("autotune report:", rep)
    assert len(rep) >= 2
    for fam, lt, ls, lanes, tpt, per_cu, ms in rep:
        assert lanes in (64, 256) and tpt in (1, 2, 4) and per_cu in (4, 8, 16) and ms > 0
    scale = np.abs(ref_u).max()
    assert np.abs(u - ref_u).max() < 1e-13 * scale
    assert np.array_equal(u, u2)
    assert np.abs(J - ref_J).max() < 1e-13 * np.abs(ref_J).max()
