"""The Gaussian-erf factor table of fluxvortex_amd/csrc/fv_gauss_table.h (the
device evaluates G1 = g/rho^3, G2 = g'/rho^4 - 3g/rho^5, G3 = g'/rho^2 -
g/rho^3 from it for rho <= 10; g as in src/fluxvortex/kernel.py:26-31)."""
import importlib.util
import math
from pathlib import Path

import numpy as np
import pytest
from scipy.special import erf

mp = pytest.importorskip("mpmath")
CSRC = Path(__file__).resolve().parent.parent / "fluxvortex_amd" / "csrc"


def _gen():
    spec = importlib.util.spec_from_file_location("gen_gauss_table", CSRC / "gen_gauss_table.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


GEN = _gen()
TAB = GEN.table()


def horner(rho, which):
    """The device's evaluation (piece, local variable, Horner) in double."""
    k = min(int(rho * 2.0), GEN.PIECES - 1) if rho < 10.0 else GEN.PIECES - 1
    t = 4.0 * rho - (2 * k + 1)
    c = TAB[k][which - 1]
    v = c[GEN.DEG]
    for j in range(GEN.DEG - 1, -1, -1):
        v = v * t + c[j]
    return v


def test_header_is_the_generators_output():
    assert (CSRC / "fv_gauss_table.h").read_text() == GEN.header(TAB)


def test_table_within_a_few_ulp_of_the_50_digit_factors():
    rng = np.random.default_rng(5)
    rhos = np.concatenate([rng.uniform(0.0, 10.0, 300), np.arange(0, 10.01, 0.5),
                           [1e-9, 1e-4, 0.249, 0.999, 1.001, 9.999999]])
    worst = 0.0
    for rho in rhos:
        for w in (1, 2, 3):
            exact = GEN.factor(w, mp.mpf(float(rho)))
            err = abs(mp.mpf(horner(float(rho), w)) - exact) / GEN.term_scale(w, mp.mpf(float(rho)))
            worst = max(worst, float(err))
    assert worst < 4 * 2.0 ** -52, worst


def test_table_matches_the_closed_forms_where_they_do_not_cancel():
    # scipy's erf and exp in double, rho in [1, 10]: an independent evaluation
    s2p = math.sqrt(2.0 / math.pi)
    for rho in np.linspace(1.0, 10.0, 37):
        g = erf(rho / math.sqrt(2.0)) - s2p * rho * math.exp(-0.5 * rho * rho)
        gp = s2p * rho * rho * math.exp(-0.5 * rho * rho)
        ref = {1: g / rho ** 3, 2: gp / rho ** 4 - 3.0 * g / rho ** 5, 3: gp / rho ** 2 - g / rho ** 3}
        scale = {1: abs(g) / rho ** 3, 2: gp / rho ** 4 + 3.0 * abs(g) / rho ** 5,
                 3: gp / rho ** 2 + abs(g) / rho ** 3}
        for w in (1, 2, 3):
            assert abs(horner(rho, w) - ref[w]) <= 2e-15 * scale[w], (rho, w)


def test_factor_identity_and_limits():
    # G3 = rho^2 G2 + 2 G1 (so that CH r^2 + 2 CG = -G3 / (4 pi sigma^3)); a
    # swapped row or factor fails it
    for rho in np.linspace(0.0, 10.0, 81):
        g1, g2, g3 = (horner(rho, w) for w in (1, 2, 3))
        assert abs(g3 - (rho * rho * g2 + 2.0 * g1)) <= 1e-15 * (abs(g3) + rho * rho * abs(g2) + 2 * abs(g1))
    s2p = math.sqrt(2.0 / math.pi)
    assert horner(0.0, 1) == pytest.approx(s2p / 3.0, rel=1e-15)
    assert horner(0.0, 2) == pytest.approx(-s2p / 5.0, rel=1e-15)
    assert horner(0.0, 3) == pytest.approx(2.0 * s2p / 3.0, rel=1e-15)
    # rho = 10: g = 1 and g' = 0 in double, the far form the kernels switch to
    assert horner(10.0, 1) == pytest.approx(1e-3, rel=2e-16)
    assert horner(10.0, 3) == pytest.approx(-1e-3, rel=2e-16)
