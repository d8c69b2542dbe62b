"""Monkey-patch surface (reference src/fluxvortex/warp_patch.py:25-143)."""
import types

import numpy as np
import pytest

from fluxvortex_amd import ptera_patch, vortex_kernels

NAMES = vortex_kernels.PTERA_FUNCTIONS


def namespace_with(fns):
    ns = types.SimpleNamespace()
    for n in NAMES:
        setattr(ns, n, fns[n])
    return ns


def oracle_originals(oracle):
    def ring(P, br, fr, fl, bl, g, rc, sing, ages=None, nu=0.0):
        return oracle.collapsed(P, br, fr, fl, bl, g, rc, ages, nu, "ring")

    def chord(P, br, fr, fl, bl, g, rc, sing, ages=None, nu=0.0):
        return oracle.collapsed(P, br, fr, fl, bl, g, rc, ages, nu, "chordwise")

    def ering(P, br, fr, fl, bl, g, rc, sing, ages=None, nu=0.0):
        return oracle.expanded(P, br, fr, fl, bl, g, rc, ages, nu, "ring")

    def hs(P, br, fr, fl, bl, g, rc, sing, nu=0.0):
        return oracle.collapsed(P, br, fr, fl, bl, g, rc, None, nu, "horseshoe")

    def ehs(P, br, fr, fl, bl, g, rc, sing, nu=0.0):
        return oracle.expanded(P, br, fr, fl, bl, g, rc, None, nu, "horseshoe")

    return dict(zip(NAMES, (ring, chord, ering, hs, ehs)))


def test_patch_unpatch_rebinds_and_restores(oracle, capsys):
    orig = oracle_originals(oracle)
    ns = namespace_with(orig)
    ptera_patch.patch(ns)
    assert ptera_patch.is_patched(ns)
    for n in NAMES:
        assert getattr(ns, n) is getattr(vortex_kernels, n)
    ptera_patch.patch(ns)                       # idempotent
    ptera_patch.unpatch(ns)
    for n in NAMES:
        assert getattr(ns, n) is orig[n]
    ptera_patch.unpatch(ns)                     # idempotent
    out = capsys.readouterr().out
    assert out.count("active") == 1 and out.count("Restored") == 1


def test_default_target_is_stand_in_without_pterasoftware():
    try:
        import pterasoftware  # noqa: F401
        pytest.skip("pterasoftware is installed")
    except ImportError:
        pass
    t = ptera_patch.default_target()
    assert t is ptera_patch.aerodynamics_functions
    with pytest.raises(RuntimeError, match="pterasoftware is not installed"):
        t.collapsed_velocities_from_ring_vortices(None, None, None, None, None, None, None, None)


def test_benchmark_needs_saved_originals():
    ns = types.SimpleNamespace()
    with pytest.raises(KeyError):
        ptera_patch.benchmark(target=ns)


@pytest.mark.gpu
def test_patched_namespace_runs_on_gpu_and_benchmark(oracle, ctx):
    orig = oracle_originals(oracle)
    ns = namespace_with(orig)
    ptera_patch.patch(ns, verbose=False)
    try:
        args = oracle.seeded_rings()
        out = ns.collapsed_velocities_from_ring_vortices(*args)
        ref = orig["collapsed_velocities_from_ring_vortices"](*args)
        assert np.max(np.abs(out - ref)) / np.max(np.abs(ref)) < 1e-12
        res = ptera_patch.benchmark(N=500, M=2000, num_runs=2, target=ns)
        assert res["max_rel_error"] < 1e-12 and res["gpu_ms"] > 0
    finally:
        ptera_patch.unpatch(ns, verbose=False)
