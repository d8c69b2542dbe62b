"""Monkey-patch surface: swap PteraSoftware's Numba Biot-Savart routines for
the MI355X kernels (reference: src/fluxvortex/warp_patch.py:25-143).

    import fluxvortex_amd as fv
    fv.patch()     # _aerodynamics_functions.* now run on the GPU
    ...            # run PteraSoftware simulations
    fv.unpatch()   # restore the originals

``patch(target)`` accepts any module/namespace that exposes the five routine
names.  Without ``target`` it patches ``pterasoftware._aerodynamics_functions``
and, when PteraSoftware is not installed, the bundled stand-in namespace
``fluxvortex_amd.aerodynamics_functions`` whose originals only raise (so the
patched surface can still be exercised).
"""
from __future__ import annotations

import time
import types

import numpy as np

from . import vortex_kernels as _vk

_NAMES = _vk.PTERA_FUNCTIONS

# per patched namespace: {"target": module, "originals": {...}, "active": bool}
_records: dict[int, dict] = {}
_default_target = None


def _missing_original(name):
    def fn(*args, **kwargs):
        raise RuntimeError(
            f"pterasoftware is not installed: its Numba original of {name} is unavailable")
    fn.__name__ = name
    return fn


def _stand_in_namespace():
    mod = types.ModuleType("fluxvortex_amd.aerodynamics_functions",
                           "Stand-in for pterasoftware._aerodynamics_functions.")
    for n in _NAMES:
        setattr(mod, n, _missing_original(n))
    return mod


aerodynamics_functions = _stand_in_namespace()


def default_target():
    """pterasoftware._aerodynamics_functions if importable, else the stand-in."""
    global _default_target
    if _default_target is None:
        try:
            from pterasoftware import _aerodynamics_functions as af  # type: ignore
            _default_target = af
        except ImportError:
            _default_target = aerodynamics_functions
    return _default_target


def _record(target):
    rec = _records.get(id(target))
    if rec is None:
        rec = {"target": target, "originals": {}, "active": False}
        _records[id(target)] = rec
    return rec


def _save_originals(target=None):
    """Saves the target's current routines once (warp_patch.py:30-40)."""
    target = default_target() if target is None else target
    rec = _record(target)
    for n in _NAMES:
        if hasattr(target, n) and n not in rec["originals"]:
            rec["originals"][n] = getattr(target, n)
    return rec


def patch(target=None, verbose=True):
    """Rebinds the five routines of ``target`` to the GPU versions (idempotent)."""
    target = default_target() if target is None else target
    rec = _save_originals(target)
    if rec["active"]:
        return
    for n in _NAMES:
        setattr(target, n, getattr(_vk, n))
    rec["active"] = True
    if verbose:
        print("[fluxvortex_amd] MI355X HIP kernels active (libfluxvortex_hip.so)")


def unpatch(target=None, verbose=True):
    """Restores the originals saved by patch() (idempotent)."""
    target = default_target() if target is None else target
    rec = _records.get(id(target))
    if rec is None or not rec["active"]:
        return
    for n, fn in rec["originals"].items():
        setattr(target, n, fn)
    rec["active"] = False
    if verbose:
        print("[fluxvortex_amd] Restored original CPU functions")


def is_patched(target=None):
    target = default_target() if target is None else target
    rec = _records.get(id(target))
    return bool(rec and rec["active"])


def originals(target=None):
    target = default_target() if target is None else target
    rec = _records.get(id(target))
    return dict(rec["originals"]) if rec else {}


def benchmark(N=500, M=2000, num_runs=5, target=None):
    """Times the saved CPU original against the GPU collapsed-ring routine on
    seeded random rings (warp_patch.py:75-143; the GPU time includes the host
    copies, as the reference's does)."""
    target = default_target() if target is None else target
    orig = originals(target).get("collapsed_velocities_from_ring_vortices")
    if orig is None:
        raise KeyError("collapsed_velocities_from_ring_vortices: call patch() or "
                       "_save_originals() first")
    rng = np.random.default_rng(42)
    points = rng.standard_normal((N, 3))
    br = rng.standard_normal((M, 3))
    fr = rng.standard_normal((M, 3))
    fl = rng.standard_normal((M, 3))
    bl = rng.standard_normal((M, 3))
    strengths = rng.standard_normal(M) * 0.5
    rc0s = np.full(M, 0.03)
    sing = np.zeros(4, dtype=np.int64)
    ages = rng.uniform(0, 1, M)
    nu = 1.5e-5
    args = (points, br, fr, fl, bl, strengths, rc0s, sing, ages, nu)
    orig(*args)
    cpu_times = []
    for _ in range(num_runs):
        sing[:] = 0
        t0 = time.perf_counter()
        cpu_result = orig(*args)
        cpu_times.append(time.perf_counter() - t0)
    gpu = _vk.collapsed_velocities_from_ring_vortices
    gpu(*args)
    gpu_times = []
    for _ in range(num_runs):
        sing[:] = 0
        t0 = time.perf_counter()
        gpu_result = gpu(*args)
        gpu_times.append(time.perf_counter() - t0)
    cpu_mean = float(np.mean(cpu_times))
    gpu_mean = float(np.mean(gpu_times))
    max_err = float(np.max(np.abs(cpu_result - gpu_result)))
    return {
        "cpu_ms": cpu_mean * 1000,
        "gpu_ms": gpu_mean * 1000,
        "speedup": cpu_mean / gpu_mean,
        "max_abs_error": max_err,
        "max_rel_error": max_err / (float(np.max(np.abs(cpu_result))) + 1e-16),
    }
