"""Vortex-particle Biot-Savart sums on the MI355X.

Same API as the reference's src/fluxvortex/kernel.py:69-190 (and its GPU twin
warp_vpm.py:52-173): velocity u(x) = -(1/4pi) sum_j g(r/sigma_j)/r^3 (x-x_j) x
alpha_j and its gradient J[t, i, j] = du_i/dx_j, with the Gaussian-erf or
Winckelmans smoothing kernel.  The self pair (r < 1e-12) is excluded, as in
warp_vpm.py:37-38,103-104.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def _kernel_id(kernel):
    try:
        return _lib.KERNELS[kernel]
    except KeyError:
        raise KeyError(f"unknown kernel {kernel!r}; expected one of {sorted(_lib.KERNELS)}")


def _call(fn_name, comp, targets, src_pos, src_gamma, src_sigma, kernel, precision):
    t = _lib.f64(targets, (3,), "targets")
    p = _lib.f64(src_pos, (3,), "src_pos")
    g = _lib.f64(src_gamma, (3,), "src_gamma")
    s = _lib.f64(src_sigma)
    n = p.shape[0]
    if g.shape[0] != n or s.shape != (n,):
        raise ValueError("src_pos, src_gamma and src_sigma must describe the same particles")
    out = np.zeros((t.shape[0],) + comp)
    if t.shape[0] == 0 or n == 0:
        return out
    ctx = _lib.default_context()
    P = _lib.ptr
    ctx.check(getattr(ctx.lib, fn_name)(ctx.handle, P(t), t.shape[0], P(p), P(g), P(s), n,
                                        _kernel_id(kernel), _lib.PRECISIONS[precision], P(out)))
    return out


def velocity_from_particles(target_points, src_pos, src_gamma, src_sigma,
                            kernel="gaussianerf", precision="fp64"):
    """Induced velocity (M, 3) at target_points from N particles (kernel.py:69-116)."""
    return _call("fv_particle_velocity_host", (3,), target_points, src_pos, src_gamma,
                 src_sigma, kernel, precision)


def jacobian_from_particles(tgt_pos, tgt_gamma, src_pos, src_gamma, src_sigma,
                            kernel="gaussianerf", precision="fp64"):
    """Velocity gradient J (M, 3, 3), J[t, i, j] = dU_i/dx_j (kernel.py:119-190).
    tgt_gamma is unused, as in the reference."""
    return _call("fv_particle_jacobian_host", (3, 3), tgt_pos, src_pos, src_gamma, src_sigma,
                 kernel, precision)
