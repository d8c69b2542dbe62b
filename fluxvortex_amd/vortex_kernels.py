"""PteraSoftware-signature line-vortex Biot-Savart functions on the MI355X.

Drop-in replacements for the five ``pterasoftware._aerodynamics_functions``
routines (reference: src/fluxvortex/warp_kernels.py:251-382).  Inputs are
caller-owned (n, 3) float64 arrays; each call returns a new array.  The work
runs in libfluxvortex_hip.so: legs are preprocessed once into a segment table
and every target's sum is one tiled, atomic-free pass (no per-leg launches).

``singularity_counts`` is accepted and left untouched, as in the reference's
GPU path (warp_kernels.py never writes it).
"""
from __future__ import annotations

import numpy as np

from . import _lib

_state = {"precision": _lib.PREC_FP64}


def set_precision(precision: str) -> None:
    """Arithmetic of the collapsed sums: "fp64" (default) or "mixed"."""
    _state["precision"] = _lib.PRECISIONS[precision]


def get_precision() -> str:
    return {v: k for k, v in _lib.PRECISIONS.items()}[_state["precision"]]


def _prep(points, br, fr, fl, bl, strengths, r_c0s, ages):
    p = _lib.f64(points, (3,), "points")
    m = int(np.shape(strengths)[0])
    corners = [_lib.f64(c, (3,), "vertices") for c in (br, fr, fl, bl)]
    for c in corners:
        if c.shape[0] != m:
            raise ValueError(f"vertex arrays must have {m} rows, got {c.shape[0]}")
    g = _lib.f64(strengths)
    rc = _lib.f64(r_c0s)
    if rc.shape != (m,):
        raise ValueError(f"r_c0s must have shape ({m},), got {rc.shape}")
    a = None if ages is None else _lib.f64(ages)
    if a is not None and a.shape != (m,):
        raise ValueError(f"ages must have shape ({m},), got {a.shape}")
    return p, corners, g, rc, a, m


def _collapsed(points, br, fr, fl, bl, strengths, r_c0s, ages, nu, legs):
    p, (cbr, cfr, cfl, cbl), g, rc, a, m = _prep(points, br, fr, fl, bl, strengths, r_c0s, ages)
    n = p.shape[0]
    out = np.zeros((n, 3))
    if n == 0 or m == 0:          # warp_kernels.py:193-194
        return out
    ctx = _lib.default_context()
    P = _lib.ptr
    ctx.check(ctx.lib.fv_vortex_collapsed_host(
        ctx.handle, P(p), n, P(cbr), P(cfr), P(cfl), P(cbl), m, P(g), P(rc), _lib.optr(a),
        float(nu), legs, _state["precision"], P(out)))
    return out


def _expanded(points, br, fr, fl, bl, strengths, r_c0s, ages, nu, legs):
    p, (cbr, cfr, cfl, cbl), g, rc, a, m = _prep(points, br, fr, fl, bl, strengths, r_c0s, ages)
    n = p.shape[0]
    out = np.zeros((n, m, 3))
    if n == 0 or m == 0:          # warp_kernels.py:224-225
        return out
    ctx = _lib.default_context()
    P = _lib.ptr
    ctx.check(ctx.lib.fv_vortex_expanded_host(
        ctx.handle, P(p), n, P(cbr), P(cfr), P(cfl), P(cbl), m, P(g), P(rc), _lib.optr(a),
        float(nu), legs, P(out)))
    return out


def collapsed_velocities_from_ring_vortices(stackP_GP1_CgP1, stackBrrvp_GP1_CgP1,
                                            stackFrrvp_GP1_CgP1, stackFlrvp_GP1_CgP1,
                                            stackBlrvp_GP1_CgP1, strengths, r_c0s,
                                            singularity_counts, ages=None, nu=0.0):
    """Sum over rings of the 4 legs Br->Fr, Fr->Fl, Fl->Bl, Bl->Br (warp_kernels.py:251-276)."""
    return _collapsed(stackP_GP1_CgP1, stackBrrvp_GP1_CgP1, stackFrrvp_GP1_CgP1,
                      stackFlrvp_GP1_CgP1, stackBlrvp_GP1_CgP1, strengths, r_c0s, ages, nu,
                      _lib.LEGS_RING)


def collapsed_velocities_from_ring_vortices_chordwise_segments(
        stackP_GP1_CgP1, stackBrrvp_GP1_CgP1, stackFrrvp_GP1_CgP1, stackFlrvp_GP1_CgP1,
        stackBlrvp_GP1_CgP1, strengths, r_c0s, singularity_counts, ages=None, nu=0.0):
    """Only the chordwise legs Br->Fr and Fl->Bl (warp_kernels.py:279-302)."""
    return _collapsed(stackP_GP1_CgP1, stackBrrvp_GP1_CgP1, stackFrrvp_GP1_CgP1,
                      stackFlrvp_GP1_CgP1, stackBlrvp_GP1_CgP1, strengths, r_c0s, ages, nu,
                      _lib.LEGS_CHORDWISE)


def expanded_velocities_from_ring_vortices(stackP_GP1_CgP1, stackBrrvp_GP1_CgP1,
                                           stackFrrvp_GP1_CgP1, stackFlrvp_GP1_CgP1,
                                           stackBlrvp_GP1_CgP1, strengths, r_c0s,
                                           singularity_counts, ages=None, nu=0.0):
    """Per (point, ring) velocity, legs summed; shape (N, M, 3) (warp_kernels.py:305-330)."""
    return _expanded(stackP_GP1_CgP1, stackBrrvp_GP1_CgP1, stackFrrvp_GP1_CgP1,
                     stackFlrvp_GP1_CgP1, stackBlrvp_GP1_CgP1, strengths, r_c0s, ages, nu,
                     _lib.LEGS_RING)


def collapsed_velocities_from_horseshoe_vortices(stackP_GP1_CgP1, stackBrhvp_GP1_CgP1,
                                                 stackFrhvp_GP1_CgP1, stackFlhvp_GP1_CgP1,
                                                 stackBlhvp_GP1_CgP1, strengths, r_c0s,
                                                 singularity_counts, nu=0.0):
    """Legs Br->Fr, Fr->Fl, Fl->Bl with zero ages (warp_kernels.py:335-357)."""
    return _collapsed(stackP_GP1_CgP1, stackBrhvp_GP1_CgP1, stackFrhvp_GP1_CgP1,
                      stackFlhvp_GP1_CgP1, stackBlhvp_GP1_CgP1, strengths, r_c0s, None, nu,
                      _lib.LEGS_HORSESHOE)


def expanded_velocities_from_horseshoe_vortices(stackP_GP1_CgP1, stackBrhvp_GP1_CgP1,
                                                stackFrhvp_GP1_CgP1, stackFlhvp_GP1_CgP1,
                                                stackBlhvp_GP1_CgP1, strengths, r_c0s,
                                                singularity_counts, nu=0.0):
    """Per (point, horseshoe) velocity, zero ages (warp_kernels.py:360-382)."""
    return _expanded(stackP_GP1_CgP1, stackBrhvp_GP1_CgP1, stackFrhvp_GP1_CgP1,
                     stackFlhvp_GP1_CgP1, stackBlhvp_GP1_CgP1, strengths, r_c0s, None, nu,
                     _lib.LEGS_HORSESHOE)


PTERA_FUNCTIONS = (
    "collapsed_velocities_from_ring_vortices",
    "collapsed_velocities_from_ring_vortices_chordwise_segments",
    "expanded_velocities_from_ring_vortices",
    "collapsed_velocities_from_horseshoe_vortices",
    "expanded_velocities_from_horseshoe_vortices",
)


def aic(colloc, normals, br, fr, fl, bl, r_c0s, nu=0.0):
    """Fused influence matrix (P, P): unit-strength rings projected on unit normals."""
    c = _lib.f64(colloc, (3,), "colloc")
    n = _lib.f64(normals, (3,), "normals")
    p = c.shape[0]
    corners = [_lib.f64(v, (3,), "vertices") for v in (br, fr, fl, bl)]
    rc = _lib.f64(r_c0s)
    if n.shape[0] != p or any(v.shape[0] != p for v in corners) or rc.shape != (p,):
        raise ValueError("aic: colloc, normals, vertices and r_c0s must all have P rows")
    out = np.zeros((p, p))
    if p == 0:
        return out
    ctx = _lib.default_context()
    P = _lib.ptr
    ctx.check(ctx.lib.fv_aic_host(ctx.handle, P(c), P(n), p, *(P(v) for v in corners), P(rc),
                                  float(nu), P(out)))
    return out


def dense_solve(A, b, precision="fp64", max_iter=8, tol=1e-12, return_iters=False):
    """x = A^-1 b by LU with partial pivoting on the GPU; A (n, n) or (batch, n, n).
    precision "mixed": FP32 factors refined with FP64 residuals until the
    relative correction is below tol (at most max_iter refinement steps)."""
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    x = np.array(b, dtype=np.float64, order="C", copy=True)
    single = A.ndim == 2
    if single:
        A = A[None]
        x = x[None]
    bsz, n, n2 = A.shape
    if n != n2 or x.shape != (bsz, n):
        raise ValueError(f"dense_solve: shapes {A.shape} and {x.shape} do not match")
    info = np.zeros(bsz, dtype=np.int32)
    iters = np.zeros(bsz, dtype=np.int32)
    if bsz and n:
        ctx = _lib.default_context()
        if precision == "fp64":
            ctx.check(ctx.lib.fv_dense_solve_host(ctx.handle, bsz, n, _lib.ptr(A), _lib.ptr(x),
                                                  info.ctypes.data_as(_lib._i32p)))
        elif precision == "mixed":
            ctx.check(ctx.lib.fv_dense_solve_mixed_host(
                ctx.handle, bsz, n, _lib.ptr(A), _lib.ptr(x), int(max_iter), float(tol),
                info.ctypes.data_as(_lib._i32p), iters.ctypes.data_as(_lib._i32p)))
        else:
            raise ValueError(f"unknown precision {precision!r}")
    if np.any(info != 0):
        raise np.linalg.LinAlgError(f"Singular matrix (zero pivot, info={info.tolist()})")
    out = x[0] if single else x
    if return_iters:
        return out, (iters[0] if single else iters)
    return out
