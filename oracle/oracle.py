"""Python access to the CPU FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.  Two independent restatements
of the reference live here:
  * ctypes bindings of oracle/_build/libfv_oracle.so (fv_oracle.c, plain C,
    no FMA contraction, OpenMP over targets);
  * small NumPy restatements (``np_*``) of the segment and particle kernels,
    written from the reference's NumPy code (newton_pc/adapters/flap.py:199-228,
    src/fluxvortex/kernel.py:69-190) and used only to cross-check the C code
    on small inputs.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from math import erf, pi, sqrt
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libfv_oracle.so"

_dp = C.POINTER(C.c_double)
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        i64 = C.c_int64
        L.orc_collapsed.argtypes = [_dp, i64, _dp, _dp, _dp, _dp, i64, _dp, _dp, _dp, C.c_double,
                                    C.c_int, _dp, C.c_int]
        L.orc_expanded.argtypes = L.orc_collapsed.argtypes
        L.orc_aic.argtypes = [_dp, _dp, i64, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp, C.c_int]
        L.orc_lu_solve.argtypes = [i64, _dp, _dp]
        L.orc_particle_vel.argtypes = [_dp, i64, _dp, _dp, _dp, i64, C.c_int, _dp, C.c_int]
        L.orc_particle_jac.argtypes = L.orc_particle_vel.argtypes
        L.orc_vpm_advect_rk3.argtypes = [i64, _dp, _dp, _dp, _dp, C.c_double, _dp, C.c_int,
                                         C.c_int, C.c_int, C.c_double, C.c_int]
        L.orc_vpm_relax.argtypes = [i64, _dp, _dp, _dp, C.c_double, C.c_int]
        L.orc_particle_vel_rule.argtypes = [_dp, i64, _dp, _dp, _dp, i64, C.c_int, C.c_int, _dp,
                                            C.c_int]
        L.orc_particle_jac_rule.argtypes = L.orc_particle_vel_rule.argtypes
        L.orc_vpm_advect_rk3_rule.argtypes = [i64, _dp, _dp, _dp, _dp, C.c_double, _dp, C.c_int,
                                              C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]
        L.orc_vpm_relax_rule.argtypes = [i64, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int]
        L.orc_vpm_advect_rk3_ext.argtypes = [i64, _dp, _dp, _dp, _dp, C.c_double, STAGE_VEL_FN,
                                             C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                             C.c_int, C.c_int]
        L.orc_erf.argtypes = [C.c_double]
        L.orc_erf.restype = C.c_double
        L.orc_convert_rings.argtypes = [i64, _dp, _dp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp]
        L.orc_convert_rings.restype = i64
        if hasattr(L, "orc_hybrid_run"):
            L.orc_hybrid_run.restype = C.c_int
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(_dp)


KERNELS = {"gaussianerf": 0, "winckelmans": 1}
LEGS = {"ring": 0xF, "chordwise": 0x5, "horseshoe": 0x7}


def collapsed(points, br, fr, fl, bl, gamma, rc0, age=None, nu=0.0, legs="ring", threads=0):
    P, c = _f(points), [_f(x) for x in (br, fr, fl, bl)]
    g, r = _f(gamma), _f(rc0)
    a = None if age is None else _f(age)
    out = np.zeros((P.shape[0], 3))
    lib().orc_collapsed(_p(P), P.shape[0], *map(_p, c), g.shape[0], _p(g), _p(r),
                        None if a is None else _p(a), nu, LEGS[legs], _p(out), threads)
    return out


def expanded(points, br, fr, fl, bl, gamma, rc0, age=None, nu=0.0, legs="ring", threads=0):
    P, c = _f(points), [_f(x) for x in (br, fr, fl, bl)]
    g, r = _f(gamma), _f(rc0)
    a = None if age is None else _f(age)
    out = np.zeros((P.shape[0], g.shape[0], 3))
    lib().orc_expanded(_p(P), P.shape[0], *map(_p, c), g.shape[0], _p(g), _p(r),
                       None if a is None else _p(a), nu, LEGS[legs], _p(out), threads)
    return out


def aic(colloc, normals, br, fr, fl, bl, rc0, nu=0.0, threads=0):
    c, n = _f(colloc), _f(normals)
    v = [_f(x) for x in (br, fr, fl, bl)]
    out = np.zeros((c.shape[0], c.shape[0]))
    lib().orc_aic(_p(c), _p(n), c.shape[0], *map(_p, v), _p(_f(rc0)), nu, _p(out), threads)
    return out


def lu_solve(A, b):
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    x = np.array(b, dtype=np.float64, order="C", copy=True)
    info = lib().orc_lu_solve(A.shape[0], _p(A), _p(x))
    if info:
        raise np.linalg.LinAlgError(f"singular (info={info})")
    return x


def particle_vel(tgt, pos, alpha, sigma, kernel="gaussianerf", threads=0, rule=0):
    """rule 0: pairs closer than 1e-12 skipped (warp_vpm.py:37); rule 1: r
    clamped to 1e-12 (kernel.py:105-106, the reference CPU path)."""
    t, p, a, s = _f(tgt), _f(pos), _f(alpha), _f(sigma)
    out = np.zeros((t.shape[0], 3))
    lib().orc_particle_vel_rule(_p(t), t.shape[0], _p(p), _p(a), _p(s), p.shape[0],
                                KERNELS[kernel], int(rule), _p(out), threads)
    return out


def particle_jac(tgt, pos, alpha, sigma, kernel="gaussianerf", threads=0, rule=0):
    """rule as particle_vel (kernel.py:156: r = sqrt(max(r^2, 1e-24)))."""
    t, p, a, s = _f(tgt), _f(pos), _f(alpha), _f(sigma)
    out = np.zeros((t.shape[0], 3, 3))
    lib().orc_particle_jac_rule(_p(t), t.shape[0], _p(p), _p(a), _p(s), p.shape[0],
                                KERNELS[kernel], int(rule), _p(out), threads)
    return out


def oracle_erf(x):
    """The oracle's erf (Cephes form below 1, as scipy.special.erf)."""
    return lib().orc_erf(float(x))


def vpm_advect_rk3(pos, gamma, sigma, age, dt, uinf, stretch, free_wake, kernel, nu=0.0,
                   rlxf=0.3, threads=0, rule=0):
    """advect_rk3 followed by _relax_pedrizzetti (particles.py:85-246); returns new arrays."""
    pos, gamma = _f(pos).copy(), _f(gamma).copy()
    sigma, age = _f(sigma).copy(), _f(age).copy()
    u = _f(uinf)
    n = pos.shape[0]
    L = lib()
    rc = L.orc_vpm_advect_rk3_rule(n, _p(pos), _p(gamma), _p(sigma), _p(age), dt, _p(u),
                                   int(stretch), int(free_wake), KERNELS[kernel], nu, int(rule),
                                   threads)
    assert rc == 0
    rc = L.orc_vpm_relax_rule(n, _p(pos), _p(gamma), _p(sigma), rlxf, int(rule), threads)
    assert rc == 0
    return pos, gamma, sigma, age


# int (*)(void *user, int64_t n, const double *pos, double *u_bound, double *u_inf)
STAGE_VEL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_double),
                           C.POINTER(C.c_double), C.POINTER(C.c_double))


def vpm_advect_rk3_ext(pos, gamma, sigma, age, dt, u_inf_func, bound_velocity_func, stretch,
                       free_wake, kernel, nu=0.0, rlxf=0.3, threads=0, rule=0):
    """advect_rk3 with the reference's callables U_inf_func(X) and
    bound_velocity_func(X) (None: zeros) evaluated at every stage's positions
    (particles.py:126-133), then _relax_pedrizzetti; returns new arrays."""
    pos, gamma = _f(pos).copy(), _f(gamma).copy()
    sigma, age = _f(sigma).copy(), _f(age).copy()
    n = pos.shape[0]
    errors = []

    def stage(_user, m, p, ub, ui):
        try:
            X = np.ctypeslib.as_array(p, shape=(m, 3)).copy()
            np.ctypeslib.as_array(ub, shape=(m, 3))[...] = (
                0.0 if bound_velocity_func is None else bound_velocity_func(X))
            np.ctypeslib.as_array(ui, shape=(m, 3))[...] = u_inf_func(X)
            return 0
        except Exception as e:   # reported after the C call returns
            errors.append(e)
            return 1

    cb = STAGE_VEL_FN(stage)
    L = lib()
    rc = L.orc_vpm_advect_rk3_ext(n, _p(pos), _p(gamma), _p(sigma), _p(age), dt, cb, None,
                                  int(stretch), int(free_wake), KERNELS[kernel], nu, int(rule),
                                  threads)
    if errors:
        raise errors[0]
    assert rc == 0
    rc = L.orc_vpm_relax_rule(n, _p(pos), _p(gamma), _p(sigma), rlxf, int(rule), threads)
    assert rc == 0
    return pos, gamma, sigma, age


def convert_rings(br, fr, fl, bl, gamma, sigma):
    c = [_f(x) for x in (br, fr, fl, bl)]
    g = _f(gamma).copy()
    nr = g.shape[0]
    pos = np.zeros((4 * nr, 3))
    al = np.zeros((4 * nr, 3))
    sg = np.zeros(4 * nr)
    k = lib().orc_convert_rings(nr, *map(_p, c), _p(g), sigma, _p(pos), _p(al), _p(sg))
    return pos[:k], al[:k], sg[:k], g


def pop_drop(pos, alpha, sigma, age, max_particles):
    """Keeps the newest max_particles (flap.py:469-473); returns copies."""
    pos, alpha, sigma, age = (_f(x).copy() for x in (pos, alpha, sigma, age))
    fn = lib().orc_pop_drop
    fn.restype = C.c_int64
    fn.argtypes = [C.c_int64, C.c_int64, _dp, _dp, _dp, _dp]
    n = fn(len(sigma), int(max_particles), _p(pos), _p(alpha), _p(sigma), _p(age))
    return pos[:n], alpha[:n], sigma[:n], age[:n]


def pop_merge(pos, alpha, sigma, age, ref, eps, vinf_norm, protect=64, protect_dist=0.0):
    """Pairwise moment-conserving merge (flap.py:478-563); returns copies and
    the number of merged pairs."""
    pos, alpha, sigma, age = (_f(x).copy() for x in (pos, alpha, sigma, age))
    ref = _f(ref)
    fn = lib().orc_pop_merge
    fn.restype = C.c_int64
    fn.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, C.c_int64, C.c_double, C.c_double,
                   C.c_int64, C.c_double, C.POINTER(C.c_int64)]
    nm = C.c_int64()
    n = fn(len(sigma), _p(pos), _p(alpha), _p(sigma), _p(age), _p(ref), ref.shape[0], eps,
           vinf_norm, int(protect), float(protect_dist), C.byref(nm))
    if n < 0:
        raise MemoryError("orc_pop_merge")
    return pos[:n], alpha[:n], sigma[:n], age[:n], nm.value


def ledger(alpha, br=None, fr=None, fl=None, bl=None, gamma=None):
    """Total vector circulation: sum of particle alpha + sum over ring legs of
    G (b - a) (flap.py:565-578)."""
    alpha = _f(alpha).reshape(-1, 3)
    out = np.zeros(3)
    fn = lib().orc_ledger
    fn.argtypes = [C.c_int64, _dp, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]
    if gamma is None:
        z = np.zeros((0, 3))
        fn(alpha.shape[0], _p(alpha), 0, _p(z), _p(z), _p(z), _p(z), _p(np.zeros(0)), _p(out))
    else:
        c = [_f(x) for x in (br, fr, fl, bl)]
        g = _f(gamma)
        fn(alpha.shape[0], _p(alpha), g.shape[0], *map(_p, c), _p(g), _p(out))
    return out


# ---------------------------------------------------------------------------
# NumPy restatements (small inputs only)
# ---------------------------------------------------------------------------
def np_segment(points, s, e, gamma, rc0, age, nu):
    """(N, M, 3) per-pair velocity of segments s->e (warp_kernels.py:46-96 semantics)."""
    p = points[:, None, :]
    r0 = (e - s)[None, :, :]
    r1 = s[None, :, :] - p
    r2 = e[None, :, :] - p
    r3 = np.cross(r1, r2)
    r0l = np.linalg.norm(r0, axis=-1)
    r1l = np.linalg.norm(r1, axis=-1)
    r2l = np.linalg.norm(r2, axis=-1)
    r3sq = np.sum(r3 * r3, axis=-1)
    r1r2 = r1l * r2l
    tol = r0l * 1e-10
    skip = (r0l < 2.22e-16) | (r1l < tol) | (r2l < tol) | (np.sqrt(r3sq) < 1e-10 * r1r2)
    rcsq = rc0 ** 2 + 5.02572 * (nu + 1e-4 * np.abs(gamma)) * age
    c1 = gamma * 0.07957747154594767
    c2 = np.sum(r0 * r0, axis=-1) * rcsq[None, :]
    c3 = np.sum(r1 * r2, axis=-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        c4 = c1[None, :] * (r1l + r2l) * (r1r2 - c3) / (r1r2 * (r3sq + c2))
    c4 = np.where(skip, 0.0, c4)
    return c4[..., None] * r3


def np_rings(points, br, fr, fl, bl, gamma, rc0, age=None, nu=0.0, legs="ring"):
    """Collapsed sum through np_segment."""
    age = np.zeros_like(gamma) if age is None else age
    c = [br, fr, fl, bl]
    mask = LEGS[legs]
    out = np.zeros((points.shape[0], 3))
    for k in range(4):
        if mask & (1 << k):
            out += np_segment(points, c[k], c[(k + 1) % 4], gamma, rc0, age, nu).sum(axis=1)
    return out


def np_g_dgdr(kernel, rb):
    if kernel == "winckelmans":
        r2 = rb * rb
        return rb ** 3 * (r2 + 2.5) / (r2 + 1.0) ** 2.5, 7.5 * r2 / (r2 + 1.0) ** 3.5
    c = sqrt(2.0 / pi)
    ev = np.exp(-0.5 * rb * rb)
    erfv = np.vectorize(erf)(rb / sqrt(2.0))
    return erfv - c * rb * ev, c * rb * rb * ev


def np_particle_vel(tgt, pos, alpha, sigma, kernel):
    dx = tgt[:, None, :] - pos[None, :, :]
    r = np.sqrt(np.sum(dx * dx, axis=-1))
    keep = r >= 1e-12
    rs = np.where(keep, r, 1.0)
    g, _ = np_g_dgdr(kernel, rs / sigma[None, :])
    coeff = np.where(keep, -g / (4 * pi * rs ** 3), 0.0)
    return np.sum(coeff[..., None] * np.cross(dx, alpha[None, :, :]), axis=1)


def np_particle_jac(tgt, pos, alpha, sigma, kernel):
    dx = tgt[:, None, :] - pos[None, :, :]
    r = np.sqrt(np.sum(dx * dx, axis=-1))
    keep = r >= 1e-12
    rs = np.where(keep, r, 1.0)
    g, dg = np_g_dgdr(kernel, rs / sigma[None, :])
    C_ = -1.0 / (4 * pi)
    kxg = C_ * np.cross(dx, alpha[None, :, :]) / rs[..., None] ** 3
    aux = np.where(keep, dg / (sigma[None, :] * rs) - 3 * g / rs ** 2, 0.0)
    aux2 = np.where(keep, C_ * g / rs ** 3, 0.0)
    J = np.einsum("tn,tni,tnj->tij", aux, kxg, dx)
    a = alpha[None, :, :]
    A = np.zeros(aux2.shape + (3, 3))
    A[..., 0, 1] = a[..., 2]
    A[..., 0, 2] = -a[..., 1]
    A[..., 1, 0] = -a[..., 2]
    A[..., 1, 2] = a[..., 0]
    A[..., 2, 0] = a[..., 1]
    A[..., 2, 1] = -a[..., 0]
    return J + np.einsum("tn,tnij->tij", aux2, A)


def seeded_rings(N=200, M=100, seed=42):
    """The reference's kernel-test input recipe (tests/test_correctness.py:19-31)."""
    rng = np.random.default_rng(seed)
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
    return points, br, fr, fl, bl, strengths, rc0s, sing, ages, nu


# ---------------------------------------------------------------------------
# Hybrid panel-particle step (fv_oracle_uvlm.c)
# ---------------------------------------------------------------------------
class _HybridDesc(C.Structure):
    _fields_ = [
        ("n_surfaces", C.c_int32), ("nc", C.POINTER(C.c_int32)), ("ns", C.POINTER(C.c_int32)),
        ("num_steps", C.c_int64), ("grids", _dp), ("vinf", C.c_double * 3),
        ("rho", C.c_double), ("dt", C.c_double), ("nu", C.c_double),
        ("rc0_bound", C.c_double), ("rc0_wake", C.c_double), ("n_keep", C.c_int32),
        ("sigma", C.c_double), ("kernel", C.c_int32), ("free_wake", C.c_int32),
        ("stretch", C.c_int32), ("rlxf", C.c_double), ("vpm_nu", C.c_double),
        ("sref", C.c_double), ("max_particles", C.c_int64), ("precision", C.c_int32),
        ("beam", C.c_void_p), ("convect_exact", C.c_int32),
        ("free_ring_wake", C.c_int32), ("truncate_x", C.c_double),
        ("pair_rule", C.c_int32),
    ]


class _OrcBeam(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("y_nodes", _dp), ("M", _dp), ("C", _dp), ("Keff_inv", _dp),
        ("d", _dp), ("v", _dp), ("a", _dp), ("beta", C.c_double), ("gamma", C.c_double),
        ("x_ea", C.c_double), ("relax", C.c_double), ("tip_w", _dp), ("tip_theta", _dp),
        ("feedback", C.c_int32), ("lift_sign", C.c_double), ("station_lift", _dp),
        ("station_moment", _dp),
    ]


def hybrid_run(case, steps=None, threads=0):
    """Runs a hybrid case (any object with the HybridCase fields) on the CPU
    oracle; returns cl, cdi, gamma, np and the particle state."""
    nc = np.ascontiguousarray(case.nc, dtype=np.int32)
    ns = np.ascontiguousarray(case.ns, dtype=np.int32)
    grids = _f(case.grids)
    d = _HybridDesc()
    d.n_surfaces = len(nc)
    d.nc = nc.ctypes.data_as(C.POINTER(C.c_int32))
    d.ns = ns.ctypes.data_as(C.POINTER(C.c_int32))
    d.num_steps = grids.shape[0]
    d.grids = _p(grids)
    for k in range(3):
        d.vinf[k] = float(case.vinf[k])
    d.rho, d.dt, d.nu = case.rho, case.dt, case.nu
    d.rc0_bound, d.rc0_wake = case.rc0_bound, case.rc0_wake
    d.n_keep = int(case.n_keep)
    d.sigma = case.sigma
    d.kernel = KERNELS[case.kernel]
    d.free_wake, d.stretch = int(bool(case.free_wake)), int(bool(case.stretch))
    d.rlxf, d.vpm_nu, d.sref = case.rlxf, case.vpm_nu, case.sref
    d.max_particles = int(case.max_particles)
    d.free_ring_wake = int(bool(getattr(case, "ring_free_wake", False)))
    d.truncate_x = float(getattr(case, "truncate_x", np.inf))
    rule = getattr(case, "pair_rule", 0)
    d.pair_rule = {"skip": 0, "clamp": 1}.get(rule, rule) if isinstance(rule, str) else int(rule)
    n = grids.shape[0]
    steps = n if steps is None else steps
    P = int(sum(int(a) * int(b) for a, b in zip(nc, ns)))
    cl, cdi, gam = np.zeros(n), np.zeros(n), np.zeros(P)
    npo = C.c_int64()
    cap = int(case.max_particles)
    pos, al, sg = np.zeros((cap, 3)), np.zeros((cap, 3)), np.zeros(cap)
    L = lib()
    fn = L.orc_hybrid_run
    fn.argtypes = [C.POINTER(_HybridDesc), C.c_int64, _dp, _dp, _dp, C.POINTER(C.c_int64),
                   _dp, _dp, _dp, C.c_int]
    rc = fn(C.byref(d), steps, _p(cl), _p(cdi), _p(gam), C.byref(npo), _p(pos), _p(al), _p(sg),
            threads)
    if rc == -2:
        raise RuntimeError("Particle field would overflow")
    if rc:
        raise RuntimeError(f"orc_hybrid_run failed ({rc})")
    k = npo.value
    return {"cl": cl, "cdi": cdi, "gamma": gam, "np": k, "pos": pos[:k].copy(),
            "gamma_p": al[:k].copy(), "sigma": sg[:k].copy()}


def aeroelastic_run(inst, threads=0):
    """One aeroelastic instance (fluxvortex_amd.aeroelastic.AeroInstance) on the
    CPU oracle: the hybrid step on the deformed lattice, the beam coupling and
    Newmark step, particles as a prescribed wake; returns cl, cdi, tip_w,
    tip_theta, gamma and np (oracle/fv_oracle_uvlm.c)."""
    from fluxvortex_amd.aeroelastic import beam_arrays
    nc = np.array([inst.nc], dtype=np.int32)
    ns = np.array([inst.ns], dtype=np.int32)
    grid = _f(inst.grid)
    if inst.beam is None:
        # without a beam the oracle reads one lattice per step
        grid = np.ascontiguousarray(np.broadcast_to(grid, (int(inst.num_steps),) + grid.shape))
    d = _HybridDesc()
    d.n_surfaces = 1
    d.nc = nc.ctypes.data_as(C.POINTER(C.c_int32))
    d.ns = ns.ctypes.data_as(C.POINTER(C.c_int32))
    d.num_steps = int(inst.num_steps)
    d.grids = _p(grid)
    for k in range(3):
        d.vinf[k] = float(inst.vinf[k])
    d.rho, d.dt, d.nu = inst.rho, inst.dt, inst.nu
    d.rc0_bound, d.rc0_wake = inst.rc0_bound, inst.rc0_wake
    d.n_keep, d.sigma = int(inst.n_keep), inst.sigma
    d.kernel = KERNELS[inst.kernel]
    d.free_wake, d.stretch, d.rlxf, d.vpm_nu = 0, 0, 0.0, 0.0
    d.sref, d.max_particles = inst.sref, int(inst.max_particles)
    d.convect_exact = 1
    d.free_ring_wake = int(bool(getattr(inst, "ring_free_wake", False)))
    d.truncate_x = float(getattr(inst, "truncate_x", np.inf))
    n = int(inst.num_steps)
    tw, tt = np.zeros(n), np.zeros(n)
    sl, sm = np.zeros((n, inst.ns)), np.zeros((n, inst.ns))
    b = None
    if inst.beam is not None:
        arr = beam_arrays(inst)
        st = {k: arr[k].copy() for k in ("d0", "v0", "a0")}
        b = _OrcBeam()
        b.n_nodes = len(arr["y"])
        b.y_nodes, b.M, b.C, b.Keff_inv = _p(arr["y"]), _p(arr["M"]), _p(arr["C"]), _p(arr["Kinv"])
        b.d, b.v, b.a = _p(st["d0"]), _p(st["v0"]), _p(st["a0"])
        b.beta, b.gamma = inst.beam.beta, inst.beam.gamma
        b.x_ea, b.relax = inst.x_ea, inst.relax
        b.tip_w, b.tip_theta = _p(tw), _p(tt)
        b.feedback = 1 if inst.feedback else 0
        b.lift_sign = float(inst.lift_sign)
        b.station_lift, b.station_moment = _p(sl), _p(sm)
        d.beam = C.cast(C.pointer(b), C.c_void_p)
    else:
        d.beam = None
    P = inst.nc * inst.ns
    cl, cdi, gam = np.zeros(n), np.zeros(n), np.zeros(P)
    npo = C.c_int64()
    cap = int(inst.max_particles)
    pos, al, sg = np.zeros((cap, 3)), np.zeros((cap, 3)), np.zeros(cap)
    fn = lib().orc_hybrid_run
    fn.argtypes = [C.POINTER(_HybridDesc), C.c_int64, _dp, _dp, _dp, C.POINTER(C.c_int64),
                   _dp, _dp, _dp, C.c_int]
    rc = fn(C.byref(d), n, _p(cl), _p(cdi), _p(gam), C.byref(npo), _p(pos), _p(al), _p(sg),
            threads)
    if rc:
        raise RuntimeError(f"orc_hybrid_run (aeroelastic) failed ({rc})")
    if b is None:
        for x in (tw, tt, sl, sm):
            x[...] = np.nan
    return {"cl": cl, "cdi": cdi, "tip_w": tw, "tip_theta": tt, "gamma": gam, "np": npo.value,
            "station_lift": sl, "station_moment": sm}


# ---------------------------------------------------------------------------
# Closure MLP (32 -> 128 -> 128 -> 128 -> 16, SiLU) in FP64, optionally with
# the device's bf16 rounding points emulated (inputs, weights, activations,
# layer errors); the parameter layout is the device's FP32 master layout.
# ---------------------------------------------------------------------------
MLP_IN, MLP_OUT = 32, 16
MLP_SHAPES = [("W1", (128, MLP_IN)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)),
              ("W3", (128, 128)), ("b3", (128,)), ("W4", (MLP_OUT, 128)), ("b4", (MLP_OUT,))]
MLP_NPARAMS = sum(int(np.prod(s)) for _, s in MLP_SHAPES)


def mlp_unpack(flat):
    out, o = {}, 0
    for name, shp in MLP_SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(flat[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def mlp_pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in MLP_SHAPES])


def bf16(x):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def mlp_forward(flat, X, emulate_bf16=True):
    p = mlp_unpack(flat)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    h = r(X)
    cache = {"h0": h}
    for k in (1, 2, 3):
        z = h @ r(p[f"W{k}"]).T + p[f"b{k}"]
        h = r(_silu(z))
        cache[f"z{k}"], cache[f"h{k}"] = z, h
    y = h @ r(p["W4"]).T + p["b4"]
    return y, cache


def mlp_gradients(flat, X, T, wout, emulate_bf16=True):
    """Gradient of L = (1/nb) sum_s sum_j w_j (Y_sj - T_sj)^2 (device layout)."""
    p = mlp_unpack(flat)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, c = mlp_forward(flat, X, emulate_bf16)
    nb = X.shape[0]
    dy = r(2.0 / nb * np.asarray(wout)[None, :] * (y - T))
    g = {"W4": dy.T @ c["h3"], "b4": dy.sum(axis=0)}
    e = dy
    for k in (3, 2, 1):
        wnext = r(p[f"W{k + 1}"])
        dz = r((e @ wnext) * _dsilu(c[f"z{k}"]))
        g[f"W{k}"] = dz.T @ c[f"h{k - 1}"]
        g[f"b{k}"] = dz.sum(axis=0)
        e = dz
    return mlp_pack(g)


def mlp_loss(flat, X, T, wout, emulate_bf16=False):
    y, _ = mlp_forward(flat, X, emulate_bf16)
    return float(np.sum(np.asarray(wout)[None, :] * (y - T) ** 2) / X.shape[0])


def adam_step(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, bias-corrected); returns new arrays."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
