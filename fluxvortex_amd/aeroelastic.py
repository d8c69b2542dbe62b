"""Aeroelastic hybrid UVLM: the cantilever beam, the Goland-wing flutter case and
the batched flutter-speed sweep on the MI355X (fv_sweep_*).

Reference: src/fluxvortex/beam_fe.py (BeamFE), src/fluxvortex/aeroelastic_solver.py
(AeroelasticSolver) and tests/benchmark_goland.py (build_goland_wing,
compute_envelope_growth, run_at_velocity, run_flutter_sweep).  The beam
matrices are host-side set-up, as in the reference; every step of every
speed runs on the device.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .geometry import freestream

# ---------------------------------------------------------------------------
# Beam: Euler-Bernoulli bending (Hermite cubic) + linear torsion, 3 DOF per
# node (w, dw/dy, theta), bending-torsion mass coupling through the CG offset,
# clamped root, Rayleigh damping C = 2 zeta / omega_1 K, Newmark average
# acceleration (beam_fe.py:16-351).
# ---------------------------------------------------------------------------


def _bending_k(EI, L):
    """Hermite cubic bending stiffness on (w1, w1', w2, w2')."""
    return EI / L ** 3 * np.array([[12.0, 6 * L, -12.0, 6 * L],
                                    [6 * L, 4 * L * L, -6 * L, 2 * L * L],
                                    [-12.0, -6 * L, 12.0, -6 * L],
                                    [6 * L, 2 * L * L, -6 * L, 4 * L * L]])


def _bending_m(m, L):
    """Consistent bending mass on (w1, w1', w2, w2')."""
    return m * L / 420.0 * np.array([[156.0, 22 * L, 54.0, -13 * L],
                                      [22 * L, 4 * L * L, 13 * L, -3 * L * L],
                                      [54.0, 13 * L, 156.0, -22 * L],
                                      [-13 * L, -3 * L * L, -22 * L, 4 * L * L]])


def _hermite(xi, L):
    return np.array([1 - 3 * xi ** 2 + 2 * xi ** 3, L * (xi - 2 * xi ** 2 + xi ** 3),
                     3 * xi ** 2 - 2 * xi ** 3, L * (-xi ** 2 + xi ** 3)])


def _coupling_m(m, xa, L):
    """m x_alpha int N_bend_i N_tors_j dy by 3-point Gauss (exact: degree 4)."""
    g = np.array([0.5 - 0.5 * math.sqrt(0.6), 0.5, 0.5 + 0.5 * math.sqrt(0.6)])
    wq = np.array([5.0, 8.0, 5.0]) / 18.0
    Nb = _hermite(g, L)                  # (4, 3)
    Nt = np.array([1 - g, g])            # (2, 3)
    out = np.zeros((6, 6))
    bidx, tidx = (0, 1, 3, 4), (2, 5)
    for a, bi in enumerate(bidx):
        for b, tj in enumerate(tidx):
            val = m * xa * np.sum(wq * Nb[a] * Nt[b]) * L
            out[bi, tj] = val
            out[tj, bi] = val
    return out


@dataclass
class Beam:
    """Cantilevered bending-torsion beam along +y from the root (y = 0)."""
    length: float
    n_elements: int
    EI: float
    GJ: float
    m: float
    Ip: float
    x_ea_cg: float = 0.0
    zeta: float = 0.0
    beta: float = 0.25
    gamma: float = 0.5
    K: np.ndarray = field(init=False, repr=False)
    M: np.ndarray = field(init=False, repr=False)
    C: np.ndarray = field(init=False, repr=False)

    def __post_init__(self):
        ne = self.n_elements
        nd = 3 * (ne + 1)
        self.K = np.zeros((nd, nd))
        self.M = np.zeros((nd, nd))
        L = self.length / ne
        for e in range(ne):
            Ke = np.zeros((6, 6))
            Me = np.zeros((6, 6))
            b = (0, 1, 3, 4)
            t = (2, 5)
            Kb, Mb = _bending_k(self.EI, L), _bending_m(self.m, L)
            for i in range(4):
                for j in range(4):
                    Ke[b[i], b[j]] += Kb[i, j]
                    Me[b[i], b[j]] += Mb[i, j]
            Kt = self.GJ / L * np.array([[1.0, -1.0], [-1.0, 1.0]])
            Mt = self.Ip * L / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]])
            for i in range(2):
                for j in range(2):
                    Ke[t[i], t[j]] += Kt[i, j]
                    Me[t[i], t[j]] += Mt[i, j]
            Me += _coupling_m(self.m, self.x_ea_cg, L)
            dofs = [3 * e + k for k in range(6)]
            for i in range(6):
                for j in range(6):
                    self.K[dofs[i], dofs[j]] += Ke[i, j]
                    self.M[dofs[i], dofs[j]] += Me[i, j]
        self.C = np.zeros_like(self.K)
        if self.zeta > 0:
            w2 = self.eigenvalues()
            omega1 = math.sqrt(max(w2[0], 1e-6))
            self.C = 2.0 * self.zeta / omega1 * self.K

    @property
    def y_nodes(self) -> np.ndarray:
        return np.linspace(0.0, self.length, self.n_elements + 1)

    @property
    def free(self) -> np.ndarray:
        """Free DOFs: every DOF except the clamped root node's three."""
        return np.arange(3, 3 * (self.n_elements + 1))

    def reduced(self):
        f = self.free
        return (self.K[np.ix_(f, f)], self.M[np.ix_(f, f)], self.C[np.ix_(f, f)])

    def eigenvalues(self) -> np.ndarray:
        """omega^2 of K x = omega^2 M x on the free DOFs, ascending."""
        from scipy.linalg import eigh
        K, M, _ = self.reduced()
        return eigh(K, M, eigvals_only=True)

    def natural_frequencies(self, n=None) -> np.ndarray:
        w2 = np.maximum(self.eigenvalues(), 0.0)
        f = np.sqrt(w2) / (2 * math.pi)
        return f if n is None else f[:n]

    def newmark_matrices(self, dt):
        """(M, C, K_eff^-1) on the free DOFs for the step size dt."""
        K, M, Cm = self.reduced()
        Keff = K + self.gamma / (self.beta * dt) * Cm + 1.0 / (self.beta * dt * dt) * M
        return M, Cm, np.linalg.inv(Keff)

    def initial_state(self, tip_w=0.0, tip_theta=0.0):
        """(d, v, a) on the free DOFs for a tip displacement at rest, with the
        consistent acceleration a = M^-1 (-K d) (benchmark_goland.py:109-119)."""
        nd = 3 * (self.n_elements + 1)
        d = np.zeros(nd)
        tip = self.n_elements
        d[3 * tip] = tip_w
        d[3 * tip + 2] = tip_theta
        K, M, _ = self.reduced()
        df = d[self.free]
        af = np.linalg.solve(M, -K @ df)
        return df, np.zeros_like(df), af


def goland_beam(semi_span=6.096, chord=1.8288, n_elements=8) -> Beam:
    """The Goland wing's beam (benchmark_goland.py:142-152)."""
    m = 35.72
    return Beam(length=semi_span, n_elements=n_elements, EI=9.773e6, GJ=0.988e6, m=m,
                Ip=m * chord ** 2 / 24.0, x_ea_cg=0.10 * chord, zeta=0.005)


# ---------------------------------------------------------------------------
# Envelope growth and flutter speed (benchmark_goland.py:66-97, 224-231)
# ---------------------------------------------------------------------------


def envelope_growth(signal, dt) -> float:
    """Exponential growth rate (1/s) of a signal's envelope: a least-squares
    line through log|peak| against time, the first peak skipped when more than
    four are found; 0 for fewer than 10 samples or 3 peaks."""
    s = np.asarray(signal, dtype=np.float64)
    if s.size < 10:
        return 0.0
    a = np.abs(s)
    idx = np.nonzero((a[1:-1] > a[:-2]) & (a[1:-1] > a[2:]))[0] + 1
    if idx.size < 3:
        return 0.0
    tp = idx * dt
    ap = np.maximum(a[idx], 1e-15)
    if tp.size > 4:
        tp, ap = tp[1:], ap[1:]
    if tp.size < 2:
        return 0.0
    return float(np.polyfit(tp, np.log(ap), 1)[0])


def flutter_speed(speeds, sigma) -> float | None:
    """The first speed where sigma_w turns from negative to positive, by
    linear interpolation between the bracketing speeds (None if none)."""
    V = np.asarray(speeds, dtype=np.float64)
    s = np.asarray(sigma, dtype=np.float64)
    for i in range(len(V) - 1):
        if s[i] < 0 and s[i + 1] > 0:
            return float(V[i] - s[i] * (V[i + 1] - V[i]) / (s[i + 1] - s[i]))
    return None


# ---------------------------------------------------------------------------
# One aeroelastic instance and the Goland case
# ---------------------------------------------------------------------------


@dataclass
class AeroInstance:
    grid: np.ndarray          # (nc+1, ns+1, 3) undeformed lattice
    nc: int
    ns: int
    num_steps: int
    vinf: np.ndarray
    rho: float
    dt: float
    nu: float
    rc0_bound: float
    rc0_wake: float
    n_keep: int
    sigma: float
    kernel: str
    sref: float
    max_particles: int
    beam: Beam | None
    x_ea: float
    relax: float = 1.0
    # False: the aerodynamics stay on the undeformed lattice and the station
    # lift is -F_z, which is what the reference computes (its vertex mutation
    # does not reach the solver; tests/golden/goland_flutter_ref.npz holds its
    # stored histories); True with lift_sign +1: the deformed lattice carries
    # the next step's aerodynamics with the lift along +z.
    feedback: bool = False
    lift_sign: float = -1.0
    tip_w: float = 0.0
    tip_theta: float = 0.0
    precision: str = "fp64"
    # True: the wake's vertex rows up to n_keep move with V + the induced
    # velocity (PteraSoftware prescribed_wake=False); False: with V (the
    # reference benchmark's prescribed wake)
    ring_free_wake: bool = False
    # rings whose corner-mean x exceeds truncate_x lose their circulation after
    # shedding (PteraSoftware's wake truncation); inf: none
    truncate_x: float = math.inf

    @property
    def speed(self) -> float:
        return float(np.linalg.norm(self.vinf))


def goland_instance(V, dt=0.003, num_chords=100.0, alpha_deg=2.0, nc=4, ns=8, n_keep=10,
                    kernel="gaussianerf", tip_w=0.05, tip_theta_deg=2.0, x_ea_chord=0.33,
                    couple=True, num_steps=None, precision="fp64", feedback=False,
                    lift_sign=None, ring_free_wake=False, truncate_x=math.inf) -> AeroInstance:
    """The Goland wing at airspeed V (benchmark_goland.py:21-63, 100-119):
    rectangular NACA 0012 wing (zero camber), chord 1.8288 m, semi-span
    6.096 m, nc x ns uniform panels from the root (y = 0) outwards, alpha 2 deg,
    dt 0.003 s, num_chords chord lengths of wake (num_steps = ceil(num_chords c
    / (V dt))), tip perturbation w = 0.05 m, theta = 2 deg, elastic axis at 33 %
    chord; the hybrid wake keeps n_keep ring rows.  feedback / lift_sign /
    ring_free_wake / truncate_x as in AeroInstance (lift_sign defaults to -1 without
    feedback, +1 with it)."""
    chord, span = 1.8288, 6.096
    xs = np.linspace(0.0, chord, nc + 1)
    ys = np.linspace(0.0, span, ns + 1)
    grid = np.zeros((nc + 1, ns + 1, 3))
    grid[:, :, 0] = xs[:, None]
    grid[:, :, 1] = ys[None, :]
    steps = int(math.ceil(num_chords * chord / (V * dt))) if num_steps is None else int(num_steps)
    return AeroInstance(grid=grid, nc=nc, ns=ns, num_steps=steps, vinf=freestream(V, alpha_deg),
                        rho=1.225, dt=dt, nu=15.06e-6, rc0_bound=0.03 * chord,
                        rc0_wake=0.03 * chord, n_keep=n_keep, sigma=V * dt, kernel=kernel,
                        sref=chord * span, max_particles=4 * ns * (steps + 1) + 64,
                        beam=goland_beam(span, chord) if couple else None,
                        x_ea=x_ea_chord * chord, tip_w=tip_w,
                        tip_theta=math.radians(tip_theta_deg), precision=precision,
                        feedback=bool(feedback),
                        lift_sign=(1.0 if feedback else -1.0) if lift_sign is None else lift_sign,
                        ring_free_wake=bool(ring_free_wake), truncate_x=float(truncate_x))


# ---------------------------------------------------------------------------
# The batched sweep on the device
# ---------------------------------------------------------------------------


class SweepDesc(C.Structure):
    _fields_ = [
        ("nc", C.c_int32), ("ns", C.c_int32), ("grid", C.POINTER(C.c_double)),
        ("num_steps", C.c_int64), ("vinf", C.c_double * 3), ("rho", C.c_double),
        ("dt", C.c_double), ("nu", C.c_double), ("rc0_bound", C.c_double),
        ("rc0_wake", C.c_double), ("n_keep", C.c_int32), ("sigma", C.c_double),
        ("kernel", C.c_int32), ("sref", C.c_double), ("max_particles", C.c_int64),
        ("couple", C.c_int32), ("n_nodes", C.c_int32), ("y_nodes", C.POINTER(C.c_double)),
        ("M", C.POINTER(C.c_double)), ("C", C.POINTER(C.c_double)),
        ("Keff_inv", C.POINTER(C.c_double)), ("d0", C.POINTER(C.c_double)),
        ("v0", C.POINTER(C.c_double)), ("a0", C.POINTER(C.c_double)), ("beta", C.c_double),
        ("gamma", C.c_double), ("x_ea", C.c_double), ("relax", C.c_double),
        ("precision", C.c_int32), ("feedback", C.c_int32), ("lift_sign", C.c_double),
        ("free_ring_wake", C.c_int32), ("truncate_wake", C.c_int32), ("truncate_x", C.c_double),
    ]


def beam_arrays(inst: AeroInstance):
    """Host arrays of an instance's beam (kept alive by the caller)."""
    b = inst.beam
    M, Cm, Kinv = b.newmark_matrices(inst.dt)
    d0, v0, a0 = b.initial_state(inst.tip_w, inst.tip_theta)
    return {"y": np.ascontiguousarray(b.y_nodes), "M": np.ascontiguousarray(M),
            "C": np.ascontiguousarray(Cm), "Kinv": np.ascontiguousarray(Kinv),
            "d0": np.ascontiguousarray(d0), "v0": np.ascontiguousarray(v0),
            "a0": np.ascontiguousarray(a0)}


class Sweep:
    """Independent aeroelastic instances stepped together on the GPU: one
    batched particle-induction launch and one workgroup-per-instance step
    launch per time step, replayed from a hipGraph."""

    def __init__(self, instances, use_graph=True, context=None):
        self._ctx = context or _lib.default_context()
        self.instances = list(instances)
        n = len(self.instances)
        descs = (SweepDesc * n)()
        self._keep = []
        for k, it in enumerate(self.instances):
            d = descs[k]
            g = np.ascontiguousarray(it.grid, dtype=np.float64)
            self._keep.append(g)
            d.nc, d.ns = it.nc, it.ns
            d.grid = _lib.ptr(g)
            d.num_steps = it.num_steps
            for c in range(3):
                d.vinf[c] = float(it.vinf[c])
            d.rho, d.dt, d.nu = it.rho, it.dt, it.nu
            d.rc0_bound, d.rc0_wake = it.rc0_bound, it.rc0_wake
            d.n_keep, d.sigma = it.n_keep, it.sigma
            d.kernel = _lib.KERNELS[it.kernel]
            d.sref, d.max_particles = it.sref, it.max_particles
            d.x_ea, d.relax = it.x_ea, it.relax
            d.precision = _lib.PRECISIONS[it.precision]
            d.feedback = 1 if it.feedback else 0
            d.lift_sign = float(it.lift_sign)
            d.free_ring_wake = 1 if it.ring_free_wake else 0
            d.truncate_wake = 1 if math.isfinite(it.truncate_x) else 0
            d.truncate_x = float(it.truncate_x) if math.isfinite(it.truncate_x) else 0.0
            if it.beam is not None:
                arr = beam_arrays(it)
                self._keep.append(arr)
                d.couple = 1
                d.n_nodes = len(arr["y"])
                d.y_nodes = _lib.ptr(arr["y"])
                d.M, d.C, d.Keff_inv = _lib.ptr(arr["M"]), _lib.ptr(arr["C"]), _lib.ptr(arr["Kinv"])
                d.d0, d.v0, d.a0 = _lib.ptr(arr["d0"]), _lib.ptr(arr["v0"]), _lib.ptr(arr["a0"])
                d.beta, d.gamma = it.beam.beta, it.beam.gamma
            else:
                y = np.array([0.0, 1.0])
                self._keep.append(y)
                d.couple, d.n_nodes, d.y_nodes = 0, 2, _lib.ptr(y)
                d.beta, d.gamma = 0.25, 0.5
        h = C.c_void_p()
        self._ctx.check(self._ctx.lib.fv_sweep_create(self._ctx.handle, descs, n,
                                                     1 if use_graph else 0, C.byref(h)))
        self._h = h
        self.max_steps = max(it.num_steps for it in self.instances)
        self.steps_done = 0

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._ctx.lib.fv_sweep_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, steps=None):
        k = self.max_steps - self.steps_done if steps is None else int(steps)
        self._ctx.check(self._ctx.lib.fv_sweep_step(self._h, k))
        self.steps_done = min(self.max_steps, self.steps_done + k)

    def read(self, i):
        it = self.instances[i]
        n = it.num_steps
        cl, cdi, tw, tt = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        gam = np.zeros(it.nc * it.ns)
        npo, st = C.c_int64(), C.c_int32()
        self._ctx.check(self._ctx.lib.fv_sweep_read(self._h, i, _lib.ptr(cl), _lib.ptr(cdi),
                                                    _lib.ptr(tw), _lib.ptr(tt), _lib.ptr(gam),
                                                    C.byref(npo), C.byref(st)))
        return {"cl": cl, "cdi": cdi, "tip_w": tw, "tip_theta": tt, "gamma": gam,
                "np": npo.value, "status": st.value}


def flutter_sweep(speeds, use_graph=True, **kw):
    """Runs the Goland case at every speed as one batched sweep; returns the
    per-speed histories, sigma_w / sigma_theta and the flutter speed."""
    insts = [goland_instance(float(V), **kw) for V in speeds]
    sw = Sweep(insts, use_graph=use_graph)
    sw.step()
    out = []
    for i, it in enumerate(insts):
        r = sw.read(i)
        if r["status"] != 0:
            raise RuntimeError(f"sweep instance {i} (V = {it.speed}) failed: status {r['status']}")
        w = r["tip_w"][~np.isnan(r["tip_w"])]
        th = r["tip_theta"][~np.isnan(r["tip_theta"])]
        r["sigma_w"] = envelope_growth(w, it.dt)
        r["sigma_theta"] = envelope_growth(th, it.dt)
        r["V"] = it.speed
        out.append(r)
    sw.close()
    Vf = flutter_speed([r["V"] for r in out], [r["sigma_w"] for r in out])
    return {"runs": out, "flutter_speed": Vf}


def main(argv=None):
    """The Goland flutter sweep from the command line (benchmark_goland.py's
    run): prints sigma_w per speed, the flutter speed and the sweep time."""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--speeds", type=int, default=64)
    ap.add_argument("--vmin", type=float, default=80.0)
    ap.add_argument("--vmax", type=float, default=180.0)
    ap.add_argument("--num-chords", type=float, default=100.0)
    ap.add_argument("--n-keep", type=int, default=10)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--feedback", action="store_true")
    ap.add_argument("--ring-free-wake", action="store_true",
                    help="wake vertex rows move with the induced velocity")
    ap.add_argument("--truncate-x", type=float, default=math.inf)
    ap.add_argument("--eager", action="store_true", help="launch without the hipGraph")
    a = ap.parse_args(argv)
    speeds = np.linspace(a.vmin, a.vmax, a.speeds)
    t0 = time.perf_counter()
    res = flutter_sweep(speeds, use_graph=not a.eager, num_chords=a.num_chords, n_keep=a.n_keep,
                        precision=a.precision, feedback=a.feedback,
                        ring_free_wake=a.ring_free_wake, truncate_x=a.truncate_x)
    sec = time.perf_counter() - t0
    print(json.dumps({"flutter_speed": res["flutter_speed"], "seconds": sec,
                      "sigma_w": {f"{r['V']:.3f}": round(r["sigma_w"], 6) for r in res["runs"]}}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
