"""Hybrid panel-particle UVLM driver on the MI355X.

The reference's HybridSolver (tests/experiment_hybrid_panel_particle.py:87-204)
subclasses PteraSoftware's unsteady ring-vortex solver: the newest ``n_keep``
wake rows stay ring panels, older rows become 4 vortex particles per ring,
the particles' induced velocity joins the wake-wing influence, and the
particle field is advected by RK3 after every wake update.  Here the whole
step -- geometry, AIC, right-hand side, dense solve, loads, wake shedding,
conversion, particle RK3 and relaxation -- runs on the device
(libfluxvortex_hip.so, fv_hybrid_*), with only the CL/CDi history read back.

Cases are built from :mod:`fluxvortex_amd.geometry` lattices; the named
builders reproduce the reference's benchmark configurations.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import _lib
from .geometry import Airplane, Motion, Surface, freestream
from .particles import VortexParticleField


@dataclass
class HybridCase:
    """Everything one hybrid run needs (all FP64, SI units)."""
    nc: list
    ns: list
    grids: np.ndarray            # (num_steps, total_vertices, 3)
    vinf: np.ndarray             # (3,)
    dt: float
    rho: float = 1.225
    nu: float = 0.0              # viscosity of the ring core growth law
    rc0_bound: float = 0.03
    rc0_wake: float = 0.03
    n_keep: int = 10
    sigma: float = 0.1
    kernel: str = "gaussianerf"
    free_wake: bool = False
    stretch: bool = False
    rlxf: float = 0.3
    vpm_nu: float = 0.0
    sref: float = 1.0
    max_particles: int = 200000
    precision: str = "fp64"
    ring_free_wake: bool = False  # ring wake moves with V + induction (prescribed: V)
    truncate_x: float = math.inf  # rings whose corner-mean x exceeds this lose gamma
    # particles closer than 1e-12 in the relaxation: "clamp" (r = 1e-12), as the
    # reference's CPU field that HybridSolver runs (kernel.py:156), or "skip"
    # (warp_vpm.py:103); the oracle reads the same field (0 skip, 1 clamp)
    pair_rule: str = "clamp"
    name: str = "case"
    meta: dict = field(default_factory=dict)

    @property
    def num_steps(self) -> int:
        return int(self.grids.shape[0])

    @property
    def panels(self) -> int:
        return int(sum(a * b for a, b in zip(self.nc, self.ns)))

    def __getitem__(self, k):      # dict-style access for scripts
        return getattr(self, k)

    def desc(self):
        """fv_hybrid_desc plus the arrays it points into (keep both alive)."""
        nc = np.ascontiguousarray(self.nc, dtype=np.int32)
        ns = np.ascontiguousarray(self.ns, dtype=np.int32)
        grids = np.ascontiguousarray(self.grids, dtype=np.float64)
        tv = sum((a + 1) * (b + 1) for a, b in zip(self.nc, self.ns))
        if grids.shape[1:] != (tv, 3):
            raise ValueError(f"grids must be (steps, {tv}, 3), got {grids.shape}")
        d = _lib.HybridDesc()
        d.n_surfaces = len(self.nc)
        d.nc = nc.ctypes.data_as(_lib._i32p)
        d.ns = ns.ctypes.data_as(_lib._i32p)
        d.num_steps = grids.shape[0]
        d.grids = grids.ctypes.data_as(_lib._dp)
        for k in range(3):
            d.vinf[k] = float(self.vinf[k])
        d.rho, d.dt, d.nu = self.rho, self.dt, self.nu
        d.rc0_bound, d.rc0_wake = self.rc0_bound, self.rc0_wake
        d.n_keep = int(self.n_keep)
        d.sigma = self.sigma
        d.kernel = _lib.KERNELS[self.kernel]
        d.free_wake = int(bool(self.free_wake))
        d.stretch = int(bool(self.stretch))
        d.rlxf, d.vpm_nu, d.sref = self.rlxf, self.vpm_nu, self.sref
        d.max_particles = int(self.max_particles)
        d.precision = _lib.PRECISIONS[self.precision]
        return d, (nc, ns, grids)


class HybridRun:
    """A device-resident hybrid case (fv_hybrid)."""

    def __init__(self, case: HybridCase, context=None, graph=False):
        self.case = case
        # graph mode changes its context's pair-sum plans: it takes a context of its own
        self._own_ctx = context is None and graph
        self._ctx = context or (_lib.Context(_lib.default_context().device) if graph
                                else _lib.default_context())
        d, self._keep = case.desc()
        h = C.c_void_p()
        self._ctx.check(self._ctx.lib.fv_hybrid_create(self._ctx.handle, C.byref(d),
                                                        C.byref(h)))
        self._h = h
        vh = self._ctx.lib.fv_hybrid_particles(self._h)
        self.particles = VortexParticleField(max_particles=case.max_particles, rlxf=case.rlxf,
                                             kernel=case.kernel, precision=case.precision,
                                             context=self._ctx, pair_rule=case.pair_rule,
                                             _handle=vh, _owner=self)
        if case.ring_free_wake or math.isfinite(case.truncate_x):
            self.set_ring_wake(case.ring_free_wake, case.truncate_x)
        if graph:
            self.set_graph(graph)

    def step(self, n: int = 1) -> None:
        rc = self._ctx.lib.fv_hybrid_step(self._h, int(n))
        if rc == _lib.FV_E_CAP:
            raise RuntimeError(self._ctx.lib.fv_last_error(self._ctx.handle).decode())
        self._ctx.check(rc)

    @property
    def current_step(self) -> int:
        s = C.c_int64()
        self._ctx.check(self._ctx.lib.fv_hybrid_current_step(self._h, C.byref(s)))
        return s.value

    def read(self):
        n = self.case.num_steps
        cl, cdi = np.zeros(n), np.zeros(n)
        gam = np.zeros(self.case.panels)
        P = _lib.ptr
        rc = self._ctx.lib.fv_hybrid_read(self._h, P(cl), P(cdi), P(gam))
        if rc == _lib.FV_E_CAP:
            raise RuntimeError(self._ctx.lib.fv_last_error(self._ctx.handle).decode())
        self._ctx.check(rc)
        return cl, cdi, gam

    def outputs(self, particles: bool = False) -> dict:
        """CL / CDi histories, circulations, particle count and pair counts
        (and the particle state with `particles`)."""
        cl, cdi, gam = self.read()
        out = {"cl": cl, "cdi": cdi, "gamma": gam, "np": self.particles.np}
        out["seg_pairs"], out["particle_pairs"] = self.pair_counts()
        if particles:
            st = self.particles.state_dict()
            out.update({"pos": st["pos"], "alpha": st["gamma"], "sigma_p": st["sigma"],
                        "age": st["age"]})
        return out

    def set_flags(self, free_wake: bool, stretch: bool) -> None:
        self._ctx.check(self._ctx.lib.fv_hybrid_set_flags(self._h, int(free_wake), int(stretch)))

    def set_precision(self, precision: str) -> None:
        self._ctx.check(self._ctx.lib.fv_hybrid_set_precision(self._h,
                                                              _lib.PRECISIONS[precision]))

    def set_solve_reuse(self, on: bool) -> None:
        """Factor reuse of the circulation solve (on by default): a step whose
        AIC equals the last factored one bit for bit skips the factorization;
        the circulation is the same bit for bit."""
        self._ctx.check(self._ctx.lib.fv_hybrid_set_solve_reuse(self._h, int(bool(on))))

    def solve_reuse_count(self) -> int:
        """Steps so far whose circulation solve took the stored factors."""
        v = C.c_int64()
        self._ctx.check(self._ctx.lib.fv_hybrid_solve_reuse_count(self._h, C.byref(v)))
        return v.value

    def set_graph(self, on) -> None:
        """Graph mode: launches sized by capacities and device counts, steady
        steps replayed from hipGraphs (one graph launch per step); 2: the same
        launches, every step eager; False / 0: off (fv_hybrid_set_graph)."""
        self._ctx.check(self._ctx.lib.fv_hybrid_set_graph(self._h, int(on)))

    def graph_stats(self) -> tuple[int, int]:
        """(steps replayed from graphs, steps run eagerly) in graph mode."""
        a, b = C.c_int64(), C.c_int64()
        self._ctx.check(self._ctx.lib.fv_hybrid_graph_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_population(self, scheme: str = "none", max_particles: int = 60000, eps: float = 1e-3,
                       protect: int = 64, protect_dist: float = 0.0) -> None:
        """Particle population control after every advect
        (newton_pc/adapters/flap.py:466-475): "none", "drop" (keep the newest
        max_particles) or "merge" (moment-conserving pairs, error eps |V_inf|
        at 9 collocation points)."""
        code = {"none": 0, "drop": 1, "merge": 2}[scheme]
        self._ctx.check(self._ctx.lib.fv_hybrid_set_population(
            self._h, code, int(max_particles), float(eps), int(protect), float(protect_dist)))

    def set_ring_wake(self, free: bool = False, truncate_x: float = math.inf) -> None:
        """Ring-wake motion: prescribed (V dt) or free (V + bound / wake-ring
        induction, one snapshot, forward Euler: PteraSoftware's
        prescribed_wake=False); rings whose corner-mean x exceeds truncate_x
        lose their circulation (warp_fsi/kernels_wake.py:110-126)."""
        self._ctx.check(self._ctx.lib.fv_hybrid_set_ring_wake(self._h, int(bool(free)),
                                                              float(truncate_x)))

    def circulation_ledger(self):
        """Total vector circulation of particles + stored wake rings
        (flap.py:565-578)."""
        out = np.zeros(3)
        self._ctx.check(self._ctx.lib.fv_hybrid_ledger_host(self._h, _lib.ptr(out)))
        return out

    def pair_counts(self):
        a, b = C.c_double(), C.c_double()
        self._ctx.check(self._ctx.lib.fv_hybrid_pair_counts(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._ctx.lib.fv_hybrid_destroy(self._h)
            self._h = C.c_void_p()
        if getattr(self, "_own_ctx", False):
            self._ctx.close()
            self._own_ctx = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HybridBatch:
    """Independent hybrid runs advanced together (fv_hybrid_batch): once every
    run is steady, one hipGraph launch per time step advances all of them.
    Each run gets a context (stream) of its own, in graph mode by default."""

    def __init__(self, cases, graph: bool = True):
        self.runs = [HybridRun(c, graph=graph) for c in cases]
        lib = self.runs[0]._ctx.lib
        arr = (C.c_void_p * len(self.runs))(*[r._h.value for r in self.runs])
        h = C.c_void_p()
        self.runs[0]._ctx.check(lib.fv_hybrid_batch_create(arr, len(self.runs), C.byref(h)))
        self._h = h
        self._lib = lib

    def step(self, n: int = 1) -> None:
        rc = self._lib.fv_hybrid_batch_step(self._h, int(n))
        if rc == _lib.FV_E_CAP:
            raise RuntimeError(self._lib.fv_last_error(self.runs[0]._ctx.handle).decode())
        self.runs[0]._ctx.check(rc)

    def stats(self) -> tuple[int, int]:
        """(graph launches, eager batch steps) so far."""
        a, b = C.c_int64(), C.c_int64()
        self.runs[0]._ctx.check(self._lib.fv_hybrid_batch_stats(self._h, C.byref(a),
                                                                C.byref(b)))
        return a.value, b.value

    def read(self):
        return [r.read() for r in self.runs]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.fv_hybrid_batch_destroy(self._h)
            self._h = C.c_void_p()
        for r in getattr(self, "runs", []):
            r.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def theodorsen_cl(k: float, h0c: float, omega: float, t, chord: float = 1.0):
    """Theodorsen's lift coefficient of an airfoil plunging h = h0 sin(omega t),
    with C(k) from Hankel functions (tests/experiment_hybrid_panel_particle.py:
    28-36)."""
    from scipy.special import hankel2
    ck = hankel2(1, k) / (hankel2(1, k) + 1j * hankel2(0, k))
    U = omega * chord / (2 * k)
    h0 = h0c * chord
    t = np.asarray(t, dtype=float)
    return np.real(np.pi * (-h0 * omega ** 2 * np.sin(omega * t)) * chord / (2 * U ** 2)
                   - 2 * np.pi * ck * h0 * omega * np.cos(omega * t) / U)


def theodorsen_ratio(case: HybridCase, cl) -> float:
    """CL amplitude after two periods over Theodorsen's, the reference's
    measure (tests/experiment_hybrid_panel_particle.py:258-262)."""
    cl = np.asarray(cl)
    t = np.arange(len(cl)) * case.dt
    m = t > 2 * case.meta["period"]
    th = theodorsen_cl(case.meta["k"], case.meta["h0"], case.meta["omega"], t)
    return float((cl[m].max() - cl[m].min()) / (th[m].max() - th[m].min()))


def plunging_sweep_cases(ks=(0.5, 0.2, 0.1), n_keeps=(10, 20, 0), **kw) -> list:
    """The reference's reduced-frequency sweep (tests/experiment_hybrid_panel_
    particle.py:206-329): the heaving wing at each reduced frequency k with
    each N_keep (its hybrid runs; N_keep = 0 converts every row), free particle
    wake, 3 cycles of 50 steps.  Every case has the same step count; the
    particle counts differ by N_keep."""
    kw.setdefault("free_wake", True)
    kw.setdefault("stretch", False)
    kw.setdefault("kernel", "gaussianerf")
    kw.setdefault("sigma_factor", 1.0)
    kw.setdefault("max_particles", 8000)
    out = []
    for k in ks:
        for nk in n_keeps:
            c = heaving_wing_case(k=k, n_keep=nk, **kw)
            c.name = f"plunge_k{k}_n{nk}"
            out.append(c)
    return out


def run_case(case: HybridCase, steps: int | None = None, particles: bool = False,
             graph=False) -> dict:
    """Runs `steps` (default all) steps on the device and returns the histories
    (graph: HybridRun's graph mode)."""
    run = HybridRun(case, graph=graph)
    try:
        run.step(case.num_steps if steps is None else steps)
        return run.outputs(particles)
    finally:
        run.close()


# ---------------------------------------------------------------- builders
def _case(airplane: Airplane, V: float, alpha_deg: float, dt: float, num_steps: int,
          **kw) -> HybridCase:
    nc, ns = airplane.sizes()
    return HybridCase(nc=nc, ns=ns, grids=airplane.grids(num_steps, dt),
                      vinf=freestream(V, alpha_deg), dt=dt, **kw)


def heaving_wing_case(nc=10, ns=6, chord=1.0, half_span=5.0, V=10.0, k=0.5, h0c=0.1,
                      cycles=3.0, steps_per_cycle=50, n_keep=10, kernel="winckelmans",
                      sigma_factor=0.5, free_wake=True, stretch=True, rlxf=0.3,
                      precision="fp64", max_particles=200000, num_steps=None,
                      pair_rule="clamp") -> HybridCase:
    """Heaving rectangular wing: NACA 0012, uniform panels mirrored into one
    lattice, h(t) = h0 sin(2 k V t / c), dt = period / 50, N_keep wake rows kept
    as rings, older rows converted with sigma = sigma_factor |V| dt
    (tests/experiment_hybrid_panel_particle.py:39-69,145-157;
    tests/experiment_winckelmans.py:75-106)."""
    omega = 2 * k * V / chord
    period = 2 * math.pi / omega
    dt = period / steps_per_cycle
    steps = int(round(cycles * steps_per_cycle)) if num_steps is None else int(num_steps)
    wing = Surface(root_chord=chord, tip_chord=chord, tip_offset=(0.0, half_span, 0.0), nc=nc,
                   ns=ns, airfoil="0012", symmetric=True)
    ap = Airplane([wing], Motion(heave_amp=h0c * chord, heave_period=period))
    return _case(ap, V, 0.0, dt, steps, rc0_bound=0.03 * chord, rc0_wake=0.03 * chord,
                 n_keep=n_keep, sigma=sigma_factor * V * dt, kernel=kernel, free_wake=free_wake,
                 stretch=stretch, rlxf=rlxf, sref=wing.planform_area(),
                 max_particles=max_particles, precision=precision, nu=15.06e-6,
                 pair_rule=pair_rule, name=f"heaving wing k={k} nc={nc} ns={2 * ns} N_keep={n_keep}",
                 meta={"period": period, "omega": omega, "h0": h0c * chord, "V": V, "k": k})


def static_wing_case(nc=7, ns=18, chord=2.0, half_span=5.0, V=10.0, alpha_deg=5.0,
                     wake_chords=6.0, symmetric=True, precision="fp64",
                     n_keep=None) -> HybridCase:
    """Static cambered wing (NACA 2412, cosine panels) with a prescribed ring
    wake grown over `wake_chords` chord lengths; dt = panel chord / V
    (tests/benchmark_vs_pterasoftware.py:30-111)."""
    dt = (chord / nc) / V
    steps = int(math.ceil(wake_chords * chord / (V * dt)))
    wing = Surface(root_chord=chord, tip_chord=chord, tip_offset=(0.0, half_span, 0.0), nc=nc,
                   ns=ns, airfoil="2412", chordwise_spacing="cosine", spanwise_spacing="cosine",
                   symmetric=symmetric)
    ap = Airplane([wing])
    return _case(ap, V, alpha_deg, dt, steps, rc0_bound=0.03 * chord, rc0_wake=0.03 * chord,
                 n_keep=steps + 1 if n_keep is None else n_keep, sigma=V * dt,
                 kernel="gaussianerf", sref=wing.planform_area(), precision=precision,
                 nu=15.06e-6, name=f"static wing {nc}x{ns * (2 if symmetric else 1)}")


def flat_plate_case(nc=4, ns=8, chord=1.0, half_span=4.0, V=10.0, alpha_deg=2.0,
                    wake_chords=20.0, symmetric=False, spacing="uniform", dt=None,
                    n_keep=None, precision="fp64") -> HybridCase:
    """Static rectangular flat plate (symmetric section, zero camber) with a
    prescribed ring wake grown over `wake_chords` chord lengths; dt defaults to
    panel chord / V.  With symmetric=False the lattice spans y in [0, half_span]
    (one semi-span, as PteraSoftware's `symmetric=False` wing)."""
    dt = (chord / nc) / V if dt is None else dt
    steps = int(math.ceil(wake_chords * chord / (V * dt)))
    wing = Surface(root_chord=chord, tip_chord=chord, tip_offset=(0.0, half_span, 0.0), nc=nc,
                   ns=ns, airfoil="0012", chordwise_spacing=spacing, spanwise_spacing=spacing,
                   symmetric=symmetric)
    ap = Airplane([wing])
    return _case(ap, V, alpha_deg, dt, steps, rc0_bound=0.0, rc0_wake=0.0,
                 n_keep=steps + 1 if n_keep is None else n_keep, sigma=V * dt,
                 kernel="gaussianerf", sref=wing.planform_area(), precision=precision,
                 nu=0.0, name=f"flat plate {nc}x{ns * (2 if symmetric else 1)}")


def flapping_case(V=10.0, alpha_deg=1.0, dt=1.0 / 43, cycles=3.0, n_keep=10,
                  kernel="winckelmans", free_wake=False, stretch=False, sigma_factor=1.0,
                  precision="fp64", max_particles=400000, num_steps=None) -> HybridCase:
    """Flapping airplane: main wing (2412, 1.75 -> 1.5 m chord, 6 m half span,
    15 deg flapping, period 1 s, root off the symmetry plane -> two lattices)
    plus a V-tail (0012, 1.5 -> 1.0 m, mirrored into one lattice): 3 surfaces,
    192 panels (tests/benchmark_flapping.py:28-146)."""
    main = Surface(root_chord=1.75, tip_chord=1.5, tip_offset=(0.75, 6.0, 1.0), nc=6, ns=8,
                   airfoil="2412", chordwise_spacing="uniform", spanwise_spacing="cosine",
                   symmetric=True, root_le=(0.0, 0.5, 0.0), tip_twist_deg=5.0, flap=True)
    tail = Surface(root_chord=1.5, tip_chord=1.0, tip_offset=(0.5, 2.0, 1.0), nc=6, ns=8,
                   airfoil="0012", chordwise_spacing="uniform", spanwise_spacing="uniform",
                   symmetric=True, root_le=(5.0, 0.0, 0.0), incidence_deg=-5.0)
    ap = Airplane([main, tail], Motion(flap_amp_deg=15.0, flap_period=1.0))
    steps = int(round(cycles / dt)) if num_steps is None else int(num_steps)
    cbar = 0.5 * (1.75 + 1.5)
    return _case(ap, V, alpha_deg, dt, steps, rc0_bound=0.03 * cbar, rc0_wake=0.03 * cbar,
                 n_keep=n_keep, sigma=sigma_factor * V * dt, kernel=kernel,
                 free_wake=free_wake, stretch=stretch, sref=main.planform_area(),
                 max_particles=max_particles, precision=precision, nu=15.06e-6,
                 name="flapping airplane 192 panels")


# ------------------------------------------------------- reference-style API
@dataclass
class UnsteadyProblem:
    """Airplane + operating point + time stepping (the pieces of PteraSoftware's
    UnsteadyProblem that the hybrid step uses)."""
    airplane: Airplane
    V: float = 10.0
    alpha_deg: float = 0.0
    rho: float = 1.225
    nu: float = 15.06e-6
    dt: float = 0.01
    num_steps: int = 100
    rc0: float | None = None
    sref: float | None = None


class HybridSolver:
    """Mirror of the reference's HybridSolver(prob, n_keep, kernel, free_vpm)
    (tests/experiment_hybrid_panel_particle.py:87-112).  ``stretch`` (default
    False, as the reference's advect_rk3 call at :140-143) and ``sigma_factor``
    (sigma = sigma_factor |V| dt, :157) are exposed."""

    def __init__(self, prob: UnsteadyProblem, n_keep=5, kernel="gaussianerf", free_vpm=False,
                 stretch=False, sigma_factor=1.0, precision="fp64", max_particles=200000,
                 rlxf=0.3, pair_rule="clamp"):
        self.problem = prob
        self._n_keep = n_keep
        self._free_vpm = free_vpm
        surf0 = prob.airplane.surfaces[0]
        cbar = 0.5 * (surf0.root_chord + surf0.tip_chord)
        rc0 = 0.03 * cbar if prob.rc0 is None else prob.rc0
        self.case = _case(prob.airplane, prob.V, prob.alpha_deg, prob.dt, prob.num_steps,
                          rho=prob.rho, nu=prob.nu, rc0_bound=rc0, rc0_wake=rc0, n_keep=n_keep,
                          sigma=sigma_factor * prob.V * prob.dt, kernel=kernel,
                          free_wake=free_vpm, stretch=stretch, rlxf=rlxf,
                          sref=prob.sref or prob.airplane.reference_area(),
                          max_particles=max_particles, precision=precision,
                          pair_rule=pair_rule)
        self._run = HybridRun(self.case)
        self._vpm = self._run.particles
        self.cl = self.cdi = self.gamma = None

    @property
    def num_steps(self):
        return self.case.num_steps

    def run(self, prescribed_wake=True, **kwargs):
        if not prescribed_wake:
            raise NotImplementedError("only the prescribed ring wake is supported")
        self._run.step(self.case.num_steps - self._run.current_step)
        self.cl, self.cdi, self.gamma = self._run.read()
        return self.cl, self.cdi


def with_precision(case: HybridCase, precision: str) -> HybridCase:
    return replace(case, precision=precision)
