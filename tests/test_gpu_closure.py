"""Learned far-field closure: cell binning, near-field exactness, convergence of
the first-order far expansion with the cutoff, the online-trained residual
network's gain over the expansion alone, and the check / rejection logic."""
import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu


def wake_sheet(nx=1600, ny=12, dx=0.125, span=10.0, seed=3):
    """Particles of a heaving wing's wake: a sinusoidal sheet of shed
    (spanwise) and trailing (streamwise) vorticity, 4 per ring as the
    conversion produces them."""
    rng = np.random.default_rng(seed)
    xs = np.arange(nx) * dx
    ys = np.linspace(-span / 2, span / 2, ny + 1)
    z = lambda x: 0.1 * np.sin(2 * np.pi * x / 6.283)
    pos, al = [], []
    for i, x in enumerate(xs):
        g = np.cos(2 * np.pi * x / 6.283) * (1 - (np.linspace(-1, 1, ny) ** 2))
        for j in range(ny):
            y0, y1 = ys[j], ys[j + 1]
            c = np.array([x, 0.5 * (y0 + y1), z(x)])
            pos.append(c)
            al.append([0.0, g[j] * (y1 - y0) * 0.05, 0.0])            # shed
            pos.append(np.array([x + dx / 2, y0, z(x + dx / 2)]))
            gl = g[j] - (g[j - 1] if j else 0.0)
            al.append([gl * dx * 0.05, 0.0, 0.0])                      # trailing
    pos = np.array(pos) + rng.normal(0, 1e-3, (len(pos), 3))
    return pos, np.array(al), np.full(len(pos), 0.0628)


@pytest.fixture(scope="module")
def field(ctx):
    from fluxvortex_amd.particles import VortexParticleField
    pos, al, sg = wake_sheet()
    f = VortexParticleField(max_particles=len(sg) + 64, rlxf=0.0, kernel="winckelmans")
    f.add_particles_batch(pos, al, sg)
    return f, pos, al, sg


def test_near_only_is_exact(field, oracle):
    f, pos, al, sg = field
    f.enable_closure(None, cell=2.0, kappa=1e3)           # every cell is near
    u, err = f.closure_velocity(use_mlp=False)
    assert err < 1e-12, err
    sub = np.arange(0, len(sg), 97)
    ref = oracle.particle_vel(pos[sub], pos, al, sg, "winckelmans")
    assert relerr(u[sub], ref) < 1e-12
    # the near-pair count: every target against every particle, per pass
    n2 = float(len(sg)) ** 2
    npairs = f.closure_stats()["near_pairs"]
    assert npairs > 0 and npairs % n2 == 0
    f.enable_closure(None, cell=2.0, kappa=2.0)
    f.closure_velocity(use_mlp=False)
    assert 0 < f.closure_stats()["near_pairs"] < n2


def test_near_only_clamped_vorticity_matches_oracle(ctx, oracle):
    """The relaxation's vorticity under the reference CPU pair rule (self pair
    kept with r = 1e-12), closure with every cell near, against the oracle's
    curl of its clamped Gaussian Jacobian (kernel.py:156)."""
    from fluxvortex_amd.particles import VortexParticleField
    pos, al, _ = wake_sheet(nx=200)
    sg = np.full(len(pos), 0.12566370614359174)   # V dt of the k = 0.5 plunging wing
    f = VortexParticleField(max_particles=len(sg) + 8, rlxf=0.3, kernel="gaussianerf")
    f.add_particles_batch(pos, al, sg)
    f.enable_closure(None, cell=2.0, kappa=1e3)
    w = f.closure_sum("vorticity_clamped", kernel="gaussianerf", use_mlp=False)
    J = oracle.particle_jac(pos, pos, al, sg, "gaussianerf", rule=1)
    ref = np.stack([J[:, 2, 1] - J[:, 1, 2], J[:, 0, 2] - J[:, 2, 0], J[:, 1, 0] - J[:, 0, 1]], 1)
    assert relerr(w, ref) < 1e-12
    # the self pair dominates: omega points against alpha for every particle
    nz = np.linalg.norm(al, axis=1) > 0
    cosang = np.sum(w[nz] * al[nz], 1) / (np.linalg.norm(w[nz], axis=1)
                                          * np.linalg.norm(al[nz], axis=1))
    assert np.all(cosang < -0.999)


def test_expansion_converges_with_cutoff(field):
    f, *_ = field
    errs = []
    for kappa in (2.0, 4.0, 8.0):
        f.enable_closure(None, cell=2.0, kappa=kappa)
        _, e = f.closure_velocity(use_mlp=False)
        errs.append(e)
    print("expansion-only error vs kappa (2, 4, 8):", errs)
    assert errs[0] > errs[1] > errs[2]
    assert errs[1] < 1e-3 and errs[2] < 2e-4


def test_learned_residual_gain(field):
    from fluxvortex_amd.mlp import ClosureMLP, closure_params
    f, *_ = field
    net = ClosureMLP(closure_params(seed=1))
    f.enable_closure(net, cell=2.0, kappa=4.0, train_samples=4096, train_iters=4, lr=2e-3)
    _, e0 = f.closure_velocity(use_mlp=False)
    _, e_untrained = f.closure_velocity(use_mlp=True)   # zero output layer: = expansion
    assert abs(e_untrained - e0) <= 1e-9 + 1e-6 * e0
    f.closure_train(rounds=200)
    _, e1 = f.closure_velocity(use_mlp=True)
    print(f"closure error: expansion only {e0:.3e}, with trained network {e1:.3e}, "
          f"gain x{e0 / e1:.2f}")
    # the network learns the quadrupole-order residual from the cells' second
    # moments; an exact residual at the target cell centres would give x3.3 here
    assert e1 < 0.8 * e0


def test_advect_with_closure_checks_and_stays_close(ctx, oracle):
    from fluxvortex_amd.mlp import ClosureMLP, closure_params
    from fluxvortex_amd.particles import VortexParticleField
    pos, al, sg = wake_sheet(nx=800)
    U = np.array([10.0, 0.0, 0.0])
    a = VortexParticleField(max_particles=len(sg) + 8, rlxf=0.3, kernel="winckelmans")
    b = VortexParticleField(max_particles=len(sg) + 8, rlxf=0.3, kernel="winckelmans")
    for f in (a, b):
        f.add_particles_batch(pos, al, sg)
    net = ClosureMLP(closure_params(seed=2))
    b.enable_closure(net, cell=2.0, kappa=6.0, check_every=2, tol=1e-3)
    for _ in range(4):
        a.advect_rk3(0.0125, U, stretch=False, free_wake=True)
        b.advect_rk3(0.0125, U, stretch=False, free_wake=True)
    st = b.closure_stats()
    assert st["checks"] == 2 and st["last_err"] >= 0.0
    if st["active"]:
        assert st["last_err"] <= 1e-3
    pa, pb = a.positions, b.positions
    assert relerr(pb - pos, pa - pos) < 1e-3


@pytest.mark.parametrize("kernel", ["winckelmans", "gaussianerf"])
def test_closure_modes_match_direct_sums(ctx, kernel):
    """Velocity, stretching J alpha and vorticity passes of the closure (near
    exact, far first-order expansion, no network) against the direct sums
    built from the GPU's exact pair kernels; the far share is small at
    kappa = 5, so each mode agrees to a few 1e-4 of its max norm."""
    from fluxvortex_amd import kernel as K
    from fluxvortex_amd.particles import VortexParticleField
    pos, al, sg = wake_sheet(nx=400)
    f = VortexParticleField(max_particles=len(sg) + 8, rlxf=0.0, kernel="winckelmans")
    f.add_particles_batch(pos, al, sg)
    f.enable_closure(None, cell=2.0, kappa=5.0)
    u = K.velocity_from_particles(pos, pos, al, sg, kernel=kernel)
    J = K.jacobian_from_particles(pos, al, pos, al, sg, kernel=kernel)
    S = np.einsum("nij,nj->ni", J, al)
    w = np.stack([J[:, 2, 1] - J[:, 1, 2], J[:, 0, 2] - J[:, 2, 0], J[:, 1, 0] - J[:, 0, 1]], 1)
    for mode, ref in (("velocity", u), ("stretching", S), ("vorticity", w)):
        got = f.closure_sum(mode, kernel=kernel, use_mlp=False)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"{kernel} {mode}: closure max-norm error {err:.2e}")
        assert err < 1e-3, (mode, err)
    # every cell near: exact up to summation order, all three modes
    f.enable_closure(None, cell=2.0, kappa=1e3)
    for mode, ref in (("velocity", u), ("stretching", S), ("vorticity", w)):
        got = f.closure_sum(mode, kernel=kernel, use_mlp=False)
        assert relerr(got, ref) < 1e-12, mode


@pytest.mark.parametrize("kernel", ["winckelmans", "gaussianerf"])
def test_mixed_near_field_matches_direct_sums(ctx, kernel):
    """Every cell near: the mixed-precision near field (FP32 kernel factors,
    FP64 geometry and sums) of all three modes within 1e-5 of the FP64 direct
    sums."""
    from fluxvortex_amd import kernel as K
    from fluxvortex_amd.particles import VortexParticleField
    pos, al, sg = wake_sheet(nx=400)
    f = VortexParticleField(max_particles=len(sg) + 8, rlxf=0.0, kernel="winckelmans",
                            precision="mixed")
    f.add_particles_batch(pos, al, sg)
    f.enable_closure(None, cell=2.0, kappa=1e3)
    u = K.velocity_from_particles(pos, pos, al, sg, kernel=kernel)
    J = K.jacobian_from_particles(pos, al, pos, al, sg, kernel=kernel)
    S = np.einsum("nij,nj->ni", J, al)
    w = np.stack([J[:, 2, 1] - J[:, 1, 2], J[:, 0, 2] - J[:, 2, 0], J[:, 1, 0] - J[:, 0, 1]], 1)
    for mode, ref in (("velocity", u), ("stretching", S), ("vorticity", w)):
        got = f.closure_sum(mode, kernel=kernel, use_mlp=False)
        e = relerr(got, ref)
        print(f"mixed near field {kernel} {mode}: {e:.2e}")
        assert e < 1e-5, (mode, e)


def test_closure_local_expansion_order():
    """Two clusters, each inside one cell: the far field of the source cell at
    the target particles enters per particle (first-order source expansion)
    or through the target cell's second-order local expansion of the same
    field.  Their difference is the local expansion's truncation error:
    relative to the far velocity it falls as (delta / D)^3 (x8 per doubling
    of D) and for the stretching (gradient) as (delta / D)^2 (x4) -- a wrong
    Hessian would leave the velocity at the lower order."""
    from fluxvortex_amd.particles import VortexParticleField
    rng = np.random.default_rng(3)
    n = 400
    errs = {}
    for D in (24.0, 48.0):
        tgt = rng.uniform(0.05, 1.95, (n, 3))
        src = rng.uniform(0.05, 1.95, (n, 3)) + [D, 0.0, 0.0]
        pos = np.concatenate([tgt, src])
        al = rng.normal(0, 0.01, (2 * n, 3))
        sg = np.full(2 * n, 0.2)
        f = VortexParticleField(max_particles=2 * n + 8, rlxf=0.0, kernel="winckelmans")
        f.add_particles_batch(pos, al, sg)
        f.enable_closure(None, cell=2.0, kappa=2.0)
        ref = {m: f.closure_sum(m, use_mlp=False)[:n] for m in ("velocity", "stretching")}
        f.set_closure_local(4.0)
        got = {m: f.closure_sum(m, use_mlp=False)[:n] for m in ("velocity", "stretching")}
        f.set_closure_local(0.0)
        assert np.array_equal(f.closure_sum("velocity", use_mlp=False)[:n], ref["velocity"])
        # the far part alone: closure minus the target cell's own exact sum
        f2 = VortexParticleField(max_particles=n + 8, rlxf=0.0, kernel="winckelmans")
        f2.add_particles_batch(tgt, al[:n], sg[:n])
        f2.enable_closure(None, cell=2.0, kappa=2.0)
        own = {m: f2.closure_sum(m, use_mlp=False) for m in ("velocity", "stretching")}
        for m in ref:
            far = ref[m] - own[m]
            errs[(m, D)] = np.abs(got[m] - ref[m]).max() / np.abs(far).max()
    print("local expansion errors", errs)
    assert errs[("velocity", 24.0)] < 1e-3
    assert 5.5 < errs[("velocity", 24.0)] / errs[("velocity", 48.0)] < 11.0
    assert 2.5 < errs[("stretching", 24.0)] / errs[("stretching", 48.0)] < 6.0


def test_closure_local_expansion_on_a_wake():
    """On the wake sheet the local expansion beyond 6 cells changes the
    closure's velocity / stretching / vorticity by under 1e-4 of their max
    norm, and the closure with it stays within 1e-3 of the direct sums."""
    from fluxvortex_amd import kernel as K
    from fluxvortex_amd.particles import VortexParticleField
    pos, al, sg = wake_sheet(nx=400)
    f = VortexParticleField(max_particles=len(sg) + 8, rlxf=0.0, kernel="winckelmans")
    f.add_particles_batch(pos, al, sg)
    f.enable_closure(None, cell=2.0, kappa=2.5)
    base = {m: f.closure_sum(m, use_mlp=False) for m in ("velocity", "stretching", "vorticity")}
    f.set_closure_local(6.0)
    u = K.velocity_from_particles(pos, pos, al, sg, kernel="winckelmans")
    for m in base:
        got = f.closure_sum(m, use_mlp=False)
        d = np.abs(got - base[m]).max() / np.abs(base[m]).max()
        print(f"local expansion {m}: change {d:.2e}")
        assert d < 1e-4, (m, d)
    got = f.closure_sum("velocity", use_mlp=False)
    assert np.abs(got - u).max() / np.abs(u).max() < 1e-3


def _flapping_at(hybrid, n_particles, prefill=129):
    """The 192-panel flapping airplane advanced three flapping periods (43
    steps each) with a free, unstretched particle wake, its particle wake then
    extended downstream by copies of the oldest period's particles shifted by
    k V T until n_particles exist.  Deterministic: every call gives the same
    state."""
    case = hybrid.flapping_case(num_steps=prefill + 8, free_wake=True, stretch=False)
    run = hybrid.HybridRun(case)
    run.set_flags(True, False)
    run.step(prefill)
    st = run.particles.state_dict()
    n0 = st["pos"].shape[0]
    block = min(n0, 43 * 128)
    parts = {k: [v] for k, v in st.items()}
    need, k = n_particles - n0, 1
    while need > 0:
        take = min(need, block)
        p = st["pos"][:take].copy()
        p[:, 0] += k * 10.0 * 1.0          # V T, V = 10 m/s, T = 1 s
        parts["pos"].append(p)
        parts["gamma"].append(st["gamma"][:take])
        parts["sigma"].append(st["sigma"][:take])
        parts["age"].append(st["age"][:take] + k * 1.0)
        need -= take
        k += 1
    run.particles.load_state_dict({k: np.concatenate(v) for k, v in parts.items()})
    return run, prefill


def test_flapping_airplane_closure_at_100k(ctx):
    """The learned closure on the multi-surface flapping configuration (three
    surfaces, 192 panels) at 100 000 particles, mixed precision: the cheapest
    (cell, cutoff) whose check passes, three free-wake steps, CL within 1e-3
    (relative to max |CL|) of the FP64 direct sums from the identical state."""
    from fluxvortex_amd import hybrid
    from fluxvortex_amd.mlp import ClosureMLP, closure_params
    n = 100_000
    ref, s0 = _flapping_at(hybrid, n)
    assert ref.particles.np == n
    ref.set_flags(True, False)
    ref.step(3)
    cl_ref = np.asarray(ref.read()[0][s0:s0 + 3]).copy()
    ref.close()
    chosen = None
    for cell, kappa in ((2.5, 2.0), (2.0, 2.5), (2.0, 3.0), (2.0, 4.0)):
        run, _ = _flapping_at(hybrid, n)
        run.set_precision("mixed")
        run.particles.enable_closure(ClosureMLP(closure_params(seed=0)), cell=cell, kappa=kappa,
                                     check_every=10, tol=1e-3, train_samples=4096,
                                     train_iters=2, lr=2e-3)
        run.particles.set_closure_local(2.0 * kappa)
        run.particles.closure_train(rounds=800)
        run.set_flags(True, False)
        run.step(3)
        st = run.particles.closure_stats()
        cl = np.asarray(run.read()[0][s0:s0 + 3]).copy()
        run.close()
        print(f"flapping 100k cell {cell} kappa {kappa}: {st}", flush=True)
        if st["active"] and st["rejections"] == 0:
            chosen = (cell, kappa, st, cl)
            break
    assert chosen is not None, "no closure candidate passed its check"
    cell, kappa, st, cl = chosen
    assert st["checks"] == 1 and st["last_err"] <= 1e-3 and st["far_evals"] > 0
    err = float(np.max(np.abs(cl - cl_ref)) / np.max(np.abs(cl_ref)))
    print(f"flapping 100k: CL error vs FP64 {err:.2e} (cell {cell}, kappa {kappa})", flush=True)
    assert np.all(np.isfinite(cl)) and err < 1e-3


def test_closure_binning_sized_on_device(ctx):
    """The binning takes its sizes from the device (launches and buffers sized
    by bounds, one host read per binning): after population control drops
    particles on the device (the host's launch bound then stays above the
    count) and after the field grows, an advect step with the closure equals,
    bit for bit, the same step of a field whose count the host knows exactly."""
    from fluxvortex_amd.mlp import ClosureMLP, closure_params
    from fluxvortex_amd.particles import VortexParticleField
    pos, al, sg = wake_sheet(nx=500)
    n0 = len(sg)
    net = ClosureMLP(closure_params(seed=4))
    U = np.array([10.0, 0.0, 0.0])

    def make(st):
        f = VortexParticleField(max_particles=n0 + 4096, rlxf=0.3, kernel="winckelmans")
        f.load_state_dict(st)
        # no training (the two fields share the network), a check that always
        # passes, the far band and local expansions on
        f.enable_closure(net, cell=2.0, kappa=3.0, check_every=1000, tol=1.0, train_iters=0)
        f.set_closure_local(6.0)
        return f

    a = VortexParticleField(max_particles=n0 + 4096, rlxf=0.3, kernel="winckelmans")
    a.add_particles_batch(pos, al, sg)
    a.enable_closure(net, cell=2.0, kappa=3.0, check_every=1000, tol=1.0, train_iters=0)
    a.set_closure_local(6.0)
    extra = wake_sheet(nx=120, seed=9)
    for change in ("drop", "grow", "drop"):
        st = a.state_dict()
        keep = len(st["sigma"]) - 700
        if change == "drop":
            a.drop_oldest(keep)            # the count changes on the device only
            c = make(st)
            c.drop_oldest(keep)
            c.load_state_dict(c.state_dict())   # c: the exact count on the host
        else:
            shift = np.array([a.positions[:, 0].max() + 0.5, 0.0, 0.0])
            a.add_particles_batch(extra[0] + shift, extra[1], extra[2])
            c = make(a.state_dict())
        a.advect_rk3(0.01, U, stretch=True, free_wake=True)
        c.advect_rk3(0.01, U, stretch=True, free_wake=True)
        sa, sc = a.state_dict(), c.state_dict()
        for k in ("pos", "gamma", "sigma"):
            assert np.array_equal(sa[k], sc[k]), (change, k)
        assert a.closure_stats()["active"] and c.closure_stats()["active"]
