"""GPU parity of the line-vortex, AIC, dense-solve and particle kernels against
the CPU oracle (FP64: <= 1e-12 normwise relative; mixed: stated bounds)."""
import numpy as np
import pytest

from helpers import particle_cloud, relerr, sheet_rings

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module")
def vk(ctx):
    from fluxvortex_amd import vortex_kernels
    vortex_kernels.set_precision("fp64")
    return vortex_kernels


def test_seeded_reference_recipe_all_five(vk, oracle):
    """tests/test_correctness.py:19-55 inputs, all five Ptera routines."""
    args = oracle.seeded_rings()
    pts, br, fr, fl, bl, g, rc, sing, ages, nu = args
    cases = [
        ("collapsed_velocities_from_ring_vortices", args,
         oracle.collapsed(pts, br, fr, fl, bl, g, rc, ages, nu, "ring")),
        ("collapsed_velocities_from_ring_vortices_chordwise_segments", args,
         oracle.collapsed(pts, br, fr, fl, bl, g, rc, ages, nu, "chordwise")),
        ("expanded_velocities_from_ring_vortices", args,
         oracle.expanded(pts, br, fr, fl, bl, g, rc, ages, nu, "ring")),
        ("collapsed_velocities_from_horseshoe_vortices", args[:8] + (nu,),
         oracle.collapsed(pts, br, fr, fl, bl, g, rc, None, nu, "horseshoe")),
        ("expanded_velocities_from_horseshoe_vortices", args[:8] + (nu,),
         oracle.expanded(pts, br, fr, fl, bl, g, rc, None, nu, "horseshoe")),
    ]
    for name, a, ref in cases:
        out = getattr(vk, name)(*a)
        assert out.shape == ref.shape, name
        assert relerr(out, ref) < TOL64, (name, relerr(out, ref))
        # the reference's own pass bar is max abs error < 1e-8 (test_correctness.py:55)
        assert np.max(np.abs(out - ref)) < 1e-12, name
    assert np.all(sing == 0)


def test_golden_fixture(vk):
    from pathlib import Path
    f = Path(__file__).parent / "golden" / "seeded_rings.npz"
    z = np.load(f)
    args = (z["points"], z["br"], z["fr"], z["fl"], z["bl"], z["strengths"], z["rc0s"],
            np.zeros(4, np.int64), z["ages"], float(z["nu"]))
    assert relerr(vk.collapsed_velocities_from_ring_vortices(*args), z["collapsed_ring"]) < TOL64
    e50 = (args[0][:50],) + args[1:]
    assert relerr(vk.expanded_velocities_from_ring_vortices(*e50), z["expanded_ring"]) < TOL64
    assert relerr(vk.collapsed_velocities_from_ring_vortices_chordwise_segments(*args),
                  z["collapsed_chordwise"]) < TOL64
    hs = args[:8] + (args[9],)
    assert relerr(vk.collapsed_velocities_from_horseshoe_vortices(*hs),
                  z["collapsed_horseshoe"]) < TOL64
    assert relerr(vk.expanded_velocities_from_horseshoe_vortices(*((hs[0][:50],) + hs[1:])),
                  z["expanded_horseshoe"]) < TOL64


@pytest.mark.parametrize("N,M", [(1, 1), (3, 257), (257, 3), (1000, 65), (4099, 1001)])
def test_ragged_sizes(vk, oracle, N, M):
    rng = np.random.default_rng(N * 7 + M)
    pts = rng.standard_normal((N, 3))
    br, fr, fl, bl = (rng.standard_normal((M, 3)) for _ in range(4))
    g = rng.standard_normal(M)
    rc = np.full(M, 0.03)
    ages = rng.uniform(0, 1, M)
    s = np.zeros(4, np.int64)
    out = vk.collapsed_velocities_from_ring_vortices(pts, br, fr, fl, bl, g, rc, s, ages, 1e-5)
    ref = oracle.collapsed(pts, br, fr, fl, bl, g, rc, ages, 1e-5)
    assert relerr(out, ref) < TOL64


def test_empty_inputs(vk):
    s = np.zeros(4, np.int64)
    z3 = np.zeros((0, 3))
    out = vk.collapsed_velocities_from_ring_vortices(np.ones((5, 3)), z3, z3, z3, z3,
                                                     np.zeros(0), np.zeros(0), s)
    assert out.shape == (5, 3) and np.all(out == 0)
    out = vk.expanded_velocities_from_ring_vortices(z3, np.ones((4, 3)), np.ones((4, 3)),
                                                    np.ones((4, 3)), np.ones((4, 3)), np.ones(4),
                                                    np.ones(4), s)
    assert out.shape == (0, 4, 3)


def test_singular_geometry(vk, oracle):
    # points on segment lines, at vertices, zero-length legs, coincident corners
    br = np.array([[1.0, 1, 0], [0, 0, 0], [2, 2, 2]])
    fr = np.array([[0.0, 1, 0], [0, 0, 0], [2, 2, 2]])
    fl = np.array([[0.0, 0, 0], [1, 0, 0], [3, 2, 2]])
    bl = np.array([[1.0, 0, 0], [1, 1, 0], [3, 3, 2]])
    pts = np.array([[0.5, 1, 0], [1, 1, 0], [0, 0, 0], [2, 2, 2], [0.5, 0.5, 0], [7, -3, 2]])
    g = np.array([1.0, -2.0, 0.5])
    rc = np.array([0.0, 0.01, 0.0])
    s = np.zeros(4, np.int64)
    out = vk.collapsed_velocities_from_ring_vortices(pts, br, fr, fl, bl, g, rc, s)
    ref = oracle.collapsed(pts, br, fr, fl, bl, g, rc)
    assert np.all(np.isfinite(out))
    assert relerr(out, ref) < TOL64


def test_sheet_sweep_large(vk, oracle):
    """Raw-kernel sweep shape: 1000 rings on a perturbed sheet, targets = ring
    centroids + 20k particles-like points."""
    br, fr, fl, bl, g, rc, age = sheet_rings(40, 25)
    cent = 0.25 * (br + fr + fl + bl)
    rng = np.random.default_rng(42)
    pts = np.concatenate([cent, rng.uniform([0, -5, -1], [20, 5, 1], (20000, 3))])
    s = np.zeros(4, np.int64)
    out = vk.collapsed_velocities_from_ring_vortices(pts, br, fr, fl, bl, g, rc, s, age, 1.5e-5)
    ref = oracle.collapsed(pts, br, fr, fl, bl, g, rc, age, 1.5e-5)
    assert relerr(out, ref) < TOL64
    # deterministic: a second call is bitwise identical
    out2 = vk.collapsed_velocities_from_ring_vortices(pts, br, fr, fl, bl, g, rc, s, age, 1.5e-5)
    assert np.array_equal(out, out2)


@pytest.mark.parametrize("m_rings,n_part,kernel", [(1000, 10_000, "gaussianerf"),
                                                   (1000, 100_000, "gaussianerf"),
                                                   (1000, 1_000_000, "gaussianerf"),
                                                   (10_000, 10_000, "winckelmans"),
                                                   (10_000, 100_000, "winckelmans"),
                                                   (10_000, 1_000_000, "winckelmans")])
def test_raw_kernel_sweep_full_sizes(vk, oracle, m_rings, n_part, kernel):
    """The seeded synthetic raw-kernel sweep at its full sizes: M rings on the
    perturbed sheet, N particles uniform in the 20 x 10 x 2 slab with
    alpha ~ N(0, 0.01^2) per component and sigma = 0.05; targets are the ring
    centroids plus every particle.  The device evaluates every target; 256
    random targets (and 256 particles for the stretching term S = J alpha)
    are checked against the oracle's direct sums over all sources: FP64 to
    1e-12, mixed to 1e-5 (normwise relative)."""
    from fluxvortex_amd import kernel as K
    nx = 40 if m_rings == 1000 else 125
    br, fr, fl, bl, g, rc, age = sheet_rings(nx, m_rings // nx, seed=11)
    rng = np.random.default_rng(n_part)
    pos = rng.uniform([0.0, -5.0, -1.0], [20.0, 5.0, 1.0], (n_part, 3))
    al = rng.normal(0.0, 0.01, (n_part, 3))
    sg = np.full(n_part, 0.05)
    tgt = np.concatenate([0.25 * (br + fr + fl + bl), pos])
    pick = np.sort(rng.choice(len(tgt), 256, replace=False))
    pick[0] = 0                                   # a collocation point is always in
    ppick = np.sort(rng.choice(n_part, 256, replace=False))
    s = np.zeros(4, np.int64)
    nu = 1.5e-5
    ref_ring = oracle.collapsed(tgt[pick], br, fr, fl, bl, g, rc, age, nu)
    ref_part = oracle.particle_vel(tgt[pick], pos, al, sg, kernel)
    ref_S = np.einsum("tij,tj->ti", oracle.particle_jac(pos[ppick], pos, al, sg, kernel),
                      al[ppick])
    errs = {}
    for prec, bar in (("fp64", TOL64), ("mixed", 1e-5)):
        vk.set_precision(prec)
        try:
            u_ring = vk.collapsed_velocities_from_ring_vortices(tgt, br, fr, fl, bl, g, rc, s,
                                                                age, nu)
        finally:
            vk.set_precision("fp64")
        u_part = K.velocity_from_particles(tgt, pos, al, sg, kernel=kernel, precision=prec)
        J = K.jacobian_from_particles(pos, al, pos, al, sg, kernel=kernel, precision=prec)
        S = np.einsum("tij,tj->ti", J[ppick], al[ppick])
        assert u_ring.shape == (len(tgt), 3) and u_part.shape == (len(tgt), 3)
        assert np.all(np.isfinite(u_part)) and np.all(np.isfinite(J))
        errs[prec] = (relerr(u_ring[pick], ref_ring), relerr(u_part[pick], ref_part),
                      relerr(S, ref_S))
        print(f"M={m_rings} N={n_part} {kernel} {prec}: ring {errs[prec][0]:.2e} "
              f"particles {errs[prec][1]:.2e} stretching {errs[prec][2]:.2e}", flush=True)
        assert max(errs[prec]) < bar, (prec, errs[prec])


def test_mixed_precision_collapsed(vk, oracle):
    br, fr, fl, bl, g, rc, age = sheet_rings(40, 25)
    rng = np.random.default_rng(1)
    pts = rng.uniform([0, -5, -1], [20, 5, 1], (5000, 3))
    s = np.zeros(4, np.int64)
    ref = oracle.collapsed(pts, br, fr, fl, bl, g, rc, age, 1.5e-5)
    vk.set_precision("mixed")
    try:
        out = vk.collapsed_velocities_from_ring_vortices(pts, br, fr, fl, bl, g, rc, s, age,
                                                         1.5e-5)
    finally:
        vk.set_precision("fp64")
    assert relerr(out, ref) < 1e-5


def test_aic_matches_oracle(vk, oracle):
    rng = np.random.default_rng(3)
    for P in (1, 24, 126, 300):
        br, fr, fl, bl = (rng.standard_normal((P, 3)) for _ in range(4))
        col = rng.standard_normal((P, 3))
        nrm = rng.standard_normal((P, 3))
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        rc = np.full(P, 0.02)
        A = vk.aic(col, nrm, br, fr, fl, bl, rc, 1e-5)
        assert relerr(A, oracle.aic(col, nrm, br, fr, fl, bl, rc, 1e-5)) < TOL64


def test_dense_solve(vk):
    rng = np.random.default_rng(5)
    for n in (1, 5, 64, 126, 141, 142, 252, 300, 1024):
        A = rng.standard_normal((n, n)) + 0.1 * n * np.eye(n)
        b = rng.standard_normal(n)
        x = vk.dense_solve(A, b)
        ref = np.linalg.solve(A, b)
        assert relerr(x, ref) < 1e-11, n
    # batched
    A = rng.standard_normal((7, 40, 40)) + 4 * np.eye(40)
    b = rng.standard_normal((7, 40))
    x = vk.dense_solve(A, b)
    assert relerr(x, np.linalg.solve(A, b[..., None])[..., 0]) < 1e-11
    with pytest.raises(np.linalg.LinAlgError):
        vk.dense_solve(np.zeros((4, 4)), np.ones(4))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dense_solve_with_row_exchanges(vk, precision):
    """Rows of a diagonally dominant matrix shuffled, so nearly every column
    pivots on another row (the register panels track row positions instead of
    moving rows), and small-integer matrices with many equal magnitudes (the
    first-maximum pivot rule): solutions within 1e-11 of numpy's, for panels
    factored in registers (<= 128 rows), through LDS and from global memory."""
    rng = np.random.default_rng(8)
    for n in (5, 17, 64, 120, 129, 142, 300):
        A = rng.standard_normal((n, n)) + 0.1 * n * np.eye(n)
        A = A[rng.permutation(n)]
        b = rng.standard_normal(n)
        x = vk.dense_solve(A, b, precision=precision)
        assert relerr(x, np.linalg.solve(A, b)) < 1e-11, (n, precision)
        B = rng.integers(-3, 4, (n, n)).astype(float) + 4.0 * np.sqrt(n) * np.eye(n)
        B = B[rng.permutation(n)]
        x = vk.dense_solve(B, b, precision=precision)
        assert relerr(x, np.linalg.solve(B, b)) < 1e-11, (n, precision, "ties")


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_particle_velocity_and_jacobian(oracle, ctx, kernel):
    from fluxvortex_amd import kernel as K
    pos, al, sg = particle_cloud(3000, seed=5, sigma=0.05)
    tgt = np.concatenate([pos, np.random.default_rng(2).uniform(-1, 1, (500, 3))])
    u = K.velocity_from_particles(tgt, pos, al, sg, kernel=kernel)
    assert relerr(u, oracle.particle_vel(tgt, pos, al, sg, kernel)) < TOL64
    J = K.jacobian_from_particles(tgt, al, pos, al, sg, kernel=kernel)
    assert relerr(J, oracle.particle_jac(tgt, pos, al, sg, kernel)) < TOL64
    um = K.velocity_from_particles(tgt, pos, al, sg, kernel=kernel, precision="mixed")
    assert relerr(um, oracle.particle_vel(tgt, pos, al, sg, kernel)) < 1e-5
    Jm = K.jacobian_from_particles(tgt, al, pos, al, sg, kernel=kernel, precision="mixed")
    assert relerr(Jm, oracle.particle_jac(tgt, pos, al, sg, kernel)) < 1e-5


def test_particle_edge_cases(oracle, ctx):
    from fluxvortex_amd import kernel as K
    z = np.zeros((0, 3))
    assert K.velocity_from_particles(np.ones((3, 3)), z, z, np.zeros(0)).shape == (3, 3)
    assert K.velocity_from_particles(z, np.ones((2, 3)), np.ones((2, 3)), np.ones(2)).shape == (0, 3)
    # coincident particles and a target on a particle
    pos = np.array([[0.0, 0, 0], [0, 0, 0], [0.1, 0, 0]])
    al = np.array([[0.0, 0, 1], [0, 1, 0], [1, 0, 0]])
    sg = np.array([0.1, 0.1, 0.1])
    for kern in ("gaussianerf", "winckelmans"):
        u = K.velocity_from_particles(pos, pos, al, sg, kernel=kern)
        J = K.jacobian_from_particles(pos, al, pos, al, sg, kernel=kern)
        assert np.all(np.isfinite(u)) and np.all(np.isfinite(J))
        assert relerr(u, oracle.particle_vel(pos, pos, al, sg, kern)) < TOL64
        assert relerr(J, oracle.particle_jac(pos, pos, al, sg, kern)) < TOL64
    # pairs just inside (r^2 = 1e-26) and just outside (r^2 = 4e-24) the
    # exclusion radius, with cores from the 1e-6 floor to 1: excluded pairs
    # contribute exactly nothing, included ones their full term.  The
    # Gaussian-erf factor of the included pair (rho <= 2e-6) is the
    # difference of two nearly equal terms in the reference formula itself
    # (g ~ rho^3 from erf - rho e^..), so that pair enters the Winckelmans
    # check only.
    base = np.array([[0.0, 0, 0], [1e-13, 0, 0], [0.3, 0.1, 0], [0.3, 0.1, 1e-13]])
    bal = np.array([[0.0, 0, 1], [0, 1, 0], [0.2, 0.1, 0.3], [-0.1, 0.4, 0]])
    for sgv in (1e-6, 0.1, 1.0):
        for kern in ("gaussianerf", "winckelmans"):
            pos, al = base, bal
            if kern == "winckelmans":
                pos = np.concatenate([base, [[0.0, 2e-12, 0.0]]])
                al = np.concatenate([bal, [[1.0, 0.0, 0.0]]])
            tgt = np.concatenate([pos, [[0.05, 0.0, 0.0]]])
            sg = np.full(len(pos), sgv)
            u = K.velocity_from_particles(tgt, pos, al, sg, kernel=kern)
            J = K.jacobian_from_particles(tgt, np.ones_like(tgt), pos, al, sg, kernel=kern)
            assert np.all(np.isfinite(u)) and np.all(np.isfinite(J)), (sgv, kern)
            assert relerr(u, oracle.particle_vel(tgt, pos, al, sg, kern)) < TOL64, (sgv, kern)
            assert relerr(J, oracle.particle_jac(tgt, pos, al, sg, kern)) < TOL64, (sgv, kern)


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_mixed_tiny_cores_stay_finite(oracle, ctx, kernel):
    """Cores at the RK3 floor sigma = 1e-6 (where a stretched wake ends up) with
    targets up to ~300 m away: 1/sigma^5 alone is 1e30, so the FP32 pair terms
    must be formed without it.  Mixed stays finite and within 1e-5 of FP64."""
    from fluxvortex_amd import kernel as K
    rng = np.random.default_rng(11)
    n = 600
    pos = rng.uniform([0, -5, -0.5], [300, 5, 0.5], (n, 3))
    al = rng.normal(0, 1.0, (n, 3))
    sg = np.where(rng.uniform(size=n) < 0.5, 1e-6, rng.uniform(0.02, 0.1, n))
    tgt = np.concatenate([pos[:200], rng.uniform([0, -5, -0.5], [300, 5, 0.5], (100, 3))])
    tal = np.concatenate([al[:200], rng.normal(0, 1.0, (100, 3))])
    um = K.velocity_from_particles(tgt, pos, al, sg, kernel=kernel, precision="mixed")
    Jm = K.jacobian_from_particles(tgt, tal, pos, al, sg, kernel=kernel, precision="mixed")
    assert np.all(np.isfinite(um)) and np.all(np.isfinite(Jm))
    assert relerr(um, oracle.particle_vel(tgt, pos, al, sg, kernel)) < 1e-5
    assert relerr(Jm, oracle.particle_jac(tgt, pos, al, sg, kernel)) < 1e-5


def test_dense_solve_mixed(vk):
    """FP32 LU + FP64-residual refinement reaches FP64 accuracy on
    well-conditioned systems (whole matrix in LDS up to n = 200, staged
    panels beyond), within the 8 refinement steps allowed."""
    rng = np.random.default_rng(6)
    for n in (1, 7, 64, 120, 200, 201, 300):
        A = rng.standard_normal((n, n)) + 0.1 * n * np.eye(n)
        b = rng.standard_normal(n)
        x, it = vk.dense_solve(A, b, precision="mixed", return_iters=True)
        ref = np.linalg.solve(A, b)
        assert relerr(x, ref) < 1e-11, (n, relerr(x, ref))
        assert it <= 8, (n, it)
    A = rng.standard_normal((5, 40, 40)) + 4 * np.eye(40)
    b = rng.standard_normal((5, 40))
    x = vk.dense_solve(A, b, precision="mixed")
    assert relerr(x, np.linalg.solve(A, b[..., None])[..., 0]) < 1e-11
    with pytest.raises(np.linalg.LinAlgError):
        vk.dense_solve(np.zeros((4, 4)), np.ones(4), precision="mixed")
