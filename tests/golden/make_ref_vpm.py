"""Writes tests/golden/ref_vpm.npz from the reference's own CPU particle code.

Runs in the build container only (the reference is not shipped anywhere):
it imports src/fluxvortex/kernel.py and src/fluxvortex/particles.py from the
reference checkout (pure NumPy + SciPy) and records, for seeded inputs,

  * kernel level: velocity_from_particles / jacobian_from_particles on a
    200-particle cloud, both smoothing kernels (kernel.py:69-190; the self
    pair is evaluated with r clamped to 1e-12, as the reference does);
  * whole trajectories of VortexParticleField.advect_rk3 + _relax_pedrizzetti
    (particles.py:85-246), five steps each:
      - cloud_w      Winckelmans, stretching + free wake, rlxf = 0
                     (pins the RK3 stage composition: position update, then J
                     at the updated positions, rVPM Z, clamps, NaN guard);
      - cloud_g      the same with the Gaussian-erf kernel;
      - cloud_w_rlx  Winckelmans, stretching + free wake, rlxf = 0.3 (the
                     relaxation always takes the Gaussian Jacobian, :219,
                     whose self pair is kept with r = 1e-12);
      - cloud_g_rlx  Gaussian, free wake, no stretching, rlxf = 0.3, sigma =
                     V dt of the k = 0.5 plunging wing (constant cores, as in
                     HybridSolver's advect);
      - sheet_rlx    particles converted from a 2 x 6 ring-wake sheet the way
                     HybridSolver converts rows (four per ring at the leg
                     midpoints, shared legs coincide exactly), Gaussian,
                     sigma = V dt, free wake, no stretching, rlxf = 0.3
                     (experiment_hybrid_panel_particle.py:111,140-143,173-203);
      - cloud_w_ext  Winckelmans, stretching + free wake, rlxf = 0, with a
                     non-uniform U_inf_func and a bound_velocity_func
                     (tests/helpers.py: shear_freestream, swirl_bound), both
                     evaluated at every stage's positions (:126-133);
      - cloud_g_ext  Gaussian, prescribed wake (free_wake False: zeros +
                     U_bound + U_inf), rlxf = 0.3, the same callables.
The fixture holds inputs and outputs only.

    python tests/golden/make_ref_vpm.py [/path/to/reference]
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
sys.path.insert(0, str(HERE.parent))
from helpers import shear_freestream, swirl_bound  # noqa: E402

STEPS = 5


def cloud(n, seed, sigma):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1.0, 1.0, (n, 3))
    alpha = rng.normal(0.0, 0.01, (n, 3))
    return pos, alpha, np.full(n, sigma)


K05_DT = (2.0 * np.pi / 10.0) / 50.0   # dt of the k = 0.5 plunging wing (period / 50)


def sheet(nrow=2, nspan=6, c=0.1, b=1.0, V=10.0, dt=K05_DT):
    """A flat 2 x 6 wake sheet of rings with a tip-decaying spanwise
    circulation, converted into particles in HybridSolver's leg order
    (front Fr->Fl, left Fl->Bl, back Bl->Br, right Br->Fr); sigma = V dt."""
    ys = np.linspace(-b / 2, b / 2, nspan + 1)
    xs = 1.0 + np.arange(nrow + 1) * c
    zs = 0.02 * np.sin(np.pi * ys)                 # slightly cambered sheet
    pos, alpha = [], []
    for i in range(nrow):
        for j in range(nspan):
            fl = np.array([xs[i], ys[j], zs[j]])
            fr = np.array([xs[i], ys[j + 1], zs[j + 1]])
            bl = np.array([xs[i + 1], ys[j], zs[j]])
            br = np.array([xs[i + 1], ys[j + 1], zs[j + 1]])
            yc = 0.5 * (ys[j] + ys[j + 1])
            g = (0.8 + 0.2 * i) * np.sqrt(max(1.0 - (2 * yc / b) ** 2, 0.0))
            for a, bb in ((fr, fl), (fl, bl), (bl, br), (br, fr)):
                al = g * (bb - a)
                if np.dot(al, al) > 1e-30:
                    pos.append(0.5 * (a + bb))
                    alpha.append(al)
    pos, alpha = np.array(pos), np.array(alpha)
    return pos, alpha, np.full(len(pos), V * dt)


def main():
    sys.path.insert(0, str(REF / "src"))
    from fluxvortex import kernel as K  # noqa: E402  (reference, this container only)
    from fluxvortex.particles import VortexParticleField  # noqa: E402

    out = {}
    pos, al, sg = cloud(200, 2024, 0.1)
    out.update(kin_pos=pos, kin_alpha=al, kin_sigma=sg)
    for kname in ("gaussianerf", "winckelmans"):
        out[f"kin_vel_{kname}"] = K.velocity_from_particles(pos, pos, al, sg, kernel=kname)
        out[f"kin_jac_{kname}"] = K.jacobian_from_particles(pos, al, pos, al, sg, kernel=kname)

    cases = {
        "cloud_w": dict(inp=cloud(200, 11, 0.1), kernel="winckelmans", rlxf=0.0, stretch=True,
                        free=True, dt=0.01, uinf=(10.0, 0.0, 0.5)),
        "cloud_g": dict(inp=cloud(200, 12, 0.1), kernel="gaussianerf", rlxf=0.0, stretch=True,
                        free=True, dt=0.01, uinf=(10.0, 0.0, 0.5)),
        "cloud_w_rlx": dict(inp=cloud(200, 13, 0.1), kernel="winckelmans", rlxf=0.3,
                            stretch=True, free=True, dt=0.01, uinf=(10.0, 0.0, 0.5)),
        "cloud_g_rlx": dict(inp=cloud(200, 14, 10.0 * K05_DT), kernel="gaussianerf", rlxf=0.3,
                            stretch=False, free=True, dt=K05_DT, uinf=(10.0, 0.0, 0.0)),
        "sheet_rlx": dict(inp=sheet(), kernel="gaussianerf", rlxf=0.3, stretch=False,
                          free=True, dt=K05_DT, uinf=(10.0, 0.0, 0.0)),
        "cloud_w_ext": dict(inp=cloud(200, 15, 0.1), kernel="winckelmans", rlxf=0.0,
                            stretch=True, free=True, dt=0.01, uinf=None),
        "cloud_g_ext": dict(inp=cloud(200, 16, 10.0 * K05_DT), kernel="gaussianerf", rlxf=0.3,
                            stretch=False, free=False, dt=K05_DT, uinf=None),
    }
    for name, c in cases.items():
        p0, a0, s0 = c["inp"]
        f = VortexParticleField(max_particles=len(p0), nu=0.0, rlxf=c["rlxf"],
                                kernel=c["kernel"])
        f.add_particles_batch(p0, a0, s0)
        ext = c["uinf"] is None
        U = np.full(3, np.nan) if ext else np.array(c["uinf"], dtype=np.float64)
        uinf_func = shear_freestream if ext else (lambda X: np.broadcast_to(U, X.shape).copy())
        bound = swirl_bound if ext else None
        traj_p, traj_g, traj_s = [], [], []
        for _ in range(STEPS):
            f.advect_rk3(c["dt"], uinf_func, bound_velocity_func=bound,
                         stretch=c["stretch"], free_wake=c["free"])
            traj_p.append(f.positions)
            traj_g.append(f.gammas)
            traj_s.append(f.sigmas)
        out[f"{name}_pos0"], out[f"{name}_alpha0"], out[f"{name}_sigma0"] = p0, a0, s0
        out[f"{name}_pos"] = np.array(traj_p)
        out[f"{name}_alpha"] = np.array(traj_g)
        out[f"{name}_sigma"] = np.array(traj_s)
        out[f"{name}_params"] = np.array([c["dt"], *U, c["rlxf"], float(c["stretch"]),
                                          float(c["free"]),
                                          1.0 if c["kernel"] == "winckelmans" else 0.0])
    np.savez_compressed(HERE / "ref_vpm.npz", **out)
    print("wrote", HERE / "ref_vpm.npz", sorted(out))


if __name__ == "__main__":
    main()
