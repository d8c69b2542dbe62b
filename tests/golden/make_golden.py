"""Writes tests/golden/*.npz: inputs from the reference's own seeded recipes
and outputs of the CPU oracle.

The reference stores no kernel-level golden arrays (its tests compare live
against PteraSoftware's Numba module, which is not installed here), so the
fixtures hold the reference's input recipes (tests/test_correctness.py:19-31,
warp_patch.py:77-88) with the oracle's FP64 outputs; the oracle itself is
pinned by tests/test_oracle.py.

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "oracle"))
sys.path.insert(0, str(HERE.parent))

import oracle  # noqa: E402
from helpers import particle_cloud  # noqa: E402


def main():
    pts, br, fr, fl, bl, g, rc, sing, ages, nu = oracle.seeded_rings()
    np.savez_compressed(
        HERE / "seeded_rings.npz",
        points=pts, br=br, fr=fr, fl=fl, bl=bl, strengths=g, rc0s=rc, ages=ages, nu=nu,
        collapsed_ring=oracle.collapsed(pts, br, fr, fl, bl, g, rc, ages, nu, "ring"),
        collapsed_chordwise=oracle.collapsed(pts, br, fr, fl, bl, g, rc, ages, nu, "chordwise"),
        # expanded outputs for the first 50 points only (size)
        expanded_ring=oracle.expanded(pts[:50], br, fr, fl, bl, g, rc, ages, nu, "ring"),
        collapsed_horseshoe=oracle.collapsed(pts, br, fr, fl, bl, g, rc, None, nu, "horseshoe"),
        expanded_horseshoe=oracle.expanded(pts[:50], br, fr, fl, bl, g, rc, None, nu,
                                           "horseshoe"),
    )
    pos, al, sg = particle_cloud(400, seed=13, sigma=0.1)
    out = {"pos": pos, "alpha": al, "sigma": sg}
    for k in ("gaussianerf", "winckelmans"):
        out[f"vel_{k}"] = oracle.particle_vel(pos, pos, al, sg, k)
        out[f"jac_{k}"] = oracle.particle_jac(pos, pos, al, sg, k)
        p2, g2, s2, a2 = oracle.vpm_advect_rk3(pos, al, sg, np.zeros(400), 0.01,
                                               np.array([10.0, 0, 0.5]), True, True, k)
        out[f"rk3_{k}_pos"], out[f"rk3_{k}_gamma"], out[f"rk3_{k}_sigma"] = p2, g2, s2
    np.savez_compressed(HERE / "particles.npz", **out)
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))


if __name__ == "__main__":
    main()
