"""Copies the reference's stored Goland aeroelastic histories into
tests/golden/goland_flutter_ref.npz.

figures/flutter_data.npz (written by tests/animate_flutter.py:124-167) holds
the nodal heave w and twist theta of the 8-element Goland beam at every
coupled step of two PteraSoftware UVLM runs (dt = 0.003 s, 60 chords of
wake, tip perturbation w = 0.05 m, theta = 2 deg): V = 100 m/s (364 steps)
and V = 160 m/s (227 steps).  The arrays are copied as they are (a safe,
pickle-free load); nothing of the reference runs.

    python tests/golden/copy_goland_fixture.py
"""
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = Path("/root/reference/figures/flutter_data.npz")


def main():
    d = np.load(SRC, allow_pickle=False)
    keys = ("w_stable", "theta_stable", "w_flutter", "theta_flutter", "y_nodes", "dt")
    np.savez_compressed(HERE / "goland_flutter_ref.npz", **{k: np.asarray(d[k]) for k in keys},
                        V_stable=100.0, V_flutter=160.0, num_chords=60.0)


if __name__ == "__main__":
    main()
