"""Lifting-surface lattices and prescribed kinematics for the hybrid driver.

Host-side setup only (no compute kernels): NACA 4-digit mean-camber sections,
uniform or cosine panel spacing, twist, sweep and dihedral through the tip
section's placement, symmetric (mirrored) wings, and per-step corner grids for
heaving and flapping motions.  Frame: x aft (chordwise), y toward the right
tip, z up; the freestream V (cos a, 0, sin a) blows along +x.

A lattice is a (nc+1, ns+1, 3) array of panel corners, i = 0 at the leading
edge, j = 0 at the left end.  Panel points are linear interpolations between
the root and tip section points, as in PteraSoftware's wing meshing (the
reference builds its cases with it: tests/experiment_hybrid_panel_particle.py:39-69,
tests/benchmark_flapping.py:28-146).  A symmetric surface whose root lies on
y = 0 becomes one lattice spanning both halves; one whose root is off the
plane becomes two lattices (right, then mirrored left).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def spacing(n: int, kind: str) -> np.ndarray:
    """n+1 stations on [0, 1]: "uniform" or "cosine" (clustered at both ends)."""
    k = np.arange(n + 1)
    if kind == "uniform":
        return k / n
    if kind == "cosine":
        return 0.5 * (1.0 - np.cos(math.pi * k / n))
    raise ValueError(f"unknown spacing {kind!r}")


def naca4_camber(x: np.ndarray, code: str) -> np.ndarray:
    """Mean-camber ordinate z/c at chord fraction x for a NACA 4-digit code."""
    code = code.lower().replace("naca", "")
    m = int(code[0]) / 100.0
    p = int(code[1]) / 10.0
    x = np.asarray(x, dtype=np.float64)
    if m == 0.0 or p == 0.0:
        return np.zeros_like(x)
    return np.where(x < p, m / p ** 2 * (2 * p * x - x * x),
                    m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * x - x * x))


def _rot_y(v: np.ndarray, deg: float) -> np.ndarray:
    """Rotation about +y; positive angles pitch the trailing edge down."""
    t = math.radians(deg)
    c, s = math.cos(t), math.sin(t)
    out = v.copy()
    out[..., 0] = v[..., 0] * c + v[..., 2] * s
    out[..., 2] = -v[..., 0] * s + v[..., 2] * c
    return out


@dataclass
class Surface:
    """One wing: root section at `root_le`, tip section at `root_le + tip_offset`."""
    root_chord: float
    tip_chord: float
    tip_offset: tuple            # (dx, dy, dz) of the tip leading edge from the root's
    nc: int
    ns: int                      # spanwise panels between root and tip
    airfoil: str = "0012"
    chordwise_spacing: str = "uniform"
    spanwise_spacing: str = "uniform"
    symmetric: bool = True
    root_le: tuple = (0.0, 0.0, 0.0)
    root_twist_deg: float = 0.0
    tip_twist_deg: float = 0.0
    incidence_deg: float = 0.0   # whole-surface rotation about y through root_le
    flap: bool = False           # follows the flapping angle of the motion

    def _section(self, chord, twist, le):
        xc = spacing(self.nc, self.chordwise_spacing)
        pts = np.zeros((self.nc + 1, 3))
        pts[:, 0] = xc * chord
        pts[:, 2] = naca4_camber(xc, self.airfoil) * chord
        return _rot_y(pts, twist) + np.asarray(le, dtype=np.float64)

    def _half(self) -> np.ndarray:
        eta = spacing(self.ns, self.spanwise_spacing)
        root = self._section(self.root_chord, self.root_twist_deg, (0.0, 0.0, 0.0))
        tip = self._section(self.tip_chord, self.tip_twist_deg, self.tip_offset)
        g = (1.0 - eta)[None, :, None] * root[:, None, :] + eta[None, :, None] * tip[:, None, :]
        return _rot_y(g, self.incidence_deg) + np.asarray(self.root_le, dtype=np.float64)

    def lattices(self):
        """[(grid, side)] with side +1 (right / single) or -1 (mirrored left)."""
        right = self._half()
        if not self.symmetric:
            return [(right, 1.0)]
        left = right[:, ::-1].copy()
        left[..., 1] *= -1.0
        if abs(self.root_le[1]) < 1e-12:
            return [(np.concatenate([left[:, :-1], right], axis=1), 0.0)]
        return [(right, 1.0), (left, -1.0)]

    def planform_area(self) -> float:
        a = 0.5 * (self.root_chord + self.tip_chord) * abs(self.tip_offset[1])
        return 2 * a if self.symmetric else a


@dataclass
class Motion:
    """Prescribed rigid motion: heave z(t) = h sin(2 pi t / T_h) of every
    surface, and flapping of the `flap` surfaces about the x axis through their
    root leading edge with phi(t) = A sin(2 pi t / T_f); mirrored halves flap
    with the opposite rotation so both tips rise together."""
    heave_amp: float = 0.0
    heave_period: float = 1.0
    flap_amp_deg: float = 0.0
    flap_period: float = 1.0

    def apply(self, grid: np.ndarray, t: float, surf: Surface, side: float) -> np.ndarray:
        g = grid.copy()
        if surf.flap and self.flap_amp_deg != 0.0:
            phi = math.radians(self.flap_amp_deg) * math.sin(2 * math.pi * t / self.flap_period)
            oy = side * surf.root_le[1] if side != 0.0 else 0.0
            oz = surf.root_le[2]
            y = g[..., 1] - oy
            z = g[..., 2] - oz
            s = np.where(y >= 0, 1.0, -1.0) if side == 0.0 else side
            ca, sa = math.cos(phi), math.sin(phi)
            g[..., 1] = oy + y * ca - s * z * sa
            g[..., 2] = oz + s * y * sa + z * ca
        if self.heave_amp != 0.0:
            g[..., 2] += self.heave_amp * math.sin(2 * math.pi * t / self.heave_period)
        return g


@dataclass
class Airplane:
    surfaces: list
    motion: Motion = field(default_factory=Motion)

    def _lattices(self):
        return [(g, side, s) for s in self.surfaces for g, side in s.lattices()]

    def grids(self, num_steps: int, dt: float) -> np.ndarray:
        """(num_steps, total_vertices, 3): every lattice's corners at t = s dt."""
        lat = self._lattices()
        out = np.empty((num_steps, sum(g.shape[0] * g.shape[1] for g, _, _ in lat), 3))
        for k in range(num_steps):
            t = k * dt
            out[k] = np.concatenate([self.motion.apply(g, t, s, side).reshape(-1, 3)
                                     for g, side, s in lat])
        return out

    def sizes(self):
        lat = self._lattices()
        return ([g.shape[0] - 1 for g, _, _ in lat], [g.shape[1] - 1 for g, _, _ in lat])

    def reference_area(self, index: int = 0) -> float:
        return self.surfaces[index].planform_area()


def freestream(V: float, alpha_deg: float = 0.0) -> np.ndarray:
    a = math.radians(alpha_deg)
    return np.array([V * math.cos(a), 0.0, V * math.sin(a)])
