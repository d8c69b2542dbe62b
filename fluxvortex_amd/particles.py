"""Device-resident vortex particle field (reference: src/fluxvortex/particles.py:15-246).

State lives on the GPU as SoA columns (positions, circulations alpha, core
sizes, ages) of a fixed capacity; ``advect_rk3`` runs the low-storage RK3
with rVPM stretching, the velocity/Gamma/sigma clamps, NaN protection and the
Pedrizzetti relaxation entirely on the device.  The host views ``_pos``,
``_gamma``, ``_sigma`` and ``_age`` are copies of the first ``np`` particles.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class VortexParticleField:
    """Structure-of-arrays vortex particle field on the MI355X."""

    def __init__(self, max_particles=50000, nu=0.0, rlxf=0.3, kernel="gaussianerf",
                 precision="fp64", context=None, pair_rule="clamp", _handle=None,
                 _owner=None):
        """pair_rule: how particles closer than 1e-12 (the self pair, coincident
        particles) enter the sums -- "clamp" (default) with r = 1e-12, as the
        reference's CPU field (kernel.py:105,156); "skip" as its GPU twin
        (warp_vpm.py:37,103).  They differ only in the relaxation, where the
        clamped Gaussian self pair contributes a round-off-sized g at r = 1e-12
        divided by r^3 (particles.py:219)."""
        self.max_particles = int(max_particles)
        self.nu = float(nu)
        self.rlxf = float(rlxf)
        self.kernel = kernel
        self.precision = precision
        self._ctx = context or _lib.default_context()
        self._owned = _handle is None
        # a borrowed field (a hybrid case's) keeps its owner alive while in use
        self._owner = _owner
        if _handle is None:
            h = C.c_void_p()
            self._ctx.check(self._ctx.lib.fv_vpm_create(
                self._ctx.handle, self.max_particles, self.nu, self.rlxf,
                _lib.KERNELS[kernel], _lib.PRECISIONS[precision], C.byref(h)))
            self._h = h
        else:
            self._h = C.c_void_p(_handle)
        self.pair_rule = pair_rule
        self._ctx.check(self._ctx.lib.fv_vpm_set_pair_rule(self._h, _lib.PAIR_RULES[pair_rule]))

    def __del__(self):
        try:
            if self._owned and self._h and self._h.value:
                self._ctx.lib.fv_vpm_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    # -- particle management (particles.py:36-69) ----------------------
    @property
    def np(self) -> int:
        n = C.c_int64()
        self._ctx.check(self._ctx.lib.fv_vpm_count(self._h, C.byref(n)))
        return n.value

    def add_particle(self, X, Gamma, sigma):
        self.add_particles_batch(np.reshape(X, (1, 3)), np.reshape(Gamma, (1, 3)),
                                 np.reshape(sigma, (1,)))

    def add_particles_batch(self, positions, gammas, sigmas):
        p = _lib.f64(positions, (3,), "positions")
        g = _lib.f64(gammas, (3,), "gammas")
        s = _lib.f64(sigmas)
        n = p.shape[0]
        if g.shape[0] != n or s.shape != (n,):
            raise ValueError("positions, gammas and sigmas must describe the same particles")
        rc = self._ctx.lib.fv_vpm_add_host(self._h, _lib.ptr(p), _lib.ptr(g), _lib.ptr(s), n)
        if rc == _lib.FV_E_CAP:
            raise RuntimeError(self._ctx.lib.fv_last_error(self._ctx.handle).decode())
        self._ctx.check(rc)

    def _read(self):
        n = self.np
        pos = np.zeros((n, 3))
        gam = np.zeros((n, 3))
        sig = np.zeros(n)
        age = np.zeros(n)
        if n:
            P = _lib.ptr
            self._ctx.check(self._ctx.lib.fv_vpm_read_host(self._h, P(pos), P(gam), P(sig),
                                                           P(age)))
        return pos, gam, sig, age

    @property
    def positions(self):
        return self._read()[0]

    @property
    def gammas(self):
        return self._read()[1]

    @property
    def sigmas(self):
        return self._read()[2]

    _pos = positions
    _gamma = gammas
    _sigma = sigmas

    @property
    def _age(self):
        return self._read()[3]

    def state_dict(self):
        """Host snapshot of the particle state (checkpoint)."""
        pos, gam, sig, age = self._read()
        return {"pos": pos, "gamma": gam, "sigma": sig, "age": age}

    def load_state_dict(self, st):
        """Restores a snapshot taken by state_dict()."""
        pos = _lib.f64(st["pos"], (3,), "pos")
        gam = _lib.f64(st["gamma"], (3,), "gamma")
        sig = _lib.f64(st["sigma"])
        age = _lib.f64(st["age"])
        P = _lib.ptr
        self._ctx.check(self._ctx.lib.fv_vpm_write_host(self._h, pos.shape[0], P(pos), P(gam),
                                                        P(sig), P(age)))

    # -- induced velocity (particles.py:72-82) ---------------------------
    def induce_velocity_at(self, target_points):
        t = _lib.f64(target_points, (3,), "target_points")
        out = np.zeros_like(t)
        if t.shape[0] == 0:
            return out
        self._ctx.check(self._ctx.lib.fv_vpm_induce_host(self._h, _lib.ptr(t), t.shape[0],
                                                         _lib.ptr(out)))
        return out

    # -- RK3 + rVPM (particles.py:85-199) + relaxation (:202-246) ----------
    def advect_rk3(self, dt, U_inf_func, bound_velocity_func=None, stretch=True,
                   free_wake=True):
        """One RK3 step followed by Pedrizzetti relaxation, on the device.

        ``U_inf_func`` is a (3,) vector or a callable X (N, 3) -> (N, 3);
        ``bound_velocity_func`` (optional) a callable X -> (N, 3), the velocity
        the bound wing vortices induce.  A uniform freestream without a bound
        velocity runs entirely on the device; otherwise the callables are
        evaluated at every RK stage's positions, as the reference does
        (particles.py:126-133: U_total = U_part + U_bound + U_inf), each stage
        reading the positions back to the host."""
        uniform = None
        if not callable(U_inf_func):
            uniform = _lib.f64(U_inf_func).reshape(3)
        elif bound_velocity_func is None:
            pos = self.positions
            if pos.shape[0] == 0:
                return
            U = np.broadcast_to(np.asarray(U_inf_func(pos), dtype=np.float64), pos.shape)
            if np.all(U == U[0]):
                uniform = np.ascontiguousarray(U[0])
        if uniform is not None and bound_velocity_func is None:
            self._ctx.check(self._ctx.lib.fv_vpm_advect_rk3(self._h, float(dt), _lib.ptr(uniform),
                                                            1 if stretch else 0,
                                                            1 if free_wake else 0))
            return
        u_inf = (lambda X: uniform) if uniform is not None else U_inf_func
        errors = []

        def stage(_user, n, p, ub, ui):
            try:
                X = np.ctypeslib.as_array(p, shape=(n, 3)).copy()
                np.ctypeslib.as_array(ub, shape=(n, 3))[...] = (
                    0.0 if bound_velocity_func is None else bound_velocity_func(X))
                np.ctypeslib.as_array(ui, shape=(n, 3))[...] = u_inf(X)
                return 0
            except Exception as e:   # raised below, after the library returns
                errors.append(e)
                return 1

        cb = _lib.STAGE_VELOCITY_FN(stage)
        rc = self._ctx.lib.fv_vpm_advect_rk3_ext(self._h, float(dt),
                                                 C.cast(cb, C.c_void_p), None,
                                                 1 if stretch else 0, 1 if free_wake else 0)
        if errors:
            raise errors[0]
        self._ctx.check(rc)

    # -- learned far-field closure -------------------------------------------
    def enable_closure(self, mlp=None, cell=2.0, kappa=4.0, check_every=10, tol=1e-3,
                       train_samples=2048, train_iters=2, lr=1e-3):
        """Replaces the self sums' far field (cells beyond kappa * cell) by the
        first-order cell expansion plus the learned residual of `mlp`
        (a fluxvortex_amd.mlp.ClosureMLP, trained online); checked against the
        exact sum every `check_every` advects and rejected above `tol`."""
        self._closure_mlp = mlp        # keep the network alive with the field
        self._closure_on = True
        self._ctx.check(self._ctx.lib.fv_vpm_set_closure(
            self._h, mlp.handle if mlp is not None else None, float(cell), float(kappa),
            int(check_every), float(tol), int(train_samples), int(train_iters), float(lr)))

    def set_closure_local(self, kappa_local):
        """Cells beyond kappa_local * cell enter through a second-order local
        expansion at each target cell (0: off; else >= 2 kappa)."""
        self._ctx.check(self._ctx.lib.fv_vpm_set_closure_local(self._h, float(kappa_local)))

    def has_closure(self) -> bool:
        return getattr(self, "_closure_on", False)

    def disable_closure(self):
        self._ctx.check(self._ctx.lib.fv_vpm_clear_closure(self._h))
        self._closure_mlp = None
        self._closure_on = False

    def closure_stats(self):
        e, fe = C.c_double(), C.c_double()
        ch, rj = C.c_int64(), C.c_int64()
        act = C.c_int32()
        self._ctx.check(self._ctx.lib.fv_vpm_closure_stats(self._h, C.byref(e), C.byref(ch),
                                                           C.byref(rj), C.byref(act),
                                                           C.byref(fe)))
        npr = C.c_double()
        self._ctx.check(self._ctx.lib.fv_vpm_closure_near_pairs(self._h, C.byref(npr)))
        return {"last_err": e.value, "checks": ch.value, "rejections": rj.value,
                "active": bool(act.value), "far_evals": fe.value, "near_pairs": npr.value}

    def closure_velocity(self, use_mlp=True):
        """(closure velocity at the particles, its normwise error vs the exact sum)."""
        out = np.zeros((self.np, 3))
        err = C.c_double()
        self._ctx.check(self._ctx.lib.fv_vpm_closure_velocity_host(
            self._h, 1 if use_mlp else 0, _lib.ptr(out), C.byref(err)))
        return out, err.value

    def closure_sum(self, mode="velocity", kernel=None, use_mlp=True):
        """One closure pass at the particles: "velocity", "stretching" (J alpha),
        "vorticity" or "vorticity_clamped" (the Gaussian vorticity under the
        clamped pair rule); kernel defaults to the field's (the relaxation uses
        "gaussianerf")."""
        m = {"velocity": 0, "stretching": 2, "vorticity": 3, "vorticity_clamped": 4}[mode]
        k = _lib.KERNELS[kernel or self.kernel]
        out = np.zeros((self.np, 3))
        self._ctx.check(self._ctx.lib.fv_vpm_closure_sum_host(self._h, m, k, 1 if use_mlp else 0,
                                                              _lib.ptr(out)))
        return out

    def closure_train(self, rounds=1):
        self._ctx.check(self._ctx.lib.fv_vpm_closure_train(self._h, int(rounds)))

    # -- population control and the circulation ledger (newton_pc/adapters/flap.py:466-578)
    def drop_oldest(self, max_particles):
        """Keeps the newest max_particles particles (flap.py:469-473)."""
        self._ctx.check(self._ctx.lib.fv_vpm_drop_oldest(self._h, int(max_particles)))

    def merge(self, ref_points, eps=1e-3, vinf_norm=1.0, protect=64, protect_dist=0.0):
        """Pairwise moment-conserving merge (flap.py:478-563): aligned particle
        pairs in one distance-graded hash cell merge when the velocity change
        at ref_points stays below eps * vinf_norm; the newest `protect`
        particles (or, with protect_dist > 0, those within protect_dist of the
        reference centre) never merge.  Returns the number of merged pairs."""
        ref = _lib.f64(ref_points, (3,), "ref_points")
        n = C.c_int64()
        self._ctx.check(self._ctx.lib.fv_vpm_merge_host(self._h, _lib.ptr(ref), ref.shape[0],
                                                        float(eps), float(vinf_norm),
                                                        int(protect), float(protect_dist),
                                                        C.byref(n)))
        return n.value

    def circulation_ledger(self):
        """Total vector circulation sum(alpha) of the particles (flap.py:565-578)."""
        out = np.zeros(3)
        self._ctx.check(self._ctx.lib.fv_vpm_ledger_host(self._h, _lib.ptr(out)))
        return out

    def convert_rings(self, br, fr, fl, bl, strengths, sigma):
        """Appends the leg particles of rings with |Gamma| >= 1e-15 and returns the
        zeroed strengths and the number added
        (tests/experiment_hybrid_panel_particle.py:173-204)."""
        corners = [_lib.f64(c, (3,), "vertices") for c in (br, fr, fl, bl)]
        g = np.array(strengths, dtype=np.float64, copy=True)
        nr = g.shape[0]
        added = C.c_int64()
        P = _lib.ptr
        rc = self._ctx.lib.fv_convert_rings_host(self._h, nr, *(P(c) for c in corners), P(g),
                                                 float(sigma), C.byref(added))
        if rc == _lib.FV_E_CAP:
            raise RuntimeError(self._ctx.lib.fv_last_error(self._ctx.handle).decode())
        self._ctx.check(rc)
        return g, added.value
