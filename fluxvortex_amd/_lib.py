"""ctypes binding of libfluxvortex_hip.so (the C-ABI declared in include/fluxvortex.h).

The library is the only compute path of this package: if it is missing, or no
HIP device is visible, every entry point raises RuntimeError.  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libfluxvortex_hip.so"

FV_OK = 0
FV_E_ARG = -1
FV_E_HIP = -2
FV_E_CAP = -3
FV_E_SING = -4

KERNEL_GAUSSIAN_ERF = 0
KERNEL_WINCKELMANS = 1
KERNELS = {"gaussianerf": KERNEL_GAUSSIAN_ERF, "winckelmans": KERNEL_WINCKELMANS}

PREC_FP64 = 0
PREC_MIXED = 1
PRECISIONS = {"fp64": PREC_FP64, "mixed": PREC_MIXED}
# particles closer than 1e-12: "skip" as warp_vpm.py:37,103; "clamp" r to 1e-12
# as the reference's CPU field (kernel.py:105,156), which HybridSolver runs
PAIR_RULES = {"skip": 0, "clamp": 1}

# fv_stage_velocity_fn: int (*)(void *user, int64_t n, const double *pos,
#                               double *u_bound, double *u_inf)
STAGE_VELOCITY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.POINTER(C.c_double))

LEGS_RING = 0xF
LEGS_CHORDWISE = 0x5
LEGS_HORSESHOE = 0x7

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p
_ctx = C.c_void_p


class HybridDesc(C.Structure):
    """Mirror of fv_hybrid_desc."""

    _fields_ = [
        ("n_surfaces", C.c_int32),
        ("nc", _i32p),
        ("ns", _i32p),
        ("num_steps", C.c_int64),
        ("grids", _dp),
        ("vinf", C.c_double * 3),
        ("rho", C.c_double),
        ("dt", C.c_double),
        ("nu", C.c_double),
        ("rc0_bound", C.c_double),
        ("rc0_wake", C.c_double),
        ("n_keep", C.c_int32),
        ("sigma", C.c_double),
        ("kernel", C.c_int32),
        ("free_wake", C.c_int32),
        ("stretch", C.c_int32),
        ("rlxf", C.c_double),
        ("vpm_nu", C.c_double),
        ("sref", C.c_double),
        ("max_particles", C.c_int64),
        ("precision", C.c_int32),
    ]


_SIGNATURES = {
    "fv_ctx_create": (C.c_int, [C.c_int, C.POINTER(_ctx)]),
    "fv_ctx_destroy": (None, [_ctx]),
    "fv_last_error": (C.c_char_p, [_ctx]),
    "fv_ctx_synchronize": (C.c_int, [_ctx]),
    "fv_ctx_stream": (_vp, [_ctx]),
    "fv_device_alloc": (C.c_int, [_ctx, C.c_size_t, C.POINTER(_vp)]),
    "fv_device_free": (C.c_int, [_ctx, _vp]),
    "fv_memcpy_h2d": (C.c_int, [_ctx, _vp, _vp, C.c_size_t]),
    "fv_memcpy_d2h": (C.c_int, [_ctx, _vp, _vp, C.c_size_t]),
    "fv_ctx_set_autotune": (C.c_int, [_ctx, C.c_int32]),
    "fv_autotune_report": (C.c_int, [_ctx, C.c_char_p, C.c_int64]),
    "fv_ctx_set_timing": (C.c_int, [_ctx, C.c_int]),
    "fv_ctx_kernel_stats": (C.c_int, [_ctx, C.c_char_p, _i64p, _dp]),
    "fv_ctx_event_record": (C.c_int, [_ctx, C.c_int32]),
    "fv_ctx_event_elapsed": (C.c_int, [_ctx, C.c_int32, C.c_int32, _dp]),
    "fv_vortex_collapsed_host": (C.c_int, [_ctx, _dp, C.c_int64, _dp, _dp, _dp, _dp, C.c_int64,
                                           _dp, _dp, _dp, C.c_double, C.c_int32, C.c_int32, _dp]),
    "fv_vortex_collapsed_dev": (C.c_int, [_ctx, _vp, C.c_int64, _vp, _vp, _vp, _vp, C.c_int64,
                                          _vp, _vp, _vp, C.c_double, C.c_int32, C.c_int32, _vp]),
    "fv_vortex_collapsed_dev_noalloc": (C.c_int, [_ctx, _vp, C.c_int64, _vp, _vp, _vp, _vp,
                                                  C.c_int64, _vp, _vp, _vp, C.c_double,
                                                  C.c_int32, C.c_int32, _vp]),
    "fv_vortex_expanded_host": (C.c_int, [_ctx, _dp, C.c_int64, _dp, _dp, _dp, _dp, C.c_int64,
                                          _dp, _dp, _dp, C.c_double, C.c_int32, _dp]),
    "fv_vortex_expanded_dev": (C.c_int, [_ctx, _vp, C.c_int64, _vp, _vp, _vp, _vp, C.c_int64,
                                         _vp, _vp, _vp, C.c_double, C.c_int32, _vp]),
    "fv_aic_host": (C.c_int, [_ctx, _dp, _dp, C.c_int64, _dp, _dp, _dp, _dp, _dp, C.c_double,
                              _dp]),
    "fv_aic_dev": (C.c_int, [_ctx, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp, _vp, C.c_double,
                             _vp]),
    "fv_dense_solve_host": (C.c_int, [_ctx, C.c_int64, C.c_int64, _dp, _dp, _i32p]),
    "fv_dense_solve_dev": (C.c_int, [_ctx, C.c_int64, C.c_int64, _vp, _vp, _vp]),
    "fv_dense_solve_mixed_dev": (C.c_int, [_ctx, C.c_int64, C.c_int64, _vp, _vp, C.c_int32,
                                           C.c_double, _vp, _vp]),
    "fv_dense_solve_mixed_host": (C.c_int, [_ctx, C.c_int64, C.c_int64, _dp, _dp, C.c_int32,
                                           C.c_double, _i32p, _i32p]),
    "fv_particle_velocity_host": (C.c_int, [_ctx, _dp, C.c_int64, _dp, _dp, _dp, C.c_int64,
                                            C.c_int32, C.c_int32, _dp]),
    "fv_particle_jacobian_host": (C.c_int, [_ctx, _dp, C.c_int64, _dp, _dp, _dp, C.c_int64,
                                            C.c_int32, C.c_int32, _dp]),
    "fv_vpm_create": (C.c_int, [_ctx, C.c_int64, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                C.POINTER(_vp)]),
    "fv_vpm_destroy": (None, [_vp]),
    "fv_vpm_set_pair_rule": (C.c_int, [_vp, C.c_int32]),
    "fv_vpm_add_host": (C.c_int, [_vp, _dp, _dp, _dp, C.c_int64]),
    "fv_vpm_count": (C.c_int, [_vp, _i64p]),
    "fv_vpm_read_host": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "fv_vpm_write_host": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp]),
    "fv_vpm_induce_host": (C.c_int, [_vp, _dp, C.c_int64, _dp]),
    "fv_vpm_advect_rk3": (C.c_int, [_vp, C.c_double, _dp, C.c_int32, C.c_int32]),
    "fv_vpm_advect_rk3_ext": (C.c_int, [_vp, C.c_double, C.c_void_p, C.c_void_p, C.c_int32,
                                        C.c_int32]),
    "fv_convert_rings_host": (C.c_int, [_vp, C.c_int64, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                        _i64p]),
    "fv_vpm_drop_oldest": (C.c_int, [_vp, C.c_int64]),
    "fv_vpm_merge_host": (C.c_int, [_vp, _dp, C.c_int32, C.c_double, C.c_double, C.c_int64,
                                    C.c_double, _i64p]),
    "fv_vpm_merge_stats": (C.c_int, [_vp, _i64p, _i64p]),
    "fv_vpm_ledger_host": (C.c_int, [_vp, _dp]),
    "fv_hybrid_create": (C.c_int, [_ctx, C.POINTER(HybridDesc), C.POINTER(_vp)]),
    "fv_hybrid_destroy": (None, [_vp]),
    "fv_hybrid_step": (C.c_int, [_vp, C.c_int64]),
    "fv_hybrid_current_step": (C.c_int, [_vp, _i64p]),
    "fv_hybrid_read": (C.c_int, [_vp, _dp, _dp, _dp]),
    "fv_hybrid_particles": (_vp, [_vp]),
    "fv_hybrid_set_population": (C.c_int, [_vp, C.c_int32, C.c_int64, C.c_double, C.c_int64,
                                           C.c_double]),
    "fv_hybrid_ledger_host": (C.c_int, [_vp, _dp]),
    "fv_hybrid_set_ring_wake": (C.c_int, [_vp, C.c_int32, C.c_double]),
    "fv_hybrid_pair_counts": (C.c_int, [_vp, _dp, _dp]),
    "fv_hybrid_set_flags": (C.c_int, [_vp, C.c_int32, C.c_int32]),
    "fv_hybrid_set_precision": (C.c_int, [_vp, C.c_int32]),
    "fv_hybrid_set_solve_reuse": (C.c_int, [_vp, C.c_int32]),
    "fv_hybrid_solve_reuse_count": (C.c_int, [_vp, _i64p]),
    "fv_hybrid_set_graph": (C.c_int, [_vp, C.c_int32]),
    "fv_hybrid_graph_stats": (C.c_int, [_vp, _i64p, _i64p]),
    "fv_hybrid_batch_create": (C.c_int, [C.POINTER(_vp), C.c_int32, C.POINTER(_vp)]),
    "fv_hybrid_batch_step": (C.c_int, [_vp, C.c_int64]),
    "fv_hybrid_batch_stats": (C.c_int, [_vp, _i64p, _i64p]),
    "fv_hybrid_batch_destroy": (None, [_vp]),
    "fv_vpm_set_closure": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_int32, C.c_double,
                                     C.c_int32, C.c_int32, C.c_double]),
    "fv_vpm_set_closure_local": (C.c_int, [_vp, C.c_double]),
    "fv_vpm_clear_closure": (C.c_int, [_vp]),
    "fv_vpm_closure_stats": (C.c_int, [_vp, _dp, _i64p, _i64p, _i32p, _dp]),
    "fv_vpm_closure_near_pairs": (C.c_int, [_vp, _dp]),
    "fv_vpm_closure_velocity_host": (C.c_int, [_vp, C.c_int32, _dp, _dp]),
    "fv_vpm_closure_sum_host": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, _dp]),
    "fv_sweep_create": (C.c_int, [_ctx, _vp, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "fv_sweep_destroy": (None, [_vp]),
    "fv_sweep_step": (C.c_int, [_vp, C.c_int64]),
    "fv_sweep_read": (C.c_int, [_vp, C.c_int32, _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_int64),
                                C.POINTER(C.c_int32)]),
    "fv_vpm_closure_train": (C.c_int, [_vp, C.c_int32]),
    "fv_mlp_create": (C.c_int, [_ctx, C.POINTER(_vp)]),
    "fv_mlp_destroy": (None, [_vp]),
    "fv_mlp_param_count": (C.c_int, []),
    "fv_mlp_set_params": (C.c_int, [_vp, _fp]),
    "fv_mlp_get_params": (C.c_int, [_vp, _fp]),
    "fv_mlp_forward_dev": (C.c_int, [_vp, _vp, C.c_int64, _vp]),
    "fv_mlp_gradients_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int64, _vp]),
    "fv_mlp_adam_dev": (C.c_int, [_vp, _vp, C.c_float]),
    "fv_mlp_train_step_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int64, C.c_float]),
    "fv_mlp_forward_host": (C.c_int, [_vp, _fp, C.c_int64, _fp]),
    "fv_mlp_gradients_host": (C.c_int, [_vp, _fp, _fp, _fp, C.c_int64, _fp]),
    "fv_mlp_adam_host": (C.c_int, [_vp, _fp, C.c_float]),
}

_lib = None
_lib_lock = threading.Lock()


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Loads the HIP library and declares the C-ABI signatures (no GPU needed)."""
    global _lib
    with _lib_lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise RuntimeError(
                f"{p.name} is not built ({p}); run `python -m fluxvortex_amd.build` "
                "(fluxvortex_amd has no CPU fallback)")
        lib = C.CDLL(str(p))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


class FluxVortexError(RuntimeError):
    """A negative status returned by libfluxvortex_hip.so."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[fv {code}] {msg}")
        self.code = code


class Context:
    """One HIP stream + workspaces on one device (fv_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _ctx()
        rc = self.lib.fv_ctx_create(int(device), C.byref(h))
        if rc != FV_OK or not h.value:
            raise RuntimeError(
                f"fv_ctx_create(device={device}) failed with {rc}: no usable HIP device "
                "(fluxvortex_amd has no CPU fallback)")
        self.handle = h
        self.device = device

    def check(self, rc: int) -> None:
        if rc != FV_OK:
            msg = self.lib.fv_last_error(self.handle)
            raise FluxVortexError(rc, msg.decode() if msg else "unknown error")

    def synchronize(self) -> None:
        self.check(self.lib.fv_ctx_synchronize(self.handle))

    def set_timing(self, enabled: bool) -> None:
        self.check(self.lib.fv_ctx_set_timing(self.handle, 1 if enabled else 0))

    def set_autotune(self, enabled: bool) -> None:
        """Pair-sum tile shapes timed at first use per kind (see fluxvortex.h)."""
        self.check(self.lib.fv_ctx_set_autotune(self.handle, 1 if enabled else 0))

    def autotune_report(self) -> list:
        """[(family, log2_targets, log2_sources, lanes, targets_per_lane,
        workgroups_per_cu, best_ms)] of the tuned kinds."""
        buf = C.create_string_buffer(1 << 16)
        self.check(self.lib.fv_autotune_report(self.handle, buf, len(buf)))
        out = []
        for line in buf.value.decode().splitlines():
            f = line.split()
            out.append((f[0], int(f[1]), int(f[2]), int(f[3]), int(f[4]), int(f[5]), float(f[6])))
        return out

    def kernel_stats(self, family: str) -> tuple[int, float]:
        n = C.c_int64()
        ms = C.c_double()
        self.check(self.lib.fv_ctx_kernel_stats(self.handle, family.encode(), C.byref(n),
                                                C.byref(ms)))
        return n.value, ms.value

    def event_record(self, which: int) -> None:
        self.check(self.lib.fv_ctx_event_record(self.handle, which))

    def event_elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_double()
        self.check(self.lib.fv_ctx_event_elapsed(self.handle, a, b, C.byref(ms)))
        return ms.value

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.fv_ctx_destroy(self.handle)
            self.handle = _ctx()


_default_ctx: Context | None = None
_default_lock = threading.Lock()


def default_context() -> Context:
    """Process-wide context on device FLUXVORTEX_DEVICE (default 0)."""
    global _default_ctx
    with _default_lock:
        if _default_ctx is None:
            _default_ctx = Context(int(os.environ.get("FLUXVORTEX_DEVICE", "0")))
        return _default_ctx


def f64(a, shape_tail: tuple[int, ...] | None = None, name: str = "array") -> np.ndarray:
    """C-contiguous float64 copy/view with an optional trailing-shape check."""
    arr = np.ascontiguousarray(a, dtype=np.float64)
    if shape_tail is not None:
        if arr.ndim != 1 + len(shape_tail) or tuple(arr.shape[1:]) != shape_tail:
            raise ValueError(f"{name} must have shape (n, {', '.join(map(str, shape_tail))}), "
                             f"got {arr.shape}")
    return arr


def ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def optr(a):
    return None if a is None else a.ctypes.data_as(_dp)
