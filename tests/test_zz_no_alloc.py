"""fv_vortex_collapsed_dev_noalloc (the collapsed segment sum with its scratch in
a file-scope device array) against fv_vortex_collapsed_dev through the C-ABI:
two calls in a row on device buffers, the second smaller; a call large enough
that the scratch caps the source split; the size limit's error."""
import ctypes as C

import numpy as np
import pytest

from helpers import relerr, sheet_rings

pytestmark = pytest.mark.gpu

FN = "fv_vortex_collapsed_dev_noalloc"


@pytest.fixture(scope="module")
def lib(ctx):
    L = ctx.lib
    dp = C.c_void_p
    getattr(L, FN).argtypes = [C.c_void_p, dp, C.c_int64, dp, dp, dp, dp, C.c_int64, dp, dp, dp,
                               C.c_double, C.c_int32, C.c_int32, dp]
    getattr(L, FN).restype = C.c_int
    L.fv_vortex_collapsed_dev.argtypes = getattr(L, FN).argtypes
    L.fv_vortex_collapsed_dev.restype = C.c_int
    return L


class Dev:
    """Device copies of host arrays through the library's own allocator."""

    def __init__(self, ctx):
        self.ctx, self.ptrs = ctx, []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        self.ctx.check(self.ctx.lib.fv_device_alloc(self.ctx.handle, max(a.nbytes, 8),
                                                    C.byref(p)))
        self.ptrs.append(p)
        if a.nbytes:
            self.ctx.check(self.ctx.lib.fv_memcpy_h2d(self.ctx.handle, p, a.ctypes.data_as(
                C.c_void_p), a.nbytes))
        return p

    def get(self, p, shape):
        out = np.zeros(shape)
        self.ctx.check(self.ctx.lib.fv_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(C.c_void_p),
                                                  p, out.nbytes))
        return out

    def free(self):
        for p in self.ptrs:
            self.ctx.lib.fv_device_free(self.ctx.handle, p)
        self.ptrs = []


def _case(n, nx, ny, seed):
    br, fr, fl, bl, g, rc, age = sheet_rings(nx, ny, seed=seed)
    pts = np.random.default_rng(seed).uniform([0, -5, -1], [20, 5, 1], (n, 3))
    return pts, br, fr, fl, bl, g, rc, age


def _both(ctx, lib, args, prec=0, legs=0xF, nu=1.5e-5):
    pts, br, fr, fl, bl, g, rc, age = args
    d = Dev(ctx)
    try:
        dp = [d.put(a) for a in (pts, br, fr, fl, bl, g, rc, age)]
        n, m = len(pts), len(g)
        o1, o2 = d.put(np.zeros((n, 3))), d.put(np.full((n, 3), np.nan))
        ctx.check(lib.fv_vortex_collapsed_dev(ctx.handle, dp[0], n, *dp[1:5], m, dp[5], dp[6],
                                              dp[7], nu, legs, prec, o1))
        ctx.check(getattr(lib, FN)(ctx.handle, dp[0], n, *dp[1:5], m, dp[5], dp[6], dp[7], nu,
                                   legs, prec, o2))
        ctx.synchronize()
        return d.get(o1, (n, 3)), d.get(o2, (n, 3))
    finally:
        d.free()


def test_noalloc_matches_allocating_entry_point_twice(ctx, lib):
    """Two calls in a row, the second smaller (the scratch keeps the first
    call's contents): each equals fv_vortex_collapsed_dev bit for bit (same
    tile shape and split), FP64 and mixed."""
    ctx.set_autotune(False)
    for prec in (0, 1):
        for n, nx, ny, seed in ((30000, 100, 50, 1), (777, 40, 25, 2)):
            print(f"[noalloc] prec={prec} n={n} m={nx * ny}", flush=True)
            ref, got = _both(ctx, lib, _case(n, nx, ny, seed), prec=prec)
            assert np.all(np.isfinite(got))
            assert np.array_equal(got, ref), relerr(got, ref)


def test_noalloc_capped_split_and_legs(ctx, lib):
    """200 000 targets: the partial sums of the default split would not fit
    the scratch, so the split is capped; the sum moves by round-off only.
    Chordwise and horseshoe leg masks, no ages."""
    ctx.set_autotune(False)
    args = _case(200000, 40, 50, 3)
    print("[noalloc] capped split n=200000 m=2000", flush=True)
    ref, got = _both(ctx, lib, args)
    assert relerr(got, ref) < 1e-13
    for legs in (0x5, 0x7):
        ref, got = _both(ctx, lib, _case(5000, 40, 25, 4), legs=legs)
        assert np.array_equal(got, ref)


def test_noalloc_rejects_oversize(ctx, lib):
    """More segments than the 32 MiB scratch holds: FV_E_ARG and a message,
    nothing launched."""
    m = 110000                       # 440 000 segments > 419 430
    d = Dev(ctx)
    try:
        z3 = d.put(np.zeros((m, 3)))
        z1 = d.put(np.zeros(m))
        pts = d.put(np.zeros((4, 3)))
        out = d.put(np.zeros((4, 3)))
        rc = getattr(lib, FN)(ctx.handle, pts, 4, z3, z3, z3, z3, m, z1, z1, None, 0.0, 0xF, 0,
                              out)
        assert rc == -1
        assert b"scratch" in ctx.lib.fv_last_error(ctx.handle)
    finally:
        d.free()
