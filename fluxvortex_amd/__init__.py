"""fluxvortex_amd -- MI355X (gfx950) hot path of FLUXVortex's hybrid
panel-particle solver, behind the reference's own surfaces:

* ``patch()`` / ``unpatch()`` / ``benchmark()``: swap PteraSoftware's five Numba
  Biot-Savart routines for HIP kernels (reference src/fluxvortex/warp_patch.py);
* :mod:`fluxvortex_amd.vortex_kernels`: the five routines themselves, plus the
  fused AIC and the dense circulation solve;
* :mod:`fluxvortex_amd.kernel` / :mod:`fluxvortex_amd.particles`: particle
  Biot-Savart sums and the device-resident ``VortexParticleField``;
* :mod:`fluxvortex_amd.hybrid`: the HybridSolver-equivalent step driver.

All compute runs in libfluxvortex_hip.so; there is no CPU fallback.
"""
from . import vortex_kernels  # noqa: F401
from .ptera_patch import benchmark, is_patched, patch, unpatch  # noqa: F401
from .vortex_kernels import get_precision, set_precision  # noqa: F401

__version__ = "0.1.0"
