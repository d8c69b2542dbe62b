"""Module-name alias of :mod:`fluxvortex_amd.ptera_patch` for scripts written
against the reference's ``fluxvortex.warp_patch`` (src/fluxvortex/warp_patch.py)."""
from .ptera_patch import _save_originals, benchmark, patch, unpatch  # noqa: F401
