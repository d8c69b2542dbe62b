"""The C-ABI library loads and exports every entry point include/fluxvortex.h
declares (no GPU needed, no compute calls)."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "fluxvortex.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fv_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_hot_path():
    names = declared_functions()
    for must in ("fv_vortex_collapsed_host", "fv_vortex_expanded_host", "fv_aic_dev",
                 "fv_dense_solve_dev", "fv_particle_velocity_host", "fv_vpm_advect_rk3",
                 "fv_convert_rings_host", "fv_hybrid_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from fluxvortex_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.fail("libfluxvortex_hip.so is not built (python -m fluxvortex_amd.build)")
    lib = _lib.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    unbound = [n for n in declared_functions() if n not in _lib._SIGNATURES]
    assert not unbound, unbound


def test_product_package_never_imports_the_oracle():
    pkg = ROOT / "fluxvortex_amd"
    for f in pkg.rglob("*.py"):
        src = f.read_text()
        assert "import oracle" not in src and "from oracle" not in src, f
        assert "fv_oracle" not in src, f
