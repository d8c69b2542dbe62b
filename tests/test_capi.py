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


def _c_layout(tmp_path, include, type_name, fields, extra=()):
    """sizeof and offsetof of each field of a C struct, from gcc (extra:
    sources linked beside the probe)."""
    import subprocess
    src = tmp_path / f"layout_{type_name}.c"
    lines = [include, "#include <stddef.h>", "#include <stdio.h>", "int main(void) {",
             f'  printf("%zu\\n", sizeof({type_name}));']
    for f in fields:
        lines.append(f'  printf("%zu\\n", offsetof({type_name}, {f}));')
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / f"layout_{type_name}"
    subprocess.run(["gcc", "-std=gnu11", "-fopenmp", "-o", str(exe), str(src),
                    *[str(e) for e in extra], "-lm"], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), [int(v) for v in out[1:]]


def _check_struct(tmp_path, include, type_name, cls, extra=()):
    import ctypes as C
    names = [f[0] for f in cls._fields_]
    size, offs = _c_layout(tmp_path, include, type_name, names, extra)
    assert C.sizeof(cls) == size, (type_name, C.sizeof(cls), size)
    for n, o in zip(names, offs):
        assert getattr(cls, n).offset == o, (type_name, n, getattr(cls, n).offset, o)


def test_ctypes_structs_match_the_c_abi(tmp_path):
    """The ctypes mirrors of the C-ABI's descriptor structs (fv_hybrid_desc,
    fv_sweep_desc) and of the oracle's run descriptors have the C layout:
    every field at the offset gcc gives it and the same size."""
    import sys
    from fluxvortex_amd import _lib
    from fluxvortex_amd.aeroelastic import SweepDesc
    inc = f'#include "{ROOT / "include" / "fluxvortex.h"}"'
    _check_struct(tmp_path, inc, "fv_hybrid_desc", _lib.HybridDesc)
    _check_struct(tmp_path, inc, "fv_sweep_desc", SweepDesc)
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    uvlm = ROOT / "oracle" / "fv_oracle_uvlm.c"
    inc = f'#include "{uvlm}"'
    rest = [c for c in sorted((ROOT / "oracle").glob("*.c")) if c != uvlm]
    _check_struct(tmp_path, inc, "orc_hybrid_desc", oracle._HybridDesc, rest)
    _check_struct(tmp_path, inc, "orc_beam", oracle._OrcBeam, rest)
