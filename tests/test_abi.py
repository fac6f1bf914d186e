"""The C-ABI library loads without a GPU and exports exactly what
include/reyn_amd.h declares; the ctypes binding declares every entry point."""

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from reyn_amd import capi

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "reyn_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:int|const char\*|void\*)\s+(reyn_[a-z0-9_]+)\s*\(", text, flags=re.M))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_library_loads_without_gpu():
    lib = capi.lib()
    assert lib.reyn_abi_version() == capi.ABI_VERSION


def test_exports_match_header():
    decl = declared_symbols()
    assert len(decl) >= 40
    exp = {s for s in exported_symbols() if s.startswith("reyn_")}
    assert decl == exp, (decl - exp, exp - decl)
    assert not {s for s in exported_symbols() if not s.startswith("reyn_")}


def test_binding_covers_every_entry_point():
    assert set(capi.EXPORTED) == declared_symbols()


def test_no_context_calls_fail_cleanly():
    import ctypes as C

    lib = capi.lib()
    assert lib.reyn_ctx_synchronize(None) == capi.ERR_INVALID_ARGUMENT
    assert lib.reyn_last_error(None) == b"no context"
    d = C.c_double()
    assert lib.reyn_pcisph_delta(7, 0.1, 0.05, C.byref(d)) == capi.ERR_INVALID_ARGUMENT
    assert lib.reyn_pcisph_delta(capi.KERNEL_W6, 0.1, 0.05, C.byref(d)) == capi.OK and d.value > 0


# ctypes mirror -> C typedef of include/reyn_amd.h
STRUCTS = {
    "Float3": "reyn_float3", "Grid": "reyn_grid", "Particles": "reyn_particles", "Boundary": "reyn_boundary",
    "SphParams": "reyn_sph_params", "SesphParams": "reyn_sesph_params", "SesphWork": "reyn_sesph_work",
    "PcisphParams": "reyn_pcisph_params", "PcisphWork": "reyn_pcisph_work", "SolveInfo": "reyn_solve_info",
    "IisphParams": "reyn_iisph_params", "IisphWork": "reyn_iisph_work", "BatchSceneInfo": "reyn_batch_scene_info",
    "BatchScene": "reyn_batch_scene", "BatchParams": "reyn_batch_params", "DamBreak": "reyn_dam_break",
}


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes struct has the size and field offsets the C compiler gives
    the header's typedef (field names are the same on both sides)."""
    import shutil
    import subprocess

    from reyn_amd import capi

    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "reyn_amd.h"', "int main(void) {"]
    for py, c in STRUCTS.items():
        lines.append(f'printf("{py} size %zu\\n", sizeof({c}));')
        for name, _ in getattr(capi, py)._fields_:
            lines.append(f'printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = Path(__file__).resolve().parent.parent / "include"
    subprocess.run(["gcc", "-std=c11", f"-I{inc}", str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        py, field, v = line.split()
        got[(py, field)] = int(v)
    for py in STRUCTS:
        cls = getattr(capi, py)
        assert got[(py, "size")] == C.sizeof(cls), py
        for name, _ in cls._fields_:
            assert got[(py, name)] == getattr(cls, name).offset, (py, name)


def test_abi_version_matches_header():
    text = (ROOT / "include" / "reyn_amd.h").read_text()
    assert int(re.search(r"#define REYN_ABI_VERSION (\d+)", text).group(1)) == capi.ABI_VERSION
