"""The C-ABI library loads without a GPU and exports exactly what
include/reyn_amd.h declares; the ctypes binding declares every entry point."""

import re
import subprocess
from pathlib import Path

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
    assert lib.reyn_abi_version() == 1


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
