"""Builds the gfx950 HIP library ``reyn_amd/libreyn_amd.so`` in-tree.

Every ``csrc/*.hip`` file is compiled to an object in parallel with hipcc,
then linked into one shared library whose only exported symbols are the
``extern "C"`` entry points declared in ``include/reyn_amd.h``.

Usage: ``python -m reyn_amd.build`` (or ``build()`` from Python).
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libreyn_amd.so"

ARCH = os.environ.get("REYN_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-fvisibility=hidden",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build reyn_amd)")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(
            "command failed:\n  " + " ".join(cmd) + "\n" + proc.stdout + proc.stderr
        )


def build(verbose: bool = False) -> Path:
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    OBJ.mkdir(exist_ok=True)
    cc = hipcc()
    jobs = []
    for src in sources:
        obj = OBJ / (src.stem + ".o")
        jobs.append([cc, *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for fut in [ex.submit(_run, j) for j in jobs]:
            fut.result()
    tmp = LIB.with_suffix(".so.tmp")
    link = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)]
    link += [str(OBJ / (s.stem + ".o")) for s in sources]
    _run(link)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
