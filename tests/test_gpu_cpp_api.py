"""Runs the C++ host-API test program (tests/cpp/test_host_api.cpp, built by
reyn_amd/examples/build.sh) and the headless dam-break example on the GPU."""

import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent.parent / "reyn_amd" / "bin"


def _run(args, timeout):
    exe = BIN / args[0]
    if not exe.exists():
        pytest.fail(f"{exe} is missing: run __graft_entry__.build()")
    return subprocess.run([str(exe), *args[1:]], capture_output=True, text=True, timeout=timeout)


def test_cpp_host_api():
    p = _run(["test_host_api"], 300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failed" in p.stdout


def test_cpp_dambreak_example():
    p = _run(["dambreak", "20", "30"], 300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "step   29" in p.stdout
