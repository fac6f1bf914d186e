"""Runs the C++ host-API test program (tests/cpp/test_host_api.cpp, built by
reyn_amd/examples/build.sh) and the headless dam-break example on the GPU."""

import subprocess

import numpy as np
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
    import os

    os.environ["REYN_TEST_GOLDEN"] = str(Path(__file__).resolve().parent / "golden")
    p = _run(["test_host_api"], 300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failed" in p.stdout


def test_cpp_dambreak_example(tmp_path):
    from reyn_amd.io import read_ply

    out = tmp_path / "run"
    p = _run(["dambreak", "20", "30", str(out)], 300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "step   29" in p.stdout
    rows = (tmp_path / "run.csv").read_text().splitlines()
    assert rows[0] == "step,dt_ms,iterations" and len(rows) == 31
    assert all(int(r.split(",")[2]) >= 3 for r in rows[1:])
    # the controller's steps: from 0.5 ms, at most 0.1 % larger each step, and
    # the fluid settles without blowing up (CFL never cuts dt)
    dts = [float(r.split(",")[1]) for r in rows[1:]]
    assert dts[0] == pytest.approx(0.5005, rel=1e-5)
    assert all(a * 0.999 * (1 - 1e-6) <= b <= a * 1.001 * (1 + 1e-6) for a, b in zip(dts, dts[1:]))
    x = read_ply(tmp_path / "run.ply")
    assert x.shape == (20 ** 3, 3) and np.isfinite(x).all()


def test_cpp_scene_main_example(tmp_path):
    """The reference's main loop on `reyn::Scene::from_obj` (examples/scene_main.cpp):
    a fluid cube inside a box cube, 8000 desired particles, 20 steps."""
    from reyn_amd.io import read_ply

    gold = Path(__file__).resolve().parent / "golden" / "cube.obj"
    lines = gold.read_text().splitlines()
    v = [ln for ln in lines if ln.startswith("v ")]
    f = [ln for ln in lines if ln.startswith("f ")]
    out = ["o box", *v, *f, "o fluid"]
    out += ["v " + " ".join(str(0.5 * float(t)) for t in ln.split()[1:]) for ln in v]
    out += ["f " + " ".join(str(int(t) + 8) for t in ln.split()[1:]) for ln in f]
    obj = tmp_path / "scene.obj"
    obj.write_text("\n".join(out) + "\n")
    res = tmp_path / "run"
    p = _run(["scene_main", str(obj), "8000", "20", str(res)], 300)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "scene: N=" in p.stdout and "step   19" in p.stdout
    n = int(p.stdout.split("scene: N=")[1].split(",")[0])
    assert 6000 < n <= 8000
    x = read_ply(tmp_path / "run.ply")
    assert x.shape == (n, 3) and np.isfinite(x).all()
    assert np.abs(x).max() <= 1.0 + 1e-6  # inside the box's boundary samples
