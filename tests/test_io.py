"""Result and checkpoint formats on the CPU: the PLY layout of the
reference's writer (src/scene/ply_io.cu:7-56) and checkpoint round trips."""

import numpy as np

from reyn_amd import scenes
from reyn_amd.io import STATE_FIELDS, Checkpoint, read_ply, write_ply


def test_ply_layout(tmp_path):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1000, 3)).astype(np.float32)
    path = tmp_path / "p.ply"
    write_ply(path, x[:, 0], x[:, 1], x[:, 2])
    raw = path.read_bytes()
    head = raw[:raw.index(b"end_header\n")].decode("ascii").splitlines()
    assert head[0] == "ply" and head[1] in ("format binary_little_endian 1.0", "format binary_big_endian 1.0")
    assert "element vertex 1000" in head
    assert [h for h in head if h.startswith("property")] == ["property float32 x", "property float32 y",
                                                            "property float32 z"]
    # body: x_i, y_i, z_i per particle, 12 bytes each
    assert len(raw) - (raw.index(b"end_header\n") + 11) == 12 * 1000
    assert np.array_equal(read_ply(path), x)


def test_checkpoint_round_trip(tmp_path):
    sc = scenes.config5_jobs(1, seed=2, n_lo=500, n_hi=900)[0][0]
    rng = np.random.default_rng(3)
    state = {k: rng.standard_normal(sc.n).astype(np.float32) for k in STATE_FIELDS}
    ck = Checkpoint(sc, state, 17, 6, 4.9e-4, 0.0083, 0.0009)
    ck.save(tmp_path / "c.npz")
    back = Checkpoint.load(tmp_path / "c.npz")
    assert (back.steps, back.iterations, back.time, back.loop_stat) == (17, 6, 0.0083, 0.0009)
    assert np.float32(back.dt) == np.float32(4.9e-4)
    for k in STATE_FIELDS:
        assert np.array_equal(back.state[k], state[k])
    assert np.array_equal(back.scene.xb, sc.xb) and back.scene.spacing == sc.spacing
    assert back.scene.mass == sc.mass and np.array_equal(back.scene.bound_min, sc.bound_min)


def test_step_trace_round_trip(tmp_path):
    from reyn_amd.telemetry import StepTrace

    t = StepTrace()
    for k in range(1, 6):
        t.record(k, 1e-3 * k, 5e-4 + 1e-7 * k, 3 + k % 2, 1e-3 / k)
    t.write_csv(tmp_path / "t.csv")
    back = StepTrace.read_csv(tmp_path / "t.csv")
    assert back == t
    s = t.summary()
    assert s["steps"] == 5 and s["iterations_max"] == 4 and s["dt_max"] == 5e-4 + 5e-7
    assert abs(t.dt_ms[0] - 0.5001) < 1e-12
