"""Triangle meshes on the CPU: the OBJ loader on the reference's own meshes
(tests/golden/cube.obj and ico_sphere.obj, copied from the reference's
scenes/), polygons, negative and slashed indices, shape filters and errors;
areas and signed volumes against closed forms; the sample count formula."""

import math
from pathlib import Path

import numpy as np
import pytest

from reyn_amd import capi
from reyn_amd.mesh import Mesh

GOLD = Path(__file__).resolve().parent / "golden"


def test_cube():
    m = Mesh.load_obj(GOLD / "cube.obj")
    inf = m.info()
    assert inf["vertices"] == 8 and inf["faces"] == 12
    assert inf["area"] == 24.0 and inf["volume"] == 8.0  # a [-1, 1]^3 cube, exactly
    v, f = m.arrays()
    assert np.array_equal(np.abs(v), np.ones((8, 3)))
    assert f.min() == 0 and f.max() == 7


def test_ico_sphere():
    m = Mesh.load_obj(GOLD / "ico_sphere.obj")
    inf = m.info()
    assert inf["vertices"] == 642 and inf["faces"] == 1280
    v, _ = m.arrays()
    r = np.linalg.norm(v, axis=1)
    assert np.allclose(r, 1.0, atol=1e-5)
    # an inscribed polyhedron: a little less area and volume than the sphere
    assert 0.98 * 4 * math.pi < inf["area"] < 4 * math.pi
    assert 0.97 * 4 / 3 * math.pi < inf["volume"] < 4 / 3 * math.pi


def test_obj_syntax(tmp_path):
    p = tmp_path / "q.obj"
    p.write_text("# quad and triangle\n"
                 "o Keep\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
                 "f 1/1/1 2/1/1 3/1/1 4/1/1\n"
                 "g Skip_me\nv 0 0 1\nf -1 1//1 2\n")
    m = Mesh.load_obj(p)
    v, f = m.arrays()
    assert v.shape == (5, 3)
    # the quad as a fan from its first corner, then the triangle with a
    # negative (relative) index
    assert f.tolist() == [[0, 1, 2], [0, 2, 3], [4, 0, 1]]
    assert Mesh.load_obj(p, ignore=("skip",)).arrays()[1].tolist() == [[0, 1, 2], [0, 2, 3]]
    assert Mesh.load_obj(p, only="skip").arrays()[1].tolist() == [[4, 0, 1]]


@pytest.mark.parametrize("text", ["v 0 0 0\nv 1 0 0\nf 1 2\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",
                                  "v 0 0\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 0\n"])
def test_obj_errors(tmp_path, text):
    p = tmp_path / "bad.obj"
    p.write_text(text)
    with pytest.raises(capi.ReynError):
        Mesh.load_obj(p)


def test_missing_file(tmp_path):
    with pytest.raises(capi.ReynError):
        Mesh.load_obj(tmp_path / "none.obj")


def test_sample_count_formula():
    m = Mesh.load_obj(GOLD / "cube.obj")
    for h, over in ((0.1, 1.0), (0.05, 2.0), (0.033, 1.5)):
        hb = np.float32(np.float64(np.float32(h)) / np.float64(np.float32(over)))
        cell = np.float32(np.float32(1.5) * np.sqrt(np.float32(3.0))) * hb * hb
        assert m.sample_count(h, over) == math.ceil(24.0 / float(cell))


def test_mesh_from_arrays_rejects_bad_index():
    with pytest.raises(capi.ReynError):
        Mesh.from_arrays(np.zeros((3, 3)), np.array([[0, 1, 3]]))


# ---- the relaxation's restatement (oracle/mesh_sampling.py) pinned -----------------


def test_hex_rest_density_shells():
    """The 25 lattice points within q, r in [-2, 2] hold every hexagonal-lattice
    point inside the support 2 h_bdy: the centre, 6 at h_bdy, 6 at sqrt(3) h_bdy
    (the 6 at 2 h_bdy sit on the support's edge, where B3 vanishes)."""
    from oracle import binding, mesh_sampling as ms

    for h in (0.01, 0.1, 0.7):
        w, _ = binding.kernel_eval(ms.B3, 2 * h, np.array([[0, 0, 0], [h, 0, 0], [0, math.sqrt(3) * h, 0],
                                                            [2 * h, 0, 0]], float))
        assert w[3] <= 1e-15 * w[0]  # zero up to the float32 support radius
        assert ms.hex_rest_density(h) == pytest.approx(w[0] + 6 * w[1] + 6 * w[2], rel=1e-13)


def test_closest_points_cube_closed_form():
    """On the [-1, 1]^3 cube the closest surface point is the clamp to the box
    outside it and, inside, the point moved along its largest coordinate to
    that face."""
    from oracle import mesh_sampling as ms

    v, f = Mesh.load_obj(GOLD / "cube.obj").arrays()
    rng = np.random.default_rng(5)
    p = rng.uniform(-1.4, 1.4, (4000, 3))
    a = np.sort(np.abs(p), axis=1)
    p = p[(a[:, 2] - a[:, 1] > 1e-6)]  # off the interior's ties
    want = np.clip(p, -1.0, 1.0)
    inner = np.abs(p).max(1) < 1.0
    k = np.abs(p[inner]).argmax(1)
    q = want[inner]
    q[np.arange(q.shape[0]), k] = np.sign(p[inner][np.arange(q.shape[0]), k])
    want[inner] = q
    got = ms.closest_points(p, v, f)
    assert np.abs(got - want).max() < 1e-12


def test_closest_points_icosphere_brute_force():
    """Against the nearest of a dense barycentric sampling of every face: never
    farther, and within the sampling's resolution."""
    from oracle import mesh_sampling as ms

    v, f = Mesh.load_obj(GOLD / "ico_sphere.obj").arrays()
    rng = np.random.default_rng(6)
    d = rng.normal(size=(40, 3))
    p = d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0.8, 1.2, (40, 1))
    got = ms.closest_points(p, v, f)
    n = 24
    s, t = np.meshgrid(np.arange(n + 1) / n, np.arange(n + 1) / n, indexing="ij")
    keep = (s + t) <= 1.0
    s, t = s[keep], t[keep]
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    dense = (a[:, None] + s[None, :, None] * (b - a)[:, None] + t[None, :, None] * (c - a)[:, None]).reshape(-1, 3)
    from scipy.spatial import cKDTree

    dist, _ = cKDTree(dense).query(p)
    mine = np.linalg.norm(got - p, axis=1)
    edge = np.linalg.norm(b - a, axis=1).max()
    assert np.all(mine <= dist + 1e-12)
    assert np.all(dist - mine <= edge / n)


def test_number_density_brute_force():
    from oracle import binding, mesh_sampling as ms

    rng = np.random.default_rng(7)
    x = rng.uniform(0, 0.5, (300, 3))
    h = 0.06
    nd, _, _ = ms.number_density(x, h)
    d = (x[:, None, :] - x[None, :, :]).reshape(-1, 3)
    w, _ = binding.kernel_eval(ms.B3, 2 * h, d)
    assert np.allclose(nd, w.reshape(300, 300).sum(1), rtol=1e-12)


def test_relax_shift_symmetry_and_direction():
    """The symmetric form conserves the mean position (the pair terms cancel),
    and on a line of three samples the middle one stays while the outer two
    move outwards, away from the denser middle."""
    from oracle import mesh_sampling as ms

    rng = np.random.default_rng(8)
    x = rng.uniform(0, 0.3, (200, 3))
    xs, _ = ms.relax_shift(x, 0.05, 1.0)
    assert np.abs((xs - x).sum(0)).max() < 1e-12
    assert np.abs(xs - x).max() > 1e-6
    line = np.array([[-0.06, 0, 0], [0, 0, 0], [0.06, 0, 0]])
    s, _ = ms.relax_shift(line, 0.05, 1.0)
    d = s - line
    assert np.abs(d[1]).max() < 1e-15 and d[0, 0] < 0 < d[2, 0]
    assert np.abs(d[:, 1:]).max() == 0.0
