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
