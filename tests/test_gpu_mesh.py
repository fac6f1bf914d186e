"""Meshes on the GPU: surface samples equal the numpy restatement bit for bit
(oracle/mesh_sampling.py) and lie on the surface, the inside test separates
the icosphere's interior from its exterior, a dam break with a mesh
obstacle steps, and the relaxation of the samples matches its float64
restatement, keeps them on the surface and evens out their number density."""

import math
from pathlib import Path

import numpy as np
import pytest

from oracle import mesh_sampling
from reyn_amd.mesh import Mesh, obstacle_scene
from reyn_amd.sim import Batch

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name,h,over", [("cube.obj", 0.05, 1.0), ("ico_sphere.obj", 0.02, 1.5)])
def test_surface_samples_match_restatement(ctx, name, h, over):
    m = Mesh.load_obj(GOLD / name)
    v, f = m.arrays()
    got = m.sample(ctx, h, over, seed=77)
    want = mesh_sampling.sample(v, f, h, over, 77)
    print(f"{name}: {got.shape[0]} samples", flush=True)
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    # stratification: every face's share of the samples is close to its share
    # of the area
    if name == "ico_sphere.obj":
        r = np.linalg.norm(got.astype(np.float64), axis=1)
        assert r.max() <= 1.0 + 1e-6 and r.min() > 0.98  # on the inscribed faces
    else:
        assert np.allclose(np.abs(got).max(axis=1), 1.0, atol=1e-6)  # on a cube face


def test_inside_icosphere(ctx):
    m = Mesh.load_obj(GOLD / "ico_sphere.obj")
    rng = np.random.default_rng(5)
    p = rng.uniform(-1.5, 1.5, size=(20000, 3)).astype(np.float32)
    r = np.linalg.norm(p.astype(np.float64), axis=1)
    clear = (r < 0.97) | (r > 1.01)  # away from the faces
    ins = m.inside(ctx, p)
    assert np.array_equal(ins[clear], (r < 1.0)[clear])


def test_mesh_obstacle_scene_steps(ctx):
    m = Mesh.load_obj(GOLD / "ico_sphere.obj")
    # a sphere of radius 0.1 inside the 0.6 x 0.4 x 0.4 fluid block
    sc = obstacle_scene(ctx, m, (30, 20, 20), 0.02, (1.6, 1.0, 1.0), offset=(0.3, 0.2, 0.2), scale=0.1)
    print(f"obstacle scene: {sc.n} fluid, {sc.n_bdy} boundary, {sc.meta}", flush=True)
    assert sc.meta["mesh_samples"] > 0
    # at least the obstacle's volume of fluid is gone, and no fluid is left
    # inside the polyhedron (its faces lie at radius 0.0995 and more)
    assert sc.meta["culled"] >= 0.97 * 4 / 3 * math.pi * 0.1 ** 3 / 0.02 ** 3
    assert (np.linalg.norm(sc.x.astype(np.float64) - (0.3, 0.2, 0.2), axis=1) > 0.098).all()
    b = Batch(ctx, [sc], solver="iisph", kernel="B3")
    b.step(5)
    d = b.download(0)
    assert all(np.isfinite(d[k]).all() for k in d)
    assert b.info(0).steps == 5
    b.close()


@pytest.mark.parametrize("name,h,over,iters,tol", [("cube.obj", 0.1, 1.0, 1, 1e-5), ("cube.obj", 0.1, 1.0, 3, 1e-4),
                                                    ("ico_sphere.obj", 0.05, 1.5, 2, 1e-4)])
def test_relax_matches_restatement(ctx, name, h, over, iters, tol):
    """`relax_sampling` against the float64 restatement: the same points (the
    library returns them in grid order; matched one to one by nearest point)
    and the same number-density ratio."""
    from scipy.spatial import cKDTree

    m = Mesh.load_obj(GOLD / name)
    v, f = m.arrays()
    pts = m.sample(ctx, h, over, seed=21)
    print(f"{name}: relax {pts.shape[0]} samples x {iters}", flush=True)
    got, ratio = m.relax(ctx, pts, h, over, iters)
    want, ratio_w = mesh_sampling.relax(pts, v, f, h, over, iters)
    d, idx = cKDTree(want).query(got.astype(np.float64))
    print(f"max deviation {d.max() / h:.2e} h, ratio {ratio} vs {ratio_w}", flush=True)
    assert np.unique(idx).size == got.shape[0]
    assert d.max() < tol * h
    assert ratio == pytest.approx(ratio_w, rel=1e-5)
    # the shift moved the samples at all
    assert cKDTree(pts).query(got)[0].max() > 1e-3 * h


def test_relax_spreads_and_stays_on_surface(ctx):
    m = Mesh.load_obj(GOLD / "cube.obj")
    pts = m.sample(ctx, 0.1, 1.0, seed=5)
    _, r1 = m.relax(ctx, pts, 0.1, 1.0, 1)
    got, r10 = m.relax(ctx, pts, 0.1, 1.0, 10)
    print(f"number-density ratio {r1:.3f} -> {r10:.3f}", flush=True)
    assert r10 < 0.8 * r1
    assert np.allclose(np.abs(got).max(axis=1), 1.0, atol=1e-6)
    assert np.abs(got).max() <= 1.0
    # deterministic: the same input gives the same bits
    again, r10b = m.relax(ctx, pts, 0.1, 1.0, 10)
    assert np.array_equal(got, again) and r10 == r10b
    # through sample(): the same as sampling, then relaxing
    assert np.array_equal(m.sample(ctx, 0.1, 1.0, seed=5, relaxation_iters=10), got)


# ---- the LBVH ---------------------------------------------------------------------


def torus(nu: int, nv: int, big: float = 1.0, small: float = 0.35) -> tuple[np.ndarray, np.ndarray]:
    """A closed, outward-oriented torus of 2 nu nv faces."""
    u = 2 * np.pi * np.arange(nu) / nu
    v = 2 * np.pi * np.arange(nv) / nv
    uu, vv = np.meshgrid(u, v, indexing="ij")
    verts = np.stack([(big + small * np.cos(vv)) * np.cos(uu), (big + small * np.cos(vv)) * np.sin(uu),
                      small * np.sin(vv)], axis=-1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    faces = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    m = Mesh.from_arrays(verts, faces)
    if m.info()["volume"] < 0:
        faces = faces[:, ::-1]
    m.close()
    return verts, np.ascontiguousarray(faces)


def _meshes():
    v, f = torus(120, 60)
    return {"cube": Mesh.load_obj(GOLD / "cube.obj"), "ico_sphere": Mesh.load_obj(GOLD / "ico_sphere.obj"),
            "torus": Mesh.from_arrays(v, f)}


def _check_tree(nodes, v, f):
    """The reference's LBVH checks (lbvh.cu:314-384): every box holds its
    children's boxes and each leaf box its face; every leaf and every internal
    node but the root is referenced exactly once."""
    left, right, leaf = nodes["left"], nodes["right"], nodes["leaf_face"]
    ib, lb = nodes["internal_box"], nodes["leaf_box"]
    nf = leaf.size
    assert np.array_equal(np.sort(leaf), np.arange(nf))
    ch = np.concatenate([left, right])
    is_int = (ch & 1) == 1
    assert np.array_equal(np.sort(ch[~is_int] >> 1), np.arange(nf))
    assert np.array_equal(np.sort(ch[is_int] >> 1), np.arange(1, nf - 1))
    for side in (left, right):
        box = np.where(((side & 1) == 1)[:, None], ib[np.minimum(side >> 1, max(nf - 2, 0))], lb[side >> 1])
        assert np.all(ib[:, :3] <= box[:, :3]) and np.all(ib[:, 3:] >= box[:, 3:])
    tri = v[f[leaf]]  # leaf k's vertices
    assert np.array_equal(lb[:, :3], tri.min(axis=1)) and np.array_equal(lb[:, 3:], tri.max(axis=1))


@pytest.mark.parametrize("name", ["cube", "ico_sphere", "torus"])
def test_lbvh_structure(ctx, name):
    from reyn_amd.mesh import LBVH

    m = _meshes()[name]
    v, f = m.arrays()
    bvh = LBVH(ctx, m)
    inf = bvh.info()
    print(f"{name}: {inf['faces']} faces, height {inf['height']}", flush=True)
    assert inf["faces"] == f.shape[0]
    assert math.ceil(math.log2(f.shape[0])) + 1 <= inf["height"] <= 64
    _check_tree(bvh.nodes(), v, f)


def test_lbvh_degenerate_trees(ctx):
    """One face, two faces, and many faces with the same centroid (equal
    Morton codes, told apart by the face index)."""
    from reyn_amd.mesh import LBVH

    tri = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
    for faces in (np.array([[0, 1, 2]]), np.array([[0, 1, 2], [0, 2, 1]]), np.tile([[0, 1, 2]], (300, 1))):
        m = Mesh.from_arrays(tri, faces)
        bvh = LBVH(ctx, m)
        if faces.shape[0] > 1:
            _check_tree(bvh.nodes(), tri, faces)
        p = np.array([[0.2, 0.2, 1.0], [2.0, 2.0, 0.0]], np.float32)
        assert np.array_equal(bvh.project(p), m.project(ctx, p))


@pytest.mark.parametrize("name", ["cube", "ico_sphere", "torus"])
def test_lbvh_project_matches_brute_force(ctx, name):
    """`LBVH::project` against every face tested in order (lbvh.cu:386-461,
    10,000 points in the mesh's box): the same bits; and against the float64
    restatement within rounding."""
    from reyn_amd.mesh import LBVH

    m = _meshes()[name]
    v, f = m.arrays()
    rng = np.random.default_rng(420)
    lo, hi = v.min(0), v.max(0)
    p = (lo + (hi - lo) * rng.uniform(-0.1, 1.1, (10000, 3))).astype(np.float32)
    bvh = LBVH(ctx, m)
    got = bvh.project(p)
    assert np.array_equal(got, m.project(ctx, p))
    k = 2000 if f.shape[0] < 2000 else 300  # the restatement loops over the faces
    want = mesh_sampling.closest_points(p[:k], v, f)
    assert np.abs(got[:k] - want).max() <= 4e-7 * np.abs(v).max()
    # distances are never larger than to any vertex
    dv = np.min(np.linalg.norm(p[:500, None, :].astype(np.float64) - v[None], axis=2), axis=1)
    assert np.all(np.linalg.norm(got[:500] - p[:500], axis=1) <= dv + 1e-6)


@pytest.mark.parametrize("name", ["cube", "ico_sphere", "torus"])
def test_lbvh_inside_matches_winding_number(ctx, name):
    from reyn_amd.mesh import LBVH

    m = _meshes()[name]
    v, f = m.arrays()
    rng = np.random.default_rng(9)
    lo, hi = v.min(0), v.max(0)
    p = (lo + (hi - lo) * rng.uniform(-0.2, 1.2, (20000, 3))).astype(np.float32)
    bvh = LBVH(ctx, m)
    near = np.linalg.norm(bvh.project(p) - p, axis=1) < 1e-5
    a, b = bvh.inside(p), m.inside(ctx, p)
    print(f"{name}: {a.sum()} of {p.shape[0]} inside, {near.sum()} on the surface", flush=True)
    assert a.sum() > 1000 and (~a).sum() > 1000
    assert np.array_equal(a[~near], b[~near])
    if name == "torus":  # the inscribed polyhedron lies within the smooth torus
        r = np.hypot(np.hypot(p[:, 0], p[:, 1]) - 1.0, p[:, 2])
        assert np.all(a[r < 0.34]) and not np.any(a[r > 0.36])


# ---- the reference's two scene kinds (Scene, Scene::from_obj) ---------------------


def _cube_in_cube(path):
    """An OBJ with a shape "box" (the [-1, 1]^3 cube) and a shape "fluid" (the
    cube scaled to [-0.5, 0.5]^3), faces as in the reference's cube.obj."""
    lines = (GOLD / "cube.obj").read_text().splitlines()
    v = [ln for ln in lines if ln.startswith("v ")]
    f = [ln for ln in lines if ln.startswith("f ")]
    out = ["o box", *v, *f, "o fluid"]
    out += ["v " + " ".join(str(0.5 * float(t)) for t in ln.split()[1:]) for ln in v]
    out += ["f " + " ".join(str(int(t) + 8) for t in ln.split()[1:]) for ln in f]
    path.write_text("\n".join(out) + "\n")
    return path


def test_scene_from_obj_and_box_scene_step(ctx, tmp_path):
    from reyn_amd.mesh import box_scene, scene_from_obj

    obj = _cube_in_cube(tmp_path / "cube_in_cube.obj")
    sc = scene_from_obj(ctx, obj, 8000, jitter_stddev=0.0, relaxation_iters=5)
    print(f"from_obj: {sc.n} fluid, {sc.n_bdy} boundary, h {sc.spacing}", flush=True)
    assert abs(sc.spacing - 0.05) < 1e-6
    assert sc.n == 19 ** 3  # lattice points on the fluid's faces count as outside
    assert sc.x.min() > -0.5 and sc.x.max() < 0.5
    assert np.abs(sc.xb).max() <= 1.0 + 1e-6 and np.allclose(np.abs(sc.xb).max(axis=1), 1.0, atol=1e-5)
    bx = box_scene(ctx, GOLD / "cube.obj", 8000, (-0.9, -0.9, -0.9), (0.1, 0.1, 0.1), relaxation_iters=5)
    print(f"box: {bx.n} fluid, {bx.n_bdy} boundary", flush=True)
    assert bx.n in (19 ** 3, 20 ** 3)
    # both step through a batch
    b = Batch(ctx, [sc, bx], solver="iisph", kernel="B3")
    b.step(3)
    for sid in b.ids():
        st = b.download(sid)
        assert all(np.isfinite(st[k]).all() for k in ("xx", "xy", "xz", "vx", "vy", "vz", "rho"))
    b.close()
