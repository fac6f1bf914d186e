"""Triangle meshes (`reyn_mesh_*`, `reyn_lbvh_*`): OBJ loading, surface
sampling and relaxation on the GPU, the LBVH with its closest-point and
inside queries, and a dam-break scene with a mesh obstacle.

The reference loads meshes with `load_mesh_from_obj` (src/scene/loader.cpp:
13-107), samples boundaries with `sample_mesh` (sample_boundary.cu:391-440),
builds an `LBVH` (src/datastructure/lbvh.cuh:614-758) and culls fluid with
its `CullPredicate` / `InsidePredicate` (scene.cu:55-140).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .device import Context
from .scenes import Scene, box_walls, lattice


class Mesh:
    """A host-side triangle mesh (`reyn_mesh`)."""

    def __init__(self, handle: C.c_void_p):
        self._lib = capi.lib()
        self.h = handle

    @staticmethod
    def load_obj(path, ignore=(), only: str | None = None) -> "Mesh":
        lib = capi.lib()
        arr = (C.c_char_p * max(1, len(ignore)))(*[s.encode() for s in ignore])
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib.reyn_mesh_load_obj(str(path).encode(), arr, len(ignore), only.encode() if only else None,
                                    C.byref(h), err, len(err))
        if rc != capi.OK:
            raise capi.ReynError(rc, f"load_obj {path}: {err.value.decode()}")
        return Mesh(h)

    @staticmethod
    def from_arrays(vertices, faces) -> "Mesh":
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        f = np.ascontiguousarray(faces, dtype=np.uint32)
        cols = [np.ascontiguousarray(v[:, a]) for a in range(3)]
        h = C.c_void_p()
        rc = capi.lib().reyn_mesh_create(*[c.ctypes.data for c in cols], v.shape[0], f.ctypes.data, f.shape[0],
                                         C.byref(h))
        if rc != capi.OK:
            raise capi.ReynError(rc, "mesh create")
        return Mesh(h)

    def info(self) -> dict:
        nv, nf, area, vol = C.c_uint32(), C.c_uint32(), C.c_double(), C.c_double()
        self._lib.reyn_mesh_info(self.h, C.byref(nv), C.byref(nf), C.byref(area), C.byref(vol))
        return {"vertices": nv.value, "faces": nf.value, "area": area.value, "volume": vol.value}

    def arrays(self) -> tuple[np.ndarray, np.ndarray]:
        inf = self.info()
        x, y, z = (np.empty(inf["vertices"]) for _ in range(3))
        f = np.empty((inf["faces"], 3), np.uint32)
        self._lib.reyn_mesh_data(self.h, x.ctypes.data, y.ctypes.data, z.ctypes.data, f.ctypes.data)
        return np.stack([x, y, z], axis=1), f

    def sample_count(self, h: float, oversampling: float = 1.0) -> int:
        n = C.c_uint32()
        rc = self._lib.reyn_mesh_sample_count(self.h, float(h), float(oversampling), C.byref(n))
        if rc != capi.OK:
            raise capi.ReynError(rc, "mesh sample count")
        return int(n.value)

    def sample(self, ctx: Context, h: float, oversampling: float = 1.0, seed: int = 1614201312,
               relaxation_iters: int = 0, relaxation_factor: float = 1.0) -> np.ndarray:
        """Boundary samples of the surface as a host (n, 3) float32 array,
        relaxed `relaxation_iters` times (then in the cell order of the last
        relaxation grid)."""
        n = self.sample_count(h, oversampling)
        xs, ys, zs = (ctx.empty(max(1, n)) for _ in range(3))
        got = C.c_uint32()
        ctx.check(ctx._lib.reyn_mesh_sample(ctx.h, self.h, float(h), float(oversampling), int(seed) & 0xFFFFFFFF,
                                            xs.ptr, ys.ptr, zs.ptr, n, C.byref(got)), "mesh sample")
        if relaxation_iters:
            ctx.check(ctx._lib.reyn_mesh_relax(ctx.h, self.h, float(h), float(oversampling), xs.ptr, ys.ptr, zs.ptr,
                                               n, int(relaxation_iters), float(relaxation_factor), None),
                      "mesh relax")
        return np.stack([a.numpy()[:n] for a in (xs, ys, zs)], axis=1)

    def relax(self, ctx: Context, points: np.ndarray, h: float, oversampling: float = 1.0, iterations: int = 1,
              factor: float = 1.0) -> tuple[np.ndarray, float]:
        """`relax_sampling` (sample_boundary.cu:100-166) of surface samples:
        returns the relaxed points (in the cell order of the last relaxation
        grid) and the max/min number density of the last iteration."""
        p = np.asarray(points, np.float32)
        n = p.shape[0]
        if n == 0:
            return p.reshape(0, 3), 1.0
        cols = [ctx.array(np.ascontiguousarray(p[:, a])) for a in range(3)]
        ratio = C.c_float(0.0)
        ctx.check(ctx._lib.reyn_mesh_relax(ctx.h, self.h, float(h), float(oversampling), *[c.ptr for c in cols], n,
                                           int(iterations), float(factor), C.byref(ratio)), "mesh relax")
        return np.stack([c.numpy() for c in cols], axis=1), float(ratio.value)

    def project(self, ctx: Context, points: np.ndarray) -> np.ndarray:
        """Closest points of the mesh, every face tested (`reyn_mesh_project`)."""
        p = np.asarray(points, np.float32)
        if p.shape[0] == 0:
            return p.reshape(0, 3)
        cols = [ctx.array(np.ascontiguousarray(p[:, a])) for a in range(3)]
        ctx.check(ctx._lib.reyn_mesh_project(ctx.h, self.h, *[c.ptr for c in cols], p.shape[0]), "mesh project")
        return np.stack([c.numpy() for c in cols], axis=1)

    def inside(self, ctx: Context, points: np.ndarray) -> np.ndarray:
        """Whether each point lies inside the (closed) mesh: the generalized
        winding number over every face."""
        p = np.asarray(points, np.float32)
        n = p.shape[0]
        if n == 0:
            return np.zeros(0, bool)
        cols = [ctx.array(np.ascontiguousarray(p[:, a])) for a in range(3)]
        out = ctx.empty(n, np.uint8)
        ctx.check(ctx._lib.reyn_mesh_inside(ctx.h, self.h, *[c.ptr for c in cols], n, out.ptr), "mesh inside")
        return out.numpy().astype(bool)

    def close(self) -> None:
        if self.h is not None and self.h.value:
            self._lib.reyn_mesh_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LBVH:
    """`LBVH` (lbvh.cuh:614-758) of a mesh on the context's GPU."""

    def __init__(self, ctx: Context, mesh: Mesh):
        self.ctx = ctx
        self._lib = ctx._lib
        self.h = C.c_void_p()
        ctx.check(self._lib.reyn_lbvh_create(ctx.h, mesh.h, C.byref(self.h)), "lbvh create")

    def info(self) -> dict:
        nf, height = C.c_uint32(), C.c_uint32()
        self._lib.reyn_lbvh_info(self.h, C.byref(nf), C.byref(height))
        return {"faces": nf.value, "height": height.value}

    def nodes(self) -> dict:
        """The tree on the host: children (index << 1 | 1 for internal nodes),
        leaf faces, and boxes (lo, hi) of the internal nodes and leaves."""
        nf = self.info()["faces"]
        ni = max(nf - 1, 0)
        left, right = np.zeros(ni, np.uint32), np.zeros(ni, np.uint32)
        leaf = np.zeros(nf, np.uint32)
        ibox, lbox = np.zeros((ni, 6)), np.zeros((nf, 6))
        self.ctx.check(self._lib.reyn_lbvh_nodes(self.ctx.h, self.h, left.ctypes.data, right.ctypes.data,
                                                 leaf.ctypes.data, ibox.ctypes.data, lbox.ctypes.data), "lbvh nodes")
        return {"left": left, "right": right, "leaf_face": leaf, "internal_box": ibox, "leaf_box": lbox}

    def project(self, points: np.ndarray) -> np.ndarray:
        """Closest points of the mesh (`LBVH::project`)."""
        p = np.asarray(points, np.float32)
        if p.shape[0] == 0:
            return p.reshape(0, 3)
        cols = [self.ctx.array(np.ascontiguousarray(p[:, a])) for a in range(3)]
        self.ctx.check(self._lib.reyn_lbvh_project(self.ctx.h, self.h, *[c.ptr for c in cols], p.shape[0]),
                       "lbvh project")
        return np.stack([c.numpy() for c in cols], axis=1)

    def inside(self, points: np.ndarray) -> np.ndarray:
        """Whether each point lies inside the (closed) mesh, from the
        pseudo-normal of its closest feature."""
        p = np.asarray(points, np.float32)
        n = p.shape[0]
        if n == 0:
            return np.zeros(0, bool)
        cols = [self.ctx.array(np.ascontiguousarray(p[:, a])) for a in range(3)]
        out = self.ctx.empty(n, np.uint8)
        self.ctx.check(self._lib.reyn_lbvh_inside(self.ctx.h, self.h, *[c.ptr for c in cols], n, out.ptr),
                       "lbvh inside")
        return out.numpy().astype(bool)

    def close(self) -> None:
        if self.h is not None and self.h.value:
            self._lib.reyn_lbvh_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def obstacle_scene(ctx: Context, mesh: Mesh, fluid_counts, spacing: float, box_hi, offset=(0.0, 0.0, 0.0),
                   scale: float = 1.0, jitter: float = 0.01, seed: int = 1, layers: int = 2, rho_0: float = 1000.0,
                   oversampling: float = 1.0, cull_radius: float = 1.0) -> Scene:
    """A dam break whose box holds the mesh (scaled, then moved by `offset`) as
    an obstacle: the walls and the mesh's surface samples form the boundary,
    and fluid points inside the mesh or within `cull_radius` spacings of a
    surface sample are removed (the reference's `InsidePredicate` and
    `CullPredicate`)."""
    from scipy.spatial import cKDTree

    v, f = mesh.arrays()
    placed = Mesh.from_arrays(v * float(scale) + np.asarray(offset, np.float64), f)
    s = float(spacing)
    surf = placed.sample(ctx, s, oversampling)
    x = lattice(fluid_counts, s, (0.0, 0.0, 0.0), jitter, seed)
    bvh = LBVH(ctx, placed)
    keep = ~bvh.inside(x)
    bvh.close()
    if surf.shape[0]:
        d, _ = cKDTree(surf.astype(np.float64)).query(x.astype(np.float64), k=1)
        keep &= d > cull_radius * s
    hi = np.asarray(box_hi, np.float64)
    xb = np.concatenate([box_walls(np.zeros(3), hi, s, layers), surf]).astype(np.float32)
    placed.close()
    return Scene(x[keep], xb, s, float(rho_0), np.zeros(3), hi,
                 meta={"fluid_counts": tuple(int(c) for c in fluid_counts), "box": tuple(float(c) for c in hi),
                       "mesh_samples": int(surf.shape[0]), "culled": int((~keep).sum())})


def _cull_and_jitter(ctx: Context, cols, n: int, xb: np.ndarray, h: float, rho_0: float, cull_bdy_radius: float,
                     jitter_stddev: float, seed: int) -> np.ndarray:
    """Removes the points within cull_bdy_radius h of a boundary sample (stable
    order) and jitters the rest; returns them as a host (n, 3) array."""
    from .sim import BoundarySamples

    bdy = BoundarySamples(ctx, xb, h, rho_0)
    keep = ctx.empty(max(n, 1), np.uint8)
    view = bdy.view()
    ctx.check(ctx._lib.reyn_boundary_cull(ctx.h, C.byref(view), float(cull_bdy_radius * h), *[c.ptr for c in cols],
                                          n, keep.ptr), "scene cull")
    arr = (C.c_void_p * 3)(*[c.ptr for c in cols])
    kept = C.c_uint32()
    ctx.check(ctx._lib.reyn_compact_f32(ctx.h, keep.ptr, n, arr, 3, C.byref(kept)), "scene cull")
    ctx.check(ctx._lib.reyn_jitter_normal(ctx.h, *[c.ptr for c in cols], kept.value, float(jitter_stddev * h),
                                          (int(seed) ^ 0x5BD1E995) & 0xFFFFFFFF), "scene jitter")
    return np.stack([c.numpy()[:kept.value] for c in cols], axis=1)


def _lattice(ctx: Context, lo, counts, h: float):
    n = int(counts[0]) * int(counts[1]) * int(counts[2])
    cols = [ctx.empty(max(n, 1)) for _ in range(3)]
    ctx.check(ctx._lib.reyn_lattice_box(ctx.h, capi.f3(np.asarray(lo, np.float32)), int(counts[0]), int(counts[1]),
                                        int(counts[2]), float(h), *[c.ptr for c in cols]), "scene lattice")
    return cols, n


def _boundary(ctx: Context, path, h: float, oversampling: float, relaxation_iters: int, seed: int) -> np.ndarray:
    m = Mesh.load_obj(path, ignore=("fluid",))
    try:
        return m.sample(ctx, h, oversampling, seed, relaxation_iters=relaxation_iters)
    finally:
        m.close()


def box_scene(ctx: Context, path, n_desired: int, lo, hi, rho_0: float = 1000.0, oversampling: float = 2.0,
              cull_bdy_radius: float = 1.0, jitter_stddev: float = 0.01, relaxation_iters: int = 50,
              seed: int = 1614201312) -> Scene:
    """The reference's `Scene(path, N_desired, min, max, ...)` (scene.cu:239-310):
    a lattice box of fluid of spacing h = cbrt(volume / n_desired) in the
    boundary sampled (and relaxed) from every shape of the OBJ file but
    "fluid"; the same set-up as the C++ `reyn::Scene`."""
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    d = hi - lo
    h = float(np.cbrt(np.float32((d[0] * d[1]) * d[2]) / np.float32(n_desired)).astype(np.float32))
    xb = _boundary(ctx, path, h, oversampling, relaxation_iters, seed)
    counts = [abs(int(np.floor(np.float32(v) / np.float32(h)))) for v in d]
    if counts[0] * counts[1] * counts[2] == 0:
        raise ValueError("Initialization of box failed, zero particles placed")
    cols, n = _lattice(ctx, lo, counts, h)
    x = _cull_and_jitter(ctx, cols, n, xb, h, rho_0, cull_bdy_radius, jitter_stddev, seed)
    return Scene(x.astype(np.float32), xb, h, float(rho_0), xb.min(axis=0), xb.max(axis=0),
                 meta={"kind": "box", "lattice": counts})


def scene_from_obj(ctx: Context, path, n_desired: int, rho_0: float = 1000.0, oversampling: float = 2.0,
                   cull_bdy_radius: float = 1.0, jitter_stddev: float = 0.01, relaxation_iters: int = 50,
                   seed: int = 1614201312) -> Scene:
    """The reference's `Scene::from_obj` (scene.cu:312-464): the volume of the
    OBJ shape named "fluid", filled on a lattice of spacing
    h = cbrt(volume / n_desired) and kept where the LBVH inside test says so,
    in the boundary sampled from the other shapes; bounds grown to hold the
    fluid's box with half a spacing to spare."""
    fluid = Mesh.load_obj(path, only="fluid")
    try:
        h = float(np.float32(np.cbrt(fluid.info()["volume"] / float(n_desired))))
        v, _ = fluid.arrays()
        flo, fhi = v.min(axis=0).astype(np.float32), v.max(axis=0).astype(np.float32)
        xb = _boundary(ctx, path, h, oversampling, relaxation_iters, seed)
        counts = [abs(int(np.floor((np.float32(b) - np.float32(a)) / np.float32(h)))) for a, b in zip(flo, fhi)]
        if counts[0] * counts[1] * counts[2] == 0:
            raise ValueError("Initialization of box failed, zero particles placed")
        cols, n = _lattice(ctx, flo, counts, h)
        bvh = LBVH(ctx, fluid)
        inside = ctx.empty(max(n, 1), np.uint8)
        ctx.check(ctx._lib.reyn_lbvh_inside(ctx.h, bvh.h, *[c.ptr for c in cols], n, inside.ptr), "fluid volume")
        arr = (C.c_void_p * 3)(*[c.ptr for c in cols])
        kept = C.c_uint32()
        ctx.check(ctx._lib.reyn_compact_f32(ctx.h, inside.ptr, n, arr, 3, C.byref(kept)), "fluid volume")
        bvh.close()
        x = _cull_and_jitter(ctx, cols, kept.value, xb, h, rho_0, cull_bdy_radius, jitter_stddev, seed)
    finally:
        fluid.close()
    half = np.float32(0.5 * h)
    lo = np.minimum(xb.min(axis=0), flo - half)
    hi = np.maximum(xb.max(axis=0), fhi + half)
    return Scene(x.astype(np.float32), xb, h, float(rho_0), lo, hi, meta={"kind": "from_obj", "lattice": counts})


__all__ = ["LBVH", "Mesh", "box_scene", "obstacle_scene", "scene_from_obj"]
