"""Restatement of the mesh surface sampling in numpy (TEST INFRASTRUCTURE).

The reference samples a mesh by stratified inversion of the face-area CDF
(src/scene/sample_boundary.cu:179-249 `_uniform_sample_tri_cdf`, 299-389
`uniform_sample_mesh`): n = ceil(area / (1.5 sqrt(3) h_bdy^2)) samples with
h_bdy = h / oversampling; sample i draws the area coordinate
(i / n) * area + (area / n) * u, takes the face whose inclusive CDF entry is
the first >= it (lower-bound binary search), and a uniform point of the face
from barycentric (s, t) in [0, 1)^2 folded into the triangle. The random
numbers here are the counter hash of (seed, sample, axis) the library uses in
place of curand. Every operation is in double, in the order the library
forms it, then rounded to float, so the GPU samples must match bit for bit.

The relaxation (`relax_sampling`, sample_boundary.cu:100-166) is restated in
float64: number densities n_i = sum_j W(x_ij) with B3 of support 2 h_bdy
(the sample itself included), the symmetric shift
-k h_bdy^2 sum_j (n_i + n_j) / n_0^2 grad W_ij with n_0 the rest density of a
hexagonal plane lattice (`_hex_plane_rest_density`, :80-98), and the
projection of every shifted sample to the closest point of the mesh. The
projection here takes a different route from the library's: the orthogonal
projection onto each face's plane when its barycentric coordinates are
non-negative, else the nearest of the three clamped edge projections.
Only tests import this module.
"""

from __future__ import annotations

import numpy as np

from reyn_amd.scenes import hash_uniform

from . import binding

B3 = 1  # kernel kind of the cubic B-spline


def face_cdf(v: np.ndarray, f: np.ndarray) -> np.ndarray:
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    e1, e2 = b - a, c - a
    cx = e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1]
    cy = e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2]
    cz = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    area = 0.5 * np.sqrt((cx * cx + cy * cy) + cz * cz)
    out = np.empty_like(area)
    run = 0.0
    for k, x in enumerate(area):  # the running sum in face order
        run = run + x
        out[k] = run
    return out


def sample_count(total: float, h: float, oversampling: float) -> int:
    hb = np.float32(np.float64(np.float32(h)) / np.float64(np.float32(oversampling)))
    cell = np.float32(np.float32(1.5) * np.sqrt(np.float32(3.0))) * hb * hb
    return int(np.ceil(total / np.float64(cell)))


def sample(v: np.ndarray, f: np.ndarray, h: float, oversampling: float, seed: int) -> np.ndarray:
    cdf = face_cdf(v, f)
    total = cdf[-1]
    n = sample_count(total, h, oversampling)
    i = np.arange(n, dtype=np.uint64)
    bin_ = total / float(n)
    a = (i.astype(np.float64) / float(n)) * total + bin_ * hash_uniform(seed, i, 0)
    tri = np.minimum(np.searchsorted(cdf, a, side="left"), len(cdf) - 1)
    u, w = hash_uniform(seed, i, 1), hash_uniform(seed, i, 2)
    inside = u + w <= 1.0
    s = np.where(inside, u, 1.0 - u)
    t = np.where(inside, w, 1.0 - w)
    p0, p1, p2 = v[f[tri, 0]], v[f[tri, 1]], v[f[tri, 2]]
    e1, e2 = p1 - p0, p2 - p0
    out = (p0 + s[:, None] * e1) + t[:, None] * e2
    return out.astype(np.float32)


def hex_rest_density(h_bdy: float) -> float:
    """Sum of W over the hexagonal lattice points q a + r b, q, r in [-2, 2],
    a = h_bdy (1, 0, 0), b = h_bdy (1/2, sqrt(3)/2, 0) (sample_boundary.cu:80-98)."""
    q, r = np.meshgrid(np.arange(-2, 3), np.arange(-2, 3), indexing="ij")
    q, r = q.ravel().astype(np.float64), r.ravel().astype(np.float64)
    d = np.stack([h_bdy * (q + 0.5 * r), h_bdy * (np.sqrt(3.0) * 0.5 * r), np.zeros_like(q)], axis=1)
    w, _ = binding.kernel_eval(B3, 2.0 * h_bdy, d)
    return float(w.sum())


def _closest_on_segment(p, a, b):
    ab = b - a
    t = np.clip(np.einsum("ij,ij->i", p - a, ab) / np.einsum("ij,ij->i", ab, ab), 0.0, 1.0)
    return a + t[:, None] * ab


def closest_points(p: np.ndarray, v: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Closest point of the mesh (v, f) to every row of p, by brute force over
    the faces (first face on ties)."""
    p = np.asarray(p, np.float64)
    best = p.copy()
    bd = np.full(p.shape[0], np.inf)
    for tri in f:
        a, b, c = (np.broadcast_to(v[k], p.shape) for k in tri)
        n = np.cross(v[tri[1]] - v[tri[0]], v[tri[2]] - v[tri[0]])
        nn = float(n @ n)
        q = p - ((p - v[tri[0]]) @ n / nn)[:, None] * n
        # barycentric coordinates of the plane projection from sub-areas
        l0 = np.cross(b - q, c - q) @ n / nn
        l1 = np.cross(c - q, a - q) @ n / nn
        l2 = np.cross(a - q, b - q) @ n / nn
        inside = (l0 >= 0) & (l1 >= 0) & (l2 >= 0)
        cand = q.copy()
        if not inside.all():
            o = ~inside
            es = [_closest_on_segment(p[o], a[o], b[o]), _closest_on_segment(p[o], b[o], c[o]),
                  _closest_on_segment(p[o], c[o], a[o])]
            ds = np.stack([((e - p[o]) ** 2).sum(1) for e in es], axis=1)
            k = ds.argmin(1)
            cand[o] = np.choose(k[:, None], es)
        d = ((cand - p) ** 2).sum(1)
        take = d < bd
        best[take], bd[take] = cand[take], d[take]
    return best


def number_density(x: np.ndarray, h_bdy: float) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """n_i and the pairs (i, j), i != j, within the support 2 h_bdy."""
    from scipy.spatial import cKDTree

    pairs = cKDTree(x).query_pairs(2.0 * h_bdy, output_type="ndarray")
    i = np.concatenate([pairs[:, 0], pairs[:, 1]])
    j = np.concatenate([pairs[:, 1], pairs[:, 0]])
    w, _ = binding.kernel_eval(B3, 2.0 * h_bdy, x[i] - x[j])
    w0, _ = binding.kernel_eval(B3, 2.0 * h_bdy, np.zeros((1, 3)))
    nd = np.full(x.shape[0], w0[0]) + np.bincount(i, weights=w, minlength=x.shape[0])
    return nd, i, j


def relax_shift(x: np.ndarray, h_bdy: float, factor: float) -> tuple[np.ndarray, np.ndarray]:
    """Shifted positions (before projection) and the number densities."""
    nd, i, j = number_density(x, h_bdy)
    inv = 1.0 / hex_rest_density(h_bdy) ** 2
    _, g = binding.kernel_eval(B3, 2.0 * h_bdy, x[i] - x[j])
    t = (inv * nd[i] + inv * nd[j])[:, None] * g
    acc = np.stack([np.bincount(i, weights=t[:, a], minlength=x.shape[0]) for a in range(3)], axis=1)
    return x - factor * h_bdy * h_bdy * acc, nd


def relax(points: np.ndarray, v: np.ndarray, f: np.ndarray, h: float, oversampling: float, iterations: int,
          factor: float = 1.0) -> tuple[np.ndarray, float]:
    """`iterations` rounds of relaxation; returns the points (input order) and
    max/min number density of the last round."""
    h_bdy = float(np.float32(np.float64(np.float32(h)) / np.float64(np.float32(oversampling))))
    x = np.asarray(points, np.float64)
    ratio = 1.0
    for _ in range(iterations):
        xs, nd = relax_shift(x, h_bdy, factor)
        ratio = float(nd.max() / nd.min())
        x = closest_points(xs, v, f)
    return x, ratio
