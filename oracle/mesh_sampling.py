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
Only tests import this module.
"""

from __future__ import annotations

import numpy as np

from reyn_amd.scenes import hash_uniform


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
