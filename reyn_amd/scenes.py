"""Seeded synthetic scenes for the SPH step: fluid lattices, wall shells and
a sphere obstacle, generated on the host in numpy.

The reference builds scenes from OBJ meshes (src/scene/scene.cu:316-464);
mesh scenes are outside the per-step path. These generators produce the
lattice scenes the workloads are defined on: a lattice of spacing s with
points at (i + 1/2) s, an optional uniform jitter, masses rho_0 s^3
(scene.cu:14-37), and boundary samples at the same spacing.

The device generator (`reyn_dam_break_generate`, reyn_amd/csrc/scene.hip)
produces the same float bits: coordinates are formed in double in the same
operations, the jitter is a counter hash of (seed, point, axis) rather than a
random stream, and the wall shell is selected by lattice index.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


_M32 = np.uint64(0xFFFFFFFF)


def _mix32(x: np.ndarray) -> np.ndarray:
    x = x & _M32
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & _M32
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & _M32
    x ^= x >> np.uint64(16)
    return x


def hash_uniform(seed: int, idx: np.ndarray, axis: int) -> np.ndarray:
    """U[0, 1) with 24 bits from a counter hash of (seed, point, axis), the
    device generator's jitter source."""
    s = _mix32(np.array([(int(seed) * 0x9E3779B9 + 0x7F4A7C15) & 0xFFFFFFFF], dtype=np.uint64))
    key = (np.asarray(idx, dtype=np.uint64) * np.uint64(3) + np.uint64(axis)) & _M32
    h = _mix32(s ^ key)
    return (h >> np.uint64(8)).astype(np.float64) * (1.0 / 16777216.0)


def lattice(counts, spacing: float, origin=(0.0, 0.0, 0.0), jitter: float = 0.0, seed: int = 0) -> np.ndarray:
    """Points origin + (i + 1/2) s for i in [0, counts) per axis, x fastest,
    each coordinate moved by U(-jitter s, +jitter s) drawn from
    `hash_uniform(seed, point, axis)`."""
    nx, ny, nz = (int(c) for c in counts)
    s = float(spacing)
    ix = np.arange(nx, dtype=np.float64)
    iy = np.arange(ny, dtype=np.float64)
    iz = np.arange(nz, dtype=np.float64)
    z, y, x = np.meshgrid(iz, iy, ix, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    pts = (pts + 0.5) * s + np.asarray(origin, dtype=np.float64)
    if jitter:
        lo, hi = -(jitter * s), jitter * s
        idx = np.arange(pts.shape[0], dtype=np.uint64)
        for a in range(3):
            pts[:, a] += lo + (hi - lo) * hash_uniform(seed, idx, a)
    return pts.astype(np.float32)


def box_walls(lo, hi, spacing: float, layers: int = 2) -> np.ndarray:
    """Lattice points of spacing s in the shell of `layers` layers around the
    box [lo, hi]: (i + 1/2) s offsets from lo - layers s on a lattice of
    rint((hi - lo) / s) + 2 layers points per axis, keeping the points outside
    its inner index box [layers, n - layers)^3 (x fastest)."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    s = float(spacing)
    start = lo - layers * s
    counts = np.rint((hi - lo) / s).astype(int) + 2 * layers
    idx = [np.arange(counts[a]) for a in range(3)]
    iz, iy, ix = np.meshgrid(idx[2], idx[1], idx[0], indexing="ij")
    ii = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)
    inside = np.all((ii >= layers) & (ii < counts - layers), axis=1)
    ii = ii[~inside]
    pts = np.stack([start[a] + (ii[:, a].astype(np.float64) + 0.5) * s for a in range(3)], axis=1)
    return pts.astype(np.float32)


def sphere_shells(center, radius: float, spacing: float, layers: int = 2) -> np.ndarray:
    """Points on `layers` concentric spheres of radii r - (k + 1/2) s, each a
    Fibonacci lattice with about one point per s^2 of area."""
    c = np.asarray(center, dtype=np.float64)
    s = float(spacing)
    out = []
    golden = np.pi * (3.0 - np.sqrt(5.0))
    for k in range(layers):
        r = radius - (k + 0.5) * s
        if r <= 0:
            break
        n = max(4, int(round(4.0 * np.pi * r * r / (s * s))))
        i = np.arange(n, dtype=np.float64)
        y = 1.0 - 2.0 * (i + 0.5) / n
        rad = np.sqrt(np.maximum(0.0, 1.0 - y * y))
        th = golden * i
        out.append(c + r * np.stack([np.cos(th) * rad, y, np.sin(th) * rad], axis=1))
    return np.concatenate(out).astype(np.float32) if out else np.zeros((0, 3), np.float32)


@dataclass
class Scene:
    """A fluid block, its boundary samples and the scene bounds (the boundary
    samples' AABB, which `hard_enforce_bounds` clamps to)."""

    x: np.ndarray
    xb: np.ndarray
    spacing: float
    rho_0: float
    box_lo: np.ndarray
    box_hi: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_bdy(self) -> int:
        return int(self.xb.shape[0])

    @property
    def mass(self) -> float:
        s = np.float32(self.spacing)
        return float(np.float32(self.rho_0) * (s * s * s))

    @property
    def bound_min(self) -> np.ndarray:
        return self.xb.min(axis=0) if self.n_bdy else self.box_lo.astype(np.float32)

    @property
    def bound_max(self) -> np.ndarray:
        return self.xb.max(axis=0) if self.n_bdy else self.box_hi.astype(np.float32)


def dam_break(fluid_counts, spacing: float, box_hi, rho_0: float = 1000.0, jitter: float = 0.01, seed: int = 1,
              sphere: tuple | None = None, layers: int = 2) -> Scene:
    """Fluid block of fluid_counts lattice points in the corner (0,0,0) of the
    box [0, box_hi], walls of `layers` layers, optional sphere obstacle
    (center, radius)."""
    x = lattice(fluid_counts, spacing, (0.0, 0.0, 0.0), jitter, seed)
    lo = np.zeros(3)
    hi = np.asarray(box_hi, dtype=np.float64)
    xb = box_walls(lo, hi, spacing, layers)
    meta = {"fluid_counts": tuple(int(c) for c in fluid_counts), "box": tuple(float(v) for v in hi),
            "jitter": jitter, "seed": seed, "layers": layers}
    if sphere is not None:
        center, radius = sphere
        sb = sphere_shells(center, radius, spacing, layers)
        # fluid points inside the obstacle are removed
        d = np.linalg.norm(x.astype(np.float64) - np.asarray(center), axis=1)
        x = x[d > radius + 0.5 * spacing]
        xb = np.concatenate([xb, sb])
        meta["sphere"] = (tuple(float(v) for v in center), float(radius))
    return Scene(x, xb, float(spacing), float(rho_0), lo, hi, meta)


def dam_break_desc(sc: Scene):
    """The `reyn_dam_break` description of a scene made by `dam_break`: the
    device generator makes the same scene from it."""
    from . import capi

    m = sc.meta
    d = capi.DamBreak()
    d.counts[:] = [int(c) for c in m["fluid_counts"]]
    d.spacing = float(sc.spacing)
    d.box[:] = [float(v) for v in m["box"]]
    d.jitter = float(m["jitter"])
    d.seed = int(m["seed"])
    d.layers = int(m.get("layers", 2))
    if "sphere" in m:
        center, radius = m["sphere"]
        d.sphere = 1
        d.center[:] = [float(v) for v in center]
        d.radius = float(radius)
    return d


def config1_sesph() -> Scene:
    """32^3 fluid at s = 0.02 in the corner of [0,1]^3, jitter +-1%, seed 1."""
    return dam_break((32, 32, 32), 0.02, (1.0, 1.0, 1.0), jitter=0.01, seed=1)


def config2_grid(side: int = 100) -> Scene:
    """side^3 lattice at s = 0.02 filling [0,2]^3, jitter +-5%, seed 2."""
    s = 2.0 / side
    return dam_break((side, side, side), s, (2.0, 2.0, 2.0), jitter=0.05, seed=2)


def config3_iisph(side: int = 128, seed: int = 3) -> Scene:
    """side^3 fluid block [0,1]^3 at s = 1/side in the box [0,2]x[0,1.25]x[0,2]
    with 2 wall layers and a sphere obstacle at (1.5, 0.25, 1.0) of radius
    U[0.1, 0.25] (seeded)."""
    s = 1.0 / side
    rng = np.random.default_rng(seed)
    radius = float(rng.uniform(0.1, 0.25))
    return dam_break((side, side, side), s, (2.0, 1.25, 2.0), jitter=0.01, seed=seed,
                     sphere=((1.5, 0.25, 1.0), radius))


def config4_batch(count: int = 64, seed: int = 4, n_lo: float = 1e4, n_hi: float = 3e5) -> list[Scene]:
    """`count` dam-break scenes: N_s = round(exp U[ln n_lo, ln n_hi]) target
    particles, a box (a, 1, 1) with a ~ U[1, 3], a fluid column of fill
    f ~ U[0.2, 0.6] along x at full height, spacing s = (f a / N_s)^(1/3)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n_s = int(round(float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))))
        a = float(rng.uniform(1.0, 3.0))
        f = float(rng.uniform(0.2, 0.6))
        s = float((f * a / n_s) ** (1.0 / 3.0))
        counts = (max(1, int(f * a / s)), max(1, int(1.0 / s)), max(1, int(1.0 / s)))
        sc = dam_break(counts, s, (a, 1.0, 1.0), jitter=0.01, seed=seed * 1000 + k)
        sc.meta.update({"target_n": n_s, "aspect": a, "fill": f})
        out.append(sc)
    return out


def config5_jobs(count: int, seed: int = 5, n_lo: float = 5e3, n_hi: float = 5e5, steps_lo: int = 200,
                 steps_hi: int = 2000, sphere_p: float = 0.5) -> list[tuple[Scene, int]]:
    """`count` serving jobs (scene, steps): N = round(exp U[ln n_lo, ln n_hi])
    target particles in a box with sides U[1, 3]^3, a dam-break block of fill
    f ~ U[0.2, 0.6] along x at 0.6 of the height and full depth, spacing
    s = (fluid volume / N)^(1/3), with probability `sphere_p` a sphere obstacle
    of radius U[0.1, 0.25] * min(height, depth) (at most half the free length)
    resting on the floor in front of the block, and steps ~ U{steps_lo ..
    steps_hi}."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n_s = int(round(float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))))
        a, b, c = (float(v) for v in rng.uniform(1.0, 3.0, 3))
        f = float(rng.uniform(0.2, 0.6))
        fluid = np.array([f * a, 0.6 * b, c])
        s = float((fluid.prod() / n_s) ** (1.0 / 3.0))
        counts = tuple(max(1, int(e / s)) for e in fluid)
        sphere = None
        has_sphere = rng.uniform() < sphere_p
        r = min(float(rng.uniform(0.1, 0.25)) * min(b, c), 0.5 * (1.0 - f) * a - 2 * s)
        u = float(rng.uniform())
        if has_sphere and r >= 4 * s:  # the sphere fits between the block and the far wall
            lo, hi = f * a + r + 2 * s, a - r - 2 * s
            sphere = ((lo + u * max(0.0, hi - lo), r + 2 * s, 0.5 * c), r)
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        sc = dam_break(counts, s, (a, b, c), jitter=0.01, seed=seed * 1000 + k, sphere=sphere)
        sc.meta.update({"target_n": n_s, "fill": f, "job": k, "steps": steps})
        out.append((sc, steps))
    return out
