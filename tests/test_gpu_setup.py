"""Scene set-up on the GPU (`reyn_lattice_box`, `reyn_jitter_normal`,
`reyn_boundary_cull`, `reyn_compact_f32`, src/scene/scene.cu:14-237) against
numpy: the lattice bit for bit in the reference's index order and float
operations, the cull against brute-force distances, the compaction against
boolean indexing (stable), and the jitter's statistics."""

import ctypes as C

import numpy as np
import pytest
from scipy.spatial import cKDTree

from reyn_amd import capi
from reyn_amd.sim import BoundarySamples

pytestmark = pytest.mark.gpu


def _arrays(ctx, p):
    return [ctx.array(np.ascontiguousarray(p[:, a], dtype=np.float32)) for a in range(3)]


def test_lattice_box(ctx):
    nx, ny, nz, h = 7, 5, 3, np.float32(0.037)
    o = np.float32([-0.3, 0.25, 1.5])
    n = nx * ny * nz
    xs, ys, zs = (ctx.empty(n) for _ in range(3))
    ctx.check(ctx._lib.reyn_lattice_box(ctx.h, capi.f3(o), nx, ny, nz, float(h), xs.ptr, ys.ptr, zs.ptr), "lattice")
    i = np.arange(n)
    ix, iy, iz = i % nx, (i // nx) % ny, i // (nx * ny)
    want = [o[0] + ix.astype(np.float32) * h, o[1] + iy.astype(np.float32) * h, o[2] + iz.astype(np.float32) * h]
    for got, w in zip((xs, ys, zs), want):
        assert np.array_equal(got.numpy(), w.astype(np.float32))


def test_jitter_statistics(ctx):
    n, sigma = 200000, 0.01
    xs, ys, zs = (ctx.full(n, 1.0) for _ in range(3))
    ctx.check(ctx._lib.reyn_jitter_normal(ctx.h, xs.ptr, ys.ptr, zs.ptr, n, sigma, 7), "jitter")
    d = np.stack([a.numpy() for a in (xs, ys, zs)], axis=1).astype(np.float64) - 1.0
    assert np.abs(d.mean(axis=0)).max() < 5 * sigma / np.sqrt(n)
    assert np.allclose(d.std(axis=0), sigma, rtol=0.01)
    # the three axes are independent
    c = np.corrcoef(d.T)
    assert np.abs(c[np.triu_indices(3, 1)]).max() < 0.01
    # deterministic in the seed
    e = ctx.full(n, 1.0)
    f2, g2 = ctx.full(n, 1.0), ctx.full(n, 1.0)
    ctx.check(ctx._lib.reyn_jitter_normal(ctx.h, e.ptr, f2.ptr, g2.ptr, n, sigma, 7), "jitter")
    assert np.array_equal(e.numpy(), xs.numpy())


@pytest.mark.parametrize("radius", [0.5, 1.0, 2.0])
def test_boundary_cull_and_compact(ctx, radius):
    h = 0.05
    rng = np.random.default_rng(3)
    xb = rng.uniform(0.0, 1.0, (3000, 3)).astype(np.float32)
    bdy = BoundarySamples(ctx, xb, h, 1000.0)
    p = rng.uniform(-0.05, 1.05, (20000, 3)).astype(np.float32)
    cols = _arrays(ctx, p)
    keep = ctx.empty(p.shape[0], np.uint8)
    view = bdy.view()
    ctx.check(ctx._lib.reyn_boundary_cull(ctx.h, C.byref(view), float(radius * h), *[c.ptr for c in cols],
                                          p.shape[0], keep.ptr), "cull")
    d, _ = cKDTree(bdy.positions().astype(np.float64)).query(p.astype(np.float64))
    want = d > radius * h
    got = keep.numpy().astype(bool)
    # the GPU tests |x_ij|^2 in float: only points at the radius itself may differ
    amb = np.abs(d - radius * h) < 1e-6
    assert np.array_equal(got[~amb], want[~amb])
    assert (~got).sum() > 100 and got.sum() > 100
    # stable compaction of three fields
    arr = (C.c_void_p * 3)(*[c.ptr for c in cols])
    cnt = C.c_uint32()
    ctx.check(ctx._lib.reyn_compact_f32(ctx.h, keep.ptr, p.shape[0], arr, 3, C.byref(cnt)), "compact")
    assert cnt.value == got.sum()
    for a, c in enumerate(cols):
        assert np.array_equal(c.numpy()[:cnt.value], p[got, a])


def test_cull_radius_beyond_grid_is_rejected(ctx):
    bdy = BoundarySamples(ctx, np.random.default_rng(1).uniform(0, 1, (100, 3)), 0.05, 1000.0)
    cols = _arrays(ctx, np.zeros((4, 3)))
    keep = ctx.empty(4, np.uint8)
    view = bdy.view()
    rc = ctx._lib.reyn_boundary_cull(ctx.h, C.byref(view), 0.2, *[c.ptr for c in cols], 4, keep.ptr)
    assert rc == capi.ERR_INVALID_ARGUMENT
