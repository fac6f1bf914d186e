"""GPU parity of the uniform grid, reorder, neighbour query and smoothing
kernels against the oracle, through the C-ABI.

Bar: grid tables (counts, prefix, sorted, keys) and neighbour counts are
bit-exact; kernel values agree with the FP64 oracle to 1e-5 relative (scaled
by W(0)) and pass the reference's own kernel-property checks
(src/kernels.cu:19-181) on the reference's own inputs.
"""

import ctypes as C

import numpy as np
import pytest

from oracle import binding as orc
from reyn_amd import capi, scenes
from reyn_amd.sim import UniformGridBuilder, grid_tables

pytestmark = pytest.mark.gpu


def _upload_xyz(ctx, x):
    return [ctx.array(np.ascontiguousarray(x[:, a])) for a in range(3)]


def _gpu_construct(ctx, x, lo, hi, r, ratio, order=0):
    b = UniformGridBuilder(ctx, lo, hi, r, ratio, order)
    xx, xy, xz = _upload_xyz(ctx, x)
    g = b.construct(r, xx, xy, xz)
    prefix, sorted_ = grid_tables(ctx, g)
    counts, keys = b.debug_tables(x.shape[0])
    return b, g, (xx, xy, xz), prefix, sorted_, counts, keys


@pytest.mark.parametrize("ratio", [0.5, 1.0, 2.0])
def test_grid_tables_bitexact_reference_points(ctx, ratio):
    """The reference grid test's 50^3 points (mt19937(161), [0,55]^3, r=2.2)."""
    x = orc.reference_grid_points(50, 1.1)
    r = np.float32(2.0 * np.float32(1.1))
    lo, hi = np.zeros(3, np.float32), np.full(3, np.float32(1.1) * 50, np.float32)
    b, g, _, prefix, sorted_, counts, keys = _gpu_construct(ctx, x, lo, hi, r, ratio)
    geom = orc.geometry(lo, hi, r, ratio)
    ref = orc.grid_build(geom, x[:, 0], x[:, 1], x[:, 2])
    assert (g.nx, g.ny, g.nz, g.range) == (geom.n[0], geom.n[1], geom.n[2], geom.range)
    assert g.cell_size == pytest.approx(geom.cell, rel=0, abs=0)
    np.testing.assert_array_equal(keys, ref.keys)
    np.testing.assert_array_equal(prefix, ref.prefix)
    np.testing.assert_array_equal(sorted_, ref.sorted)
    # counts are consumed by the scan and reset to zero for the next build
    assert not counts.any()
    np.testing.assert_array_equal(np.diff(prefix), ref.counts)


@pytest.mark.parametrize("ratio", [0.5, 1.0, 2.0])
def test_neighbour_query_matches_brute_force(ctx, ratio):
    """Reference test (uniformgrid.cu:247-346): grid and O(N^2) brute force
    agree on every neighbour count exactly (20k-point subset) and on the sums
    of |x_ij|^2 and x_ij approximately."""
    x = orc.reference_grid_points(50, 1.1)
    r = np.float32(2.0 * np.float32(1.1))
    lo, hi = np.zeros(3, np.float32), np.full(3, np.float32(1.1) * 50, np.float32)
    b, g, (xx, xy, xz), *_ = _gpu_construct(ctx, x, lo, hi, r, ratio)
    n = x.shape[0]
    cnt, l2, v = (ctx.empty(n, np.uint32), ctx.empty(n), [ctx.empty(n) for _ in range(3)])
    ctx.check(ctx._lib.reyn_grid_query_stats(ctx.h, C.byref(g), xx._ptr, xy._ptr, xz._ptr, n, cnt._ptr, l2._ptr,
                                             v[0]._ptr, v[1]._ptr, v[2]._ptr))
    cnt_h, l2_h = cnt.numpy(), l2.numpy()
    idx = np.random.default_rng(0).choice(n, 20000, replace=False).astype(np.uint32)
    bc, bl2, bv = orc.brute_stats_subset(r * r, x[:, 0], x[:, 1], x[:, 2], idx)
    np.testing.assert_array_equal(cnt_h[idx], bc)
    np.testing.assert_allclose(l2_h[idx], bl2, rtol=1e-4, atol=1e-4)
    vv = np.stack([a.numpy() for a in v], axis=1)
    np.testing.assert_allclose(vv[idx], bv, rtol=1e-3, atol=1e-3)


def test_reorder_is_the_sorted_gather(ctx):
    sc = scenes.config2_grid(side=40)
    x = sc.x
    r = 2.0 * sc.spacing
    b = UniformGridBuilder(ctx, sc.bound_min, sc.bound_max, r, 0.5)

    class Owner:
        pass

    o = Owner()
    o.xx, o.xy, o.xz = _upload_xyz(ctx, x)
    extra = np.arange(x.shape[0], dtype=np.float32) * 0.5
    o.e = ctx.array(extra)
    g = b.construct_and_reorder(r, [(o, "xx"), (o, "xy"), (o, "xz"), (o, "e")])
    assert not g.sorted
    geom = orc.geometry(sc.bound_min, sc.bound_max, r, 0.5)
    ref = orc.grid_build(geom, x[:, 0], x[:, 1], x[:, 2])
    np.testing.assert_array_equal(o.xx.numpy(), x[ref.sorted, 0])
    np.testing.assert_array_equal(o.xz.numpy(), x[ref.sorted, 2])
    np.testing.assert_array_equal(o.e.numpy(), extra[ref.sorted])
    prefix, _ = grid_tables(ctx, g)
    np.testing.assert_array_equal(prefix, ref.prefix)


def test_grid_full_size_properties(ctx):
    """1M-point config-2 scene: counting-sort invariants at full size."""
    sc = scenes.config2_grid(side=100)
    x = sc.x
    r = 2.0 * sc.spacing
    for ratio in (0.5, 1.0, 2.0):
        b, g, _, prefix, sorted_, counts, keys = _gpu_construct(ctx, x, sc.bound_min, sc.bound_max, r, ratio)
        n = x.shape[0]
        assert prefix[-1] == n
        assert np.all(np.diff(prefix.astype(np.int64)) >= 0)
        assert np.array_equal(np.sort(sorted_), np.arange(n, dtype=np.uint32))  # permutation
        ks = keys[sorted_]
        assert np.all(np.diff(ks.astype(np.int64)) >= 0)  # cell order
        same = ks[1:] == ks[:-1]
        assert np.all(sorted_[1:][same] > sorted_[:-1][same])  # ascending index inside a cell
        geom = orc.geometry(sc.bound_min, sc.bound_max, r, ratio)
        ref = orc.grid_build(geom, x[:, 0], x[:, 1], x[:, 2])
        np.testing.assert_array_equal(sorted_, ref.sorted)
        np.testing.assert_array_equal(prefix, ref.prefix)


def test_out_of_bounds_points_are_clamped_and_reported(ctx):
    x = scenes.lattice((8, 8, 8), 0.1)
    x[3] = (np.nan, 0.2, 0.2)
    x[7] = (50.0, 0.2, 0.2)
    b = UniformGridBuilder(ctx, (0, 0, 0), (0.8, 0.8, 0.8), 0.2, 0.5)
    xx, xy, xz = _upload_xyz(ctx, x)
    g = capi.Grid()
    rc = ctx._lib.reyn_grid_construct(ctx.h, b.h, C.c_float(0.2), xx._ptr, xy._ptr, xz._ptr, x.shape[0], C.byref(g))
    assert rc == capi.ERR_OUT_OF_BOUNDS
    assert "2 points" in ctx.last_error()
    prefix, sorted_ = grid_tables(ctx, g)
    assert prefix[-1] == x.shape[0]
    assert np.array_equal(np.sort(sorted_), np.arange(x.shape[0]))


def test_deferred_check_reports_what_construct_would(ctx):
    """Deferred build checks: construct returns at once and the check returns
    the error the waiting call returns (clamped points), once; a construct
    after an unchecked failed build returns that build's error and builds
    nothing; switching deferral off checks a pending build; the tables are
    those of a waiting build."""
    bad = scenes.lattice((8, 8, 8), 0.1)
    bad[3] = (np.nan, 0.2, 0.2)
    bad[7] = (50.0, 0.2, 0.2)
    good = scenes.lattice((8, 8, 8), 0.1) + 0.01
    n = bad.shape[0]
    b = UniformGridBuilder(ctx, (0, 0, 0), (0.8, 0.8, 0.8), 0.2, 0.5)
    ref = UniformGridBuilder(ctx, (0, 0, 0), (0.8, 0.8, 0.8), 0.2, 0.5)
    assert ctx._lib.reyn_grid_builder_check(ctx.h, b.h) == capi.OK  # nothing pending
    ctx.check(ctx._lib.reyn_grid_builder_defer_check(b.h, 1))
    xb, xg = _upload_xyz(ctx, bad), _upload_xyz(ctx, good)

    def construct(xs):
        g = capi.Grid()
        rc = ctx._lib.reyn_grid_construct(ctx.h, b.h, C.c_float(0.2), xs[0]._ptr, xs[1]._ptr, xs[2]._ptr, n,
                                          C.byref(g))
        return rc, g

    rc, g = construct(xb)
    assert rc == capi.OK
    assert ctx._lib.reyn_grid_builder_check(ctx.h, b.h) == capi.ERR_OUT_OF_BOUNDS
    assert "2 points" in ctx.last_error()
    assert ctx._lib.reyn_grid_builder_check(ctx.h, b.h) == capi.OK  # reported once
    prefix, sorted_ = grid_tables(ctx, g)
    g_ref = capi.Grid()
    assert ctx._lib.reyn_grid_construct(ctx.h, ref.h, C.c_float(0.2), xb[0]._ptr, xb[1]._ptr, xb[2]._ptr, n,
                                        C.byref(g_ref)) == capi.ERR_OUT_OF_BOUNDS
    prefix_ref, sorted_ref = grid_tables(ctx, g_ref)
    np.testing.assert_array_equal(prefix, prefix_ref)
    np.testing.assert_array_equal(sorted_, sorted_ref)
    # an unchecked failed build: the next construct reports it and builds nothing
    assert construct(xb)[0] == capi.OK
    assert construct(xg)[0] == capi.ERR_OUT_OF_BOUNDS
    assert ctx._lib.reyn_grid_builder_check(ctx.h, b.h) == capi.OK
    rc, g = construct(xg)
    assert rc == capi.OK and ctx._lib.reyn_grid_builder_check(ctx.h, b.h) == capi.OK
    ref_g = ref.construct(0.2, *xg)
    for a, r in zip(grid_tables(ctx, g), grid_tables(ctx, ref_g)):
        np.testing.assert_array_equal(a, r)
    # switching deferral off checks what is pending, and construct waits again
    assert construct(xb)[0] == capi.OK
    assert ctx._lib.reyn_grid_builder_defer_check(b.h, 0) == capi.ERR_OUT_OF_BOUNDS
    assert construct(xb)[0] == capi.ERR_OUT_OF_BOUNDS


def test_grid_keeps_its_builder_alive(ctx):
    """A grid from a builder that nothing else references still points at live
    tables: the returned grid holds its builder (checked on the host before
    the grid is used on the device)."""
    import gc

    x = scenes.lattice((8, 8, 8), 0.1) + 0.01
    xx, xy, xz = _upload_xyz(ctx, x)
    g = UniformGridBuilder(ctx, (0, 0, 0), (0.8, 0.8, 0.8), 0.2, 0.5).construct(0.2, xx, xy, xz)
    gc.collect()
    assert g._builder.h, "the builder was destroyed while its grid lives"
    prefix, sorted_ = grid_tables(ctx, g)
    assert prefix[-1] == x.shape[0] and np.array_equal(np.sort(sorted_), np.arange(x.shape[0]))


def test_builder_rejects_bad_geometry(ctx):
    h = C.c_void_p()
    rc = ctx._lib.reyn_grid_builder_create(ctx.h, capi.f3((1, 1, 1)), capi.f3((0, 0, 0)), C.c_float(0.1),
                                           C.c_float(0.5), C.byref(h))
    assert rc == capi.ERR_INVALID_ARGUMENT
    assert "Negative number of grid cells" in ctx.last_error()


@pytest.mark.parametrize("kernel", ["B3", "C2", "W6", "COS"])
def test_kernel_properties_reference_inputs(ctx, kernel):
    """src/kernels.cu:19-181 on its own inputs (h_bar = 25, 25^3 samples,
    srand(16142069)), plus FP64-oracle parity."""
    kind = capi.KERNELS[kernel]
    h_bar = 25.0
    d = orc.reference_kernel_points(h_bar)
    n = d.shape[0]
    dx, dy, dz = (ctx.array(np.ascontiguousarray(d[:, a])) for a in range(3))
    w, gx, gy, gz = (ctx.empty(n) for _ in range(4))
    ctx.check(ctx._lib.reyn_kernel_eval(ctx.h, kind, C.c_float(h_bar), dx._ptr, dy._ptr, dz._ptr, n, w._ptr,
                                        gx._ptr, gy._ptr, gz._ptr))
    W = w.numpy().astype(np.float64)
    G = np.stack([gx.numpy(), gy.numpy(), gz.numpy()], axis=1).astype(np.float64)
    w64, g64 = orc.kernel_eval(kind, h_bar, d.astype(np.float64))
    norm = np.linalg.norm(d.astype(np.float64), axis=1)
    # positivity, compact support
    assert np.all(W >= -1e-12)
    out = norm >= h_bar
    gscale = np.abs(g64[np.isfinite(g64)]).max()
    assert np.all(np.abs(W[out]) < 1e-9)
    assert np.all(np.linalg.norm(G[out], axis=1) < 1e-6 * gscale)
    # FP64 parity, relative to the kernel's scale
    assert np.max(np.abs(W - w64)) <= 1e-5 * w64.max()
    assert np.max(np.abs(G - g64)) <= 1e-5 * gscale
    assert W[0] > 0 and np.all(G[0] == 0)
    # regular lattice: antisymmetry, sum W = 1/spacing^3, sum grad W = 0
    side = 25
    sp = h_bar * 2.0 / side
    i = np.arange(side, dtype=np.float32)
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    grid = (np.float32(-h_bar + sp / 2) + np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) * np.float32(sp))
    grid = grid.astype(np.float32)
    dx.upload(grid[:, 0]), dy.upload(grid[:, 1]), dz.upload(grid[:, 2])
    ctx.check(ctx._lib.reyn_kernel_eval(ctx.h, kind, C.c_float(h_bar), dx._ptr, dy._ptr, dz._ptr, n, w._ptr,
                                        gx._ptr, gy._ptr, gz._ptr))
    Wg = w.numpy()
    Gg = np.stack([gx.numpy(), gy.numpy(), gz.numpy()], axis=1)
    np.testing.assert_allclose(-Gg, Gg[::-1], rtol=1e-4, atol=1e-9)
    assert Wg.astype(np.float64).sum() == pytest.approx(1.0 / sp ** 3, rel=0.005)
    assert np.all(np.abs(Gg.astype(np.float64).sum(axis=0)) <= 0.01)


@pytest.mark.parametrize("ratio", [0.5, 1.0])
def test_boundary_occupancy(ctx, ratio):
    """near[c] is 1 exactly when a boundary sample lies in the (2R+1)^3 cells
    around c (numpy over the same cell starts), and a density pass with the
    flags gives the same bits as one that walks the boundary grid for every
    particle."""
    import ctypes as C

    from oracle.parity import GpuCase, make_case
    from reyn_amd.sim import sph_params

    sc, v = make_case(side=12, seed=3)
    case = GpuCase(ctx, sc, v, ratio)
    g = case.bdy.grid
    nx, ny, nz, R = g.nx, g.ny, g.nz, g.range
    prefix, _ = grid_tables(ctx, g)
    counts = np.diff(prefix.astype(np.int64)).reshape(nz, ny, nx) > 0
    pad = np.pad(counts, R)
    want = np.zeros_like(counts)
    for dz in range(2 * R + 1):
        for dy in range(2 * R + 1):
            for dx in range(2 * R + 1):
                want |= pad[dz:dz + nz, dy:dy + ny, dx:dx + nx]
    near = case.bdy.near.numpy().reshape(nz, ny, nx).astype(bool)
    assert np.array_equal(near, want)
    assert (~near).any() and near.any()
    gf = case.grid()
    prm = sph_params("B3", sc.spacing, 0.001, sc.rho_0)
    out = []
    for use_near in (True, False):
        view = case.bdy.view()
        if not use_near:
            view.near = None
        rho = ctx.empty(sc.n)
        fl = case.state.view()
        ctx.check(ctx._lib.reyn_density(ctx.h, C.byref(prm), C.byref(fl), C.byref(gf), C.byref(view), rho.ptr),
                  "density")
        out.append(rho.numpy())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("ratio", [0.5, 1.0, 2.0])
def test_brick_order_tables_bitexact_and_counts_equal_brute_force(ctx, ratio):
    """The brick cell order on the reference test's points: keys, prefix and
    sorted indices bit-exact with the oracle's brick mode, and every neighbour
    count equal to brute force (the XYZ counts)."""
    x = orc.reference_grid_points(50, 1.1)
    r = np.float32(2.0 * np.float32(1.1))
    lo, hi = np.zeros(3, np.float32), np.full(3, np.float32(1.1) * 50, np.float32)
    b, g, (xx, xy, xz), prefix, sorted_, counts, keys = _gpu_construct(ctx, x, lo, hi, r, ratio, "brick")
    geom = orc.geometry(lo, hi, r, ratio, order=1)
    ref = orc.grid_build(geom, x[:, 0], x[:, 1], x[:, 2])
    assert g.order == capi.CELL_ORDER_BRICK and (g.nx, g.ny, g.nz) == tuple(geom.n)
    assert g.nx % 2 == 0 and g.ny % 2 == 0 and g.nz % 2 == 0
    np.testing.assert_array_equal(keys, ref.keys)
    np.testing.assert_array_equal(prefix, ref.prefix)
    np.testing.assert_array_equal(sorted_, ref.sorted)
    n = x.shape[0]
    cnt, l2, v = (ctx.empty(n, np.uint32), ctx.empty(n), [ctx.empty(n) for _ in range(3)])
    ctx.check(ctx._lib.reyn_grid_query_stats(ctx.h, C.byref(g), xx._ptr, xy._ptr, xz._ptr, n, cnt._ptr, l2._ptr,
                                             v[0]._ptr, v[1]._ptr, v[2]._ptr))
    idx = np.random.default_rng(4).choice(n, 20000, replace=False).astype(np.uint32)
    bc, _, _ = orc.brute_stats_subset(r * r, x[:, 0], x[:, 1], x[:, 2], idx)
    np.testing.assert_array_equal(cnt.numpy()[idx], bc)
    # the reordered query over the brick order equals the oracle's brick walk
    o_cnt, _, _ = orc.query_stats(ref, x[:, 0], x[:, 1], x[:, 2])
    np.testing.assert_array_equal(cnt.numpy(), o_cnt)


def test_brick_order_occupancy(ctx):
    """Occupancy flags over a brick-ordered grid: near[key] is 1 exactly when a
    point lies in the (2R+1)^3 cells around that key's cell."""
    rng = np.random.default_rng(8)
    x = rng.uniform(0.2, 0.5, size=(3000, 3)).astype(np.float32)
    r = np.float32(0.05)
    b, g, _, prefix, _, _, _ = _gpu_construct(ctx, x, (0, 0, 0), (1, 1, 1), r, 0.5, "brick")
    nx, ny, nz, R = g.nx, g.ny, g.nz, g.range
    near = ctx.empty(nx * ny * nz, np.uint8)
    ctx.check(ctx._lib.reyn_grid_occupancy(ctx.h, C.byref(g), near._ptr), "occupancy")
    cz, cy, cx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    key = (((cz >> 1) * (ny // 2) + (cy >> 1)) * (nx // 2) + (cx >> 1)) * 8 + (cz & 1) * 4 + (cy & 1) * 2 + (cx & 1)
    occ = (np.diff(prefix.astype(np.int64))[key] > 0)
    pad = np.pad(occ, R)
    want = np.zeros_like(occ)
    for dz in range(2 * R + 1):
        for dy in range(2 * R + 1):
            for dx in range(2 * R + 1):
                want |= pad[dz:dz + nz, dy:dy + ny, dx:dx + nx]
    got = near.numpy()[key].astype(bool)
    assert np.array_equal(got, want) and want.any() and (~want).any()
