"""Pins of the oracle (oracle/reyn_oracle.cpp), each by something other than
its own formula: numpy restatements with library routines, brute force,
exact rational known answers, closed forms and invariants. CPU only."""

import json
from pathlib import Path

import numpy as np
import pytest
from scipy import integrate

from oracle import binding as orc

GOLDEN = Path(__file__).resolve().parent / "golden"
KINDS = {"C2": 0, "B3": 1, "W6": 2, "COS": 3}


# ---- grid ------------------------------------------------------------------------

def numpy_grid(x, lo, hi, r, ratio):
    """The grid in numpy: float32 geometry and keys, np.bincount counts,
    np.cumsum prefix and a stable argsort for the in-cell order."""
    f32 = np.float32
    cell = f32(r) * f32(ratio)
    pad = f32(2.0 * float(r))
    blo = np.asarray(lo, f32) - pad
    bhi = np.asarray(hi, f32) + pad
    n = np.ceil((bhi - blo) / cell).astype(np.int64)
    with np.errstate(invalid="ignore"):
        q = np.floor((x.astype(f32) - blo) / cell)
    q = np.where(np.isnan(q), 0, q)
    q = np.clip(q, 0, n - 1).astype(np.int64)
    keys = q[:, 2] * n[0] * n[1] + q[:, 1] * n[0] + q[:, 0]
    counts = np.bincount(keys, minlength=int(np.prod(n)))
    prefix = np.concatenate([[0], np.cumsum(counts)])
    order = np.argsort(keys, kind="stable")
    return n, blo, cell, keys, counts, prefix, order


@pytest.mark.parametrize("ratio", [0.5, 1.0, 2.0])
def test_grid_build_matches_numpy_restatement(ratio):
    rng = np.random.default_rng(5)
    x = rng.uniform(-0.05, 1.05, size=(20000, 3)).astype(np.float32)
    x[:50] = x[:50].round(1)  # points on cell faces
    x[60] = (np.nan, 0.5, 0.5)
    x[61] = (1e30, -1e30, 0.5)
    lo, hi, r = (0, 0, 0), (1, 1, 1), np.float32(0.05)
    n, blo, cell, keys, counts, prefix, order = numpy_grid(x, lo, hi, r, ratio)
    geom = orc.geometry(lo, hi, r, ratio)
    assert tuple(geom.n) == tuple(n) and geom.cell == cell and np.allclose(geom.bmin, blo, rtol=0, atol=0)
    assert geom.range == int(np.ceil(1 / ratio - 1e-9))
    g = orc.grid_build(geom, x[:, 0], x[:, 1], x[:, 2])
    np.testing.assert_array_equal(g.keys, keys)
    np.testing.assert_array_equal(g.counts, counts)
    np.testing.assert_array_equal(g.prefix, prefix)
    np.testing.assert_array_equal(g.sorted, order)
    assert g.outside >= 2


def test_reference_points_brute_force_agreement():
    """The reference's grid test on its own points: the grid walk and O(N^2)
    brute force give equal neighbour counts (uniformgrid.cu:340)."""
    x = orc.reference_grid_points(50, 1.1)
    r = np.float32(2.0 * np.float32(1.1))
    lo, hi = np.zeros(3, np.float32), np.full(3, np.float32(1.1) * 50, np.float32)
    geom = orc.geometry(lo, hi, r, 0.5)
    g = orc.grid_build(geom, x[:, 0], x[:, 1], x[:, 2])
    cnt, l2, v = orc.query_stats(g, x[:, 0], x[:, 1], x[:, 2])
    idx = np.random.default_rng(1).choice(x.shape[0], 3000, replace=False).astype(np.uint32)
    bc, bl2, bv = orc.brute_stats_subset(r * r, x[:, 0], x[:, 1], x[:, 2], idx)
    np.testing.assert_array_equal(cnt[idx], bc)
    np.testing.assert_allclose(l2[idx], bl2, rtol=1e-4)
    # every point lies in the box, as the reference test asserts (uniformgrid.cu:332-337)
    assert x.min() >= 0 and x.max() <= np.float32(1.1) * 50


def test_brute_force_matches_float64_emulated_fma():
    """The neighbour test fmaf(x,x, fmaf(y,y, z*z)) <= r^2 in FP32, emulated in
    numpy float64 (products of floats are exact in double)."""
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 3, size=(1500, 3)).astype(np.float32)
    r2 = np.float32(0.3) * np.float32(0.3)
    cnt, _, _ = orc.brute_stats(r2, x[:, 0], x[:, 1], x[:, 2])
    d = (x[:, None, :] - x[None, :, :]).astype(np.float32)  # FP32 differences
    zz = (d[..., 2] * d[..., 2]).astype(np.float32)
    yy = (d[..., 1].astype(np.float64) ** 2 + zz).astype(np.float32)
    d2 = (d[..., 0].astype(np.float64) ** 2 + yy).astype(np.float32)
    ref = (d2 <= r2).sum(axis=1) - 1
    np.testing.assert_array_equal(cnt, ref)


@pytest.mark.parametrize("ratio", [0.5, 1.0, 2.0])
def test_neighbour_lists_are_the_brute_force_sets(ratio):
    rng = np.random.default_rng(9)
    x = rng.uniform(0, 1, size=(4000, 3)).astype(np.float32)
    r = np.float32(0.08)
    g = orc.grid_build(orc.geometry((0, 0, 0), (1, 1, 1), r, ratio), x[:, 0], x[:, 1], x[:, 2])
    off, lst = orc.neighbour_lists(g, x[:, 0], x[:, 1], x[:, 2])
    cnt, _, _ = orc.brute_stats(r * r, x[:, 0], x[:, 1], x[:, 2])
    np.testing.assert_array_equal(np.diff(off.astype(np.int64)) - 1, cnt)  # self included in the lists
    for i in rng.choice(4000, 50, replace=False):
        d = x - x[i]
        brute = set(np.nonzero(np.einsum("ij,ij->i", d.astype(np.float64), d.astype(np.float64)) <= r * r * (1 + 1e-6))[0])
        got = set(lst[off[i]:off[i + 1]].tolist())
        assert got <= brute and len(got) >= len(brute) - 2  # edge pairs may round either way in FP64


# ---- smoothing kernels ----------------------------------------------------------------

def test_shape_functions_match_exact_known_answers():
    ka = json.loads((GOLDEN / "kernel_known_answers.json").read_text())
    qs = [float(q) for q in ka["q"]]
    for name, kind in KINDS.items():
        w = [float(v) for v in ka["kernels"][name]["w"]]
        dw = [float(v) for v in ka["kernels"][name]["dw"]]
        for q, a, b in zip(qs, w, dw):
            assert orc.shape_w(kind, q) == pytest.approx(a, rel=1e-14, abs=1e-14), (name, q)
            assert orc.shape_dw(kind, q) == pytest.approx(b, rel=1e-12, abs=1e-13), (name, q)


@pytest.mark.parametrize("name", list(KINDS))
def test_kernel_is_normalised_and_gradient_is_derivative(name):
    kind = KINDS[name]
    h = 0.7
    # integral over R^3 of W = 4 pi int_0^h r^2 W(r) dr = 1
    w = lambda r: orc.kernel_eval(kind, h, np.array([[r, 0.0, 0.0]]))[0][0]
    val, _ = integrate.quad(lambda r: 4 * np.pi * r * r * w(r), 0, h, limit=200)
    assert val == pytest.approx(1.0, rel=1e-9)
    # grad W = dW/dr x/|x|: central differences along a random direction
    rng = np.random.default_rng(kind)
    for _ in range(20):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        r = rng.uniform(0.05, 0.95) * h
        eps = 1e-6
        wp = orc.kernel_eval(kind, h, np.array([(r + eps) * u]))[0][0]
        wm = orc.kernel_eval(kind, h, np.array([(r - eps) * u]))[0][0]
        _, g = orc.kernel_eval(kind, h, np.array([r * u]))
        np.testing.assert_allclose(g[0], (wp - wm) / (2 * eps) * u, rtol=1e-6, atol=1e-9 * abs(wp) / h)


@pytest.mark.parametrize("name", list(KINDS))
def test_kernel_closed_form_values_and_support(name):
    kind = KINDS[name]
    h = 2.0
    alpha = {"C2": 21 / (2 * np.pi), "B3": 16 / np.pi, "W6": 1365 / (64 * np.pi),
             "COS": np.pi / (4 * np.pi ** 2 - 30)}[name]
    w0 = {"C2": 1.0, "B3": 0.5, "W6": 1.0, "COS": 8.0}[name]
    W, G = orc.kernel_eval(kind, h, np.array([[0.0, 0, 0], [h, 0, 0], [0, 0, 1.5 * h], [1e30, 0, 0]]))
    assert W[0] == pytest.approx(alpha * w0 / h ** 3, rel=1e-14)
    assert np.all(W[1:] == 0) and np.all(G[0] == 0)
    assert np.all(np.abs(G[1:]) < 1e-12)


@pytest.mark.parametrize("name", list(KINDS))
def test_kernel_reference_property_test_float_port(name):
    """src/kernels.cu:19-181 on its own inputs, for the float port: host
    (float) and FP64 agree within the reference's 0.005; lattice sums."""
    kind = KINDS[name]
    hb = 25.0
    d = orc.reference_kernel_points(hb).astype(np.float64)
    w32, g32 = orc.kernel_eval(kind, hb, d.astype(np.float32), dtype=np.float32)
    w64, g64 = orc.kernel_eval(kind, hb, d)
    assert np.all(w32 >= 0)
    # the reference's Approx(..).epsilon(0.005) has an absolute floor; here it
    # is 1e-5 of W(0), which covers FP32 cancellation near q = 1 (COS)
    ok = np.abs(w32 - w64) <= 0.005 * np.abs(w64) + 1e-5 * w64.max()
    assert ok.all()
    side, sp = 25, hb * 2.0 / 25
    i = np.arange(side)
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    lat = -hb + sp / 2 + np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) * sp
    W, G = orc.kernel_eval(kind, hb, lat)
    assert W.sum() == pytest.approx(1 / sp ** 3, rel=0.005)
    np.testing.assert_allclose(G, -G[::-1], atol=1e-15)
    assert np.all(np.abs(G.sum(axis=0)) <= 0.01)


# ---- phases -------------------------------------------------------------------------------

def small_grid(x, r, ratio=0.5, lo=None, hi=None):
    lo = x.min(axis=0) if lo is None else lo
    hi = x.max(axis=0) if hi is None else hi
    g = orc.grid_build(orc.geometry(lo, hi, r, ratio), x[:, 0], x[:, 1], x[:, 2])
    return g


def state_of(x, v=None, m=1.0):
    n = x.shape[0]
    v = np.zeros_like(x) if v is None else v
    return orc.State({"x": x[:, 0], "y": x[:, 1], "z": x[:, 2], "vx": v[:, 0], "vy": v[:, 1], "vz": v[:, 2],
                      "m": np.full(n, m)})


def test_density_two_particles_closed_form():
    h = 0.1
    x = np.array([[0.5, 0.5, 0.5], [0.5 + 0.06, 0.5, 0.5]], np.float32)
    g = small_grid(x, 2 * h, lo=np.zeros(3), hi=np.ones(3))
    rho = orc.density(g, state_of(x, m=2.0), None, 1, 2 * h)
    # B3: W(0) = 8 / (pi h_bar^3), W(d) = 16/(pi h_bar^3) ((1-q)^3 - 4 (1/2-q)_+^3)
    hb = float(np.float32(2 * h))  # the oracle takes h_bar as a float
    W0 = 16 / np.pi / hb ** 3 * 0.5
    d = np.float64(np.float32(0.56) - np.float32(0.5))
    Wd = 16 / np.pi / hb ** 3 * ((1 - d / hb) ** 3 - 4 * max(0.5 - d / hb, 0) ** 3)
    np.testing.assert_allclose(rho, 2.0 * (W0 + Wd), rtol=1e-12)


def test_density_lattice_interior_is_rest_density():
    s = 0.05
    from reyn_amd import scenes
    x = scenes.lattice((20, 20, 20), s)
    g = small_grid(x, 2 * s)
    rho0 = 1000.0
    rho = orc.density(g, state_of(x, m=rho0 * s ** 3), None, 1, 2 * s)
    centre = np.argmin(np.linalg.norm(x - x.mean(axis=0), axis=1))
    assert rho[centre] == pytest.approx(rho0, rel=0.01)


def test_sesph_free_fall_and_pair_momentum():
    h, dt = 0.05, 1e-3
    prm = orc.sph(0, 2 * h, h, 0.001, 1000.0)
    # isolated particle: a = g, v = g dt, x += dt v
    x = np.array([[0.5, 0.5, 0.5]], np.float32)
    st = state_of(x, m=0.125)
    g = small_grid(x, 2 * h, lo=np.zeros(3), hi=np.ones(3))
    rho, a = orc.sesph_step(g, st, None, prm, 1000.0, dt)
    g32, dt32 = float(np.float32(-9.81)), float(np.float32(dt))  # the oracle takes both as floats
    np.testing.assert_allclose(a[0], [0, g32, 0], rtol=1e-15)
    assert st.vy[0] == pytest.approx(g32 * dt32, rel=1e-15)
    assert st.y[0] == pytest.approx(0.5 + dt32 * dt32 * g32, rel=1e-14)
    # compressed pair with opposite velocities: internal forces cancel
    x = np.array([[0.5, 0.5, 0.5], [0.53, 0.52, 0.49]], np.float32)
    v = np.array([[0.3, 0.0, 0.1], [-0.2, 0.1, 0.0]], np.float32)
    st = state_of(x, v, m=0.6)
    g = small_grid(x, 2 * h, lo=np.zeros(3), hi=np.ones(3))
    rho, a = orc.sesph_step(g, st, None, prm, 1.0e5, dt)
    assert rho[0] > 1000.0  # compressed, so the pressure term is active
    np.testing.assert_allclose(a.sum(axis=0), [0, 2 * -9.81, 0], atol=1e-9 * np.abs(a).max())
    assert np.abs(a[0] - np.array([0, -9.81, 0])).max() > 1.0


def test_pcisph_delta_closed_form():
    """On the symmetric 5^3 template lattice sum grad W = 0, so
    delta = 1 / (2 V0^2 sum |grad W|^2) (PCISPH.cuh:49-77)."""
    import ctypes as C

    from reyn_amd import capi

    h = float(np.float32(0.02))  # the oracle takes h and h_bar as floats
    hb = float(np.float32(0.04))
    for kind in range(4):
        lat = np.array([[-x * h, -y * h, -z * h] for x in range(-2, 3) for y in range(-2, 3) for z in range(-2, 3)])
        _, G = orc.kernel_eval(kind, hb, lat)
        assert np.abs(G.sum(axis=0)).max() < 1e-9 * np.abs(G).max()
        expect = 1.0 / (2 * (h ** 3) ** 2 * (G ** 2).sum())
        assert orc.pcisph_delta_f64(kind, hb, h) == pytest.approx(expect, rel=1e-10)
        # the library's host-side stiffness (float arithmetic, as the reference)
        d = C.c_double()
        assert capi.lib().reyn_pcisph_delta(kind, hb, h, C.byref(d)) == capi.OK
        assert d.value == pytest.approx(expect, rel=2e-6)


def test_iisph_pair_diagonal_closed_form():
    """Two particles, no boundary: A = -(2 dt m^2 |grad W|^2 / rho0^2)[[1,1],[1,1]],
    so a_ii = -2 dt m^2 |grad W_ij|^2 / rho0^2 (IISPH.cu:10-66)."""
    h, dt, m, rho0 = 0.05, 1e-3, 0.125, 1000.0
    x = np.array([[0.5, 0.5, 0.5], [0.57, 0.5, 0.5]], np.float32)
    st = state_of(x, m=m)
    g = small_grid(x, 2 * h, lo=np.zeros(3), hi=np.ones(3))
    prm = orc.sph(1, 2 * h, h, 0.0, rho0, g=(0.0, 0.0, 0.0))
    ip = orc.OrcIisph(0.5, 0.005, 3, 0, 0, 0)
    res = orc.iisph_step(g, st, None, prm, ip, np.zeros(2), dt, fixed_iters=3)
    d = float(np.float32(0.57) - np.float32(0.5))
    _, G = orc.kernel_eval(1, 2 * h, np.array([[-d, 0, 0]]))
    dt32 = float(np.float32(dt))  # dt crosses the oracle's interface as a float
    expect = -2 * dt32 * m * m * (G[0] ** 2).sum() / rho0 ** 2
    np.testing.assert_allclose(res["a_ii"], [expect, expect], rtol=1e-12)
    # at rest with rho < rho0 the source term is positive and the pressure stays 0
    assert np.all(res["s_i"] > 0) and np.all(res["p"] == 0)


def test_iisph_float_port_tracks_fp64():
    """The FP32 port (the CPU baseline) agrees with the FP64 oracle on one step."""
    from reyn_amd import scenes
    sc = scenes.dam_break((10, 10, 10), 0.05, (1.0, 1.0, 1.0), jitter=0.05, seed=4)
    r = 2 * sc.spacing
    g = small_grid(sc.x, r, lo=sc.bound_min, hi=sc.bound_max)
    x = sc.x[g.sorted]
    g = small_grid(x, r, lo=sc.bound_min, hi=sc.bound_max).resorted()
    xb = sc.xb
    bg = orc.grid_build(orc.geometry(xb.min(0), xb.max(0), r, 0.5), xb[:, 0], xb[:, 1], xb[:, 2])
    xb = xb[bg.sorted]
    bg = orc.grid_build(orc.geometry(xb.min(0), xb.max(0), r, 0.5), xb[:, 0], xb[:, 1], xb[:, 2]).resorted()
    mb = orc.boundary_masses(bg, xb[:, 0], xb[:, 1], xb[:, 2], sc.spacing, sc.rho_0, 5)
    prm = orc.sph(1, r, sc.spacing, 0.001, sc.rho_0)
    ip = orc.OrcIisph(0.5, 0.005, 3, 0, 0, 1)
    out = {}
    for dt_ in (np.float64, np.float32):
        st = state_of(x, m=sc.mass)
        st = orc.State({k: getattr(st, k) for k in orc.State.FIELDS}, dtype=dt_)
        b = orc.Bdy(bg, xb[:, 0], xb[:, 1], xb[:, 2], mb, dtype=dt_)
        out[dt_] = orc.iisph_step(g, st, b, prm, ip, np.zeros(sc.n), 2e-3, fixed_iters=5)
        out[dt_]["y"] = st.y.copy()
    for k in ("rho", "p"):
        a, b = out[np.float64][k], out[np.float32][k]
        assert np.abs(a - b).max() <= 1e-4 * np.abs(a).max(), k


def test_iisph_pair_pressure_iteration_closed_form():
    """A compressed pair at rest, no gravity, viscosity or boundary, cold
    start: by symmetry p_0 = p_1 = p, a_0 = -a_1 = -(2 m p / rho0^2) grad W_01,
    (Ap)_i = 2 a_ii p, so the relaxed Jacobi iterate is p <- p + max(0,
    omega (s / a_ii - 2 p)) from p = omega s / a_ii (IISPH.cu:68-162,
    :199-215), and the step ends with x_0 += dt^2 a_0 (solver.cu:15-32)."""
    h, dt, m, rho0, omega = 0.05, 1e-3, 0.6, 1000.0, 0.3
    x = np.array([[0.5, 0.5, 0.5], [0.56, 0.5, 0.5]], np.float32)
    st = state_of(x, m=m)
    g = small_grid(x, 2 * h, lo=np.zeros(3), hi=np.ones(3))
    prm = orc.sph(1, 2 * h, h, 0.0, rho0, g=(0.0, 0.0, 0.0))
    ip = orc.OrcIisph(omega, 0.005, 3, 0, 0, 0)
    res = orc.iisph_step(g, st, None, prm, ip, np.zeros(2), dt, fixed_iters=3)
    f32 = lambda v: float(np.float32(v))  # parameters cross the interface as floats
    dt, omega = f32(dt), f32(omega)
    d = float(np.float32(0.56) - np.float32(0.5))
    W, G = orc.kernel_eval(1, 2 * h, np.array([[0.0, 0, 0], [-d, 0, 0]]))
    gw2 = (G[1] ** 2).sum()
    rho = m * (W[0] + W[1])
    a_ii = -2 * dt * m * m * gw2 / rho0 ** 2
    c = ((rho0 - rho) / dt) / a_ii
    assert c > 0  # compressed: a positive pressure solves the pair
    p = omega * c
    for _ in range(3):
        p = p + max(0.0, omega * (c - 2 * p))
    np.testing.assert_allclose(res["rho"], [rho, rho], rtol=1e-12)
    np.testing.assert_allclose(res["p"], [p, p], rtol=1e-10)
    a0 = -(2 * m * p / rho0 ** 2) * G[1]
    np.testing.assert_allclose(st.x[0] - 0.5, dt * dt * a0[0], rtol=1e-6)
    np.testing.assert_allclose(st.x[1] - float(np.float32(0.56)), -dt * dt * a0[0], rtol=1e-6)


def test_pcisph_free_fall_and_pair_momentum():
    """PCISPH step (PCISPH.cu:58-111): gravity and viscosity enter the
    velocity in the first loop pass and `a` ends as the last pressure
    acceleration. An isolated particle is below rest density, so its pressure
    is 0: after exactly min_iter passes a = 0, v = g dt, y += dt^2 g. A
    compressed pair's pressure and viscosity forces are equal and opposite:
    the final pressure accelerations cancel and over all passes the momentum
    gains exactly 2 m g dt."""
    h, dt = 0.05, 1e-3
    prm = orc.sph(2, 2 * h, h, 0.001, 1000.0)
    delta = orc.pcisph_delta_f64(2, 2 * h, h)
    g32, dt32 = float(np.float32(-9.81)), float(np.float32(dt))
    x = np.array([[0.5, 0.5, 0.5]], np.float32)
    st = state_of(x, m=0.125)
    g = small_grid(x, 2 * h, lo=np.zeros(3), hi=np.ones(3))
    it, rho, a, avg = orc.pcisph_step(g, st, None, prm, delta, dt)
    assert it == 3
    np.testing.assert_array_equal(a[0], [0, 0, 0])
    assert st.vy[0] == pytest.approx(g32 * dt32, rel=1e-15)
    assert st.y[0] == pytest.approx(float(np.float32(0.5)) + dt32 * dt32 * g32, rel=1e-14)
    x = np.array([[0.5, 0.5, 0.5], [0.53, 0.52, 0.49]], np.float32)
    v = np.array([[0.3, 0.0, 0.1], [-0.2, 0.1, 0.0]], np.float32)
    st = state_of(x, v, m=0.6)
    g = small_grid(x, 2 * h, lo=np.zeros(3), hi=np.ones(3))
    p0 = v.astype(np.float64).sum(axis=0)
    it, rho, a, avg = orc.pcisph_step(g, st, None, prm, delta, dt, max_iter=20)
    assert abs(st.vx[0] - st.vx[1]) > 1.0  # the pressure pushed the pair apart
    np.testing.assert_allclose(a.sum(axis=0), 0.0, atol=1e-12)
    p1 = np.array([st.vx.sum(), st.vy.sum(), st.vz.sum()])
    np.testing.assert_allclose(p1 - p0, [0, 2 * g32 * dt32, 0], atol=1e-12)


def test_gather_is_the_permutation():
    """The reorder gather (buffer.cuh:184-209) is numpy's fancy indexing."""
    rng = np.random.default_rng(5)
    src = rng.normal(size=1000).astype(np.float32)
    perm = rng.permutation(1000).astype(np.uint32)
    np.testing.assert_array_equal(orc.gather(perm, src), src[perm])
    assert orc.gather(np.zeros(0, np.uint32), np.zeros(0, np.float32)).size == 0


def test_boundary_masses_fixed_point():
    """On a large flat 2-layer wall the refined mass of a central sample tends
    to rho0 / sum_j W_bj (the Akinci fixed point, sample_boundary.cu:251-297)."""
    from reyn_amd import scenes
    s = 0.05
    xb = scenes.lattice((40, 2, 40), s)
    r = 2 * s
    g = orc.grid_build(orc.geometry(xb.min(0), xb.max(0), r, 0.5), xb[:, 0], xb[:, 1], xb[:, 2])
    xb = xb[g.sorted]
    g = orc.grid_build(orc.geometry(xb.min(0), xb.max(0), r, 0.5), xb[:, 0], xb[:, 1], xb[:, 2]).resorted()
    m = orc.boundary_masses(g, xb[:, 0], xb[:, 1], xb[:, 2], s, 1000.0, 30)
    c = np.argmin(np.linalg.norm(xb - xb.mean(0), axis=1))
    W, _ = orc.kernel_eval(1, r, (xb - xb[c]).astype(np.float64))
    assert m[c] == pytest.approx(1000.0 / W.sum(), rel=1e-6)


def test_cfl_closed_forms():
    h, lam = 0.02, 0.4
    dt, vmax = orc.cfl_time_step(lam, h, np.zeros((10, 3)))
    assert vmax == 0 and dt == pytest.approx(np.sqrt(2 * lam * h / 9.81), rel=1e-6)
    v = np.zeros((10, 3), np.float32)
    v[3] = (3.0, 4.0, 0.0)
    dt, vmax = orc.cfl_time_step(lam, h, v)
    assert vmax == 5.0 and dt == pytest.approx(lam * h / 5.0, rel=1e-6)


def test_bounds_hand_cases():
    lo, hi = (0, 0, 0), (1, 1, 1)
    x = np.array([[0.5, 0.5, 0.5], [0.0, 0.5, 0.5], [1.2, -0.1, 0.5], [0.5, 1.0, 0.2]], np.float32)
    v = np.ones((4, 3), np.float32)
    nx, nv = orc.hard_enforce_bounds(lo, hi, x, v)
    np.testing.assert_array_equal(nx, np.array([[0.5, 0.5, 0.5], [0.0, 0.5, 0.5], [1.0, 0.0, 0.5], [0.5, 1.0, 0.2]],
                                               np.float32))
    c = np.float32(-0.1)
    np.testing.assert_array_equal(nv, [[1, 1, 1], [c, c, c], [c, c, c], [c, c, c]])


def test_walk_capacity_rule_library_and_oracle():
    """The most points per cell a grid build accepts: (2^15 candidates) /
    (2 range + 1)^3, the same in the library (host-only call) and the oracle,
    whose build rejects one point more."""
    import ctypes as C

    from reyn_amd import capi

    for rng in (1, 2, 3, 4):
        cap = C.c_uint32()
        assert capi.lib().reyn_grid_cell_capacity(rng, C.byref(cap)) == capi.OK
        w = (2 * rng + 1) ** 3
        assert cap.value == orc.grid_cell_capacity(rng) == (1 << 15) // w
        assert cap.value * w <= 1 << 15 < (cap.value + 1) * w
    assert capi.lib().reyn_grid_cell_capacity(0, C.byref(cap)) == capi.ERR_INVALID_ARGUMENT
    r = np.float32(0.1)
    geom = orc.geometry((0, 0, 0), (1, 1, 1), r, 1.0)
    cap = orc.grid_cell_capacity(geom.range)
    pts = np.random.default_rng(2).uniform(0.5, 0.501, size=(cap + 1, 3)).astype(np.float32)
    orc.grid_build(geom, pts[:cap, 0], pts[:cap, 1], pts[:cap, 2])
    with pytest.raises(ValueError):
        orc.grid_build(geom, pts[:, 0], pts[:, 1], pts[:, 2])


def numpy_brick_grid(x, lo, hi, r, ratio):
    """The brick-ordered grid in numpy: the XYZ geometry with every cell count
    rounded up to even, key 8 brick + 4 (z & 1) + 2 (y & 1) + (x & 1) with the
    bricks in XYZ order, np.bincount counts and a stable argsort."""
    f32 = np.float32
    cell = f32(r) * f32(ratio)
    pad = f32(2.0 * float(r))
    blo = np.asarray(lo, f32) - pad
    bhi = np.asarray(hi, f32) + pad
    n = np.ceil((bhi - blo) / cell).astype(np.int64)
    n += n & 1
    with np.errstate(invalid="ignore"):
        q = np.floor((x.astype(f32) - blo) / cell)
    q = np.where(np.isnan(q), 0, q)
    q = np.clip(q, 0, n - 1).astype(np.int64)
    b = q >> 1
    nb = n // 2
    brick = (b[:, 2] * nb[1] + b[:, 1]) * nb[0] + b[:, 0]
    keys = brick * 8 + (q[:, 2] & 1) * 4 + (q[:, 1] & 1) * 2 + (q[:, 0] & 1)
    counts = np.bincount(keys, minlength=int(np.prod(n)))
    prefix = np.concatenate([[0], np.cumsum(counts)])
    return n, keys, counts, prefix, np.argsort(keys, kind="stable")


@pytest.mark.parametrize("ratio", [0.5, 1.0, 2.0])
def test_brick_grid_build_matches_numpy_restatement(ratio):
    rng = np.random.default_rng(15)
    x = rng.uniform(-0.05, 1.05, size=(20000, 3)).astype(np.float32)
    x[:50] = x[:50].round(1)
    lo, hi, r = (0, 0, 0), (1, 1, 1), np.float32(0.05)
    n, keys, counts, prefix, order = numpy_brick_grid(x, lo, hi, r, ratio)
    geom = orc.geometry(lo, hi, r, ratio, order=1)
    assert tuple(geom.n) == tuple(n) and geom.order == 1
    g = orc.grid_build(geom, x[:, 0], x[:, 1], x[:, 2])
    np.testing.assert_array_equal(g.keys, keys)
    np.testing.assert_array_equal(g.counts, counts)
    np.testing.assert_array_equal(g.prefix, prefix)
    np.testing.assert_array_equal(g.sorted, order)
    # the keys are a permutation of the cells: every cell has exactly one key
    cells = np.stack(np.meshgrid(*[np.arange(int(v)) for v in n], indexing="ij"), -1).reshape(-1, 3)
    b = cells >> 1
    nb = n // 2
    all_keys = ((b[:, 2] * nb[1] + b[:, 1]) * nb[0] + b[:, 0]) * 8 + (cells[:, 2] & 1) * 4 + (cells[:, 1] & 1) * 2 \
        + (cells[:, 0] & 1)
    np.testing.assert_array_equal(np.sort(all_keys), np.arange(int(np.prod(n))))


@pytest.mark.parametrize("ratio", [0.5, 1.0])
def test_brick_order_neighbour_sets_equal_xyz_order(ratio):
    """The walk over the brick-ordered grid finds, for every point, exactly the
    neighbours the XYZ walk finds (the sets; the visiting order differs)."""
    rng = np.random.default_rng(19)
    x = rng.uniform(0, 1, size=(4000, 3)).astype(np.float32)
    r = np.float32(0.08)
    out = {}
    for order in (0, 1):
        g = orc.grid_build(orc.geometry((0, 0, 0), (1, 1, 1), r, ratio, order=order), x[:, 0], x[:, 1], x[:, 2])
        off, lst = orc.neighbour_lists(g, x[:, 0], x[:, 1], x[:, 2])
        out[order] = [set(lst[off[i]:off[i + 1]].tolist()) for i in range(x.shape[0])]
    assert out[0] == out[1]
    cnt, _, _ = orc.brute_stats(r * r, x[:, 0], x[:, 1], x[:, 2])
    assert [len(s) - 1 for s in out[1]] == cnt.tolist()
