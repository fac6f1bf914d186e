"""Hand-derived closed forms for the parts of the FP64 oracle that the pair and
lattice pins in test_oracle.py leave open: the boundary coupling of every
solver, the viscosity coefficient, the SESPH and PCISPH equations of state and
the velocity term of the PCISPH density prediction. CPU only.

Each expected value is computed here from the published kernel forms (the
numpy functions below, themselves checked against the exact rational known
answers) and the reference's equations, for inputs small enough that the
solver reduces to a formula: one fluid particle next to two boundary samples,
or one pair. None of them goes through the oracle's own kernel helpers. Every
parameter that crosses the oracle's interface as a float (spacing, support,
nu, rho_0, g, stiffness, dt) is rounded to float here as well.

`oracle/mutations.py` plants one mistake at a time in the oracle source (a
dropped boundary term, a 1/2 factor taken away, an unnormalised EOS, 10 nu
changed, ...) and checks that this file fails for each.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import binding as orc

GOLDEN = Path(__file__).resolve().parent / "golden"

C2, B3, W6 = 0, 1, 2
ALPHA = {C2: 21.0 / (2.0 * np.pi), B3: 16.0 / np.pi, W6: 1365.0 / (64.0 * np.pi)}


def f32(v):
    return float(np.float32(v))


def shape(kind, q):
    """w(q) and dw/dq of the published forms (src/kernels.cuh:75-145)."""
    t = max(0.0, 1.0 - q)
    if kind == C2:
        return t ** 4 * (4 * q + 1), -20 * q * t ** 3
    if kind == B3:
        u = min(max(0.5 - q, 0.0), 1.0)
        return t ** 3 - 4 * u ** 3, -3 * t ** 2 + 12 * u ** 2
    p = 32 * q ** 3 + 25 * q ** 2 + 8 * q + 1
    return t ** 8 * p, -22 * q * (16 * q * q + 7 * q + 1) * t ** 7


def W(kind, hb, d):
    r = float(np.linalg.norm(d))
    return ALPHA[kind] / hb ** 3 * shape(kind, min(r / hb, 1.0))[0]


def gradW(kind, hb, d):
    d = np.asarray(d, np.float64)
    r = float(np.linalg.norm(d))
    if r == 0.0:
        return np.zeros(3)
    return ALPHA[kind] / hb ** 4 * shape(kind, r / hb)[1] / r * d


def test_numpy_shapes_match_exact_known_answers():
    ka = json.loads((GOLDEN / "kernel_known_answers.json").read_text())
    for name, kind in (("C2", C2), ("B3", B3), ("W6", W6)):
        for q, w, dw in zip(ka["q"], ka["kernels"][name]["w"], ka["kernels"][name]["dw"]):
            got = shape(kind, float(q))
            assert got[0] == pytest.approx(float(w), rel=1e-14, abs=1e-15)
            assert got[1] == pytest.approx(float(dw), rel=1e-12, abs=1e-13)


# ---- scenes -------------------------------------------------------------------------------

H = 0.05  # particle spacing; kernel support 2h
HB = f32(2 * H)
X0 = np.array([0.5, 0.5, 0.5], np.float32)
# two boundary samples below and beside the fluid particle, within its support
XB = np.array([[0.5, 0.47, 0.5], [0.54, 0.455, 0.52]], np.float32)
MB = np.array([0.4, 0.3])


def fluid_grid(x):
    return orc.grid_build(orc.geometry((0, 0, 0), (1, 1, 1), 2 * H, 0.5), x[:, 0], x[:, 1], x[:, 2])


def boundary(xb=XB, mb=MB):
    """The boundary samples in the cell order of their own grid, as the
    solvers take them (a Resort::yes boundary grid)."""
    r = 2 * H
    g = orc.grid_build(orc.geometry(xb.min(0), xb.max(0), r, 0.5), xb[:, 0], xb[:, 1], xb[:, 2])
    xs, ms = xb[g.sorted], np.asarray(mb, np.float64)[g.sorted]
    g = orc.grid_build(orc.geometry(xs.min(0), xs.max(0), r, 0.5), xs[:, 0], xs[:, 1], xs[:, 2]).resorted()
    return orc.Bdy(g, xs[:, 0], xs[:, 1], xs[:, 2], ms)


def state(x, v=None, m=0.125):
    x = np.asarray(x, np.float32)
    v = np.zeros_like(x) if v is None else np.asarray(v, np.float32)
    m = np.broadcast_to(np.asarray(m, np.float64), (x.shape[0],))
    return orc.State({"x": x[:, 0], "y": x[:, 1], "z": x[:, 2], "vx": v[:, 0], "vy": v[:, 1], "vz": v[:, 2],
                      "m": m})


def xij(a, b):
    # positions are float32 values; their difference is exact in float64
    return np.asarray(a, np.float64) - np.asarray(b, np.float64)


def bsums(kind, x):
    """sum_k m_k W_ik and sum_k m_k grad W_ik of the two boundary samples."""
    rb = sum(MB[k] * W(kind, HB, xij(x, XB[k])) for k in range(2))
    gb = sum(MB[k] * gradW(kind, HB, xij(x, XB[k])) for k in range(2))
    return rb, gb


# ---- density with a boundary (solver.cuh:18-38) ----------------------------------------------

@pytest.mark.parametrize("kind", [C2, B3, W6])
def test_boundary_density_closed_form(kind):
    """rho_i = m_i W(0) + sum_k m_k W_ik (compute_densities, solver.cuh:18-38)."""
    m = 0.125
    x = X0[None, :]
    rho = orc.density(fluid_grid(x), state(x, m=m), boundary(), kind, HB)
    rb, _ = bsums(kind, X0)
    assert rb > 0.2 * m * W(kind, HB, np.zeros(3))  # both samples contribute visibly
    np.testing.assert_allclose(rho, [m * W(kind, HB, np.zeros(3)) + rb], rtol=1e-12)


# ---- viscosity (write_non_prs_acc_ff, solver.cuh:40-69) -------------------------------------

def test_viscosity_of_a_moving_pair_closed_form():
    """a_i = g + sum_j 10 nu (m_j / rho_j) (v_ij . x_ij) / (|x_ij|^2 + 0.01 h^2)
    grad W_ij, with h the particle spacing; unequal masses and densities tell
    m_j / rho_j from m_i / rho_i."""
    kind, nu = C2, 0.05
    x = np.array([[0.5, 0.5, 0.5], [0.53, 0.52, 0.49]], np.float32)
    v = np.array([[0.3, -0.2, 0.1], [-0.4, 0.1, 0.25]], np.float32)
    m = np.array([0.5, 0.7])
    rho = np.array([900.0, 1150.0])
    prm = orc.sph(kind, HB, H, nu, 1000.0)
    a = orc.nonpressure(fluid_grid(x), state(x, v, m), prm, rho)
    g = np.array([0.0, f32(-9.81), 0.0])
    c, eps = 10.0 * f32(nu), 0.01 * f32(H) ** 2
    for i, j in ((0, 1), (1, 0)):
        d = xij(x[i], x[j])
        vij = xij(v[i], v[j])
        s = c * m[j] / rho[j] * vij.dot(d) / (d.dot(d) + eps)
        np.testing.assert_allclose(a[i], g + s * gradW(kind, HB, d), rtol=1e-12)
    # the viscous term is far above rounding, and 0.01 h^2 is not negligible
    assert np.abs(a[0] - g).max() > 1e-3
    d = xij(x[0], x[1])
    assert eps / d.dot(d) > 1e-3


# ---- SESPH (SESPH.cu:6-95) ---------------------------------------------------------------------

def sesph_eos(rho, k, rho0):
    return max(0.0, k * (rho / rho0 - 1.0))


def test_sesph_boundary_pressure_closed_form():
    """One fluid particle pressed onto two boundary samples: rho = m W(0) +
    sum_k m_k W_ik, p = max(0, k (rho / rho0 - 1)) (SESPH.cu:9) and
    a = g - (p / rho^2) sum_k m_k grad W_ik (SESPH.cu:57-61); then
    v += dt a, x += dt v."""
    kind, m, rho0, k, dt = C2, 0.125, 1000.0, 1000.0, 1e-3
    v0 = np.array([0.1, -0.3, 0.05], np.float32)
    x = X0[None, :]
    st = state(x, v0[None, :], m)
    prm = orc.sph(kind, HB, H, 0.001, rho0)
    rho, a = orc.sesph_step(fluid_grid(x), st, boundary(), prm, k, dt)
    rb, gb = bsums(kind, X0)
    r = m * W(kind, HB, np.zeros(3)) + rb
    p = sesph_eos(r, k, rho0)
    assert r > 1.05 * rho0 and p > 0
    g = np.array([0.0, f32(-9.81), 0.0])
    expect = g - p / r ** 2 * gb
    np.testing.assert_allclose(rho, [r], rtol=1e-12)
    np.testing.assert_allclose(a[0], expect, rtol=1e-12)
    assert np.abs(expect - g).max() > 1.0  # the wall pushes visibly
    dt = f32(dt)
    v = v0.astype(np.float64) + dt * expect
    np.testing.assert_allclose([st.vx[0], st.vy[0], st.vz[0]], v, rtol=1e-12)
    np.testing.assert_allclose([st.x[0], st.y[0], st.z[0]], X0.astype(np.float64) + dt * v, rtol=1e-12)


def test_sesph_pair_pressure_and_viscosity_closed_form():
    """A compressed, moving pair without boundary: rho = m (W(0) + W_01) for
    both, p = max(0, k (rho / rho0 - 1)), and
    a_i = g + (visc_ij - m (p / rho^2 + p / rho^2)) grad W_ij
    (SESPH.cu:14-65, the normalised EOS of :9)."""
    kind, m, rho0, k, nu = C2, 0.6, 1000.0, 2.0e4, 0.01
    x = np.array([[0.5, 0.5, 0.5], [0.53, 0.52, 0.49]], np.float32)
    v = np.array([[0.3, 0.0, 0.1], [-0.2, 0.1, 0.0]], np.float32)
    prm = orc.sph(kind, HB, H, nu, rho0)
    rho, a = orc.sesph_step(fluid_grid(x), state(x, v, m), None, prm, k, 1e-4)
    d = xij(x[0], x[1])
    r = m * (W(kind, HB, np.zeros(3)) + W(kind, HB, d))
    p = sesph_eos(r, k, rho0)
    assert r > rho0 and p > 0
    np.testing.assert_allclose(rho, [r, r], rtol=1e-12)
    g = np.array([0.0, f32(-9.81), 0.0])
    c, eps = 10.0 * f32(nu), 0.01 * f32(H) ** 2
    for i, j in ((0, 1), (1, 0)):
        dij = xij(x[i], x[j])
        visc = c * m / r * xij(v[i], v[j]).dot(dij) / (dij.dot(dij) + eps)
        expect = g + (visc - m * (p / r ** 2 + p / r ** 2)) * gradW(kind, HB, dij)
        np.testing.assert_allclose(a[i], expect, rtol=1e-12)


# ---- PCISPH (PCISPH.cu:15-111, solver.cuh:71-105) -------------------------------------------

def test_pcisph_boundary_prediction_and_pressure_closed_form():
    """One fluid particle at the wall, two loop passes (no fluid neighbour:
    the non-pressure acceleration is g). Pass l: v += dt a (a = g, then the
    previous pressure acceleration); rho* = m W(0) + sum_k m_k (W_ik + dt
    grad W_ik . v) with the boundary at rest (solver.cuh:71-105);
    p = max(0, kk (rho* - rho0)) with kk = (delta / dt) / dt (PCISPH.cu:15-19,
    :65); a = -1/2 (p / rho*^2 + p / rho0^2) sum_k m_k grad W_ik
    (PCISPH.cu:49-53). The step ends with v += dt a, x += dt v."""
    kind, m, rho0, dt, delta = W6, 0.125, 1000.0, 5e-4, 1.0e-4
    v0 = np.array([0.05, -0.8, 0.02], np.float32)
    x = X0[None, :]
    st = state(x, v0[None, :], m)
    prm = orc.sph(kind, HB, H, 0.001, rho0)
    it, rho, a, avg = orc.pcisph_step(fluid_grid(x), st, boundary(), prm, delta, dt, fixed_iters=2)
    dt = f32(dt)
    kk = (delta / dt) / dt
    rb, gb = bsums(kind, X0)
    w0 = m * W(kind, HB, np.zeros(3))
    g = np.array([0.0, f32(-9.81), 0.0])
    v = v0.astype(np.float64)
    acc = g
    for _ in range(2):
        v = v + dt * acc
        r = w0 + rb + dt * gb.dot(v)
        p = max(0.0, kk * (r - rho0))
        acc = -0.5 * (p / r ** 2 + p / rho0 ** 2) * gb
    assert it == 2 and p > 0 and abs(dt * gb.dot(v)) > 1e-3 * r
    np.testing.assert_allclose(rho, [r], rtol=1e-11)
    assert avg == pytest.approx(r, rel=1e-11)
    np.testing.assert_allclose(a[0], acc, rtol=1e-11)
    v = v + dt * acc
    np.testing.assert_allclose([st.vx[0], st.vy[0], st.vz[0]], v, rtol=1e-11)
    np.testing.assert_allclose([st.x[0], st.y[0], st.z[0]], X0.astype(np.float64) + dt * v, rtol=1e-11)


def test_pcisph_pair_prediction_closed_form():
    """An approaching pair, no boundary, nu = 0 and g = 0 (so the first pass
    leaves v unchanged): rho*_i = m (W(0) + W_ij + dt grad W_ij . v_ij)
    (solver.cuh:88-104) and a_i = -m (p / rho*^2 + p / rho*^2) grad W_ij with
    the unnormalised EOS p = max(0, kk (rho* - rho0)) (PCISPH.cu:15-19)."""
    kind, m, rho0, dt, delta = W6, 0.6, 1000.0, 1e-3, 2.0e-3
    x = np.array([[0.5, 0.5, 0.5], [0.56, 0.51, 0.5]], np.float32)
    v = np.array([[0.9, 0.2, 0.0], [-0.7, 0.0, 0.1]], np.float32)
    prm = orc.sph(kind, HB, H, 0.0, rho0, g=(0.0, 0.0, 0.0))
    it, rho, a, _ = orc.pcisph_step(fluid_grid(x), state(x, v, m), None, prm, delta, dt, fixed_iters=1)
    dt = f32(dt)
    kk = (delta / dt) / dt
    d = xij(x[0], x[1])
    vij = xij(v[0], v[1])
    gw = gradW(kind, HB, d)
    r = m * (W(kind, HB, np.zeros(3)) + W(kind, HB, d) + dt * gw.dot(vij))
    p = max(0.0, kk * (r - rho0))
    assert it == 1 and p > 0 and dt * gw.dot(vij) > 0.01 * W(kind, HB, d)
    np.testing.assert_allclose(rho, [r, r], rtol=1e-12)
    np.testing.assert_allclose(a[0], -m * (2 * p / r ** 2) * gw, rtol=1e-12)
    np.testing.assert_allclose(a[1], m * (2 * p / r ** 2) * gw, rtol=1e-12)


# ---- IISPH (IISPH.cu:10-262) -------------------------------------------------------------------

def test_iisph_boundary_terms_closed_form():
    """One fluid particle moving onto two boundary samples, cold start, three
    Jacobi iterations. With no fluid neighbour: v = v0 + dt g after the
    non-pressure pass; G = sum_k m_k grad W_ik; a_ii = -dt |G|^2 / rho0^2 and
    s = (rho0 - rho) / dt - v . G (IISPH.cu:48-65); the pressure acceleration
    is a = -(p / rho0^2) G (:84-95) and Ap = dt a . G = a_ii p (:136-147), so
    the relaxed Jacobi iterate is p <- p + max(0, omega (s / a_ii - p)) from
    p = max(0, omega s / a_ii) (:108-117, :199-215); the step ends with
    v += dt a(p), x += dt v."""
    kind, m, rho0, omega, dt = B3, 0.125, 1000.0, 0.5, 1e-3
    v0 = np.array([0.05, -0.6, 0.0], np.float32)
    x = X0[None, :]
    st = state(x, v0[None, :], m)
    prm = orc.sph(kind, HB, H, 0.001, rho0)
    ip = orc.OrcIisph(omega, 0.005, 3, 0, 0, 0)
    res = orc.iisph_step(fluid_grid(x), st, boundary(), prm, ip, np.zeros(1), dt, fixed_iters=3)
    dt, omega = f32(dt), f32(omega)
    rb, gb = bsums(kind, X0)
    r = m * W(kind, HB, np.zeros(3)) + rb
    v = v0.astype(np.float64) + dt * np.array([0.0, f32(-9.81), 0.0])
    a_ii = -dt * gb.dot(gb) / rho0 ** 2
    s = (rho0 - r) / dt - v.dot(gb)
    assert r > rho0 and s < 0 and abs(v.dot(gb)) > 0.05 * abs(s)
    p = max(0.0, omega * s / a_ii)
    err = 0.0
    for _ in range(3):
        err = (a_ii * p - s) * dt
        p = p + max(0.0, omega * (s / a_ii - p))
    np.testing.assert_allclose(res["rho"], [r], rtol=1e-12)
    np.testing.assert_allclose(res["a_ii"], [a_ii], rtol=1e-12)
    np.testing.assert_allclose(res["s_i"], [s], rtol=1e-12)
    np.testing.assert_allclose(res["p"], [p], rtol=1e-10)
    assert res["mean_rel_error"] == pytest.approx(max(err, 0.0) / rho0, rel=1e-9)
    acc = -(p / rho0 ** 2) * gb
    np.testing.assert_allclose(res["a"][0], acc, rtol=1e-10)
    v = v + dt * acc
    np.testing.assert_allclose([st.vx[0], st.vy[0], st.vz[0]], v, rtol=1e-10)
    np.testing.assert_allclose([st.x[0], st.y[0], st.z[0]], X0.astype(np.float64) + dt * v, rtol=1e-10)


def test_iisph_pair_at_the_wall_diagonal_and_source_closed_form():
    """A fluid pair next to the boundary samples: the diagonal uses the square
    of the combined gradient sum G_i = m grad W_ij + sum_k m_k grad W_ik plus
    m_i m_j |grad W_ij|^2 (IISPH.cu:26-61), and the source term the fluid and
    boundary velocity divergences (:63-65); nu = 0 so v = v0 + dt g."""
    kind, m, rho0, dt = B3, 0.125, 1000.0, 1e-3
    x = np.array([X0, [0.52, 0.51, 0.49]], np.float32)
    v0 = np.array([[0.1, -0.5, 0.0], [-0.2, 0.3, 0.1]], np.float32)
    prm = orc.sph(kind, HB, H, 0.0, rho0)
    ip = orc.OrcIisph(0.5, 0.005, 3, 0, 0, 0)
    res = orc.iisph_step(fluid_grid(x), state(x, v0, m), boundary(), prm, ip, np.zeros(2), dt, fixed_iters=1)
    dt = f32(dt)
    v = v0.astype(np.float64) + dt * np.array([0.0, f32(-9.81), 0.0])
    for i, j in ((0, 1), (1, 0)):
        d = xij(x[i], x[j])
        gw = gradW(kind, HB, d)
        rb, gb = bsums(kind, x[i])
        r = m * (W(kind, HB, np.zeros(3)) + W(kind, HB, d)) + rb
        G = m * gw + gb
        a_ii = -dt * (G.dot(G) + m * m * gw.dot(gw)) / rho0 ** 2
        s = (rho0 - r) / dt - (m * gw.dot(v[i] - v[j]) + v[i].dot(gb))
        assert abs(2 * (m * gw).dot(gb)) > 0.05 * G.dot(G)  # the cross term matters
        np.testing.assert_allclose(res["rho"][i], r, rtol=1e-12)
        np.testing.assert_allclose(res["a_ii"][i], a_ii, rtol=1e-12)
        np.testing.assert_allclose(res["s_i"][i], s, rtol=1e-12)
