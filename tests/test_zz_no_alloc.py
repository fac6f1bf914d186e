"""reyn_sesph_step_noalloc, the SESPH step that allocates nothing, through
the C-ABI: on two calls in a row (the second with a smaller scene, so the
static scratch is reused) its results equal reyn_sesph_step's on the same
input, rho and x within 1e-5 and a and v within 1e-4 of each field's largest
value (the two order their neighbour sums differently); inputs above its
capacity are refused with REYN_ERR_CAPACITY."""

import ctypes as C

import numpy as np
import pytest

from oracle.parity import make_case, rel_err
from reyn_amd import capi
from reyn_amd.sim import SESPH, BoundarySamples, Particles, UniformGridBuilder

pytestmark = pytest.mark.gpu

TOL = 1e-5
TOL_DIFF = 1e-4


def one_step(ctx, sc, v, entry):
    h, r = sc.spacing, 2.0 * sc.spacing
    state = Particles(ctx, sc.x, v, sc.mass, sc.rho_0)
    bdy = BoundarySamples(ctx, sc.xb, h, sc.rho_0)
    builder = UniformGridBuilder(ctx, sc.bound_min, sc.bound_max, r, 1.0)
    g = builder.construct_and_reorder(r, [(state, f) for f in Particles.FIELDS])
    solver = SESPH(ctx, "C2", sc.n, 0.001, h, sc.rho_0, k=1000.0)
    work = capi.SesphWork(solver.rho.ptr, solver.ax.ptr, solver.ay.ptr, solver.az.ptr)
    view, bview = state.view(), bdy.view()
    fn = getattr(ctx._lib, entry)
    ctx.check(fn(ctx.h, C.byref(solver.prm), C.byref(view), C.byref(g), C.byref(bview), C.byref(work), 1e-3), entry)
    ctx.synchronize()
    out = state.host()
    res = {"rho": solver.rho.numpy(), "a": np.stack([solver.ax.numpy(), solver.ay.numpy(), solver.az.numpy()], 1),
           "x": np.stack([out["xx"], out["xy"], out["xz"]], 1), "v": np.stack([out["vx"], out["vy"], out["vz"]], 1)}
    builder.close()
    return res


@pytest.mark.parametrize("sides", [(16, 10)])
def test_noalloc_sesph_step_equals_sesph_step(ctx, sides):
    for side in sides:
        sc, v = make_case(side=side, seed=31 + side)
        print(f"no-alloc SESPH step: {sc.n} fluid, {sc.n_bdy} boundary", flush=True)
        ref = one_step(ctx, sc, v, "reyn_sesph_step")
        got = one_step(ctx, sc, v, "reyn_sesph_step_noalloc")
        errs = {k: rel_err(got[k], ref[k]) for k in ref}
        print(f"  errors against reyn_sesph_step: {errs}", flush=True)
        assert errs["rho"] <= TOL and errs["x"] <= TOL, errs
        assert errs["a"] <= TOL_DIFF and errs["v"] <= TOL_DIFF, errs


def test_noalloc_rejects_inputs_above_capacity(ctx):
    n = (1 << 20) + 1
    x = np.random.default_rng(1).uniform(0.1, 0.9, size=(n, 3)).astype(np.float32)
    state = Particles(ctx, x, None, 1e-3, 1000.0)
    r = 0.05
    builder = UniformGridBuilder(ctx, (0, 0, 0), (1, 1, 1), r, 1.0)
    g = builder.construct_and_reorder(r, [(state, f) for f in Particles.FIELDS])
    solver = SESPH(ctx, "C2", n, 0.001, r / 2, 1000.0, k=1000.0)
    work = capi.SesphWork(solver.rho.ptr, solver.ax.ptr, solver.ay.ptr, solver.az.ptr)
    view = state.view()
    rc = ctx._lib.reyn_sesph_step_noalloc(ctx.h, C.byref(solver.prm), C.byref(view), C.byref(g), None,
                                          C.byref(work), 1e-3)
    assert rc == capi.ERR_CAPACITY
    assert b"static scratch" in ctx._lib.reyn_last_error(ctx.h)
    builder.close()
