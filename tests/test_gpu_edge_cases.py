"""Edge cases through the C ABI: empty fluid, a single particle, a scene
without boundary samples next to ordinary scenes in a batch, and batch input
that must be rejected."""

import math

import numpy as np
import pytest

from reyn_amd import capi, scenes
from reyn_amd.sim import (IISPH, PCISPH, SESPH, Batch, Particles, UniformGridBuilder, cfl_time_step,
                          hard_enforce_bounds)

pytestmark = pytest.mark.gpu

G = 9.81


def test_empty_fluid(ctx):
    """Zero particles: every call succeeds and changes nothing; the CFL step is
    the gravity limit alone (max |v| = 0)."""
    h, r = 0.05, 0.1
    state = Particles(ctx, np.zeros((0, 3), np.float32), None, 1.0, 1000.0)
    builder = UniformGridBuilder(ctx, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), r)
    builder.construct(r, state.xx, state.xy, state.xz)
    g = builder.construct_and_reorder(r, [(state, f) for f in Particles.FIELDS])
    SESPH(ctx, "C2", 0, 0.001, h, 1000.0).step(state, g, None, 1e-3)
    assert PCISPH(ctx, "W6", 0, 0.001, h, 1000.0).step(state, g, None, 1e-3) == 0
    assert IISPH(ctx, "B3", 0, 0.001, h, 1000.0).step(state, g, None, 1e-3) == 0
    dt, vmax = cfl_time_step(ctx, 0.4, h, state)
    assert vmax == 0.0
    assert math.isclose(dt, math.sqrt(2.0 * 0.4 * h / G), rel_tol=1e-6)
    hard_enforce_bounds(ctx, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), state)
    builder.close()


def _free_particle(box=4.0, at=(2.0, 3.0, 2.0), spacing=0.05):
    x = np.array([at], np.float32)
    return scenes.Scene(x, np.zeros((0, 3), np.float32), spacing, 1000.0, np.zeros(3), np.full(3, box))


@pytest.mark.parametrize("solver,kernel", [("pcisph", "W6"), ("iisph", "B3")])
def test_free_particle_in_a_batch(ctx, solver, kernel):
    """A single particle with no neighbours and no boundary falls freely: each
    step adds dt g to its velocity and dt v to its position (the pressure is
    zero, the viscosity has no pair), in a batch beside ordinary scenes, which
    do not disturb it. Tolerance: one rounding per update (the device may fuse
    the multiply-add)."""
    lone = _free_particle()
    others = scenes.config5_jobs(2, seed=31, n_lo=1500, n_hi=3000)
    b = Batch(ctx, [others[0][0], lone, others[1][0]], solver=solver, kernel=kernel, rho_0=1000.0)
    y = np.float32(lone.x[0, 1])
    vy = np.float32(0.0)
    for _ in range(6):
        b.step(1)
        dt = np.float32(b.info(1).dt)
        vy = np.float32(vy + dt * np.float32(-G))
        y = np.float32(y + dt * vy)
    got = b.download(1)
    assert math.isclose(got["vy"][0], vy, rel_tol=1e-6) and math.isclose(got["xy"][0], y, rel_tol=1e-6)
    assert got["xx"][0] == lone.x[0, 0] and got["vx"][0] == 0.0 and got["vz"][0] == 0.0
    assert b.info(1).iterations == 3  # the minimum: nothing to correct
    b.close()


def test_batch_rejects_empty_and_bad_scenes(ctx):
    empty = scenes.Scene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), 0.05, 1000.0,
                         np.zeros(3), np.ones(3))
    with pytest.raises(capi.ReynError):
        Batch(ctx, [empty], solver="pcisph", kernel="W6")
    b = Batch(ctx, [_free_particle()], solver="iisph", kernel="B3", rho_0=1000.0)
    with pytest.raises(capi.ReynError):
        b.add([empty])
    # the failed add leaves the batch as it was
    assert b.ids() == [0]
    b.step(2)
    assert b.info(0).steps == 2
    b.close()


@pytest.mark.parametrize("ratio", [1.0, 0.5])
def test_cell_above_the_walk_capacity_is_rejected(ctx, ratio):
    """A cell holding more points than a neighbour walk may visit within its
    budget fails the build with REYN_ERR_CAPACITY; at the limit it builds."""
    import ctypes as C

    r = 0.1
    builder = UniformGridBuilder(ctx, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), r, ratio)
    rng = int(builder.geometry().range)
    cap = C.c_uint32()
    assert capi.lib().reyn_grid_cell_capacity(rng, C.byref(cap)) == capi.OK
    cap = int(cap.value)
    # points packed into one small cube inside one cell
    pts = np.random.default_rng(3).uniform(0.5, 0.5 + 0.01 * r, size=(cap + 1, 3)).astype(np.float32)
    print(f"ratio {ratio}: range {rng}, cell capacity {cap}", flush=True)
    at_cap = Particles(ctx, pts[:cap], None, 1.0, 1000.0)
    builder.construct(r, at_cap.xx, at_cap.xy, at_cap.xz)
    over = Particles(ctx, pts, None, 1.0, 1000.0)
    with pytest.raises(capi.ReynError) as e:
        builder.construct(r, over.xx, over.xy, over.xz)
    assert e.value.code == capi.ERR_CAPACITY
    builder.close()
