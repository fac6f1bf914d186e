"""Scene batches (reyn_batch_*): a scene advanced inside a batch gives the
same bits as the same scene alone, and a batch of one gives the same bits as
the single-scene entry points driven by the reference's main loop (grid +
reorder, controller time step, solver step, bounds)."""

import numpy as np
import pytest

from reyn_amd import capi, scenes
from reyn_amd.sim import IISPH, PCISPH, Batch, BoundarySamples, Particles, Simulation

pytestmark = pytest.mark.gpu

FIELDS = ("xx", "xy", "xz", "vx", "vy", "vz", "rho")


def small_scenes(count=4, seed=7):
    return scenes.config4_batch(count=count, seed=seed, n_lo=2000, n_hi=8000)


def assert_same(a, b, what):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), f"{what}: field {k} differs (max |d| {np.abs(a[k] - b[k]).max()})"


@pytest.mark.parametrize("solver,kernel,order", [("pcisph", "W6", "xyz"), ("iisph", "B3", "xyz"),
                                                 ("iisph", "B3", "brick")])
def test_scene_alone_equals_scene_in_batch(ctx, solver, kernel, order):
    scs = small_scenes()
    print(f"{solver}, {order} order: scenes of {[sc.n for sc in scs]} particles", flush=True)
    opts = dict(solver=solver, kernel=kernel, cell_order=order, cell_ratio=1.0 if order == "brick" else 0.5)
    mix = Batch(ctx, scs, **opts)
    mix.step(12)
    for s, sc in enumerate(scs):
        alone = Batch(ctx, [sc], **opts)
        alone.step(12)
        a, b = alone.download(0), mix.download(s)
        assert_same(a, b, f"scene {s}")
        ia, ib = alone.info(0), mix.info(s)
        assert (ia.steps, ia.iterations, ia.dt, ia.time) == (ib.steps, ib.iterations, ib.dt, ib.time)
        alone.close()
    mix.close()


@pytest.mark.parametrize("solver", ["pcisph", "iisph"])
def test_batch_of_one_equals_single_scene_path(ctx, solver):
    """PCISPH with the batch defaults; IISPH with the serving options (cell
    ratio 1, mean stop rule, warm start with the pressure reordered, check
    every 2), the equivalence the serving leg's plain baseline rests on."""
    sc = small_scenes(count=1, seed=11)[0]
    h = sc.spacing
    ratio = 0.5 if solver == "pcisph" else 1.0
    state = Particles(ctx, sc.x, None, sc.mass, sc.rho_0)
    bdy = BoundarySamples(ctx, sc.xb, h, sc.rho_0, cell_ratio=ratio)
    if solver == "pcisph":
        sol = PCISPH(ctx, "W6", sc.n, 0.001, h, sc.rho_0, eta_rho_max=0.001, min_iter=3, max_iter=100)
        extra, opts = (), dict(solver="pcisph", kernel="W6")
    else:
        sol = IISPH(ctx, "B3", sc.n, 0.001, h, sc.rho_0, omega=0.5, eta_rho_max=0.001, min_iter=3, max_iter=100,
                    stop_rule=capi.IISPH_STOP_MEAN, warm_start=True, check_every=2)
        extra, opts = ("p",), dict(solver="iisph", kernel="B3", cell_ratio=1.0, check_every=2)
    sim = Simulation(ctx, sol, state, bdy, h, sc.bound_min, sc.bound_max, cfl_lambda=1.0, cell_ratio=ratio,
                     extra_reorder=extra, dt_policy="controller", dt0=0.0005, iter_target=10, change_factor=0.001)
    b = Batch(ctx, [sc], **opts)
    for k in range(8):
        st = sim.step()
        b.step(1)
        info = b.info(0)
        assert (info.iterations, info.dt) == (st.iterations, np.float32(st.dt)), k
    got = b.download(0)
    ref = state.host()
    ref["rho"] = sol.rho.numpy()
    assert_same(got, ref, "batch of one vs single-scene path")
    if solver == "iisph":
        assert np.array_equal(got["p"], sol.p.numpy()), "pressure"
    b.close()


@pytest.mark.parametrize("solver,kernel", [("pcisph", "W6"), ("iisph", "B3")])
def test_graph_replays_match_direct_launches(ctx, solver, kernel):
    """The captured step (graphs for both state buffers, replayed every step
    with host checks in between) against the same batch launched directly,
    compared after every step."""
    scs = small_scenes(count=3, seed=5)
    direct = Batch(ctx, scs, solver=solver, kernel=kernel, graph=False)
    graphed = Batch(ctx, scs, solver=solver, kernel=kernel, graph=True)
    for k in range(9):
        direct.step(1)
        graphed.step(1)
        for s in range(len(scs)):
            assert_same(graphed.download(s), direct.download(s), f"step {k} scene {s}")
            a, b = graphed.info(s), direct.info(s)
            assert (a.iterations, a.dt) == (b.iterations, b.dt), (k, s)
    direct.close()
    graphed.close()


def test_batch_rejects_bad_input(ctx):
    sc = small_scenes(count=1)[0]
    with pytest.raises(capi.ReynError):
        Batch(ctx, [sc], solver="pcisph", kernel="W6", dt0=0.0)


@pytest.mark.parametrize("solver,kernel,graph", [("pcisph", "W6", False), ("iisph", "B3", False),
                                                 ("iisph", "B3", True)])
def test_scenes_joining_and_leaving_match_scenes_alone(ctx, solver, kernel, graph):
    """Continuous batching: scenes join and leave a running batch between
    steps; each scene's state after its own step count equals the scene run
    alone for that many steps (the layout is rebuilt at every change, and
    resident scenes carry their state, time step, warm-start pressure and
    iteration history across it)."""
    scs = small_scenes(count=4, seed=13)
    total = {0: 10, 1: 6, 2: 8, 3: 5}
    # (steps to take, scenes that join before them)
    plan = [(3, [0, 1]), (3, [2]), (2, []), (2, [3]), (1, []), (2, [])]
    mix = Batch(ctx, [], solver=solver, kernel=kernel, graph=graph, rho_0=scs[0].rho_0)
    ids, taken, results = {}, {}, {}
    for steps, join in plan:
        new = mix.add([scs[k] for k in join])
        for k, i in zip(join, new):
            ids[k] = i
            taken[k] = 0
        assert sorted(mix.ids()) == sorted(ids[k] for k in taken)
        mix.step(steps)
        for k in list(taken):
            taken[k] += steps
            if taken[k] == total[k]:
                results[k] = (mix.download(ids[k]), mix.info(ids[k]))
                mix.remove([ids[k]])
                del taken[k]
    assert sorted(results) == [0, 1, 2, 3] and not mix.ids()
    mix.step(1)  # an empty batch steps as a no-op
    for k, sc in enumerate(scs):
        alone = Batch(ctx, [sc], solver=solver, kernel=kernel)
        alone.step(total[k])
        got, info = results[k]
        assert_same(alone.download(0), got, f"scene {k}")
        ia = alone.info(0)
        assert (ia.steps, ia.iterations, ia.dt, ia.time) == (info.steps, info.iterations, info.dt, info.time)
        alone.close()
    mix.close()


def test_batch_remove_rejects_unknown_ids(ctx):
    scs = small_scenes(count=2, seed=3)
    b = Batch(ctx, scs, solver="pcisph", kernel="W6")
    with pytest.raises(capi.ReynError):
        b.remove([7])
    with pytest.raises(capi.ReynError):
        b.remove([0, 0])
    b.remove([0])
    assert b.ids() == [1]
    b.step(1)
    assert b.info(1).steps == 1
    b.close()


@pytest.mark.parametrize("solver,kernel", [("pcisph", "W6"), ("iisph", "B3")])
def test_checkpoint_resume_continues_exactly(ctx, tmp_path, solver, kernel):
    """A scene checkpointed after 5 steps (saved to disk and loaded back) and
    resumed in another batch, next to another scene, ends where the scene run
    for 11 steps without interruption ends, bit for bit."""
    from reyn_amd.io import Checkpoint

    scs = small_scenes(count=2, seed=17)
    straight = Batch(ctx, [scs[0]], solver=solver, kernel=kernel)
    straight.step(11)
    first = Batch(ctx, [scs[0]], solver=solver, kernel=kernel)
    first.step(5)
    ck = first.checkpoint(0)
    first.close()
    ck.save(tmp_path / "scene.npz")
    back = Checkpoint.load(tmp_path / "scene.npz")
    second = Batch(ctx, [scs[1]], solver=solver, kernel=kernel)
    second.step(2)
    sid = second.add([back])[0]
    assert second.info(sid).steps == 5
    second.step(6)
    assert_same(second.download(sid), straight.download(0), "resumed vs straight")
    a, b = second.info(sid), straight.info(0)
    assert (a.steps, a.iterations, a.dt, a.time) == (b.steps, b.iterations, b.dt, b.time)
    second.close()
    straight.close()


def test_step_traces(ctx):
    """The telemetry traces: one record per step with the step's time step and
    iteration count, for a single-scene simulation and for batch scenes."""
    from reyn_amd.telemetry import StepTrace, record_batch_step

    sc = small_scenes(count=1, seed=19)[0]
    h = sc.spacing
    state = Particles(ctx, sc.x, None, sc.mass, sc.rho_0)
    bdy = BoundarySamples(ctx, sc.xb, h, sc.rho_0)
    solver = PCISPH(ctx, "W6", sc.n, 0.001, h, sc.rho_0, eta_rho_max=0.001, min_iter=3, max_iter=100)
    trace = StepTrace()
    sim = Simulation(ctx, solver, state, bdy, h, sc.bound_min, sc.bound_max, cfl_lambda=1.0,
                     dt_policy="controller", dt0=0.0005, iter_target=10, change_factor=0.001, trace=trace)
    stats = [sim.step() for _ in range(5)]
    assert trace.steps == [1, 2, 3, 4, 5]
    assert trace.iterations == [s.iterations for s in stats] and trace.dt == [s.dt for s in stats]
    b = Batch(ctx, [sc], solver="pcisph", kernel="W6")
    traces = {}
    for _ in range(5):
        b.step(1)
        record_batch_step(b, traces)
    assert traces[0].steps == [1, 2, 3, 4, 5]
    # the batch of one takes the same steps as the single-scene path
    assert traces[0].iterations == trace.iterations
    assert [np.float32(v) for v in traces[0].dt] == [np.float32(v) for v in trace.dt]
    b.close()
