"""Python mirror of reyn's host API for the SPH step, over the C-ABI.

The names and call shapes follow the reference's C++ classes so that tests
and drivers read like the reference's own code:

* ``Particles``            src/particles.cuh:25-75
* ``UniformGridBuilder``   src/datastructure/uniformgrid.cuh:200-489
* ``BoundarySamples``      src/scene/scene.cuh:29-46
* ``SESPH``/``PCISPH``/``IISPH``  src/solvers/{SESPH,PCISPH,IISPH}.cuh
* ``cfl_time_step``/``simple_dt_controller``  src/timestep/cfl.cuh:13-32
* ``Scene.hard_enforce_bounds``/``Scene.get_grid``  src/scene/scene.cuh:136-156

Every call goes to hand-written HIP in ``libreyn_amd.so``; nothing here
computes on the CPU beyond argument marshalling.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi
from .device import Context, DeviceArray

GRAVITY = (0.0, -9.81, 0.0)


class Particles:
    """SoA fluid state: xx, xy, xz, vx, vy, vz, m (+ the reference's rho_0)."""

    FIELDS = ("xx", "xy", "xz", "vx", "vy", "vz", "m")

    def __init__(self, ctx: Context, x: np.ndarray, v: np.ndarray | None, m, rho_0: float):
        x = np.asarray(x, dtype=np.float32)
        n = x.shape[0]
        v = np.zeros((n, 3), np.float32) if v is None else np.asarray(v, dtype=np.float32)
        m = np.broadcast_to(np.asarray(m, dtype=np.float32), (n,))
        self.ctx = ctx
        self.n = n
        self.rho_0 = float(rho_0)
        cols = [x[:, 0], x[:, 1], x[:, 2], v[:, 0], v[:, 1], v[:, 2], m]
        for name, col in zip(self.FIELDS, cols):
            setattr(self, name, ctx.array(np.ascontiguousarray(col)))

    def view(self) -> capi.Particles:
        return capi.Particles(*(getattr(self, f).ptr for f in self.FIELDS), self.n)

    def positions(self) -> np.ndarray:
        return np.stack([self.xx.numpy(), self.xy.numpy(), self.xz.numpy()], axis=1)

    def velocities(self) -> np.ndarray:
        return np.stack([self.vx.numpy(), self.vy.numpy(), self.vz.numpy()], axis=1)

    def host(self) -> dict:
        return {f: getattr(self, f).numpy() for f in self.FIELDS}


class UniformGridBuilder:
    """Owns counts/prefix/sorted; ``construct`` and ``construct_and_reorder``
    return a ``capi.Grid`` view (uniformgrid.cu:65-181). A returned grid holds
    a reference to its builder, whose tables it points into; it stays valid
    until the next construct call on the same builder."""

    def __init__(self, ctx: Context, bound_min, bound_max, search_radius: float, cell_ratio: float = 0.5,
                 order: str | int = "xyz"):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(
            ctx._lib.reyn_grid_builder_create_ordered(
                ctx.h, capi.f3(bound_min), capi.f3(bound_max), float(search_radius), float(cell_ratio),
                cell_order_id(order), C.byref(h)
            ),
            "UniformGridBuilder",
        )
        self.h = h
        self.search_radius = float(search_radius)
        self._tmp: dict[int, list[DeviceArray]] = {}
        self.deferred = False
        ctx.adopt(self)

    def defer_check(self, on: bool = True) -> None:
        """Deferred build checks: construct calls return before the build's
        status is read, and ``check`` raises what they would have raised."""
        self.ctx.check(self.ctx._lib.reyn_grid_builder_defer_check(self.h, int(bool(on))), "defer_check")
        self.deferred = bool(on)

    def check(self) -> None:
        """Raises the error of the last deferred build, if it had one."""
        self.ctx.check(self.ctx._lib.reyn_grid_builder_check(self.ctx.h, self.h), "construct (deferred check)")

    def geometry(self) -> capi.Grid:
        g = capi.Grid()
        self.ctx.check(self.ctx._lib.reyn_grid_builder_geometry(self.h, C.byref(g)))
        return g

    def construct(self, search_radius: float, xx: DeviceArray, xy: DeviceArray, xz: DeviceArray) -> capi.Grid:
        if self.deferred:
            self.check()
        g = capi.Grid()
        self.ctx.check(
            self.ctx._lib.reyn_grid_construct(
                self.ctx.h, self.h, float(search_radius), xx._ptr, xy._ptr, xz._ptr, xx.n, C.byref(g)
            ),
            "construct",
        )
        g._builder = self  # the grid points into this builder's tables: keep it alive with the grid
        return g

    def construct_and_reorder(self, search_radius: float, fields) -> capi.Grid:
        """Reorders every ``(owner, name)`` array in ``fields`` (x, y, z first)
        and swaps the reordered arrays in, as the reference swaps storage with
        its tmp buffer (src/buffer.cuh:184-209, src/main.cu:68)."""
        if self.deferred:
            self.check()  # an earlier build's error, before anything is swapped
        arrays = [getattr(o, nm) for o, nm in fields]
        n = arrays[0].n
        tmps = self._tmp.setdefault(n, [])
        while len(tmps) < len(arrays):
            tmps.append(self.ctx.empty(n))
        fptr = (C.c_void_p * len(arrays))(*[a.ptr for a in arrays])
        tptr = (C.c_void_p * len(arrays))(*[t.ptr for t in tmps[: len(arrays)]])
        g = capi.Grid()
        rc = self.ctx._lib.reyn_grid_construct_and_reorder(
            self.ctx.h, self.h, float(search_radius), fptr, tptr, len(arrays), n, C.byref(g)
        )
        if rc in (capi.OK, capi.ERR_OUT_OF_BOUNDS, capi.ERR_CAPACITY):
            # the gather ran: swap storage with the tmp arrays
            for k, (o, nm) in enumerate(fields):
                setattr(o, nm, tmps[k])
                tmps[k] = arrays[k]
        self.ctx.check(rc, "construct_and_reorder")
        g._builder = self  # the grid points into this builder's tables: keep it alive with the grid
        return g

    def debug_tables(self, n: int) -> tuple[np.ndarray, np.ndarray]:
        counts = C.c_void_p()
        keys = C.c_void_p()
        self.ctx.check(self.ctx._lib.reyn_grid_debug_tables(self.h, C.byref(counts), C.byref(keys)))
        g = self.geometry()
        c = np.empty(g.ncells, np.uint32)
        k = np.empty(n, np.uint32)
        self.ctx.check(self.ctx._lib.reyn_memcpy(self.ctx.h, c.ctypes.data_as(C.c_void_p), counts, c.nbytes, 1))
        if n:
            self.ctx.check(self.ctx._lib.reyn_memcpy(self.ctx.h, k.ctypes.data_as(C.c_void_p), keys, k.nbytes, 1))
        return c, k

    def close(self) -> None:
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx._lib.reyn_grid_builder_destroy(self.h)
        self.h = None

    release = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def grid_tables(ctx: Context, g: capi.Grid) -> tuple[np.ndarray, np.ndarray | None]:
    """Host copies of a grid's prefix (C+1) and sorted (n) tables."""
    prefix = np.empty(g.ncells + 1, np.uint32)
    ctx.check(ctx._lib.reyn_memcpy(ctx.h, prefix.ctypes.data_as(C.c_void_p), g.prefix, prefix.nbytes, 1))
    sorted_ = None
    if g.sorted:
        sorted_ = np.empty(g.n, np.uint32)
        if g.n:
            ctx.check(ctx._lib.reyn_memcpy(ctx.h, sorted_.ctypes.data_as(C.c_void_p), g.sorted, sorted_.nbytes, 1))
    return prefix, sorted_


class BoundarySamples:
    """Boundary samples with Akinci masses and their own cell-ordered grid
    (scene.cuh:29-46; masses: sample_boundary.cu:251-297)."""

    def __init__(self, ctx: Context, xb: np.ndarray, h: float, rho_0: float, cell_ratio: float = 0.5,
                 mass_iterations: int = 5):
        xb = np.asarray(xb, dtype=np.float32)
        self.ctx = ctx
        self.n = xb.shape[0]
        self.xs = ctx.array(np.ascontiguousarray(xb[:, 0]))
        self.ys = ctx.array(np.ascontiguousarray(xb[:, 1]))
        self.zs = ctx.array(np.ascontiguousarray(xb[:, 2]))
        self.m = ctx.full(self.n, 1.0)
        self.bound_min = xb.min(axis=0).astype(np.float32)
        self.bound_max = xb.max(axis=0).astype(np.float32)
        self.cell_ratio = float(cell_ratio)
        self.builder = UniformGridBuilder(ctx, self.bound_min, self.bound_max, 2.0 * h, cell_ratio)
        # the boundary grid is built once and the samples are stored in cell order
        self.grid = self.builder.construct_and_reorder(2.0 * h, [(self, "xs"), (self, "ys"), (self, "zs")])
        g = self.grid
        self.near = ctx.empty(int(g.nx) * int(g.ny) * int(g.nz), np.uint8)
        ctx.check(ctx._lib.reyn_grid_occupancy(ctx.h, C.byref(self.grid), self.near._ptr), "boundary occupancy")
        view = self.view()
        ctx.check(ctx._lib.reyn_boundary_masses(ctx.h, C.byref(view), self.m._ptr, float(h), float(rho_0),
                                                int(mass_iterations)), "boundary masses")

    def view(self) -> capi.Boundary:
        return capi.Boundary(self.xs.ptr, self.ys.ptr, self.zs.ptr, self.m.ptr, self.n, self.grid, self.near.ptr)

    def positions(self) -> np.ndarray:
        return np.stack([self.xs.numpy(), self.ys.numpy(), self.zs.numpy()], axis=1)


def sph_params(kernel: str | int, h: float, nu: float, rho_0: float, g=GRAVITY, support: float = 2.0) -> capi.SphParams:
    k = capi.KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    return capi.SphParams(k, float(support * h), float(h), float(nu), float(rho_0), capi.f3(g))


class _Solver:
    def __init__(self, ctx: Context, n: int):
        self.ctx = ctx
        self.n = n
        self.rho = ctx.full(n, 0.0)
        self.ax = ctx.full(n, 0.0)
        self.ay = ctx.full(n, 0.0)
        self.az = ctx.full(n, 0.0)

    @staticmethod
    def _bdy(bdy):
        if bdy is None:
            return None
        return C.byref(bdy.view() if hasattr(bdy, "view") else bdy)


class SESPH(_Solver):
    """src/solvers/SESPH.cuh:11-70 (k is a parameter; the reference fixes 500)."""

    def __init__(self, ctx, kernel, n, nu, h, rho_0, k=500.0, g=GRAVITY):
        super().__init__(ctx, n)
        self.prm = capi.SesphParams(sph_params(kernel, h, nu, rho_0, g), float(k))

    def step(self, state: Particles, grid: capi.Grid, bdy, dt: float) -> None:
        work = capi.SesphWork(self.rho.ptr, self.ax.ptr, self.ay.ptr, self.az.ptr)
        view = state.view()
        self.ctx.check(self.ctx._lib.reyn_sesph_step(self.ctx.h, C.byref(self.prm), C.byref(view), C.byref(grid),
                                                     self._bdy(bdy), C.byref(work), float(dt)), "SESPH.step")


class PCISPH(_Solver):
    """src/solvers/PCISPH.cuh:8-111."""

    def __init__(self, ctx, kernel, n, nu, h, rho_0, g=GRAVITY, eta_rho_max=0.001, min_iter=3, max_iter=0,
                 check_every=4):
        super().__init__(ctx, n)
        sph = sph_params(kernel, h, nu, rho_0, g)
        delta = C.c_double()
        capi.lib().reyn_pcisph_delta(sph.kernel, sph.h_bar, sph.h, C.byref(delta))
        self.prm = capi.PcisphParams(sph, delta.value, float(eta_rho_max), int(min_iter), int(max_iter),
                                     int(check_every))
        self.info = capi.SolveInfo()

    def step(self, state: Particles, grid: capi.Grid, bdy, dt: float) -> int:
        work = capi.PcisphWork(self.rho.ptr, self.ax.ptr, self.ay.ptr, self.az.ptr)
        view = state.view()
        self.ctx.check(self.ctx._lib.reyn_pcisph_step(self.ctx.h, C.byref(self.prm), C.byref(view), C.byref(grid),
                                                      self._bdy(bdy), C.byref(work), float(dt), C.byref(self.info)),
                       "PCISPH.step")
        return int(self.info.iterations)


class IISPH(_Solver):
    """src/solvers/IISPH.cuh:8-93; ``stop_rule`` REYN_IISPH_STOP_ANY is the
    reference's rule, STOP_MEAN the mean-error rule."""

    def __init__(self, ctx, kernel, n, nu, h, rho_0, g=GRAVITY, omega=0.5, eta_rho_max=0.005, min_iter=3,
                 max_iter=0, stop_rule=capi.IISPH_STOP_ANY, warm_start=True, check_every=4):
        super().__init__(ctx, n)
        self.p = ctx.full(n, 0.0)
        self.a_ii = ctx.full(n, 0.0)
        self.s_i = ctx.full(n, 0.0)
        self.prm = capi.IisphParams(sph_params(kernel, h, nu, rho_0, g), float(omega), float(eta_rho_max),
                                    int(min_iter), int(max_iter), int(stop_rule), int(bool(warm_start)),
                                    int(check_every))
        self.info = capi.SolveInfo()

    def step(self, state: Particles, grid: capi.Grid, bdy, dt: float) -> int:
        work = capi.IisphWork(self.rho.ptr, self.p.ptr, self.a_ii.ptr, self.s_i.ptr, self.ax.ptr, self.ay.ptr,
                              self.az.ptr)
        view = state.view()
        self.ctx.check(self.ctx._lib.reyn_iisph_step(self.ctx.h, C.byref(self.prm), C.byref(view), C.byref(grid),
                                                     self._bdy(bdy), C.byref(work), float(dt), C.byref(self.info)),
                       "IISPH.step")
        return int(self.info.iterations)


def cfl_time_step(ctx: Context, lam: float, h: float, state: Particles, g=GRAVITY) -> tuple[float, float]:
    dt = C.c_float()
    vmax = C.c_double()
    view = state.view()
    ctx.check(ctx._lib.reyn_cfl_time_step(ctx.h, float(lam), float(h), C.byref(view), capi.f3(g), C.byref(dt),
                                          C.byref(vmax)), "cfl_time_step")
    return float(dt.value), float(vmax.value)


def simple_dt_controller(ctx, iter_last, dt_last, lam, h, state, g=GRAVITY, iter_target=5, change_factor=0.01):
    dt = C.c_float()
    view = state.view()
    ctx.check(ctx._lib.reyn_simple_dt_controller(ctx.h, int(iter_last), float(dt_last), float(lam), float(h),
                                                 C.byref(view), capi.f3(g), int(iter_target), float(change_factor),
                                                 C.byref(dt)), "simple_dt_controller")
    return float(dt.value)


def hard_enforce_bounds(ctx: Context, bound_min, bound_max, state: Particles) -> None:
    view = state.view()
    ctx.check(ctx._lib.reyn_hard_enforce_bounds(ctx.h, capi.f3(bound_min), capi.f3(bound_max), C.byref(view)),
              "hard_enforce_bounds")


def reduce(ctx: Context, a: DeviceArray, op: str) -> float:
    ops = {"sum": 0, "min": 1, "max": 2, "avg": 3}
    out = C.c_double()
    ctx.check(ctx._lib.reyn_reduce_f32(ctx.h, a._ptr, a.n, ops[op], C.byref(out)), "reduce")
    return float(out.value)


@dataclass
class StepStats:
    dt: float
    iterations: int
    mean_rel_error: float
    n_violating: int


class Simulation:
    """The reference's per-step loop (src/main.cu:60-78): time step, grid
    build + reorder (every ``reorder_every`` steps, ``Resort::no`` between),
    solver step, bounds. The grid is built first and the time step taken
    after it (the same dt: the velocity maximum does not depend on the
    order), so that one wait for the GPU serves both.

    Time step: ``dt_policy="cfl"`` takes ``cfl_time_step(cfl_lambda)`` every
    step; ``dt_policy="controller"`` is the reference's main loop
    (main.cu:59-65): dt starts at ``dt0`` and follows
    ``simple_dt_controller(iters, dt, cfl_lambda, h, ..., iter_target,
    change_factor)``."""

    def __init__(self, ctx: Context, solver, state: Particles, bdy: BoundarySamples | None, h: float,
                 bound_min, bound_max, cfl_lambda: float = 0.4, cell_ratio: float = 0.5, reorder_every: int = 1,
                 extra_reorder: tuple[str, ...] = (), dt_policy: str = "cfl", dt0: float = 0.0005,
                 iter_target: int = 10, change_factor: float = 0.001, trace=None, cell_order: str | int = "xyz"):
        if dt_policy not in ("cfl", "controller"):
            raise ValueError(f"unknown dt policy {dt_policy!r}")
        self.ctx = ctx
        self.solver = solver
        self.state = state
        self.bdy = bdy
        self.h = float(h)
        self.bound_min = np.asarray(bound_min, np.float32)
        self.bound_max = np.asarray(bound_max, np.float32)
        self.cfl_lambda = cfl_lambda
        self.cell_ratio = float(cell_ratio)
        self.cell_order = cell_order_id(cell_order)
        self.builder = UniformGridBuilder(ctx, self.bound_min, self.bound_max, 2.0 * h, cell_ratio, cell_order)
        # the build's status is read after the time step's read-back: one wait
        # per step instead of two
        self.builder.defer_check(True)
        self.reorder_every = max(1, int(reorder_every))
        self.extra_reorder = extra_reorder
        self.dt_policy = dt_policy
        self.dt = float(dt0)
        self.iter_target = int(iter_target)
        self.change_factor = float(change_factor)
        self.last_iters = 0
        self.steps = 0
        self.time = 0.0
        self.trace = trace  # a telemetry.StepTrace, or None

    def get_grid(self) -> capi.Grid:
        """Scene::get_grid (src/scene/scene.cuh:149-156): build the grid and,
        on reorder steps, reorder the state plus ``extra_reorder`` solver
        fields (the IISPH warm-start pressure, src/main.cu:68)."""
        r = 2.0 * self.h
        if self.steps % self.reorder_every == 0:
            fields = [(self.state, f) for f in Particles.FIELDS]
            fields += [(self.solver, nm) for nm in self.extra_reorder]
            return self.builder.construct_and_reorder(r, fields)
        return self.builder.construct(r, self.state.xx, self.state.xy, self.state.xz)

    def next_dt(self) -> float:
        if self.dt_policy == "cfl":
            self.dt, _ = cfl_time_step(self.ctx, self.cfl_lambda, self.h, self.state)
        else:
            self.dt = simple_dt_controller(self.ctx, self.last_iters, self.dt, self.cfl_lambda, self.h, self.state,
                                           GRAVITY, self.iter_target, self.change_factor)
        return self.dt

    def step(self) -> StepStats:
        # the grid is built before the time step is read back (the reference
        # takes dt first, main.cu:59-68): the CFL maximum over the reordered
        # velocities is the same number, and the read-back's wait covers the
        # build, whose status is checked right after it
        grid = self.get_grid()
        dt = self.next_dt()
        self.builder.check()
        iters = self.solver.step(self.state, grid, self.bdy, dt) or 0
        hard_enforce_bounds(self.ctx, self.bound_min, self.bound_max, self.state)
        self.steps += 1
        self.time += dt
        self.last_iters = int(iters)
        info = getattr(self.solver, "info", None)
        stats = StepStats(dt, int(iters), float(info.mean_rel_error) if info else 0.0,
                          int(info.n_violating) if info else 0)
        if self.trace is not None:
            self.trace.record(self.steps, self.time, dt, stats.iterations, stats.mean_rel_error)
        return stats


def cell_order_id(order: str | int) -> int:
    """REYN_CELL_ORDER_* from "xyz" / "brick" or the number itself."""
    return capi.CELL_ORDERS[order] if isinstance(order, str) else int(order)


def batch_params(scenes=(), solver: str = "pcisph", kernel: str | int = "W6", nu: float = 0.001,
                 cell_ratio: float = 0.5, dt0: float = 0.0005, cfl_lambda: float = 1.0,
                 change_factor: float = 0.001, iter_target: int = 10, eta_rho_max: float | None = None,
                 omega: float = 0.5, min_iter: int = 3, max_iter: int = 100,
                 stop_rule: int = capi.IISPH_STOP_MEAN, warm_start: bool = True, g=GRAVITY, graph: bool = False,
                 check_every: int = 4, rho_0: float | None = None, cell_order: str | int = "xyz") -> capi.BatchParams:
    """`reyn_batch_params` from keyword options (rho_0 defaults to the first
    scene's)."""
    solver_id = {"pcisph": capi.SOLVER_PCISPH, "iisph": capi.SOLVER_IISPH}[solver]
    if eta_rho_max is None:
        eta_rho_max = 0.001
    if rho_0 is None:
        rho_0 = scenes[0].rho_0 if len(scenes) else 1000.0
    k = capi.KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    return capi.BatchParams(solver_id, k, float(nu), float(rho_0), capi.f3(g),
                            float(cell_ratio), float(dt0), float(cfl_lambda), float(change_factor),
                            int(iter_target), float(eta_rho_max), float(omega), int(min_iter),
                            int(max_iter), int(stop_rule), int(bool(warm_start)), int(bool(graph)),
                            int(check_every), cell_order_id(cell_order))


def batch_scenes(scenes, keep: list):
    """A `reyn_batch_scene` array over the scenes' host arrays (held in
    `keep`); an item is a Scene (starts at rest) or an `io.Checkpoint`
    (continues where it was taken)."""
    from .io import Checkpoint

    def ptr(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        keep.append(a)
        return a.ctypes.data

    arr = (capi.BatchScene * max(1, len(scenes)))()
    for i, item in enumerate(scenes):
        sc = item.scene if isinstance(item, Checkpoint) else item
        xb = np.asarray(sc.xb, np.float32)
        bdy = (ptr(xb[:, 0]), ptr(xb[:, 1]), ptr(xb[:, 2])) if sc.n_bdy else (None, None, None)
        if isinstance(item, Checkpoint):
            st = item.state
            pos = [ptr(st[k]) for k in ("xx", "xy", "xz")]
            vel = [ptr(st[k]) for k in ("vx", "vy", "vz")]
            prs = ptr(st["p"])
            info = capi.BatchSceneInfo(item.steps, item.iterations, item.dt, item.time, item.loop_stat)
            keep.append(info)
            resume = C.pointer(info)
            n = item.n
        else:
            x = np.asarray(sc.x, np.float32)
            pos = [ptr(x[:, 0]), ptr(x[:, 1]), ptr(x[:, 2])]
            vel, prs, resume, n = [None, None, None], None, None, sc.n
        arr[i] = capi.BatchScene(n, *pos, *vel, float(sc.mass), sc.n_bdy, *bdy, capi.f3(sc.bound_min),
                                 capi.f3(sc.bound_max), float(sc.spacing), prs, resume)
    return arr


def dam_break_sizes(desc: capi.DamBreak) -> tuple[int, int]:
    """Lattice size and boundary count of a `reyn_dam_break` (host only)."""
    nl, nb = C.c_uint32(), C.c_uint32()
    rc = capi.lib().reyn_dam_break_sizes(C.byref(desc), C.byref(nl), C.byref(nb))
    if rc != 0:
        raise capi.ReynError(rc, "dam break sizes")
    return int(nl.value), int(nb.value)


def generate_dam_break(ctx: Context, desc: capi.DamBreak) -> tuple[np.ndarray, np.ndarray]:
    """Generates a dam break on the device; returns its fluid and boundary
    positions as host (n, 3) float32 arrays."""
    nl, nb = dam_break_sizes(desc)
    fl = [ctx.empty(max(1, nl)) for _ in range(3)]
    bd = [ctx.empty(max(1, nb)) for _ in range(3)]
    nf = C.c_uint32()
    ctx.check(ctx._lib.reyn_dam_break_generate(ctx.h, C.byref(desc), *[a.ptr for a in fl], C.byref(nf),
                                               *[a.ptr for a in bd]), "dam break generate")
    n = int(nf.value)
    x = np.stack([a.numpy()[:n] for a in fl], axis=1)
    xb = np.stack([a.numpy()[:nb] for a in bd], axis=1) if nb else np.zeros((0, 3), np.float32)
    return x, xb


def scene_bytes(ctx: Context, prm: capi.BatchParams, sc) -> int:
    """Device bytes the scene occupies while resident in a batch (`reyn_batch_scene_bytes`)."""
    keep = []
    arr = batch_scenes([sc], keep)
    out = C.c_uint64()
    ctx.check(ctx._lib.reyn_batch_scene_bytes(C.byref(prm), arr, C.byref(out)), "batch scene bytes")
    return int(out.value)


class Batch:
    """Many scenes advanced together (`reyn_batch_*`): one launch per phase over
    all scenes, each scene with its own grid, boundary, spacing, time step
    (the reference main loop's controller) and pressure-loop control. Scenes
    are addressed by id (0 .. len(scenes)-1 for the initial ones); `add` and
    `remove` change the resident set between steps."""

    def __init__(self, ctx: Context, scenes, **options):
        self.ctx = ctx
        self.h = None
        scenes = list(scenes)
        self.prm = batch_params(scenes, **options)
        self.scenes = {}
        keep = []
        arr = batch_scenes(scenes, keep)
        self.h = C.c_void_p()
        ctx.check(ctx._lib.reyn_batch_create(ctx.h, C.byref(self.prm), arr, len(scenes), C.byref(self.h)),
                  "batch create")
        self.scenes = dict(enumerate(scenes))
        ctx.adopt(self)

    @property
    def n(self) -> int:
        return sum(sc.n for sc in self.scenes.values())

    def add(self, scenes) -> list[int]:
        scenes = list(scenes)
        if not scenes:
            return []
        keep = []
        arr = batch_scenes(scenes, keep)
        from .io import Checkpoint

        ids = (C.c_uint32 * len(scenes))()
        self.ctx.check(self.ctx._lib.reyn_batch_add(self.h, arr, len(scenes), ids), "batch add")
        out = list(ids)
        self.scenes.update(zip(out, [s.scene if isinstance(s, Checkpoint) else s for s in scenes]))
        return out

    def add_dam_breaks(self, scenes) -> list[int]:
        """Adds scenes made by `scenes.dam_break`, generated on the device
        from their descriptions (`reyn_batch_add_dam_breaks`)."""
        from .scenes import dam_break_desc

        scenes = list(scenes)
        if not scenes:
            return []
        arr = (capi.DamBreak * len(scenes))(*[dam_break_desc(sc) for sc in scenes])
        ids = (C.c_uint32 * len(scenes))()
        self.ctx.check(self.ctx._lib.reyn_batch_add_dam_breaks(self.h, arr, len(scenes), ids), "batch add dam breaks")
        out = list(ids)
        self.scenes.update(zip(out, scenes))
        return out

    def checkpoint(self, scene: int):
        """The scene's state and step record (`io.Checkpoint`); adding it to
        any batch continues the scene exactly."""
        from .io import STATE_FIELDS, Checkpoint

        d = self.download(scene)
        inf = self.info(scene)
        return Checkpoint(self.scenes[scene], {k: d[k] for k in STATE_FIELDS}, int(inf.steps), int(inf.iterations),
                          float(inf.dt), float(inf.time), float(inf.loop_stat))

    def remove(self, ids) -> None:
        ids = [int(i) for i in ids]
        if not ids:
            return
        arr = (C.c_uint32 * len(ids))(*ids)
        self.ctx.check(self.ctx._lib.reyn_batch_remove(self.h, arr, len(ids)), "batch remove")
        for i in ids:
            self.scenes.pop(i, None)

    def ids(self) -> list[int]:
        cnt = C.c_uint32()
        self.ctx.check(self.ctx._lib.reyn_batch_scenes(self.h, None, 0, C.byref(cnt)), "batch scenes")
        arr = (C.c_uint32 * max(1, cnt.value))()
        self.ctx.check(self.ctx._lib.reyn_batch_scenes(self.h, arr, cnt.value, C.byref(cnt)), "batch scenes")
        return list(arr)[:cnt.value]

    def step(self, steps: int = 1) -> None:
        self.ctx.check(self.ctx._lib.reyn_batch_step(self.h, int(steps)), "batch step")

    def info(self, scene: int) -> capi.BatchSceneInfo:
        out = capi.BatchSceneInfo()
        self.ctx.check(self.ctx._lib.reyn_batch_info(self.h, int(scene), C.byref(out)), "batch info")
        return out

    def download(self, scene: int) -> dict:
        n = self.scenes[scene].n
        out = {k: np.empty(n, np.float32) for k in ("xx", "xy", "xz", "vx", "vy", "vz", "rho", "p")}
        ptrs = [out[k].ctypes.data for k in ("xx", "xy", "xz", "vx", "vy", "vz", "rho", "p")]
        self.ctx.check(self.ctx._lib.reyn_batch_download(self.h, int(scene), *ptrs), "batch download")
        return out

    def release(self) -> None:
        if self.h is not None and self.h.value and self.ctx.h:
            self.ctx._lib.reyn_batch_destroy(self.h)
        self.h = None

    close = release

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

