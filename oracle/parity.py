"""GPU-versus-oracle comparison of one solver step (TEST INFRASTRUCTURE).

Builds a seeded scene, runs one step through the C-ABI on the GPU and the
same step in the FP64 oracle on identical FP32 inputs (the GPU's reordered
state, the GPU's boundary masses and the same iteration count), and returns
per-field relative errors max|gpu - ref| / max|ref|. Used by the GPU tests
and by ``__graft_entry__.smoke()``.
"""

from __future__ import annotations

import numpy as np

from oracle import binding as orc
from reyn_amd import capi, scenes
from reyn_amd.sim import IISPH, PCISPH, SESPH, BoundarySamples, Particles, UniformGridBuilder, grid_tables

NU = 0.001


def rel_err(gpu, ref) -> float:
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = np.abs(ref).max()
    return float(np.abs(gpu - ref).max() / scale) if scale > 0 else float(np.abs(gpu).max())


def elem_rel_err(gpu, ref, floor: float = 1e-3) -> float:
    """max_i |gpu_i - ref_i| / max(|ref_i|, floor max_j |ref_j|): the error of
    each value relative to itself, with a floor for values near zero (rows of
    vectors use their largest component)."""
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    if ref.size == 0:
        return 0.0
    mag = np.abs(ref) if ref.ndim == 1 else np.abs(ref).max(axis=1)
    diff = np.abs(gpu - ref) if ref.ndim == 1 else np.abs(gpu - ref).max(axis=1)
    scale = np.maximum(mag, floor * mag.max())
    return float((diff / np.where(scale > 0, scale, 1.0)).max())


def oracle_boundary(bdy, r: float):
    """The oracle's view of a GPU boundary (its samples are stored in cell
    order with the GPU's Akinci masses)."""
    xb = bdy.positions()
    bgeom = orc.geometry(bdy.bound_min, bdy.bound_max, r, bdy.cell_ratio)
    bgrid = orc.grid_build(bgeom, xb[:, 0], xb[:, 1], xb[:, 2])
    assert np.array_equal(bgrid.sorted, np.arange(xb.shape[0])), "boundary is not in cell order"
    return orc.Bdy(bgrid.resorted(), xb[:, 0], xb[:, 1], xb[:, 2], bdy.m.numpy())


def iisph_sim_step_parity(sim, ob=None):
    """One step of an IISPH `Simulation` (time step, grid + reorder of the
    state and the pressure, solver step, bounds) on the GPU, and the same step
    in the FP64 oracle from the GPU's state before it, with the GPU's time step,
    iteration count and solver parameters. Returns (errors, gpu, ref): per
    field the error relative to the field's largest value and per element
    (`elem_rel_err`)."""
    ctx, solver = sim.ctx, sim.solver
    h = sim.h
    r = 2.0 * h
    hs = sim.state.host()
    p_in = solver.p.numpy()
    ob = ob if ob is not None else oracle_boundary(sim.bdy, r)
    s = sim.step()
    out = sim.state.host()
    gpu = {"rho": solver.rho.numpy(), "p": solver.p.numpy(), "a_ii": solver.a_ii.numpy(),
           "s_i": solver.s_i.numpy(), "x": np.stack([out["xx"], out["xy"], out["xz"]], 1),
           "v": np.stack([out["vx"], out["vy"], out["vz"]], 1)}
    geom = orc.geometry(sim.bound_min, sim.bound_max, r, sim.cell_ratio, sim.cell_order)
    g = orc.grid_build(geom, hs["xx"], hs["xy"], hs["xz"])
    order = g.sorted
    st = orc.State({"x": hs["xx"][order], "y": hs["xy"][order], "z": hs["xz"][order], "vx": hs["vx"][order],
                    "vy": hs["vy"][order], "vz": hs["vz"][order], "m": hs["m"][order]})
    x0 = np.stack([st.x, st.y, st.z], 1).astype(np.float32)
    sp, q = solver.prm.sph, solver.prm
    prm = orc.sph(sp.kernel, sp.h_bar, sp.h, sp.nu, sp.rho_0, g=(sp.g.x, sp.g.y, sp.g.z))
    ip = orc.OrcIisph(q.omega, q.eta_rho_max, q.min_iter, q.max_iter, q.stop_rule, q.warm_start)
    ref = orc.iisph_step(g.resorted(), st, ob, prm, ip, p_in[order], s.dt, fixed_iters=s.iterations)
    x, v = orc.hard_enforce_bounds(sim.bound_min, sim.bound_max, np.stack([st.x, st.y, st.z], 1),
                                   np.stack([st.vx, st.vy, st.vz], 1))
    ref["x"], ref["v"] = x, v
    errs = {}
    for k in ("rho", "p", "a_ii", "s_i", "x", "v"):
        errs[k] = rel_err(gpu[k], ref[k])
        errs[k + "_elem"] = elem_rel_err(gpu[k], ref[k])
    errs["dx"] = rel_err(gpu["x"] - x0, ref["x"] - x0)
    errs["dt"] = s.dt
    errs["iterations"] = s.iterations
    return errs, gpu, ref


def make_case(side=14, spacing=0.05, compress=0.97, vel=0.2, seed=7, box=(1.4, 1.2, 1.0), sphere=None):
    """A fluid block slightly compressed (lattice spacing compress*s) with
    random velocities, in a walled box; returns (scene, velocities)."""
    sc = scenes.dam_break((side, side, side), spacing, box, jitter=0.05, seed=seed, sphere=sphere)
    sc.x = (sc.x * np.float32(compress) + np.float32(0.5 * spacing)).astype(np.float32)
    rng = np.random.default_rng(seed + 1)
    v = rng.normal(0.0, vel, size=sc.x.shape).astype(np.float32)
    return sc, v


class GpuCase:
    """Uploads a scene, builds the boundary (GPU Akinci masses) and the
    fluid grid with reordering, and keeps host copies of every input."""

    def __init__(self, ctx, sc, v, cell_ratio=0.5, with_bdy=True, masses=None, order=0):
        self.ctx = ctx
        self.sc = sc
        h = sc.spacing
        self.h = h
        self.r = 2.0 * h
        self.state = Particles(ctx, sc.x, v, sc.mass if masses is None else masses, sc.rho_0)
        self.bdy = BoundarySamples(ctx, sc.xb, h, sc.rho_0, cell_ratio) if with_bdy else None
        self.builder = UniformGridBuilder(ctx, sc.bound_min, sc.bound_max, self.r, cell_ratio, order)
        self.cell_ratio = cell_ratio
        self.order = order
        self.extra = []

    def grid(self, extra_owner=None, extra=()):
        fields = [(self.state, f) for f in Particles.FIELDS] + [(extra_owner, nm) for nm in extra]
        self.g = self.builder.construct_and_reorder(self.r, fields)
        self.host = self.state.host()
        self.prefix, _ = grid_tables(self.ctx, self.g)
        return self.g

    def oracle_inputs(self):
        """Oracle grid/boundary/state built from the GPU's reordered inputs."""
        sc = self.sc
        hs = self.host
        geom = orc.geometry(sc.bound_min, sc.bound_max, self.r, self.cell_ratio, self.order)
        og = orc.grid_build(geom, hs["xx"], hs["xy"], hs["xz"])
        # the reordered state is in cell order: its build is the identity
        assert np.array_equal(og.sorted, np.arange(og.sorted.size)), "GPU reorder is not in cell order"
        assert np.array_equal(og.prefix, self.prefix), "GPU prefix differs from the oracle's"
        fg = og.resorted()
        st = orc.State({"x": hs["xx"], "y": hs["xy"], "z": hs["xz"], "vx": hs["vx"], "vy": hs["vy"],
                        "vz": hs["vz"], "m": hs["m"]})
        ob = None
        if self.bdy is not None:
            xb = self.bdy.positions()
            mb = self.bdy.m.numpy()
            bgeom = orc.geometry(self.bdy.bound_min, self.bdy.bound_max, self.r, self.cell_ratio)
            bgrid = orc.grid_build(bgeom, xb[:, 0], xb[:, 1], xb[:, 2])
            assert np.array_equal(bgrid.sorted, np.arange(xb.shape[0])), "boundary is not in cell order"
            ob = orc.Bdy(bgrid.resorted(), xb[:, 0], xb[:, 1], xb[:, 2], mb)
        return fg, st, ob


def sph_oracle(kernel: int, h: float, rho0: float):
    return orc.sph(kernel, 2.0 * h, h, NU, rho0)


def run_iisph_case(ctx, kernel="B3", dt=2e-3, side=14, warm=True, stop_rule=capi.IISPH_STOP_ANY, eta=0.005,
                   min_iter=3, max_iter=0, seed=7, p0_scale=200.0, cell_ratio=0.5, compress=0.97, mass_spread=0.0,
                   order=0):
    """One IISPH step on the GPU against the FP64 oracle from the same inputs;
    `mass_spread` > 0 gives every particle its own mass, m (1 + U(-spread,
    spread))."""
    sc, v = make_case(side=side, seed=seed, compress=compress)
    masses = None
    if mass_spread:
        u = np.random.default_rng(seed + 5).uniform(-mass_spread, mass_spread, sc.n)
        masses = (sc.mass * (1.0 + u)).astype(np.float32)
    case = GpuCase(ctx, sc, v, cell_ratio, masses=masses, order=order)
    solver = IISPH(ctx, kernel, sc.n, NU, sc.spacing, sc.rho_0, eta_rho_max=eta, min_iter=min_iter,
                   max_iter=max_iter, stop_rule=stop_rule, warm_start=warm)
    p0 = np.abs(np.random.default_rng(seed + 2).normal(0.0, p0_scale, sc.n)).astype(np.float32)
    solver.p.upload(p0)
    case.grid(solver, ("p",))
    p_in = solver.p.numpy()
    fg, st, ob = case.oracle_inputs()
    iters = solver.step(case.state, case.g, case.bdy, dt)
    out = case.state.host()
    gpu = {"rho": solver.rho.numpy(), "p": solver.p.numpy(), "a_ii": solver.a_ii.numpy(), "s_i": solver.s_i.numpy(),
           "x": np.stack([out["xx"], out["xy"], out["xz"]], 1), "v": np.stack([out["vx"], out["vy"], out["vz"]], 1),
           "iterations": iters, "mean_rel_error": solver.info.mean_rel_error, "n_violating": solver.info.n_violating}
    ip = orc.OrcIisph(0.5, eta, min_iter, max_iter, stop_rule, 1 if warm else 0)
    x0 = np.stack([st.x, st.y, st.z], 1).copy()
    ref = orc.iisph_step(fg, st, ob, sph_oracle(capi.KERNELS[kernel], sc.spacing, sc.rho_0), ip, p_in, dt,
                         fixed_iters=iters)
    ref["x"] = np.stack([st.x, st.y, st.z], 1)
    ref["v"] = np.stack([st.vx, st.vy, st.vz], 1)
    errs = {k: rel_err(gpu[k], ref[k]) for k in ("rho", "p", "a_ii", "s_i", "x", "v")}
    errs["dx"] = rel_err(gpu["x"] - x0, ref["x"] - x0)
    return errs, gpu, ref


def run_sesph_case(ctx, kernel="C2", dt=1e-3, side=14, k=1000.0, seed=11, cell_ratio=0.5, compress=0.97):
    sc, v = make_case(side=side, seed=seed, compress=compress)
    case = GpuCase(ctx, sc, v, cell_ratio)
    solver = SESPH(ctx, kernel, sc.n, NU, sc.spacing, sc.rho_0, k=k)
    case.grid()
    fg, st, ob = case.oracle_inputs()
    x0 = np.stack([st.x, st.y, st.z], 1).copy()
    solver.step(case.state, case.g, case.bdy, dt)
    out = case.state.host()
    gpu = {"rho": solver.rho.numpy(), "a": np.stack([solver.ax.numpy(), solver.ay.numpy(), solver.az.numpy()], 1),
           "x": np.stack([out["xx"], out["xy"], out["xz"]], 1), "v": np.stack([out["vx"], out["vy"], out["vz"]], 1)}
    rho, a = orc.sesph_step(fg, st, ob, sph_oracle(capi.KERNELS[kernel], sc.spacing, sc.rho_0), k, dt)
    ref = {"rho": rho, "a": a, "x": np.stack([st.x, st.y, st.z], 1), "v": np.stack([st.vx, st.vy, st.vz], 1)}
    errs = {q: rel_err(gpu[q], ref[q]) for q in ("rho", "a", "x", "v")}
    errs["dx"] = rel_err(gpu["x"] - x0, ref["x"] - x0)
    return errs, gpu, ref


def run_pcisph_case(ctx, kernel="W6", dt=1e-3, side=14, seed=13, min_iter=3, max_iter=0, cell_ratio=0.5,
                    compress=0.97, mass_spread=0.0):
    """One PCISPH step on the GPU against the FP64 oracle from the same
    inputs; `mass_spread` > 0 gives every particle its own mass, as in
    `run_iisph_case`."""
    sc, v = make_case(side=side, seed=seed, compress=compress)
    masses = None
    if mass_spread:
        u = np.random.default_rng(seed + 5).uniform(-mass_spread, mass_spread, sc.n)
        masses = (sc.mass * (1.0 + u)).astype(np.float32)
    case = GpuCase(ctx, sc, v, cell_ratio, masses=masses)
    solver = PCISPH(ctx, kernel, sc.n, NU, sc.spacing, sc.rho_0, min_iter=min_iter, max_iter=max_iter)
    case.grid()
    fg, st, ob = case.oracle_inputs()
    x0 = np.stack([st.x, st.y, st.z], 1).copy()
    iters = solver.step(case.state, case.g, case.bdy, dt)
    out = case.state.host()
    gpu = {"rho": solver.rho.numpy(), "a": np.stack([solver.ax.numpy(), solver.ay.numpy(), solver.az.numpy()], 1),
           "x": np.stack([out["xx"], out["xy"], out["xz"]], 1), "v": np.stack([out["vx"], out["vy"], out["vz"]], 1),
           "iterations": iters, "avg": solver.info.avg_density}
    it, rho, a, avg = orc.pcisph_step(fg, st, ob, sph_oracle(capi.KERNELS[kernel], sc.spacing, sc.rho_0),
                                      solver.prm.delta, dt, fixed_iters=iters)
    ref = {"rho": rho, "a": a, "x": np.stack([st.x, st.y, st.z], 1), "v": np.stack([st.vx, st.vy, st.vz], 1),
           "avg": avg}
    errs = {q: rel_err(gpu[q], ref[q]) for q in ("rho", "a", "x", "v")}
    errs["dx"] = rel_err(gpu["x"] - x0, ref["x"] - x0)
    errs["avg"] = abs(gpu["avg"] - avg) / avg
    return errs, gpu, ref


def batch_scene_step_parity(batch, sid: int, bdy_iterations: int = 5):
    """One step of a `Batch` (every resident scene advances) and the same step
    of scene `sid` in the FP64 oracle from that scene's state before it: the
    grid over the scene's domain at the batch's cell ratio, the reorder, the
    batch's PCISPH or IISPH step with the scene's time step and iteration
    count, the bounds. The oracle's boundary is the scene's samples with FP64
    Akinci masses. Returns per-field errors relative to the largest value and
    per element (`elem_rel_err`)."""
    import ctypes as C

    sc = batch.scenes[sid]
    prm = batch.prm
    before = batch.download(sid)
    batch.step(1)
    after = batch.download(sid)
    inf = batch.info(sid)
    h = float(sc.spacing)
    r = 2.0 * h
    geom = orc.geometry(sc.bound_min, sc.bound_max, r, prm.cell_ratio, prm.cell_order)
    g = orc.grid_build(geom, before["xx"], before["xy"], before["xz"])
    o = g.sorted
    st = orc.State({"x": before["xx"][o], "y": before["xy"][o], "z": before["xz"][o], "vx": before["vx"][o],
                    "vy": before["vy"][o], "vz": before["vz"][o], "m": np.full(sc.n, np.float32(sc.mass))})
    xb = np.asarray(sc.xb, np.float32)
    bg = orc.grid_build(orc.geometry(xb.min(0), xb.max(0), r, prm.cell_ratio), xb[:, 0], xb[:, 1], xb[:, 2])
    xb = xb[bg.sorted]
    bg = orc.grid_build(orc.geometry(xb.min(0), xb.max(0), r, prm.cell_ratio), xb[:, 0], xb[:, 1],
                        xb[:, 2]).resorted()
    mb = orc.boundary_masses(bg, xb[:, 0], xb[:, 1], xb[:, 2], h, prm.rho_0, bdy_iterations)
    ob = orc.Bdy(bg, xb[:, 0], xb[:, 1], xb[:, 2], mb)
    sph = orc.sph(prm.kernel, r, h, prm.nu, prm.rho_0, g=(prm.g.x, prm.g.y, prm.g.z))
    if prm.solver == capi.SOLVER_PCISPH:
        delta = C.c_double()
        capi.lib().reyn_pcisph_delta(prm.kernel, np.float32(r), np.float32(h), C.byref(delta))
        _, rho, _, _ = orc.pcisph_step(g.resorted(), st, ob, sph, delta.value, inf.dt, eta=prm.eta_rho_max,
                                       min_iter=prm.min_iter, max_iter=prm.max_iter, fixed_iters=inf.iterations)
        ref = {"rho": rho}
    else:
        ip = orc.OrcIisph(prm.omega, prm.eta_rho_max, prm.min_iter, prm.max_iter, prm.stop_rule, prm.warm_start)
        res = orc.iisph_step(g.resorted(), st, ob, sph, ip, before["p"][o], inf.dt, fixed_iters=inf.iterations)
        ref = {"rho": res["rho"], "p": res["p"]}
    x, v = orc.hard_enforce_bounds(sc.bound_min, sc.bound_max, np.stack([st.x, st.y, st.z], 1),
                                   np.stack([st.vx, st.vy, st.vz], 1))
    ref["x"], ref["v"] = x, v
    gpu = {"rho": after["rho"], "p": after["p"], "x": np.stack([after["xx"], after["xy"], after["xz"]], 1),
           "v": np.stack([after["vx"], after["vy"], after["vz"]], 1)}
    x0 = np.stack([st.x, st.y, st.z], 1).astype(np.float32)
    errs = {}
    for k in ref:
        errs[k] = rel_err(gpu[k], ref[k])
        errs[k + "_elem"] = elem_rel_err(gpu[k], ref[k])
    errs["dx"] = rel_err(gpu["x"] - x0, ref["x"] - x0)
    errs["dt"] = float(inf.dt)
    errs["iterations"] = int(inf.iterations)
    return errs
