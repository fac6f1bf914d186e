#!/usr/bin/env python3
"""Benchmark: IISPH dam break, 2,097,152 FP32 fluid particles, one MI355X.

Workload: a 128^3 fluid block [0,1]^3 at spacing s = 1/128 in the box
[0,2]x[0,1.25]x[0,2] with 2 wall layers (~600k boundary samples) and a seeded
sphere obstacle; cubic B-spline kernel with support 2s, rho_0 = 1000,
nu = 0.001, relaxed Jacobi omega = 0.5, stop at mean compression error
<= 0.1% or 100 iterations, the reference main loop's time-step controller
(lambda 1, 10 target iterations, 0.1% change per step, from dt0 = 5e-4). One
step is the reference's per-step loop (src/main.cu:60-78): time step, grid
build and reorder of every particle field (and the warm-start pressure), the
IISPH step, bounds enforcement.

Prints one JSON line with particle-steps/s, the roofline of the dominant
kernel (algorithmic bytes per launch / its average HIP-event duration, timed
live on the launch stream over the timed steps) and the CPU baseline (the
oracle's FP32 OpenMP port of the same step, on the same state, on this host).
`config.phases` gives every kernel's HBM fraction and its VALU fraction:
VALU instructions per launch (rocprofv3 SQ_INSTS_VALU, profiles/valu_insts.json)
over its live launch time, against one wave64 instruction per CU per clock.
Two more objects in the same line:

- "batch": 64 PCISPH scenes of 10k-300k particles (seeded) advanced in one
  batch with the autotuned cell ratio and check interval, against the same
  64 scenes solved one at a time through the single-scene entry points.
- "serving": 32 seeded IISPH jobs (5k-500k particles, 200-2000 steps) with
  Poisson arrivals at 80% of the measured batch capacity, served by
  continuous batching under the HBM budget, against one job at a time in
  arrival order (FIFO); p50/p99 time-to-result, particle-steps/s, and whether
  every served result equals its alone result bit for bit.
- "grid": the 1M-particle grid and SESPH configuration at cell ratios 0.5, 1
  and 2: grid build, build + reorder, neighbour query and SESPH step times.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HBM_PEAK_GBS = 8000.0
# VALU issue peak: one wave64 FP32 instruction per CU per clock (256 CUs at
# 2.4 GHz; packed FP32 is not used), i.e. 39.3 T lane-operations/s
VALU_PEAK_INSTS = 256 * 2.4e9
# the texture-address units: one per CU, busy cycles counted per CU
# (TA_TA_BUSY summed over the 256 CUs)
TA_CYCLES_PER_S = 256 * 2.4e9
METRIC = "particle-steps/sec + % HBM roofline (IISPH 2M); serving p50/p99 vs plain"


def algorithmic_bytes(name: str, n: int, nb: int, ncells: int, nfields: int) -> float | None:
    """Compulsory DRAM bytes of one launch: every array it reads once plus
    every array it writes once (neighbour re-reads count zero)."""
    table = {
        # x,y,z m + boundary x,y,z,m -> rho, grad-sum(3), boundary density
        "k_density": 16 * n + 16 * nb + 20 * n,
        # x v m rho -> a
        "k_nonpressure": 32 * n + 12 * n,
        # v a -> v
        "k_integrate_v": 24 * n + 12 * n,
        # x v a -> x v
        "k_integrate": 36 * n + 24 * n,
        # x v m rho grad-sum p -> a_ii s_i p
        "k_iisph_aii_s": 48 * n + 12 * n,
        # x p m grad-sum -> a
        "k_iisph_pacc": 32 * n + 12 * n,
        # x a m a_ii s_i p grad-sum -> p
        "k_iisph_jacobi": 52 * n + 4 * n,
        # x -> keys slots (+ counts)
        "k_count": 12 * n + 8 * n,
        "k_scan_reduce": 4 * ncells,
        "k_scan_down": 12 * ncells,
        "k_scatter": 12 * n + 4 * n,
        "k_stabilize": 16 * n + 4 * n,
        "k_gather": 4 * n + 8 * nfields * n,
        # x -> x where clamped (the velocity of the few particles at a face)
        "k_bounds": 12 * n,
    }
    return float(table[name]) if name in table else None


def step_bytes(n: int, nb: int, ncells: int, iters: float) -> float:
    """Compulsory bytes of one IISPH step under the reference's phase
    structure: 372N + 48N_b + 28C + l (72N + 32N_b)."""
    return 372.0 * n + 48.0 * nb + 28.0 * ncells + iters * (72.0 * n + 32.0 * nb)


def measured_traffic(kernel: str) -> float | None:
    """HBM bytes per working launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (scaled by the copy-kernel calibration it
    records), or None when the summary does not cover the kernel."""
    path = Path(__file__).resolve().parent / "profiles" / "hbm_traffic.json"
    try:
        table = json.loads(path.read_text())["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None
    v = table.get(kernel)
    return float(v) if v is not None else None


def measured_valu(kernel: str) -> float | None:
    """VALU instructions per working launch of `kernel` from the committed
    rocprofv3 SQ_INSTS_VALU summary, or None when it does not cover the
    kernel."""
    path = Path(__file__).resolve().parent / "profiles" / "valu_insts.json"
    try:
        table = json.loads(path.read_text())["valu_insts_per_launch"]
    except (OSError, ValueError, KeyError):
        return None
    v = table.get(kernel)
    return float(v) if v is not None else None


def measured_ta_busy(kernel: str) -> float | None:
    """Texture-address busy cycles (summed over the CUs) per working launch
    of `kernel` from the committed rocprofv3 TA_TA_BUSY summary, or None."""
    path = Path(__file__).resolve().parent / "profiles" / "address_path.json"
    try:
        table = json.loads(path.read_text())["ta_busy_cycles_per_launch"]
    except (OSError, ValueError, KeyError):
        return None
    v = table.get(kernel)
    return float(v) if v is not None else None


def parse_profile(text: str) -> dict:
    """Kernel name -> (launches, total ms, working launches, working ms) from
    reyn_profile_summary (working launches: those that took at least a fifth
    of the kernel's longest launch; a solver-loop launch that finds its loop
    converged returns at once and is not one)."""
    out = {}
    for line in text.strip().splitlines():
        name, cnt, ms, wcnt, wms = line.split()
        out[name] = (int(cnt), float(ms), int(wcnt), float(wms))
    return out


def profile_summary(ctx) -> dict:
    import ctypes as C
    need = C.c_size_t()
    ctx.check(ctx._lib.reyn_profile_summary(ctx.h, None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value + 16)
    ctx.check(ctx._lib.reyn_profile_summary(ctx.h, buf, len(buf), None))
    return parse_profile(buf.value.decode())


# ---- rocprofv3 summaries ------------------------------------------------------
# `bench.py --rocprof-summary KIND DIR...` reads the CSV output of rocprofv3
# runs of this script (`--output-format csv`) and prints one JSON object. The
# window is the timed steps of the profiled run: from its (warmup)-th k_vmax
# dispatch (0-based; the first kernel of every step) to its last k_bounds
# dispatch (the last kernel of every step); run the profiled command with
# --no-cpu-baseline so nothing follows the timed steps but a final reduction.
# A launch is a working launch when it took (trace) or counted (counters) at
# least a fifth of the kernel's largest figure in the window: a solver-loop
# launch that finds its loop converged returns at once and is not one.

def _kernel_base(name: str) -> str:
    n = name.replace("(anonymous namespace)", "").split("(")[0].split("<")[0]
    return n.split("::")[-1].replace("void ", "").strip()


def _csv_rows(d: str, suffix: str) -> list[dict]:
    import csv
    files = sorted(Path(d).rglob(f"*{suffix}"))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    rows = []
    for f in files:
        with open(f, newline="") as fh:
            rows += list(csv.DictReader(fh))
    return rows


def _window(names: list[str], warmup: int) -> tuple[int, int]:
    vm = [k for k, n in enumerate(names) if _kernel_base(n) == "k_vmax"]
    bd = [k for k, n in enumerate(names) if _kernel_base(n) == "k_bounds"]
    if len(vm) <= warmup or not bd:
        raise SystemExit(f"{len(vm)} k_vmax dispatches for a warm-up of {warmup} steps")
    return vm[warmup], bd[-1]


def _working(vals: list[float]) -> list[float]:
    top = max(vals)
    return [v for v in vals if v >= 0.2 * top]


def rocprof_summary(kind: str, dirs: list[str], warmup: int) -> dict:
    from collections import defaultdict
    if kind == "trace":
        rows = sorted(_csv_rows(dirs[0], "kernel_trace.csv"), key=lambda r: int(r["Dispatch_Id"]))
        a, b = _window([r["Kernel_Name"] for r in rows], warmup)
        per = defaultdict(list)
        for r in rows[a:b + 1]:
            per[_kernel_base(r["Kernel_Name"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        span = (int(rows[b]["End_Timestamp"]) - int(rows[a]["Start_Timestamp"])) / 1e6
        out = {"window_dispatches": b - a + 1, "window_ms": round(span, 3),
               "kernels_ms": round(sum(sum(v) for v in per.values()) / 1e3, 3), "kernels": {}}
        for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
            w = _working(v)
            out["kernels"][k] = {"launches": len(v), "total_ms": round(sum(v) / 1e3, 4),
                                 "working_launches": len(w), "working_avg_us": round(sum(w) / len(w), 2)}
        return out

    def counters(d):
        rows = _csv_rows(d, "counter_collection.csv")
        vals, names, meta = defaultdict(lambda: defaultdict(float)), {}, {}
        for r in rows:
            i = int(r["Dispatch_Id"])
            vals[i][r["Counter_Name"]] += float(r["Counter_Value"])
            names[i] = r["Kernel_Name"]
            meta[i] = {k: int(r[k]) for k in ("VGPR_Count", "SGPR_Count", "LDS_Block_Size") if r.get(k)}
        order = sorted(names)
        return order, vals, names, meta

    def per_kernel(d, counter):
        order, vals, names, meta = counters(d)
        a, b = _window([names[i] for i in order], warmup)
        per, m = defaultdict(list), {}
        for i in order[a:b + 1]:
            k = _kernel_base(names[i])
            per[k].append(vals[i])
            m[k] = meta[i]
        out = {}
        for k, v in per.items():
            key = [x[counter] for x in v]
            top = max(key)
            w = [x for x in v if x[counter] >= 0.2 * top] if top > 0 else v
            out[k] = {"launches": len(v), "working_launches": len(w), **m[k],
                      **{c: sum(x[c] for x in w) / len(w) for c in v[0]}}
        return out

    def calib(d, counter):
        order, vals, names, _ = counters(d)
        got = [vals[i][counter] for i in order if _kernel_base(names[i]) == "k_gather"]
        if not got:
            raise SystemExit(f"no k_gather dispatch under {d}")
        return sum(got) / len(got)

    if kind == "valu":
        per = per_kernel(dirs[0], "SQ_INSTS_VALU")
        return {"peak_wave_insts_per_s": VALU_PEAK_INSTS,
                "valu_insts_per_launch": {k: round(v["SQ_INSTS_VALU"]) for k, v in per.items()
                                          if v["SQ_INSTS_VALU"] > 0},
                "kernels": per}
    if kind == "address":
        per = per_kernel(dirs[0], "TA_TA_BUSY_sum")
        return {"cycles_per_s": TA_CYCLES_PER_S,
                "ta_busy_cycles_per_launch": {k: round(v["TA_TA_BUSY_sum"]) for k, v in per.items()
                                              if v.get("TA_TA_BUSY_sum", 0) > 0},
                "ta_stalled_by_l1_cycles_per_launch": {k: round(v.get("TA_ADDR_STALLED_BY_TC_CYCLES_sum", 0))
                                                       for k, v in per.items() if v.get("TA_TA_BUSY_sum", 0) > 0},
                "kernels": per}
    if kind == "traffic":
        n = 1 << 26  # the copy calibration: reads 8n bytes, writes 4n bytes
        cf, cw = calib(dirs[2], "FETCH_SIZE"), calib(dirs[3], "WRITE_SIZE")
        fs, ws = 8 * n / (cf * 1024), 4 * n / (cw * 1024)
        fe, wr = per_kernel(dirs[0], "FETCH_SIZE"), per_kernel(dirs[1], "WRITE_SIZE")
        out = {"fetch_scale": fs, "write_scale": ws,
               "calibration_kib": {"FETCH_SIZE": cf, "WRITE_SIZE": cw, "read_bytes": 8 * n, "write_bytes": 4 * n},
               "bytes_per_launch": {}, "kernels": {}}
        for k in sorted(fe, key=lambda k: -fe[k]["FETCH_SIZE"]):
            f = fe[k]["FETCH_SIZE"]
            w = wr.get(k, {}).get("WRITE_SIZE", 0.0)
            out["bytes_per_launch"][k] = round(1024 * (f * fs + w * ws))
            out["kernels"][k] = {"working_launches": fe[k]["working_launches"], "fetch_kib_raw": round(f, 1),
                                 "write_kib_raw": round(w, 1), "fetch_bytes": round(1024 * f * fs),
                                 "write_bytes": round(1024 * w * ws)}
        return out
    raise SystemExit(f"unknown summary kind {kind}")


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(ctx, sim, sc, args, log) -> tuple[dict, dict, dict | None]:
    """The oracle's part of the bench, run after the timed steps:

    - the FP64 check: one more GPU step from the GPU's state against the same
      step in the FP64 oracle (per-field errors);
    - the CPU baseline: the oracle's float port of the same main loop (time-step
      controller, parallel counting sort with rocThrust's OpenMP scan, OpenMP
      gathers, the IISPH step, bounds) from the GPU's state after that step, on
      all host threads: 5 warm-up steps, then the median of 3 runs of
      `cpu_steps` steps (std::chrono around whole steps);
    - configuration 1 on the host: the SESPH dam break (32^3, CFL 0.4, C2, cell
      = support radius) for its 1000 steps in the float port, particle-steps/s
      over steps 5-999, beside the same 1000 steps on the GPU."""
    from oracle import binding as orc
    from oracle import parity
    from reyn_amd import capi, scenes
    from reyn_amd.sim import SESPH, BoundarySamples, Particles, Simulation

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    orc.set_threads(threads)
    model = cpu_model()
    h, r = sc.spacing, 2.0 * sc.spacing
    ob = parity.oracle_boundary(sim.bdy, r)
    t = time.perf_counter()
    errs, _, _ = parity.iisph_sim_step_parity(sim, ob)
    log(f"fp64 check of step {sim.steps}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items())
        + f" ({time.perf_counter() - t:.1f} s)")
    fp64 = {"step": sim.steps, **{k: (float(v) if k != "iterations" else int(v)) for k, v in errs.items()}}

    hs = sim.state.host()
    port = orc.PortState(np.stack([hs["xx"], hs["xy"], hs["xz"]], 1), np.stack([hs["vx"], hs["vy"], hs["vz"]], 1),
                         hs["m"], sim.solver.p.numpy())
    bdy = orc.Bdy(ob.grid, ob.x, ob.y, ob.z, ob.m, dtype=np.float32)
    geom = orc.geometry(sc.bound_min, sc.bound_max, r, sim.cell_ratio)
    prm = orc.sph(capi.KERNEL_B3, r, h, 0.001, sc.rho_0)
    ip = orc.OrcIisph(0.5, 0.001, 3, 100, capi.IISPH_STOP_MEAN, 1)
    dt, it = sim.dt, sim.last_iters

    def run(steps):
        nonlocal dt, it
        secs, its, dt, it = orc.port_iisph_run(geom, port, bdy, prm, ip, sc.bound_min, sc.bound_max,
                                               sim.cfl_lambda, sim.iter_target, sim.change_factor, dt, it, steps)
        return secs, its
    warm, warm_its = run(5)
    runs = [run(args.cpu_steps) for _ in range(3)]
    per_run = sorted(float(secs.sum()) for secs, _ in runs)
    median = per_run[1]
    its_all = [int(v) for _, its in runs for v in its]
    value = sc.n * args.cpu_steps / median
    log(f"cpu baseline: 5 warm-up steps {warm.sum():.1f} s, runs of {args.cpu_steps} steps {per_run} s on "
        f"{threads} threads ({model}), iterations {its_all}")
    cpu = {"value": value, "unit": "particle-steps/s", "cores": threads, "kind": "port", "cpu_model": model,
           "sample": f"the same {sc.n}-particle IISPH main loop from the GPU's state after step {sim.steps}: 5 "
                     f"warm-up steps, then the median of 3 runs of {args.cpu_steps} steps (Jacobi iterations "
                     f"{its_all}), oracle FP32 OpenMP port with a parallel counting sort and rocThrust OpenMP "
                     f"scan and gathers, {median:.2f} s per run",
           "runs_s": per_run, "warmup_s": float(warm.sum())}

    c1 = None
    if not args.no_config1:
        s1 = scenes.config1_sesph()
        h1, r1 = s1.spacing, 2.0 * s1.spacing
        st = Particles(ctx, s1.x, None, s1.mass, s1.rho_0)
        b1 = BoundarySamples(ctx, s1.xb, h1, s1.rho_0, 1.0)
        solver = SESPH(ctx, "C2", s1.n, 0.001, h1, s1.rho_0, k=1000.0)
        sim1 = Simulation(ctx, solver, st, b1, h1, s1.bound_min, s1.bound_max, cfl_lambda=0.4, cell_ratio=1.0)
        ob1 = parity.oracle_boundary(b1, r1)
        for _ in range(5):
            sim1.step()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(995):
            sim1.step()
        ctx.synchronize()
        gpu_s = time.perf_counter() - t
        port1 = orc.PortState(s1.x, None, s1.mass)
        secs, simt = orc.port_sesph_run(orc.geometry(s1.bound_min, s1.bound_max, r1, 1.0), port1,
                                        orc.Bdy(ob1.grid, ob1.x, ob1.y, ob1.z, ob1.m, dtype=np.float32),
                                        orc.sph(capi.KERNEL_C2, r1, h1, 0.001, s1.rho_0), 1000.0, s1.bound_min,
                                        s1.bound_max, 0.4, 1000)
        cpu_s = float(secs[5:].sum())
        c1 = {"workload": f"SESPH dam break, {s1.n} fluid particles (32^3 at s = 0.02, jitter 1%, seed 1) + "
                          f"{s1.n_bdy} boundary samples in [0,1]^3, C2, k = 1000, CFL lambda 0.4 every step, cell = "
                          "support radius, 1000 steps; rates over steps 5-999",
              "gpu_particle_steps_per_s": round(s1.n * 995 / gpu_s, 1),
              "cpu_particle_steps_per_s": round(s1.n * 995 / cpu_s, 1), "cpu_threads": threads,
              "cpu_model": model, "cpu_full_run_s": round(float(secs.sum()), 3),
              "cpu_simulated_time": simt, "gpu_simulated_time": sim1.time}
        log(f"configuration 1: GPU {c1['gpu_particle_steps_per_s'] / 1e6:.1f} M, CPU "
            f"{c1['cpu_particle_steps_per_s'] / 1e6:.2f} M particle-steps/s over steps 5-999")
        for o in (solver, b1, st, sim1.builder):
            getattr(o, "close", lambda: None)()
    return cpu, fp64, c1


def grid_leg(ctx, args, log) -> dict:
    """The grid and SESPH configuration: 1,000,000 particles (100^3 at s = 0.02,
    jitter 5%, seed 2) in [0,2]^3, at cell ratios 0.5, 1 and 2: grid build
    alone, build + reorder of the 7 state fields, the neighbour query of the
    reference's micro-benchmark (sum_j x_i x_j W_ij with W6, uniformgrid.cu:231-245)
    over the reordered and the unreordered grid, and SESPH steps with the grid
    rebuilt every step and the state reordered every 10 steps. Wall clock per
    call, `grid_reps` calls each after one untimed call."""
    import ctypes as C

    from reyn_amd import capi, scenes
    from reyn_amd.sim import SESPH, BoundarySamples, Particles, Simulation, UniformGridBuilder

    sc = scenes.config2_grid()
    h = sc.spacing
    r = 2.0 * h
    K = args.grid_reps
    out = {"workload": f"{sc.n} particles (100^3 lattice, s = 0.02, jitter 5%, seed 2) in [0,2]^3, "
                       f"{sc.n_bdy} boundary samples; W6 query; SESPH k = 1000, C2",
           "particles": sc.n, "ratios": {}}

    def timed(fn):
        fn()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(K):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t) / K

    for ratio in (0.5, 1.0, 2.0):
        state = Particles(ctx, sc.x, None, sc.mass, sc.rho_0)
        builder = UniformGridBuilder(ctx, sc.bound_min, sc.bound_max, r, ratio)
        ncells = builder.geometry().ncells
        t_build = timed(lambda: builder.construct(r, state.xx, state.xy, state.xz))
        fields = [(state, f) for f in Particles.FIELDS]
        t_both = timed(lambda: builder.construct_and_reorder(r, fields))
        res = ctx.empty(sc.n)
        g_sorted = builder.construct_and_reorder(r, fields)

        def query(g):
            ctx.check(ctx._lib.reyn_grid_query_bench(ctx.h, C.byref(g), capi.KERNEL_W6, float(r), state.xx.ptr,
                                                     state.xy.ptr, state.xz.ptr, sc.n, res.ptr), "query")
        t_q_sorted = timed(lambda: query(g_sorted))
        # the unreordered view: a shuffled copy of the state, built without reorder
        perm = np.random.default_rng(2).permutation(sc.n)
        shuffled = Particles(ctx, state.positions()[perm], None, sc.mass, sc.rho_0)
        g_unsorted = builder.construct(r, shuffled.xx, shuffled.xy, shuffled.xz)
        t_q_unsorted = timed(lambda: ctx.check(ctx._lib.reyn_grid_query_bench(
            ctx.h, C.byref(g_unsorted), capi.KERNEL_W6, float(r), shuffled.xx.ptr, shuffled.xy.ptr,
            shuffled.xz.ptr, sc.n, res.ptr), "query"))
        builder.close()
        st2 = Particles(ctx, sc.x, None, sc.mass, sc.rho_0)
        bdy = BoundarySamples(ctx, sc.xb, h, sc.rho_0, ratio)
        solver = SESPH(ctx, "C2", sc.n, 0.001, h, sc.rho_0, k=1000.0)
        sim = Simulation(ctx, solver, st2, bdy, h, sc.bound_min, sc.bound_max, cfl_lambda=0.4, cell_ratio=ratio,
                         reorder_every=10, dt_policy="cfl")
        for _ in range(2):
            sim.step()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(2 * K):
            sim.step()
        ctx.synchronize()
        t_step = (time.perf_counter() - t) / (2 * K)
        sim.builder.close()
        row = {"cells": int(ncells), "build_us": round(t_build * 1e6, 1),
               "build_and_reorder_us": round(t_both * 1e6, 1),
               "query_reordered_us": round(t_q_sorted * 1e6, 1), "query_unreordered_us": round(t_q_unsorted * 1e6, 1),
               "sesph_step_us": round(t_step * 1e6, 1)}
        out["ratios"][str(ratio)] = row
        log(f"grid ratio {ratio}: {row}")
    return out


def batch_leg(ctx, args, log) -> dict:
    """64 PCISPH scenes in one batch (autotuned cell ratio and check interval)
    with every step after the first replayed as captured HIP graphs, and the
    same batch launched directly, against 64 sequential plain solves (the
    reference's main loop per scene through the single-scene entry points) at
    the same cell ratio and check interval; all timed over `batch_steps` steps
    after 2 untimed ones, set-up excluded."""
    from reyn_amd import scenes
    from reyn_amd.sim import PCISPH, Batch, BoundarySamples, Particles, Simulation
    from reyn_amd.tune import autotune_batch

    scs = scenes.config4_batch()
    n = sum(sc.n for sc in scs)
    opts = dict(solver="pcisph", kernel="W6", eta_rho_max=0.001, min_iter=3, max_iter=100)
    best, trials = autotune_batch(ctx, scs, log=log, **opts)

    def run_batch(graph):
        b = Batch(ctx, scs, graph=graph, **opts, **best)
        b.step(2)
        ctx.synchronize()
        t = time.perf_counter()
        b.step(args.batch_steps)
        ctx.synchronize()
        dt = time.perf_counter() - t
        its = [b.info(k).iterations for k in range(len(scs))]
        b.close()
        return dt, its
    t_graph, it_b = run_batch(True)
    t_direct, it_d = run_batch(False)
    assert it_b == it_d, "captured and direct batches took different iterations"
    t_seq = 0.0
    for sc in scs:
        h = sc.spacing
        state = Particles(ctx, sc.x, None, sc.mass, sc.rho_0)
        bdy = BoundarySamples(ctx, sc.xb, h, sc.rho_0)
        solver = PCISPH(ctx, "W6", sc.n, 0.001, h, sc.rho_0, eta_rho_max=0.001, min_iter=3, max_iter=100,
                        check_every=best["check_every"])
        sim = Simulation(ctx, solver, state, bdy, h, sc.bound_min, sc.bound_max, cfl_lambda=1.0,
                         cell_ratio=best["cell_ratio"], dt_policy="controller", dt0=0.0005, iter_target=10,
                         change_factor=0.001, cell_order=best["cell_order"])
        sim.step()
        sim.step()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(args.batch_steps):
            sim.step()
        ctx.synchronize()
        t_seq += time.perf_counter() - t
        for o in (solver, bdy, state, sim.builder):
            getattr(o, "close", lambda: None)()
    v_b = n * args.batch_steps / t_graph
    v_d = n * args.batch_steps / t_direct
    v_s = n * args.batch_steps / t_seq
    log(f"batch: {v_b / 1e6:.1f} M particle-steps/s batched with graphs, {v_d / 1e6:.1f} M launched directly, "
        f"{v_s / 1e6:.1f} M sequential plain at the same settings, x{v_b / v_s:.2f}; iterations "
        f"{min(it_b)}..{max(it_b)}")
    return {
        "workload": f"64 PCISPH dam-break scenes of 10k-300k particles (log-uniform, seed 4; {n} in total), "
                    "W6 kernel, eta 0.1%, min 3 iterations, time-step controller, grid + reorder every step",
        "particles": n, "steps": args.batch_steps,
        "batched_particle_steps_per_s": round(v_b, 1),
        "batched_direct_particle_steps_per_s": round(v_d, 1),
        "sequential_plain_particle_steps_per_s": round(v_s, 1),
        "speedup": round(v_b / v_s, 3),
        "speedup_direct": round(v_d / v_s, 3),
        "settings": f"cell ratio {best['cell_ratio']}, {best['cell_order']} cell order and check interval "
                    f"{best['check_every']} (autotuned) for the batch and the sequential solves alike; the batch "
                    "replays each step after the first as captured HIP graphs (direct launches reported beside)",
        "batched_s": round(t_graph, 4), "batched_direct_s": round(t_direct, 4), "sequential_s": round(t_seq, 4),
        "last_iterations_range": [min(it_b), max(it_b)],
        "autotune": {"chosen": best, "trials": [tr.as_dict() for tr in trials]},
    }


def serve_leg(ctx, args, log) -> dict:
    """Continuous batching of seeded IISPH jobs against FIFO one at a time."""
    from reyn_amd import scenes
    from reyn_amd.serve import (Server, SizeClassServer, make_jobs, poisson_arrivals, run_fifo, run_fifo_plain,
                                same_as_plain, same_results)
    from reyn_amd.sim import Batch

    spec = scenes.config5_jobs(args.serve_jobs)
    opts = dict(solver="iisph", kernel="B3", rho_0=1000.0, eta_rho_max=0.001, omega=0.5, min_iter=3,
                max_iter=100, cell_ratio=1.0, check_every=2)
    # capacity: every job resident at once, 10 steps after 2
    b = Batch(ctx, [sc for sc, _ in spec], **opts)
    b.step(2)
    ctx.synchronize()
    t = time.perf_counter()
    b.step(10)
    ctx.synchronize()
    cap = sum(sc.n for sc, _ in spec) * 10 / (time.perf_counter() - t)
    b.close()
    mean_work = float(np.mean([sc.n * st for sc, st in spec]))
    rate = args.serve_load * cap / mean_work
    arrivals = poisson_arrivals(len(spec), rate, seed=5)
    log(f"serving: capacity {cap / 1e6:.1f} M particle-steps/s, {rate:.2f} arrivals/s")
    # three size classes split at 20k and 150k particles: of the splits tried
    # on this workload (50k; 100k; 23k and 108k; 20k and 150k; 15k, 50k and
    # 160k; 12k, 30k, 80k and 200k) the one with the lowest p99 at the same p50
    thresholds = (20000, 150000)
    served = make_jobs(spec, arrivals)
    srv = SizeClassServer(ctx, thresholds=thresholds, device_setup=True, **opts)
    rep = srv.run(served)
    srv.close()
    single = make_jobs(spec, arrivals)
    srv1 = Server(ctx, device_setup=True, **opts)
    rep1 = srv1.run(single)
    srv1.close()
    alone = make_jobs(spec, arrivals)
    fifo = run_fifo(ctx, alone, device_setup=True, **opts)
    same, diff = same_results(served, alone)
    same1, diff1 = same_results(single, alone)
    # the plain one-at-a-time baseline: every job through the single-scene
    # entry points driven by the reference's main loop, set-up included
    plain = make_jobs(spec, arrivals)
    fifo_plain = run_fifo_plain(ctx, plain, **opts)
    same_p, diff_p = same_as_plain(served, plain)
    work = sum(j.particle_steps for j in plain)
    plain_cap = work / fifo_plain["busy_s"]
    # saturated arrivals: every job waiting at the start, so the served rate
    # is the server's capacity on this job mix rather than the offered load
    sat = make_jobs(spec, np.zeros(len(spec)))
    srv2 = SizeClassServer(ctx, thresholds=thresholds, device_setup=True, **opts)
    rep_sat = srv2.run(sat)
    srv2.close()
    same_s, _ = same_results(sat, alone)
    log(f"serving: p50 {rep['p50_s']:.2f} s / p99 {rep['p99_s']:.2f} s in size classes {thresholds}, "
        f"p50 {rep1['p50_s']:.2f} s / p99 {rep1['p99_s']:.2f} s in one batch, p50 {fifo['p50_s']:.2f} s / "
        f"p99 {fifo['p99_s']:.2f} s FIFO; {rep['particle_steps_per_s'] / 1e6:.1f} vs "
        f"{fifo['particle_steps_per_s'] / 1e6:.1f} M particle-steps/s; {same}/{len(spec)} and "
        f"{same1}/{len(spec)} results identical")
    log(f"serving: plain one at a time p50 {fifo_plain['p50_s']:.2f} s / p99 {fifo_plain['p99_s']:.2f} s, "
        f"{fifo_plain['particle_steps_per_s'] / 1e6:.1f} M particle-steps/s (capacity {plain_cap / 1e6:.1f} M); "
        f"served/plain x{rep['particle_steps_per_s'] / fifo_plain['particle_steps_per_s']:.2f}; saturated: "
        f"{rep_sat['particle_steps_per_s'] / 1e6:.1f} M, x{rep_sat['particle_steps_per_s'] / plain_cap:.2f} the "
        f"plain capacity; {same_p}/{len(spec)} served results equal the plain results, {same_s}/{len(spec)} "
        "saturated")
    return {
        "workload": f"{len(spec)} IISPH dam-break jobs (seed 5): N log-uniform in [5k, 500k], box sides U[1,3], "
                    "sphere obstacle with p 0.5, steps U{200..2000}, B3, stop at mean error <= 0.1% or 100 "
                    f"iterations; Poisson arrivals at {args.serve_load:.0%} of the measured batch capacity; each "
                    "job is built on the GPU from its description when admitted (in its time-to-result)",
        "capacity_particle_steps_per_s": round(cap, 1), "arrival_rate_per_s": round(rate, 4),
        "server": f"size classes below {thresholds[0]}, below {thresholds[1]} and above particles, each a "
                  "continuously batched batch on its own stream and host thread, one HBM budget",
        "served": rep, "served_single_batch": rep1, "fifo_one_at_a_time": fifo,
        "p50_speedup": round(fifo["p50_s"] / rep["p50_s"], 3),
        "p99_speedup": round(fifo["p99_s"] / rep["p99_s"], 3),
        "throughput_speedup": round(rep["particle_steps_per_s"] / fifo["particle_steps_per_s"], 3),
        "results_identical": f"{same}/{len(spec)}", "jobs_differing": diff,
        "single_batch_results_identical": f"{same1}/{len(spec)}",
        "fifo_plain_one_at_a_time": fifo_plain,
        "p50_speedup_vs_plain": round(fifo_plain["p50_s"] / rep["p50_s"], 3),
        "p99_speedup_vs_plain": round(fifo_plain["p99_s"] / rep["p99_s"], 3),
        "throughput_speedup_vs_plain": round(rep["particle_steps_per_s"] / fifo_plain["particle_steps_per_s"], 3),
        "plain_results_identical": f"{same_p}/{len(spec)}", "jobs_differing_from_plain": diff_p,
        "saturated": {"arrivals": "all jobs waiting at the start", "served": rep_sat,
                      "results_identical": f"{same_s}/{len(spec)}",
                      "plain_capacity_particle_steps_per_s": round(plain_cap, 1),
                      "throughput_vs_plain_capacity": round(rep_sat["particle_steps_per_s"] / plain_cap, 3),
                      "throughput_vs_batch_of_one_fifo": round(rep_sat["particle_steps_per_s"]
                                                               / (work / fifo["busy_s"]), 3)},
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=480,
                    help="untimed steps of the rollout before the timed window (480: the window is steps 480-499)")
    ap.add_argument("--early-window", type=int, nargs=2, default=(5, 25), metavar=("FIRST", "END"),
                    help="steps [FIRST, END) of the warm-up timed on their own (the early, near-static state)")
    ap.add_argument("--side", type=int, default=128, help="fluid lattice side (128 -> 2,097,152 particles)")
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="IISPH steps per run of the CPU baseline sample (about 1-2 s each on 16 cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the oracle leg (FP64 check of one step, CPU baseline, configuration 1)")
    ap.add_argument("--no-config1", action="store_true", help="skip configuration 1 (GPU and CPU) in the oracle leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch leg")
    ap.add_argument("--batch-steps", type=int, default=8)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-jobs", type=int, default=32)
    ap.add_argument("--serve-load", type=float, default=0.8)
    ap.add_argument("--only", choices=("batch", "serve", "grid"), default=None,
                    help="run only the batch, serving or grid leg and print its object as one JSON line")
    ap.add_argument("--grid-reps", type=int, default=10)
    ap.add_argument("--no-grid", action="store_true", help="skip the grid and SESPH leg")
    ap.add_argument("--no-profile", action="store_true", help="time without per-launch events")
    ap.add_argument("--dt-policy", choices=("controller", "cfl"), default="controller",
                    help="controller: the reference main loop's simple_dt_controller from dt0 = 5e-4 "
                         "(lambda 1, 10 target iterations, 0.1%% change per step); cfl: cfl_time_step(lambda)")
    ap.add_argument("--cfl-lambda", type=float, default=None)
    ap.add_argument("--cell-ratio", type=float, default=None,
                    help="grid cell size / search radius (default: autotuned over 0.5 and 1.0, with the order)")
    ap.add_argument("--cell-order", choices=("xyz", "brick"), default=None,
                    help="cell order of the fluid grid (default: autotuned with the ratio)")
    ap.add_argument("--check-every", type=int, default=2,
                    help="pressure-loop iterations between host checks after the first chunk (the first chunk is "
                         "one less than the previous step's iteration count, at least min_iter)")
    ap.add_argument("--bdy-cell-ratio", type=float, default=1.0,
                    help="cell size / search radius of the boundary grid (1.0: 9 strips per boundary walk)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run an identity gather of 2^26 floats (reads 512 MiB, writes 256 MiB), the "
                         "plain copy kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes")
    ap.add_argument("--rocprof-summary", nargs="+", metavar=("KIND", "DIR"), default=None,
                    help="summarise rocprofv3 CSV output of a run of this script over its timed window and print "
                         "JSON: trace DIR | valu DIR | address DIR (TA_TA_BUSY_sum, TA_ADDR_STALLED_BY_TC_CYCLES_sum) | "
                         "traffic FETCH_DIR WRITE_DIR CALIB_FETCH_DIR CALIB_WRITE_DIR "
                         "(the calibration dirs from --copy-calibration runs); uses --warmup")
    args = ap.parse_args()
    if args.rocprof_summary:
        print(json.dumps(rocprof_summary(args.rocprof_summary[0], args.rocprof_summary[1:], args.warmup), indent=1))
        return 0

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.copy_calibration:
        import ctypes as C

        from reyn_amd.device import Context
        n = 1 << 26
        with Context(0) as ctx:
            src = ctx.array(np.arange(n, dtype=np.float32))
            dst = ctx.empty(n)
            idx = ctx.array(np.arange(n, dtype=np.uint32))
            sp = (C.c_void_p * 1)(src.ptr)
            dp = (C.c_void_p * 1)(dst.ptr)
            for _ in range(3):
                ctx.check(ctx._lib.reyn_gather_f32(ctx.h, idx._ptr, n, sp, dp, 1))
            ctx.synchronize()
            ok = bool(np.array_equal(dst.numpy()[:: 1 << 16], np.arange(n, dtype=np.float32)[:: 1 << 16]))
        print(json.dumps({"copy_calibration": {"kernel": "k_gather", "launches": 3, "read_bytes": 8 * n,
                                               "write_bytes": 4 * n, "ok": ok}}))
        return 0

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the path is built and measured "
              "on a single GPU; scenes and batches would shard by scene, one large scene by z-slabs of the cell "
              "order", file=sys.stderr)
        return 2

    from reyn_amd import capi, scenes
    from reyn_amd.device import Context
    from reyn_amd.sim import IISPH, BoundarySamples, Particles, Simulation, reduce

    if args.only:
        legs = {"batch": ("batch", batch_leg), "serve": ("serving", serve_leg), "grid": ("grid", grid_leg)}
        key, fn = legs[args.only]
        with Context(0) as ctx:
            leg = fn(ctx, args, log)
        print(json.dumps({key: leg}), flush=True)
        return 0

    t_setup = time.perf_counter()
    sc = scenes.config3_iisph(side=args.side)
    log(f"scene: {sc.n} fluid, {sc.n_bdy} boundary, spacing {sc.spacing:.6f}, sphere {sc.meta.get('sphere')}")
    ctx = Context(0)
    lam = args.cfl_lambda if args.cfl_lambda is not None else (1.0 if args.dt_policy == "controller" else 0.4)

    def make(ratio, order="xyz"):
        state = Particles(ctx, sc.x, None, sc.mass, sc.rho_0)
        bdy = BoundarySamples(ctx, sc.xb, sc.spacing, sc.rho_0, args.bdy_cell_ratio)
        solver = IISPH(ctx, "B3", sc.n, 0.001, sc.spacing, sc.rho_0, omega=0.5, eta_rho_max=0.001, min_iter=3,
                       max_iter=100, stop_rule=capi.IISPH_STOP_MEAN, warm_start=True,
                       check_every=args.check_every)
        sim = Simulation(ctx, solver, state, bdy, sc.spacing, sc.bound_min, sc.bound_max, cfl_lambda=lam,
                         cell_ratio=ratio, extra_reorder=("p",), dt_policy=args.dt_policy, dt0=0.0005,
                         iter_target=10, change_factor=0.001, cell_order=order)

        def close():
            for o in (solver, bdy, state, sim.builder):
                getattr(o, "close", lambda: None)()
        sim.close = close
        return sim

    tune = None
    ratio, order = args.cell_ratio, args.cell_order
    if ratio is None:
        from reyn_amd.tune import autotune_step
        (ratio, order), trials = autotune_step(ctx, make, orders=(order,) if order else ("xyz", "brick"), log=log)
        tune = {"chosen_cell_ratio": ratio, "chosen_cell_order": order, "trials_s": trials,
                "trial": "fresh scene, 3 steps after 2"}
    order = order or "xyz"
    sim = make(ratio, order)
    state, bdy, solver = sim.state, sim.bdy, sim.solver
    ncells = sim.builder.geometry().ncells
    log(f"setup {time.perf_counter() - t_setup:.1f} s; fluid grid {ncells} cells")

    # the rollout: `warmup` untimed steps (the early window of steps
    # early[0] .. early[1] - 1 is timed by itself on the way), then the timed
    # window of `steps` steps
    early = tuple(args.early_window)
    # the last `steps` warm-up steps are the phase window: every launch timed
    # with HIP events, for the per-kernel table and to find the dominant
    # kernel; the timed window after it records events around that kernel's
    # launches only (events around every launch add about 6% to a step)
    phase_first = max(args.warmup - args.steps, 0) if not args.no_profile else args.warmup
    hist = {}
    t_roll = time.perf_counter()
    early_s = None
    phase_iters = []
    for k in range(args.warmup):
        if k == early[0]:
            ctx.synchronize()
            t_e = time.perf_counter()
        if k == phase_first:
            ctx.synchronize()
            ctx.check(ctx._lib.reyn_profile_select(ctx.h, None))
            ctx.check(ctx._lib.reyn_profile_enable(ctx.h, 1))
            t_phase = time.perf_counter()
        s = sim.step()
        hist[s.iterations] = hist.get(s.iterations, 0) + 1
        if k >= phase_first:
            phase_iters.append(s.iterations)
        if k + 1 == early[1]:
            ctx.synchronize()
            early_s = time.perf_counter() - t_e
        if k % 40 == 0 or k + 1 == args.warmup:
            log(f"warmup {k}: dt {s.dt:.3e} iters {s.iterations} mean err {s.mean_rel_error:.2e}")

    ctx.synchronize()
    phase_ms = (time.perf_counter() - t_phase) * 1e3 if phase_iters else 0.0
    phase_prof = {}
    dominant = None
    if phase_iters:
        phase_prof = profile_summary(ctx)
        ranked = sorted((v[1], name) for name, v in phase_prof.items()
                        if algorithmic_bytes(name, sc.n, sc.n_bdy, 1, 8) is not None)
        dominant = ranked[-1][1] if ranked else None
    ctx.check(ctx._lib.reyn_profile_enable(ctx.h, 1 if dominant else 0))
    ctx.check(ctx._lib.reyn_profile_select(ctx.h, dominant.encode() if dominant else None))
    iters = []
    errs = []
    dts = []
    t0 = time.perf_counter()
    ctx.timer_start()
    for k in range(args.steps):
        s = sim.step()
        iters.append(s.iterations)
        errs.append(s.mean_rel_error)
        dts.append(s.dt)
        hist[s.iterations] = hist.get(s.iterations, 0) + 1
    gpu_ms = ctx.timer_stop()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    roll_wall = time.perf_counter() - t_roll
    log(f"timed steps {args.warmup}..{args.warmup + args.steps - 1}: {gpu_ms:.1f} ms (events), {wall * 1e3:.1f} ms "
        f"(wall); iterations {iters}; simulated time {sim.time:.4f} s")

    prof = {}
    if dominant:
        prof = profile_summary(ctx)
        ctx.check(ctx._lib.reyn_profile_enable(ctx.h, 0))
        ctx.check(ctx._lib.reyn_profile_select(ctx.h, None))

    ms_per_step = wall * 1e3 / args.steps
    value = sc.n * args.steps / wall
    nfields = 8
    roofline = None
    # the committed rocprofv3 counters (traffic, VALU and texture-address
    # cycles per launch) were taken on the 128^3 scene; other sizes report
    # none of them rather than a figure of another workload
    profiled = args.side == 128
    phases = []  # per kernel: HBM fraction (algorithmic bytes) and VALU fraction (counted instructions)
    step_valu_frac = None
    if phase_prof:
        # per-kernel table of the phase window (every launch timed): working
        # launches only, so a loop launch that returned at once (its loop had
        # converged) does not dilute a kernel's mean
        log(f"phase window, steps {phase_first}..{args.warmup - 1} ({phase_ms:.1f} ms, iterations {phase_iters}):")
        log(f"{'kernel':<18}{'launches':>9}{'working':>9}{'total ms':>11}{'work us':>10}{'GB/s':>9}{'HBM %':>7}"
            f"{'VALU %':>8}")
        valu_total = 0.0
        for name, (cnt, tot, wcnt, wtot) in sorted(phase_prof.items(), key=lambda kv: -kv[1][1]):
            by = algorithmic_bytes(name, sc.n, sc.n_bdy, ncells, nfields)
            avg_us = wtot * 1e3 / wcnt
            gbs = by / (avg_us * 1e-6) / 1e9 if by else float("nan")
            vi = measured_valu(name) if profiled else None
            vf = vi / (avg_us * 1e-6) / VALU_PEAK_INSTS if vi else float("nan")
            ta = measured_ta_busy(name) if profiled else None
            log(f"{name:<18}{cnt:>9}{wcnt:>9}{tot:>11.3f}{avg_us:>10.1f}{gbs:>9.0f}{100 * gbs / HBM_PEAK_GBS:>7.1f}"
                f"{100 * vf:>8.1f}")
            if vi:
                valu_total += vi * wcnt
            phases.append({"kernel": name, "launches": cnt, "working_launches": wcnt,
                           "working_avg_us": round(avg_us, 2), "total_ms": round(tot, 3),
                           "hbm_frac": round(gbs / HBM_PEAK_GBS, 5) if by else None,
                           "valu_frac": round(vf, 4) if vi else None,
                           "ta_busy_frac": round(ta / (avg_us * 1e-6) / TA_CYCLES_PER_S, 4) if ta else None})
        step_valu_frac = valu_total / (phase_ms * 1e-3) / VALU_PEAK_INSTS
    if prof and dominant in prof:
        # the dominant kernel over the timed window (events around its launches only)
        cnt, tot, wcnt, wtot = prof[dominant]
        by = algorithmic_bytes(dominant, sc.n, sc.n_bdy, ncells, nfields)
        avg_us = wtot * 1e3 / wcnt
        gbs = by / (avg_us * 1e-6) / 1e9
        vi = measured_valu(dominant) if profiled else None
        tr = measured_traffic(dominant) if profiled else None
        ta = measured_ta_busy(dominant) if profiled else None
        roofline = {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": tr, "kernel": dominant,
                    "bytes_per_launch": by, "avg_launch_us": round(avg_us, 2), "launches": cnt,
                    "working_launches": wcnt, "noop_launches_ms": round(tot - wtot, 3),
                    "share_of_step_time": round(tot / gpu_ms, 4),
                    "traffic_over_algorithmic": round(tr / by, 3) if tr else None,
                    "valu_frac": round(vi / (avg_us * 1e-6) / VALU_PEAK_INSTS, 4) if vi else None,
                    # the unit that binds the listed loops: the texture-address
                    # path, busy this fraction of the launch (counted cycles
                    # from the committed profile over the live launch time)
                    "ta_busy_frac": round(ta / (avg_us * 1e-6) / TA_CYCLES_PER_S, 4) if ta else None}
        log(f"timed window: {dominant} {wcnt} working launches of {cnt} (sum of iterations {sum(iters)}), "
            f"{avg_us:.1f} us each, {gbs:.0f} GB/s algorithmic ({100 * gbs / HBM_PEAK_GBS:.2f}% of HBM peak)")
    mean_iters = float(np.mean(iters))
    sim_time_end = sim.time
    sbytes = step_bytes(sc.n, sc.n_bdy, ncells, mean_iters)
    step_frac = sbytes / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9)

    rho_err = abs(reduce(ctx, solver.rho, "avg") / sc.rho_0 - 1.0)
    cpu, fp64, config1 = None, None, None
    if not args.no_cpu_baseline:
        cpu, fp64, config1 = cpu_leg(ctx, sim, sc, args, log)
    sim.close()
    batch = None if args.no_batch else batch_leg(ctx, args, log)
    serving = None if args.no_serve else serve_leg(ctx, args, log)
    grid = None if args.no_grid else grid_leg(ctx, args, log)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"IISPH dam break, {sc.n} FP32 fluid particles (128^3 block at s=1/128) + {sc.n_bdy} "
                        "boundary samples (2 wall layers of box [0,2]x[0,1.25]x[0,2] + sphere obstacle), B3 "
                        "kernel, omega 0.5, stop at mean error <= 0.1% or 100 iterations, grid rebuild + "
                        f"reorder every step, dt policy {args.dt_policy}",
            "fluid_particles": sc.n,
            "boundary_particles": sc.n_bdy,
            "grid_cells": ncells,
            "cell_ratio": ratio,
            "cell_order": order,
            "boundary_cell_ratio": args.bdy_cell_ratio,
            "cell_ratio_autotune": tune,
            "timed_steps": [args.warmup, args.warmup + args.steps - 1],
            "jacobi_iterations_mean": round(mean_iters, 2),
            "jacobi_iterations": iters,
            "rollout_jacobi_iterations_histogram": {str(k): hist[k] for k in sorted(hist)},
            "rollout_steps": args.warmup + args.steps,
            "rollout_particle_steps_per_s": round(sc.n * (args.warmup + args.steps) / roll_wall, 1),
            "early_window": ({"steps": [early[0], early[1] - 1],
                              "particle_steps_per_s": round(sc.n * (early[1] - early[0]) / early_s, 1),
                              "ms_per_step": round(early_s * 1e3 / (early[1] - early[0]), 4)}
                             if early_s else None),
            "simulated_time_end": sim_time_end,
            "fp64_step_errors": fp64,
            "mean_density_error_max": max(errs) if errs else None,
            "dt_policy": args.dt_policy,
            "cfl_lambda": lam,
            "dt_first_last": [dts[0], dts[-1]] if dts else None,
            "simulated_time_timed": float(sum(dts)),
            "density_mean_rel_dev": rho_err,
            "step_hbm_frac": round(step_frac, 5),
            "step_valu_frac": round(step_valu_frac, 4) if step_valu_frac else None,
            "phases_window": ([phase_first, args.warmup - 1] if phase_iters else None),
            "phases_window_jacobi_iterations": phase_iters,
            "phases": phases,
            "step_bytes": sbytes,
            "event_ms_timed": round(gpu_ms, 3),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "config1": config1,
        "batch": batch,
        "serving": serving,
        "grid": grid,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
