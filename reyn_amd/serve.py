"""Serving: a request loop that keeps simulation jobs resident in one
continuously batched ``reyn_batch``.

Jobs arrive over time; each is a scene and a number of steps. Waiting jobs
are admitted in arrival order whenever the device bytes of the resident jobs
plus the next job fit the HBM budget (``reyn_batch_scene_bytes`` per job).
Every batch step advances all resident jobs by one step of their own (own
grid, boundary, time step and pressure loop, in one launch per phase), a job
that has taken its steps is downloaded and leaves, and jobs that arrived
meanwhile join before the next step. A job's result does not depend on which
jobs shared its steps: it is the same, bit for bit, as the job run alone.

The reference has no serving layer (one scene per process, src/main.cu:46-78);
each job's step is the reference's IISPH or PCISPH step with the main loop's
grid, time-step controller and bounds (main.cu:59-78).

`run_fifo` is the one-job-at-a-time baseline on the same arrivals: each job
alone through the same kernels, in arrival order, with its service time
measured and its time-to-result taken on the arrival clock.

With `device_setup`, a job is a description (`scenes.dam_break_desc` of its
scene) and both the server and the baseline build it on the GPU when it is
admitted, so its set-up is part of its time-to-result.
"""

from __future__ import annotations

import ctypes as C
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .device import Context
from .sim import Batch, batch_params, scene_bytes

#: Device memory held back from the admission budget (runtime, the grid and
#: layout rebuilds that run while jobs join and leave).
RESERVE_BYTES = 16 << 30


@dataclass
class Job:
    """One request: a scene advanced `steps` steps from rest. Times are
    seconds on the run's clock (0 = the start of the run)."""

    jid: int
    scene: object
    steps: int
    arrival: float = 0.0
    bytes: int = 0
    admitted: float | None = None
    finished: float | None = None
    taken: int = 0
    result: dict | None = None
    info: dict = field(default_factory=dict)

    @property
    def particle_steps(self) -> int:
        return int(self.scene.n) * int(self.steps)

    @property
    def time_to_result(self) -> float:
        return self.finished - self.arrival


def device_memory(ctx: Context) -> tuple[int, int]:
    free, total = C.c_uint64(), C.c_uint64()
    ctx.check(ctx._lib.reyn_device_memory(ctx.h, C.byref(free), C.byref(total)), "device memory")
    return int(free.value), int(total.value)


def poisson_arrivals(count: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times of a Poisson process of `rate` per second (seeded)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


def summary(jobs: list[Job], wall: float) -> dict:
    """p50/p99 time-to-result, mean queueing delay and particle-steps/s of a run
    (all particle-steps over the span from the first arrival to the last
    result)."""
    ttr = np.array([j.time_to_result for j in jobs])
    span = max(j.finished for j in jobs) - min(j.arrival for j in jobs)
    work = sum(j.particle_steps for j in jobs)
    wait = [j.admitted - j.arrival for j in jobs if j.admitted is not None]
    return {
        "jobs": len(jobs),
        "p50_s": float(np.percentile(ttr, 50)),
        "p99_s": float(np.percentile(ttr, 99)),
        "mean_s": float(ttr.mean()),
        "max_s": float(ttr.max()),
        "mean_wait_s": float(np.mean(wait)) if wait else None,
        "particle_steps": int(work),
        "span_s": float(span),
        "wall_s": float(wall),
        "particle_steps_per_s": float(work / span),
    }


class Server:
    """The request loop over one batch. `budget_bytes` defaults to the
    device's total memory less `reserve_bytes`, capped by what is free when the
    server starts; `max_resident` bounds the resident jobs."""

    def __init__(self, ctx: Context, budget_bytes: int | None = None, reserve_bytes: int = RESERVE_BYTES,
                 max_resident: int = 512, keep_results: bool = True, device_setup: bool = False,
                 **batch_options):
        self.ctx = ctx
        self.device_setup = device_setup
        self.options = dict(batch_options)
        self.batch = Batch(ctx, [], **self.options)
        self.prm = self.batch.prm
        free, total = device_memory(ctx)
        limit = min(total - reserve_bytes, free - reserve_bytes)
        self.budget = int(limit if budget_bytes is None else min(budget_bytes, limit))
        if self.budget <= 0:
            raise RuntimeError(f"serve: no device memory left for jobs (free {free}, reserve {reserve_bytes})")
        self.max_resident = int(max_resident)
        self.keep_results = keep_results
        self.used = 0
        self.peak_used = 0
        self.peak_resident = 0
        self.batch_steps = 0
        self.layout_changes = 0

    def job_bytes(self, scene) -> int:
        return scene_bytes(self.ctx, self.prm, scene)

    def close(self) -> None:
        self.batch.close()

    def run(self, jobs: list[Job], clock=time.perf_counter, log=None) -> dict:
        """Serves `jobs` (arrival times relative to the start) in real time and
        returns the run's summary; each job gets its admission and finish
        times, its step record and (with keep_results) its final state."""
        jobs = sorted(jobs, key=lambda j: (j.arrival, j.jid))
        for j in jobs:
            j.bytes = self.job_bytes(j.scene)
            if j.bytes > self.budget:
                raise RuntimeError(f"serve: job {j.jid} needs {j.bytes} bytes, more than the budget {self.budget}")
        waiting: deque[Job] = deque()
        resident: dict[int, Job] = {}
        nxt = 0
        t0 = clock()
        while nxt < len(jobs) or waiting or resident:
            now = clock() - t0
            while nxt < len(jobs) and jobs[nxt].arrival <= now:
                waiting.append(jobs[nxt])
                nxt += 1
            admit = []
            while (waiting and len(resident) + len(admit) < self.max_resident
                   and self.used + waiting[0].bytes <= self.budget):
                j = waiting.popleft()
                self.used += j.bytes
                admit.append(j)
            if admit:
                scs = [j.scene for j in admit]
                ids = self.batch.add_dam_breaks(scs) if self.device_setup else self.batch.add(scs)
                self.layout_changes += 1
                t_adm = clock() - t0
                for j, bid in zip(admit, ids):
                    j.admitted = t_adm
                    resident[bid] = j
                self.peak_used = max(self.peak_used, self.used)
                self.peak_resident = max(self.peak_resident, len(resident))
                if log:
                    log(f"serve t={t_adm:7.3f}s admitted {[j.jid for j in admit]}, resident {len(resident)}, "
                        f"{self.used / 2**30:.2f} GiB")
            if not resident:
                if nxt < len(jobs):
                    time.sleep(max(0.0, jobs[nxt].arrival - (clock() - t0)))
                continue
            self.batch.step(1)
            self.batch_steps += 1
            done = []
            for bid, j in resident.items():
                j.taken += 1
                if j.taken >= j.steps:
                    done.append(bid)
            if done:
                fin = []
                for bid in done:
                    j = resident.pop(bid)
                    if self.keep_results:
                        j.result = self.batch.download(bid)
                    inf = self.batch.info(bid)
                    j.info = {"steps": inf.steps, "iterations": inf.iterations, "dt": inf.dt, "time": inf.time}
                    self.used -= j.bytes
                    fin.append(j)
                self.batch.remove(done)
                self.layout_changes += 1
                t_fin = clock() - t0
                for j in fin:
                    j.finished = t_fin
                if log:
                    log(f"serve t={t_fin:7.3f}s finished {len(done)}, resident {len(resident)}")
        wall = clock() - t0
        out = summary(jobs, wall)
        out.update({"batch_steps": self.batch_steps, "layout_changes": self.layout_changes,
                    "peak_resident": self.peak_resident, "peak_bytes": int(self.peak_used),
                    "budget_bytes": int(self.budget)})
        return out


class SizeClassServer:
    """Jobs split by particle count into size classes (`thresholds`, ascending:
    class k holds jobs with thresholds[k-1] <= n < thresholds[k]), each class
    served by its own batch on its own context (its own HIP stream at default
    priority) from its own host thread. A small job then steps at the pace of
    the small jobs, not of the largest resident job, while the kernels of the
    classes share the GPU. Admission is in arrival order within a class, under
    one HBM budget for all classes. Results are the same as alone, bit for bit,
    as in `Server`."""

    def __init__(self, ctx: Context, thresholds=(50000,), budget_bytes: int | None = None,
                 reserve_bytes: int = RESERVE_BYTES, keep_results: bool = True, device_setup: bool = False,
                 **batch_options):
        self.thresholds = tuple(sorted(int(t) for t in thresholds))
        self.device_setup = device_setup
        self.keep_results = keep_results
        self.ctxs = [ctx] + [Context(ctx.device) for _ in self.thresholds]
        self.batches = [Batch(c, [], **batch_options) for c in self.ctxs]
        self.prm = self.batches[0].prm
        free, total = device_memory(ctx)
        limit = min(total - reserve_bytes, free - reserve_bytes)
        self.budget = int(limit if budget_bytes is None else min(budget_bytes, limit))
        if self.budget <= 0:
            raise RuntimeError(f"serve: no device memory left for jobs (free {free}, reserve {reserve_bytes})")
        self.used = 0
        self.peak_used = 0

    def job_bytes(self, scene) -> int:
        return scene_bytes(self.ctxs[0], self.prm, scene)

    def class_of(self, job: Job) -> int:
        k = 0
        while k < len(self.thresholds) and job.scene.n >= self.thresholds[k]:
            k += 1
        return k

    def close(self) -> None:
        for b in self.batches:
            b.close()
        for c in self.ctxs[1:]:
            c.close()

    def run(self, jobs: list[Job], clock=time.perf_counter) -> dict:
        jobs = sorted(jobs, key=lambda j: (j.arrival, j.jid))
        for j in jobs:
            j.bytes = self.job_bytes(j.scene)
            if j.bytes > self.budget:
                raise RuntimeError(f"serve: job {j.jid} needs {j.bytes} bytes, more than the budget {self.budget}")
        lock = threading.Lock()
        waiting = [deque() for _ in self.batches]
        state = {"next": 0, "finished": 0}
        stats = [{"batch_steps": 0, "layout_changes": 0, "peak_resident": 0, "jobs": 0} for _ in self.batches]
        errors = []
        t0 = clock()

        def arrivals(now):  # under the lock
            while state["next"] < len(jobs) and jobs[state["next"]].arrival <= now:
                j = jobs[state["next"]]
                waiting[self.class_of(j)].append(j)
                state["next"] += 1

        def lane(k):
            batch, resident, st = self.batches[k], {}, stats[k]
            try:
                while True:
                    admit = []
                    with lock:
                        arrivals(clock() - t0)
                        q = waiting[k]
                        while q and self.used + q[0].bytes <= self.budget:
                            j = q.popleft()
                            self.used += j.bytes
                            admit.append(j)
                        self.peak_used = max(self.peak_used, self.used)
                        finished_all = state["finished"] == len(jobs)
                    if admit:
                        scs = [j.scene for j in admit]
                        ids = batch.add_dam_breaks(scs) if self.device_setup else batch.add(scs)
                        st["layout_changes"] += 1
                        t_adm = clock() - t0
                        for j, bid in zip(admit, ids):
                            j.admitted = t_adm
                            resident[bid] = j
                        st["peak_resident"] = max(st["peak_resident"], len(resident))
                        st["jobs"] += len(admit)
                    if not resident:
                        if finished_all:
                            return
                        time.sleep(0.0005)
                        continue
                    batch.step(1)
                    st["batch_steps"] += 1
                    done = []
                    for bid, j in resident.items():
                        j.taken += 1
                        if j.taken >= j.steps:
                            done.append(bid)
                    if done:
                        fin = []
                        for bid in done:
                            j = resident.pop(bid)
                            if self.keep_results:
                                j.result = batch.download(bid)
                            inf = batch.info(bid)
                            j.info = {"steps": inf.steps, "iterations": inf.iterations, "dt": inf.dt,
                                      "time": inf.time}
                            fin.append(j)
                        batch.remove(done)
                        st["layout_changes"] += 1
                        t_fin = clock() - t0
                        with lock:
                            for j in fin:
                                j.finished = t_fin
                                self.used -= j.bytes
                            state["finished"] += len(fin)
            except Exception as e:  # surfaced after the join
                errors.append(e)
                with lock:
                    state["finished"] = len(jobs)

        threads = [threading.Thread(target=lane, args=(k,)) for k in range(len(self.batches))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        out = summary(jobs, clock() - t0)
        out.update({"classes": [{"max_particles": (self.thresholds[k] if k < len(self.thresholds) else None), **stats[k]}
                                for k in range(len(self.batches))],
                    "peak_bytes": int(self.peak_used), "budget_bytes": int(self.budget)})
        return out


def run_fifo(ctx: Context, jobs: list[Job], keep_results: bool = True, clock=time.perf_counter,
             device_setup: bool = False, **batch_options) -> dict:
    """One job at a time in arrival order, each alone (a batch of one: the
    same bits as the single-scene entry points) for all its steps. Service
    times are measured; a job starts when it has arrived and the previous job
    has finished."""
    jobs = sorted(jobs, key=lambda j: (j.arrival, j.jid))
    t_free = 0.0
    busy = 0.0
    for j in jobs:
        t = clock()
        if device_setup:
            b = Batch(ctx, [], rho_0=j.scene.rho_0, **{k: v for k, v in batch_options.items() if k != "rho_0"})
            b.add_dam_breaks([j.scene])
        else:
            b = Batch(ctx, [j.scene], **batch_options)
        b.step(j.steps)
        if keep_results:
            j.result = b.download(0)
        inf = b.info(0)
        j.info = {"steps": inf.steps, "iterations": inf.iterations, "dt": inf.dt, "time": inf.time}
        b.close()
        service = clock() - t
        busy += service
        j.admitted = max(t_free, j.arrival)
        j.finished = j.admitted + service
        j.taken = j.steps
        t_free = j.finished
    out = summary(jobs, busy)
    out["busy_s"] = float(busy)
    return out


def run_fifo_plain(ctx: Context, jobs: list[Job], keep_results: bool = True, clock=time.perf_counter,
                   solver: str = "iisph", kernel: str = "B3", nu: float = 0.001, rho_0: float | None = None,
                   cell_ratio: float = 0.5, dt0: float = 0.0005, cfl_lambda: float = 1.0,
                   change_factor: float = 0.001, iter_target: int = 10, eta_rho_max: float = 0.001,
                   omega: float = 0.5, min_iter: int = 3, max_iter: int = 100, stop_rule: int | None = None,
                   warm_start: bool = True, check_every: int = 4, cell_order: str = "xyz", **_unused) -> dict:
    """One job at a time in arrival order through the plain single-scene
    path: the reference's main loop (src/main.cu:59-78: time-step
    controller, grid with reorder, solver step, bounds) on the single-scene
    entry points, with the batch's options. A job's set-up (upload of its
    scene, boundary grid and masses) is part of its service time. Results are
    the same bits as the job alone in a batch."""
    from . import capi
    from .sim import IISPH, PCISPH, BoundarySamples, Particles, Simulation

    if stop_rule is None:
        stop_rule = capi.IISPH_STOP_MEAN
    jobs = sorted(jobs, key=lambda j: (j.arrival, j.jid))
    t_free = 0.0
    busy = 0.0
    for j in jobs:
        sc = j.scene
        t = clock()
        h = sc.spacing
        rho = sc.rho_0 if rho_0 is None else rho_0
        state = Particles(ctx, sc.x, None, sc.mass, rho)
        bdy = BoundarySamples(ctx, sc.xb, h, rho, cell_ratio=cell_ratio)
        if solver == "iisph":
            sol = IISPH(ctx, kernel, sc.n, nu, h, rho, omega=omega, eta_rho_max=eta_rho_max, min_iter=min_iter,
                        max_iter=max_iter, stop_rule=stop_rule, warm_start=warm_start, check_every=check_every)
            extra = ("p",)
        else:
            sol = PCISPH(ctx, kernel, sc.n, nu, h, rho, eta_rho_max=eta_rho_max, min_iter=min_iter,
                         max_iter=max_iter, check_every=check_every)
            extra = ()
        sim = Simulation(ctx, sol, state, bdy, h, sc.bound_min, sc.bound_max, cfl_lambda=cfl_lambda,
                         cell_ratio=cell_ratio, extra_reorder=extra, dt_policy="controller", dt0=dt0,
                         iter_target=iter_target, change_factor=change_factor, cell_order=cell_order)
        st = None
        for _ in range(j.steps):
            st = sim.step()
        if keep_results:
            res = state.host()
            res = {k: res[k] for k in ("xx", "xy", "xz", "vx", "vy", "vz")}
            res["rho"] = sol.rho.numpy()
            if solver == "iisph":
                res["p"] = sol.p.numpy()
            j.result = res
        else:
            ctx.synchronize()
        j.info = {"steps": j.steps, "iterations": st.iterations if st else 0}
        for o in (sol, bdy, state, sim.builder):
            getattr(o, "close", lambda: None)()
        service = clock() - t
        busy += service
        j.admitted = max(t_free, j.arrival)
        j.finished = j.admitted + service
        j.taken = j.steps
        t_free = j.finished
    out = summary(jobs, busy)
    out["busy_s"] = float(busy)
    return out


def same_as_plain(a: list[Job], plain: list[Job]) -> tuple[int, list[int]]:
    """Jobs (by id) whose final fields or last iteration count differ between
    a batched run and the plain single-scene run."""
    pp = {j.jid: j for j in plain}
    diff = []
    for j in a:
        o = pp[j.jid]
        same = all(np.array_equal(j.result[k], o.result[k]) for k in o.result)
        if not same or j.info.get("iterations") != o.info.get("iterations"):
            diff.append(j.jid)
    return len(a) - len(diff), diff


def make_jobs(scene_steps, arrivals) -> list[Job]:
    return [Job(k, sc, int(st), float(a)) for k, ((sc, st), a) in enumerate(zip(scene_steps, arrivals))]


def same_results(a: list[Job], b: list[Job]) -> tuple[int, list[int]]:
    """Jobs (by id) whose final states differ in any bit between two runs."""
    bb = {j.jid: j for j in b}
    diff = []
    for j in a:
        o = bb[j.jid]
        if any(not np.array_equal(j.result[k], o.result[k]) for k in j.result) or j.info != o.info:
            diff.append(j.jid)
    return len(a) - len(diff), diff


__all__ = ["Job", "Server", "SizeClassServer", "run_fifo", "run_fifo_plain", "same_as_plain", "make_jobs",
           "poisson_arrivals", "summary", "same_results", "device_memory", "batch_params", "RESERVE_BYTES"]
