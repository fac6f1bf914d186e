"""Autotuner for scene batches: picks the grid cell-size ratio and the
pressure loop's host-check interval for a given set of scenes by timing
short runs of fresh batches.

Every trial starts from the same initial state and runs the same steps, so
the trials compare equal work: the check interval does not change results at
all (the loop control runs on the device and each scene stops at its own
iteration), and the cell ratio changes only the order in which neighbour
terms are summed (the neighbour sets are the same).

The workgroup size is not a tuning knob: every kernel runs 256 lanes (four
waves), the size the density pass's per-wave neighbour-list staging in LDS
and its ballot flush are laid out for, and workgroups are aligned to scene
starts at that size.

The reference has no tuner (its cell size is fixed at half the search radius,
src/scene/uniformgrid.cu:69).
"""

from __future__ import annotations

import time
from dataclasses import dataclass

from .device import Context
from .sim import Batch

#: cell ratios (cell size / search radius) and host-check intervals tried
RATIOS = (0.5, 1.0)
CHECKS = (1, 2, 4)


@dataclass
class Trial:
    cell_ratio: float
    check_every: int
    seconds: float
    particle_steps_per_s: float

    def as_dict(self) -> dict:
        return {"cell_ratio": self.cell_ratio, "check_every": self.check_every, "seconds": round(self.seconds, 5),
                "particle_steps_per_s": round(self.particle_steps_per_s, 1)}


def autotune_batch(ctx: Context, scenes, ratios=RATIOS, checks=CHECKS, warmup: int = 2, steps: int = 4,
                   log=None, **options) -> tuple[dict, list[Trial]]:
    """Times `steps` steps (after `warmup`) of a fresh batch of `scenes` for
    every (cell ratio, check interval) pair; returns the fastest pair as batch
    options and every trial."""
    n = sum(sc.n for sc in scenes)
    trials = []
    for r in ratios:
        for ce in checks:
            b = Batch(ctx, scenes, cell_ratio=r, check_every=ce, **options)
            try:
                b.step(warmup)
                ctx.synchronize()
                t = time.perf_counter()
                b.step(steps)
                ctx.synchronize()
                dt = time.perf_counter() - t
            finally:
                b.close()
            tr = Trial(float(r), int(ce), dt, n * steps / dt)
            trials.append(tr)
            if log:
                log(f"autotune: cell ratio {r}, check every {ce}: {dt * 1e3:.2f} ms for {steps} steps, "
                    f"{tr.particle_steps_per_s / 1e6:.1f} M particle-steps/s")
    best = min(trials, key=lambda t: t.seconds)
    return {"cell_ratio": best.cell_ratio, "check_every": best.check_every}, trials


def autotune_step(ctx: Context, make, ratios=RATIOS, warmup: int = 2, steps: int = 3, log=None):
    """The single-scene form: `make(cell_ratio)` returns a fresh simulation
    (an object with `step()` and `close()`) of the same scene; times `steps`
    steps after `warmup` per ratio and returns the fastest ratio and the
    trials as (ratio, seconds) pairs."""
    trials = []
    for r in ratios:
        sim = make(r)
        try:
            for _ in range(warmup):
                sim.step()
            ctx.synchronize()
            t = time.perf_counter()
            for _ in range(steps):
                sim.step()
            ctx.synchronize()
            dt = time.perf_counter() - t
        finally:
            sim.close()
        trials.append((float(r), dt))
        if log:
            log(f"autotune: cell ratio {r}: {dt * 1e3:.2f} ms for {steps} steps")
    best = min(trials, key=lambda t: t[1])
    return best[0], trials
