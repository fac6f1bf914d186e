"""Autotuner for scene batches and single scenes: picks the grid cell-size
ratio, the cell order (the reference's XYZ keys or 2 x 2 x 2 bricks) and, for
batches, the pressure loop's host-check interval, by timing short runs of
fresh batches or simulations.

Every trial starts from the same initial state and runs the same steps, so
the trials compare equal work: the check interval does not change results at
all (the loop control runs on the device and each scene stops at its own
iteration), and the cell ratio and order change only the order in which
neighbour terms are summed (the neighbour sets are the same).

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

#: cell ratios (cell size / search radius), cell orders and host-check
#: intervals tried (brick order keeps neighbour lists at ratio 1 only, so it
#: is tried there)
RATIOS = (0.5, 1.0)
ORDERS = ("xyz", "brick")
CHECKS = (1, 2, 4)


def ratio_orders(ratios=RATIOS, orders=ORDERS):
    return [(r, o) for r in ratios for o in orders if o == "xyz" or r >= 1.0]


@dataclass
class Trial:
    cell_ratio: float
    check_every: int
    seconds: float
    particle_steps_per_s: float
    cell_order: str = "xyz"

    def as_dict(self) -> dict:
        return {"cell_ratio": self.cell_ratio, "cell_order": self.cell_order, "check_every": self.check_every,
                "seconds": round(self.seconds, 5), "particle_steps_per_s": round(self.particle_steps_per_s, 1)}


def autotune_batch(ctx: Context, scenes, ratios=RATIOS, checks=CHECKS, orders=ORDERS, warmup: int = 2,
                   steps: int = 4, log=None, **options) -> tuple[dict, list[Trial]]:
    """Times `steps` steps (after `warmup`) of a fresh batch of `scenes` for
    every (cell ratio, cell order, check interval); returns the fastest as
    batch options and every trial."""
    n = sum(sc.n for sc in scenes)
    trials = []
    for r, order in ratio_orders(ratios, orders):
        for ce in checks:
            b = Batch(ctx, scenes, cell_ratio=r, check_every=ce, cell_order=order, **options)
            try:
                b.step(warmup)
                ctx.synchronize()
                t = time.perf_counter()
                b.step(steps)
                ctx.synchronize()
                dt = time.perf_counter() - t
            finally:
                b.close()
            tr = Trial(float(r), int(ce), dt, n * steps / dt, order)
            trials.append(tr)
            if log:
                log(f"autotune: cell ratio {r}, {order} order, check every {ce}: {dt * 1e3:.2f} ms for {steps} "
                    f"steps, {tr.particle_steps_per_s / 1e6:.1f} M particle-steps/s")
    best = min(trials, key=lambda t: t.seconds)
    return {"cell_ratio": best.cell_ratio, "check_every": best.check_every, "cell_order": best.cell_order}, trials


def autotune_step(ctx: Context, make, ratios=RATIOS, orders=ORDERS, warmup: int = 2, steps: int = 3, log=None):
    """The single-scene form: `make(cell_ratio, cell_order)` returns a fresh
    simulation (an object with `step()` and `close()`) of the same scene;
    times `steps` steps after `warmup` per (ratio, order) and returns the
    fastest (ratio, order) and the trials as (ratio, order, seconds)."""
    trials = []
    for r, order in ratio_orders(ratios, orders):
        sim = make(r, order)
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
        trials.append((float(r), order, dt))
        if log:
            log(f"autotune: cell ratio {r}, {order} order: {dt * 1e3:.2f} ms for {steps} steps")
    best = min(trials, key=lambda t: t[2])
    return (best[0], best[1]), trials
