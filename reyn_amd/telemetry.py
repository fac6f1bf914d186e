"""Step telemetry without a GUI: the time step and pressure-iteration traces
the reference plots in its GUI (`plot_dts`, `plot_iters`, src/gui.cu:956-957,
1163-1178), recorded per step and written as CSV or JSON.

`Simulation(..., trace=StepTrace())` records every step; for a batch,
`record_batch_step(batch, traces)` records every resident scene after a
step.
"""

from __future__ import annotations

import csv
import json
from dataclasses import dataclass, field
from pathlib import Path


@dataclass
class StepTrace:
    """Per step: index, simulated time after the step, time step (s), pressure
    iterations and the loop statistic (PCISPH mean predicted density, IISPH
    mean compression error; 0 when the solver has none)."""

    steps: list = field(default_factory=list)
    time: list = field(default_factory=list)
    dt: list = field(default_factory=list)
    iterations: list = field(default_factory=list)
    loop_stat: list = field(default_factory=list)

    def record(self, step: int, time: float, dt: float, iterations: int, loop_stat: float = 0.0) -> None:
        self.steps.append(int(step))
        self.time.append(float(time))
        self.dt.append(float(dt))
        self.iterations.append(int(iterations))
        self.loop_stat.append(float(loop_stat))

    def __len__(self) -> int:
        return len(self.steps)

    @property
    def dt_ms(self) -> list:
        """The reference's plotted quantity: dt in milliseconds."""
        return [1000.0 * v for v in self.dt]

    def summary(self) -> dict:
        n = len(self)
        if n == 0:
            return {"steps": 0}
        return {"steps": n, "time": self.time[-1], "dt_min": min(self.dt), "dt_max": max(self.dt),
                "dt_mean": sum(self.dt) / n, "iterations_mean": sum(self.iterations) / n,
                "iterations_max": max(self.iterations), "iterations_total": sum(self.iterations)}

    def write_csv(self, path) -> None:
        with open(path, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["step", "time", "dt", "iterations", "loop_stat"])
            for row in zip(self.steps, self.time, self.dt, self.iterations, self.loop_stat):
                w.writerow([row[0], repr(row[1]), repr(row[2]), row[3], repr(row[4])])

    @staticmethod
    def read_csv(path) -> "StepTrace":
        t = StepTrace()
        with open(path, newline="") as f:
            r = csv.reader(f)
            next(r)
            for step, time, dt, it, st in r:
                t.record(int(step), float(time), float(dt), int(it), float(st))
        return t

    def to_json(self) -> str:
        return json.dumps({"step": self.steps, "time": self.time, "dt": self.dt, "iterations": self.iterations,
                           "loop_stat": self.loop_stat})

    def write_json(self, path) -> None:
        Path(path).write_text(self.to_json())


def record_batch_step(batch, traces: dict) -> None:
    """Appends the last step of every resident scene of `batch` to
    traces[scene id] (a new StepTrace for a scene not seen before)."""
    for sid in batch.ids():
        inf = batch.info(sid)
        traces.setdefault(sid, StepTrace()).record(inf.steps, inf.time, inf.dt, inf.iterations, inf.loop_stat)


__all__ = ["StepTrace", "record_batch_step"]
