"""Result and checkpoint formats.

- PLY: particle positions as a binary PLY file in the host's byte order
  (`save_to_ply`, src/scene/ply_io.cu:7-56: a header naming one vertex element
  of float32 x, y, z, then x_i, y_i, z_i per particle).
- Checkpoint: a scene's full state (positions, velocities, pressure in the
  order the batch holds them), its description (boundary samples, spacing,
  rest density, box) and its step record (steps, last time step and
  iteration count, simulated time). Adding a checkpoint to any batch continues
  the scene exactly as if it had stayed in its old batch. Stored with
  `numpy.savez` and read back without pickle. The reference has no
  checkpoint.
"""

from __future__ import annotations

import sys
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .scenes import Scene

STATE_FIELDS = ("xx", "xy", "xz", "vx", "vy", "vz", "p")


def write_ply(path, x, y, z) -> None:
    """Binary PLY of the positions (x, y, z arrays of one length)."""
    x, y, z = (np.asarray(a, dtype=np.float32).ravel() for a in (x, y, z))
    if not (x.size == y.size == z.size):
        raise ValueError("write_ply: x, y and z differ in length")
    order = "little" if sys.byteorder == "little" else "big"
    header = ("ply\n"
              f"format binary_{order}_endian 1.0\n"
              "comment particle positions written by reyn_amd\n"
              f"element vertex {x.size}\n"
              "property float32 x\nproperty float32 y\nproperty float32 z\n"
              "end_header\n")
    data = np.stack([x, y, z], axis=1)  # x_i, y_i, z_i per particle
    with open(path, "wb") as f:
        f.write(header.encode("ascii"))
        f.write(np.ascontiguousarray(data).tobytes())


def read_ply(path) -> np.ndarray:
    """The (n, 3) float32 positions of a binary PLY file of one vertex element
    with float32 x, y, z properties."""
    raw = Path(path).read_bytes()
    end = raw.index(b"end_header\n") + len(b"end_header\n")
    lines = raw[:end].decode("ascii").splitlines()
    if lines[0] != "ply":
        raise ValueError("read_ply: not a PLY file")
    fmt = lines[1].split()[1]
    if fmt not in ("binary_little_endian", "binary_big_endian"):
        raise ValueError(f"read_ply: unsupported format {fmt}")
    n = next(int(ln.split()[2]) for ln in lines if ln.startswith("element vertex"))
    props = [ln.split()[2] for ln in lines if ln.startswith("property")]
    if props != ["x", "y", "z"]:
        raise ValueError(f"read_ply: expected float32 x, y, z, got {props}")
    dt = np.dtype("<f4" if fmt == "binary_little_endian" else ">f4")
    return np.frombuffer(raw, dtype=dt, count=3 * n, offset=end).reshape(n, 3).astype(np.float32)


@dataclass
class Checkpoint:
    """A scene's state and step record, taken from a batch."""

    scene: Scene
    state: dict  # STATE_FIELDS -> float32 arrays of n, in the batch's order
    steps: int
    iterations: int
    dt: float
    time: float
    loop_stat: float

    @property
    def n(self) -> int:
        return int(self.state["xx"].size)

    def save(self, path) -> None:
        sc = self.scene
        np.savez(path, x0=sc.x, xb=sc.xb, spacing=sc.spacing, rho_0=sc.rho_0, box_lo=sc.box_lo, box_hi=sc.box_hi,
                 steps=self.steps, iterations=self.iterations, dt=np.float32(self.dt), time=self.time,
                 loop_stat=self.loop_stat, **{f"s_{k}": self.state[k] for k in STATE_FIELDS})

    @staticmethod
    def load(path) -> "Checkpoint":
        with np.load(path, allow_pickle=False) as z:
            sc = Scene(z["x0"], z["xb"], float(z["spacing"]), float(z["rho_0"]), z["box_lo"], z["box_hi"])
            state = {k: z[f"s_{k}"].astype(np.float32) for k in STATE_FIELDS}
            return Checkpoint(sc, state, int(z["steps"]), int(z["iterations"]), float(np.float32(z["dt"])),
                              float(z["time"]), float(z["loop_stat"]))
