"""ctypes binding of the oracle library (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle_reyn.so`` (built from ``reyn_oracle.cpp`` by
``make -C oracle``). Only the tests, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` use this module, as the checker or the timed
baseline; the product path never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# REYN_ORACLE_LIB names another build of the same source (oracle/mutations.py
# loads deliberately broken builds this way to check that the pins fail)
LIB_PATH = Path(os.environ.get("REYN_ORACLE_LIB", HERE / "liboracle_reyn.so"))

u32, i32, f32, f64, vp = C.c_uint32, C.c_int32, C.c_float, C.c_double, C.c_void_p


class OrcGrid(C.Structure):
    _fields_ = [
        ("bmin", f32 * 3),
        ("cell", f32),
        ("r2", f32),
        ("n", i32 * 3),
        ("range", i32),
        ("prefix", vp),
        ("sorted", vp),
        ("order", i32),
    ]

    @property
    def ncells(self) -> int:
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])


class OrcSph(C.Structure):
    _fields_ = [("kernel", i32), ("h_bar", f32), ("h", f32), ("nu", f32), ("rho0", f32), ("g", f32 * 3)]


class OrcIisph(C.Structure):
    _fields_ = [("omega", f32), ("eta", f32), ("min_iter", u32), ("max_iter", u32), ("stop_rule", i32),
                ("warm_start", i32)]


def build() -> Path:
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _lib.orc_grid_build.restype = C.c_long
        _lib.orc_pcisph_step_f64.restype = u32
        _lib.orc_pcisph_step_f32.restype = u32
        _lib.orc_iisph_step_f64.restype = u32
        _lib.orc_iisph_step_f32.restype = u32
        _lib.orc_cfl_time_step.restype = f32
        _lib.orc_shape_w.restype = f64
        _lib.orc_shape_w.argtypes = [C.c_int, f64]
        _lib.orc_shape_dw.restype = f64
        _lib.orc_shape_dw.argtypes = [C.c_int, f64]
        _lib.orc_pcisph_delta_f64.restype = f64
        _lib.orc_pcisph_delta_f64.argtypes = [C.c_int, f32, f32]
        _lib.orc_num_threads.restype = C.c_int
        _lib.orc_grid_cell_capacity.restype = u32
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(vp)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


class Grid:
    """Geometry plus host tables; keeps the arrays alive for the struct."""

    def __init__(self, geom: OrcGrid, prefix=None, sorted_=None, counts=None, keys=None, outside=0):
        self.geom = geom
        self.prefix = prefix
        self.sorted = sorted_
        self.counts = counts
        self.keys = keys
        self.outside = outside
        geom.prefix = _p(prefix) if prefix is not None else None
        geom.sorted = _p(sorted_) if sorted_ is not None else None

    def resorted(self) -> "Grid":
        """The same tables viewed as Resort::yes (arrays already in cell order)."""
        g = OrcGrid()
        C.pointer(g)[0] = self.geom
        return Grid(g, self.prefix, None, self.counts, self.keys, self.outside)


def geometry(bmin, bmax, r: float, ratio: float = 0.5, order: int = 0) -> OrcGrid:
    """Grid geometry; order 0 is the reference's XYZ cell order, 1 the
    2 x 2 x 2 brick order."""
    g = OrcGrid()
    lo = _c(bmin, np.float32)
    hi = _c(bmax, np.float32)
    rc = lib().orc_grid_geometry_ordered(_p(lo), _p(hi), f32(r), f32(ratio), C.c_int(int(order)), C.byref(g))
    if rc != 0:
        raise ValueError("invalid grid geometry")
    return g


def grid_build(geom: OrcGrid, x, y, z) -> Grid:
    x, y, z = (_c(a, np.float32) for a in (x, y, z))
    n = x.size
    C_ = geom.ncells
    counts = np.empty(C_, np.uint32)
    prefix = np.empty(C_ + 1, np.uint32)
    sorted_ = np.empty(n, np.uint32)
    keys = np.empty(n, np.uint32)
    outside = lib().orc_grid_build(C.byref(geom), _p(x), _p(y), _p(z), u32(n), _p(counts), _p(prefix),
                                   _p(sorted_), _p(keys))
    if outside < 0:
        raise ValueError(f"a grid cell holds more than {grid_cell_capacity(geom.range)} points")
    g = OrcGrid()
    C.pointer(g)[0] = geom
    return Grid(g, prefix, sorted_, counts, keys, int(outside))


def grid_cell_capacity(range_: int) -> int:
    """The most points per cell a grid build accepts at this search range."""
    return int(lib().orc_grid_cell_capacity(int(range_)))


def query_stats(grid: Grid, x, y, z):
    x, y, z = (_c(a, np.float32) for a in (x, y, z))
    n = x.size
    cnt = np.empty(n, np.uint32)
    l2, vx, vy, vz = (np.empty(n, np.float32) for _ in range(4))
    lib().orc_query_stats(C.byref(grid.geom), _p(x), _p(y), _p(z), u32(n), _p(cnt), _p(l2), _p(vx), _p(vy), _p(vz))
    return cnt, l2, np.stack([vx, vy, vz], axis=1)


def brute_stats(r2: float, x, y, z):
    x, y, z = (_c(a, np.float32) for a in (x, y, z))
    n = x.size
    cnt = np.empty(n, np.uint32)
    l2, vx, vy, vz = (np.empty(n, np.float32) for _ in range(4))
    lib().orc_brute_stats(f32(r2), _p(x), _p(y), _p(z), u32(n), _p(cnt), _p(l2), _p(vx), _p(vy), _p(vz))
    return cnt, l2, np.stack([vx, vy, vz], axis=1)


def brute_stats_subset(r2: float, x, y, z, idx):
    x, y, z = (_c(a, np.float32) for a in (x, y, z))
    idx = _c(idx, np.uint32)
    nq = idx.size
    cnt = np.empty(nq, np.uint32)
    l2, vx, vy, vz = (np.empty(nq, np.float32) for _ in range(4))
    lib().orc_brute_stats_subset(f32(r2), _p(x), _p(y), _p(z), u32(x.size), _p(idx), u32(nq), _p(cnt), _p(l2),
                                 _p(vx), _p(vy), _p(vz))
    return cnt, l2, np.stack([vx, vy, vz], axis=1)


def reference_grid_points(side: int = 50, h: float = 1.1):
    """The reference grid test's points (uniformgrid.cu:257-275)."""
    n = side ** 3
    x, y, z = (np.empty(n, np.float32) for _ in range(3))
    lib().orc_reference_grid_points(u32(side), f32(h), _p(x), _p(y), _p(z))
    return np.stack([x, y, z], axis=1)


def reference_kernel_points(h_bar: float = 25.0, n: int = 25 ** 3):
    """The reference kernel test's offsets (kernels.cu:31-52)."""
    x, y, z = (np.empty(n, np.float32) for _ in range(3))
    lib().orc_reference_kernel_points(f32(h_bar), u32(n), _p(x), _p(y), _p(z))
    return np.stack([x, y, z], axis=1)


def neighbour_lists(grid: Grid, x, y, z):
    x, y, z = (_c(a, np.float32) for a in (x, y, z))
    n = x.size
    off = np.zeros(n + 1, np.uint32)
    lib().orc_neighbour_lists(C.byref(grid.geom), _p(x), _p(y), _p(z), u32(n), _p(off), None)
    lst = np.empty(int(off[-1]), np.uint32)
    lib().orc_neighbour_lists(C.byref(grid.geom), _p(x), _p(y), _p(z), u32(n), _p(off), _p(lst))
    return off, lst


def kernel_eval(kind: int, h_bar: float, d, dtype=np.float64):
    d = np.asarray(d)
    dx, dy, dz = (_c(d[:, a], dtype) for a in range(3))
    n = dx.size
    w, gx, gy, gz = (np.empty(n, dtype) for _ in range(4))
    fn = lib().orc_kernel_eval_f64 if dtype == np.float64 else lib().orc_kernel_eval_f32
    fn(C.c_int(kind), f32(h_bar), _p(dx), _p(dy), _p(dz), u32(n), _p(w), _p(gx), _p(gy), _p(gz))
    return w, np.stack([gx, gy, gz], axis=1)


def shape_w(kind: int, q: float) -> float:
    return float(lib().orc_shape_w(kind, q))


def shape_dw(kind: int, q: float) -> float:
    return float(lib().orc_shape_dw(kind, q))


def pcisph_delta_f64(kind: int, h_bar: float, h: float) -> float:
    return float(lib().orc_pcisph_delta_f64(kind, h_bar, h))


def sph(kernel: int, h_bar: float, h: float, nu: float, rho0: float, g=(0.0, -9.81, 0.0)) -> OrcSph:
    return OrcSph(int(kernel), float(h_bar), float(h), float(nu), float(rho0), (f32 * 3)(*g))


class State:
    """Fluid state in the oracle's arithmetic type (float64 or float32)."""

    FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m")

    def __init__(self, fields: dict, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        for f in self.FIELDS:
            setattr(self, f, _c(fields[f], dtype).copy())
        self.n = self.x.size

    def ptrs(self):
        return [_p(getattr(self, f)) for f in self.FIELDS]


class Bdy:
    def __init__(self, grid: Grid | None, x, y, z, m, dtype=np.float64):
        self.grid = grid
        self.x, self.y, self.z, self.m = (_c(a, dtype).copy() for a in (x, y, z, m))
        self.n = self.x.size

    def args(self):
        if self.grid is None or self.n == 0:
            return [None, None, None, None, None, u32(0)]
        return [C.byref(self.grid.geom), _p(self.x), _p(self.y), _p(self.z), _p(self.m), u32(self.n)]


NO_BDY = [None, None, None, None, None, u32(0)]


def _suffix(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def density(grid: Grid, st: State, bdy: Bdy | None, kernel: int, h_bar: float):
    rho = np.empty(st.n, st.dtype)
    fn = getattr(lib(), "orc_density_" + _suffix(st.dtype))
    x, y, z, _, _, _, m = st.ptrs()
    b = bdy.args() if bdy else NO_BDY
    fn(C.byref(grid.geom), x, y, z, m, u32(st.n), *b, C.c_int(kernel), f32(h_bar), _p(rho))
    return rho


def nonpressure(grid: Grid, st: State, prm: OrcSph, rho):
    rho = _c(rho, st.dtype)
    a = [np.empty(st.n, st.dtype) for _ in range(3)]
    fn = getattr(lib(), "orc_nonpressure_" + _suffix(st.dtype))
    fn(C.byref(grid.geom), *st.ptrs(), u32(st.n), C.byref(prm), _p(rho), *(_p(v) for v in a))
    return np.stack(a, axis=1)


def sesph_step(grid: Grid, st: State, bdy: Bdy | None, prm: OrcSph, k: float, dt: float):
    rho = np.empty(st.n, st.dtype)
    a = [np.empty(st.n, st.dtype) for _ in range(3)]
    fn = getattr(lib(), "orc_sesph_step_" + _suffix(st.dtype))
    b = bdy.args() if bdy else NO_BDY
    fn(C.byref(grid.geom), *st.ptrs(), u32(st.n), *b, C.byref(prm), f32(k), f32(dt), _p(rho), *(_p(v) for v in a))
    return rho, np.stack(a, axis=1)


def pcisph_step(grid: Grid, st: State, bdy: Bdy | None, prm: OrcSph, delta: float, dt: float, eta=0.001,
                min_iter=3, max_iter=0, fixed_iters=0):
    rho = np.empty(st.n, st.dtype)
    a = [np.empty(st.n, st.dtype) for _ in range(3)]
    avg = C.c_double()
    fn = getattr(lib(), "orc_pcisph_step_" + _suffix(st.dtype))
    b = bdy.args() if bdy else NO_BDY
    it = fn(C.byref(grid.geom), *st.ptrs(), u32(st.n), *b, C.byref(prm), f64(delta), f32(eta), u32(min_iter),
            u32(max_iter), u32(fixed_iters), f32(dt), _p(rho), *(_p(v) for v in a), C.byref(avg))
    return int(it), rho, np.stack(a, axis=1), float(avg.value)


def iisph_step(grid: Grid, st: State, bdy: Bdy | None, prm: OrcSph, ip: OrcIisph, p, dt: float, fixed_iters=0):
    p = _c(p, st.dtype).copy()
    rho, aii, si = (np.empty(st.n, st.dtype) for _ in range(3))
    a = [np.empty(st.n, st.dtype) for _ in range(3)]
    mean = C.c_double()
    viol = u32()
    fn = getattr(lib(), "orc_iisph_step_" + _suffix(st.dtype))
    b = bdy.args() if bdy else NO_BDY
    it = fn(C.byref(grid.geom), *st.ptrs(), u32(st.n), *b, C.byref(prm), C.byref(ip), u32(fixed_iters), f32(dt),
            _p(rho), _p(p), _p(aii), _p(si), *(_p(v) for v in a), C.byref(mean), C.byref(viol))
    return {"iterations": int(it), "rho": rho, "p": p, "a_ii": aii, "s_i": si, "a": np.stack(a, axis=1),
            "mean_rel_error": float(mean.value), "n_violating": int(viol.value)}


def boundary_masses(grid: Grid, x, y, z, h: float, rho0: float, iters: int = 5, dtype=np.float64):
    x, y, z = (_c(a, dtype) for a in (x, y, z))
    m = np.empty(x.size, dtype)
    fn = getattr(lib(), "orc_boundary_masses_" + _suffix(dtype))
    fn(C.byref(grid.geom), _p(x), _p(y), _p(z), u32(x.size), _p(m), f32(h), f32(rho0), u32(iters))
    return m


def cfl_time_step(lam: float, h: float, v, g=(0.0, -9.81, 0.0)):
    v = np.asarray(v, np.float32)
    vx, vy, vz = (_c(v[:, a], np.float32) for a in range(3))
    gg = _c(g, np.float32)
    vmax = C.c_double()
    dt = lib().orc_cfl_time_step(f32(lam), f32(h), _p(vx), _p(vy), _p(vz), u32(vx.size), _p(gg), C.byref(vmax))
    return float(dt), float(vmax.value)


def hard_enforce_bounds(lo, hi, x, v):
    x = np.array(x, np.float32)
    v = np.array(v, np.float32)
    cols = [np.ascontiguousarray(x[:, a]) for a in range(3)] + [np.ascontiguousarray(v[:, a]) for a in range(3)]
    lo = _c(lo, np.float32)
    hi = _c(hi, np.float32)
    lib().orc_hard_enforce_bounds(_p(lo), _p(hi), *(_p(c) for c in cols), u32(x.shape[0]))
    return np.stack(cols[:3], axis=1), np.stack(cols[3:], axis=1)


def gather(sorted_, src):
    src = _c(src, np.float32)
    dst = np.empty_like(src)
    s = _c(sorted_, np.uint32)
    lib().orc_gather_f32(_p(s), u32(s.size), _p(src), _p(dst))
    return dst


def grid_build_par(geom: OrcGrid, x, y, z):
    """The CPU baseline's parallel counting sort: (prefix, sorted)."""
    x, y, z = (_c(a, np.float32) for a in (x, y, z))
    prefix = np.empty(geom.ncells + 1, np.uint32)
    sorted_ = np.empty(x.size, np.uint32)
    lib().orc_grid_build_par(C.byref(geom), _p(x), _p(y), _p(z), u32(x.size), _p(prefix), _p(sorted_))
    return prefix, sorted_


class PortState:
    """Float32 SoA state of the CPU baseline loops, updated in place."""

    FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m")

    def __init__(self, x, v, m, p=None):
        x = np.asarray(x, np.float32)
        v = np.zeros_like(x) if v is None else np.asarray(v, np.float32)
        self.x, self.y, self.z = (np.ascontiguousarray(x[:, a]) for a in range(3))
        self.vx, self.vy, self.vz = (np.ascontiguousarray(v[:, a]) for a in range(3))
        self.m = np.ascontiguousarray(np.broadcast_to(np.asarray(m, np.float32), (x.shape[0],)))
        self.p = np.zeros(x.shape[0], np.float32) if p is None else _c(p, np.float32).copy()
        self.n = x.shape[0]

    def ptrs(self):
        return [_p(getattr(self, f)) for f in self.FIELDS]


def port_iisph_run(geom: OrcGrid, st: PortState, bdy: Bdy | None, prm: OrcSph, ip: OrcIisph, lo, hi, lam: float,
                   iter_target: int, change_factor: float, dt: float, iters: int, steps: int):
    """`steps` steps of the float IISPH main loop (controller, parallel grid and
    reorder, step, bounds); returns (per-step seconds, per-step iterations,
    dt, iterations) and updates `st`."""
    lo, hi = _c(lo, np.float32), _c(hi, np.float32)
    secs = np.zeros(steps)
    its = np.zeros(steps, np.uint32)
    dt_c, it_c = f32(dt), u32(iters)
    b = bdy.args() if bdy else NO_BDY
    lib().orc_port_iisph_run(C.byref(geom), *st.ptrs(), _p(st.p), u32(st.n), *b, C.byref(prm), C.byref(ip), _p(lo),
                             _p(hi), f32(lam), u32(iter_target), f32(change_factor), C.byref(dt_c), C.byref(it_c),
                             u32(steps), _p(secs), _p(its))
    return secs, its, float(dt_c.value), int(it_c.value)


def port_sesph_run(geom: OrcGrid, st: PortState, bdy: Bdy | None, prm: OrcSph, k: float, lo, hi, lam: float,
                   steps: int):
    """`steps` steps of the float SESPH main loop with dt = cfl_time_step(lam);
    returns (per-step seconds, simulated time) and updates `st`."""
    lo, hi = _c(lo, np.float32), _c(hi, np.float32)
    secs = np.zeros(steps)
    t = C.c_double()
    b = bdy.args() if bdy else NO_BDY
    lib().orc_port_sesph_run(C.byref(geom), *st.ptrs(), u32(st.n), *b, C.byref(prm), f32(k), _p(lo), _p(hi),
                             f32(lam), u32(steps), _p(secs), C.byref(t))
    return secs, float(t.value)
