"""ctypes binding of the C-ABI in ``include/reyn_amd.h``.

Loads the in-tree ``libreyn_amd.so`` (built by ``reyn_amd.build``) and
declares every entry point's signature. There is no fallback: if the library
is missing, importing a function from here raises.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libreyn_amd.so"
#: the C ABI version this binding is written for (REYN_ABI_VERSION of include/reyn_amd.h)
ABI_VERSION = 3

# status codes
OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_HIP = -2
ERR_OUT_OF_MEMORY = -3
ERR_OUT_OF_BOUNDS = -4
ERR_NOT_FINITE = -5
ERR_NOT_CONVERGED = -6
ERR_CAPACITY = -7

KERNEL_C2, KERNEL_B3, KERNEL_W6, KERNEL_COS = 0, 1, 2, 3
KERNELS = {"C2": KERNEL_C2, "B3": KERNEL_B3, "W6": KERNEL_W6, "COS": KERNEL_COS}
IISPH_STOP_ANY, IISPH_STOP_MEAN = 0, 1
CELL_ORDER_XYZ, CELL_ORDER_BRICK = 0, 1
CELL_ORDERS = {"xyz": CELL_ORDER_XYZ, "brick": CELL_ORDER_BRICK}

u32 = C.c_uint32
u64 = C.c_uint64
i32 = C.c_int32
f32 = C.c_float
f64 = C.c_double
vp = C.c_void_p
fp = C.POINTER(C.c_float)
up = C.POINTER(C.c_uint32)


class Float3(C.Structure):
    _fields_ = [("x", f32), ("y", f32), ("z", f32)]


class Grid(C.Structure):
    _fields_ = [
        ("bound_min", Float3),
        ("cell_size", f32),
        ("r_c_2", f32),
        ("nx", i32),
        ("ny", i32),
        ("nz", i32),
        ("range", i32),
        ("prefix", vp),
        ("sorted", vp),
        ("n", u32),
        ("order", i32),
    ]

    @property
    def ncells(self) -> int:
        return int(self.nx) * int(self.ny) * int(self.nz)


class Particles(C.Structure):
    _fields_ = [
        ("xx", vp), ("xy", vp), ("xz", vp),
        ("vx", vp), ("vy", vp), ("vz", vp),
        ("m", vp),
        ("n", u32),
    ]


class Boundary(C.Structure):
    _fields_ = [
        ("xx", vp), ("xy", vp), ("xz", vp), ("m", vp),
        ("n", u32),
        ("grid", Grid),
        ("near", vp),
    ]


class SphParams(C.Structure):
    _fields_ = [
        ("kernel", i32),
        ("h_bar", f32),
        ("h", f32),
        ("nu", f32),
        ("rho_0", f32),
        ("g", Float3),
    ]


class SesphParams(C.Structure):
    _fields_ = [("sph", SphParams), ("k", f32)]


class SesphWork(C.Structure):
    _fields_ = [("rho", vp), ("ax", vp), ("ay", vp), ("az", vp)]


class PcisphParams(C.Structure):
    _fields_ = [
        ("sph", SphParams),
        ("delta", f64),
        ("eta_rho_max", f32),
        ("min_iter", u32),
        ("max_iter", u32),
        ("check_every", u32),
    ]


class PcisphWork(C.Structure):
    _fields_ = [("rho", vp), ("ax", vp), ("ay", vp), ("az", vp)]


class SolveInfo(C.Structure):
    _fields_ = [
        ("iterations", u32),
        ("n_violating", u32),
        ("mean_rel_error", f64),
        ("avg_density", f64),
    ]


class IisphParams(C.Structure):
    _fields_ = [
        ("sph", SphParams),
        ("omega", f32),
        ("eta_rho_max", f32),
        ("min_iter", u32),
        ("max_iter", u32),
        ("stop_rule", i32),
        ("warm_start", i32),
        ("check_every", u32),
    ]


class IisphWork(C.Structure):
    _fields_ = [
        ("rho", vp), ("p", vp), ("a_ii", vp), ("s_i", vp),
        ("ax", vp), ("ay", vp), ("az", vp),
    ]


SOLVER_PCISPH, SOLVER_IISPH = 1, 2


class DamBreak(C.Structure):
    _fields_ = [
        ("counts", u32 * 3), ("spacing", f64), ("box", f64 * 3), ("jitter", f64), ("seed", u32), ("layers", u32),
        ("sphere", i32), ("center", f64 * 3), ("radius", f64),
    ]


class BatchSceneInfo(C.Structure):
    _fields_ = [("steps", u32), ("iterations", u32), ("dt", f32), ("time", f64), ("loop_stat", f64)]

class BatchScene(C.Structure):
    _fields_ = [
        ("n", u32), ("x", vp), ("y", vp), ("z", vp), ("vx", vp), ("vy", vp), ("vz", vp), ("mass", f32),
        ("n_bdy", u32), ("bx", vp), ("by", vp), ("bz", vp), ("bound_min", Float3), ("bound_max", Float3),
        ("h", f32), ("p", vp), ("resume", C.POINTER(BatchSceneInfo)), ("device", i32),
    ]


class BatchParams(C.Structure):
    _fields_ = [
        ("solver", i32), ("kernel", i32), ("nu", f32), ("rho_0", f32), ("g", Float3), ("cell_ratio", f32),
        ("dt0", f32), ("cfl_lambda", f32), ("change_factor", f32), ("iter_target", u32),
        ("eta_rho_max", f32), ("omega", f32), ("min_iter", u32), ("max_iter", u32), ("stop_rule", i32),
        ("warm_start", i32), ("graph", i32), ("check_every", u32), ("cell_order", i32),
    ]



P = C.POINTER
_SIGS = {
    "reyn_abi_version": (i32, []),
    "reyn_grid_cell_capacity": (i32, [i32, P(u32)]),
    "reyn_device_count": (i32, [P(C.c_int)]),
    "reyn_ctx_create": (i32, [C.c_int, P(vp)]),
    "reyn_ctx_destroy": (i32, [vp]),
    "reyn_last_error": (C.c_char_p, [vp]),
    "reyn_ctx_synchronize": (i32, [vp]),
    "reyn_ctx_stream": (vp, [vp]),
    "reyn_malloc": (i32, [vp, P(vp), C.c_size_t]),
    "reyn_free": (i32, [vp, vp]),
    "reyn_memcpy": (i32, [vp, vp, vp, C.c_size_t, C.c_int]),
    "reyn_fill_u32": (i32, [vp, vp, u32, C.c_size_t]),
    "reyn_fill_f32": (i32, [vp, vp, f32, C.c_size_t]),
    "reyn_timer_start": (i32, [vp]),
    "reyn_timer_stop": (i32, [vp, P(f32)]),
    "reyn_profile_enable": (i32, [vp, C.c_int]),
    "reyn_profile_select": (i32, [vp, C.c_char_p]),
    "reyn_profile_summary": (i32, [vp, C.c_char_p, C.c_size_t, P(C.c_size_t)]),
    "reyn_grid_builder_create": (i32, [vp, Float3, Float3, f32, f32, P(vp)]),
    "reyn_grid_builder_create_ordered": (i32, [vp, Float3, Float3, f32, f32, i32, P(vp)]),
    "reyn_grid_builder_destroy": (i32, [vp]),
    "reyn_grid_builder_geometry": (i32, [vp, P(Grid)]),
    "reyn_grid_builder_defer_check": (i32, [vp, i32]),
    "reyn_grid_builder_check": (i32, [vp, vp]),
    "reyn_grid_construct": (i32, [vp, vp, f32, vp, vp, vp, u32, P(Grid)]),
    "reyn_grid_construct_and_reorder": (i32, [vp, vp, f32, P(vp), P(vp), u32, u32, P(Grid)]),
    "reyn_grid_debug_tables": (i32, [vp, P(vp), P(vp)]),
    "reyn_gather_f32": (i32, [vp, vp, u32, P(vp), P(vp), u32]),
    "reyn_grid_query_stats": (i32, [vp, P(Grid), vp, vp, vp, u32, vp, vp, vp, vp, vp]),
    "reyn_grid_query_bench": (i32, [vp, P(Grid), i32, f32, vp, vp, vp, u32, vp]),
    "reyn_kernel_eval": (i32, [vp, i32, f32, vp, vp, vp, u32, vp, vp, vp, vp]),
    "reyn_density": (i32, [vp, P(SphParams), P(Particles), P(Grid), P(Boundary), vp]),
    "reyn_nonpressure_acc": (i32, [vp, P(SphParams), P(Particles), P(Grid), vp, vp, vp, vp]),
    "reyn_predict_rho": (i32, [vp, P(SphParams), P(Particles), P(Grid), P(Boundary), f32, vp]),
    "reyn_integrate": (i32, [vp, P(Particles), vp, vp, vp, f32]),
    "reyn_integrate_a_v_only": (i32, [vp, P(Particles), vp, vp, vp, f32]),
    "reyn_sesph_step": (i32, [vp, P(SesphParams), P(Particles), P(Grid), P(Boundary), P(SesphWork), f32]),
    "reyn_sesph_step_noalloc": (i32, [vp, P(SesphParams), P(Particles), P(Grid), P(Boundary), P(SesphWork), f32]),
    "reyn_pcisph_delta": (i32, [i32, f32, f32, P(f64)]),
    "reyn_pcisph_step": (i32, [vp, P(PcisphParams), P(Particles), P(Grid), P(Boundary), P(PcisphWork), f32, P(SolveInfo)]),
    "reyn_iisph_step": (i32, [vp, P(IisphParams), P(Particles), P(Grid), P(Boundary), P(IisphWork), f32, P(SolveInfo)]),
    "reyn_boundary_masses": (i32, [vp, P(Boundary), vp, f32, f32, u32]),
    "reyn_cfl_time_step": (i32, [vp, f32, f32, P(Particles), Float3, P(f32), P(f64)]),
    "reyn_simple_dt_controller": (i32, [vp, u32, f32, f32, f32, P(Particles), Float3, u32, f32, P(f32)]),
    "reyn_hard_enforce_bounds": (i32, [vp, Float3, Float3, P(Particles)]),
    "reyn_reduce_f32": (i32, [vp, vp, u32, i32, P(f64)]),
    "reyn_batch_create": (i32, [vp, P(BatchParams), P(BatchScene), u32, P(vp)]),
    "reyn_batch_destroy": (i32, [vp]),
    "reyn_batch_add": (i32, [vp, P(BatchScene), u32, P(u32)]),
    "reyn_batch_add_dam_breaks": (i32, [vp, P(DamBreak), u32, P(u32)]),
    "reyn_dam_break_sizes": (i32, [P(DamBreak), P(u32), P(u32)]),
    "reyn_dam_break_generate": (i32, [vp, P(DamBreak), vp, vp, vp, P(u32), vp, vp, vp]),
    "reyn_batch_remove": (i32, [vp, P(u32), u32]),
    "reyn_batch_scenes": (i32, [vp, P(u32), u32, P(u32)]),
    "reyn_batch_scene_size": (i32, [vp, u32, P(u32)]),
    "reyn_batch_scene_bytes": (i32, [P(BatchParams), P(BatchScene), P(u64)]),
    "reyn_batch_step": (i32, [vp, u32]),
    "reyn_mesh_load_obj": (i32, [C.c_char_p, P(C.c_char_p), u32, C.c_char_p, P(vp), C.c_char_p, C.c_size_t]),
    "reyn_mesh_create": (i32, [vp, vp, vp, u32, vp, u32, P(vp)]),
    "reyn_mesh_destroy": (i32, [vp]),
    "reyn_mesh_info": (i32, [vp, P(u32), P(u32), P(f64), P(f64)]),
    "reyn_mesh_data": (i32, [vp, vp, vp, vp, vp]),
    "reyn_mesh_sample_count": (i32, [vp, f32, f32, P(u32)]),
    "reyn_mesh_sample": (i32, [vp, vp, f32, f32, u32, vp, vp, vp, u32, P(u32)]),
    "reyn_mesh_inside": (i32, [vp, vp, vp, vp, vp, u32, vp]),
    "reyn_mesh_relax": (i32, [vp, vp, f32, f32, vp, vp, vp, u32, u32, f32, P(f32)]),
    "reyn_mesh_project": (i32, [vp, vp, vp, vp, vp, u32]),
    "reyn_lattice_box": (i32, [vp, Float3, u32, u32, u32, f32, vp, vp, vp]),
    "reyn_jitter_normal": (i32, [vp, vp, vp, vp, u32, f32, u32]),
    "reyn_boundary_cull": (i32, [vp, vp, f32, vp, vp, vp, u32, vp]),
    "reyn_compact_f32": (i32, [vp, vp, u32, vp, u32, P(u32)]),
    "reyn_lbvh_create": (i32, [vp, vp, P(vp)]),
    "reyn_lbvh_destroy": (i32, [vp]),
    "reyn_lbvh_info": (i32, [vp, P(u32), P(u32)]),
    "reyn_lbvh_nodes": (i32, [vp, vp, vp, vp, vp, vp, vp]),
    "reyn_lbvh_project": (i32, [vp, vp, vp, vp, vp, u32]),
    "reyn_lbvh_inside": (i32, [vp, vp, vp, vp, vp, u32, vp]),
    "reyn_grid_occupancy": (i32, [vp, P(Grid), vp]),
    "reyn_device_memory": (i32, [vp, P(u64), P(u64)]),
    "reyn_batch_info": (i32, [vp, u32, P(BatchSceneInfo)]),
    "reyn_batch_download": (i32, [vp, u32, vp, vp, vp, vp, vp, vp, vp, vp]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def lib() -> C.CDLL:
    """The loaded library with all signatures declared (raises if missing)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m reyn_amd.build` "
                "(there is no CPU fallback)"
            )
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        if handle.reyn_abi_version() != ABI_VERSION:
            raise RuntimeError("reyn_amd ABI version mismatch")
        _lib = handle
    return _lib


class ReynError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"reyn error {code}: {message}")
        self.code = code


def f3(v) -> Float3:
    return Float3(float(v[0]), float(v[1]), float(v[2]))
