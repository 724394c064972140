"""ctypes binding of ``libapple_amd.so`` (declared in ``include/apple_amd.h``).

The library is built in-tree by ``apple_amd.build`` (or ``__graft_entry__.build``)
into ``apple_amd/_lib/libapple_amd.so``.  There is no fallback: if the
library is missing every entry point raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libapple_amd.so")

APL_F32, APL_F64 = 0, 1
APL_SNH, APL_ARAP, APL_SNH_MUSCLE, APL_MODEL = 0, 1, 2, 3
APL_ENERGY, APL_GRAD, APL_HESS_DIAG, APL_HESS_PROD, APL_HESS_QUAD, APL_MIXED = 1, 2, 4, 8, 16, 32
APL_ACCUMULATE = 1 << 16

# layout arrays (apl_mesh_layout): name -> (id, numpy dtype, entries as a function of the mesh info)
LAYOUT_ARRAYS = {
    "corner": (0, "uint16"),
    "blk_tet_off": (1, "int32"),
    "blk_loc_off": (2, "int32"),
    "blk_inst": (3, "int32"),
    "inst_blk_off": (4, "int64"),
    "loc_gid": (5, "int32"),
    "blk_co_off": (6, "int32"),
    "chunk_off": (7, "uint16"),
    "vtx_slot_off": (8, "int32"),
    "loc_dst": (9, "int32"),
    "vblk_off": (10, "int32"),
    "vblk_inst": (11, "int32"),
    "inst_vblk_off": (12, "int64"),
}
LAYOUT_DEVICE, LAYOUT_HOST = 0, 1

# solver result codes
APL_RUNNING, APL_PRIMARY_SUCCESS, APL_MAX_STEPS_REACHED, APL_NON_FINITE, APL_STAGNATION = 0, 1, 2, 3, 4

_lock = threading.Lock()
_lib = None

c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_vp = ctypes.c_void_p
c_dp = ctypes.c_void_p  # device pointer to double


class MeshInfo(ctypes.Structure):
    _fields_ = [
        ("n_tets", c_i64),
        ("n_verts", c_i64),
        ("n_inst", c_i64),
        ("n_blocks", c_i64),
        ("n_partials", c_i64),
        ("n_shared_verts", c_i64),
        ("n_vertex_blocks", c_i64),
        ("nloc_max", c_i32),
        ("chunks", c_i32),
        ("dtype", c_i32),
        ("elem_grid_snh", c_i32),
        ("elem_grid_arap", c_i32),
        ("device_bytes", c_i64),
    ]


class PncgParams(ctypes.Structure):
    _fields_ = [
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("max_steps", c_i32),
        ("max_backtracking", c_i32),
        ("overstep", ctypes.c_double),
        ("c1", ctypes.c_double),
        ("wolfe_delta", ctypes.c_double),
        ("wolfe_eps", ctypes.c_double),
        ("beta_max", ctypes.c_double),
        ("iters_per_launch", c_i32),
        ("use_graph", c_i32),
        ("grid_override", c_i32),
        ("stall_steps", c_i32),
    ]


class PncgStats(ctypes.Structure):
    _fields_ = [
        ("status", c_i32),
        ("n_steps", c_i32),
        ("n_evals", c_i32),
        ("n_backtracks", c_i32),
        ("f", ctypes.c_double),
        ("grad_norm", ctypes.c_double),
        ("grad_norm_first", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("f_first", ctypes.c_double),
    ]


class PncgState(ctypes.Structure):
    _fields_ = [
        ("status", c_i32),
        ("n_steps", c_i32),
        ("n_evals", c_i32),
        ("ls_steps", c_i32),
        ("ok", c_i32),
        ("alpha", ctypes.c_double),
        ("f0", ctypes.c_double),
        ("f_alpha", ctypes.c_double),
        ("grad_norm", ctypes.c_double),
        ("grad_norm_first", ctypes.c_double),
        ("hess_diag_mean", ctypes.c_double),
        ("beta", ctypes.c_double),
    ]


def _declare(lib):
    lib.apl_last_error.restype = ctypes.c_char_p
    lib.apl_version.restype = ctypes.c_int
    lib.apl_mesh_create.argtypes = [
        ctypes.c_int, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, ctypes.c_int, ctypes.c_int,
        ctypes.POINTER(c_vp),
    ]
    lib.apl_mesh_create_from_points.argtypes = [
        ctypes.c_int, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, ctypes.c_int, ctypes.c_int,
        ctypes.POINTER(c_vp),
    ]
    lib.apl_model_create.argtypes = [
        ctypes.c_int, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, ctypes.c_int, c_vp, c_vp, ctypes.c_int,
        ctypes.c_int, ctypes.POINTER(c_vp),
    ]
    lib.apl_mesh_region.argtypes = [c_vp, c_vp, c_vp, c_vp]
    lib.apl_mesh_destroy.argtypes = [c_vp]
    lib.apl_mesh_info.argtypes = [c_vp, ctypes.POINTER(MeshInfo)]
    lib.apl_mesh_set_schedule.argtypes = [c_vp, ctypes.c_int]
    lib.apl_layout_mode.argtypes = [ctypes.c_int]
    lib.apl_mesh_layout.argtypes = [c_vp, ctypes.c_int, c_vp, c_i64]
    lib.apl_morton_order.argtypes = [c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, ctypes.c_int, c_vp]
    lib.apl_mesh_lumped_volume.argtypes = [c_vp, c_vp, c_vp, c_vp]
    lib.apl_elem_eval.argtypes = [
        c_vp, ctypes.c_int, c_vp, c_vp, c_vp, c_vp, ctypes.c_uint, c_dp, c_vp, c_vp, c_vp, c_dp, c_vp, c_vp, c_vp,
    ]
    lib.apl_elem_eval_static.argtypes = lib.apl_elem_eval.argtypes
    lib.apl_static_scratch_bytes.restype = ctypes.c_size_t
    lib.apl_mesh_set_activation.argtypes = [c_vp, c_vp]
    lib.apl_elem_mixed_activation.argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_vp]
    lib.apl_ext_force_eval.argtypes = [c_vp, c_i64, c_vp, c_vp, c_vp, c_dp, c_vp, ctypes.c_int, c_vp]
    lib.apl_bench_elem_pass.argtypes = [
        c_vp, ctypes.c_int, ctypes.c_uint, c_vp, c_vp, c_vp, c_vp, c_dp, c_vp, c_vp, ctypes.c_int,
        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), c_vp,
    ]
    lib.apl_bench_elem_graph.argtypes = [
        c_vp, ctypes.c_int, ctypes.c_uint, c_vp, c_vp, c_vp, c_vp, c_dp, c_vp, c_vp, ctypes.c_int,
        ctypes.POINTER(ctypes.c_double), c_vp,
    ]
    if hasattr(lib, "apl_solver_create"):
        lib.apl_solver_create.argtypes = [c_vp, ctypes.c_int, ctypes.POINTER(c_vp)]
        lib.apl_solver_destroy.argtypes = [c_vp]
        lib.apl_pncg_solve.argtypes = [
            c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(PncgParams), ctypes.POINTER(PncgStats), c_vp,
        ]
        lib.apl_pncg_begin.argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(PncgParams), c_vp]
        lib.apl_pncg_advance.argtypes = [c_vp, c_i32, c_i64, c_vp, ctypes.POINTER(PncgState), c_vp]
        lib.apl_pncg_end.argtypes = [c_vp, ctypes.POINTER(PncgStats), c_vp]
        lib.apl_cg_solve.argtypes = [
            c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(PncgStats),
            c_vp,
        ]
        lib.apl_cg_solve_normal.argtypes = lib.apl_cg_solve.argtypes
    return lib


def lib():
    """Load the library once (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"{LIB_PATH} is missing: build it with `python -m apple_amd.build` "
                        "(apple_amd has no CPU fallback)"
                    )
                _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError("apple_amd: " + lib().apl_last_error().decode())


def exported_symbols() -> list[str]:
    """Names declared in include/apple_amd.h (used by the ABI export test)."""
    import re

    hdr = os.path.join(os.path.dirname(_HERE), "include", "apple_amd.h")
    with open(hdr) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(apl_[a-z0-9_]+)\s*\(", text)))
