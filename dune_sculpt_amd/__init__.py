"""MI355X sculpt brush engine -- Python host mirror of the C sculpt/PBVH interface.

The product is libsculpt_gpu.so (HIP kernels for gfx950 behind the C ABI declared in
include/sculpt_gpu.h).  This module only binds that library with ctypes: it mirrors
the reference's PBVH operator interface (BKE_pbvh_build_mesh, BKE_pbvh_search_gather,
BKE_pbvh_update_bounds, BKE_pbvh_update_normals ... kernel/intern/pbvh.c) and the brush
entry points that replace the absent do_brush_action.  There is no CPU fallback: if the
library is missing, importing the engine raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libsculpt_gpu.so"

TOOL_DRAW, TOOL_SMOOTH, TOOL_INFLATE, TOOL_GRAB, TOOL_ELASTIC_DEFORM = 1, 2, 4, 5, 21
ELASTIC_GRAB, ELASTIC_GRAB_BISCALE, ELASTIC_GRAB_TRISCALE, ELASTIC_SCALE, ELASTIC_TWIST = range(5)
CURVE_CUSTOM, CURVE_SMOOTH, CURVE_SPHERE, CURVE_ROOT, CURVE_SHARP = 0, 1, 2, 3, 4
CURVE_LIN, CURVE_POW4, CURVE_INVSQUARE, CURVE_CONSTANT, CURVE_SMOOTHER = 5, 6, 7, 8, 9
BRUSH_DIR_IN = 1 << 9
CM_TABLE = 256
STROKE_GRAPH, STROKE_REFIT_EACH, STROKE_LOG = 1, 2, 4

PBVH_Leaf = 1 << 0
PBVH_UpdateNormals = 1 << 1
PBVH_UpdateBB = 1 << 2
PBVH_UpdateOriginalBB = 1 << 3
PBVH_UpdateRedraw = 1 << 5
PBVH_FullyHidden = 1 << 9
ME_HIDE = 1 << 4
ME_SMOOTH = 1 << 0

TOOL_NAMES = {TOOL_DRAW: "draw", TOOL_SMOOTH: "smooth", TOOL_INFLATE: "inflate", TOOL_GRAB: "grab",
              TOOL_ELASTIC_DEFORM: "elastic_deform"}


class MVert(C.Structure):  # types_meshdata.h:13-17
    _fields_ = [("co", C.c_float * 3), ("flag", C.c_char), ("bweight", C.c_char), ("_pad", C.c_char * 2)]


class MPoly(C.Structure):  # types_meshdata.h:50-57
    _fields_ = [("loopstart", C.c_int), ("totloop", C.c_int), ("mat_nr", C.c_short), ("flag", C.c_char),
                ("_pad", C.c_char)]


class MLoop(C.Structure):  # types_meshdata.h:69-74
    _fields_ = [("v", C.c_uint), ("e", C.c_uint)]


class MLoopTri(C.Structure):  # types_meshdata.h:178-181
    _fields_ = [("tri", C.c_uint * 3), ("poly", C.c_uint)]


class SgDabParams(C.Structure):
    _fields_ = [
        ("sculpt_tool", C.c_int),
        ("curve_preset", C.c_int),
        ("flag", C.c_int),
        ("smooth_iterations", C.c_int),
        ("location", C.c_float * 3),
        ("radius", C.c_float),
        ("strength", C.c_float),
        ("hardness", C.c_float),
        ("normal_radius_factor", C.c_float),
        ("grab_delta", C.c_float * 3),
        ("curve_table", C.POINTER(C.c_float)),
        ("elastic_deform_type", C.c_int),
        ("elastic_deform_volume_preservation", C.c_float),
        ("surface_normal", C.c_float * 3),
    ]


class SgRay(C.Structure):
    _fields_ = [("start", C.c_float * 3), ("normal", C.c_float * 3), ("depth", C.c_float), ("reserved", C.c_int)]


class SgRayHit(C.Structure):
    _fields_ = [("hit", C.c_int), ("depth", C.c_float), ("location", C.c_float * 3), ("face_normal", C.c_float * 3),
                ("active_vertex", C.c_int), ("active_face", C.c_int), ("looptri", C.c_int), ("node", C.c_int),
                ("status", C.c_int), ("reserved", C.c_int)]


RAY_OVERFLOW = -2
FLT_MAX = float(np.finfo(np.float32).max)

SearchCB = C.CFUNCTYPE(C.c_bool, C.c_void_p, C.c_void_p)
HitCB = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)
OccludedCB = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.POINTER(C.c_float))


class IsectRayPrecalc(C.Structure):
    _fields_ = [("kx", C.c_int), ("ky", C.c_int), ("kz", C.c_int), ("sx", C.c_float), ("sy", C.c_float),
                ("sz", C.c_float)]

_lib = None

# every symbol include/sculpt_gpu.h declares
EXPORTS = [
    "BKE_pbvh_new", "BKE_pbvh_free", "BKE_pbvh_build_mesh", "BKE_pbvh_search_gather",
    "BKE_pbvh_search_callback", "BKE_pbvh_update_bounds", "BKE_pbvh_update_normals",
    "BKE_pbvh_vert_coords_apply", "BKE_pbvh_vert_coords_alloc", "BKE_pbvh_node_mark_update",
    "BKE_pbvh_node_mark_normals_update", "BKE_pbvh_node_mark_redraw", "BKE_pbvh_vert_mark_update",
    "BKE_pbvh_node_get_BB", "BKE_pbvh_node_get_original_BB", "BKE_pbvh_node_num_verts",
    "BKE_pbvh_node_get_verts", "BKE_pbvh_node_fully_hidden_get", "BKE_pbvh_node_add_proxy",
    "BKE_pbvh_node_free_proxies", "BKE_pbvh_gather_proxies", "BKE_pbvh_node_frustum_contain_AABB",
    "BKE_pbvh_node_frustum_exclude_AABB", "BKE_pbvh_get_verts", "DUNE_pbvh_get_vert_normals",
    "DUNE_pbvh_respect_hide_set", "BKE_pbvh_respect_hide_set", "BKE_pbvh_type",
    "sg_dab_params_default", "sg_pbvh_leaf_limit_set", "sg_dab_apply", "sg_dab_apply_stroke",
    "sg_dab_stroke_log", "sg_sync", "sg_last_error", "sg_dab_last_affected", "sg_dab_last_stats",
    "sg_dab_last_gathered", "sg_pbvh_set_stats", "sg_pbvh_set_grid", "sg_pbvh_read_positions",
    "sg_pbvh_read_normals", "sg_pbvh_set_mask", "sg_pbvh_tree_info", "sg_pbvh_node_index",
    "sg_pbvh_node_at", "sg_pbvh_node_flag", "sg_pbvh_set_fused", "sg_pbvh_halo_info", "sg_pbvh_prim_indices", "sg_mesh_icosphere_size", "sg_mesh_icosphere",
    "sg_mesh_torus", "sg_mesh_noise_displace", "sg_pbvh_set_profile", "sg_pbvh_profile_read",
    "sg_stroke_prepare", "sg_stroke_launch", "sg_stroke_free", "sg_pbvh_set_kernel_grids",
    "sg_pbvh_get_kernel_grids", "sg_pbvh_snapshot_save", "sg_pbvh_snapshot_restore",
    "sg_pbvh_autotune", "sg_mlp_new", "sg_mlp_free", "sg_mlp_get_params", "sg_mlp_set_params",
    "sg_mlp_set_norm", "sg_mlp_get_norm", "sg_mlp_set_decision", "sg_mlp_get_decision", "sg_mlp_set_data",
    "sg_mlp_gradient", "sg_mlp_train", "sg_mlp_predict", "sg_mfma_probe", "sg_pbvh_set_learned",
    "sg_pbvh_learned_stats", "sg_learned_collect", "sg_pbvh_timer_start", "sg_pbvh_timer_stop",
    "sg_learned_judge", "sg_batch_new", "sg_batch_free", "sg_batch_info", "sg_batch_load", "sg_batch_launch",
    "sg_batch_rewind", "sg_batch_run", "sg_batch_sync", "sg_batch_totals", "sg_batch_read_positions",
    "sg_batch_read_normals", "sg_batch_snapshot_save", "sg_batch_snapshot_restore", "sg_batch_set_launch",
    "sg_batch_get_launch", "sg_batch_autotune", "sg_batch_set_learned", "sg_batch_set_persistent", "sg_batch_timer_start",
    "sg_batch_timer_stop", "sg_hbm_copy_probe", "sg_raycast", "sg_raycast_device", "sg_raycast_stats", "sg_raycast_stats_noalloc",
    "sg_dab_apply_stroke_placed", "sg_stroke_prepare_placed", "sg_stroke_placements", "BKE_pbvh_raycast",
    "BKE_pbvh_find_nearest_to_ray", "BKE_pbvh_node_get_tmin", "BKE_pbvh_node_raycast",
    "BKE_pbvh_node_find_nearest_to_ray", "BKE_pbvh_raycast_project_ray_root", "sg_isect_ray_tri_precalc",
    "sg_pbvh_draw_buffers_build", "sg_pbvh_draw_buffers_update", "sg_pbvh_draw_buffers_info",
    "sg_pbvh_draw_buffer_node", "sg_pbvh_draw_buffer_read", "sg_pbvh_draw_buffers_free",
    "sg_undo_begin", "sg_undo_push_node", "sg_undo_end", "sg_undo", "sg_redo", "sg_undo_info", "sg_undo_clear",
    "sg_undo_cancel", "sg_undo_original", "sg_pbvh_combine_proxies", "sg_undo_restore_original",
]
SMOOTH_LEARNED = 1 << 30
DAB_ORIGINAL = 1 << 28  # Grab from the original coordinates (SG_DAB_ORIGINAL)
NCTR = 24

KERNEL_NAMES = ["k_search", "k_area", "k_brush", "k_smooth_iter", "k_post", "k_flush_sub", "k_flush_top"]


def lib():
    """Load libsculpt_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback exists)")
    L = C.CDLL(str(LIB_PATH))
    vp, i32, f32p, i32p = C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int)
    sig = {
        "BKE_pbvh_new": (vp, []),
        "BKE_pbvh_free": (None, [vp]),
        "BKE_pbvh_build_mesh": (None, [vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32]),
        "BKE_pbvh_search_gather": (None, [vp, SearchCB, vp, C.POINTER(C.POINTER(vp)), i32p]),
        "BKE_pbvh_search_callback": (None, [vp, SearchCB, vp, HitCB, vp]),
        "BKE_pbvh_update_bounds": (None, [vp, i32]),
        "BKE_pbvh_update_normals": (None, [vp, vp]),
        "BKE_pbvh_vert_coords_apply": (None, [vp, vp, i32]),
        "BKE_pbvh_vert_coords_alloc": (vp, [vp]),
        "BKE_pbvh_node_mark_update": (None, [vp]),
        "BKE_pbvh_node_mark_normals_update": (None, [vp]),
        "BKE_pbvh_node_mark_redraw": (None, [vp]),
        "BKE_pbvh_vert_mark_update": (None, [vp, i32]),
        "BKE_pbvh_node_get_BB": (None, [vp, f32p, f32p]),
        "BKE_pbvh_node_get_original_BB": (None, [vp, f32p, f32p]),
        "BKE_pbvh_node_num_verts": (None, [vp, vp, i32p, i32p]),
        "BKE_pbvh_node_get_verts": (None, [vp, vp, C.POINTER(i32p), C.POINTER(vp)]),
        "BKE_pbvh_node_fully_hidden_get": (C.c_bool, [vp]),
        "BKE_pbvh_node_add_proxy": (vp, [vp, vp]),
        "BKE_pbvh_node_free_proxies": (None, [vp]),
        "BKE_pbvh_gather_proxies": (None, [vp, C.POINTER(C.POINTER(vp)), i32p]),
        "BKE_pbvh_node_frustum_contain_AABB": (C.c_bool, [vp, vp]),
        "BKE_pbvh_node_frustum_exclude_AABB": (C.c_bool, [vp, vp]),
        "BKE_pbvh_get_verts": (vp, [vp]),
        "DUNE_pbvh_get_vert_normals": (vp, [vp]),
        "DUNE_pbvh_respect_hide_set": (None, [vp, C.c_bool]),
        "BKE_pbvh_type": (i32, [vp]),
        "sg_dab_params_default": (None, [C.POINTER(SgDabParams), i32]),
        "sg_pbvh_leaf_limit_set": (i32, [vp, i32]),
        "sg_dab_apply": (i32, [vp, C.POINTER(SgDabParams)]),
        "sg_dab_apply_stroke": (i32, [vp, C.POINTER(SgDabParams), i32, i32]),
        "sg_dab_stroke_log": (i32, [vp, C.POINTER(C.c_longlong), i32]),
        "sg_sync": (i32, [vp]),
        "sg_last_error": (C.c_char_p, []),
        "sg_dab_last_affected": (i32, [vp, i32p, i32]),
        "sg_dab_last_stats": (i32, [vp, C.POINTER(C.c_int64)]),
        "sg_dab_last_gathered": (i32, [vp, i32p, i32]),
        "sg_pbvh_set_stats": (i32, [vp, i32]),
        "sg_pbvh_set_grid": (i32, [vp, i32]),
        "sg_pbvh_read_positions": (i32, [vp, f32p]),
        "sg_pbvh_read_normals": (i32, [vp, f32p]),
        "sg_pbvh_set_mask": (i32, [vp, f32p]),
        "sg_pbvh_tree_info": (i32, [vp, i32p]),
        "sg_pbvh_node_index": (i32, [vp, vp]),
        "sg_pbvh_node_at": (vp, [vp, i32]),
        "sg_pbvh_node_flag": (i32, [vp]),
        "sg_pbvh_set_fused": (i32, [vp, i32]),
        "sg_pbvh_halo_info": (i32, [vp, C.POINTER(C.c_longlong)]),
        "sg_pbvh_prim_indices": (i32, [vp, i32p, i32]),
        "sg_mesh_icosphere_size": (i32, [i32, i32p, i32p]),
        "sg_mesh_icosphere": (i32, [i32, C.c_float, f32p, i32p]),
        "sg_mesh_torus": (i32, [i32, i32, C.c_float, C.c_float, f32p, i32p]),
        "sg_mesh_noise_displace": (i32, [f32p, f32p, i32, C.c_uint32, C.c_float, C.c_float, i32]),
        "sg_pbvh_set_profile": (i32, [vp, i32]),
        "sg_stroke_prepare": (vp, [vp, C.POINTER(SgDabParams), i32, i32]),
        "sg_pbvh_set_kernel_grids": (i32, [vp, i32, i32, i32]),
        "sg_pbvh_get_kernel_grids": (i32, [vp, i32p]),
        "sg_pbvh_snapshot_save": (i32, [vp]),
        "sg_pbvh_autotune": (i32, [vp, C.POINTER(SgDabParams), i32, C.POINTER(C.c_double)]),
        "sg_mlp_new": (vp, [C.c_uint64]),
        "sg_mlp_free": (None, [vp]),
        "sg_mlp_get_params": (i32, [vp, f32p]),
        "sg_mlp_set_params": (i32, [vp, f32p]),
        "sg_mlp_set_norm": (i32, [vp, f32p, f32p, f32p]),
        "sg_mlp_get_norm": (i32, [vp, f32p]),
        "sg_mlp_set_decision": (i32, [vp, C.c_float, C.c_float]),
        "sg_mlp_get_decision": (i32, [vp, f32p]),
        "sg_mlp_set_data": (i32, [vp, f32p, f32p, i32]),
        "sg_mlp_gradient": (i32, [vp, i32, f32p, f32p]),
        "sg_mlp_train": (i32, [vp, i32, i32, C.c_float, C.c_uint64, f32p]),
        "sg_mlp_predict": (i32, [vp, f32p, i32, f32p]),
        "sg_mfma_probe": (i32, [C.POINTER(C.c_uint16), C.POINTER(C.c_uint16), f32p]),
        "sg_pbvh_set_learned": (i32, [vp, vp]),
        "sg_pbvh_learned_stats": (i32, [vp, C.POINTER(C.c_longlong)]),
        "sg_learned_collect": (i32, [vp, C.POINTER(SgDabParams), f32p, f32p, f32p, i32]),
        "sg_learned_judge": (i32, [vp, C.POINTER(SgDabParams), C.POINTER(C.c_double)]),
        "sg_hbm_copy_probe": (i32, [C.c_longlong, i32, f32p]),
        "sg_batch_new": (vp, [i32, C.POINTER(f32p), C.POINTER(i32p), i32p, i32p, i32]),
        "sg_batch_free": (None, [vp]),
        "sg_batch_info": (i32, [vp, C.POINTER(C.c_longlong), i32]),
        "sg_batch_load": (i32, [vp, C.POINTER(SgDabParams), i32]),
        "sg_batch_launch": (i32, [vp, i32, i32]),
        "sg_batch_rewind": (i32, [vp]),
        "sg_batch_run": (i32, [vp, C.POINTER(SgDabParams), i32, i32]),
        "sg_batch_sync": (i32, [vp]),
        "sg_batch_totals": (i32, [vp, C.POINTER(C.c_longlong), i32]),
        "sg_batch_read_positions": (i32, [vp, i32, f32p]),
        "sg_batch_read_normals": (i32, [vp, i32, f32p]),
        "sg_batch_snapshot_save": (i32, [vp]),
        "sg_batch_snapshot_restore": (i32, [vp]),
        "sg_batch_set_launch": (i32, [vp, i32p]),
        "sg_batch_get_launch": (i32, [vp, i32p]),
        "sg_batch_autotune": (i32, [vp, C.POINTER(SgDabParams), i32, C.POINTER(C.c_double)]),
        "sg_batch_set_learned": (i32, [vp, vp]),
        "sg_batch_set_persistent": (i32, [vp, i32]),
        "sg_batch_timer_start": (i32, [vp]),
        "sg_batch_timer_stop": (i32, [vp, f32p]),
        "sg_pbvh_timer_start": (i32, [vp]),
        "sg_pbvh_timer_stop": (i32, [vp, f32p]),
        "sg_pbvh_snapshot_restore": (i32, [vp]),
        "sg_stroke_launch": (i32, [vp, i32]),
        "sg_stroke_free": (None, [vp]),
        "sg_pbvh_profile_read": (i32, [vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong), i32, i32]),
        "sg_raycast": (i32, [vp, C.POINTER(SgRay), i32, i32, C.POINTER(SgRayHit)]),
        "sg_raycast_device": (i32, [vp, vp, i32, i32, vp]),
        "sg_raycast_stats": (i32, [vp, C.POINTER(SgRay), i32, i32, C.POINTER(C.c_longlong)]),
        "sg_raycast_stats_noalloc": (i32, [vp, C.POINTER(SgRay), i32, i32, C.POINTER(C.c_longlong)]),
        "sg_dab_apply_stroke_placed": (i32, [vp, C.POINTER(SgDabParams), C.POINTER(SgRay), i32, i32]),
        "sg_stroke_prepare_placed": (vp, [vp, C.POINTER(SgDabParams), C.POINTER(SgRay), i32, i32]),
        "sg_stroke_placements": (i32, [vp, f32p, i32]),
        "BKE_pbvh_raycast": (None, [vp, OccludedCB, vp, f32p, f32p, C.c_bool]),
        "BKE_pbvh_find_nearest_to_ray": (None, [vp, OccludedCB, vp, f32p, f32p, C.c_bool]),
        "BKE_pbvh_node_get_tmin": (C.c_float, [vp]),
        "BKE_pbvh_node_raycast": (C.c_bool, [vp, vp, vp, C.c_bool, f32p, f32p, C.POINTER(IsectRayPrecalc), f32p,
                                             i32p, i32p, f32p]),
        "BKE_pbvh_node_find_nearest_to_ray": (C.c_bool, [vp, vp, vp, C.c_bool, f32p, f32p, f32p, f32p]),
        "BKE_pbvh_raycast_project_ray_root": (None, [vp, C.c_bool, f32p, f32p, f32p]),
        "sg_isect_ray_tri_precalc": (None, [C.POINTER(IsectRayPrecalc), f32p]),
        "sg_pbvh_draw_buffers_build": (i32, [vp, i32p, i32, i32, i32]),
        "sg_pbvh_draw_buffers_update": (i32, [vp, i32]),
        "sg_pbvh_draw_buffers_info": (i32, [vp, C.POINTER(C.c_void_p), C.POINTER(C.c_longlong), i32p]),
        "sg_pbvh_draw_buffer_node": (i32, [vp, vp, C.POINTER(C.c_longlong), i32p, i32p, i32p]),
        "sg_pbvh_draw_buffer_read": (C.c_longlong, [vp, vp, vp, C.c_longlong]),
        "sg_pbvh_draw_buffers_free": (None, [vp]),
        "sg_undo_begin": (i32, [vp, C.c_longlong]),
        "sg_undo_push_node": (i32, [vp, vp]),
        "sg_undo_end": (i32, [vp]),
        "sg_undo": (i32, [vp]),
        "sg_redo": (i32, [vp]),
        "sg_undo_info": (i32, [vp, C.POINTER(C.c_longlong)]),
        "sg_undo_clear": (None, [vp]),
        "sg_undo_cancel": (i32, [vp]),
        "sg_undo_restore_original": (i32, [vp]),
        "sg_undo_original": (i32, [vp, vp, vp]),
        "sg_pbvh_combine_proxies": (i32, [vp, i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def last_error() -> str:
    return (lib().sg_last_error() or b"").decode()


def _check(rc, what):
    if rc < 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")
    return rc


# ---------------------------------------------------------------------------
# synthetic meshes (C generators in the library)
# ---------------------------------------------------------------------------

def icosphere(level: int, radius: float = 1.0):
    L = lib()
    V, T = C.c_int(), C.c_int()
    _check(L.sg_mesh_icosphere_size(level, C.byref(V), C.byref(T)), "icosphere size")
    co = np.empty((V.value, 3), np.float32)
    tris = np.empty((T.value, 3), np.int32)
    _check(L.sg_mesh_icosphere(level, radius, _fp(co), _ip(tris)), "icosphere")
    return co, tris


def torus(nu: int, nv: int, major_r: float = 1.0, minor_r: float = 0.3):
    co = np.empty((nu * nv, 3), np.float32)
    tris = np.empty((2 * nu * nv, 3), np.int32)
    _check(lib().sg_mesh_torus(nu, nv, major_r, minor_r, _fp(co), _ip(tris)), "torus")
    return co, tris


def torus_normals(nu: int, nv: int):
    """Analytic outward normals of torus(nu, nv) (direction for noise displacement)."""
    u = 2 * np.pi * np.arange(nu)[:, None] / nu
    v = 2 * np.pi * np.arange(nv)[None, :] / nv
    n = np.stack([np.cos(v) * np.cos(u), np.cos(v) * np.sin(u), np.sin(v) + 0 * u], -1)
    return n.reshape(-1, 3).astype(np.float32)


def noise_displace(co, direction, seed: int, amplitude: float, frequency: float = 4.0, octaves: int = 4):
    co = np.ascontiguousarray(co, np.float32)
    direction = np.ascontiguousarray(direction, np.float32)
    _check(lib().sg_mesh_noise_displace(_fp(co), _fp(direction), co.shape[0], seed, amplitude, frequency, octaves),
           "noise")
    return co


def noisy_icosphere(level: int, seed: int, amplitude: float):
    co, tris = icosphere(level)
    if amplitude:
        d = co / np.linalg.norm(co, axis=1, keepdims=True)
        co = noise_displace(co, d.astype(np.float32), seed, amplitude)
    return co, tris


def noisy_torus(nu: int, nv: int, seed: int, amplitude: float, major_r=1.0, minor_r=0.3):
    co, tris = torus(nu, nv, major_r, minor_r)
    if amplitude:
        co = noise_displace(co, torus_normals(nu, nv), seed, amplitude)
    return co, tris


# ---------------------------------------------------------------------------
# strokes
# ---------------------------------------------------------------------------

def great_circle_centres(seed: int, radii, spacing_frac: float = 0.1, sphere_radius: float = 1.0):
    """Dab centres along a seeded random great circle (axis uniform on S^2), arc step
    spacing_frac * r_i between dab i-1 and dab i."""
    rng = np.random.default_rng(seed)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    a = np.cross(axis, [1.0, 0.0, 0.0])
    if np.linalg.norm(a) < 1e-6:
        a = np.cross(axis, [0.0, 1.0, 0.0])
    a /= np.linalg.norm(a)
    b = np.cross(axis, a)
    theta0 = rng.uniform(0, 2 * np.pi)
    radii = np.asarray(radii, np.float64)
    steps = spacing_frac * radii / sphere_radius
    theta = theta0 + np.concatenate([[0.0], np.cumsum(steps[1:])])
    pts = sphere_radius * (np.cos(theta)[:, None] * a + np.sin(theta)[:, None] * b)
    return pts.astype(np.float32)


def make_dab(tool=TOOL_DRAW, location=(0, 0, 1), radius=0.1, strength=0.5, falloff=CURVE_SMOOTH, hardness=0.0,
             smooth_iterations=1, grab_delta=(0, 0, 0), flag=0, normal_radius_factor=0.5, curve_table=None,
             elastic_type=0, volume_preservation=0.4, surface_normal=(0, 0, 1)):
    d = SgDabParams()
    d.elastic_deform_type = elastic_type
    d.elastic_deform_volume_preservation = volume_preservation
    d.surface_normal[:] = [float(x) for x in surface_normal]
    d.sculpt_tool = tool
    d.curve_preset = falloff
    d.flag = flag
    d.smooth_iterations = smooth_iterations
    d.location[:] = [float(x) for x in location]
    d.radius = radius
    d.strength = strength
    d.hardness = hardness
    d.normal_radius_factor = normal_radius_factor
    d.grab_delta[:] = [float(x) for x in grab_delta]
    if curve_table is not None:
        tab = np.ascontiguousarray(curve_table, np.float32)
        assert tab.shape == (CM_TABLE + 1,)
        d.curve_table = _fp(tab)
        d._keep = tab  # keep the table alive with the struct
    return d


def stroke_config(tools, centres, radii, strengths, smooth_iterations=4, falloff=CURVE_SMOOTH):
    """Dab list: tool i, centre i, radius i, strength i; Grab delta = next centre - centre."""
    dabs = []
    n = len(centres)
    for i in range(n):
        nxt = centres[i + 1] if i + 1 < n else centres[i]
        dabs.append(make_dab(tool=int(tools[i]), location=centres[i], radius=float(radii[i]),
                             strength=float(strengths[i]), falloff=falloff, smooth_iterations=smooth_iterations,
                             grab_delta=(np.asarray(nxt) - np.asarray(centres[i]))))
    return dabs


# ---------------------------------------------------------------------------
# PBVH handle
# ---------------------------------------------------------------------------

class MLP:
    """Learned Smooth model: 28 -> 64 -> 64 -> 3 (SiLU), FP32 master weights on the GPU,
    BF16 MFMA inference."""

    NPARAM = 6211

    def __init__(self, seed=0):
        self.h = lib().sg_mlp_new(seed)
        if not self.h:
            raise RuntimeError("sg_mlp_new failed: " + last_error())

    def close(self):
        if getattr(self, "h", None):
            lib().sg_mlp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.NPARAM, np.float32)
        _check(lib().sg_mlp_get_params(self.h, _fp(out)), "get params")
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.shape == (self.NPARAM,)
        _check(lib().sg_mlp_set_params(self.h, _fp(w)), "set params")

    def set_norm(self, mu, isig, out_scale):
        a, b, c = (np.ascontiguousarray(x, np.float32) for x in (mu, isig, out_scale))
        _check(lib().sg_mlp_set_norm(self.h, _fp(a), _fp(b), _fp(c)), "set norm")

    def norm(self):
        out = np.zeros(64, np.float32)
        lib().sg_mlp_get_norm(self.h, _fp(out))
        return out[:28].copy(), out[28:56].copy(), out[56:59].copy()

    def set_decision(self, alpha, tol):
        lib().sg_mlp_set_decision(self.h, alpha, tol)

    def decision(self):
        out = np.zeros(2, np.float32)
        lib().sg_mlp_get_decision(self.h, _fp(out))
        return float(out[0]), float(out[1])

    def set_data(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        self._X, self._T = X, T
        _check(lib().sg_mlp_set_data(self.h, _fp(X), _fp(T), X.shape[0]), "set data")

    def gradient(self, n):
        g = np.zeros(self.NPARAM, np.float32)
        loss = C.c_float()
        _check(lib().sg_mlp_gradient(self.h, n, _fp(g), C.byref(loss)), "gradient")
        return g, loss.value

    def train(self, epochs, batch_wg=256, lr=1e-3, seed=0):
        curve = np.zeros(epochs, np.float32)
        _check(lib().sg_mlp_train(self.h, epochs, batch_wg, lr, seed, _fp(curve)), "train")
        return curve

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], 3), np.float32)
        _check(lib().sg_mlp_predict(self.h, _fp(X), X.shape[0], _fp(Y)), "predict")
        return Y

    def save(self, path):
        mu, isig, osc = self.norm()
        alpha, tol = self.decision()
        np.savez(path, w=self.params(), mu=mu, isig=isig, out_scale=osc, alpha=alpha, tol=tol)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        m = cls()
        m.set_norm(z["mu"], z["isig"], z["out_scale"])
        m.set_params(z["w"])
        m.set_decision(float(z["alpha"]), float(z["tol"]))
        return m


def hbm_copy_probe(nbytes, reps=5):
    """Best time (ms) of a device-to-device copy of nbytes; returns (ms, GB/s counting
    read + write)."""
    ms = C.c_float()
    _check(lib().sg_hbm_copy_probe(int(nbytes), reps, C.byref(ms)), "copy probe")
    return ms.value, 2 * nbytes / (ms.value * 1e-3) / 1e9


def mfma_probe(A, B):
    """Run the 16x16x16 bf16 MFMA once with the engine's operand layout."""
    A = np.ascontiguousarray(A, np.uint16)
    B = np.ascontiguousarray(B, np.uint16)
    D = np.zeros((16, 16), np.float32)
    _check(lib().sg_mfma_probe(A.ctypes.data_as(C.POINTER(C.c_uint16)), B.ctypes.data_as(C.POINTER(C.c_uint16)),
                               _fp(D)), "mfma probe")
    return D


class PreparedStroke:
    """A dab list captured into a hipGraph; freed before its PBVH."""

    def __init__(self, h, owner):
        self.h = h
        self.owner = owner

    def launch(self, sync=True):
        if not self.h:
            raise RuntimeError("stroke already freed")
        _check(lib().sg_stroke_launch(self.h, 1 if sync else 0), "sg_stroke_launch")

    def close(self):
        if self.h and getattr(self.owner, "h", None):
            lib().sg_stroke_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PBVH:
    """Owner of one GPU-resident PBVH (BKE_pbvh_new/build_mesh ... BKE_pbvh_free)."""

    def __init__(self, co, tris, leaf_limit: int = 128, hidden=None, mat=None, polys=None):
        """polys (optional): (poly_off, poly_verts, smooth flags, mat_nr) -- the mesh's
        polygons, fan-triangulated into looptris; then `tris` is ignored."""
        L = lib()
        co = np.ascontiguousarray(co, np.float32)
        if polys is not None:
            return self._init_polys(co, polys, leaf_limit, hidden)
        tris = np.ascontiguousarray(tris, np.int32)
        self.totvert = co.shape[0]
        self.tottri = tris.shape[0]
        self._strokes = []
        self.h = L.BKE_pbvh_new()
        if not self.h:
            raise RuntimeError("BKE_pbvh_new failed: " + last_error())
        _check(L.sg_pbvh_leaf_limit_set(self.h, leaf_limit), "leaf limit")
        # reference mesh arrays: MVert, one MLoop per corner, one tri poly per face
        self._mvert = (MVert * self.totvert)()
        mv = np.frombuffer(self._mvert, dtype=np.uint8).reshape(self.totvert, 16)
        mv[:, :12] = co.view(np.uint8).reshape(self.totvert, 12)
        if hidden is not None:
            mv[:, 12] = np.where(np.asarray(hidden, bool), ME_HIDE, 0).astype(np.uint8)
        self._mloop = np.zeros((self.tottri * 3, 2), np.uint32)
        self._mloop[:, 0] = tris.reshape(-1)
        self._mpoly = np.zeros(self.tottri, dtype=[("loopstart", "<i4"), ("totloop", "<i4"), ("mat_nr", "<i2"),
                                                   ("flag", "u1"), ("_pad", "u1")])
        self._mpoly["loopstart"] = np.arange(self.tottri) * 3
        self._mpoly["totloop"] = 3
        if mat is not None:
            self._mpoly["mat_nr"] = np.asarray(mat, np.int16)
        # the PBVH takes ownership of looptri (freed with free()): allocate with malloc
        libc = C.CDLL(None)
        libc.malloc.restype = C.c_void_p
        libc.malloc.argtypes = [C.c_size_t]
        lt_ptr = libc.malloc(max(self.tottri, 1) * 16)
        lt = np.frombuffer((C.c_uint32 * (max(self.tottri, 1) * 4)).from_address(lt_ptr), np.uint32).reshape(-1, 4)
        lt[: self.tottri, :3] = np.arange(self.tottri * 3, dtype=np.uint32).reshape(-1, 3)
        lt[: self.tottri, 3] = np.arange(self.tottri, dtype=np.uint32)
        L.BKE_pbvh_build_mesh(self.h, None, self._mpoly.ctypes.data, self._mloop.ctypes.data,
                              C.cast(self._mvert, C.c_void_p), self.totvert, None, None, None, lt_ptr, self.tottri)
        err = last_error()
        info = self.tree_info()
        if info["nodes"] == 0 and self.tottri > 0:
            raise RuntimeError("BKE_pbvh_build_mesh failed: " + err)

    def _init_polys(self, co, polys, leaf_limit, hidden):
        L = lib()
        off, pv, smooth, mat = (np.asarray(x) for x in polys)
        off = off.astype(np.int64)
        P = len(off) - 1
        self.totvert = co.shape[0]
        tri_poly, tri_loops = fan_looptris(off)
        self.tottri = len(tri_poly)
        self.tri_poly = tri_poly
        self._strokes = []
        self.h = L.BKE_pbvh_new()
        if not self.h:
            raise RuntimeError("BKE_pbvh_new failed: " + last_error())
        _check(L.sg_pbvh_leaf_limit_set(self.h, leaf_limit), "leaf limit")
        self._mvert = (MVert * self.totvert)()
        mv = np.frombuffer(self._mvert, dtype=np.uint8).reshape(self.totvert, 16)
        mv[:, :12] = co.view(np.uint8).reshape(self.totvert, 12)
        if hidden is not None:
            mv[:, 12] = np.where(np.asarray(hidden, bool), ME_HIDE, 0).astype(np.uint8)
        self._mloop = np.zeros((len(pv), 2), np.uint32)
        self._mloop[:, 0] = np.asarray(pv, np.uint32)
        self._mpoly = np.zeros(P, dtype=[("loopstart", "<i4"), ("totloop", "<i4"), ("mat_nr", "<i2"),
                                          ("flag", "u1"), ("_pad", "u1")])
        self._mpoly["loopstart"] = off[:-1]
        self._mpoly["totloop"] = np.diff(off)
        self._mpoly["mat_nr"] = np.asarray(mat, np.int16)
        self._mpoly["flag"] = np.where(np.asarray(smooth, bool), ME_SMOOTH, 0).astype(np.uint8)
        libc = C.CDLL(None)
        libc.malloc.restype = C.c_void_p
        libc.malloc.argtypes = [C.c_size_t]
        lt_ptr = libc.malloc(max(self.tottri, 1) * 16)
        lt = np.frombuffer((C.c_uint32 * (max(self.tottri, 1) * 4)).from_address(lt_ptr), np.uint32).reshape(-1, 4)
        lt[: self.tottri, :3] = tri_loops
        lt[: self.tottri, 3] = tri_poly
        L.BKE_pbvh_build_mesh(self.h, None, self._mpoly.ctypes.data, self._mloop.ctypes.data,
                              C.cast(self._mvert, C.c_void_p), self.totvert, None, None, None, lt_ptr, self.tottri)
        self.tris = np.asarray(pv, np.int32)[tri_loops]
        if self.tree_info()["nodes"] == 0 and self.tottri > 0:
            raise RuntimeError("BKE_pbvh_build_mesh failed: " + last_error())

    # -- draw buffers
    def draw_build(self, face_sets, seed=0, default=1):
        fs = np.ascontiguousarray(face_sets, np.int32)
        _check(lib().sg_pbvh_draw_buffers_build(self.h, _ip(fs), len(fs), seed, default), "draw build")

    def draw_update(self, all=False):
        _check(lib().sg_pbvh_draw_buffers_update(self.h, 1 if all else 0), "draw update")

    def draw_info(self):
        ptr, nb, nt = C.c_void_p(), C.c_longlong(), C.c_int()
        _check(lib().sg_pbvh_draw_buffers_info(self.h, C.byref(ptr), C.byref(nb), C.byref(nt)), "draw info")
        return {"bytes": nb.value, "visible_tris": nt.value}

    def draw_leaf(self, index):
        """(records as a (verts, 36) uint8 array, smooth, material) of one leaf."""
        L = lib()
        node = self.node(index)
        off, nv, sm, mt = C.c_longlong(), C.c_int(), C.c_int(), C.c_int()
        _check(L.sg_pbvh_draw_buffer_node(self.h, node, C.byref(off), C.byref(nv), C.byref(sm), C.byref(mt)),
               "draw node")
        out = np.zeros((nv.value, 36), np.uint8)
        n = L.sg_pbvh_draw_buffer_read(self.h, node, out.ctypes.data, out.nbytes)
        if n < 0:
            raise RuntimeError("draw read failed: " + last_error())
        return out, sm.value, mt.value

    # -- undo steps
    def undo_begin(self, max_verts=0):
        _check(lib().sg_undo_begin(self.h, int(max_verts)), "undo begin")

    def undo_push_node(self, index):
        _check(lib().sg_undo_push_node(self.h, self.node(index)), "undo push")

    def undo_end(self):
        _check(lib().sg_undo_end(self.h), "undo end")

    def undo(self):
        """True after undoing a step, False when there is none."""
        return _check(lib().sg_undo(self.h), "undo") == 0

    def redo(self):
        return _check(lib().sg_redo(self.h), "redo") == 0

    def undo_info(self):
        out = (C.c_longlong * 4)()
        _check(lib().sg_undo_info(self.h, out), "undo info")
        return {"steps": out[0], "applied": out[1], "bytes": out[2], "leaves": out[3]}

    def undo_clear(self):
        lib().sg_undo_clear(self.h)

    def undo_cancel(self):
        _check(lib().sg_undo_cancel(self.h), "undo cancel")

    def undo_restore_original(self):
        """The open step's leaves back to their state at begin; the step stays open."""
        _check(lib().sg_undo_restore_original(self.h), "undo restore original")

    def undo_original(self):
        """(positions, normals) at the begin of the open undo step (current ones when
        none is open), in the caller's vertex order."""
        co = np.empty((self.totvert, 3), np.float32)
        no = np.empty((self.totvert, 3), np.float32)
        _check(lib().sg_undo_original(self.h, co.ctypes.data, no.ctypes.data), "undo original")
        return co, no

    # -- proxies
    def add_proxy(self, index):
        """BKE_pbvh_node_add_proxy on node `index`: a zeroed (uniq_verts, 3) float32
        array backed by the proxy's host buffer (write displacements into it)."""
        node = self.node(index)
        px = lib().BKE_pbvh_node_add_proxy(self.h, node)
        if not px:
            raise RuntimeError("add proxy failed: " + last_error())
        uniq = self.node_verts(index)[1]
        co_ptr = C.cast(px, C.POINTER(C.POINTER(C.c_float))).contents
        return np.ctypeslib.as_array(co_ptr, shape=(max(uniq, 1), 3))[:uniq]

    def combine_proxies(self, use_orco=False):
        """sg_pbvh_combine_proxies: folds every node's proxies into the mesh on the GPU;
        returns the number of leaves changed."""
        return _check(lib().sg_pbvh_combine_proxies(self.h, 1 if use_orco else 0), "combine proxies")

    def close(self):
        for st in getattr(self, "_strokes", []):
            st.close()
        self._strokes = []
        if getattr(self, "h", None):
            lib().BKE_pbvh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- brush
    def dab(self, d: SgDabParams):
        _check(lib().sg_dab_apply(self.h, C.byref(d)), "sg_dab_apply")

    def stroke(self, dabs, graph=True, refit_each=False, log=False):
        arr = (SgDabParams * len(dabs))(*dabs)
        flags = (STROKE_GRAPH if graph else 0) | (STROKE_REFIT_EACH if refit_each else 0) | (STROKE_LOG if log else 0)
        _check(lib().sg_dab_apply_stroke(self.h, arr, len(dabs), flags), "sg_dab_apply_stroke")

    def prepare(self, dabs, refit_each=False, log=False):
        """Capture the dab list into a hipGraph; returns a PreparedStroke."""
        arr = (SgDabParams * len(dabs))(*dabs)
        flags = (STROKE_REFIT_EACH if refit_each else 0) | (STROKE_LOG if log else 0)
        h = lib().sg_stroke_prepare(self.h, arr, len(dabs), flags)
        if not h:
            raise RuntimeError("sg_stroke_prepare failed: " + last_error())
        st = PreparedStroke(h, self)
        self._strokes.append(st)
        return st

    def stroke_placed(self, dabs, starts, normals, depths=None, graph=True, refit_each=False, log=False):
        """Dab i centred where ray i hits the mesh the previous dabs left (cast on the GPU)."""
        arr = (SgDabParams * len(dabs))(*dabs)
        rays = make_rays(starts, normals, depths)
        flags = (STROKE_GRAPH if graph else 0) | (STROKE_REFIT_EACH if refit_each else 0) | (STROKE_LOG if log else 0)
        _check(lib().sg_dab_apply_stroke_placed(self.h, arr, rays, len(dabs), flags), "sg_dab_apply_stroke_placed")

    def prepare_placed(self, dabs, starts, normals, depths=None, refit_each=False, log=False):
        arr = (SgDabParams * len(dabs))(*dabs)
        rays = make_rays(starts, normals, depths)
        flags = (STROKE_REFIT_EACH if refit_each else 0) | (STROKE_LOG if log else 0)
        h = lib().sg_stroke_prepare_placed(self.h, arr, rays, len(dabs), flags)
        if not h:
            raise RuntimeError("sg_stroke_prepare_placed failed: " + last_error())
        st = PreparedStroke(h, self)
        self._strokes.append(st)
        return st

    def placements(self, n):
        """(n, 8) placements of the last placed stroke: centre, Grab translation, hit, 0."""
        out = np.zeros((n, 8), np.float32)
        m = _check(lib().sg_stroke_placements(self.h, _fp(out), n), "placements")
        return out[:m]

    def stroke_log(self, n):
        out = np.zeros((n, NCTR), np.int64)
        _check(lib().sg_dab_stroke_log(self.h, out.ctypes.data_as(C.POINTER(C.c_longlong)), n), "stroke log")
        return out

    def set_stats(self, on=True):
        lib().sg_pbvh_set_stats(self.h, 1 if on else 0)

    def judge(self, d):
        """Learned vs plain Smooth dab from the current state (the plain result stays
        applied): (RMS error / r, RMS plain displacement / r, accepted, affected)."""
        out = (C.c_double * 4)()
        _check(lib().sg_learned_judge(self.h, C.byref(d), out), "judge")
        return float(out[0]), float(out[1]), bool(out[2]), int(out[3])

    def timer_start(self):
        """Record a HIP event on the engine's stream (start of a timed region)."""
        _check(lib().sg_pbvh_timer_start(self.h), "timer start")

    def timer_stop(self):
        """Record the closing event, wait for it; milliseconds since timer_start."""
        ms = C.c_float()
        _check(lib().sg_pbvh_timer_stop(self.h, C.byref(ms)), "timer stop")
        return ms.value

    def set_profile(self, on=True):
        _check(lib().sg_pbvh_set_profile(self.h, 1 if on else 0), "profile")

    def profile_read(self, reset=True):
        ms = np.zeros(16, np.float64)
        cnt = np.zeros(16, np.int64)
        n = _check(lib().sg_pbvh_profile_read(self.h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                              cnt.ctypes.data_as(C.POINTER(C.c_longlong)), 16, 1 if reset else 0),
                   "profile read")
        return {KERNEL_NAMES[k]: (float(ms[k]), int(cnt[k])) for k in range(n)}

    def set_fused(self, on=True):
        """Fused dabs on or off; returns whether the mesh has the halo tables."""
        return bool(_check(lib().sg_pbvh_set_fused(self.h, 1 if on else 0), "set fused"))

    def halo_info(self):
        out = (C.c_longlong * 5)()
        _check(lib().sg_pbvh_halo_info(self.h, out), "halo info")
        return {"built": bool(out[0]), "max_locals": int(out[1]), "max_locals_ring0": int(out[2]),
                "max_neighbour_entries": int(out[3]), "bytes": int(out[4])}

    def set_kernel_grids(self, brush=0, post=0, smooth=0):
        _check(lib().sg_pbvh_set_kernel_grids(self.h, brush, post, smooth), "grids")

    def kernel_grids(self):
        out = np.zeros(4, np.int32)
        lib().sg_pbvh_get_kernel_grids(self.h, _ip(out))
        return {"common": int(out[0]), "brush": int(out[1]), "post": int(out[2]), "smooth": int(out[3])}

    def autotune(self, dabs):
        """Pick per-kernel workgroup counts on a sample of dabs; returns (before, after) ms."""
        arr = (SgDabParams * len(dabs))(*dabs)
        out = np.zeros(2, np.float64)
        _check(lib().sg_pbvh_autotune(self.h, arr, len(dabs), out.ctypes.data_as(C.POINTER(C.c_double))),
               "autotune")
        return float(out[0]), float(out[1])

    def set_learned(self, mlp):
        self._mlp = mlp
        _check(lib().sg_pbvh_set_learned(self.h, mlp.h if mlp is not None else None), "set learned")

    def learned_stats(self):
        out = np.zeros(2, np.int64)
        lib().sg_pbvh_learned_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_longlong)))
        return {"accepted": int(out[0]), "fallbacks": int(out[1])}

    def collect(self, d, cap=None):
        """Plain Smooth dab recording learned-model inputs and targets per affected vertex."""
        cap = cap or min(self.totvert, 1 << 18)
        X = np.zeros((cap, 28), np.float32)
        T = np.zeros((cap, 3), np.float32)
        ok = np.zeros(cap, np.float32)
        n = _check(lib().sg_learned_collect(self.h, C.byref(d), _fp(X), _fp(T), _fp(ok), cap), "collect")
        return X[:n], T[:n], ok[:n] > 0.5

    def snapshot(self):
        _check(lib().sg_pbvh_snapshot_save(self.h), "snapshot")

    def restore(self):
        _check(lib().sg_pbvh_snapshot_restore(self.h), "restore")

    def set_grid(self, wg):
        _check(lib().sg_pbvh_set_grid(self.h, wg), "grid")

    def sync(self):
        _check(lib().sg_sync(self.h), "sg_sync")

    # -- readback
    def positions(self):
        out = np.empty((self.totvert, 3), np.float32)
        _check(lib().sg_pbvh_read_positions(self.h, _fp(out)), "read positions")
        return out

    def normals(self):
        out = np.empty((self.totvert, 3), np.float32)
        _check(lib().sg_pbvh_read_normals(self.h, _fp(out)), "read normals")
        return out

    def set_mask(self, mask):
        if mask is None:
            _check(lib().sg_pbvh_set_mask(self.h, None), "mask")
        else:
            m = np.ascontiguousarray(mask, np.float32)
            _check(lib().sg_pbvh_set_mask(self.h, _fp(m)), "mask")

    def last_affected(self):
        cap = self.totvert
        out = np.empty(cap, np.int32)
        n = _check(lib().sg_dab_last_affected(self.h, _ip(out), cap), "affected")
        return out[:n].copy()

    def last_gathered(self):
        out = np.empty(max(self.tree_info()["leaves"], 1), np.int32)
        n = _check(lib().sg_dab_last_gathered(self.h, _ip(out), out.size), "gathered")
        return out[:n].copy()

    def last_stats(self):
        out = np.zeros(NCTR, np.int64)
        _check(lib().sg_dab_last_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))), "stats")
        return out

    def tree_info(self):
        out = np.zeros(4, np.int32)
        lib().sg_pbvh_tree_info(self.h, _ip(out))
        return {"nodes": int(out[0]), "leaves": int(out[1]), "cuts": int(out[2]), "sub_maxn": int(out[3])}

    def node(self, index):
        return lib().sg_pbvh_node_at(self.h, index)

    def node_bb(self, index, original=False):
        mn = np.zeros(3, np.float32)
        mx = np.zeros(3, np.float32)
        fn = lib().BKE_pbvh_node_get_original_BB if original else lib().BKE_pbvh_node_get_BB
        fn(self.node(index), _fp(mn), _fp(mx))
        return np.concatenate([mn, mx])

    def node_verts(self, index):
        L = lib()
        node = self.node(index)
        u, t = C.c_int(), C.c_int()
        L.BKE_pbvh_node_num_verts(self.h, node, C.byref(u), C.byref(t))
        ptr = C.POINTER(C.c_int)()
        L.BKE_pbvh_node_get_verts(self.h, node, C.byref(ptr), None)
        return np.ctypeslib.as_array(ptr, shape=(t.value,)).copy() if t.value else np.zeros(0, np.int32), u.value

    def prim_indices(self):
        out = np.empty(self.tottri, np.int32)
        lib().sg_pbvh_prim_indices(self.h, _ip(out), self.tottri)
        return out

    def raycast(self, starts, normals, depths=None, original=False):
        """Rays (n x 3 starts, n x 3 directions, optional initial depths) -> structured
        array of hits (sg_raycast)."""
        rays = make_rays(starts, normals, depths)
        hits = (SgRayHit * len(rays))()
        _check(lib().sg_raycast(self.h, rays, len(rays), 1 if original else 0, hits), "sg_raycast")
        return hits_to_numpy(hits)

    def raycast_stats(self, starts, normals, depths=None, original=False):
        """Work counters of a ray batch: boxes tested, leaves scanned, tris tested."""
        rays = make_rays(starts, normals, depths)
        out = (C.c_longlong * 3)()
        _check(lib().sg_raycast_stats(self.h, rays, len(rays), 1 if original else 0, out), "sg_raycast_stats")
        return {"boxes": int(out[0]), "leaves": int(out[1]), "tris": int(out[2])}

    def raycast_stats_noalloc(self, starts, normals, depths=None, original=False):
        """raycast_stats through the entry point that allocates nothing (static scratch)."""
        rays = make_rays(starts, normals, depths)
        out = (C.c_longlong * 3)()
        _check(lib().sg_raycast_stats_noalloc(self.h, rays, len(rays), 1 if original else 0, out),
               "sg_raycast_stats_noalloc")
        return {"boxes": int(out[0]), "leaves": int(out[1]), "tris": int(out[2])}

    def occluded_order(self, start, direction, mode=0, original=False):
        """BKE_pbvh_raycast (mode 0) / BKE_pbvh_find_nearest_to_ray (mode 1) with a
        recording callback: [(node index, tmin)] in visiting order."""
        L = lib()
        seen = []

        def rec(node, data, tmin):
            seen.append((L.sg_pbvh_node_index(self.h, node), L.BKE_pbvh_node_get_tmin(node)))

        cb = OccludedCB(rec)
        a = np.ascontiguousarray(start, np.float32)
        b = np.ascontiguousarray(direction, np.float32)
        fn = L.BKE_pbvh_raycast if mode == 0 else L.BKE_pbvh_find_nearest_to_ray
        fn(self.h, cb, None, _fp(a), _fp(b), bool(original))
        return seen

    def raycast_callbacks(self, start, direction, depth=None, original=False, origco=False):
        """The stroke-location ray cast through the reference's callback API: the hit
        callback tests a leaf while its tmin is below the bound and lowers the bound to
        the hit depth.  origco=True passes the leaf's current positions as original
        coordinates (same result).  Returns (hit, depth, active vertex, face, normal)."""
        L = lib()
        a = np.ascontiguousarray(start, np.float32)
        b = np.ascontiguousarray(direction, np.float32)
        pre = IsectRayPrecalc()
        L.sg_isect_ray_tri_precalc(C.byref(pre), _fp(b))
        st = {"hit": False, "depth": C.c_float(FLT_MAX if depth is None else depth), "av": C.c_int(-1),
              "af": C.c_int(-1), "fn": (C.c_float * 3)()}
        pos = self.positions() if origco else None

        def hit_cb(node, data, tmin):
            if L.BKE_pbvh_node_get_tmin(node) < tmin[0]:
                oc = None
                if origco:
                    verts, _ = self.node_verts(L.sg_pbvh_node_index(self.h, node))
                    buf = np.ascontiguousarray(pos[verts], np.float32)
                    st.setdefault("bufs", []).append(buf)
                    oc = buf.ctypes.data
                if L.BKE_pbvh_node_raycast(self.h, node, oc, False, _fp(a), _fp(b), C.byref(pre),
                                           C.byref(st["depth"]), C.byref(st["av"]), C.byref(st["af"]), st["fn"]):
                    st["hit"] = True
                    tmin[0] = st["depth"].value

        L.BKE_pbvh_raycast(self.h, OccludedCB(hit_cb), None, _fp(a), _fp(b), bool(original))
        return st["hit"], st["depth"].value, st["av"].value, st["af"].value, np.array(st["fn"][:], np.float32)

    def nearest_callbacks(self, start, direction, original=False):
        """BKE_pbvh_find_nearest_to_ray with the nearest callback: (hit, depth, dist_sq)."""
        L = lib()
        a = np.ascontiguousarray(start, np.float32)
        b = np.ascontiguousarray(direction, np.float32)
        st = {"hit": False, "depth": C.c_float(FLT_MAX), "dsq": C.c_float(FLT_MAX)}

        def near_cb(node, data, tmin):
            if L.BKE_pbvh_node_get_tmin(node) < tmin[0]:
                if L.BKE_pbvh_node_find_nearest_to_ray(self.h, node, None, False, _fp(a), _fp(b),
                                                       C.byref(st["depth"]), C.byref(st["dsq"])):
                    st["hit"] = True
                    tmin[0] = st["dsq"].value

        L.BKE_pbvh_find_nearest_to_ray(self.h, OccludedCB(near_cb), None, _fp(a), _fp(b), bool(original))
        return st["hit"], st["depth"].value, st["dsq"].value

    def project_ray_root(self, start, end, direction, original=False):
        a, e, d = (np.array(x, np.float32) for x in (start, end, direction))
        lib().BKE_pbvh_raycast_project_ray_root(self.h, bool(original), _fp(a), _fp(e), _fp(d))
        return a, e

    # -- the drop-in update API (what the reference's callers run after host edits)
    def mverts(self):
        """The MVert array BKE_pbvh_get_verts returns, as an (V, 3) float32 view of its
        coordinates (writable: edits go through it, as the reference's callers do)."""
        ptr = lib().BKE_pbvh_get_verts(self.h)
        raw = np.ctypeslib.as_array((C.c_uint8 * (16 * self.totvert)).from_address(ptr))
        return raw.view(np.float32).reshape(self.totvert, 4)[:, :3]

    def vert_mark_update(self, v):
        lib().BKE_pbvh_vert_mark_update(self.h, int(v))

    def node_mark_update(self, index):
        lib().BKE_pbvh_node_mark_update(self.node(index))

    def node_flag(self, index):
        return lib().sg_pbvh_node_flag(self.node(index))

    def update_bounds(self, flag):
        lib().BKE_pbvh_update_bounds(self.h, int(flag))

    def update_normals(self):
        lib().BKE_pbvh_update_normals(self.h, None)

    def coords_apply(self, co):
        c = np.ascontiguousarray(co, np.float32)
        lib().BKE_pbvh_vert_coords_apply(self.h, c.ctypes.data, c.shape[0])

    def coords_alloc(self):
        """BKE_pbvh_vert_coords_alloc: a malloc'd (V, 3) copy, freed here after copying."""
        ptr = lib().BKE_pbvh_vert_coords_alloc(self.h)
        if not ptr:
            raise RuntimeError("BKE_pbvh_vert_coords_alloc failed: " + last_error())
        out = np.ctypeslib.as_array((C.c_float * (3 * self.totvert)).from_address(ptr)).reshape(-1, 3).copy()
        C.CDLL(None).free(C.c_void_p(ptr))
        return out

    def frustum_gather(self, planes, exclude=False):
        """BKE_pbvh_search_gather with the reference's frustum callbacks and a
        PBVHFrustumPlanes of the given (n, 4) planes: node indices, left to right."""
        L = lib()
        pl = np.ascontiguousarray(planes, np.float32)

        class Frustum(C.Structure):
            _fields_ = [("planes", C.POINTER(C.c_float)), ("num_planes", C.c_int)]

        fr = Frustum(_fp(pl), pl.shape[0])
        fn = L.BKE_pbvh_node_frustum_exclude_AABB if exclude else L.BKE_pbvh_node_frustum_contain_AABB
        cb = C.cast(fn, SearchCB)
        arr = C.POINTER(C.c_void_p)()
        tot = C.c_int()
        L.BKE_pbvh_search_gather(self.h, cb, C.byref(fr), C.byref(arr), C.byref(tot))
        res = [L.sg_pbvh_node_index(self.h, arr[i]) for i in range(tot.value)]
        if tot.value:
            C.CDLL(None).free(arr)
        return res

    def search_gather(self, callback=None):
        """BKE_pbvh_search_gather with a Python callback(node_index) -> bool."""
        L = lib()
        if callback is None:
            cb = SearchCB()
        else:
            cb = SearchCB(lambda node, data: bool(callback(L.sg_pbvh_node_index(self.h, node))))
        arr = C.POINTER(C.c_void_p)()
        tot = C.c_int()
        L.BKE_pbvh_search_gather(self.h, cb, None, C.byref(arr), C.byref(tot))
        res = [L.sg_pbvh_node_index(self.h, arr[i]) for i in range(tot.value)]
        if tot.value:
            C.CDLL(None).free(arr)
        return res


def fan_looptris(poly_off):
    """Fan triangulation of polygons given by corner offsets: (tri -> poly, the three
    loop indices of each tri)."""
    off = np.asarray(poly_off, np.int64)
    n = np.diff(off)
    tri_poly = np.repeat(np.arange(len(n)), np.maximum(n - 2, 0)).astype(np.uint32)
    loops = []
    for q in range(len(n)):
        for k in range(1, n[q] - 1):
            loops.append((off[q], off[q] + k, off[q] + k + 1))
    return tri_poly, np.array(loops, np.uint32).reshape(-1, 3)


def make_rays(starts, normals, depths=None):
    starts = np.asarray(starts, np.float32).reshape(-1, 3)
    normals = np.asarray(normals, np.float32).reshape(-1, 3)
    n = starts.shape[0]
    dep = np.full(n, FLT_MAX, np.float32) if depths is None else np.broadcast_to(np.asarray(depths, np.float32), (n,))
    rays = (SgRay * n)()
    a = np.frombuffer(rays, np.float32).reshape(n, 8)
    a[:, 0:3] = starts
    a[:, 3:6] = normals
    a[:, 6] = dep
    a[:, 7] = 0.0
    return rays


RAY_HIT_DTYPE = np.dtype([("hit", "<i4"), ("depth", "<f4"), ("location", "<f4", 3), ("face_normal", "<f4", 3),
                          ("active_vertex", "<i4"), ("active_face", "<i4"), ("looptri", "<i4"), ("node", "<i4"),
                          ("status", "<i4"), ("reserved", "<i4")])


def hits_to_numpy(hits):
    return np.frombuffer(hits, RAY_HIT_DTYPE).copy()


# ---------------------------------------------------------------------------
# many meshes, one dab per mesh per step
# ---------------------------------------------------------------------------

BATCH_TOTALS = ["gathered", "affected", "tests", "candidates", "deg_t", "smooth_nbr_reads", "shared", "area_verts",
                "smooth_records_k", "learned_records", "", "", "learned_accepted", "learned_rejected", "steps"]
BATCH_LAUNCH = ["search", "brush", "smooth", "post", "learned", "brush_threads", "brush_vpl"]


class Batch:
    """Owner of an SgBatch: many meshes in one device arena, stepped together (one dab
    per mesh per step, one captured hipGraph launch per step)."""

    def __init__(self, meshes, leaf_limit: int = 128):
        L = lib()
        self.M = len(meshes)
        cos = [np.ascontiguousarray(c, np.float32) for c, _ in meshes]
        trs = [np.ascontiguousarray(t, np.int32) for _, t in meshes]
        self.nverts = [c.shape[0] for c in cos]
        co_p = (C.POINTER(C.c_float) * self.M)(*[_fp(c) for c in cos])
        tr_p = (C.POINTER(C.c_int) * self.M)(*[_ip(t) for t in trs])
        nv = (C.c_int * self.M)(*self.nverts)
        nt = (C.c_int * self.M)(*[t.shape[0] for t in trs])
        self.h = L.sg_batch_new(self.M, co_p, tr_p, nv, nt, leaf_limit)
        if not self.h:
            raise RuntimeError("sg_batch_new failed: " + last_error())
        self._mlp = None
        self._rows = 0

    def close(self):
        if getattr(self, "h", None):
            lib().sg_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = (C.c_longlong * (6 + self.M))()
        _check(lib().sg_batch_info(self.h, out, 6 + self.M), "batch info")
        return {"meshes": out[0], "verts": out[1], "tris": out[2], "leaves": out[3], "cuts": out[4],
                "vbase": [out[5 + i] for i in range(self.M + 1)]}

    def _table(self, table):
        steps = len(table)
        arr = (SgDabParams * (steps * self.M))()
        for s, row in enumerate(table):
            assert len(row) == self.M
            for m, d in enumerate(row):
                if d is not None:
                    arr[s * self.M + m] = d
        return arr, steps

    def load(self, table):
        """table[step][mesh] = SgDabParams or None (no dab)."""
        arr, steps = self._table(table)
        _check(lib().sg_batch_load(self.h, arr, steps), "sg_batch_load")
        self._rows = steps

    def launch(self, steps=None, graph=True, sync=True):
        steps = self._rows if steps is None else steps
        flags = (0 if graph else 1) | (0 if sync else 2)
        _check(lib().sg_batch_launch(self.h, steps, flags), "sg_batch_launch")

    def rewind(self):
        _check(lib().sg_batch_rewind(self.h), "rewind")

    def run(self, table, graph=True):
        self.load(table)
        self.launch(graph=graph)

    def sync(self):
        _check(lib().sg_batch_sync(self.h), "sync")

    def totals(self, reset=False):
        out = (C.c_longlong * 16)()
        _check(lib().sg_batch_totals(self.h, out, 1 if reset else 0), "totals")
        return {k: int(out[i]) for i, k in enumerate(BATCH_TOTALS) if k}

    def positions(self, m):
        out = np.empty((self.nverts[m], 3), np.float32)
        _check(lib().sg_batch_read_positions(self.h, m, _fp(out)), "read positions")
        return out

    def normals(self, m):
        out = np.empty((self.nverts[m], 3), np.float32)
        _check(lib().sg_batch_read_normals(self.h, m, _fp(out)), "read normals")
        return out

    def snapshot(self):
        _check(lib().sg_batch_snapshot_save(self.h), "snapshot")

    def restore(self):
        _check(lib().sg_batch_snapshot_restore(self.h), "restore")

    def set_launch(self, **kw):
        g = (C.c_int * 7)(*[int(kw.get(k, 0)) for k in BATCH_LAUNCH])
        _check(lib().sg_batch_set_launch(self.h, g), "set launch")

    def launch_shape(self):
        g = (C.c_int * 7)()
        _check(lib().sg_batch_get_launch(self.h, g), "get launch")
        return {k: int(g[i]) for i, k in enumerate(BATCH_LAUNCH)}

    def autotune(self, table):
        arr, steps = self._table(table)
        out = (C.c_double * 2)()
        _check(lib().sg_batch_autotune(self.h, arr, steps, out), "batch autotune")
        self._rows = 0  # the tuner loaded its own sample
        return float(out[0]), float(out[1])

    def set_persistent(self, on=True):
        """Persistent step (a workgroup per mesh, every row in one launch) on or off;
        returns whether the batch fits it."""
        return bool(_check(lib().sg_batch_set_persistent(self.h, 1 if on else 0), "set persistent"))

    def set_learned(self, mlp):
        self._mlp = mlp
        _check(lib().sg_batch_set_learned(self.h, mlp.h if mlp is not None else None), "set learned")

    def timer_start(self):
        _check(lib().sg_batch_timer_start(self.h), "timer start")

    def timer_stop(self):
        ms = C.c_float()
        _check(lib().sg_batch_timer_stop(self.h, C.byref(ms)), "timer stop")
        return ms.value
