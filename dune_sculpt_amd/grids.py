"""Multires grids (PBVH_GRIDS) through the C ABI of libsculpt_gpu.so.

`Grids` owns a SubdivCCG (sg_subdiv_ccg_new: grids of a coarse polygon mesh) and
`GridPBVH` a PBVH built over them (BKE_pbvh_build_grids + DUNE_pbvh_subdiv_cgg_set),
mirroring the reference's call sequence (kernel/intern/pbvh.c:598-643, :4973).
Element arrays are grid after grid, row-major (x fastest).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import STROKE_GRAPH, STROKE_LOG, PreparedStroke, SgDabParams, _check, _fp, _ip, last_error, lib

EXPORTS = [
    "sg_subdiv_ccg_new", "KERNEL_subdiv_ccg_destroy", "KERNEL_subdiv_ccg_key_top_level",
    "KERNEL_subdiv_ccg_recalc_normals", "KERNEL_subdiv_ccg_average_grids", "KERNEL_subdiv_ccg_update_normals",
    "KERNEL_subdiv_ccg_neighbor_coords_get", "sg_subdiv_ccg_neighbors_free", "KERNEL_subdiv_ccg_grid_to_face_index",
    "sg_subdiv_ccg_grids", "sg_subdiv_ccg_num_grids", "sg_subdiv_ccg_grid_faces", "sg_subdiv_ccg_faces",
    "sg_subdiv_ccg_flag_mats", "sg_subdiv_ccg_grid_hidden", "sg_subdiv_ccg_get_co", "sg_subdiv_ccg_set_co",
    "sg_subdiv_ccg_get_normals", "sg_subdiv_ccg_get_mask", "sg_subdiv_ccg_set_mask", "BKE_pbvh_build_grids",
    "DUNE_pbvh_subdiv_cgg_set", "BKE_pbvh_subdiv_cgg_set", "BKE_subdiv_ccg_key_top_level", "BKE_pbvh_node_get_grids", "BKE_pbvh_grid_hidden", "BKE_pbvh_get_grid_key",
    "BKE_pbvh_get_grids", "BKE_pbvh_get_grid_num_vertices", "BKE_pbvh_get_grid_num_faces",
    "BKE_pbvh_get_grid_updates", "BKE_pbvh_count_grid_quads", "sg_pbvh_grids_sync_host", "sg_pbvh_grids_upload",
]


class CCGKey(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("level", "elem_size", "grid_size", "grid_area", "grid_bytes",
                                        "normal_offset", "mask_offset", "has_normals", "has_mask")]


class DMFlagMat(C.Structure):
    _fields_ = [("mat_nr", C.c_short), ("flag", C.c_char)]


class SubdivCCGFace(C.Structure):
    _fields_ = [("num_grids", C.c_int), ("start_grid_index", C.c_int)]


class SubdivCCGCoord(C.Structure):
    _fields_ = [("grid_index", C.c_int), ("x", C.c_short), ("y", C.c_short)]


class SubdivCCGNeighbors(C.Structure):
    _fields_ = [("coords", C.POINTER(SubdivCCGCoord)), ("size", C.c_int), ("num_duplicates", C.c_int),
                ("coords_fixed", SubdivCCGCoord * 256)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    vp, i32, f32p, i32p = C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int)
    sig = {
        "sg_subdiv_ccg_new": (vp, [f32p, i32, i32p, i32p, i32, i32, i32, i32]),
        "KERNEL_subdiv_ccg_destroy": (None, [vp]),
        "KERNEL_subdiv_ccg_key_top_level": (None, [C.POINTER(CCGKey), vp]),
        "KERNEL_subdiv_ccg_recalc_normals": (None, [vp]),
        "KERNEL_subdiv_ccg_average_grids": (None, [vp]),
        "KERNEL_subdiv_ccg_update_normals": (None, [vp, C.POINTER(vp), i32]),
        "KERNEL_subdiv_ccg_neighbor_coords_get": (None, [vp, C.POINTER(SubdivCCGCoord), C.c_bool,
                                                         C.POINTER(SubdivCCGNeighbors)]),
        "sg_subdiv_ccg_neighbors_free": (None, [C.POINTER(SubdivCCGNeighbors)]),
        "KERNEL_subdiv_ccg_grid_to_face_index": (i32, [vp, i32]),
        "sg_subdiv_ccg_grids": (vp, [vp]),
        "sg_subdiv_ccg_num_grids": (i32, [vp]),
        "sg_subdiv_ccg_grid_faces": (vp, [vp]),
        "sg_subdiv_ccg_faces": (C.POINTER(SubdivCCGFace), [vp, i32p]),
        "sg_subdiv_ccg_flag_mats": (C.POINTER(DMFlagMat), [vp]),
        "sg_subdiv_ccg_grid_hidden": (C.POINTER(C.POINTER(C.c_uint32)), [vp]),
        "sg_subdiv_ccg_get_co": (i32, [vp, f32p]),
        "sg_subdiv_ccg_set_co": (i32, [vp, f32p]),
        "sg_subdiv_ccg_get_normals": (i32, [vp, f32p]),
        "sg_subdiv_ccg_get_mask": (i32, [vp, f32p]),
        "sg_subdiv_ccg_set_mask": (i32, [vp, f32p]),
        "BKE_pbvh_build_grids": (None, [vp, vp, i32, C.POINTER(CCGKey), vp, vp, vp]),
        "DUNE_pbvh_subdiv_cgg_set": (None, [vp, vp]),
        "BKE_pbvh_node_get_grids": (None, [vp, vp, C.POINTER(i32p), i32p, i32p, i32p, C.POINTER(vp)]),
        "BKE_pbvh_grid_hidden": (vp, [vp]),
        "BKE_pbvh_get_grid_key": (C.POINTER(CCGKey), [vp]),
        "BKE_pbvh_get_grids": (vp, [vp]),
        "BKE_pbvh_get_grid_num_vertices": (i32, [vp]),
        "BKE_pbvh_get_grid_num_faces": (i32, [vp]),
        "BKE_pbvh_get_grid_updates": (None, [vp, C.c_bool, C.POINTER(C.POINTER(vp)), i32p]),
        "BKE_pbvh_count_grid_quads": (i32, [vp, i32p, i32, i32]),
        "sg_pbvh_grids_sync_host": (i32, [vp]),
        "sg_pbvh_grids_upload": (i32, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _bound = True
    return L


def quad_sphere(n: int):
    """A cube with n x n quads per side projected to the unit sphere: (co, poly_off,
    poly_verts).  Valence-3 vertices at the cube's corners, 4 elsewhere."""
    idx = {}
    co = []

    def vid(p):
        key = tuple(int(round(v * 2 * n)) for v in p)
        if key not in idx:
            idx[key] = len(co)
            v = np.asarray(p, np.float64)
            co.append(v / np.linalg.norm(v))
        return idx[key]

    polys = []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            u_ax, v_ax = [a for a in range(3) if a != axis]
            for i in range(n):
                for j in range(n):
                    corners = []
                    for di, dj in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = [0.0, 0.0, 0.0]
                        p[axis] = sign
                        p[u_ax] = -1.0 + 2.0 * (i + di) / n
                        p[v_ax] = -1.0 + 2.0 * (j + dj) / n
                        corners.append(vid(p))
                    if sign < 0:  # outward winding on both sides
                        corners = corners[::-1]
                    if axis == 1:
                        corners = corners[::-1]
                    polys.append(corners)
    off = np.zeros(len(polys) + 1, np.int32)
    off[1:] = np.cumsum([len(p) for p in polys])
    return np.asarray(co, np.float32), off, np.asarray([v for p in polys for v in p], np.int32)


class Grids:
    """A SubdivCCG over a coarse polygon mesh (sg_subdiv_ccg_new)."""

    def __init__(self, co, poly_off, poly_verts, level: int, normals=True, mask=False):
        L = _bind()
        co = np.ascontiguousarray(co, np.float32)
        self.poly_off = np.ascontiguousarray(poly_off, np.int32)
        self.poly_verts = np.ascontiguousarray(poly_verts, np.int32)
        self.totvert = co.shape[0]
        self.h = L.sg_subdiv_ccg_new(_fp(co), self.totvert, _ip(self.poly_off), _ip(self.poly_verts),
                                     len(self.poly_off) - 1, level, 1 if normals else 0, 1 if mask else 0)
        if not self.h:
            raise RuntimeError("sg_subdiv_ccg_new failed: " + last_error())
        self.key = CCGKey()
        L.KERNEL_subdiv_ccg_key_top_level(C.byref(self.key), self.h)
        self.G = L.sg_subdiv_ccg_num_grids(self.h)
        self.gs = self.key.grid_size
        self.ga = self.key.grid_area
        self.E = self.G * self.ga

    def close(self):
        if getattr(self, "h", None):
            lib().KERNEL_subdiv_ccg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def co(self):
        out = np.empty((self.E, 3), np.float32)
        _check(lib().sg_subdiv_ccg_get_co(self.h, _fp(out)), "get_co")
        return out

    def set_co(self, co):
        c = np.ascontiguousarray(co, np.float32)
        _check(lib().sg_subdiv_ccg_set_co(self.h, _fp(c)), "set_co")

    def normals(self):
        out = np.empty((self.E, 3), np.float32)
        _check(lib().sg_subdiv_ccg_get_normals(self.h, _fp(out)), "get_normals")
        return out

    def mask(self):
        out = np.empty(self.E, np.float32)
        _check(lib().sg_subdiv_ccg_get_mask(self.h, _fp(out)), "get_mask")
        return out

    def set_mask(self, m):
        a = np.ascontiguousarray(m, np.float32)
        _check(lib().sg_subdiv_ccg_set_mask(self.h, _fp(a)), "set_mask")

    def recalc_normals(self):
        lib().KERNEL_subdiv_ccg_recalc_normals(self.h)

    def average(self):
        lib().KERNEL_subdiv_ccg_average_grids(self.h)

    def faces(self):
        n = C.c_int()
        ptr = lib().sg_subdiv_ccg_faces(self.h, C.byref(n))
        return ptr, n.value

    def update_normals(self, faces):
        """KERNEL_subdiv_ccg_update_normals on the host for face indices `faces`."""
        ptr, _ = self.faces()
        base = C.cast(ptr, C.c_void_p).value
        arr = (C.c_void_p * max(len(faces), 1))(*[base + 8 * int(f) for f in faces])
        lib().KERNEL_subdiv_ccg_update_normals(self.h, arr, len(faces))

    def neighbors(self, g, x, y, include_duplicates=False):
        nb = SubdivCCGNeighbors()
        c = SubdivCCGCoord(g, x, y)
        lib().KERNEL_subdiv_ccg_neighbor_coords_get(self.h, C.byref(c), include_duplicates, C.byref(nb))
        out = [(nb.coords[i].grid_index, nb.coords[i].x, nb.coords[i].y) for i in range(nb.size)]
        nd = nb.num_duplicates
        lib().sg_subdiv_ccg_neighbors_free(C.byref(nb))
        return out, nd

    def grid_to_face(self, g):
        return lib().KERNEL_subdiv_ccg_grid_to_face_index(self.h, g)

    def set_hidden(self, hidden):
        """hidden: bool per element -> the grids' hidden bitmaps."""
        hid = np.asarray(hidden, bool).reshape(self.G, self.ga)
        bm = lib().sg_subdiv_ccg_grid_hidden(self.h)
        words = (self.ga + 31) // 32
        weights = np.uint64(1) << np.arange(32, dtype=np.uint64)
        for g in range(self.G):
            bits = np.zeros(words * 32, np.uint64)
            bits[: self.ga] = hid[g]
            w32 = (bits.reshape(words, 32) * weights).sum(axis=1)  # bit i -> word i >> 5, bit i & 31
            for k in range(words):
                bm[g][k] = int(w32[k])

    def set_flag_mats(self, mat_nr, smooth):
        fm = lib().sg_subdiv_ccg_flag_mats(self.h)
        for g in range(self.G):
            fm[g].mat_nr = int(mat_nr[g])
            fm[g].flag = bytes([1 if smooth[g] else 0])


class GridPBVH:
    """A PBVH over the grids of `grids` (BKE_pbvh_build_grids + DUNE_pbvh_subdiv_cgg_set)."""

    def __init__(self, grids: Grids, leaf_limit: int = 10000, with_flagmats=True, with_hidden=True):
        L = _bind()
        self.grids = grids
        self.h = L.BKE_pbvh_new()
        if not self.h:
            raise RuntimeError("BKE_pbvh_new failed: " + last_error())
        _check(L.sg_pbvh_leaf_limit_set(self.h, leaf_limit), "leaf limit")
        L.BKE_pbvh_build_grids(self.h, L.sg_subdiv_ccg_grids(grids.h), grids.G, C.byref(grids.key),
                               L.sg_subdiv_ccg_grid_faces(grids.h),
                               C.cast(L.sg_subdiv_ccg_flag_mats(grids.h), C.c_void_p) if with_flagmats else None,
                               C.cast(L.sg_subdiv_ccg_grid_hidden(grids.h), C.c_void_p) if with_hidden else None)
        info = np.zeros(4, np.int32)
        L.sg_pbvh_tree_info(self.h, _ip(info))
        if info[0] == 0:
            raise RuntimeError("BKE_pbvh_build_grids failed: " + last_error())
        self.nodes = int(info[0])
        self.leaves = int(info[1])
        L.DUNE_pbvh_subdiv_cgg_set(self.h, grids.h)

    def close(self):
        for st in getattr(self, "_strokes", []):
            st.close()
        if getattr(self, "h", None):
            lib().BKE_pbvh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stroke(self, dabs, graph=True, log=False):
        arr = (SgDabParams * len(dabs))(*dabs)
        flags = (STROKE_GRAPH if graph else 0) | (STROKE_LOG if log else 0)
        _check(lib().sg_dab_apply_stroke(self.h, arr, len(dabs), flags), "sg_dab_apply_stroke")

    def prepare(self, dabs, log=False):
        """Capture the dab list into a hipGraph; returns a PreparedStroke."""
        arr = (SgDabParams * len(dabs))(*dabs)
        h = lib().sg_stroke_prepare(self.h, arr, len(dabs), STROKE_LOG if log else 0)
        if not h:
            raise RuntimeError("sg_stroke_prepare failed: " + last_error())
        st = PreparedStroke(h, self)
        self._strokes = getattr(self, "_strokes", []) + [st]
        return st

    def stroke_log(self, n):
        out = np.zeros((n, 24), np.int64)
        _check(lib().sg_dab_stroke_log(self.h, out.ctypes.data_as(C.POINTER(C.c_longlong)), n), "stroke log")
        return out

    def snapshot(self):
        _check(lib().sg_pbvh_snapshot_save(self.h), "snapshot")

    def restore(self):
        _check(lib().sg_pbvh_snapshot_restore(self.h), "restore")

    def timer_start(self):
        _check(lib().sg_pbvh_timer_start(self.h), "timer start")

    def timer_stop(self):
        ms = C.c_float()
        _check(lib().sg_pbvh_timer_stop(self.h, C.byref(ms)), "timer stop")
        return ms.value

    def dab(self, d: SgDabParams):
        _check(lib().sg_dab_apply(self.h, C.byref(d)), "sg_dab_apply")

    def sync_host(self):
        _check(lib().sg_pbvh_grids_sync_host(self.h), "grids sync")

    def upload(self):
        _check(lib().sg_pbvh_grids_upload(self.h), "grids upload")

    def state(self):
        """(co, normals, mask or None) of the device state, by grid id."""
        self.sync_host()
        return self.grids.co(), self.grids.normals(), (self.grids.mask() if self.grids.key.has_mask else None)

    def last_affected(self):
        out = np.empty(self.grids.E, np.int32)
        n = _check(lib().sg_dab_last_affected(self.h, _ip(out), out.size), "affected")
        return out[:n].copy()

    def last_gathered(self):
        out = np.empty(max(self.leaves, 1), np.int32)
        n = _check(lib().sg_dab_last_gathered(self.h, _ip(out), out.size), "gathered")
        return out[:n].copy()

    def last_stats(self):
        out = np.zeros(24, np.int64)
        _check(lib().sg_dab_last_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))), "stats")
        return out

    def node(self, i):
        return lib().sg_pbvh_node_at(self.h, i)

    def node_info(self, i):
        L = lib()
        nd = self.node(i)
        ptr = C.POINTER(C.c_int)()
        tot = C.c_int()
        L.BKE_pbvh_node_get_grids(self.h, nd, C.byref(ptr), C.byref(tot), None, None, None)
        grids = np.ctypeslib.as_array(ptr, shape=(tot.value,)).copy() if tot.value else np.zeros(0, np.int32)
        return {"flag": L.sg_pbvh_node_flag(nd), "grids": grids}

    def node_bb(self, i, original=False):
        mn = np.zeros(3, np.float32)
        mx = np.zeros(3, np.float32)
        fn = lib().BKE_pbvh_node_get_original_BB if original else lib().BKE_pbvh_node_get_BB
        fn(self.node(i), _fp(mn), _fp(mx))
        return np.concatenate([mn, mx])

    def boxes(self):
        return np.stack([self.node_bb(i) for i in range(self.nodes)]), \
            np.stack([self.node_bb(i, True) for i in range(self.nodes)])

    def prim_indices(self):
        out = np.empty(self.grids.G, np.int32)
        lib().sg_pbvh_prim_indices(self.h, _ip(out), out.size)
        return out

    def mark_update(self, i):
        lib().BKE_pbvh_node_mark_update(self.node(i))

    def update_normals(self):
        lib().BKE_pbvh_update_normals(self.h, self.grids.h)

    def update_bounds(self, flag):
        lib().BKE_pbvh_update_bounds(self.h, flag)

    def grid_updates(self, clear=True):
        arr = C.POINTER(C.c_void_p)()
        n = C.c_int()
        lib().BKE_pbvh_get_grid_updates(self.h, clear, C.byref(arr), C.byref(n))
        ptr, _ = self.grids.faces()
        base = C.cast(ptr, C.c_void_p).value
        out = [(arr[i] - base) // 8 for i in range(n.value)]
        if n.value:
            libc = C.CDLL(None)
            libc.free.argtypes = [C.c_void_p]
            libc.free(C.cast(arr, C.c_void_p))
        return out
