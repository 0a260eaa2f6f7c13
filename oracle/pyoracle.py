"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module.  It is the checker: the product (dune_sculpt_amd / libsculpt_gpu.so) never
imports, links or calls it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

TOOL_DRAW, TOOL_SMOOTH, TOOL_INFLATE, TOOL_GRAB = 1, 2, 4, 5


class OrcDab(C.Structure):
    _fields_ = [
        ("tool", C.c_int), ("falloff", C.c_int), ("flag", C.c_int), ("smooth_iters", C.c_int),
        ("center", C.c_float * 3), ("radius", C.c_float), ("strength", C.c_float), ("hardness", C.c_float),
        ("normal_radius_factor", C.c_float), ("grab_delta", C.c_float * 3), ("curve_table", C.POINTER(C.c_float)),
        ("elastic_type", C.c_int), ("poisson_ratio", C.c_float), ("surface_normal", C.c_float * 3),
    ]


class OrcRayHit(C.Structure):
    _fields_ = [("hit", C.c_int), ("depth", C.c_float), ("location", C.c_float * 3),
                ("face_normal", C.c_float * 3), ("tri", C.c_int), ("active_vertex", C.c_int), ("leaf", C.c_int)]


class OrcRayNearest(C.Structure):
    _fields_ = [("hit", C.c_int), ("depth", C.c_float), ("dist_sq", C.c_float), ("location", C.c_float * 3),
                ("tri", C.c_int), ("leaf", C.c_int)]


class OrcPolys(C.Structure):
    _fields_ = [("tri_poly", C.POINTER(C.c_int)), ("poly_off", C.POINTER(C.c_int)),
                ("poly_verts", C.POINTER(C.c_int)), ("poly_smooth", C.POINTER(C.c_uint8)),
                ("poly_mat", C.POINTER(C.c_int))]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    vp, i32, f32p, i32p = C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int)
    sig = {
        "orc_build": (vp, [f32p, i32, i32p, i32, i32p, C.POINTER(C.c_uint8), i32, i32, i32]),
        "orc_free": (None, [vp]),
        "orc_totnode": (i32, [vp]),
        "orc_totleaf": (i32, [vp]),
        "orc_node_info": (None, [vp, i32, i32p]),
        "orc_node_bb": (None, [vp, i32, f32p, f32p]),
        "orc_node_verts": (i32, [vp, i32, i32p, i32]),
        "orc_prim_indices": (None, [vp, i32p]),
        "orc_leaves_dfs": (i32, [vp, i32p, i32]),
        "orc_get_positions": (None, [vp, f32p]),
        "orc_get_normals": (None, [vp, f32p]),
        "orc_get_positions64": (None, [vp, C.POINTER(C.c_double)]),
        "orc_get_normals64": (None, [vp, C.POINTER(C.c_double)]),
        "orc_set_mask": (None, [vp, f32p]),
        "orc_set_normals": (None, [vp, f32p]),
        "orc_set_positions": (None, [vp, f32p]),
        "orc_combine_proxies": (None, [vp, i32p, i32, i32p, f32p, f32p]),
        "orc_dab_grab_orig": (i32, [vp, C.POINTER(OrcDab), f32p, i32p, i32]),
        "orc_dab_elastic": (i32, [vp, C.POINTER(OrcDab), f32p, i32p, i32]),
        "orc_kelvinlet": (None, [C.POINTER(OrcDab), f32p, f32p, f32p]),
        "orc_search_sphere": (i32, [vp, f32p, C.c_float, i32, i32p, i32]),
        "orc_dab": (i32, [vp, C.POINTER(OrcDab), i32p, i32]),
        "orc_falloff": (C.c_float, [C.POINTER(OrcDab), C.c_float]),
        "orc_tri_normal": (None, [f32p, f32p, f32p, f32p]),
        "orc_calc_initial_normals": (None, [f32p, i32, i32p, i32, f32p]),
        "orc_calc_initial_normals_poly": (None, [f32p, i32, i32p, i32p, i32, f32p]),
        "orc_poly_normal": (None, [f32p, i32p, i32, f32p]),
        "orc_set_polys": (None, [vp, i32p, i32p, i32p, i32]),
        "orc_set_vert": (None, [vp, i32, f32p]),
        "orc_vert_mark_update": (None, [vp, i32]),
        "orc_node_mark_update": (None, [vp, i32]),
        "orc_update_bounds": (None, [vp, i32]),
        "orc_update_normals": (None, [vp]),
        "orc_node_flag": (i32, [vp, i32]),
        "orc_vert_coords_apply": (None, [vp, f32p]),
        "orc_set_threads": (None, [i32]),
        "orc_get_threads": (i32, []),
        "orc_last_stats": (None, [vp, C.POINTER(C.c_int64)]),
        "orc_raycast": (i32, [vp, f32p, f32p, C.c_float, i32, C.POINTER(OrcRayHit), C.POINTER(C.c_int64)]),
        "orc_find_nearest_to_ray": (i32, [vp, f32p, f32p, i32, C.POINTER(OrcRayNearest)]),
        "orc_project_ray_root": (i32, [vp, i32, f32p, f32p, f32p]),
        "orc_occluded_order": (i32, [vp, f32p, f32p, i32, i32, i32p, f32p, i32]),
        "orc_draw_leaf": (i32, [vp, i32, C.POINTER(OrcPolys), f32p, i32p, i32, i32, C.POINTER(C.c_uint8), i32,
                                i32p, i32p]),
        "orc_face_set_color": (None, [i32, i32, C.POINTER(C.c_uint8)]),
        "orc_lookup3_final": (None, [C.POINTER(C.c_uint32)]),
        "orc_hash_int_01": (C.c_float, [C.c_uint32]),
        "orc_isect_ray_tri": (i32, [f32p, f32p, f32p, f32p, f32p, f32p]),
        "orc_isect_ray_box": (i32, [f32p, f32p, f32p, f32p, f32p]),
        "orc_dist_sq_ray_box": (C.c_float, [f32p, f32p, f32p, f32p, f32p]),
        "orc_dist_sq_ray_tri": (C.c_float, [f32p, f32p, f32p, f32p, f32p, f32p]),
        "orc_grids_new": (vp, [i32, i32p, i32p, i32, i32, f32p, f32p, f32p, C.POINTER(C.c_uint8), i32p, i32, i32]),
        "orc_grids_free": (None, [vp]),
        "orc_grids_totnode": (i32, [vp]),
        "orc_grids_node_info": (None, [vp, i32, i32p]),
        "orc_grids_node_bb": (None, [vp, i32, f32p, f32p]),
        "orc_grids_prim_indices": (None, [vp, i32p]),
        "orc_grids_get": (None, [vp, f32p, f32p, f32p]),
        "orc_grids_set_co": (None, [vp, f32p]),
        "orc_grids_last_stats": (None, [vp, C.POINTER(C.c_int64)]),
        "orc_grids_edge_count": (i32, [vp]),
        "orc_grids_count_quads": (i32, [vp, i32p, i32]),
        "orc_grids_neighbors": (i32, [vp, i32, i32, i32, i32, i32p, i32, i32p]),
        "orc_grids_average": (None, [vp]),
        "orc_grids_recalc_normals": (None, [vp]),
        "orc_grids_update_normals_faces": (None, [vp, i32p, i32]),
        "orc_grids_search": (i32, [vp, f32p, C.c_float, i32, i32p, i32]),
        "orc_grids_dab": (i32, [vp, C.POINTER(OrcDab), i32p, i32]),
        "orc_grids_mark": (None, [vp, i32p, i32, i32]),
        "orc_grids_update_normals": (None, [vp]),
        "orc_grids_update_bounds": (None, [vp]),
        "orc_dyn_new": (vp, [f32p, i32, i32p, i32, f32p, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), i32, i32, i32,
                             i32, C.c_char_p, i32]),
        "orc_dyn_free": (None, [vp]),
        "orc_dyn_detail_size_set": (None, [vp, C.c_float]),
        "orc_dyn_topology_set": (None, [vp, i32, f32p, i32, i32]),
        "orc_dyn_set_normals": (None, [vp, f32p]),
        "orc_dyn_search": (i32, [vp, f32p, C.c_float, i32p, i32]),
        "orc_dyn_update_topology": (i32, [vp, i32, f32p, C.c_float]),
        "orc_dyn_subdivide_sequential": (i32, [vp, f32p, C.c_float]),
        "orc_dyn_dab": (i32, [vp, C.POINTER(OrcDab), C.POINTER(C.c_uint64), i32]),
        "orc_dyn_after_stroke": (None, [vp]),
        "orc_dyn_mark": (None, [vp, i32p, i32, i32]),
        "orc_dyn_update_normals": (None, [vp]),
        "orc_dyn_update_bounds": (None, [vp]),
        "orc_dyn_counts": (None, [vp, i32p]),
        "orc_dyn_export_verts": (None, [vp, C.POINTER(C.c_uint64), f32p, f32p, f32p, i32p]),
        "orc_dyn_export_faces": (None, [vp, C.POINTER(C.c_uint64), i32p, f32p, C.POINTER(C.c_uint8)]),
        "orc_dyn_node": (None, [vp, i32, i32p, f32p, f32p]),
        "orc_dyn_last_stats": (None, [vp, C.POINTER(C.c_int64)]),
        "orc_dyn_check": (i32, [vp]),
    }
    for k, (r, a) in sig.items():
        f = getattr(L, k)
        f.restype = r
        f.argtypes = a
    _lib = L
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def to_orc(d) -> OrcDab:
    """Convert an SgDabParams-like object (same field meaning) into an OrcDab."""
    o = OrcDab()
    o.tool = d.sculpt_tool
    o.falloff = d.curve_preset
    o.flag = d.flag
    o.smooth_iters = d.smooth_iterations
    o.center[:] = list(d.location)
    o.radius = d.radius
    o.strength = d.strength
    o.hardness = d.hardness
    o.normal_radius_factor = d.normal_radius_factor
    o.grab_delta[:] = list(d.grab_delta)
    if d.curve_table:
        o.curve_table = d.curve_table
    o.elastic_type = getattr(d, "elastic_deform_type", 0)
    o.poisson_ratio = getattr(d, "elastic_deform_volume_preservation", 0.0)
    o.surface_normal[:] = list(getattr(d, "surface_normal", (0.0, 0.0, 1.0)))
    return o


class Oracle:
    def __init__(self, co, tris, leaf_limit=128, fp64=True, hidden=None, mat=None, respect_hide=True, polys=None):
        """polys (optional): (tri_poly, poly_off, poly_verts) -- the mesh's polygons, for
        the polygon normal rules; `tris` are then its looptris."""
        L = lib()
        co = np.ascontiguousarray(co, np.float32)
        tris = np.ascontiguousarray(tris, np.int32)
        self.totvert, self.tottri = co.shape[0], tris.shape[0]
        hp = None
        if hidden is not None:
            self._hide = np.ascontiguousarray(np.asarray(hidden, np.uint8))
            hp = self._hide.ctypes.data_as(C.POINTER(C.c_uint8))
        mp = None
        if mat is not None:
            self._mat = np.ascontiguousarray(np.asarray(mat, np.int32))
            mp = _ip(self._mat)
        self.h = L.orc_build(_fp(co), self.totvert, _ip(tris), self.tottri, mp, hp, leaf_limit,
                             1 if respect_hide else 0, 1 if fp64 else 0)
        if polys is not None:
            tp, off, pv = (np.ascontiguousarray(x, np.int32) for x in polys)
            L.orc_set_polys(self.h, _ip(tp), _ip(off), _ip(pv), len(off) - 1)

    def close(self):
        if self.h:
            lib().orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dab(self, d):
        o = to_orc(d)
        out = np.empty(self.totvert, np.int32)
        n = lib().orc_dab(self.h, C.byref(o), _ip(out), self.totvert)
        return out[:n].copy()

    def positions(self):
        out = np.empty((self.totvert, 3), np.float32)
        lib().orc_get_positions(self.h, _fp(out))
        return out

    def normals(self):
        out = np.empty((self.totvert, 3), np.float32)
        lib().orc_get_normals(self.h, _fp(out))
        return out

    def positions64(self):
        out = np.empty((self.totvert, 3), np.float64)
        lib().orc_get_positions64(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def normals64(self):
        out = np.empty((self.totvert, 3), np.float64)
        lib().orc_get_normals64(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def set_normals(self, no):
        self._no = np.ascontiguousarray(no, np.float32)
        lib().orc_set_normals(self.h, _fp(self._no))

    def set_positions(self, co):
        c = np.ascontiguousarray(co, np.float32)
        lib().orc_set_positions(self.h, _fp(c))

    def dab_grab_orig(self, d, orco):
        """Grab from the original coordinates orco; returns the written vertices."""
        od = to_orc(d)
        oc = np.ascontiguousarray(orco, np.float32)
        out = np.empty(self.totvert + 1, np.int32)
        n = lib().orc_dab_grab_orig(self.h, C.byref(od), _fp(oc), _ip(out), out.size)
        return out[:n].copy()

    def dab_elastic(self, d, orco):
        """Elastic deform from the original coordinates orco; returns the written vertices."""
        od = to_orc(d)
        oc = np.ascontiguousarray(orco, np.float32)
        out = np.empty(self.totvert + 1, np.int32)
        n = lib().orc_dab_elastic(self.h, C.byref(od), _fp(oc), _ip(out), out.size)
        return out[:n].copy()

    def combine_proxies(self, nodes, proxies, orco=None):
        """proxies[j]: list of (uniq_verts, 3) arrays for nodes[j] (added order)."""
        nd = np.ascontiguousarray(nodes, np.int32)
        npx = np.array([len(pl) for pl in proxies], np.int32)
        flat = [np.ascontiguousarray(a, np.float32).reshape(-1) for pl in proxies for a in pl]
        pk = np.concatenate(flat) if flat else np.zeros(1, np.float32)
        oc = None if orco is None else np.ascontiguousarray(orco, np.float32)
        lib().orc_combine_proxies(self.h, _ip(nd), len(nd), _ip(npx), _fp(pk), _fp(oc) if oc is not None else None)

    def set_mask(self, mask):
        if mask is None:
            lib().orc_set_mask(self.h, None)
        else:
            self._mask = np.ascontiguousarray(mask, np.float32)
            lib().orc_set_mask(self.h, _fp(self._mask))

    def totnode(self):
        return lib().orc_totnode(self.h)

    def totleaf(self):
        return lib().orc_totleaf(self.h)

    def node_info(self, i):
        out = np.zeros(6, np.int32)
        lib().orc_node_info(self.h, i, _ip(out))
        return out

    def node_bb(self, i):
        vb = np.zeros(6, np.float32)
        ob = np.zeros(6, np.float32)
        lib().orc_node_bb(self.h, i, _fp(vb), _fp(ob))
        return vb, ob

    def node_verts(self, i):
        cap = 4096
        out = np.empty(cap, np.int32)
        n = lib().orc_node_verts(self.h, i, _ip(out), cap)
        return out[:n].copy()

    def prim_indices(self):
        out = np.empty(self.tottri, np.int32)
        lib().orc_prim_indices(self.h, _ip(out))
        return out

    def leaves_dfs(self):
        out = np.empty(self.tottri + 1, np.int32)
        n = lib().orc_leaves_dfs(self.h, _ip(out), out.size)
        return out[:n].copy()

    def search_sphere(self, center, radius, use_orig=False):
        c = np.asarray(center, np.float32)
        out = np.empty(self.tottri + 1, np.int32)
        n = lib().orc_search_sphere(self.h, _fp(c), radius, 1 if use_orig else 0, _ip(out), out.size)
        return out[:n].copy()

    def raycast(self, start, direction, depth0=3.4028234663852886e38, original=False, stats=None):
        """One ray (ray.c): returns an OrcRayHit; stats (int64[4], optional) gets node
        box tests, leaves passing, leaves ray-tested, tris tested."""
        a = np.ascontiguousarray(start, np.float32)
        b = np.ascontiguousarray(direction, np.float32)
        h = OrcRayHit()
        sp = None if stats is None else stats.ctypes.data_as(C.POINTER(C.c_int64))
        lib().orc_raycast(self.h, _fp(a), _fp(b), depth0, 1 if original else 0, C.byref(h), sp)
        return h

    def nearest_to_ray(self, start, direction, original=False):
        a = np.ascontiguousarray(start, np.float32)
        b = np.ascontiguousarray(direction, np.float32)
        r = OrcRayNearest()
        lib().orc_find_nearest_to_ray(self.h, _fp(a), _fp(b), 1 if original else 0, C.byref(r))
        return r

    def occluded_order(self, start, direction, mode=0, original=False):
        """(node ids, tmins) in the occluded search's visiting order (0 ray cast, 1 nearest)."""
        a = np.ascontiguousarray(start, np.float32)
        b = np.ascontiguousarray(direction, np.float32)
        cap = self.totleaf()
        nodes = np.zeros(cap, np.int32)
        tm = np.zeros(cap, np.float32)
        n = lib().orc_occluded_order(self.h, _fp(a), _fp(b), 1 if original else 0, mode, _ip(nodes), _fp(tm), cap)
        return nodes[:n].copy(), tm[:n].copy()

    def draw_leaf(self, node, tri_poly, poly_off, poly_verts, poly_smooth, poly_mat, face_sets, mask=None, seed=0,
                  default=1):
        """The reference's draw records of one leaf: ((verts, 36) uint8, smooth, material)."""
        keep = [np.ascontiguousarray(x, np.int32) for x in (tri_poly, poly_off, poly_verts, poly_mat, face_sets)]
        sm = np.ascontiguousarray(poly_smooth, np.uint8)
        pl = OrcPolys(_ip(keep[0]), _ip(keep[1]), _ip(keep[2]), sm.ctypes.data_as(C.POINTER(C.c_uint8)), _ip(keep[3]))
        cap = 3 * self.tottri
        out = np.zeros((cap, 36), np.uint8)
        r_sm, r_mat = C.c_int(), C.c_int()
        mp = None if mask is None else _fp(np.ascontiguousarray(mask, np.float32))
        mkeep = None if mask is None else np.ascontiguousarray(mask, np.float32)
        if mkeep is not None:
            mp = _fp(mkeep)
        n = lib().orc_draw_leaf(self.h, node, C.byref(pl), mp, _ip(keep[4]), seed, default,
                                out.ctypes.data_as(C.POINTER(C.c_uint8)), cap, C.byref(r_sm), C.byref(r_mat))
        return out[:max(n, 0)].copy(), r_sm.value, r_mat.value

    def project_ray_root(self, start, end, direction, original=False):
        a = np.array(start, np.float32)
        e = np.array(end, np.float32)
        d = np.array(direction, np.float32)
        moved = lib().orc_project_ray_root(self.h, 1 if original else 0, _fp(a), _fp(e), _fp(d))
        return bool(moved), a, e

    # -- the drop-in update API
    def set_vert(self, v, co):
        lib().orc_set_vert(self.h, int(v), _fp(np.ascontiguousarray(co, np.float32)))

    def vert_mark_update(self, v):
        lib().orc_vert_mark_update(self.h, int(v))

    def node_mark_update(self, node):
        lib().orc_node_mark_update(self.h, int(node))

    def update_bounds(self, flag):
        lib().orc_update_bounds(self.h, int(flag))

    def update_normals(self):
        lib().orc_update_normals(self.h)

    def node_flag(self, node):
        return lib().orc_node_flag(self.h, int(node))

    def vert_coords_apply(self, co):
        lib().orc_vert_coords_apply(self.h, _fp(np.ascontiguousarray(co, np.float32)))

    def last_stats(self):
        out = np.zeros(8, np.int64)
        lib().orc_last_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_int64)))
        return out


def falloff(d, d2):
    o = to_orc(d)
    return lib().orc_falloff(C.byref(o), d2)


def poly_normal(co, corners):
    """BKE_mesh_calc_poly_normal of the polygon with the given corner ids."""
    c = np.ascontiguousarray(co, np.float32)
    k = np.ascontiguousarray(corners, np.int32)
    out = np.zeros(3, np.float32)
    lib().orc_poly_normal(_fp(c), _ip(k), len(k), _fp(out))
    return out


def initial_normals_poly(co, poly_off, poly_verts):
    c = np.ascontiguousarray(co, np.float32)
    off = np.ascontiguousarray(poly_off, np.int32)
    pv = np.ascontiguousarray(poly_verts, np.int32)
    out = np.zeros_like(c)
    lib().orc_calc_initial_normals_poly(_fp(c), c.shape[0], _ip(off), _ip(pv), len(off) - 1, _fp(out))
    return out


def tri_normal(a, b, c):
    a, b, c = (np.ascontiguousarray(x, np.float32) for x in (a, b, c))
    out = np.zeros(3, np.float32)
    lib().orc_tri_normal(_fp(a), _fp(b), _fp(c), _fp(out))
    return out


def initial_normals(co, tris):
    co = np.ascontiguousarray(co, np.float32)
    tris = np.ascontiguousarray(tris, np.int32)
    out = np.zeros_like(co)
    lib().orc_calc_initial_normals(_fp(co), co.shape[0], _ip(tris), tris.shape[0], _fp(out))
    return out


def set_threads(n):
    lib().orc_set_threads(n)


def _f3(v):
    return np.ascontiguousarray(v, np.float32)


def isect_ray_tri(start, direction, v0, v1, v2):
    lam = np.zeros(1, np.float32)
    ok = lib().orc_isect_ray_tri(_fp(_f3(start)), _fp(_f3(direction)), _fp(_f3(v0)), _fp(_f3(v1)),
                                 _fp(_f3(v2)), _fp(lam))
    return bool(ok), float(lam[0])


def isect_ray_box(start, direction, bmin, bmax):
    t = np.zeros(1, np.float32)
    ok = lib().orc_isect_ray_box(_fp(_f3(start)), _fp(_f3(direction)), _fp(_f3(bmin)), _fp(_f3(bmax)), _fp(t))
    return bool(ok), float(t[0])


def dist_sq_ray_box(start, direction, bmin, bmax):
    dep = np.zeros(1, np.float32)
    d = lib().orc_dist_sq_ray_box(_fp(_f3(start)), _fp(_f3(direction)), _fp(_f3(bmin)), _fp(_f3(bmax)), _fp(dep))
    return float(d), float(dep[0])


def dist_sq_ray_tri(start, direction, v0, v1, v2):
    dep = np.zeros(1, np.float32)
    d = lib().orc_dist_sq_ray_tri(_fp(_f3(start)), _fp(_f3(direction)), _fp(_f3(v0)), _fp(_f3(v1)),
                                  _fp(_f3(v2)), _fp(dep))
    return float(d), float(dep[0])


def kelvinlet(d, co, bstrength):
    """Displacement of one original coordinate co by the elastic deform dab d (before the
    mask), with the dab's bstrength given."""
    od = to_orc(d)
    c = np.ascontiguousarray(co, np.float32)
    b = np.array([bstrength], np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_kelvinlet(C.byref(od), _fp(c), _fp(b), _fp(out))
    return out


def lookup3_final(a, b, c):
    abc = (C.c_uint32 * 3)(a & 0xFFFFFFFF, b & 0xFFFFFFFF, c & 0xFFFFFFFF)
    lib().orc_lookup3_final(abc)
    return abc[0], abc[1], abc[2]


def hash_int_01(k):
    return float(lib().orc_hash_int_01(k & 0xFFFFFFFF))


def face_set_color(fs, seed):
    out = (C.c_uint8 * 3)()
    lib().orc_face_set_color(fs, seed, out)
    return bytes(out)


class OracleGrids:
    """The oracle's PBVH over multires grids (grids.c)."""

    def __init__(self, totvert, poly_off, poly_verts, grid_size, co, no=None, mask=None, hidden=None,
                 grid_mat=None, leaf_limit=10000, respect_hide=True):
        L = lib()
        self.poly_off = np.ascontiguousarray(poly_off, np.int32)
        self.poly_verts = np.ascontiguousarray(poly_verts, np.int32)
        self.gs = int(grid_size)
        self.ga = self.gs * self.gs
        self.G = int(self.poly_off[-1] - self.poly_off[0])
        self.E = self.G * self.ga
        f32 = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        co, no, mask = f32(co), f32(no), f32(mask)
        hid = None if hidden is None else np.ascontiguousarray(hidden, np.uint8)
        gm = None if grid_mat is None else np.ascontiguousarray(grid_mat, np.int32)
        fp = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))
        self.has_normal = no is not None
        self.has_mask = mask is not None
        self.h = L.orc_grids_new(int(totvert), self.poly_off.ctypes.data_as(C.POINTER(C.c_int)),
                                 self.poly_verts.ctypes.data_as(C.POINTER(C.c_int)), len(self.poly_off) - 1,
                                 self.gs, fp(co), fp(no), fp(mask),
                                 None if hid is None else hid.ctypes.data_as(C.POINTER(C.c_uint8)),
                                 None if gm is None else gm.ctypes.data_as(C.POINTER(C.c_int)), int(leaf_limit),
                                 1 if respect_hide else 0)

    def close(self):
        if getattr(self, "h", None):
            lib().orc_grids_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def totnode(self):
        return lib().orc_grids_totnode(self.h)

    def node_info(self, i):
        out = np.zeros(4, np.int32)
        lib().orc_grids_node_info(self.h, i, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def node_bb(self, i):
        vb = np.zeros(6, np.float32)
        ob = np.zeros(6, np.float32)
        lib().orc_grids_node_bb(self.h, i, vb.ctypes.data_as(C.POINTER(C.c_float)),
                                ob.ctypes.data_as(C.POINTER(C.c_float)))
        return vb, ob

    def boxes(self):
        b = [self.node_bb(i) for i in range(self.totnode())]
        return np.stack([x[0] for x in b]), np.stack([x[1] for x in b])

    def prim_indices(self):
        out = np.empty(self.G, np.int32)
        lib().orc_grids_prim_indices(self.h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def state(self):
        co = np.empty((self.E, 3), np.float32)
        no = np.empty((self.E, 3), np.float32)
        mk = np.empty(self.E, np.float32)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        lib().orc_grids_get(self.h, fp(co), fp(no), fp(mk))
        return co, no, (mk if self.has_mask else None)

    def set_co(self, co):
        c = np.ascontiguousarray(co, np.float32)
        lib().orc_grids_set_co(self.h, c.ctypes.data_as(C.POINTER(C.c_float)))

    def neighbors(self, g, x, y, include_duplicates=False):
        out = np.zeros(3 * 512, np.int32)
        nd = C.c_int()
        n = lib().orc_grids_neighbors(self.h, g, x, y, 1 if include_duplicates else 0,
                                      out.ctypes.data_as(C.POINTER(C.c_int)), 512, C.byref(nd))
        return [tuple(int(v) for v in out[3 * i: 3 * i + 3]) for i in range(n)], nd.value

    def average(self):
        lib().orc_grids_average(self.h)

    def recalc_normals(self):
        lib().orc_grids_recalc_normals(self.h)

    def update_normals_faces(self, faces):
        f = np.ascontiguousarray(faces, np.int32)
        lib().orc_grids_update_normals_faces(self.h, f.ctypes.data_as(C.POINTER(C.c_int)), f.size)

    def count_quads(self, grids):
        g = np.ascontiguousarray(grids, np.int32)
        return lib().orc_grids_count_quads(self.h, g.ctypes.data_as(C.POINTER(C.c_int)), g.size)

    def search(self, center, radius, use_orig=False):
        c = np.ascontiguousarray(center, np.float32)
        out = np.empty(self.totnode() + 1, np.int32)
        n = lib().orc_grids_search(self.h, c.ctypes.data_as(C.POINTER(C.c_float)), float(radius),
                                   1 if use_orig else 0, out.ctypes.data_as(C.POINTER(C.c_int)), out.size)
        return out[:n].copy()

    def dab(self, d):
        out = np.empty(self.E, np.int32)
        n = lib().orc_grids_dab(self.h, C.byref(d), out.ctypes.data_as(C.POINTER(C.c_int)), out.size)
        return out[:n].copy()

    def last_stats(self):
        out = np.zeros(4, np.int64)
        lib().orc_grids_last_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_int64)))
        return out

    def mark(self, nodes, flag):
        a = np.ascontiguousarray(nodes, np.int32)
        lib().orc_grids_mark(self.h, a.ctypes.data_as(C.POINTER(C.c_int)), a.size, int(flag))

    def update_normals(self):
        lib().orc_grids_update_normals(self.h)

    def update_bounds(self):
        lib().orc_grids_update_bounds(self.h)


SUBDIVIDE, COLLAPSE = 1, 2
UPDATE_TOPOLOGY = 1 << 13


class OracleDyn:
    """The oracle's dynamic-topology PBVH (dyntopo.c).  State is read by vertex label:
    `state()` gives live vertices sorted by label and faces as label triples rotated
    to start at their smallest label, sorted."""

    def __init__(self, co, tris, mask=None, vhide=None, fhide=None, vert_cap=None, tri_cap=None, leaf_limit=100,
                 respect_hide=True):
        L = lib()
        co = np.ascontiguousarray(co, np.float32)
        tris = np.ascontiguousarray(tris, np.int32)
        V, T = len(co), len(tris)
        self.has_mask = mask is not None
        m = None if mask is None else np.ascontiguousarray(mask, np.float32)
        vh = None if vhide is None else np.ascontiguousarray(vhide, np.uint8)
        fh = None if fhide is None else np.ascontiguousarray(fhide, np.uint8)
        u8 = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint8))
        err = C.create_string_buffer(256)
        self.h = L.orc_dyn_new(_fp(co), V, _ip(tris), T, None if m is None else _fp(m), u8(vh), u8(fh),
                               int(vert_cap or 4 * V), int(tri_cap or 4 * T), int(leaf_limit),
                               1 if respect_hide else 0, err, 256)
        if not self.h:
            raise RuntimeError("orc_dyn_new: " + err.value.decode())

    def close(self):
        if getattr(self, "h", None):
            lib().orc_dyn_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def detail_size_set(self, d):
        lib().orc_dyn_detail_size_set(self.h, float(d))

    def topology_set(self, mode=SUBDIVIDE | COLLAPSE, view_normal=(0.0, 0.0, 1.0), use_frontface=False,
                     use_projected=False):
        vn = np.ascontiguousarray(view_normal, np.float32)
        lib().orc_dyn_topology_set(self.h, int(mode), _fp(vn), 1 if use_frontface else 0, 1 if use_projected else 0)

    def counts(self):
        out = np.zeros(6, np.int32)
        lib().orc_dyn_counts(self.h, _ip(out))
        return {"verts": int(out[0]), "tris": int(out[1]), "nodes": int(out[2]), "vert_slots": int(out[3]),
                "tri_slots": int(out[4]), "slab": int(out[5])}

    def search(self, center, radius):
        c = np.ascontiguousarray(center, np.float32)
        n = self.counts()["nodes"]
        out = np.empty(n + 1, np.int32)
        k = lib().orc_dyn_search(self.h, _fp(c), float(radius), _ip(out), out.size)
        return out[:k].copy()

    def mark(self, nodes, flag):
        a = np.ascontiguousarray(nodes, np.int32)
        lib().orc_dyn_mark(self.h, _ip(a), a.size, int(flag))

    def update_topology(self, mode, center, radius):
        c = np.ascontiguousarray(center, np.float32)
        return bool(lib().orc_dyn_update_topology(self.h, int(mode), _fp(c), float(radius)))

    def subdivide_sequential(self, center, radius):
        c = np.ascontiguousarray(center, np.float32)
        return int(lib().orc_dyn_subdivide_sequential(self.h, _fp(c), float(radius)))

    def dab(self, d):
        cap = self.counts()["vert_slots"] + 1
        out = np.empty(cap, np.uint64)
        n = lib().orc_dyn_dab(self.h, C.byref(d), out.ctypes.data_as(C.POINTER(C.c_uint64)), cap)
        return out[:n].copy()

    def after_stroke(self):
        lib().orc_dyn_after_stroke(self.h)

    def update_normals(self):
        lib().orc_dyn_update_normals(self.h)

    def update_bounds(self):
        lib().orc_dyn_update_bounds(self.h)

    def last_stats(self):
        out = np.zeros(12, np.int64)
        lib().orc_dyn_last_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_int64)))
        return out

    def check(self):
        return int(lib().orc_dyn_check(self.h))

    def raw_verts(self):
        n = self.counts()["verts"]
        lab = np.empty(n, np.uint64)
        co = np.empty((n, 3), np.float32)
        no = np.empty((n, 3), np.float32)
        mk = np.empty(n, np.float32)
        node = np.empty(n, np.int32)
        lib().orc_dyn_export_verts(self.h, lab.ctypes.data_as(C.POINTER(C.c_uint64)), _fp(co), _fp(no), _fp(mk),
                                   _ip(node))
        return lab, co, no, mk, node

    def raw_faces(self):
        n = self.counts()["tris"]
        labs = np.empty((n, 3), np.uint64)
        node = np.empty(n, np.int32)
        fno = np.empty((n, 3), np.float32)
        hide = np.empty(n, np.uint8)
        lib().orc_dyn_export_faces(self.h, labs.ctypes.data_as(C.POINTER(C.c_uint64)), _ip(node), _fp(fno),
                                   hide.ctypes.data_as(C.POINTER(C.c_uint8)))
        return labs, node, fno, hide

    def state(self):
        return canonical_state(self.raw_verts(), self.raw_faces())

    def nodes(self):
        n = self.counts()["nodes"]
        info = np.zeros((n, 3), np.int32)
        vb = np.zeros((n, 6), np.float32)
        ob = np.zeros((n, 6), np.float32)
        for i in range(n):
            lib().orc_dyn_node(self.h, i, _ip(info[i]), _fp(vb[i]), _fp(ob[i]))
        return info, vb, ob


def canonical_state(verts, faces):
    """(vertex arrays sorted by label, faces rotated to their smallest label and sorted)."""
    lab, co, no, mk, node = verts
    o = np.argsort(lab, kind="stable")
    v = {"lab": lab[o], "co": co[o], "no": no[o], "mask": mk[o], "node": node[o]}
    labs, fnode, fno, hide = faces
    r = np.argmin(labs, axis=1)
    idx = (r[:, None] + np.arange(3)[None, :]) % 3
    rot = np.take_along_axis(labs, idx, axis=1)
    fo = np.lexsort((rot[:, 2], rot[:, 1], rot[:, 0]))
    f = {"labs": rot[fo], "node": fnode[fo], "fno": fno[fo], "hide": hide[fo]}
    return v, f
