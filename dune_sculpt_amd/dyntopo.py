"""Dynamic topology (PBVH_BMESH) through the C ABI of libsculpt_gpu.so.

`DynMesh` owns an SgDynMesh (the stand-in for the reference's BMesh: a manifold
triangle mesh with room to grow) and `DynPBVH` a PBVH built over it with
DUNE_pbvh_build_bmesh (kernel/intern/pbvh_bmesh.c:1854-1929).  Vertices are followed
by their 64-bit labels (input index, or a hash for vertices a split creates); `state()`
gives the mesh sorted by label with faces as label triples, the form the oracle's
OracleDyn uses too.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check, _fp, _ip, last_error, lib

SUBDIVIDE, COLLAPSE = 1, 2
PBVH_UpdateTopology = 1 << 13

EXPORTS = [
    "sg_dynmesh_new", "sg_dynmesh_free", "sg_dynmesh_counts", "sg_dynmesh_read", "DUNE_pbvh_build_bmesh",
    "BKE_pbvh_build_bmesh", "DUNE_pbvh_bmesh_update_topology", "BKE_pbvh_bmesh_update_topology",
    "DUNE_pbvh_bmesh_after_stroke", "BKE_pbvh_bmesh_after_stroke", "DUNE_pbvh_bmesh_detail_size_set",
    "BKE_pbvh_bmesh_detail_size_set", "DUNE_pbvh_node_mark_topology_update", "BKE_pbvh_node_mark_topology_update",
    "sg_pbvh_bmesh_node_faces", "sg_pbvh_bmesh_node_unique_verts", "sg_pbvh_bmesh_node_other_verts",
    "sg_pbvh_bmesh_topology_set", "sg_pbvh_bmesh_last_stats", "sg_pbvh_bmesh_last_affected",
    "sg_pbvh_bmesh_node_info",
]

_bound = False


def _bind():
    global _bound
    L = lib()
    if _bound:
        return L
    vp, i32, f32p, i32p = C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int)
    u8p, u64p = C.POINTER(C.c_uint8), C.POINTER(C.c_uint64)
    sig = {
        "sg_dynmesh_new": (vp, [f32p, i32, i32p, i32, f32p, u8p, u8p, i32, i32]),
        "sg_dynmesh_free": (None, [vp]),
        "sg_dynmesh_counts": (i32, [vp, i32p, i32p, i32p, i32p]),
        "sg_dynmesh_read": (i32, [vp, u64p, f32p, f32p, f32p, i32p, u64p, i32p, f32p, u8p]),
        "DUNE_pbvh_build_bmesh": (None, [vp, vp, C.c_bool, vp, i32, i32]),
        "DUNE_pbvh_bmesh_update_topology": (C.c_bool, [vp, i32, f32p, f32p, C.c_float, C.c_bool, C.c_bool]),
        "DUNE_pbvh_bmesh_after_stroke": (None, [vp]),
        "DUNE_pbvh_bmesh_detail_size_set": (None, [vp, C.c_float]),
        "DUNE_pbvh_node_mark_topology_update": (None, [vp]),
        "sg_pbvh_bmesh_node_faces": (i32, [vp, i32p, i32]),
        "sg_pbvh_bmesh_node_unique_verts": (i32, [vp, i32p, i32]),
        "sg_pbvh_bmesh_node_other_verts": (i32, [vp, i32p, i32]),
        "sg_pbvh_bmesh_topology_set": (i32, [vp, i32, f32p, i32, i32]),
        "sg_pbvh_bmesh_last_stats": (i32, [vp, C.POINTER(C.c_longlong)]),
        "sg_pbvh_bmesh_last_affected": (i32, [vp, u64p, i32]),
        "sg_pbvh_bmesh_node_info": (i32, [vp, i32, i32p, f32p, f32p]),
    }
    for name, v in sig.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = v
    _bound = True
    return L


class DynMesh:
    """sg_dynmesh_new: co [V, 3], tris [T, 3] (outward winding); caps default 4x."""

    def __init__(self, co, tris, mask=None, vhide=None, fhide=None, vert_cap=0, tri_cap=0):
        L = _bind()
        co = np.ascontiguousarray(co, np.float32)
        tris = np.ascontiguousarray(tris, np.int32)
        self.has_mask = mask is not None
        m = None if mask is None else np.ascontiguousarray(mask, np.float32)
        vh = None if vhide is None else np.ascontiguousarray(vhide, np.uint8)
        fh = None if fhide is None else np.ascontiguousarray(fhide, np.uint8)
        u8 = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint8))
        self.h = L.sg_dynmesh_new(_fp(co), len(co), _ip(tris), len(tris), None if m is None else _fp(m), u8(vh),
                                  u8(fh), int(vert_cap), int(tri_cap))
        if not self.h:
            raise RuntimeError("sg_dynmesh_new: " + last_error())

    def close(self):
        if getattr(self, "h", None):
            lib().sg_dynmesh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def counts(self):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().sg_dynmesh_counts(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)), "counts")
        return {"verts": a.value, "tris": b.value, "vert_slots": c.value, "tri_slots": d.value}

    def raw(self):
        c = self.counts()
        nv, nt = c["verts"], c["tris"]
        lab = np.empty(nv, np.uint64)
        co = np.empty((nv, 3), np.float32)
        no = np.empty((nv, 3), np.float32)
        mk = np.empty(nv, np.float32)
        vn = np.empty(nv, np.int32)
        fl = np.empty((nt, 3), np.uint64)
        fn = np.empty(nt, np.int32)
        fno = np.empty((nt, 3), np.float32)
        fh = np.empty(nt, np.uint8)
        u64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint64))
        n = lib().sg_dynmesh_read(self.h, u64(lab), _fp(co), _fp(no), _fp(mk), _ip(vn), u64(fl), _ip(fn), _fp(fno),
                                  fh.ctypes.data_as(C.POINTER(C.c_uint8)))
        _check(n if n < 0 else 0, "read")
        return (lab, co, no, mk, vn), (fl, fn, fno, fh)


class DynPBVH:
    """BKE_pbvh_new + DUNE_pbvh_build_bmesh over a DynMesh (leaf limit 100, detail 0.75)."""

    def __init__(self, mesh: DynMesh):
        L = _bind()
        self.mesh = mesh
        self.h = L.BKE_pbvh_new()
        if not self.h:
            raise RuntimeError("BKE_pbvh_new: " + last_error())
        L.DUNE_pbvh_build_bmesh(self.h, mesh.h, True, None, -1, -1)
        if L.BKE_pbvh_type(self.h) != 2:
            raise RuntimeError("DUNE_pbvh_build_bmesh: " + last_error())

    def close(self):
        if getattr(self, "h", None):
            lib().BKE_pbvh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def detail_size_set(self, d):
        lib().DUNE_pbvh_bmesh_detail_size_set(self.h, float(d))

    def topology_set(self, mode=SUBDIVIDE | COLLAPSE, view_normal=(0.0, 0.0, 1.0), use_frontface=False,
                     use_projected=False):
        vn = np.ascontiguousarray(view_normal, np.float32)
        _check(lib().sg_pbvh_bmesh_topology_set(self.h, int(mode), _fp(vn), 1 if use_frontface else 0,
                                                 1 if use_projected else 0), "topology_set")

    def dab(self, d):
        _check(lib().sg_dab_apply(self.h, C.byref(d)), "dab")

    def last_stats(self):
        out = (C.c_longlong * 12)()
        _check(lib().sg_pbvh_bmesh_last_stats(self.h, out), "stats")
        return np.array(out[:], np.int64)

    def last_affected(self):
        n = int(self.last_stats()[1])
        out = np.empty(max(n, 1), np.uint64)
        k = lib().sg_pbvh_bmesh_last_affected(self.h, out.ctypes.data_as(C.POINTER(C.c_uint64)), out.size)
        _check(k if k < 0 else 0, "last_affected")
        return out[:k].copy()

    def last_gathered(self):
        n = int(self.last_stats()[0])
        out = np.empty(max(n, 1), np.int32)
        k = lib().sg_dab_last_gathered(self.h, _ip(out), out.size)
        _check(k if k < 0 else 0, "last_gathered")
        return out[:k].copy()

    def _nodes_count(self):
        n = 0
        info = np.zeros(3, np.int32)
        vb = np.zeros(6, np.float32)
        while lib().sg_pbvh_bmesh_node_info(self.h, n, _ip(info), _fp(vb), _fp(vb)) == 0:
            n += 1
        return n

    def nodes(self):
        n = self._nodes_count()
        info = np.zeros((n, 3), np.int32)
        vb = np.zeros((n, 6), np.float32)
        ob = np.zeros((n, 6), np.float32)
        for i in range(n):
            _check(lib().sg_pbvh_bmesh_node_info(self.h, i, _ip(info[i]), _fp(vb[i]), _fp(ob[i])), "node_info")
        return info, vb, ob

    def search(self, center, radius):
        """Leaves whose box meets the sphere (host tree; DFS order)."""
        info, vb, _ = self.nodes()
        c = np.asarray(center, np.float32)
        r2 = np.float32(radius) * np.float32(radius)
        out, st = [], [0]
        while st:
            i = st.pop()
            if (info[i, 0] & 1) and (info[i, 0] & (1 << 9)):
                continue
            near = np.where(vb[i, :3] > c, vb[i, :3], np.where(vb[i, 3:] < c, vb[i, 3:], c)).astype(np.float32)
            t = (c - near).astype(np.float32)
            if not (np.float32(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) < r2):
                continue
            if info[i, 0] & 1:
                out.append(i)
                continue
            st += [info[i, 1] + 1, info[i, 1]]
        return np.array(out, np.int32)

    def node(self, i):
        """The PBVHNode handle of node i (host mirror)."""
        lib().sg_pbvh_bmesh_node_info(self.h, int(i), _ip(np.zeros(3, np.int32)), _fp(np.zeros(6, np.float32)),
                                      _fp(np.zeros(6, np.float32)))  # refresh the mirror
        return lib().sg_pbvh_node_at(self.h, int(i))

    def mark_topology(self, nodes):
        for i in nodes:
            lib().DUNE_pbvh_node_mark_topology_update(self.node(i))

    def mark_update(self, nodes):
        for i in nodes:
            lib().BKE_pbvh_node_mark_update(self.node(i))

    def update_normals(self):
        lib().BKE_pbvh_update_normals(self.h, None)

    def update_bounds(self, flag=(1 << 2) | (1 << 3) | (1 << 5)):
        lib().BKE_pbvh_update_bounds(self.h, int(flag))

    def node_lists(self, i):
        """(faces, unique vertex slots, other vertex slots) of node i."""
        L = lib()
        nd = self.node(i)
        out = []
        for fn in (L.sg_pbvh_bmesh_node_faces, L.sg_pbvh_bmesh_node_unique_verts, L.sg_pbvh_bmesh_node_other_verts):
            n = fn(nd, None, 0)
            _check(n if n < 0 else 0, "node lists")
            a = np.empty(max(n, 1), np.int32)
            fn(nd, _ip(a), a.size)
            out.append(a[:n].copy())
        return tuple(out)

    def update_topology(self, mode, center, radius):
        c = np.ascontiguousarray(center, np.float32)
        return bool(lib().DUNE_pbvh_bmesh_update_topology(self.h, int(mode), _fp(c), None, float(radius), False, False))

    def after_stroke(self):
        lib().DUNE_pbvh_bmesh_after_stroke(self.h)

    def state(self):
        v, f = self.mesh.raw()
        return canonical_state(v, f)


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
