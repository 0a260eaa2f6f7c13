"""CPU tests of the product's host side: the C ABI loads and exports every declared
symbol, the host PBVH build reproduces the oracle's tree exactly, and the synthetic
mesh generators have the closed-form counts."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from meshes import edges_of

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol(sg):
    header = (ROOT / "include" / "sculpt_gpu.h").read_text()
    body = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    names = set(re.findall(r"\b((?:BKE|DUNE|KERNEL|sg)_[A-Za-z0-9_]+)\s*\(", body))
    names |= set(re.findall(r"\(\*((?:BKE|DUNE|KERNEL|sg)_[A-Za-z0-9_]+)\(", body))
    assert len(names) > 40
    L = C.CDLL(str(sg.LIB_PATH))
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(sg.EXPORTS) <= names
    from dune_sculpt_amd import grids as gr

    assert set(gr.EXPORTS) <= names


def test_library_is_gfx950_hip(sg):
    data = sg.LIB_PATH.read_bytes()
    assert b"gfx950" in data
    assert b"k_brush" in data and b"k_search" in data


def _host_tree(sg, co, tris, leaf):
    p = sg.PBVH.__new__(sg.PBVH)
    try:
        sg.PBVH.__init__(p, co, tris, leaf_limit=leaf)
    except RuntimeError:
        pass
    return p


@pytest.mark.parametrize("level,leaf", [(2, 8), (4, 32), (6, 128)])
def test_host_tree_matches_oracle(sg, po, level, leaf):
    co, tris = sg.noisy_icosphere(level, seed=7, amplitude=0.02)
    p = _host_tree(sg, co, tris, leaf)
    o = po.Oracle(co, tris, leaf_limit=leaf)
    assert p.tree_info()["nodes"] == o.totnode()
    assert np.array_equal(p.prim_indices(), o.prim_indices())
    for i in range(o.totnode()):
        vb, ob = o.node_bb(i)
        assert np.array_equal(p.node_bb(i).view(np.uint32), vb.view(np.uint32))
        info = o.node_info(i)
        if info[0] & 1:
            pv, pu = p.node_verts(i)
            assert pu == info[4]
            assert np.array_equal(pv, o.node_verts(i))
    # gather with no callback returns every leaf, left to right
    assert p.search_gather(None) == list(o.leaves_dfs())
    p.close()


def test_host_tree_torus_and_materials(sg, po):
    co, tris = sg.noisy_torus(40, 12, seed=3, amplitude=0.01)
    mat = (np.arange(tris.shape[0]) // 97) % 3
    p = sg.PBVH.__new__(sg.PBVH)
    try:
        sg.PBVH.__init__(p, co, tris, leaf_limit=64, mat=mat)
    except RuntimeError:
        pass
    o = po.Oracle(co, tris, leaf_limit=64, mat=mat)
    assert p.tree_info()["nodes"] == o.totnode()
    assert np.array_equal(p.prim_indices(), o.prim_indices())
    for node in o.leaves_dfs():  # material split: one material per leaf
        off, tot = o.node_info(node)[2:4]
        assert len(set(mat[o.prim_indices()[off:off + tot]].tolist())) == 1


def test_host_search_callback(sg, po):
    co, tris = sg.icosphere(4)
    p = _host_tree(sg, co, tris, 32)
    o = po.Oracle(co, tris, leaf_limit=32)
    c, r = co[3], 0.3

    def cb(i):
        bb = p.node_bb(i)
        nearest = np.clip(c, bb[:3], bb[3:])
        t = (c - nearest).astype(np.float32)
        return float((t[0] * t[0] + t[1] * t[1]) + t[2] * t[2]) < float(np.float32(r) * np.float32(r))

    assert p.search_gather(cb) == list(o.search_sphere(c, r))


@pytest.mark.parametrize("level", [0, 1, 3, 5])
def test_icosphere_counts(sg, level):
    co, tris = sg.icosphere(level)
    V, T = 10 * 4 ** level + 2, 20 * 4 ** level
    assert co.shape == (V, 3) and tris.shape == (T, 3)
    assert np.abs(np.linalg.norm(co, axis=1) - 1).max() < 1e-6
    E = edges_of(tris)
    assert V - E.shape[0] + T == 2  # Euler
    val = np.bincount(E.reshape(-1), minlength=V)
    assert (val == 5).sum() == 12 and (val[val != 5] == 6).all()
    P = np.float64(co)[tris]  # outward winding
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    assert ((n * P.mean(1)).sum(1) > 0).all()


def test_torus_counts_and_noise(sg):
    co, tris = sg.torus(30, 9)
    assert co.shape == (270, 3) and tris.shape == (540, 3)
    E = edges_of(tris)
    assert co.shape[0] - E.shape[0] + tris.shape[0] == 0  # genus 1
    a, _ = sg.noisy_icosphere(3, seed=5, amplitude=0.02)
    b, _ = sg.noisy_icosphere(3, seed=5, amplitude=0.02)
    c, _ = sg.noisy_icosphere(3, seed=6, amplitude=0.02)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    r = np.linalg.norm(a, axis=1)
    assert 0.97 < r.min() and r.max() < 1.03


def test_stroke_spacing(sg):
    radii = np.full(50, 0.1)
    c = sg.great_circle_centres(3, radii)
    step = np.linalg.norm(np.diff(np.float64(c), axis=0), axis=1)
    assert np.allclose(step, 2 * np.sin(0.01 / 2), rtol=1e-5)
    assert np.allclose(np.linalg.norm(c, axis=1), 1.0, atol=1e-6)


def test_host_proxies_worklist(sg, po):
    """BKE_pbvh_node_add_proxy / _gather_proxies / _free_proxies on the host tree: zeroed
    buffers of one float[3] per uniq vertex, gathered in node order whatever order they
    were added in (the combine's worklist), none after a free."""
    co, tris = sg.noisy_icosphere(3, seed=2, amplitude=0.02)
    p = _host_tree(sg, co, tris, 16)
    o = po.Oracle(co, tris, leaf_limit=16)
    leaves = [int(x) for x in o.leaves_dfs()]
    picks = [leaves[5], leaves[1], leaves[9], leaves[1]]  # leaf 1 twice: two proxies
    for i in picks:
        a = p.add_proxy(i)
        assert a.shape == (o.node_info(i)[4], 3) and not a.any()
        a[:] = 1.0
    L = sg.lib()
    arr = sg.C.POINTER(sg.C.c_void_p)()
    tot = sg.C.c_int()
    L.BKE_pbvh_gather_proxies(p.h, sg.C.byref(arr), sg.C.byref(tot))
    got = [L.sg_pbvh_node_index(p.h, arr[k]) for k in range(tot.value)]
    sg.C.CDLL(None).free(arr)
    assert got == sorted(set(picks))
    for i in set(picks):
        L.BKE_pbvh_node_free_proxies(p.node(i))
    L.BKE_pbvh_gather_proxies(p.h, sg.C.byref(arr), sg.C.byref(tot))
    assert tot.value == 0
    p.close()


def test_frustum_search_host_tree(sg, po):
    """BKE_pbvh_search_gather with the frustum callbacks (pbvh.c:4507-4557) on the host
    tree: six axis-aligned planes bound lo <= x <= hi, so a box is outside when some
    bmax < lo or bmin > hi and inside when every bmin > lo and bmax < hi; boxes nest, so
    the search returns exactly the passing leaves, left to right."""
    co, tris = sg.noisy_icosphere(4, seed=2, amplitude=0.02)
    p = _host_tree(sg, co, tris, 32)
    o = po.Oracle(co, tris, leaf_limit=32)
    lo, hi = np.float32([-0.4, -0.1, -0.2]), np.float32([0.6, 0.7, 0.9])
    planes = np.array([[1, 0, 0, -lo[0]], [0, 1, 0, -lo[1]], [0, 0, 1, -lo[2]],
                       [-1, 0, 0, hi[0]], [0, -1, 0, hi[1]], [0, 0, -1, hi[2]]], np.float32)
    for exclude in (False, True):
        want = []
        for n in o.leaves_dfs():
            vb, _ = o.node_bb(n)
            outside = (vb[3:] < lo).any() or (vb[:3] > hi).any()
            inside = (vb[:3] > lo).all() and (vb[3:] < hi).all()
            if (not inside) if exclude else (not outside):
                want.append(int(n))
        got = p.frustum_gather(planes, exclude=exclude)
        assert got == want and 0 < len(got) < len(o.leaves_dfs())
    p.close()
    o.close()
