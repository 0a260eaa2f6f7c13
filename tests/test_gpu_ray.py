"""GPU ray cast (sg_raycast) against the oracle's sequential BKE_pbvh_raycast.

Bar: bit-exact.  Hit flag, depth, hit location, face normal, active vertex, looptri
and leaf node of every ray equal the oracle's bits, on the same positions and tree
(the oracle takes the GPU's positions after a stroke, then refits its bounds).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32
FLT_MAX = float(np.finfo(np.float32).max)


def _unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(F32)


def _rays(seed, n, radius=1.0):
    """Rays from outside and inside a mesh of about `radius`: hits, misses, bounded
    depths, axis-aligned and subnormal directions (the ancestor-walk path)."""
    rng = np.random.default_rng(seed)
    starts, dirs, depths = [], [], []
    for k in range(n):
        if k % 5 == 0:
            s = rng.normal(size=3) * 0.3 * radius
        else:
            s = _unit(rng.normal(size=3)) * 2.5 * radius
        aim = _unit(rng.normal(size=3)) * rng.uniform(0.0, 1.4) * radius
        d = _unit(aim - s)
        if k % 11 == 0:
            d = np.zeros(3, F32)
            d[k % 3] = -1.0 if (k // 3) % 2 else 1.0
            s = np.array(s, F32)
            s[k % 3] = -2.5 * radius * d[k % 3]
        if k % 13 == 0:
            d = np.array(d, F32)
            d[(k + 1) % 3] = F32(1e-39) if k % 2 else F32(0.0)
        starts.append(s)
        dirs.append(d)
        depths.append(FLT_MAX if k % 4 else float(rng.uniform(0.5, 3.0) * radius))
    return np.array(starts, F32), np.array(dirs, F32), np.array(depths, F32)


def _compare(po_o, hits, starts, dirs, depths, original=False):
    nhit = 0
    for i in range(len(starts)):
        w = po_o.raycast(starts[i], dirs[i], float(depths[i]), original=original)
        g = hits[i]
        ctx = f"ray {i}: start {starts[i]} dir {dirs[i]} depth0 {depths[i]}"
        assert g["status"] == 0, ctx
        assert int(g["hit"]) == int(w.hit), ctx
        assert np.float32(g["depth"]).view(np.uint32) == np.float32(w.depth).view(np.uint32), ctx
        if w.hit:
            nhit += 1
            assert np.array_equal(g["location"].view(np.uint32), np.array(w.location, F32).view(np.uint32)), ctx
            assert np.array_equal(g["face_normal"].view(np.uint32), np.array(w.face_normal, F32).view(np.uint32)), ctx
            assert g["active_vertex"] == w.active_vertex, ctx
            assert g["looptri"] == w.tri and g["active_face"] == w.tri, ctx
            assert g["node"] == w.leaf, ctx
    return nhit


def test_raycast_after_stroke_bit_identical(sg, po):
    co, tris = sg.noisy_icosphere(5, seed=21, amplitude=0.03)
    p = sg.PBVH(co, tris, leaf_limit=32)
    o = po.Oracle(co, tris, leaf_limit=32)
    n = 16
    rng = np.random.default_rng(2)
    radii = rng.uniform(0.08, 0.2, n).astype(F32)
    centres = sg.great_circle_centres(7, radii)
    dabs = sg.stroke_config([1, 4, 2, 5] * 4, centres, radii, rng.uniform(0.3, 1.0, n).astype(F32))
    p.stroke(dabs)
    o.set_positions(p.positions())
    for i in range(o.totnode()):
        vb, ob = o.node_bb(i)
        assert np.array_equal(p.node_bb(i).view(np.uint32), vb.view(np.uint32)), i
    starts, dirs, depths = _rays(5, 400)
    print(f"[ray] {len(starts)} rays, leaf 32, after a 16-dab stroke", flush=True)
    hits = p.raycast(starts, dirs, depths)
    nhit = _compare(o, hits, starts, dirs, depths)
    assert nhit > 150
    hits = p.raycast(starts, dirs, depths, original=True)
    _compare(o, hits, starts, dirs, depths, original=True)
    st = p.raycast_stats(starts, dirs, depths)
    assert st["boxes"] > len(starts) and st["tris"] > 0
    p.close()
    o.close()


def test_raycast_hidden_faces(sg, po):
    co, tris = sg.icosphere(4)
    hidden = (co[:, 2] > 0.2) | (co[:, 0] < -0.6)
    p = sg.PBVH(co, tris, leaf_limit=16, hidden=hidden)
    o = po.Oracle(co, tris, leaf_limit=16, hidden=hidden.astype(np.uint8))
    starts, dirs, depths = _rays(9, 200)
    print("[ray] hidden faces", flush=True)
    nhit = _compare(o, p.raycast(starts, dirs, depths), starts, dirs, depths)
    assert nhit > 50
    p.close()
    o.close()


def test_raycast_coincident_copies_tie_break(sg, po):
    """Two coincident copies of a sphere: every hit ties in depth across leaves, and
    the reference's visit order (tmin, then left->right) picks the winner."""
    co1, tris1 = sg.icosphere(3)
    co = np.concatenate([co1, co1])
    tris = np.concatenate([tris1, tris1 + len(co1)])
    p = sg.PBVH(co, tris, leaf_limit=8)
    o = po.Oracle(co, tris, leaf_limit=8)
    starts, dirs, depths = _rays(13, 200)
    print("[ray] coincident copies", flush=True)
    nhit = _compare(o, p.raycast(starts, dirs, depths), starts, dirs, depths)
    assert nhit > 50
    p.close()
    o.close()


def test_raycast_grazing_many_leaves(sg, po):
    """Nearly tangent rays along the diagonal of two coincident copies of a bumpy plane
    pass thousands of leaf boxes (several rounds of the leaf pass) and tie across the
    copies."""
    n = 240
    h = n // 2
    ij = np.stack(np.meshgrid(np.arange(n) - h, np.arange(n) - h, indexing="ij"), -1).reshape(-1, 2)
    rng = np.random.default_rng(4)
    co1 = np.zeros((n * n, 3), F32)
    co1[:, 0] = ij[:, 0] / h
    co1[:, 1] = ij[:, 1] / h
    co1[:, 2] = rng.uniform(-0.01, 0.01, n * n)
    q = (np.arange(n - 1)[:, None] * n + np.arange(n - 1)[None, :]).reshape(-1)
    t1 = np.concatenate([np.stack([q, q + n, q + n + 1], 1), np.stack([q, q + n + 1, q + 1], 1)]).astype(np.int32)
    copies = 2
    co = np.concatenate([co1] * copies)
    tris = np.concatenate([t1 + k * len(co1) for k in range(copies)])
    p = sg.PBVH(co, tris, leaf_limit=8)
    o = po.Oracle(co, tris, leaf_limit=8)
    starts, dirs = [], []
    for k in range(32):
        ang = np.pi / 4 + (k % 4) * np.pi / 2 + rng.uniform(-0.05, 0.05)
        s = np.array([1.2 * np.cos(ang), 1.2 * np.sin(ang), rng.uniform(0.001, 0.006)], F32)
        tgt = np.array([-1.2 * np.cos(ang), -1.2 * np.sin(ang), -rng.uniform(0.001, 0.006)], F32)
        starts.append(s)
        dirs.append(_unit(tgt - s))
    starts, dirs = np.array(starts, F32), np.array(dirs, F32)
    depths = np.full(len(starts), FLT_MAX, F32)
    print("[ray] grazing rays over coincident bumpy planes", flush=True)
    st = p.raycast_stats(starts, dirs, depths)
    nhit = _compare(o, p.raycast(starts, dirs, depths), starts, dirs, depths)
    print(st, nhit, p.tree_info())
    assert st["boxes"] > len(starts) * 2048  # more than one round of leaves per ray on average
    assert nhit > 20
    p.close()
    o.close()


def test_raycast_device_pointers(sg, po):
    """sg_raycast_device with device buffers equals sg_raycast."""
    import ctypes as C

    co, tris = sg.noisy_icosphere(4, seed=2, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    starts, dirs, depths = _rays(17, 64)
    want = p.raycast(starts, dirs, depths)
    rays = sg.make_rays(starts, dirs, depths)
    hip = C.CDLL("libamdhip64.so")
    dr, dh = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dr), C.c_size_t(C.sizeof(rays))) == 0
    assert hip.hipMalloc(C.byref(dh), C.c_size_t(64 * C.sizeof(sg.SgRayHit))) == 0
    assert hip.hipMemcpy(dr, rays, C.c_size_t(C.sizeof(rays)), 1) == 0  # host -> device
    L = sg.lib()
    assert L.sg_raycast_device(p.h, dr, 64, 0, dh) == 0
    assert L.sg_sync(p.h) == 0
    out = (sg.SgRayHit * 64)()
    assert hip.hipMemcpy(out, dh, C.c_size_t(C.sizeof(out)), 2) == 0  # device -> host
    got = sg.hits_to_numpy(out)
    assert got.tobytes() == want.tobytes()
    hip.hipFree(dr)
    hip.hipFree(dh)
    p.close()


def _camera_rays(sg, n, seed, radii, miss_every=9):
    """A stroke seen from a camera: rays from 2.5 units out aimed at points of a great
    circle on the unit sphere (every miss_every-th aimed past the mesh)."""
    rng = np.random.default_rng(seed)
    centres = sg.great_circle_centres(seed, radii)
    starts, dirs = [], []
    for i, c in enumerate(centres):
        c = np.asarray(c, np.float64)
        s = 2.5 * c / np.linalg.norm(c) + rng.normal(size=3) * 0.2
        aim = c * (1.8 if (miss_every and i % miss_every == miss_every - 1) else 1.0)
        if miss_every and i % miss_every == miss_every - 1:
            aim = aim + np.cross(c, [0.0, 0.0, 1.0]) * 2.0
        starts.append(s)
        dirs.append(_unit(aim - s))
    return np.array(starts, F32), np.array(dirs, F32)


def _oracle_placed(sg, po, o, dabs, starts, dirs):
    """The oracle's stroke: ray i cast on its current state gives dab i's centre (and
    ray i+1 the Grab translation); a miss skips the dab.  Returns the placements."""
    out = np.zeros((len(dabs), 8), F32)
    for i, d in enumerate(dabs):
        h = o.raycast(starts[i], dirs[i])
        g = np.zeros(3, F32)
        if d.sculpt_tool == sg.TOOL_GRAB and i + 1 < len(dabs):
            h2 = o.raycast(starts[i + 1], dirs[i + 1])
            if h.hit and h2.hit:
                g = np.array(h2.location, F32) - np.array(h.location, F32)
        if h.hit:
            out[i, :3] = np.array(h.location, F32)
            out[i, 6] = 1.0
        out[i, 3:6] = g
        if h.hit:
            q = sg.make_dab(tool=d.sculpt_tool, location=out[i, :3], radius=d.radius, strength=d.strength,
                            smooth_iterations=d.smooth_iterations, grab_delta=g)
            o.dab(q)
    return out


@pytest.mark.parametrize("graph", [True, False])
def test_placed_stroke_matches_oracle(sg, po, graph):
    from parity_util import REL_TOL, n_bit_diff, rel_err

    co, tris = sg.noisy_icosphere(5, seed=8, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=32)
    o = po.Oracle(co, tris, leaf_limit=32)
    o.set_normals(p.normals())
    n = 36
    rng = np.random.default_rng(6)
    radii = rng.uniform(0.06, 0.15, n).astype(F32)
    starts, dirs = _camera_rays(sg, n, 6, radii)
    tools = [1, 4, 2, 5]
    dabs = [sg.make_dab(tool=tools[i % 4], location=(0, 0, 0), radius=float(radii[i]),
                        strength=float(rng.uniform(0.3, 1.0)), smooth_iterations=3) for i in range(n)]
    print(f"[ray] placed stroke of {n} dabs, graph={graph}", flush=True)
    p.stroke_placed(dabs, starts, dirs, graph=graph)
    got = p.placements(n)
    want = _oracle_placed(sg, po, o, dabs, starts, dirs)
    assert want[:, 6].sum() >= n - n // 9 - 1 and want[:, 6].sum() < n  # hits and misses both present
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    pp, op = p.positions(), o.positions()
    print("position bits differing:", n_bit_diff(pp, op))
    assert rel_err(pp, op) <= REL_TOL and rel_err(p.normals(), o.normals()) <= REL_TOL
    p.close()
    o.close()


def test_prepared_placed_stroke_replays(sg, po):
    """A captured placed stroke launched twice from a checkpoint gives the same bits as
    the direct call; the placements follow the mesh it runs on."""
    co, tris = sg.noisy_icosphere(4, seed=9, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    n = 20
    radii = np.full(n, 0.12, F32)
    starts, dirs = _camera_rays(sg, n, 3, radii, miss_every=0)
    dabs = [sg.make_dab(tool=[1, 2, 4, 5][i % 4], location=(0, 0, 0), radius=0.12, strength=0.6) for i in range(n)]
    p.snapshot()
    p.stroke_placed(dabs, starts, dirs)
    ref_pos, ref_pl = p.positions(), p.placements(n)
    st = p.prepare_placed(dabs, starts, dirs)
    for _ in range(2):
        p.restore()
        st.launch()
        assert np.array_equal(p.positions().view(np.uint32), ref_pos.view(np.uint32))
        assert np.array_equal(p.placements(n).view(np.uint32), ref_pl.view(np.uint32))
    st.launch()  # on the already sculpted mesh: same rays, new surface, new centres
    assert not np.array_equal(p.placements(n), ref_pl)
    p.close()


def _same_tmin(a, b):
    a, b = np.float32(a), np.float32(b)
    if np.isnan(a) or np.isnan(b):
        return bool(np.isnan(a) and np.isnan(b))  # NaN payloads differ between x86 and the GPU
    return a.view(np.uint32) == b.view(np.uint32)


def test_callback_api_matches_oracle(sg, po):
    """BKE_pbvh_raycast / BKE_pbvh_find_nearest_to_ray with the sculpt operator's
    callbacks (BKE_pbvh_node_get_tmin, BKE_pbvh_node_raycast,
    BKE_pbvh_node_find_nearest_to_ray): the visiting order and tmins, and the results,
    equal the oracle's; the ray cast also equals sg_raycast.  Rays along an axis whose
    start lies on leaf box faces give NaN tmins (placed as node_tree_insert places them)."""
    co, tris = sg.noisy_icosphere(4, seed=3, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=16)
    o = po.Oracle(co, tris, leaf_limit=16)
    starts, dirs, _ = _rays(23, 48)
    rng = np.random.default_rng(1)
    extra_s, extra_d = [], []
    for v in rng.choice(len(co), 16, replace=False):
        extra_s.append(np.array([co[v, 0], co[v, 1], 3.0], F32))
        extra_d.append(np.array([0.0, 0.0, -1.0], F32))
    starts = np.concatenate([starts, np.array(extra_s)])
    dirs = np.concatenate([dirs, np.array(extra_d)])
    batch = p.raycast(starts, dirs)
    n_nan = 0
    print("[ray] callback API", flush=True)
    for i in range(len(starts)):
        s, d = starts[i], dirs[i]
        for mode in (0, 1):
            got = p.occluded_order(s, d, mode=mode)
            wn, wt = o.occluded_order(s, d, mode=mode)
            assert [g[0] for g in got] == list(wn), (i, mode)
            assert all(_same_tmin(g[1], t) for g, t in zip(got, wt)), (i, mode)
            n_nan += int(np.isnan(wt).sum())
        hit, depth, av, af, fn = p.raycast_callbacks(s, d)
        w = o.raycast(s, d)
        assert hit == bool(w.hit) and np.float32(depth).view(np.uint32) == np.float32(w.depth).view(np.uint32), i
        assert hit == bool(batch[i]["hit"]) and np.float32(depth) == batch[i]["depth"]
        if hit:
            assert av == w.active_vertex and af == w.tri and av == batch[i]["active_vertex"]
            assert np.array_equal(fn.view(np.uint32), np.array(w.face_normal, F32).view(np.uint32))
            assert p.raycast_callbacks(s, d, origco=True)[:4] == (hit, depth, av, af)
        nh, ndep, ndsq = p.nearest_callbacks(s, d)
        wnear = o.nearest_to_ray(s, d)
        assert nh == bool(wnear.hit), i
        if nh:
            assert np.float32(ndsq).view(np.uint32) == np.float32(wnear.dist_sq).view(np.uint32), i
            assert np.float32(ndep).view(np.uint32) == np.float32(wnear.depth).view(np.uint32), i
    assert n_nan > 0
    for s, e, d in (([0.1, 0.2, 5.0], [0.1, 0.2, -5.0], [0, 0, -1.0]), ([3.0, 0.1, 0.2], [-3.0, 0.1, 0.2], [-1.0, 0, 0])):
        gs, ge = p.project_ray_root(s, e, d)
        moved, ws, we = o.project_ray_root(s, e, d)
        assert moved and np.array_equal(gs, ws) and np.array_equal(ge, we)
    p.close()
    o.close()
