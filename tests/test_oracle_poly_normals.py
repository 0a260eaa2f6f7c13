"""The oracle's polygon normal rules, pinned by closed forms and FP64 restatements.

- Face normal of a polygon (BKE_mesh_calc_poly_normal, mesh_evaluate.c:62-84): a
  triangle is normal_tri_v3, a quad the normalised cross product of its diagonals
  (normal_quad_v3, math_geom.cc:48-68), more corners Newell's method (:39-60), fewer
  than three +z.
- After a dab each marked vertex gets normalize(sum over its incident looptris of the
  looptri's polygon normal) (pbvh.c:1041-1058, 1083-1087): a quad's normal is added
  once per looptri of the quad that holds the vertex.
- Initial normals (mesh_normals.c:225-299): each polygon's Newell normal weighted by
  the angle at each of its corners; a zero sum falls back to the normalised position.
- Smooth's 1-ring on a polygon mesh is the polygon edges.
"""
import numpy as np
import pytest

F64 = np.float64


def _quad_grid(n, spacing=0.25, noise=0.0, seed=0):
    """(n x n) vertex grid with (n-1)^2 quads (counter-clockwise seen from +z)."""
    ij = np.stack(np.meshgrid(np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 2)
    co = np.zeros((n * n, 3), np.float32)
    co[:, :2] = (ij - (n - 1) / 2) * spacing
    if noise:
        co[:, 2] = np.random.default_rng(seed).uniform(-noise, noise, n * n)
    pv = []
    for i in range(n - 1):
        for j in range(n - 1):
            pv += [i * n + j, (i + 1) * n + j, (i + 1) * n + j + 1, i * n + j + 1]
    return co, np.arange(0, len(pv) + 1, 4, dtype=np.int32), np.array(pv, np.int32)


def _newell64(P):
    n = np.zeros(3)
    prev = P[-1]
    for cur in P:
        n += [(prev[1] - cur[1]) * (prev[2] + cur[2]), (prev[2] - cur[2]) * (prev[0] + cur[0]),
              (prev[0] - cur[0]) * (prev[1] + cur[1])]
        prev = cur
    return n / np.linalg.norm(n)


def _poly_normal64(co, corners):
    P = co[corners].astype(F64)
    if len(P) == 4:
        n = np.cross(P[0] - P[2], P[1] - P[3])
        return n / np.linalg.norm(n)
    if len(P) == 3:
        n = np.cross(P[0] - P[1], P[1] - P[2])
        return n / np.linalg.norm(n)
    return _newell64(P)


def test_poly_normal_closed_forms(po):
    co = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)], np.float32)
    assert po.poly_normal(co, [0, 1, 2, 3]).tolist() == [0.0, 0.0, 1.0]  # planar quad: the plane normal
    assert po.poly_normal(co, [3, 2, 1, 0]).tolist() == [0.0, 0.0, -1.0]
    # a non-planar quad: the diagonal cross, which differs from the mean of its two
    # triangles' normals
    co = np.array([(0, 0, 0), (1, 0, 0.5), (1, 1, 0), (0, 1, 0.25)], np.float32)
    got = po.poly_normal(co, [0, 1, 2, 3]).astype(F64)
    d = np.cross(F64([-1, -1, 0]), F64([1, -1, 0.25]))
    assert np.abs(got - d / np.linalg.norm(d)).max() < 1e-7
    t = po.tri_normal(co[0], co[1], co[2]).astype(F64) + po.tri_normal(co[0], co[2], co[3]).astype(F64)
    assert np.abs(got - t / np.linalg.norm(t)).max() > 1e-3
    # degenerate polygons
    assert po.poly_normal(co, [0, 1]).tolist() == [0.0, 0.0, 1.0]
    flat = np.zeros((5, 3), np.float32)
    assert po.poly_normal(flat, [0, 1, 2, 3, 4]).tolist() == [0.0, 0.0, 1.0]  # zero Newell sum -> +z
    # regular hexagon tilted out of its plane: Newell against FP64
    ang = np.arange(6) * np.pi / 3
    hexa = np.stack([np.cos(ang), np.sin(ang), 0.1 * np.cos(3 * ang)], -1).astype(np.float32)
    assert np.abs(po.poly_normal(hexa, range(6)) - _newell64(hexa.astype(F64))).max() < 1e-6


def _dab_normals_check(sg, po, co, off, pv, leaf):
    tri_poly, loops = sg.fan_looptris(off)
    tris = pv[loops]
    o = po.Oracle(co, tris, leaf_limit=leaf, polys=(tri_poly, off, pv))
    n0 = o.normals()
    aff = o.dab(sg.make_dab(tool=5, location=(0.1, -0.05, 0.0), radius=0.6, strength=1.0,
                            grab_delta=(0.0, 0.0, 0.2)))
    assert len(aff) > 10
    pos, nor = o.positions(), o.normals()
    want = []
    tri_normal_rule = []
    for v in aff:
        ts = np.nonzero((tris == v).any(1))[0]
        s = sum(_poly_normal64(pos, pv[off[tri_poly[t]]:off[tri_poly[t] + 1]]) for t in ts)
        want.append(s / np.linalg.norm(s))
        s2 = sum(np.cross(*(pos[tris[t][:2]].astype(F64) - pos[tris[t][1:]].astype(F64))) /
                 np.linalg.norm(np.cross(*(pos[tris[t][:2]].astype(F64) - pos[tris[t][1:]].astype(F64))))
                 for t in ts)
        tri_normal_rule.append(s2 / np.linalg.norm(s2))
    want = np.array(want)
    assert np.abs(nor[aff] - want).max() < 2e-6
    # untouched vertices keep their normals
    rest = np.setdiff1d(np.arange(co.shape[0]), aff)
    assert np.array_equal(nor[rest], n0[rest])
    return np.abs(np.array(tri_normal_rule) - want).max()


def test_update_normals_add_polygon_normal_per_looptri(sg, po):
    """A Grab dab on a noisy quad grid: every affected vertex's new normal equals the
    FP64 sum of its incident looptris' quad normals (a quad counted once per looptri
    holding the vertex); the per-triangle rule gives visibly different normals."""
    co, off, pv = _quad_grid(12, noise=0.05, seed=1)
    diff = _dab_normals_check(sg, po, co, off, pv, leaf=8)
    assert diff > 1e-3


def test_update_normals_ngons(sg, po):
    """Pentagons and triangles mixed: Newell for the pentagons."""
    # a strip of pentagons: quads with an extra vertex on the top edge's midpoint
    co, off4, pv4 = _quad_grid(6, noise=0.04, seed=2)
    extra = []
    off, pv = [0], []
    base = co.shape[0]
    for q in range(len(off4) - 1):
        c = pv4[off4[q]:off4[q + 1]]
        if q % 3 == 0:  # split into two triangles
            pv += [c[0], c[1], c[2]]
            off.append(len(pv))
            pv += [c[0], c[2], c[3]]
            off.append(len(pv))
            continue
        m = (co[c[2]] + co[c[3]]) * 0.5
        m[2] += 0.03
        extra.append(m)
        # the new vertex is shared only by this polygon (a T-junction is fine here)
        pv += [c[0], c[1], c[2], base + len(extra) - 1, c[3]]
        off.append(len(pv))
    co = np.concatenate([co, np.array(extra, np.float32)])
    _dab_normals_check(sg, po, co, np.array(off, np.int32), np.array(pv, np.int32), leaf=8)


def _initial64(co, off, pv):
    out = np.zeros((co.shape[0], 3))
    for p in range(len(off) - 1):
        c = pv[off[p]:off[p + 1]]
        P = co[c].astype(F64)
        n = _newell64(P)
        k = len(c)
        for i in range(k):
            e_prev = P[i - 1] - P[i]
            e_next = P[i] - P[(i + 1) % k]
            e_prev /= np.linalg.norm(e_prev)
            e_next /= np.linalg.norm(e_next)
            out[c[i]] += n * np.arccos(np.clip(-e_prev @ e_next, -1, 1))
    return out / np.linalg.norm(out, axis=1, keepdims=True)


@pytest.mark.parametrize("noise", [0.0, 0.05])
def test_initial_normals_polygon_corner_angles(po, noise):
    co, off, pv = _quad_grid(9, noise=noise, seed=3)
    got = po.initial_normals_poly(co, off, pv)
    want = _initial64(co, off, pv)
    assert np.abs(got - want).max() < 2e-6
    if noise:
        # the same mesh as triangles gives different normals (angles and normals per tri)
        tris = np.concatenate([pv.reshape(-1, 4)[:, [0, 1, 2]], pv.reshape(-1, 4)[:, [0, 2, 3]]])
        assert np.abs(po.initial_normals(co, tris) - got).max() > 1e-3
    else:
        assert np.abs(got - F64([0, 0, 1])).max() < 1e-7


def test_smooth_ring_is_polygon_edges(sg, po):
    """Smooth's 1-ring on a polygon mesh is the polygon edges (vertex -> edge map,
    mesh_mapping.c:297), not the fan triangulation's diagonals.  Grid columns x = 0, 1, 2
    and rows y = 0, 1, 3: the centre vertex's edge neighbours average to (1, 1.25) exactly;
    Smooth with factor 1 (constant falloff, strength 1) moves it there."""
    xs, ys = [0.0, 1.0, 2.0], [0.0, 1.0, 3.0]
    co = np.array([(x, y, 0.0) for x in xs for y in ys], np.float32)  # id = 3 i + j
    pv = []
    for i in range(2):
        for j in range(2):
            pv += [3 * i + j, 3 * (i + 1) + j, 3 * (i + 1) + j + 1, 3 * i + j + 1]
    off = np.arange(0, len(pv) + 1, 4, dtype=np.int32)
    pv = np.array(pv, np.int32)
    tri_poly, loops = sg.fan_looptris(off)
    o = po.Oracle(co, pv[loops], leaf_limit=8, polys=(tri_poly, off, pv))
    got = o.dab(sg.make_dab(tool=2, location=(1.0, 1.0, 0.0), radius=0.5, strength=1.0, falloff=8,
                            smooth_iterations=1))
    assert got.tolist() == [4]
    assert o.positions()[4].tolist() == [1.0, 1.25, 0.0]
    # the same mesh given as its triangles only: the diagonals join the ring
    o2 = po.Oracle(co, pv[loops], leaf_limit=8)
    o2.dab(sg.make_dab(tool=2, location=(1.0, 1.0, 0.0), radius=0.5, strength=1.0, falloff=8, smooth_iterations=1))
    assert o2.positions()[4].tolist() != [1.0, 1.25, 0.0]
