"""Known-answer tests that pin the oracle's ray cast (oracle/ray.c).

The reference has no tests for its ray functions, so each piece is pinned against an
independent FP64 computation: Moller-Trumbore for the watertight ray/triangle test, a
line/slab test for the ray/box test, convex minimisation for the ray/box and ray/edge
distances, brute force over every triangle for the traversal (visit order, skip rule,
depth bound, hidden faces), and the analytic sphere for the hit points.  A dropped term
or flipped comparison in any piece moves a hit or a depth and fails one of these.
"""
import numpy as np
import pytest
from scipy.optimize import minimize_scalar

F32 = np.float32
FLT_MAX = float(np.finfo(np.float32).max)


def _mt64(o, d, v0, v1, v2):
    """Moller-Trumbore in FP64: barycentrics (u, v, w) and the line parameter t."""
    o, d, v0, v1, v2 = (np.asarray(x, np.float64) for x in (o, d, v0, v1, v2))
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-12:
        return None
    s = o - v0
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    t = (e2 @ q) / det
    return u, v, 1.0 - u - v, t


def _unit(rng, n):
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(F32)


def test_ray_tri_against_fp64(po):
    rng = np.random.default_rng(3)
    checked = hits = 0
    for _ in range(4000):
        tri = rng.normal(size=(3, 3)).astype(F32)
        o = (rng.normal(size=3) * 2).astype(F32)
        # aim near the triangle so about half the rays hit
        target = tri.mean(0) + rng.normal(size=3) * 0.5
        d = (target - o) / np.linalg.norm(target - o)
        d = d.astype(F32)
        if rng.random() < 0.2:
            d = -d  # behind the origin: a line hit at t < 0 is a miss
        r = _mt64(o, d, *tri)
        if r is None:
            continue
        u, v, w, t = r
        if min(abs(u), abs(v), abs(w)) < 1e-5 or abs(t) < 1e-6:
            continue  # too close to an edge or the origin for an FP32 verdict
        want = u >= 0 and v >= 0 and w >= 0 and t >= 0
        ok, lam = po.isect_ray_tri(o, d, *tri)
        assert ok == want
        # two-sided: the other winding gives the same answer and depth
        ok2, lam2 = po.isect_ray_tri(o, d, tri[0], tri[2], tri[1])
        assert ok2 == want
        if want:
            assert abs(lam - t) <= 1e-5 * (1 + abs(t))
            assert abs(lam2 - lam) <= 1e-6 * (1 + abs(lam))  # same hit, other summation order
            hits += 1
        checked += 1
    assert checked > 3000 and hits > 800


def test_ray_tri_closed_forms(po):
    tri = np.array([[0, 0, 0.5], [1, 0, 0.5], [0, 1, 0.5]], F32)
    # straight down onto the plane z = 0.5 from z = 2.25: depth 1.75, exact in FP32
    ok, lam = po.isect_ray_tri([0.25, 0.25, 2.25], [0, 0, -1], *tri)
    assert ok and lam == F32(1.75)
    ok, lam = po.isect_ray_tri([0.25, 0.25, -1.0], [0, 0, 1], *tri)
    assert ok and lam == F32(1.5)
    assert not po.isect_ray_tri([0.75, 0.75, 2.0], [0, 0, -1], *tri)[0]  # outside the hypotenuse
    assert not po.isect_ray_tri([0.25, 0.25, 2.0], [0, 0, 1], *tri)[0]   # plane behind the origin
    assert not po.isect_ray_tri([0.25, 0.25, 2.0], [1, 0, 0], *tri)[0]   # parallel to the plane
    # a ray through a shared edge hits both triangles (watertight)
    t2 = np.array([[1, 0, 0.5], [1, 1, 0.5], [0, 1, 0.5]], F32)
    assert po.isect_ray_tri([0.5, 0.5, 2.0], [0, 0, -1], *tri)[0]
    assert po.isect_ray_tri([0.5, 0.5, 2.0], [0, 0, -1], *t2)[0]


def _slab64(o, d, bmin, bmax):
    lo, hi = -np.inf, np.inf
    for i in range(3):
        if d[i] == 0:
            if o[i] < bmin[i] or o[i] > bmax[i]:
                return None
            continue
        t1, t2 = (bmin[i] - o[i]) / d[i], (bmax[i] - o[i]) / d[i]
        lo, hi = max(lo, min(t1, t2)), min(hi, max(t1, t2))
    return lo, hi


def test_ray_box_against_fp64(po):
    rng = np.random.default_rng(5)
    n = hits = 0
    for _ in range(4000):
        c, h = rng.normal(size=3), rng.uniform(0.05, 1.0, 3)
        bmin, bmax = (c - h).astype(F32), (c + h).astype(F32)
        o = (rng.normal(size=3) * 3).astype(F32)
        aim = c + rng.normal(size=3) * 0.8 - o  # about half of the lines meet the box
        d = (aim / np.linalg.norm(aim)).astype(F32)
        r = _slab64(o.astype(np.float64), d.astype(np.float64), bmin.astype(np.float64), bmax.astype(np.float64))
        want = r is not None and r[0] <= r[1]
        if r is not None and abs(r[1] - r[0]) < 1e-4:
            continue  # grazing
        ok, tmin = po.isect_ray_box(o, d, bmin, bmax)
        assert ok == want  # a line test: boxes behind the origin pass too
        if want:
            assert abs(tmin - r[0]) <= 1e-5 * (1 + abs(r[0]))
            hits += 1
        n += 1
    assert n > 3500 and hits > 1000
    # an axis-parallel ray outside the slab of a zero component misses; inside passes
    assert not po.isect_ray_box([2, 0, 5], [0, 0, -1], [-1, -1, -1], [1, 1, 1])[0]
    ok, t = po.isect_ray_box([0.5, 0, 5], [0, 0, -1], [-1, -1, -1], [1, 1, 1])
    assert ok and t == 4.0


def _box_d2(p, bmin, bmax):
    q = np.clip(p, bmin, bmax)
    return float(((p - q) ** 2).sum())


def test_ray_box_distance_against_minimisation(po):
    rng = np.random.default_rng(7)
    inter = 0
    for _ in range(1500):
        c, h = rng.normal(size=3), rng.uniform(0.1, 1.0, 3)
        bmin, bmax = (c - h).astype(F32), (c + h).astype(F32)
        o = (rng.normal(size=3) * 3).astype(F32)
        d = _unit(rng, 1)[0]
        o64, d64 = o.astype(np.float64), d.astype(np.float64)
        b0, b1 = bmin.astype(np.float64), bmax.astype(np.float64)
        res = minimize_scalar(lambda t: _box_d2(o64 + t * d64, b0, b1), bounds=(-60, 60), method="bounded",
                              options={"xatol": 1e-12})
        dsq, depth = po.dist_sq_ray_box(o, d, bmin, bmax)
        assert abs(dsq - res.fun) <= 1e-4 * (1 + res.fun)
        if res.fun < 1e-9:
            # the line meets the box: depth of the far corner along the ray
            far = np.where(d64 < 0, b0, b1)
            assert dsq == 0.0 and abs(depth - (far - o64) @ d64) <= 1e-5 * (1 + abs(depth))
            inter += 1
        else:
            assert abs(depth - res.x) <= 1e-3 * (1 + abs(res.x))
    assert inter > 20


def _seg_line64(o, d, a, b):
    """Squared distance from segment ab to the line o + t d and the line parameter of
    the segment's closest point (FP64, closed form)."""
    ab = b - a
    w = a - o
    # |w + s ab|^2 - ((w + s ab).d)^2, a quadratic in s, minimised over [0, 1]
    A = ab @ ab - (ab @ d) ** 2
    B = w @ ab - (w @ d) * (ab @ d)
    s = 0.0 if A <= 0 else min(1.0, max(0.0, -B / A))
    p = a + s * ab - o
    return p @ p - (p @ d) ** 2, p @ d


def test_ray_tri_edge_distance_against_fp64(po):
    rng = np.random.default_rng(9)
    for _ in range(2000):
        tri = rng.normal(size=(3, 3)).astype(F32)
        o = (rng.normal(size=3) * 3).astype(F32)
        d = _unit(rng, 1)[0]
        dsq, depth = po.dist_sq_ray_tri(o, d, *tri)
        T = tri.astype(np.float64)
        cands = [_seg_line64(o.astype(np.float64), d.astype(np.float64), T[i], T[j]) for i, j in ((0, 2), (1, 0), (2, 1))]
        best = min(cands, key=lambda x: x[0])
        assert abs(dsq - best[0]) <= 1e-4 * (1 + best[0])
        second = sorted(c[0] for c in cands)[1]
        if second - best[0] > 1e-3:
            assert abs(depth - best[1]) <= 1e-3 * (1 + abs(best[1]))


def _brute_cast(po, co, tris, o, d, depth0, hidden=None):
    """Nearest hit over every triangle (FP32 watertight test, pinned above)."""
    best, who = depth0, []
    for t, (a, b, c) in enumerate(tris):
        if hidden is not None and (hidden[a] or hidden[b] or hidden[c]):
            continue
        ok, lam = po.isect_ray_tri(o, d, co[a], co[b], co[c])
        if ok and lam < best:
            best, who = lam, [t]
        elif ok and lam == best and who:
            who.append(t)
    return best, who


@pytest.mark.parametrize("leaf", [8, 64])
def test_raycast_traversal_against_brute_force(sg, po, leaf):
    co, tris = sg.noisy_icosphere(3, seed=4, amplitude=0.05)
    o = po.Oracle(co, tris, leaf_limit=leaf)
    rng = np.random.default_rng(11)
    n_hit = 0
    for k in range(120):
        if k % 3 == 0:  # from inside the sphere
            start = (rng.normal(size=3) * 0.3).astype(F32)
        else:
            start = (_unit(rng, 1)[0] * 2.5).astype(F32)
        target = _unit(rng, 1)[0] * rng.uniform(0.0, 1.3)
        d = (target - start)
        d = (d / np.linalg.norm(d)).astype(F32)
        depth0 = FLT_MAX if k % 4 else F32(rng.uniform(0.5, 3.0))
        h = o.raycast(start, d, depth0)
        best, who = _brute_cast(po, co, tris, start, d, depth0)
        assert bool(h.hit) == bool(who)
        if who:
            n_hit += 1
            assert F32(h.depth) == F32(best) and h.tri in who
            loc = start + d * F32(h.depth)
            assert np.array_equal(np.array(h.location, F32), loc.astype(F32))
            c = co[tris[h.tri]].astype(np.float64)
            dist = ((c - loc.astype(np.float64)) ** 2).sum(1)
            srt = np.sort(dist)
            if srt[1] - srt[0] > 1e-9:
                assert h.active_vertex == tris[h.tri][int(np.argmin(dist))]
            fn = np.cross(c[0] - c[1], c[1] - c[2])
            assert np.allclose(np.array(h.face_normal), fn / np.linalg.norm(fn), atol=1e-5)
        else:
            assert h.depth == F32(depth0)
    assert n_hit > 60
    o.close()


def test_raycast_hidden_faces_are_skipped(sg, po):
    co, tris = sg.icosphere(3)
    hidden = np.zeros(len(co), bool)
    hidden[co[:, 2] > 0.3] = True
    o = po.Oracle(co, tris, leaf_limit=16, hidden=hidden.astype(np.uint8))
    # straight down from above: the top cap is hidden, so the hit is on the bottom
    h = o.raycast([0.01, 0.02, 3.0], [0, 0, -1])
    best, who = _brute_cast(po, co, tris, np.array([0.01, 0.02, 3.0], F32), np.array([0, 0, -1], F32), FLT_MAX, hidden)
    assert h.hit and F32(h.depth) == F32(best) and h.location[2] < -0.9
    o.close()


def test_raycast_hits_the_sphere(sg, po):
    co, tris = sg.icosphere(5)
    o = po.Oracle(co, tris, leaf_limit=128)
    rng = np.random.default_rng(13)
    for _ in range(100):
        start = _unit(rng, 1)[0] * 3.0
        aim = _unit(rng, 1)[0] * 0.7
        d = (aim - start) / np.linalg.norm(aim - start)
        h = o.raycast(start.astype(F32), d.astype(F32))
        # analytic: |s + t d| = 1; the mesh is inscribed, chord sag < 1e-3 at level 5
        b = start @ d
        disc = b * b - (start @ start - 1.0)
        assert h.hit and disc > 0
        t = -b - np.sqrt(disc)
        assert abs(h.depth - t) < 2e-3
    o.close()


def test_nearest_to_ray_against_brute_force(sg, po):
    co, tris = sg.icosphere(3)
    o = po.Oracle(co, tris, leaf_limit=16)
    rng = np.random.default_rng(17)
    for _ in range(60):
        # rays from far away that pass beside the sphere: every box lies ahead
        start = (_unit(rng, 1)[0] * 8.0).astype(F32)
        side = np.cross(start, _unit(rng, 1)[0])
        side = side / np.linalg.norm(side) * rng.uniform(1.05, 1.5)
        d = (side - start) / np.linalg.norm(side - start)
        d = d.astype(F32)
        r = o.nearest_to_ray(start, d)
        ds = [po.dist_sq_ray_tri(start, d, *co[t]) for t in tris]
        k = int(np.argmin([x[0] for x in ds]))
        assert r.hit and F32(r.dist_sq) == F32(ds[k][0])
        assert abs(r.depth - ds[k][1]) < 1e-4 * (1 + abs(ds[k][1])) or ds[k][0] == sorted(x[0] for x in ds)[1]
    o.close()


def test_project_ray_root(sg, po):
    co, tris = sg.icosphere(2)
    o = po.Oracle(co, tris, leaf_limit=16)
    vb, _ = o.node_bb(0)
    moved, s, e = o.project_ray_root([0.0, 0.0, 10.0], [0.0, 0.0, -10.0], [0.0, 0.0, -1.0])
    assert moved
    c = 0.5 * (vb[5] + vb[2])
    half = (vb[5] - c + 1e-3) * (1 + 1e-3)
    assert abs(s[2] - (c + half)) < 1e-6 and abs(e[2] - (c - half)) < 1e-6
    moved, s, e = o.project_ray_root([5.0, 5.0, 10.0], [5.0, 5.0, -10.0], [0.0, 0.0, -1.0])
    assert not moved and s[2] == 10.0
    o.close()
