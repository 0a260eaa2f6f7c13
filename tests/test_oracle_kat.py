"""Known-answer tests that pin the CPU oracle.

The reference holds no tests, fixtures or golden vectors for this path (its brush code
is absent), so each oracle function is pinned by closed forms, invariants and brute
force on small meshes, chosen so that a dropped term, a flipped sign/comparison or a
wrong index fails at least one check.
"""
import numpy as np
import pytest

from meshes import edges_of, fan, grid

F32 = np.float32


def dab(sg, **kw):
    return sg.make_dab(**kw)


# --- S3 falloff ---------------------------------------------------------------

PRESET_FORMS = {
    1: lambda p: 3 * p * p - 2 * p ** 3,              # SMOOTH
    2: lambda p: np.sqrt(2 * p - p * p),              # SPHERE
    3: lambda p: np.sqrt(p),                          # ROOT
    4: lambda p: p * p,                               # SHARP
    5: lambda p: p,                                   # LIN
    6: lambda p: p ** 4,                              # POW4
    7: lambda p: p * (2 - p),                         # INVSQUARE
    8: lambda p: 1.0 + 0 * p,                         # CONSTANT
    9: lambda p: p ** 3 * (6 * p * p - 15 * p + 10),  # SMOOTHER
}


@pytest.mark.parametrize("preset", sorted(PRESET_FORMS))
def test_falloff_closed_forms(sg, po, preset):
    d = dab(sg, radius=1.0, falloff=preset)
    for p in [0.0, 0.25, 0.5, 0.75, 1.0]:
        dist = 1.0 - p  # exact in FP32
        got = po.falloff(d, F32(dist * dist))
        want = PRESET_FORMS[preset](p) if dist < 1.0 else 0.0
        assert abs(got - want) <= 2e-7, (preset, p, got, want)


def test_falloff_smooth_symmetry_and_range(sg, po):
    d = dab(sg, radius=0.37, falloff=1)
    rng = np.random.default_rng(0)
    for x in rng.uniform(0, 0.37, 200).astype(F32):
        f = po.falloff(d, F32(x * x))
        g = po.falloff(d, F32((F32(0.37) - x) ** 2))
        assert 0.0 <= f <= 1.0
        assert abs(f + g - 1.0) < 5e-6  # 3p^2-2p^3 + 3(1-p)^2-2(1-p)^3 = 1
    assert po.falloff(d, F32(0.37 * 0.37) * F32(1.0001)) == 0.0  # outside


def test_falloff_hardness(sg, po):
    d = dab(sg, radius=1.0, falloff=1, hardness=0.5)
    assert po.falloff(d, F32(0.4 ** 2)) == 1.0          # inside the hard core
    assert abs(po.falloff(d, F32(0.75 ** 2)) - 0.5) < 1e-7  # (0.75-0.5)/(1-0.5) -> p=0.5


def test_falloff_custom_table(sg, po):
    x = np.linspace(0, 1, sg.CM_TABLE + 1)
    table = (x ** 3).astype(F32)
    d = dab(sg, radius=1.0, falloff=0, curve_table=table)
    # table is evaluated at 1 - p = dist / r
    for k in [0, 17, 128, 200, 255]:
        dist = k / 256.0
        assert abs(po.falloff(d, F32(dist * dist)) - table[k]) < 1e-6
    dist = (100 + 0.5) / 256.0
    assert abs(po.falloff(d, F32(dist) ** 2) - 0.5 * (table[100] + table[101])) < 1e-5


# --- tri normal / normalize --------------------------------------------------

def test_tri_normal_orientation_and_unit(po):
    n = po.tri_normal([0, 0, 0], [1, 0, 0], [0, 1, 0])
    assert np.array_equal(n, F32([0, 0, 1]))
    n = po.tri_normal([0, 0, 0], [0, 1, 0], [1, 0, 0])
    assert np.array_equal(n, F32([0, 0, -1]))
    rng = np.random.default_rng(3)
    for _ in range(50):
        a, b, c = rng.normal(size=(3, 3)).astype(F32)
        n = po.tri_normal(a, b, c)
        want = np.cross(np.float64(a) - b, np.float64(b) - c)
        want /= np.linalg.norm(want)
        assert np.abs(n - want).max() < 1e-5
    assert np.array_equal(po.tri_normal([1, 1, 1], [1, 1, 1], [2, 2, 2]), F32([0, 0, 0]))  # degenerate -> zero


def test_initial_normals_flat_and_sphere(sg, po):
    co, tris = grid(9)
    n = po.initial_normals(co, tris)
    assert np.array_equal(n, np.tile(F32([0, 0, 1]), (co.shape[0], 1)))
    co, tris = sg.noisy_icosphere(3, seed=4, amplitude=0.05)
    n = po.initial_normals(co, tris)
    radial = co / np.linalg.norm(co, axis=1, keepdims=True)
    assert np.abs(n - radial).max() < 0.2
    # independent FP64 statement: normalize(sum over tris of corner angle * unit face normal)
    P = np.float64(co)[tris]
    fn = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    acc = np.zeros((co.shape[0], 3))
    for k in range(3):
        e1 = P[:, (k + 1) % 3] - P[:, k]
        e2 = P[:, (k + 2) % 3] - P[:, k]
        cosang = (e1 * e2).sum(1) / np.linalg.norm(e1, axis=1) / np.linalg.norm(e2, axis=1)
        np.add.at(acc, tris[:, k], fn * np.arccos(np.clip(cosang, -1, 1))[:, None])
    want = acc / np.linalg.norm(acc, axis=1, keepdims=True)
    assert np.abs(n - want).max() < 1e-5
    # unweighted sum differs measurably (so the weighting is pinned)
    un = np.zeros((co.shape[0], 3))
    for k in range(3):
        np.add.at(un, tris[:, k], fn)
    un /= np.linalg.norm(un, axis=1, keepdims=True)
    assert np.abs(un - want).max() > 1e-4
    # angle weighting: at a valence-5 vertex all 5 corner angles are equal, so the
    # normal is exactly symmetric; shuffling tri order changes only rounding
    perm = np.random.default_rng(1).permutation(tris.shape[0])
    assert np.abs(po.initial_normals(co, tris[perm]) - n).max() < 1e-6


# --- build invariants ---------------------------------------------------------

@pytest.mark.parametrize("level,leaf", [(2, 8), (3, 16), (4, 128)])
def test_build_invariants(sg, po, level, leaf):
    co, tris = sg.noisy_icosphere(level, seed=2, amplitude=0.02)
    o = po.Oracle(co, tris, leaf_limit=leaf)
    prim = o.prim_indices()
    assert np.array_equal(np.sort(prim), np.arange(tris.shape[0]))
    seen_tri = np.zeros(tris.shape[0], int)
    uniq_owner = np.full(co.shape[0], -1)
    leaves = o.leaves_dfs()
    last_off = -1
    for li, node in enumerate(leaves):
        flag, child, off, totprim, nu, nf = o.node_info(node)
        assert flag & 1 and child == -1
        assert 1 <= totprim <= leaf
        assert off > last_off  # DFS leaf order == increasing prim offset
        last_off = off
        seen_tri[prim[off:off + totprim]] += 1
        vi = o.node_verts(node)
        ltris = tris[prim[off:off + totprim]]
        assert set(vi.tolist()) == set(ltris.reshape(-1).tolist())
        for v in vi[:nu]:
            assert uniq_owner[v] == -1
            uniq_owner[v] = li
        vb, ob = o.node_bb(node)
        pts = co[vi]
        assert np.array_equal(vb[:3], pts.min(0)) and np.array_equal(vb[3:], pts.max(0))
        assert np.array_equal(vb, ob)
    assert (seen_tri == 1).all()
    assert (uniq_owner >= 0).all()
    for li, node in enumerate(leaves):  # shared verts are uniq in an earlier leaf
        _, _, _, _, nu, nf = o.node_info(node)
        for v in o.node_verts(node)[nu:]:
            assert uniq_owner[v] < li
    for i in range(o.totnode()):
        flag, child = o.node_info(i)[:2]
        if not flag & 1:
            vb = o.node_bb(i)[0]
            a, b = o.node_bb(child)[0], o.node_bb(child + 1)[0]
            assert np.array_equal(vb[:3], np.minimum(a[:3], b[:3]))
            assert np.array_equal(vb[3:], np.maximum(a[3:], b[3:]))


# --- S1/S2 search and affected set vs brute force ----------------------------

def _flat_search(o, centre, r):
    """Brute force: every leaf box against the S1 predicate in FP32 (numpy float32
    does not fuse), in DFS order."""
    c = np.asarray(centre, F32)
    r2 = F32(r) * F32(r)
    out = []
    for node in o.leaves_dfs():
        vb = o.node_bb(node)[0]
        nearest = np.where(vb[:3] > c, vb[:3], np.where(vb[3:] < c, vb[3:], c))
        t = c - nearest
        if (t[0] * t[0] + t[1] * t[1]) + t[2] * t[2] < r2:
            out.append(node)
    return out


def _brute_affected(o, co_now, centre, r):
    c = np.asarray(centre, F32)
    r2 = F32(r) * F32(r)
    aff = []
    for node in _flat_search(o, centre, r):
        nu = o.node_info(node)[4]
        for v in o.node_verts(node)[:nu]:
            d = c - co_now[v]
            if (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2] <= r2:
                aff.append(v)
    return np.array(sorted(aff), np.int32)


def test_search_and_affected_brute_force(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=6, amplitude=0.02)
    o = po.Oracle(co, tris, leaf_limit=16)
    rng = np.random.default_rng(5)
    for _ in range(12):
        c = co[rng.integers(co.shape[0])] * F32(1.01)
        r = float(rng.uniform(0.05, 0.5))
        assert list(o.search_sphere(c, r)) == _flat_search(o, c, r)
        before = o.positions()
        want = _brute_affected(o, before, c, r)
        got = o.dab(sg.make_dab(tool=1, location=c, radius=r, strength=0.3))
        assert np.array_equal(got, want)
        # tree independence except exact ties: {d2 < r2} <= A <= {d2 <= r2}
        dd = np.float32(c) - before
        d2 = (dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]) + dd[:, 2] * dd[:, 2]
        r2 = F32(r) * F32(r)
        assert set(np.nonzero(d2 < r2)[0]) <= set(got) <= set(np.nonzero(d2 <= r2)[0])


def test_affected_lattice_ties(sg, po):
    """Plane grid with spacing 1/8 and r = 3/8: squared distances are exact, so the
    vertices at exactly r (4 of the 29 lattice points within r) exercise '<=' and the
    strict leaf test."""
    co, tris = grid(17)
    for leaf in (4, 8, 2048):
        o = po.Oracle(co, tris, leaf_limit=leaf)
        got = o.dab(sg.make_dab(tool=1, location=(0, 0, 0), radius=0.375, strength=0.5))
        assert np.array_equal(got, _brute_affected(po.Oracle(co, tris, leaf_limit=leaf), co, (0, 0, 0), 0.375))
        ij = np.rint(co[got, :2] / 0.125).astype(int)
        assert ((ij ** 2).sum(1) <= 9).all()
        if leaf == 2048:  # one leaf holding everything: every tie is inside
            assert len(got) == 29


# --- S6-S9 displacement closed forms ------------------------------------------

def _f64_smooth(dist, r):
    p = 1.0 - dist / r
    return np.where(dist < r, 3 * p * p - 2 * p ** 3, 0.0)


def test_draw_flat_grid_closed_form(sg, po):
    co, tris = grid(33, spacing=1 / 32)
    o = po.Oracle(co, tris, leaf_limit=64)
    r, s = 0.3, 0.6
    got = o.dab(sg.make_dab(tool=1, location=(0.01, -0.02, 0.0), radius=r, strength=s))
    new = o.positions()
    dist = np.linalg.norm(np.float64(co[got]) - [0.01, -0.02, 0.0], axis=1)
    want_dz = np.float64(F32(r)) * np.float64(F32(s)) ** 2 * _f64_smooth(dist, np.float64(F32(r)))
    assert np.abs(new[got, 2] - want_dz).max() < 1e-6
    assert np.array_equal(new[:, :2], co[:, :2])
    # BRUSH_DIR_IN negates
    o2 = po.Oracle(co, tris, leaf_limit=64)
    o2.dab(sg.make_dab(tool=1, location=(0.01, -0.02, 0.0), radius=r, strength=s, flag=sg.BRUSH_DIR_IN))
    assert np.abs(o2.positions()[got, 2] + want_dz).max() < 1e-6


def test_inflate_is_radial_on_sphere(sg, po):
    co, tris = sg.icosphere(4)
    o = po.Oracle(co, tris, leaf_limit=64)
    n0 = o.normals()
    r, s = 0.4, 0.8
    c = co[10]
    got = o.dab(sg.make_dab(tool=4, location=c, radius=r, strength=s))
    disp = np.float64(o.positions()[got]) - co[got]
    dist = np.linalg.norm(np.float64(co[got]) - c, axis=1)
    want = 0.25 * s * s * r * _f64_smooth(dist, r)
    assert np.abs(np.linalg.norm(disp, axis=1) - want).max() < 1e-6
    cross = np.linalg.norm(np.cross(disp, n0[got]), axis=1)
    assert cross.max() < 1e-6


def test_grab_closed_form(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=3, amplitude=0.02)
    o = po.Oracle(co, tris, leaf_limit=64)
    r, s, delta = 0.3, 0.7, np.array([0.01, -0.02, 0.005])
    c = co[100]
    got = o.dab(sg.make_dab(tool=5, location=c, radius=r, strength=s, grab_delta=delta))
    dist = np.linalg.norm(np.float64(co[got]) - c, axis=1)
    want = co[got] + np.outer(np.float64(F32(s)) * _f64_smooth(dist, np.float64(F32(r))), delta)
    assert np.abs(o.positions()[got] - want).max() < 1e-6


def test_smooth_fixed_point_and_fan(sg, po):
    co, tris = grid(17)
    o = po.Oracle(co, tris, leaf_limit=32)
    got = o.dab(sg.make_dab(tool=2, location=(0, 0, 0), radius=0.5, strength=1.0, smooth_iterations=3))
    new = o.positions()
    interior = (np.abs(co[:, :2]) < 0.99).all(1)
    moved = np.abs(new - co).max(1)
    assert moved[np.intersect1d(got, np.nonzero(interior)[0])].max() < 1e-6  # regular interior: fixed point
    # n-gon fan: the centre moves toward the ring centroid by fade per step
    co, tris = fan(8, centre=(0.1, 0.0, 0.0))
    o = po.Oracle(co, tris, leaf_limit=64)
    s = 0.5
    o.dab(sg.make_dab(tool=2, location=(0.1, 0, 0), radius=0.05, strength=s, falloff=8, smooth_iterations=1))
    assert abs(o.positions()[0, 0] - 0.1 * (1 - s)) < 1e-7
    o.dab(sg.make_dab(tool=2, location=(0.05, 0, 0), radius=0.05, strength=s, falloff=8, smooth_iterations=2))
    assert abs(o.positions()[0, 0] - 0.1 * (1 - s) ** 3) < 1e-7


OCTA = np.array([(1, 0, 0), (-0.5, 0, 0), (0, 0.5, 0), (0, -1, 0), (0, 0, 0.25), (0, 0, -0.75)], np.float32)
OCTA_TRIS = np.array([(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)],
                     np.int32)


def _smooth_variants(k, s, r):
    """Exact rational Smooth(k) on the octahedron with LIN falloff around the origin, in
    three readings of spec S8: Jacobi with the factor fixed from the pre-dab positions
    (the spec), in-place (Gauss-Seidel, vertex order), and Jacobi with the factor
    recomputed from each iterate.  Every vertex has 4 edge neighbours (all but its
    opposite), so every quantity of the spec's reading is a short dyadic rational and FP32
    holds it exactly (the refreshed factors are not exact: that reading is only compared
    for being different)."""
    from fractions import Fraction as Fr

    P0 = [[Fr(float(x)) for x in v] for v in OCTA]
    nbr = [[u for u in range(6) if u != v and u != (v ^ 1)] for v in range(6)]

    def fade(P, v, exact=True):
        d = sum(c * c for c in P[v])  # squared distance: a dyadic square at the start
        dist = Fr(float(d) ** 0.5)
        assert dist * dist == d or not exact
        return Fr(float(s)) * (1 - dist / Fr(float(r)))

    out = {}
    for name in ("jacobi", "inplace", "refade"):
        P = [list(v) for v in P0]
        f0 = [fade(P0, v) for v in range(6)]
        for _ in range(k):
            src = P if name == "inplace" else [list(v) for v in P]
            f = [fade(src, v, exact=False) for v in range(6)] if name == "refade" else f0
            for v in range(6):
                avg = [sum(src[u][a] for u in nbr[v]) / 4 for a in range(3)]
                P[v] = [src[v][a] + (avg[a] - src[v][a]) * f[v] for a in range(3)]
        out[name] = np.array([[float(c) for c in v] for v in P])
    return out


@pytest.mark.parametrize("k", [2, 3])
def test_smooth_jacobi_closed_form(sg, po, k):
    """Spec S8 on a skewed octahedron, every vertex affected (r = 2, LIN falloff, s = 0.5:
    factors 0.25 .. 0.4375): the oracle equals exact rational Jacobi with the factor
    fixed at the dab's start, bit for bit, and the in-place and refreshed-factor readings
    give different values, so either mistake fails here."""
    want = _smooth_variants(k, 0.5, 2.0)
    assert np.abs(want["jacobi"] - want["inplace"]).max() > 1e-3
    assert np.abs(want["jacobi"] - want["refade"]).max() > 1e-3
    o = po.Oracle(OCTA, OCTA_TRIS, leaf_limit=8)
    got = o.dab(sg.make_dab(tool=2, location=(0, 0, 0), radius=2.0, strength=0.5, falloff=5, smooth_iterations=k))
    assert got.tolist() == list(range(6))
    assert np.array_equal(o.positions(), want["jacobi"].astype(np.float32))
    o.close()


def test_mask_scales_factor(sg, po):
    co, tris = grid(17, spacing=1 / 16)
    mask = np.full(co.shape[0], 0.25, np.float32)
    o1 = po.Oracle(co, tris, leaf_limit=64)
    o2 = po.Oracle(co, tris, leaf_limit=64)
    o2.set_mask(mask)
    d = sg.make_dab(tool=5, location=(0, 0, 0), radius=0.4, strength=1.0, grab_delta=(0, 0, 1))
    a1 = o1.dab(d)
    a2 = o2.dab(d)
    assert np.array_equal(a1, a2)
    assert np.abs(o2.positions()[:, 2] - 0.75 * o1.positions()[:, 2]).max() < 1e-6


# --- S10 normals after a dab, refit, replay -----------------------------------

def test_normals_after_dab_flat_and_translation(sg, po):
    co, tris = grid(17)
    o = po.Oracle(co, tris, leaf_limit=32)
    o.dab(sg.make_dab(tool=5, location=(0, 0, 0), radius=0.3, strength=1.0, grab_delta=(0.01, 0.0, 0.0)))
    assert np.array_equal(o.normals()[:, :2], np.zeros((co.shape[0], 2), np.float32))  # still planar
    co, tris = sg.noisy_icosphere(4, seed=1, amplitude=0.02)
    shift = np.float32([0.5, -0.25, 0.125])
    o1 = po.Oracle(co, tris, leaf_limit=64)
    o2 = po.Oracle(co + shift, tris, leaf_limit=64)
    o2.set_normals(o1.normals())
    d1 = sg.make_dab(tool=1, location=co[7], radius=0.3, strength=0.5)
    d2 = sg.make_dab(tool=1, location=co[7] + shift, radius=0.3, strength=0.5)
    a1, a2 = o1.dab(d1), o2.dab(d2)
    assert np.array_equal(a1, a2)
    assert np.abs(o1.normals() - o2.normals()).max() < 1e-5


def test_refit_brute_force_after_stroke(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=9, amplitude=0.02)
    o = po.Oracle(co, tris, leaf_limit=32)
    cs = sg.great_circle_centres(3, np.full(15, 0.3))
    for i, c in enumerate(cs):
        o.dab(sg.make_dab(tool=[1, 4, 2, 5][i % 4], location=c, radius=0.3, strength=0.7, smooth_iterations=2,
                          grab_delta=(0.01, 0, 0)))
    pos = o.positions()
    for node in o.leaves_dfs():
        vb = o.node_bb(node)[0]
        pts = pos[o.node_verts(node)]
        assert np.array_equal(vb[:3], pts.min(0)) and np.array_equal(vb[3:], pts.max(0))
    root = o.node_bb(0)[0]
    assert np.array_equal(root[:3], pos.min(0)) and np.array_equal(root[3:], pos.max(0))


def test_fp64_replay_close(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=2, amplitude=0.02)
    o = po.Oracle(co, tris, leaf_limit=64)
    cs = sg.great_circle_centres(8, np.full(30, 0.2))
    for i, c in enumerate(cs):
        o.dab(sg.make_dab(tool=[1, 4, 2, 5][i % 4], location=c, radius=0.2, strength=0.6, smooth_iterations=4,
                          grab_delta=(0.005, 0, 0)))
    assert np.abs(o.positions() - o.positions64()).max() < 1e-6
    assert np.abs(o.normals() - o.normals64()).max() < 1e-5
