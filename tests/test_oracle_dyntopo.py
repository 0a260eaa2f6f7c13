"""The dynamic-topology oracle (oracle/dyntopo.c) pinned without the product:

- the fast build's numbering (children allocated in pairs, depth first) and its
  midpoint partition, on the exported tree;
- splits: a lone long edge and far-apart long edges give exactly what the reference's
  one-heap order gives (orc_dyn_subdivide_sequential, a separate restatement with its
  own recursive even-subdivision walk); new vertices at 0.5f * (a + b) bit for bit;
  the fixpoint (no long edge left on a face in range), the rules for masked, hidden
  and back-facing edges and the view-projected range; counts within a few percent of
  the sequential order on whole dabs;
- collapses: the kept vertex at the midpoint, the lower mask removed, no short edge
  left unless the surface would stop being manifold;
- every operation keeps the corner table, leaf lists and owners sound, the Euler
  characteristic and (splits) the area and the old vertices;
- the dab with topology off equals the faces oracle's dab (affected sets, Draw /
  Inflate / Grab positions bit for bit, Smooth to 1e-6), its normals equal the
  angle-weighted normals of the faces oracle to 1e-5;
- after_stroke leaves every leaf within the limit.
"""
import numpy as np
import pytest


def euler(o):
    c = o.counts()
    labs, _, _, _ = o.raw_faces()
    e = set()
    for t in labs:
        for k in range(3):
            a, b = int(t[k]), int(t[(k + 1) % 3])
            e.add((min(a, b), max(a, b)))
    return c["verts"] - len(e) + c["tris"]


def area(o):
    lab, co, _, _, _ = o.raw_verts()
    pos = {int(l): co[i].astype(np.float64) for i, l in enumerate(lab)}
    labs = o.raw_faces()[0]
    s = 0.0
    for t in labs:
        a, b, c = (pos[int(x)] for x in t)
        s += 0.5 * np.linalg.norm(np.cross(b - a, c - a))
    return s


def geometry(o):
    """Canonical geometry without labels: sorted vertex positions and sorted faces as
    position triples rotated to their smallest position."""
    lab, co, _, _, _ = o.raw_verts()
    pos = {int(l): tuple(co[i].view(np.uint32).tolist()) for i, l in enumerate(lab)}
    faces = []
    for t in o.raw_faces()[0]:
        p = [pos[int(x)] for x in t]
        k = min(range(3), key=lambda i: p[i])
        faces.append(tuple(p[(k + i) % 3] for i in range(3)))
    return sorted(pos.values()), sorted(faces)


def plane(n, size=1.0, z=0.0):
    xs = np.linspace(0.0, size, n + 1, dtype=np.float32)
    co = np.array([(x, y, z) for y in xs for x in xs], np.float32)
    tris = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            b, c, d = a + 1, a + n + 1, a + n + 2
            tris += [(a, b, d), (a, d, c)]
    return co, np.array(tris, np.int32)


@pytest.fixture(scope="module")
def ico(sg):
    return sg.noisy_icosphere(4, seed=1, amplitude=0.02)


def all_leaves(o):
    info, _, _ = o.nodes()
    return np.nonzero(info[:, 0] & 1)[0].astype(np.int32)


def test_build_numbering_and_midpoint_partition(sg, po, ico):
    co, tris = ico
    o = po.OracleDyn(co, tris)
    info, vb, _ = o.nodes()
    # children allocated in pairs as the depth-first walk reaches them
    nxt = [1]

    def visit(n):
        if info[n, 0] & 1:
            return
        assert info[n, 1] == nxt[0]
        nxt[0] += 2
        visit(info[n, 1])
        visit(info[n, 1] + 1)

    visit(0)
    assert nxt[0] == len(info)
    # leaf lists partition the faces, each leaf within the limit 100
    labs, fnode, _, _ = o.raw_faces()
    assert (info[fnode, 0] & 1).all()
    assert info[info[:, 0] & 1 == 1, 2].max() <= 100
    assert info[info[:, 0] & 1 == 1, 2].sum() == len(tris)
    # midpoint partition: every face of the left subtree has centroid <= mid on the
    # widest axis of the node's centroid box, every face of the right one > mid
    cen = {}
    pos = co
    for f, t in enumerate(tris):
        p = pos[t]
        cen[f] = ((p.min(0) + p.max(0)) * np.float32(0.5)).astype(np.float32)
    face_of_label = {tuple(int(x) for x in tris[f]): f for f in range(len(tris))}
    leaf_faces = {}
    for i, t in enumerate(labs):
        leaf_faces.setdefault(int(fnode[i]), []).append(face_of_label[tuple(int(x) for x in t)])

    def faces(n):
        if info[n, 0] & 1:
            return leaf_faces.get(n, [])
        return faces(info[n, 1]) + faces(info[n, 1] + 1)

    for n in range(len(info)):
        if info[n, 0] & 1:
            continue
        fl, fr = faces(info[n, 1]), faces(info[n, 1] + 1)
        cs = np.array([cen[f] for f in fl + fr])
        lo, hi = cs.min(0), cs.max(0)
        dim = hi - lo
        axis = (0 if dim[0] > dim[2] else 2) if dim[0] > dim[1] else (1 if dim[1] > dim[2] else 2)
        mid = np.float32((hi[axis] + lo[axis]) * np.float32(0.5))
        if len(fl) > 1 and len(fr) > 1:
            assert all(cen[f][axis] <= mid for f in fl) and all(cen[f][axis] > mid for f in fr)
    assert o.check() == 0
    o.close()


def test_lone_long_edge_matches_sequential_and_midpoint_bits(sg, po):
    # a square of two triangles: only the diagonal (1.13) is past the limit 0.9
    co = np.array([(0, 0, 0), (0.8, 0, 0), (0.8, 0.8, 0), (0, 0.8, 0)], np.float32)
    tris = np.array([(0, 1, 2), (0, 2, 3)], np.int32)
    outs = []
    for seq in (False, True):
        o = po.OracleDyn(co, tris)
        o.detail_size_set(0.9)
        o.topology_set(po.SUBDIVIDE)
        o.mark(all_leaves(o), po.UPDATE_TOPOLOGY)
        if seq:
            o.subdivide_sequential((0.5, 0.1, 0.0), 2.0)
        else:
            o.update_topology(po.SUBDIVIDE, (0.5, 0.1, 0.0), 2.0)
        assert o.check() == 0
        outs.append(geometry(o))
        lab, cov, _, _, _ = o.raw_verts()
        labs = o.raw_faces()[0]
        o.close()
    assert outs[0] == outs[1]
    assert len(outs[0][0]) == 5 and len(outs[0][1]) == 4
    # the diagonal split at its midpoint, bit for bit
    mid = (np.float32(0.5) * (co[0] + co[2])).astype(np.float32)
    assert any(np.array_equal(c.view(np.uint32), mid.view(np.uint32)) for c in cov)


def test_far_apart_long_edges_match_sequential(sg, po):
    co, tris = plane(12, 1.2)
    rng = np.random.default_rng(4)
    co[:, 2] = rng.uniform(-0.002, 0.002, len(co)).astype(np.float32)
    # spacing 0.1, diagonals 0.141 < 0.15: nothing is long.  Moving a vertex by
    # (-0.03, -0.03) makes exactly one of its edges long (the diagonal to (i+1, j+1),
    # 0.184), and none of that split's new edges; three such places far apart
    for v in (2 * 13 + 2, 6 * 13 + 9, 10 * 13 + 3):
        co[v, :2] -= np.float32(0.03)
    res = []
    for seq in (False, True):
        o = po.OracleDyn(co, tris)
        o.detail_size_set(0.15)
        o.mark(all_leaves(o), po.UPDATE_TOPOLOGY)
        if seq:
            n = o.subdivide_sequential((0.6, 0.6, 0.0), 5.0)
        else:
            o.update_topology(po.SUBDIVIDE, (0.6, 0.6, 0.0), 5.0)
            n = o.last_stats()[2]
        assert o.check() == 0
        res.append((n, geometry(o)))
        o.close()
    assert res[0][0] == res[1][0] == 3
    assert res[0][1] == res[1][1]


def fixpoint_violations(o, center, radius, limit2, long=True):
    """Faces in range of marked leaves with an edge past the limit (long) or under it."""
    lab, co, _, mk, _ = o.raw_verts()
    pos = {int(l): co[i] for i, l in enumerate(lab)}
    bad = 0
    c = np.asarray(center, np.float32)
    for t in o.raw_faces()[0]:
        p = [pos[int(x)] for x in t]
        # closest point test done coarsely: every corner inside the sphere
        if max(float(((q - c) ** 2).sum()) for q in p) > radius * radius:
            continue
        for k in range(3):
            d = p[(k + 1) % 3] - p[k]
            l2 = np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
            bad += (l2 > limit2) if long else (l2 < limit2)
    return bad


def test_subdivide_fixpoint_invariants_and_sequential_count(sg, po, ico):
    co, tris = ico
    counts = []
    for seq in (False, True):
        o = po.OracleDyn(co, tris)
        o.detail_size_set(0.04)
        a0, e0 = area(o), euler(o)
        lab0, co0, _, _, _ = o.raw_verts()
        c = (0.0, 0.0, 1.0)
        o.mark(o.search(c, 0.5), po.UPDATE_TOPOLOGY)
        if seq:
            counts.append(o.subdivide_sequential(c, 0.4))
        else:
            o.update_topology(po.SUBDIVIDE, c, 0.4)
            counts.append(int(o.last_stats()[2]))
            assert o.last_stats()[4] < 20  # rounds
        assert o.check() == 0
        assert euler(o) == e0 == 2
        assert abs(area(o) - a0) < 1e-5 * a0
        lab, cov, _, _, _ = o.raw_verts()
        old = np.isin(lab, lab0)
        assert np.array_equal(cov[old].view(np.uint32), co0[np.searchsorted(lab0, lab[old])].view(np.uint32))
        assert fixpoint_violations(o, c, 0.3, np.float32(0.04) ** 2) == 0
        o.close()
    assert counts[0] > 100
    assert abs(counts[0] - counts[1]) <= 0.1 * counts[1]


def test_masked_hidden_backfacing_and_projected(sg, po):
    co, tris = plane(8, 1.0)
    n = len(co)
    mask = np.zeros(n, np.float32)
    mask[co[:, 0] < 0.5] = 1.0  # left half fully masked
    o = po.OracleDyn(co, tris, mask=mask)
    o.detail_size_set(0.1)
    o.mark(all_leaves(o), po.UPDATE_TOPOLOGY)
    o.update_topology(po.SUBDIVIDE, (0.5, 0.5, 0.0), 5.0)
    lab, cov, _, mk, _ = o.raw_verts()
    new = lab >= (1 << 63)
    # a new vertex only on edges with some vertex below full mask: never at x < 0.5
    # between two masked vertices
    assert new.any() and (cov[new, 0] >= 0.5 - 0.0626).all()
    assert o.check() == 0
    o.close()
    # hidden vertices: none of their edges split
    hid = np.zeros(n, np.uint8)
    hid[co[:, 1] > 0.5] = 1
    o = po.OracleDyn(co, tris, vhide=hid)
    o.detail_size_set(0.1)
    o.mark(all_leaves(o), po.UPDATE_TOPOLOGY)
    o.update_topology(po.SUBDIVIDE, (0.5, 0.5, 0.0), 5.0)
    lab, cov, _, _, _ = o.raw_verts()
    assert (cov[lab >= (1 << 63), 1] <= 0.5).all()
    o.close()
    # front faces only: the plane faces +z; viewed from -z nothing splits
    o = po.OracleDyn(co, tris)
    o.detail_size_set(0.1)
    o.topology_set(po.SUBDIVIDE, view_normal=(0.0, 0.0, -1.0), use_frontface=True)
    o.mark(all_leaves(o), po.UPDATE_TOPOLOGY)
    assert not o.update_topology(po.SUBDIVIDE, (0.5, 0.5, 0.0), 5.0)
    o.close()
    # projected range: a centre far above the plane reaches it only as a circle
    for projected, want in ((False, False), (True, True)):
        o = po.OracleDyn(co, tris)
        o.detail_size_set(0.1)
        o.topology_set(po.SUBDIVIDE, view_normal=(0.0, 0.0, 1.0), use_projected=projected)
        o.mark(all_leaves(o), po.UPDATE_TOPOLOGY)
        assert o.update_topology(po.SUBDIVIDE, (0.5, 0.5, 3.0), 0.3) == want
        o.close()


def test_collapse_midpoint_mask_rule_and_fixpoint(sg, po, ico):
    co, tris = ico
    rng = np.random.default_rng(2)
    mask = rng.uniform(0.0, 0.9, len(co)).astype(np.float32)
    o = po.OracleDyn(co, tris, mask=mask)
    lab0, co0, _, mk0, _ = o.raw_verts()
    pos0 = {int(l): co0[i] for i, l in enumerate(lab0)}
    mk = {int(l): mk0[i] for i, l in enumerate(lab0)}
    o.detail_size_set(0.5)  # min 0.2: every edge near the top is short
    c = (0.0, 0.0, 1.0)
    o.mark(o.search(c, 0.4), po.UPDATE_TOPOLOGY)
    assert o.update_topology(po.COLLAPSE, c, 0.3)
    assert o.check() == 0 and euler(o) == 2
    st = o.last_stats()
    assert st[3] > 10
    lab, cov, _, _, _ = o.raw_verts()
    gone = set(pos0) - {int(x) for x in lab}
    assert len(gone) == st[3]
    moved = [int(l) for i, l in enumerate(lab) if not np.array_equal(cov[i], pos0[int(l)])]
    # every moved vertex is the midpoint (0.5f*(a+b)) of itself and a removed vertex
    # with a mask no lower than the removed one's
    for l in moved:
        cands = [g for g in gone if np.array_equal((np.float32(0.5) * (pos0[l] + pos0[g])).astype(np.float32),
                                                   cov[list(lab).index(l)])]
        if cands:
            assert any(mk[l] >= mk[g] for g in cands)
    o.close()


def test_dab_without_topology_equals_faces_oracle(sg, po):
    import ctypes as C

    co, tris = sg.noisy_icosphere(4, seed=3, amplitude=0.02)
    f = po.Oracle(co, tris, leaf_limit=100)
    o = po.OracleDyn(co, tris)
    o.topology_set(0)
    tools = [sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_GRAB, sg.TOOL_SMOOTH]
    for i, tool in enumerate(tools):
        # both start from the faces oracle's normals (labels are the input indices)
        nrm = np.ascontiguousarray(f.normals(), np.float32)
        po.lib().orc_dyn_set_normals(o.h, nrm.ctypes.data_as(C.POINTER(C.c_float)))
        d = sg.make_dab(tool=tool, location=(0.1 * i, 0.2, 0.97), radius=0.25, strength=0.6,
                        smooth_iterations=3, grab_delta=(0.01, 0.02, -0.01))
        want = f.dab(d)
        got = o.dab(po.to_orc(d))
        assert len(want) > 20
        assert np.array_equal(got.astype(np.int64), np.sort(want).astype(np.int64)), tool
        lab, cov, nov, _, _ = o.raw_verts()
        fp = f.positions()[lab.astype(np.int64)]
        if tool == sg.TOOL_SMOOTH:
            assert np.abs(cov - fp).max() < 1e-6
        else:
            assert np.array_equal(cov.view(np.uint32), fp.view(np.uint32)), tool
        f.set_positions(cov[np.argsort(lab)])
    # the dyntopo normal rule (corner angle x face normal) equals the angle-weighted
    # initial normals of the faces oracle on the same positions
    o.mark(all_leaves(o), 2)  # UpdateNormals
    o.update_normals()
    lab, cov, nov, _, _ = o.raw_verts()
    order = np.argsort(lab)
    ref = po.initial_normals(cov[order], tris)
    assert np.abs(nov[order] - ref).max() < 1e-5
    f.close()
    o.close()


def test_stroke_invariants_and_after_stroke(sg, po, ico):
    co, tris = ico
    o = po.OracleDyn(co, tris)
    o.detail_size_set(0.03)
    rng = np.random.default_rng(1)
    radii = rng.uniform(0.1, 0.2, 32).astype(np.float32)
    centres = sg.great_circle_centres(3, radii, spacing_frac=0.5)
    tools = [sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH, sg.TOOL_GRAB]
    splits = collapses = 0
    for i in range(32):
        d = sg.make_dab(tool=tools[i % 4], location=centres[i], radius=float(radii[i]), strength=0.5,
                        smooth_iterations=2, grab_delta=(0.01, 0, 0))
        aff = o.dab(po.to_orc(d))
        st = o.last_stats()
        splits += st[2]
        collapses += st[3]
        assert o.check() == 0, i
        assert st[8] == 0  # no capacity stop
        if tools[i % 4] == sg.TOOL_GRAB:
            assert st[2] == st[3] == 0
        assert len(aff) == st[1]
    assert splits > 1000 and collapses > 100
    assert euler(o) == 2
    o.after_stroke()
    info, vb, _ = o.nodes()
    leaves = info[:, 0] & 1 == 1
    assert info[leaves, 2].max() <= 100 and info[leaves, 2].sum() == o.counts()["tris"]
    assert o.check() == 0
    # every leaf box holds its faces' vertices
    lab, cov, _, _, _ = o.raw_verts()
    pos = {int(l): cov[i] for i, l in enumerate(lab)}
    labs, fnode, _, _ = o.raw_faces()
    for t, n in zip(labs, fnode):
        for x in t:
            p = pos[int(x)]
            assert (p >= vb[n, :3]).all() and (p <= vb[n, 3:]).all()
    o.close()
