"""PBVH_BMESH on the GPU against the oracle (oracle/dyntopo.c), bit for bit through
vertex labels: per dab the affected vertices, gathered leaves and counters (splits,
collapses, rounds, candidates, leaves re-normalled and refit); at checkpoints every
vertex (position, normal, mask, owner), every face (labels, leaf, normal) and every
node (flags, children, face count, boxes); the drop-in update_topology / normals /
bounds path; after_stroke."""
import numpy as np
import pytest

from dune_sculpt_amd import dyntopo as dy

pytestmark = pytest.mark.gpu


def same_state(p, o, tag):
    pv, pf = p.state()
    ov, of = o.state()
    assert np.array_equal(pv["lab"], ov["lab"]), f"{tag}: vertex labels"
    for k in ("co", "no", "mask"):
        assert np.array_equal(pv[k].view(np.uint32), ov[k].view(np.uint32)), f"{tag}: {k}"
    assert np.array_equal(pv["node"], ov["node"]), f"{tag}: owners"
    assert np.array_equal(pf["labs"], of["labs"]), f"{tag}: faces"
    assert np.array_equal(pf["node"], of["node"]), f"{tag}: face leaves"
    assert np.array_equal(pf["fno"].view(np.uint32), of["fno"].view(np.uint32)), f"{tag}: face normals"
    assert np.array_equal(pf["hide"], of["hide"]), f"{tag}: face hidden"
    info, vb, ob = p.nodes()
    oinfo, ovb, oob = o.nodes()
    assert np.array_equal(info[:, 0] & ~(1 << 9), oinfo[:, 0] & ~(1 << 10)), f"{tag}: node flags"
    assert np.array_equal((info[:, 0] >> 9) & 1, (oinfo[:, 0] >> 10) & 1), f"{tag}: fully hidden"
    assert np.array_equal(info[:, 1:], oinfo[:, 1:]), f"{tag}: children / face counts"
    assert np.array_equal(vb.view(np.uint32), ovb.view(np.uint32)), f"{tag}: boxes"
    assert np.array_equal(ob.view(np.uint32), oob.view(np.uint32)), f"{tag}: original boxes"


def stroke(sg, seed, n, rmin, rmax):
    rng = np.random.default_rng(seed)
    radii = rng.uniform(rmin, rmax, n).astype(np.float32)
    strengths = rng.uniform(0.2, 1.0, n)
    centres = sg.great_circle_centres(seed, radii, spacing_frac=0.5)
    tools = [sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH, sg.TOOL_GRAB]
    out = []
    for i in range(n):
        nxt = centres[i + 1] if i + 1 < n else centres[i]
        out.append(sg.make_dab(tool=tools[i % 4], location=centres[i], radius=float(radii[i]),
                               strength=float(strengths[i]), smooth_iterations=1 + (i // 4) % 3,
                               grab_delta=(nxt - centres[i]) * 0.5))
    return out


def pair(sg, po, co, tris, detail, **kw):
    m = dy.DynMesh(co, tris, **kw)
    p = dy.DynPBVH(m)
    o = po.OracleDyn(co, tris, **kw)
    p.detail_size_set(detail)
    o.detail_size_set(detail)
    return m, p, o


def run(sg, po, p, o, dabs, check_every):
    splits = collapses = 0
    for i, d in enumerate(dabs):
        print(f"dab {i} tool {d.sculpt_tool} r {d.radius:.3f}", flush=True)
        p.dab(d)
        want = o.dab(po.to_orc(d))
        got = p.last_affected()
        ps, os_ = p.last_stats(), o.last_stats()
        assert np.array_equal(ps, os_), f"dab {i}: counters {ps.tolist()} vs {os_.tolist()}"
        assert np.array_equal(got, want), f"dab {i}: affected"
        assert sorted(p.last_gathered().tolist()) == sorted(o.search(d.location, d.radius).tolist()), f"dab {i}"
        splits += ps[2]
        collapses += ps[3]
        if (i + 1) % check_every == 0 or i + 1 == len(dabs):
            same_state(p, o, f"after dab {i}")
    return splits, collapses


def test_stroke_matches_oracle(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=1, amplitude=0.02)
    m, p, o = pair(sg, po, co, tris, 0.03)
    same_state(p, o, "built")
    splits, collapses = run(sg, po, p, o, stroke(sg, 3, 24, 0.1, 0.2), 4)
    assert splits > 500 and collapses > 50
    p.after_stroke()
    o.after_stroke()
    same_state(p, o, "after stroke")
    # the stroke goes on over the split leaves
    run(sg, po, p, o, stroke(sg, 4, 8, 0.1, 0.2), 4)
    p.close()
    m.close()
    o.close()


def test_mask_hidden_frontface_projected(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=5, amplitude=0.02)
    rng = np.random.default_rng(6)
    c64 = co.astype(np.float64)
    mask = (0.5 + 0.5 * np.sin(6 * c64[:, 0])).astype(np.float32)
    mask[c64[:, 1] > 0.5] = 1.0
    vh = (c64[:, 2] < -0.6).astype(np.uint8)
    fh = (rng.uniform(0, 1, len(tris)) < 0.03).astype(np.uint8)
    m, p, o = pair(sg, po, co, tris, 0.035, mask=mask, vhide=vh, fhide=fh)
    for proj, front in ((False, True), (True, False), (True, True)):
        p.topology_set(dy.SUBDIVIDE | dy.COLLAPSE, view_normal=(0.0, 0.6, 0.8), use_frontface=front,
                       use_projected=proj)
        o.topology_set(po.SUBDIVIDE | po.COLLAPSE, view_normal=(0.0, 0.6, 0.8), use_frontface=front,
                       use_projected=proj)
        run(sg, po, p, o, stroke(sg, 7 + proj + 2 * front, 8, 0.12, 0.25), 4)
    p.close()
    m.close()
    o.close()


def test_drop_in_update_topology_normals_bounds(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=8, amplitude=0.02)
    m, p, o = pair(sg, po, co, tris, 0.04)
    c = (0.0, 0.0, 1.0)
    leaves = o.search(c, 0.45)
    assert sorted(p.search(c, 0.45).tolist()) == sorted(leaves.tolist())
    p.mark_topology(leaves)
    o.mark(leaves, po.UPDATE_TOPOLOGY)
    assert p.update_topology(dy.SUBDIVIDE | dy.COLLAPSE, c, 0.35)
    assert o.update_topology(po.SUBDIVIDE | po.COLLAPSE, c, 0.35)
    ps, os_ = p.last_stats(), o.last_stats()
    assert np.array_equal(ps[2:10], os_[2:10]), (ps.tolist(), os_.tolist())
    # the caller refreshes normals and bounds of the flagged leaves
    p.update_normals()
    o.update_normals()
    p.update_bounds()
    o.update_bounds()
    same_state(p, o, "drop-in")
    p.close()
    m.close()
    o.close()


def test_larger_mesh_long_stroke(sg, po):
    """icosphere 6 (41k vertices), 48 dabs with subdivision and collapse."""
    co, tris = sg.noisy_icosphere(6, seed=9, amplitude=0.01)
    m, p, o = pair(sg, po, co, tris, 0.012)
    splits, collapses = run(sg, po, p, o, stroke(sg, 11, 48, 0.05, 0.12), 16)
    assert splits > 1000
    p.close()
    m.close()
    o.close()


def test_capacity_stop_and_node_lists(sg, po):
    """Tiny capacities: the split phase stops at the same round on both sides (counter
    [8] bit 1) and the mesh stays sound; the node accessors list each leaf's live faces,
    its unique and other vertices disjoint."""
    co, tris = sg.noisy_icosphere(3, seed=12, amplitude=0.02)
    V, T = len(co), len(tris)
    m = dy.DynMesh(co, tris, vert_cap=V + 40, tri_cap=T + 80)
    p = dy.DynPBVH(m)
    o = po.OracleDyn(co, tris, vert_cap=V + 40, tri_cap=T + 80)
    p.detail_size_set(0.02)
    o.detail_size_set(0.02)
    stopped = 0
    for i, d in enumerate(stroke(sg, 13, 6, 0.2, 0.3)):
        p.dab(d)
        o.dab(po.to_orc(d))
        ps, os_ = p.last_stats(), o.last_stats()
        assert np.array_equal(ps, os_), (i, ps.tolist(), os_.tolist())
        stopped += int(ps[8] & 1)
    assert stopped > 0
    same_state(p, o, "capacity")
    assert o.check() == 0
    # node lists: the leaf's live faces, and unique / other vertices disjoint
    info, _, _ = p.nodes()
    for leaf in np.nonzero(info[:, 0] & 1)[0][:8]:
        faces, uniq, other = p.node_lists(leaf)
        assert len(faces) == info[leaf, 2]
        assert len(np.intersect1d(uniq, other)) == 0
    p.close()
    m.close()
    o.close()


def test_bench_size_stroke_final_state(sg, po):
    """The bench's dyntopo workload (icosphere-8, 1.31 M triangles, detail 0.003, 200
    dabs): every dab's counters, then every vertex, face and node bit-identical."""
    co, tris = sg.noisy_icosphere(8, seed=1, amplitude=0.02)
    m, p, o = pair(sg, po, co, tris, 0.003)
    rng = np.random.default_rng(5)
    n = 200
    radii = rng.uniform(0.02, 0.05, n).astype(np.float32)
    strengths = rng.uniform(0.2, 1.0, n)
    centres = sg.great_circle_centres(5, radii, spacing_frac=0.3)
    tools = [sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH, sg.TOOL_GRAB]
    for i in range(n):
        nxt = centres[i + 1] if i + 1 < n else centres[i]
        d = sg.make_dab(tool=tools[i % 4], location=centres[i], radius=float(radii[i]),
                        strength=float(strengths[i]), smooth_iterations=2, grab_delta=(nxt - centres[i]) * 0.5)
        p.dab(d)
        o.dab(po.to_orc(d))
        assert np.array_equal(p.last_stats(), o.last_stats()), i
    same_state(p, o, "bench stroke")
    p.after_stroke()
    o.after_stroke()
    same_state(p, o, "bench stroke, after_stroke")
    p.close()
    m.close()
    o.close()


def polar_disk(rings, seg, seed, amplitude=0.01):
    """An open disk: a centre vertex of valence `seg` (past the fans the kernels hold in
    registers), `rings` rings of `seg` vertices, the outer ring a boundary."""
    co = [(0.0, 0.0, 0.0)]
    for r in range(1, rings + 1):
        rad = r / rings
        for k in range(seg):
            a = 2.0 * np.pi * k / seg
            co.append((rad * np.cos(a), rad * np.sin(a), 0.0))
    co = np.array(co, np.float32)
    co[:, 2] = np.random.default_rng(seed).uniform(-amplitude, amplitude, len(co)).astype(np.float32)
    tris = [(0, 1 + k, 1 + (k + 1) % seg) for k in range(seg)]
    for r in range(1, rings):
        b0, b1 = 1 + (r - 1) * seg, 1 + r * seg
        for k in range(seg):
            a, b = b0 + k, b0 + (k + 1) % seg
            c, d = b1 + k, b1 + (k + 1) % seg
            tris += [(a, c, d), (a, d, b)]
    return co, np.array(tris, np.int32)


def test_open_disk_boundary_and_high_valence(sg, po):
    """Boundary fans (open rings, boundary edges, boundary collapse rules) and a centre of
    valence 32 (walks past the register-held fans) through a stroke that crosses the rim
    and the centre."""
    co, tris = polar_disk(16, 32, seed=21)
    m, p, o = pair(sg, po, co, tris, 0.12)
    rng = np.random.default_rng(22)
    tools = [sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH, sg.TOOL_GRAB]
    dabs = []
    n = 24
    for i in range(n):
        t = i / (n - 1)
        rad = 1.1 * (1.0 - t)  # from past the rim to the centre
        a = 3.0 * t
        c = np.array([rad * np.cos(a), rad * np.sin(a), 0.0], np.float32)
        dabs.append(sg.make_dab(tool=tools[i % 4], location=c, radius=float(rng.uniform(0.2, 0.35)),
                                strength=float(rng.uniform(0.3, 0.9)), smooth_iterations=2,
                                grab_delta=np.array([0.01, -0.02, 0.005], np.float32)))
    splits, collapses = run(sg, po, p, o, dabs, 4)
    assert splits > 100 and collapses > 10, (splits, collapses)
    assert o.check() == 0
    p.after_stroke()
    o.after_stroke()
    same_state(p, o, "disk after stroke")
    p.close()
    m.close()
    o.close()


def test_whole_mesh_dabs_refit_every_ancestor(sg, po):
    """Dabs covering the whole mesh with topology off: every leaf refits, so the bounds
    flush lists every internal node (more than it stages in LDS) and runs over HBM;
    every node box bit-identical to the oracle."""
    co, tris = sg.noisy_icosphere(6, seed=31, amplitude=0.02)
    m, p, o = pair(sg, po, co, tris, 0.02)
    p.topology_set(0)
    o.topology_set(0)
    info, _, _ = p.nodes()
    assert (info[:, 0] & 1).sum() > 600  # leaves; their ancestors number about as many
    tools = [sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH, sg.TOOL_GRAB]
    dabs = [sg.make_dab(tool=tools[i], location=np.array([0.1, 0.2, 0.3], np.float32), radius=1.6,
                        strength=0.3, smooth_iterations=1, grab_delta=np.array([0.002, 0.0, -0.001], np.float32))
            for i in range(4)]
    run(sg, po, p, o, dabs, 2)
    p.close()
    m.close()
    o.close()
