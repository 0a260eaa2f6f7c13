"""PBVH_GRIDS on the GPU against the oracle (oracle/grids.c), bit for bit: per dab the
affected elements and gathered leaves, and at checkpoints every element's position,
normal and mask and every node box; the drop-in normals / bounds path after host
edits; the entry points that are faces only refuse grids."""
import numpy as np
import pytest

from dune_sculpt_amd import grids as gr

pytestmark = pytest.mark.gpu

TOOLS = None


def noisy_sphere_grids(sg, n, level, seed=1, amp=0.02, mask=False):
    co, off, pv = gr.quad_sphere(n)
    g = gr.Grids(co, off, pv, level=level, mask=mask)
    c = g.co()
    d = (c / np.linalg.norm(c, axis=1, keepdims=True)).astype(np.float32)
    c = sg.noise_displace(d.copy(), d, seed, amp)  # a function of position: element copies agree
    g.set_co(c)
    g.recalc_normals()
    return g, (co, off, pv)


def oracle_of(po, g, mesh, leaf_limit, hidden=None):
    co, off, pv = mesh
    c, n = g.co(), g.normals()
    m = g.mask() if g.key.has_mask else None
    return po.OracleGrids(len(co), off, pv, g.gs, c, n, mask=m, hidden=hidden, leaf_limit=leaf_limit)


def stroke(sg, seed, ndabs, rmin=0.05, rmax=0.2):
    rng = np.random.default_rng(seed)
    radii = rng.uniform(rmin, rmax, ndabs)
    strengths = rng.uniform(0.2, 1.0, ndabs)
    centres = sg.great_circle_centres(seed, radii, spacing_frac=0.5)
    tools = [sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH, sg.TOOL_GRAB]
    dabs = []
    for i in range(ndabs):
        nxt = centres[i + 1] if i + 1 < ndabs else centres[i]
        dabs.append(sg.make_dab(tool=tools[i % 4], location=centres[i], radius=float(radii[i]),
                                strength=float(strengths[i]), smooth_iterations=1 + (i // 4) % 4,
                                grab_delta=(nxt - centres[i]) * 0.5))
    return dabs


def assert_same_state(p, o, tag):
    co, no, mk = p.state()
    oc, on, om = o.state()
    assert np.array_equal(co.view(np.uint32), oc.view(np.uint32)), f"{tag}: positions"
    assert np.array_equal(no.view(np.uint32), on.view(np.uint32)), f"{tag}: normals"
    if mk is not None:
        assert np.array_equal(mk.view(np.uint32), om.view(np.uint32)), f"{tag}: mask"
    vb, ob = p.boxes()
    ovb, oob = o.boxes()
    assert np.array_equal(vb, ovb), f"{tag}: node boxes"
    assert np.array_equal(ob, oob), f"{tag}: original node boxes"


def run_parity(sg, po, p, o, dabs, check_every):
    start = o.state()[0].copy()
    total = 0
    for i, d in enumerate(dabs):
        od = po.to_orc(d)
        gathered = o.search(d.location, d.radius, use_orig=d.sculpt_tool == sg.TOOL_GRAB)
        p.dab(d)
        want = o.dab(od)
        got = p.last_affected()
        assert np.array_equal(got, want), f"dab {i}: affected elements"
        assert np.array_equal(p.last_gathered(), gathered), f"dab {i}: gathered leaves"
        st = p.last_stats()
        assert st[0] == len(gathered) and st[1] == len(want)
        total += len(want)
        if (i + 1) % check_every == 0 or i + 1 == len(dabs):
            assert_same_state(p, o, f"after dab {i}")
    assert total > 10 * len(dabs)  # the stroke touched the surface
    assert np.abs(o.state()[0] - start).max() > 1e-3


@pytest.mark.parametrize("leaf_limit", [10000, 2000])
def test_stroke_matches_oracle(sg, po, leaf_limit):
    g, mesh = noisy_sphere_grids(sg, 16, 4)
    o = oracle_of(po, g, mesh, leaf_limit)
    p = gr.GridPBVH(g, leaf_limit=leaf_limit)
    assert_same_state(p, o, "built")
    run_parity(sg, po, p, o, stroke(sg, 11, 48), 8)
    p.close()
    o.close()


def test_mask_hidden_and_custom_curve(sg, po):
    g, mesh = noisy_sphere_grids(sg, 12, 4, seed=3, mask=True)
    c = g.co().astype(np.float64)
    g.set_mask((0.5 + 0.5 * np.sin(4 * c[:, 0]) * np.cos(3 * c[:, 1])).astype(np.float32))
    face = np.array([g.grid_to_face(i) for i in range(g.G)])
    rng = np.random.default_rng(2)
    hid_faces = rng.choice(len(mesh[1]) - 1, 40, replace=False)
    hidden = np.isin(face, hid_faces).repeat(g.ga)
    g.set_hidden(hidden)
    o = oracle_of(po, g, mesh, 3000, hidden=hidden.astype(np.uint8))
    p = gr.GridPBVH(g, leaf_limit=3000)
    dabs = stroke(sg, 5, 24, 0.1, 0.3)
    table = np.linspace(0.0, 1.0, sg.CM_TABLE + 1) ** 2
    for i, d in enumerate(dabs):
        if i % 3 == 1:
            dabs[i] = sg.make_dab(tool=d.sculpt_tool, location=d.location, radius=d.radius, strength=d.strength,
                                  smooth_iterations=d.smooth_iterations, grab_delta=d.grab_delta,
                                  falloff=sg.CURVE_CUSTOM, curve_table=table)
    run_parity(sg, po, p, o, dabs, 6)
    p.close()
    o.close()


def test_mixed_polygons(sg, po):
    from test_oracle_grids import mixed_patch

    co, off, pv = mixed_patch()
    g = gr.Grids(co, off, pv, level=5)
    o = po.OracleGrids(len(co), off, pv, g.gs, g.co(), g.normals(), leaf_limit=4 * g.ga)
    p = gr.GridPBVH(g, leaf_limit=4 * g.ga)
    centre = np.array([1.0, 1.0, 0.05], np.float32)
    dabs = []
    for i, tool in enumerate([sg.TOOL_DRAW, sg.TOOL_SMOOTH, sg.TOOL_INFLATE, sg.TOOL_GRAB] * 3):
        loc = centre + np.array([0.1 * i - 0.5, 0.05 * i, 0.0], np.float32)
        dabs.append(sg.make_dab(tool=tool, location=loc, radius=0.5, strength=0.6, smooth_iterations=2,
                                grab_delta=(0.02, 0.01, 0.03)))
    run_parity(sg, po, p, o, dabs, 1)
    p.close()
    o.close()


def test_host_edit_update_normals_and_bounds(sg, po):
    g, mesh = noisy_sphere_grids(sg, 10, 4, seed=4)
    o = oracle_of(po, g, mesh, 2000)
    p = gr.GridPBVH(g, leaf_limit=2000)
    p.dab(sg.make_dab(tool=sg.TOOL_DRAW, location=(0.0, 0.0, 1.0), radius=0.3, strength=0.8))
    o.dab(po.to_orc(sg.make_dab(tool=sg.TOOL_DRAW, location=(0.0, 0.0, 1.0), radius=0.3, strength=0.8)))
    p.sync_host()
    leaf = next(i for i in range(p.nodes) if p.node_info(i)["flag"] & 1)
    grids = p.node_info(leaf)["grids"]
    c = g.co()
    for gi in grids[:3]:  # lift every element of three grids of one leaf (as a caller editing grids would)
        c[gi * g.ga:(gi + 1) * g.ga] *= 1.02
    g.set_co(c)
    p.upload()
    o.set_co(c)
    p.mark_update(leaf)
    faces = sorted({g.grid_to_face(int(x)) for x in grids})
    assert sorted(p.grid_updates(clear=False)) == faces
    p.update_normals()
    p.update_bounds(4 | 8 | 32)  # UpdateBB | UpdateOriginalBB | UpdateRedraw
    o.mark([leaf], 2 | 4 | 8)
    o.update_normals()
    o.update_bounds()
    assert_same_state(p, o, "host edit")
    assert p.grid_updates(clear=False) == []
    p.close()
    o.close()


def test_graph_stroke_log_and_snapshot(sg, po):
    g, mesh = noisy_sphere_grids(sg, 12, 4, seed=6)
    o = oracle_of(po, g, mesh, 4000)
    p = gr.GridPBVH(g, leaf_limit=4000)
    dabs = stroke(sg, 8, 40)
    p.snapshot()
    st = p.prepare(dabs, log=True)
    st.launch()
    counts = [len(o.dab(po.to_orc(d))) for d in dabs]
    assert_same_state(p, o, "graph stroke")
    log = p.stroke_log(len(dabs))
    assert log[:, 1].tolist() == counts
    after = p.state()
    p.restore()
    st.launch()  # the same stroke again from the restored state gives the same bits
    again = p.state()
    for a, b in zip(after[:2], again[:2]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    p.close()
    o.close()


def test_faces_only_entry_points_refuse_grids(sg):
    g, mesh = noisy_sphere_grids(sg, 4, 3)
    p = gr.GridPBVH(g)
    with pytest.raises(RuntimeError):
        p.dab(sg.make_dab(tool=sg.TOOL_ELASTIC_DEFORM, location=(0, 0, 1), radius=0.3))
    L = sg.lib()
    assert L.sg_pbvh_read_positions(p.h, None) < 0
    assert "faces only" in sg.last_error()
    p.close()


def test_four_million_elements(sg, po):
    """4.5M elements (55,296 grids of 9 x 9): 64 dabs, every affected set, final state."""
    g, mesh = noisy_sphere_grids(sg, 48, 4, seed=9)
    assert g.E == 6 * 48 * 48 * 4 * 81
    o = oracle_of(po, g, mesh, 10000)
    p = gr.GridPBVH(g)
    run_parity(sg, po, p, o, stroke(sg, 21, 64, 0.02, 0.1), 64)
    # every element inside its leaf's box
    co = p.state()[0]
    vb, _ = p.boxes()
    for i in range(p.nodes):
        info = p.node_info(i)
        if not info["flag"] & 1:
            continue
        els = (info["grids"][:, None] * g.ga + np.arange(g.ga)).reshape(-1)
        assert (co[els] >= vb[i, :3]).all() and (co[els] <= vb[i, 3:]).all()
    p.close()
    o.close()
