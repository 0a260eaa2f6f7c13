"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bar: affected-vertex sets and gathered leaves bit-exact per dab; positions and normals
within 1e-5 relative (REL_TOL); node bounds bit-exact after the refit.
"""
import numpy as np
import pytest

from parity_util import REL_TOL, rel_err, run_pair

pytestmark = pytest.mark.gpu


def _c1_stroke(sg, n=200, radius=0.1, strength=0.5, seed=1, tool=1, **kw):
    radii = np.full(n, radius, np.float32)
    centres = sg.great_circle_centres(seed, radii)
    return [sg.make_dab(tool=tool, location=c, radius=radius, strength=strength, **kw) for c in centres]


def test_c1_draw_stroke_icosphere6(sg, po):
    """Config 1: icosphere-6, Draw r=0.1 s=0.5 smooth falloff, 200 dabs, leaf 128."""
    co, tris = sg.icosphere(6)
    dabs = _c1_stroke(sg)
    rep = run_pair(sg, po, co, tris, dabs, leaf_limit=128, check_each=False)
    print(rep)
    assert rep["affected"] > 200 * 50


@pytest.mark.parametrize("tool", [1, 2, 4, 5])
def test_each_tool_noisy_icosphere5(sg, po, tool):
    co, tris = sg.noisy_icosphere(5, seed=3, amplitude=0.02)
    rng = np.random.default_rng(tool)
    n = 24
    radii = rng.uniform(0.05, 0.2, n).astype(np.float32)
    centres = sg.great_circle_centres(10 + tool, radii)
    strengths = rng.uniform(0.2, 1.0, n).astype(np.float32)
    dabs = sg.stroke_config([tool] * n, centres, radii, strengths, smooth_iterations=4)
    rep = run_pair(sg, po, co, tris, dabs, leaf_limit=128)
    print(rep)


def test_mixed_cycle_with_flags(sg, po):
    """Tools cycled D/I/S(4)/G with every falloff preset, hardness and BRUSH_DIR_IN."""
    co, tris = sg.noisy_icosphere(4, seed=5, amplitude=0.02)
    n = 40
    rng = np.random.default_rng(9)
    radii = rng.uniform(0.1, 0.4, n).astype(np.float32)
    centres = sg.great_circle_centres(4, radii)
    dabs = []
    tools = [1, 4, 2, 5]
    for i in range(n):
        nxt = centres[min(i + 1, n - 1)]
        dabs.append(sg.make_dab(tool=tools[i % 4], location=centres[i], radius=float(radii[i]),
                                strength=float(rng.uniform(0.2, 1.0)), falloff=int(1 + i % 9),
                                hardness=float(0.3 if i % 5 == 0 else 0.0), smooth_iterations=1 + i % 4,
                                flag=(sg.BRUSH_DIR_IN if i % 7 == 3 else 0), grab_delta=nxt - centres[i]))
    run_pair(sg, po, co, tris, dabs, leaf_limit=64)


def test_custom_curve_and_mask(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=8, amplitude=0.01)
    x = np.linspace(0.0, 1.0, sg.CM_TABLE + 1)
    table = (1.0 - x) ** 2
    rng = np.random.default_rng(2)
    mask = rng.uniform(0.0, 1.0, co.shape[0]).astype(np.float32)
    centres = sg.great_circle_centres(6, np.full(20, 0.3))
    dabs = [sg.make_dab(tool=[1, 4][i % 2], location=c, radius=0.3, strength=0.7, falloff=sg.CURVE_CUSTOM,
                        curve_table=table) for i, c in enumerate(centres)]
    run_pair(sg, po, co, tris, dabs, leaf_limit=32, mask=mask)


def test_torus_smooth(sg, po):
    co, tris = sg.noisy_torus(60, 18, seed=2, amplitude=0.01)
    rng = np.random.default_rng(4)
    pts = co[rng.integers(0, co.shape[0], 20)]
    dabs = [sg.make_dab(tool=2, location=p, radius=0.15, strength=0.8, smooth_iterations=4) for p in pts]
    run_pair(sg, po, co, tris, dabs, leaf_limit=128)


def test_stroke_api_matches_single_dabs(sg, po):
    """sg_dab_apply_stroke (graph captured, deferred internal refit) == the oracle."""
    co, tris = sg.noisy_icosphere(5, seed=11, amplitude=0.02)
    n = 60
    rng = np.random.default_rng(12)
    radii = rng.uniform(0.05, 0.15, n).astype(np.float32)
    centres = sg.great_circle_centres(13, radii)
    tools = np.array([1, 4, 2, 5] * (n // 4))
    dabs = sg.stroke_config(tools, centres, radii, rng.uniform(0.2, 1.0, n), smooth_iterations=4)
    o = po.Oracle(co, tris, leaf_limit=128)
    for graph in (True, False):
        p = sg.PBVH(co, tris, leaf_limit=128)
        o2 = po.Oracle(co, tris, leaf_limit=128)
        o2.set_normals(p.normals())
        p.stroke(dabs, graph=graph)
        for d in dabs:
            o2.dab(d)
        assert rel_err(p.positions(), o2.positions()) <= REL_TOL
        assert rel_err(p.normals(), o2.normals()) <= REL_TOL
        for i in range(o2.totnode()):
            vb, _ = o2.node_bb(i)
            assert np.array_equal(p.node_bb(i).view(np.uint32), vb.view(np.uint32)), i
        p.close()


def test_determinism_bitwise(sg):
    co, tris = sg.noisy_icosphere(5, seed=1, amplitude=0.02)
    n = 40
    radii = np.full(n, 0.12, np.float32)
    centres = sg.great_circle_centres(2, radii)
    dabs = sg.stroke_config([1, 4, 2, 5] * (n // 4), centres, radii, np.full(n, 0.6), smooth_iterations=4)
    outs = []
    for _ in range(2):
        p = sg.PBVH(co, tris, leaf_limit=128)
        p.stroke(dabs)
        outs.append((p.positions(), p.normals()))
        p.close()
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert np.array_equal(outs[0][1].view(np.uint32), outs[1][1].view(np.uint32))


def test_initial_normals_match_oracle(sg, po):
    co, tris = sg.noisy_icosphere(5, seed=4, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=128)
    want = po.initial_normals(co, tris)
    assert rel_err(p.normals(), want) <= REL_TOL


def test_prepared_stroke_replay(sg, po):
    """sg_stroke_prepare captures once; each launch applies the dabs again."""
    co, tris = sg.noisy_icosphere(5, seed=21, amplitude=0.02)
    n = 32
    rng = np.random.default_rng(22)
    radii = rng.uniform(0.05, 0.15, n).astype(np.float32)
    centres = sg.great_circle_centres(23, radii)
    dabs = sg.stroke_config([1, 4, 2, 5] * (n // 4), centres, radii, rng.uniform(0.2, 1.0, n), smooth_iterations=4)
    p = sg.PBVH(co, tris, leaf_limit=128)
    o = po.Oracle(co, tris, leaf_limit=128)
    o.set_normals(p.normals())
    st = p.prepare(dabs, log=True)
    want_aff = []
    for rep in range(2):
        st.launch()
        for d in dabs:
            want_aff.append(len(o.dab(d)))
    logrows = p.stroke_log(n)
    assert list(logrows[:, 1]) == want_aff[n:]
    assert rel_err(p.positions(), o.positions()) <= REL_TOL
    assert rel_err(p.normals(), o.normals()) <= REL_TOL
    for i in range(o.totnode()):
        assert np.array_equal(p.node_bb(i).view(np.uint32), o.node_bb(i)[0].view(np.uint32)), i
    st.close()
    p.close()


def test_snapshot_restore_and_autotune_keep_results(sg):
    """Checkpoint restore is exact; the autotuner changes launch shapes, never results."""
    co, tris = sg.noisy_icosphere(5, seed=31, amplitude=0.02)
    n = 40
    rng = np.random.default_rng(32)
    radii = rng.uniform(0.05, 0.15, n).astype(np.float32)
    centres = sg.great_circle_centres(33, radii)
    dabs = sg.stroke_config([1, 4, 2, 5] * (n // 4), centres, radii, rng.uniform(0.2, 1.0, n), smooth_iterations=4)
    p = sg.PBVH(co, tris, leaf_limit=128)
    p0, n0 = p.positions(), p.normals()
    bb0 = [p.node_bb(i) for i in range(p.tree_info()["nodes"])]
    p.snapshot()
    p.stroke(dabs)
    ref_pos, ref_nor = p.positions(), p.normals()
    p.restore()
    assert np.array_equal(p.positions().view(np.uint32), p0.view(np.uint32))
    assert np.array_equal(p.normals().view(np.uint32), n0.view(np.uint32))
    for i, b in enumerate(bb0):
        assert np.array_equal(p.node_bb(i).view(np.uint32), b.view(np.uint32))
    before, after = p.autotune(dabs[:20])
    assert after <= before * 1.05
    assert np.array_equal(p.positions().view(np.uint32), p0.view(np.uint32))  # tuner restores
    for grids in [(32, 64, 128), (512, 32, 64)]:
        p.restore()
        p.set_kernel_grids(*grids)
        p.stroke(dabs)
        assert np.array_equal(p.positions().view(np.uint32), ref_pos.view(np.uint32))
        assert np.array_equal(p.normals().view(np.uint32), ref_nor.view(np.uint32))
    p.close()
