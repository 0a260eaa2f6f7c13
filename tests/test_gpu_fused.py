"""Fused dabs (search + one kernel per gathered leaf over its halo) against the separate
kernels (brush, Smooth steps, normals and bounds) and against the oracle.

Bar: for every dab the affected set and gathered leaves are identical, and the whole
state -- positions, normals, every node box (current and original) -- is bit-identical
between the two GPU paths, and within 1e-5 relative of the oracle (bit-identical in
practice).  Covers every fused tool, Smooth with k = 1..4 (and k = 5, which takes the
separate kernels), masks, hidden vertices, custom falloff, BRUSH_DIR_IN, strokes as a
captured graph and one dab at a time, and an undo step recorded around fused dabs.
"""
import numpy as np
import pytest

from parity_util import REL_TOL, n_bit_diff, rel_err

pytestmark = pytest.mark.gpu


def _state(p):
    n = p.tree_info()["nodes"]
    boxes = np.array([np.concatenate([p.node_bb(i), p.node_bb(i, original=True)]) for i in range(n)], np.float32)
    return p.positions(), p.normals(), boxes


def _same(a, b):
    return all(n_bit_diff(x, y) == 0 for x, y in zip(a, b))


def _dabs(sg, co, n, seed, tools=(1, 4, 2, 5), ks=(4,), **kw):
    rng = np.random.default_rng(seed)
    pts = co[rng.integers(0, co.shape[0], n)]
    out = []
    for i in range(n):
        out.append(sg.make_dab(tool=tools[i % len(tools)], location=pts[i], radius=float(rng.uniform(0.05, 0.3)),
                               strength=float(rng.uniform(0.2, 1.0)), smooth_iterations=int(ks[i % len(ks)]),
                               grab_delta=rng.normal(size=3) * 0.02, **kw))
    return out


def _pair_run(sg, po, co, tris, dabs, leaf=64, mask=None, hidden=None, graph_too=True):
    a = sg.PBVH(co, tris, leaf_limit=leaf, hidden=hidden)
    b = sg.PBVH(co, tris, leaf_limit=leaf, hidden=hidden)
    assert a.set_fused(True), "halo tables missing"
    b.set_fused(False)
    o = po.Oracle(co, tris, leaf_limit=leaf, hidden=None if hidden is None else np.asarray(hidden, np.uint8))
    o.set_normals(a.normals())
    for q in (a, b):
        if mask is not None:
            q.set_mask(mask)
    if mask is not None:
        o.set_mask(mask)
    if graph_too:
        a.snapshot()
    counts = []
    for i, d in enumerate(dabs):
        want = o.dab(d)
        counts.append(len(want))
        a.dab(d)
        b.dab(d)
        ga, gb = a.last_affected(), b.last_affected()
        assert np.array_equal(ga, want) and np.array_equal(gb, want), f"dab {i}: affected set"
        assert np.array_equal(a.last_gathered(), b.last_gathered()), f"dab {i}: gathered leaves"
        assert int(a.last_stats()[1]) == len(want), f"dab {i}: affected count"
        sa, sb = _state(a), _state(b)
        assert _same(sa, sb), f"dab {i} (tool {d.sculpt_tool}, k {d.smooth_iterations}): fused and separate differ"
    pos, nor, _ = _state(a)
    assert rel_err(pos, o.positions()) <= REL_TOL and rel_err(nor, o.normals()) <= REL_TOL
    for i in range(o.totnode()):
        vb, ob = o.node_bb(i)
        assert n_bit_diff(a.node_bb(i), vb) == 0 and n_bit_diff(a.node_bb(i, original=True), ob) == 0, i
    if graph_too:
        # the same dabs as one captured stroke, with the per-dab counts logged
        final = _state(a)
        a.restore()
        a.stroke(dabs, graph=True, log=True)
        assert _same(_state(a), final)
        assert a.stroke_log(len(dabs))[:, 1].tolist() == counts
    a.close()
    b.close()
    o.close()


def test_fused_each_tool_and_k(sg, po):
    co, tris = sg.noisy_icosphere(5, seed=21, amplitude=0.02)
    _pair_run(sg, po, co, tris, _dabs(sg, co, 36, 1, tools=(1, 4, 2, 5, 2, 2), ks=(1, 2, 3, 4, 5)))


def test_fused_mask_hidden_custom_dir_in(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=22, amplitude=0.02)
    rng = np.random.default_rng(5)
    mask = rng.uniform(0, 1, co.shape[0]).astype(np.float32)
    hidden = rng.random(co.shape[0]) < 0.03
    x = np.linspace(0.0, 1.0, sg.CM_TABLE + 1)
    table = (1.0 - x) ** 3
    dabs = _dabs(sg, co, 16, 2, falloff=sg.CURVE_CUSTOM, curve_table=table)
    dabs += _dabs(sg, co, 16, 3, flag=sg.BRUSH_DIR_IN)
    _pair_run(sg, po, co, tris, dabs, leaf=32, mask=mask, hidden=hidden, graph_too=False)


def test_fused_stroke_counts_match_oracle(sg, po):
    """A captured stroke: each dab's logged affected count (summed by the next dab's
    search or, for the last dab, the finishing kernel) equals the oracle's."""
    co, tris = sg.noisy_icosphere(5, seed=23, amplitude=0.02)
    radii = np.random.default_rng(1).uniform(0.03, 0.2, 40).astype(np.float32)
    dabs = sg.stroke_config([1, 4, 2, 5] * 10, sg.great_circle_centres(9, radii), radii, np.full(40, 0.7))
    p = sg.PBVH(co, tris, leaf_limit=64)
    o = po.Oracle(co, tris, leaf_limit=64)
    o.set_normals(p.normals())
    p.stroke(dabs, graph=True, log=True)
    log = p.stroke_log(len(dabs))
    counts = [len(o.dab(d)) for d in dabs]
    assert log[:, 1].tolist() == counts
    assert rel_err(p.positions(), o.positions()) <= REL_TOL and rel_err(p.normals(), o.normals()) <= REL_TOL
    p.close()
    o.close()


def test_fused_undo_step(sg):
    co, tris = sg.noisy_icosphere(4, seed=24, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    before = _state(p)
    p.undo_begin()
    for d in _dabs(sg, co, 12, 4):
        p.dab(d)
    p.undo_end()
    after = _state(p)
    p.undo()
    assert _same(_state(p), before)
    p.redo()
    assert _same(_state(p), after)
    p.close()
