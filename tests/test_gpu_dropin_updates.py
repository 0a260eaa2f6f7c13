"""The drop-in update entry points the reference's callers run, on the GPU against the
oracle's restatement of the same calls.

- Host edits (paint.c:1762-1765, pbvh.c:1401, :4559): vertices edited through
  BKE_pbvh_get_verts, marked with BKE_pbvh_vert_mark_update, their leaves with
  BKE_pbvh_node_mark_update, then BKE_pbvh_update_bounds and BKE_pbvh_update_normals.
- Partial bounds flags (UpdateBB during a stroke, UpdateOriginalBB at its end): the
  per-node rules of pbvh_update_BB_redraw and pbvh_flush_bb (pbvh.c:1206-1242,
  :1369-1399), including an original box that lags the current one.
- BKE_pbvh_vert_coords_apply (pbvh.c:4707-4749): positions, marks, every node box; the
  vertex normals stay until the next BKE_pbvh_update_normals, which then updates the
  marked vertices; BKE_pbvh_vert_coords_alloc returns the positions.
- Frustum searches (pbvh.c:4507-4557) through BKE_pbvh_search_gather after a stroke,
  against a closed form on the oracle's boxes.

Bar: positions and node boxes bit-exact, node flags equal, normals within 1e-5 relative
(they come out bit-identical: the same sums in the same order).
"""
import numpy as np
import pytest

from parity_util import REL_TOL, n_bit_diff, rel_err

pytestmark = pytest.mark.gpu

UPD_NORMALS, UPD_BB, UPD_OBB, UPD_DRAW, UPD_REDRAW = 2, 4, 8, 16, 32
LEAF = 1


def _pair(sg, po, level=5, leaf=64):
    co, tris = sg.noisy_icosphere(level, seed=9, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=leaf)
    o = po.Oracle(co, tris, leaf_limit=leaf)
    o.set_normals(p.normals())
    return co, p, o


def _boxes_equal(p, o):
    bad = 0
    for i in range(o.totnode()):
        vb, ob = o.node_bb(i)
        bad += n_bit_diff(p.node_bb(i), vb) != 0 or n_bit_diff(p.node_bb(i, original=True), ob) != 0
    return bad


def _flags_equal(p, o, mask=UPD_NORMALS | UPD_BB | UPD_OBB | UPD_REDRAW | LEAF):
    return [i for i in range(o.totnode()) if (p.node_flag(i) & mask) != (o.node_flag(i) & mask)]


def _edit(p, o, co, verts, offset):
    mv = p.mverts()
    for v in verts:
        new = (co[v] + offset).astype(np.float32)
        mv[v] = new
        o.set_vert(v, new)
        p.vert_mark_update(v)
        o.vert_mark_update(v)


def _leaves_with(o, verts):
    s = set(int(v) for v in verts)
    return [n for n in o.leaves_dfs() if s & set(o.node_verts(n).tolist())]


def test_host_edit_loop(sg, po):
    co, p, o = _pair(sg, po)
    n_before = p.normals()
    # a patch of vertices around one vertex, moved off the surface
    d = np.linalg.norm(co - co[123], axis=1)
    verts = np.nonzero(d < 0.15)[0]
    _edit(p, o, co, verts, np.float32([0.01, -0.02, 0.03]))
    for n in _leaves_with(o, verts):
        p.node_mark_update(n)
        o.node_mark_update(n)
    flag = UPD_BB | UPD_OBB | UPD_REDRAW
    p.update_bounds(flag)
    o.update_bounds(flag)
    assert _boxes_equal(p, o) == 0
    p.update_normals()
    o.update_normals()
    assert n_bit_diff(p.positions(), o.positions()) == 0
    e = rel_err(p.normals(), o.normals())
    assert e <= REL_TOL, e
    assert not _flags_equal(p, o)
    # only marked vertices get new normals
    moved = np.nonzero(np.abs(p.normals() - n_before).max(1) > 0)[0]
    assert len(moved) and set(moved.tolist()) <= set(verts.tolist())
    print(f"[dropin] {len(verts)} edited vertices, {len(moved)} new normals, normals rel {e:.3g}")
    # a dab afterwards sees the edited mesh: same affected set and result
    dab = sg.make_dab(tool=sg.TOOL_DRAW, location=co[123], radius=0.2, strength=0.8)
    p.dab(dab)
    assert np.array_equal(p.last_affected(), o.dab(dab))
    assert rel_err(p.positions(), o.positions()) <= REL_TOL
    p.close()
    o.close()


def test_marks_outside_flagged_leaves_wait(sg, po):
    """pbvh.c:1083-1087: only the marked uniq vertices of UpdateNormals leaves are
    recomputed; a mark in another leaf waits for that leaf to be flagged."""
    co, p, o = _pair(sg, po)
    leaves = o.leaves_dfs()
    a, b = leaves[3], leaves[40]
    va, vb_ = o.node_verts(a)[:5], o.node_verts(b)[:5]
    _edit(p, o, co, np.concatenate([va, vb_]), np.float32([0.0, 0.0, 0.05]))
    for q in (p, o):
        q.node_mark_update(a)
        q.update_bounds(UPD_BB | UPD_OBB)
        q.update_normals()
    e1 = rel_err(p.normals(), o.normals())
    # leaf b's vertices moved but keep their old normals until b is flagged
    for q in (p, o):
        q.node_mark_update(b)
        q.update_bounds(UPD_BB | UPD_OBB)
        q.update_normals()
    e2 = rel_err(p.normals(), o.normals())
    assert e1 <= REL_TOL and e2 <= REL_TOL and _boxes_equal(p, o) == 0 and not _flags_equal(p, o)
    p.close()
    o.close()


def test_partial_bounds_flags(sg, po):
    """UpdateBB alone refits the current boxes and leaves the original ones; a later
    UpdateOriginalBB copies the current boxes into them (Blender's stroke step / stroke
    end pattern)."""
    co, p, o = _pair(sg, po)
    verts = np.nonzero(np.linalg.norm(co - co[777], axis=1) < 0.2)[0]
    _edit(p, o, co, verts, np.float32([0.05, 0.05, 0.05]))
    for n in _leaves_with(o, verts):
        p.node_mark_update(n)
        o.node_mark_update(n)
    p.update_bounds(UPD_BB)
    o.update_bounds(UPD_BB)
    assert _boxes_equal(p, o) == 0 and not _flags_equal(p, o)
    lagging = sum(n_bit_diff(*o.node_bb(i)) != 0 for i in range(o.totnode()))
    assert lagging > 0  # current and original boxes now differ somewhere
    p.update_bounds(UPD_OBB | UPD_REDRAW)
    o.update_bounds(UPD_OBB | UPD_REDRAW)
    assert _boxes_equal(p, o) == 0 and not _flags_equal(p, o)
    print(f"[dropin] partial flags: {lagging} nodes had lagging original boxes")
    p.close()
    o.close()


def test_vert_coords_apply(sg, po):
    co, p, o = _pair(sg, po)
    # a stroke first, so the replaced coordinates start from a sculpted state
    dabs = sg.stroke_config([1, 4, 2, 5] * 3, sg.great_circle_centres(3, np.full(12, 0.2)), np.full(12, 0.2),
                            np.full(12, 0.6))
    for d in dabs:
        o.dab(d)
        p.dab(d)
    before = p.normals()
    obefore = o.normals()
    rng = np.random.default_rng(4)
    new = o.positions()
    sel = rng.random(len(new)) < 0.5
    new[sel] += rng.normal(scale=0.003, size=(sel.sum(), 3)).astype(np.float32)
    p.coords_apply(new)
    o.vert_coords_apply(new)
    assert n_bit_diff(p.positions(), new) == 0 and n_bit_diff(o.positions(), new) == 0
    assert n_bit_diff(p.normals(), before) == 0 and n_bit_diff(o.normals(), obefore) == 0  # not yet updated
    assert _boxes_equal(p, o) == 0 and not _flags_equal(p, o)
    assert n_bit_diff(p.coords_alloc(), new) == 0
    p.update_normals()
    o.update_normals()
    e = rel_err(p.normals(), o.normals())
    assert e <= REL_TOL
    changed = np.nonzero(np.abs(p.normals() - before).max(1) > 0)[0]
    assert set(changed.tolist()) <= set(np.nonzero(sel)[0].tolist())  # only marked vertices change
    # dabs on the replaced mesh
    for d in dabs[:6]:
        p.dab(d)
        assert np.array_equal(p.last_affected(), o.dab(d))
    assert rel_err(p.positions(), o.positions()) <= REL_TOL and _boxes_equal(p, o) == 0
    print(f"[dropin] coords apply: {int(sel.sum())} vertices replaced, normals after update rel {e:.3g}, "
          f"{len(changed)} normals changed")
    p.close()
    o.close()


def _frustum_closed_form(o, lo, hi, exclude):
    """Axis-aligned region lo <= x <= hi as six planes: a box is outside when some
    bmax < lo or bmin > hi, inside when every bmin > lo and bmax < hi (the plane tests'
    strict / non-strict sides); boxes nest, so the pruned search returns exactly the
    leaves that pass, left to right."""
    out = []
    for n in o.leaves_dfs():
        vb, _ = o.node_bb(n)
        outside = (vb[3:] < lo).any() or (vb[:3] > hi).any()
        inside = (vb[:3] > lo).all() and (vb[3:] < hi).all()
        if (not inside) if exclude else (not outside):
            out.append(int(n))
    return out


def test_frustum_search_after_stroke(sg, po):
    co, p, o = _pair(sg, po)
    dabs = sg.stroke_config([1, 4] * 6, sg.great_circle_centres(8, np.full(12, 0.25)), np.full(12, 0.25),
                            np.full(12, 0.9))
    for d in dabs:
        o.dab(d)
        p.dab(d)
    lo, hi = np.float32([-0.3, -0.2, 0.1]), np.float32([0.5, 0.6, 1.2])
    planes = np.array([[1, 0, 0, -lo[0]], [0, 1, 0, -lo[1]], [0, 0, 1, -lo[2]],
                       [-1, 0, 0, hi[0]], [0, -1, 0, hi[1]], [0, 0, -1, hi[2]]], np.float32)
    for exclude in (False, True):
        want = _frustum_closed_form(o, lo, hi, exclude)
        got = p.frustum_gather(planes, exclude=exclude)
        assert got == want and 0 < len(got) < len(o.leaves_dfs())
    p.close()
    o.close()
