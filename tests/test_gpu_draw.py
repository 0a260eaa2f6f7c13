"""GPU draw buffers (sg_pbvh_draw_buffers_*) against the oracle's restatement of the
reference's per-leaf draw records.  Bar: bit-exact bytes for every leaf (all 36 bytes
of every vertex record; padding and the unused colour attribute are zero on both)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32


def _torus_quads(sg, nu, nv, seed, amp):
    co, _ = sg.noisy_torus(nu, nv, seed=seed, amplitude=amp)
    pv = []
    for i in range(nu):
        for j in range(nv):
            pv += [i * nv + j, ((i + 1) % nu) * nv + j, ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv]
    off = np.arange(0, len(pv) + 1, 4)
    return co, off, np.array(pv, np.int32)


def _compare_all(p, o, tri_poly, off, pv, smooth, mat, fs, mask, seed, default):
    n = 0
    for leaf in o.leaves_dfs():
        want, wsm, wmat = o.draw_leaf(leaf, tri_poly, off, pv, smooth, mat, fs, mask=mask, seed=seed, default=default)
        got, gsm, gmat = p.draw_leaf(leaf)
        assert (gsm, gmat) == (wsm, wmat), leaf
        assert got.shape == want.shape, (leaf, got.shape, want.shape)
        assert np.array_equal(got, want), leaf
        n += len(got)
    return n


def test_draw_buffers_quads_flat_smooth_hidden_mask_face_sets(sg, po):
    co, off, pv = _torus_quads(sg, 96, 32, seed=3, amp=0.01)
    P = len(off) - 1
    rng = np.random.default_rng(1)
    smooth = (np.arange(P) // 512 % 2).astype(np.uint8)  # regions of smooth and flat faces
    mat = (np.arange(P) // 700).astype(np.int32)
    fs = rng.choice([1, 1, 1, 1, 2, 7, 0, -3], P).astype(np.int32)
    hidden = rng.random(len(co)) < 0.02
    mask = rng.random(len(co)).astype(F32)
    p = sg.PBVH(co, None, leaf_limit=32, hidden=hidden, polys=(off, pv, smooth, mat))
    # the tree splits leaves by material and smooth flag (the reference's material key)
    key = (mat | (smooth.astype(np.int32) << 16))[p.tri_poly]
    o = po.Oracle(co, p.tris, leaf_limit=32, hidden=hidden.astype(np.uint8), mat=key)
    o.set_normals(p.normals())
    p.set_mask(mask)
    print("[draw] quad torus build", flush=True)
    p.draw_build(fs, seed=5, default=1)
    info = p.draw_info()
    tri_poly = p.tri_poly.astype(np.int32)
    n = _compare_all(p, o, tri_poly, off, pv, smooth, mat, fs, mask, 5, 1)
    assert n == 3 * info["visible_tris"] and info["bytes"] == 36 * n
    # after a stroke every leaf is refilled from the current mesh
    radii = np.full(12, 0.15, F32)
    centres = sg.great_circle_centres(2, radii)
    p.stroke(sg.stroke_config([1, 4, 2, 5] * 3, centres, radii, np.full(12, 0.6, F32)))
    p.draw_update(all=True)
    o.set_positions(p.positions())
    o.set_normals(p.normals())
    _compare_all(p, o, tri_poly, off, pv, smooth, mat, fs, mask, 5, 1)
    p.close()
    o.close()


def test_draw_buffers_incremental_after_dabs(sg, po):
    """Triangle polygons: the leaves a stroke's dabs gather are the only ones whose
    records change, and an update refreshes exactly those to the oracle's bytes."""
    co, tris = sg.noisy_icosphere(5, seed=4, amplitude=0.02)
    P = len(tris)
    off = np.arange(0, 3 * P + 1, 3)
    pv = tris.reshape(-1).astype(np.int32)
    smooth = np.ones(P, np.uint8)
    mat = np.zeros(P, np.int32)
    fs = np.ones(P, np.int32)
    p = sg.PBVH(co, None, leaf_limit=64, polys=(off, pv, smooth, mat))
    o = po.Oracle(co, p.tris, leaf_limit=64)
    p.draw_build(fs)
    before = {l: p.draw_leaf(l)[0] for l in o.leaves_dfs()}
    radii = np.full(8, 0.1, F32)
    centres = sg.great_circle_centres(9, radii)
    print("[draw] incremental", flush=True)
    p.stroke(sg.stroke_config([1, 4, 2, 5] * 2, centres, radii, np.full(8, 0.7, F32)), graph=True)
    untouched = {l: p.draw_leaf(l)[0] for l in o.leaves_dfs()}
    assert all(np.array_equal(untouched[l], before[l]) for l in before)  # nothing refilled yet
    p.draw_update()
    o.set_positions(p.positions())
    o.set_normals(p.normals())
    changed = _compare_all(p, o, p.tri_poly.astype(np.int32), off, pv, smooth, mat, fs, None, 0, 1)
    assert changed > 0
    assert sum(not np.array_equal(p.draw_leaf(l)[0], before[l]) for l in before) > 0
    # host marks: a vertex-coordinate replace marks every node; the update refills all
    new = p.positions() * F32(1.01)
    sg.lib().BKE_pbvh_vert_coords_apply(p.h, new.ctypes.data, len(new))
    p.draw_update()
    o.set_positions(p.positions())
    o.set_normals(p.normals())
    _compare_all(p, o, p.tri_poly.astype(np.int32), off, pv, smooth, mat, fs, None, 0, 1)
    p.close()
    o.close()
