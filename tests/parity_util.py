"""Helpers shared by the GPU parity tests: run the same dabs through the HIP engine
(C ABI) and the CPU oracle and compare."""
from __future__ import annotations

import numpy as np

REL_TOL = 1e-5  # positions and normals: relative to the largest coordinate magnitude


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return float(np.abs(a - b).max() / scale)


def n_bit_diff(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return int((a.view(np.uint32) != b.view(np.uint32)).sum())


def run_pair(sg, po, co, tris, dabs, leaf_limit=128, check_each=True, check_bb=True, mask=None,
             same_start_normals=True, hidden=None, mat=None):
    """Apply dabs one by one on both; assert affected sets and gathered leaves are
    identical per dab and positions/normals agree.  Returns a report dict."""
    p = sg.PBVH(co, tris, leaf_limit=leaf_limit, hidden=hidden, mat=mat)
    o = po.Oracle(co, tris, leaf_limit=leaf_limit, fp64=True, hidden=hidden, mat=mat)
    if mask is not None:
        p.set_mask(mask)
        o.set_mask(mask)
    rep = {"dabs": 0, "affected": 0, "max_pos_rel": 0.0, "max_nor_rel": 0.0, "pos_bits": 0, "nor_bits": 0}
    n0 = p.normals()
    rep["init_nor_rel"] = rel_err(n0, o.normals())
    assert rep["init_nor_rel"] <= REL_TOL
    if same_start_normals:
        # start the dab comparison from bit-identical state (initial normals use acosf,
        # whose last bit may differ between the GPU and glibc)
        o.set_normals(n0)
    for i, d in enumerate(dabs):
        want_leaves = o.search_sphere(list(d.location), d.radius, use_orig=(d.sculpt_tool == sg.TOOL_GRAB))
        want = o.dab(d)
        p.dab(d)
        got = p.last_affected()
        assert np.array_equal(got, want), f"dab {i}: affected set differs ({len(got)} vs {len(want)})"
        got_leaves = p.last_gathered()
        assert np.array_equal(got_leaves, want_leaves), f"dab {i}: gathered leaves differ"
        rep["dabs"] += 1
        rep["affected"] += len(want)
        if check_each or i == len(dabs) - 1:
            pp, op = p.positions(), o.positions()
            pn, on = p.normals(), o.normals()
            e = rel_err(pp, op)
            assert e <= REL_TOL, f"dab {i}: positions rel err {e}"
            rep["max_pos_rel"] = max(rep["max_pos_rel"], e)
            # normals are only recomputed for affected vertices; untouched vertices
            # keep their initial (angle-weighted) normals on both sides
            en = rel_err(pn, on)
            assert en <= REL_TOL, f"dab {i}: normals rel err {en}"
            rep["max_nor_rel"] = max(rep["max_nor_rel"], en)
            rep["pos_bits"] = n_bit_diff(pp, op)
            rep["nor_bits"] = n_bit_diff(pn, on)
    if check_bb:
        bad = 0
        for i in range(o.totnode()):
            vb, ob = o.node_bb(i)
            if n_bit_diff(p.node_bb(i), vb) or n_bit_diff(p.node_bb(i, original=True), ob):
                bad += 1
        rep["bb_mismatch"] = bad
        assert bad == 0, f"{bad} node bounds differ"
    rep["fp64_pos_rel_gpu"] = rel_err(p.positions(), o.positions64())
    rep["fp64_pos_rel_oracle"] = rel_err(o.positions(), o.positions64())
    rep["fp64_nor_rel_gpu"] = rel_err(p.normals(), o.normals64())
    p.close()
    o.close()
    return rep
