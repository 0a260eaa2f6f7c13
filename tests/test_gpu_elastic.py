"""GPU elastic deform (SCULPT_TOOL_ELASTIC_DEFORM, the Kelvinlets of
kernel/intern/kelvinlet.c) against the oracle: every mode, with a mask and hidden
vertices, from the open undo step's original coordinates and from the current ones.
Bar: bit-exact positions, normals and node boxes; the written set equals the oracle's."""
import numpy as np
import pytest

from parity_util import n_bit_diff

pytestmark = pytest.mark.gpu

F32 = np.float32


def _bbs_equal(p, o):
    for i in range(o.totnode()):
        vb, ob = o.node_bb(i)
        if n_bit_diff(p.node_bb(i), vb) or n_bit_diff(p.node_bb(i, original=True), ob):
            return False
    return True


@pytest.mark.parametrize("etype", [0, 1, 2, 3, 4])
def test_elastic_deform_matches_oracle(sg, po, etype):
    co, tris = sg.noisy_icosphere(5, seed=41 + etype, amplitude=0.02)
    hidden = np.zeros(len(co), np.uint8)
    hidden[::23] = 1
    mask = (np.random.default_rng(etype).random(len(co)) * 0.8).astype(F32)
    p = sg.PBVH(co, tris, leaf_limit=64, hidden=hidden.astype(bool))
    p.set_mask(mask)
    o = po.Oracle(co, tris, leaf_limit=64, fp64=False, hidden=hidden)
    o.set_mask(mask)
    # a plain stroke first, so the original and current coordinates differ
    radii = np.full(6, 0.2, F32)
    p.stroke(sg.stroke_config([1, 4, 2, 5, 1, 4], sg.great_circle_centres(3, radii), radii, np.full(6, 0.6, F32)))
    s0 = p.positions()
    dabs = [sg.make_dab(tool=sg.TOOL_ELASTIC_DEFORM, location=s0[100 * (k + 1)], radius=0.3, strength=0.8,
                        grab_delta=(0.01 * (k + 1), -0.005, 0.004), elastic_type=etype, volume_preservation=0.3,
                        surface_normal=(0.0, 0.6, 0.8), flag=sg.BRUSH_DIR_IN if k == 1 else 0) for k in range(3)]
    for in_step in (True, False):
        o.set_positions(p.positions())
        o.set_normals(p.normals())
        begin = p.positions()
        if in_step:
            p.undo_begin()
        print(f"[elastic] mode {etype} in_step={in_step}", flush=True)
        for i, d in enumerate(dabs):
            orco = begin if in_step else p.positions()
            want = o.dab_elastic(d, orco)
            p.dab(d)
            assert np.array_equal(p.last_affected(), want), (i, in_step)
            assert n_bit_diff(p.positions(), o.positions()) == 0, (i, in_step)
        assert n_bit_diff(p.normals(), o.normals()) == 0
        assert _bbs_equal(p, o)
        if in_step:
            p.undo_end()
            after = p.positions()
            assert p.undo()
            assert np.array_equal(p.positions(), begin)
            assert p.redo()
            assert np.array_equal(p.positions(), after)
    p.close()
    o.close()


def test_elastic_in_a_captured_stroke(sg):
    """Elastic dabs between plain dabs in one captured stroke give the same bits as the
    same dabs one call at a time."""
    co, tris = sg.noisy_icosphere(5, seed=47, amplitude=0.02)
    dabs = []
    for k in range(6):
        if k % 2:
            dabs.append(sg.make_dab(tool=sg.TOOL_ELASTIC_DEFORM, location=co[50 * k], radius=0.25, strength=0.6,
                                    grab_delta=(0.0, 0.01, 0.0), elastic_type=k % 5))
        else:
            dabs.append(sg.make_dab(tool=[1, 4, 5][k // 2], location=co[50 * k], radius=0.15, strength=0.6,
                                    grab_delta=(0.01, 0.0, 0.0)))
    a = sg.PBVH(co, tris, leaf_limit=64)
    a.stroke(dabs)
    b = sg.PBVH(co, tris, leaf_limit=64)
    for d in dabs:
        b.dab(d)
    assert n_bit_diff(a.positions(), b.positions()) == 0
    assert n_bit_diff(a.normals(), b.normals()) == 0
    a.close()
    b.close()
