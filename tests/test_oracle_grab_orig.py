"""Known-answer tests for the oracle's Grab from the original coordinates
(orc_dab_grab_orig): it reduces to the plain Grab when the original coordinates are
the current ones, it resets every visible uniq vertex of a gathered leaf to its
original coordinate when the delta is zero, and a stroke of such dabs at one anchored
centre ends at original + last delta * fade whatever the dabs before it did."""
import numpy as np

F32 = np.float32


def _setup(sg, po, seed=7, level=4, leaf=16):
    co, tris = sg.noisy_icosphere(level, seed=seed, amplitude=0.02)
    return co, tris, po.Oracle(co, tris, leaf_limit=leaf, fp64=False)


def _written(o, d):
    """Visible uniq vertices of the leaves the Grab search gathers (brute force)."""
    out = []
    for nd in o.search_sphere(list(d.location), d.radius, use_orig=True):
        out += o.node_verts(int(nd))[: o.node_info(int(nd))[4]].tolist()
    return np.array(sorted(out), np.int32)


def test_reduces_to_plain_grab_from_current_coordinates(sg, po):
    co, tris, a = _setup(sg, po)
    b = po.Oracle(co, tris, leaf_limit=16, fp64=False)
    b.set_normals(a.normals())
    d = sg.make_dab(tool=sg.TOOL_GRAB, location=co[17], radius=0.3, strength=0.8, grab_delta=(0.01, -0.02, 0.005))
    want_written = _written(a, d)
    aff_plain = b.dab(d)
    got = a.dab_grab_orig(d, a.positions())
    assert np.array_equal(got, want_written)
    assert set(aff_plain.tolist()) <= set(got.tolist()) and len(aff_plain) < len(got)
    assert np.array_equal(a.positions().view(np.uint32), b.positions().view(np.uint32))
    assert np.array_equal(a.normals()[aff_plain].view(np.uint32), b.normals()[aff_plain].view(np.uint32))
    a.close()
    b.close()


def test_zero_delta_resets_written_vertices_to_original(sg, po):
    co, tris, o = _setup(sg, po)
    orco = co.copy()
    moved = (co * F32(1.01)).astype(F32)
    o.set_positions(moved)
    d = sg.make_dab(tool=sg.TOOL_GRAB, location=moved[40], radius=0.25, strength=1.0, grab_delta=(0, 0, 0))
    want = _written(o, d)
    got = o.dab_grab_orig(d, orco)
    assert np.array_equal(got, want)
    pos = o.positions()
    assert np.array_equal(pos[got], orco[got])
    rest = np.setdiff1d(np.arange(len(co)), got)
    assert np.array_equal(pos[rest], moved[rest])
    o.close()


def test_anchored_stroke_ends_at_original_plus_last_delta(sg, po):
    co, tris, o = _setup(sg, po)
    orco = co.copy()
    c = co[123].copy()
    r = F32(0.3)
    deltas = [np.array([0.002 * k, -0.001 * k, 0.0015 * k], F32) for k in range(1, 7)]
    for g in deltas:
        d = sg.make_dab(tool=sg.TOOL_GRAB, location=c, radius=float(r), strength=0.7, grab_delta=tuple(g))
        o.dab_grab_orig(d, orco)
    pos = o.positions()
    dist = np.sqrt(((c - orco).astype(np.float64) ** 2).sum(1))
    inner = np.nonzero(dist < r - 0.03)[0]
    assert len(inner) > 20
    for v in inner:
        e = (c - orco[v]).astype(F32)
        d2 = F32(e[0] * e[0] + e[1] * e[1] + e[2] * e[2])
        fade = F32(0.7) * F32(po.falloff(d, d2))
        assert np.array_equal(pos[v], (orco[v] + deltas[-1] * fade).astype(F32)), v
    o.close()
