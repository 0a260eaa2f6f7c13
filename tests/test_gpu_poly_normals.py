"""Normals of polygon meshes on the GPU against the oracle (no normals copied between
the two): the initial normals weight each polygon's Newell normal by its corner angles
(mesh_normals.c:225-299), and after each dab an affected vertex sums the normal of each
incident looptri's polygon (BKE_mesh_calc_poly_normal, pbvh.c:1041-1058).  Bar:
initial normals within 1e-5 relative (acosf may differ in the last bit), affected sets
bit-exact, positions, normals and node boxes within 1e-5 relative after every dab; and,
with the oracle started from the GPU's initial normals, positions and every node box
bit-exact."""
import numpy as np
import pytest

from parity_util import REL_TOL, n_bit_diff, rel_err

pytestmark = pytest.mark.gpu


def _torus_quads(sg, nu, nv, seed, amp):
    co, _ = sg.noisy_torus(nu, nv, seed=seed, amplitude=amp)
    pv = []
    for i in range(nu):
        for j in range(nv):
            pv += [i * nv + j, ((i + 1) % nu) * nv + j, ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv]
    return co, np.arange(0, len(pv) + 1, 4), np.array(pv, np.int32)


def _mixed(sg, co, off, pv, every=5):
    """Every `every`-th quad becomes two triangles and every other one a pentagon with an
    extra vertex lifted off the middle of its third edge."""
    extra = []
    noff, npv = [0], []
    base = co.shape[0]
    for q in range(len(off) - 1):
        c = pv[off[q]:off[q + 1]]
        if q % every == 0:
            npv += [c[0], c[1], c[2]]
            noff.append(len(npv))
            npv += [c[0], c[2], c[3]]
            noff.append(len(npv))
        elif q % every == 2:
            m = (co[c[2]] + co[c[3]]) * 0.5 * 1.01
            extra.append(m)
            npv += [c[0], c[1], c[2], base + len(extra) - 1, c[3]]
            noff.append(len(npv))
        else:
            npv += list(c)
            noff.append(len(npv))
    co2 = np.concatenate([co, np.array(extra, np.float32)]) if extra else co
    return co2, np.array(noff), np.array(npv, np.int32)


def _run(sg, po, co, off, pv, dabs, leaf, seed_normals=False):
    """seed_normals=False: each side keeps its own initial normals (acosf may differ in
    the last bit, and Inflate carries that into positions, so boxes are compared within
    the tolerance); True: the oracle starts from the GPU's normals and every node box
    must be bit-identical."""
    P = len(off) - 1
    p = sg.PBVH(co, None, leaf_limit=leaf, polys=(off, pv, np.ones(P, np.uint8), np.zeros(P, np.int32)))
    o = po.Oracle(co, p.tris, leaf_limit=leaf, polys=(p.tri_poly.astype(np.int32), off, pv))
    e0 = rel_err(p.normals(), o.normals())
    assert e0 <= REL_TOL, f"initial normals {e0}"
    if seed_normals:
        o.set_normals(p.normals())
    worst_p = worst_n = 0.0
    for i, d in enumerate(dabs):
        want = o.dab(d)
        p.dab(d)
        assert np.array_equal(p.last_affected(), want), f"dab {i}: affected set differs"
        worst_p = max(worst_p, rel_err(p.positions(), o.positions()))
        worst_n = max(worst_n, rel_err(p.normals(), o.normals()))
        assert worst_p <= REL_TOL and worst_n <= REL_TOL, (i, worst_p, worst_n)
    nbad = 0
    for i in range(o.totnode()):
        vb, ob = o.node_bb(i)
        gb, go = p.node_bb(i), p.node_bb(i, original=True)
        nbad += n_bit_diff(gb, vb) + n_bit_diff(go, ob)
        assert rel_err(gb, vb) <= REL_TOL and rel_err(go, ob) <= REL_TOL, i
    if seed_normals:
        assert nbad == 0 and n_bit_diff(p.positions(), o.positions()) == 0
    print(f"[poly] {P} polygons, {len(dabs)} dabs, seeded normals {seed_normals}: initial normals rel {e0:.3g}, "
          f"positions rel {worst_p:.3g}, normals rel {worst_n:.3g}, box words differing {nbad}")
    p.close()
    o.close()


def _dabs(sg, co, n, seed):
    rng = np.random.default_rng(seed)
    pts = co[rng.integers(0, co.shape[0], n)]
    tools = [sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH, sg.TOOL_GRAB]
    return [sg.make_dab(tool=tools[i % 4], location=pts[i], radius=float(rng.uniform(0.08, 0.2)),
                        strength=float(rng.uniform(0.3, 1.0)), smooth_iterations=3,
                        grab_delta=rng.normal(size=3) * 0.01) for i in range(n)]


@pytest.mark.parametrize("seed_normals", [False, True])
def test_quad_torus_normals(sg, po, seed_normals):
    co, off, pv = _torus_quads(sg, 120, 36, seed=3, amp=0.01)
    _run(sg, po, co, off, pv, _dabs(sg, co, 24, 1), leaf=64, seed_normals=seed_normals)


@pytest.mark.parametrize("seed_normals", [False, True])
def test_mixed_polygons_normals(sg, po, seed_normals):
    co, off, pv = _torus_quads(sg, 90, 30, seed=5, amp=0.015)
    co, off, pv = _mixed(sg, co, off, pv)
    assert (np.diff(off) == 5).any() and (np.diff(off) == 3).any() and (np.diff(off) == 4).any()
    _run(sg, po, co, off, pv, _dabs(sg, co, 20, 2), leaf=32, seed_normals=seed_normals)
