"""Known-answer tests for the oracle's elastic deform (oracle/elastic.c, the Kelvinlets
of kernel/intern/kelvinlet.c): closed forms at the brush centre, the far-field decay,
the geometry of scale and twist, the reference's fourth integration stage, the mask and
the written set."""
import math

import numpy as np
import pytest

F32 = np.float32


def _dab(sg, etype, loc=(0.0, 0.0, 0.0), r=0.3, s=0.7, delta=(0.02, -0.01, 0.03), nu=0.4, n=(0, 0, 1), flag=0):
    return sg.make_dab(tool=sg.TOOL_ELASTIC_DEFORM, location=loc, radius=r, strength=s, grab_delta=delta,
                       elastic_type=etype, volume_preservation=nu, surface_normal=n, flag=flag)


def _abc(nu):
    a = 1.0 / (4.0 * math.pi)
    b = a / (4.0 * (1.0 - nu))
    return a, b, 2.0 * (3.0 * a - 2.0 * b)


def test_grab_modes_at_the_centre(sg, po):
    """Formula (6) at distance 0: fade = c (a - b + a/2) / r; the bi- and tri-scale
    normalisations make their fade c (a - b + a/2), i.e. r times the single-scale one."""
    r, s, nu = 0.3, 0.7, 0.4
    a, b, c = _abc(nu)
    delta = np.array([0.02, -0.01, 0.03])
    want_grab = delta * (c * (1.5 * a - b) / r) * s * 20.0
    got = po.kelvinlet(_dab(sg, 0, r=r, s=s, nu=nu), (0, 0, 0), s)
    assert np.allclose(got, want_grab, rtol=2e-6, atol=0)
    for mode in (1, 2):
        got = po.kelvinlet(_dab(sg, mode, r=r, s=s, nu=nu), (0, 0, 0), s)
        assert np.allclose(got, want_grab * r, rtol=2e-5, atol=0), mode


def test_grab_far_field_decays_as_one_over_distance(sg, po):
    d = _dab(sg, 0, r=0.05)
    u1 = np.linalg.norm(po.kelvinlet(d, (40.0, 0, 0), 0.7))
    u2 = np.linalg.norm(po.kelvinlet(d, (80.0, 0, 0), 0.7))
    assert abs(u1 / u2 - 2.0) < 1e-3
    # the displacement is along the drag
    g = po.kelvinlet(d, (0.1, 0.2, -0.05), 0.7)
    assert np.allclose(np.cross(g, [0.02, -0.01, 0.03]), 0, atol=1e-7 * np.linalg.norm(g) + 1e-12)


def test_scale_is_radial_and_odd_twist_is_tangential(sg, po):
    rng = np.random.default_rng(1)
    n = np.array([0.3, -0.4, 0.866], np.float32)
    for _ in range(20):
        p = rng.normal(size=3).astype(F32) * F32(0.2)
        sc = po.kelvinlet(_dab(sg, 3), p, 0.7)
        assert np.array_equal(po.kelvinlet(_dab(sg, 3), -p, 0.7), -sc)  # exact odd symmetry about loc = 0
        assert np.linalg.norm(np.cross(sc, p)) <= 1e-6 * np.linalg.norm(sc) * np.linalg.norm(p)
        tw = po.kelvinlet(_dab(sg, 4, n=tuple(n)), p, 0.7)
        assert abs(np.dot(tw, n)) <= 1e-6 * np.linalg.norm(tw)
    # the direction flag reverses the force: the displacement points the other way (its
    # size differs, the integration stages being evaluated at the displaced points)
    d1, d2 = _dab(sg, 3), _dab(sg, 3, flag=sg.BRUSH_DIR_IN)
    p = np.array([0.1, 0.05, 0.0], F32)
    g1, g2 = po.kelvinlet(d1, p, 0.7), po.kelvinlet(d2, p, 0.7)
    assert np.dot(g1, g2) < -0.99 * np.linalg.norm(g1) * np.linalg.norm(g2)


def _integrate64(p, loc, nu, r, force, twist, nor, quirk):
    a, b, c = _abc(nu)

    def f(x):
        rv = x - loc
        rad = np.linalg.norm(rv)
        re = math.sqrt(rad * rad + r * r)
        lead = -a if twist else (2.0 * b - a)
        u = lead / re ** 3 + 3.0 * r * r / (2.0 * re ** 5)
        v = np.cross(nor, rv) if twist else rv
        return v * (u * c * force)

    k0 = f(p)
    k1 = f(p + 0.5 * k0)
    k2 = f(p + 0.5 * k1)
    k3 = f(p + k2 - loc) if quirk else f(p + k2)
    return (k0 + 2 * k1 + 2 * k2 + k3) / 6.0


def test_scale_fourth_stage_is_offset_by_the_location(sg, po):
    """The reference evaluates the fourth stage at vertex + k2 - location: the oracle
    follows that, not the textbook stage at vertex + k2 (the two differ off the origin)."""
    loc = np.array([0.4, -0.2, 0.1])
    nu, r, s = 0.3, 0.25, 0.9
    dab = _dab(sg, 3, loc=tuple(loc), r=r, s=s, nu=nu, delta=(0.5, 0, 0))
    force = 0.5 * s
    p = np.array([0.55, -0.1, 0.2])
    got = po.kelvinlet(dab, p.astype(F32), s).astype(np.float64)
    quirk = _integrate64(p, loc, nu, r, force, False, None, True)
    plain = _integrate64(p, loc, nu, r, force, False, None, False)
    assert np.allclose(got, quirk, rtol=1e-4)
    assert np.abs(quirk - plain).max() > 100 * np.abs(got - quirk).max()


def test_elastic_dab_writes_every_visible_uniq_vertex(sg, po):
    co, tris = sg.noisy_icosphere(3, seed=4, amplitude=0.02)
    hidden = np.zeros(len(co), np.uint8)
    hidden[::17] = 1
    o = po.Oracle(co, tris, leaf_limit=16, fp64=False, hidden=hidden)
    mask = np.zeros(len(co), F32)
    mask[::5] = 1.0
    o.set_mask(mask)
    orco = (co * F32(1.001)).astype(F32)
    d = _dab(sg, 0, loc=tuple(co[10]))
    got = o.dab_elastic(d, orco)
    want = np.array(sorted(v for nd in o.leaves_dfs() for v in o.node_verts(int(nd))[: o.node_info(int(nd))[4]]
                           if not hidden[v]), np.int32)
    assert np.array_equal(got, want)
    pos = o.positions()
    full = mask[got] == 1.0
    assert np.array_equal(pos[got[full]], orco[got[full]])  # mask 1: back to the original
    v = got[~full][3]
    assert np.array_equal(pos[v], (orco[v] + po.kelvinlet(d, orco[v], 0.7)).astype(F32))
    hid = np.nonzero(hidden)[0]
    assert np.array_equal(pos[hid], co[hid])  # hidden vertices keep their positions
    o.close()
