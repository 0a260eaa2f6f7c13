"""Known-answer tests for the oracle's proxy combine (orc_combine_proxies): which vertex
a proxy entry lands on, the order of the additions, the original-coordinate base, and
the normal and bounds updates that follow, against closed forms and brute force."""
import numpy as np

F32 = np.float32


def _mesh(sg, po, level=3, leaf=16):
    co, tris = sg.noisy_icosphere(level, seed=5, amplitude=0.03)
    o = po.Oracle(co, tris, leaf_limit=leaf, fp64=False)
    return co, tris, o


def _uniq(o, node):
    return o.node_verts(node)[: o.node_info(node)[4]]


def test_proxy_entry_lands_on_its_vertex(sg, po):
    co, tris, o = _mesh(sg, po)
    leaves = o.leaves_dfs()
    node = int(leaves[len(leaves) // 2])
    uv = _uniq(o, node)
    # displacement i along x encodes the proxy index: vertex uv[i] must move by (i+1) * 2^-10
    d = np.zeros((len(uv), 3), F32)
    d[:, 0] = (np.arange(len(uv)) + 1) * F32(2.0 ** -10)
    o.combine_proxies([node], [[d]])
    got = o.positions()
    moved = np.nonzero(np.any(got != co, axis=1))[0]
    assert set(moved.tolist()) == set(uv.tolist())
    for i, v in enumerate(uv):
        assert got[v, 0] == F32(co[v, 0] + d[i, 0]) and got[v, 1] == co[v, 1] and got[v, 2] == co[v, 2]
    o.close()


def test_proxies_add_left_to_right_from_the_chosen_base(sg, po):
    """(base + p0) + p1 in float32, not base + (p0 + p1): with base 1, p0 = p1 = 2^-24 the
    first rounds to 1 twice, the second gives 1 + 2^-23."""
    co, tris, o = _mesh(sg, po)
    node = int(o.leaves_dfs()[0])
    uv = _uniq(o, node)
    orco = co.copy()
    orco[uv] = 1.0
    e = np.full((len(uv), 3), F32(2.0 ** -24))
    o.combine_proxies([node], [[e, e]], orco=orco)
    got = o.positions()
    assert np.all(got[uv] == F32(1.0))
    orco[uv] = 1.0
    o.combine_proxies([node], [[2 * e]], orco=orco)
    assert np.all(got[uv] != o.positions()[uv]) and np.all(o.positions()[uv] == F32(1.0 + 2.0 ** -23))
    # without orco the base is the current position
    cur = o.positions()
    z = np.zeros((len(uv), 3), F32)
    z[:, 2] = 0.5
    o.combine_proxies([node], [[z]])
    assert np.all(o.positions()[uv, 2] == cur[uv, 2] + F32(0.5))
    o.close()


def test_normals_and_bounds_after_combine(sg, po):
    co, tris, o = _mesh(sg, po, level=3, leaf=8)
    leaves = o.leaves_dfs()
    rng = np.random.default_rng(3)
    nodes = [int(x) for x in rng.choice(leaves, 5, replace=False)]
    props = []
    for nd in nodes:
        uv = _uniq(o, nd)
        props.append([rng.normal(0, 0.01, (len(uv), 3)).astype(F32) for _ in range(rng.integers(1, 4))])
    no0 = o.normals()
    o.combine_proxies(nodes, props)
    pos, no = o.positions(), o.normals()
    # normals: the uniq vertices of the combined nodes get normalize(sum of the unit
    # normals of their incident tris); every other vertex keeps its normal
    changed = np.concatenate([_uniq(o, nd) for nd in nodes])
    keep = np.ones(len(co), bool)
    keep[changed] = False
    assert np.array_equal(no[keep], no0[keep])
    p64 = pos.astype(np.float64)
    fn = np.cross(p64[tris[:, 0]] - p64[tris[:, 1]], p64[tris[:, 1]] - p64[tris[:, 2]])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    acc = np.zeros((len(co), 3))
    for k in range(3):
        np.add.at(acc, tris[:, k], fn)
    want = acc / np.linalg.norm(acc, axis=1, keepdims=True)
    assert np.abs(no[changed] - want[changed]).max() < 2e-6
    # bounds: a changed leaf's box is the box of all its vertices (uniq and shared);
    # every internal node contains its children
    for nd in nodes:
        vb, ob = o.node_bb(nd)
        allv = o.node_verts(nd)
        assert np.array_equal(vb[:3], pos[allv].min(0)) and np.array_equal(vb[3:], pos[allv].max(0))
        assert np.array_equal(vb, ob)
    root_vb, _ = o.node_bb(0)
    assert np.all(root_vb[:3] <= pos.min(0)) and np.all(root_vb[3:] >= pos.max(0))
    o.close()
