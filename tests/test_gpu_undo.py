"""GPU undo steps (sg_undo_*): snapshots of the leaves an operation touches.

Bar: bit-exact.  Undoing a step gives back the positions, normals and every node box
(current and original) the mesh had before the step; redoing it gives back the state
after the step.  The leaves a step saves are the union of the leaves the oracle's
sphere search gathers for the step's dabs (the leaves the reference would push with
SCULPT_undo_push_node), and the step's size follows from their unique vertices."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32


def _state(p):
    n = p.tree_info()["nodes"]
    bb = np.array([p.node_bb(i) for i in range(n)], F32)
    obb = np.array([p.node_bb(i, original=True) for i in range(n)], F32)
    return p.positions(), p.normals(), bb, obb


def _same(a, b):
    return all(np.array_equal(np.ascontiguousarray(x, F32).view(np.uint32),
                              np.ascontiguousarray(y, F32).view(np.uint32)) for x, y in zip(a, b))


def _stroke_dabs(sg, seed, n=16, r=0.15, tools=(1, 4, 2, 5)):
    radii = np.full(n, r, F32)
    centres = sg.great_circle_centres(seed, radii)
    rng = np.random.default_rng(seed)
    dabs = []
    for i in range(n):
        t = tools[i % len(tools)]
        g = tuple(rng.normal(0, 0.01, 3)) if t == sg.TOOL_GRAB else (0, 0, 0)
        dabs.append(sg.make_dab(tool=t, location=centres[i], radius=r, strength=0.6, smooth_iterations=3,
                                grab_delta=g))
    return dabs


def test_undo_redo_stroke_bit_exact_and_saved_leaves(sg, po):
    from parity_util import REL_TOL, rel_err

    co, tris = sg.noisy_icosphere(5, seed=21, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    o = po.Oracle(co, tris, leaf_limit=64)
    o.set_normals(p.normals())
    dabs = _stroke_dabs(sg, 3)
    # the leaves the stroke gathers, from the oracle run dab by dab
    want = set()
    for d in dabs:
        want |= set(o.search_sphere(list(d.location), d.radius, use_orig=(d.sculpt_tool == sg.TOOL_GRAB)).tolist())
        o.dab(d)
    s0 = _state(p)
    print("[undo] stroke of", len(dabs), "dabs in one step", flush=True)
    p.undo_begin()
    p.stroke(dabs)
    p.undo_end()
    s1 = _state(p)
    assert rel_err(s1[0], o.positions()) <= REL_TOL and rel_err(s1[1], o.normals()) <= REL_TOL
    info = p.undo_info()
    n_v = sum(p.node_verts(int(l))[1] for l in want)
    assert info["steps"] == 1 and info["applied"] == 1 and info["leaves"] == len(want)
    assert info["bytes"] == 24 * n_v + 80 * len(want)
    assert p.undo()
    assert _same(_state(p), s0)
    assert not p.undo()  # nothing left
    assert p.redo()
    assert _same(_state(p), s1)
    assert not p.redo()
    # undo, then a dab outside any step: the mesh continues from the undone state
    assert p.undo()
    p.dab(dabs[0])
    q = sg.PBVH(co, tris, leaf_limit=64)
    q.dab(dabs[0])
    assert _same(_state(p), _state(q))
    q.close()
    p.close()
    o.close()


def test_undo_several_steps_and_dropped_redo(sg):
    co, tris = sg.noisy_icosphere(5, seed=22, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=32)
    states = [_state(p)]
    for k in range(3):
        print("[undo] step", k, flush=True)
        p.undo_begin()
        p.stroke(_stroke_dabs(sg, 10 + k, n=12))
        p.undo_end()
        states.append(_state(p))
    assert p.undo_info()["steps"] == 3
    for k in (2, 1, 0):
        assert p.undo()
        assert _same(_state(p), states[k]), k
    assert not p.undo()
    for k in (1, 2):
        assert p.redo()
        assert _same(_state(p), states[k]), k
    # a new step drops the step still undone
    p.undo_begin()
    p.stroke(_stroke_dabs(sg, 40, n=8))
    p.undo_end()
    s_new = _state(p)
    info = p.undo_info()
    assert info["steps"] == 3 and info["applied"] == 3
    assert not p.redo()
    assert p.undo()
    assert _same(_state(p), states[2])
    assert p.redo()
    assert _same(_state(p), s_new)
    p.undo_clear()
    assert p.undo_info()["steps"] == 0 and not p.undo()
    p.close()


def test_counters_and_log_unchanged_while_recording(sg):
    """The open-step flag the dabs carry for the brush kernels does not leak into the
    gathered-leaf counts the host reads (per-dab log and last-dab statistics)."""
    co, tris = sg.noisy_icosphere(4, seed=27, amplitude=0.02)
    dabs = _stroke_dabs(sg, 4, n=8)
    p = sg.PBVH(co, tris, leaf_limit=64)
    q = sg.PBVH(co, tris, leaf_limit=64)
    p.undo_begin()
    p.stroke(dabs, log=True)
    lp, sp = p.stroke_log(len(dabs)), p.last_stats()
    p.undo_end()
    q.stroke(dabs, log=True)
    lq, sq = q.stroke_log(len(dabs)), q.last_stats()
    assert lp[:, 0].max() > 0 and lp[:, 0].max() < 1 << 32
    assert np.array_equal(lp, lq) and np.array_equal(sp, sq)
    assert _same(_state(p), _state(q))
    p.undo_begin()
    p.dab(dabs[0])
    assert p.last_stats()[0] == len(p.last_gathered()) > 0
    p.undo_end()
    p.close()
    q.close()


def test_undo_overflow_drops_step(sg):
    co, tris = sg.noisy_icosphere(4, seed=23, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    s0 = _state(p)
    p.undo_begin(max_verts=50)  # far less than one leaf set of a stroke
    p.stroke(_stroke_dabs(sg, 5, n=8))
    with pytest.raises(RuntimeError, match="overflow"):
        p.undo_end()
    assert p.undo_info()["steps"] == 0
    s1 = _state(p)
    # a later step with room records normally
    p.undo_begin()
    p.stroke(_stroke_dabs(sg, 6, n=8))
    p.undo_end()
    assert p.undo()
    assert _same(_state(p), s1)
    assert not _same(s1, s0)
    # the limit holds for the step that asks for it, whatever arena earlier steps left
    p.undo_begin(max_verts=50)
    p.stroke(_stroke_dabs(sg, 7, n=8))
    with pytest.raises(RuntimeError, match="overflow"):
        p.undo_end()
    p.close()


def test_undo_placed_stroke_in_graph_and_host_push(sg, po):
    """A captured placed stroke records its leaves at run time (the step opened after the
    graph was built), and a host edit is saved by pushing nodes explicitly."""
    co, tris = sg.noisy_icosphere(5, seed=24, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    n = 16
    rng = np.random.default_rng(7)
    ang = np.linspace(0, 1.0, n)
    starts = np.stack([np.cos(ang) * 3, np.sin(ang) * 3, 0.2 * np.ones(n)], 1).astype(F32)
    dirs = (-starts / np.linalg.norm(starts, axis=1, keepdims=True)).astype(F32)
    dabs = [sg.make_dab(tool=[1, 2, 4, 5][i % 4], location=(0, 0, 0), radius=0.12,
                        strength=float(rng.uniform(0.3, 0.9))) for i in range(n)]
    s0 = _state(p)
    st = p.prepare_placed(dabs, starts, dirs)
    p.undo_begin()
    print("[undo] prepared placed stroke", flush=True)
    st.launch()
    p.undo_end()
    s1 = _state(p)
    assert not _same(s1, s0)
    assert p.undo()
    assert _same(_state(p), s0)
    assert p.redo()
    assert _same(_state(p), s1)
    # host edit of every vertex (BKE_pbvh_vert_coords_apply): push every leaf first
    o = po.Oracle(co, tris, leaf_limit=64)
    leaves = o.leaves_dfs()
    o.close()
    p.undo_begin()
    with pytest.raises(RuntimeError, match="not a leaf"):
        p.undo_push_node(0)  # the root
    for i in leaves:
        p.undo_push_node(int(i))
    moved = np.ascontiguousarray(s1[0] * F32(1.01), F32)
    print("[undo] host coordinate edit", flush=True)
    sg.lib().BKE_pbvh_vert_coords_apply(p.h, moved.ctypes.data, len(moved))
    assert np.array_equal(p.positions(), moved)
    p.undo_end()
    assert p.undo_info()["leaves"] == len(leaves)
    assert p.undo()
    assert _same(_state(p), s1)
    p.close()


def test_undo_refreshes_draw_buffers_and_ray_cast(sg, po):
    co, tris = sg.noisy_icosphere(5, seed=25, amplitude=0.02)
    P = len(tris)
    off = np.arange(0, 3 * P + 1, 3)
    p = sg.PBVH(co, None, leaf_limit=64,
                polys=(off, tris.reshape(-1).astype(np.int32), np.ones(P, np.uint8), np.zeros(P, np.int32)))
    p.draw_build(np.ones(P, np.int32))
    o = po.Oracle(co, p.tris, leaf_limit=64)
    leaves = [int(i) for i in o.leaves_dfs()]
    o.close()

    def records():
        return [p.draw_leaf(i)[0] for i in leaves]

    rng = np.random.default_rng(2)
    d = rng.normal(size=(256, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    starts, dirs = (3 * d).astype(F32), (-d).astype(F32)
    r0, h0 = records(), p.raycast(starts, dirs)
    p.undo_begin()
    print("[undo] stroke with draw buffers", flush=True)
    p.stroke(_stroke_dabs(sg, 8, n=20))
    p.undo_end()
    p.draw_update()
    r1, h1 = records(), p.raycast(starts, dirs)
    assert any(not np.array_equal(a, b) for a, b in zip(r0, r1))
    assert p.undo()
    p.draw_update()
    assert all(np.array_equal(a, b) for a, b in zip(records(), r0))
    assert p.raycast(starts, dirs).tobytes() == h0.tobytes()
    assert p.redo()
    p.draw_update()
    assert all(np.array_equal(a, b) for a, b in zip(records(), r1))
    assert p.raycast(starts, dirs).tobytes() == h1.tobytes()
    p.close()


def test_undo_learned_dabs(sg):
    """Learned Smooth dabs save their leaves in the learned brush kernel."""
    co, tris = sg.noisy_icosphere(5, seed=26, amplitude=0.02)
    m = sg.MLP(seed=9)
    alpha, _ = m.decision()
    m.set_decision(alpha, 1e30)  # accept every learned dab
    p = sg.PBVH(co, tris, leaf_limit=64)
    p.set_learned(m)
    radii = np.full(10, 0.15, F32)
    c = sg.great_circle_centres(12, radii)
    dabs = [sg.make_dab(tool=sg.TOOL_SMOOTH, location=ci, radius=0.15, strength=0.6, smooth_iterations=2,
                        flag=sg.SMOOTH_LEARNED) for ci in c]
    s0 = _state(p)
    p.undo_begin()
    print("[undo] learned dabs", flush=True)
    p.stroke(dabs)
    p.undo_end()
    assert p.learned_stats()["accepted"] > 0
    s1 = _state(p)
    assert not _same(s1, s0)
    assert p.undo()
    assert _same(_state(p), s0)
    assert p.redo()
    assert _same(_state(p), s1)
    p.close()
    m.close()


def test_undo_cancel_and_original_data(sg):
    co, tris = sg.noisy_icosphere(5, seed=28, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    s0 = _state(p)
    assert _same(p.undo_original(), s0[:2])  # no open step: the current state
    p.undo_begin()
    print("[undo] stroke then cancel", flush=True)
    p.stroke(_stroke_dabs(sg, 14, n=12))
    assert not _same(_state(p), s0)
    assert _same(p.undo_original(), s0[:2])  # the state at begin
    p.undo_cancel()
    assert _same(_state(p), s0)
    assert p.undo_info()["steps"] == 0
    # a cancel keeps the steps before it
    p.undo_begin()
    p.stroke(_stroke_dabs(sg, 15, n=12))
    p.undo_end()
    s2 = _state(p)
    p.undo_begin()
    p.stroke(_stroke_dabs(sg, 16, n=12))
    p.undo_cancel()
    assert _same(_state(p), s2) and p.undo_info()["steps"] == 1
    assert p.undo()
    assert _same(_state(p), s0)
    with pytest.raises(RuntimeError, match="no undo step is open"):
        p.undo_cancel()
    p.close()


def test_combine_proxies_matches_oracle(sg, po):
    """Proxies on 40 leaves (1-3 each) folded in on the GPU, first from the original
    coordinates of the open step, then from the current ones: positions, normals and
    every node box equal the oracle's; undoing the step undoes the stroke and the
    combine."""
    from parity_util import n_bit_diff

    co, tris = sg.noisy_icosphere(5, seed=29, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    o = po.Oracle(co, tris, leaf_limit=64, fp64=False)
    s0 = _state(p)
    rng = np.random.default_rng(5)
    leaves = o.leaves_dfs()
    assert p.combine_proxies() == 0
    p.undo_begin()
    p.stroke(_stroke_dabs(sg, 17, n=10))
    o.set_positions(p.positions())
    o.set_normals(p.normals())
    for use_orco in (True, False):
        nodes = sorted(int(x) for x in rng.choice(leaves, 40, replace=False))
        props = []
        for nd in nodes:
            lst = []
            for _ in range(int(rng.integers(1, 4))):
                a = p.add_proxy(nd)
                assert a.shape == (p.node_verts(nd)[1], 3) and not a.any()
                a[:] = rng.normal(0, 0.01, a.shape)
                lst.append(a.copy())
            props.append(lst)
        print(f"[undo] combine proxies use_orco={use_orco}", flush=True)
        assert p.combine_proxies(use_orco=use_orco) == len(nodes)
        o.combine_proxies(nodes, props, orco=s0[0] if use_orco else None)
        pp, op = p.positions(), o.positions()
        assert n_bit_diff(pp, op) == 0
        assert n_bit_diff(p.normals(), o.normals()) == 0
        bad = 0
        for i in range(o.totnode()):
            vb, ob = o.node_bb(i)
            bad += n_bit_diff(p.node_bb(i), vb) != 0 or n_bit_diff(p.node_bb(i, original=True), ob) != 0
        assert bad == 0
        assert p.combine_proxies() == 0  # the proxies were freed
    p.undo_end()
    assert p.undo()
    assert _same(_state(p), s0)
    p.close()
    o.close()


def test_combine_proxies_across_arena_chunks(sg, po):
    """Proxies are carved from 16 MiB pinned chunks: 9 proxies on every leaf of a 164k-
    vertex mesh (1.5 M proxy vertices, two chunks), added round-robin over the leaves so
    each leaf's proxies are scattered over both chunks, fold in proxy order, bit-equal
    to the oracle's; a second, smaller combine reuses the reset arena."""
    from parity_util import n_bit_diff

    co, tris = sg.noisy_icosphere(7, seed=31, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    o = po.Oracle(co, tris, leaf_limit=64, fp64=False)
    leaves = [int(x) for x in o.leaves_dfs()]
    rng = np.random.default_rng(8)
    for rnd, (nprox, nodes) in enumerate(((9, leaves), (2, leaves[::7]))):
        props = {nd: [] for nd in nodes}
        for _ in range(nprox):
            for nd in nodes:
                a = p.add_proxy(nd)
                assert not a.any()
                a[:] = rng.normal(0, 1e-3, a.shape)
                props[nd].append(a.copy())
        print(f"[undo] combine across arena chunks round {rnd}: {len(nodes)} leaves x {nprox}", flush=True)
        assert p.combine_proxies() == len(nodes)
        o.combine_proxies(nodes, [props[nd] for nd in nodes])
        assert n_bit_diff(p.positions(), o.positions()) == 0
        assert n_bit_diff(p.normals(), o.normals()) == 0
        vb, _ = o.node_bb(0)
        assert n_bit_diff(p.node_bb(0), vb) == 0
    p.close()
    o.close()


@pytest.mark.parametrize("in_step", [True, False])
def test_grab_from_original_coordinates(sg, po, in_step):
    """SG_DAB_ORIGINAL Grab dabs at an anchored centre with growing deltas: inside an
    undo step they work from the step's original coordinates, outside one from the
    positions before each dab; every dab's written set, gathered leaves and positions
    equal the oracle's, and at the end the normals and every node box too."""
    from parity_util import n_bit_diff

    co, tris = sg.noisy_icosphere(5, seed=31, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    o = po.Oracle(co, tris, leaf_limit=64, fp64=False)
    o.set_normals(p.normals())
    c = co[777].copy()
    dabs = [sg.make_dab(tool=sg.TOOL_GRAB, location=c, radius=0.25, strength=0.8,
                        grab_delta=(0.004 * k, -0.003 * k, 0.002 * k), flag=sg.DAB_ORIGINAL) for k in range(1, 9)]
    s0 = _state(p)
    if in_step:
        p.undo_begin()
    print(f"[undo] original-coordinate grab, in_step={in_step}", flush=True)
    for i, d in enumerate(dabs):
        orco = s0[0] if in_step else o.positions()
        want_leaves = o.search_sphere(list(d.location), d.radius, use_orig=True)
        want = o.dab_grab_orig(d, orco)
        p.dab(d)
        assert np.array_equal(p.last_gathered(), want_leaves), i
        assert np.array_equal(p.last_affected(), want), i
        assert n_bit_diff(p.positions(), o.positions()) == 0, i
    assert n_bit_diff(p.normals(), o.normals()) == 0
    bad = 0
    for i in range(o.totnode()):
        vb, ob = o.node_bb(i)
        bad += n_bit_diff(p.node_bb(i), vb) != 0 or n_bit_diff(p.node_bb(i, original=True), ob) != 0
    assert bad == 0
    if in_step:
        # the same dabs as one captured stroke inside a new step give the same bits
        end = _state(p)
        p.undo_end()
        assert p.undo()
        assert _same(_state(p), s0)
        p.undo_begin()
        p.stroke(dabs)
        p.undo_end()
        assert _same(_state(p), end)
    p.close()
    o.close()


def test_restore_original_for_anchored_strokes(sg):
    """Inside a step, restoring the original mesh gives back the begin state bit for
    bit and keeps the step open; a dab after it equals the same dab on a fresh mesh,
    and undoing the step returns to the begin state."""
    co, tris = sg.noisy_icosphere(5, seed=32, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    s0 = _state(p)
    c = co[2024].copy()
    grow = [sg.make_dab(tool=sg.TOOL_DRAW, location=c, radius=0.05 * k, strength=0.7) for k in range(1, 5)]
    p.undo_begin()
    print("[undo] anchored stroke", flush=True)
    for d in grow[:-1]:
        p.dab(d)
        p.undo_restore_original()
        assert _same(_state(p), s0)
    p.dab(grow[-1])
    p.undo_end()
    q = sg.PBVH(co, tris, leaf_limit=64)
    q.dab(grow[-1])
    assert _same(_state(p), _state(q))
    assert p.undo_info()["steps"] == 1
    assert p.undo()
    assert _same(_state(p), s0)
    with pytest.raises(RuntimeError, match="no undo step is open"):
        p.undo_restore_original()
    q.close()
    p.close()
