"""GPU parity at the full sizes of the four GPU workloads, against the C oracle.

- plain strokes: the 10,485,762-vertex noisy icosphere and the benchmark's 1000-dab
  Draw / Inflate / Smooth(4) / Grab stroke.  The stroke runs twice on the GPU: as the
  captured graph the benchmark times (per-dab counters logged) and dab by dab.  Every
  dab's affected count equals the oracle's; every 50th dab's affected set and gathered
  leaves are identical; the final positions and normals are within 1e-5 relative and
  every node box (current and original) is bit-identical; both GPU runs end bit-identical.
- learned Smooth: 500 Smooth(4) dabs of radius 0.05 on the same mesh; the plain brush
  is checked against the oracle dab by dab, and every learned dab is judged against
  the plain dab from the same state (RMS error / r <= 1e-3 when accepted), with the
  error also stated relative to the plain displacement (a no-op scores 1).
- many small meshes (256 meshes of 5k-500k vertices, 64 Draw / Smooth steps): eight
  sampled meshes (smallest, largest, icospheres and tori) against the oracle.
- many large meshes (64 meshes of 1-4M vertices, 128 Draw / Inflate / Smooth steps):
  four sampled meshes of the plain table against the oracle; the learned table's
  accept / reject counts, one sampled mesh bit-identical to the single-mesh engine, and
  the judge on that mesh's Smooth dabs.

The sizes are the workloads' own; each test prints what it measured.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from parity_util import REL_TOL, n_bit_diff, rel_err

pytestmark = pytest.mark.gpu

CHECK_EVERY = 50


def _log(msg):
    print(f"[configs] {msg}", flush=True)


def _node_boxes(obj, n, gpu):
    """All node boxes (current, original) as two (n, 6) float32 arrays."""
    cur = np.zeros((n, 6), np.float32)
    org = np.zeros((n, 6), np.float32)
    for i in range(n):
        if gpu:
            cur[i] = obj.node_bb(i)
            org[i] = obj.node_bb(i, original=True)
        else:
            cur[i], org[i] = obj.node_bb(i)
    return cur, org


def _build_pair(sg, po, co, tris, fp64=True):
    """The GPU PBVH and the oracle for the same mesh, built side by side."""
    with ThreadPoolExecutor(2) as ex:
        fo = ex.submit(po.Oracle, co, tris, 128, fp64)
        p = sg.PBVH(co, tris, leaf_limit=128)
        o = fo.result()
    return p, o


@pytest.fixture(scope="module")
def big_mesh(sg):
    t = time.perf_counter()
    co, tris = sg.noisy_icosphere(10, seed=1, amplitude=0.02)
    _log(f"icosphere-10 noisy: {co.shape[0]} V, {tris.shape[0]} T in {time.perf_counter() - t:.1f}s")
    return co, tris


def test_plain_stroke_full_size(sg, po, big_mesh):
    from dune_sculpt_amd import workloads

    co, tris = big_mesh
    assert co.shape[0] == 10_485_762 and tris.shape[0] == 20_971_520
    dabs, tools = workloads.plain_stroke(sg, 1000)
    t = time.perf_counter()
    p, o = _build_pair(sg, po, co, tris)
    _log(f"built GPU PBVH + oracle in {time.perf_counter() - t:.1f}s: {p.tree_info()}")
    n0 = p.normals()
    e0 = rel_err(n0, o.normals())
    assert e0 <= REL_TOL, f"initial normals {e0}"
    o.set_normals(n0)  # start from bit-identical state
    p.snapshot()

    # 1) the captured stroke graph the benchmark times, per-dab counters logged
    t = time.perf_counter()
    p.stroke(dabs, graph=True, log=True)
    logrows = p.stroke_log(len(dabs))
    g_pos, g_nor = p.positions(), p.normals()
    _log(f"graph stroke in {time.perf_counter() - t:.2f}s")
    p.restore()

    # 2) dab by dab beside the oracle
    t = time.perf_counter()
    total = 0
    for i, d in enumerate(dabs):
        full = i % CHECK_EVERY == 0 or i == len(dabs) - 1
        if full:
            want_leaves = o.search_sphere(list(d.location), d.radius, use_orig=(d.sculpt_tool == sg.TOOL_GRAB))
        want = o.dab(d)
        p.dab(d)
        total += len(want)
        assert int(logrows[i, 1]) == len(want), f"dab {i} (tool {tools[i]}): graph affected count " \
                                                f"{int(logrows[i, 1])} vs oracle {len(want)}"
        assert int(p.last_stats()[1]) == len(want), f"dab {i}: affected count differs"
        if full:
            assert np.array_equal(p.last_affected(), want), f"dab {i}: affected set differs"
            assert np.array_equal(p.last_gathered(), want_leaves), f"dab {i}: gathered leaves differ"
        if i % 250 == 249:
            _log(f"{i + 1} dabs checked ({time.perf_counter() - t:.1f}s)")
    _log(f"1000 dabs beside the oracle in {time.perf_counter() - t:.1f}s, {total} vertex updates")
    pos, nor = p.positions(), p.normals()
    opos, onor = o.positions(), o.normals()
    ep, en = rel_err(pos, opos), rel_err(nor, onor)
    assert ep <= REL_TOL and en <= REL_TOL, f"positions {ep}, normals {en}"
    # the graph run and the dab-by-dab run give the same bits
    assert n_bit_diff(pos, g_pos) == 0 and n_bit_diff(nor, g_nor) == 0
    nn = o.totnode()
    gc, go = _node_boxes(p, nn, True)
    oc, oo = _node_boxes(o, nn, False)
    assert n_bit_diff(gc, oc) == 0 and n_bit_diff(go, oo) == 0, "node boxes differ"
    e64 = rel_err(pos, o.positions64())
    _log(f"positions rel {ep:.3g} ({n_bit_diff(pos, opos)} words differ), normals rel {en:.3g} "
         f"({n_bit_diff(nor, onor)} differ), {nn} node boxes bit-exact; FP64 replay: GPU positions rel {e64:.3g}, "
         f"normals rel {rel_err(nor, o.normals64()):.3g}")
    p.close()
    o.close()


def test_learned_smooth_full_size(sg, po, big_mesh):
    from dune_sculpt_amd import learned as L
    from dune_sculpt_amd import workloads

    if not L.MODEL_PATH.exists():
        pytest.skip("no trained model")
    co, tris = big_mesh
    plain = workloads.smooth_stroke(sg, 500, 0.05)
    p, o = _build_pair(sg, po, co, tris, fp64=False)
    o.set_normals(p.normals())
    m = sg.MLP.load(L.MODEL_PATH)
    alpha, tol = m.decision()
    p.set_learned(m)
    p.snapshot()
    # the learned stroke as one captured graph (the benchmark's path): every dab is
    # accepted or redone by the plain brush
    st0 = p.learned_stats()
    p.stroke(L.learned_dabs(plain), graph=True)
    st1 = p.learned_stats()
    acc_graph = st1["accepted"] - st0["accepted"]
    fb_graph = st1["fallbacks"] - st0["fallbacks"]
    assert acc_graph + fb_graph == len(plain)
    p.restore()
    # each learned dab judged against the plain dab from the same state; the plain dab
    # stays applied, so the mesh follows the plain stroke, which the oracle checks
    errs, disp, acc = [], [], []
    for i, d in enumerate(plain):
        e, dm, a, n = p.judge(d)
        want = o.dab(d)
        assert n == len(want), f"dab {i}: plain affected count {n} vs oracle {len(want)}"
        if i % CHECK_EVERY == 0:
            assert np.array_equal(p.last_affected(), want), f"dab {i}: affected set differs"
        errs.append(e)
        disp.append(dm)
        acc.append(a)
    errs, disp, acc = np.array(errs), np.array(disp), np.array(acc, bool)
    ep, en = rel_err(p.positions(), o.positions()), rel_err(p.normals(), o.normals())
    assert ep <= REL_TOL and en <= REL_TOL, f"plain Smooth stroke: positions {ep}, normals {en}"
    assert acc.any()
    assert errs[acc].max() <= 1e-3, f"accepted learned dab with RMS error / r {errs[acc].max()}"
    ratio = float(errs.mean() / disp.mean())
    assert ratio < 0.5, f"learned error is {ratio:.2f} of the plain displacement (a no-op scores 1)"
    _log(f"learned graph stroke: {acc_graph} accepted, {fb_graph} fallbacks; judge over {len(errs)} dabs: RMS err/r "
         f"mean {errs.mean():.3g} max {errs.max():.3g}; plain displacement/r mean {disp.mean():.3g}; error / "
         f"displacement {ratio:.3f} (no-op 1.0); accepted {acc.mean():.3f} (alpha {alpha:.3f}, tol {tol:g}); plain "
         f"stroke vs oracle: positions rel {ep:.3g}, normals rel {en:.3g}")
    p.set_learned(None)
    m.close()
    p.close()
    o.close()


def _gen_meshes(sg, shapes, amplitude):
    from dune_sculpt_amd import workloads

    with ThreadPoolExecutor(16) as ex:
        res = list(ex.map(lambda a: workloads.batch_mesh(sg, a[1], a[0], amplitude), enumerate(shapes)))
    return [r[0] for r in res], [r[1] for r in res]


def _oracle_mesh(po, mesh, dabs, n0):
    co, tris = mesh
    o = po.Oracle(co, tris, 128, False)
    o.set_normals(n0)
    aff = 0
    for d in dabs:
        if d is not None:
            aff += len(o.dab(d))
    out = (o.positions(), o.normals(), aff)
    o.close()
    return out


def _check_sampled(sg, po, B, meshes, table, sample, init_normals):
    """Sampled meshes of a batch against the oracle run on the same dabs."""
    with ThreadPoolExecutor(4) as ex:
        futs = {m: ex.submit(_oracle_mesh, po, meshes[m], [row[m] for row in table], init_normals[m])
                for m in sample}
        rep = {}
        for m, f in futs.items():
            opos, onor, aff = f.result()
            pos, nor = B.positions(m), B.normals(m)
            ep, en = rel_err(pos, opos), rel_err(nor, onor)
            assert ep <= REL_TOL and en <= REL_TOL, f"mesh {m}: positions {ep}, normals {en}"
            rep[m] = (meshes[m][0].shape[0], aff, ep, en, n_bit_diff(pos, opos), n_bit_diff(nor, onor))
    return rep


def _sample(meshes, specs, extra, seed):
    V = np.array([c.shape[0] for c, _ in meshes])
    tori = [i for i, s in enumerate(specs) if s[0] == "torus"]
    spheres = [i for i, s in enumerate(specs) if s[0] == "sphere"]
    pick = [int(V.argmin()), int(V.argmax()), min(tori, key=lambda i: V[i]), max(spheres, key=lambda i: V[i])]
    rng = np.random.default_rng(seed)
    rest = [i for i in range(len(meshes)) if i not in pick]
    pick += [int(x) for x in rng.choice(rest, extra, replace=False)]
    return sorted(set(pick))


def test_many_small_meshes_full_size(sg, po):
    from dune_sculpt_amd import workloads

    shapes, amp, targs = workloads.many_small(sg)
    t = time.perf_counter()
    meshes, specs = _gen_meshes(sg, shapes, amp)
    table = workloads.batch_table(sg, meshes, specs, **targs)
    B = sg.Batch(meshes, leaf_limit=128)
    info = B.info()
    _log(f"256-mesh batch: {info['verts']} V, {info['leaves']} leaves, built in {time.perf_counter() - t:.1f}s")
    assert info["meshes"] == 256 and len(table) == 64
    sample = _sample(meshes, specs, 4, 31)
    n0 = {m: B.normals(m) for m in sample}
    t = time.perf_counter()
    B.run(table)
    tot = B.totals()
    rep = _check_sampled(sg, po, B, meshes, table, sample, n0)
    assert sum(r[1] for r in rep.values()) > 0
    _log(f"{tot['steps'] if 'steps' in tot else 64} steps, {tot['affected']} vertex updates; sampled meshes "
         f"(V, affected, pos rel, nor rel, pos words differing, nor differing): {rep} "
         f"({time.perf_counter() - t:.1f}s)")
    B.close()


def test_many_large_meshes_full_size(sg, po):
    from dune_sculpt_amd import learned as L
    from dune_sculpt_amd import workloads

    shapes, amp, targs = workloads.many_large(sg)
    t = time.perf_counter()
    meshes, specs = _gen_meshes(sg, shapes, amp)
    plain = workloads.batch_table(sg, meshes, specs, **targs)
    B = sg.Batch(meshes, leaf_limit=128)
    info = B.info()
    _log(f"64-mesh batch: {info['verts']} V, {info['leaves']} leaves, built in {time.perf_counter() - t:.1f}s")
    assert info["meshes"] == 64 and len(plain) == 128
    assert all(1_000_000 <= c.shape[0] <= 4_200_000 for c, _ in meshes)
    sample = _sample(meshes, specs, 0, 41)
    n0 = {m: B.normals(m) for m in sample}
    B.snapshot()
    B.run(plain)
    rep = _check_sampled(sg, po, B, meshes, plain, sample, n0)
    _log(f"plain table: sampled meshes (V, affected, pos rel, nor rel, words differing): {rep}")
    if not L.MODEL_PATH.exists():
        B.close()
        pytest.skip("no trained model")
    mlp = sg.MLP.load(L.MODEL_PATH)
    learned = workloads.batch_table(sg, meshes, specs, learned_smooth=True, **targs)
    B.restore()
    B.rewind()
    B.set_learned(mlp)
    B.totals(reset=True)
    B.run(learned)
    tot = B.totals()
    n_smooth = sum(1 for row in learned for d in row if d is not None and d.sculpt_tool == sg.TOOL_SMOOTH)
    assert tot["learned_accepted"] + tot["learned_rejected"] == n_smooth
    # one sampled mesh: the batch equals the single-mesh engine on the same learned dabs,
    # and the single-mesh judge scores its Smooth dabs against the plain brush
    m = sample[0]
    co, tris = meshes[m]
    p = sg.PBVH(co, tris, leaf_limit=128)
    p.set_learned(mlp)
    p.snapshot()
    p.stroke([row[m] for row in learned])
    assert n_bit_diff(p.positions(), B.positions(m)) == 0 and n_bit_diff(p.normals(), B.normals(m)) == 0
    p.restore()
    errs, disp, acc = [], [], []
    for row in plain:
        d = row[m]
        if d.sculpt_tool == sg.TOOL_SMOOTH:
            e, dm, a, n = p.judge(d)
            if n:
                errs.append(e)
                disp.append(dm)
                acc.append(a)
        else:
            p.dab(d)
    errs, disp, acc = np.array(errs), np.array(disp), np.array(acc, bool)
    assert len(errs) and (not acc.any() or errs[acc].max() <= 1e-3)
    _log(f"learned table: {tot['learned_accepted']} accepted, {tot['learned_rejected']} rejected of {n_smooth} "
         f"Smooth dabs; mesh {m} bit-identical to the single-mesh engine; judge on its {len(errs)} Smooth dabs: "
         f"RMS err/r mean {errs.mean():.3g} max {errs.max():.3g}, plain displacement/r mean {disp.mean():.3g}, "
         f"accepted {acc.mean():.2f}")
    p.close()
    B.set_learned(None)
    mlp.close()
    B.close()
