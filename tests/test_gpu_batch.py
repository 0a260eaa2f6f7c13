"""GPU tests of the batched engine (many meshes, one dab per mesh per step).

Bar: every mesh of a batch ends bit-identical (positions and normals) to the same dabs
applied to that mesh alone by the single-mesh engine -- which the parity tests pin to
the C oracle -- for the step graph and for stream launches, with steps where a mesh has
no dab, after snapshot/restore and the autotuner, and for learned Smooth dabs both
accepted and rejected (redone by the plain brush inside the step).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _meshes(sg):
    return [sg.noisy_icosphere(3, seed=5, amplitude=0.02),
            sg.noisy_torus(40, 12, seed=6, amplitude=0.01),
            sg.noisy_icosphere(4, seed=7, amplitude=0.02),
            sg.icosphere(2),
            sg.noisy_torus(60, 18, seed=8, amplitude=0.01),
            sg.noisy_icosphere(5, seed=9, amplitude=0.02)]


def _table(sg, meshes, steps, seed, tools=(1, 2, 4, 5), skip=0.2, flag=0):
    rng = np.random.default_rng(seed)
    table = []
    for s in range(steps):
        row = []
        for m, (co, _) in enumerate(meshes):
            if rng.random() < skip:
                row.append(None)
                continue
            tool = tools[(s + m) % len(tools)]
            c = co[rng.integers(len(co))]
            row.append(sg.make_dab(tool=tool, location=c, radius=float(rng.uniform(0.15, 0.4)),
                                   strength=float(rng.uniform(0.2, 1.0)), falloff=int(1 + (s + 2 * m) % 9),
                                   smooth_iterations=int(1 + (s + m) % 4),
                                   grab_delta=rng.normal(size=3) * 0.02,
                                   flag=flag if tool == sg.TOOL_SMOOTH else 0))
        table.append(row)
    return table


def _single(sg, meshes, table, mlp=None):
    out = []
    for m, (co, tris) in enumerate(meshes):
        p = sg.PBVH(co, tris, leaf_limit=128)
        if mlp is not None:
            p.set_learned(mlp)
        dabs = [row[m] for row in table if row[m] is not None]
        if dabs:
            p.stroke(dabs)
        out.append((p.positions(), p.normals()))
        p.close()
    return out


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _check(batch, ref, meshes):
    for m in range(len(meshes)):
        assert _same(batch.positions(m), ref[m][0]), f"positions of mesh {m}"
        assert _same(batch.normals(m), ref[m][1]), f"normals of mesh {m}"


def test_batch_matches_single_mesh_engine(sg):
    meshes = _meshes(sg)
    table = _table(sg, meshes, 16, seed=1)
    ref = _single(sg, meshes, table)
    b = sg.Batch(meshes, leaf_limit=128)
    info = b.info()
    assert info["meshes"] == len(meshes) and info["verts"] == sum(len(c) for c, _ in meshes)
    # untouched build: initial normals equal the single-mesh build's
    p = sg.PBVH(*meshes[1], leaf_limit=128)
    assert _same(b.normals(1), p.normals())
    p.close()
    b.run(table)
    tot = b.totals()
    print(info, tot)
    assert tot["steps"] == 16 and tot["affected"] > 0
    _check(b, ref, meshes)
    b.close()


def test_batch_stream_mode_and_split_launches(sg):
    meshes = _meshes(sg)[:4]
    table = _table(sg, meshes, 10, seed=2)
    ref = _single(sg, meshes, table)
    b = sg.Batch(meshes, leaf_limit=64)
    b.load(table)
    b.launch(4, graph=False)
    b.launch(6, graph=True)
    _check(b, ref, meshes)
    b.close()


def test_batch_snapshot_and_autotune_keep_results(sg):
    meshes = _meshes(sg)
    table = _table(sg, meshes, 12, seed=3)
    ref = _single(sg, meshes, table)
    b = sg.Batch(meshes, leaf_limit=128)
    b.snapshot()
    b.run(table[:5])
    b.restore()
    before, after = b.autotune(table[:6])
    print("autotune", before, after, b.launch_shape())
    b.run(table)
    _check(b, ref, meshes)
    b.close()


def test_batch_learned_accept_and_reject(sg):
    from dune_sculpt_amd import learned as L

    X, T = L.collect_pairs(30000, mesh_seeds=(401,), level=6, amplitude=0.02, radius=(0.15, 0.3), seed=5)
    m, _ = L.train_model(X, T, X[:4000], T[:4000], epochs=10, batch_wg=64, lr=2e-3)
    alpha, _ = m.decision()
    meshes = _meshes(sg)
    table = _table(sg, meshes, 10, seed=4, tools=(2, 1, 2, 4), flag=sg.SMOOTH_LEARNED)
    for tol, want in ((1.0, "accepted"), (0.0, "rejected")):
        m.set_decision(alpha, tol)
        ref = _single(sg, meshes, table, mlp=m)
        b = sg.Batch(meshes, leaf_limit=128)
        b.set_learned(m)
        b.run(table)
        tot = b.totals()
        print(tol, tot)
        assert tot["learned_" + want] > 0
        assert tot["learned_" + ("rejected" if want == "accepted" else "accepted")] == 0
        _check(b, ref, meshes)
        b.close()
    m.close()


@pytest.mark.parametrize("leaf_limit", [128, 16])
def test_persistent_step_matches_single_mesh_engine(sg, leaf_limit):
    """The persistent step (a workgroup per mesh, every row in one launch): each mesh
    bit-identical to the single-mesh engine and to the step graph, with rows where a mesh
    has no dab, Smooth k = 1..4, every tool, totals equal to the step graph's; small leaves
    give many gathered leaves and search cuts per dab."""
    meshes = _meshes(sg)
    table = _table(sg, meshes, 16, seed=11)
    ref = _single(sg, meshes, table)
    b = sg.Batch(meshes, leaf_limit=leaf_limit)
    assert b.set_persistent(True), "the batch should fit the persistent step"
    b.snapshot()
    b.run(table)
    _check(b, ref, meshes)
    tot_p = b.totals(reset=True)
    b.restore()
    b.set_persistent(False)
    b.run(table)
    _check(b, ref, meshes)
    tot_g = b.totals()
    print("persistent", tot_p, "graph", tot_g)
    assert tot_p == tot_g
    b.close()


def test_persistent_step_split_launches_and_large_dabs(sg):
    """Rows run in pieces (the launch takes its first row from the host); dabs of radius
    up to 0.6 on a 40,962-vertex icosphere, so some affected lists exceed what the
    workgroup holds in registers (2048 vertices) and take the round-by-round steps."""
    meshes = [sg.noisy_icosphere(6, seed=31, amplitude=0.02), sg.noisy_torus(200, 60, seed=32, amplitude=0.01)]
    rng = np.random.default_rng(12)
    table = []
    for s in range(9):
        row = []
        for m, (co, _) in enumerate(meshes):
            row.append(sg.make_dab(tool=[1, 2, 4, 5][(s + m) % 4], location=co[rng.integers(len(co))],
                                   radius=float(rng.uniform(0.3, 0.6)), strength=0.7,
                                   smooth_iterations=int(1 + s % 4), grab_delta=rng.normal(size=3) * 0.02))
        table.append(row)
    ref = _single(sg, meshes, table)
    b = sg.Batch(meshes, leaf_limit=128)
    b.set_persistent(True)
    b.load(table)
    b.launch(2)
    b.launch(3)
    b.launch(4)
    _check(b, ref, meshes)
    tot = b.totals()
    print(tot)
    assert tot["steps"] == 9 and tot["affected"] > 0
    b.close()
