"""CPU tests of the host-side tools around the engine: the product fails loudly with
no GPU (no CPU fallback), the learned-Smooth feature baselines and normalisation, the
benchmark's workload generators and its byte model."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the behaviour without a GPU")
def test_product_fails_loudly_without_gpu(sg):
    co, tris = sg.icosphere(2)
    # the tree builds on the host, but a dab has nowhere to run: an error, not a CPU path
    L = sg.lib()
    h = L.BKE_pbvh_new()
    assert h
    d = sg.make_dab(tool=sg.TOOL_DRAW, location=co[0], radius=0.3)
    rc = L.sg_dab_apply(h, d)
    assert rc != 0 and b"device" in L.sg_last_error().lower()
    L.BKE_pbvh_free(h)
    with pytest.raises(RuntimeError):
        sg.Batch([(co, tris)])
    with pytest.raises(RuntimeError):
        sg.MLP(0)


def test_learned_baselines_and_norm():
    from dune_sculpt_amd import learned as L

    rng = np.random.default_rng(0)
    n = 50
    m = rng.integers(3, 9, n)
    offs = rng.normal(size=(n, 8, 3)).astype(np.float32)
    for i in range(n):
        offs[i, m[i]:] = 0.0
    f = rng.uniform(0, 1, n).astype(np.float32)
    X = np.zeros((n, 28), np.float32)
    X[:, :24] = offs.reshape(n, 24)
    X[:, 24] = m / 8.0
    X[:, 25] = f
    X[:, 26] = f * f
    X[:, 27] = 1 - (1 - f) ** 4
    b = L.baselines(X)
    lap = np.stack([offs[i, :m[i]].mean(0) for i in range(n)])
    assert np.allclose(b["one_step"], f[:, None] * lap, atol=1e-6)
    assert np.allclose(b["scaled_step"], X[:, 27:28] * lap, atol=1e-6)
    assert np.all(b["none"] == 0)
    T = rng.normal(size=(n, 3)).astype(np.float32)
    assert L.rel_rms(T, T) == 0.0 and abs(L.rel_rms(np.zeros_like(T), T) - 1.0) < 1e-6
    X[:, 5] = 2.0  # a constant feature gets a zero inverse scale, not a division by zero
    mu, isig, osc = L.fit_norm(X, T)
    assert isig[5] == 0.0 and np.all(np.isfinite(isig)) and np.all(osc > 0)


def test_bench_workload_generators(sg):
    sys.path.insert(0, str(ROOT))
    from dune_sculpt_amd import workloads

    meshes, specs = workloads.batch_meshes(sg, 6, seed=11, vmin=5000, vmax=500000, amplitude=0.01)
    for (co, tris), (kind, seed, amp) in zip(meshes, specs):
        V = co.shape[0]
        if kind == "sphere":
            k = round(np.log((V - 2) / 10) / np.log(4))
            assert V == 10 * 4 ** k + 2 and tris.shape[0] == 20 * 4 ** k
        else:
            assert tris.shape[0] == 2 * V  # closed torus grid
        assert tris.min() >= 0 and tris.max() < V
    table = workloads.batch_table(sg, meshes, specs, 8, seed=12, tools=(sg.TOOL_DRAW, sg.TOOL_SMOOTH),
                              rmin=0.02, rmax=0.1)
    assert len(table) == 8 and all(len(r) == len(meshes) for r in table)
    for m, (co, _) in enumerate(meshes):
        cs = np.array([list(table[s][m].location) for s in range(8)])
        # centres lie on the noisy surface: within a few edge lengths of a vertex
        d = np.sqrt(((cs[:, None, :] - co[None, ::7, :]) ** 2).sum(-1)).min(1)
        assert d.max() < 0.15
        # consecutive centres 0.1 r apart along the path
        steps = np.linalg.norm(np.diff(cs, axis=0), axis=1)
        radii = np.array([table[s][m].radius for s in range(8)])
        assert np.all(steps < 0.2 * radii[1:] + 0.05)
        assert {table[s][m].sculpt_tool for s in range(8)} == {sg.TOOL_DRAW, sg.TOOL_SMOOTH}


def test_bench_byte_models(sg):
    sys.path.insert(0, str(ROOT))
    import bench

    c = np.zeros(24, np.int64)
    c[0], c[1], c[6], c[7], c[8], c[9], c[10], c[11], c[12], c[13], c[14] = 20, 1000, 700, 1500, 200, 6000, 6000, 400, 2000, 1100, 300
    smooth = bench.dab_bytes(sg, sg.TOOL_SMOOTH, 4, c)
    draw = bench.dab_bytes(sg, sg.TOOL_DRAW, 4, c)
    assert smooth["total"] == sum(smooth[k] for k in ("search", "brush", "post", "smooth_total"))
    assert draw["total"] == sum(draw[k] for k in ("search", "brush", "post", "area"))
    assert smooth["smooth_total"] > 0 and "area" not in smooth and draw["area"] == 12 * 200
    t = {"tests": 10, "candidates": 100, "affected": 50, "area_verts": 5, "smooth_records_k": 40,
         "smooth_nbr_reads": 240, "learned_records": 0, "deg_t": 300, "shared": 30, "gathered": 3}
    assert bench.batch_bytes(t) == (24 * 10 + 12 * 100 + 12 * 50 + 12 * 5 + 24 * 40 + 4 * 240 + 16 * 300
                                    + 24 * 50 + 12 * 30 + 24 * 3)
