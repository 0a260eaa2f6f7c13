"""GPU parity on the edge cases of the PBVH path, against the C oracle: hidden
vertices (and fully hidden leaves), material splits, dabs that touch nothing, a mesh
that is one leaf, loose vertices referenced by no triangle, tiny and large radii."""
import numpy as np
import pytest

from parity_util import run_pair

pytestmark = pytest.mark.gpu


def _stroke(sg, n, seed, tools=(1, 4, 2, 5), rmin=0.1, rmax=0.4):
    rng = np.random.default_rng(seed)
    radii = rng.uniform(rmin, rmax, n).astype(np.float32)
    centres = sg.great_circle_centres(seed, radii)
    return sg.stroke_config([tools[i % len(tools)] for i in range(n)], centres, radii,
                            rng.uniform(0.2, 1.0, n).astype(np.float32), smooth_iterations=3)


def test_hidden_vertices_and_fully_hidden_leaves(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=21, amplitude=0.02)
    rng = np.random.default_rng(1)
    hidden = rng.random(co.shape[0]) < 0.15
    hidden |= co[:, 2] > 0.7  # a cap: whole leaves hidden
    rep = run_pair(sg, po, co, tris, _stroke(sg, 24, 3), leaf_limit=32, hidden=hidden)
    print(rep)


def test_material_split(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=22, amplitude=0.02)
    cen = co[tris].mean(1)
    mat = (cen[:, 0] > 0).astype(np.int32) + 2 * (cen[:, 1] > 0.3).astype(np.int32)
    run_pair(sg, po, co, tris, _stroke(sg, 20, 4), leaf_limit=64, mat=mat)


def test_dabs_that_touch_nothing(sg, po):
    co, tris = sg.noisy_icosphere(4, seed=23, amplitude=0.02)
    far = [sg.make_dab(tool=t, location=(5.0, 0.0, 0.0), radius=0.2, strength=0.7, smooth_iterations=2)
           for t in (1, 2, 4, 5)]
    p = sg.PBVH(co, tris, leaf_limit=64)
    before = p.positions()
    p.stroke(far)
    for d in far:
        p.dab(d)
        assert len(p.last_affected()) == 0 and len(p.last_gathered()) == 0
    assert np.array_equal(p.positions().view(np.uint32), before.view(np.uint32))
    p.close()
    # mixed with real dabs, against the oracle
    dabs = _stroke(sg, 8, 5)
    run_pair(sg, po, co, tris, dabs[:4] + far + dabs[4:], leaf_limit=64)


def test_single_leaf_mesh_default_leaf_limit(sg, po):
    co, tris = sg.icosphere(1)  # 42 vertices, 80 tris: one leaf at the reference's 10000
    run_pair(sg, po, co, tris, _stroke(sg, 12, 6, rmin=0.3, rmax=0.9), leaf_limit=10000)


def test_loose_vertices(sg, po):
    co, tris = sg.noisy_icosphere(3, seed=24, amplitude=0.02)
    extra = np.random.default_rng(2).normal(size=(37, 3)).astype(np.float32) * 0.5
    co2 = np.concatenate([co, extra])  # referenced by no triangle
    run_pair(sg, po, co2, tris, _stroke(sg, 16, 7), leaf_limit=32)


@pytest.mark.parametrize("radius", [1e-3, 2.5])
def test_extreme_radii(sg, po, radius):
    co, tris = sg.noisy_icosphere(4, seed=25, amplitude=0.02)
    rng = np.random.default_rng(8)
    dabs = [sg.make_dab(tool=t, location=co[rng.integers(len(co))], radius=radius, strength=0.5,
                        smooth_iterations=2, grab_delta=(0.01, 0.0, 0.0)) for t in (1, 2, 4, 5, 1, 2)]
    run_pair(sg, po, co, tris, dabs, leaf_limit=64)
