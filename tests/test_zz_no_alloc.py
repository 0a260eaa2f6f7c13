"""sg_raycast_stats_noalloc (static device scratch, no allocation) against
sg_raycast_stats on the same rays: equal work counters on two calls in a row, the
second with fewer rays, and a clear error beyond the scratch's ray limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    starts = (d * 2.5).astype(np.float32)
    aims = rng.normal(size=(n, 3)) * 0.4
    dirs = aims - starts
    return starts, (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)


def test_raycast_stats_noalloc_matches(sg):
    co, tris = sg.noisy_icosphere(5, seed=2, amplitude=0.02)
    p = sg.PBVH(co, tris, leaf_limit=64)
    s, d = _rays(8, 1)
    with pytest.raises(RuntimeError):  # no ray layout yet: refused, nothing allocated
        p.raycast_stats_noalloc(s, d)
    p.raycast(s, d)  # builds the ray layout
    # a stroke first, so subtree boxes are stale and the internal-node refit is pending
    radii = np.full(8, 0.2, np.float32)
    p.stroke(sg.stroke_config([1, 4, 2, 5] * 2, sg.great_circle_centres(3, radii), radii, np.full(8, 0.6)))
    for n, seed in ((3000, 2), (500, 3)):  # the second call smaller than the first
        s, d = _rays(n, seed)
        for original in (False, True):
            want = p.raycast_stats(s, d, original=original)
            got = p.raycast_stats_noalloc(s, d, original=original)
            assert got == want and want["boxes"] > 0, (n, original, got, want)
    s, d = _rays(4097, 4)  # beyond SG_NOALLOC_RAYS
    with pytest.raises(RuntimeError):
        p.raycast_stats_noalloc(s, d)
    p.close()
