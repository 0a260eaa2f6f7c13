"""CPU: pin the FP64 MLP oracle (its analytic backward against central finite
differences, and the BF16 rounding helper against exact cases)."""
import numpy as np

import mlp_oracle as mo


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, mo.NPARAM)
    X = rng.normal(size=(17, 28))
    T = rng.normal(size=(17, 3))
    loss, g = mo.loss_and_grad(w, X, T)
    idx = rng.choice(mo.NPARAM, 200, replace=False)
    h = 1e-6
    for i in idx:
        wp = w.copy()
        wm = w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mo.loss_and_grad(wp, X, T)[0] - mo.loss_and_grad(wm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # a dropped activation derivative or a transposed weight would fail here
    assert loss > 0


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0], np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # halfway rounds to even (down)
    assert r[2] == np.float32(1.0 + 2 ** -7)
    assert r[3] == np.float32(1.0 + 2 ** -6)  # halfway rounds to even (up)
    assert r[4] == -2.5 and r[5] == 0.0
