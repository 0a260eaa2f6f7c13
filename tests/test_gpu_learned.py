"""GPU tests of the learned Smooth operator: MFMA operand layout, training gradients
against the FP64 oracle, BF16 inference against its emulation, training on pairs
from the plain brush (with the gain over the methods without the network), the
plain-brush fallback (forced by tolerance 0 and by corrupted weights), determinism."""
import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def test_mfma_layout(sg):
    rng = np.random.default_rng(1)
    A = mo.bf16(rng.normal(size=(16, 16)).astype(np.float32))
    B = mo.bf16(rng.normal(size=(16, 16)).astype(np.float32))
    D = sg.mfma_probe(A.view(np.uint32) >> 16, B.view(np.uint32) >> 16)
    want = A.astype(np.float64) @ B.astype(np.float64)
    assert np.abs(D - want).max() < 1e-4 * np.abs(want).max()


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 28)).astype(np.float32)
    T = (np.tanh(X[:, :3]) * 0.5 + 0.1 * X[:, 3:6] * X[:, 25:26]).astype(np.float32)
    return X, T


def test_gradient_matches_oracle(sg):
    X, T = _data(1000)
    m = sg.MLP(seed=3)
    mu, isig, osc = X.mean(0), 1 / X.std(0), T.std(0)
    m.set_norm(mu, isig, osc)
    m.set_data(X, T)
    g, loss = m.gradient(1000)
    w = m.params().astype(np.float64)
    Xn, Tn = mo.normalize(X, T, (mu.astype(np.float64), isig.astype(np.float64), osc.astype(np.float64)))
    lo, go = mo.loss_and_grad(w, Xn, Tn)
    assert abs(loss - lo) <= 1e-4 * lo
    assert np.abs(g - go).max() <= 2e-4 * np.abs(go).max()


def test_predict_matches_bf16_emulation(sg):
    X, T = _data(777, seed=2)
    m = sg.MLP(seed=5)
    mu, isig, osc = X.mean(0), 1 / X.std(0), T.std(0)
    m.set_norm(mu, isig, osc)
    Y = m.predict(X)
    Xn = (X - mu) * isig
    want = mo.predict_bf16(m.params(), Xn) * osc
    assert np.abs(Y - want).max() <= 2e-2 * np.abs(want).max()


def test_training_gains_over_methods_without_network(sg):
    from dune_sculpt_amd import learned as L

    X, T = L.collect_pairs(60000, mesh_seeds=(201,), level=7, amplitude=0.02, radius=(0.15, 0.3), seed=1)
    Xv, Tv = L.collect_pairs(15000, mesh_seeds=(202,), level=7, amplitude=0.02, radius=(0.15, 0.3), seed=2)
    m, info = L.train_model(X, T, Xv, Tv, epochs=30, batch_wg=64, lr=2e-3)
    print(info["val_rel_rms"], info["baselines_rel_rms"], info["loss_curve"][:3], info["loss_curve"][-3:])
    assert info["loss_curve"][-1] < 0.5 * info["loss_curve"][0]
    for name, b in info["baselines_rel_rms"].items():
        assert info["val_rel_rms"] < b, (name, info)


def _small_setup(sg, seed=301):
    from dune_sculpt_amd import learned as L

    X, T = L.collect_pairs(40000, mesh_seeds=(seed,), level=7, amplitude=0.02, radius=(0.15, 0.3), seed=3)
    m, info = L.train_model(X, T, X[:5000], T[:5000], epochs=15, batch_wg=64, lr=2e-3)
    co, tris = sg.noisy_icosphere(7, seed=seed + 50, amplitude=0.02)
    n = 24
    radii = np.full(n, 0.2, np.float32)
    c = L.random_surface_centres(n, np.random.default_rng(4), seed + 50, 0.02)
    dabs = [sg.make_dab(tool=sg.TOOL_SMOOTH, location=ci, radius=0.2, strength=0.6, smooth_iterations=4) for ci in c]
    return L, m, co, tris, dabs


def test_learned_dabs_fallback_and_determinism(sg):
    L, m, co, tris, dabs = _small_setup(sg)
    ld = L.learned_dabs(dabs)
    # plain reference
    pp = sg.PBVH(co, tris, leaf_limit=128)
    pp.stroke(dabs)
    plain = pp.positions()
    # tolerance 0: every learned dab is rejected and redone by the plain brush -> bitwise equal
    alpha, tol = m.decision()
    m.set_decision(alpha, 0.0)
    p = sg.PBVH(co, tris, leaf_limit=128)
    p.set_learned(m)
    p.stroke(ld)
    assert np.array_equal(p.positions().view(np.uint32), plain.view(np.uint32))
    st = p.learned_stats()
    assert st["fallbacks"] == len(dabs) and st["accepted"] == 0
    p.close()
    # corrupted weights (NaN): rejected, plain result
    m2 = sg.MLP(seed=9)
    w = m.params()
    w[5] = np.nan
    m2.set_norm(*m.norm())
    m2.set_params(w)
    m2.set_decision(alpha, 1.0)
    p = sg.PBVH(co, tris, leaf_limit=128)
    p.set_learned(m2)
    p.stroke(ld)
    assert np.array_equal(p.positions().view(np.uint32), plain.view(np.uint32))
    p.close()
    # accepting everything: runs the network; deterministic run to run
    m.set_decision(alpha, 1.0)
    outs = []
    for _ in range(2):
        p = sg.PBVH(co, tris, leaf_limit=128)
        p.set_learned(m)
        p.stroke(ld)
        outs.append(p.positions())
        assert p.learned_stats()["accepted"] == len(dabs)
        p.close()
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    assert not np.array_equal(outs[0].view(np.uint32), plain.view(np.uint32))
    # the learned stroke (a small model: 40k pairs, 15 epochs, errors compounding over
    # 24 overlapping dabs) stays clearly closer to the plain stroke than not smoothing,
    # whose error is the displacement itself
    mv = np.any(plain != co, axis=1)
    disp = np.sqrt(((plain[mv] - co[mv]).astype(np.float64) ** 2).sum(1).mean())
    err = np.sqrt(((outs[0][mv] - plain[mv]).astype(np.float64) ** 2).sum(1).mean())
    print("stroke RMS error / RMS displacement", err / disp)
    assert err < 0.75 * disp


def test_judge_against_plain_brush(sg):
    """sg_learned_judge: the learned dab compared with the plain dab from the same state;
    the plain result is what stays applied; a rejected dab has zero error."""
    L, m, co, tris, dabs = _small_setup(sg)
    alpha, _ = m.decision()
    p = sg.PBVH(co, tris, leaf_limit=128)
    q = sg.PBVH(co, tris, leaf_limit=128)
    p.set_learned(m)
    m.set_decision(alpha, 1.0)
    errs = []
    for d in dabs[:6]:
        e, disp, acc, n = p.judge(d)
        q.dab(d)
        assert acc and n > 0 and disp > 0
        errs.append(e)
        # the plain result stays applied: bitwise the plain stroke
        assert np.array_equal(p.positions().view(np.uint32), q.positions().view(np.uint32))
    assert 0 < max(errs) < 1e-1
    m.set_decision(alpha, 0.0)
    e, disp, acc, n = p.judge(dabs[6])
    assert not acc and e == 0.0 and n > 0
    p.close()
    q.close()


def test_learned_dab_reports_its_affected_set(sg):
    """An accepted learned dab stamps its vertices instead of listing them; the API
    still returns the affected set, which equals the plain brush's from the same state."""
    L, m, co, tris, dabs = _small_setup(sg)
    alpha, _ = m.decision()
    m.set_decision(alpha, 1.0)
    ld = L.learned_dabs(dabs[:3])
    for d, l in zip(dabs[:3], ld):
        p = sg.PBVH(co, tris, leaf_limit=128)
        q = sg.PBVH(co, tris, leaf_limit=128)
        p.set_learned(m)
        p.dab(l)
        q.dab(d)
        assert p.learned_stats()["accepted"] == 1
        a, b = p.last_affected(), q.last_affected()
        assert len(a) > 0 and np.array_equal(a, b)
        assert np.array_equal(p.last_gathered(), q.last_gathered())
        p.close()
        q.close()
