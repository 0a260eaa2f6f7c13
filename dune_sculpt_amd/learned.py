"""Learned Smooth: pair collection from the plain brush, training on the GPU with the
engine's own kernels, calibration of the fallback estimate, and the per-dab judge.

Everything runs through libsculpt_gpu.so (collection: a plain Smooth dab that also
records features/targets; training: FP32 forward/backward/Adam kernels; inference:
BF16 MFMA inside the learned brush kernel).
"""
from __future__ import annotations

import time
from pathlib import Path

import numpy as np

from . import (MLP, PBVH, SMOOTH_LEARNED, TOOL_SMOOTH, great_circle_centres, make_dab, noise_displace,
               noisy_icosphere)

MODEL_PATH = Path(__file__).resolve().parent / "models" / "learned_smooth_k4.npz"


def random_surface_centres(n, rng, mesh_seed, amplitude):
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u = u.astype(np.float32)
    return noise_displace(u.copy(), u, mesh_seed, amplitude)


def collect_pairs(n_pairs, mesh_seeds=(101,), level=10, amplitude=0.02, radius=(0.03, 0.08), strength=(0.1, 1.0),
                  k=4, seed=0, log=None):
    """Rows (features, targets) from plain Smooth(k) dabs at random places on seeded
    noisy icospheres; rows whose vertex is outside the model's range are dropped."""
    rng = np.random.default_rng(seed)
    Xs, Ts = [], []
    total = 0
    for ms in mesh_seeds:
        co, tris = noisy_icosphere(level, seed=ms, amplitude=amplitude)
        p = PBVH(co, tris, leaf_limit=128)
        per_mesh = (n_pairs - total) // max(1, len(mesh_seeds) - mesh_seeds.index(ms))
        got = 0
        while got < per_mesh:
            c = random_surface_centres(1, rng, ms, amplitude)[0]
            d = make_dab(tool=TOOL_SMOOTH, location=c, radius=float(rng.uniform(*radius)),
                         strength=float(rng.uniform(*strength)), smooth_iterations=k)
            X, T, ok = p.collect(d)
            X, T = X[ok], T[ok]
            Xs.append(X)
            Ts.append(T)
            got += X.shape[0]
        total += got
        p.close()
        if log:
            log(f"collected {got} pairs on mesh seed {ms}")
    X = np.concatenate(Xs)[:n_pairs]
    T = np.concatenate(Ts)[:n_pairs]
    return X, T


def fit_norm(X, T):
    mu = X.mean(0)
    sd = X.std(0)
    isig = np.where(sd > 1e-6, 1.0 / np.maximum(sd, 1e-6), 0.0)
    osc = T.std(0)
    osc = np.where(osc > 0, osc, 1.0)
    return mu.astype(np.float32), isig.astype(np.float32), osc.astype(np.float32)


def baselines(X):
    """Predictions of the methods without the network, in the same (local frame / h)
    units as the targets: no smoothing, one Jacobi step, and the one-step Laplacian
    scaled by 1 - (1 - f)^4 (what 4 steps give on a locally linear surface)."""
    m = np.rint(X[:, 24] * 8).astype(int)
    offs = X[:, :24].reshape(-1, 8, 3)
    mask = (np.arange(8)[None, :] < m[:, None]).astype(np.float32)
    lap = (offs * mask[:, :, None]).sum(1) / np.maximum(m, 1)[:, None]
    f = X[:, 25:26]
    return {"none": np.zeros_like(lap), "one_step": f * lap, "scaled_step": X[:, 27:28] * lap}


def rel_rms(pred, T):
    return float(np.sqrt(((pred - T) ** 2).sum(1).mean()) / np.sqrt((T ** 2).sum(1).mean()))


def train_model(X, T, Xv, Tv, epochs=40, batch_wg=256, lr=2e-3, seed=0, log=None):
    mu, isig, osc = fit_norm(X, T)
    m = MLP(seed)
    m.set_norm(mu, isig, osc)
    m.set_data(X, T)
    t = time.perf_counter()
    curve = m.train(epochs, batch_wg=batch_wg, lr=lr, seed=seed)
    train_s = time.perf_counter() - t
    pred = m.predict(Xv)
    err = rel_rms(pred, Tv)
    base = {k: rel_rms(v, Tv) for k, v in baselines(Xv).items()}
    # fallback calibration: expected RMS error = alpha * RMS(predicted displacement)
    alpha = float(np.sqrt(((pred - Tv) ** 2).sum(1).mean()) / np.sqrt((pred ** 2).sum(1).mean()))
    m.set_decision(alpha, 1e-3)
    info = {"train_rows": int(X.shape[0]), "val_rows": int(Xv.shape[0]), "epochs": epochs, "train_s": train_s,
            "loss_curve": [float(x) for x in curve], "val_rel_rms": err, "baselines_rel_rms": base, "alpha": alpha}
    if log:
        log(f"trained {epochs} epochs in {train_s:.1f}s: val rel RMS {err:.4f}, baselines {base}")
    return m, info


def learned_dabs(dabs):
    out = []
    for d in dabs:
        e = make_dab(tool=d.sculpt_tool, location=list(d.location), radius=d.radius, strength=d.strength,
                     falloff=d.curve_preset, smooth_iterations=d.smooth_iterations,
                     flag=d.flag | SMOOTH_LEARNED)
        out.append(e)
    return out


def judge(pbvh, dabs, mlp):
    """Per-dab RMS displacement error of the learned dab against the plain brush from
    the same state, relative to the radius (the plain result is kept, so the mesh
    follows the plain stroke).  Returns (errors, plain displacements, accepted flags),
    each relative to r."""
    errs, disp, acc = [], [], []
    pbvh.set_learned(mlp)
    for d in dabs:
        e, m, a, n = pbvh.judge(d)
        if n:
            errs.append(e)
            disp.append(m)
            acc.append(a)
    return np.array(errs), np.array(disp), np.array(acc, bool)
