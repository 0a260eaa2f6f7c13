#!/usr/bin/env python3
"""Train the learned Smooth(4) model on the GPU and write it with its report.

Pairs come from plain Smooth(4) dabs (the engine's own brush) at random places on
seeded noisy icospheres; the network is trained with the engine's FP32 training
kernels; the fallback estimate's scale (alpha) is calibrated on held-out meshes.

    python -m dune_sculpt_amd.train_learned --out dune_sculpt_amd/models/learned_smooth_k4.npz
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np


def main(argv=None):
    from . import learned as L

    ap = argparse.ArgumentParser()
    ap.add_argument("--level", type=int, default=10)
    ap.add_argument("--amplitude", type=float, default=0.02)
    ap.add_argument("--pairs", type=int, default=2_000_000)
    ap.add_argument("--val", type=int, default=200_000)
    ap.add_argument("--train-seeds", default="11,12,13,14")
    ap.add_argument("--val-seed", type=int, default=21)
    ap.add_argument("--radius", default="0.02,0.1")
    ap.add_argument("--strength", default="0.2,1.0")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch-wg", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--out", default=str(L.MODEL_PATH))
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[train] {msg}", flush=True)

    radius = tuple(float(x) for x in args.radius.split(","))
    strength = tuple(float(x) for x in args.strength.split(","))
    seeds = tuple(int(s) for s in args.train_seeds.split(","))
    t = time.perf_counter()
    X, T = L.collect_pairs(args.pairs, mesh_seeds=seeds, level=args.level, amplitude=args.amplitude, radius=radius,
                           strength=strength, seed=args.seed + 1, log=log)
    Xv, Tv = L.collect_pairs(args.val, mesh_seeds=(args.val_seed,), level=args.level, amplitude=args.amplitude,
                             radius=radius, strength=strength, seed=args.seed + 2, log=log)
    collect_s = time.perf_counter() - t
    log(f"{X.shape[0]} train / {Xv.shape[0]} validation pairs in {collect_s:.1f}s")
    m, info = L.train_model(X, T, Xv, Tv, epochs=args.epochs, batch_wg=args.batch_wg, lr=args.lr, seed=args.seed,
                            log=log)
    alpha, _ = m.decision()
    m.set_decision(alpha, args.tol)
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    m.save(out)
    info.update({"collect_s": collect_s, "level": args.level, "amplitude": args.amplitude, "train_seeds": seeds,
                 "val_seed": args.val_seed, "radius": radius, "strength": strength, "tol": args.tol,
                 "batch_rows": args.batch_wg * 32, "lr": args.lr})
    out.with_suffix(".json").write_text(json.dumps(info, indent=1) + "\n")
    log(f"wrote {out} (val rel RMS {info['val_rel_rms']:.4f}, alpha {alpha:.4f})")
    m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
