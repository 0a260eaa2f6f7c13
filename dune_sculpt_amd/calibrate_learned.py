#!/usr/bin/env python3
"""Calibrate the learned Smooth fallback on held-out meshes of several resolutions.

The learned dab is accepted when alpha * RMS(|d| / r) <= tol, where d is the network's
displacement of each affected vertex (the learned brush sums it on the GPU).  Here
alpha is set from held-out dabs so that the estimate bounds the actual error: for each
dab the learned result and the plain brush's result from the same state give the
RMS error / r (the judge), and alpha = margin * max over dabs of error / RMS(|d| / r).
The meshes span the resolutions and radii the workloads use (noisy icospheres of
levels 9 and 10, a noisy torus; r ~ U[0.02, 0.1], s ~ U[0.2, 1.0]).

    python -m dune_sculpt_amd.calibrate_learned --out dune_sculpt_amd/models/learned_smooth_k4.npz
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

import numpy as np

CTR_NAFF, CTR_LERR = 1, 15


def main(argv=None):
    import dune_sculpt_amd as sg
    from . import learned as L

    ap = argparse.ArgumentParser()
    ap.add_argument("--model", default=str(L.MODEL_PATH))
    ap.add_argument("--out", default=str(L.MODEL_PATH))
    ap.add_argument("--dabs", type=int, default=120, help="held-out dabs per mesh")
    ap.add_argument("--margin", type=float, default=1.1)
    ap.add_argument("--tol", type=float, default=1e-3)
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[calibrate] {msg}", flush=True)

    m = sg.MLP.load(args.model)
    alpha0, _ = m.decision()
    m.set_decision(alpha0, 1e30)  # accept every learned dab while measuring
    meshes = [("icosphere-9", sg.noisy_icosphere(9, seed=31, amplitude=0.02)),
              ("icosphere-10", sg.noisy_icosphere(10, seed=32, amplitude=0.02)),
              ("torus 2400x720", sg.noisy_torus(2400, 720, seed=33, amplitude=0.01))]
    rng = np.random.default_rng(7)
    rows = []
    for name, (co, tris) in meshes:
        p = sg.PBVH(co, tris, leaf_limit=128)
        p.set_learned(m)
        idx = rng.integers(0, co.shape[0], args.dabs)
        for i in range(args.dabs):
            d = sg.make_dab(tool=sg.TOOL_SMOOTH, location=co[idx[i]], radius=float(rng.uniform(0.02, 0.1)),
                            strength=float(rng.uniform(0.2, 1.0)), smooth_iterations=4)
            ld = L.learned_dabs([d])[0]
            p.snapshot()
            p.dab(ld)
            st = p.last_stats()
            pred = float(np.sqrt(st[CTR_LERR] / 4294967296.0 / st[CTR_NAFF])) if st[CTR_NAFF] else 0.0
            p.restore()
            err, disp, _, n = p.judge(d)
            if n and pred > 0:
                rows.append((name, err, pred, disp))
        p.close()
        log(f"{name}: {args.dabs} dabs measured")
    ratio = np.array([r[1] / r[2] for r in rows])
    alpha = float(args.margin * ratio.max())
    m.set_decision(alpha, args.tol)
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    m.save(out)
    info_path = Path(args.model).with_suffix(".json")
    info = json.loads(info_path.read_text()) if info_path.exists() else {}
    info["alpha_validation_rms"] = info.get("alpha", alpha0)
    info["alpha"] = alpha
    info["calibration"] = {
        "meshes": [n for n, _ in meshes], "dabs": len(rows), "radius": [0.02, 0.1], "strength": [0.2, 1.0],
        "margin": args.margin, "ratio_max": float(ratio.max()), "ratio_p99": float(np.quantile(ratio, 0.99)),
        "ratio_median": float(np.median(ratio)),
        "err_rel_r_max": float(max(r[1] for r in rows)),
        "accepted_at_tol": float(np.mean([alpha * r[2] <= args.tol for r in rows])),
        "accepted_err_rel_r_max": float(max([r[1] for r in rows if alpha * r[2] <= args.tol], default=0.0)),
    }
    out.with_suffix(".json").write_text(json.dumps(info, indent=1) + "\n")
    log(f"alpha {alpha0:.4f} -> {alpha:.4f}: {info['calibration']}")
    m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
