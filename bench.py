#!/usr/bin/env python3
"""Sculpt brush throughput on one MI355X.

Workload (the plain-brush scale configuration): icosphere level 10 (10,485,762 verts,
20,971,520 tris, FP32) displaced along the normal by seeded 3D value noise (amplitude
0.02, 4 octaves); one step = one stroke of 1000 dabs cycling Draw / Inflate /
Smooth(k=4) / Grab with radius ~ U[0.02, 0.1] and strength ~ U[0.2, 1.0] along a
seeded great circle (spacing 10% of the radius), PBVH leaves of <= 128 tris.  Each dab
runs the whole hot path on the GPU: sphere search, falloff and strength, displacement,
normals of the affected vertices, leaf and search-level bounds; the internal-node
refit runs at the end of every stroke.  The stroke is captured once into a hipGraph
and launched per step (the mesh keeps evolving from step to step).

value = vertex-updates/s = sum over timed dabs of |affected set| / wall time.
Algorithmic HBM bytes per dab follow the byte model in DESIGN.md (counted exactly by
an untimed replay of the same strokes with the kernels' statistics counters on).
The CPU baseline is the C oracle (OpenMP over gathered leaves) on a bounded sample of
the same stroke on the same mesh.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0
# HBM bytes per dab from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (summed over the
# kernels of the stroke, scaled by the copy-kernel calibration); None until measured.
TRAFFIC_BYTES_PER_DAB = None
METRIC = "vertex-updates/s + HBM GB/s; learned vs plain Smooth dabs/s at ≤1e-3 RMS"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def surface_centres(sg, centres, mesh_seed=1, amplitude=0.02):
    """Project great-circle points onto the noisy surface: the mesh was displaced
    radially by the same seeded noise field, so the surface point above a unit
    direction u is u displaced by that field."""
    u = centres / np.linalg.norm(centres, axis=1, keepdims=True)
    return sg.noise_displace(u.astype(np.float32), u.astype(np.float32), mesh_seed, amplitude)


def make_c2_stroke(sg, n_dabs=1000, seed=3):
    rng = np.random.default_rng(seed)
    radii = rng.uniform(0.02, 0.1, n_dabs).astype(np.float32)
    strengths = rng.uniform(0.2, 1.0, n_dabs).astype(np.float32)
    centres = surface_centres(sg, sg.great_circle_centres(seed + 1, radii))
    tools = np.array([sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH, sg.TOOL_GRAB] * (n_dabs // 4 + 1))[:n_dabs]
    return sg.stroke_config(tools, centres, radii, strengths, smooth_iterations=4), tools


def dab_bytes(sg, tool, k, c):
    """Algorithmic HBM bytes of one dab from its statistics counters (byte model:
    SoA FP32 = 12 B per vec3, int32 topology, every item counted once per dab)."""
    ng, a, tests, ncand, narea, nbr, tris, r1x, fa, vfa, shared = (int(c[i]) for i in (0, 1, 6, 7, 8, 9, 10, 11, 12, 13, 14))
    out = {}
    out["search"] = 24 * tests
    brush = 12 * ncand
    if tool == sg.TOOL_DRAW:
        out["area"] = 12 * narea
    if tool == sg.TOOL_INFLATE:
        brush += 12 * a
    if tool == sg.TOOL_SMOOTH:
        r1_it = (r1x / (k - 1)) if k > 1 else 0.0
        out["smooth_step"] = 12 * (a + r1_it) + 4 * (a + 1) + 4 * (nbr / k) + 12 * a
        smooth_total = k * (12 * (a + r1_it) + 4 * (a + 1) + 12 * a) + 4 * nbr
        out["smooth_total"] = smooth_total
    else:
        brush += 12 * a
    out["brush"] = brush
    out["normals"] = 4 * (a + 1) + 4 * tris + 12 * fa + 12 * vfa + 12 * a
    out["refit"] = 12 * shared + 24 * ng
    out["post"] = out["normals"] + out["refit"]
    total = out["search"] + out["brush"] + out["post"] + out.get("area", 0) + out.get("smooth_total", 0)
    out["total"] = total
    return out


def run_learned(sg, A, W, args):
    """Learned Smooth on the same mesh: 500 Smooth(4) dabs, r = 0.05, s ~ U[0.2, 1.0]
    along a seeded great circle; plain vs learned strokes timed from the pristine mesh
    (HIP events around each stroke; rejected learned dabs are redone by the plain brush
    inside the timed region), then every dab judged against the plain brush."""
    from dune_sculpt_amd import learned as L

    if not L.MODEL_PATH.exists():
        return {"skipped": f"no trained model at {L.MODEL_PATH.name}"}
    m = sg.MLP.load(L.MODEL_PATH)
    alpha, tol = m.decision()
    n = args.learned_dabs
    rng = np.random.default_rng(5)
    radii = np.full(n, 0.05, np.float32)
    strengths = rng.uniform(0.2, 1.0, n).astype(np.float32)
    centres = surface_centres(sg, sg.great_circle_centres(6, radii))
    plain = sg.stroke_config(np.full(n, sg.TOOL_SMOOTH), centres, radii, strengths, smooth_iterations=4)
    learned = L.learned_dabs(plain)
    A.set_learned(m)
    out = {"dabs": n, "radius": 0.05, "k": 4, "model": L.MODEL_PATH.name, "alpha": alpha, "tol": tol}
    K = args.learned_steps
    for name, dabs in (("plain", plain), ("learned", learned)):
        prep = A.prepare(dabs) if args.mode == "graph" else None
        st0 = A.learned_stats()
        times = []
        for s in range(W + K):
            A.restore()
            A.sync()
            A.timer_start()
            if prep is not None:
                prep.launch(sync=False)
            else:
                A.stroke(dabs, graph=False)
            ms = A.timer_stop()
            if s >= W:
                times.append(ms)
        st1 = A.learned_stats()
        if prep is not None:
            prep.close()
        ms = float(np.mean(times))
        out[name] = {"ms_per_stroke": ms, "dabs_per_s": n / (ms * 1e-3), "us_per_dab": ms * 1e3 / n}
        if name == "learned":
            out[name]["accepted_per_stroke"] = (st1["accepted"] - st0["accepted"]) / (W + K)
            out[name]["fallbacks_per_stroke"] = (st1["fallbacks"] - st0["fallbacks"]) / (W + K)
    out["speedup"] = out["learned"]["dabs_per_s"] / out["plain"]["dabs_per_s"]
    if not args.no_judge:
        A.restore()
        t = time.perf_counter()
        errs, disp, acc = L.judge(A, plain, m)
        out["judge"] = {"dabs": int(len(errs)), "seconds": round(time.perf_counter() - t, 2),
                        "rms_err_rel_r_mean": float(errs.mean()), "rms_err_rel_r_median": float(np.median(errs)),
                        "rms_err_rel_r_max": float(errs.max()),
                        "frac_dabs_within_1e-3": float((errs <= 1e-3).mean()),
                        "rms_plain_disp_rel_r_mean": float(disp.mean()),
                        "accepted_frac": float(acc.mean()),
                        "rms_err_rel_r_accepted_max": float(errs[acc].max()) if acc.any() else None}
    A.set_learned(None)
    m.close()
    log(f"learned smooth: {out}")
    return out


def cpu_baseline(sg, co, tris, dabs, budget_s, threads):
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po

    po.set_threads(threads)
    o = po.Oracle(co, tris, leaf_limit=128, fp64=False)
    used = po.lib().orc_get_threads()
    n = 0
    aff = 0
    t0 = time.perf_counter()
    for d in dabs:
        aff += len(o.dab(d))
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    o.close()
    return {"dabs": n, "affected": aff, "seconds": dt, "threads": used}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=10)
    ap.add_argument("--dabs", type=int, default=1000)
    ap.add_argument("--mode", choices=["graph", "stream"], default="graph")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of oracle work for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stats", action="store_true")
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tune-dabs", type=int, default=200)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-dabs", type=int, default=500)
    ap.add_argument("--learned-steps", type=int, default=3)
    ap.add_argument("--no-judge", action="store_true")
    ap.add_argument("--only-learned", action="store_true",
                    help="run only the learned Smooth section (profiling); prints its JSON")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one mesh stroke per GPU; only --gpus 1 is supported "
                                   "(multi-GPU sharding is a design note, not built)", "n_gpus": args.gpus}))
        sys.exit(2)

    import dune_sculpt_amd as sg

    t = time.perf_counter()
    co, tris = sg.noisy_icosphere(args.level, seed=1, amplitude=0.02)
    log(f"mesh icosphere-{args.level}: {co.shape[0]} verts {tris.shape[0]} tris in {time.perf_counter() - t:.1f}s")
    dabs, tools = make_c2_stroke(sg, args.dabs)

    t = time.perf_counter()
    A = sg.PBVH(co, tris, leaf_limit=128)
    build_s = time.perf_counter() - t
    info = A.tree_info()
    log(f"PBVH build {build_s:.1f}s: {info}")
    if args.grid:
        A.set_grid(args.grid)

    W, K = args.warmup, args.steps
    A.snapshot()  # the pristine mesh: every step applies the stroke to it
    if args.only_learned:
        print(json.dumps({"learned_smooth": run_learned(sg, A, W, args)}))
        A.close()
        return
    tune = None
    if not args.no_tune:
        t = time.perf_counter()
        before, after = A.autotune(dabs[:args.tune_dabs])
        tune = {"sample_dabs": args.tune_dabs, "before_ms": round(before, 3), "after_ms": round(after, 3),
                "grids": A.kernel_grids(), "seconds": round(time.perf_counter() - t, 2)}
        log(f"autotune: {tune}")
    t = time.perf_counter()
    prepared = A.prepare(dabs, log=True) if args.mode == "graph" else None
    prep_s = time.perf_counter() - t
    log(f"stroke capture+instantiate {prep_s * 1e3:.1f} ms")

    def step():
        if prepared is not None:
            prepared.launch(sync=False)
        else:
            A.stroke(dabs, graph=False, log=True)

    step_times, event_ms = [], []
    for s in range(W + K):
        A.restore()  # untimed: back to the pristine mesh
        A.sync()
        t0 = time.perf_counter()
        A.timer_start()
        step()
        ms = A.timer_stop()  # HIP events on the engine's stream around the whole stroke
        if s >= W:
            step_times.append(time.perf_counter() - t0)
            event_ms.append(ms)
    elapsed = sum(step_times)
    event_s = sum(event_ms) / 1e3
    logrows = A.stroke_log(args.dabs)
    log(f"timed {K} strokes in {elapsed:.3f}s ({elapsed / K * 1e3:.2f} ms/stroke, "
        f"{elapsed / K / args.dabs * 1e6:.2f} us/dab)")

    # exact byte accounting: the same stroke from the same mesh with the kernels'
    # statistics counters on (untimed); every timed step repeats it exactly
    B = sg.PBVH(co, tris, leaf_limit=128)
    B.set_stats(not args.no_stats)
    B.stroke(dabs, graph=False, log=True)
    rows = B.stroke_log(args.dabs)
    B.close()
    per_stroke_aff = int(rows[:, 1].sum())
    assert int(logrows[:, 1].sum()) == per_stroke_aff, "stats replay diverged from the timed run"
    kern_bytes = {}
    stroke_bytes = 0
    for i in range(args.dabs):
        bb = dab_bytes(sg, int(tools[i]), 4, rows[i])
        stroke_bytes += bb["total"]
        for key, val in bb.items():
            kern_bytes[key] = kern_bytes.get(key, 0) + val
    timed_aff = per_stroke_aff * K
    timed_bytes = stroke_bytes * K
    vups = timed_aff / elapsed
    gbps = timed_bytes / elapsed / 1e9

    # Roofline of the dab step: the stroke graph is one dependent chain of small kernels
    # (k_search -> [k_area] -> k_brush -> [k_smooth_iter x3] -> k_post per dab), so
    # its unit is the dab: algorithmic bytes per dab over event time per dab.
    us_per_dab_ev = event_s / (K * args.dabs) * 1e6
    bytes_per_dab = stroke_bytes / args.dabs
    achieved = bytes_per_dab / (us_per_dab_ev * 1e-6) / 1e9
    roofline = {"bound": "hbm", "kernel": "dab step (stroke hipGraph, all kernels of one dab)",
                "achieved": round(achieved, 3), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 6), "traffic": TRAFFIC_BYTES_PER_DAB,
                "alg_bytes_per_launch": round(bytes_per_dab, 1), "avg_launch_us": round(us_per_dab_ev, 3)}
    kernel_bytes_per_dab = {k: round(v / args.dabs, 1) for k, v in kern_bytes.items()}

    c3 = None
    if not args.no_learned:
        c3 = run_learned(sg, A, W, args)

    cpu = None
    if not args.no_cpu:
        threads = os.cpu_count() or 1
        log(f"CPU baseline: oracle, {threads} threads, budget {args.cpu_budget}s")
        r = cpu_baseline(sg, co, tris, dabs, args.cpu_budget, threads)
        cpu = {"value": r["affected"] / r["seconds"], "unit": "vertex-updates/s", "cores": r["threads"],
               "kind": "port",
               "sample": f"first {r['dabs']} dabs of the same stroke on the same icosphere-10 mesh "
                         f"({r['seconds']:.1f}s, C oracle -O2 -ffp-contract=off, OpenMP over gathered leaves)"}
    A.close()

    line = {
        "metric": METRIC,
        "value": vups,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"icosphere-{args.level} noisy 0.02x4oct ({co.shape[0]} V, {tris.shape[0]} T), "
                               f"{args.dabs}-dab stroke cycling Draw/Inflate/Smooth(4)/Grab, r~U[0.02,0.1], "
                               f"s~U[0.2,1.0], leaf<=128 tris",
                   "dabs_per_step": args.dabs, "mode": args.mode, "leaves": info["leaves"],
                   "search_subtrees": info["cuts"]},
        "hbm_gbps": gbps,
        "hbm_frac": gbps / HBM_PEAK_GBPS,
        "us_per_dab": elapsed / K / args.dabs * 1e6,
        "affected_per_dab": timed_aff / (K * args.dabs),
        "alg_bytes_per_dab": timed_bytes / (K * args.dabs),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "event_s": event_s,
        "alg_bytes_per_dab_by_kernel": kernel_bytes_per_dab,
        "learned_smooth": c3,

        "build_s": build_s,
        "autotune": tune,
        "step_ms": [round(t * 1e3, 3) for t in step_times],
        "stroke_prepare_ms": prep_s * 1e3,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
