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
and launched per step; every step starts from the same pristine mesh (a device
snapshot restored, untimed, before each step).

value = vertex-updates/s = sum over timed dabs of |affected set| / wall time.
Algorithmic HBM bytes per dab follow the byte model in `dab_bytes` below (FP32 12 B
per vec3, int32 topology, each item counted once per dab), counted exactly by an
untimed replay of the same strokes with the kernels' statistics counters on.
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

from dune_sculpt_amd.workloads import batch_meshes, batch_table, plain_stroke, smooth_stroke  # noqa: E402

HBM_PEAK_GBPS = 8000.0
# HBM bytes per dab of the 1000-dab stroke on the 10.5M-vertex mesh (fused dabs), from
# two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) summed over the dab kernels,
# FETCH_SIZE scaled by 2.00 and WRITE_SIZE by 1.00 (what they read for a 1 GiB device
# copy); profiles/dab_step_c2.md holds the commands and the per-kernel figures.
TRAFFIC_BYTES_PER_DAB = 11_878_502
METRIC = "vertex-updates/s + HBM GB/s; learned vs plain Smooth dabs/s at ≤1e-3 RMS"
NO_PERSISTENT = False  # set from --no-persistent


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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
    plain = smooth_stroke(sg, n)
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


def batch_bytes(t):
    """Lower bound on the HBM bytes of the batch's dab work from its exact totals (SoA
    FP32 12 B per vec3, int32 topology, 16 B per flattened incidence): box tests 24 B,
    candidate positions 12 B, affected writes 12 B, Draw area normals 12 B, Smooth k
    passes over the affected records (12 B read + 12 B write + 4 B per neighbour index;
    neighbour positions counted only as the record itself, a lower bound), learned
    records (one neighbour pass + write), normals (16 B per incidence + 12 B read + 12 B
    write per affected vertex), refit (12 B per shared vertex + 24 B per leaf)."""
    b = 24 * t["tests"] + 12 * t["candidates"] + 12 * t["affected"] + 12 * t["area_verts"]
    b += 24 * t["smooth_records_k"] + 4 * t["smooth_nbr_reads"]
    b += 24 * t["learned_records"]
    b += 16 * t["deg_t"] + 24 * t["affected"]
    b += 12 * t["shared"] + 24 * t["gathered"]
    return b


def run_batch(sg, name, meshes, table, W, K, tune_steps, learned_model=None, plain_table=None):
    """Time the [steps][mesh] table on one batch from a pristine snapshot: W warmup + K
    timed runs, HIP events around each run (one graph launch per step)."""
    t = time.perf_counter()
    B = sg.Batch(meshes, leaf_limit=128)
    build_s = time.perf_counter() - t
    info = B.info()
    log(f"{name}: batch of {info['meshes']} meshes, {info['verts']} V, {info['leaves']} leaves, "
        f"{info['cuts']} cuts, built in {build_s:.1f}s")
    B.snapshot()
    out = {"meshes": info["meshes"], "verts": info["verts"], "tris": info["tris"], "leaves": info["leaves"],
           "cuts": info["cuts"], "steps": len(table), "build_s": round(build_s, 2)}
    if learned_model is not None:
        B.set_learned(learned_model)
    out["persistent"] = bool(B.set_persistent(not NO_PERSISTENT)) and not NO_PERSISTENT
    if tune_steps:
        t = time.perf_counter()
        before, after = B.autotune(table[:tune_steps])
        out["autotune"] = {"sample_steps": tune_steps, "before_ms": round(before, 3), "after_ms": round(after, 3),
                           "launch": B.launch_shape(), "seconds": round(time.perf_counter() - t, 2)}
        log(f"{name}: autotune {out['autotune']}")

    def timed(tab):
        B.load(tab)
        times = []
        tot = None
        for s in range(W + K):
            B.restore()
            B.rewind()
            B.totals(reset=True)
            B.timer_start()
            B.launch(sync=False)
            ms = B.timer_stop()
            if s >= W:
                times.append(ms)
                tot = B.totals()
        ms = float(np.mean(times))
        n_dabs = sum(1 for row in tab for d in row if d is not None)
        r = {"ms_per_run": ms, "us_per_step": ms * 1e3 / len(tab), "dabs_per_s": n_dabs / (ms * 1e-3),
             "vertex_updates_per_s": tot["affected"] / (ms * 1e-3), "affected_per_run": tot["affected"],
             "alg_bytes_per_run": batch_bytes(tot), "totals": tot}
        r["hbm_gbps"] = r["alg_bytes_per_run"] / (ms * 1e-3) / 1e9
        return r

    out["run"] = timed(table)
    if plain_table is not None:
        out["plain"] = timed(plain_table)
        out["learned_speedup_dabs_per_s"] = out["run"]["dabs_per_s"] / out["plain"]["dabs_per_s"]
    B.close()
    log(f"{name}: {json.dumps(out)[:600]}")
    return out


def run_batch_workloads(sg, args, W):
    """Many meshes: 256 meshes of 5k-500k vertices (log-uniform; icospheres and tori,
    noise 0.01), 64 steps of Draw / Smooth(4) with C2's dab distribution; and 64
    meshes of 1M-4M vertices, 128 steps cycling Draw / Inflate / learned Smooth(4),
    r ~ U[0.02, 0.08], timed against the same table with plain Smooth(4)."""
    res = {}
    if not args.no_c4:
        t = time.perf_counter()
        meshes, kinds = batch_meshes(sg, args.batch_meshes, seed=11, vmin=5000, vmax=500000, amplitude=0.01)
        table = batch_table(sg, meshes, kinds, args.batch_steps, seed=12, tools=(sg.TOOL_DRAW, sg.TOOL_SMOOTH),
                            rmin=0.02, rmax=0.1)
        log(f"c4 meshes and strokes in {time.perf_counter() - t:.1f}s")
        res["c4"] = run_batch(sg, "c4", meshes, table, W, 3, tune_steps=0 if args.no_tune else 8)
        res["c4"]["workload"] = (f"{len(meshes)} meshes, V log-uniform [5k, 500k] (icospheres / tori 1.0/0.3, noise "
                                 f"0.01), {args.batch_steps} Draw/Smooth(4) dabs each, r~U[0.02,0.1], s~U[0.2,1.0]")
        del meshes, table
    if not args.no_c5:
        from dune_sculpt_amd import learned as L

        mlp = sg.MLP.load(L.MODEL_PATH) if L.MODEL_PATH.exists() else None
        t = time.perf_counter()
        meshes, kinds = batch_meshes(sg, args.c5_meshes, seed=21, vmin=1e6, vmax=4e6, amplitude=0.02,
                                     log_uniform=False, levels=(9, 9))
        tools = (sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH)
        tl = batch_table(sg, meshes, kinds, args.c5_steps, seed=22, tools=tools, rmin=0.02, rmax=0.08,
                         learned_smooth=mlp is not None)
        tp = batch_table(sg, meshes, kinds, args.c5_steps, seed=22, tools=tools, rmin=0.02, rmax=0.08)
        log(f"c5 meshes and strokes in {time.perf_counter() - t:.1f}s")
        res["c5"] = run_batch(sg, "c5", meshes, tl, W, 2, tune_steps=0 if args.no_tune else 8, learned_model=mlp,
                              plain_table=tp)
        res["c5"]["workload"] = (f"{len(meshes)} meshes of 1M-4M V (icosphere-9 / tori, noise 0.02), "
                                 f"{args.c5_steps} dabs each cycling Draw/Inflate/learned Smooth(4), r~U[0.02,0.08]")
        if mlp is not None:
            mlp.close()
    return res


def camera_rays(n_side, eye_dist=3.0, seed=7):
    """A pinhole camera eye_dist from the origin looking at it; one ray per pixel of an
    n_side x n_side image whose field of view just covers a unit sphere (the image
    corners miss).  Returns starts and unit directions (FP32)."""
    rng = np.random.default_rng(seed)
    fwd = rng.normal(size=3)
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, [0.0, 0.0, 1.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    half = np.tan(np.arcsin(1.05 / eye_dist))
    u = (np.arange(n_side) + 0.5) / n_side * 2.0 - 1.0
    uu, vv = np.meshgrid(u * half, u * half, indexing="ij")
    d = fwd[None, :] + uu.reshape(-1, 1) * right[None, :] + vv.reshape(-1, 1) * up[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    starts = np.repeat((-fwd * eye_dist)[None, :], d.shape[0], axis=0)
    return starts.astype(np.float32), d.astype(np.float32)


def _hip():
    import ctypes as C

    return C.CDLL("libamdhip64.so")


def run_raycast(sg, A, W, K, n_side):
    """Ray cast throughput on the mesh as it is: one camera image of n_side^2 rays,
    rays and hits resident in HBM, HIP events around each batch on the engine's
    stream.  Algorithmic bytes per ray of this kernel's algorithm: 24 per box tested
    (search-cut and leaf boxes), 52 per triangle tested (corner ids 16 + three corner
    positions 36), 32 per ray read and 56 per hit written, counted exactly by an
    untimed pass with the work counters on."""
    import ctypes as C

    starts, dirs = camera_rays(n_side)
    n = starts.shape[0]
    rays = sg.make_rays(starts, dirs)
    hip = _hip()
    dr, dh = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(dr), C.c_size_t(C.sizeof(rays))) or \
            hip.hipMalloc(C.byref(dh), C.c_size_t(n * C.sizeof(sg.SgRayHit))):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemcpy(dr, rays, C.c_size_t(C.sizeof(rays)), 1)
    L = sg.lib()
    times = []
    for s in range(W + K):
        A.sync()
        A.timer_start()
        if L.sg_raycast_device(A.h, dr, n, 0, dh) != 0:
            raise RuntimeError(sg.last_error())
        ms = A.timer_stop()
        if s >= W:
            times.append(ms)
    out = (sg.SgRayHit * n)()
    hip.hipMemcpy(out, dh, C.c_size_t(C.sizeof(out)), 2)
    hits = sg.hits_to_numpy(out)
    hip.hipFree(dr)
    hip.hipFree(dh)
    st = A.raycast_stats(starts, dirs)
    ms = float(np.mean(times))
    alg = 24 * st["boxes"] + 52 * st["tris"] + (32 + 56) * n
    gbps = alg / (ms * 1e-3) / 1e9
    res = {"rays": n, "image": f"{n_side}x{n_side} pinhole camera at distance 3",
           "hit_frac": float(hits["hit"].mean()), "ms_per_batch": ms, "rays_per_s": n / (ms * 1e-3),
           "us_per_ray_amortized": ms * 1e3 / n, "boxes_per_ray": st["boxes"] / n, "tris_per_ray": st["tris"] / n,
           "leaves_scanned_per_ray": st["leaves"] / n,
           "roofline": {"bound": "hbm", "kernel": "k_raycast (one workgroup per ray)", "achieved": round(gbps, 3),
                        "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 6),
                        "alg_bytes_per_launch": alg, "avg_launch_us": round(ms * 1e3, 3)}}
    log(f"raycast: {res}")
    return res, starts, dirs


def placed_rays(sg, centres, seed=11, eye=3.0, jitter=0.05):
    """Stroke input as a camera would give it: for each dab, a ray from eye units out
    along the centre's direction (plus jitter) aimed at the centre."""
    rng = np.random.default_rng(seed)
    c = np.asarray(centres, np.float64)
    u = c / np.linalg.norm(c, axis=1, keepdims=True)
    starts = u * eye + rng.normal(size=c.shape) * jitter
    d = c - starts
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return starts.astype(np.float32), d.astype(np.float32)


def run_placed(sg, A, dabs, W, K, only_placed=False, stream=False):
    """The headline stroke with every dab placed by a GPU ray cast inside the stroke
    graph (dab i's centre = where ray i hits the mesh the previous dabs left) against
    the same dabs at fixed centres; both captured once and launched per step from the
    pristine mesh."""
    centres = np.array([list(d.location) for d in dabs], np.float32)
    starts, dirs = placed_rays(sg, centres)
    res = {}
    for name in (("placed",) if only_placed else ("fixed", "placed")):
        prep = None
        if not stream:
            prep = A.prepare(dabs) if name == "fixed" else A.prepare_placed(dabs, starts, dirs)
        times = []
        for s in range(W + K):
            A.restore()
            A.sync()
            A.timer_start()
            if prep is not None:
                prep.launch(sync=False)
            elif name == "fixed":
                A.stroke(dabs, graph=False)
            else:
                A.stroke_placed(dabs, starts, dirs, graph=False)
            ms = A.timer_stop()
            if s >= W:
                times.append(ms)
        if name == "placed":
            pl = A.placements(len(dabs))
            res["hit_frac"] = float(pl[:, 6].mean())
        if prep is not None:
            prep.close()
        ms = float(np.mean(times))
        res[name] = {"ms_per_stroke": ms, "us_per_dab": ms * 1e3 / len(dabs)}
    if "fixed" in res:
        res["placement_us_per_dab"] = res["placed"]["us_per_dab"] - res["fixed"]["us_per_dab"]
    log(f"placed stroke: {res}")
    return res


def run_draw(sg, A, dabs, W, K):
    """Draw buffers of the whole mesh in the reference's 36-byte vertex layout, built in
    HBM (one face set): a full refill timed with HIP events, then the incremental
    refill of the leaves one stroke gathered.  Algorithmic bytes of a full refill: 108
    written per visible tri, 28 read per tri (corner ids, slot, polygon, face set) and
    24 per vertex (position, normal) once."""
    A.restore()
    P = A.tottri
    A.draw_build(np.ones(P, np.int32))
    info = A.draw_info()
    times = []
    for s in range(W + K):
        A.sync()
        A.timer_start()
        sg.lib().sg_pbvh_draw_buffers_update(A.h, 1)
        ms = A.timer_stop()
        if s >= W:
            times.append(ms)
    full_ms = float(np.mean(times))
    alg = info["visible_tris"] * (108 + 28) + A.totvert * 24
    A.stroke(dabs, graph=False)  # the dabs mark the leaves they gather
    A.sync()
    A.timer_start()
    sg.lib().sg_pbvh_draw_buffers_update(A.h, 0)
    inc_ms = A.timer_stop()
    gbps = alg / (full_ms * 1e-3) / 1e9
    res = {"visible_tris": info["visible_tris"], "vbo_bytes": info["bytes"], "full_refill_ms": full_ms,
           "incremental_refill_ms_after_stroke": inc_ms,
           "roofline": {"bound": "hbm", "kernel": "k_draw_fill (full refill)", "achieved": round(gbps, 3),
                        "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 6),
                        "alg_bytes_per_launch": alg, "avg_launch_us": round(full_ms * 1e3, 3)}}
    sg.lib().sg_pbvh_draw_buffers_free(A.h)
    log(f"draw buffers: {res}")
    return res


def run_undo(sg, A, launch, base_ms, n_dabs, W, K):
    """Undo steps of the headline stroke: the stroke timed with a step open (every dab
    saves the leaves it gathers the first time it writes them), then undo and redo of
    the step.  An undo or redo swaps each saved vertex's position, normal and record
    (read and write of position, normal and record, write of the shadow copy: 108 bytes
    per vertex) and refits the touched subtrees."""
    rec_ms = []
    info = None
    for s in range(W + K):
        A.restore()
        A.undo_begin()
        A.sync()
        A.timer_start()
        launch()
        ms = A.timer_stop()
        A.undo_end()
        info = A.undo_info()
        if s >= W:
            rec_ms.append(ms)
        if s < W + K - 1:
            A.undo_clear()
    undo_ms, redo_ms = [], []
    for _ in range(max(K, 3)):
        A.sync()
        A.timer_start()
        assert A.undo()
        undo_ms.append(A.timer_stop())
        A.timer_start()
        assert A.redo()
        redo_ms.append(A.timer_stop())
    A.undo_clear()
    n_v = (info["bytes"] - 80 * info["leaves"]) // 24
    rec = float(np.mean(rec_ms))
    um, rm = float(np.mean(undo_ms)), float(np.mean(redo_ms))
    res = {"step_leaves": info["leaves"], "step_vertices": int(n_v), "step_bytes": info["bytes"],
           "stroke_ms_recording": rec, "stroke_ms_plain": base_ms,
           "recording_us_per_dab": (rec - base_ms) * 1e3 / n_dabs,
           "undo_ms": um, "redo_ms": rm,
           "undo_gbps": round(108 * n_v / (um * 1e-3) / 1e9, 3)}
    log(f"undo: {res}")
    return res


def run_proxies(sg, A, reps):
    """Proxies combined on the GPU: one proxy on every tenth leaf (~10% of the mesh),
    filled on the host, then sg_pbvh_combine_proxies (pack + host-to-device copy +
    combine kernel + normals + leaf and tree refit), timed by wall clock."""
    A.restore()
    leaves = A.search_gather()[::10]
    nverts = sum(A.node_verts(l)[1] for l in leaves)
    times = []
    for r in range(reps + 1):
        for l in leaves:
            A.add_proxy(l)[:] = 1e-4
        A.sync()
        t0 = time.perf_counter()
        n = A.combine_proxies()
        dt = time.perf_counter() - t0
        assert n == len(leaves)
        if r:
            times.append(dt)
    ms = float(np.mean(times)) * 1e3
    res = {"leaves": len(leaves), "vertices": int(nverts), "combine_ms": ms,
           "vertices_per_s": nverts / (ms * 1e-3), "host_bytes_uploaded": 12 * int(nverts)}
    log(f"proxies: {res}")
    A.restore()
    return res


def run_elastic(sg, A, W, K):
    """Elastic deform dabs on the whole mesh (every leaf; each visible vertex moved by a
    Kelvinlet of the dab), one per mode, from the current coordinates.  k_elastic is
    timed by the engine's per-kernel events; its algorithmic bytes: 36 per vertex (read
    the position, write it and its shadow copy) and 8 per leaf (its vertex range)."""
    A.restore()
    info = A.tree_info()
    nbytes = 36 * A.totvert + 8 * info["leaves"]
    res = {"vertices": A.totvert, "alg_bytes_per_launch": nbytes, "modes": {}}
    names = ["grab", "grab_biscale", "grab_triscale", "scale", "twist"]
    co = A.positions()
    for mode in range(5):
        d = sg.make_dab(tool=sg.TOOL_ELASTIC_DEFORM, location=co[12345], radius=0.1, strength=0.5,
                        grab_delta=(0.001, 0.0, 0.0), elastic_type=mode, volume_preservation=0.4,
                        surface_normal=(0.0, 0.0, 1.0))
        for _ in range(W):
            A.dab(d)
        A.set_profile(True)
        A.profile_read(reset=True)
        walls = []
        for _ in range(K):
            A.sync()
            A.timer_start()
            A.dab(d)
            walls.append(A.timer_stop())
        prof = A.profile_read(reset=True)
        A.set_profile(False)
        kms, kn = prof["k_brush"]
        us = kms / max(kn, 1) * 1e3
        gbps = nbytes / (us * 1e-6) / 1e9
        res["modes"][names[mode]] = {"dab_ms": float(np.mean(walls)), "k_elastic_us": round(us, 2),
                                     "gbps": round(gbps, 1)}
    best = max(res["modes"].values(), key=lambda m: m["gbps"])
    res["roofline"] = {"bound": "hbm", "kernel": "k_elastic (grab)", "achieved": res["modes"]["grab"]["gbps"],
                       "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                       "frac": round(res["modes"]["grab"]["gbps"] / HBM_PEAK_GBPS, 4),
                       "avg_launch_us": res["modes"]["grab"]["k_elastic_us"], "best_mode_gbps": best["gbps"]}
    log(f"elastic: {res}")
    A.restore()
    return res


def cpu_raycast(o, starts, dirs, budget_s):
    """The oracle's sequential ray cast (the reference's traversal and callback) on the
    same camera rays, one host thread, until the budget is spent."""
    n = 0
    t0 = time.perf_counter()
    for i in range(len(starts)):
        o.raycast(starts[i], dirs[i])
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "rays/s", "cores": 1, "kind": "port",
            "sample": f"first {n} rays of the same camera image ({dt:.1f}s, C oracle, one thread)"}


def cpu_draw(o, tris, budget_s):
    """The oracle's per-leaf draw fill (the reference's gpu_pbvh_mesh_buffers_update
    restated) on the same mesh, one host thread, leaf after leaf until the budget."""
    T = len(tris)
    tri_poly = np.arange(T, dtype=np.int32)
    off = np.arange(0, 3 * T + 1, 3, dtype=np.int32)
    pv = np.ascontiguousarray(tris.reshape(-1), np.int32)
    sm = np.zeros(T, np.uint8)
    mt = np.zeros(T, np.int32)
    fs = np.ones(T, np.int32)
    n = 0
    t0 = time.perf_counter()
    for leaf in o.leaves_dfs():
        recs, _, _ = o.draw_leaf(int(leaf), tri_poly, off, pv, sm, mt, fs)
        n += len(recs) // 3
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "tris/s", "cores": 1, "kind": "port",
            "sample": f"{n} tris of the same mesh's leaves ({dt:.1f}s, C oracle, one thread, Python per leaf)"}


def run_c1(sg, W, K):
    """The reference's CPU-sized case: icosphere-6 (40,962 V), 200 Draw dabs, r = 0.1,
    s = 0.5, smooth falloff, great circle with spacing 0.1 r, leaves <= 128 tris.  GPU:
    the stroke as a captured graph from the pristine mesh (HIP events around it); CPU:
    the oracle on the same dabs with 1 thread and with every host thread."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po

    co, tris = sg.icosphere(6)
    n = 200
    radii = np.full(n, 0.1, np.float32)
    dabs = [sg.make_dab(tool=sg.TOOL_DRAW, location=c, radius=0.1, strength=0.5)
            for c in sg.great_circle_centres(1, radii)]
    P = sg.PBVH(co, tris, leaf_limit=128)
    P.snapshot()
    prep = P.prepare(dabs, log=True)
    times = []
    for s in range(W + max(K, 5)):
        P.restore()
        P.sync()
        P.timer_start()
        prep.launch(sync=False)
        ms = P.timer_stop()
        if s >= W:
            times.append(ms)
    aff = int(P.stroke_log(n)[:, 1].sum())
    prep.close()
    P.close()
    gpu_ms = float(np.median(times))
    out = {"workload": "icosphere-6 (40962 V, 81920 T), 200 Draw dabs r=0.1 s=0.5 smooth falloff, leaf<=128 tris",
           "affected": aff, "gpu_ms_per_stroke": gpu_ms, "gpu_us_per_dab": gpu_ms * 1e3 / n,
           "gpu_vertex_updates_per_s": aff / (gpu_ms * 1e-3)}
    for threads in (1, os.cpu_count() or 1):
        po.set_threads(threads)
        o = po.Oracle(co, tris, leaf_limit=128, fp64=False)
        used = po.lib().orc_get_threads()
        t0 = time.perf_counter()
        a = sum(len(o.dab(d)) for d in dabs)
        dt = time.perf_counter() - t0
        o.close()
        assert a == aff, "CPU and GPU affected counts differ"
        out[f"cpu_{used}t_ms_per_stroke"] = dt * 1e3
        out[f"cpu_{used}t_vertex_updates_per_s"] = a / dt
    log(f"c1: {out}")
    return out


def run_grids(sg, W, K, args):
    """Multires grids (PBVH_GRIDS): a cube of 64 x 64 quads per side on the unit sphere,
    subdivision level 4 (98,304 grids of 9 x 9 = 7,962,624 elements) displaced by the
    same seeded noise as the mesh workload (amplitude 0.02), the same 1000-dab stroke
    (Draw / Inflate / Smooth(4) / Grab), leaves of max(10000 / 81, 1) = 123 grids.  GPU:
    the stroke as a captured graph from the restored state (HIP events around it);
    CPU: the oracle (one thread) on the first dabs of the same stroke within a budget,
    its affected counts checked against the GPU's."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    from dune_sculpt_amd import grids as gr

    t = time.perf_counter()
    co, off, pv = gr.quad_sphere(args.grids_n)
    g = gr.Grids(co, off, pv, level=args.grids_level)
    c = g.co()
    d = (c / np.linalg.norm(c, axis=1, keepdims=True)).astype(np.float32)
    g.set_co(sg.noise_displace(d.copy(), d, 1, 0.02))
    g.recalc_normals()
    P = gr.GridPBVH(g)
    build_s = time.perf_counter() - t
    dabs, _ = plain_stroke(sg, args.dabs)
    P.snapshot()
    prep = None if args.grids_stream else P.prepare(dabs, log=True)
    times = []
    for s in range(W + max(K, 3)):
        P.restore()
        P.timer_start()
        if prep:
            prep.launch(sync=False)
        else:
            P.stroke(dabs, graph=False, log=True)
        ms = P.timer_stop()
        if s >= W:
            times.append(ms)
    counts = P.stroke_log(len(dabs))[:, 1]
    aff = int(counts.sum())
    if prep:
        prep.close()
    gpu_ms = float(np.median(times))
    out = {"workload": f"quad sphere {args.grids_n}x{args.grids_n} per side, level {args.grids_level} "
                       f"({g.G} grids of {g.gs}x{g.gs}, {g.E} elements), noise 0.02x4oct, "
                       f"{len(dabs)}-dab Draw/Inflate/Smooth(4)/Grab stroke, leaf limit 10000 elements",
           "mode": "stream" if args.grids_stream else "graph",
           "build_s": build_s, "leaves": P.leaves, "affected": aff,
           "gpu_ms_per_stroke": gpu_ms, "gpu_us_per_dab": gpu_ms * 1e3 / len(dabs),
           "gpu_element_updates_per_s": aff / (gpu_ms * 1e-3)}
    if not args.no_cpu:
        P.restore()
        P.sync_host()
        o = po.OracleGrids(len(co), off, pv, g.gs, g.co(), g.normals())
        t0 = time.perf_counter()
        n = a = 0
        for i, dab in enumerate(dabs):
            k = len(o.dab(po.to_orc(dab)))
            assert k == counts[i], f"grids dab {i}: oracle {k} vs GPU {counts[i]} affected"
            a += k
            n += 1
            if time.perf_counter() - t0 > args.cpu_budget / 2:
                break
        dt = time.perf_counter() - t0
        o.close()
        out["cpu_baseline"] = {"value": a / dt, "unit": "element-updates/s", "cores": 1, "kind": "port",
                               "sample": f"first {n} dabs ({dt:.1f}s, C oracle, one thread)"}
    P.close()
    g.close()
    log(f"grids: {out}")
    return out


def run_dyntopo(sg, W, K, args):
    """Dynamic topology (PBVH_BMESH): a noisy icosphere (level args.dyn_level, noise
    0.02 x 4 octaves), detail size args.dyn_detail (splits past it, collapses under 0.4 of
    it, on the leaves within 1.25 r of each Draw / Inflate / Smooth dab), a stroke of
    args.dyn_dabs dabs cycling Draw / Inflate / Smooth(2) / Grab, r ~ U[0.02, 0.05].  Each
    step rebuilds the mesh and PBVH (untimed) and times the stroke (the host's clock around
    it: each dab waits once, at its end, for its counters).  CPU: the oracle, one thread, on the
    first dabs of the same stroke within a budget, every dab's counters checked against
    the GPU's."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    from dune_sculpt_amd import dyntopo as dy

    co, tris = sg.noisy_icosphere(args.dyn_level, seed=1, amplitude=0.02)
    rng = np.random.default_rng(5)
    n = args.dyn_dabs
    radii = rng.uniform(0.02, 0.05, n).astype(np.float32)
    strengths = rng.uniform(0.2, 1.0, n)
    centres = sg.great_circle_centres(5, radii, spacing_frac=0.3)
    tools = [sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH, sg.TOOL_GRAB]
    dabs = []
    for i in range(n):
        nxt = centres[i + 1] if i + 1 < n else centres[i]
        dabs.append(sg.make_dab(tool=tools[i % 4], location=centres[i], radius=float(radii[i]),
                                strength=float(strengths[i]), smooth_iterations=2, grab_delta=(nxt - centres[i]) * 0.5))
    times, stats, build = [], None, []
    for s in range(W + K):
        t = time.perf_counter()
        m = dy.DynMesh(co, tris)
        P = dy.DynPBVH(m)
        P.detail_size_set(args.dyn_detail)
        build.append(time.perf_counter() - t)
        st = []
        t = time.perf_counter()
        for d in dabs:
            P.dab(d)
            st.append(P.last_stats())
        dt = time.perf_counter() - t
        if s >= W:
            times.append(dt)
        stats = np.array(st)
        cnt = m.counts()
        P.close()
        m.close()
    sec = float(np.median(times))
    aff = int(stats[:, 1].sum())
    out = {"workload": f"icosphere-{args.dyn_level} noisy 0.02x4oct ({len(co)} V, {len(tris)} T), detail size "
                       f"{args.dyn_detail}, {n}-dab Draw/Inflate/Smooth(2)/Grab stroke, r~U[0.02,0.05]",
           "build_s": float(np.median(build)), "ms_per_stroke": sec * 1e3, "us_per_dab": sec * 1e6 / n,
           "dabs_per_s": n / sec, "vertex_updates_per_s": aff / sec,
           "per_dab": {"affected": aff / n, "splits": float(stats[:, 2].mean()),
                       "collapses": float(stats[:, 3].mean()), "split_rounds": float(stats[:, 4].mean()),
                       "collapse_rounds": float(stats[:, 5].mean()), "leaves_refit": float(stats[:, 11].mean())},
           "final_counts": cnt, "capacity_stops": int((stats[:, 8] & 1).sum())}
    if not args.no_cpu:
        o = po.OracleDyn(co, tris)
        o.detail_size_set(args.dyn_detail)
        t0 = time.perf_counter()
        k = 0
        for i, d in enumerate(dabs):
            o.dab(po.to_orc(d))
            assert np.array_equal(o.last_stats(), stats[i]), f"dyntopo dab {i}: oracle counters differ from the GPU's"
            k += 1
            if time.perf_counter() - t0 > args.cpu_budget / 2:
                break
        dt = time.perf_counter() - t0
        o.close()
        out["cpu_baseline"] = {"value": k / dt, "unit": "dabs/s", "cores": 1, "kind": "port",
                               "vertex_updates_per_s": float(stats[:k, 1].sum()) / dt,
                               "sample": f"first {k} dabs ({dt:.1f}s, C oracle, one thread)"}
    log(f"dyntopo: {out}")
    return out


def cpu_model():
    """The host CPU's model name, for the record beside the CPU baseline."""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sg, co, tris, dabs, budget_s, threads, before=None):
    """The oracle's dab loop on the same mesh; `before(oracle)` runs first on the fresh
    mesh (the ray cast sample shares the build)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po

    po.set_threads(threads)
    o = po.Oracle(co, tris, leaf_limit=128, fp64=False)
    if before is not None:
        before(o)
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
    ap.add_argument("--no-batch", action="store_true", help="skip the many-mesh workloads")
    ap.add_argument("--copy-mb", type=int, default=1024, help="size of the measured HBM copy (0: skip)")
    ap.add_argument("--only-batch", action="store_true", help="run only the many-mesh workloads")
    ap.add_argument("--no-persistent", action="store_true",
                    help="many-mesh workloads with one captured graph per step instead of the persistent step (A/B)")
    ap.add_argument("--batch-meshes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=64)
    ap.add_argument("--no-c4", action="store_true", help="skip the 256-mesh batch")
    ap.add_argument("--no-c5", action="store_true", help="skip the 64 x 1-4M-vertex learned batch")
    ap.add_argument("--c5-meshes", type=int, default=64)
    ap.add_argument("--c5-steps", type=int, default=128)
    ap.add_argument("--no-ray", action="store_true", help="skip the ray cast and placed-stroke measurements")
    ap.add_argument("--ray-side", type=int, default=128, help="camera image side (rays per batch = side^2)")
    ap.add_argument("--only-ray", action="store_true", help="run only the ray cast measurements")
    ap.add_argument("--no-placed", action="store_true", help="skip the placed-stroke measurement")
    ap.add_argument("--no-draw", action="store_true", help="skip the draw-buffer measurement")
    ap.add_argument("--no-undo", action="store_true", help="skip the undo-step measurement")
    ap.add_argument("--only-undo", action="store_true", help="run only the undo-step measurement")
    ap.add_argument("--no-elastic", action="store_true", help="skip the elastic deform measurement")
    ap.add_argument("--no-c1", action="store_true", help="skip the icosphere-6 CPU-reference case")
    ap.add_argument("--no-grids", action="store_true", help="skip the multires grids (PBVH_GRIDS) stroke")
    ap.add_argument("--only-grids", action="store_true", help="run only the multires grids stroke")
    ap.add_argument("--grids-n", type=int, default=64, help="coarse quads per cube side of the grids mesh")
    ap.add_argument("--grids-level", type=int, default=4, help="subdivision level of the grids (grid size 2^(l-1)+1)")
    ap.add_argument("--grids-stream", action="store_true", help="grids stroke as stream launches instead of a graph")
    ap.add_argument("--no-dyn", action="store_true", help="skip the dynamic topology (PBVH_BMESH) stroke")
    ap.add_argument("--only-dyn", action="store_true", help="run only the dynamic topology stroke")
    ap.add_argument("--dyn-level", type=int, default=8, help="icosphere level of the dynamic topology mesh")
    ap.add_argument("--dyn-detail", type=float, default=0.003, help="dynamic topology detail size (max edge length)")
    ap.add_argument("--dyn-dabs", type=int, default=200, help="dabs in the dynamic topology stroke")
    ap.add_argument("--no-fused", action="store_true", help="separate brush / Smooth / post kernels (A/B)")
    ap.add_argument("--only-elastic", action="store_true", help="run only the elastic deform measurement")
    ap.add_argument("--placed-only", action="store_true",
                    help="placed stroke without the fixed-centre reference stroke (one captured graph)")
    args = ap.parse_args()
    global NO_PERSISTENT
    NO_PERSISTENT = args.no_persistent

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one mesh stroke per GPU; only --gpus 1 is supported "
                                   "(multi-GPU sharding is a design note, not built)", "n_gpus": args.gpus}))
        sys.exit(2)

    import dune_sculpt_amd as sg

    if args.only_batch:
        print(json.dumps({"batch": run_batch_workloads(sg, args, args.warmup)}))
        return
    if args.only_grids:
        print(json.dumps({"grids": run_grids(sg, args.warmup, args.steps, args)}))
        return
    if args.only_dyn:
        print(json.dumps({"dyntopo": run_dyntopo(sg, args.warmup, min(args.steps, 3), args)}))
        return

    t = time.perf_counter()
    co, tris = sg.noisy_icosphere(args.level, seed=1, amplitude=0.02)
    log(f"mesh icosphere-{args.level}: {co.shape[0]} verts {tris.shape[0]} tris in {time.perf_counter() - t:.1f}s")
    dabs, tools = plain_stroke(sg, args.dabs)

    t = time.perf_counter()
    A = sg.PBVH(co, tris, leaf_limit=128)
    build_s = time.perf_counter() - t
    info = A.tree_info()
    halo = A.halo_info()
    if args.no_fused:
        A.set_fused(False)
    log(f"PBVH build {build_s:.1f}s: {info}, halo {halo}, fused {halo['built'] and not args.no_fused}")
    if args.grid:
        A.set_grid(args.grid)

    W, K = args.warmup, args.steps
    A.snapshot()  # the pristine mesh: every step applies the stroke to it
    if args.only_ray:
        out = {}
        if args.copy_mb:
            ms, gbps_copy = sg.hbm_copy_probe(args.copy_mb << 20)
            out["hbm_copy"] = {"bytes": args.copy_mb << 20, "ms": ms, "gbps_read_plus_write": gbps_copy}
        out["raycast"], _, _ = run_raycast(sg, A, W, max(K, 5), args.ray_side)
        if not args.no_placed:
            out["placed_stroke"] = run_placed(sg, A, dabs, W, K, args.placed_only, args.mode == "stream")
        if not args.no_draw:
            out["draw_buffers"] = run_draw(sg, A, dabs, W, max(K, 3))
        print(json.dumps(out))
        A.close()
        return
    if args.only_elastic:
        print(json.dumps({"elastic_deform": run_elastic(sg, A, W, max(K, 3))}))
        A.close()
        return
    if args.only_undo:
        out = {}
        if args.copy_mb:
            ms, gbps_copy = sg.hbm_copy_probe(args.copy_mb << 20)
            out["hbm_copy"] = {"bytes": args.copy_mb << 20, "ms": ms, "gbps_read_plus_write": gbps_copy}
        prepared = A.prepare(dabs) if args.mode == "graph" else None

        def launch():
            if prepared is not None:
                prepared.launch(sync=False)
            else:
                A.stroke(dabs, graph=False)

        base = []
        for s in range(W + K):
            A.restore()
            A.sync()
            A.timer_start()
            launch()
            ms = A.timer_stop()
            if s >= W:
                base.append(ms)
        out["undo"] = run_undo(sg, A, launch, float(np.mean(base)), args.dabs, W, K)
        out["proxies"] = run_proxies(sg, A, max(K, 2))
        print(json.dumps(out))
        A.close()
        return
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
    kern_bytes = {}
    stroke_bytes = 0
    if args.no_stats:  # profiling runs: no replay (its launches would join the profile)
        per_stroke_aff = int(logrows[:, 1].sum())
    else:
        B = sg.PBVH(co, tris, leaf_limit=128)
        B.set_stats(True)
        B.stroke(dabs, graph=False, log=True)
        rows = B.stroke_log(args.dabs)
        B.close()
        per_stroke_aff = int(rows[:, 1].sum())
        assert int(logrows[:, 1].sum()) == per_stroke_aff, "stats replay diverged from the timed run"
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
    # (k_search -> [k_area] -> k_dab_fused per dab), so
    # its unit is the dab: algorithmic bytes per dab over event time per dab.
    us_per_dab_ev = event_s / (K * args.dabs) * 1e6
    bytes_per_dab = stroke_bytes / args.dabs
    achieved = bytes_per_dab / (us_per_dab_ev * 1e-6) / 1e9
    roofline = {"bound": "hbm", "kernel": "dab step (stroke hipGraph, all kernels of one dab)",
                "achieved": round(achieved, 3), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 6),
                "traffic": TRAFFIC_BYTES_PER_DAB if (args.level == 10 and args.dabs == 1000) else None,
                "alg_bytes_per_launch": round(bytes_per_dab, 1), "avg_launch_us": round(us_per_dab_ev, 3)}
    kernel_bytes_per_dab = {k: round(v / args.dabs, 1) for k, v in kern_bytes.items()}

    undo = None if args.no_undo else run_undo(sg, A, step, float(np.mean(event_ms)), args.dabs, W, K)
    proxies = None if args.no_undo else run_proxies(sg, A, max(K, 2))
    elastic = None if args.no_elastic else run_elastic(sg, A, W, max(K, 3))

    copy = None
    if args.copy_mb:
        ms, gbps_copy = sg.hbm_copy_probe(args.copy_mb << 20)
        copy = {"bytes": args.copy_mb << 20, "ms": ms, "gbps_read_plus_write": gbps_copy}
        log(f"HBM copy probe: {copy}")

    ray = placed = None
    ray_starts = ray_dirs = None
    if not args.no_ray:
        A.restore()
        ray, ray_starts, ray_dirs = run_raycast(sg, A, W, max(K, 5), args.ray_side)
        placed = None if args.no_placed else run_placed(sg, A, dabs, W, K)
    draw = None if (args.no_ray or args.no_draw) else run_draw(sg, A, dabs, W, max(K, 3))

    c1 = None if args.no_c1 else run_c1(sg, W, K)
    grids = None if args.no_grids else run_grids(sg, W, K, args)
    dyn = None if args.no_dyn else run_dyntopo(sg, 1, 2, args)
    c3 = None
    if not args.no_learned:
        c3 = run_learned(sg, A, W, args)
    batch = None
    if not args.no_batch:
        batch = run_batch_workloads(sg, args, W)

    cpu = None
    if not args.no_cpu:
        threads = os.cpu_count() or 1
        log(f"CPU baseline: oracle, {threads} threads, budget {args.cpu_budget}s")
        ray_cpu, draw_cpu = {}, {}
        before = None
        if ray is not None or draw is not None:
            def before(o):
                if ray is not None:
                    ray_cpu.update(cpu_raycast(o, ray_starts, ray_dirs, args.cpu_budget / 4))
                if draw is not None:
                    draw_cpu.update(cpu_draw(o, tris, args.cpu_budget / 4))
        r = cpu_baseline(sg, co, tris, dabs, args.cpu_budget, threads, before)
        cpu = {"value": r["affected"] / r["seconds"], "unit": "vertex-updates/s", "cores": r["threads"],
               "kind": "port", "host": cpu_model(),
               "sample": f"first {r['dabs']} dabs of the same stroke on the same icosphere-10 mesh "
                         f"({r['seconds']:.1f}s, C oracle -O2 -ffp-contract=off, OpenMP over gathered leaves)"}
        r1 = cpu_baseline(sg, co, tris, dabs, args.cpu_budget / 2, 1)
        cpu["one_thread"] = {"value": r1["affected"] / r1["seconds"], "cores": r1["threads"],
                             "sample": f"first {r1['dabs']} dabs ({r1['seconds']:.1f}s)"}
        log(f"CPU baseline: {cpu}")
        if ray is not None:
            ray["cpu_baseline"] = ray_cpu
            log(f"CPU ray cast: {ray_cpu}")
        if draw is not None:
            draw["cpu_baseline"] = draw_cpu
            log(f"CPU draw fill: {draw_cpu}")
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
                   "search_subtrees": info["cuts"], "fused": bool(halo["built"] and not args.no_fused)},
        "halo": halo,
        "hbm_gbps": gbps,
        "hbm_frac": gbps / HBM_PEAK_GBPS,
        "us_per_dab": elapsed / K / args.dabs * 1e6,
        "affected_per_dab": timed_aff / (K * args.dabs),
        "alg_bytes_per_dab": timed_bytes / (K * args.dabs),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "event_s": event_s,
        "alg_bytes_per_dab_by_kernel": kernel_bytes_per_dab,
        "hbm_copy": copy,
        "c1": c1,
        "grids": grids,
        "dyntopo": dyn,
        "learned_smooth": c3,
        "batch": batch,
        "raycast": ray,
        "placed_stroke": placed,
        "draw_buffers": draw,
        "undo": undo,
        "proxies": proxies,
        "elastic_deform": elastic,

        "build_s": build_s,
        "autotune": tune,
        "step_ms": [round(t * 1e3, 3) for t in step_times],
        "stroke_prepare_ms": prep_s * 1e3,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
