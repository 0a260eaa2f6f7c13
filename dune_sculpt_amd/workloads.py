"""Seeded synthetic workloads of the sculpt engine: the meshes and dab strokes that
the benchmark times and the full-size parity tests check.

- plain strokes on one mesh: icosphere level 10 (10,485,762 verts) displaced by seeded
  value noise (amplitude 0.02, 4 octaves), 1000 dabs cycling Draw / Inflate /
  Smooth(k=4) / Grab, r ~ U[0.02, 0.1], s ~ U[0.2, 1.0], along a seeded great circle
  with spacing 10% of the radius (`plain_stroke`);
- learned Smooth on the same mesh: 500 Smooth(4) dabs, r = 0.05 (`smooth_stroke`);
- many meshes: seeded icospheres and tori (major / minor radius 1.0 / 0.3) of given
  vertex counts (`batch_meshes`) with a [step][mesh] table of dabs along each mesh's
  own path (`batch_table`).
"""
from __future__ import annotations

import numpy as np


def surface_centres(sg, centres, mesh_seed=1, amplitude=0.02):
    """Project great-circle points onto the noisy surface: the mesh was displaced
    radially by the same seeded noise field, so the surface point above a unit
    direction u is u displaced by that field."""
    u = centres / np.linalg.norm(centres, axis=1, keepdims=True)
    return sg.noise_displace(u.astype(np.float32), u.astype(np.float32), mesh_seed, amplitude)


def plain_stroke(sg, n_dabs=1000, seed=3):
    """The plain-brush stroke on the noisy level-10 icosphere (mesh seed 1): returns
    (dabs, tools)."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(0.02, 0.1, n_dabs).astype(np.float32)
    strengths = rng.uniform(0.2, 1.0, n_dabs).astype(np.float32)
    centres = surface_centres(sg, sg.great_circle_centres(seed + 1, radii))
    tools = np.array([sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH, sg.TOOL_GRAB] * (n_dabs // 4 + 1))[:n_dabs]
    return sg.stroke_config(tools, centres, radii, strengths, smooth_iterations=4), tools


def smooth_stroke(sg, n_dabs=500, radius=0.05):
    """Plain Smooth(4) dabs of radius 0.05, s ~ U[0.2, 1.0] (seed 5), along a great
    circle (seed 6) on the noisy level-10 icosphere: the learned Smooth workload."""
    rng = np.random.default_rng(5)
    radii = np.full(n_dabs, radius, np.float32)
    strengths = rng.uniform(0.2, 1.0, n_dabs).astype(np.float32)
    centres = surface_centres(sg, sg.great_circle_centres(6, radii))
    return sg.stroke_config(np.full(n_dabs, sg.TOOL_SMOOTH), centres, radii, strengths, smooth_iterations=4)


def path_centres(sg, spec, n, radii, rng):
    """Dab centres along a seeded path on the mesh's noisy surface, arc step 0.1 r: a
    great circle on an icosphere, a ring at a seeded tube angle on a torus (major /
    minor radius 1.0 / 0.3).  Each point of the smooth surface is displaced along its
    normal by the mesh's own seeded noise field, as the mesh vertices were."""
    kind, seed, amplitude = spec
    if kind == "sphere":
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        e1 = np.cross(axis, [1.0, 0.0, 0.0] if abs(axis[0]) < 0.9 else [0.0, 1.0, 0.0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(axis, e1)
        t = np.cumsum(0.1 * radii) - 0.1 * radii[0] + rng.uniform(0, 2 * np.pi)
        pts = np.cos(t)[:, None] * e1 + np.sin(t)[:, None] * e2
        nrm = pts
    else:
        v0 = rng.uniform(0, 2 * np.pi)
        ring = 1.0 + 0.3 * np.cos(v0)
        u = np.cumsum(0.1 * radii / ring) + rng.uniform(0, 2 * np.pi)
        pts = np.stack([ring * np.cos(u), ring * np.sin(u), np.full(n, 0.3 * np.sin(v0))], -1)
        nrm = np.stack([np.cos(v0) * np.cos(u), np.cos(v0) * np.sin(u), np.full(n, np.sin(v0))], -1)
    return sg.noise_displace(pts.astype(np.float32), nrm.astype(np.float32), seed, amplitude)


def batch_specs(n, seed, vmin, vmax, log_uniform=True, levels=(4, 8)):
    """The mesh list of a batch without building the meshes: per mesh ("sphere",
    level) or ("torus", (nu, nv)), vertex counts drawn from [vmin, vmax] (log-uniform
    or uniform), alternating between the nearest icosphere level (within `levels`) and
    a torus with n_u / n_v = 10 / 3 (near-equilateral triangles)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        V = float(np.exp(rng.uniform(np.log(vmin), np.log(vmax)))) if log_uniform else float(rng.uniform(vmin, vmax))
        if i % 2 == 0:
            level = int(np.clip(np.round(np.log(V / 10.0) / np.log(4.0)), levels[0], levels[1]))
            out.append(("sphere", level))
        else:
            nv = max(3, int(round(np.sqrt(V * 3.0 / 10.0))))
            nu = max(3, int(round(V / nv)))
            out.append(("torus", (nu, nv)))
    return out


def batch_mesh(sg, shape, i, amplitude):
    """Mesh i of a batch (noise seed 1000 + i) and its path spec."""
    kind, arg = shape
    if kind == "sphere":
        return sg.noisy_icosphere(arg, seed=1000 + i, amplitude=amplitude), ("sphere", 1000 + i, amplitude)
    return sg.noisy_torus(arg[0], arg[1], seed=1000 + i, amplitude=amplitude), ("torus", 1000 + i, amplitude)


def batch_meshes(sg, n, seed, vmin, vmax, amplitude, log_uniform=True, levels=(4, 8)):
    """n seeded meshes (see `batch_specs`): returns (meshes, path specs)."""
    meshes, specs = [], []
    for i, shape in enumerate(batch_specs(n, seed, vmin, vmax, log_uniform, levels)):
        m, s = batch_mesh(sg, shape, i, amplitude)
        meshes.append(m)
        specs.append(s)
    return meshes, specs


def batch_table(sg, meshes, specs, steps, seed, tools, rmin, rmax, learned_smooth=False):
    """table[step][mesh]: mesh m's dab `step` cycles `tools` (offset by m), r ~ U[rmin,
    rmax], s ~ U[0.2, 1.0], Smooth with k = 4, along the mesh's path."""
    rng = np.random.default_rng(seed)
    table = [[None] * len(specs) for _ in range(steps)]
    for m, spec in enumerate(specs):
        radii = rng.uniform(rmin, rmax, steps)
        strengths = rng.uniform(0.2, 1.0, steps)
        centres = path_centres(sg, spec, steps, radii, rng)
        for s in range(steps):
            tool = tools[(s + m) % len(tools)]
            flag = sg.SMOOTH_LEARNED if (learned_smooth and tool == sg.TOOL_SMOOTH) else 0
            table[s][m] = sg.make_dab(tool=tool, location=centres[s], radius=float(radii[s]),
                                      strength=float(strengths[s]), smooth_iterations=4, flag=flag)
    return table


# The many-mesh configurations: (meshes, mesh seed, V range, noise, log-uniform,
# icosphere levels, table seed, tools, radius range, steps)
def many_small(sg):
    """256 meshes of 5k-500k vertices (log-uniform, noise 0.01), 64 Draw / Smooth(4)
    steps, r ~ U[0.02, 0.1]: (shapes, amplitude, table arguments)."""
    return (batch_specs(256, 11, 5000, 500000, True, (4, 8)), 0.01,
            dict(steps=64, seed=12, tools=(sg.TOOL_DRAW, sg.TOOL_SMOOTH), rmin=0.02, rmax=0.1))


def many_large(sg):
    """64 meshes of 1M-4M vertices (uniform, icosphere level 9 or tori, noise 0.02),
    128 steps cycling Draw / Inflate / Smooth(4), r ~ U[0.02, 0.08]."""
    return (batch_specs(64, 21, 1e6, 4e6, False, (9, 9)), 0.02,
            dict(steps=128, seed=22, tools=(sg.TOOL_DRAW, sg.TOOL_INFLATE, sg.TOOL_SMOOTH), rmin=0.02, rmax=0.08))
