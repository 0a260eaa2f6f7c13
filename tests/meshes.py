"""Small hand-made meshes for known-answer tests (numpy)."""
import numpy as np


def grid(n, spacing=0.125, z=0.0):
    """(n x n) vertex plane grid centred on the origin, quads split along the same
    diagonal; coordinates are exact dyadic rationals."""
    h = n // 2
    ij = np.stack(np.meshgrid(np.arange(n) - h, np.arange(n) - h, indexing="ij"), -1).reshape(-1, 2)
    co = np.zeros((n * n, 3), np.float32)
    co[:, 0] = ij[:, 0] * spacing
    co[:, 1] = ij[:, 1] * spacing
    co[:, 2] = z
    tris = []
    for i in range(n - 1):
        for j in range(n - 1):
            a, b, c, d = i * n + j, (i + 1) * n + j, (i + 1) * n + j + 1, i * n + j + 1
            tris += [(a, b, c), (a, c, d)]
    return co, np.array(tris, np.int32)


def fan(n, rho=1.0, centre=(0.1, 0.0, 0.0)):
    """A regular n-gon of radius rho around a centre vertex (index 0)."""
    ang = 2 * np.pi * np.arange(n) / n
    co = np.zeros((n + 1, 3), np.float32)
    co[0] = centre
    co[1:, 0] = rho * np.cos(ang)
    co[1:, 1] = rho * np.sin(ang)
    tris = np.array([(0, 1 + k, 1 + (k + 1) % n) for k in range(n)], np.int32)
    return co, tris


def edges_of(tris):
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    e.sort(axis=1)
    return np.unique(e, axis=0)
