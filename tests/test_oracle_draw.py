"""Known-answer tests for the oracle's draw records (oracle/draw.c): the reference's
vertex layout, the normal and mask conversions, visibility and flat/smooth normals,
against closed forms and brute force.  The face-set overlay colour hashes with
BLI_hash_int_01, whose header (BLI_hash.h) is absent from the reference: it is Bob
Jenkins' lookup3 hashword of (k, 0) with initial value 13, so the oracle's lookup3 mix is
pinned by lookup3's own published hash vectors, and the hash by its identity with
lookup3's byte hash of the same eight bytes."""
import struct

import numpy as np

F32 = np.float32


def _quad_grid(n, spacing=0.25):
    """(n x n) vertex plane z = 0 with (n-1)^2 quads as polygons."""
    ij = np.stack(np.meshgrid(np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 2)
    co = np.zeros((n * n, 3), F32)
    co[:, :2] = ij * spacing
    pv = []
    for i in range(n - 1):
        for j in range(n - 1):
            pv += [i * n + j, (i + 1) * n + j, (i + 1) * n + j + 1, i * n + j + 1]
    off = np.arange(0, len(pv) + 1, 4)
    return co, off, np.array(pv, np.int32)


def _setup(sg, po, co, off, pv, smooth, leaf=8, hidden=None):
    tri_poly, loops = sg.fan_looptris(off)
    tris = pv[loops]
    o = po.Oracle(co, tris, leaf_limit=leaf, hidden=None if hidden is None else hidden.astype(np.uint8))
    return o, tri_poly.astype(np.int32), tris


def _pack(attrs):
    """VertexFormat_pack's rule: each attribute aligned to its component size, 3-component
    8/16-bit attributes to 4x that; the stride padded to the first attribute's alignment."""
    offs, off = [], 0
    aligns = []
    for comp, n in attrs:
        a = 4 * comp if (n == 3 and comp <= 2) else comp
        aligns.append(a)
        off += (-off) % a
        offs.append(off)
        off += comp * n
    return offs, off + (-off) % aligns[0]


def test_vertex_layout_and_conversions(sg, po):
    offs, stride = _pack([(4, 3), (2, 3), (1, 1), (2, 4), (1, 3)])
    assert offs == [0, 16, 22, 24, 32] and stride == 36
    co, off, pv = _quad_grid(5)
    P = len(off) - 1
    smooth = np.zeros(P, np.uint8)
    o, tri_poly, tris = _setup(sg, po, co, off, pv, smooth)
    fs = np.ones(P, np.int32)
    mask = np.linspace(0, 1, len(co)).astype(F32)
    total = 0
    for leaf in o.leaves_dfs():
        recs, sm, mt = o.draw_leaf(leaf, tri_poly, off, pv, smooth, np.zeros(P, np.int32), fs, mask=mask)
        assert sm == 0 and mt == 0
        total += len(recs)
        pos = recs[:, 0:12].copy().view(F32)
        nor = recs[:, 16:22].copy().view(np.int16)
        # flat shading of a planar grid: every face normal is +z, 32767 after the cast
        assert np.all(nor == np.array([0, 0, 32767], np.int16))
        # positions are vertex coordinates; the mask byte is the face average * 255, truncated
        idx = [int(np.flatnonzero((co == p).all(1))[0]) for p in pos]
        for k in range(0, len(idx), 3):
            m = (mask[idx[k]] + mask[idx[k + 1]] + mask[idx[k + 2]]) / F32(3.0)
            assert recs[k, 22] == int(F32(m) * F32(255))
        assert np.all(recs[:, 32:35] == 255) and np.all(recs[:, 12:16] == 0) and np.all(recs[:, 23:32] == 0)
    assert total == 3 * len(tris)


def test_smooth_normals_and_visibility(sg, po):
    co, tris0 = sg.icosphere(2)
    # tri polygons: poly t = looptri t
    P = len(tris0)
    off = np.arange(0, 3 * P + 1, 3)
    pv = tris0.reshape(-1).astype(np.int32)
    smooth = np.ones(P, np.uint8)
    hidden = co[:, 2] > 0.7
    o, tri_poly, tris = _setup(sg, po, co, off, pv, smooth, hidden=hidden)
    rng = np.random.default_rng(2)
    fs = rng.choice([1, 1, 1, 2, 0, -4], P).astype(np.int32)
    no = o.normals()
    total = 0
    for leaf in o.leaves_dfs():
        recs, sm, _ = o.draw_leaf(leaf, tri_poly, off, pv, smooth, np.zeros(P, np.int32), fs)
        assert sm == 1
        total += len(recs)
        for r in recs:
            p = r[0:12].copy().view(F32)
            v = int(np.flatnonzero((co == p).all(1))[0])
            want = (no[v] * F32(32767.0)).astype(np.int16)  # truncation toward zero
            assert np.array_equal(r[16:22].copy().view(np.int16), want)
            assert r[22] == 0  # no mask layer
    vis = [t for t in range(P) if fs[t] > 0 and not hidden[tris[t]].any()]
    assert total == 3 * len(vis)


def test_face_set_colours(sg, po):
    cols = {po.face_set_color(f, 3) for f in range(2, 40)}
    assert len(cols) > 30  # distinct ids give distinct colours
    assert po.face_set_color(5, 3) == po.face_set_color(-5, 3)  # abs
    assert po.face_set_color(5, 3) == po.face_set_color(5, 3)
    co, off, pv = _quad_grid(4)
    P = len(off) - 1
    o, tri_poly, _ = _setup(sg, po, co, off, pv, np.zeros(P, np.uint8))
    fs = np.full(P, 2, np.int32)
    for leaf in o.leaves_dfs():
        recs, _, _ = o.draw_leaf(leaf, tri_poly, off, pv, np.zeros(P, np.uint8), np.zeros(P, np.int32), fs,
                                 seed=3, default=2)
        assert np.all(recs[:, 32:35] == 255)  # the default face set draws white
        recs, _, _ = o.draw_leaf(leaf, tri_poly, off, pv, np.zeros(P, np.uint8), np.zeros(P, np.int32), fs,
                                 seed=3, default=1)
        assert all(bytes(r[32:35]) == po.face_set_color(2, 3) for r in recs)


def test_quad_and_ngon_normals(sg, po):
    """Flat normals of a tilted quad and a regular hexagon equal the plane normal."""
    t = np.array([0.0, 0.6, 0.8])
    u = np.cross(t, [1.0, 0, 0])
    u /= np.linalg.norm(u)
    w = np.cross(t, u)
    ang = np.arange(6) * np.pi / 3
    hexv = (np.cos(ang)[:, None] * u + np.sin(ang)[:, None] * w).astype(F32)
    quad = (np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]])[:, :1] * u + np.array([[-1], [-1], [1], [1]]) * w + 3 * t)
    co = np.concatenate([hexv, quad.astype(F32)]).astype(F32)
    off = np.array([0, 6, 10])
    pv = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9], np.int32)
    o, tri_poly, _ = _setup(sg, po, co, off, pv, np.zeros(2, np.uint8), leaf=100)
    recs, _, _ = o.draw_leaf(o.leaves_dfs()[0], tri_poly, off, pv, np.zeros(2, np.uint8), np.zeros(2, np.int32),
                             np.ones(2, np.int32))
    nor = recs[:, 16:22].copy().view(np.int16).astype(np.float64) / 32767
    assert np.abs(np.abs(nor @ t) - 1).max() < 1e-3


def _rot(x, k):
    return ((x << k) | (x >> (32 - k))) & 0xFFFFFFFF


def _mix(a, b, c):
    """lookup3.c `mix` (the block mix; the oracle holds only `final`)."""
    M = 0xFFFFFFFF
    a = (a - c) & M; a ^= _rot(c, 4); c = (c + b) & M
    b = (b - a) & M; b ^= _rot(a, 6); a = (a + c) & M
    c = (c - b) & M; c ^= _rot(b, 8); b = (b + a) & M
    a = (a - c) & M; a ^= _rot(c, 16); c = (c + b) & M
    b = (b - a) & M; b ^= _rot(a, 19); a = (a + c) & M
    c = (c - b) & M; c ^= _rot(b, 4); b = (b + a) & M
    return a, b, c


def _hashlittle2(po, data, pc, pb):
    """lookup3.c hashlittle2 (little-endian byte hash) with the oracle's final mix."""
    M = 0xFFFFFFFF
    n = len(data)
    a = b = c = (0xDEADBEEF + n + pc) & M
    c = (c + pb) & M
    while n > 12:
        w = struct.unpack_from("<3I", data, len(data) - n)
        a, b, c = (a + w[0]) & M, (b + w[1]) & M, (c + w[2]) & M
        a, b, c = _mix(a, b, c)
        n -= 12
    if n == 0:
        return c, b
    tail = data[len(data) - n:] + bytes(12 - n)  # the masked tail words read zero bytes
    w = struct.unpack("<3I", tail)
    a, b, c = (a + w[0]) & M, (b + w[1]) & M, (c + w[2]) & M
    a, b, c = po.lookup3_final(a, b, c)
    return c, b


def test_lookup3_published_vectors(po):
    """The driver vectors printed by lookup3.c (driver5): the 30-byte string runs two
    block mixes and the oracle's final mix; the empty-key cases pin the initial value."""
    s = b"Four score and seven years ago"
    assert _hashlittle2(po, b"", 0, 0) == (0xDEADBEEF, 0xDEADBEEF)
    assert _hashlittle2(po, b"", 0, 0xDEADBEEF) == (0xBD5B7DDE, 0xDEADBEEF)
    assert _hashlittle2(po, b"", 0xDEADBEEF, 0xDEADBEEF) == (0x9C093CCD, 0xBD5B7DDE)
    assert _hashlittle2(po, s, 0, 0) == (0x17770551, 0xCE7226E6)
    assert _hashlittle2(po, s, 0, 1) == (0xE3607CAE, 0xBD371DE4)
    assert _hashlittle2(po, s, 1, 0) == (0xCD628161, 0x6CBEA4B3)


def test_hash_int_01_is_lookup3_of_two_words(po):
    """BLI_hash_int_2d(k, 0) = hashword({k, 0}, 2, 13) = hashlittle of the same 8 bytes
    with initial value 13 (both start from 0xdeadbeef + 8 + 13 and add the two words);
    BLI_hash_int_01 scales it by 1 / 0xFFFFFFFF in float32."""
    scale = np.float32(1.0) / np.float32(0xFFFFFFFF)
    for k in [0, 1, 2, 3, 7, 42, 1000, 65535, 123456789, 0x7FFFFFFF, 0xFFFFFFFF]:
        c, _ = _hashlittle2(po, struct.pack("<2I", k, 0), 13, 0)
        assert np.float32(po.hash_int_01(k)) == np.float32(c) * scale, k
