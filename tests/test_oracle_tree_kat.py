"""Known-answer trees: the PBVH build (pbvh.c:152-170 partition_indices, :172-214
partition_indices_material, :223-244 map_insert_vert, :256-320 build_mesh_leaf_node,
:454-507 build_sub, :534-596 BKE_pbvh_build_mesh) worked by hand on a seven-triangle
strip, checked against the oracle and against the product's host build.

The strip (z = 0): bottom vertices B0..B4 = ids 0..4 at x = 0, 1, 2, 3, 6 (y = 0), top
vertices T0..T3 = ids 5..8 at x = 0.5, 1.5, 2.5, 3.5 (y = 1).  Triangle ids in a
scrambled order, with the centroid of each triangle's box ((min + max) * 0.5, so a
triangle's centroid is its box centre, not its vertex mean):

  tri  corners        x range      centroid x
  0    (6, 2, 7)      [1.5, 2.5]   2.0
  1    (3, 4, 8)      [3.0, 6.0]   4.5     (B4 at x = 6: the box is not centred on the strip)
  2    (0, 1, 5)      [0.0, 1.0]   0.5
  3    (7, 3, 8)      [2.5, 3.5]   3.0
  4    (2, 3, 7)      [2.0, 3.0]   2.5
  5    (5, 1, 6)      [0.5, 1.5]   1.0
  6    (1, 2, 6)      [1.0, 2.0]   1.5

Every centroid has y = 0.5, z = 0, so the widest centroid axis is x everywhere.  The
root's centroid box is [0.5, 4.5]: the split value is 2.5, which triangle 4's centroid
equals (the partition's tie); the root's vertex box [0, 6] would give 3.0 instead.  A
`<=` in either scan, a split at the vertex box's midpoint, or a wrong axis each move a
triangle to the other side, so each changes the expected trees below.

Hand trace of the root split (prim_indices [0..6], mid 2.5): i stops at 1 (4.5), j at 6
(1.5): swap -> [0,6,2,3,4,5,1]; i stops at 3 (3.0), j at 5 (1.0): swap ->
[0,6,2,5,4,3,1]; i stops at 4 (2.5 is not < 2.5), j passes 3 (3.0) and stops at 4 (2.5
is not > 2.5): i >= j, the right side starts at 4.  Left [0,6,2,5] (centroids 2, 1.5,
0.5, 1: box [0.5, 2], mid 1.25) splits into [5,2] | [6,0]; right [4,3,1] (box [2.5,
4.5], mid 3.5) into [4,3] | [1].
"""
import numpy as np
import pytest

B = [(0, 0), (1, 0), (2, 0), (3, 0), (6, 0)]
T = [(0.5, 1), (1.5, 1), (2.5, 1), (3.5, 1)]
TRIS = np.array([(6, 2, 7), (3, 4, 8), (0, 1, 5), (7, 3, 8), (2, 3, 7), (5, 1, 6), (1, 2, 6)], np.int32)


def strip():
    co = np.zeros((9, 3), np.float32)
    for i, (x, y) in enumerate(B + T):
        co[i] = (x, y, 0.0)
    return co, TRIS


# Expected trees: node -> ("leaf", prims, vert_indices, uniq) or ("inner", children_offset).
# Node ids follow build_sub: a node's two children get the next two free ids when the
# node is split, and the left subtree is built completely before the right one.
KAT = {
    # leaf limit 2
    2: {
        "prim_indices": [5, 2, 6, 0, 4, 3, 1],
        "nodes": {0: ("inner", 1), 1: ("inner", 3), 2: ("inner", 5),
                  # leaves in build order 3, 4, 5, 6; a vertex is uniq in the first leaf
                  # that uses it, and both lists keep first-insertion order
                  3: ("leaf", [5, 2], [5, 1, 6, 0], 4),
                  4: ("leaf", [6, 0], [2, 7, 1, 6], 2),
                  5: ("leaf", [4, 3], [3, 8, 2, 7], 2),
                  6: ("leaf", [1], [4, 3, 8], 1)},
        "dfs": [3, 4, 5, 6],
    },
    # leaf limit 3: the root's right child (3 triangles) is a leaf
    3: {
        "prim_indices": [5, 2, 6, 0, 4, 3, 1],
        "nodes": {0: ("inner", 1), 1: ("inner", 3),
                  3: ("leaf", [5, 2], [5, 1, 6, 0], 4),
                  4: ("leaf", [6, 0], [2, 7, 1, 6], 2),
                  2: ("leaf", [4, 3, 1], [3, 8, 4, 2, 7], 3)},
        "dfs": [3, 4, 2],
    },
    # leaf limit 1: [5,2] (centroids 1, 0.5; mid 0.75) -> swap -> [2] | [5];
    # [6,0] (1.5, 2; mid 1.75) -> [6] | [0]; [4,3] (2.5, 3; mid 2.75) -> [4] | [3]
    1: {
        "prim_indices": [2, 5, 6, 0, 4, 3, 1],
        "nodes": {0: ("inner", 1), 1: ("inner", 3), 3: ("inner", 5), 4: ("inner", 7), 2: ("inner", 9),
                  9: ("inner", 11),
                  5: ("leaf", [2], [0, 1, 5], 3),
                  6: ("leaf", [5], [6, 5, 1], 1),
                  7: ("leaf", [6], [2, 1, 6], 1),
                  8: ("leaf", [0], [7, 6, 2], 1),
                  11: ("leaf", [4], [3, 2, 7], 1),
                  12: ("leaf", [3], [8, 7, 3], 1),
                  10: ("leaf", [1], [4, 3, 8], 1)},
        "dfs": [5, 6, 7, 8, 11, 12, 10],
    },
}

# Material split (pbvh.c:172-214): leaf limit 100 holds all 7, but triangles 1 and 4
# carry another material.  First = triangle 0's material; i stops at 1 (mat 1), j at 6:
# swap -> [0,6,2,3,4,5,1]; i stops at 4, j passes 6 (mat 1), stops at 5: swap ->
# [0,6,2,3,5,4,1]; i stops at 5, j passes 5 and stops at 4: right side starts at 5.
MAT = np.array([0, 1, 0, 0, 1, 0, 0], np.int32)
KAT_MAT = {
    "prim_indices": [0, 6, 2, 3, 5, 4, 1],
    "nodes": {0: ("inner", 1),
              1: ("leaf", [0, 6, 2, 3, 5], [6, 2, 7, 1, 0, 5, 3, 8], 8),
              2: ("leaf", [4, 1], [4, 2, 3, 7, 8], 1)},
    "dfs": [1, 2],
}


def _check_oracle(o, want):
    assert o.totnode() == len(want["nodes"])
    assert o.prim_indices().tolist() == want["prim_indices"]
    for node, spec in want["nodes"].items():
        info = o.node_info(node)
        if spec[0] == "inner":
            assert not (info[0] & 1) and info[1] == spec[1], (node, info)
        else:
            _, prims, verts, uniq = spec
            assert info[0] & 1, (node, info)
            off, tot = info[2], info[3]
            assert o.prim_indices()[off:off + tot].tolist() == prims, node
            assert o.node_verts(node).tolist() == verts, node
            assert info[4] == uniq and info[5] == len(verts) - uniq, node
    assert o.leaves_dfs().tolist() == want["dfs"]


def _check_host(sg, co, tris, leaf, want, mat=None):
    p = sg.PBVH.__new__(sg.PBVH)
    try:
        sg.PBVH.__init__(p, co, tris, leaf_limit=leaf, mat=mat)
    except RuntimeError:
        pass
    assert p.tree_info()["nodes"] == len(want["nodes"])
    assert p.prim_indices().tolist() == want["prim_indices"]
    for node, spec in want["nodes"].items():
        if spec[0] == "leaf":
            verts, uniq = p.node_verts(node)
            assert verts.tolist() == spec[2] and uniq == spec[3], node
    assert p.search_gather(None) == want["dfs"]
    p.close()


@pytest.mark.parametrize("leaf", [1, 2, 3])
def test_strip_tree_oracle(po, leaf):
    co, tris = strip()
    o = po.Oracle(co, tris, leaf_limit=leaf)
    _check_oracle(o, KAT[leaf])
    o.close()


@pytest.mark.parametrize("leaf", [1, 2, 3])
def test_strip_tree_host_build(sg, leaf):
    co, tris = strip()
    _check_host(sg, co, tris, leaf, KAT[leaf])


def test_strip_material_split(sg, po):
    co, tris = strip()
    o = po.Oracle(co, tris, leaf_limit=100, mat=MAT)
    _check_oracle(o, KAT_MAT)
    o.close()
    _check_host(sg, co, tris, 100, KAT_MAT, mat=MAT)


def test_strip_leaf_boxes_from_triangle_boxes(po):
    """update_vb (pbvh.c:322-329): a node's box at build is the union of its triangles'
    boxes; leaf 3 of the leaf-limit-2 tree holds triangles 5 and 2: x [0, 1.5], y [0, 1]."""
    co, tris = strip()
    o = po.Oracle(co, tris, leaf_limit=2)
    vb, ob = o.node_bb(3)
    assert vb.tolist() == [0.0, 0.0, 0.0, 1.5, 1.0, 0.0] and ob.tolist() == vb.tolist()
    vb, _ = o.node_bb(0)
    assert vb.tolist() == [0.0, 0.0, 0.0, 6.0, 1.0, 0.0]
    o.close()
