// Multires grids (PBVH_GRIDS): the host SubdivCCG container and the device state of a
// PBVH built over grids.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "sculpt_gpu.h"

// The SubdivCCG fields the grids path uses (kernel/intern/subdiv_ccg.c: layers
// :62-90, storage :104-140, faces :312-323, edge / vertex neighbourhoods :361-530).
// The coarse topology OpenSubdiv's refiner answers in the reference (face vertices,
// face edges, edge vertices, vertex edges) is held here directly.
struct SubdivCCG {
  int level = 0, grid_size = 0, grid_area = 0;
  int num_grids = 0, num_faces = 0, num_coarse_verts = 0, num_edges = 0;
  bool has_normal = false, has_mask = false;
  int normal_offset = -1, mask_offset = -1, elem_size = 12;  // bytes
  std::vector<unsigned char> storage;       // num_grids * grid_area * elem_size
  std::vector<CCGElem *> grids;
  std::vector<SubdivCCGFace> faces;
  std::vector<SubdivCCGFace *> grid_faces;  // handed out as void **gridfaces
  std::vector<DMFlagMat> flag_mats;
  std::vector<BLI_bitmap> hidden_words;
  std::vector<BLI_bitmap *> grid_hidden;
  std::vector<int> face_verts, face_edges;  // per grid (= per face corner)
  std::vector<int> edge_v0, edge_v1;        // edge vertices, as first used
  std::vector<int> edge_face_off;           // [num_edges + 1]: adjacent faces of each edge
  std::vector<SubdivCCGCoord> edge_coords;  // 2 * grid_size boundary coords per (edge, face)
  std::vector<int> vert_face_off;           // [num_coarse_verts + 1]
  std::vector<SubdivCCGCoord> vert_coords;  // the corner coord of each adjacent face
  std::vector<int> vert_edge_off, vert_edges;
  // a counter bumped by every change of the tables above (the PBVH re-uploads them)
  int topo_version = 0;
};

inline float *ccg_elem(const SubdivCCG *s, int grid, int x, int y) {
  return (float *)((unsigned char *)s->grids[grid] + ((size_t)y * s->grid_size + x) * s->elem_size);
}
inline float *ccg_co(const SubdivCCG *s, int grid, int x, int y) { return ccg_elem(s, grid, x, y); }
inline float *ccg_no(const SubdivCCG *s, int grid, int x, int y) {
  return (float *)((unsigned char *)ccg_elem(s, grid, x, y) + s->normal_offset);
}
inline float *ccg_mask(const SubdivCCG *s, int grid, int x, int y) {
  return (float *)((unsigned char *)ccg_elem(s, grid, x, y) + s->mask_offset);
}

// Neighbours of every element (include_duplicates = false) as a CSR over element ids
// grid * grid_area + y * grid_size + x, in KERNEL_subdiv_ccg_neighbor_coords_get order.
void sg_ccg_neighbor_csr(const SubdivCCG *s, std::vector<long long> &off, std::vector<int> &idx);

// Device state of a PBVH over grids (HBM).  Elements are numbered grid by grid in leaf
// (DFS, left->right) order, each grid row-major, so a leaf's elements are one range and
// an element id is grid_pos * grid_area + y * grid_size + x.
struct SgGridDev {
  int G = 0, gs = 0, ga = 0, L = 0, N = 0, H = 0, maxleafg = 0;
  long long E = 0;
  float *px = nullptr, *py = nullptr, *pz = nullptr;  // positions [E]
  float *qx = nullptr, *qy = nullptr, *qz = nullptr;  // shadow copy (== positions between dabs)
  float *nx = nullptr, *ny = nullptr, *nz = nullptr;  // normals [E]
  float *mask = nullptr;                              // [E] when the key has a mask
  unsigned char *hide = nullptr;                      // [E] when any element is hidden
  int *stamp = nullptr;                               // [E] serial of the last dab that affected it
  float *fade = nullptr;                              // [E] Smooth factor of the affected ones
  int *grid_pos = nullptr, *grid_old = nullptr;       // [G] grid -> position, position -> grid
  int *grid_leaf = nullptr;                           // [G] by position
  int *leaf_gbeg = nullptr;                           // [L + 1] grid positions of each leaf
  int *leaf_flags = nullptr;                          // [L] bit0 fully hidden
  int *leaf_mark = nullptr;                           // [L] serial of the last refit request
  float *leaf_bb = nullptr, *leaf_obb = nullptr;      // [L * 8]
  int *node_child = nullptr, *node_leaf = nullptr, *node_mark = nullptr;  // [N]
  float *node_bb = nullptr, *node_obb = nullptr;      // [N * 8] internal nodes
  int *lvl_off = nullptr, *lvl_node = nullptr;        // internal nodes by height 1..H
  long long *nb_off = nullptr;                        // [E + 1] neighbour CSR
  int *nb_idx = nullptr;
  // topology of the attached SubdivCCG
  int F = 0, Ed = 0, Vc = 0;
  int *face_gbeg = nullptr;   // [F + 1] grid ids (consecutive per face)
  int *grid_face = nullptr;   // [G] by grid id
  int *face_edge = nullptr, *face_vert = nullptr;  // [G] per face corner
  int *edge_foff = nullptr;   // [Ed + 1]
  int *edge_el = nullptr;     // 2 * gs element ids per (edge, face)
  int *vert_foff = nullptr;   // [Vc + 1]
  int *vert_el = nullptr;     // element id per (vertex, face)
  int *face_stamp = nullptr, *edge_stamp = nullptr, *vert_stamp = nullptr;
  int *gathered = nullptr;    // [L] the dab's gathered leaves, then the leaves its averages moved
  int *dfaces = nullptr, *ditems = nullptr;  // dirty faces [F], dirty edges then vertices [Ed + Vc]
  long long *ctr = nullptr;   // live counters
  long long *ctr_last = nullptr;  // the last dab's counters
  float *curve = nullptr;     // CUSTOM falloff table
  int *hlist = nullptr;       // [L] leaf lists from the host (drop-in updates)
  void *snap = nullptr;       // checkpoint of the mutable state (sg_pbvh_snapshot_save)
  int *dser = nullptr;        // [4] device serials: the running operation's, the last dab's; blocks done
  int *leaf_dirty = nullptr, *node_dirty = nullptr;  // refit since the last internal-node flush
  float *parts = nullptr;     // [L * 8 * 8] boxes of the parts of the leaves being refit
  bool ctr_unclean = false;   // a dab's launches failed midway: clear the counters first
  const SubdivCCG *topo_of = nullptr;
  int topo_version = -1;
};

// counters (ctr / ctr_last)
enum {
  SGG_NGATHER = 0, SGG_NAFF = 1, SGG_AREA_X = 2, SGG_AREA_Y = 3, SGG_AREA_Z = 4, SGG_NAREA = 5,
  SGG_NFACES = 6, SGG_NITEMS = 7, SGG_NMOVED = 8, SGG_NLEAVES = 9, SGG_CTR_COUNT = 12
};

struct PBVH;
struct SgDabK;
int sg_grids_build_device(PBVH *p);
int sg_grids_attach_topology(PBVH *p, const SubdivCCG *s);
void sg_grids_free_device(PBVH *p);
int sg_grids_enqueue_dab(PBVH *p, const SgDabK &k, const float *curve_dev, long long *log = nullptr,
                         int log_index = 0);
int sg_grids_snapshot(PBVH *p, int restore);
int sg_grids_enqueue_update_normals(PBVH *p, const int *leaves_dev, int n);
int sg_grids_enqueue_refit(PBVH *p, const int *leaves_dev, int n, int set_orig);
int sg_grids_enqueue_flush(PBVH *p);  // internal nodes above the leaves refit since the last flush
int sg_grids_download(PBVH *p, float *co, float *no, float *mask);  // element order of the grids (by grid id)
int sg_grids_upload(PBVH *p, const float *co, const float *no, const float *mask);
int sg_grids_read_boxes(PBVH *p, float *leaf_bb, float *node_bb);
int sg_grids_read_stamps(PBVH *p, int *stamp);
// grid branches of the entry points in api.cpp (grids_api.cpp)
struct SgDabParams;
int sg_grids_dab_apply(PBVH *p, const SgDabParams *dab);
int sg_grids_fetch_last(PBVH *p);
int sg_grids_last_affected(PBVH *p, int *out, int cap);
int sg_grids_last_gathered(PBVH *p, int *out, int cap);
int sg_grids_sync_bbs(PBVH *p);
void sg_grids_update_bounds(PBVH *p, int flag);
void sg_grids_update_normals(PBVH *p, SubdivCCG *s);
