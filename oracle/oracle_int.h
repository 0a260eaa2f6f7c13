/*
 * Internal types of the CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Shared by oracle.c (build, search, brush, normals, refit) and ray.c (ray cast).
 */
#pragma once
#include <stdint.h>

/* BB / BBC: pbvh_intern.h:3-11. */
typedef struct {
  float bmin[3], bmax[3];
} OBB;
typedef struct {
  float bmin[3], bmax[3], bcentroid[3];
} OBBC;

/* ------------------------------------------------------------------------- */
/* Tree (PBVHNode / PBVH: pbvh_intern.h:15-90, :98-162, mesh fields only).    */
/* ------------------------------------------------------------------------- */

typedef struct {
  OBB vb, orig_vb;
  int children_offset;
  int prim_off, totprim;
  int *vert_indices;
  int uniq_verts, face_verts;
  int (*face_vert_indices)[3];
  int flag;
} ONode;

struct OrcPBVH {
  int totvert, tottri, leaf_limit, respect_hide, with_fp64;
  float *co, *no, *mask;
  double *co64, *no64;
  int *tris, *tri_mat;
  /* polygons (NULL: each tri is its own polygon): looptri -> polygon, each polygon's
   * corner range and corner vertex ids (MPoly / MLoop, types_meshdata.h:50-74) */
  int *tri_poly, *poly_off, *poly_verts;
  int npoly;
  uint8_t *vhide;
  ONode *nodes;
  int totnode, node_mem;
  int *prim_indices;
  uint8_t *vert_bitmap;
  /* vertex -> incident tris (ascending tri id) and vertex -> edge neighbours
   * (ascending vertex id): the 1-ring sources (mesh_mapping.c:182-295). */
  int *vt_off, *vt_idx, *vv_off, *vv_idx;
  /* build scratch */
  int *map_local, *map_touched;
  int64_t stats[8];
};

int64_t orc_update_normals_bounds(OrcPBVH *o);

/* Spec S5 bstrength (oracle.c), shared with grids.c. */
struct OrcDab;
float orc_bstrength(const struct OrcDab *dab);
