/*
 * CPU oracle for the sculpt brush hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product.  Only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it.  The product path
 * (libsculpt_gpu.so) never links, loads or calls anything here.
 *
 * It restates, in plain C (FP32 with no FMA contraction, like the reference's
 * SSE2-only build, CMakeLists.txt:703-768), the reference's present PBVH code
 * and the project's written spec for the brush code the reference lacks:
 *
 *   build      kernel/intern/pbvh.c:152-214, :256-320, :322-329, :391-407,
 *              :454-532, :534-596
 *   search     kernel/intern/pbvh.c:704-787, :818-849
 *   normals    kernel/intern/pbvh.c:994-1118 (summed in vertex->tri CSR order
 *              instead of float atomics, see orc_dab), math_geom.cc:31-46,
 *              math_vector_inline.c:1165-1180
 *   initial    kernel/intern/mesh_normals.c:225-299 (angle weighted)
 *   refit      kernel/intern/pbvh.c:108-128, :1206-1242, :1369-1421
 *   brushes    project spec S1-S11 (the reference's sculpt.c/brush.c are absent)
 *   grids      kernel/intern/subdiv_ccg.c and the PBVH_GRIDS parts of pbvh.c
 *              (grids.c, cited there)
 *
 * Parity status of each function is stated in its own comment in the .c files.
 * An FP64 replay runs beside the FP32 state (same affected sets, every other
 * arithmetic step in double) so error against exact-ish arithmetic can be
 * reported.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Node flags: names follow the reference's PBVHNodeFlags uses (pbvh.c:114, :1003,
 * :1215, :1725); values are this project's (the enum is absent upstream). */
enum {
  ORC_Leaf = 1 << 0,
  ORC_UpdateNormals = 1 << 1,
  ORC_UpdateBB = 1 << 2,
  ORC_UpdateOriginalBB = 1 << 3,
  ORC_UpdateDrawBuffers = 1 << 4,
  ORC_UpdateRedraw = 1 << 5,
  ORC_FullyHidden = 1 << 10,
  ORC_UpdateTopology = 1 << 13,
};

/* Tools (types_brush_enums.h:410-415) and falloff presets (:176-187). */
enum { ORC_TOOL_DRAW = 1, ORC_TOOL_SMOOTH = 2, ORC_TOOL_INFLATE = 4, ORC_TOOL_GRAB = 5, ORC_TOOL_ELASTIC = 21 };
enum {
  ORC_CURVE_CUSTOM = 0, ORC_CURVE_SMOOTH = 1, ORC_CURVE_SPHERE = 2, ORC_CURVE_ROOT = 3,
  ORC_CURVE_SHARP = 4, ORC_CURVE_LIN = 5, ORC_CURVE_POW4 = 6, ORC_CURVE_INVSQUARE = 7,
  ORC_CURVE_CONSTANT = 8, ORC_CURVE_SMOOTHER = 9,
};
#define ORC_BRUSH_DIR_IN (1 << 9) /* types_brush_enums.h:347 */
#define ORC_CM_TABLE 256          /* types_color.h:12 */

typedef struct OrcDab {
  int tool;
  int falloff;
  int flag;          /* ORC_BRUSH_DIR_IN */
  int smooth_iters;  /* Smooth(k) */
  float center[3];
  float radius;
  float strength;    /* Brush.alpha */
  float hardness;
  float normal_radius_factor;
  float grab_delta[3];
  const float *curve_table; /* ORC_CM_TABLE+1 samples of y over x in [0,1] (CUSTOM) */
  int elastic_type;         /* eBrushElasticDeformType */
  float poisson_ratio;      /* elastic_deform_volume_preservation */
  float surface_normal[3];  /* the stroke's sculpt normal (Scale, Twist) */
} OrcDab;

typedef struct OrcPBVH OrcPBVH;

/* co: totvert*3 floats (MVert.co order); tris: tottri*3 vertex ids (looptri corners
 * resolved through MLoop.v).  tri_mat (may be NULL): per-tri material key
 * (mat_nr | smooth<<16) used by the material split.  vhide (may be NULL): ME_HIDE. */
OrcPBVH *orc_build(const float *co, int totvert, const int *tris, int tottri, const int *tri_mat,
                   const uint8_t *vhide, int leaf_limit, int respect_hide, int with_fp64);
void orc_free(OrcPBVH *o);

int orc_totnode(const OrcPBVH *o);
int orc_totleaf(const OrcPBVH *o);
/* Per node: flag, children_offset, prim offset, totprim, uniq_verts, face_verts. */
void orc_node_info(const OrcPBVH *o, int node, int out[6]);
void orc_node_bb(const OrcPBVH *o, int node, float vb[6], float orig_vb[6]);
/* Copies uniq+face verts of a leaf; returns count. */
int orc_node_verts(const OrcPBVH *o, int node, int *out, int cap);
/* prim_indices array (tottri). */
void orc_prim_indices(const OrcPBVH *o, int *out);
/* Leaves in DFS left->right order (node ids). */
int orc_leaves_dfs(const OrcPBVH *o, int *out, int cap);

void orc_get_positions(const OrcPBVH *o, float *out);
void orc_get_normals(const OrcPBVH *o, float *out);
void orc_get_positions64(const OrcPBVH *o, double *out);
void orc_get_normals64(const OrcPBVH *o, double *out);
void orc_set_mask(OrcPBVH *o, const float *mask); /* per-vertex paint mask or NULL */
/* Replace every position and refit all node bounds (orig_vb follows vb), e.g. to
 * continue from another engine's state. */
void orc_set_positions(OrcPBVH *o, const float *co);
/* Replace the vertex normals (e.g. to start two runs from identical state). */
void orc_set_normals(OrcPBVH *o, const float *no);
/* Grab from the original coordinates orco (totvert*3): tested and displaced from orco,
 * every visible uniq vertex of a gathered leaf written (orco + displacement, zero
 * outside the sphere) and returned as affected (ascending). */
int orc_dab_grab_orig(OrcPBVH *o, const OrcDab *dab, const float *orco, int *out_affected, int cap);
/* Elastic deform (tool 21) from the original coordinates orco (totvert*3): every leaf,
 * each visible uniq vertex = orco + Kelvinlet displacement; written vertices returned
 * (ascending). */
int orc_dab_elastic(OrcPBVH *o, const OrcDab *dab, const float *orco, int *out_affected, int cap);
/* Proxy combine over the given nodes (node order): uniq vertex i of node j becomes
 * (orco or co)[v] + its proxies in order; prox packed [nprox[j]][uniq][3] node after
 * node; then normals of those vertices and the bounds update. */
void orc_combine_proxies(OrcPBVH *o, const int *nodes, int n, const int *nprox, const float *prox,
                         const float *orco);

/* Sphere search (spec S1, pbvh.c:818-849 traversal).  Returns leaf node ids in
 * traversal (left->right) order. */
int orc_search_sphere(const OrcPBVH *o, const float center[3], float radius, int use_orig,
                      int *out, int cap);

/* One dab: search, affected set, displacement, normals, refit.  Writes the affected
 * vertex ids (ascending) when out != NULL.  Returns |A|. */
int orc_dab(OrcPBVH *o, const OrcDab *dab, int *out_affected, int cap);

/* Falloff value as the brush computes it (spec S3), for known-answer tests. */
float orc_falloff(const OrcDab *dab, float d2);

/* Reference-style whole-mesh pieces, exposed for tests. */
void orc_tri_normal(const float v1[3], const float v2[3], const float v3[3], float out[3]);
void orc_calc_initial_normals(const float *co, int totvert, const int *tris, int tottri,
                              float *out);
void orc_calc_initial_normals_poly(const float *co, int totvert, const int *poly_off, const int *poly_verts,
                                   int npoly, float *out);
void orc_poly_normal(const float *co, const int *corners, int n, float out[3]);
/* The mesh's polygons: looptri -> polygon, corner offsets [npoly + 1], corner vertex
 * ids.  Normals then follow the polygon rules; the initial normals are recomputed. */
void orc_set_polys(OrcPBVH *o, const int *tri_poly, const int *poly_off, const int *poly_verts, int npoly);

/* The drop-in update API (host edits): set a vertex, the mark helpers, bounds with a
 * flag, normals of the marked vertices of UpdateNormals leaves, node flags, and the
 * whole-mesh coordinate replace. */
void orc_set_vert(OrcPBVH *o, int v, const float co[3]);
void orc_vert_mark_update(OrcPBVH *o, int v);
void orc_node_mark_update(OrcPBVH *o, int node);
void orc_update_bounds(OrcPBVH *o, int flag);
void orc_update_normals(OrcPBVH *o);
int orc_node_flag(const OrcPBVH *o, int node);
void orc_vert_coords_apply(OrcPBVH *o, const float *vert_cos);

/* Ray cast (ray.c).  One ray against the current positions; depth0 bounds the hit
 * (hits at depth >= depth0 are ignored).  Returns hit.  stats (may be NULL): node
 * box tests, leaves passing the traversal, leaves ray-tested, tris tested. */
typedef struct OrcRayHit {
  int hit;
  float depth;
  float location[3];
  float face_normal[3];
  int tri;           /* looptri index of the hit tri */
  int active_vertex; /* corner of that tri nearest the hit point */
  int leaf;          /* node id of its leaf */
} OrcRayHit;
int orc_raycast(const OrcPBVH *o, const float start[3], const float dir[3], float depth0,
                int original, OrcRayHit *out, int64_t stats[4]);

typedef struct OrcRayNearest {
  int hit;
  float depth;
  float dist_sq;
  float location[3];
  int tri;
  int leaf;
} OrcRayNearest;
int orc_find_nearest_to_ray(const OrcPBVH *o, const float start[3], const float dir[3],
                            int original, OrcRayNearest *out);
/* Leaves in the occluded search's visiting order with their tmin (mode 0 ray cast,
 * 1 nearest to ray); returns the count. */
int orc_occluded_order(const OrcPBVH *o, const float start[3], const float dir[3], int original,
                       int mode, int *nodes, float *tmin, int cap);
/* BKE_pbvh_raycast_project_ray_root; returns 1 when start/end were moved. */
int orc_project_ray_root(const OrcPBVH *o, int original, float start[3], float end[3], float dir[3]);
/* The pieces, for known-answer tests. */
int orc_isect_ray_tri(const float start[3], const float dir[3], const float v0[3], const float v1[3],
                      const float v2[3], float *lambda);
int orc_isect_ray_box(const float start[3], const float dir[3], const float bmin[3],
                      const float bmax[3], float *tmin);
float orc_dist_sq_ray_box(const float start[3], const float dir[3], const float bmin[3],
                          const float bmax[3], float *depth);
float orc_dist_sq_ray_tri(const float start[3], const float dir[3], const float v0[3],
                          const float v1[3], const float v2[3], float *depth);

/* Draw buffer of one leaf (draw.c): the visible tris' vertices in the reference's
 * 36-byte vertex layout.  Polygons: tri -> poly, poly corner lists (CSR, vertex ids),
 * smooth flag and material per poly.  mask may be NULL; face_sets per poly (NULL:
 * nothing is visible, as in the reference).  Returns the vertex count (-1 if cap_verts
 * is too small). */
typedef struct OrcPolys {
  const int *tri_poly;
  const int *poly_off;
  const int *poly_verts;
  const uint8_t *poly_smooth;
  const int *poly_mat;
} OrcPolys;
int orc_draw_leaf(const OrcPBVH *o, int node, const OrcPolys *polys, const float *mask,
                  const int *face_sets, int seed, int fset_default, unsigned char *out, int cap_verts,
                  int *r_smooth, int *r_material);
/* BKE_paint_face_set_overlay_color_get (rgb) */
void orc_face_set_color(int face_set, int seed, unsigned char out[3]);
void orc_lookup3_final(uint32_t abc[3]);
float orc_hash_int_01(uint32_t k);

/* Multires grids (grids.c): grids of a coarse polygon mesh (grid_size^2 elements per
 * polygon corner, element (g, x, y) at g * grid_size^2 + y * grid_size + x), the tree
 * over them, the seam averages and normals, and the grid dab.  co [G*ga*3]; no
 * (NULL: no normal layer), mask (NULL: none), hidden bytes per element (NULL: none),
 * grid_mat material key per grid (NULL: one material); leaf_limit before the
 * division by grid_size^2. */
typedef struct OrcGrids OrcGrids;
OrcGrids *orc_grids_new(int totvert, const int *poly_off, const int *poly_verts, int totpoly, int grid_size,
                        const float *co, const float *no, const float *mask, const uint8_t *hidden,
                        const int *grid_mat, int leaf_limit, int respect_hide);
void orc_grids_free(OrcGrids *o);
int orc_grids_totnode(const OrcGrids *o);
/* flag, children_offset (-1 for leaves), prim offset, prim count */
void orc_grids_node_info(const OrcGrids *o, int node, int out[4]);
void orc_grids_node_bb(const OrcGrids *o, int node, float vb[6], float orig_vb[6]);
void orc_grids_prim_indices(const OrcGrids *o, int *out);
void orc_grids_get(const OrcGrids *o, float *co, float *no, float *mask);
void orc_grids_set_co(OrcGrids *o, const float *co);
void orc_grids_last_stats(const OrcGrids *o, int64_t out[4]);
int orc_grids_edge_count(const OrcGrids *o);
int orc_grids_count_quads(const OrcGrids *o, const int *grids, int n);
/* neighbours of (g, x, y) as (grid, x, y) triples; *r_ndup of them (the last) are
 * duplicates of the element itself; returns the count */
int orc_grids_neighbors(const OrcGrids *o, int g, int x, int y, int include_duplicates, int *out, int cap,
                        int *r_ndup);
void orc_grids_average(OrcGrids *o);
void orc_grids_recalc_normals(OrcGrids *o);
void orc_grids_update_normals_faces(OrcGrids *o, const int *faces, int nfaces);
int orc_grids_search(const OrcGrids *o, const float center[3], float radius, int use_orig, int *out, int cap);
/* returns |A|; out: affected element ids, ascending */
int orc_grids_dab(OrcGrids *o, const OrcDab *dab, int *out_affected, int cap);
void orc_grids_mark(OrcGrids *o, const int *nodes, int n, int flag);
void orc_grids_update_normals(OrcGrids *o);
void orc_grids_update_bounds(OrcGrids *o);

/* Threads used by the OpenMP loops (0 = runtime default). */
void orc_set_threads(int n);
int orc_get_threads(void);

/* Stats of the last dab for the byte model (gathered leaves etc.). */
void orc_last_stats(const OrcPBVH *o, int64_t out[8]);

/* ------------------------------------------------------------------------- */
/* Dynamic topology, PBVH_BMESH (dyntopo.c).                                  */
/* ------------------------------------------------------------------------- */
enum { ORC_Subdivide = 1, ORC_Collapse = 2 }; /* PBVHTopologyUpdateMode */
typedef struct OrcDyn OrcDyn;
OrcDyn *orc_dyn_new(const float *co, int totvert, const int *tris, int tottri, const float *mask,
                    const uint8_t *vhide, const uint8_t *fhide, int vert_cap, int tri_cap, int leaf_limit,
                    int respect_hide, char *err, int errlen);
void orc_dyn_free(OrcDyn *o);
void orc_dyn_detail_size_set(OrcDyn *o, float detail_size);
void orc_dyn_topology_set(OrcDyn *o, int mode, const float view_normal[3], int use_frontface, int use_projected);
void orc_dyn_set_normals(OrcDyn *o, const float *no);
int orc_dyn_search(const OrcDyn *o, const float center[3], float radius, int *out, int cap);
int orc_dyn_update_topology(OrcDyn *o, int mode, const float center[3], float radius);
int orc_dyn_subdivide_sequential(OrcDyn *o, const float center[3], float radius);
int orc_dyn_dab(OrcDyn *o, const OrcDab *dab, uint64_t *out_affected, int cap);
void orc_dyn_after_stroke(OrcDyn *o);
void orc_dyn_mark(OrcDyn *o, const int *nodes, int n, int flag);
void orc_dyn_update_normals(OrcDyn *o);
void orc_dyn_update_bounds(OrcDyn *o);
void orc_dyn_counts(const OrcDyn *o, int out[6]);
void orc_dyn_export_verts(const OrcDyn *o, uint64_t *lab, float *co, float *no, float *mask, int *node);
void orc_dyn_export_faces(const OrcDyn *o, uint64_t *labs, int *node, float *fno, uint8_t *hide);
void orc_dyn_node(const OrcDyn *o, int node, int out[3], float vb[6], float orig_vb[6]);
void orc_dyn_last_stats(const OrcDyn *o, int64_t out[12]);
int orc_dyn_check(const OrcDyn *o);
#ifdef __cplusplus
}
#endif
