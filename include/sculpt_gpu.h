/*
 * sculpt_gpu.h -- C ABI of the MI355X sculpt brush engine (libsculpt_gpu.so).
 *
 * Two groups of entry points:
 *
 *  1. Drop-in replacements for the reference's PBVH C API (PBVH_FACES; PBVH_GRIDS in
 *     section 3, PBVH_BMESH in section 4).  The reference's public header (BKE_pbvh.h / DUNE_pbvh.h /
 *     KERNEL_pbvh.h) is absent,
 *     so each signature is taken from its definition in kernel/intern/pbvh.c and its
 *     call sites (paint.c:1676-1771, :2084-2191; editors/sculpt_paint/paint_hide.c).
 *     Each declaration cites the reference function it replaces.  Mesh data stays in
 *     HBM (SoA, leaf-ordered); host views are synchronised lazily.
 *
 *  2. New sg_* entry points replacing the reference's absent brush action
 *     (do_brush_action / SCULPT_OT_brush_stroke, named at paint_curve.c:671):
 *     one dab or a whole stroke of dabs on the GPU, plus batched multi-mesh steps.
 *     These return int status codes (0 = ok, < 0 = error, text in sg_last_error()).
 *
 * Struct layouts of MVert/MPoly/MLoop/MLoopTri match types/types_meshdata.h:13-17,
 * :50-57, :69-74, :178-181 so reference mesh arrays can be passed as they are.
 * Threading: calls from one host thread per PBVH; work is queued on the PBVH's own
 * HIP stream and every drop-in call returns with the stream drained.
 */
#ifndef SCULPT_GPU_H
#define SCULPT_GPU_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mesh data layouts (types/types_meshdata.h) ---------------------------- */
typedef struct MVert { /* types_meshdata.h:13-17 */
  float co[3];
  char flag, bweight;
  char _pad[2];
} MVert;
enum { ME_HIDE = (1 << 4) }; /* types_meshdata.h:20-24 */

typedef struct MPoly { /* types_meshdata.h:50-57 */
  int loopstart;
  int totloop;
  short mat_nr;
  char flag, _pad;
} MPoly;
enum { ME_SMOOTH = (1 << 0) }; /* types_meshdata.h:60-64 */

typedef struct MLoop { /* types_meshdata.h:69-74 */
  unsigned int v;
  unsigned int e;
} MLoop;

typedef struct MLoopTri { /* types_meshdata.h:178-181 */
  unsigned int tri[3];
  unsigned int poly;
} MLoopTri;

/* Opaque reference types accepted (and ignored) by the drop-in signatures. */
typedef struct Mesh Mesh;
typedef struct CustomData CustomData;
typedef struct SubdivCCG SubdivCCG;

typedef struct PBVH PBVH;
typedef struct PBVHNode PBVHNode;

/* PBVHType / PBVHNodeFlags are used but never defined by the reference
 * (pbvh.c:550, :114, :1003, :1215, :1725); the names follow those uses. */
typedef enum { PBVH_FACES = 0, PBVH_GRIDS = 1, PBVH_BMESH = 2 } PBVHType;
typedef enum {
  PBVH_Leaf = 1 << 0,
  PBVH_UpdateNormals = 1 << 1,
  PBVH_UpdateBB = 1 << 2,
  PBVH_UpdateOriginalBB = 1 << 3,
  PBVH_UpdateDrawBuffers = 1 << 4,
  PBVH_UpdateRedraw = 1 << 5,
  PBVH_UpdateMask = 1 << 6,
  PBVH_UpdateVisibility = 1 << 7,
  PBVH_RebuildDrawBuffers = 1 << 8,
  PBVH_FullyHidden = 1 << 9,
  PBVH_FullyMasked = 1 << 10,
  PBVH_FullyUnmasked = 1 << 11,
  PBVH_UpdateColor = 1 << 12,
  PBVH_UpdateTopology = 1 << 13,
} PBVHNodeFlags;

/* PBVHTopologyUpdateMode (pbvh_bmesh.c:1932, :1956, :1976) */
typedef enum { PBVH_Subdivide = 1, PBVH_Collapse = 2 } PBVHTopologyUpdateMode;

typedef struct PBVHProxyNode { /* used at pbvh.c:4757-4789 */
  float (*co)[3];
} PBVHProxyNode;

typedef struct PBVHFrustumPlanes { /* used at pbvh.c:4507-4557 */
  float (*planes)[4];
  int num_planes;
} PBVHFrustumPlanes;

typedef bool (*BKE_pbvh_SearchCallback)(PBVHNode *node, void *data); /* pbvh.c:769 */
typedef void (*BKE_pbvh_HitCallback)(PBVHNode *node, void *data);    /* pbvh.c:864 */
/* the occluded searches' hit callbacks, called with traverse_tree's bound (pbvh.c:898-912) */
typedef void (*BKE_pbvh_HitOccludedCallback)(PBVHNode *node, void *data, float *tmin);
typedef void (*BKE_pbvh_SearchNearestCallback)(PBVHNode *node, void *data, float *tmin);
/* Watertight ray/triangle precalc; fields as math_geom.cc:1755-1855 uses them (the lib
 * header is absent). */
typedef struct IsectRayPrecalc {
  int kx, ky, kz;
  float sx, sy, sz;
} IsectRayPrecalc;

/* ---- 1. drop-in PBVH API ---------------------------------------------------- */

/* pbvh.c:645 BKE_pbvh_new */
PBVH *BKE_pbvh_new(void);
/* pbvh.c:652 BKE_pbvh_free (also frees the looptri array it took over, pbvh.c:687) */
void BKE_pbvh_free(PBVH *pbvh);
/* pbvh.c:534 BKE_pbvh_build_mesh.  Builds the tree on the host exactly as the
 * reference does, then uploads a leaf-ordered SoA copy of the mesh to HBM and
 * computes the angle-weighted vertex normals on the GPU (mesh_normals.c:225-299).
 * Takes ownership of looptri (freed with free() by BKE_pbvh_free, as pbvh.c:687). */
void BKE_pbvh_build_mesh(PBVH *pbvh, Mesh *mesh, const MPoly *mpoly, const MLoop *mloop,
                         MVert *verts, int totvert, CustomData *vdata, CustomData *ldata,
                         CustomData *pdata, const MLoopTri *looptri, int looptri_num);
/* pbvh.c:818 BKE_pbvh_search_gather: result array is malloc'ed, caller frees. */
void BKE_pbvh_search_gather(PBVH *pbvh, BKE_pbvh_SearchCallback scb, void *search_data,
                            PBVHNode ***r_array, int *r_tot);
/* pbvh.c:851 BKE_pbvh_search_callback */
void BKE_pbvh_search_callback(PBVH *pbvh, BKE_pbvh_SearchCallback scb, void *search_data,
                              BKE_pbvh_HitCallback hcb, void *hit_data);
/* pbvh.c:1401 BKE_pbvh_update_bounds: refits flagged leaves and their ancestors on
 * the GPU (host-side edits made through BKE_pbvh_get_verts are uploaded for the
 * flagged leaves first; proxies are folded in by sg_pbvh_combine_proxies). */
void BKE_pbvh_update_bounds(PBVH *pbvh, int flag);
/* pbvh.c:4559 BKE_pbvh_update_normals: marked vertices of UpdateNormals leaves get
 * normalize(sum of incident unit tri normals) -- deterministic CSR gather on GPU. */
void BKE_pbvh_update_normals(PBVH *pbvh, SubdivCCG *subdiv_ccg);
/* pbvh.c:4707 BKE_pbvh_vert_coords_apply */
void BKE_pbvh_vert_coords_apply(PBVH *pbvh, const float (*vertCos)[3], int totvert);
/* pbvh.c:4689 BKE_pbvh_vert_coords_alloc (malloc'ed, caller frees) */
float (*BKE_pbvh_vert_coords_alloc(PBVH *pbvh))[3];
/* pbvh.c:1723 / :1811 / :1753 marking */
void BKE_pbvh_node_mark_update(PBVHNode *node);
void BKE_pbvh_node_mark_normals_update(PBVHNode *node);
void BKE_pbvh_node_mark_redraw(PBVHNode *node);
void BKE_pbvh_vert_mark_update(PBVH *pbvh, int index);
/* pbvh.c:3830-3840 node bounds (synchronised from HBM on first read after a change) */
void BKE_pbvh_node_get_BB(PBVHNode *node, float bb_min[3], float bb_max[3]);
void BKE_pbvh_node_get_original_BB(PBVHNode *node, float bb_min[3], float bb_max[3]);
/* pbvh.c:1831 / :1817 node vertex access (original vertex indices; uniq first) */
void BKE_pbvh_node_num_verts(PBVH *pbvh, PBVHNode *node, int *r_uniquevert, int *r_totvert);
void BKE_pbvh_node_get_verts(PBVH *pbvh, PBVHNode *node, const int **r_vert_indices,
                             MVert **r_verts);
/* pbvh.c:1761 / :1766 */
bool BKE_pbvh_node_fully_hidden_get(PBVHNode *node);
/* pbvh.c:4757 / :4778 / :4791 proxies.  add_proxy returns a zeroed buffer carved from
 * pinned host memory the PBVH owns (uploaded by sg_pbvh_combine_proxies without a
 * copy on the host); release it only through BKE_pbvh_node_free_proxies. */
PBVHProxyNode *BKE_pbvh_node_add_proxy(PBVH *pbvh, PBVHNode *node);
void BKE_pbvh_node_free_proxies(PBVHNode *node);
void BKE_pbvh_gather_proxies(PBVH *pbvh, PBVHNode ***r_array, int *r_tot);
/* pbvh.c:4539-4557 frustum tests (search callbacks taking PBVHFrustumPlanes*) */
bool BKE_pbvh_node_frustum_contain_AABB(PBVHNode *node, void *frustum);
bool BKE_pbvh_node_frustum_exclude_AABB(PBVHNode *node, void *frustum);
/* pbvh.c:4961 / :4967 host views, synchronised from HBM when called */
MVert *BKE_pbvh_get_verts(const PBVH *pbvh);
const float (*DUNE_pbvh_get_vert_normals(const PBVH *pbvh))[3];
/* pbvh.c:4983 */
void DUNE_pbvh_respect_hide_set(PBVH *pbvh, bool respect_hide);
/* paint.c:2089 calls the setter by this name; same function. */
void BKE_pbvh_respect_hide_set(PBVH *pbvh, bool respect_hide);
/* pbvh.c:4953 PBVHType accessor (BKE_pbvh_type is called throughout paint.c) */
PBVHType BKE_pbvh_type(const PBVH *pbvh);
/* pbvh.c:3915 BKE_pbvh_raycast / :4339 BKE_pbvh_find_nearest_to_ray: the leaves the
 * occluded search passes (every ancestor's box too) are found on the GPU, then cb is
 * called for each in the reference's order (ascending tmin, ties left->right, the
 * node_tree_insert placement of NaN) with the bound starting at FLT_MAX. */
void BKE_pbvh_raycast(PBVH *pbvh, BKE_pbvh_HitOccludedCallback cb, void *data, const float ray_start[3],
                      const float ray_normal[3], bool original);
void BKE_pbvh_find_nearest_to_ray(PBVH *pbvh, BKE_pbvh_SearchNearestCallback cb, void *data,
                                  const float ray_start[3], const float ray_normal[3], bool original);
/* pbvh.c:929 BKE_pbvh_node_get_tmin: the tmin the last occluded search gave the node */
float BKE_pbvh_node_get_tmin(PBVHNode *node);
/* pbvh.c:4203 BKE_pbvh_node_raycast / :4463 BKE_pbvh_node_find_nearest_to_ray: one
 * leaf's triangles on the GPU (origco: the node's original coordinates, indexed by its
 * face_vert_indices, as the reference; use_origco is for dyntopo and ignored). */
bool BKE_pbvh_node_raycast(PBVH *pbvh, PBVHNode *node, float (*origco)[3], bool use_origco,
                           const float ray_start[3], const float ray_normal[3],
                           IsectRayPrecalc *isect_precalc, float *depth, int *active_vertex_index,
                           int *active_face_grid_index, float *face_normal);
bool BKE_pbvh_node_find_nearest_to_ray(PBVH *pbvh, PBVHNode *node, float (*origco)[3],
                                       bool use_origco, const float ray_start[3],
                                       const float ray_normal[3], float *depth, float *dist_sq);
/* pbvh.c:4262 BKE_pbvh_raycast_project_ray_root */
void BKE_pbvh_raycast_project_ray_root(PBVH *pbvh, bool original, float ray_start[3],
                                       float ray_end[3], float ray_normal[3]);
/* isect_ray_tri_watertight_v3_precalc (math_geom.cc:1755-1780) under this library's
 * prefix, for callers without the reference's math library. */
void sg_isect_ray_tri_precalc(IsectRayPrecalc *out, const float ray_direction[3]);

/* ---- 2. brush engine entry points (new; replace the absent do_brush_action) ---- */

/* Tools (types_brush_enums.h:410-415) and falloff presets (:176-187). */
enum { SG_TOOL_DRAW = 1, SG_TOOL_SMOOTH = 2, SG_TOOL_INFLATE = 4, SG_TOOL_GRAB = 5, SG_TOOL_ELASTIC_DEFORM = 21 };
/* Elastic deform modes (eBrushElasticDeformType, types_brush_enums.h:195-199). */
enum {
  SG_ELASTIC_GRAB = 0, SG_ELASTIC_GRAB_BISCALE = 1, SG_ELASTIC_GRAB_TRISCALE = 2, SG_ELASTIC_SCALE = 3,
  SG_ELASTIC_TWIST = 4,
};
enum {
  SG_CURVE_CUSTOM = 0, SG_CURVE_SMOOTH = 1, SG_CURVE_SPHERE = 2, SG_CURVE_ROOT = 3,
  SG_CURVE_SHARP = 4, SG_CURVE_LIN = 5, SG_CURVE_POW4 = 6, SG_CURVE_INVSQUARE = 7,
  SG_CURVE_CONSTANT = 8, SG_CURVE_SMOOTHER = 9,
};
#define SG_BRUSH_DIR_IN (1 << 9)    /* BRUSH_DIR_IN, types_brush_enums.h:347 */
#define SG_SMOOTH_LEARNED (1 << 30) /* Smooth through the attached learned model */
/* Grab from the original coordinates (single-mesh entry points): each uniq vertex of a
 * gathered leaf is tested and displaced from its coordinate at sg_undo_begin (the
 * current one when no undo step is open) and every one of them becomes original +
 * displacement, zero outside the sphere -- the reference's Grab over SculptOrigVertData
 * followed by the proxy combine from the original coordinates.  All written vertices
 * are the dab's affected set. */
#define SG_DAB_ORIGINAL (1 << 28)
#define SG_CM_TABLE 256             /* CM_TABLE, types_color.h:12 */

/* One dab.  Field names mirror the Brush RNA properties (api_brush.c:2281-2880) and
 * OperatorStrokeElement (api_brush.c:3526-3576) they come from. */
typedef struct SgDabParams {
  int sculpt_tool;            /* SG_TOOL_* (api_brush.c:2281) */
  int curve_preset;           /* SG_CURVE_* (api_brush.c:2364) */
  int flag;                   /* SG_BRUSH_DIR_IN */
  int smooth_iterations;      /* Laplacian Smooth(k); 0 -> 1 */
  float location[3];          /* dab centre (OperatorStrokeElement.location) */
  float radius;               /* world-space radius (unprojected_radius, :2463) */
  float strength;             /* Brush.alpha (:2536) */
  float hardness;             /* (:2420) */
  float normal_radius_factor; /* area-normal sphere for Draw (:2714 area), default 0.5 */
  float grab_delta[3];        /* Grab translation for this dab */
  const float *curve_table;   /* SG_CM_TABLE+1 samples over [0,1] for SG_CURVE_CUSTOM */
  /* Elastic deform (single-mesh entry points): mode (elastic_deform_type, :2375),
   * Poisson ratio (elastic_deform_volume_preservation, :2722, range 0..0.9) and the
   * stroke's sculpt normal, which Scale and Twist use. */
  int elastic_deform_type;
  float elastic_deform_volume_preservation;
  float surface_normal[3];
} SgDabParams;

/* Fills defaults (types_brush_defaults.h:8-90: SMOOTH falloff, hardness 0,
 * normal_radius_factor 0.5, strength 0.5). */
void sg_dab_params_default(SgDabParams *p, int tool);

/* Set the leaf size (triangles per leaf) used by the next BKE_pbvh_build_mesh.
 * The reference hard-codes LEAF_LIMIT 10000 (pbvh.c:32); 0 restores that. */
int sg_pbvh_leaf_limit_set(PBVH *pbvh, int leaf_limit);

/* Apply one dab: search -> falloff/strength -> displacement -> normals -> leaf and
 * node bounds refit, all on the GPU, returning when done. */
int sg_dab_apply(PBVH *pbvh, const SgDabParams *dab);
/* Apply n dabs in order on the PBVH's stream without host synchronisation, then
 * refit every dirty internal node and drain.  flags: SG_STROKE_GRAPH captures the
 * whole stroke into one hipGraph; SG_STROKE_REFIT_EACH refits internal nodes after
 * every dab (leaf and search-level bounds are always refit per dab);
 * SG_STROKE_LOG keeps per-dab counters for sg_dab_stroke_log. */
enum { SG_STROKE_GRAPH = 1, SG_STROKE_REFIT_EACH = 2, SG_STROKE_LOG = 4 };
int sg_dab_apply_stroke(PBVH *pbvh, const SgDabParams *dabs, int n, int flags);
/* A stroke captured once into a hipGraph (all its dabs + the final refit) and
 * launched any number of times on the PBVH's stream (each launch applies the dabs
 * again to the current mesh).  sync != 0 waits for completion. */
typedef struct SgStroke SgStroke;
SgStroke *sg_stroke_prepare(PBVH *pbvh, const SgDabParams *dabs, int n, int flags);
int sg_stroke_launch(SgStroke *stroke, int sync);
void sg_stroke_free(SgStroke *stroke);
/* Per-dab counters (24 per dab, layout of sg_dab_last_stats) of the last stroke. */
int sg_dab_stroke_log(PBVH *pbvh, long long *out, int n);
/* Block until queued work is done; returns the first recorded device error. */
int sg_sync(PBVH *pbvh);
const char *sg_last_error(void);

/* Affected vertices (original indices, ascending) of the most recent dab. */
int sg_dab_last_affected(PBVH *pbvh, int *out, int cap);
/* Counters of the most recent dab:
 * [0] gathered leaves [1] affected verts [2..4] area-normal sums (fixed point)
 * [5] cut subtrees touched [6] boxes tested; with statistics on (sg_pbvh_set_stats):
 * [7] candidate verts [8] area-normal verts [9] smooth neighbour reads (steps >= 2
 * plus step 1) [10] normal tri reads [11] distinct non-affected 1-ring verts (last
 * Jacobi steps) [12] distinct tris incident to A [13] distinct verts of those tris
 * [14] shared verts read by the leaf refit; learned Smooth: [15] sum (|d|/r)^2 in 2^-40
 * units, [16] out-of-range vertices, [17] 1 when accepted.  PBVH_GRIDS: [0], [1]
 * (affected elements), [2..4], [8], and [20] leaves refit, [21] faces whose normals
 * were redone, [22] edges and coarse vertices averaged. */
int sg_dab_last_stats(PBVH *pbvh, int64_t out[24]);
/* Gathered leaves (node indices, left->right) of the most recent dab. */
int sg_dab_last_gathered(PBVH *pbvh, int *out, int cap);
/* Exact byte-accounting counters in the brush kernels (adds atomics; off by default). */
int sg_pbvh_set_stats(PBVH *pbvh, int on);
/* Workgroups launched by the per-dab kernels (autotuned by the host layer). */
int sg_pbvh_set_grid(PBVH *pbvh, int workgroups);
/* Workgroups of the brush / post / smooth-step kernels (0 = the common grid). */
int sg_pbvh_set_kernel_grids(PBVH *pbvh, int brush, int post, int smooth);
int sg_pbvh_get_kernel_grids(PBVH *pbvh, int out[4]);
/* Fused dabs: on (default) a plain Draw / Inflate / Grab / Smooth(k <= 4) dab on a
 * triangle mesh runs as the search plus one kernel per gathered leaf over the leaf's
 * halo; off, the separate brush / Smooth-step / normals-and-bounds kernels.  Both give
 * the same bits.  Returns whether the mesh has the halo tables the fused path needs. */
int sg_pbvh_set_fused(PBVH *pbvh, int on);
/* Halo tables of the fused dabs: {built, most locals of a leaf (all rings), most
 * locals in ring 0, most neighbour entries of a leaf, device bytes of the tables}. */
int sg_pbvh_halo_info(PBVH *pbvh, long long out[5]);
/* Autotuner: times a sample of dabs from the current mesh state (restored after each
 * trial) for candidate workgroup counts of each per-dab kernel and keeps the fastest.
 * out_ms (may be NULL) receives {time before, time after} of the sample in ms. */
int sg_pbvh_autotune(PBVH *pbvh, const SgDabParams *dabs, int n, double *out_ms);
/* Checkpoint of the mutable device state (positions, normals, all node bounds):
 * save copies it in HBM, restore puts the mesh back exactly (undo / stroke cancel /
 * replaying a stroke on the same starting mesh). */
int sg_pbvh_snapshot_save(PBVH *pbvh);
int sg_pbvh_snapshot_restore(PBVH *pbvh);
/* Per-kernel timing with HIP events on the PBVH's stream (stream launches only).
 * sg_pbvh_profile_read returns the number of kernel kinds and fills total ms and
 * launch counts for: 0 search, 1 area normal, 2 brush, 3 smooth step, 4 post
 * (normals + leaf bounds), 5 subtree refit, 6 top refit. */
int sg_pbvh_set_profile(PBVH *pbvh, int on);
int sg_pbvh_profile_read(PBVH *pbvh, double *ms, long long *count, int n, int reset);
/* Wall time of everything enqueued on the PBVH's stream between start and stop (a HIP
 * event pair on that stream; stop waits for the second event). */
int sg_pbvh_timer_start(PBVH *pbvh);
/* Device-to-device copy of `bytes` (best of `reps`, HIP events): the measured copy
 * bandwidth, and a kernel of known traffic for calibrating hardware byte counters. */
int sg_hbm_copy_probe(long long bytes, int reps, float *ms);
int sg_pbvh_timer_stop(PBVH *pbvh, float *ms);

/* Copies of the device state in original vertex order (AoS float[3]). */
int sg_pbvh_read_positions(PBVH *pbvh, float *out);
int sg_pbvh_read_normals(PBVH *pbvh, float *out);
/* Per-vertex paint mask (CD_PAINT_MASK); NULL clears it. */
int sg_pbvh_set_mask(PBVH *pbvh, const float *mask);
/* Tree shape for inspection: node count, leaf count, cut count. */
int sg_pbvh_tree_info(PBVH *pbvh, int out[4]);
/* Node index of a handle (its slot in the reference's nodes array order). */
int sg_pbvh_node_index(PBVH *pbvh, PBVHNode *node);
PBVHNode *sg_pbvh_node_at(PBVH *pbvh, int index);
/* A node's PBVHNodeFlags (the reference's callers read node->flag directly,
 * pbvh_intern.h:69; the handle is opaque here). */
int sg_pbvh_node_flag(PBVHNode *node);
/* prim_indices array (looptri order of the leaves) */
int sg_pbvh_prim_indices(PBVH *pbvh, int *out, int cap);

/* ---- ray cast (stroke placement) ----------------------------------------------
 * One ray = BKE_pbvh_raycast (pbvh.c:3915-3928) with the sculpt operator's hit
 * callback (sculpt.c is absent: leaves visited in ascending box-entry tmin, each
 * tested while its tmin is below the running bound, FLT_MAX then the hit depth) and
 * BKE_pbvh_node_raycast -> pbvh_faces_node_raycast (:4041-4107, :4203-4260) on the
 * current positions; the hit point is ray_start + ray_normal * depth (the stroke's
 * location, SCULPT_stroke_get_location in the absent sculpt.c).  Bit-identical to
 * that sequential visit. */
typedef struct SgRay {
  float start[3];  /* object space */
  float normal[3]; /* direction (unit in the reference's callers) */
  float depth;     /* initial depth: only hits strictly nearer count (FLT_MAX: none) */
  int reserved;
} SgRay;
typedef struct SgRayHit {
  int hit;              /* a triangle was hit */
  float depth;          /* ray parameter of the hit (the initial depth when none) */
  float location[3];    /* start + normal * depth */
  float face_normal[3]; /* normal_tri_v3 of the hit triangle (pbvh.c:4073) */
  int active_vertex;    /* its corner nearest the hit (pbvh.c:4076-4101), mesh index */
  int active_face;      /* MLoopTri.poly of the hit triangle */
  int looptri;          /* looptri index of the hit triangle */
  int node;             /* PBVH node index of its leaf */
  int status;           /* 0, or SG_RAY_OVERFLOW (more than 512 leaves hit: no result) */
  int reserved;
} SgRayHit;
enum { SG_RAY_OVERFLOW = -2 };
/* Rays from host memory, results to host memory; original != 0 tests orig_vb
 * (BKE_pbvh_raycast's `original`). */
int sg_raycast(PBVH *pbvh, const SgRay *rays, int n, int original, SgRayHit *hits);
/* The same with device pointers, queued on the PBVH's stream (sg_sync waits). */
int sg_raycast_device(PBVH *pbvh, const SgRay *rays_dev, int n, int original, SgRayHit *hits_dev);
/* Work counters of a ray batch: [0] boxes tested [1] leaves scanned [2] tris tested. */
int sg_raycast_stats(PBVH *pbvh, const SgRay *rays, int n, int original, long long out[3]);
/* No counterpart in the reference.  sg_raycast_stats for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the same results with no hipMalloc / hipFree.  Its scratch (counters, staged rays and
 * hits) is a static device array holding at most SG_NOALLOC_RAYS rays; more is an error.
 * The ray layout must exist already (one sg_raycast on the PBVH builds it; otherwise an
 * error).  Calls that overlap on one GPU share that array and race. */
enum { SG_NOALLOC_RAYS = 4096 };
int sg_raycast_stats_noalloc(PBVH *pbvh, const SgRay *rays, int n, int original, long long out[3]);
/* A stroke whose dab centres come from rays cast on the GPU, each on the mesh as the
 * previous dab left it (the stroke location of each OperatorStrokeElement,
 * api_brush.c:3526-3576, found by SCULPT_stroke_get_location in the absent sculpt.c):
 * dab i's location is ray i's hit, and a dab whose ray misses does nothing; a Grab
 * dab's translation is ray i+1's hit minus ray i's (zero for the last dab or a miss).
 * Other SgDabParams fields and flags as for sg_dab_apply_stroke / sg_stroke_prepare. */
int sg_dab_apply_stroke_placed(PBVH *pbvh, const SgDabParams *dabs, const SgRay *rays, int n, int flags);
SgStroke *sg_stroke_prepare_placed(PBVH *pbvh, const SgDabParams *dabs, const SgRay *rays, int n, int flags);
/* Placements of the last ray-placed stroke, 8 floats per dab: centre xyz, Grab
 * translation xyz, 1 when the ray hit (else 0), 0.  Returns the number of dabs. */
int sg_stroke_placements(PBVH *pbvh, float *out, int n);

/* ---- draw buffers (the draw code's per-leaf vertex data, kept in HBM) -----------
 * What GPU_pbvh_mesh_buffers_build / gpu_pbvh_mesh_buffers_update
 * (gpu/intern/gpu_buffers.c:174-350, driven by pbvh_update_draw_buffers,
 * pbvh.c:1251-1367) write for every leaf, in the same vertex layout (36-byte stride:
 * pos float3 @0, nor short3 @16, msk uchar @22, ac ushort4 @24 (zero), fset uchar3 @32),
 * written on the GPU into one device arena, leaf after leaf, so a renderer on the same
 * GPU reads it in place.  face_sets: one int per polygon (CD_SCULPT_FACE_SETS; a tri
 * is drawn when no corner is hidden and its face set is > 0, as the reference).  */
int sg_pbvh_draw_buffers_build(PBVH *pbvh, const int *face_sets, int npoly, int seed, int face_set_default);
/* Refill leaves the dabs gathered since the last update and leaves marked
 * PBVH_UpdateDrawBuffers / PBVH_RebuildDrawBuffers on the host (all != 0: every leaf). */
int sg_pbvh_draw_buffers_update(PBVH *pbvh, int all);
int sg_pbvh_draw_buffers_info(PBVH *pbvh, void **vbo_dev, long long *bytes, int *visible_tris);
/* One leaf's records in the arena; smooth and material as the reference's buffers hold them. */
int sg_pbvh_draw_buffer_node(PBVH *pbvh, PBVHNode *node, long long *byte_offset, int *vert_count,
                             int *smooth, int *material);
long long sg_pbvh_draw_buffer_read(PBVH *pbvh, PBVHNode *node, void *out, long long cap);
void sg_pbvh_draw_buffers_free(PBVH *pbvh);

/* ---- undo steps (snapshots of the leaves an operation touches, kept in HBM) ------
 * Replaces SCULPT_undo_push_begin / SCULPT_undo_push_node / SCULPT_undo_push_end as
 * the reference calls them around an operation (editors/sculpt_paint/paint_hide.c:
 * 350-376 begin/end, :78/:126/:255 node; sculpt_undo.c itself is absent): between
 * begin and end every dab saves each leaf it gathers, once, with its positions,
 * normals and box as they were before the step; push_node saves a leaf the host is
 * about to change.  max_verts bounds the saved vertices of one step (<= 0: the whole
 * mesh); a step that would exceed it is dropped at end with an error.  sg_undo /
 * sg_redo return 0 after applying a step, 1 when there is none; they mark the touched
 * subtrees for the bounds refit and the draw-buffer refill.  Beginning a step drops
 * the steps undone so far.  info: steps held, steps applied, device bytes held,
 * leaves in the newest step. */
int sg_undo_begin(PBVH *pbvh, long long max_verts);
int sg_undo_push_node(PBVH *pbvh, PBVHNode *node);
int sg_undo_end(PBVH *pbvh);
int sg_undo(PBVH *pbvh);
int sg_redo(PBVH *pbvh);
int sg_undo_info(PBVH *pbvh, long long out[4]);
void sg_undo_clear(PBVH *pbvh);
/* Cancel of the open step (a cancelled stroke): its leaves return to their state at
 * begin and the step is dropped. */
int sg_undo_cancel(PBVH *pbvh);
/* Original data: positions and normals at the begin of the open step (current ones when
 * none is open), float[3] per vertex in the caller's order; either may be NULL.  What
 * the reference reads from its undo nodes (SculptOrigVertData: Grab, anchored strokes,
 * BKE_pbvh_node_raycast's origco). */
int sg_undo_original(PBVH *pbvh, float *co, float *no);
/* The open step's saved leaves back to their state at begin, the step staying open
 * (the sculpt code restores the original mesh from its undo nodes before each dab of
 * an anchored stroke; sculpt.c is absent from the reference). */
int sg_undo_restore_original(PBVH *pbvh);

/* ---- proxies combined on the GPU -------------------------------------------------
 * Folds the proxies added with BKE_pbvh_node_add_proxy (host buffers, one float[3] per
 * uniq vertex, in the order of BKE_pbvh_node_get_verts) into the mesh on the GPU:
 * vertex = (original coordinate when use_orco, else current) + each proxy in order;
 * then its normal, the leaf boxes and the tree; frees the proxies (the sculpt code's
 * combine step over BKE_pbvh_gather_proxies, pbvh.c:4791-4818; sculpt.c is absent).
 * Original coordinates are those of the open undo step.  Returns the leaves changed. */
int sg_pbvh_combine_proxies(PBVH *pbvh, int use_orco);

/* ---- learned Smooth (MLP over 1-ring features; BF16 MFMA inference) ----------
 * A Smooth dab with SG_SMOOTH_LEARNED in SgDabParams.flag goes through the model
 * attached to the PBVH; when its error estimate (alpha * RMS(|d|)/r) exceeds the
 * tolerance, or a vertex is outside the trained range, the dab is redone with the
 * plain brush (sg_dab_apply*, sg_stroke_launch handle this). */
typedef struct SgMlp SgMlp;
SgMlp *sg_mlp_new(uint64_t seed);
void sg_mlp_free(SgMlp *mlp);
int sg_mlp_get_params(SgMlp *mlp, float *out);           /* 6,211 FP32 master weights */
int sg_mlp_set_params(SgMlp *mlp, const float *w);
int sg_mlp_set_norm(SgMlp *mlp, const float *in_mu, const float *in_inv_sigma, const float *out_scale);
int sg_mlp_get_norm(SgMlp *mlp, float *out64);
int sg_mlp_set_decision(SgMlp *mlp, float alpha, float tol);
int sg_mlp_get_decision(SgMlp *mlp, float *out2);
int sg_mlp_set_data(SgMlp *mlp, const float *X, const float *T, int n);  /* rows of 28 / 3 */
int sg_mlp_gradient(SgMlp *mlp, int n, float *grad_out, float *loss_out);
int sg_mlp_train(SgMlp *mlp, int epochs, int batch_wg, float lr, uint64_t seed, float *loss_curve);
int sg_mlp_predict(SgMlp *mlp, const float *X, int n, float *Y);
int sg_mfma_probe(const uint16_t *A, const uint16_t *B, float *D);  /* 16x16 bf16 layout check */
int sg_pbvh_set_learned(PBVH *pbvh, SgMlp *mlp);
int sg_pbvh_learned_stats(PBVH *pbvh, long long out[2]);  /* accepted, fallbacks */
/* The plain brush as the judge of a learned Smooth dab, from the current state: runs
 * the learned dab, returns to the state, runs the plain dab (which stays applied).
 * out = {RMS |learned - plain| / r, RMS |plain - before| / r over the affected
 * vertices, 1 if the learned dab was accepted (0: it fell back to the plain brush),
 * affected count}. */
int sg_learned_judge(PBVH *pbvh, const SgDabParams *dab, double out[4]);
/* A plain Smooth dab recording features (X[i][28]) before and targets (T[i][3]) after
 * it for each affected vertex; ok[i] = 0 for vertices outside the model's range.
 * Returns the number of rows. */
int sg_learned_collect(PBVH *pbvh, const SgDabParams *dab, float *X, float *T, float *ok, int cap);

/* ---- many meshes, one dab per mesh per step (no counterpart in the reference) ----
 * The meshes are concatenated into one device arena and every kernel of a step covers
 * all of them; one captured hipGraph is launched per step.  Each mesh ends
 * bit-identical to the same dabs applied to it alone with sg_dab_apply_stroke.
 * Tools: Draw, Smooth (plain or learned), Inflate, Grab; preset falloff curves only. */
typedef struct SgBatch SgBatch;
SgBatch *sg_batch_new(int n_meshes, const float *const *co, const int *const *tris, const int *totvert,
                      const int *tottri, int leaf_limit);
void sg_batch_free(SgBatch *batch);
/* out[0..4] = meshes, vertices, tris, leaves, search cuts; then the first vertex of
 * each mesh in the arena (n_meshes + 1 entries) while out_n allows. */
int sg_batch_info(SgBatch *batch, long long *out, int out_n);
/* dabs[step * n_meshes + mesh], sculpt_tool 0 = no dab for that mesh in that step.
 * load uploads the table and captures the step graph; launch runs the next `steps`
 * rows (flags bit 0: enqueue kernels instead of the graph, bit 1: do not wait);
 * rewind restarts at row 0; run = load + launch of all rows. */
int sg_batch_load(SgBatch *batch, const SgDabParams *dabs, int steps);
int sg_batch_launch(SgBatch *batch, int steps, int flags);
int sg_batch_rewind(SgBatch *batch);
int sg_batch_run(SgBatch *batch, const SgDabParams *dabs, int steps, int flags);
int sg_batch_sync(SgBatch *batch);
/* Totals since the last reset: [0] gathered leaves, [1] affected vertices, [2] box
 * tests, [3] candidate vertices, [4] vertex->tri degrees of affected vertices, [5]
 * Smooth neighbour reads, [6] shared vertices refit, [7] area-normal vertices, [8]
 * plain Smooth vertices times k, [9] learned Smooth vertices, [12] learned dabs
 * accepted, [13] rejected (redone by the plain brush in the same step), [14] steps. */
int sg_batch_totals(SgBatch *batch, long long out[16], int reset);
int sg_batch_read_positions(SgBatch *batch, int mesh, float *out);
int sg_batch_read_normals(SgBatch *batch, int mesh, float *out);
int sg_batch_snapshot_save(SgBatch *batch);
int sg_batch_snapshot_restore(SgBatch *batch);
/* launch shape {search, brush, smooth, post, learned-brush workgroups, brush threads
 * per block (64/128/256), brush vertices per lane (1/2)}; <= 0 keeps a value */
int sg_batch_set_launch(SgBatch *batch, const int g[7]);
int sg_batch_get_launch(SgBatch *batch, int g[7]);
/* Autotuner over the launch shape on a sample of steps (restored afterwards). */
int sg_batch_autotune(SgBatch *batch, const SgDabParams *dabs, int steps, double out_ms[2]);
int sg_batch_set_learned(SgBatch *batch, SgMlp *mlp);
/* Persistent step on or off (default: on for batches of at least 128 meshes; the
 * autotuner times both): graph launches (sg_batch_launch flags bit 0 clear) of a table
 * without learned Smooth dabs run every row in one kernel, a workgroup per mesh
 * advancing its own dabs with workgroup barriers only; otherwise one captured step
 * graph per row.  Both give the same bits.  Returns whether the batch fits the
 * persistent step (every mesh at most 1024 search cuts). */
int sg_batch_set_persistent(SgBatch *batch, int on);
int sg_batch_timer_start(SgBatch *batch);
int sg_batch_timer_stop(SgBatch *batch, float *ms);

/* ---- 3. multires grids (PBVH_GRIDS) ------------------------------------------------
 * The CCG types live in headers the reference does not ship (BKE_ccg.h,
 * BKE_subdiv_ccg.h); their fields follow the reference's uses: the key
 * (kernel/intern/subdiv_ccg.c:633-646), coordinates (:368-379), neighbour sets
 * (:1365-1380), faces (:312-323).  An element is co float[3], then the mask float
 * (has_mask), then the normal float[3] (has_normals) (:62-90); a grid is grid_size^2
 * elements, row-major (x fastest).  In a face corner's grid, (0,0) is the face centre,
 * (grid_size-1, grid_size-1) the corner's coarse vertex, the column x = grid_size-1 runs
 * along the corner's edge and the row y = grid_size-1 along the previous corner's
 * edge (:397-468, :951-984). */
typedef struct CCGElem CCGElem;
typedef struct CCGKey {
  int level;
  int elem_size;
  int grid_size;
  int grid_area;
  int grid_bytes;
  int normal_offset;
  int mask_offset;
  int has_normals;
  int has_mask;
} CCGKey;
typedef struct DMFlagMat { /* grid_materials_match (pbvh.c:145-148) reads flag and mat_nr */
  short mat_nr;
  char flag;
} DMFlagMat;
typedef unsigned int BLI_bitmap; /* one bit per grid element, 32 per word */
typedef struct CCGFace CCGFace;  /* a SubdivCCGFace, as gridfaces hands it around */
typedef struct SubdivCCGFace {
  int num_grids;
  int start_grid_index;
} SubdivCCGFace;
typedef struct SubdivCCGCoord {
  int grid_index;
  short x, y;
} SubdivCCGCoord;
typedef struct SubdivCCGNeighbors {
  SubdivCCGCoord *coords; /* coords_fixed, or malloc'ed when size >= 256 */
  int size;
  int num_duplicates;
  SubdivCCGCoord coords_fixed[256];
} SubdivCCGNeighbors;

/* Grids of a coarse polygon mesh at subdivision `level` (grid_size = 2^(level-1) + 1,
 * subdiv_inline.h:26-29): a grid per polygon corner (subdiv_ccg.c:104-140), the face,
 * edge and vertex neighbourhoods (:312-530), normals and mask layers on request.  The
 * reference evaluates elements on the OpenSubdiv limit surface (KERNEL_subdiv_to_ccg,
 * :541-562; OpenSubdiv is a third-party dependency it does not vendor); here each
 * grid is the bilinear patch over (face centre, the corner's edge midpoint, the
 * corner vertex, the previous edge's midpoint), after which the grids are averaged
 * (KERNEL_subdiv_ccg_average_grids) so every duplicated element holds the same bits,
 * and normals computed (KERNEL_subdiv_ccg_recalc_normals).  Edges are numbered in
 * order of first use by the polygons, each oriented as first used; a vertex's edges
 * are listed by ascending edge number.  NULL on bad input (text in sg_last_error). */
SubdivCCG *sg_subdiv_ccg_new(const float (*coarse_co)[3], int totvert, const int *poly_offsets,
                             const int *poly_verts, int totpoly, int level, int need_normal, int need_mask);
/* subdiv_ccg.c:595 */
void KERNEL_subdiv_ccg_destroy(SubdivCCG *subdiv_ccg);
/* subdiv_ccg.c:648; paint.c:2126 calls it by the BKE_ name: same function */
void KERNEL_subdiv_ccg_key_top_level(CCGKey *key, const SubdivCCG *subdiv_ccg);
void BKE_subdiv_ccg_key_top_level(CCGKey *key, const SubdivCCG *subdiv_ccg);
/* subdiv_ccg.c:781 every grid's normals, then the grids averaged (host) */
void KERNEL_subdiv_ccg_recalc_normals(SubdivCCG *subdiv_ccg);
/* subdiv_ccg.c:1171 seams inside faces, then edges, then corners (host) */
void KERNEL_subdiv_ccg_average_grids(SubdivCCG *subdiv_ccg);
/* subdiv_ccg.c:847 normals of the given faces' grids and the seams around them (host;
 * a PBVH built on these grids runs the same step on the GPU in BKE_pbvh_update_normals) */
void KERNEL_subdiv_ccg_update_normals(SubdivCCG *subdiv_ccg, CCGFace **effected_faces, int num_effected_faces);
/* subdiv_ccg.c:1882 */
void KERNEL_subdiv_ccg_neighbor_coords_get(const SubdivCCG *subdiv_ccg, const SubdivCCGCoord *coord,
                                           bool include_duplicates, SubdivCCGNeighbors *r_neighbors);
/* frees r_neighbors->coords when it is not coords_fixed */
void sg_subdiv_ccg_neighbors_free(SubdivCCGNeighbors *neighbors);
/* subdiv_ccg.c:1911 */
int KERNEL_subdiv_ccg_grid_to_face_index(const SubdivCCG *subdiv_ccg, int grid_index);
/* The SubdivCCG fields the reference's callers read (paint.c passes them to
 * BKE_pbvh_build_grids): grids, num_grids, grid_faces (as void **gridfaces),
 * faces, grid_flag_mats, grid_hidden (one bitmap per grid, all clear at creation). */
CCGElem **sg_subdiv_ccg_grids(SubdivCCG *subdiv_ccg);
int sg_subdiv_ccg_num_grids(const SubdivCCG *subdiv_ccg);
void **sg_subdiv_ccg_grid_faces(SubdivCCG *subdiv_ccg);
SubdivCCGFace *sg_subdiv_ccg_faces(SubdivCCG *subdiv_ccg, int *r_num_faces);
DMFlagMat *sg_subdiv_ccg_flag_mats(SubdivCCG *subdiv_ccg);
BLI_bitmap **sg_subdiv_ccg_grid_hidden(SubdivCCG *subdiv_ccg);
/* Element copies, grid after grid, row-major: co / normals 3 floats, mask 1 float
 * (-1 when the layer is absent). */
int sg_subdiv_ccg_get_co(const SubdivCCG *subdiv_ccg, float *out);
int sg_subdiv_ccg_set_co(SubdivCCG *subdiv_ccg, const float *in);
int sg_subdiv_ccg_get_normals(const SubdivCCG *subdiv_ccg, float *out);
int sg_subdiv_ccg_get_mask(const SubdivCCG *subdiv_ccg, float *out);
int sg_subdiv_ccg_set_mask(SubdivCCG *subdiv_ccg, const float *in);

/* pbvh.c:598 BKE_pbvh_build_grids: the tree over grids exactly as the reference builds
 * it (a grid's box over all its elements, leaf_limit = max(limit / grid_size^2, 1) with
 * limit the handle's leaf limit, 10000 by default as pbvh.c:32, material split on
 * DMFlagMat, leaves fully hidden when every quad has a hidden corner), then the
 * elements uploaded to HBM leaf by leaf.  The grids stay the caller's (borrowed);
 * sg_pbvh_grids_sync_host writes the device state back into them.  Dabs, normals and
 * the grid-seam averaging also need the SubdivCCG the grids belong to:
 * DUNE_pbvh_subdiv_cgg_set. */
void BKE_pbvh_build_grids(PBVH *pbvh, CCGElem **grids, int totgrid, CCGKey *key, void **gridfaces,
                          DMFlagMat *flagmats, BLI_bitmap **grid_hidden);
/* pbvh.c:4973 (uploads the face, edge, vertex and neighbour tables of the grids);
 * paint.c:1680 calls it by the BKE_ name: same function */
void DUNE_pbvh_subdiv_cgg_set(PBVH *pbvh, SubdivCCG *subdiv_ccg);
void BKE_pbvh_subdiv_cgg_set(PBVH *pbvh, SubdivCCG *subdiv_ccg);
/* pbvh.c:1865 */
void BKE_pbvh_node_get_grids(PBVH *pbvh, PBVHNode *node, int **r_grid_indices, int *r_totgrid, int *r_maxgrid,
                             int *r_gridsize, CCGElem ***r_griddata);
/* pbvh.c:1679 / :1685 / :1691 / :1703 / :1709 */
BLI_bitmap **BKE_pbvh_grid_hidden(const PBVH *pbvh);
const CCGKey *BKE_pbvh_get_grid_key(const PBVH *pbvh);
CCGElem **BKE_pbvh_get_grids(const PBVH *pbvh);
int BKE_pbvh_get_grid_num_vertices(const PBVH *pbvh);
int BKE_pbvh_get_grid_num_faces(const PBVH *pbvh);
/* pbvh.c:1605: faces (SubdivCCGFace*) of the grids of UpdateNormals nodes, each once,
 * in order of first appearance; malloc'ed, caller frees. */
void BKE_pbvh_get_grid_updates(PBVH *pbvh, bool clear, void ***r_gridfaces, int *r_totface);
/* pbvh.c:331 (paint_is_grid_face_hidden, absent from the reference, taken as upstream
 * Blender's: a quad is hidden when any of its four corners is) */
int BKE_pbvh_count_grid_quads(BLI_bitmap **grid_hidden, const int *grid_indices, int totgrid, int gridsize);
/* Device state -> the caller's grids (co, normals, mask), and back after host edits. */
int sg_pbvh_grids_sync_host(PBVH *pbvh);
int sg_pbvh_grids_upload(PBVH *pbvh);

/* ---- 4. dynamic topology (PBVH_BMESH) ----------------------------------------------
 * The reference builds this variant over a BMesh (kernel/intern/pbvh_bmesh.c), whose
 * library is outside the path.  SgDynMesh stands in for it: a triangle mesh, manifold
 * with boundary, held as a corner table with room to grow (vert_cap, tri_cap).  Once
 * a PBVH is built over it, the PBVH's HBM copy is the mesh; sg_dynmesh_read fetches
 * it.  Vertices carry 64-bit labels (the input index, or a hash of the split edge's
 * labels for vertices a split creates) so callers can follow them across topology
 * changes.  Topology updates run on the GPU in rounds of non-conflicting edge splits
 * or collapses, each edge treated by the reference's rules (pbvh_bmesh.c:773-1484).
 * A leaf holds at most SG_DYN_SLAB faces between DUNE_pbvh_bmesh_after_stroke calls. */
typedef struct SgDynMesh SgDynMesh;
enum { SG_DYN_SLAB = 4096 };
/* New in this build (the BMesh it stands for is built by the absent bmesh library):
 * co[totvert][3], tris[tottri][3] (outward winding), optional mask[totvert],
 * vert_hidden[totvert], face_hidden[tottri]; caps >= the counts (0: 4x).  NULL on a
 * non-manifold or degenerate mesh (sg_last_error). */
SgDynMesh *sg_dynmesh_new(const float *co, int totvert, const int *tris, int tottri, const float *mask,
                          const uint8_t *vert_hidden, const uint8_t *face_hidden, int vert_cap, int tri_cap);
void sg_dynmesh_free(SgDynMesh *bm);
/* Live vertex and face counts, and the slots in use. */
int sg_dynmesh_counts(SgDynMesh *bm, int *r_totvert, int *r_tottri, int *r_vert_slots, int *r_tri_slots);
/* Live vertices in slot order (label, position, normal, mask, owner node) and live
 * faces in slot order (corner labels, leaf node, face normal, hidden); any output may
 * be NULL.  Returns the vertex count. */
int sg_dynmesh_read(SgDynMesh *bm, uint64_t *vert_labels, float *co, float *no, float *mask, int *vert_node,
                    uint64_t *face_labels, int *face_node, float *face_no, uint8_t *face_hidden);
/* pbvh_bmesh.c:1854-1929 DUNE_pbvh_build_bmesh: leaf limit 100 faces, detail size
 * 0.75.  The reference's BMLog and custom-data offsets are accepted and unused (node
 * indices live in the PBVH). */
void DUNE_pbvh_build_bmesh(PBVH *pbvh, SgDynMesh *bm, bool smooth_shading, void *log, int cd_vert_node_offset,
                           int cd_face_node_offset);
void BKE_pbvh_build_bmesh(PBVH *pbvh, SgDynMesh *bm, bool smooth_shading, void *log, int cd_vert_node_offset,
                          int cd_face_node_offset);
/* pbvh_bmesh.c:1931-2005: over the leaves flagged PBVH_UpdateTopology. */
bool DUNE_pbvh_bmesh_update_topology(PBVH *pbvh, PBVHTopologyUpdateMode mode, const float center[3],
                                     const float view_normal[3], float radius, bool use_frontface,
                                     bool use_projected);
bool BKE_pbvh_bmesh_update_topology(PBVH *pbvh, PBVHTopologyUpdateMode mode, const float center[3],
                                    const float view_normal[3], float radius, bool use_frontface, bool use_projected);
/* pbvh_bmesh.c:2065-2078 (leaf lists compacted, leaves above the limit split) */
void DUNE_pbvh_bmesh_after_stroke(PBVH *pbvh);
void BKE_pbvh_bmesh_after_stroke(PBVH *pbvh);
/* pbvh_bmesh.c:2080-2084 */
void DUNE_pbvh_bmesh_detail_size_set(PBVH *pbvh, float detail_size);
void BKE_pbvh_bmesh_detail_size_set(PBVH *pbvh, float detail_size);
/* pbvh_bmesh.c:2086-2089 */
void DUNE_pbvh_node_mark_topology_update(PBVHNode *node);
void BKE_pbvh_node_mark_topology_update(PBVHNode *node);
/* pbvh_bmesh.c:2091-2104 return GSets; here the node's face slots / unique and other
 * vertex slots are written to out (up to cap), returning the count. */
int sg_pbvh_bmesh_node_faces(PBVHNode *node, int *out, int cap);
int sg_pbvh_bmesh_node_unique_verts(PBVHNode *node, int *out, int cap);
int sg_pbvh_bmesh_node_other_verts(PBVHNode *node, int *out, int cap);
/* New: what a dab on a PBVH_BMESH does before its brush (the reference's sculpt code
 * that calls update_topology is absent): Draw, Inflate and Smooth dabs update the
 * topology of the leaves within 1.25 r with this mode (0: none), view normal and
 * front-face / projected options; Grab keeps the topology. */
int sg_pbvh_bmesh_topology_set(PBVH *pbvh, int mode, const float view_normal[3], int use_frontface, int use_projected);
/* Counters of the last dab or topology update on a PBVH_BMESH: [0] gathered leaves
 * [1] affected verts [2] splits [3] collapses [4] split rounds [5] collapse rounds
 * [6] split candidates [7] collapse candidates [8] 1: capacity stop, 2: walk stack
 * [9] leaves updated [10] faces re-normalled [11] leaves refit. */
int sg_pbvh_bmesh_last_stats(PBVH *pbvh, long long out[12]);
/* Affected vertex labels of the last dab, ascending. */
int sg_pbvh_bmesh_last_affected(PBVH *pbvh, uint64_t *out, int cap);
/* Node i: flag, children_offset, live faces; current and original box (6 each). */
int sg_pbvh_bmesh_node_info(PBVH *pbvh, int node, int out[3], float vb[6], float orig_vb[6]);

/* ---- synthetic meshes and strokes (deterministic; for tests and benchmarks) ---- */
int sg_mesh_icosphere_size(int level, int *r_totvert, int *r_tottri);
int sg_mesh_icosphere(int level, float radius, float *co, int *tris);
int sg_mesh_torus(int nu, int nv, float major_r, float minor_r, float *co, int *tris);
/* co += dir * amplitude * fbm(co * frequency); value noise (lattice hash, trilinear
 * with smoothstep), octaves with lacunarity 2 and gain 0.5, normalised to [-1,1]. */
int sg_mesh_noise_displace(float *co, const float *dir, int totvert, uint32_t seed,
                           float amplitude, float frequency, int octaves);

#ifdef __cplusplus
}
#endif
#endif /* SCULPT_GPU_H */
