// Internal structures of the sculpt engine (host side + device buffer handles).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sculpt_gpu.h"

#define SG_HALO_RINGS 4  // Smooth(k) runs fused for k <= this

// Cut of the tree used by the two-level sphere search: maximal subtrees holding at
// most cut_leaf_cap leaves.  Leaves are numbered in DFS (left->right) order, so every
// cut subtree owns a contiguous leaf range.
struct SgCut {
  std::vector<int> node;       // node id of each cut subtree root (DFS order)
  std::vector<int> lbeg, lend; // leaf range [lbeg, lend)
};

// Host mirror of one tree node (the PBVHNode handle handed to callers).
struct PBVHNode {
  float vb[6];      // bmin xyz, bmax xyz
  float orig_vb[6];
  int children_offset;  // -1 for leaves
  int flag;
  int index;            // slot in nodes[]
  int leaf;             // DFS leaf number or -1
  int prim_off, totprim;
  int uniq_verts, face_verts;
  int *vert_indices;    // original vertex ids: uniq first, then shared
  int (*face_vert_indices)[3];
  PBVHNode *self_check;
  float tmin;           // set by the occluded searches (BKE_pbvh_node_get_tmin)
  int proxy_count;
  PBVHProxyNode *proxies;
  struct PBVH *owner;
};

// Device-side mesh and tree (all HBM, SoA, leaf-ordered vertex numbering).
struct SgUndoCtl;  // device control block of the undo recording (sg_kernels.h)
// One recorded undo step (device arrays, exact size) and the recording arena.
struct SgUndoStep {
  float *pos = nullptr, *nor = nullptr, *bbs = nullptr;
  int4 *tab = nullptr;
  int n_e = 0;
  long long n_v = 0;
};
struct SgUndoState {
  std::vector<SgUndoStep> steps;
  int cursor = 0;               // steps [0, cursor) are applied
  bool recording = false;
  int step_id = 0;
  float *rec_pos = nullptr, *rec_nor = nullptr, *rec_bbs = nullptr;
  int4 *rec_tab = nullptr;
  long long rec_cap_v = 0;
  int rec_cap_e = 0;
  long long limit_v = 0;        // the open step's max_verts
};

struct SgDev {
  int V = 0, T = 0, L = 0, N = 0, C = 0;
  long long nuniq = 0, nuniq_visible = 0;  // uniq vertices of the leaves, and the visible ones
  int nhidden_leaves = 0;                   // fully hidden leaves (fixed at build)
  float *px = nullptr, *py = nullptr, *pz = nullptr; // positions [V]
  float *qx = nullptr, *qy = nullptr, *qz = nullptr; // shadow copy (== positions between dabs)
  float *nx = nullptr, *ny = nullptr, *nz = nullptr; // normals [V]
  float *mask = nullptr;                             // [V] or null
  int *vt_off = nullptr;        // [V+1] vertex -> incident tris, ascending original tri id
  int4 *vt4 = nullptr;          // [3T] per incidence: tri corners (new ids), .w = tri | corner << 30
  int *vv_off = nullptr, *vv_idx = nullptr; // vertex -> edge neighbours, asc. orig id
  int *leaf_vbeg = nullptr;     // [L+1] uniq vertex range of each leaf
  int *leaf_sh_off = nullptr;   // [L+1]
  int *leaf_sh = nullptr;       // shared (face) vertices of each leaf
  int *leaf_flags = nullptr;    // [L] bit0 fully hidden
  int *leaf_cut = nullptr;      // [L] cut index of each leaf
  float *leaf_bb = nullptr;     // [L*8] vb   (min xyz, 0, max xyz, 0)
  float *leaf_obb = nullptr;    // [L*8] orig_vb
  float *node_bb = nullptr;     // [N*8] internal nodes (and cut leaves)
  float *node_obb = nullptr;
  int *cut_node = nullptr, *cut_lbeg = nullptr, *cut_lend = nullptr; // [C]
  int *cut_flags = nullptr;     // [C] bit0: single hidden leaf
  float *cut_bb = nullptr, *cut_obb = nullptr; // [C*8] search boxes of the cut subtrees
  // deferred refit tables: per cut, its internal nodes by height (int4: node id,
  // local child 0, local child 1, height) and the top tree above the cut
  int *sub_off = nullptr;       // [C+1]
  int4 *sub_tab = nullptr;
  int *top_tab_n = nullptr;
  int4 *top_tab = nullptr;
  int top_n = 0, top_maxh = 0, sub_maxn = 0;
  int *cut_dirty = nullptr;     // [C]
  int *cut_touch = nullptr;     // [C] serial of the last dab that gathered a leaf of it
  int *halt = nullptr;          // [4] set by a rejected learned dab: its serial + 1
  int *seq = nullptr;           // [1] dab counter (each dab's search adds one): the learned dab's stamp
  int *vstamp = nullptr;        // [V] stamp of the last learned dab that affected the vertex
  // polygons of a mesh that has any non-triangle polygon (null for triangle meshes):
  // corner offsets and corner vertices (new ids), the polygon of each tri (leaf order),
  // and each vertex's polygon corners (polygon, corner) in ascending polygon order
  int *poly_off = nullptr, *poly_v = nullptr, *poly_of_tri = nullptr, *vc_off = nullptr;
  int2 *vc = nullptr;
  int *vhide = nullptr;         // [V] or null
  uint8_t *vmark = nullptr;     // [V] host-driven normal update marks
  // per-dab scratch
  int *gathered = nullptr;      // [L] gathered leaves (ascending)
  int2 *gl_v = nullptr, *gl_s = nullptr; // [L] their uniq / shared vertex ranges
  // affected vertices of the running dab: gathered leaf g owns the slots
  // [g << aff_shift, (g << aff_shift) + aff_cnt[g]) (2^aff_shift >= its uniq vertices)
  int *aff_v = nullptr;         // [L << aff_shift]
  int2 *aff_vt = nullptr, *aff_vv = nullptr; // their vertex->tri / vertex->vertex ranges
  float *aff_fade = nullptr;    // Smooth factor
  int *aff_cnt = nullptr;       // [L] affected vertices per gathered leaf
  int aff_shift = 6;
  int *prev_cuts = nullptr;     // [C] cuts refit by the previous dab (+ count in ctr)
  long long *ctr = nullptr;     // two counter sets (even / odd dabs), see SG_CTR_*
  int *err = nullptr;           // device error word
  int *vperm_old = nullptr;     // [V] new -> old id (device copy)
  int *stamp_v = nullptr, *stamp_v2 = nullptr, *stamp_t = nullptr, *stamp_a = nullptr; // stats mode
  float *curve = nullptr;       // custom falloff tables for queued dabs
  int curve_cap = 0;
  long long *log = nullptr;     // per-dab counter log for strokes
  int log_cap = 0;
  // ray cast (built on first use): tris in leaf order (new corner ids, .w = looptri |
  // hidden << 31), each leaf's range, looptri -> poly, tree links
  int4 *ltri = nullptr;
  int *leaf_toff = nullptr, *tri_poly = nullptr, *node_parent = nullptr, *leaf_node = nullptr;
  SgRay *ray_in = nullptr;      // staging for host-side ray lists
  SgRayHit *ray_out = nullptr;
  int ray_cap = 0;
  int2 *rl_list = nullptr;      // [L] leaves an occluded search passes (callback API)
  float *rl_args = nullptr;     // [64] one ray's arguments, results and list count
  // draw buffers (built on request): visible tris per leaf, their vertex records
  int *draw_dirty = nullptr;    // [C] subtrees whose leaves' draw records are stale (set by the search; lives with the PBVH)
  int *draw_tri = nullptr;      // [visible tris] position in ltri, leaf after leaf
  int *draw_off = nullptr;      // [L+1] each leaf's first visible tri
  int *draw_leaf = nullptr;     // [L] bit0 smooth (first face), bits 8+ material
  int *draw_fset = nullptr;     // [polys] face sets (copy)
  int *draw_poly_off = nullptr, *draw_poly_v = nullptr;  // polygon corners (new ids)
  uint8_t *draw_vbo = nullptr;  // 36 bytes per vertex, 3 per visible tri
  long long draw_bytes = 0;
  int draw_ntri = 0, draw_seed = 0, draw_fset_default = 0;
  SgUndoCtl *undo = nullptr;    // undo recording control (lives with the PBVH)
  int *undo_stamp = nullptr;    // [L]
  int *undo_slot = nullptr;     // [L]
  float *place = nullptr;       // [8 per dab] placements of a ray-placed stroke
  int *place_pair = nullptr;    // [16] a Grab placement's two rays meet here (count, then results)
  SgRay *place_rays = nullptr;  // the stroke's rays
  int place_cap = 0;
  // proxy combine staging (grow-only): pinned host pack of heap-allocated proxies,
  // device copy of all proxies, leaf table, per-proxy offsets
  float *prox_host = nullptr;   // pinned
  size_t prox_host_cap = 0;     // floats
  float *prox_dev = nullptr;
  size_t prox_cap = 0;          // floats
  int4 *prox_tab = nullptr;
  int *prox_leaves = nullptr;
  int prox_tab_cap = 0;
  long long *prox_off = nullptr;  // float offset of each proxy in prox_dev
  int prox_off_cap = 0;
  // grow-only staging for the host-driven entry points (lists, records, coordinate
  // uploads and readbacks): allocated on first use, reused after
  void *stage = nullptr;
  size_t stage_cap = 0;         // bytes
  // Per-leaf halos of the fused dab kernel (triangle meshes; see sg_build_halo).  A
  // leaf's local vertices are its uniq range, then ring 0 (its shared vertices and the
  // edge neighbours of its uniq vertices), then rings 1..SG_HALO_RINGS around those.
  bool halo_ok = false;
  int halo_maxloc = 0, halo_maxloc0 = 0;   // most local vertices of a leaf: all rings / ring 0
  // [2L] leaf record: {uniq begin, shared begin, ext offset, neighbour-offset block
  // start}, {nu | ns << 16, e0 | e1 << 16, e2 | e3 << 16, e4} with e_j the ext count at
  // the end of ring j (every count below 65536)
  int4 *hl_rec = nullptr;
  int2 *hl_v = nullptr;         // [ext] vertex id, owner leaf (the leaf holding it as uniq)
  int *hl_nboff = nullptr;      // per leaf: offsets into hl_nbl of its locals before the last ring (+1)
  unsigned short *hl_nbl = nullptr;  // local ids of those locals' edge neighbours, vertex->vertex order
  unsigned short *hl_ti = nullptr;   // [3T] parallel to vt4: the incidence's face as a local face index of
                                     // its vertex's leaf; the leaf's faces (corners as local ids, c0 | c1 << 16,
                                     // c2) follow its neighbour offsets in hl_nboff
  unsigned short *hl_sh = nullptr;   // parallel to leaf_sh: local id of each shared vertex
  int *leaf_gstamp = nullptr;   // [L] dab stamp of the last search that gathered the leaf
  int *dstamp = nullptr;        // [2] dab stamps by counter-set parity
  float *mx = nullptr, *my = nullptr, *mz = nullptr;  // [V] normals the fused kernel wrote, copied in by the next search
  int *pend = nullptr;          // [4] the fused dab whose results wait for that copy: gathered count, stamp
  int4 *gl_rec = nullptr;                 // [2L] the gathered leaves' records (fused dabs)
  int halo_maxnbl = 0, leaf_maxsh = 0;    // most neighbour entries / shared vertices of a leaf
  int halo_maxlt = 0;                     // most faces incident to a leaf's uniq vertices
  int halo_maxlocr[SG_HALO_RINGS + 1] = {};  // most locals of a leaf up to ring j (uniq + rings 0..j)
  long long halo_bytes = 0;               // device bytes of the halo tables
};

// Pinned host chunk the proxies are carved from: the caller fills them in place and
// the combine uploads each chunk's used part in one copy.
struct SgProxChunk {
  float *host;
  size_t cap, used;  // floats
};

enum {
  SG_CTR_NGATHER = 0,
  SG_CTR_NAFF = 1,
  SG_CTR_AREA_X = 2,
  SG_CTR_AREA_Y = 3,
  SG_CTR_AREA_Z = 4,
  SG_CTR_NPREV = 5,
  SG_CTR_TESTS = 6,
  SG_CTR_NCAND = 7,
  SG_CTR_NAREA = 8,
  SG_CTR_NBR = 9,
  SG_CTR_TRIS = 10,
  SG_CTR_COUNT = 24,
};

struct SgGridDev;
struct SubdivCCG;
struct SgDynState;
struct PBVH {
  PBVHType type = PBVH_FACES;
  // PBVH_GRIDS (pbvh_intern.h:129-134, :161): the caller's grids (borrowed), their key,
  // faces, material flags and hidden bitmaps, the SubdivCCG they belong to, and the
  // device state
  CCGElem **grids = nullptr;
  CCGKey gridkey{};
  void **gridfaces = nullptr;
  DMFlagMat *grid_flag_mats = nullptr;
  BLI_bitmap **grid_hidden = nullptr;
  int totgrid = 0;
  SubdivCCG *subdiv_ccg = nullptr;
  SgGridDev *gd = nullptr;
  SgDynState *dyn = nullptr;         // PBVH_BMESH (sg_dyn.h)
  int leaf_limit = 10000;
  bool respect_hide = true;
  bool no_device = false;            // no HIP device: host tree only
  bool build_ok = false;             // device layout built and within kernel limits
  int totvert = 0, totprim = 0;
  std::vector<PBVHNode> nodes;
  std::vector<int> prim_indices;
  std::vector<int> leaf_node;        // DFS leaf -> node
  std::vector<int> perm_new_of_old;  // vertex renumbering
  std::vector<int> perm_old_of_new;
  std::vector<int> tris;             // original corner vertex ids [T*3]
  std::vector<int> tri_mat;          // material key per tri (empty = one material)
  std::vector<int> poly_off, poly_verts;  // polygons: corner lists (original vertex ids)
  std::vector<uint8_t> poly_smooth;       // ME_SMOOTH per polygon
  std::vector<int> poly_mat;              // mat_nr per polygon
  bool poly_mode = false;                 // some polygon is not the triangle of its looptri: polygon normal rules
  std::vector<uint8_t> vhide;
  SgCut cut;
  MVert *host_verts = nullptr;       // caller's array (borrowed; duplicated on deform)
  bool verts_owned = false;
  MLoopTri *looptri = nullptr;       // owned (freed at BKE_pbvh_free)
  std::vector<float> host_normals;   // [V*3] original order view
  std::vector<uint8_t> host_vmark;   // BKE_pbvh_vert_mark_update marks
  std::vector<int> host_mark_list;   // the marked vertices (original ids), each once
  std::vector<int> leaf_vbeg;        // [L+1] each leaf's uniq range in the new numbering
  bool host_bb_stale = false;
  bool host_verts_stale = false;
  bool host_normals_stale = false;
  bool nodes_dirty_deferred = false; // internal-node refit pending on the device
  hipStream_t stream = nullptr;
  SgDev d;
  // last dab
  int last_tool = 0;
  int64_t last_stats[8] = {0};
  long long last_ctr[24] = {0};
  struct SgMlp *mlp = nullptr;       // learned Smooth model (borrowed)
  long long learned_accepted = 0, learned_fallbacks = 0;
  bool last_learned = false;         // the last dab was an accepted learned dab (affected = stamped vertices)
  int dab_count = 0;
  int grid_wg = 256;                 // workgroups of the per-dab kernels (tunable)
  int grid_brush = 0, grid_post = 0, grid_smooth = 0;  // per-kernel overrides (autotuner)
  int grid_fused = 1024;             // workgroups of the fused dab kernel (a leaf each)
  bool fused_enabled = true;
  bool fused_pending = false;        // the last dab enqueued was fused and no finish followed yet
  // device snapshot (checkpoint) of the mutable mesh state
  // slot 0: the caller's (sg_pbvh_snapshot_save/restore); slot 1: internal (autotuner, judge)
  std::vector<void *> snap[2];
  bool snap_valid[2] = {false, false};
  float *judge_buf = nullptr;        // [3V] positions of the learned result (sg_learned_judge)
  double *judge_acc = nullptr;       // [2] device sums
  // per-kernel event timing (sg_pbvh_set_profile)
  bool profile = false;
  bool capturing = false;            // a stroke is being captured (profiling uses event nodes)
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<int> ev_kid;           // kernel id per recorded pair (events 2i, 2i+1)
  double prof_ms[16] = {0};
  long long prof_n[16] = {0};
  int stamp = 0;                     // stats-mode stamp generator
  hipEvent_t timer[2] = {nullptr, nullptr};  // sg_pbvh_timer_start/stop
  bool halt_live = false;            // the work being enqueued contains learned dabs (kernels test the halt flag)
  bool draw_marks_pending = false;   // some node carries a host-side draw-buffer mark
  std::vector<int> proxy_nodes;      // nodes given a proxy since the last combine (may repeat)
  std::vector<SgProxChunk> prox_chunks;  // proxy arena (pinned; reset once no proxy is left)
  SgUndoState undo;                  // undo steps recorded on the GPU (undo.hip)
  const float *last_place = nullptr; // placement slots of the last ray-placed stroke
  int last_place_n = 0;
};

// Host copies of the leaf-ordered device arrays of one mesh (pbvh_build.cpp).
struct SgLayout {
  int V = 0, T = 0, L = 0, N = 0;
  std::vector<float> x, y, z;                     // positions, new numbering
  std::vector<int4> vt4;                          // [3T]
  std::vector<int> vt_off, vv_off, vv_idx;        // [V+1], [V+1], neighbours
  std::vector<int> leaf_vbeg, leaf_sh_off, leaf_sh, leaf_flags;
  std::vector<float> leaf_bb, leaf_obb;           // [L*8]
  // polygon meshes only: corner vertices (new ids), each vertex's (polygon, corner)
  // list in ascending polygon order
  std::vector<int> poly_v, vc_off;
  std::vector<int2> vc;
};


// host build (pbvh_build.cpp)
void sg_build_tree(PBVH *p, const float *co, const int *tris, int T, int V);
void sg_build_layout(PBVH *p, const float *co, SgLayout &h);
void sg_build_cut(PBVH *p, std::vector<int> &lbeg, std::vector<int> &lend, std::vector<int> &height,
                  std::vector<int> &leaf_cut);
void sg_build_device(PBVH *p, const float *co);
void sg_free_device(PBVH *p);

// undo (undo.hip)
void sg_undo_release(PBVH *p);  // every step and the arena

// ray cast (ray.hip)
int sg_ray_ensure(PBVH *p);
int sg_sync_host_bbs(PBVH *p);  // api.cpp: host node bounds from HBM
int sg_enqueue_raycast(PBVH *p, const SgRay *rays_dev, int n, int original, int stale, SgRayHit *out_dev,
                       float *place_dev, int grab_pair, long long *stats_dev);
bool sg_ray_host_needs_refit(const SgRay *rays, int n);

// errors (kernels.hip)
int sg_set_error(const char *fmt, ...);
int sg_check_hip(hipError_t e, const char *what);
