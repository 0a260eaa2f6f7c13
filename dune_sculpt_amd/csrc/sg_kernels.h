// Kernel-side views and launch entry points shared by the host API and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "sg_internal.h"

enum {
  SG_CTR_R1X = 11,    // distinct non-affected 1-ring vertices read (stats mode)
  SG_CTR_FA = 12,     // distinct tris incident to the affected set (stats mode)
  SG_CTR_VFA = 13,    // distinct vertices of those tris (stats mode)
  SG_CTR_SHARED = 14, // shared vertices read by the leaf refit
  SG_CTR_LERR = 15,   // learned Smooth: sum (|d|/r)^2 in 2^-40 fixed point
  SG_CTR_LOOD = 16,   // learned Smooth: vertices outside the trained range
  SG_CTR_LACC = 17,   // learned Smooth: 1 when the dab was accepted
};

// High half of the gathered-leaf counter: an undo step is open.  k_search sets it for
// the brush kernels, which load that counter anyway (a load of its own would add a
// dependent memory round trip to their start); the count is the low half.
#define SG_NGATHER_UNDO (1ull << 32)

// pair collection for the learned Smooth (device buffers)
struct SgCollect {
  float *X, *frame, *T;
  int cap;
};

// Undo recording (undo.hip).  Lives as long as the PBVH so captured strokes can read it
// at run time: while `active`, the first kernel that writes a leaf in the open step
// copies the leaf's uniq positions and normals as they were before the step, and its
// two boxes, into the recording arena.
struct SgUndoCtl {
  int active, step, overflow, n_e;  // n_e: entries used
  unsigned long long n_v;           // vertices used
  unsigned long long cap_v;
  int cap_e, pad;
  int *stamp;   // [L] step that saved the leaf
  int *slot;    // [L] its entry in tab (valid when stamp == step)
  float *pos;   // [cap_v * 3]
  float *nor;   // [cap_v * 3]
  int4 *tab;    // [cap_e]: leaf, count, vertex offset low, high
  float *bbs;   // [cap_e * 16]: vb, orig_vb (8 floats each)
};

// Plain-pointer view of the device mesh passed to kernels by value.
struct SgDevView {
  int V, T, L, N, C, respect_hide;
  float *px, *py, *pz, *qx, *qy, *qz, *nx, *ny, *nz, *mask;
  const int4 *vt4;
  const int *vt_off, *vv_off, *vv_idx;
  // polygons (null for triangle meshes): corner offsets, corner vertices, polygon of each
  // tri, each vertex's (polygon, corner) list
  const int *poly_off, *poly_v, *poly_of_tri, *vc_off;
  const int2 *vc;
  const int *leaf_vbeg, *leaf_sh_off, *leaf_sh, *leaf_flags, *leaf_cut;
  float *leaf_bb, *leaf_obb, *node_bb, *node_obb, *cut_bb, *cut_obb;
  const int *cut_node, *cut_lbeg, *cut_lend, *cut_flags, *sub_off;
  const int4 *sub_tab, *top_tab;
  int *cut_dirty;
  const int *vhide;
  uint8_t *vmark;
  int *gathered;
  int *draw_dirty;
  SgUndoCtl *undo;
  int *place_pair;  // handoff of a Grab placement's two rays (ray.hip)
  int2 *gl_v, *gl_s;
  int *aff_v;  // gathered leaf g: slots (g << aff_shift) + [0, aff_cnt[g])
  int2 *aff_vt, *aff_vv;
  float *aff_fade;
  int *aff_cnt;
  int aff_shift;
  int *prev_cuts;
  int *cut_touch;
  int *halt;
  int *seq, *vstamp;
  long long *ctr, *ctr_next;
  const int *vperm_old;
  int *stamp_v, *stamp_v2, *stamp_t, *stamp_a;
  // fused dabs: per-leaf halos, the gathered leaves' halo headers, gather stamps, dab
  // stamps, the normals the fused kernel wrote and its pending record
  const int4 *hl_rec;
  const int2 *hl_v;
  const int *hl_nboff;
  const unsigned short *hl_nbl, *hl_sh;
  const unsigned short *hl_ti;
  int4 *gl_rec;
  int *leaf_gstamp, *dstamp, *pend;
  float *mx, *my, *mz;
};

// Per-dab constants (spec S1-S9), computed on the host in the same FP32 order as the
// CPU statement of the spec.
struct SgDabK {
  float cx, cy, cz, r, r2, nr2, bstr, hardness;
  float gx, gy, gz;
  int tool, falloff, iters, use_orig, stats, learned;
  int orig_co;  // Grab from the original coordinates (SG_DAB_ORIGINAL)
  int hide_skip;  // 1: the search skips fully hidden subtrees and leaves; 0: every leaf (elastic deform)
  int etype;      // elastic deform mode
  float epois, enx, eny, enz, eforce;  // Poisson ratio, sculpt normal, force (Kelvinlet parameters)
  int fused;      // the search fills the halo headers and gather stamps for the fused kernel
};

// kernel ids for the per-kernel profile
enum { SG_K_SEARCH = 0, SG_K_AREA = 1, SG_K_BRUSH = 2, SG_K_SMOOTH = 3, SG_K_POST = 4, SG_K_FLUSH_SUB = 5,
       SG_K_FLUSH_TOP = 6, SG_K_COUNT = 7 };

SgDevView sg_view(PBVH *p);
SgDevView sg_view_of(SgDev &d, bool respect_hide, bool halt_live);
int sg_enqueue_full_normals_view(const SgDevView &v, hipStream_t s);
int sg_enqueue_gather_raw(hipStream_t s, const float *x, const float *y, const float *z, const int *perm_old, int V,
                          float *out_dev);
int sg_profile_collect(PBVH *p);
SgDabK sg_make_dab(const SgDabParams *dab, int stats);
int sg_enqueue_dab(PBVH *p, const SgDabK &k, const float *curve_dev, long long *log, int log_index,
                   int stamp, int serial, const SgCollect *col = nullptr, const float *place = nullptr);
int sg_enqueue_clear_ctr(PBVH *p);
int sg_enqueue_flush(PBVH *p);
int sg_enqueue_log_row(PBVH *p, long long *log, int row);
int sg_enqueue_full_normals(PBVH *p);
int sg_enqueue_refit_list(PBVH *p, const int *leaves_dev, int n);
int sg_enqueue_normals_list(PBVH *p, const int *leaves_dev, int n);
int sg_enqueue_coords_apply(PBVH *p, const float *src_dev);
int sg_enqueue_leaf_boxes(PBVH *p, const int *leaves_dev, int n, float *out_dev);
int sg_enqueue_scatter_boxes(PBVH *p, const void *recs_dev, int n);
int sg_enqueue_set_marks(PBVH *p, const int *verts_dev, int n);
int sg_enqueue_finish(PBVH *p, int serial_last);
int sg_enqueue_gather(PBVH *p, int which, float *out_dev);
const char *sg_last_error_str();
