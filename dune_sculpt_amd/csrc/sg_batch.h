// Many independent meshes stepped together (one dab per mesh per step).
//
// The meshes are concatenated into one device arena: every per-vertex, per-tri,
// per-leaf and per-cut array of the single-mesh layout, with indices offset to the
// arena, so a batch is a forest of PBVHs over one disconnected mesh.  Each kernel of a
// step covers all meshes at once: work items (cuts, gathered leaves) carry their
// mesh, whose dab constants come from a [step][mesh] table in HBM.
// Affected vertices are not listed: the brush stamps each with the step's serial and
// later kernels walk the gathered leaves' vertex ranges testing the stamp, so no
// kernel reserves list slots with atomics.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "sg_internal.h"
#include "sg_kernels.h"

// Step counters (one set, cleared by the step-end kernel).  Counters that every wave
// would bump are kept per mesh instead: returning atomics on one address serialise
// across the XCDs (~12 ns each), per-mesh addresses spread them out.
enum {
  SB_NGATHER = 0,   // gathered leaves (all meshes)
  SB_TESTS = 1,     // cut + leaf box tests (one add per search wave)
  SB_NPC = 2,       // listed (passing) cuts
  SB_NGL = 3,       // gathered leaves of learned-Smooth meshes (the learned brush's list)
  SB_COUNT = 4,
};
// per-mesh counters, SB_PM entries per mesh after the global ones
enum {
  SBM_AREA_X = 0, SBM_AREA_Y = 1, SBM_AREA_Z = 2,  // Draw area normal, 2^-32 fixed point
  SBM_LERR = 3,     // learned Smooth: sum (|d|/r)^2, 2^-32 fixed point
  SBM_LOOD = 4,     // learned Smooth: vertices outside the trained range
  SBM_LNAFF = 5,    // learned Smooth: affected vertices
  SBM_NAFF = 6,     // affected vertices
  SBM_CAND = 7,     // uniq vertices of gathered leaves
  SBM_DEGT = 8,     // vertex->tri degrees of affected vertices
  SBM_DEG = 9,      // vertex->vertex degrees of Smooth affected vertices (one step)
  SBM_SHARED = 10,  // shared vertices read by the leaf refit
  SBM_NAREA = 11,   // vertices inside the Draw area-normal sphere
  SB_PM = 16,       // 128 bytes per mesh
};
// running totals (step-end kernel)
enum {
  SBT_GATHER = 0, SBT_AFF = 1, SBT_TESTS = 2, SBT_CAND = 3, SBT_DEGT = 4, SBT_NBR = 5, SBT_SHARED = 6,
  SBT_NAREA = 7, SBT_AK = 8, SBT_ALRN = 9, SBT_LACC = 12, SBT_LREJ = 13, SBT_STEPS = 14, SBT_COUNT = 16,
};

// kernel-side view of a batch
struct SgBatchView {
  SgDevView d;              // arena arrays (global ids)
  int M;                    // meshes
  const int2 *leaf_vr;      // [L] uniq vertex range of each leaf
  const int2 *leaf_shr;     // [L] shared vertex range (into leaf_sh)
  const int *cut_mesh;      // [C]
  int *gl_mesh;             // [L] mesh of each gathered entry
  int *vstamp;              // [V] serial of the last step that affected the vertex
  float *vfade;             // [V] Smooth factor of that step
  int4 *pc_list;            // [C] cuts whose box passed this step's sphere test: cut, leaf range, mesh | orig << 30
  float4 *pc_sph;           // [C] their dab sphere (centre, r^2)
  int *gl_lrn;              // [L] positions in the gathered list of leaves of learned-Smooth meshes
  unsigned *done;           // [1] blocks of the step-end kernel that finished
  long long *ctr;           // [SB_COUNT + SB_PM * M]
  long long *tot;           // [SBT_COUNT]
  int *step;                // [2]: serial (monotonic, for the search's touch test), table row
  const SgDabK *tab;        // [rows][M]
  float alpha, tol;         // learned decision
};

struct SgBatch {
  int M = 0;
  hipStream_t stream = nullptr;
  SgDev d;
  int2 *leaf_vr = nullptr, *leaf_shr = nullptr;
  int *cut_mesh = nullptr, *gl_mesh = nullptr, *vstamp = nullptr;
  int4 *pc_list = nullptr;
  float4 *pc_sph = nullptr;
  int *gl_lrn = nullptr;
  float *vfade = nullptr;
  unsigned *done = nullptr;
  long long *ctr = nullptr, *tot = nullptr;
  int *step = nullptr;
  SgDabK *tab = nullptr;
  size_t tab_cap = 0;       // rows allocated
  int rows = 0, row_next = 0;  // rows loaded, next row to run
  int maxk = 0;             // largest Smooth iteration count of the loaded table
  bool learned = false;     // the loaded table has learned Smooth dabs
  bool odd_k = false;       // ... Smooth dabs with an odd iteration count
  std::vector<long long> vbase, lbase, cbase, tbase;  // [M+1]
  struct SgMlp *mlp = nullptr;
  // launch shape (autotuned): workgroups of search / brush / smooth+post / learned brush,
  // threads per block of the brush, vertices per lane of the brush
  int g_search = 1024, g_brush = 1024, g_smooth = 1024, g_post = 1024, g_learned = 1024, g_end = 256;
  int brush_threads = 256, brush_vpl = 1;
  // captured step graph and the configuration it was captured for
  hipGraphExec_t exec = nullptr;
  long long exec_key = -1;
  // checkpoint
  std::vector<void *> snap;
  bool snap_valid = false;
  hipEvent_t timer[2] = {nullptr, nullptr};
  float *gather_buf = nullptr;
  size_t gather_cap = 0;
  // persistent step (one workgroup per mesh runs every row in one launch): per-mesh
  // first cut / leaf / vertex, the running dab's affected list and gathered-leaf
  // candidate offsets, an error flag
  int *mesh_cut = nullptr, *mesh_leaf = nullptr, *mesh_vert = nullptr;
  int *pm_aff_v = nullptr, *pm_g_off = nullptr, *pm_err = nullptr;
  float *pm_aff_f = nullptr;
  bool pm_ok = false;       // every mesh fits the persistent step (at most 1024 search cuts)
  bool persistent = false;  // graph launches of tables without learned dabs use it (default: >= 128 meshes)
};

// A learned Smooth dab of mesh m is rejected (the plain brush redoes it in the same
// step) when a vertex left the model's range or the error estimate exceeds tolerance:
// alpha * RMS(|d| / r) > tol, from the per-mesh sums of the learned brush.
__device__ inline bool sb_learned_rejected(const SgBatchView &b, int mesh) {
  const long long *mc = b.ctr + SB_COUNT + SB_PM * mesh;
  const long long na = mc[SBM_LNAFF];
  const double rms = na > 0 ? sqrt((double)mc[SBM_LERR] / 4294967296.0 / (double)na) : 0.0;
  return !(mc[SBM_LOOD] == 0 && (double)b.alpha * rms <= (double)b.tol);
}

SgBatchView sg_batch_view(SgBatch *b);
// enqueue one step of the batch; maxk = largest Smooth iteration count in the run,
// learned = the run has learned Smooth dabs (model set)
int sg_batch_enqueue_step(SgBatch *b, int maxk, bool learned);
// enqueue rows [row0, row0 + steps) as one persistent launch (batch_mesh.hip)
int sg_batch_enqueue_persistent(SgBatch *b, int row0, int steps);
int sg_batch_enqueue_learned_brush(SgBatch *b, const SgBatchView &v);
int sg_batch_enqueue_decide(SgBatch *b, const SgBatchView &v);
