// Many independent meshes stepped together (one dab per mesh per step).
//
// The meshes are concatenated into one device arena: every per-vertex, per-tri,
// per-leaf and per-cut array of the single-mesh layout, with indices offset to the
// arena, so a batch is a forest of PBVHs over one disconnected mesh.  Each kernel of a
// step covers all meshes at once: work items (cuts, gathered leaves, affected records)
// carry their mesh, whose dab constants come from a [step][mesh] table in HBM.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "sg_internal.h"
#include "sg_kernels.h"

// step counters (one set, cleared by the step-end kernel)
enum {
  SB_NGATHER = 0,   // gathered leaves (all meshes)
  SB_NAFF = 1,      // affected records
  SB_TESTS = 2,     // cut + leaf box tests
  SB_CAND = 3,      // uniq vertices of gathered leaves
  SB_DEGT = 4,      // sum of vertex->tri degrees of affected vertices
  SB_DEG = 5,       // sum of vertex->vertex degrees of Smooth records, times k
  SB_SHARED = 6,    // shared vertices read by the leaf refit
  SB_NAREA = 7,     // vertices inside the Draw area-normal sphere
  SB_AK = 8,        // plain Smooth records times their iteration count
  SB_ALRN = 9,      // learned Smooth records
  SB_COUNT = 12,
};
// per-mesh counters, SB_PM entries per mesh after the global ones
enum {
  SBM_AREA_X = 0, SBM_AREA_Y = 1, SBM_AREA_Z = 2,  // Draw area normal, 2^-32 fixed point
  SBM_LERR = 3,     // learned Smooth: sum (|d|/r)^2, 2^-32 fixed point
  SBM_LOOD = 4,     // learned Smooth: vertices outside the trained range
  SBM_LNAFF = 5,    // learned Smooth: affected vertices
  SB_PM = 8,
};
// running totals
// running totals: entries 0..SB_COUNT-1 accumulate the step counters of the same index
enum { SBT_LACC = 12, SBT_LREJ = 13, SBT_STEPS = 14, SBT_COUNT = 16 };

// kernel-side view of a batch
struct SgBatchView {
  SgDevView d;              // arena arrays (global ids)
  int M;                    // meshes
  const int2 *leaf_vr;      // [L] uniq vertex range of each leaf
  const int2 *leaf_shr;     // [L] shared vertex range (into leaf_sh)
  const int *cut_mesh;      // [C]
  int *gl_mesh;             // [L] mesh of each gathered entry
  int *aff_mesh;            // [V] mesh of each affected record
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
  int *cut_mesh = nullptr, *gl_mesh = nullptr, *aff_mesh = nullptr;
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
  int g_search = 1024, g_brush = 1024, g_smooth = 1024, g_post = 1024, g_learned = 512;
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
int sg_batch_enqueue_learned_brush(SgBatch *b, const SgBatchView &v);
int sg_batch_enqueue_decide(SgBatch *b, const SgBatchView &v);
