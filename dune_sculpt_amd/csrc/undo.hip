// Undo steps recorded on the GPU: snapshots of the leaves a stroke touches.
//
// The reference's sculpt undo (sculpt_undo.c) is absent; its call sites fix the shape:
// SCULPT_undo_push_begin / SCULPT_undo_push_node / SCULPT_undo_push_end around an
// operation (editors/sculpt_paint/paint_hide.c:347-376, :78, :126, :255), a node pushed
// at most once per step with its state before the operation.  Here the brushes push
// themselves: while a step is open, the first brush kernel that writes a leaf copies
// the leaf's uniq positions and normals and its current and original boxes into a
// recording arena in HBM (undo_save_leaf); a vertex is affected only through the leaf
// that owns it and every gathered leaf is saved, so the step holds exactly what the
// operation changed.  Closing the step copies the used part into an exact-size
// allocation.  Undo and redo swap a step's records with the mesh (so the same records
// serve both directions), then mark the touched subtrees for the bounds refit and the
// draw-buffer refill.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "sg_internal.h"
#include "sg_kernels.h"
#include "sg_device.h"

// One wave per record: swap positions (and the shadow copy), normals and the leaf box.
// keep != 0 copies the record into the mesh and leaves the record as it is.
__global__ __launch_bounds__(256) void k_undo_swap(SgDevView d, const int4 *tab, float *pos, float *nor, float *bbs,
                                                   int n_e, int keep) {
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  for (int e = gw; e < n_e; e += nw) {
    const int4 t = tab[e];
    const int leaf = t.x, n = t.y;
    const unsigned long long vb = ((unsigned long long)(unsigned)t.w << 32) | (unsigned)t.z;
    const int v0 = d.leaf_vbeg[leaf];
    for (int i = lane; i < n; i += SG_WAVE) {
      const int v = v0 + i;
      const size_t o = 3 * (size_t)(vb + i);
      const float sx = pos[o], sy = pos[o + 1], sz = pos[o + 2];
      const float tx = nor[o], ty = nor[o + 1], tz = nor[o + 2];
      if (!keep) {
        pos[o] = d.px[v];
        pos[o + 1] = d.py[v];
        pos[o + 2] = d.pz[v];
        nor[o] = d.nx[v];
        nor[o + 1] = d.ny[v];
        nor[o + 2] = d.nz[v];
      }
      d.px[v] = sx;
      d.py[v] = sy;
      d.pz[v] = sz;
      d.qx[v] = sx;
      d.qy[v] = sy;
      d.qz[v] = sz;
      d.nx[v] = tx;
      d.ny[v] = ty;
      d.nz[v] = tz;
    }
    if (lane < 8) {
      // the restored box is also the original box (the leaf is updated with
      // PBVH_UpdateBB | PBVH_UpdateOriginalBB, BKE_pbvh_node_mark_update kernel/intern/pbvh.c:1723-1727,
      // which update_node_vb turns into orig_vb = vb, pbvh.c:1221-1223)
      float *cur = d.leaf_bb + 8 * (size_t)leaf + lane;
      const float saved = bbs[16 * (size_t)e + lane];
      if (!keep) bbs[16 * (size_t)e + lane] = *cur;
      *cur = saved;
      d.leaf_obb[8 * (size_t)leaf + lane] = saved;
    }
    if (lane == 0) {
      const int c = d.leaf_cut[leaf];
      d.cut_dirty[c] = 1;   // the deferred refit rebuilds the subtree and its box
      d.draw_dirty[c] = 1;  // the draw records of its leaves are stale
    }
  }
}

// Original data of the open step: the saved leaves' positions and normals over the
// current ones (copied in first), leaf-major SoA.
__global__ __launch_bounds__(256) void k_undo_orig(SgDevView d, const int4 *tab, const float *pos, const float *nor,
                                                   int n_e, float *ox, float *oy, float *oz, float *onx, float *ony,
                                                   float *onz) {
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  for (int e = gw; e < n_e; e += nw) {
    const int4 t = tab[e];
    const unsigned long long vb = ((unsigned long long)(unsigned)t.w << 32) | (unsigned)t.z;
    const int v0 = d.leaf_vbeg[t.x];
    for (int i = lane; i < t.y; i += SG_WAVE) {
      const size_t o = 3 * (size_t)(vb + i);
      if (ox) {
        ox[v0 + i] = pos[o];
        oy[v0 + i] = pos[o + 1];
        oz[v0 + i] = pos[o + 2];
      }
      if (onx) {
        onx[v0 + i] = nor[o];
        ony[v0 + i] = nor[o + 1];
        onz[v0 + i] = nor[o + 2];
      }
    }
  }
}

// Host-side push of one leaf (SCULPT_undo_push_node for host edits).
__global__ void k_undo_push_leaf(SgDevView d, int leaf, int2 vr) {
  if (undo_active(d)) undo_save_leaf(d, leaf, vr, d.px, d.py, d.pz);
}

// ---------------------------------------------------------------------------

static void free_step(SgUndoStep &s) {
  void *q[] = {s.pos, s.nor, s.bbs, s.tab};
  for (void *x : q)
    if (x) (void)hipFree(x);
  s = SgUndoStep();
}

static void free_arena(SgUndoState &u) {
  void *q[] = {u.rec_pos, u.rec_nor, u.rec_bbs, u.rec_tab};
  for (void *x : q)
    if (x) (void)hipFree(x);
  u.rec_pos = u.rec_nor = u.rec_bbs = nullptr;
  u.rec_tab = nullptr;
  u.rec_cap_v = 0;
  u.rec_cap_e = 0;
}

void sg_undo_release(PBVH *p) {
  SgUndoState &u = p->undo;
  for (SgUndoStep &s : u.steps) free_step(s);
  u.steps.clear();
  u.cursor = 0;
  u.recording = false;
  free_arena(u);
}

static int write_ctl(PBVH *p, bool active, int step) {
  SgUndoState &u = p->undo;
  SgUndoCtl c;
  memset(&c, 0, sizeof(c));
  c.active = active ? 1 : 0;
  c.step = step;
  c.cap_v = (unsigned long long)std::min(u.rec_cap_v, u.limit_v);  // the open step's limit
  c.cap_e = u.rec_cap_e;
  c.stamp = p->d.undo_stamp;
  c.slot = p->d.undo_slot;
  c.pos = u.rec_pos;
  c.nor = u.rec_nor;
  c.tab = u.rec_tab;
  c.bbs = u.rec_bbs;
  if (int rc = sg_check_hip(hipMemcpyAsync(p->d.undo, &c, sizeof(c), hipMemcpyHostToDevice, p->stream), "copy"))
    return rc;
  return sg_check_hip(hipStreamSynchronize(p->stream), "undo control");
}

static int swap_step(PBVH *p, SgUndoStep &s, int keep = 0) {
  if (s.n_e > 0) {
    const int blocks = std::max(1, std::min((s.n_e + 3) / 4, 8192));
    hipLaunchKernelGGL(k_undo_swap, dim3(blocks), dim3(256), 0, p->stream, sg_view(p), s.tab, s.pos, s.nor, s.bbs,
                       s.n_e, keep);
    if (int rc = sg_check_hip(hipGetLastError(), "undo launch")) return rc;
  }
  if (int rc = sg_enqueue_flush(p)) return rc;  // internal nodes and subtree boxes from the leaf boxes
  p->nodes_dirty_deferred = false;
  p->host_bb_stale = p->host_verts_stale = p->host_normals_stale = true;
  return sg_check_hip(hipStreamSynchronize(p->stream), "undo");
}

extern "C" {

int sg_undo_begin(PBVH *p, long long max_verts) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: undo steps of PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p || p->no_device || !p->build_ok) return sg_set_error("no device mesh");
  SgUndoState &u = p->undo;
  if (u.recording) return sg_set_error("an undo step is already open");
  if (max_verts <= 0) max_verts = p->d.V;
  // a new step drops the steps undone so far (no redo past a new operation)
  for (size_t i = u.cursor; i < u.steps.size(); i++) free_step(u.steps[i]);
  u.steps.resize(u.cursor);
  if (max_verts > u.rec_cap_v || p->d.L > u.rec_cap_e) {
    free_arena(u);
    const int cap_e = p->d.L;  // every leaf at most once
    if (sg_check_hip(hipMalloc(&u.rec_pos, (size_t)max_verts * 3 * sizeof(float)), "hipMalloc") ||
        sg_check_hip(hipMalloc(&u.rec_nor, (size_t)max_verts * 3 * sizeof(float)), "hipMalloc") ||
        sg_check_hip(hipMalloc(&u.rec_bbs, (size_t)cap_e * 16 * sizeof(float)), "hipMalloc") ||
        sg_check_hip(hipMalloc(&u.rec_tab, (size_t)cap_e * sizeof(int4)), "hipMalloc")) {
      free_arena(u);
      return -1;
    }
    u.rec_cap_v = max_verts;  // grow-only: a larger arena serves smaller limits
    u.rec_cap_e = cap_e;
  }
  u.limit_v = max_verts;
  u.step_id++;
  if (int rc = write_ctl(p, true, u.step_id)) return rc;
  u.recording = true;
  return 0;
}

int sg_undo_push_node(PBVH *p, PBVHNode *node) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: undo steps of PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p || !node || node < p->nodes.data() || node >= p->nodes.data() + p->nodes.size())
    return sg_set_error("not a node of this PBVH");
  if (node->leaf < 0 || node->leaf >= p->d.L) return sg_set_error("not a leaf");
  if (!p->undo.recording) return sg_set_error("no undo step is open");
  int2 vr;
  if (int rc = sg_check_hip(hipMemcpy(&vr, p->d.leaf_vbeg + node->leaf, sizeof(int2), hipMemcpyDeviceToHost), "copy"))
    return rc;
  hipLaunchKernelGGL(k_undo_push_leaf, dim3(1), dim3(64), 0, p->stream, sg_view(p), node->leaf, vr);
  if (int rc = sg_check_hip(hipGetLastError(), "undo push launch")) return rc;
  return sg_check_hip(hipStreamSynchronize(p->stream), "undo push");
}

// Closes the step: its records move to an exact-size allocation.  Fails (and keeps no
// step) when the arena overflowed.
int sg_undo_end(PBVH *p) {
  if (!p) return sg_set_error("bad arguments");
  SgUndoState &u = p->undo;
  if (!u.recording) return sg_set_error("no undo step is open");
  SgUndoCtl c;
  int rc = sg_check_hip(hipMemcpyAsync(&c, p->d.undo, sizeof(c), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = write_ctl(p, false, u.step_id);
  u.recording = false;
  if (rc) return rc;
  if (c.overflow) return sg_set_error("undo step overflowed its arena (%lld vertices): step dropped", u.limit_v);
  SgUndoStep s;
  s.n_e = c.n_e;
  s.n_v = (long long)c.n_v;
  const size_t nv = (size_t)std::max<long long>(s.n_v, 1), ne = (size_t)std::max(s.n_e, 1);
  rc = sg_check_hip(hipMalloc(&s.pos, nv * 3 * sizeof(float)), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&s.nor, nv * 3 * sizeof(float)), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&s.bbs, ne * 16 * sizeof(float)), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&s.tab, ne * sizeof(int4)), "hipMalloc");
  if (!rc && s.n_v > 0) {
    rc = sg_check_hip(hipMemcpyAsync(s.pos, u.rec_pos, (size_t)s.n_v * 3 * sizeof(float), hipMemcpyDeviceToDevice,
                                     p->stream),
                      "copy");
    if (!rc)
      rc = sg_check_hip(hipMemcpyAsync(s.nor, u.rec_nor, (size_t)s.n_v * 3 * sizeof(float), hipMemcpyDeviceToDevice,
                                       p->stream),
                        "copy");
  }
  if (!rc && s.n_e > 0) {
    rc = sg_check_hip(
        hipMemcpyAsync(s.bbs, u.rec_bbs, (size_t)s.n_e * 16 * sizeof(float), hipMemcpyDeviceToDevice, p->stream), "copy");
    if (!rc)
      rc = sg_check_hip(hipMemcpyAsync(s.tab, u.rec_tab, (size_t)s.n_e * sizeof(int4), hipMemcpyDeviceToDevice, p->stream),
                        "copy");
  }
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "undo end");
  if (rc) {
    free_step(s);
    return rc;
  }
  u.steps.push_back(s);
  u.cursor = (int)u.steps.size();
  return 0;
}

// Cancels the open step: its leaves go back to their state at sg_undo_begin and the
// step is dropped (a cancelled stroke).  Fails, changing nothing, when the arena
// overflowed (some leaves were not saved).
int sg_undo_cancel(PBVH *p) {
  if (!p) return sg_set_error("bad arguments");
  SgUndoState &u = p->undo;
  if (!u.recording) return sg_set_error("no undo step is open");
  SgUndoCtl c;
  int rc = sg_check_hip(hipMemcpyAsync(&c, p->d.undo, sizeof(c), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = write_ctl(p, false, u.step_id);
  u.recording = false;
  if (rc) return rc;
  if (c.overflow) return sg_set_error("undo step overflowed its arena: cannot cancel");
  SgUndoStep s;  // the arena itself, not owned
  s.pos = u.rec_pos;
  s.nor = u.rec_nor;
  s.bbs = u.rec_bbs;
  s.tab = u.rec_tab;
  s.n_e = c.n_e;
  s.n_v = (long long)c.n_v;
  return swap_step(p, s);
}

// Puts the open step's saved leaves back to their state at sg_undo_begin and keeps the
// step open with its records: what the sculpt code (absent from the reference) does
// before each dab of an anchored stroke, restoring the original mesh from its undo
// nodes; a later dab of the step saves nothing again.
int sg_undo_restore_original(PBVH *p) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: undo steps of PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p) return sg_set_error("bad arguments");
  SgUndoState &u = p->undo;
  if (!u.recording) return sg_set_error("no undo step is open");
  SgUndoCtl c;
  int rc = sg_check_hip(hipMemcpyAsync(&c, p->d.undo, sizeof(c), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "undo restore");
  if (rc) return rc;
  if (c.overflow) return sg_set_error("undo step overflowed its arena: original data incomplete");
  SgUndoStep s;  // the arena itself, not owned
  s.pos = u.rec_pos;
  s.nor = u.rec_nor;
  s.bbs = u.rec_bbs;
  s.tab = u.rec_tab;
  s.n_e = c.n_e;
  s.n_v = (long long)c.n_v;
  return swap_step(p, s, 1);
}

// Positions and normals at the begin of the open step (the current ones when no step
// is open), AoS in the caller's vertex order; either output may be null.  The original
// coordinates the reference reads from its undo nodes (SculptOrigVertData) for Grab,
// anchored strokes and BKE_pbvh_node_raycast's origco.
int sg_undo_original(PBVH *p, float *co, float *no) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: undo steps of PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p || p->no_device || !p->build_ok) return sg_set_error("no device mesh");
  SgDev &d = p->d;
  const size_t V = (size_t)d.V;
  SgUndoCtl c;
  memset(&c, 0, sizeof(c));
  if (p->undo.recording) {
    if (int rc = sg_check_hip(hipMemcpyAsync(&c, d.undo, sizeof(c), hipMemcpyDeviceToHost, p->stream), "copy"))
      return rc;
    if (int rc = sg_check_hip(hipStreamSynchronize(p->stream), "undo original")) return rc;
    if (c.overflow) return sg_set_error("undo step overflowed its arena: original data incomplete");
  }
  float *soa = nullptr, *aos = nullptr;
  int rc = sg_check_hip(hipMalloc(&soa, V * 6 * sizeof(float)), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&aos, V * 3 * sizeof(float)), "hipMalloc");
  float *o[6] = {soa, soa + V, soa + 2 * V, soa + 3 * V, soa + 4 * V, soa + 5 * V};
  const float *cur[6] = {d.px, d.py, d.pz, d.nx, d.ny, d.nz};
  for (int k = 0; k < 6 && !rc; k++)
    rc = sg_check_hip(hipMemcpyAsync(o[k], cur[k], V * sizeof(float), hipMemcpyDeviceToDevice, p->stream), "copy");
  if (!rc && c.n_e > 0) {
    const int blocks = std::max(1, std::min((c.n_e + 3) / 4, 8192));
    hipLaunchKernelGGL(k_undo_orig, dim3(blocks), dim3(256), 0, p->stream, sg_view(p), p->undo.rec_tab,
                       p->undo.rec_pos, p->undo.rec_nor, c.n_e, o[0], o[1], o[2], o[3], o[4], o[5]);
    rc = sg_check_hip(hipGetLastError(), "undo original launch");
  }
  for (int k = 0; k < 2 && !rc; k++) {
    float *out = k == 0 ? co : no;
    if (!out) continue;
    rc = sg_enqueue_gather_raw(p->stream, o[3 * k], o[3 * k + 1], o[3 * k + 2], d.vperm_old, d.V, aos);
    if (!rc)
      rc = sg_check_hip(hipMemcpyAsync(out, aos, V * 3 * sizeof(float), hipMemcpyDeviceToHost, p->stream), "copy");
    if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "undo original");
  }
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "undo original");
  if (soa) (void)hipFree(soa);
  if (aos) (void)hipFree(aos);
  return rc;
}

// Returns 0 after undoing a step, 1 when there is nothing to undo.
int sg_undo(PBVH *p) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: undo steps of PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p) return sg_set_error("bad arguments");
  SgUndoState &u = p->undo;
  if (u.recording) return sg_set_error("an undo step is open");
  if (u.cursor == 0) return 1;
  if (int rc = swap_step(p, u.steps[u.cursor - 1])) return rc;
  u.cursor--;
  return 0;
}

int sg_redo(PBVH *p) {
  if (!p) return sg_set_error("bad arguments");
  SgUndoState &u = p->undo;
  if (u.recording) return sg_set_error("an undo step is open");
  if (u.cursor >= (int)u.steps.size()) return 1;
  if (int rc = swap_step(p, u.steps[u.cursor])) return rc;
  u.cursor++;
  return 0;
}

// out: steps held, steps applied, bytes held on the device, leaves in the newest step.
int sg_undo_info(PBVH *p, long long out[4]) {
  if (!p || !out) return sg_set_error("bad arguments");
  const SgUndoState &u = p->undo;
  long long bytes = 0;
  for (const SgUndoStep &s : u.steps) bytes += s.n_v * 24 + (long long)s.n_e * (16 * 4 + 16);
  out[0] = (long long)u.steps.size();
  out[1] = u.cursor;
  out[2] = bytes;
  out[3] = u.steps.empty() ? 0 : u.steps.back().n_e;
  return 0;
}

void sg_undo_clear(PBVH *p) {
  if (!p) return;
  if (p->undo.recording) write_ctl(p, false, p->undo.step_id);
  sg_undo_release(p);
}

}  // extern "C"
