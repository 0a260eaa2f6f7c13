// C ABI: drop-in PBVH API (kernel/intern/pbvh.c) and the sg_* brush entry points.
// All mesh state lives in HBM; the host keeps the tree shape, node flags and lazily
// synchronised views (positions, normals, node bounds) for callers that read them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <cstring>

#include "sg_internal.h"
#include "sg_grids.h"
#include "sg_dyn.h"
#include "sg_kernels.h"
#include "sg_learned.h"

namespace {

bool ensure_stats_buffers(PBVH *p) {
  SgDev &d = p->d;
  if (d.stamp_v) return true;
  const size_t V = (size_t)std::max(d.V, 1), T = (size_t)std::max(d.T, 1);
  if (sg_check_hip(hipMalloc(&d.stamp_v, V * sizeof(int)), "hipMalloc")) return false;
  if (sg_check_hip(hipMalloc(&d.stamp_v2, V * sizeof(int)), "hipMalloc")) return false;
  if (sg_check_hip(hipMalloc(&d.stamp_t, T * sizeof(int)), "hipMalloc")) return false;
  if (sg_check_hip(hipMalloc(&d.stamp_a, V * sizeof(int)), "hipMalloc")) return false;
  hipMemsetAsync(d.stamp_a, 0, V * sizeof(int), p->stream);
  hipMemsetAsync(d.stamp_v, 0, V * sizeof(int), p->stream);
  hipMemsetAsync(d.stamp_v2, 0, V * sizeof(int), p->stream);
  hipMemsetAsync(d.stamp_t, 0, T * sizeof(int), p->stream);
  return true;
}

int ensure_curves(PBVH *p, int n) {
  SgDev &d = p->d;
  if (n <= d.curve_cap) return 0;
  if (d.curve) hipFree(d.curve);
  d.curve = nullptr;
  d.curve_cap = 0;
  if (sg_check_hip(hipMalloc(&d.curve, (size_t)n * (SG_CM_TABLE + 1) * sizeof(float)), "hipMalloc")) return -1;
  d.curve_cap = n;
  return 0;
}

int ensure_log(PBVH *p, int n) {
  SgDev &d = p->d;
  if (n <= d.log_cap) return 0;
  if (d.log) hipFree(d.log);
  d.log = nullptr;
  d.log_cap = 0;
  if (sg_check_hip(hipMalloc(&d.log, (size_t)n * SG_CTR_COUNT * sizeof(long long)), "hipMalloc")) return -1;
  d.log_cap = n;
  return 0;
}

void invalidate_host_views(PBVH *p) {
  p->host_bb_stale = true;
  p->host_verts_stale = true;
  p->host_normals_stale = true;
}

int sync_bbs(PBVH *p) {
  if (!p->host_bb_stale || p->nodes.empty() || p->no_device) return 0;
  if (p->type == PBVH_GRIDS) return p->gd ? sg_grids_sync_bbs(p) : 0;
  if (p->type == PBVH_BMESH) return sg_dyn_refresh(p);
  if (p->nodes_dirty_deferred) {
    if (sg_enqueue_flush(p)) return -1;
    p->nodes_dirty_deferred = false;
  }
  const SgDev &d = p->d;
  std::vector<float> lbb((size_t)d.L * 8), lobb((size_t)d.L * 8), nbb((size_t)d.N * 8), nobb((size_t)d.N * 8);
  hipMemcpyAsync(lbb.data(), d.leaf_bb, lbb.size() * 4, hipMemcpyDeviceToHost, p->stream);
  hipMemcpyAsync(lobb.data(), d.leaf_obb, lobb.size() * 4, hipMemcpyDeviceToHost, p->stream);
  hipMemcpyAsync(nbb.data(), d.node_bb, nbb.size() * 4, hipMemcpyDeviceToHost, p->stream);
  hipMemcpyAsync(nobb.data(), d.node_obb, nobb.size() * 4, hipMemcpyDeviceToHost, p->stream);
  if (sg_check_hip(hipStreamSynchronize(p->stream), "sync bounds")) return -1;
  for (PBVHNode &n : p->nodes) {
    const float *src = (n.leaf >= 0) ? &lbb[8 * (size_t)n.leaf] : &nbb[8 * (size_t)n.index];
    const float *osrc = (n.leaf >= 0) ? &lobb[8 * (size_t)n.leaf] : &nobb[8 * (size_t)n.index];
    for (int k = 0; k < 3; k++) {
      n.vb[k] = src[k];
      n.vb[3 + k] = src[4 + k];
      n.orig_vb[k] = osrc[k];
      n.orig_vb[3 + k] = osrc[4 + k];
    }
  }
  p->host_bb_stale = false;
  return 0;
}

// Grow-only device staging of the host-driven entry points: allocated on first use
// (and when a larger size is asked for), then reused, so those calls allocate nothing
// in steady state.  Freed with the PBVH.
void *stage(PBVH *p, size_t bytes) {
  SgDev &d = p->d;
  bytes = std::max<size_t>(bytes, 256);
  if (bytes <= d.stage_cap) return d.stage;
  if (d.stage) {
    hipStreamSynchronize(p->stream);
    hipFree(d.stage);
  }
  d.stage = nullptr;
  d.stage_cap = 0;
  const size_t cap = std::max(bytes, (size_t)1 << 20);
  if (sg_check_hip(hipMalloc(&d.stage, cap), "hipMalloc")) return nullptr;
  d.stage_cap = cap;
  return d.stage;
}

int read_aos(PBVH *p, int which, float *out) {
  if (p->no_device) return sg_set_error("no HIP device");
  const size_t n = 3 * (size_t)p->d.V;
  float *tmp = (float *)stage(p, n * sizeof(float));
  if (!tmp) return -1;
  int rc = sg_enqueue_gather(p, which, tmp);
  if (!rc) rc = sg_check_hip(hipMemcpyAsync(out, tmp, n * sizeof(float), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "sync");
  return rc;
}

int sync_verts(PBVH *p) {
  if (!p->host_verts || !p->host_verts_stale || p->no_device) return 0;
  std::vector<float> co(3 * (size_t)p->totvert);
  if (read_aos(p, 0, co.data())) return -1;
  for (int v = 0; v < p->totvert; v++) memcpy(p->host_verts[v].co, &co[3 * (size_t)v], 12);
  p->host_verts_stale = false;
  return 0;
}

int sync_normals(PBVH *p) {
  if (!p->host_normals_stale) return 0;
  p->host_normals.resize(3 * (size_t)p->totvert);
  if (read_aos(p, 1, p->host_normals.data())) return -1;
  p->host_normals_stale = false;
  return 0;
}

// Upload the host view of the vertices of the given leaves (edits made through
// BKE_pbvh_get_verts before the caller flagged the nodes).
int upload_host_leaf_verts(PBVH *p, const std::vector<int> &leaves) {
  if (!p->host_verts || p->host_verts_stale) return 0;
  std::vector<int> idx;
  for (int l : leaves) {
    const PBVHNode &n = p->nodes[p->leaf_node[l]];
    for (int k = 0; k < n.uniq_verts + n.face_verts; k++) idx.push_back(p->perm_new_of_old[n.vert_indices[k]]);
  }
  std::sort(idx.begin(), idx.end());
  idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
  // contiguous runs -> few copies
  size_t i = 0;
  std::vector<float> x, y, z;
  while (i < idx.size()) {
    size_t j = i;
    while (j + 1 < idx.size() && idx[j + 1] == idx[j] + 1) j++;
    const int b = idx[i], n = (int)(j - i + 1);
    x.resize(n);
    y.resize(n);
    z.resize(n);
    for (int q = 0; q < n; q++) {
      const float *co = p->host_verts[p->perm_old_of_new[b + q]].co;
      x[q] = co[0];
      y[q] = co[1];
      z[q] = co[2];
    }
    float *dst[6] = {p->d.px, p->d.py, p->d.pz, p->d.qx, p->d.qy, p->d.qz};
    const float *src[6] = {x.data(), y.data(), z.data(), x.data(), y.data(), z.data()};
    for (int a = 0; a < 6; a++)
      if (int rc = sg_check_hip(hipMemcpy(dst[a] + b, src[a], n * 4, hipMemcpyHostToDevice), "copy")) return rc;
    i = j + 1;
  }
  return 0;
}

std::vector<int> flagged_leaves(PBVH *p, int flag) {
  std::vector<int> out;
  for (size_t l = 0; l < p->leaf_node.size(); l++) {
    if (p->nodes[p->leaf_node[l]].flag & flag) out.push_back((int)l);
  }
  return out;
}

}  // namespace

int sg_sync_host_bbs(PBVH *p) { return sync_bbs(p); }

extern "C" {

const char *sg_last_error(void) { return sg_last_error_str(); }

PBVH *BKE_pbvh_new(void) {
  PBVH *p = new PBVH();
  p->respect_hide = true;  // pbvh.c:648
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    // No GPU: the handle can still build and expose the host tree (for inspection);
    // every call that needs the device reports an error.
    (void)hipGetLastError();
    p->no_device = true;
    return p;
  }
  if (sg_check_hip(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate")) {
    delete p;
    return nullptr;
  }
  return p;
}

void BKE_pbvh_free(PBVH *p) {
  if (!p) return;
  if (!p->no_device) hipStreamSynchronize(p->stream);
  for (PBVHNode &n : p->nodes) {
    if (n.flag & PBVH_Leaf) {
      free(n.vert_indices);
      free(n.face_vert_indices);
      if (n.proxies) BKE_pbvh_node_free_proxies(&n);
    }
  }
  for (SgProxChunk &c : p->prox_chunks) (void)hipHostFree(c.host);
  p->prox_chunks.clear();
  if (!p->no_device) sg_free_device(p);
  if (!p->no_device) sg_grids_free_device(p);
  sg_dyn_release(p);
  for (int sl = 0; sl < 2; sl++)
    for (void *q : p->snap[sl]) hipFree(q);
  if (p->judge_buf) hipFree(p->judge_buf);
  if (p->judge_acc) hipFree(p->judge_acc);
  if (p->verts_owned) free(p->host_verts);
  free(p->looptri);
  for (int i = 0; i < 2; i++)
    if (p->timer[i]) hipEventDestroy(p->timer[i]);
  if (!p->no_device) hipStreamDestroy(p->stream);
  delete p;
}

int sg_pbvh_leaf_limit_set(PBVH *p, int leaf_limit) {
  if (!p) return sg_set_error("null PBVH");
  p->leaf_limit = leaf_limit > 0 ? leaf_limit : 10000;
  return 0;
}

void DUNE_pbvh_respect_hide_set(PBVH *p, bool respect_hide) {
  if (p) p->respect_hide = respect_hide;
}

void BKE_pbvh_respect_hide_set(PBVH *p, bool respect_hide) { DUNE_pbvh_respect_hide_set(p, respect_hide); }

PBVHType BKE_pbvh_type(const PBVH *p) { return p ? p->type : PBVH_FACES; }

void BKE_pbvh_build_mesh(PBVH *p, Mesh *mesh, const MPoly *mpoly, const MLoop *mloop, MVert *verts,
                         int totvert, CustomData *vdata, CustomData *ldata, CustomData *pdata,
                         const MLoopTri *looptri, int looptri_num) {
  (void)mesh;
  (void)vdata;
  (void)ldata;
  (void)pdata;
  if (!p || !mloop || !verts || totvert <= 0 || looptri_num < 0) {
    sg_set_error("BKE_pbvh_build_mesh: invalid arguments");
    return;
  }
  if (!p->nodes.empty()) {
    for (PBVHNode &n : p->nodes) {
      if (n.flag & PBVH_Leaf) {
        free(n.vert_indices);
        free(n.face_vert_indices);
        if (n.proxies) BKE_pbvh_node_free_proxies(&n);
      }
    }
    for (SgProxChunk &c : p->prox_chunks) c.used = 0;
    if (!p->no_device) sg_free_device(p);
  }
  p->type = PBVH_FACES;
  p->totvert = totvert;
  p->totprim = looptri_num;
  p->host_verts = verts;
  p->verts_owned = false;
  p->looptri = const_cast<MLoopTri *>(looptri);
  p->tris.resize(3 * (size_t)looptri_num);
  for (int t = 0; t < looptri_num; t++) {
    for (int j = 0; j < 3; j++) p->tris[3 * (size_t)t + j] = (int)mloop[looptri[t].tri[j]].v;
  }
  // polygons as the draw buffers need them (corner lists, smooth flag, material);
  // without mpoly each looptri is its own flat polygon
  p->poly_off.clear();
  p->poly_verts.clear();
  p->poly_smooth.clear();
  p->poly_mat.clear();
  {
    int npoly = 0;
    for (int t = 0; t < looptri_num; t++) npoly = std::max(npoly, (int)looptri[t].poly + 1);
    p->poly_off.assign(npoly + 1, 0);
    p->poly_smooth.assign(npoly, 0);
    p->poly_mat.assign(npoly, 0);
    if (mpoly) {
      for (int q = 0; q < npoly; q++) {
        p->poly_off[q + 1] = p->poly_off[q] + mpoly[q].totloop;
        p->poly_smooth[q] = (mpoly[q].flag & ME_SMOOTH) ? 1 : 0;
        p->poly_mat[q] = mpoly[q].mat_nr;
      }
      p->poly_verts.resize(p->poly_off[npoly]);
      for (int q = 0; q < npoly; q++)
        for (int k = 0; k < mpoly[q].totloop; k++)
          p->poly_verts[p->poly_off[q] + k] = (int)mloop[mpoly[q].loopstart + k].v;
    } else {
      std::vector<int> first(npoly, -1);
      for (int t = 0; t < looptri_num; t++)
        if (first[looptri[t].poly] < 0) first[looptri[t].poly] = t;
      for (int q = 0; q < npoly; q++) p->poly_off[q + 1] = p->poly_off[q] + (first[q] >= 0 ? 3 : 0);
      p->poly_verts.resize(p->poly_off[npoly]);
      for (int q = 0; q < npoly; q++)
        if (first[q] >= 0)
          for (int k = 0; k < 3; k++) p->poly_verts[p->poly_off[q] + k] = p->tris[3 * (size_t)first[q] + k];
    }
  }
  // Polygon normal rules are needed when some looptri is not exactly its polygon's
  // three corners in loop order (quads, n-gons); a triangle mesh keeps the tri paths,
  // which give the same bits there.
  p->poly_mode = false;
  if (mpoly) {
    for (int t = 0; t < looptri_num && !p->poly_mode; t++) {
      const int q = (int)looptri[t].poly;
      const int b = p->poly_off[q];
      if (p->poly_off[q + 1] - b != 3) {
        p->poly_mode = true;
        break;
      }
      for (int j = 0; j < 3; j++)
        if (p->tris[3 * (size_t)t + j] != p->poly_verts[b + j]) p->poly_mode = true;
    }
  }
  p->tri_mat.clear();
  if (mpoly && looptri_num > 0) {
    std::vector<int> key(looptri_num);
    bool uniform = true;
    for (int t = 0; t < looptri_num; t++) {
      const MPoly &mp = mpoly[looptri[t].poly];
      key[t] = (int)(unsigned short)mp.mat_nr | ((mp.flag & ME_SMOOTH) << 16);  // face_materials_match
      if (key[t] != key[0]) uniform = false;
    }
    if (!uniform) p->tri_mat.swap(key);
  }
  p->vhide.clear();
  bool any_hidden = false;
  for (int v = 0; v < totvert && !any_hidden; v++) any_hidden = (verts[v].flag & ME_HIDE) != 0;
  if (any_hidden) {
    p->vhide.resize(totvert);
    for (int v = 0; v < totvert; v++) p->vhide[v] = (verts[v].flag & ME_HIDE) ? 1 : 0;
  }
  std::vector<float> co(3 * (size_t)totvert);
  for (int v = 0; v < totvert; v++) memcpy(&co[3 * (size_t)v], verts[v].co, 12);
  sg_build_tree(p, co.data(), p->tris.data(), looptri_num, totvert);
  if (p->no_device) {
    p->host_bb_stale = false;
    p->host_verts_stale = false;
    sg_set_error("no HIP device: tree built on the host only");
    return;
  }
  p->build_ok = false;
  sg_build_device(p, co.data());
  if (p->cut.node.size() > 1024) {
    sg_set_error("mesh too large for the two-level search (%zu cut subtrees)", p->cut.node.size());
    return;
  }
  if ((size_t)p->d.sub_maxn * 24 > 65536 || (size_t)(p->d.C + p->d.top_n) * 24 > 65536) {
    sg_set_error("refit tables exceed LDS");
    return;
  }
  p->build_ok = true;
  sg_enqueue_full_normals(p);
  hipStreamSynchronize(p->stream);
  p->host_vmark.assign(totvert, 0);
  p->host_mark_list.clear();
  invalidate_host_views(p);
  p->host_verts_stale = false;  // the caller's array is the source
}

void sg_dab_params_default(SgDabParams *d, int tool) {
  memset(d, 0, sizeof(*d));
  d->sculpt_tool = tool;
  d->curve_preset = SG_CURVE_SMOOTH;
  d->smooth_iterations = 1;
  d->radius = 0.1f;
  d->strength = 0.5f;
  d->hardness = 0.0f;
  d->normal_radius_factor = 0.5f;
  d->elastic_deform_volume_preservation = 0.4f;
  d->surface_normal[2] = 1.0f;
}

static int check_dab(PBVH *p, const SgDabParams *dab) {
  if (!p || !dab) return sg_set_error("null argument");
  if (p->no_device) return sg_set_error("no HIP device");
  if (p->nodes.empty() || !p->build_ok) return sg_set_error("PBVH has no usable mesh");
  const int t = dab->sculpt_tool;
  if (t != SG_TOOL_DRAW && t != SG_TOOL_SMOOTH && t != SG_TOOL_INFLATE && t != SG_TOOL_GRAB &&
      t != SG_TOOL_ELASTIC_DEFORM)
    return sg_set_error("unsupported sculpt_tool %d", t);
  if (t == SG_TOOL_ELASTIC_DEFORM &&
      (dab->elastic_deform_type < SG_ELASTIC_GRAB || dab->elastic_deform_type > SG_ELASTIC_TWIST))
    return sg_set_error("bad elastic_deform_type");
  if (t == SG_TOOL_ELASTIC_DEFORM && (dab->flag & (SG_SMOOTH_LEARNED | SG_DAB_ORIGINAL)))
    return sg_set_error("flags not valid for elastic deform");
  if (!(dab->radius > 0.0f)) return sg_set_error("radius must be > 0");
  if (dab->curve_preset < 0 || dab->curve_preset > SG_CURVE_SMOOTHER) return sg_set_error("bad curve_preset");
  if (t == SG_TOOL_SMOOTH && dab->smooth_iterations > 64) return sg_set_error("smooth_iterations > 64");
  return 0;
}

static bool g_stats = false;

int sg_pbvh_set_stats(PBVH *p, int on) {
  (void)p;
  g_stats = on != 0;
  return 0;
}

int sg_pbvh_set_grid(PBVH *p, int workgroups) {
  if (!p || workgroups < 1 || workgroups > 65536) return sg_set_error("bad grid");
  p->grid_wg = workgroups;
  p->grid_brush = p->grid_post = p->grid_smooth = 0;
  return 0;
}

int sg_pbvh_set_fused(PBVH *p, int on) {
  if (!p) return sg_set_error("null PBVH");
  p->fused_enabled = on != 0;
  return p->d.halo_ok ? 1 : 0;
}

int sg_pbvh_halo_info(PBVH *p, long long out[5]) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: halo tables");
  if (!p || !out) return sg_set_error("bad arguments");
  const SgDev &d = p->d;
  out[0] = d.halo_ok ? 1 : 0;
  out[1] = d.halo_maxloc;
  out[2] = d.halo_maxloc0;
  out[3] = d.halo_maxnbl;
  out[4] = d.halo_bytes;
  return 0;
}

int sg_pbvh_set_kernel_grids(PBVH *p, int brush, int post, int smooth) {
  if (!p) return sg_set_error("null PBVH");
  for (int g : {brush, post, smooth})
    if (g < 0 || g > 65536) return sg_set_error("bad grid");
  p->grid_brush = brush;
  p->grid_post = post;
  p->grid_smooth = smooth;
  return 0;
}

int sg_pbvh_get_kernel_grids(PBVH *p, int out[4]) {
  if (!p) return sg_set_error("null PBVH");
  out[0] = p->grid_wg;
  out[1] = p->grid_brush;
  out[2] = p->grid_post;
  out[3] = p->grid_smooth;
  return 0;
}

// ---- checkpoint: device copy of the mutable state (positions, shadow copy,
// normals, leaf / subtree / node boxes).  Restore returns the mesh to it exactly.

namespace {
struct SnapItem {
  void *dev;
  size_t bytes;
};
std::vector<SnapItem> snap_items(PBVH *p) {
  SgDev &d = p->d;
  const size_t V = (size_t)d.V, L = (size_t)d.L, N = (size_t)d.N, C = (size_t)d.C;
  return {{d.px, V * 4},       {d.py, V * 4},       {d.pz, V * 4},           {d.qx, V * 4},
          {d.qy, V * 4},       {d.qz, V * 4},       {d.nx, V * 4},           {d.ny, V * 4},
          {d.nz, V * 4},       {d.leaf_bb, L * 32}, {d.leaf_obb, L * 32},    {d.node_bb, N * 32},
          {d.node_obb, N * 32}, {d.cut_bb, C * 32}, {d.cut_obb, C * 32}};
}
}  // namespace

static int snap_save(PBVH *p, int slot) {
  if (!p || p->no_device || !p->build_ok) return sg_set_error("no usable PBVH");
  auto items = snap_items(p);
  std::vector<void *> &sn = p->snap[slot];
  if (sn.size() != items.size()) {
    for (void *q : sn) hipFree(q);
    sn.assign(items.size(), nullptr);
    for (size_t i = 0; i < items.size(); i++)
      if (sg_check_hip(hipMalloc(&sn[i], std::max<size_t>(items[i].bytes, 4)), "hipMalloc")) return -1;
  }
  if (p->nodes_dirty_deferred) sg_enqueue_flush(p);
  for (size_t i = 0; i < items.size(); i++)
    hipMemcpyAsync(sn[i], items[i].dev, items[i].bytes, hipMemcpyDeviceToDevice, p->stream);
  p->snap_valid[slot] = true;
  return sg_check_hip(hipStreamSynchronize(p->stream), "snapshot");
}

static int snap_restore(PBVH *p, int slot) {
  if (!p || !p->snap_valid[slot]) return sg_set_error("no snapshot");
  auto items = snap_items(p);
  for (size_t i = 0; i < items.size(); i++)
    hipMemcpyAsync(items[i].dev, p->snap[slot][i], items[i].bytes, hipMemcpyDeviceToDevice, p->stream);
  invalidate_host_views(p);
  return sg_check_hip(hipStreamSynchronize(p->stream), "restore");
}

int sg_pbvh_snapshot_save(PBVH *p) {
  if (p && p->type == PBVH_BMESH) return sg_set_error("PBVH_BMESH: no snapshots (the topology changes)");
  if (p && p->type == PBVH_GRIDS) return p->gd ? sg_grids_snapshot(p, 0) : sg_set_error("no grids");
  return snap_save(p, 0);
}

int sg_pbvh_snapshot_restore(PBVH *p) {
  if (p && p->type == PBVH_BMESH) return sg_set_error("PBVH_BMESH: no snapshots (the topology changes)");
  if (p && p->type == PBVH_GRIDS) {
    if (!p->gd) return sg_set_error("no grids");
    p->host_bb_stale = true;
    return sg_grids_snapshot(p, 1);
  }
  return snap_restore(p, 0);
}

int sg_dab_apply(PBVH *p, const SgDabParams *dab) {
  if (int rc = check_dab(p, dab)) return rc;
  if (p->type == PBVH_GRIDS) return sg_grids_dab_apply(p, dab);
  if (p->type == PBVH_BMESH) return sg_dyn_dab_apply(p, dab);
  const float *curve = nullptr;
  if (dab->curve_preset == SG_CURVE_CUSTOM && dab->curve_table) {
    if (ensure_curves(p, 1)) return -1;
    hipMemcpyAsync(p->d.curve, dab->curve_table, (SG_CM_TABLE + 1) * sizeof(float), hipMemcpyHostToDevice,
                   p->stream);
    curve = p->d.curve;
  }
  if (g_stats && !ensure_stats_buffers(p)) return -1;
  SgDabK k = sg_make_dab(dab, g_stats ? 1 : 0);
  const int stamp = (++p->stamp) * 128;
  p->halt_live = k.learned && p->mlp;
  int rc = sg_enqueue_clear_ctr(p);
  if (!rc) rc = sg_enqueue_dab(p, k, curve, nullptr, 0, stamp, 0);
  if (!rc) rc = sg_enqueue_finish(p, 0);
  if (!rc) rc = sg_enqueue_flush(p);
  long long ctr[SG_CTR_COUNT];
  int halt = 0;
  if (!rc) rc = sg_check_hip(hipMemcpyAsync(ctr, p->d.ctr, sizeof(ctr), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipMemcpyAsync(&halt, p->d.halt, 4, hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "dab");
  if (rc) return rc;
  if (k.learned && p->mlp) {
    if (halt) {  // rejected: the plain brush does this dab
      p->learned_fallbacks++;
      SgDabParams plain = *dab;
      plain.flag &= ~SG_SMOOTH_LEARNED;
      return sg_dab_apply(p, &plain);
    }
    p->learned_accepted++;
  }
  for (int i = 0; i < 8; i++) p->last_stats[i] = 0;
  p->last_tool = dab->sculpt_tool;
  p->last_learned = k.learned && p->mlp;
  p->dab_count++;
  p->nodes_dirty_deferred = false;
  invalidate_host_views(p);
  memcpy(p->last_ctr, ctr, sizeof(ctr));
  return 0;
}

// Device buffers of a ray-placed stroke: its rays and one 8-float placement slot per
// dab (cx cy cz gx gy gz hit 0).
static int ensure_place(float **place, SgRay **rays, int *cap, int n) {
  if (n <= *cap) return 0;
  if (*place) hipFree(*place);
  if (*rays) hipFree(*rays);
  *place = nullptr;
  *rays = nullptr;
  *cap = 0;
  if (sg_check_hip(hipMalloc(place, (size_t)n * 8 * sizeof(float)), "hipMalloc")) return -1;
  if (sg_check_hip(hipMalloc(rays, ((size_t)n + 1) * sizeof(SgRay)), "hipMalloc")) return -1;
  *cap = n;
  return 0;
}

// Dab i's placement: a ray cast on the mesh as the previous dab left it (serial - 1 is
// the dab whose touched search boxes are not refreshed yet); Grab also casts the next
// dab's ray for its translation.  Rays along a coordinate plane need current internal
// boxes, so the deferred refit runs first for them.
static int enqueue_place(PBVH *p, const SgDabParams *dabs, const SgRay *rays_host, const SgRay *rays_dev,
                         float *place_dev, int i, int n, int serial) {
  const bool grab = dabs[i].sculpt_tool == SG_TOOL_GRAB && i + 1 < n;
  if (sg_ray_host_needs_refit(rays_host + i, grab ? 2 : 1) && sg_enqueue_flush(p)) return -1;
  return sg_enqueue_raycast(p, rays_dev + i, grab ? 2 : 1, 0, serial - 1, nullptr, place_dev + 8 * (size_t)i,
                            grab ? 1 : 0, nullptr);
}

static int check_rays(const SgRay *rays, int n) {
  for (int i = 0; i < n; i++) {
    const float *d = rays[i].normal;
    if (!(d[0] == d[0] && d[1] == d[1] && d[2] == d[2]) || (d[0] == 0.0f && d[1] == 0.0f && d[2] == 0.0f))
      return sg_set_error("ray %d has no direction", i);
  }
  return 0;
}

// A rejected learned dab of a placed stroke is redone by the plain brush at the centre
// its ray cast found (or not at all when the ray missed).
static int plain_fallback(PBVH *p, const SgDabParams &dab, const float *place_dev, int i) {
  SgDabParams plain = dab;
  plain.flag &= ~SG_SMOOTH_LEARNED;
  if (place_dev) {
    float pl[8];
    if (int rc = sg_check_hip(hipMemcpy(pl, place_dev + 8 * (size_t)i, sizeof(pl), hipMemcpyDeviceToHost), "copy"))
      return rc;
    if (pl[6] == 0.0f) return 0;
    memcpy(plain.location, pl, 3 * sizeof(float));
    memcpy(plain.grab_delta, pl + 3, 3 * sizeof(float));
  }
  return sg_dab_apply(p, &plain);
}

static int stroke_apply(PBVH *p, const SgDabParams *dabs, const SgRay *rays, int n, int flags) {
  if (!p || (!dabs && n > 0) || n < 0) return sg_set_error("bad arguments");
  if (p->no_device) return sg_set_error("no HIP device");
  for (int i = 0; i < n; i++) {
    if (int rc = check_dab(p, &dabs[i])) return rc;
  }
  if (rays && check_rays(rays, n)) return -1;
  if (n == 0) return 0;
  if (rays) {
    if (sg_ray_ensure(p)) return -1;  // builds its tables with synchronous copies: not under capture
    if (ensure_place(&p->d.place, &p->d.place_rays, &p->d.place_cap, n)) return -1;
    if (sg_check_hip(hipMemcpyAsync(p->d.place_rays, rays, (size_t)n * sizeof(SgRay), hipMemcpyHostToDevice, p->stream),
                     "copy"))
      return -1;
    p->last_place = p->d.place;
    p->last_place_n = n;
  }
  const bool use_graph = (flags & 1) != 0;
  const bool flush_each = (flags & 2) != 0;
  const bool log = (flags & 4) != 0;
  int ncustom = 0;
  for (int i = 0; i < n; i++) ncustom += (dabs[i].curve_preset == SG_CURVE_CUSTOM && dabs[i].curve_table) ? 1 : 0;
  if (ncustom && ensure_curves(p, ncustom)) return -1;
  if (log && ensure_log(p, n)) return -1;
  if (g_stats && !ensure_stats_buffers(p)) return -1;
  {
    int ci = 0;
    for (int i = 0; i < n; i++) {
      if (dabs[i].curve_preset == SG_CURVE_CUSTOM && dabs[i].curve_table) {
        hipMemcpyAsync(p->d.curve + (size_t)ci * (SG_CM_TABLE + 1), dabs[i].curve_table,
                       (SG_CM_TABLE + 1) * sizeof(float), hipMemcpyHostToDevice, p->stream);
        ci++;
      }
    }
  }
  // Dabs [a, n) as one segment (serials restart at 0); returns the position of a
  // rejected learned dab (halt) or -1.
  auto run_segment = [&](int a0, int *halted) -> int {
    auto enqueue_all = [&]() -> int {
      int ci = 0;
      for (int i = 0; i < a0; i++)
        ci += (dabs[i].curve_preset == SG_CURVE_CUSTOM && dabs[i].curve_table) ? 1 : 0;
      if (sg_enqueue_clear_ctr(p)) return -1;
      for (int i = a0; i < n; i++) {
        const float *curve = nullptr;
        if (dabs[i].curve_preset == SG_CURVE_CUSTOM && dabs[i].curve_table)
          curve = p->d.curve + (size_t)(ci++) * (SG_CM_TABLE + 1);
        SgDabK k = sg_make_dab(&dabs[i], g_stats ? 1 : 0);
        const int stamp = (++p->stamp) * 128;
        const float *place = nullptr;
        if (rays) {
          if (int rc = enqueue_place(p, dabs, rays, p->d.place_rays, p->d.place, i, n, i - a0)) return rc;
          place = p->d.place + 8 * (size_t)i;
        }
        if (int rc = sg_enqueue_dab(p, k, curve, log ? p->d.log : nullptr, i - a0, stamp, i - a0, nullptr, place))
          return rc;
        if (flush_each && sg_enqueue_flush(p)) return -1;
      }
      if (sg_enqueue_finish(p, n - 1 - a0)) return -1;
      if (!flush_each && sg_enqueue_flush(p)) return -1;
      if (log && sg_enqueue_log_row(p, p->d.log, n - 1 - a0)) return -1;
      return 0;
    };
    int rc = 0;
    if (use_graph) {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      p->capturing = true;
      rc = sg_check_hip(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal), "begin capture");
      if (!rc) {
        const int erc = enqueue_all();
        const hipError_t e = hipStreamEndCapture(p->stream, &graph);
        rc = erc ? erc : sg_check_hip(e, "end capture");
      }
      p->capturing = false;
      if (!rc) rc = sg_check_hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
      if (!rc) rc = sg_check_hip(hipGraphLaunch(exec, p->stream), "graph launch");
      if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "stroke");
      if (exec) hipGraphExecDestroy(exec);
      if (graph) hipGraphDestroy(graph);
    } else {
      rc = enqueue_all();
      if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "stroke");
    }
    int h = 0;
    if (!rc) rc = sg_check_hip(hipMemcpy(&h, p->d.halt, 4, hipMemcpyDeviceToHost), "copy");
    *halted = h ? a0 + h - 1 : -1;
    return rc;
  };
  int rc = 0;
  int a0 = 0;
  const int nlearned = [&]() {
    int c = 0;
    for (int i = 0; i < n; i++) c += (dabs[i].sculpt_tool == SG_TOOL_SMOOTH && (dabs[i].flag & SG_SMOOTH_LEARNED)) ? 1 : 0;
    return p->mlp ? c : 0;
  }();
  int nfallback = 0;
  p->halt_live = nlearned > 0;
  while (a0 < n) {
    int halted = -1;
    rc = run_segment(a0, &halted);
    if (rc) return rc;
    if (halted < 0) break;
    // the rejected learned dab runs with the plain brush, then the stroke resumes
    if ((rc = plain_fallback(p, dabs[halted], rays ? p->d.place : nullptr, halted))) return rc;
    nfallback++;
    a0 = halted + 1;
    p->halt_live = true;  // sg_dab_apply cleared it for the plain dab
  }
  p->learned_fallbacks += nfallback;
  p->learned_accepted += nlearned - nfallback;
  if (a0 < n) {  // otherwise the last dab was a plain fallback and set last_ctr itself
    long long ctr[SG_CTR_COUNT];
    hipMemcpy(ctr, p->d.ctr + (size_t)((n - 1 - a0) & 1) * SG_CTR_COUNT, sizeof(ctr), hipMemcpyDeviceToHost);
    memcpy(p->last_ctr, ctr, sizeof(ctr));
    p->last_learned = p->mlp && dabs[n - 1].sculpt_tool == SG_TOOL_SMOOTH && (dabs[n - 1].flag & SG_SMOOTH_LEARNED);
    p->last_tool = dabs[n - 1].sculpt_tool;
  }
  p->dab_count += n - nfallback;  // the fallbacks counted themselves
  p->nodes_dirty_deferred = false;
  invalidate_host_views(p);
  return 0;
}

int sg_dab_apply_stroke(PBVH *p, const SgDabParams *dabs, int n, int flags) {
  if (p && p->type == PBVH_BMESH) {  // dab by dab: each topology round waits for its counters
    for (int i = 0; i < n; i++)
      if (int rc = sg_dab_apply(p, &dabs[i])) return rc;
    return 0;
  }
  if (p && p->type == PBVH_GRIDS) {
    if (flags & SG_STROKE_GRAPH) {  // one captured graph
      SgStroke *s = sg_stroke_prepare(p, dabs, n, flags);
      if (!s) return -1;
      const int rc = sg_stroke_launch(s, 1);
      sg_stroke_free(s);
      return rc;
    }
    // stream launches: every dab enqueued, one wait at the end
    for (int i = 0; i < n; i++)
      if (int rc = check_dab(p, &dabs[i])) return rc;
    if (!p->gd->topo_of && p->subdiv_ccg)
      if (int rc = sg_grids_attach_topology(p, p->subdiv_ccg)) return rc;
    const bool log = (flags & SG_STROKE_LOG) != 0;
    if (log && ensure_log(p, n)) return -1;
    int ncustom = 0;
    for (int i = 0; i < n; i++) ncustom += (dabs[i].curve_preset == SG_CURVE_CUSTOM && dabs[i].curve_table) ? 1 : 0;
    float *curves = nullptr;
    if (ncustom && sg_check_hip(hipMalloc(&curves, (size_t)ncustom * (SG_CM_TABLE + 1) * sizeof(float)), "hipMalloc"))
      return -1;
    int rc = 0, ci = 0;
    for (int i = 0; i < n && !rc; i++) {
      const float *curve = nullptr;
      if (dabs[i].curve_preset == SG_CURVE_CUSTOM && dabs[i].curve_table) {
        float *c = curves + (size_t)(ci++) * (SG_CM_TABLE + 1);
        rc = sg_check_hip(hipMemcpyAsync(c, dabs[i].curve_table, (SG_CM_TABLE + 1) * sizeof(float),
                                         hipMemcpyHostToDevice, p->stream),
                          "copy");
        curve = c;
      }
      if (!rc) rc = sg_grids_enqueue_dab(p, sg_make_dab(&dabs[i], 0), curve, log ? p->d.log : nullptr, i);
    }
    if (!rc) rc = sg_grids_enqueue_flush(p);
    if (!rc) rc = sg_grids_fetch_last(p);  // waits for the stream
    if (curves) hipFree(curves);
    p->last_tool = dabs[n - 1].sculpt_tool;
    p->dab_count += n;
    p->host_bb_stale = true;
    return rc;
  }
  return stroke_apply(p, dabs, nullptr, n, flags);
}

int sg_dab_apply_stroke_placed(PBVH *p, const SgDabParams *dabs, const SgRay *rays, int n, int flags) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: PBVH_GRIDS has no ray-placed strokes");
  if (n > 0 && !rays) return sg_set_error("bad arguments");
  return stroke_apply(p, dabs, rays, n, flags);
}

int sg_stroke_placements(PBVH *p, float *out, int n) {
  if (!p || !out || n < 0) return sg_set_error("bad arguments");
  if (!p->last_place) return sg_set_error("no placed stroke has run");
  const int m = n < p->last_place_n ? n : p->last_place_n;
  if (int rc = sg_check_hip(hipStreamSynchronize(p->stream), "sync")) return rc;
  if (int rc = sg_check_hip(hipMemcpy(out, p->last_place, (size_t)m * 8 * sizeof(float), hipMemcpyDeviceToHost), "copy"))
    return rc;
  return m;
}

// ---- prepared strokes: one hipGraph per dab list, replayed on demand ----------

struct SgStroke {
  PBVH *p = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  float *curves = nullptr;
  int n = 0;
  int flags = 0;
  bool learned = false;           // contains learned Smooth dabs: launch checks for a halt
  std::vector<SgDabParams> dabs;  // (custom curve pointers stay the caller's)
  std::vector<SgRay> rays;        // ray-placed stroke: one ray per dab
  float *place = nullptr;         // its placement slots and device rays
  SgRay *place_rays = nullptr;
  int place_cap = 0;
};

static void stroke_release(SgStroke *s) {
  if (s->exec) hipGraphExecDestroy(s->exec);
  if (s->graph) hipGraphDestroy(s->graph);
  if (s->curves) hipFree(s->curves);
  if (s->place) hipFree(s->place);
  if (s->place_rays) hipFree(s->place_rays);
  delete s;
}

static SgStroke *stroke_prepare(PBVH *p, const SgDabParams *dabs, const SgRay *rays, int n, int flags) {
  if (!p || n <= 0 || !dabs) {
    sg_set_error("bad arguments");
    return nullptr;
  }
  if (rays && (check_rays(rays, n) || sg_ray_ensure(p))) return nullptr;
  if (p->no_device) {
    sg_set_error("no HIP device");
    return nullptr;
  }
  for (int i = 0; i < n; i++) {
    if (check_dab(p, &dabs[i])) return nullptr;
  }
  const bool flush_each = (flags & 2) != 0;
  const bool log = (flags & 4) != 0;
  if (log && ensure_log(p, n)) return nullptr;
  SgStroke *s = new SgStroke();
  s->p = p;
  s->n = n;
  s->flags = flags;
  s->dabs.assign(dabs, dabs + n);
  if (rays) {
    s->rays.assign(rays, rays + n);
    if (ensure_place(&s->place, &s->place_rays, &s->place_cap, n) ||
        sg_check_hip(hipMemcpy(s->place_rays, rays, (size_t)n * sizeof(SgRay), hipMemcpyHostToDevice), "copy")) {
      stroke_release(s);
      return nullptr;
    }
  }
  for (int i = 0; i < n; i++)
    if (dabs[i].sculpt_tool == SG_TOOL_SMOOTH && (dabs[i].flag & SG_SMOOTH_LEARNED) && p->mlp) s->learned = true;
  int ncustom = 0;
  for (int i = 0; i < n; i++) ncustom += (dabs[i].curve_preset == SG_CURVE_CUSTOM && dabs[i].curve_table) ? 1 : 0;
  if (ncustom) {
    if (sg_check_hip(hipMalloc(&s->curves, (size_t)ncustom * (SG_CM_TABLE + 1) * sizeof(float)), "hipMalloc")) {
      delete s;
      return nullptr;
    }
    int ci = 0;
    for (int i = 0; i < n; i++) {
      if (dabs[i].curve_preset == SG_CURVE_CUSTOM && dabs[i].curve_table)
        hipMemcpy(s->curves + (size_t)(ci++) * (SG_CM_TABLE + 1), dabs[i].curve_table,
                  (SG_CM_TABLE + 1) * sizeof(float), hipMemcpyHostToDevice);
    }
  }
  p->halt_live = s->learned;
  if (p->type == PBVH_GRIDS && !p->gd->topo_of && p->subdiv_ccg && sg_grids_attach_topology(p, p->subdiv_ccg)) {
    stroke_release(s);
    return nullptr;
  }
  p->capturing = true;
  int rc = sg_check_hip(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal), "begin capture");
  int erc = 0;
  if (!rc && p->type == PBVH_GRIDS) {  // each grid dab is self-contained (counters, normals, refit)
    int ci = 0;
    for (int i = 0; i < n && !erc; i++) {
      const float *curve = nullptr;
      if (dabs[i].curve_preset == SG_CURVE_CUSTOM && dabs[i].curve_table)
        curve = s->curves + (size_t)(ci++) * (SG_CM_TABLE + 1);
      erc = sg_grids_enqueue_dab(p, sg_make_dab(&dabs[i], 0), curve, log ? p->d.log : nullptr, i);
    }
    if (!erc) erc = sg_grids_enqueue_flush(p);
    const hipError_t e = hipStreamEndCapture(p->stream, &s->graph);
    rc = erc ? erc : sg_check_hip(e, "end capture");
  } else if (!rc) {
    int ci = 0;
    erc = sg_enqueue_clear_ctr(p);
    for (int i = 0; i < n && !erc; i++) {
      const float *curve = nullptr;
      if (dabs[i].curve_preset == SG_CURVE_CUSTOM && dabs[i].curve_table)
        curve = s->curves + (size_t)(ci++) * (SG_CM_TABLE + 1);
      SgDabK k = sg_make_dab(&dabs[i], 0);
      const float *place = nullptr;
      if (rays) {
        erc = enqueue_place(p, dabs, rays, s->place_rays, s->place, i, n, i);
        place = s->place + 8 * (size_t)i;
      }
      if (!erc) erc = sg_enqueue_dab(p, k, curve, log ? p->d.log : nullptr, i, 0, i, nullptr, place);
      if (!erc && flush_each) erc = sg_enqueue_flush(p);
    }
    if (!erc) erc = sg_enqueue_finish(p, n - 1);
    if (!erc && !flush_each) erc = sg_enqueue_flush(p);
    if (!erc && log) erc = sg_enqueue_log_row(p, p->d.log, n - 1);
    const hipError_t e = hipStreamEndCapture(p->stream, &s->graph);
    rc = erc ? erc : sg_check_hip(e, "end capture");
  }
  p->capturing = false;
  if (!rc) rc = sg_check_hip(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0), "instantiate");
  if (rc) {
    stroke_release(s);
    return nullptr;
  }
  return s;
}

SgStroke *sg_stroke_prepare(PBVH *p, const SgDabParams *dabs, int n, int flags) {
  if (p && p->type == PBVH_BMESH) {
    sg_set_error("PBVH_BMESH: no captured strokes (each topology round waits for its counters)");
    return nullptr;
  }
  return stroke_prepare(p, dabs, nullptr, n, flags);
}

SgStroke *sg_stroke_prepare_placed(PBVH *p, const SgDabParams *dabs, const SgRay *rays, int n, int flags) {
  if (p && p->type != PBVH_FACES) { sg_set_error("faces only: PBVH_GRIDS has no captured strokes"); return nullptr; }
  if (!rays) {
    sg_set_error("bad arguments");
    return nullptr;
  }
  return stroke_prepare(p, dabs, rays, n, flags);
}

int sg_stroke_launch(SgStroke *s, int sync) {
  if (!s) return sg_set_error("null stroke");
  PBVH *p = s->p;
  if (s->place) {
    p->last_place = s->place;
    p->last_place_n = s->n;
  }
  int rc = sg_check_hip(hipGraphLaunch(s->exec, p->stream), "graph launch");
  if (!rc && (sync || s->learned)) rc = sg_check_hip(hipStreamSynchronize(p->stream), "stroke");
  if (!rc && s->learned) {
    int h = 0;
    rc = sg_check_hip(hipMemcpy(&h, p->d.halt, 4, hipMemcpyDeviceToHost), "copy");
    const int ran = h ? h - 1 : s->n;  // dabs the graph completed
    if (p->mlp)
      for (int i = 0; i < ran; i++)
        p->learned_accepted +=
            (s->dabs[i].sculpt_tool == SG_TOOL_SMOOTH && (s->dabs[i].flag & SG_SMOOTH_LEARNED)) ? 1 : 0;
    if (!rc && h) {  // dab h-1 was rejected: plain brush for it, then the rest
      const int f = h - 1;
      p->dab_count += f;  // the graph's dabs before it; the calls below count themselves
      rc = plain_fallback(p, s->dabs[f], s->place, f);
      p->learned_fallbacks++;
      if (!rc && f + 1 < s->n)
        rc = stroke_apply(p, s->dabs.data() + f + 1, s->rays.empty() ? nullptr : s->rays.data() + f + 1,
                          s->n - f - 1, s->flags & ~1);
      invalidate_host_views(p);
      return rc;
    }
  }
  p->dab_count += s->n;
  if (!rc && sync && p->type == PBVH_GRIDS) {
    rc = sg_grids_fetch_last(p);
    p->last_tool = s->dabs[s->n - 1].sculpt_tool;
    p->last_learned = false;
    p->host_bb_stale = true;
    return rc;
  }
  if (!rc && (sync || s->learned)) {  // the stream is idle: record the last dab's counters
    const SgDabParams &last = s->dabs[s->n - 1];
    rc = sg_check_hip(hipMemcpy(p->last_ctr, p->d.ctr + (size_t)((s->n - 1) & 1) * SG_CTR_COUNT,
                                SG_CTR_COUNT * sizeof(long long), hipMemcpyDeviceToHost),
                      "copy");
    p->last_tool = last.sculpt_tool;
    p->last_learned = s->learned && last.sculpt_tool == SG_TOOL_SMOOTH && (last.flag & SG_SMOOTH_LEARNED);
  }
  invalidate_host_views(p);
  return rc;
}

void sg_stroke_free(SgStroke *s) {
  if (!s) return;
  hipStreamSynchronize(s->p->stream);
  if (s->p->last_place == s->place) {
    s->p->last_place = nullptr;
    s->p->last_place_n = 0;
  }
  stroke_release(s);
}

// ---- autotuner: workgroup counts of the per-dab kernels --------------------------
// Times a sample stroke (captured graph, launched from the same checkpoint each
// trial) for each candidate grid of one kernel at a time (coordinate descent over
// brush, post, smooth step), keeps the fastest, and restores the mesh afterwards.
static double time_stroke(PBVH *p, const SgDabParams *dabs, int n, int reps) {
  SgStroke *st = sg_stroke_prepare(p, dabs, n, 0);
  if (!st) return -1.0;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  double best = 1e30;
  for (int r = 0; r < reps + 1; r++) {
    snap_restore(p, 1);
    hipEventRecord(e0, p->stream);
    sg_stroke_launch(st, 0);
    hipEventRecord(e1, p->stream);
    hipEventSynchronize(e1);
    float ms = 0.0f;
    hipEventElapsedTime(&ms, e0, e1);
    if (r > 0 && ms < best) best = ms;  // first launch warms up
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  sg_stroke_free(st);
  return best;
}

int sg_pbvh_autotune(PBVH *p, const SgDabParams *dabs, int n, double *out_ms) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: no autotuner for PBVH_GRIDS");
  if (!p || !dabs || n <= 0) return sg_set_error("bad arguments");
  if (int rc = snap_save(p, 1)) return rc;
  const int cand[] = {32, 64, 128, 192, 256, 384, 512};
  int *slots[3] = {&p->grid_brush, &p->grid_post, &p->grid_smooth};
  for (int *sl : slots) *sl = p->grid_wg;
  double best = time_stroke(p, dabs, n, 2);
  if (best < 0) return -1;
  const double start = best;
  for (int pass = 0; pass < 2; pass++) {
    for (int *sl : slots) {
      const int keep = *sl;
      int choice = keep;
      for (int g : cand) {
        if (g == keep) continue;
        *sl = g;
        const double t = time_stroke(p, dabs, n, 2);
        if (t > 0 && t < best * 0.98) {
          best = t;
          choice = g;
        }
      }
      *sl = choice;
    }
  }
  snap_restore(p, 1);
  if (out_ms) {
    out_ms[0] = start;
    out_ms[1] = best;
  }
  return 0;
}

// The plain brush as the judge of a learned Smooth dab: from the current state run the
// learned dab and keep its positions, return to the state, run the plain dab (which
// stays applied) and compare the two on the affected vertices.
int sg_learned_judge(PBVH *p, const SgDabParams *dab, double out[4]) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: learned Smooth is faces only");
  if (int rc = check_dab(p, dab)) return rc;
  if (!p->mlp || dab->sculpt_tool != SG_TOOL_SMOOTH) return sg_set_error("judge needs a model and a Smooth dab");
  const size_t V = (size_t)p->d.V;
  if (!p->judge_buf && sg_check_hip(hipMalloc(&p->judge_buf, std::max<size_t>(V, 1) * 12), "hipMalloc")) return -1;
  if (!p->judge_acc && sg_check_hip(hipMalloc(&p->judge_acc, 16), "hipMalloc")) return -1;
  if (int rc = snap_save(p, 1)) return rc;
  SgDabParams ld = *dab;
  ld.flag |= SG_SMOOTH_LEARNED;
  const long long acc0 = p->learned_accepted;
  if (int rc = sg_dab_apply(p, &ld)) return rc;
  const bool accepted = p->learned_accepted > acc0;
  hipMemcpyAsync(p->judge_buf, p->d.px, V * 4, hipMemcpyDeviceToDevice, p->stream);
  hipMemcpyAsync(p->judge_buf + V, p->d.py, V * 4, hipMemcpyDeviceToDevice, p->stream);
  hipMemcpyAsync(p->judge_buf + 2 * V, p->d.pz, V * 4, hipMemcpyDeviceToDevice, p->stream);
  if (int rc = snap_restore(p, 1)) return rc;
  SgDabParams pd = *dab;
  pd.flag &= ~SG_SMOOTH_LEARNED;
  if (int rc = sg_dab_apply(p, &pd)) return rc;
  const int na = (int)p->last_ctr[SG_CTR_NAFF];
  const int nslot = (int)p->last_ctr[SG_CTR_NGATHER] << p->d.aff_shift;  // the list's slots
  double acc[2] = {0.0, 0.0};
  if (na > 0) {
    // snapshot slot 1 holds the positions before the dab (items 0..2)
    int rc = sg_enqueue_judge(p, nslot, p->judge_buf, (const float *)p->snap[1][0], (const float *)p->snap[1][1],
                              (const float *)p->snap[1][2], p->judge_acc);
    if (!rc) rc = sg_check_hip(hipMemcpyAsync(acc, p->judge_acc, 16, hipMemcpyDeviceToHost, p->stream), "copy");
    if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "judge");
    if (rc) return rc;
  }
  const double r = dab->radius;
  out[0] = na > 0 ? sqrt(acc[0] / na) / r : 0.0;
  out[1] = na > 0 ? sqrt(acc[1] / na) / r : 0.0;
  out[2] = accepted ? 1.0 : 0.0;
  out[3] = (double)na;
  return 0;
}

int sg_pbvh_set_learned(PBVH *p, SgMlp *m) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: learned Smooth is faces only");
  if (!p) return sg_set_error("null PBVH");
  p->mlp = m;
  return 0;
}

int sg_pbvh_learned_stats(PBVH *p, long long out[2]) {
  if (!p) return sg_set_error("null PBVH");
  out[0] = p->learned_accepted;
  out[1] = p->learned_fallbacks;
  return 0;
}

// One plain Smooth dab that also records, for each affected vertex, the learned
// model's input features before the dab and the displacement target after it.
int sg_learned_collect(PBVH *p, const SgDabParams *dab, float *X, float *T, float *ok, int cap) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: learned Smooth is faces only");
  if (int rc = check_dab(p, dab)) return rc;
  if (dab->sculpt_tool != SG_TOOL_SMOOTH) return sg_set_error("collection needs a Smooth dab");
  if (cap <= 0 || !X || !T || !ok) return sg_set_error("bad arguments");
  SgCollect col;
  col.cap = cap;
  col.X = col.frame = col.T = nullptr;
  int arc = sg_check_hip(hipMalloc(&col.X, (size_t)cap * 28 * 4), "hipMalloc");
  if (!arc) arc = sg_check_hip(hipMalloc(&col.frame, (size_t)cap * 16 * 4), "hipMalloc");
  if (!arc) arc = sg_check_hip(hipMalloc(&col.T, (size_t)cap * 3 * 4), "hipMalloc");
  if (arc) {
    if (col.X) hipFree(col.X);
    if (col.frame) hipFree(col.frame);
    return arc;
  }
  SgDabParams plain = *dab;
  plain.flag &= ~SG_SMOOTH_LEARNED;
  SgDabK k = sg_make_dab(&plain, 0);
  p->halt_live = false;
  int rc = sg_enqueue_clear_ctr(p);
  if (!rc) rc = sg_enqueue_dab(p, k, nullptr, nullptr, 0, 0, 0, &col);
  if (!rc) rc = sg_enqueue_flush(p);
  long long ctr[SG_CTR_COUNT];
  if (!rc) rc = sg_check_hip(hipMemcpyAsync(ctr, p->d.ctr, sizeof(ctr), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "collect");
  int na = 0;
  if (!rc) {
    // one row per slot of the affected list (rows without a record are marked not valid)
    na = (int)std::min<long long>((long long)(int)ctr[SG_CTR_NGATHER] << p->d.aff_shift, cap);
    std::vector<float> fr((size_t)std::max(na, 1) * 16);
    rc = sg_check_hip(hipMemcpy(X, col.X, (size_t)na * 28 * 4, hipMemcpyDeviceToHost), "copy");
    if (!rc) rc = sg_check_hip(hipMemcpy(T, col.T, (size_t)na * 3 * 4, hipMemcpyDeviceToHost), "copy");
    if (!rc) rc = sg_check_hip(hipMemcpy(fr.data(), col.frame, (size_t)na * 16 * 4, hipMemcpyDeviceToHost), "copy");
    for (int i = 0; i < na; i++) ok[i] = fr[(size_t)i * 16 + 13];
    memcpy(p->last_ctr, ctr, sizeof(ctr));
    p->last_tool = SG_TOOL_SMOOTH;
    p->last_learned = false;
    p->dab_count++;
  }
  hipFree(col.X);
  hipFree(col.frame);
  hipFree(col.T);
  invalidate_host_views(p);
  return rc ? rc : na;
}

int sg_pbvh_set_profile(PBVH *p, int on) {
  if (!p || p->no_device) return sg_set_error("no device");
  if (p->profile && !on) sg_profile_collect(p);
  p->profile = on != 0;
  return 0;
}

int sg_pbvh_profile_read(PBVH *p, double *ms, long long *count, int n, int reset) {
  if (!p || p->no_device) return sg_set_error("no device");
  if (sg_profile_collect(p)) return -1;
  for (int i = 0; i < n && i < 16; i++) {
    ms[i] = p->prof_ms[i];
    count[i] = p->prof_n[i];
  }
  if (reset) {
    for (int i = 0; i < 16; i++) {
      p->prof_ms[i] = 0.0;
      p->prof_n[i] = 0;
    }
  }
  return SG_K_COUNT;
}

// Device-to-device copy of `bytes` between two fresh buffers, best of `reps` copies
// timed with HIP events: the GPU's measured copy bandwidth, and a kernel of known
// traffic (bytes read, bytes written) for calibrating the hardware byte counters.
int sg_hbm_copy_probe(long long bytes, int reps, float *ms_out) {
  if (bytes <= 0 || reps <= 0 || !ms_out) return sg_set_error("bad arguments");
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = sg_check_hip(hipMalloc(&a, (size_t)bytes), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&b, (size_t)bytes), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
  if (!rc) rc = sg_check_hip(hipEventCreate(&e0), "event");
  if (!rc) rc = sg_check_hip(hipEventCreate(&e1), "event");
  if (!rc) rc = sg_check_hip(hipMemsetAsync(a, 1, (size_t)bytes, s), "memset");
  float best = 1e30f;
  for (int r = 0; r < reps && !rc; r++) {
    hipEventRecord(e0, s);
    rc = sg_check_hip(hipMemcpyAsync(b, a, (size_t)bytes, hipMemcpyDeviceToDevice, s), "copy");
    hipEventRecord(e1, s);
    if (!rc) rc = sg_check_hip(hipEventSynchronize(e1), "copy sync");
    float ms = 0.0f;
    if (!rc) hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) best = ms;
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (s) hipStreamDestroy(s);
  if (a) hipFree(a);
  if (b) hipFree(b);
  *ms_out = best;
  return rc;
}

int sg_pbvh_timer_start(PBVH *p) {
  if (!p || p->no_device) return sg_set_error("no device");
  for (int i = 0; i < 2; i++)
    if (!p->timer[i] && sg_check_hip(hipEventCreate(&p->timer[i]), "event create")) return -1;
  return sg_check_hip(hipEventRecord(p->timer[0], p->stream), "event record");
}

int sg_pbvh_timer_stop(PBVH *p, float *ms) {
  if (!p || p->no_device || !p->timer[0]) return sg_set_error("timer not started");
  if (sg_check_hip(hipEventRecord(p->timer[1], p->stream), "event record")) return -1;
  if (sg_check_hip(hipEventSynchronize(p->timer[1]), "event sync")) return -1;
  return sg_check_hip(hipEventElapsedTime(ms, p->timer[0], p->timer[1]), "event elapsed");
}

int sg_dab_stroke_log(PBVH *p, long long *out, int n) {
  if (!p || !p->d.log || n > p->d.log_cap) return sg_set_error("no stroke log");
  return sg_check_hip(hipMemcpy(out, p->d.log, (size_t)n * SG_CTR_COUNT * sizeof(long long), hipMemcpyDeviceToHost),
                      "copy log");
}

int sg_sync(PBVH *p) {
  if (!p) return sg_set_error("null PBVH");
  if (p->no_device) return sg_set_error("no HIP device");
  return sg_check_hip(hipStreamSynchronize(p->stream), "sync");
}

int sg_dab_last_affected(PBVH *p, int *out, int cap) {
  if (!p) return sg_set_error("null PBVH");
  if (p->no_device) return sg_set_error("no HIP device");
  if (p->type == PBVH_GRIDS) return p->gd ? sg_grids_last_affected(p, out, cap) : sg_set_error("no grids");
  if (p->type == PBVH_BMESH) return sg_set_error("PBVH_BMESH: vertices are labels (sg_pbvh_bmesh_last_affected)");
  int na = (int)p->last_ctr[SG_CTR_NAFF];
  std::vector<int> tmp(std::max(na, 1));
  if (p->last_tool == SG_TOOL_ELASTIC_DEFORM) {
    // an elastic deform writes every visible uniq vertex (it keeps no list)
    tmp.clear();
    for (const PBVHNode &n : p->nodes) {
      if (n.leaf < 0) continue;
      for (int k = 0; k < n.uniq_verts; k++) {
        const int v = n.vert_indices[k];
        if (!(p->respect_hide && !p->vhide.empty() && p->vhide[v])) tmp.push_back(p->perm_new_of_old[v]);
      }
    }
    na = (int)tmp.size();
    tmp.resize(std::max(na, 1));
  } else if (p->last_learned) {
    // an accepted learned dab stamps its vertices instead of listing them: the
    // stamped vertices of the gathered leaves' ranges
    const int ng = (int)p->last_ctr[SG_CTR_NGATHER];
    std::vector<int2> gv(std::max(ng, 1));
    std::vector<int> stamp((size_t)std::max(p->d.V, 1));
    int seq = 0;
    if (ng > 0 && sg_check_hip(hipMemcpy(gv.data(), p->d.gl_v, ng * sizeof(int2), hipMemcpyDeviceToHost), "copy"))
      return -1;
    if (sg_check_hip(hipMemcpy(stamp.data(), p->d.vstamp, (size_t)p->d.V * sizeof(int), hipMemcpyDeviceToHost),
                     "copy") ||
        sg_check_hip(hipMemcpy(&seq, p->d.seq, sizeof(int), hipMemcpyDeviceToHost), "copy"))
      return -1;
    tmp.clear();
    for (int g = 0; g < ng; g++)
      for (int v = gv[g].x; v < gv[g].y; v++)
        if (stamp[v] == seq) tmp.push_back(v);
    na = (int)tmp.size();
    tmp.resize(std::max(na, 1));
  } else {
    // the list: gathered leaf g's records in slots (g << aff_shift) + [0, aff_cnt[g])
    const int ng = (int)p->last_ctr[SG_CTR_NGATHER];
    const size_t nslot = (size_t)ng << p->d.aff_shift;
    std::vector<int> cnt(std::max(ng, 1)), lst(std::max<size_t>(nslot, 1));
    if (ng > 0 &&
        (sg_check_hip(hipMemcpy(cnt.data(), p->d.aff_cnt, ng * sizeof(int), hipMemcpyDeviceToHost), "copy") ||
         sg_check_hip(hipMemcpy(lst.data(), p->d.aff_v, nslot * sizeof(int), hipMemcpyDeviceToHost), "copy")))
      return -1;
    tmp.clear();
    for (int g = 0; g < ng; g++)
      for (int i = 0; i < cnt[g]; i++) tmp.push_back(lst[((size_t)g << p->d.aff_shift) + i]);
    na = (int)tmp.size();
    tmp.resize(std::max(na, 1));
  }
  for (int i = 0; i < na; i++) tmp[i] = p->perm_old_of_new[tmp[i]];
  std::sort(tmp.begin(), tmp.begin() + na);
  for (int i = 0; i < na && i < cap; i++) out[i] = tmp[i];
  return na;
}

int sg_dab_last_stats(PBVH *p, int64_t out[24]) {
  if (!p) return sg_set_error("null PBVH");
  if (p->type == PBVH_BMESH) {  // [0] gathered [1] affected, [12..23] the dyntopo counters
    long long st[12];
    if (sg_pbvh_bmesh_last_stats(p, st)) return -1;
    for (int i = 0; i < 24; i++) out[i] = 0;
    out[0] = st[0];
    out[1] = st[1];
    for (int i = 0; i < 12; i++) out[12 + i] = st[i];
    return 0;
  }
  for (int i = 0; i < 24; i++) out[i] = p->last_ctr[i];
  out[SG_CTR_NGATHER] &= 0xffffffffll;  // the high half flags an open undo step
  return 0;
}

int sg_dab_last_gathered(PBVH *p, int *out, int cap) {
  if (!p) return sg_set_error("null PBVH");
  if (p->no_device) return sg_set_error("no HIP device");
  if (p->type == PBVH_GRIDS) return p->gd ? sg_grids_last_gathered(p, out, cap) : sg_set_error("no grids");
  if (p->type == PBVH_BMESH) return sg_dyn_last_gathered(p, out, cap);
  const int ng = (int)p->last_ctr[SG_CTR_NGATHER];
  std::vector<int> tmp(std::max(ng, 1));
  if (ng > 0 && sg_check_hip(hipMemcpy(tmp.data(), p->d.gathered, ng * sizeof(int), hipMemcpyDeviceToHost), "copy"))
    return -1;
  std::sort(tmp.begin(), tmp.begin() + ng);  // DFS leaf order = the reference's gather order
  for (int i = 0; i < ng && i < cap; i++) out[i] = p->leaf_node[tmp[i]];
  return ng;
}

int sg_pbvh_read_positions(PBVH *p, float *out) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: read grids with sg_pbvh_grids_sync_host");
  if (!p) return sg_set_error("null PBVH");
  return read_aos(p, 0, out);
}

int sg_pbvh_read_normals(PBVH *p, float *out) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: read grids with sg_pbvh_grids_sync_host");
  if (!p) return sg_set_error("null PBVH");
  return read_aos(p, 1, out);
}

int sg_pbvh_set_mask(PBVH *p, const float *mask) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: grid masks live in the grids (sg_pbvh_grids_upload)");
  if (!p) return sg_set_error("null PBVH");
  if (p->no_device) return sg_set_error("no HIP device");
  SgDev &d = p->d;
  if (!mask) {
    if (d.mask) hipFree(d.mask);
    d.mask = nullptr;
    return 0;
  }
  if (!d.mask && sg_check_hip(hipMalloc(&d.mask, (size_t)std::max(d.V, 1) * sizeof(float)), "hipMalloc")) return -1;
  std::vector<float> m(d.V);
  for (int i = 0; i < d.V; i++) m[i] = mask[p->perm_old_of_new[i]];
  return sg_check_hip(hipMemcpy(d.mask, m.data(), (size_t)d.V * sizeof(float), hipMemcpyHostToDevice), "copy");
}

int sg_pbvh_tree_info(PBVH *p, int out[4]) {
  if (!p) return sg_set_error("null PBVH");
  out[0] = (int)p->nodes.size();
  out[1] = (int)p->leaf_node.size();
  out[2] = (int)p->cut.node.size();
  out[3] = p->d.sub_maxn;
  return 0;
}

int sg_pbvh_node_index(PBVH *p, PBVHNode *node) {
  if (!p || !node) return -1;
  return node->index;
}

PBVHNode *sg_pbvh_node_at(PBVH *p, int index) {
  if (!p || index < 0 || index >= (int)p->nodes.size()) return nullptr;
  return &p->nodes[index];
}

int sg_pbvh_node_flag(PBVHNode *node) { return node ? node->flag : 0; }

int sg_pbvh_prim_indices(PBVH *p, int *out, int cap) {
  if (!p) return sg_set_error("null PBVH");
  const int n = (int)p->prim_indices.size();
  for (int i = 0; i < n && i < cap; i++) out[i] = p->prim_indices[i];
  return n;
}

// ---- search ----------------------------------------------------------------

// pbvh_iter_next order (pbvh.c:746-787): children pushed right then left.
static void host_search(PBVH *p, BKE_pbvh_SearchCallback scb, void *data, std::vector<PBVHNode *> &out) {
  out.clear();
  if (p->nodes.empty()) return;
  sync_bbs(p);
  std::vector<int> st{0};
  while (!st.empty()) {
    const int ni = st.back();
    st.pop_back();
    PBVHNode *n = &p->nodes[ni];
    if (scb && !scb(n, data)) continue;
    if (n->flag & PBVH_Leaf) {
      out.push_back(n);
      continue;
    }
    st.push_back(n->children_offset + 1);
    st.push_back(n->children_offset);
  }
}

void BKE_pbvh_search_gather(PBVH *p, BKE_pbvh_SearchCallback scb, void *search_data, PBVHNode ***r_array,
                            int *r_tot) {
  *r_array = nullptr;
  *r_tot = 0;
  if (!p) return;
  std::vector<PBVHNode *> v;
  host_search(p, scb, search_data, v);
  if (v.empty()) return;  // NULL when nothing found (pbvh.c:842-845)
  PBVHNode **arr = (PBVHNode **)malloc(sizeof(PBVHNode *) * v.size());
  memcpy(arr, v.data(), sizeof(PBVHNode *) * v.size());
  *r_array = arr;
  *r_tot = (int)v.size();
}

void BKE_pbvh_search_callback(PBVH *p, BKE_pbvh_SearchCallback scb, void *search_data, BKE_pbvh_HitCallback hcb,
                              void *hit_data) {
  if (!p) return;
  std::vector<PBVHNode *> v;
  host_search(p, scb, search_data, v);
  for (PBVHNode *n : v) hcb(n, hit_data);
}

// ---- marking ---------------------------------------------------------------

void BKE_pbvh_node_mark_update(PBVHNode *node) {  // pbvh.c:1723-1727
  const int bits = PBVH_UpdateNormals | PBVH_UpdateBB | PBVH_UpdateOriginalBB | PBVH_UpdateDrawBuffers | PBVH_UpdateRedraw;
  node->flag |= bits;
  sg_dyn_note_flags(node, bits);
  if (node->owner) node->owner->draw_marks_pending = true;
}
void BKE_pbvh_node_mark_normals_update(PBVHNode *node) {
  node->flag |= PBVH_UpdateNormals;
  sg_dyn_note_flags(node, PBVH_UpdateNormals);
}
void BKE_pbvh_node_mark_redraw(PBVHNode *node) {
  node->flag |= PBVH_UpdateDrawBuffers | PBVH_UpdateRedraw;
  sg_dyn_note_flags(node, PBVH_UpdateDrawBuffers | PBVH_UpdateRedraw);
  if (node->owner) node->owner->draw_marks_pending = true;
}
void BKE_pbvh_vert_mark_update(PBVH *p, int index) {  // pbvh.c:1811-1815
  if (!p || index < 0 || index >= p->totvert || p->host_vmark.empty()) return;
  if (!p->host_vmark[index]) {
    p->host_vmark[index] = 1;
    p->host_mark_list.push_back(index);
  }
}

// ---- updates ---------------------------------------------------------------

// pbvh_flush_bb (pbvh.c:1369-1399) over the host mirror: leaves report and clear
// their bounds flags, internal nodes take the union of their children's boxes (UpdateBB)
// and copy it into orig_vb (UpdateOriginalBB); changed internal nodes are collected.
static int host_flush_bb(PBVH *p, int ni, int flag, std::vector<int> &changed) {
  PBVHNode &n = p->nodes[ni];
  int update = 0;
  if (n.flag & PBVH_Leaf) {
    if (flag & PBVH_UpdateBB) {
      update |= n.flag & PBVH_UpdateBB;
      n.flag &= ~PBVH_UpdateBB;
    }
    if (flag & PBVH_UpdateOriginalBB) {
      update |= n.flag & PBVH_UpdateOriginalBB;
      n.flag &= ~PBVH_UpdateOriginalBB;
    }
    return update;
  }
  update |= host_flush_bb(p, n.children_offset, flag, changed);
  update |= host_flush_bb(p, n.children_offset + 1, flag, changed);
  if (update & PBVH_UpdateBB) {
    const PBVHNode &a = p->nodes[n.children_offset], &b = p->nodes[n.children_offset + 1];
    // update_node_vb (pbvh.c:108-128): BB_reset, then BB_expand_with_bb by both
    // children with min_ff / max_ff
    auto mn = [](float x, float y) { return x < y ? x : y; };
    auto mx = [](float x, float y) { return x > y ? x : y; };
    for (int k = 0; k < 3; k++) {
      n.vb[k] = mn(mn(FLT_MAX, a.vb[k]), b.vb[k]);
      n.vb[3 + k] = mx(mx(-FLT_MAX, a.vb[3 + k]), b.vb[3 + k]);
    }
  }
  if (update & PBVH_UpdateOriginalBB) memcpy(n.orig_vb, n.vb, sizeof(n.vb));
  if (update) changed.push_back(ni);
  return update;
}

void BKE_pbvh_update_bounds(PBVH *p, int flag) {  // pbvh.c:1401-1421
  if (!p || p->nodes.empty() || p->no_device || !p->build_ok) return;
  if (p->type == PBVH_GRIDS) {
    sg_grids_update_bounds(p, flag);
    return;
  }
  if (p->type == PBVH_BMESH) {
    if (!sg_dyn_sync_flags_down(p)) sg_dyn_update_normals_bounds(p, false, true);
    return;
  }
  // the leaves update_search_cb gathers (any bit of `flag`), with the bounds bits each
  // carries
  std::vector<int> leaves, modes;
  for (size_t l = 0; l < p->leaf_node.size(); l++) {
    PBVHNode &n = p->nodes[p->leaf_node[l]];
    if (!(n.flag & flag)) continue;
    if ((flag & PBVH_UpdateRedraw) && (n.flag & PBVH_UpdateRedraw)) n.flag &= ~PBVH_UpdateRedraw;
    const int m = n.flag & flag & (PBVH_UpdateBB | PBVH_UpdateOriginalBB);
    if (m) {
      leaves.push_back((int)l);
      modes.push_back(m);
    }
  }
  if (!(flag & (PBVH_UpdateBB | PBVH_UpdateOriginalBB)) || leaves.empty()) return;
  upload_host_leaf_verts(p, leaves);
  const int both = PBVH_UpdateBB | PBVH_UpdateOriginalBB;
  const bool all_both = std::all_of(modes.begin(), modes.end(), [&](int m) { return m == both; });
  const size_t nl = leaves.size();
  if (all_both) {
    // the brush's case: each listed leaf's box from its vertices into vb and orig_vb, the
    // tree above on the device
    int *dl = (int *)stage(p, nl * sizeof(int));
    if (!dl) return;
    hipMemcpyAsync(dl, leaves.data(), nl * sizeof(int), hipMemcpyHostToDevice, p->stream);
    sg_enqueue_refit_list(p, dl, (int)nl);
    sg_enqueue_flush(p);
    if (sg_check_hip(hipStreamSynchronize(p->stream), "update bounds")) return;
    std::vector<int> changed;
    host_flush_bb(p, 0, flag, changed);  // clears the leaves' flags as the reference does
    p->host_bb_stale = true;
    return;
  }
  // partial flags: the reference's per-node rules on the host mirror, the new vertex
  // boxes computed on the device, the changed boxes written back
  if (sync_bbs(p)) return;
  const size_t off_out = (nl * sizeof(int) + 255) & ~(size_t)255;
  char *st = (char *)stage(p, off_out + nl * 32);
  if (!st) return;
  hipMemcpyAsync(st, leaves.data(), nl * sizeof(int), hipMemcpyHostToDevice, p->stream);
  std::vector<float> boxes(nl * 8);
  if (sg_enqueue_leaf_boxes(p, (const int *)st, (int)nl, (float *)(st + off_out))) return;
  hipMemcpyAsync(boxes.data(), st + off_out, nl * 32, hipMemcpyDeviceToHost, p->stream);
  if (sg_check_hip(hipStreamSynchronize(p->stream), "leaf boxes")) return;
  struct Rec {
    int kind, idx, pad0, pad1;
    float vb[8], ob[8];
  };
  std::vector<Rec> recs;
  auto rec_of = [&](int kind, int idx, const PBVHNode &n) {
    Rec r;
    memset(&r, 0, sizeof(r));
    r.kind = kind;
    r.idx = idx;
    for (int k = 0; k < 3; k++) {
      r.vb[k] = n.vb[k];
      r.vb[4 + k] = n.vb[3 + k];
      r.ob[k] = n.orig_vb[k];
      r.ob[4 + k] = n.orig_vb[3 + k];
    }
    return r;
  };
  for (size_t i = 0; i < nl; i++) {  // pbvh_update_BB_redraw_task_cb (pbvh.c:1206-1227)
    PBVHNode &n = p->nodes[p->leaf_node[leaves[i]]];
    if (modes[i] & PBVH_UpdateBB)
      for (int k = 0; k < 3; k++) {
        n.vb[k] = boxes[8 * i + k];
        n.vb[3 + k] = boxes[8 * i + 4 + k];
      }
    if (modes[i] & PBVH_UpdateOriginalBB) memcpy(n.orig_vb, n.vb, sizeof(n.vb));
    recs.push_back(rec_of(0, leaves[i], n));
  }
  std::vector<int> changed;
  host_flush_bb(p, 0, flag, changed);
  std::vector<char> is_changed(p->nodes.size(), 0);
  for (int ni : changed) {
    is_changed[ni] = 1;
    recs.push_back(rec_of(1, ni, p->nodes[ni]));
  }
  for (int l : leaves) is_changed[p->leaf_node[l]] = 1;
  for (size_t c = 0; c < p->cut.node.size(); c++) {  // the search's subtree boxes are their roots' boxes
    const int root = p->cut.node[c];
    if (!is_changed[root]) continue;
    recs.push_back(rec_of(2, (int)c, p->nodes[root]));
    if (p->nodes[root].flag & PBVH_Leaf) recs.push_back(rec_of(1, root, p->nodes[root]));
  }
  Rec *dr = (Rec *)stage(p, recs.size() * sizeof(Rec));
  if (!dr) return;
  hipMemcpyAsync(dr, recs.data(), recs.size() * sizeof(Rec), hipMemcpyHostToDevice, p->stream);
  sg_enqueue_scatter_boxes(p, dr, (int)recs.size());
  sg_check_hip(hipStreamSynchronize(p->stream), "update bounds");
}

void BKE_pbvh_update_normals(PBVH *p, SubdivCCG *subdiv_ccg) {  // pbvh.c:4559-4587
  if (!p || p->nodes.empty() || p->no_device || !p->build_ok) return;
  if (p->type == PBVH_GRIDS) {
    sg_grids_update_normals(p, subdiv_ccg);
    return;
  }
  if (p->type == PBVH_BMESH) {  // pbvh_bmesh_normals_update (pbvh_bmesh.c:1635-1656)
    if (!sg_dyn_sync_flags_down(p)) sg_dyn_update_normals_bounds(p, true, false);
    return;
  }
  std::vector<int> leaves = flagged_leaves(p, PBVH_UpdateNormals);
  if (leaves.empty()) return;
  upload_host_leaf_verts(p, leaves);
  // the vertices marked on the host join the device's marks (those of
  // BKE_pbvh_vert_coords_apply); only the marks are uploaded, not a V-sized array
  std::vector<int> marks(p->host_mark_list.size());
  for (size_t i = 0; i < marks.size(); i++) marks[i] = p->perm_new_of_old[p->host_mark_list[i]];
  const size_t nm = marks.size(), nl = leaves.size();
  const size_t off_l = (nm * sizeof(int) + 255) & ~(size_t)255;
  char *st = (char *)stage(p, off_l + nl * sizeof(int));
  if (!st) return;
  if (nm) hipMemcpyAsync(st, marks.data(), nm * sizeof(int), hipMemcpyHostToDevice, p->stream);
  hipMemcpyAsync(st + off_l, leaves.data(), nl * sizeof(int), hipMemcpyHostToDevice, p->stream);
  sg_enqueue_set_marks(p, (const int *)st, (int)nm);
  sg_enqueue_normals_list(p, (const int *)(st + off_l), (int)nl);
  if (sg_check_hip(hipStreamSynchronize(p->stream), "update normals")) return;
  // marks of uniq verts of the processed leaves are consumed (pbvh.c:1084-1087)
  std::vector<char> done(p->leaf_node.size(), 0);
  for (int l : leaves) {
    done[l] = 1;
    p->nodes[p->leaf_node[l]].flag &= ~PBVH_UpdateNormals;
  }
  std::vector<int> keep;
  for (size_t i = 0; i < nm; i++) {
    const int leaf = (int)(std::upper_bound(p->leaf_vbeg.begin(), p->leaf_vbeg.end(), marks[i]) -
                           p->leaf_vbeg.begin()) - 1;
    if (done[leaf])
      p->host_vmark[p->host_mark_list[i]] = 0;
    else
      keep.push_back(p->host_mark_list[i]);
  }
  p->host_mark_list.swap(keep);
  p->host_normals_stale = true;
}

// pbvh.c:4707-4749.  Positions that differ bitwise from the new ones are replaced and
// their vertices marked for the next normal update; every node is marked
// (BKE_pbvh_node_mark_update) and the bounds refit with UpdateBB | UpdateOriginalBB.
// The reference's BKE_mesh_calc_normals_looptri call here accumulates into a
// temporary that it frees (mesh_normals.c:455-507): the vertex normals themselves
// change only at the next BKE_pbvh_update_normals, as here.
void BKE_pbvh_vert_coords_apply(PBVH *p, const float (*vertCos)[3], int totvert) {
  if (p && p->type != PBVH_FACES) { sg_set_error("faces only: BKE_pbvh_vert_coords_apply"); return; }
  if (!p || p->nodes.empty() || p->no_device || !p->build_ok) return;
  if (totvert != p->totvert) {
    sg_set_error("PBVH: Given deforming vcos number does not match PBVH vertex number!");
    return;
  }
  if (!p->verts_owned && p->host_verts) {  // duplicate on first deform (pbvh.c:4714-4726)
    sync_verts(p);
    MVert *dup = (MVert *)malloc(sizeof(MVert) * (size_t)totvert);
    memcpy(dup, p->host_verts, sizeof(MVert) * (size_t)totvert);
    p->host_verts = dup;
    p->verts_owned = true;
  }
  float *src = (float *)stage(p, 3 * (size_t)totvert * sizeof(float));
  if (!src) return;
  hipMemcpyAsync(src, vertCos, 3 * (size_t)totvert * sizeof(float), hipMemcpyHostToDevice, p->stream);
  sg_enqueue_coords_apply(p, src);
  sg_enqueue_flush(p);
  if (sg_check_hip(hipStreamSynchronize(p->stream), "coords apply")) return;
  if (p->host_verts && !p->host_verts_stale)
    for (int v = 0; v < totvert; v++) memcpy(p->host_verts[v].co, vertCos[v], 12);
  for (PBVHNode &n : p->nodes) BKE_pbvh_node_mark_update(&n);
  std::vector<int> changed;
  host_flush_bb(p, 0, PBVH_UpdateBB | PBVH_UpdateOriginalBB, changed);  // the flags the refit consumed
  p->host_bb_stale = true;
}

float (*BKE_pbvh_vert_coords_alloc(PBVH *p))[3] {
  if (!p || p->totvert == 0) return nullptr;
  float(*out)[3] = (float(*)[3])calloc((size_t)p->totvert, sizeof(float[3]));
  if (read_aos(p, 0, &out[0][0])) {
    free(out);
    return nullptr;
  }
  return out;
}

// ---- node access -------------------------------------------------------------

void BKE_pbvh_node_get_BB(PBVHNode *node, float bb_min[3], float bb_max[3]) {
  sync_bbs(node->owner);
  memcpy(bb_min, node->vb, 12);
  memcpy(bb_max, node->vb + 3, 12);
}

void BKE_pbvh_node_get_original_BB(PBVHNode *node, float bb_min[3], float bb_max[3]) {
  sync_bbs(node->owner);
  memcpy(bb_min, node->orig_vb, 12);
  memcpy(bb_max, node->orig_vb + 3, 12);
}

void BKE_pbvh_node_num_verts(PBVH *p, PBVHNode *node, int *r_uniquevert, int *r_totvert) {
  if (p && p->type == PBVH_BMESH) {  // pbvh.c:1846-1851: unique + other vertices
    const int u = std::max(sg_pbvh_bmesh_node_unique_verts(node, nullptr, 0), 0);
    const int o = std::max(sg_pbvh_bmesh_node_other_verts(node, nullptr, 0), 0);
    if (r_totvert) *r_totvert = u + o;
    if (r_uniquevert) *r_uniquevert = u;
    return;
  }
  if (p && p->type == PBVH_GRIDS) {  // pbvh.c:1836-1844: every grid element is its node's own
    const int tot = node->totprim * p->gridkey.grid_area;
    if (r_totvert) *r_totvert = tot;
    if (r_uniquevert) *r_uniquevert = tot;
    return;
  }
  if (r_totvert) *r_totvert = node->uniq_verts + node->face_verts;
  if (r_uniquevert) *r_uniquevert = node->uniq_verts;
}

void BKE_pbvh_node_get_verts(PBVH *p, PBVHNode *node, const int **r_vert_indices, MVert **r_verts) {
  if (r_vert_indices) *r_vert_indices = node->vert_indices;
  if (r_verts) *r_verts = BKE_pbvh_get_verts(p);
}

bool BKE_pbvh_node_fully_hidden_get(PBVHNode *node) {
  return (node->flag & PBVH_Leaf) && (node->flag & PBVH_FullyHidden);
}

MVert *BKE_pbvh_get_verts(const PBVH *cp) {
  PBVH *p = const_cast<PBVH *>(cp);
  if (!p) return nullptr;
  sync_verts(p);
  return p->host_verts;
}

const float (*DUNE_pbvh_get_vert_normals(const PBVH *cp))[3] {
  PBVH *p = const_cast<PBVH *>(cp);
  if (!p || p->type != PBVH_FACES) return nullptr;
  sync_normals(p);
  return (const float(*)[3])p->host_normals.data();
}

// ---- proxies (pbvh.c:4757-4818) ----------------------------------------------

// A zeroed block of nf floats for a proxy of a node in p->nodes: carved from the
// pinned proxy arena (a new 16 MiB chunk when the last one is full), so the combine
// uploads it without a host pack; from the heap when pinned memory is unavailable.
static float *proxy_alloc(PBVH *p, size_t nf, bool in_tree) {
  const size_t nfa = (nf + 15) & ~(size_t)15;  // 64-byte blocks
  if (in_tree && !p->no_device) {
    if (p->prox_chunks.empty() || p->prox_chunks.back().cap - p->prox_chunks.back().used < nfa) {
      const size_t cap = std::max<size_t>(nfa, (size_t)1 << 22);
      void *h = nullptr;
      if (hipHostMalloc(&h, cap * sizeof(float), hipHostMallocDefault) == hipSuccess)
        p->prox_chunks.push_back({(float *)h, cap, 0});
      else
        (void)hipGetLastError();
    }
    if (!p->prox_chunks.empty() && p->prox_chunks.back().cap - p->prox_chunks.back().used >= nfa) {
      SgProxChunk &c = p->prox_chunks.back();
      float *r = c.host + c.used;
      c.used += nfa;
      memset(r, 0, nf * sizeof(float));
      return r;
    }
  }
  return (float *)calloc(nf, sizeof(float));
}

bool sg_proxy_in_arena(const PBVH *p, const void *ptr) {
  for (const SgProxChunk &c : p->prox_chunks)
    if (ptr >= (const void *)c.host && ptr < (const void *)(c.host + c.cap)) return true;
  return false;
}

PBVHProxyNode *BKE_pbvh_node_add_proxy(PBVH *p, PBVHNode *node) {
  if (p && p->type != PBVH_FACES) { sg_set_error("faces only: proxies"); return nullptr; }
  const int index = node->proxy_count;
  const bool in_tree = node >= p->nodes.data() && node < p->nodes.data() + p->nodes.size();
  if (index == 0 && in_tree) p->proxy_nodes.push_back((int)(node - p->nodes.data()));  // the combine's worklist
  node->proxy_count++;
  node->proxies = (PBVHProxyNode *)realloc(node->proxies, node->proxy_count * sizeof(PBVHProxyNode));
  int totverts = 0;
  BKE_pbvh_node_num_verts(p, node, &totverts, nullptr);
  node->proxies[index].co = (float(*)[3])proxy_alloc(p, 3 * (size_t)std::max(totverts, 1), in_tree);
  return node->proxies + index;
}

void BKE_pbvh_node_free_proxies(PBVHNode *node) {
  for (int i = 0; i < node->proxy_count; i++) {
    // arena blocks are reclaimed together once no proxy is left (the combine)
    if (!(node->owner && sg_proxy_in_arena(node->owner, node->proxies[i].co))) free(node->proxies[i].co);
    node->proxies[i].co = nullptr;
  }
  free(node->proxies);
  node->proxies = nullptr;
  node->proxy_count = 0;
}

void BKE_pbvh_gather_proxies(PBVH *p, PBVHNode ***r_array, int *r_tot) {
  // the nodes holding proxies in node order (the worklist add_proxy keeps, instead of
  // a walk over every node)
  std::vector<PBVHNode *> v;
  std::sort(p->proxy_nodes.begin(), p->proxy_nodes.end());
  p->proxy_nodes.erase(std::unique(p->proxy_nodes.begin(), p->proxy_nodes.end()), p->proxy_nodes.end());
  for (int i : p->proxy_nodes) {
    if (p->nodes[i].proxy_count > 0) v.push_back(&p->nodes[i]);
  }
  *r_array = nullptr;
  *r_tot = 0;
  if (v.empty()) return;
  *r_array = (PBVHNode **)malloc(sizeof(PBVHNode *) * v.size());
  memcpy(*r_array, v.data(), sizeof(PBVHNode *) * v.size());
  *r_tot = (int)v.size();
}

// ---- frustum (pbvh.c:4507-4557) -----------------------------------------------

static int frustum_aabb(const float *bmin, const float *bmax, const PBVHFrustumPlanes *f) {
  int ret = 0;  // 0 inside, 1 intersect, 2 outside
  for (int i = 0; i < f->num_planes; i++) {
    float vmin[3], vmax[3];
    for (int a = 0; a < 3; a++) {
      if (f->planes[i][a] < 0) {
        vmin[a] = bmin[a];
        vmax[a] = bmax[a];
      } else {
        vmin[a] = bmax[a];
        vmax[a] = bmin[a];
      }
    }
    const float *pl = f->planes[i];
    if (pl[0] * vmin[0] + pl[1] * vmin[1] + pl[2] * vmin[2] + pl[3] < 0) return 2;
    if (pl[0] * vmax[0] + pl[1] * vmax[1] + pl[2] * vmax[2] + pl[3] <= 0) ret = 1;
  }
  return ret;
}

bool BKE_pbvh_node_frustum_contain_AABB(PBVHNode *node, void *data) {
  sync_bbs(node->owner);
  return frustum_aabb(node->vb, node->vb + 3, (const PBVHFrustumPlanes *)data) != 2;
}

bool BKE_pbvh_node_frustum_exclude_AABB(PBVHNode *node, void *data) {
  sync_bbs(node->owner);
  return frustum_aabb(node->vb, node->vb + 3, (const PBVHFrustumPlanes *)data) != 0;
}

}  // extern "C"
