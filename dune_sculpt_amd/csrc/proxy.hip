// Proxies combined on the GPU.
//
// A proxy is a per-leaf buffer of displacements, one float3 per uniq vertex of the
// leaf, zeroed when added (BKE_pbvh_node_add_proxy, kernel/intern/pbvh.c:4757-4776),
// which a brush written against the reference's API fills on the host.  The sculpt
// code that folds them into the mesh (sculpt_combine_proxies) is absent from the
// reference; Blender's published version does, for every node BKE_pbvh_gather_proxies
// returns (pbvh.c:4791-4818), for each uniq vertex: val = original coordinate (when the
// tool works from the original mesh) or the current one, val += each proxy in order
// (add_v3_v3), then SCULPT_clip (a plain copy without mirror clipping), the vertex is
// flagged for the normal update, and the node's proxies are freed
// (BKE_pbvh_node_free_proxies, :4778-4789).  Here the proxies live in pinned host
// chunks (BKE_pbvh_node_add_proxy carves them there), so the caller's fill is already
// in memory the copy engine reads: each chunk goes to HBM in one copy, and one kernel
// writes the vertices, marks them for normals and saves each leaf for the open undo
// step first; the normals, leaf boxes and search boxes follow on the GPU.  The original
// coordinates are those of the open undo step (the state at sg_undo_begin): the
// reference reads them from its undo nodes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sg_internal.h"
#include "sg_kernels.h"
#include "sg_device.h"

// One wave per leaf holding proxies.  tab: leaf, proxy count, index of its first
// proxy in poff; poff: each proxy's float offset in prox ([uniq][3] per proxy).
__global__ __launch_bounds__(256) void k_combine_proxies(SgDevView d, const int4 *tab, const long long *poff,
                                                         const float *prox, int n, int use_orco) {
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  for (int e = gw; e < n; e += nw) {
    const int4 t = tab[e];
    const int leaf = t.x, np = t.y, pi = t.z;
    const int2 vr = make_int2(d.leaf_vbeg[leaf], d.leaf_vbeg[leaf + 1]);
    const int nu = vr.y - vr.x;
    // original coordinates: the step's record when an earlier kernel of the open step
    // saved this leaf; otherwise the leaf is unchanged since the step began
    const float *orig = nullptr;
    const bool rec = undo_active(d);
    if (rec && use_orco) {
      const SgUndoCtl *u = d.undo;
      if (u->stamp[leaf] == u->step) {
        const int4 r = u->tab[u->slot[leaf]];
        orig = u->pos + 3 * (((size_t)(unsigned)r.w << 32) | (unsigned)r.z);
      }
    }
    if (rec) undo_save_leaf(d, leaf, vr, d.px, d.py, d.pz);  // before this wave writes the leaf
    for (int i = lane; i < nu; i += SG_WAVE) {
      const int v = vr.x + i;
      float x, y, z;
      if (orig) {
        x = orig[3 * (size_t)i];
        y = orig[3 * (size_t)i + 1];
        z = orig[3 * (size_t)i + 2];
      } else {
        x = d.px[v];
        y = d.py[v];
        z = d.pz[v];
      }
      for (int q = 0; q < np; q++) {
        const float *pq = prox + poff[pi + q] + 3 * (size_t)i;
        x = x + pq[0];
        y = y + pq[1];
        z = z + pq[2];
      }
      d.px[v] = x;
      d.py[v] = y;
      d.pz[v] = z;
      d.qx[v] = x;
      d.qy[v] = y;
      d.qz[v] = z;
      d.vmark[v] = 1;  // ME_VERT_PBVH_UPDATE: the normal update takes it
    }
    if (lane == 0) d.draw_dirty[d.leaf_cut[leaf]] = 1;
  }
}

int sg_enqueue_normals_list(PBVH *p, const int *leaves_dev, int n);
int sg_enqueue_refit_list(PBVH *p, const int *leaves_dev, int n);
bool sg_proxy_in_arena(const PBVH *p, const void *ptr);

extern "C" {

// Folds every node's proxies into the mesh and frees them.  use_orco != 0 starts from
// the original coordinates (those of the open undo step), else from the current ones.
// Returns the number of leaves changed.
int sg_pbvh_combine_proxies(PBVH *p, int use_orco) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: proxies");
  if (!p || p->no_device || !p->build_ok) return sg_set_error("no device mesh");
  // nodes holding proxies, in node order (BKE_pbvh_gather_proxies)
  std::vector<PBVHNode *> nodes;
  {
    PBVHNode **arr = nullptr;
    int tot = 0;
    BKE_pbvh_gather_proxies(p, &arr, &tot);
    nodes.assign(arr, arr + tot);
    free(arr);
  }
  if (nodes.empty()) return 0;
  for (PBVHNode *np : nodes)
    if (np->leaf < 0) return sg_set_error("a proxy on an internal node");
  // device layout: the used part of each arena chunk, in chunk order, then the proxies
  // that came from the heap, packed
  std::vector<size_t> cbase(p->prox_chunks.size());
  size_t total = 0;
  for (size_t c = 0; c < p->prox_chunks.size(); c++) {
    cbase[c] = total;
    total += p->prox_chunks[c].used;
  }
  const size_t arena_total = total;
  std::vector<int> leaves;
  std::vector<int4> tab;
  std::vector<long long> poff;
  std::vector<const float *> heap_src;
  std::vector<size_t> heap_len;
  for (PBVHNode *np : nodes) {
    const PBVHNode &n = *np;
    tab.push_back(make_int4(n.leaf, n.proxy_count, (int)poff.size(), 0));
    leaves.push_back(n.leaf);
    const size_t len = (size_t)n.uniq_verts * 3;
    for (int q = 0; q < n.proxy_count; q++) {
      const float *co = &n.proxies[q].co[0][0];
      long long off = -1;
      for (size_t c = 0; c < p->prox_chunks.size() && off < 0; c++) {
        const SgProxChunk &ch = p->prox_chunks[c];
        if (co >= ch.host && co < ch.host + ch.cap) off = (long long)(cbase[c] + (size_t)(co - ch.host));
      }
      if (off < 0) {  // a heap block: packed after the chunks
        off = (long long)total;
        heap_src.push_back(co);
        heap_len.push_back(len);
        total += len;
      }
      poff.push_back(off);
    }
  }
  SgDev &d = p->d;
  const int nl = (int)leaves.size(), npx = (int)poff.size();
  const size_t need = std::max<size_t>(total, 1), heap_need = total - arena_total;
  int rc = 0;
  if (need > d.prox_cap) {  // grow-only device copy of the proxies
    if (d.prox_dev) (void)hipFree(d.prox_dev);
    d.prox_dev = nullptr;
    d.prox_cap = 0;
    const size_t cap = need + need / 2;
    rc = sg_check_hip(hipMalloc(&d.prox_dev, cap * sizeof(float)), "hipMalloc");
    if (rc) return rc;
    d.prox_cap = cap;
  }
  if (heap_need > d.prox_host_cap) {  // grow-only pinned pack of the heap proxies
    if (d.prox_host) (void)hipHostFree(d.prox_host);
    d.prox_host = nullptr;
    d.prox_host_cap = 0;
    const size_t cap = heap_need + heap_need / 2;
    rc = sg_check_hip(hipHostMalloc(&d.prox_host, cap * sizeof(float), hipHostMallocDefault), "hipHostMalloc");
    if (rc) return rc;
    d.prox_host_cap = cap;
  }
  if (nl > d.prox_tab_cap) {
    if (d.prox_tab) (void)hipFree(d.prox_tab);
    if (d.prox_leaves) (void)hipFree(d.prox_leaves);
    d.prox_tab = nullptr;
    d.prox_leaves = nullptr;
    d.prox_tab_cap = 0;
    const int cap = nl + nl / 2;
    rc = sg_check_hip(hipMalloc(&d.prox_tab, (size_t)cap * sizeof(int4)), "hipMalloc");
    if (!rc) rc = sg_check_hip(hipMalloc(&d.prox_leaves, (size_t)cap * sizeof(int)), "hipMalloc");
    if (rc) return rc;
    d.prox_tab_cap = cap;
  }
  if (npx > d.prox_off_cap) {
    if (d.prox_off) (void)hipFree(d.prox_off);
    d.prox_off = nullptr;
    d.prox_off_cap = 0;
    const int cap = npx + npx / 2;
    rc = sg_check_hip(hipMalloc(&d.prox_off, (size_t)cap * sizeof(long long)), "hipMalloc");
    if (rc) return rc;
    d.prox_off_cap = cap;
  }
  for (size_t c = 0; c < p->prox_chunks.size() && !rc; c++) {
    const SgProxChunk &ch = p->prox_chunks[c];
    if (ch.used)
      rc = sg_check_hip(hipMemcpyAsync(d.prox_dev + cbase[c], ch.host, ch.used * sizeof(float),
                                       hipMemcpyHostToDevice, p->stream),
                        "copy");
  }
  if (!rc && heap_need) {
    size_t w = 0;
    for (size_t h = 0; h < heap_src.size(); h++) {
      memcpy(d.prox_host + w, heap_src[h], heap_len[h] * sizeof(float));
      w += heap_len[h];
    }
    rc = sg_check_hip(hipMemcpyAsync(d.prox_dev + arena_total, d.prox_host, heap_need * sizeof(float),
                                     hipMemcpyHostToDevice, p->stream),
                      "copy");
  }
  if (!rc)
    rc = sg_check_hip(hipMemcpyAsync(d.prox_tab, tab.data(), tab.size() * sizeof(int4), hipMemcpyHostToDevice,
                                     p->stream),
                      "copy");
  if (!rc)
    rc = sg_check_hip(hipMemcpyAsync(d.prox_off, poff.data(), poff.size() * sizeof(long long),
                                     hipMemcpyHostToDevice, p->stream),
                      "copy");
  if (!rc)
    rc = sg_check_hip(hipMemcpyAsync(d.prox_leaves, leaves.data(), leaves.size() * sizeof(int),
                                     hipMemcpyHostToDevice, p->stream),
                      "copy");
  if (!rc) {
    const int blocks = std::max(1, std::min((nl + 3) / 4, 8192));
    hipLaunchKernelGGL(k_combine_proxies, dim3(blocks), dim3(256), 0, p->stream, sg_view(p), d.prox_tab, d.prox_off,
                       d.prox_dev, nl, use_orco ? 1 : 0);
    rc = sg_check_hip(hipGetLastError(), "combine launch");
  }
  // normals of the marked vertices, then leaf boxes and search boxes, then the tree
  if (!rc) rc = sg_enqueue_normals_list(p, d.prox_leaves, nl);
  if (!rc) rc = sg_enqueue_refit_list(p, d.prox_leaves, nl);
  if (!rc) rc = sg_enqueue_flush(p);
  // the copies have read the arena once the stream is idle
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "combine proxies");
  if (rc) return rc;
  p->nodes_dirty_deferred = false;
  p->host_bb_stale = p->host_verts_stale = p->host_normals_stale = true;
  for (PBVHNode *np : nodes) BKE_pbvh_node_free_proxies(np);
  p->proxy_nodes.clear();
  // every proxy of the tree was gathered and is freed: the arena starts over
  for (SgProxChunk &c : p->prox_chunks) c.used = 0;
  return nl;
}

}  // extern "C"
