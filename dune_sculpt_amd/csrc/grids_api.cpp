// Host entry points of PBVH_GRIDS: BKE_pbvh_build_grids and the grid accessors
// (kernel/intern/pbvh.c:331-361, :598-643, :1605-1712, :1831-1910, :4973), and the
// grid branches of the dab and drop-in update entry points, which api.cpp calls when
// the PBVH's type is PBVH_GRIDS.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "sg_grids.h"
#include "sg_kernels.h"

void sg_build_tree_grids(PBVH *p);  // pbvh_build.cpp

namespace {

// paint_is_grid_face_hidden: a quad is hidden when any of its four corners is (absent
// from the reference; upstream Blender's rule)
inline bool grid_quad_hidden(const BLI_bitmap *gh, int gs, int x, int y) {
  auto test = [&](int i) { return (gh[i >> 5] >> (i & 31)) & 1u; };
  return test(y * gs + x) || test(y * gs + x + 1) || test((y + 1) * gs + x + 1) || test((y + 1) * gs + x);
}

std::vector<int> flagged_leaves(PBVH *p, int flag) {
  std::vector<int> out;
  for (size_t l = 0; l < p->leaf_node.size(); l++)
    if (p->nodes[p->leaf_node[l]].flag & flag) out.push_back((int)l);
  return out;
}

int upload_list(PBVH *p, const std::vector<int> &v) {
  if (v.empty()) return 0;
  return sg_check_hip(hipMemcpyAsync(p->gd->hlist, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, p->stream),
                      "copy");
}

}  // namespace

extern "C" {

int BKE_pbvh_count_grid_quads(BLI_bitmap **grid_hidden, const int *grid_indices, int totgrid, int gridsize) {
  const int gridarea = (gridsize - 1) * (gridsize - 1);
  int totquad = 0;
  for (int i = 0; i < totgrid; i++) {
    const BLI_bitmap *gh = grid_hidden ? grid_hidden[grid_indices[i]] : nullptr;
    if (gh) {
      for (int y = 0; y < gridsize - 1; y++)
        for (int x = 0; x < gridsize - 1; x++)
          if (!grid_quad_hidden(gh, gridsize, x, y)) totquad++;
    } else {
      totquad += gridarea;
    }
  }
  return totquad;
}

void BKE_pbvh_build_grids(PBVH *p, CCGElem **grids, int totgrid, CCGKey *key, void **gridfaces, DMFlagMat *flagmats,
                          BLI_bitmap **grid_hidden) {
  if (!p || !grids || !key || totgrid <= 0 || key->grid_size < 2) {
    sg_set_error("BKE_pbvh_build_grids: empty grids");
    return;
  }
  if (!p->nodes.empty()) {
    sg_set_error("BKE_pbvh_build_grids: the PBVH is already built");
    return;
  }
  const int gs = key->grid_size;
  p->type = PBVH_GRIDS;
  p->grids = grids;
  p->gridfaces = gridfaces;
  p->grid_flag_mats = flagmats;
  p->totgrid = totgrid;
  p->gridkey = *key;
  p->grid_hidden = grid_hidden;
  p->leaf_limit = std::max(p->leaf_limit / (gs * gs), 1);
  p->totprim = totgrid;
  // material keys for partition_indices_material / leaf_needs_material_split
  p->tri_mat.clear();
  if (flagmats) {
    std::vector<int> keys(totgrid);
    bool differ = false;
    for (int g = 0; g < totgrid; g++) {
      keys[g] = ((flagmats[g].flag & ME_SMOOTH) << 16) | (flagmats[g].mat_nr & 0xffff);
      differ |= keys[g] != keys[0];
    }
    if (differ) p->tri_mat.swap(keys);
  }
  sg_build_tree_grids(p);
  p->build_ok = false;
  if (!p->no_device) p->build_ok = sg_grids_build_device(p) == 0;
  p->host_bb_stale = false;
}

void DUNE_pbvh_subdiv_cgg_set(PBVH *p, SubdivCCG *s) {
  if (!p) return;
  p->subdiv_ccg = s;
  if (s && p->type == PBVH_GRIDS && p->gd) sg_grids_attach_topology(p, s);
}

void BKE_pbvh_subdiv_cgg_set(PBVH *p, SubdivCCG *s) { DUNE_pbvh_subdiv_cgg_set(p, s); }

void BKE_pbvh_node_get_grids(PBVH *p, PBVHNode *node, int **r_grid_indices, int *r_totgrid, int *r_maxgrid,
                             int *r_gridsize, CCGElem ***r_griddata) {
  const bool grids = p && node && p->type == PBVH_GRIDS;
  if (r_grid_indices) *r_grid_indices = grids ? &p->prim_indices[node->prim_off] : nullptr;
  if (r_totgrid) *r_totgrid = grids ? node->totprim : 0;
  if (r_maxgrid) *r_maxgrid = grids ? p->totgrid : 0;
  if (r_gridsize) *r_gridsize = grids ? p->gridkey.grid_size : 0;
  if (r_griddata) *r_griddata = grids ? p->grids : nullptr;
}

BLI_bitmap **BKE_pbvh_grid_hidden(const PBVH *p) { return p && p->type == PBVH_GRIDS ? p->grid_hidden : nullptr; }
const CCGKey *BKE_pbvh_get_grid_key(const PBVH *p) { return p && p->type == PBVH_GRIDS ? &p->gridkey : nullptr; }
CCGElem **BKE_pbvh_get_grids(const PBVH *p) { return p && p->type == PBVH_GRIDS ? p->grids : nullptr; }
int BKE_pbvh_get_grid_num_vertices(const PBVH *p) {
  return p && p->type == PBVH_GRIDS ? p->totgrid * p->gridkey.grid_area : 0;
}
int BKE_pbvh_get_grid_num_faces(const PBVH *p) {
  return p && p->type == PBVH_GRIDS ? p->totgrid * (p->gridkey.grid_size - 1) * (p->gridkey.grid_size - 1) : 0;
}

void BKE_pbvh_get_grid_updates(PBVH *p, bool clear, void ***r_gridfaces, int *r_totface) {
  *r_gridfaces = nullptr;
  *r_totface = 0;
  if (!p || p->type != PBVH_GRIDS || p->nodes.empty() || !p->gridfaces) return;
  std::vector<void *> faces;
  // pbvh_iter order: depth first, left child first
  std::vector<int> st{0};
  while (!st.empty()) {
    const int ni = st.back();
    st.pop_back();
    PBVHNode &n = p->nodes[ni];
    if (n.flag & PBVH_UpdateNormals) {
      for (int i = 0; i < n.totprim; i++) faces.push_back(p->gridfaces[p->prim_indices[n.prim_off + i]]);
      if (clear) n.flag &= ~PBVH_UpdateNormals;
    }
    if (!(n.flag & PBVH_Leaf)) {
      st.push_back(n.children_offset + 1);
      st.push_back(n.children_offset);
    }
  }
  std::vector<void *> uniq;
  {
    std::vector<void *> sorted = faces;
    std::sort(sorted.begin(), sorted.end());
    std::vector<uint8_t> taken(sorted.size(), 0);
    for (void *f : faces) {
      const size_t i = std::lower_bound(sorted.begin(), sorted.end(), f) - sorted.begin();
      if (!taken[i]) {
        taken[i] = 1;
        uniq.push_back(f);
      }
    }
  }
  if (uniq.empty()) return;
  void **out = (void **)malloc(sizeof(void *) * uniq.size());
  memcpy(out, uniq.data(), sizeof(void *) * uniq.size());
  *r_gridfaces = out;
  *r_totface = (int)uniq.size();
}

int sg_pbvh_grids_sync_host(PBVH *p) {
  if (!p || p->type != PBVH_GRIDS || !p->gd) return sg_set_error("not a PBVH over grids on the device");
  const CCGKey &key = p->gridkey;
  const size_t E = (size_t)p->gd->E;
  std::vector<float> co(3 * E), no(key.has_normals ? 3 * E : 0), mk(key.has_mask ? E : 0);
  if (sg_grids_download(p, co.data(), key.has_normals ? no.data() : nullptr, key.has_mask ? mk.data() : nullptr))
    return -1;
  for (int g = 0; g < p->totgrid; g++) {
    unsigned char *base = (unsigned char *)p->grids[g];
    for (int j = 0; j < key.grid_area; j++) {
      const size_t v = (size_t)g * key.grid_area + j;
      unsigned char *el = base + (size_t)j * key.elem_size;
      memcpy(el, &co[3 * v], 12);
      if (key.has_normals) memcpy(el + key.normal_offset, &no[3 * v], 12);
      if (key.has_mask) memcpy(el + key.mask_offset, &mk[v], 4);
    }
  }
  return 0;
}

int sg_pbvh_grids_upload(PBVH *p) {
  if (!p || p->type != PBVH_GRIDS || !p->gd) return sg_set_error("not a PBVH over grids on the device");
  const CCGKey &key = p->gridkey;
  const size_t E = (size_t)p->gd->E;
  std::vector<float> co(3 * E), no(key.has_normals ? 3 * E : 0), mk(key.has_mask ? E : 0);
  for (int g = 0; g < p->totgrid; g++) {
    const unsigned char *base = (const unsigned char *)p->grids[g];
    for (int j = 0; j < key.grid_area; j++) {
      const size_t v = (size_t)g * key.grid_area + j;
      const unsigned char *el = base + (size_t)j * key.elem_size;
      memcpy(&co[3 * v], el, 12);
      if (key.has_normals) memcpy(&no[3 * v], el + key.normal_offset, 12);
      if (key.has_mask) memcpy(&mk[v], el + key.mask_offset, 4);
    }
  }
  return sg_grids_upload(p, co.data(), key.has_normals ? no.data() : nullptr, key.has_mask ? mk.data() : nullptr);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// grid branches of api.cpp's entry points
// ---------------------------------------------------------------------------
int sg_grids_dab_apply(PBVH *p, const SgDabParams *dab) {
  if (!p->gd) return sg_set_error("grids: no device layout");
  if (!p->gd->topo_of && p->subdiv_ccg) {
    if (int rc = sg_grids_attach_topology(p, p->subdiv_ccg)) return rc;
  }
  const SgDabK k = sg_make_dab(dab, 0);
  const float *curve = nullptr;
  if (dab->curve_preset == SG_CURVE_CUSTOM && dab->curve_table) {
    if (sg_check_hip(hipMemcpyAsync(p->gd->curve, dab->curve_table, (SG_CM_TABLE + 1) * sizeof(float),
                                    hipMemcpyHostToDevice, p->stream),
                     "copy"))
      return -1;
    curve = p->gd->curve;
  }
  if (int rc = sg_grids_enqueue_dab(p, k, curve)) return rc;
  if (int rc = sg_grids_enqueue_flush(p)) return rc;
  if (int rc = sg_grids_fetch_last(p)) return rc;
  p->last_tool = dab->sculpt_tool;
  p->last_learned = false;
  p->dab_count++;
  p->host_bb_stale = true;
  return 0;
}

// the last dab's counters (after the stream drains), in the faces counters' layout
int sg_grids_fetch_last(PBVH *p) {
  long long ctr[SGG_CTR_COUNT];
  if (sg_check_hip(hipMemcpyAsync(ctr, p->gd->ctr_last, sizeof(ctr), hipMemcpyDeviceToHost, p->stream), "copy"))
    return -1;
  if (sg_check_hip(hipStreamSynchronize(p->stream), "grid dab")) return -1;
  for (int i = 0; i < SG_CTR_COUNT; i++) p->last_ctr[i] = 0;
  p->last_ctr[SG_CTR_NGATHER] = ctr[SGG_NGATHER];
  p->last_ctr[SG_CTR_NAFF] = ctr[SGG_NAFF];
  p->last_ctr[SG_CTR_AREA_X] = ctr[SGG_AREA_X];
  p->last_ctr[SG_CTR_AREA_Y] = ctr[SGG_AREA_Y];
  p->last_ctr[SG_CTR_AREA_Z] = ctr[SGG_AREA_Z];
  p->last_ctr[SG_CTR_NAREA] = ctr[SGG_NAREA];
  p->last_ctr[20] = ctr[SGG_NLEAVES];
  p->last_ctr[21] = ctr[SGG_NFACES];
  p->last_ctr[22] = ctr[SGG_NITEMS];
  return 0;
}

int sg_grids_last_affected(PBVH *p, int *out, int cap) {
  const SgGridDev &d = *p->gd;
  std::vector<int> stamp((size_t)d.E), grid_old(d.G);
  int ser[2] = {0, 0};
  if (sg_grids_read_stamps(p, stamp.data())) return -1;
  if (sg_check_hip(hipMemcpy(ser, d.dser, sizeof(ser), hipMemcpyDeviceToHost), "copy")) return -1;
  if (sg_check_hip(hipMemcpy(grid_old.data(), d.grid_old, d.G * sizeof(int), hipMemcpyDeviceToHost), "copy"))
    return -1;
  std::vector<int> ids;
  for (int pos = 0; pos < d.G; pos++)
    for (int j = 0; j < d.ga; j++)
      if (stamp[(size_t)pos * d.ga + j] == ser[1]) ids.push_back(grid_old[pos] * d.ga + j);
  std::sort(ids.begin(), ids.end());
  for (size_t i = 0; i < ids.size() && (int)i < cap; i++) out[i] = ids[i];
  return (int)ids.size();
}

int sg_grids_last_gathered(PBVH *p, int *out, int cap) {
  const int ng = (int)p->last_ctr[SG_CTR_NGATHER];
  std::vector<int> tmp(std::max(ng, 1));
  if (ng > 0 && sg_check_hip(hipMemcpy(tmp.data(), p->gd->gathered, ng * sizeof(int), hipMemcpyDeviceToHost), "copy"))
    return -1;
  std::sort(tmp.begin(), tmp.begin() + ng);
  for (int i = 0; i < ng && i < cap; i++) out[i] = p->leaf_node[tmp[i]];
  return ng;
}

int sg_grids_sync_bbs(PBVH *p) {
  const SgGridDev &d = *p->gd;
  std::vector<float> lbb((size_t)d.L * 16), nbb((size_t)d.N * 16);
  if (sg_grids_read_boxes(p, lbb.data(), nbb.data())) return -1;
  for (PBVHNode &n : p->nodes) {
    const float *src = (n.leaf >= 0) ? &lbb[8 * (size_t)n.leaf] : &nbb[8 * (size_t)n.index];
    const float *osrc = (n.leaf >= 0) ? &lbb[8 * ((size_t)d.L + n.leaf)] : &nbb[8 * ((size_t)d.N + n.index)];
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

// BKE_pbvh_update_bounds (pbvh.c:1401-1421) on grids: leaves flagged with a bit of
// `flag` among UpdateBB / UpdateOriginalBB are refit from their elements (orig_vb
// following vb when `flag` holds UpdateOriginalBB), their ancestors rebuilt; the
// flags are cleared.
void sg_grids_update_bounds(PBVH *p, int flag) {
  const int f = flag & (PBVH_UpdateBB | PBVH_UpdateOriginalBB);
  if (!f || !p->gd) return;
  std::vector<int> leaves = flagged_leaves(p, f);
  if (!leaves.empty()) {
    if (upload_list(p, leaves) ||
        sg_grids_enqueue_refit(p, p->gd->hlist, (int)leaves.size(), (flag & PBVH_UpdateOriginalBB) ? 1 : 0))
      return;
    if (sg_check_hip(hipStreamSynchronize(p->stream), "grid refit")) return;
    p->host_bb_stale = true;
  }
  for (PBVHNode &n : p->nodes) {
    n.flag &= ~f;
    if (flag & PBVH_UpdateRedraw) n.flag &= ~PBVH_UpdateRedraw;
  }
}

// BKE_pbvh_update_normals (pbvh.c:4559-4587) on grids: the faces of the grids of
// UpdateNormals leaves get new normals and averaged seams (KERNEL_subdiv_ccg_update_normals
// on the GPU); the flag is cleared.
void sg_grids_update_normals(PBVH *p, SubdivCCG *s) {
  if (!p->gd) return;
  if (s && (p->gd->topo_of != s || p->gd->topo_version != s->topo_version)) {
    p->subdiv_ccg = s;
    if (sg_grids_attach_topology(p, s)) return;
  }
  std::vector<int> leaves = flagged_leaves(p, PBVH_UpdateNormals);
  if (leaves.empty()) return;
  if (upload_list(p, leaves) || sg_grids_enqueue_update_normals(p, p->gd->hlist, (int)leaves.size())) return;
  if (sg_check_hip(hipStreamSynchronize(p->stream), "grid normals")) return;
  for (int l : leaves) p->nodes[p->leaf_node[l]].flag &= ~PBVH_UpdateNormals;
}
