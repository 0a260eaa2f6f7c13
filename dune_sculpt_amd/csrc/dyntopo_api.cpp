// Host side of PBVH_BMESH: the mesh container (stand-in for BMesh), its corner table
// and normals, the fast tree build (pbvh_bmesh.c:1670-1928), the leaf split after a
// stroke (pbvh_bmesh.c:179-400, :2065-2078) and the C entry points of section 4 of
// sculpt_gpu.h.  The per-dab work (topology rounds, brush, normals, bounds) runs on the
// GPU (dyntopo.hip).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sg_dyn.h"

namespace {

inline int c_next(int c) { return (c % 3 == 2) ? c - 2 : c + 1; }
inline int c_prev(int c) { return (c % 3 == 0) ? c + 2 : c - 1; }
inline float fmin_ff(float a, float b) { return (a < b) ? a : b; }
inline float fmax_ff(float a, float b) { return (a > b) ? a : b; }

struct Box {
  float mn[3], mx[3];
  void reset() {
    mn[0] = mn[1] = mn[2] = FLT_MAX;
    mx[0] = mx[1] = mx[2] = -FLT_MAX;
  }
  void add(const float *p) {
    for (int i = 0; i < 3; i++) {
      mn[i] = fmin_ff(mn[i], p[i]);
      mx[i] = fmax_ff(mx[i], p[i]);
    }
  }
  void add(const Box &b) {
    for (int i = 0; i < 3; i++) {
      mn[i] = fmin_ff(mn[i], b.mn[i]);
      mx[i] = fmax_ff(mx[i], b.mx[i]);
    }
  }
  // BB_widest_axis
  int widest() const {
    float d[3];
    for (int i = 0; i < 3; i++) d[i] = mx[i] - mn[i];
    if (d[0] > d[1]) return (d[0] > d[2]) ? 0 : 2;
    return (d[1] > d[2]) ? 1 : 2;
  }
  void store(float *o) const {
    for (int i = 0; i < 3; i++) {
      o[i] = mn[i];
      o[4 + i] = mx[i];
    }
    o[3] = o[7] = 0.0f;
  }
};

struct FaceBox {
  Box b;
  float c[3];
};

// normalize_v3 (math_vector_inline.c:1165-1180)
float normalize3(float *n) {
  float d = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  if (d > 1.0e-35f) {
    d = sqrtf(d);
    const float inv = 1.0f / d;
    n[0] *= inv;
    n[1] *= inv;
    n[2] *= inv;
  } else {
    n[0] = n[1] = n[2] = 0.0f;
    d = 0.0f;
  }
  return d;
}

// The corners around v (a boundary vertex from its fan's first face); -1 past cap.
int ring(const SgDynMesh &m, int v, int *out, int cap, int *bnd) {
  *bnd = 0;
  const int c0 = m.vcorner[v];
  if (c0 < 0) return 0;
  int c = c0, steps = 0;
  for (;;) {
    const int x = m.topp[c_next(c)];
    if (x < 0) {
      *bnd = 1;
      break;
    }
    c = c_next(x);
    if (c == c0) break;
    if (++steps > cap) return -1;
  }
  const int start = c;
  int n = 0;
  for (;;) {
    if (n >= cap) return -1;
    out[n++] = c;
    const int x = m.topp[c_prev(c)];
    if (x < 0) break;
    c = c_prev(x);
    if (c == start) break;
  }
  return n;
}

// normal_tri_v3 (math_geom.cc:31-46)
void face_normal(SgDynMesh &m, int f) {
  const float *a = &m.co[3 * (size_t)m.tv[3 * f]], *b = &m.co[3 * (size_t)m.tv[3 * f + 1]],
              *c = &m.co[3 * (size_t)m.tv[3 * f + 2]];
  float n1[3], n2[3], *n = &m.fno[3 * (size_t)f];
  for (int i = 0; i < 3; i++) {
    n1[i] = a[i] - b[i];
    n2[i] = b[i] - c[i];
  }
  n[0] = n1[1] * n2[2] - n1[2] * n2[1];
  n[1] = n1[2] * n2[0] - n1[0] * n2[2];
  n[2] = n1[0] * n2[1] - n1[1] * n2[0];
  normalize3(n);
}

// mesh_vert_calc_normal (mesh/tools/intern/mesh_polygon.c:591-634): corner angle x
// face normal over the faces around v, normalized; an exact 2^-32 fixed-point sum.
void vert_normal(SgDynMesh &m, int v) {
  int cs[SG_DYN_RING], bnd;
  const int n = ring(m, v, cs, SG_DYN_RING, &bnd);
  if (n <= 0) return;
  long long s[3] = {0, 0, 0};
  const float *pv = &m.co[3 * (size_t)v];
  for (int i = 0; i < n; i++) {
    const int c = cs[i];
    const float *pp = &m.co[3 * (size_t)m.tv[c_prev(c)]], *pn = &m.co[3 * (size_t)m.tv[c_next(c)]];
    float e1[3] = {pv[0] - pp[0], pv[1] - pp[1], pv[2] - pp[2]};
    float e2[3] = {pn[0] - pv[0], pn[1] - pv[1], pn[2] - pv[2]};
    normalize3(e1);
    normalize3(e2);
    const float dot = -(e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]);
    const float fac = (dot <= -1.0f) ? (float)M_PI : ((dot >= 1.0f) ? 0.0f : (float)acos((double)dot));
    const float *fn = &m.fno[3 * (size_t)(c / 3)];
    for (int q = 0; q < 3; q++) s[q] += llrint((double)(fn[q] * fac) * 4294967296.0);
  }
  float r[3];
  for (int q = 0; q < 3; q++) r[q] = (float)((double)s[q] * (1.0 / 4294967296.0));
  normalize3(r);
  memcpy(&m.no[3 * (size_t)v], r, 12);
}

void face_box(const SgDynMesh &m, int f, FaceBox &fb) {
  fb.b.reset();
  for (int k = 0; k < 3; k++) fb.b.add(&m.co[3 * (size_t)m.tv[3 * f + k]]);
  for (int i = 0; i < 3; i++) fb.c[i] = (fb.b.mn[i] + fb.b.mx[i]) * 0.5f;
}

struct FastInfo {
  int totface, start, c1, c2;
};

// pbvh_bmesh_node_limit_ensure_fast (pbvh_bmesh.c:1670-1763)
void limit_ensure_fast(int leaf_limit, std::vector<int> &ni, const std::vector<FaceBox> &fb,
                       std::vector<FastInfo> &info, int i) {
  if (info[i].totface <= leaf_limit) return;
  Box cb;
  cb.reset();
  for (int k = 0; k < info[i].totface; k++) cb.add(fb[ni[k + info[i].start]].c);
  const int axis = cb.widest();
  const float mid = (cb.mx[axis] + cb.mn[axis]) * 0.5f;
  int num1 = 0, num2 = 0;
  const int end = info[i].start + info[i].totface;
  for (int k = info[i].start; k < end - num2; k++) {
    if (fb[ni[k]].c[axis] > mid) {
      int it = end - num2 - 1, candidate = -1;
      for (; it > k; it--) {
        if (fb[ni[it]].c[axis] <= mid) {
          candidate = it;
          break;
        }
        num2++;
      }
      if (candidate != -1) {
        std::swap(ni[k], ni[candidate]);
        num1++;
        num2++;
      } else {
        num2++;
      }
    } else {
      num1++;
    }
  }
  if (num2 == 0) {
    num2++;
    num1--;
  } else if (num1 == 0) {
    num1++;
    num2--;
  }
  const int c1 = (int)info.size(), c2 = c1 + 1;
  info.push_back({num1, info[i].start, -1, -1});
  info.push_back({num2, info[i].start + num1, -1, -1});
  info[i].c1 = c1;
  info[i].c2 = c2;
  limit_ensure_fast(leaf_limit, ni, fb, info, c1);
  limit_ensure_fast(leaf_limit, ni, fb, info, c2);
}

void grow_nodes(SgDynTree &t, int n) {
  const size_t old = t.child.size();
  if ((size_t)n <= old) return;
  t.child.resize(n, -1);
  t.flag.resize(n, 0);
  t.scount.resize(n, 0);
  t.slab.resize((size_t)n * SG_DYN_SLAB, -1);
  t.bb.resize((size_t)n * 8, 0.0f);
  t.obb.resize((size_t)n * 8, 0.0f);
}

// pbvh_bmesh_create_nodes_fast_recursive (pbvh_bmesh.c:1765-1852)
void create_nodes_fast(const SgDynMesh &m, const std::vector<int> &ni, const std::vector<FaceBox> &fb,
                       const std::vector<FastInfo> &info, int ii, int node, SgDynTree &t, std::vector<int> &vnode,
                       std::vector<int> &tnode, std::vector<int> &tslot) {
  if (info[ii].c1 >= 0) {
    const int children = (int)t.child.size();
    t.child[node] = children;
    grow_nodes(t, children + 2);
    create_nodes_fast(m, ni, fb, info, info[ii].c1, children, t, vnode, tnode, tslot);
    create_nodes_fast(m, ni, fb, info, info[ii].c2, children + 1, t, vnode, tnode, tslot);
    Box b;
    b.reset();
    for (int k = 0; k < 2; k++) {
      Box c;
      for (int i = 0; i < 3; i++) {
        c.mn[i] = t.bb[8 * (size_t)(children + k) + i];
        c.mx[i] = t.bb[8 * (size_t)(children + k) + 4 + i];
      }
      b.add(c);
    }
    b.store(&t.bb[8 * (size_t)node]);
    b.store(&t.obb[8 * (size_t)node]);
    return;
  }
  t.flag[node] = PBVH_Leaf;
  t.child[node] = -1;
  Box b;
  b.reset();
  bool has_visible = false;
  t.scount[node] = 0;
  for (int k = info[ii].start; k < info[ii].start + info[ii].totface; k++) {
    const int f = ni[k];
    const int pos = t.scount[node]++;
    t.slab[(size_t)node * SG_DYN_SLAB + pos] = f;
    tnode[f] = node;
    tslot[f] = pos;
    for (int c = 0; c < 3; c++) {
      const int v = m.tv[3 * f + c];
      if (vnode[v] < 0) vnode[v] = node;
    }
    if (!m.fhide[f]) has_visible = true;
    b.add(fb[f].b);
  }
  b.store(&t.bb[8 * (size_t)node]);
  b.store(&t.obb[8 * (size_t)node]);
  t.flag[node] |= PBVH_UpdateNormals | PBVH_UpdateDrawBuffers;
  if (!has_visible) t.flag[node] |= PBVH_FullyHidden;
}

// Faces of a leaf in the canonical order: by labels from the corner with the
// smallest label (rotation kept).
struct FaceKey {
  uint64_t k[3];
  int f;
  bool operator<(const FaceKey &o) const {
    for (int i = 0; i < 3; i++)
      if (k[i] != o.k[i]) return k[i] < o.k[i];
    return false;
  }
};
FaceKey face_key(const SgDynMesh &m, int f) {
  int s = 0;
  for (int k = 1; k < 3; k++)
    if (m.lab[m.tv[3 * f + k]] < m.lab[m.tv[3 * f + s]]) s = k;
  FaceKey fk;
  for (int k = 0; k < 3; k++) fk.k[k] = m.lab[m.tv[3 * f + (s + k) % 3]];
  fk.f = f;
  return fk;
}

// pbvh_bmesh_node_finalize (pbvh_bmesh.c:179-235)
void node_finalize(SgDynMesh &m, SgDynTree &t, int ni) {
  Box b;
  b.reset();
  bool has_visible = false;
  for (int i = 0; i < t.scount[ni]; i++) {
    const int f = t.slab[(size_t)ni * SG_DYN_SLAB + i];
    m.tnode[f] = ni;
    m.tslot[f] = i;
    for (int k = 0; k < 3; k++) {
      const int v = m.tv[3 * f + k];
      if (m.vnode[v] < 0) m.vnode[v] = ni;
      b.add(&m.co[3 * (size_t)v]);
    }
    if (!m.fhide[f]) has_visible = true;
  }
  b.store(&t.bb[8 * (size_t)ni]);
  b.store(&t.obb[8 * (size_t)ni]);
  t.flag[ni] |= PBVH_UpdateDrawBuffers | PBVH_UpdateNormals;
  if (has_visible)
    t.flag[ni] &= ~PBVH_FullyHidden;
  else
    t.flag[ni] |= PBVH_FullyHidden;
}

// pbvh_bmesh_node_split (pbvh_bmesh.c:237-361), faces in the canonical order
void node_split(SgDynMesh &m, SgDynTree &t, const std::vector<FaceBox> &fb, int leaf_limit, int ni) {
  if (t.scount[ni] <= leaf_limit) {
    node_finalize(m, t, ni);
    return;
  }
  const int cnt = t.scount[ni];
  std::vector<int> fl(t.slab.begin() + (size_t)ni * SG_DYN_SLAB, t.slab.begin() + (size_t)ni * SG_DYN_SLAB + cnt);
  Box cb;
  cb.reset();
  for (int f : fl) cb.add(fb[f].c);
  const int axis = cb.widest();
  const float mid = (cb.mx[axis] + cb.mn[axis]) * 0.5f;
  const int children = (int)t.child.size();
  t.child[ni] = children;
  grow_nodes(t, children + 2);
  for (int k = 0; k < 2; k++) {
    t.flag[children + k] = PBVH_Leaf;
    t.child[children + k] = -1;
    t.scount[children + k] = 0;
  }
  int *s1 = &t.slab[(size_t)children * SG_DYN_SLAB], *s2 = &t.slab[(size_t)(children + 1) * SG_DYN_SLAB];
  for (int f : fl) {
    if (fb[f].c[axis] < mid)
      s1[t.scount[children]++] = f;
    else
      s2[t.scount[children + 1]++] = f;
  }
  if (t.scount[children] == 0 || t.scount[children + 1] == 0) {
    const int e = t.scount[children] == 0 ? children : children + 1, other = e == children ? children + 1 : children;
    int *se = &t.slab[(size_t)e * SG_DYN_SLAB], *so = &t.slab[(size_t)other * SG_DYN_SLAB];
    se[t.scount[e]++] = so[0];
    memmove(so, so + 1, sizeof(int) * (size_t)(t.scount[other] - 1));
    t.scount[other]--;
  }
  for (int f : fl)
    for (int k = 0; k < 3; k++) {
      const int v = m.tv[3 * f + k];
      if (m.vnode[v] == ni) m.vnode[v] = -1;
    }
  t.scount[ni] = 0;
  t.flag[ni] &= ~PBVH_Leaf;
  node_split(m, t, fb, leaf_limit, children);
  node_split(m, t, fb, leaf_limit, children + 1);
  Box b;
  b.reset();
  for (int k = 0; k < 2; k++) {
    Box c;
    for (int i = 0; i < 3; i++) {
      c.mn[i] = t.bb[8 * (size_t)(children + k) + i];
      c.mx[i] = t.bb[8 * (size_t)(children + k) + 4 + i];
    }
    b.add(c);
  }
  b.store(&t.bb[8 * (size_t)ni]);
  b.store(&t.obb[8 * (size_t)ni]);
}

SgDynState *dyn_of(PBVH *p) { return (p && p->type == PBVH_BMESH) ? p->dyn : nullptr; }

int node_index(PBVHNode *n) { return n && n->owner ? n->index : -1; }

// The host node handles follow the tree mirror.
void sync_host_nodes(PBVH *p) {
  SgDynState *s = p->dyn;
  const SgDynTree &t = s->t;
  const int N = (int)t.child.size();
  p->nodes.resize(N);
  p->leaf_node.clear();
  for (int i = 0; i < N; i++) {
    PBVHNode &n = p->nodes[i];
    n.index = i;
    n.owner = p;
    n.flag = t.flag[i];
    n.children_offset = t.child[i];
    n.leaf = -1;
    for (int k = 0; k < 3; k++) {
      n.vb[k] = t.bb[8 * (size_t)i + k];
      n.vb[3 + k] = t.bb[8 * (size_t)i + 4 + k];
      n.orig_vb[k] = t.obb[8 * (size_t)i + k];
      n.orig_vb[3 + k] = t.obb[8 * (size_t)i + 4 + k];
    }
    n.totprim = (t.flag[i] & PBVH_Leaf) ? t.scount[i] : 0;
  }
  std::vector<int> st{0};
  while (!st.empty()) {
    const int i = st.back();
    st.pop_back();
    if (t.flag[i] & PBVH_Leaf) {
      p->nodes[i].leaf = (int)p->leaf_node.size();
      p->leaf_node.push_back(i);
      continue;
    }
    st.push_back(t.child[i] + 1);
    st.push_back(t.child[i]);
  }
}

int refresh(PBVH *p, bool mesh) {
  SgDynState *s = p->dyn;
  if (p->no_device || !p->build_ok) return 0;  // the host copy is the mesh
  if (!s->host_stale && !mesh) return 0;
  if (sg_dyn_download(p, mesh)) return -1;
  s->host_stale = false;
  sync_host_nodes(p);
  return 0;
}

}  // namespace

int sg_dyn_refresh(PBVH *p) { return (p && p->dyn) ? refresh(p, false) : 0; }

void sg_dyn_note_flags(PBVHNode *node, int bits) {
  SgDynState *s = node ? dyn_of(node->owner) : nullptr;
  if (!s || node->index < 0) return;
  if ((int)s->pend.size() <= node->index) s->pend.resize(node->index + 1, 0);
  s->pend[node->index] |= bits;
}

void sg_dyn_release(PBVH *p) {
  if (!p || !p->dyn) return;
  if (!p->no_device) sg_dyn_free_device(p);
  if (p->dyn->bm) p->dyn->bm->owner = nullptr;
  delete p->dyn;
  p->dyn = nullptr;
}

int sg_dyn_last_gathered(PBVH *p, int *out, int cap) {
  SgDynState *s = dyn_of(p);
  if (!s) return sg_set_error("not a PBVH_BMESH");
  const int n = (int)s->stats[0];
  std::vector<int> g(std::max(n, 1));
  if (n && sg_check_hip(hipMemcpy(g.data(), s->d.gathered, n * 4, hipMemcpyDeviceToHost), "copy")) return -1;
  std::sort(g.begin(), g.begin() + n);
  for (int i = 0; i < n && i < cap; i++) out[i] = g[i];
  return n;
}

int sg_dynmesh_build_tables(SgDynMesh *m) {
  const int V = m->V, T = m->T;
  for (int c = 0; c < 3 * T; c++) {
    if (m->tv[c] < 0 || m->tv[c] >= V) return sg_set_error("sg_dynmesh_new: corner %d: vertex out of range", c);
    m->topp[c] = -1;
  }
  for (int f = 0; f < T; f++) {
    const int a = m->tv[3 * f], b = m->tv[3 * f + 1], c = m->tv[3 * f + 2];
    if (a == b || b == c || a == c) return sg_set_error("sg_dynmesh_new: triangle %d is degenerate", f);
  }
  // opposite corners: a directed edge used twice is non-manifold or flipped
  std::vector<std::pair<uint64_t, int>> er(3 * (size_t)T);
  for (int c = 0; c < 3 * T; c++)
    er[c] = {((uint64_t)(uint32_t)m->tv[c_next(c)] << 32) | (uint32_t)m->tv[c_prev(c)], c};
  std::sort(er.begin(), er.end());
  for (size_t i = 1; i < er.size(); i++)
    if (er[i].first == er[i - 1].first)
      return sg_set_error("sg_dynmesh_new: edge %u-%u used twice in one direction (non-manifold or flipped faces)",
                          (unsigned)(er[i].first >> 32), (unsigned)(er[i].first & 0xffffffffu));
  for (int c = 0; c < 3 * T; c++) {
    const uint64_t tw = ((uint64_t)(uint32_t)m->tv[c_prev(c)] << 32) | (uint32_t)m->tv[c_next(c)];
    auto it = std::lower_bound(er.begin(), er.end(), std::make_pair(tw, INT32_MIN));
    if (it != er.end() && it->first == tw) m->topp[c] = it->second;
  }
  std::vector<int> cnt(V, 0);
  for (int c = 0; c < 3 * T; c++) {
    cnt[m->tv[c]]++;
    if (m->vcorner[m->tv[c]] < 0) m->vcorner[m->tv[c]] = c;
  }
  for (int v = 0; v < V; v++) {
    if (!cnt[v]) continue;
    int cs[SG_DYN_RING], bnd;
    const int n = ring(*m, v, cs, SG_DYN_RING, &bnd);
    if (n != cnt[v])
      return sg_set_error("sg_dynmesh_new: vertex %d is not manifold (%d faces, %d around one fan)", v, cnt[v], n);
  }
  for (int f = 0; f < T; f++) face_normal(*m, f);
  for (int v = 0; v < V; v++) vert_normal(*m, v);
  return 0;
}

void sg_dyn_build_tree(const SgDynMesh &m, int leaf_limit, SgDynTree &t, std::vector<int> &vnode,
                       std::vector<int> &tnode, std::vector<int> &tslot) {
  std::vector<FaceBox> fb(m.T);
  std::vector<int> ni(m.T);
  for (int f = 0; f < m.T; f++) {
    face_box(m, f, fb[f]);
    ni[f] = f;
  }
  std::fill(vnode.begin(), vnode.end(), -1);
  std::vector<FastInfo> info;
  info.reserve(2 * (size_t)m.T + 2);
  info.push_back({m.T, 0, -1, -1});
  limit_ensure_fast(leaf_limit, ni, fb, info, 0);
  t = SgDynTree();
  grow_nodes(t, 1);
  create_nodes_fast(m, ni, fb, info, 0, 0, t, vnode, tnode, tslot);
}

void sg_dyn_after_stroke_host(SgDynMesh &m, int leaf_limit, SgDynTree &t) {
  const int N = (int)t.child.size();
  for (int ni = 0; ni < N; ni++) {
    if (!(t.flag[ni] & PBVH_Leaf)) continue;
    int *sl = &t.slab[(size_t)ni * SG_DYN_SLAB];
    std::vector<FaceKey> fk;
    for (int i = 0; i < t.scount[ni]; i++)
      if (sl[i] >= 0) fk.push_back(face_key(m, sl[i]));
    std::sort(fk.begin(), fk.end());
    for (size_t i = 0; i < fk.size(); i++) {
      sl[i] = fk[i].f;
      m.tslot[fk[i].f] = (int)i;
    }
    for (int i = (int)fk.size(); i < t.scount[ni]; i++) sl[i] = -1;
    t.scount[ni] = (int)fk.size();
  }
  std::vector<FaceBox> fb(m.T);
  for (int ni = 0; ni < (int)t.child.size(); ni++) {
    if (!(t.flag[ni] & PBVH_Leaf) || t.scount[ni] <= leaf_limit) continue;
    for (int i = 0; i < t.scount[ni]; i++) {
      const int f = t.slab[(size_t)ni * SG_DYN_SLAB + i];
      face_box(m, f, fb[f]);
    }
    node_split(m, t, fb, leaf_limit, ni);
  }
}

extern "C" {

SgDynMesh *sg_dynmesh_new(const float *co, int totvert, const int *tris, int tottri, const float *mask,
                          const uint8_t *vert_hidden, const uint8_t *face_hidden, int vert_cap, int tri_cap) {
  if (!co || !tris || totvert <= 0 || tottri <= 0) {
    sg_set_error("sg_dynmesh_new: empty mesh");
    return nullptr;
  }
  if (vert_cap == 0) vert_cap = 4 * totvert;
  if (tri_cap == 0) tri_cap = 4 * tottri;
  if (vert_cap < totvert || tri_cap < tottri || (long long)tri_cap * 3 > INT32_MAX) {
    sg_set_error("sg_dynmesh_new: capacities below the counts or above 2^31 corners");
    return nullptr;
  }
  SgDynMesh *m = new SgDynMesh();
  m->V = totvert;
  m->T = tottri;
  m->Vcap = vert_cap;
  m->Tcap = tri_cap;
  m->has_mask = mask != nullptr;
  m->co.assign(co, co + 3 * (size_t)totvert);
  m->co.resize(3 * (size_t)vert_cap, 0.0f);
  m->no.assign(3 * (size_t)vert_cap, 0.0f);
  if (mask) {
    m->mask.assign(mask, mask + totvert);
    m->mask.resize(vert_cap, 0.0f);
  }
  m->vhide.assign(vert_cap, 0);
  if (vert_hidden) std::copy(vert_hidden, vert_hidden + totvert, m->vhide.begin());
  m->fhide.assign(tri_cap, 0);
  if (face_hidden) std::copy(face_hidden, face_hidden + tottri, m->fhide.begin());
  m->lab.resize(vert_cap);
  for (int v = 0; v < vert_cap; v++) m->lab[v] = (uint64_t)v;
  m->vnode.assign(vert_cap, -1);
  m->vcorner.assign(vert_cap, -1);
  m->tv.assign(tris, tris + 3 * (size_t)tottri);
  m->tv.resize(3 * (size_t)tri_cap, 0);
  m->topp.assign(3 * (size_t)tri_cap, -1);
  m->tnode.assign(tri_cap, -1);
  m->tslot.assign(tri_cap, -1);
  m->fno.assign(3 * (size_t)tri_cap, 0.0f);
  if (sg_dynmesh_build_tables(m)) {
    delete m;
    return nullptr;
  }
  return m;
}

void sg_dynmesh_free(SgDynMesh *m) {
  if (!m) return;
  if (m->owner && m->owner->dyn) m->owner->dyn->bm = nullptr;
  delete m;
}

int sg_dynmesh_counts(SgDynMesh *m, int *r_totvert, int *r_tottri, int *r_vert_slots, int *r_tri_slots) {
  if (!m) return sg_set_error("null mesh");
  if (m->owner && refresh(m->owner, true)) return -1;
  int nv = 0, nt = 0;
  for (int v = 0; v < m->V; v++) nv += m->vcorner[v] >= 0;
  for (int f = 0; f < m->T; f++) nt += m->owner ? m->tnode[f] >= 0 : 1;
  if (r_totvert) *r_totvert = nv;
  if (r_tottri) *r_tottri = nt;
  if (r_vert_slots) *r_vert_slots = m->V;
  if (r_tri_slots) *r_tri_slots = m->T;
  return 0;
}

int sg_dynmesh_read(SgDynMesh *m, uint64_t *vert_labels, float *co, float *no, float *mask, int *vert_node,
                    uint64_t *face_labels, int *face_node, float *face_no, uint8_t *face_hidden) {
  if (!m) return sg_set_error("null mesh");
  if (m->owner && refresh(m->owner, true)) return -1;
  int w = 0;
  for (int v = 0; v < m->V; v++) {
    if (m->vcorner[v] < 0) continue;
    if (vert_labels) vert_labels[w] = m->lab[v];
    if (co) memcpy(co + 3 * (size_t)w, &m->co[3 * (size_t)v], 12);
    if (no) memcpy(no + 3 * (size_t)w, &m->no[3 * (size_t)v], 12);
    if (mask) mask[w] = m->has_mask ? m->mask[v] : 0.0f;
    if (vert_node) vert_node[w] = m->vnode[v];
    w++;
  }
  int u = 0;
  for (int f = 0; f < m->T; f++) {
    if (m->owner && m->tnode[f] < 0) continue;
    for (int k = 0; k < 3; k++)
      if (face_labels) face_labels[3 * (size_t)u + k] = m->lab[m->tv[3 * f + k]];
    if (face_node) face_node[u] = m->tnode[f];
    if (face_no) memcpy(face_no + 3 * (size_t)u, &m->fno[3 * (size_t)f], 12);
    if (face_hidden) face_hidden[u] = m->fhide[f];
    u++;
  }
  return w;
}

void DUNE_pbvh_build_bmesh(PBVH *p, SgDynMesh *bm, bool smooth_shading, void *log, int cd_vert_node_offset,
                           int cd_face_node_offset) {
  (void)smooth_shading;
  (void)log;
  (void)cd_vert_node_offset;
  (void)cd_face_node_offset;
  if (!p || !bm) {
    sg_set_error("DUNE_pbvh_build_bmesh: null argument");
    return;
  }
  if (!p->nodes.empty()) {
    sg_set_error("DUNE_pbvh_build_bmesh: the PBVH is already built");
    return;
  }
  if (bm->owner) {
    sg_set_error("DUNE_pbvh_build_bmesh: the mesh already has a PBVH");
    return;
  }
  p->type = PBVH_BMESH;
  p->dyn = new SgDynState();
  p->dyn->bm = bm;
  p->leaf_limit = 100;  // pbvh_bmesh.c:1868
  bm->owner = p;
  sg_dyn_build_tree(*bm, p->leaf_limit, p->dyn->t, bm->vnode, bm->tnode, bm->tslot);
  sync_host_nodes(p);
  p->build_ok = false;
  if (!p->no_device) p->build_ok = sg_dyn_upload(p) == 0;
}

void BKE_pbvh_build_bmesh(PBVH *p, SgDynMesh *bm, bool smooth_shading, void *log, int cd_vert_node_offset,
                          int cd_face_node_offset) {
  DUNE_pbvh_build_bmesh(p, bm, smooth_shading, log, cd_vert_node_offset, cd_face_node_offset);
}

bool DUNE_pbvh_bmesh_update_topology(PBVH *p, PBVHTopologyUpdateMode mode, const float center[3],
                                     const float view_normal[3], float radius, bool use_frontface,
                                     bool use_projected) {
  SgDynState *s = dyn_of(p);
  if (!s || !p->build_ok || !center) {
    sg_set_error("DUNE_pbvh_bmesh_update_topology: no PBVH_BMESH on the device");
    return false;
  }
  if (view_normal) memcpy(s->view_normal, view_normal, 12);
  s->use_frontface = use_frontface;
  s->use_projected = use_projected;
  if (sg_dyn_sync_flags_down(p)) return false;
  const int rc = sg_dyn_update_topology(p, (int)mode, center, radius);
  s->host_stale = true;
  refresh(p, false);
  return rc > 0;
}

bool BKE_pbvh_bmesh_update_topology(PBVH *p, PBVHTopologyUpdateMode mode, const float center[3],
                                    const float view_normal[3], float radius, bool use_frontface, bool use_projected) {
  return DUNE_pbvh_bmesh_update_topology(p, mode, center, view_normal, radius, use_frontface, use_projected);
}

void DUNE_pbvh_bmesh_after_stroke(PBVH *p) {
  SgDynState *s = dyn_of(p);
  if (!s || !s->bm || !p->build_ok) return;
  if (sg_dyn_download(p, true)) return;
  sg_dyn_after_stroke_host(*s->bm, p->leaf_limit, s->t);
  if (sg_dyn_upload(p)) {
    p->build_ok = false;
    return;
  }
  s->host_stale = false;
  sync_host_nodes(p);
}

void BKE_pbvh_bmesh_after_stroke(PBVH *p) { DUNE_pbvh_bmesh_after_stroke(p); }

void DUNE_pbvh_bmesh_detail_size_set(PBVH *p, float detail_size) {
  SgDynState *s = dyn_of(p);
  if (!s) return;
  s->max_len = detail_size;
  s->min_len = s->max_len * 0.4f;
}

void BKE_pbvh_bmesh_detail_size_set(PBVH *p, float detail_size) { DUNE_pbvh_bmesh_detail_size_set(p, detail_size); }

void DUNE_pbvh_node_mark_topology_update(PBVHNode *node) {
  if (!node) return;
  node->flag |= PBVH_UpdateTopology;
  sg_dyn_note_flags(node, PBVH_UpdateTopology);
}

void BKE_pbvh_node_mark_topology_update(PBVHNode *node) { DUNE_pbvh_node_mark_topology_update(node); }

int sg_pbvh_bmesh_topology_set(PBVH *p, int mode, const float view_normal[3], int use_frontface, int use_projected) {
  SgDynState *s = dyn_of(p);
  if (!s) return sg_set_error("not a PBVH_BMESH");
  s->mode = mode & (PBVH_Subdivide | PBVH_Collapse);
  if (view_normal) memcpy(s->view_normal, view_normal, 12);
  s->use_frontface = use_frontface != 0;
  s->use_projected = use_projected != 0;
  return 0;
}

int sg_pbvh_bmesh_last_stats(PBVH *p, long long out[12]) {
  SgDynState *s = dyn_of(p);
  if (!s) return sg_set_error("not a PBVH_BMESH");
  memcpy(out, s->stats, sizeof(s->stats));
  return 0;
}

int sg_pbvh_bmesh_last_affected(PBVH *p, uint64_t *out, int cap) {
  if (!dyn_of(p)) return sg_set_error("not a PBVH_BMESH");
  return sg_dyn_last_affected(p, out, cap);
}

int sg_pbvh_bmesh_node_info(PBVH *p, int node, int out[3], float vb[6], float orig_vb[6]) {
  SgDynState *s = dyn_of(p);
  if (!s) return sg_set_error("not a PBVH_BMESH");
  if (refresh(p, false)) return -1;
  const SgDynTree &t = s->t;
  if (node < 0 || node >= (int)t.child.size()) return sg_set_error("node out of range");
  out[0] = t.flag[node];
  out[1] = t.child[node];
  int nf = 0;
  if (t.flag[node] & PBVH_Leaf)
    for (int i = 0; i < t.scount[node]; i++) nf += t.slab[(size_t)node * SG_DYN_SLAB + i] >= 0;
  out[2] = nf;
  for (int k = 0; k < 3; k++) {
    vb[k] = t.bb[8 * (size_t)node + k];
    vb[3 + k] = t.bb[8 * (size_t)node + 4 + k];
    orig_vb[k] = t.obb[8 * (size_t)node + k];
    orig_vb[3 + k] = t.obb[8 * (size_t)node + 4 + k];
  }
  return 0;
}

static int node_lists(PBVHNode *node, int *out, int cap, int which) {
  PBVH *p = node ? node->owner : nullptr;
  SgDynState *s = dyn_of(p);
  if (!s || !s->bm) return sg_set_error("not a PBVH_BMESH node");
  if (refresh(p, true)) return -1;
  const int ni = node_index(node);
  const SgDynTree &t = s->t;
  const SgDynMesh &m = *s->bm;
  if (!(t.flag[ni] & PBVH_Leaf)) return 0;
  std::vector<int> vs;
  int n = 0;
  for (int i = 0; i < t.scount[ni]; i++) {
    const int f = t.slab[(size_t)ni * SG_DYN_SLAB + i];
    if (f < 0) continue;
    if (which == 0) {
      if (n < cap && out) out[n] = f;
      n++;
      continue;
    }
    for (int k = 0; k < 3; k++) {
      const int v = m.tv[3 * f + k];
      if ((m.vnode[v] == ni) == (which == 1)) vs.push_back(v);
    }
  }
  if (which == 0) return n;
  std::sort(vs.begin(), vs.end());
  vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
  for (size_t i = 0; i < vs.size() && (int)i < cap && out; i++) out[i] = vs[i];
  return (int)vs.size();
}

int sg_pbvh_bmesh_node_faces(PBVHNode *node, int *out, int cap) { return node_lists(node, out, cap, 0); }
int sg_pbvh_bmesh_node_unique_verts(PBVHNode *node, int *out, int cap) { return node_lists(node, out, cap, 1); }
int sg_pbvh_bmesh_node_other_verts(PBVHNode *node, int *out, int cap) { return node_lists(node, out, cap, 2); }

}  // extern "C"
