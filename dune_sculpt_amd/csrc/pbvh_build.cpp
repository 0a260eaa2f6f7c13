// Host-side PBVH build and the leaf-ordered SoA layout uploaded to HBM.
//
// The tree is built exactly as the reference builds it (kernel/intern/pbvh.c:534-596
// BKE_pbvh_build_mesh -> :509 pbvh_build -> :454 build_sub -> :152 partition_indices /
// :172 partition_indices_material -> :391 build_leaf -> :256 build_mesh_leaf_node), so
// node numbering, leaf membership and the uniq/shared vertex split are identical.  The
// recursion is replaced by an explicit stack that visits nodes in the same order.
//
// On top of the tree this file derives the GPU layout:
//  * vertices renumbered leaf by leaf in DFS (left->right) order, uniq vertices of a
//    leaf contiguous, so a leaf's vertex reads are one coalesced range;
//  * triangles in prim_indices order (each leaf's tris contiguous);
//  * vertex->tri and vertex->vertex CSR in the new numbering, each list ordered by
//    ascending ORIGINAL tri / vertex id (the order the normal and smooth sums use);
//  * a cut of the tree into subtrees of <= K leaves for the two-level sphere search,
//    with height-ordered tables for the deferred refit of internal nodes.
#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <thread>
#include <unordered_map>

#include "sg_internal.h"
#include "sg_kernels.h"

namespace {

struct BBC {
  float bmin[3], bmax[3], bcentroid[3];
};

inline float fmin_ref(float a, float b) { return (a < b) ? a : b; }  // min_ff
inline float fmax_ref(float a, float b) { return (a > b) ? a : b; }  // max_ff

inline void bb_reset(float *bb) {
  bb[0] = bb[1] = bb[2] = FLT_MAX;
  bb[3] = bb[4] = bb[5] = -FLT_MAX;
}
inline void bb_expand(float *bb, const float *co) {
  for (int i = 0; i < 3; i++) {
    bb[i] = fmin_ref(bb[i], co[i]);
    bb[3 + i] = fmax_ref(bb[3 + i], co[i]);
  }
}
inline void bb_expand_bb(float *bb, const float *b2) {
  for (int i = 0; i < 3; i++) {
    bb[i] = fmin_ref(bb[i], b2[i]);
    bb[3 + i] = fmax_ref(bb[3 + i], b2[3 + i]);
  }
}
inline int bb_widest_axis(const float *bb) {  // BB_widest_axis pbvh.c:78-96
  float d[3];
  for (int i = 0; i < 3; i++) d[i] = bb[3 + i] - bb[i];
  if (d[0] > d[1]) return (d[0] > d[2]) ? 0 : 2;
  return (d[1] > d[2]) ? 1 : 2;
}

template <class F>
void parallel_for(int n, F f) {
  int nt = (int)std::thread::hardware_concurrency();
  if (nt > 16) nt = 16;
  if (nt < 1 || n < 4096) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  std::vector<std::thread> th;
  const int chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    const int b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=, &f]() {
      for (int i = b; i < e; i++) f(i);
    });
  }
  for (auto &x : th) x.join();
}

struct Builder {
  PBVH *p;
  const float *co;
  const int *tris;
  std::vector<BBC> bbc;
  std::vector<uint8_t> vert_bitmap;
  std::vector<int> map_local;
  std::vector<int> touched;
  bool grids = false;  // PBVH_GRIDS: prims are grids, leaves take build_grid_leaf

  bool mat_match(int t1, int t2) const {
    if (p->tri_mat.empty()) return true;
    return p->tri_mat[t1] == p->tri_mat[t2];
  }

  void update_vb(PBVHNode &n, int offset, int count) {  // pbvh.c:322-329
    bb_reset(n.vb);
    for (int i = offset + count - 1; i >= offset; i--) {
      bb_expand_bb(n.vb, bbc[p->prim_indices[i]].bmin);
    }
    memcpy(n.orig_vb, n.vb, sizeof(n.vb));
  }

  int partition(int lo, int hi, int axis, float mid) {  // pbvh.c:152-170
    int *prim = p->prim_indices.data();
    int i = lo, j = hi;
    for (;;) {
      while (bbc[prim[i]].bcentroid[axis] < mid) i++;
      while (mid < bbc[prim[j]].bcentroid[axis]) j--;
      if (!(i < j)) return i;
      std::swap(prim[i], prim[j]);
      i++;
    }
  }

  int partition_material(int lo, int hi) {  // pbvh.c:172-212
    int *ind = p->prim_indices.data();
    const int first = ind[lo];
    int i = lo, j = hi;
    for (;;) {
      while (mat_match(first, ind[i])) i++;
      while (!mat_match(first, ind[j])) j--;
      if (!(i < j)) return i;
      std::swap(ind[i], ind[j]);
      i++;
    }
  }

  bool needs_material_split(int offset, int count) const {  // pbvh.c:409-440
    if (count <= 1 || p->tri_mat.empty()) return false;
    const int first = p->prim_indices[offset];
    for (int i = offset + count - 1; i > offset; i--) {
      if (!mat_match(first, p->prim_indices[i])) return true;
    }
    return false;
  }

  // build_mesh_leaf_node pbvh.c:256-320 with map_insert_vert :230-254.  The GHash
  // value is the insertion counter, so vert_indices is in first-insertion order.
  // build_leaf (pbvh.c:391-407) -> build_grid_leaf_node (:383-389): fully hidden when
  // no quad of the leaf's grids is visible; draw buffers to be rebuilt.
  void build_grid_leaf(int ni, int offset, int count) {
    PBVHNode &n = p->nodes[ni];
    n.flag |= PBVH_Leaf;
    n.children_offset = -1;
    n.prim_off = offset;
    n.totprim = count;
    update_vb(n, offset, count);
    n.uniq_verts = n.face_verts = 0;
    n.vert_indices = nullptr;
    n.face_vert_indices = nullptr;
    const int totquads =
        BKE_pbvh_count_grid_quads(p->grid_hidden, &p->prim_indices[offset], count, p->gridkey.grid_size);
    if (totquads == 0) n.flag |= PBVH_FullyHidden;
    n.flag |= PBVH_RebuildDrawBuffers | PBVH_UpdateDrawBuffers | PBVH_UpdateRedraw;
  }

  void build_leaf(int ni, int offset, int count) {
    if (grids) {
      build_grid_leaf(ni, offset, count);
      return;
    }
    PBVHNode &n = p->nodes[ni];
    n.flag |= PBVH_Leaf;
    n.children_offset = -1;
    n.prim_off = offset;
    n.totprim = count;
    update_vb(n, offset, count);
    n.uniq_verts = n.face_verts = 0;
    n.face_vert_indices = (int(*)[3])malloc(sizeof(int[3]) * (size_t)std::max(count, 1));
    touched.clear();
    bool has_visible = !p->respect_hide;
    for (int i = 0; i < count; i++) {
      const int t = p->prim_indices[offset + i];
      bool hidden = false;
      for (int j = 0; j < 3; j++) {
        const int v = tris[3 * t + j];
        int val = map_local[v];
        if (val == INT32_MIN) {
          if (!vert_bitmap[v]) {
            vert_bitmap[v] = 1;
            val = n.uniq_verts++;
          } else {
            val = ~n.face_verts;
            n.face_verts++;
          }
          map_local[v] = val;
          touched.push_back(v);
        }
        n.face_vert_indices[i][j] = val;
        if (!p->vhide.empty() && p->vhide[v]) hidden = true;
      }
      if (!hidden) has_visible = true;
    }
    const int tot = n.uniq_verts + n.face_verts;
    n.vert_indices = (int *)calloc((size_t)std::max(tot, 1), sizeof(int));
    for (int v : touched) {
      int ndx = map_local[v];
      if (ndx < 0) ndx = -ndx + n.uniq_verts - 1;
      n.vert_indices[ndx] = v;
      map_local[v] = INT32_MIN;
    }
    for (int i = 0; i < count; i++) {
      for (int j = 0; j < 3; j++) {
        if (n.face_vert_indices[i][j] < 0) {
          n.face_vert_indices[i][j] = -n.face_vert_indices[i][j] + n.uniq_verts - 1;
        }
      }
    }
    n.flag |= PBVH_RebuildDrawBuffers | PBVH_UpdateDrawBuffers | PBVH_UpdateRedraw;
    if (!has_visible) n.flag |= PBVH_FullyHidden;
  }

  void new_node() {
    PBVHNode n;
    memset(&n, 0, sizeof(n));
    n.children_offset = -1;
    n.leaf = -1;
    n.owner = p;
    p->nodes.push_back(n);
  }

  void build(int T) {
    struct Task {
      int node, offset, count;
      bool root;
    };
    p->prim_indices.resize(T);
    for (int i = 0; i < T; i++) p->prim_indices[i] = i;
    // root centroid bounds, ascending prim order (pbvh.c:572-586)
    float cb_root[6];
    bb_reset(cb_root);
    for (int i = 0; i < T; i++) bb_expand(cb_root, bbc[i].bcentroid);

    p->nodes.clear();
    p->proxy_nodes.clear();
    p->nodes.reserve((size_t)std::max(2 * (T / std::max(p->leaf_limit / 2, 1)) + 16, 16));
    new_node();
    std::vector<Task> stack;
    stack.push_back({0, 0, T, true});
    while (!stack.empty()) {
      Task tk = stack.back();
      stack.pop_back();
      const bool below = tk.count <= p->leaf_limit;
      if (below && !needs_material_split(tk.offset, tk.count)) {
        build_leaf(tk.node, tk.offset, tk.count);
        continue;
      }
      const int c = (int)p->nodes.size();
      new_node();
      new_node();
      p->nodes[tk.node].children_offset = c;
      update_vb(p->nodes[tk.node], tk.offset, tk.count);
      int end;
      if (!below) {
        float cbb[6];
        const float *cb = cb_root;
        if (!tk.root) {
          bb_reset(cbb);
          for (int i = tk.offset + tk.count - 1; i >= tk.offset; i--) {
            bb_expand(cbb, bbc[p->prim_indices[i]].bcentroid);
          }
          cb = cbb;
        }
        const int axis = bb_widest_axis(cb);
        end = partition(tk.offset, tk.offset + tk.count - 1, axis,
                        (cb[3 + axis] + cb[axis]) * 0.5f);
      } else {
        end = partition_material(tk.offset, tk.offset + tk.count - 1);
      }
      // right pushed first so the left subtree is built (and numbered) first
      stack.push_back({c + 1, end, tk.offset + tk.count - end, false});
      stack.push_back({c, tk.offset, end - tk.offset, false});
    }
    for (size_t i = 0; i < p->nodes.size(); i++) p->nodes[i].index = (int)i;
  }
};

}  // namespace

void sg_build_tree(PBVH *p, const float *co, const int *tris, int T, int V) {
  Builder b;
  b.p = p;
  b.co = co;
  b.tris = tris;
  b.bbc.resize((size_t)T);
  parallel_for(T, [&](int i) {  // per-prim BB + centroid (pbvh.c:572-586, :98-103)
    BBC &x = b.bbc[i];
    bb_reset(x.bmin);
    for (int j = 0; j < 3; j++) bb_expand(x.bmin, co + 3 * (size_t)tris[3 * (size_t)i + j]);
    for (int k = 0; k < 3; k++) x.bcentroid[k] = (x.bmin[k] + x.bmax[k]) * 0.5f;
  });
  b.vert_bitmap.assign((size_t)V, 0);
  b.map_local.assign((size_t)V, INT32_MIN);
  if (T > 0) b.build(T);

  // DFS leaf order (pbvh_iter_next pushes right then left: leaves come out left->right)
  p->leaf_node.clear();
  if (!p->nodes.empty()) {
    std::vector<int> st{0};
    while (!st.empty()) {
      const int ni = st.back();
      st.pop_back();
      PBVHNode &n = p->nodes[ni];
      if (n.flag & PBVH_Leaf) {
        n.leaf = (int)p->leaf_node.size();
        p->leaf_node.push_back(ni);
      } else {
        st.push_back(n.children_offset + 1);
        st.push_back(n.children_offset);
      }
    }
  }
  // vertex renumbering: uniq verts leaf by leaf, then unreferenced verts
  p->perm_new_of_old.assign((size_t)V, -1);
  p->perm_old_of_new.clear();
  p->perm_old_of_new.reserve((size_t)V);
  for (int ni : p->leaf_node) {
    const PBVHNode &n = p->nodes[ni];
    for (int k = 0; k < n.uniq_verts; k++) {
      const int v = n.vert_indices[k];
      p->perm_new_of_old[v] = (int)p->perm_old_of_new.size();
      p->perm_old_of_new.push_back(v);
    }
  }
  for (int v = 0; v < V; v++) {
    if (p->perm_new_of_old[v] < 0) {
      p->perm_new_of_old[v] = (int)p->perm_old_of_new.size();
      p->perm_old_of_new.push_back(v);
    }
  }
}

// BKE_pbvh_build_grids (pbvh.c:598-643): a grid's box over all its elements in element
// order, its centroid, the root centroid box in grid order, then the same top-down
// build as for triangles (leaf_limit already divided by grid_size^2 by the caller).
// Fills prim_indices, nodes and leaf_node (DFS leaf order).
void sg_build_tree_grids(PBVH *p) {
  Builder b;
  b.p = p;
  b.co = nullptr;
  b.tris = nullptr;
  b.grids = true;
  const int G = p->totgrid, ga = p->gridkey.grid_area;
  b.bbc.resize((size_t)G);
  parallel_for(G, [&](int i) {
    BBC &x = b.bbc[i];
    bb_reset(x.bmin);
    const unsigned char *base = (const unsigned char *)p->grids[i];
    for (int j = 0; j < ga; j++) bb_expand(x.bmin, (const float *)(base + (size_t)j * p->gridkey.elem_size));
    for (int k = 0; k < 3; k++) x.bcentroid[k] = (x.bmin[k] + x.bmax[k]) * 0.5f;
  });
  if (G > 0) b.build(G);
  p->leaf_node.clear();
  if (!p->nodes.empty()) {
    std::vector<int> st{0};
    while (!st.empty()) {
      const int ni = st.back();
      st.pop_back();
      PBVHNode &n = p->nodes[ni];
      if (n.flag & PBVH_Leaf) {
        n.leaf = (int)p->leaf_node.size();
        p->leaf_node.push_back(ni);
      } else {
        st.push_back(n.children_offset + 1);
        st.push_back(n.children_offset);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Device layout.
// ---------------------------------------------------------------------------

namespace {

template <class T>
int upload(T **dst, const T *src, size_t n, hipStream_t s) {
  if (n == 0) n = 1;
  if (sg_check_hip(hipMalloc((void **)dst, n * sizeof(T)), "hipMalloc")) return -1;
  if (src) {
    if (sg_check_hip(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s),
                     "hipMemcpyAsync"))
      return -1;
  }
  return 0;
}

template <class T>
int alloc_fill(T **dst, size_t n, int byte, hipStream_t s) {
  if (n == 0) n = 1;
  if (sg_check_hip(hipMalloc((void **)dst, n * sizeof(T)), "hipMalloc")) return -1;
  return sg_check_hip(hipMemsetAsync(*dst, byte, n * sizeof(T), s), "hipMemsetAsync");
}

}  // namespace

void sg_build_layout(PBVH *p, const float *co, SgLayout &h) {
  const int V = p->totvert, T = p->totprim;
  const int L = (int)p->leaf_node.size();
  const int N = (int)p->nodes.size();
  h.V = V;
  h.T = T;
  h.L = L;
  h.N = N;

  // positions SoA in the new numbering
  std::vector<float> &x = h.x, &y = h.y, &z = h.z;
  x.resize(V);
  y.resize(V);
  z.resize(V);
  parallel_for(V, [&](int i) {
    const int o = p->perm_old_of_new[i];
    x[i] = co[3 * (size_t)o];
    y[i] = co[3 * (size_t)o + 1];
    z[i] = co[3 * (size_t)o + 2];
  });

  // triangles in prim_indices order
  std::vector<int> tri_new_of_old(T);
  for (int i = 0; i < T; i++) tri_new_of_old[p->prim_indices[i]] = i;
  std::vector<int4> tri(T);
  parallel_for(T, [&](int i) {
    const int o = p->prim_indices[i];
    tri[i] = make_int4(p->perm_new_of_old[p->tris[3 * (size_t)o]],
                       p->perm_new_of_old[p->tris[3 * (size_t)o + 1]],
                       p->perm_new_of_old[p->tris[3 * (size_t)o + 2]], o);
  });

  // vertex -> tri CSR (ascending original tri id), entries tri | corner << 30
  std::vector<int> &vt_off = h.vt_off, vt_idx(3 * (size_t)T);
  vt_off.assign(V + 1, 0);
  for (size_t k = 0; k < 3 * (size_t)T; k++) vt_off[p->perm_new_of_old[p->tris[k]] + 1]++;
  for (int v = 0; v < V; v++) vt_off[v + 1] += vt_off[v];
  {
    std::vector<int> fill(vt_off.begin(), vt_off.end() - 1);
    for (int t = 0; t < T; t++) {
      for (int j = 0; j < 3; j++) {
        const int v = p->perm_new_of_old[p->tris[3 * (size_t)t + j]];
        vt_idx[fill[v]++] = tri_new_of_old[t] | (j << 30);
      }
    }
  }
  // polygon meshes: corners in new ids and each vertex's (polygon, corner) list
  if (p->poly_mode) {
    const int P = (int)p->poly_off.size() - 1;
    h.poly_v.resize(p->poly_verts.size());
    for (size_t k = 0; k < h.poly_v.size(); k++) h.poly_v[k] = p->perm_new_of_old[p->poly_verts[k]];
    h.vc_off.assign(V + 1, 0);
    for (int k : h.poly_v) h.vc_off[k + 1]++;
    for (int v = 0; v < V; v++) h.vc_off[v + 1] += h.vc_off[v];
    h.vc.resize((size_t)std::max(h.vc_off[V], 1));
    std::vector<int> fill(h.vc_off.begin(), h.vc_off.end() - 1);
    for (int q = 0; q < P; q++)
      for (int k = p->poly_off[q]; k < p->poly_off[q + 1]; k++)
        h.vc[fill[h.poly_v[k]]++] = make_int2(q, k - p->poly_off[q]);
  }
  // vertex -> vertex CSR: distinct edge neighbours, ascending original id.  The edges
  // are the mesh's: the triangles' edges, or for polygons the polygon edges (the
  // corners before and after the vertex in each of its polygons).
  std::vector<int> vv_cnt(V, 0);
  std::vector<std::vector<int>> tmp_tls;
  auto neighbours = [&](int v, std::vector<int> &buf) {
    buf.clear();
    if (p->poly_mode) {
      for (int q = h.vc_off[v]; q < h.vc_off[v + 1]; q++) {
        const int2 pc = h.vc[q];
        const int b = p->poly_off[pc.x], n = p->poly_off[pc.x + 1] - b;
        const int up = p->poly_verts[b + (pc.y + n - 1) % n], un = p->poly_verts[b + (pc.y + 1) % n];
        const int self = p->perm_old_of_new[v];
        if (up != self) buf.push_back(up);
        if (un != self) buf.push_back(un);
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      return;
    }
    for (int q = vt_off[v]; q < vt_off[v + 1]; q++) {
      const int t = vt_idx[q] & 0x3fffffff;
      const int4 c = tri[t];
      const int cs[3] = {c.x, c.y, c.z};
      for (int j = 0; j < 3; j++) {
        if (cs[j] != v) buf.push_back(p->perm_old_of_new[cs[j]]);
      }
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  {
    const int nt = 16;
    std::vector<std::thread> th;
    const int chunk = (V + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
      th.emplace_back([&, t]() {
        std::vector<int> buf;
        for (int v = t * chunk; v < std::min(V, (t + 1) * chunk); v++) {
          neighbours(v, buf);
          vv_cnt[v] = (int)buf.size();
        }
      });
    }
    for (auto &h : th) h.join();
  }
  std::vector<int> &vv_off = h.vv_off, &vv_idx = h.vv_idx;
  vv_off.assign(V + 1, 0);
  for (int v = 0; v < V; v++) vv_off[v + 1] = vv_off[v] + vv_cnt[v];
  vv_idx.assign((size_t)vv_off[V] + 1, 0);
  {
    const int nt = 16;
    std::vector<std::thread> th;
    const int chunk = (V + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
      th.emplace_back([&, t]() {
        std::vector<int> buf;
        for (int v = t * chunk; v < std::min(V, (t + 1) * chunk); v++) {
          neighbours(v, buf);
          for (size_t k = 0; k < buf.size(); k++) vv_idx[vv_off[v] + k] = p->perm_new_of_old[buf[k]];
        }
      });
    }
    for (auto &h : th) h.join();
  }
  {
    // flattened incidences: the corners of each incident tri, in CSR order
    std::vector<int4> &vt4 = h.vt4;
    vt4.resize(3 * (size_t)T);
    parallel_for((int)vt4.size(), [&](int q) {
      const int e = vt_idx[q];
      const int4 c = tri[e & 0x3fffffff];
      vt4[q] = make_int4(c.x, c.y, c.z, e);
    });
  }

  // leaves
  std::vector<int> &leaf_vbeg = h.leaf_vbeg, &leaf_sh_off = h.leaf_sh_off, &leaf_flags = h.leaf_flags;
  std::vector<int> &leaf_sh = h.leaf_sh;
  std::vector<float> &leaf_bb = h.leaf_bb, &leaf_obb = h.leaf_obb;
  leaf_vbeg.assign(L + 1, 0);
  leaf_sh_off.assign(L + 1, 0);
  leaf_flags.assign(L, 0);
  leaf_sh.clear();
  leaf_bb.assign((size_t)L * 8, 0.0f);
  leaf_obb.assign((size_t)L * 8, 0.0f);
  for (int l = 0; l < L; l++) {
    const PBVHNode &n = p->nodes[p->leaf_node[l]];
    leaf_vbeg[l + 1] = leaf_vbeg[l] + n.uniq_verts;
    for (int k = 0; k < n.face_verts; k++) leaf_sh.push_back(p->perm_new_of_old[n.vert_indices[n.uniq_verts + k]]);
    leaf_sh_off[l + 1] = (int)leaf_sh.size();
    leaf_flags[l] = (n.flag & PBVH_FullyHidden) ? 1 : 0;
    for (int k = 0; k < 3; k++) {
      leaf_bb[8 * (size_t)l + k] = n.vb[k];
      leaf_bb[8 * (size_t)l + 4 + k] = n.vb[3 + k];
      leaf_obb[8 * (size_t)l + k] = n.orig_vb[k];
      leaf_obb[8 * (size_t)l + 4 + k] = n.orig_vb[3 + k];
    }
  }
}

// Leaf ranges and heights per node (children have larger indices than parents) and
// the search cut: maximal subtrees with <= K leaves, K the smallest power of two >= 64
// that keeps the cut at <= 1024 subtrees (search cost C + passing*K; LDS tables).
// Fills p->cut and leaf_cut[L].
void sg_build_cut(PBVH *p, std::vector<int> &lbeg, std::vector<int> &lend, std::vector<int> &height,
                  std::vector<int> &leaf_cut) {
  const int N = (int)p->nodes.size();
  const int L = (int)p->leaf_node.size();
  lbeg.assign(N, 0);
  lend.assign(N, 0);
  height.assign(N, 0);
  for (int i = N - 1; i >= 0; i--) {
    const PBVHNode &n = p->nodes[i];
    if (n.flag & PBVH_Leaf) {
      lbeg[i] = n.leaf;
      lend[i] = n.leaf + 1;
      height[i] = 0;
    } else {
      const int c = n.children_offset;
      lbeg[i] = lbeg[c];
      lend[i] = lend[c + 1];
      height[i] = 1 + std::max(height[c], height[c + 1]);
    }
  }
  int K = 64;
  std::vector<int> cut_nodes;
  for (;;) {
    cut_nodes.clear();
    if (N > 0) {
      std::vector<int> st{0};
      while (!st.empty()) {
        const int ni = st.back();
        st.pop_back();
        if (lend[ni] - lbeg[ni] <= K) {
          cut_nodes.push_back(ni);
        } else {
          st.push_back(p->nodes[ni].children_offset + 1);
          st.push_back(p->nodes[ni].children_offset);
        }
      }
    }
    if ((int)cut_nodes.size() <= 1024 || K >= 1024) break;
    K *= 2;
  }
  const int C = (int)cut_nodes.size();
  p->cut.node = cut_nodes;
  p->cut.lbeg.resize(C);
  p->cut.lend.resize(C);
  leaf_cut.assign(L, 0);
  for (int c = 0; c < C; c++) {
    p->cut.lbeg[c] = lbeg[cut_nodes[c]];
    p->cut.lend[c] = lend[cut_nodes[c]];
    for (int l = p->cut.lbeg[c]; l < p->cut.lend[c]; l++) leaf_cut[l] = c;
  }
}

// Halo tables of the fused dab kernel.  For leaf l with uniq range U (new ids):
//   ring 0 = (shared vertices of l) and (edge neighbours of U), minus U;
//   ring j = edge neighbours of ring j-1 (ring 0 counting U) not yet listed, j <= R.
// Local ids: U first (its own order), then ring 0, ring 1, ..., each ring by ascending
// vertex id.  A dab's results for l need final positions of U and ring 0 only (the
// normals' corners and the leaf box's shared vertices are there); Smooth(k) needs rings
// up to k, and a vertex of ring j is exact after k - j Jacobi steps.  Every vertex of a
// ring below R has all its edge neighbours among the locals; their local ids are listed
// in vertex->vertex order, so the local Jacobi sums in the same order as the global one.
// Returns false (the fused path off) when a leaf has more than 65535 locals.
static bool build_halo(PBVH *p, const SgLayout &h, hipStream_t s) {
  SgDev &d = p->d;
  const int L = h.L, V = h.V, R = SG_HALO_RINGS;
  struct Leaf {
    std::vector<int2> ext;
    int ring_end[SG_HALO_RINGS + 1];  // ext count at the end of ring j
    std::vector<int> nboff;           // relative to the leaf's neighbour block
    std::vector<unsigned short> nbl;
    std::vector<uint2> lt;            // faces incident to the uniq vertices: corners as local ids
    bool ok = true;
  };
  std::vector<Leaf> lv(L);
  auto owner_of = [&](int v) {
    return (int)(std::upper_bound(h.leaf_vbeg.begin(), h.leaf_vbeg.end(), v) - h.leaf_vbeg.begin()) - 1;
  };
  std::vector<unsigned short> ti(h.vt4.size());  // per incidence: local face index in its vertex's leaf
  std::vector<unsigned short> shl(h.leaf_sh.size());
  std::atomic<int> bad{0};
  const int nt = 16;
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) {
    th.emplace_back([&, t]() {
      // open addressing vertex -> local id, cleared after each leaf
      std::vector<long long> tab(1 << 12, -1);
      std::vector<int> used;
      auto find = [&](int v) -> int {
        size_t m = tab.size() - 1, i = ((size_t)v * 2654435761u) & m;
        while (tab[i] >= 0) {
          if ((int)(tab[i] >> 32) == v) return (int)(tab[i] & 0xffffffff);
          i = (i + 1) & m;
        }
        return -1;
      };
      auto insert = [&](int v, int loc) {
        size_t m = tab.size() - 1, i = ((size_t)v * 2654435761u) & m;
        while (tab[i] >= 0) i = (i + 1) & m;
        tab[i] = ((long long)v << 32) | (unsigned)loc;
        used.push_back((int)i);
      };
      std::vector<int> ring, next, all;
      std::unordered_map<int, int> face;  // face id -> local face index (per leaf)
      for (int l = t; l < L; l += nt) {
        Leaf &o = lv[l];
        const int vb = h.leaf_vbeg[l], ve = h.leaf_vbeg[l + 1], nu = ve - vb;
        all.clear();
        for (int v = vb; v < ve; v++) all.push_back(v);
        // ring 0
        next.clear();
        for (int k = h.leaf_sh_off[l]; k < h.leaf_sh_off[l + 1]; k++) next.push_back(h.leaf_sh[k]);
        for (int v = vb; v < ve; v++)
          for (int q = h.vv_off[v]; q < h.vv_off[v + 1]; q++) next.push_back(h.vv_idx[q]);
        std::sort(next.begin(), next.end());
        next.erase(std::unique(next.begin(), next.end()), next.end());
        ring.clear();
        for (int v : next)
          if (v < vb || v >= ve) ring.push_back(v);
        o.ring_end[0] = (int)ring.size();
        for (int v : ring) all.push_back(v);
        std::vector<int> prev = ring;
        for (int j = 1; j <= R; j++) {
          // grow the table when it would pass half full
          while (tab.size() < 2 * (all.size() + 8 * prev.size() + 64)) {
            tab.assign(tab.size() * 2, -1);
            used.clear();
          }
          for (size_t i = 0; i < used.size(); i++) tab[used[i]] = -1;
          used.clear();
          for (size_t i = 0; i < all.size(); i++) insert(all[i], (int)i);
          next.clear();
          const std::vector<int> &src = (j == 1) ? all : prev;  // ring 1 grows from U and ring 0
          for (int u : src)
            for (int q = h.vv_off[u]; q < h.vv_off[u + 1]; q++) {
              const int w = h.vv_idx[q];
              if (find(w) < 0) next.push_back(w);
            }
          std::sort(next.begin(), next.end());
          next.erase(std::unique(next.begin(), next.end()), next.end());
          for (int v : next) all.push_back(v);
          o.ring_end[j] = o.ring_end[j - 1] + (int)next.size();
          prev = next;
        }
        if (all.size() > 65535) {
          o.ok = false;
          bad++;
          continue;
        }
        for (size_t i = 0; i < used.size(); i++) tab[used[i]] = -1;
        used.clear();
        while (tab.size() < 2 * all.size() + 64) tab.assign(tab.size() * 2, -1);
        for (size_t i = 0; i < all.size(); i++) insert(all[i], (int)i);
        o.ext.resize(all.size() - nu);
        for (size_t i = nu; i < all.size(); i++) o.ext[i - nu] = make_int2(all[i], owner_of(all[i]));
        // neighbour lists of the locals before the last ring
        const int inner = nu + o.ring_end[R - 1];
        o.nboff.assign(inner + 1, 0);
        // every id below must be a local of the right rings; a miss turns the fused path off
        const int nl0 = nu + o.ring_end[0], nlR = (int)all.size();
        bool miss = false;
        for (int i = 0; i < inner; i++) {
          const int u = all[i];
          for (int q = h.vv_off[u]; q < h.vv_off[u + 1]; q++) {
            const int w = find(h.vv_idx[q]);
            miss |= w < 0 || w >= nlR;
            o.nbl.push_back((unsigned short)w);
          }
          o.nboff[i + 1] = (int)o.nbl.size();
        }
        // faces incident to the uniq vertices, numbered in order of first incidence
        // (vertices in order, each in vertex->tri order); corners as local ids
        {
          face.clear();
          for (int v = vb; v < ve; v++)
            for (int q = h.vt_off[v]; q < h.vt_off[v + 1]; q++) {
              const int4 c = h.vt4[q];
              const int t = c.w & 0x3fffffff;
              auto it = face.find(t);
              int j;
              if (it == face.end()) {
                j = (int)o.lt.size();
                face.emplace(t, j);
                const int a = find(c.x), b = find(c.y), e = find(c.z);
                miss |= a < 0 || b < 0 || e < 0 || a >= nl0 || b >= nl0 || e >= nl0;
                o.lt.push_back(make_uint2((unsigned)a | ((unsigned)b << 16), (unsigned)e));
              } else {
                j = it->second;
              }
              miss |= j > 65535;
              ti[q] = (unsigned short)j;
            }
        }
        for (int k = h.leaf_sh_off[l]; k < h.leaf_sh_off[l + 1]; k++) {
          const int w = find(h.leaf_sh[k]);
          miss |= w < 0 || w >= nl0;
          shl[k] = (unsigned short)w;
        }
        if (miss) {
          o.ok = false;
          bad++;
        }
        for (size_t i = 0; i < used.size(); i++) tab[used[i]] = -1;
        used.clear();
      }
    });
  }
  for (auto &x : th) x.join();
  if (bad) return false;
  // concatenate
  std::vector<int4> rec(2 * (size_t)std::max(L, 1));
  size_t next_ext = 0, next_nb = 0, next_nbl = 0;
  int maxloc = 0, maxloc0 = 0, maxnbl = 0, maxsh = 0, maxlt = 0;
  for (int l = 0; l < L; l++) {
    const Leaf &o = lv[l];
    const int nu = h.leaf_vbeg[l + 1] - h.leaf_vbeg[l];
    maxnbl = std::max(maxnbl, (int)o.nbl.size());
    maxlt = std::max(maxlt, (int)o.lt.size());
    maxsh = std::max(maxsh, h.leaf_sh_off[l + 1] - h.leaf_sh_off[l]);
    const int ns = h.leaf_sh_off[l + 1] - h.leaf_sh_off[l];
    if (nu > 65535 || ns > 65535 || o.ring_end[R] > 65535 || o.lt.size() > 65535)
      return false;  // the record's 16-bit fields
    rec[2 * l] = make_int4(h.leaf_vbeg[l], h.leaf_sh_off[l], (int)next_ext, (int)next_nb);
    rec[2 * l + 1] = make_int4(nu | (ns << 16), o.ring_end[0] | (o.ring_end[1] << 16),
                               o.ring_end[2] | (o.ring_end[3] << 16), o.ring_end[4] | ((int)o.lt.size() << 16));
    next_ext += o.ext.size();
    next_nb += o.nboff.size() + 2 * o.lt.size();
    next_nbl += o.nbl.size();
    maxloc = std::max(maxloc, nu + o.ring_end[R]);
    maxloc0 = std::max(maxloc0, nu + o.ring_end[0]);
    for (int j = 0; j <= R; j++) d.halo_maxlocr[j] = std::max(d.halo_maxlocr[j], nu + o.ring_end[j]);
  }
  if (next_ext >= (1ull << 31) || next_nbl >= (1ull << 31)) return false;
  std::vector<int2> ext(std::max<size_t>(next_ext, 1));
  std::vector<int> nboff(std::max<size_t>(next_nb, 1));
  std::vector<unsigned short> nbl(std::max<size_t>(next_nbl, 1));
  {
    size_t e = 0, b = 0, n = 0;
    for (int l = 0; l < L; l++) {
      const Leaf &o = lv[l];
      std::copy(o.ext.begin(), o.ext.end(), ext.begin() + e);
      for (size_t i = 0; i < o.nboff.size(); i++) nboff[b + i] = (int)n + o.nboff[i];
      for (size_t i = 0; i < o.lt.size(); i++) {
        nboff[b + o.nboff.size() + 2 * i] = (int)o.lt[i].x;
        nboff[b + o.nboff.size() + 2 * i + 1] = (int)o.lt[i].y;
      }
      std::copy(o.nbl.begin(), o.nbl.end(), nbl.begin() + n);
      e += o.ext.size();
      b += o.nboff.size() + 2 * o.lt.size();
      n += o.nbl.size();
    }
  }
  lv.clear();
  lv.shrink_to_fit();
  int rc = 0;
  rc |= upload(&d.hl_rec, rec.data(), rec.size(), s);
  rc |= upload(&d.hl_v, ext.data(), ext.size(), s);
  rc |= upload(&d.hl_nboff, nboff.data(), nboff.size(), s);
  rc |= upload(&d.hl_nbl, nbl.data(), nbl.size(), s);
  rc |= upload(&d.hl_ti, ti.data(), std::max<size_t>(ti.size(), 1), s);
  rc |= upload(&d.hl_sh, shl.data(), std::max<size_t>(shl.size(), 1), s);
  rc |= alloc_fill(&d.leaf_gstamp, std::max(L, 1), 0xFF, s);  // -1: no dab's stamp (stamps start at 1)
  rc |= alloc_fill(&d.dstamp, 2, 0, s);
  rc |= alloc_fill(&d.pend, 4, 0xFF, s);  // no fused dab pending
  rc |= alloc_fill(&d.mx, V, 0, s);
  rc |= alloc_fill(&d.my, V, 0, s);
  rc |= alloc_fill(&d.mz, V, 0, s);
  rc |= alloc_fill(&d.gl_rec, 2 * std::max(L, 1), 0, s);
  hipStreamSynchronize(s);  // the host tables are freed on return
  d.halo_maxloc = maxloc;
  d.halo_maxloc0 = maxloc0;
  d.halo_maxnbl = maxnbl;
  d.leaf_maxsh = maxsh;
  d.halo_maxlt = maxlt;
  d.halo_bytes = (long long)(rec.size() * 16 + ext.size() * 8 + nboff.size() * 4 + nbl.size() * 2 + ti.size() * 2 +
                             shl.size() * 2);
  return rc == 0;
}

void sg_build_device(PBVH *p, const float *co) {
  SgDev &d = p->d;
  hipStream_t s = p->stream;
  SgLayout h;
  sg_build_layout(p, co, h);
  const int V = h.V, T = h.T, L = h.L, N = h.N;
  (void)T;
  d.V = V;
  d.T = h.T;
  d.L = L;
  d.N = N;
  upload(&d.px, h.x.data(), V, s);
  upload(&d.py, h.y.data(), V, s);
  upload(&d.pz, h.z.data(), V, s);
  upload(&d.qx, h.x.data(), V, s);
  upload(&d.qy, h.y.data(), V, s);
  upload(&d.qz, h.z.data(), V, s);
  alloc_fill(&d.nx, V, 0, s);
  alloc_fill(&d.ny, V, 0, s);
  alloc_fill(&d.nz, V, 0, s);
  upload(&d.vt4, h.vt4.data(), h.vt4.size(), s);
  upload(&d.vt_off, h.vt_off.data(), V + 1, s);
  upload(&d.vv_off, h.vv_off.data(), V + 1, s);
  upload(&d.vv_idx, h.vv_idx.data(), h.vv_idx.size(), s);
  upload(&d.leaf_vbeg, h.leaf_vbeg.data(), L + 1, s);
  p->leaf_vbeg = h.leaf_vbeg;
  upload(&d.leaf_sh_off, h.leaf_sh_off.data(), L + 1, s);
  upload(&d.leaf_sh, h.leaf_sh.data(), h.leaf_sh.size(), s);
  upload(&d.leaf_flags, h.leaf_flags.data(), L, s);
  upload(&d.leaf_bb, h.leaf_bb.data(), (size_t)L * 8, s);
  upload(&d.leaf_obb, h.leaf_obb.data(), (size_t)L * 8, s);
  hipStreamSynchronize(s);  // the layout's host arrays are freed on return

  // nodes
  std::vector<float> node_bb((size_t)N * 8, 0.0f), node_obb((size_t)N * 8, 0.0f);
  for (int i = 0; i < N; i++) {
    const PBVHNode &n = p->nodes[i];
    for (int k = 0; k < 3; k++) {
      node_bb[8 * (size_t)i + k] = n.vb[k];
      node_bb[8 * (size_t)i + 4 + k] = n.vb[3 + k];
      node_obb[8 * (size_t)i + k] = n.orig_vb[k];
      node_obb[8 * (size_t)i + 4 + k] = n.orig_vb[3 + k];
    }
  }
  upload(&d.node_bb, node_bb.data(), (size_t)N * 8, s);
  upload(&d.node_obb, node_obb.data(), (size_t)N * 8, s);

  std::vector<int> lbeg, lend, height, leaf_cut;
  sg_build_cut(p, lbeg, lend, height, leaf_cut);
  const std::vector<int> &cut_nodes = p->cut.node;
  const int C = (int)cut_nodes.size();
  d.C = C;
  upload(&d.cut_node, cut_nodes.data(), C, s);
  {
    std::vector<int> cflags(C, 0);
    std::vector<float> cbb((size_t)C * 8, 0.0f);
    for (int c = 0; c < C; c++) {
      const PBVHNode &n = p->nodes[cut_nodes[c]];
      if ((n.flag & PBVH_Leaf) && (n.flag & PBVH_FullyHidden)) cflags[c] = 1;
      for (int k = 0; k < 3; k++) {
        cbb[8 * (size_t)c + k] = n.vb[k];
        cbb[8 * (size_t)c + 4 + k] = n.vb[3 + k];
      }
    }
    upload(&d.cut_flags, cflags.data(), C, s);
    upload(&d.cut_bb, cbb.data(), cbb.size(), s);
    upload(&d.cut_obb, cbb.data(), cbb.size(), s);
    hipStreamSynchronize(s);
  }
  upload(&d.cut_lbeg, p->cut.lbeg.data(), C, s);
  upload(&d.cut_lend, p->cut.lend.data(), C, s);
  upload(&d.leaf_cut, leaf_cut.data(), L, s);

  // deferred refit tables.  Local numbering inside a cut: its leaves 0..nl-1 in DFS
  // order, then its internal nodes by ascending height.
  std::vector<int> sub_off(C + 1, 0);
  std::vector<int4> sub_tab;
  int sub_maxn = 1;
  {
    std::vector<int> local(N, -1);
    for (int c = 0; c < C; c++) {
      const int root = cut_nodes[c];
      std::vector<int> inner;
      std::vector<int> st{root};
      while (!st.empty()) {
        const int ni = st.back();
        st.pop_back();
        const PBVHNode &n = p->nodes[ni];
        if (n.flag & PBVH_Leaf) {
          local[ni] = n.leaf - p->cut.lbeg[c];
        } else {
          inner.push_back(ni);
          st.push_back(n.children_offset);
          st.push_back(n.children_offset + 1);
        }
      }
      std::stable_sort(inner.begin(), inner.end(), [&](int a, int b) { return height[a] < height[b]; });
      const int nl = p->cut.lend[c] - p->cut.lbeg[c];
      for (size_t k = 0; k < inner.size(); k++) local[inner[k]] = nl + (int)k;
      for (int ni : inner) {
        const int ch = p->nodes[ni].children_offset;
        sub_tab.push_back(make_int4(ni, local[ch], local[ch + 1], height[ni]));
      }
      sub_off[c + 1] = (int)sub_tab.size();
      sub_maxn = std::max(sub_maxn, nl + (int)inner.size());
    }
    // top tree: strict ancestors of cut roots; local 0..C-1 = cut roots
    std::vector<int> top_local(N, -1);
    for (int c = 0; c < C; c++) top_local[cut_nodes[c]] = c;
    std::vector<int> top;
    if (N > 0 && top_local[0] < 0) {
      std::vector<int> st{0};
      while (!st.empty()) {
        const int ni = st.back();
        st.pop_back();
        if (top_local[ni] >= 0) continue;
        top.push_back(ni);
        st.push_back(p->nodes[ni].children_offset);
        st.push_back(p->nodes[ni].children_offset + 1);
      }
    }
    // height within the top tree (cut roots = 0)
    std::vector<int> th(N, 0);
    std::vector<int> order = top;
    std::sort(order.begin(), order.end(), [](int a, int b) { return a > b; });  // children first
    for (int ni : order) {
      const int ch = p->nodes[ni].children_offset;
      const int h0 = top_local[ch] >= 0 && top_local[ch] < C ? 0 : th[ch];
      const int h1 = top_local[ch + 1] >= 0 && top_local[ch + 1] < C ? 0 : th[ch + 1];
      th[ni] = 1 + std::max(h0, h1);
    }
    std::stable_sort(top.begin(), top.end(), [&](int a, int b) { return th[a] < th[b]; });
    for (size_t k = 0; k < top.size(); k++) top_local[top[k]] = C + (int)k;
    std::vector<int4> top_tab;
    int maxh = 0;
    for (int ni : top) {
      const int ch = p->nodes[ni].children_offset;
      top_tab.push_back(make_int4(ni, top_local[ch], top_local[ch + 1], th[ni]));
      maxh = std::max(maxh, th[ni]);
    }
    d.top_n = (int)top.size();
    d.top_maxh = maxh;
    d.sub_maxn = sub_maxn;
    upload(&d.sub_off, sub_off.data(), C + 1, s);
    upload(&d.sub_tab, sub_tab.data(), sub_tab.size(), s);
    upload(&d.top_tab, top_tab.data(), top_tab.size(), s);
  }
  alloc_fill(&d.cut_dirty, C, 0, s);
  alloc_fill(&d.draw_dirty, C, 0, s);  // subtrees the dabs touched since the last draw refill
  alloc_fill(&d.undo_stamp, L, 0xFF, s);  // -1: saved by no undo step
  alloc_fill(&d.undo_slot, L, 0, s);
  alloc_fill(&d.undo, 1, 0, s);          // undo recording off
  alloc_fill(&d.place_pair, 16, 0, s);   // count 0: no ray of a Grab placement has arrived
  alloc_fill(&d.cut_touch, C, 0x80, s);  // INT_MIN-ish: never "touched by the previous dab"
  alloc_fill(&d.prev_cuts, C, 0, s);

  d.nuniq = d.nuniq_visible = h.leaf_vbeg[L];
  d.nhidden_leaves = 0;
  for (int l = 0; l < L; l++) d.nhidden_leaves += h.leaf_flags[l] ? 1 : 0;
  if (!p->vhide.empty()) {
    std::vector<int> vh(V);
    for (int i = 0; i < V; i++) vh[i] = p->vhide[p->perm_old_of_new[i]];
    for (int i = 0; i < h.leaf_vbeg[L]; i++) d.nuniq_visible -= vh[i] ? 1 : 0;
    upload(&d.vhide, vh.data(), V, s);
  }
  alloc_fill(&d.vmark, V, 0, s);
  // scratch
  alloc_fill(&d.gathered, L, 0, s);

  alloc_fill(&d.gl_v, L, 0, s);
  alloc_fill(&d.gl_s, L, 0, s);
  {
    int maxu = 1;
    for (int l = 0; l < L; l++) maxu = std::max(maxu, h.leaf_vbeg[l + 1] - h.leaf_vbeg[l]);
    d.aff_shift = 6;
    while ((1 << d.aff_shift) < maxu) d.aff_shift++;
  }
  const size_t na = std::max<size_t>((size_t)L << d.aff_shift, 1);
  alloc_fill(&d.aff_v, na, 0, s);
  alloc_fill(&d.aff_vt, na, 0, s);
  alloc_fill(&d.aff_vv, na, 0, s);
  alloc_fill(&d.aff_fade, na, 0, s);
  alloc_fill(&d.aff_cnt, std::max(L, 1), 0, s);
  alloc_fill(&d.ctr, 2 * SG_CTR_COUNT, 0, s);
  alloc_fill(&d.halt, 4, 0, s);
  alloc_fill(&d.seq, 1, 0, s);
  alloc_fill(&d.vstamp, V, 0xFF, s);  // -1: never a dab's stamp (stamps start at 1)
  alloc_fill(&d.err, 4, 0, s);
  upload(&d.vperm_old, p->perm_old_of_new.data(), V, s);
  if (p->poly_mode) {
    // polygons for the polygon normal rules: corners (new ids), the polygon of each
    // tri in leaf order, each vertex's (polygon, corner) list
    const int P = (int)p->poly_off.size() - 1;
    std::vector<int> pt((size_t)std::max(p->totprim, 1));
    for (int i = 0; i < p->totprim; i++) pt[i] = (int)p->looptri[p->prim_indices[i]].poly;
    upload(&d.poly_off, p->poly_off.data(), P + 1, s);
    upload(&d.poly_v, h.poly_v.data(), std::max<size_t>(h.poly_v.size(), 1), s);
    upload(&d.poly_of_tri, pt.data(), pt.size(), s);
    upload(&d.vc_off, h.vc_off.data(), V + 1, s);
    upload(&d.vc, h.vc.data(), h.vc.size(), s);
    hipStreamSynchronize(s);  // the host arrays are freed on return
  } else {
    d.halo_ok = build_halo(p, h, s);
  }
  hipStreamSynchronize(s);
}

void sg_free_device(PBVH *p) {
  sg_undo_release(p);
  SgDev &d = p->d;
  void *ptrs[] = {d.px, d.py, d.pz, d.qx, d.qy, d.qz, d.nx, d.ny, d.nz, d.mask, d.vt4,
                  d.vt_off, d.vv_off, d.vv_idx, d.leaf_vbeg, d.leaf_sh_off, d.leaf_sh,
                  d.leaf_flags, d.leaf_cut, d.leaf_bb, d.leaf_obb, d.node_bb, d.node_obb,
                  d.cut_bb, d.cut_obb, d.cut_flags, d.cut_node, d.cut_lbeg, d.cut_lend,
                  d.sub_off, d.sub_tab, d.top_tab_n, d.top_tab, d.cut_dirty, d.vhide, d.vmark,
                  d.gathered, d.gl_v, d.gl_s, d.aff_v, d.aff_vt, d.aff_vv, d.aff_fade, d.aff_cnt,
                  d.prev_cuts, d.cut_touch, d.halt, d.seq, d.vstamp, d.ctr, d.err, d.vperm_old, d.stamp_v, d.stamp_v2, d.stamp_t,
                  d.stamp_a, d.curve, d.log, d.ltri, d.leaf_toff, d.tri_poly, d.node_parent, d.leaf_node,
                  d.ray_in, d.ray_out, d.place, d.place_rays, d.rl_list, d.rl_args,
                  d.draw_dirty, d.draw_tri, d.draw_off, d.draw_leaf, d.draw_fset, d.draw_poly_off,
                  d.draw_poly_v, d.draw_vbo, d.undo, d.undo_stamp, d.undo_slot, d.place_pair,
                  d.prox_dev, d.prox_tab, d.prox_leaves, d.prox_off, d.poly_off, d.poly_v,
                  d.poly_of_tri, d.vc_off, d.vc, d.stage, d.hl_rec, d.hl_v, d.hl_nboff,
                  d.hl_nbl, d.hl_ti, d.hl_sh, d.leaf_gstamp, d.dstamp, d.mx, d.my, d.mz, d.pend,
                  d.gl_rec};
  for (void *q : ptrs) {
    if (q) hipFree(q);
  }
  if (d.prox_host) (void)hipHostFree(d.prox_host);
  d = SgDev();
}
