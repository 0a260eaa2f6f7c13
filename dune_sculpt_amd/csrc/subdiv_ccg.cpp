// Host SubdivCCG: grids of a coarse polygon mesh, their seam averaging and normals,
// and neighbour coordinates.  Follows kernel/intern/subdiv_ccg.c (cited per function).
// These run when grids are made or when a caller asks for them on the host; the
// per-dab normals and seam averaging of a PBVH over grids run on the GPU (grids.hip).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <unordered_map>

#include "sg_grids.h"
#include "sg_internal.h"

namespace {

// normalize_v3 (math_vector_inline.c:1165-1180): zero below 1e-35, else multiply by
// the rounded reciprocal of the correctly rounded sqrt.
inline float normalize3(float n[3]) {
  float d = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  if (d > 1.0e-35f) {
    d = sqrtf(d);
    const float inv = 1.0f / d;
    n[0] = n[0] * inv;
    n[1] = n[1] * inv;
    n[2] = n[2] * inv;
  } else {
    n[0] = n[1] = n[2] = 0.0f;
    d = 0.0f;
  }
  return d;
}

// normal_quad_v3 (math_geom.cc:48-67): the cross product of the two diagonals.
inline void quad_normal(float n[3], const float *v1, const float *v2, const float *v3, const float *v4) {
  const float a0 = v1[0] - v3[0], a1 = v1[1] - v3[1], a2 = v1[2] - v3[2];
  const float b0 = v2[0] - v4[0], b1 = v2[1] - v4[1], b2 = v2[2] - v4[2];
  n[0] = a1 * b2 - a2 * b1;
  n[1] = a2 * b0 - a0 * b2;
  n[2] = a0 * b1 - a1 * b0;
  normalize3(n);
}

inline SubdivCCGCoord coord(int g, int x, int y) {
  SubdivCCGCoord c;
  c.grid_index = g;
  c.x = (short)x;
  c.y = (short)y;
  return c;
}

// Element values a seam pass carries: co, normal, mask (subdiv_ccg.c:902-949).
struct Acc {
  float co[3], no[3], mask;
};
inline void acc_init(Acc &a) {
  a.co[0] = a.co[1] = a.co[2] = 0.0f;
  a.no[0] = a.no[1] = a.no[2] = 0.0f;
  a.mask = 0.0f;
}
inline void acc_add(const SubdivCCG *s, Acc &a, const SubdivCCGCoord &c) {
  const float *co = ccg_co(s, c.grid_index, c.x, c.y);
  a.co[0] += co[0];
  a.co[1] += co[1];
  a.co[2] += co[2];
  if (s->has_normal) {
    const float *no = ccg_no(s, c.grid_index, c.x, c.y);
    a.no[0] += no[0];
    a.no[1] += no[1];
    a.no[2] += no[2];
  }
  if (s->has_mask) a.mask += *ccg_mask(s, c.grid_index, c.x, c.y);
}
inline void acc_mul(Acc &a, float f) {
  for (int i = 0; i < 3; i++) {
    a.co[i] *= f;
    a.no[i] *= f;
  }
  a.mask *= f;
}
inline void acc_copy(const SubdivCCG *s, const SubdivCCGCoord &c, const Acc &a) {
  memcpy(ccg_co(s, c.grid_index, c.x, c.y), a.co, sizeof(a.co));
  if (s->has_normal) memcpy(ccg_no(s, c.grid_index, c.x, c.y), a.no, sizeof(a.no));
  if (s->has_mask) *ccg_mask(s, c.grid_index, c.x, c.y) = a.mask;
}

// average_grid_element (subdiv_ccg.c:876-900): a = (a + b) * 0.5, b = a.
inline void avg_pair3(float *a, float *b) {
  a[0] += b[0];
  a[1] += b[1];
  a[2] += b[2];
  a[0] *= 0.5f;
  a[1] *= 0.5f;
  a[2] *= 0.5f;
  b[0] = a[0];
  b[1] = a[1];
  b[2] = a[2];
}
void avg_pair(const SubdivCCG *s, const SubdivCCGCoord &ca, const SubdivCCGCoord &cb) {
  avg_pair3(ccg_co(s, ca.grid_index, ca.x, ca.y), ccg_co(s, cb.grid_index, cb.x, cb.y));
  if (s->has_normal) avg_pair3(ccg_no(s, ca.grid_index, ca.x, ca.y), ccg_no(s, cb.grid_index, cb.x, cb.y));
  if (s->has_mask) {
    float *ma = ccg_mask(s, ca.grid_index, ca.x, ca.y), *mb = ccg_mask(s, cb.grid_index, cb.x, cb.y);
    const float m = (*ma + *mb) * 0.5f;
    *ma = m;
    *mb = m;
  }
}

// subdiv_ccg_recalc_inner_face_normals + subdiv_ccg_average_inner_face_normals
// (subdiv_ccg.c:670-740): unit quad normals, then each element gets their mean over
// the up to four quads around it (not renormalised).
void grid_normals(SubdivCCG *s, int g, std::vector<float> &fn) {
  const int gs = s->grid_size, g1 = gs - 1;
  fn.resize((size_t)g1 * g1 * 3);
  for (int y = 0; y < g1; y++)
    for (int x = 0; x < g1; x++)
      quad_normal(&fn[3 * ((size_t)y * g1 + x)], ccg_co(s, g, x, y + 1), ccg_co(s, g, x + 1, y + 1),
                  ccg_co(s, g, x + 1, y), ccg_co(s, g, x, y));
  for (int y = 0; y < gs; y++) {
    for (int x = 0; x < gs; x++) {
      float acc[3] = {0.0f, 0.0f, 0.0f};
      int counter = 0;
      auto add = [&](int qx, int qy) {
        const float *q = &fn[3 * ((size_t)qy * g1 + qx)];
        acc[0] += q[0];
        acc[1] += q[1];
        acc[2] += q[2];
        counter++;
      };
      if (x < g1 && y < g1) add(x, y);
      if (x >= 1) {
        if (y < g1) add(x - 1, y);
        if (y >= 1) add(x - 1, y - 1);
      }
      if (y >= 1 && x < g1) add(x, y - 1);
      const float f = 1.0f / counter;
      float *no = ccg_no(s, g, x, y);
      no[0] = acc[0] * f;
      no[1] = acc[1] * f;
      no[2] = acc[2] * f;
    }
  }
}

// subdiv_ccg_average_inner_face_grids (subdiv_ccg.c:951-984): the seams between a
// face's grids, then the shared centre.
void average_inner_face(SubdivCCG *s, int f) {
  const SubdivCCGFace &face = s->faces[f];
  const int n = face.num_grids, gs = s->grid_size;
  int prev = face.start_grid_index + n - 1;
  for (int c = 0; c < n; c++) {
    const int g = face.start_grid_index + c;
    for (int i = 1; i < gs; i++) avg_pair(s, coord(prev, i, 0), coord(g, 0, i));
    prev = g;
  }
  Acc a;
  acc_init(a);
  for (int c = 0; c < n; c++) acc_add(s, a, coord(face.start_grid_index + c, 0, 0));
  acc_mul(a, 1.0f / n);
  for (int c = 0; c < n; c++) acc_copy(s, coord(face.start_grid_index + c, 0, 0), a);
}

// subdiv_ccg_average_grids_boundary (subdiv_ccg.c:1010-1048)
void average_edge(SubdivCCG *s, int e, std::vector<Acc> &acc) {
  const int nf = s->edge_face_off[e + 1] - s->edge_face_off[e];
  if (nf == 1) return;
  const int n2 = 2 * s->grid_size;
  acc.resize(n2);
  for (int i = 1; i < n2 - 1; i++) acc_init(acc[i]);
  const SubdivCCGCoord *bc = &s->edge_coords[(size_t)s->edge_face_off[e] * n2];
  for (int k = 0; k < nf; k++)
    for (int i = 1; i < n2 - 1; i++) acc_add(s, acc[i], bc[(size_t)k * n2 + i]);
  for (int i = 1; i < n2 - 1; i++) acc_mul(acc[i], 1.0f / nf);
  for (int k = 0; k < nf; k++)
    for (int i = 1; i < n2 - 1; i++) acc_copy(s, bc[(size_t)k * n2 + i], acc[i]);
}

// subdiv_ccg_average_grids_corners (subdiv_ccg.c:1081-1104)
void average_corner(SubdivCCG *s, int v) {
  const int b = s->vert_face_off[v], nf = s->vert_face_off[v + 1] - b;
  if (nf == 1) return;
  Acc a;
  acc_init(a);
  for (int k = 0; k < nf; k++) acc_add(s, a, s->vert_coords[b + k]);
  acc_mul(a, 1.0f / nf);
  for (int k = 0; k < nf; k++) acc_copy(s, s->vert_coords[b + k], a);
}

int grid_size_from_level(int level) { return (1 << (level - 1)) + 1; }  // subdiv_inline.h:26-29

}  // namespace

// Neighbours of one coordinate, the duplicates last when asked for (subdiv_ccg.c:
// 1365-1909).  Appends to out.
static void neighbors_of(const SubdivCCG *s, const SubdivCCGCoord &c, bool dup, std::vector<SubdivCCGCoord> &out,
                         int *r_num_dup);

extern "C" {

SubdivCCG *sg_subdiv_ccg_new(const float (*coarse_co)[3], int totvert, const int *poly_off, const int *poly_verts,
                             int totpoly, int level, int need_normal, int need_mask) {
  if (!coarse_co || !poly_off || !poly_verts || totvert <= 0 || totpoly <= 0) {
    sg_set_error("sg_subdiv_ccg_new: empty mesh");
    return nullptr;
  }
  if (level < 1 || level > 12) {
    sg_set_error("sg_subdiv_ccg_new: level %d outside [1, 12]", level);
    return nullptr;
  }
  for (int f = 0; f < totpoly; f++) {
    const int n = poly_off[f + 1] - poly_off[f];
    if (n < 3) {
      sg_set_error("sg_subdiv_ccg_new: polygon %d has %d corners", f, n);
      return nullptr;
    }
    for (int c = 0; c < n; c++) {
      const int v = poly_verts[poly_off[f] + c];
      if (v < 0 || v >= totvert) {
        sg_set_error("sg_subdiv_ccg_new: polygon %d corner %d: vertex %d out of range", f, c, v);
        return nullptr;
      }
    }
  }
  SubdivCCG *s = new SubdivCCG();
  s->level = level;
  s->grid_size = grid_size_from_level(level);
  s->grid_area = s->grid_size * s->grid_size;
  // layers (subdiv_ccg.c:62-90): co, then mask, then normal
  int off = 12;
  s->has_mask = need_mask != 0;
  s->mask_offset = s->has_mask ? off : -1;
  if (s->has_mask) off += 4;
  s->has_normal = need_normal != 0;
  s->normal_offset = s->has_normal ? off : -1;
  if (s->has_normal) off += 12;
  s->elem_size = off;
  // faces and grids (:104-140, :312-323)
  s->num_faces = totpoly;
  s->num_coarse_verts = totvert;
  s->num_grids = poly_off[totpoly] - poly_off[0];
  s->faces.resize(totpoly);
  for (int f = 0; f < totpoly; f++) {
    s->faces[f].num_grids = poly_off[f + 1] - poly_off[f];
    s->faces[f].start_grid_index = poly_off[f] - poly_off[0];
  }
  const size_t gbytes = (size_t)s->grid_area * s->elem_size;
  s->storage.assign((size_t)s->num_grids * gbytes, 0);
  s->grids.resize(s->num_grids);
  s->grid_faces.resize(s->num_grids);
  for (int g = 0; g < s->num_grids; g++) s->grids[g] = (CCGElem *)&s->storage[(size_t)g * gbytes];
  for (int f = 0; f < totpoly; f++)
    for (int c = 0; c < s->faces[f].num_grids; c++) s->grid_faces[s->faces[f].start_grid_index + c] = &s->faces[f];
  s->flag_mats.assign(s->num_grids, DMFlagMat{0, 0});
  const int words = (s->grid_area + 31) / 32;
  s->hidden_words.assign((size_t)s->num_grids * words, 0u);
  s->grid_hidden.resize(s->num_grids);
  for (int g = 0; g < s->num_grids; g++) s->grid_hidden[g] = &s->hidden_words[(size_t)g * words];
  // coarse topology: edges in order of first use, oriented as first used
  s->face_verts.resize(s->num_grids);
  s->face_edges.resize(s->num_grids);
  std::unordered_map<unsigned long long, int> edge_of;
  edge_of.reserve((size_t)s->num_grids);
  for (int f = 0; f < totpoly; f++) {
    const int n = s->faces[f].num_grids, g0 = s->faces[f].start_grid_index;
    for (int c = 0; c < n; c++) {
      const int a = poly_verts[poly_off[f] + c], b = poly_verts[poly_off[f] + (c + 1) % n];
      s->face_verts[g0 + c] = a;
      const unsigned long long key =
          ((unsigned long long)(unsigned)std::min(a, b) << 32) | (unsigned)std::max(a, b);
      auto it = edge_of.find(key);
      int e;
      if (it == edge_of.end()) {
        e = (int)s->edge_v0.size();
        edge_of.emplace(key, e);
        s->edge_v0.push_back(a);
        s->edge_v1.push_back(b);
      } else {
        e = it->second;
      }
      s->face_edges[g0 + c] = e;
    }
  }
  s->num_edges = (int)s->edge_v0.size();
  // vertex edges, ascending edge number
  s->vert_edge_off.assign(totvert + 1, 0);
  for (int e = 0; e < s->num_edges; e++) {
    s->vert_edge_off[s->edge_v0[e] + 1]++;
    s->vert_edge_off[s->edge_v1[e] + 1]++;
  }
  for (int v = 0; v < totvert; v++) s->vert_edge_off[v + 1] += s->vert_edge_off[v];
  s->vert_edges.resize(s->vert_edge_off[totvert]);
  {
    std::vector<int> fill(s->vert_edge_off.begin(), s->vert_edge_off.end() - 1);
    for (int e = 0; e < s->num_edges; e++) {
      s->vert_edges[fill[s->edge_v0[e]]++] = e;
      s->vert_edges[fill[s->edge_v1[e]]++] = e;
    }
  }
  // edge neighbourhood (:397-468): per edge, per adjacent face in face order, the
  // 2 * grid_size elements along it
  const int gs = s->grid_size, n2 = 2 * gs;
  s->edge_face_off.assign(s->num_edges + 1, 0);
  for (int g = 0; g < s->num_grids; g++) s->edge_face_off[s->face_edges[g] + 1]++;
  for (int e = 0; e < s->num_edges; e++) s->edge_face_off[e + 1] += s->edge_face_off[e];
  s->edge_coords.resize((size_t)s->edge_face_off[s->num_edges] * n2);
  {
    std::vector<int> fill(s->edge_face_off.begin(), s->edge_face_off.end() - 1);
    for (int f = 0; f < totpoly; f++) {
      const int n = s->faces[f].num_grids, g0 = s->faces[f].start_grid_index;
      for (int c = 0; c < n; c++) {
        const int v = s->face_verts[g0 + c], e = s->face_edges[g0 + c];
        const bool flipped = s->edge_v0[e] != v;
        const int cur = g0 + c, next = g0 + (c + 1) % n;
        SubdivCCGCoord *bc = &s->edge_coords[(size_t)(fill[e]++) * n2];
        int k = 0;
        if (flipped) {
          for (int i = 0; i < gs; i++) bc[k++] = coord(next, gs - i - 1, gs - 1);
          for (int i = 0; i < gs; i++) bc[k++] = coord(cur, gs - 1, i);
        } else {
          for (int i = 0; i < gs; i++) bc[k++] = coord(cur, gs - 1, gs - i - 1);
          for (int i = 0; i < gs; i++) bc[k++] = coord(next, i, gs - 1);
        }
      }
    }
  }
  // vertex neighbourhood (:492-530)
  s->vert_face_off.assign(totvert + 1, 0);
  for (int g = 0; g < s->num_grids; g++) s->vert_face_off[s->face_verts[g] + 1]++;
  for (int v = 0; v < totvert; v++) s->vert_face_off[v + 1] += s->vert_face_off[v];
  s->vert_coords.resize(s->vert_face_off[totvert]);
  {
    std::vector<int> fill(s->vert_face_off.begin(), s->vert_face_off.end() - 1);
    for (int g = 0; g < s->num_grids; g++) s->vert_coords[fill[s->face_verts[g]]++] = coord(g, gs - 1, gs - 1);
  }
  // elements: the bilinear patch of each corner quad
  const double inv = 1.0 / (gs - 1);
  for (int f = 0; f < totpoly; f++) {
    const int n = s->faces[f].num_grids, g0 = s->faces[f].start_grid_index;
    double cen[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < n; c++)
      for (int q = 0; q < 3; q++) cen[q] += coarse_co[s->face_verts[g0 + c]][q];
    for (int q = 0; q < 3; q++) cen[q] /= n;
    for (int c = 0; c < n; c++) {
      const float *vc = coarse_co[s->face_verts[g0 + c]];
      const float *vn = coarse_co[s->face_verts[g0 + (c + 1) % n]];
      const float *vp = coarse_co[s->face_verts[g0 + (c + n - 1) % n]];
      for (int y = 0; y < gs; y++) {
        const double b = y * inv;
        for (int x = 0; x < gs; x++) {
          const double a = x * inv;
          float *co = ccg_co(s, g0 + c, x, y);
          for (int q = 0; q < 3; q++) {
            const double mc = 0.5 * ((double)vc[q] + vn[q]), mp = 0.5 * ((double)vp[q] + vc[q]);
            co[q] = (float)((1 - a) * (1 - b) * cen[q] + a * (1 - b) * mc + a * b * vc[q] + (1 - a) * b * mp);
          }
        }
      }
    }
  }
  KERNEL_subdiv_ccg_average_grids(s);
  KERNEL_subdiv_ccg_recalc_normals(s);
  return s;
}

void KERNEL_subdiv_ccg_destroy(SubdivCCG *s) { delete s; }

void KERNEL_subdiv_ccg_key_top_level(CCGKey *key, const SubdivCCG *s) {  // subdiv_ccg.c:633-651
  if (!key || !s) return;
  key->level = s->level;
  key->elem_size = s->elem_size;
  key->grid_size = s->grid_size;
  key->grid_area = s->grid_area;
  key->grid_bytes = s->elem_size * s->grid_area;
  key->normal_offset = s->normal_offset;
  key->mask_offset = s->mask_offset;
  key->has_normals = s->has_normal;
  key->has_mask = s->has_mask;
}

void BKE_subdiv_ccg_key_top_level(CCGKey *key, const SubdivCCG *s) { KERNEL_subdiv_ccg_key_top_level(key, s); }

void KERNEL_subdiv_ccg_average_grids(SubdivCCG *s) {  // subdiv_ccg.c:1171-1189
  if (!s) return;
  for (int f = 0; f < s->num_faces; f++) average_inner_face(s, f);
  std::vector<Acc> acc;
  for (int e = 0; e < s->num_edges; e++) average_edge(s, e, acc);
  for (int v = 0; v < s->num_coarse_verts; v++) average_corner(s, v);
}

void KERNEL_subdiv_ccg_recalc_normals(SubdivCCG *s) {  // subdiv_ccg.c:781-795
  if (!s || !s->has_normal) return;
  std::vector<float> fn;
  for (int g = 0; g < s->num_grids; g++) grid_normals(s, g, fn);
  KERNEL_subdiv_ccg_average_grids(s);
}

void KERNEL_subdiv_ccg_update_normals(SubdivCCG *s, CCGFace **faces, int num) {  // subdiv_ccg.c:847-874
  if (!s || !s->has_normal || num == 0) return;
  std::vector<float> fn;
  std::vector<uint8_t> edge_seen(s->num_edges, 0), vert_seen(s->num_coarse_verts, 0);
  std::vector<int> edges, verts;
  for (int i = 0; i < num; i++) {
    const SubdivCCGFace *face = (const SubdivCCGFace *)faces[i];
    const int f = (int)(face - s->faces.data());
    for (int c = 0; c < face->num_grids; c++) grid_normals(s, face->start_grid_index + c, fn);
    average_inner_face(s, f);
  }
  // subdiv_ccg_affected_face_adjacency (:1191-1235): the faces' edges and vertices,
  // each once (their passes touch disjoint elements, so the order is free)
  for (int i = 0; i < num; i++) {
    const SubdivCCGFace *face = (const SubdivCCGFace *)faces[i];
    for (int c = 0; c < face->num_grids; c++) {
      const int g = face->start_grid_index + c;
      if (!edge_seen[s->face_edges[g]]) {
        edge_seen[s->face_edges[g]] = 1;
        edges.push_back(s->face_edges[g]);
      }
      if (!vert_seen[s->face_verts[g]]) {
        vert_seen[s->face_verts[g]] = 1;
        verts.push_back(s->face_verts[g]);
      }
    }
  }
  std::vector<Acc> acc;
  for (int e : edges) average_edge(s, e, acc);
  for (int v : verts) average_corner(s, v);
}

int KERNEL_subdiv_ccg_grid_to_face_index(const SubdivCCG *s, int grid_index) {  // subdiv_ccg.c:1911-1916
  if (!s || grid_index < 0 || grid_index >= s->num_grids) return -1;
  return (int)(s->grid_faces[grid_index] - s->faces.data());
}

void KERNEL_subdiv_ccg_neighbor_coords_get(const SubdivCCG *s, const SubdivCCGCoord *c, bool include_duplicates,
                                           SubdivCCGNeighbors *r) {
  if (!s || !c || !r) return;
  std::vector<SubdivCCGCoord> out;
  int ndup = 0;
  neighbors_of(s, *c, include_duplicates, out, &ndup);
  r->size = (int)out.size();
  r->num_duplicates = ndup;
  if (r->size < 256) {
    r->coords = r->coords_fixed;
  } else {
    r->coords = (SubdivCCGCoord *)malloc(sizeof(SubdivCCGCoord) * out.size());
  }
  if (!out.empty()) memcpy(r->coords, out.data(), sizeof(SubdivCCGCoord) * out.size());
}

void sg_subdiv_ccg_neighbors_free(SubdivCCGNeighbors *r) {
  if (r && r->coords && r->coords != r->coords_fixed) free(r->coords);
  if (r) r->coords = nullptr;
}

CCGElem **sg_subdiv_ccg_grids(SubdivCCG *s) { return s ? s->grids.data() : nullptr; }
int sg_subdiv_ccg_num_grids(const SubdivCCG *s) { return s ? s->num_grids : 0; }
void **sg_subdiv_ccg_grid_faces(SubdivCCG *s) { return s ? (void **)s->grid_faces.data() : nullptr; }
SubdivCCGFace *sg_subdiv_ccg_faces(SubdivCCG *s, int *r_num) {
  if (r_num) *r_num = s ? s->num_faces : 0;
  return s ? s->faces.data() : nullptr;
}
DMFlagMat *sg_subdiv_ccg_flag_mats(SubdivCCG *s) { return s ? s->flag_mats.data() : nullptr; }
BLI_bitmap **sg_subdiv_ccg_grid_hidden(SubdivCCG *s) { return s ? s->grid_hidden.data() : nullptr; }

static int copy_layer(const SubdivCCG *s, float *buf, int which, bool to_grids) {
  if (!s || !buf) return sg_set_error("null SubdivCCG or buffer");
  const int w = which == 2 ? 1 : 3;
  if (which == 1 && !s->has_normal) return sg_set_error("the grids have no normals");
  if (which == 2 && !s->has_mask) return sg_set_error("the grids have no mask");
  const int offb = which == 0 ? 0 : (which == 1 ? s->normal_offset : s->mask_offset);
  size_t i = 0;
  for (int g = 0; g < s->num_grids; g++) {
    const unsigned char *base = (const unsigned char *)s->grids[g];
    for (int e = 0; e < s->grid_area; e++, i++) {
      float *el = (float *)(base + (size_t)e * s->elem_size + offb);
      if (to_grids)
        memcpy(el, buf + i * w, sizeof(float) * w);
      else
        memcpy(buf + i * w, el, sizeof(float) * w);
    }
  }
  return 0;
}
int sg_subdiv_ccg_get_co(const SubdivCCG *s, float *out) { return copy_layer(s, out, 0, false); }
int sg_subdiv_ccg_set_co(SubdivCCG *s, const float *in) { return copy_layer(s, (float *)in, 0, true); }
int sg_subdiv_ccg_get_normals(const SubdivCCG *s, float *out) { return copy_layer(s, out, 1, false); }
int sg_subdiv_ccg_get_mask(const SubdivCCG *s, float *out) { return copy_layer(s, out, 2, false); }
int sg_subdiv_ccg_set_mask(SubdivCCG *s, const float *in) { return copy_layer(s, (float *)in, 2, true); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Neighbour coordinates (subdiv_ccg.c:1365-1909).
// ---------------------------------------------------------------------------
namespace {

inline bool is_corner(const SubdivCCG *s, const SubdivCCGCoord &c) {  // :1382-1389
  const int g1 = s->grid_size - 1;
  return (c.x == 0 && c.y == 0) || (c.x == 0 && c.y == g1) || (c.x == g1 && c.y == g1) || (c.x == g1 && c.y == 0);
}
inline bool is_boundary(const SubdivCCG *s, const SubdivCCGCoord &c) {  // :1391-1396
  const int g1 = s->grid_size - 1;
  return c.x == 0 || c.y == 0 || c.x == g1 || c.y == g1;
}
inline bool is_inner_edge(const SubdivCCG *s, const SubdivCCGCoord &c) {  // :1398-1409
  const int g1 = s->grid_size - 1;
  if (c.x == 0) return c.y > 0 && c.y < g1;
  if (c.y == 0) return c.x > 0 && c.x < g1;
  return false;
}
inline SubdivCCGCoord step_inside(const SubdivCCG *s, SubdivCCGCoord c) {  // :1448-1471
  const int g1 = s->grid_size - 1;
  if (c.x == g1)
    c.x--;
  else if (c.y == g1)
    c.y--;
  else if (c.x == 0)
    c.x++;
  else if (c.y == 0)
    c.y++;
  return c;
}
inline int next_grid(const SubdivCCG *s, int g) {  // :1473-1481
  const SubdivCCGFace *f = s->grid_faces[g];
  int k = g + 1 - f->start_grid_index;
  if (k == f->num_grids) k = 0;
  return f->start_grid_index + k;
}
inline int prev_grid(const SubdivCCG *s, int g) {  // :1483-1494
  const SubdivCCGFace *f = s->grid_faces[g];
  int k = g - 1 - f->start_grid_index;
  if (k < 0) k = f->num_grids - 1;
  return f->start_grid_index + k;
}

// the corner at the face centre (:1496-1521)
void corner_center(const SubdivCCG *s, const SubdivCCGCoord &c, bool dup, std::vector<SubdivCCGCoord> &out,
                   int *ndup) {
  const SubdivCCGFace *f = s->grid_faces[c.grid_index];
  const int n = f->num_grids;
  std::vector<SubdivCCGCoord> d;
  for (int k = 0; k < n; k++) {
    const int g = f->start_grid_index + k;
    out.push_back(coord(g, 1, 0));
    if (dup && g != c.grid_index) d.push_back(coord(g, 0, 0));
  }
  out.insert(out.end(), d.begin(), d.end());
  *ndup = (int)d.size();
}

// the corner at a coarse vertex (:1523-1610)
void corner_vertex(const SubdivCCG *s, const SubdivCCGCoord &c, bool dup, std::vector<SubdivCCGCoord> &out,
                   int *ndup) {
  const int v = s->face_verts[c.grid_index];
  const int n2 = 2 * s->grid_size;
  for (int q = s->vert_edge_off[v]; q < s->vert_edge_off[v + 1]; q++) {
    const int e = s->vert_edges[q];
    int point;
    if (s->edge_v0[e] == v) {
      point = 1;
    } else {
      point = n2 - 2;
    }
    out.push_back(s->edge_coords[(size_t)s->edge_face_off[e] * n2 + point]);  // the edge's first face
  }
  int nd = 0;
  if (dup) {
    for (int q = s->vert_face_off[v]; q < s->vert_face_off[v + 1]; q++) {
      if (s->vert_coords[q].grid_index != c.grid_index) {
        out.push_back(s->vert_coords[q]);
        nd++;
      }
    }
  }
  *ndup = nd;
}

int edge_of_coord(const SubdivCCG *s, const SubdivCCGCoord &c) {  // :1612-1641
  const SubdivCCGFace *f = s->grid_faces[c.grid_index];
  const int k = c.grid_index - f->start_grid_index;
  if (c.x == s->grid_size - 1) return s->face_edges[c.grid_index];
  return s->face_edges[f->start_grid_index + (k == 0 ? f->num_grids - 1 : k - 1)];
}

int edge_point_of_coord(const SubdivCCG *s, const SubdivCCGCoord &c, int e) {  // :1643-1683
  const int v = s->face_verts[c.grid_index];
  const int g1 = s->grid_size - 1;
  int point, dir_v;
  if (c.x == g1) {
    point = s->grid_size - c.y - 1;
    dir_v = s->edge_v0[e];
  } else {
    point = s->grid_size + c.x;
    dir_v = s->edge_v1[e];
  }
  if (v != dir_v) point = 2 * s->grid_size - point - 1;
  return point;
}

// a coordinate on a coarse edge, not at a coarse vertex (:1714-1774)
void edge_nbrs(const SubdivCCG *s, const SubdivCCGCoord &c, bool dup, std::vector<SubdivCCGCoord> &out, int *ndup) {
  const bool corner = is_corner(s, c);
  const int e = edge_of_coord(s, c);
  const int n2 = 2 * s->grid_size, gs = s->grid_size;
  const int nf = s->edge_face_off[e + 1] - s->edge_face_off[e];
  const int point = edge_point_of_coord(s, c, e);
  const int point_dup = (point == gs) ? point - 1 : point + 1;   // :1703-1712
  const int next_point = (point == gs - 1) ? point + 2 : point + 1;  // :1685-1690
  const int prev_point = (point == gs) ? point - 2 : point - 1;      // :1692-1697
  const size_t base = out.size();
  out.resize(base + 2 + nf);
  std::vector<SubdivCCGCoord> d;
  for (int k = 0; k < nf; k++) {
    const SubdivCCGCoord *bc = &s->edge_coords[(size_t)(s->edge_face_off[e] + k) * n2];
    const SubdivCCGCoord gc = bc[point];
    out[base + 2 + k] = step_inside(s, gc);
    if (gc.grid_index == c.grid_index) {
      out[base + 0] = bc[prev_point];
      out[base + 1] = bc[next_point];
    } else if (dup) {
      d.push_back(gc);
    }
    if (dup && corner) d.push_back(bc[point_dup]);
  }
  out.insert(out.end(), d.begin(), d.end());
  *ndup = (int)d.size();
}

// a coordinate on a seam between two grids of one face (:1807-1843)
void boundary_inner(const SubdivCCG *s, const SubdivCCGCoord &c, bool dup, std::vector<SubdivCCGCoord> &out,
                    int *ndup) {
  SubdivCCGCoord n3;
  if (c.x == 0) {
    out.push_back(coord(c.grid_index, c.x, c.y - 1));
    out.push_back(coord(c.grid_index, c.x, c.y + 1));
    out.push_back(coord(c.grid_index, c.x + 1, c.y));
    n3 = coord(prev_grid(s, c.grid_index), c.y, 1);
    out.push_back(n3);
    if (dup) out.push_back(coord(n3.grid_index, n3.x, 0));
  } else {
    out.push_back(coord(c.grid_index, c.x - 1, c.y));
    out.push_back(coord(c.grid_index, c.x + 1, c.y));
    out.push_back(coord(c.grid_index, c.x, c.y + 1));
    n3 = coord(next_grid(s, c.grid_index), 1, c.x);
    out.push_back(n3);
    if (dup) out.push_back(coord(n3.grid_index, 0, n3.y));
  }
  *ndup = dup ? 1 : 0;
}

}  // namespace

static void neighbors_of(const SubdivCCG *s, const SubdivCCGCoord &c, bool dup, std::vector<SubdivCCGCoord> &out,
                         int *ndup) {
  *ndup = 0;
  const int g1 = s->grid_size - 1;
  if (is_corner(s, c)) {  // :1785-1805
    if (c.x == 0 && c.y == 0)
      corner_center(s, c, dup, out, ndup);
    else if (c.x == g1 && c.y == g1)
      corner_vertex(s, c, dup, out, ndup);
    else
      edge_nbrs(s, c, dup, out, ndup);
  } else if (is_boundary(s, c)) {  // :1856-1868
    if (is_inner_edge(s, c))
      boundary_inner(s, c, dup, out, ndup);
    else
      edge_nbrs(s, c, dup, out, ndup);
  } else {  // :1870-1880
    out.push_back(coord(c.grid_index, c.x, c.y - 1));
    out.push_back(coord(c.grid_index, c.x, c.y + 1));
    out.push_back(coord(c.grid_index, c.x - 1, c.y));
    out.push_back(coord(c.grid_index, c.x + 1, c.y));
  }
}

void sg_ccg_neighbor_csr(const SubdivCCG *s, std::vector<long long> &off, std::vector<int> &idx) {
  const long long E = (long long)s->num_grids * s->grid_area;
  off.assign((size_t)E + 1, 0);
  idx.clear();
  idx.reserve((size_t)E * 4 + 64);
  std::vector<SubdivCCGCoord> out;
  long long e = 0;
  for (int g = 0; g < s->num_grids; g++) {
    for (int y = 0; y < s->grid_size; y++) {
      for (int x = 0; x < s->grid_size; x++, e++) {
        out.clear();
        int nd = 0;
        neighbors_of(s, coord(g, x, y), false, out, &nd);
        for (const SubdivCCGCoord &c : out)
          idx.push_back(c.grid_index * s->grid_area + c.y * s->grid_size + c.x);
        off[e + 1] = (long long)idx.size();
      }
    }
  }
}
