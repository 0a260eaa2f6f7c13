/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).  Never part of the
 * product: only tests/, smoke() and bench.py's cpu_baseline leg load it.
 *
 * Built with -O2 -ffp-contract=off so every FP32 expression rounds exactly as
 * written (the reference build is SSE2-only and never fuses, CMakeLists.txt:703-768).
 */
#include "oracle.h"
#include "oracle_int.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif


/* ------------------------------------------------------------------------- */
/* FP32 primitives (restated from the reference's math library).             */
/* ------------------------------------------------------------------------- */

/* min_ff / max_ff: math_base_inline.c:402-409 -- ties return the second operand. */
static inline float o_min(float a, float b) { return (a < b) ? a : b; }
static inline float o_max(float a, float b) { return (a > b) ? a : b; }

/* dot_v3v3: math_vector_inline.c:895-898, summed left to right. */
static inline float o_dot(const float a[3], const float b[3])
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* normalize_v3: math_vector_inline.c:1165-1180 via :1242-1245.  Zero below 1e-35,
 * otherwise multiply by the rounded reciprocal of the correctly rounded sqrt. */
static inline float o_normalize(float n[3])
{
  float d = o_dot(n, n);
  if (d > 1.0e-35f) {
    d = sqrtf(d);
    const float inv = 1.0f / d;
    n[0] = n[0] * inv;
    n[1] = n[1] * inv;
    n[2] = n[2] * inv;
  }
  else {
    n[0] = n[1] = n[2] = 0.0f;
    d = 0.0f;
  }
  return d;
}

static inline double o_normalize64(double n[3])
{
  double d = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  if (d > 1.0e-35) {
    d = sqrt(d);
    n[0] /= d;
    n[1] /= d;
    n[2] /= d;
  }
  else {
    n[0] = n[1] = n[2] = 0.0;
    d = 0.0;
  }
  return d;
}

/* normal_tri_v3: math_geom.cc:31-46. */
void orc_tri_normal(const float v1[3], const float v2[3], const float v3[3], float n[3])
{
  float n1[3], n2[3];
  n1[0] = v1[0] - v2[0];
  n2[0] = v2[0] - v3[0];
  n1[1] = v1[1] - v2[1];
  n2[1] = v2[1] - v3[1];
  n1[2] = v1[2] - v2[2];
  n2[2] = v2[2] - v3[2];
  n[0] = n1[1] * n2[2] - n1[2] * n2[1];
  n[1] = n1[2] * n2[0] - n1[0] * n2[2];
  n[2] = n1[0] * n2[1] - n1[1] * n2[0];
  o_normalize(n);
}

static void tri_normal64(const double v1[3], const double v2[3], const double v3[3], double n[3])
{
  double n1[3], n2[3];
  for (int i = 0; i < 3; i++) {
    n1[i] = v1[i] - v2[i];
    n2[i] = v2[i] - v3[i];
  }
  n[0] = n1[1] * n2[2] - n1[2] * n2[1];
  n[1] = n1[2] * n2[0] - n1[0] * n2[2];
  n[2] = n1[0] * n2[1] - n1[1] * n2[0];
  o_normalize64(n);
}

/* ------------------------------------------------------------------------- */
/* Bounding boxes: pbvh.c:55-105.                                            */
/* ------------------------------------------------------------------------- */


static inline void bb_reset(OBB *bb)
{
  bb->bmin[0] = bb->bmin[1] = bb->bmin[2] = FLT_MAX;
  bb->bmax[0] = bb->bmax[1] = bb->bmax[2] = -FLT_MAX;
}
static inline void bb_expand(OBB *bb, const float co[3])
{
  for (int i = 0; i < 3; i++) {
    bb->bmin[i] = o_min(bb->bmin[i], co[i]);
    bb->bmax[i] = o_max(bb->bmax[i], co[i]);
  }
}
static inline void bb_expand_bb(OBB *bb, const OBB *b2)
{
  for (int i = 0; i < 3; i++) {
    bb->bmin[i] = o_min(bb->bmin[i], b2->bmin[i]);
    bb->bmax[i] = o_max(bb->bmax[i], b2->bmax[i]);
  }
}
static int bb_widest_axis(const OBB *bb)
{
  float dim[3];
  for (int i = 0; i < 3; i++) {
    dim[i] = bb->bmax[i] - bb->bmin[i];
  }
  if (dim[0] > dim[1]) {
    return (dim[0] > dim[2]) ? 0 : 2;
  }
  return (dim[1] > dim[2]) ? 1 : 2;
}

/* ------------------------------------------------------------------------- */
/* Tree.                                                                     */
/* ------------------------------------------------------------------------- */


static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void)
{
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

static void grow_nodes(OrcPBVH *o, int totnode)
{
  /* pbvh_grow_nodes: pbvh.c:214-226 (growth policy irrelevant to results). */
  if (totnode > o->node_mem) {
    int nm = o->node_mem + o->node_mem / 3;
    if (nm < totnode) {
      nm = totnode;
    }
    o->nodes = (ONode *)realloc(o->nodes, sizeof(ONode) * (size_t)nm);
    memset(o->nodes + o->node_mem, 0, sizeof(ONode) * (size_t)(nm - o->node_mem));
    o->node_mem = nm;
  }
  o->totnode = totnode;
}

/* partition_indices: pbvh.c:152-170 (Hoare, first element of the right side). */
static int partition_indices(int *prim, int lo, int hi, int axis, float mid, const OBBC *bbc)
{
  int i = lo, j = hi;
  for (;;) {
    for (; bbc[prim[i]].bcentroid[axis] < mid; i++) {
    }
    for (; mid < bbc[prim[j]].bcentroid[axis]; j--) {
    }
    if (!(i < j)) {
      return i;
    }
    int t = prim[i];
    prim[i] = prim[j];
    prim[j] = t;
    i++;
  }
}

static inline int mat_match(const OrcPBVH *o, int t1, int t2)
{
  if (!o->tri_mat) {
    return 1;
  }
  return o->tri_mat[t1] == o->tri_mat[t2];
}

/* partition_indices_material: pbvh.c:172-212 (mesh branch). */
static int partition_indices_material(OrcPBVH *o, int lo, int hi)
{
  const int *ind = o->prim_indices;
  const int first = ind[lo];
  int i = lo, j = hi;
  for (;;) {
    for (; mat_match(o, first, ind[i]); i++) {
    }
    for (; !mat_match(o, first, ind[j]); j--) {
    }
    if (!(i < j)) {
      return i;
    }
    int t = o->prim_indices[i];
    o->prim_indices[i] = o->prim_indices[j];
    o->prim_indices[j] = t;
    i++;
  }
}

/* leaf_needs_material_split: pbvh.c:409-440 (mesh branch). */
static int leaf_needs_material_split(const OrcPBVH *o, int offset, int count)
{
  if (count <= 1 || !o->tri_mat) {
    return 0;
  }
  const int first = o->prim_indices[offset];
  for (int i = offset + count - 1; i > offset; i--) {
    if (!mat_match(o, first, o->prim_indices[i])) {
      return 1;
    }
  }
  return 0;
}

/* update_vb: pbvh.c:322-329 -- reverse prim order. */
static void update_vb(OrcPBVH *o, ONode *node, const OBBC *bbc, int offset, int count)
{
  bb_reset(&node->vb);
  for (int i = offset + count - 1; i >= offset; i--) {
    bb_expand_bb(&node->vb, (const OBB *)&bbc[o->prim_indices[i]]);
  }
  node->orig_vb = node->vb;
}

/* map_insert_vert: pbvh.c:230-254.  Values >= 0 are uniq slots, < 0 shared
 * (~index).  The reference's GHash stores the insertion counter as the value and
 * vert_indices is filled by value, so the list order is first-insertion order. */
static int map_insert_vert(OrcPBVH *o, int *ntouched, int *face_verts, int *uniq_verts, int v)
{
  int val = o->map_local[v];
  if (val != INT32_MIN) {
    return val;
  }
  if (!o->vert_bitmap[v]) {
    o->vert_bitmap[v] = 1;
    val = (*uniq_verts)++;
  }
  else {
    val = ~(*face_verts);
    (*face_verts)++;
  }
  o->map_local[v] = val;
  o->map_touched[(*ntouched)++] = v;
  return val;
}

/* build_mesh_leaf_node: pbvh.c:256-320. */
static void build_mesh_leaf_node(OrcPBVH *o, ONode *node)
{
  const int totface = node->totprim;
  int ntouched = 0;
  node->uniq_verts = node->face_verts = 0;
  node->face_vert_indices = (int(*)[3])malloc(sizeof(int[3]) * (size_t)(totface > 0 ? totface : 1));
  int has_visible = !o->respect_hide;

  for (int i = 0; i < totface; i++) {
    const int t = o->prim_indices[node->prim_off + i];
    int hidden = 0;
    for (int j = 0; j < 3; j++) {
      const int v = o->tris[3 * t + j];
      node->face_vert_indices[i][j] = map_insert_vert(o, &ntouched, &node->face_verts,
                                                      &node->uniq_verts, v);
      if (o->vhide && o->vhide[v]) {
        hidden = 1; /* paint_is_face_hidden: any hidden corner hides the face */
      }
    }
    if (!has_visible && !hidden) {
      has_visible = 1;
    }
  }

  const int tot = node->uniq_verts + node->face_verts;
  node->vert_indices = (int *)calloc((size_t)(tot > 0 ? tot : 1), sizeof(int));
  for (int k = 0; k < ntouched; k++) {
    const int v = o->map_touched[k];
    int ndx = o->map_local[v];
    if (ndx < 0) {
      ndx = -ndx + node->uniq_verts - 1;
    }
    node->vert_indices[ndx] = v;
    o->map_local[v] = INT32_MIN;
  }
  for (int i = 0; i < totface; i++) {
    for (int j = 0; j < 3; j++) {
      if (node->face_vert_indices[i][j] < 0) {
        node->face_vert_indices[i][j] = -node->face_vert_indices[i][j] + node->uniq_verts - 1;
      }
    }
  }
  /* BKE_pbvh_node_mark_rebuild_draw + fully hidden (pbvh.c:316-318) */
  node->flag |= ORC_UpdateDrawBuffers | ORC_UpdateRedraw;
  if (!has_visible) {
    node->flag |= ORC_FullyHidden;
  }
  else {
    node->flag &= ~ORC_FullyHidden;
  }
}

/* build_leaf: pbvh.c:391-407. */
static void build_leaf(OrcPBVH *o, int node_index, const OBBC *bbc, int offset, int count)
{
  ONode *n = &o->nodes[node_index];
  n->flag |= ORC_Leaf;
  n->prim_off = offset;
  n->totprim = count;
  update_vb(o, n, bbc, offset, count);
  build_mesh_leaf_node(o, n);
}

/* build_sub: pbvh.c:454-507. */
static void build_sub(OrcPBVH *o, int node_index, OBB *cb, const OBBC *bbc, int offset, int count)
{
  OBB cb_backing;
  const int below = count <= o->leaf_limit;
  if (below) {
    if (!leaf_needs_material_split(o, offset, count)) {
      build_leaf(o, node_index, bbc, offset, count);
      return;
    }
  }
  o->nodes[node_index].children_offset = o->totnode;
  grow_nodes(o, o->totnode + 2);
  update_vb(o, &o->nodes[node_index], bbc, offset, count);

  int end;
  if (!below) {
    if (!cb) {
      cb = &cb_backing;
      bb_reset(cb);
      for (int i = offset + count - 1; i >= offset; i--) {
        bb_expand(cb, bbc[o->prim_indices[i]].bcentroid);
      }
    }
    const int axis = bb_widest_axis(cb);
    end = partition_indices(o->prim_indices, offset, offset + count - 1, axis,
                            (cb->bmax[axis] + cb->bmin[axis]) * 0.5f, bbc);
  }
  else {
    end = partition_indices_material(o, offset, offset + count - 1);
  }
  const int c = o->nodes[node_index].children_offset;
  build_sub(o, c, NULL, bbc, offset, end - offset);
  build_sub(o, c + 1, NULL, bbc, end, offset + count - end);
}

static void build_csr(OrcPBVH *o)
{
  const int V = o->totvert, T = o->tottri;
  o->vt_off = (int *)calloc((size_t)V + 1, sizeof(int));
  for (int t = 0; t < T; t++) {
    for (int j = 0; j < 3; j++) {
      o->vt_off[o->tris[3 * t + j] + 1]++;
    }
  }
  for (int v = 0; v < V; v++) {
    o->vt_off[v + 1] += o->vt_off[v];
  }
  o->vt_idx = (int *)malloc(sizeof(int) * (size_t)(3 * (size_t)T + 1));
  int *fill = (int *)malloc(sizeof(int) * (size_t)(V + 1));
  memcpy(fill, o->vt_off, sizeof(int) * (size_t)V);
  for (int t = 0; t < T; t++) {
    for (int j = 0; j < 3; j++) {
      const int v = o->tris[3 * t + j];
      /* a degenerate tri listing v twice is counted once per corner, as the
       * reference's per-corner accumulation does */
      o->vt_idx[fill[v]++] = t;
    }
  }
  /* vertex -> vertex: distinct edge neighbours, ascending */
  o->vv_off = (int *)calloc((size_t)V + 1, sizeof(int));
  int *tmp = (int *)malloc(sizeof(int) * 64);
  int tmpcap = 64;
  int *cnt = (int *)calloc((size_t)V, sizeof(int));
  /* two passes: count distinct, then fill */
  for (int pass = 0; pass < 2; pass++) {
    for (int v = 0; v < V; v++) {
      int n = 0;
      for (int k = o->vt_off[v]; k < o->vt_off[v + 1]; k++) {
        const int t = o->vt_idx[k];
        for (int j = 0; j < 3; j++) {
          const int u = o->tris[3 * t + j];
          if (u == v) {
            continue;
          }
          int dup = 0;
          for (int q = 0; q < n; q++) {
            if (tmp[q] == u) {
              dup = 1;
              break;
            }
          }
          if (!dup) {
            if (n == tmpcap) {
              tmpcap *= 2;
              tmp = (int *)realloc(tmp, sizeof(int) * (size_t)tmpcap);
            }
            tmp[n++] = u;
          }
        }
      }
      if (pass == 0) {
        cnt[v] = n;
      }
      else {
        /* insertion sort ascending */
        for (int a = 1; a < n; a++) {
          int x = tmp[a], b = a - 1;
          while (b >= 0 && tmp[b] > x) {
            tmp[b + 1] = tmp[b];
            b--;
          }
          tmp[b + 1] = x;
        }
        memcpy(o->vv_idx + o->vv_off[v], tmp, sizeof(int) * (size_t)n);
      }
    }
    if (pass == 0) {
      for (int v = 0; v < V; v++) {
        o->vv_off[v + 1] = o->vv_off[v] + cnt[v];
      }
      o->vv_idx = (int *)malloc(sizeof(int) * (size_t)(o->vv_off[V] + 1));
    }
  }
  free(cnt);
  free(tmp);
  free(fill);
}

/* BKE_mesh_calc_poly_normal (mesh_evaluate.c:62-84): more than 4 corners ->
 * mesh_calc_ngon_normal (:39-60, Newell's method from the last corner,
 * add_newell_cross_v3_v3v3 math_vector_inline.c:961-966, zero -> +z); 3 ->
 * normal_tri_v3; 4 -> normal_quad_v3 (math_geom.cc:48-68, cross of the diagonals);
 * fewer -> +z.  `c` holds the polygon's corner positions in loop order. */
static void poly_normal(const float *co, const int *corners, int n, float r[3])
{
  if (n > 4) {
    r[0] = r[1] = r[2] = 0.0f;
    const float *v_prev = co + 3 * corners[n - 1];
    for (int i = 0; i < n; i++) {
      const float *v_curr = co + 3 * corners[i];
      r[0] += (v_prev[1] - v_curr[1]) * (v_prev[2] + v_curr[2]);
      r[1] += (v_prev[2] - v_curr[2]) * (v_prev[0] + v_curr[0]);
      r[2] += (v_prev[0] - v_curr[0]) * (v_prev[1] + v_curr[1]);
      v_prev = v_curr;
    }
    if (o_normalize(r) == 0.0f) {
      r[2] = 1.0f;
    }
  }
  else if (n == 3) {
    orc_tri_normal(co + 3 * corners[0], co + 3 * corners[1], co + 3 * corners[2], r);
  }
  else if (n == 4) {
    const float *v1 = co + 3 * corners[0], *v2 = co + 3 * corners[1], *v3 = co + 3 * corners[2],
                *v4 = co + 3 * corners[3];
    float n1[3], n2[3];
    n1[0] = v1[0] - v3[0];
    n1[1] = v1[1] - v3[1];
    n1[2] = v1[2] - v3[2];
    n2[0] = v2[0] - v4[0];
    n2[1] = v2[1] - v4[1];
    n2[2] = v2[2] - v4[2];
    r[0] = n1[1] * n2[2] - n1[2] * n2[1];
    r[1] = n1[2] * n2[0] - n1[0] * n2[2];
    r[2] = n1[0] * n2[1] - n1[1] * n2[0];
    o_normalize(r);
  }
  else {
    r[0] = 0.0f;
    r[1] = 0.0f;
    r[2] = 1.0f;
  }
}

/* The same in FP64 (the replay). */
static void poly_normal64(const double *co, const int *corners, int n, double r[3])
{
  if (n > 4) {
    r[0] = r[1] = r[2] = 0.0;
    const double *v_prev = co + 3 * corners[n - 1];
    for (int i = 0; i < n; i++) {
      const double *v_curr = co + 3 * corners[i];
      r[0] += (v_prev[1] - v_curr[1]) * (v_prev[2] + v_curr[2]);
      r[1] += (v_prev[2] - v_curr[2]) * (v_prev[0] + v_curr[0]);
      r[2] += (v_prev[0] - v_curr[0]) * (v_prev[1] + v_curr[1]);
      v_prev = v_curr;
    }
    if (o_normalize64(r) == 0.0) {
      r[2] = 1.0;
    }
  }
  else if (n == 3) {
    tri_normal64(co + 3 * corners[0], co + 3 * corners[1], co + 3 * corners[2], r);
  }
  else if (n == 4) {
    const double *v1 = co + 3 * corners[0], *v2 = co + 3 * corners[1], *v3 = co + 3 * corners[2],
                 *v4 = co + 3 * corners[3];
    const double n1[3] = {v1[0] - v3[0], v1[1] - v3[1], v1[2] - v3[2]};
    const double n2[3] = {v2[0] - v4[0], v2[1] - v4[1], v2[2] - v4[2]};
    r[0] = n1[1] * n2[2] - n1[2] * n2[1];
    r[1] = n1[2] * n2[0] - n1[0] * n2[2];
    r[2] = n1[0] * n2[1] - n1[1] * n2[0];
    o_normalize64(r);
  }
  else {
    r[0] = 0.0;
    r[1] = 0.0;
    r[2] = 1.0;
  }
}

void orc_poly_normal(const float *co, const int *corners, int n, float r[3]) { poly_normal(co, corners, n, r); }

/* saacos (absent upstream; clamped acos as its name and use say). */
static inline float o_saacos(float f)
{
  return (f <= -1.0f) ? (float)M_PI : ((f >= 1.0f) ? 0.0f : acosf(f));
}

/* KERNEL_mesh_calc_normals_poly_and_vertex: mesh_normals.c:225-299 (accumulation
 * :226-285, finalize :287-299).  Polygons are given by their corner offsets and corner
 * vertex ids (poly_off == NULL: tris, each a 3-corner polygon).  Accumulated serially
 * in polygon order (the reference adds with a spin-locked vector add in task order,
 * mesh_normals.c:54-81). */
void orc_calc_initial_normals_poly(const float *co, int totvert, const int *poly_off, const int *poly_verts,
                                   int npoly, float *out)
{
  memset(out, 0, sizeof(float) * 3 * (size_t)totvert);
  for (int p = 0; p < npoly; p++) {
    const int *ml = poly_verts + (poly_off ? poly_off[p] : 3 * p);
    const int i_end = (poly_off ? poly_off[p + 1] - poly_off[p] : 3) - 1;
    float pnor[3] = {0.0f, 0.0f, 0.0f};
    {
      const float *v_curr = co + 3 * ml[i_end];
      for (int i_next = 0; i_next <= i_end; i_next++) {
        const float *v_next = co + 3 * ml[i_next];
        /* add_newell_cross_v3_v3v3: math_vector_inline.c:961-966 */
        pnor[0] += (v_curr[1] - v_next[1]) * (v_curr[2] + v_next[2]);
        pnor[1] += (v_curr[2] - v_next[2]) * (v_curr[0] + v_next[0]);
        pnor[2] += (v_curr[0] - v_next[0]) * (v_curr[1] + v_next[1]);
        v_curr = v_next;
      }
      if (o_normalize(pnor) == 0.0f) {
        pnor[2] = 1.0f;
      }
    }
    {
      float ep[3], en[3], ee[3];
      const float *v_curr = co + 3 * ml[i_end];
      const float *vp = co + 3 * ml[i_end - 1];
      for (int k = 0; k < 3; k++) {
        ep[k] = vp[k] - v_curr[k];
      }
      o_normalize(ep);
      memcpy(ee, ep, sizeof(ee));
      for (int i_next = 0, i_curr = i_end; i_next <= i_end; i_curr = i_next++) {
        const float *v_next = co + 3 * ml[i_next];
        if (i_next != i_end) {
          for (int k = 0; k < 3; k++) {
            en[k] = v_curr[k] - v_next[k];
          }
          o_normalize(en);
        }
        else {
          memcpy(en, ee, sizeof(en));
        }
        const float fac = o_saacos(-o_dot(ep, en));
        float *vn = out + 3 * ml[i_curr];
        vn[0] += pnor[0] * fac;
        vn[1] += pnor[1] * fac;
        vn[2] += pnor[2] * fac;
        v_curr = v_next;
        memcpy(ep, en, sizeof(ep));
      }
    }
  }
  for (int v = 0; v < totvert; v++) {
    float *no = out + 3 * v;
    if (o_normalize(no) == 0.0f) {
      no[0] = co[3 * v + 0];
      no[1] = co[3 * v + 1];
      no[2] = co[3 * v + 2];
      o_normalize(no);
    }
  }
}

/* The triangle-mesh form: every tri its own 3-corner polygon. */
void orc_calc_initial_normals(const float *co, int totvert, const int *tris, int tottri, float *out)
{
  orc_calc_initial_normals_poly(co, totvert, NULL, tris, tottri, out);
}

/* The mesh's polygons (MPoly/MLoop): looptri -> polygon, polygon corner offsets and
 * corner vertex ids.  Normals then follow the reference's polygon rules: the initial
 * normals weight each polygon's Newell normal by its corner angles, and the update
 * after a dab adds BKE_mesh_calc_poly_normal of each incident looptri's polygon
 * (pbvh.c:1041-1058).  The initial normals are recomputed here, and Smooth's 1-ring
 * becomes the polygon-edge neighbours. */
void orc_set_polys(OrcPBVH *o, const int *tri_poly, const int *poly_off, const int *poly_verts, int npoly)
{
  free(o->tri_poly);
  free(o->poly_off);
  free(o->poly_verts);
  o->npoly = npoly;
  o->tri_poly = (int *)malloc(sizeof(int) * (size_t)(o->tottri > 0 ? o->tottri : 1));
  memcpy(o->tri_poly, tri_poly, sizeof(int) * (size_t)o->tottri);
  o->poly_off = (int *)malloc(sizeof(int) * (size_t)(npoly + 1));
  memcpy(o->poly_off, poly_off, sizeof(int) * (size_t)(npoly + 1));
  o->poly_verts = (int *)malloc(sizeof(int) * (size_t)(poly_off[npoly] > 0 ? poly_off[npoly] : 1));
  memcpy(o->poly_verts, poly_verts, sizeof(int) * (size_t)poly_off[npoly]);
  orc_calc_initial_normals_poly(o->co, o->totvert, o->poly_off, o->poly_verts, npoly, o->no);
  /* vertex -> vertex from the polygon edges (BKE_mesh_vert_edge_map_create,
   * mesh_mapping.c:297: the edges of a polygon join consecutive corners; a fan
   * triangulation's diagonals are not edges): distinct, ascending */
  {
    const int V = o->totvert;
    int *cnt = (int *)calloc((size_t)V + 1, sizeof(int));
    for (int q = 0; q < npoly; q++) {
      for (int k = poly_off[q]; k < poly_off[q + 1]; k++) {
        cnt[poly_verts[k] + 1]++;
      }
    }
    for (int v = 0; v < V; v++) {
      cnt[v + 1] += cnt[v];
    }
    int *fill = (int *)malloc(sizeof(int) * (size_t)(V + 1));
    memcpy(fill, cnt, sizeof(int) * (size_t)(V + 1));
    int *nb = (int *)malloc(sizeof(int) * (size_t)(2 * cnt[V] + 1));
    for (int q = 0; q < npoly; q++) {
      const int b = poly_off[q], n = poly_off[q + 1] - b;
      for (int k = 0; k < n; k++) {
        const int v = poly_verts[b + k];
        const int f = fill[v]++;
        nb[2 * f] = poly_verts[b + (k + n - 1) % n];
        nb[2 * f + 1] = poly_verts[b + (k + 1) % n];
      }
    }
    free(fill);
    int *vv_off = (int *)calloc((size_t)V + 1, sizeof(int));
    int *vv_idx = (int *)malloc(sizeof(int) * (size_t)(2 * cnt[V] + 1));
    for (int v = 0; v < V; v++) {
      int *lst = vv_idx + vv_off[v];
      int n = 0;
      for (int f = 2 * cnt[v]; f < 2 * cnt[v + 1]; f++) {
        const int u = nb[f];
        int dup = u == v;
        for (int a = 0; a < n && !dup; a++) {
          dup = lst[a] == u;
        }
        if (!dup) {
          int a = n++;
          while (a > 0 && lst[a - 1] > u) {
            lst[a] = lst[a - 1];
            a--;
          }
          lst[a] = u;
        }
      }
      vv_off[v + 1] = vv_off[v] + n;
    }
    free(nb);
    free(cnt);
    free(o->vv_off);
    free(o->vv_idx);
    o->vv_off = vv_off;
    o->vv_idx = vv_idx;
  }
  if (o->no64) {
    for (size_t i = 0; i < 3 * (size_t)o->totvert; i++) {
      o->no64[i] = o->no[i];
    }
  }
}

/* BKE_pbvh_build_mesh: pbvh.c:534-596. */
OrcPBVH *orc_build(const float *co, int totvert, const int *tris, int tottri, const int *tri_mat,
                   const uint8_t *vhide, int leaf_limit, int respect_hide, int with_fp64)
{
  OrcPBVH *o = (OrcPBVH *)calloc(1, sizeof(OrcPBVH));
  o->totvert = totvert;
  o->tottri = tottri;
  o->leaf_limit = leaf_limit > 0 ? leaf_limit : 10000; /* LEAF_LIMIT pbvh.c:32 */
  o->respect_hide = respect_hide;
  o->with_fp64 = with_fp64;
  o->co = (float *)malloc(sizeof(float) * 3 * (size_t)totvert);
  memcpy(o->co, co, sizeof(float) * 3 * (size_t)totvert);
  o->no = (float *)malloc(sizeof(float) * 3 * (size_t)totvert);
  o->tris = (int *)malloc(sizeof(int) * 3 * (size_t)(tottri > 0 ? tottri : 1));
  memcpy(o->tris, tris, sizeof(int) * 3 * (size_t)tottri);
  if (tri_mat) {
    o->tri_mat = (int *)malloc(sizeof(int) * (size_t)tottri);
    memcpy(o->tri_mat, tri_mat, sizeof(int) * (size_t)tottri);
  }
  if (vhide) {
    o->vhide = (uint8_t *)malloc((size_t)totvert);
    memcpy(o->vhide, vhide, (size_t)totvert);
  }
  o->vert_bitmap = (uint8_t *)calloc((size_t)totvert + 1, 1);
  o->map_local = (int *)malloc(sizeof(int) * (size_t)(totvert + 1));
  for (int v = 0; v < totvert; v++) {
    o->map_local[v] = INT32_MIN;
  }
  o->map_touched = (int *)malloc(sizeof(int) * (size_t)(3 * (size_t)(leaf_limit > 0 ? leaf_limit : 10000) + 3 * (size_t)tottri + 8));

  orc_calc_initial_normals(o->co, totvert, o->tris, tottri, o->no);
  build_csr(o);

  OBBC *bbc = (OBBC *)malloc(sizeof(OBBC) * (size_t)(tottri > 0 ? tottri : 1));
  OBB cb;
  bb_reset(&cb);
  for (int i = 0; i < tottri; i++) {
    OBB *b = (OBB *)&bbc[i];
    bb_reset(b);
    for (int j = 0; j < 3; j++) {
      bb_expand(b, o->co + 3 * o->tris[3 * i + j]);
    }
    for (int k = 0; k < 3; k++) {
      bbc[i].bcentroid[k] = (bbc[i].bmin[k] + bbc[i].bmax[k]) * 0.5f; /* pbvh.c:98-103 */
    }
    bb_expand(&cb, bbc[i].bcentroid);
  }
  if (tottri) {
    /* pbvh_build: pbvh.c:509-532 */
    o->prim_indices = (int *)malloc(sizeof(int) * (size_t)tottri);
    for (int i = 0; i < tottri; i++) {
      o->prim_indices[i] = i;
    }
    o->node_mem = 100;
    o->nodes = (ONode *)calloc(100, sizeof(ONode));
    o->totnode = 1;
    build_sub(o, 0, &cb, bbc, 0, tottri);
  }
  free(bbc);
  memset(o->vert_bitmap, 0, (size_t)totvert);

  if (with_fp64) {
    o->co64 = (double *)malloc(sizeof(double) * 3 * (size_t)totvert);
    o->no64 = (double *)malloc(sizeof(double) * 3 * (size_t)totvert);
    for (size_t i = 0; i < 3 * (size_t)totvert; i++) {
      o->co64[i] = o->co[i];
      o->no64[i] = o->no[i];
    }
  }
  return o;
}

void orc_free(OrcPBVH *o)
{
  if (!o) {
    return;
  }
  for (int i = 0; i < o->totnode; i++) {
    if (o->nodes[i].flag & ORC_Leaf) {
      free(o->nodes[i].vert_indices);
      free(o->nodes[i].face_vert_indices);
    }
  }
  free(o->nodes);
  free(o->prim_indices);
  free(o->co);
  free(o->no);
  free(o->co64);
  free(o->no64);
  free(o->mask);
  free(o->tris);
  free(o->tri_mat);
  free(o->tri_poly);
  free(o->poly_off);
  free(o->poly_verts);
  free(o->vhide);
  free(o->vert_bitmap);
  free(o->vt_off);
  free(o->vt_idx);
  free(o->vv_off);
  free(o->vv_idx);
  free(o->map_local);
  free(o->map_touched);
  free(o);
}

int orc_totnode(const OrcPBVH *o) { return o->totnode; }
int orc_totleaf(const OrcPBVH *o)
{
  int n = 0;
  for (int i = 0; i < o->totnode; i++) {
    n += (o->nodes[i].flag & ORC_Leaf) != 0;
  }
  return n;
}
void orc_node_info(const OrcPBVH *o, int node, int out[6])
{
  const ONode *n = &o->nodes[node];
  out[0] = n->flag;
  out[1] = (n->flag & ORC_Leaf) ? -1 : n->children_offset;
  out[2] = (n->flag & ORC_Leaf) ? n->prim_off : -1;
  out[3] = (n->flag & ORC_Leaf) ? n->totprim : 0;
  out[4] = n->uniq_verts;
  out[5] = n->face_verts;
}
void orc_node_bb(const OrcPBVH *o, int node, float vb[6], float orig_vb[6])
{
  const ONode *n = &o->nodes[node];
  if (vb) {
    memcpy(vb, &n->vb, sizeof(float) * 6);
  }
  if (orig_vb) {
    memcpy(orig_vb, &n->orig_vb, sizeof(float) * 6);
  }
}
int orc_node_verts(const OrcPBVH *o, int node, int *out, int cap)
{
  const ONode *n = &o->nodes[node];
  if (!(n->flag & ORC_Leaf)) {
    return 0;
  }
  const int tot = n->uniq_verts + n->face_verts;
  for (int i = 0; i < tot && i < cap; i++) {
    out[i] = n->vert_indices[i];
  }
  return tot;
}
void orc_prim_indices(const OrcPBVH *o, int *out)
{
  if (o->prim_indices) {
    memcpy(out, o->prim_indices, sizeof(int) * (size_t)o->tottri);
  }
}
void orc_get_positions(const OrcPBVH *o, float *out)
{
  memcpy(out, o->co, sizeof(float) * 3 * (size_t)o->totvert);
}
void orc_get_normals(const OrcPBVH *o, float *out)
{
  memcpy(out, o->no, sizeof(float) * 3 * (size_t)o->totvert);
}
void orc_get_positions64(const OrcPBVH *o, double *out)
{
  if (o->co64) {
    memcpy(out, o->co64, sizeof(double) * 3 * (size_t)o->totvert);
  }
}
void orc_get_normals64(const OrcPBVH *o, double *out)
{
  if (o->no64) {
    memcpy(out, o->no64, sizeof(double) * 3 * (size_t)o->totvert);
  }
}
void orc_set_normals(OrcPBVH *o, const float *no)
{
  memcpy(o->no, no, sizeof(float) * 3 * (size_t)o->totvert);
  if (o->no64) {
    for (size_t i = 0; i < 3 * (size_t)o->totvert; i++) {
      o->no64[i] = no[i];
    }
  }
}
void orc_set_mask(OrcPBVH *o, const float *mask)
{
  free(o->mask);
  o->mask = NULL;
  if (mask) {
    o->mask = (float *)malloc(sizeof(float) * (size_t)o->totvert);
    memcpy(o->mask, mask, sizeof(float) * (size_t)o->totvert);
  }
}
void orc_last_stats(const OrcPBVH *o, int64_t out[8]) { memcpy(out, o->stats, sizeof(o->stats)); }

/* ------------------------------------------------------------------------- */
/* Traversal: pbvh_iter_begin/next (pbvh.c:704-787).                         */
/* ------------------------------------------------------------------------- */

typedef int (*search_cb)(const OrcPBVH *o, const ONode *n, const void *data);

typedef struct {
  int node;
  int revisiting;
} OStack;

/* Returns leaves the callback accepts, in the order pbvh_iter_next yields them
 * (children pushed right then left, so leaves come out left to right).  Also
 * counts BB tests in *ntests when given. */
static int search_gather(const OrcPBVH *o, search_cb scb, const void *data, int **r_arr,
                         int64_t *ntests)
{
  if (!o->nodes || o->totnode == 0) {
    *r_arr = NULL;
    return 0;
  }
  int cap = 64, tot = 0;
  int *arr = (int *)malloc(sizeof(int) * (size_t)cap);
  int scap = 128, ssz = 0;
  OStack *st = (OStack *)malloc(sizeof(OStack) * (size_t)scap);
  st[ssz].node = 0;
  st[ssz].revisiting = 0;
  ssz++;
  while (ssz) {
    ssz--;
    const int ni = st[ssz].node;
    const ONode *n = &o->nodes[ni];
    if (st[ssz].revisiting) {
      continue; /* internal node after its children: not a leaf, not gathered */
    }
    if (scb) {
      if (ntests) {
        (*ntests)++;
      }
      if (!scb(o, n, data)) {
        continue;
      }
    }
    if (n->flag & ORC_Leaf) {
      if (tot == cap) {
        cap *= 2;
        arr = (int *)realloc(arr, sizeof(int) * (size_t)cap);
      }
      arr[tot++] = ni;
      continue;
    }
    if (ssz + 3 > scap) {
      scap *= 2;
      st = (OStack *)realloc(st, sizeof(OStack) * (size_t)scap);
    }
    st[ssz].node = ni;
    st[ssz].revisiting = 1;
    ssz++;
    st[ssz].node = n->children_offset + 1;
    st[ssz].revisiting = 0;
    ssz++;
    st[ssz].node = n->children_offset;
    st[ssz].revisiting = 0;
    ssz++;
  }
  free(st);
  *r_arr = arr;
  return tot;
}

typedef struct {
  float center[3];
  float radius_squared;
  int original;
} SphereData;

/* Spec S1 (the reference's SCULPT_search_sphere_cb is absent): nearest point of
 * the node box, strict "<" against r*r; fully hidden leaves are skipped. */
static int sphere_cb(const OrcPBVH *o, const ONode *n, const void *vd)
{
  const SphereData *d = (const SphereData *)vd;
  (void)o;
  if ((n->flag & ORC_Leaf) && (n->flag & ORC_FullyHidden)) {
    return 0;
  }
  const OBB *bb = d->original ? &n->orig_vb : &n->vb;
  float nearest[3], t[3];
  for (int i = 0; i < 3; i++) {
    if (bb->bmin[i] > d->center[i]) {
      nearest[i] = bb->bmin[i];
    }
    else if (bb->bmax[i] < d->center[i]) {
      nearest[i] = bb->bmax[i];
    }
    else {
      nearest[i] = d->center[i];
    }
  }
  t[0] = d->center[0] - nearest[0];
  t[1] = d->center[1] - nearest[1];
  t[2] = d->center[2] - nearest[2];
  return (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) < d->radius_squared;
}

/* update_search_cb: pbvh.c:973-982. */
static int update_cb(const OrcPBVH *o, const ONode *n, const void *vd)
{
  (void)o;
  const int flag = *(const int *)vd;
  if (n->flag & ORC_Leaf) {
    return (n->flag & flag) != 0;
  }
  return 1;
}

int orc_search_sphere(const OrcPBVH *o, const float center[3], float radius, int use_orig,
                      int *out, int cap)
{
  SphereData d;
  memcpy(d.center, center, sizeof(d.center));
  d.radius_squared = radius * radius;
  d.original = use_orig;
  int *arr;
  const int tot = search_gather(o, sphere_cb, &d, &arr, NULL);
  for (int i = 0; i < tot && i < cap; i++) {
    out[i] = arr[i];
  }
  free(arr);
  return tot;
}

/* ------------------------------------------------------------------------- */
/* Brush arithmetic: project spec S3-S9.                                     */
/* ------------------------------------------------------------------------- */

/* KERNEL_curvemap_evaluateF: colortools.c:942-962, table over [0,1] with
 * CM_TABLE+1 samples (mintable 0, range CM_TABLE). */
static float curve_table_eval(const float *table, float value)
{
  float fi = value * (float)ORC_CM_TABLE;
  int i = (int)fi;
  if (fi < 0.0f || fi > (float)ORC_CM_TABLE) {
    /* extrapolation: hold the end value (the CUSTOM table's default extend) */
    return fi < 0.0f ? table[0] : table[ORC_CM_TABLE];
  }
  if (i < 0) {
    return table[0];
  }
  if (i >= ORC_CM_TABLE) {
    return table[ORC_CM_TABLE];
  }
  fi = fi - (float)i;
  return (1.0f - fi) * table[i] + (fi)*table[i + 1];
}

/* Spec S3: hardness remap then KERNEL_brush_curve_strength_clamped (brush.h:84-91,
 * body absent; presets types_brush_enums.h:176-187). */
static float falloff32(const OrcDab *dab, float dist)
{
  const float r = dab->radius;
  const float h = dab->hardness;
  {
    float pr = dist / r;
    if (pr < h) {
      dist = 0.0f;
    }
    else if (h == 1.0f) {
      dist = r;
    }
    else {
      pr = (pr - h) / (1.0f - h);
      dist = pr * r;
    }
  }
  if (dist >= r) {
    return 0.0f;
  }
  float p = dist / r;
  p = 1.0f - p;
  float s;
  switch (dab->falloff) {
    case ORC_CURVE_CUSTOM:
      s = dab->curve_table ? curve_table_eval(dab->curve_table, 1.0f - p) : 1.0f;
      break;
    case ORC_CURVE_SHARP:
      s = p * p;
      break;
    case ORC_CURVE_SMOOTHER:
      s = ((p * p) * p) * (6.0f * p * p - 15.0f * p + 10.0f);
      break;
    case ORC_CURVE_ROOT:
      s = sqrtf(p);
      break;
    case ORC_CURVE_LIN:
      s = p;
      break;
    case ORC_CURVE_CONSTANT:
      s = 1.0f;
      break;
    case ORC_CURVE_SPHERE:
      s = sqrtf(2.0f * p - p * p);
      break;
    case ORC_CURVE_POW4:
      s = p * p * p * p;
      break;
    case ORC_CURVE_INVSQUARE:
      s = p * (2.0f - p);
      break;
    case ORC_CURVE_SMOOTH:
    default:
      s = 3.0f * p * p - 2.0f * p * p * p;
      break;
  }
  /* _clamped */
  if (s < 0.0f) {
    s = 0.0f;
  }
  else if (s > 1.0f) {
    s = 1.0f;
  }
  return s;
}

static double falloff64(const OrcDab *dab, double dist)
{
  const double r = dab->radius;
  const double h = dab->hardness;
  {
    double pr = dist / r;
    if (pr < h) {
      dist = 0.0;
    }
    else if (h == 1.0) {
      dist = r;
    }
    else {
      dist = ((pr - h) / (1.0 - h)) * r;
    }
  }
  if (dist >= r) {
    return 0.0;
  }
  double p = 1.0 - dist / r, s;
  switch (dab->falloff) {
    case ORC_CURVE_CUSTOM: {
      if (!dab->curve_table) {
        s = 1.0;
        break;
      }
      double fi = (1.0 - p) * ORC_CM_TABLE;
      int i = (int)fi;
      if (i < 0) {
        s = dab->curve_table[0];
      }
      else if (i >= ORC_CM_TABLE) {
        s = dab->curve_table[ORC_CM_TABLE];
      }
      else {
        fi -= i;
        s = (1.0 - fi) * dab->curve_table[i] + fi * dab->curve_table[i + 1];
      }
      break;
    }
    case ORC_CURVE_SHARP: s = p * p; break;
    case ORC_CURVE_SMOOTHER: s = p * p * p * (6.0 * p * p - 15.0 * p + 10.0); break;
    case ORC_CURVE_ROOT: s = sqrt(p); break;
    case ORC_CURVE_LIN: s = p; break;
    case ORC_CURVE_CONSTANT: s = 1.0; break;
    case ORC_CURVE_SPHERE: s = sqrt(2.0 * p - p * p); break;
    case ORC_CURVE_POW4: s = p * p * p * p; break;
    case ORC_CURVE_INVSQUARE: s = p * (2.0 - p); break;
    case ORC_CURVE_SMOOTH:
    default: s = 3.0 * p * p - 2.0 * p * p * p; break;
  }
  return s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
}

float orc_falloff(const OrcDab *dab, float d2) { return falloff32(dab, sqrtf(d2)); }

/* Spec S5: per-tool bstrength from Brush.alpha (s). */
static float bstrength32(const OrcDab *dab)
{
  const float s = dab->strength;
  const float flip = (dab->flag & ORC_BRUSH_DIR_IN) ? -1.0f : 1.0f;
  const float alpha = s * s;
  switch (dab->tool) {
    case ORC_TOOL_DRAW: return alpha * flip;
    case ORC_TOOL_INFLATE: return 0.25f * alpha * flip;
    case ORC_TOOL_SMOOTH: return s < 0.0f ? 0.0f : (s > 1.0f ? 1.0f : s);
    case ORC_TOOL_GRAB: return s;
    case ORC_TOOL_ELASTIC: return s;
  }
  return 0.0f;
}
float orc_bstrength(const OrcDab *dab) { return bstrength32(dab); }

/* Exact order-independent sum for the Draw area normal (spec S6): each FP32
 * component is rounded to a multiple of 2^-32 and summed in int64, so any
 * summation order gives the same bits. */
#define AREA_SCALE 4294967296.0
static inline int64_t area_fix(float x) { return (int64_t)llrint((double)x * AREA_SCALE); }

/* ------------------------------------------------------------------------- */
/* One dab.                                                                  */
/* ------------------------------------------------------------------------- */

static int cmp_int(const void *pa, const void *pb)
{
  const int a = *(const int *)pa, b = *(const int *)pb;
  return (a > b) - (a < b);
}

static inline float mask_of(const OrcPBVH *o, int v) { return o->mask ? o->mask[v] : 0.0f; }

/* BKE_pbvh_update_bounds(flag): pbvh.c:1401-1421 -> pbvh_update_BB_redraw :1206-1242 ->
 * update_node_vb :108-128 -> pbvh_flush_bb :1369-1399. */
static void update_bounds_flag(OrcPBVH *o, int flag, int nthreads)
{
  int *upd;
  const int nupd = search_gather(o, update_cb, &flag, &upd, NULL);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int li = 0; li < nupd; li++) {
    ONode *n = &o->nodes[upd[li]];
    if ((flag & ORC_UpdateBB) && (n->flag & ORC_UpdateBB)) {
      OBB vb;
      bb_reset(&vb);
      const int tot = n->uniq_verts + n->face_verts;
      for (int k = 0; k < tot; k++) {
        bb_expand(&vb, o->co + 3 * n->vert_indices[k]);
      }
      n->vb = vb;
    }
    if ((flag & ORC_UpdateOriginalBB) && (n->flag & ORC_UpdateOriginalBB)) {
      n->orig_vb = n->vb;
    }
    if ((flag & ORC_UpdateRedraw) && (n->flag & ORC_UpdateRedraw)) {
      n->flag &= ~ORC_UpdateRedraw;
    }
  }
  free(upd);
  if (!(flag & (ORC_UpdateBB | ORC_UpdateOriginalBB))) {
    return;
  }
  /* pbvh_flush_bb: serial recursion over the whole tree (explicit stack,
   * children before parents). */
  int *st = (int *)malloc(sizeof(int) * 2 * (size_t)(o->totnode + 2));
  int *upd_of = (int *)calloc((size_t)o->totnode, sizeof(int));
  int ssz = 0;
  st[ssz++] = 0;
  st[ssz++] = 0;
  while (ssz) {
    const int stage = st[--ssz];
    const int ni = st[--ssz];
    ONode *n = &o->nodes[ni];
    if (n->flag & ORC_Leaf) {
      int u = 0;
      if (flag & ORC_UpdateBB) {
        u |= n->flag & ORC_UpdateBB;
        n->flag &= ~ORC_UpdateBB;
      }
      if (flag & ORC_UpdateOriginalBB) {
        u |= n->flag & ORC_UpdateOriginalBB;
        n->flag &= ~ORC_UpdateOriginalBB;
      }
      upd_of[ni] = u;
      continue;
    }
    if (stage == 0) {
      st[ssz++] = ni;
      st[ssz++] = 1;
      st[ssz++] = n->children_offset + 1;
      st[ssz++] = 0;
      st[ssz++] = n->children_offset;
      st[ssz++] = 0;
      continue;
    }
    const int u = upd_of[n->children_offset] | upd_of[n->children_offset + 1];
    if (u & ORC_UpdateBB) {
      OBB vb;
      bb_reset(&vb);
      bb_expand_bb(&vb, &o->nodes[n->children_offset].vb);
      bb_expand_bb(&vb, &o->nodes[n->children_offset + 1].vb);
      n->vb = vb;
    }
    if (u & ORC_UpdateOriginalBB) {
      n->orig_vb = n->vb;
    }
    upd_of[ni] = u;
  }
  free(st);
  free(upd_of);
}

/* The bounds update a brush step ends with (UpdateBB | UpdateOriginalBB | UpdateRedraw). */
static void update_bounds(OrcPBVH *o, int nthreads)
{
  update_bounds_flag(o, ORC_UpdateBB | ORC_UpdateOriginalBB | ORC_UpdateRedraw, nthreads);
}

/* Normals of the flagged vertices of UpdateNormals leaves (pbvh_faces_update_normals
 * pbvh.c:1093-1118 structure; each vertex sums its incident tris in vertex->tri CSR
 * order, spec S10).  Returns the tri reads. */
static int64_t update_normals(OrcPBVH *o, int nthreads)
{
  const int fp64 = o->with_fp64;
  int *upd;
  int flag = ORC_UpdateNormals;
  const int nupd = search_gather(o, update_cb, &flag, &upd, NULL);
  int64_t tri_reads = 0;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) reduction(+ : tri_reads)
  for (int li = 0; li < nupd; li++) {
    ONode *n = &o->nodes[upd[li]];
    for (int k = 0; k < n->uniq_verts; k++) {
      const int v = n->vert_indices[k];
      if (!o->vert_bitmap[v]) {
        continue;
      }
      float acc[3] = {0.0f, 0.0f, 0.0f};
      double acc64[3] = {0.0, 0.0, 0.0};
      tri_reads += o->vt_off[v + 1] - o->vt_off[v];
      for (int q = o->vt_off[v]; q < o->vt_off[v + 1]; q++) {
        /* the face normal of the looptri's polygon (BKE_mesh_calc_poly_normal), once
         * per incident looptri */
        const int *t = o->tris + 3 * o->vt_idx[q];
        const int *pc = t;
        int pn = 3;
        if (o->tri_poly) {
          const int pi = o->tri_poly[o->vt_idx[q]];
          pc = o->poly_verts + o->poly_off[pi];
          pn = o->poly_off[pi + 1] - o->poly_off[pi];
        }
        float fn[3];
        poly_normal(o->co, pc, pn, fn);
        acc[0] = acc[0] + fn[0];
        acc[1] = acc[1] + fn[1];
        acc[2] = acc[2] + fn[2];
        if (fp64) {
          double fn64[3];
          poly_normal64(o->co64, pc, pn, fn64);
          acc64[0] += fn64[0];
          acc64[1] += fn64[1];
          acc64[2] += fn64[2];
        }
      }
      o_normalize(acc);
      memcpy(o->no + 3 * v, acc, sizeof(acc));
      if (fp64) {
        o_normalize64(acc64);
        memcpy(o->no64 + 3 * v, acc64, sizeof(acc64));
      }
      o->vert_bitmap[v] = 0;
    }
    n->flag &= ~ORC_UpdateNormals;
  }
  free(upd);
  return tri_reads;
}

int orc_dab(OrcPBVH *o, const OrcDab *dab, int *out_affected, int cap)
{
  memset(o->stats, 0, sizeof(o->stats));
  if (!o->nodes || o->totnode == 0) {
    return 0;
  }
  const int tool = dab->tool;
  const float r = dab->radius;
  const float r2 = r * r;
  const float *c = dab->center;
  const float bstr = bstrength32(dab);
  const double bstr64 = (double)bstr;
  const int fp64 = o->with_fp64;
  const int nthreads = orc_get_threads();

  SphereData sd;
  memcpy(sd.center, c, sizeof(sd.center));
  sd.radius_squared = r2;
  sd.original = (tool == ORC_TOOL_GRAB);
  int *leaves;
  int64_t ntests = 0;
  const int nleaf = search_gather(o, sphere_cb, &sd, &leaves, &ntests);
  o->stats[0] = nleaf;
  o->stats[1] = ntests;

  /* candidate offsets: per gathered leaf, its uniq verts */
  int *cand_off = (int *)malloc(sizeof(int) * (size_t)(nleaf + 1));
  cand_off[0] = 0;
  for (int i = 0; i < nleaf; i++) {
    cand_off[i + 1] = cand_off[i] + o->nodes[leaves[i]].uniq_verts;
  }
  const int ncand = cand_off[nleaf];
  o->stats[2] = ncand;
  /* per-candidate: affected flag, strength factor, new position */
  uint8_t *aff = (uint8_t *)calloc((size_t)ncand + 1, 1);
  float *fac = (float *)malloc(sizeof(float) * (size_t)(ncand + 1));
  float *newco = (float *)malloc(sizeof(float) * 3 * (size_t)(ncand + 1));
  double *newco64 = fp64 ? (double *)malloc(sizeof(double) * 3 * (size_t)(ncand + 1)) : NULL;
  double *fac64 = fp64 ? (double *)malloc(sizeof(double) * (size_t)(ncand + 1)) : NULL;

  /* Pass 1: sphere test (spec S2) and strength factor (S3/S4). */
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int li = 0; li < nleaf; li++) {
    const ONode *n = &o->nodes[leaves[li]];
    for (int k = 0; k < n->uniq_verts; k++) {
      const int v = n->vert_indices[k];
      const int ci = cand_off[li] + k;
      aff[ci] = 0;
      if (o->respect_hide && o->vhide && o->vhide[v]) {
        continue;
      }
      const float *co = o->co + 3 * v;
      float d[3];
      d[0] = c[0] - co[0];
      d[1] = c[1] - co[1];
      d[2] = c[2] - co[2];
      const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (d2 <= r2) {
        aff[ci] = 1;
        const float f = falloff32(dab, sqrtf(d2)) * (1.0f - mask_of(o, v));
        fac[ci] = f;
        if (fp64) {
          const double *c64 = o->co64 + 3 * v;
          double e[3] = {c[0] - c64[0], c[1] - c64[1], c[2] - c64[2]};
          fac64[ci] = falloff64(dab, sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2])) *
                      (1.0 - (double)mask_of(o, v));
        }
      }
    }
  }

  /* Draw: area normal over uniq verts within normal_radius_factor*r (S6). */
  float area_no[3] = {0.0f, 0.0f, 0.0f};
  double area_no64[3] = {0.0, 0.0, 0.0};
  if (tool == ORC_TOOL_DRAW) {
    const float nr = r * dab->normal_radius_factor;
    const float nr2 = nr * nr;
    int64_t s[3] = {0, 0, 0};
    double s64[3] = {0.0, 0.0, 0.0};
    int64_t na = 0;
    for (int li = 0; li < nleaf; li++) {
      const ONode *n = &o->nodes[leaves[li]];
      for (int k = 0; k < n->uniq_verts; k++) {
        const int v = n->vert_indices[k];
        if (o->respect_hide && o->vhide && o->vhide[v]) {
          continue;
        }
        const float *co = o->co + 3 * v;
        float d[3];
        d[0] = c[0] - co[0];
        d[1] = c[1] - co[1];
        d[2] = c[2] - co[2];
        const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (d2 <= nr2) {
          na++;
          for (int q = 0; q < 3; q++) {
            s[q] += area_fix(o->no[3 * v + q]);
            if (fp64) {
              s64[q] += o->no64[3 * v + q];
            }
          }
        }
      }
    }
    o->stats[3] = na;
    for (int q = 0; q < 3; q++) {
      area_no[q] = (float)((double)s[q] * (1.0 / AREA_SCALE));
      area_no64[q] = s64[q];
    }
    o_normalize(area_no);
    if (fp64) {
      o_normalize64(area_no64);
    }
  }

  /* Per-dab offset for Draw: (N * r) * bstrength. */
  float offset[3];
  double offset64[3];
  for (int q = 0; q < 3; q++) {
    offset[q] = (area_no[q] * r) * bstr;
    offset64[q] = area_no64[q] * (double)r * bstr64;
  }

  /* Pass 2: displacement into per-candidate proxies (reads before writes). */
  const int iters = (tool == ORC_TOOL_SMOOTH) ? (dab->smooth_iters > 0 ? dab->smooth_iters : 1) : 1;
  if (tool == ORC_TOOL_SMOOTH) {
    /* Spec S8: k Jacobi steps over the affected set; the affected set and the
     * factors are fixed from pre-dab positions.  vstate maps a vertex to its
     * candidate slot while the dab runs. */
    int *slot = (int *)malloc(sizeof(int) * (size_t)o->totvert);
    for (int v = 0; v < o->totvert; v++) {
      slot[v] = -1;
    }
    for (int li = 0; li < nleaf; li++) {
      const ONode *n = &o->nodes[leaves[li]];
      for (int k = 0; k < n->uniq_verts; k++) {
        const int ci = cand_off[li] + k;
        if (aff[ci]) {
          slot[n->vert_indices[k]] = ci;
        }
      }
    }
    float *prev = (float *)malloc(sizeof(float) * 3 * (size_t)(ncand + 1));
    double *prev64 = fp64 ? (double *)malloc(sizeof(double) * 3 * (size_t)(ncand + 1)) : NULL;
    for (int li = 0; li < nleaf; li++) {
      const ONode *n = &o->nodes[leaves[li]];
      for (int k = 0; k < n->uniq_verts; k++) {
        const int ci = cand_off[li] + k, v = n->vert_indices[k];
        memcpy(prev + 3 * ci, o->co + 3 * v, sizeof(float) * 3);
        if (fp64) {
          memcpy(prev64 + 3 * ci, o->co64 + 3 * v, sizeof(double) * 3);
        }
      }
    }
    const float bs = bstr;
    int64_t nbr_reads = 0;
    for (int it = 0; it < iters; it++) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) reduction(+ : nbr_reads)
      for (int li = 0; li < nleaf; li++) {
        const ONode *n = &o->nodes[leaves[li]];
        for (int k = 0; k < n->uniq_verts; k++) {
          const int ci = cand_off[li] + k;
          if (!aff[ci]) {
            continue;
          }
          const int v = n->vert_indices[k];
          const float *self = prev + 3 * ci;
          float sum[3] = {0.0f, 0.0f, 0.0f};
          double sum64[3] = {0.0, 0.0, 0.0};
          const int b = o->vv_off[v], e = o->vv_off[v + 1];
          nbr_reads += e - b;
          for (int q = b; q < e; q++) {
            const int u = o->vv_idx[q];
            const int su = slot[u];
            const float *pu = (su >= 0) ? prev + 3 * su : o->co + 3 * u;
            sum[0] = sum[0] + pu[0];
            sum[1] = sum[1] + pu[1];
            sum[2] = sum[2] + pu[2];
            if (fp64) {
              const double *pu64 = (su >= 0) ? prev64 + 3 * su : o->co64 + 3 * u;
              sum64[0] += pu64[0];
              sum64[1] += pu64[1];
              sum64[2] += pu64[2];
            }
          }
          const float fade = bs * fac[ci];
          float *dst = newco + 3 * ci;
          if (e > b) {
            const float inv = 1.0f / (float)(e - b);
            for (int q = 0; q < 3; q++) {
              const float avg = sum[q] * inv;
              dst[q] = self[q] + (avg - self[q]) * fade;
            }
          }
          else {
            memcpy(dst, self, sizeof(float) * 3);
          }
          if (fp64) {
            const double *s64 = prev64 + 3 * ci;
            double *d64 = newco64 + 3 * ci;
            const double fade64 = bstr64 * fac64[ci];
            for (int q = 0; q < 3; q++) {
              d64[q] = (e > b) ? s64[q] + (sum64[q] / (double)(e - b) - s64[q]) * fade64 : s64[q];
            }
          }
        }
      }
      /* Jacobi: swap the iterate into prev (only affected slots changed) */
      for (int ci = 0; ci < ncand; ci++) {
        if (aff[ci]) {
          memcpy(prev + 3 * ci, newco + 3 * ci, sizeof(float) * 3);
          if (fp64) {
            memcpy(prev64 + 3 * ci, newco64 + 3 * ci, sizeof(double) * 3);
          }
        }
      }
    }
    o->stats[4] = nbr_reads;
    free(prev);
    free(prev64);
    free(slot);
  }
  else {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int li = 0; li < nleaf; li++) {
      const ONode *n = &o->nodes[leaves[li]];
      for (int k = 0; k < n->uniq_verts; k++) {
        const int ci = cand_off[li] + k;
        if (!aff[ci]) {
          continue;
        }
        const int v = n->vert_indices[k];
        const float *co = o->co + 3 * v;
        float *dst = newco + 3 * ci;
        const float f = fac[ci];
        if (tool == ORC_TOOL_DRAW) {
          for (int q = 0; q < 3; q++) {
            dst[q] = co[q] + offset[q] * f; /* S6 */
          }
        }
        else if (tool == ORC_TOOL_INFLATE) {
          const float s = (bstr * f) * r; /* S7 */
          for (int q = 0; q < 3; q++) {
            dst[q] = co[q] + o->no[3 * v + q] * s;
          }
        }
        else { /* GRAB, S9: orig = pre-dab position */
          const float fade = bstr * f;
          for (int q = 0; q < 3; q++) {
            dst[q] = co[q] + dab->grab_delta[q] * fade;
          }
        }
        if (fp64) {
          const double *c64 = o->co64 + 3 * v;
          double *d64 = newco64 + 3 * ci;
          const double f64 = fac64[ci];
          for (int q = 0; q < 3; q++) {
            if (tool == ORC_TOOL_DRAW) {
              d64[q] = c64[q] + offset64[q] * f64;
            }
            else if (tool == ORC_TOOL_INFLATE) {
              d64[q] = c64[q] + o->no64[3 * v + q] * (bstr64 * f64 * (double)r);
            }
            else {
              d64[q] = c64[q] + (double)dab->grab_delta[q] * (bstr64 * f64);
            }
          }
        }
      }
    }
  }

  /* Apply proxies, mark verts (BKE_pbvh_vert_mark_update pbvh.c:1811) and nodes
   * (BKE_pbvh_node_mark_update pbvh.c:1723 on every gathered leaf). */
  int naff = 0;
  for (int li = 0; li < nleaf; li++) {
    ONode *n = &o->nodes[leaves[li]];
    n->flag |= ORC_UpdateNormals | ORC_UpdateBB | ORC_UpdateOriginalBB | ORC_UpdateDrawBuffers |
               ORC_UpdateRedraw;
    for (int k = 0; k < n->uniq_verts; k++) {
      const int ci = cand_off[li] + k;
      if (!aff[ci]) {
        continue;
      }
      const int v = n->vert_indices[k];
      memcpy(o->co + 3 * v, newco + 3 * ci, sizeof(float) * 3);
      if (fp64) {
        memcpy(o->co64 + 3 * v, newco64 + 3 * ci, sizeof(double) * 3);
      }
      o->vert_bitmap[v] = 1;
      naff++;
    }
  }
  o->stats[5] = naff;

  /* Normals (pbvh_faces_update_normals pbvh.c:1093-1118 structure; the accumulate
   * step sums every incident tri in vertex->tri CSR order, spec S10). */
  o->stats[6] = update_normals(o, nthreads);

  /* Bounds: BKE_pbvh_update_bounds(UpdateBB|UpdateOriginalBB|UpdateRedraw). */
  update_bounds(o, nthreads);

  /* output */
  if (out_affected) {
    int w = 0;
    for (int li = 0; li < nleaf && w < cap; li++) {
      const ONode *n = &o->nodes[leaves[li]];
      for (int k = 0; k < n->uniq_verts && w < cap; k++) {
        if (aff[cand_off[li] + k]) {
          out_affected[w++] = n->vert_indices[k];
        }
      }
    }
    qsort(out_affected, (size_t)w, sizeof(int), cmp_int); /* ascending vertex id */
  }
  o->stats[7] = iters;

  free(leaves);
  free(cand_off);
  free(aff);
  free(fac);
  free(newco);
  free(newco64);
  free(fac64);
  return naff;
}

/* Grab from the original coordinates: the Grab brush task over SculptOrigVertData
 * (sculpt.c, absent from the reference; Blender's published do_grab_brush_task_cb_ex:
 * test and falloff on the original coordinate, proxy = grab_delta * fade) followed by
 * the proxy combine with use_orco (every visible uniq vertex of a gathered node becomes
 * original + proxy, zero outside the sphere).  Search on orig_vb as the Grab dab; the
 * node and vertex marks, normals and bounds as orc_dab.  orco: totvert*3 original
 * coordinates.  out_affected: every written vertex, ascending. */
int orc_dab_grab_orig(OrcPBVH *o, const OrcDab *dab, const float *orco, int *out_affected, int cap)
{
  memset(o->stats, 0, sizeof(o->stats));
  if (!o->nodes || o->totnode == 0 || dab->tool != ORC_TOOL_GRAB) {
    return 0;
  }
  const float r = dab->radius;
  const float r2 = r * r;
  const float *c = dab->center;
  const float bstr = bstrength32(dab);
  SphereData sd;
  memcpy(sd.center, c, sizeof(sd.center));
  sd.radius_squared = r2;
  sd.original = 1;
  int *leaves;
  int64_t ntests = 0;
  const int nleaf = search_gather(o, sphere_cb, &sd, &leaves, &ntests);
  o->stats[0] = nleaf;
  o->stats[1] = ntests;
  int naff = 0, w = 0;
  for (int li = 0; li < nleaf; li++) {
    ONode *n = &o->nodes[leaves[li]];
    n->flag |= ORC_UpdateNormals | ORC_UpdateBB | ORC_UpdateOriginalBB | ORC_UpdateDrawBuffers |
               ORC_UpdateRedraw;
    for (int k = 0; k < n->uniq_verts; k++) {
      const int v = n->vert_indices[k];
      if (o->respect_hide && o->vhide && o->vhide[v]) {
        continue; /* the vertex iterator skips hidden vertices */
      }
      const float *oc = orco + 3 * (size_t)v;
      float d[3];
      d[0] = c[0] - oc[0];
      d[1] = c[1] - oc[1];
      d[2] = c[2] - oc[2];
      const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      float val[3] = {oc[0], oc[1], oc[2]};
      if (d2 <= r2) {
        const float f = falloff32(dab, sqrtf(d2)) * (1.0f - mask_of(o, v));
        const float fade = bstr * f;
        for (int q = 0; q < 3; q++) {
          val[q] = oc[q] + dab->grab_delta[q] * fade;
        }
      }
      memcpy(o->co + 3 * (size_t)v, val, sizeof(val));
      if (o->co64) {
        for (int q = 0; q < 3; q++) {
          o->co64[3 * (size_t)v + q] = val[q];
        }
      }
      o->vert_bitmap[v] = 1;
      naff++;
      if (out_affected && w < cap) {
        out_affected[w++] = v;
      }
    }
  }
  o->stats[5] = naff;
  o->stats[6] = update_normals(o, orc_get_threads());
  update_bounds(o, orc_get_threads());
  if (out_affected) {
    qsort(out_affected, (size_t)w, sizeof(int), cmp_int);
  }
  free(leaves);
  return naff;
}

int orc_leaves_dfs(const OrcPBVH *o, int *out, int cap)
{
  int *arr;
  const int tot = search_gather(o, NULL, NULL, &arr, NULL);
  for (int i = 0; i < tot && i < cap; i++) {
    out[i] = arr[i];
  }
  free(arr);
  return tot;
}

/* Replace every position (e.g. with another engine's result) and refit all node
 * bounds as a dab does (orig_vb follows vb). */
void orc_set_positions(OrcPBVH *o, const float *co)
{
  memcpy(o->co, co, sizeof(float) * 3 * (size_t)o->totvert);
  if (o->co64) {
    for (size_t i = 0; i < 3 * (size_t)o->totvert; i++) {
      o->co64[i] = co[i];
    }
  }
  for (int i = 0; i < o->totnode; i++) {
    if (o->nodes[i].flag & ORC_Leaf) {
      o->nodes[i].flag |= ORC_UpdateBB | ORC_UpdateOriginalBB | ORC_UpdateRedraw;
    }
  }
  update_bounds(o, orc_get_threads());
}

/* The sculpt code's proxy combine (sculpt_combine_proxies, absent from the reference;
 * Blender's published version) over the nodes BKE_pbvh_gather_proxies returns
 * (pbvh.c:4791-4818), in node order: every uniq vertex i of a node becomes
 * base + proxy_0[i] + proxy_1[i] + ... (add_v3_v3, component by component, left to
 * right), base = orco[v] (the original coordinate) when orco is given, else co[v];
 * SCULPT_clip without mirror clipping copies it; the vertex is flagged
 * (ME_VERT_PBVH_UPDATE) and the node marked as the brush marks it
 * (BKE_pbvh_node_mark_update, pbvh.c:1723-1727); then the normal and bounds updates.
 * prox: the nodes' proxies packed [nprox[j]][uniq_verts][3], node after node. */
void orc_combine_proxies(OrcPBVH *o, const int *nodes, int n, const int *nprox, const float *prox,
                         const float *orco)
{
  size_t off = 0;
  for (int j = 0; j < n; j++) {
    ONode *node = &o->nodes[nodes[j]];
    const int nu = node->uniq_verts;
    for (int i = 0; i < nu; i++) {
      const int v = node->vert_indices[i];
      const float *b = orco ? orco + 3 * (size_t)v : o->co + 3 * (size_t)v;
      float val[3] = {b[0], b[1], b[2]};
      for (int q = 0; q < nprox[j]; q++) {
        const float *d = prox + off + 3 * ((size_t)q * nu + i);
        val[0] = val[0] + d[0];
        val[1] = val[1] + d[1];
        val[2] = val[2] + d[2];
      }
      memcpy(o->co + 3 * (size_t)v, val, sizeof(val));
      if (o->co64) {
        o->co64[3 * (size_t)v] = val[0];
        o->co64[3 * (size_t)v + 1] = val[1];
        o->co64[3 * (size_t)v + 2] = val[2];
      }
      o->vert_bitmap[v] = 1;
    }
    off += (size_t)nprox[j] * nu * 3;
    node->flag |= ORC_UpdateNormals | ORC_UpdateBB | ORC_UpdateOriginalBB | ORC_UpdateDrawBuffers |
                  ORC_UpdateRedraw;
  }
  update_normals(o, orc_get_threads());
  update_bounds(o, orc_get_threads());
}

/* The normal and bounds updates a dab ends with, for the other oracle files. */
int64_t orc_update_normals_bounds(OrcPBVH *o)
{
  const int64_t t = update_normals(o, orc_get_threads());
  update_bounds(o, orc_get_threads());
  return t;
}

/* ---- the drop-in update API, restated for the host-edit tests -------------- */

/* A caller's edit of one vertex through the mesh arrays (BKE_pbvh_get_verts). */
void orc_set_vert(OrcPBVH *o, int v, const float co[3])
{
  memcpy(o->co + 3 * (size_t)v, co, sizeof(float) * 3);
  if (o->co64) {
    for (int k = 0; k < 3; k++) {
      o->co64[3 * (size_t)v + k] = co[k];
    }
  }
}

/* BKE_pbvh_vert_mark_update: pbvh.c:1811-1815. */
void orc_vert_mark_update(OrcPBVH *o, int v) { o->vert_bitmap[v] = 1; }

/* BKE_pbvh_node_mark_update: pbvh.c:1723-1727. */
void orc_node_mark_update(OrcPBVH *o, int node)
{
  o->nodes[node].flag |= ORC_UpdateNormals | ORC_UpdateBB | ORC_UpdateOriginalBB | ORC_UpdateDrawBuffers |
                         ORC_UpdateRedraw;
}

void orc_update_bounds(OrcPBVH *o, int flag) { update_bounds_flag(o, flag, orc_get_threads()); }

void orc_update_normals(OrcPBVH *o) { update_normals(o, orc_get_threads()); }

int orc_node_flag(const OrcPBVH *o, int node) { return o->nodes[node].flag; }

/* BKE_pbvh_vert_coords_apply: pbvh.c:4707-4749.  Each position that differs bitwise is
 * replaced and its vertex marked; BKE_mesh_calc_normals_looptri computes into a
 * temporary it frees (mesh_normals.c:455-507), so the vertex normals are unchanged;
 * every node is marked and the bounds updated with UpdateBB | UpdateOriginalBB. */
void orc_vert_coords_apply(OrcPBVH *o, const float *vert_cos)
{
  for (int v = 0; v < o->totvert; v++) {
    if (memcmp(o->co + 3 * (size_t)v, vert_cos + 3 * (size_t)v, sizeof(float) * 3) != 0) {
      orc_set_vert(o, v, vert_cos + 3 * (size_t)v);
      o->vert_bitmap[v] = 1;
    }
  }
  for (int i = 0; i < o->totnode; i++) {
    orc_node_mark_update(o, i);
  }
  update_bounds_flag(o, ORC_UpdateBB | ORC_UpdateOriginalBB, orc_get_threads());
}
