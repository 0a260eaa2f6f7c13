/*
 * Dynamic topology (PBVH_BMESH) in the CPU oracle -- TEST INFRASTRUCTURE ONLY (see
 * oracle.h).  Plain C, FP32 without contraction, restating:
 *
 *   tree build (fast build, leaf limit 100)   kernel/intern/pbvh_bmesh.c:1670-1928
 *   leaf split after a stroke                 pbvh_bmesh.c:179-235, :237-400, :2065-2078
 *   edge queues: faces in range (sphere or    pbvh_bmesh.c:773-1090
 *     view-projected circle), front faces,
 *     mask / hidden vertices, the recursive
 *     "even subdivision" neighbours
 *   edge split                                pbvh_bmesh.c:1093-1209
 *   edge collapse                             pbvh_bmesh.c:1264-1429
 *   update topology (collapse, then split)    pbvh_bmesh.c:1931-2005
 *   normals of flagged leaves                 pbvh_bmesh.c:1635-1656 with
 *                                             mesh/tools/intern/mesh_polygon.c:591-693, :722
 *   the dab                                   the faces dab's spec over each gathered
 *                                             leaf's own (unique) vertices, after the
 *                                             topology update of the leaves within
 *                                             1.25 r (the reference's sculpt code that
 *                                             calls it is absent)
 *
 * The mesh is a corner table (the BMesh substrate is absent from the reference's
 * boundary): triangle t has corners 3t..3t+2, tv[c] is the corner's vertex and topp[c]
 * the corner across the edge opposite c (-1 on a boundary); the surface must be
 * manifold with boundary.  Every vertex carries a 64-bit label (its input index, or a
 * hash of its edge's labels when a split creates it), so results never depend on slot
 * numbers; the product (dyntopo.hip) allocates slots in any order and compares equal
 * through the labels.
 *
 * Where this departs from the reference, and why:
 *  - Edge operations run in rounds instead of one heap: each round collects the
 *    queue's edges (the reference's rules for which edges enter it), gives each a key
 *    (a band of its length first -- longest for splits, shortest for collapses, four
 *    bands per octave of squared length -- then a hash of its two labels), and applies every edge that holds the largest key on all faces it
 *    touches (its two faces for a split, whose neighbours are joined in a second
 *    phase; every face around both vertices for a collapse); when two edges share the
 *    largest key of a face, both wait.  Rounds
 *    repeat until no edge qualifies.  The reference pops one edge at a time; its order
 *    among equal lengths follows pointer hashing (GSet iteration), so it has no
 *    reproducible order to match.  A single edge, or edges that touch no common face,
 *    are split exactly as the reference splits them (tests pin this against a
 *    sequential heap restatement, orc_dyn_subdivide_sequential).
 *  - A collapse that would make the surface non-manifold (link condition), pinch two
 *    boundaries, or remove an ear is skipped; the reference collapses and deletes the
 *    duplicate faces.  Faces a collapse re-points keep their corner order (the
 *    reference re-creates them starting at the kept vertex).
 *  - Sums over a vertex's neighbours (Smooth) and faces (vertex normals), and the Draw
 *    area normal, are exact 2^-32 fixed-point sums (order free; the reference's order
 *    is its disk-cycle history).  saacos is evaluated in double and rounded.
 *  - New owners of vertices: the new vertex goes to the leaf of the split face whose
 *    opposite vertex has the smaller label; a vertex whose owner leaf loses its last
 *    face with it goes to the smallest-numbered leaf among its faces.
 *  - Each leaf holds at most DYN_SLAB face slots between after_stroke calls; a split
 *    touching a leaf already past half of that waits for the next after_stroke.
 *  - Leaf boxes are min/max over the leaf faces' vertices with zeros taken as +0.
 *  - Grab keeps the topology (SCULPT_TOOL_HAS_DYNTOPO excludes it) and reads current
 *    positions (no original data on this path).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_int.h"

#define DYN_SLAB 4096      /* face slots per leaf */
#define DYN_MAX_ROUNDS 64  /* rounds per phase of one topology update */
#define DYN_RING 128       /* faces around a vertex (larger: the vertex is skipped) */
#define DYN_STACK 32       /* pending edges of the even-subdivision walk */

struct OrcDyn {
  int V, Vcap, T, Tcap;
  float *co, *no, *mask; /* mask NULL: no mask layer */
  uint8_t *vhide;
  uint64_t *lab;
  int *vnode, *vcorner; /* owner node (-1: none), a corner at the vertex (-1: dead) */
  int *tv, *topp, *tnode, *tslot;
  float *fno;
  uint8_t *fhide;
  ONode *nodes;
  int totnode, node_mem;
  int *slab, *scount; /* [node_mem * DYN_SLAB] face ids (-1: hole), appended count */
  int leaf_limit, respect_hide;
  float max_len, min_len;
  int serial; /* topology updates so far (labels of new vertices) */
  int mode, use_frontface, use_projected;
  float view_normal[3];
  int64_t stats[12];
  /* scratch kept across calls, valid by stamp (sized by the capacities) */
  int stamp;
  int *estamp, *remap, *remap_ser; /* per corner */
  int *wstamp, *wlist;             /* per face */
  uint64_t *best, *tie;            /* per face: largest key, largest tied key */
  int *vst, *vslot;                /* per vertex */
  int *cand, cand_mem;             /* grown on demand */
};

/* stats: [0] gathered [1] affected [2] splits [3] collapses [4] split rounds
 * [5] collapse rounds [6] split candidates [7] collapse candidates [8] capacity stop
 * [9] topology leaves [10] faces re-normalled [11] leaves refit */

static inline int c_next(int c) { return (c % 3 == 2) ? c - 2 : c + 1; }
static inline int c_prev(int c) { return (c % 3 == 0) ? c + 2 : c - 1; }
static inline float d_min(float a, float b) { return (a < b) ? a : b; }
static inline float d_max(float a, float b) { return (a > b) ? a : b; }

static void dbb_reset(OBB *b)
{
  b->bmin[0] = b->bmin[1] = b->bmin[2] = FLT_MAX;
  b->bmax[0] = b->bmax[1] = b->bmax[2] = -FLT_MAX;
}
static void dbb_expand(OBB *b, const float *co)
{
  for (int i = 0; i < 3; i++) {
    b->bmin[i] = d_min(b->bmin[i], co[i]);
    b->bmax[i] = d_max(b->bmax[i], co[i]);
  }
}
static void dbb_expand_bb(OBB *b, const OBB *c)
{
  for (int i = 0; i < 3; i++) {
    b->bmin[i] = d_min(b->bmin[i], c->bmin[i]);
    b->bmax[i] = d_max(b->bmax[i], c->bmax[i]);
  }
}
/* BB_widest_axis (pbvh.c:108-128 in this reference's copy) */
static int dbb_widest_axis(const OBB *bb)
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

/* normalize_v3, math_vector_inline.c:1165-1180 */
static float d_normalize(float n[3])
{
  float d = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  if (d > 1.0e-35f) {
    d = sqrtf(d);
    const float inv = 1.0f / d;
    n[0] *= inv;
    n[1] *= inv;
    n[2] *= inv;
  }
  else {
    n[0] = n[1] = n[2] = 0.0f;
    d = 0.0f;
  }
  return d;
}

/* len_squared_v3v3 */
static inline float d_len2(const float *a, const float *b)
{
  const float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
  return x * x + y * y + z * z;
}

/* normal_tri_v3, math_geom.cc:31-46 */
static void d_tri_normal(float n[3], const float *v1, const float *v2, const float *v3)
{
  float n1[3], n2[3];
  for (int i = 0; i < 3; i++) {
    n1[i] = v1[i] - v2[i];
    n2[i] = v2[i] - v3[i];
  }
  n[0] = n1[1] * n2[2] - n1[2] * n2[1];
  n[1] = n1[2] * n2[0] - n1[0] * n2[2];
  n[2] = n1[0] * n2[1] - n1[1] * n2[0];
  d_normalize(n);
}

/* closest_on_tri_to_point_v3, math_geom.cc:966-1044 */
static void d_closest_on_tri(float r[3], const float p[3], const float v1[3], const float v2[3], const float v3[3])
{
  float ab[3], ac[3], ap[3], bp[3], cp[3];
  for (int i = 0; i < 3; i++) {
    ab[i] = v2[i] - v1[i];
    ac[i] = v3[i] - v1[i];
    ap[i] = p[i] - v1[i];
  }
  const float d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
  const float d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
  if (d1 <= 0.0f && d2 <= 0.0f) {
    memcpy(r, v1, sizeof(float) * 3);
    return;
  }
  for (int i = 0; i < 3; i++) {
    bp[i] = p[i] - v2[i];
  }
  const float d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
  const float d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
  if (d3 >= 0.0f && d4 <= d3) {
    memcpy(r, v2, sizeof(float) * 3);
    return;
  }
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float v = d1 / (d1 - d3);
    for (int i = 0; i < 3; i++) {
      r[i] = v1[i] + ab[i] * v;
    }
    return;
  }
  for (int i = 0; i < 3; i++) {
    cp[i] = p[i] - v3[i];
  }
  const float d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
  const float d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
  if (d6 >= 0.0f && d5 <= d6) {
    memcpy(r, v3, sizeof(float) * 3);
    return;
  }
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = d2 / (d2 - d6);
    for (int i = 0; i < 3; i++) {
      r[i] = v1[i] + ac[i] * w;
    }
    return;
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int i = 0; i < 3; i++) {
      r[i] = (v3[i] - v2[i]) * w + v2[i];
    }
    return;
  }
  const float denom = 1.0f / (va + vb + vc);
  const float v = vb * denom;
  const float w = vc * denom;
  for (int i = 0; i < 3; i++) {
    r[i] = (v1[i] + ab[i] * v) + ac[i] * w;
  }
}

/* project_plane_normalized_v3_v3v3, math_vector.c:635-641 */
static void d_project_plane(float out[3], const float p[3], const float n[3])
{
  const float mul = p[0] * n[0] + p[1] * n[1] + p[2] * n[2];
  for (int i = 0; i < 3; i++) {
    out[i] = p[i] + n[i] * -mul;
  }
}

/* splitmix64's finalizer (Steele, Lea, Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014) */
static inline uint64_t d_mix(uint64_t x)
{
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}
static inline uint64_t d_new_label(uint64_t a, uint64_t b, int serial, int round)
{
  const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
  return d_mix(lo ^ d_mix(hi ^ d_mix(((uint64_t)(uint32_t)serial << 20) ^ (uint64_t)(uint32_t)round))) |
         (1ULL << 63);
}
static inline uint64_t d_edge_hash33(uint64_t a, uint64_t b)
{
  const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
  return d_mix(lo * 0x9E3779B97F4A7C15ULL ^ hi) >> 31;
}
static inline uint32_t f_bits(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

/* ------------------------------------------------------------------------- */
/* Corner-table queries.                                                      */
/* ------------------------------------------------------------------------- */

/* The corners at v, face by face around it (a boundary vertex from the fan's first
 * face); -1 when more than cap.  *bnd: the vertex is on a boundary. */
static int d_ring(const OrcDyn *o, int v, int *out, int cap, int *bnd)
{
  *bnd = 0;
  const int c0 = o->vcorner[v];
  if (c0 < 0) {
    return 0;
  }
  int c = c0, steps = 0;
  for (;;) {
    const int x = o->topp[c_next(c)];
    if (x < 0) {
      *bnd = 1;
      break;
    }
    c = c_next(x);
    if (c == c0) {
      break;
    }
    if (++steps > cap) {
      return -1;
    }
  }
  const int start = c;
  int n = 0;
  for (;;) {
    if (n >= cap) {
      return -1;
    }
    out[n++] = c;
    const int x = o->topp[c_prev(c)];
    if (x < 0) {
      break;
    }
    c = c_prev(x);
    if (c == start) {
      break;
    }
  }
  return n;
}

/* The vertices joined to v by an edge (each once). */
static int d_neighbors(const OrcDyn *o, int v, int *out, int cap)
{
  int cs[DYN_RING];
  int bnd;
  const int n = d_ring(o, v, cs, DYN_RING, &bnd);
  if (n < 0) {
    return -1;
  }
  int k = 0;
  if (bnd && n > 0 && k < cap) {
    out[k++] = o->tv[c_prev(cs[0])];
  }
  for (int i = 0; i < n && k < cap; i++) {
    out[k++] = o->tv[c_next(cs[i])];
  }
  return k;
}

static int d_valence(const OrcDyn *o, int v)
{
  int cs[DYN_RING];
  int bnd;
  const int n = d_ring(o, v, cs, DYN_RING, &bnd);
  return n < 0 ? DYN_RING + 1 : n + bnd;
}

static inline float edge_len2(const OrcDyn *o, int c)
{
  return d_len2(o->co + 3 * o->tv[c_next(c)], o->co + 3 * o->tv[c_prev(c)]);
}

/* edge_queue_tri_in_sphere / _in_circle (pbvh_bmesh.c:773-805) */
static int tri_in_range(const OrcDyn *o, int f, const float center[3], const float center_proj[3], float r2,
                        int projected)
{
  const float *v0 = o->co + 3 * o->tv[3 * f], *v1 = o->co + 3 * o->tv[3 * f + 1], *v2 = o->co + 3 * o->tv[3 * f + 2];
  float c[3];
  if (projected) {
    float p0[3], p1[3], p2[3];
    d_project_plane(p0, v0, o->view_normal);
    d_project_plane(p1, v1, o->view_normal);
    d_project_plane(p2, v2, o->view_normal);
    d_closest_on_tri(c, center_proj, p0, p1, p2);
    return d_len2(center_proj, c) <= r2;
  }
  d_closest_on_tri(c, center, v0, v1, v2);
  return d_len2(center, c) <= r2;
}

static inline int front_ok(const OrcDyn *o, int f)
{
  if (!o->use_frontface) {
    return 1;
  }
  const float *n = o->fno + 3 * f;
  return !(n[0] * o->view_normal[0] + n[1] * o->view_normal[1] + n[2] * o->view_normal[2] < 0.0f);
}

/* edge_queue_insert's filter (pbvh_bmesh.c:812-833): some vertex below full mask,
 * neither hidden */
static inline int edge_insert_ok(const OrcDyn *o, int c)
{
  const int a = o->tv[c_next(c)], b = o->tv[c_prev(c)];
  if (o->mask && !(o->mask[a] < 1.0f || o->mask[b] < 1.0f)) {
    return 0;
  }
  return !(o->vhide && (o->vhide[a] || o->vhide[b]));
}

static inline int canon(const OrcDyn *o, int c)
{
  const int x = o->topp[c];
  return (x >= 0 && x < c) ? x : c;
}

/* ------------------------------------------------------------------------- */
/* Tree and leaf face lists.                                                  */
/* ------------------------------------------------------------------------- */

static void d_grow_nodes(OrcDyn *o, int totnode)
{
  if (totnode > o->node_mem) {
    int mem = o->node_mem ? o->node_mem : 16;
    while (mem < totnode) {
      mem *= 2;
    }
    o->nodes = (ONode *)realloc(o->nodes, sizeof(ONode) * (size_t)mem);
    memset(o->nodes + o->node_mem, 0, sizeof(ONode) * (size_t)(mem - o->node_mem));
    o->slab = (int *)realloc(o->slab, sizeof(int) * (size_t)mem * DYN_SLAB);
    o->scount = (int *)realloc(o->scount, sizeof(int) * (size_t)mem);
    memset(o->scount + o->node_mem, 0, sizeof(int) * (size_t)(mem - o->node_mem));
    o->node_mem = mem;
  }
  o->totnode = totnode;
}

static void slab_append(OrcDyn *o, int node, int f)
{
  const int pos = o->scount[node]++;
  o->slab[(size_t)node * DYN_SLAB + pos] = f;
  o->tnode[f] = node;
  o->tslot[f] = pos;
}

static void face_bbc(const OrcDyn *o, int f, OBBC *b)
{
  OBB *bb = (OBB *)b;
  dbb_reset(bb);
  for (int k = 0; k < 3; k++) {
    dbb_expand(bb, o->co + 3 * o->tv[3 * f + k]);
  }
  for (int i = 0; i < 3; i++) {
    b->bcentroid[i] = (b->bmin[i] + b->bmax[i]) * 0.5f;
  }
}

typedef struct {
  int totface, start, c1, c2;
} FastInfo;

/* pbvh_bmesh_node_limit_ensure_fast (pbvh_bmesh.c:1670-1763) */
static void limit_ensure_fast(OrcDyn *o, int *nodeinfo, const OBBC *bbc, FastInfo *info, int *ninfo, int ni)
{
  if (info[ni].totface <= o->leaf_limit) {
    return;
  }
  OBB cb;
  dbb_reset(&cb);
  for (int i = 0; i < info[ni].totface; i++) {
    dbb_expand(&cb, bbc[nodeinfo[i + info[ni].start]].bcentroid);
  }
  const int axis = dbb_widest_axis(&cb);
  const float mid = (cb.bmax[axis] + cb.bmin[axis]) * 0.5f;
  int num1 = 0, num2 = 0;
  const int end = info[ni].start + info[ni].totface;
  for (int i = info[ni].start; i < end - num2; i++) {
    if (bbc[nodeinfo[i]].bcentroid[axis] > mid) {
      int i_iter = end - num2 - 1, candidate = -1;
      for (; i_iter > i; i_iter--) {
        if (bbc[nodeinfo[i_iter]].bcentroid[axis] <= mid) {
          candidate = i_iter;
          break;
        }
        num2++;
      }
      if (candidate != -1) {
        const int tmp = nodeinfo[i];
        nodeinfo[i] = nodeinfo[candidate];
        nodeinfo[candidate] = tmp;
        num1++;
        num2++;
      }
      else {
        num2++;
      }
    }
    else {
      num1++;
    }
  }
  if (num2 == 0) {
    num2++;
    num1--;
  }
  else if (num1 == 0) {
    num1++;
    num2--;
  }
  const int c1 = (*ninfo)++, c2 = (*ninfo)++;
  info[c1].totface = num1;
  info[c1].start = info[ni].start;
  info[c2].totface = num2;
  info[c2].start = info[ni].start + num1;
  info[c1].c1 = info[c1].c2 = info[c2].c1 = info[c2].c2 = -1;
  info[ni].c1 = c1;
  info[ni].c2 = c2;
  limit_ensure_fast(o, nodeinfo, bbc, info, ninfo, c1);
  limit_ensure_fast(o, nodeinfo, bbc, info, ninfo, c2);
}

/* pbvh_bmesh_create_nodes_fast_recursive (pbvh_bmesh.c:1765-1852) */
static void create_nodes_fast(OrcDyn *o, const int *nodeinfo, const OBBC *bbc, const FastInfo *info, int ii,
                              int node_index)
{
  if (info[ii].c1 >= 0) {
    const int children = o->totnode;
    o->nodes[node_index].children_offset = children;
    d_grow_nodes(o, o->totnode + 2);
    create_nodes_fast(o, nodeinfo, bbc, info, info[ii].c1, children);
    create_nodes_fast(o, nodeinfo, bbc, info, info[ii].c2, children + 1);
    ONode *n = &o->nodes[node_index];
    dbb_reset(&n->vb);
    dbb_expand_bb(&n->vb, &o->nodes[children].vb);
    dbb_expand_bb(&n->vb, &o->nodes[children + 1].vb);
    n->orig_vb = n->vb;
    return;
  }
  ONode *n = &o->nodes[node_index];
  n->flag = ORC_Leaf;
  n->children_offset = -1;
  dbb_reset(&n->vb);
  int has_visible = 0;
  o->scount[node_index] = 0;
  for (int i = info[ii].start; i < info[ii].start + info[ii].totface; i++) {
    const int f = nodeinfo[i];
    slab_append(o, node_index, f);
    for (int k = 0; k < 3; k++) {
      const int v = o->tv[3 * f + k];
      if (o->vnode[v] < 0) {
        o->vnode[v] = node_index;
      }
    }
    if (!(o->fhide && o->fhide[f])) {
      has_visible = 1;
    }
    dbb_expand_bb(&n->vb, (const OBB *)&bbc[f]);
  }
  n->orig_vb = n->vb;
  n->flag |= ORC_UpdateNormals | ORC_UpdateDrawBuffers;
  if (!has_visible) {
    n->flag |= ORC_FullyHidden;
  }
}

static void build_tree(OrcDyn *o)
{
  OBBC *bbc = (OBBC *)malloc(sizeof(OBBC) * (size_t)(o->T + 1));
  int *nodeinfo = (int *)malloc(sizeof(int) * (size_t)(o->T + 1));
  for (int f = 0; f < o->T; f++) {
    face_bbc(o, f, &bbc[f]);
    nodeinfo[f] = f;
  }
  for (int v = 0; v < o->V; v++) {
    o->vnode[v] = -1;
  }
  FastInfo *info = (FastInfo *)malloc(sizeof(FastInfo) * (size_t)(2 * o->T + 2));
  int ninfo = 1;
  info[0].totface = o->T;
  info[0].start = 0;
  info[0].c1 = info[0].c2 = -1;
  limit_ensure_fast(o, nodeinfo, bbc, info, &ninfo, 0);
  o->totnode = 0;
  d_grow_nodes(o, 1);
  memset(&o->nodes[0], 0, sizeof(ONode));
  create_nodes_fast(o, nodeinfo, bbc, info, 0, 0);
  free(info);
  free(nodeinfo);
  free(bbc);
}

/* ------------------------------------------------------------------------- */
/* Creation.                                                                  */
/* ------------------------------------------------------------------------- */

typedef struct {
  uint64_t key;
  int c;
} EdgeRec;
static int cmp_edge(const void *a, const void *b)
{
  const uint64_t x = ((const EdgeRec *)a)->key, y = ((const EdgeRec *)b)->key;
  return (x > y) - (x < y);
}

void orc_dyn_free(OrcDyn *o)
{
  if (!o) {
    return;
  }
  free(o->co);
  free(o->no);
  free(o->mask);
  free(o->vhide);
  free(o->lab);
  free(o->vnode);
  free(o->vcorner);
  free(o->tv);
  free(o->topp);
  free(o->tnode);
  free(o->tslot);
  free(o->fno);
  free(o->fhide);
  free(o->nodes);
  free(o->slab);
  free(o->scount);
  free(o->estamp);
  free(o->remap);
  free(o->remap_ser);
  free(o->wstamp);
  free(o->wlist);
  free(o->best);
  free(o->tie);
  free(o->vst);
  free(o->vslot);
  free(o->cand);
  free(o);
}

/* Vertex normals (mesh_vert_calc_normal: corner angle x face normal over the faces
 * around v, normalized; exact fixed-point sum). */
static void vert_normal(OrcDyn *o, int v)
{
  int cs[DYN_RING];
  int bnd;
  const int n = d_ring(o, v, cs, DYN_RING, &bnd);
  if (n <= 0) {
    return;
  }
  int64_t s[3] = {0, 0, 0};
  const float *pv = o->co + 3 * v;
  for (int i = 0; i < n; i++) {
    const int c = cs[i];
    const float *pp = o->co + 3 * o->tv[c_prev(c)], *pn = o->co + 3 * o->tv[c_next(c)];
    float e1[3] = {pv[0] - pp[0], pv[1] - pp[1], pv[2] - pp[2]};
    float e2[3] = {pn[0] - pv[0], pn[1] - pv[1], pn[2] - pv[2]};
    d_normalize(e1);
    d_normalize(e2);
    const float dot = -(e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]);
    const float fac = (dot <= -1.0f) ? (float)M_PI : ((dot >= 1.0f) ? 0.0f : (float)acos((double)dot));
    const float *fn = o->fno + 3 * (c / 3);
    for (int q = 0; q < 3; q++) {
      s[q] += (int64_t)llrint((double)(fn[q] * fac) * 4294967296.0);
    }
  }
  float r[3];
  for (int q = 0; q < 3; q++) {
    r[q] = (float)((double)s[q] * (1.0 / 4294967296.0));
  }
  d_normalize(r);
  memcpy(o->no + 3 * v, r, sizeof(float) * 3);
}

static void face_normal(OrcDyn *o, int f)
{
  d_tri_normal(o->fno + 3 * f, o->co + 3 * o->tv[3 * f], o->co + 3 * o->tv[3 * f + 1], o->co + 3 * o->tv[3 * f + 2]);
}

OrcDyn *orc_dyn_new(const float *co, int totvert, const int *tris, int tottri, const float *mask,
                    const uint8_t *vhide, const uint8_t *fhide, int vert_cap, int tri_cap, int leaf_limit,
                    int respect_hide, char *err, int errlen)
{
  if (totvert <= 0 || tottri <= 0 || vert_cap < totvert || tri_cap < tottri || leaf_limit > DYN_SLAB / 2) {
    snprintf(err, (size_t)errlen, "bad sizes");
    return NULL;
  }
  OrcDyn *o = (OrcDyn *)calloc(1, sizeof(OrcDyn));
  o->V = totvert;
  o->Vcap = vert_cap;
  o->T = tottri;
  o->Tcap = tri_cap;
  o->leaf_limit = leaf_limit > 0 ? leaf_limit : 100;
  o->respect_hide = respect_hide;
  o->co = (float *)calloc((size_t)vert_cap * 3, sizeof(float));
  o->no = (float *)calloc((size_t)vert_cap * 3, sizeof(float));
  if (mask) {
    o->mask = (float *)calloc((size_t)vert_cap, sizeof(float));
    memcpy(o->mask, mask, sizeof(float) * (size_t)totvert);
  }
  o->vhide = (uint8_t *)calloc((size_t)vert_cap, 1);
  if (vhide) {
    memcpy(o->vhide, vhide, (size_t)totvert);
  }
  o->lab = (uint64_t *)calloc((size_t)vert_cap, sizeof(uint64_t));
  o->vnode = (int *)malloc(sizeof(int) * (size_t)vert_cap);
  o->vcorner = (int *)malloc(sizeof(int) * (size_t)vert_cap);
  o->tv = (int *)malloc(sizeof(int) * 3 * (size_t)tri_cap);
  o->topp = (int *)malloc(sizeof(int) * 3 * (size_t)tri_cap);
  o->tnode = (int *)malloc(sizeof(int) * (size_t)tri_cap);
  o->tslot = (int *)malloc(sizeof(int) * (size_t)tri_cap);
  o->fno = (float *)calloc((size_t)tri_cap * 3, sizeof(float));
  o->fhide = (uint8_t *)calloc((size_t)tri_cap, 1);
  if (fhide) {
    memcpy(o->fhide, fhide, (size_t)tottri);
  }
  o->estamp = (int *)calloc(3 * (size_t)tri_cap, sizeof(int));
  o->remap = (int *)malloc(sizeof(int) * 3 * (size_t)tri_cap);
  o->remap_ser = (int *)calloc(3 * (size_t)tri_cap, sizeof(int));
  o->wstamp = (int *)calloc((size_t)tri_cap, sizeof(int));
  o->wlist = (int *)malloc(sizeof(int) * (size_t)tri_cap);
  o->best = (uint64_t *)calloc((size_t)tri_cap, sizeof(uint64_t));
  o->tie = (uint64_t *)calloc((size_t)tri_cap, sizeof(uint64_t));
  o->vst = (int *)calloc((size_t)vert_cap, sizeof(int));
  o->vslot = (int *)malloc(sizeof(int) * (size_t)vert_cap);
  memcpy(o->co, co, sizeof(float) * 3 * (size_t)totvert);
  for (int v = 0; v < vert_cap; v++) {
    o->lab[v] = (uint64_t)v;
    o->vnode[v] = -1;
    o->vcorner[v] = -1;
  }
  memcpy(o->tv, tris, sizeof(int) * 3 * (size_t)tottri);
  for (int c = 0; c < 3 * tottri; c++) {
    if (o->tv[c] < 0 || o->tv[c] >= totvert) {
      snprintf(err, (size_t)errlen, "corner %d: vertex out of range", c);
      orc_dyn_free(o);
      return NULL;
    }
    o->topp[c] = -1;
  }
  for (int f = 0; f < tottri; f++) {
    const int a = o->tv[3 * f], b = o->tv[3 * f + 1], c = o->tv[3 * f + 2];
    if (a == b || b == c || a == c) {
      snprintf(err, (size_t)errlen, "triangle %d is degenerate", f);
      orc_dyn_free(o);
      return NULL;
    }
  }
  /* opposite corners: the edge opposite c runs tv[next c] -> tv[prev c]; its twin runs
   * the other way.  A directed edge used twice is non-manifold (or misoriented). */
  EdgeRec *er = (EdgeRec *)malloc(sizeof(EdgeRec) * 3 * (size_t)tottri);
  for (int c = 0; c < 3 * tottri; c++) {
    er[c].key = ((uint64_t)(uint32_t)o->tv[c_next(c)] << 32) | (uint32_t)o->tv[c_prev(c)];
    er[c].c = c;
  }
  qsort(er, 3 * (size_t)tottri, sizeof(EdgeRec), cmp_edge);
  for (size_t i = 1; i < 3 * (size_t)tottri; i++) {
    if (er[i].key == er[i - 1].key) {
      snprintf(err, (size_t)errlen, "edge %u-%u used twice in one direction (non-manifold or flipped faces)",
               (unsigned)(er[i].key >> 32), (unsigned)(er[i].key & 0xffffffffu));
      free(er);
      orc_dyn_free(o);
      return NULL;
    }
  }
  for (int c = 0; c < 3 * tottri; c++) {
    const uint64_t tw = ((uint64_t)(uint32_t)o->tv[c_prev(c)] << 32) | (uint32_t)o->tv[c_next(c)];
    size_t lo = 0, hi = 3 * (size_t)tottri;
    while (lo < hi) {
      const size_t m = (lo + hi) / 2;
      if (er[m].key < tw) {
        lo = m + 1;
      }
      else {
        hi = m;
      }
    }
    if (lo < 3 * (size_t)tottri && er[lo].key == tw) {
      o->topp[c] = er[lo].c;
    }
  }
  free(er);
  int *cnt = (int *)calloc((size_t)totvert, sizeof(int));
  for (int c = 0; c < 3 * tottri; c++) {
    cnt[o->tv[c]]++;
    if (o->vcorner[o->tv[c]] < 0) {
      o->vcorner[o->tv[c]] = c;
    }
  }
  for (int v = 0; v < totvert; v++) {
    if (!cnt[v]) {
      continue;
    }
    int cs[DYN_RING], bnd;
    const int n = d_ring(o, v, cs, DYN_RING, &bnd);
    if (n != cnt[v]) {
      snprintf(err, (size_t)errlen, "vertex %d is not manifold (%d faces, %d around one fan)", v, cnt[v], n);
      free(cnt);
      orc_dyn_free(o);
      return NULL;
    }
  }
  free(cnt);
  for (int f = 0; f < tottri; f++) {
    face_normal(o, f);
  }
  for (int v = 0; v < totvert; v++) {
    vert_normal(o, v);
  }
  o->max_len = 0.75f;
  o->min_len = o->max_len * 0.4f;
  o->mode = 3;
  o->view_normal[2] = 1.0f;
  build_tree(o);
  return o;
}

/* BKE_pbvh_bmesh_detail_size_set (pbvh_bmesh.c:2080-2084) */
void orc_dyn_detail_size_set(OrcDyn *o, float detail_size)
{
  o->max_len = detail_size;
  o->min_len = o->max_len * 0.4f;
}

void orc_dyn_topology_set(OrcDyn *o, int mode, const float view_normal[3], int use_frontface, int use_projected)
{
  o->mode = mode;
  memcpy(o->view_normal, view_normal, sizeof(float) * 3);
  o->use_frontface = use_frontface;
  o->use_projected = use_projected;
}

void orc_dyn_set_normals(OrcDyn *o, const float *no) { memcpy(o->no, no, sizeof(float) * 3 * (size_t)o->V); }

/* ------------------------------------------------------------------------- */
/* Search, normals, bounds.                                                   */
/* ------------------------------------------------------------------------- */

static int d_sphere_test(const OBB *bb, const float c[3], float r2)
{
  float t[3];
  for (int i = 0; i < 3; i++) {
    float nearest;
    if (bb->bmin[i] > c[i]) {
      nearest = bb->bmin[i];
    }
    else if (bb->bmax[i] < c[i]) {
      nearest = bb->bmax[i];
    }
    else {
      nearest = c[i];
    }
    t[i] = c[i] - nearest;
  }
  return (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) < r2;
}

/* Leaves whose box meets the sphere, depth first, left first; fully hidden leaves
 * skipped (spec S1). */
int orc_dyn_search(const OrcDyn *o, const float c[3], float radius, int *out, int cap)
{
  const float r2 = radius * radius;
  int *st = (int *)malloc(sizeof(int) * (size_t)(o->totnode + 2));
  int ssz = 0, tot = 0;
  st[ssz++] = 0;
  while (ssz) {
    const int ni = st[--ssz];
    const ONode *n = &o->nodes[ni];
    if ((n->flag & ORC_Leaf) && (n->flag & ORC_FullyHidden)) {
      continue;
    }
    if (!d_sphere_test(&n->vb, c, r2)) {
      continue;
    }
    if (n->flag & ORC_Leaf) {
      if (tot < cap) {
        out[tot] = ni;
      }
      tot++;
      continue;
    }
    st[ssz++] = n->children_offset + 1;
    st[ssz++] = n->children_offset;
  }
  free(st);
  return tot;
}

/* pbvh_bmesh_normals_update (pbvh_bmesh.c:1635-1656): the faces of UpdateNormals
 * leaves, then the vertices of those faces (all face normals first). */
static void d_update_normals(OrcDyn *o)
{
  const int seen_stamp = ++o->stamp;
  int *vl = (int *)malloc(sizeof(int) * ((size_t)o->V + 1));
  int nv = 0;
  int64_t nf = 0;
  for (int ni = 0; ni < o->totnode; ni++) {
    ONode *n = &o->nodes[ni];
    if (!(n->flag & ORC_Leaf) || !(n->flag & ORC_UpdateNormals)) {
      continue;
    }
    const int *sl = o->slab + (size_t)ni * DYN_SLAB;
    for (int i = 0; i < o->scount[ni]; i++) {
      const int f = sl[i];
      if (f < 0) {
        continue;
      }
      face_normal(o, f);
      nf++;
      for (int k = 0; k < 3; k++) {
        const int v = o->tv[3 * f + k];
        if (o->vst[v] != seen_stamp) {
          o->vst[v] = seen_stamp;
          vl[nv++] = v;
        }
      }
    }
    n->flag &= ~ORC_UpdateNormals;
  }
  for (int i = 0; i < nv; i++) {
    vert_normal(o, vl[i]);
  }
  o->stats[10] = nf;
  free(vl);
}

/* UpdateBB leaves: box over their faces' vertices (zeros as +0); orig_vb = vb with
 * UpdateOriginalBB; then every ancestor of an updated leaf (pbvh_flush_bb). */
static void d_update_bounds(OrcDyn *o)
{
  int64_t nr = 0;
  for (int ni = 0; ni < o->totnode; ni++) {
    ONode *n = &o->nodes[ni];
    if (!(n->flag & ORC_Leaf) || !(n->flag & ORC_UpdateBB)) {
      continue;
    }
    OBB vb;
    dbb_reset(&vb);
    const int *sl = o->slab + (size_t)ni * DYN_SLAB;
    for (int i = 0; i < o->scount[ni]; i++) {
      const int f = sl[i];
      if (f < 0) {
        continue;
      }
      for (int k = 0; k < 3; k++) {
        const float *p = o->co + 3 * o->tv[3 * f + k];
        const float q[3] = {p[0] + 0.0f, p[1] + 0.0f, p[2] + 0.0f};
        dbb_expand(&vb, q);
      }
    }
    n->vb = vb;
    if (n->flag & ORC_UpdateOriginalBB) {
      n->orig_vb = n->vb;
    }
    nr++;
  }
  o->stats[11] = nr;
  int *upd = (int *)calloc((size_t)o->totnode, sizeof(int));
  int *st = (int *)malloc(sizeof(int) * 2 * (size_t)(o->totnode + 2));
  int ssz = 0;
  st[ssz++] = 0;
  st[ssz++] = 0;
  while (ssz) {
    const int stage = st[--ssz];
    const int ni = st[--ssz];
    ONode *n = &o->nodes[ni];
    if (n->flag & ORC_Leaf) {
      upd[ni] = n->flag & (ORC_UpdateBB | ORC_UpdateOriginalBB);
      n->flag &= ~(ORC_UpdateBB | ORC_UpdateOriginalBB | ORC_UpdateRedraw);
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
    const int u = upd[n->children_offset] | upd[n->children_offset + 1];
    if (u & ORC_UpdateBB) {
      OBB vb;
      dbb_reset(&vb);
      dbb_expand_bb(&vb, &o->nodes[n->children_offset].vb);
      dbb_expand_bb(&vb, &o->nodes[n->children_offset + 1].vb);
      n->vb = vb;
    }
    if (u & ORC_UpdateOriginalBB) {
      n->orig_vb = n->vb;
    }
    upd[ni] = u;
  }
  free(st);
  free(upd);
}

void orc_dyn_mark(OrcDyn *o, const int *nodes, int n, int flag)
{
  for (int i = 0; i < n; i++) {
    o->nodes[nodes[i]].flag |= flag;
  }
}
void orc_dyn_update_normals(OrcDyn *o) { d_update_normals(o); }
void orc_dyn_update_bounds(OrcDyn *o) { d_update_bounds(o); }

/* ------------------------------------------------------------------------- */
/* Edge operations.                                                           */
/* ------------------------------------------------------------------------- */

static const int TOUCH = ORC_UpdateNormals | ORC_UpdateBB | ORC_UpdateDrawBuffers;

/* Split the edge opposite corner cA (pbvh_bmesh_split_edge, pbvh_bmesh.c:1093-1209)
 * at the new vertex m: each adjacent face (v1, v2, v_opp) in its own winding becomes
 * (v1, m, v_opp), kept in its slot, and (m, v2, v_opp), a new slot in the same leaf.
 * First phase: the edge's own faces are rewritten; each new outer corner holds the
 * old id of its neighbour corner and the old outer corners record where they went
 * (remap, stamped with the round).  The second phase (split_relink) joins the outer
 * edges once every split of the round has moved its corners. */
static int split_edge(OrcDyn *o, int cA, int serial, int round, int stamp, int *new_faces, int *opps, int *outer)
{
  const int cB = o->topp[cA];
  const int A = cA / 3, B = cB >= 0 ? cB / 3 : -1;
  const int p = o->tv[c_next(cA)], q = o->tv[c_prev(cA)], oA = o->tv[cA];
  const int oB = cB >= 0 ? o->tv[cB] : -1;
  const int m = o->V++;
  const int yA = o->topp[c_prev(cA)]; /* across (oA, p) */
  const int xA = o->topp[c_next(cA)]; /* across (q, oA) */
  const int yB = cB >= 0 ? o->topp[c_prev(cB)] : -1; /* across (oB, q) */
  const int xB = cB >= 0 ? o->topp[c_next(cB)] : -1; /* across (p, oB) */
  const int oldA_y = c_prev(cA), oldA_x = c_next(cA);
  const int oldB_y = cB >= 0 ? c_prev(cB) : -1, oldB_x = cB >= 0 ? c_next(cB) : -1;
  const float *pp = o->co + 3 * p, *pq = o->co + 3 * q;
  for (int i = 0; i < 3; i++) {
    o->co[3 * m + i] = 0.5f * (pp[i] + pq[i]);
  }
  float nm[3];
  for (int i = 0; i < 3; i++) {
    nm[i] = 0.5f * (o->no[3 * p + i] + o->no[3 * q + i]);
  }
  d_normalize(nm);
  memcpy(o->no + 3 * m, nm, sizeof(float) * 3);
  if (o->mask) {
    o->mask[m] = 0.5f * (o->mask[p] + o->mask[q]);
  }
  o->vhide[m] = 0;
  o->lab[m] = d_new_label(o->lab[p], o->lab[q], serial, round);
  const int leafA = o->tnode[A], leafB = B >= 0 ? o->tnode[B] : -1;
  o->vnode[m] = (B >= 0 && o->lab[oB] < o->lab[oA]) ? leafB : leafA;
  const int NA = o->T++;
  const int NB = B >= 0 ? o->T++ : -1;
  int *tv = o->tv, *tp = o->topp;
  /* A' = (p, m, oA), NA = (m, q, oA) */
  tv[3 * A] = p;
  tv[3 * A + 1] = m;
  tv[3 * A + 2] = oA;
  tv[3 * NA] = m;
  tv[3 * NA + 1] = q;
  tv[3 * NA + 2] = oA;
  tp[3 * A] = 3 * NA + 1; /* (m, oA) */
  tp[3 * NA + 1] = 3 * A;
  tp[3 * A + 1] = yA; /* (oA, p): old neighbour id, joined by split_relink */
  tp[3 * NA] = xA;    /* (q, oA) */
  o->remap[oldA_y] = 3 * A + 1;
  o->remap_ser[oldA_y] = stamp;
  o->remap[oldA_x] = 3 * NA;
  o->remap_ser[oldA_x] = stamp;
  outer[0] = 3 * A + 1;
  outer[1] = 3 * NA;
  outer[2] = outer[3] = -1;
  if (B >= 0) {
    /* B' = (q, m, oB), NB = (m, p, oB) */
    tv[3 * B] = q;
    tv[3 * B + 1] = m;
    tv[3 * B + 2] = oB;
    tv[3 * NB] = m;
    tv[3 * NB + 1] = p;
    tv[3 * NB + 2] = oB;
    tp[3 * B] = 3 * NB + 1;
    tp[3 * NB + 1] = 3 * B;
    tp[3 * B + 1] = yB; /* (oB, q) */
    tp[3 * NB] = xB;    /* (p, oB) */
    o->remap[oldB_y] = 3 * B + 1;
    o->remap_ser[oldB_y] = stamp;
    o->remap[oldB_x] = 3 * NB;
    o->remap_ser[oldB_x] = stamp;
    outer[2] = 3 * B + 1;
    outer[3] = 3 * NB;
    /* the halves: A' (p, m) with NB (m, p); NA (m, q) with B' (q, m) */
    tp[3 * A + 2] = 3 * NB + 2;
    tp[3 * NB + 2] = 3 * A + 2;
    tp[3 * NA + 2] = 3 * B + 2;
    tp[3 * B + 2] = 3 * NA + 2;
  }
  else {
    tp[3 * A + 2] = -1;
    tp[3 * NA + 2] = -1;
  }
  /* attributes of the new faces follow the split face (face attrs copy: hflag, normal) */
  memcpy(o->fno + 3 * NA, o->fno + 3 * A, sizeof(float) * 3);
  o->fhide[NA] = o->fhide[A];
  slab_append(o, leafA, NA);
  o->nodes[leafA].flag |= TOUCH;
  o->nodes[leafA].flag &= ~ORC_FullyHidden;
  if (B >= 0) {
    memcpy(o->fno + 3 * NB, o->fno + 3 * B, sizeof(float) * 3);
    o->fhide[NB] = o->fhide[B];
    slab_append(o, leafB, NB);
    o->nodes[leafB].flag |= TOUCH;
    o->nodes[leafB].flag &= ~ORC_FullyHidden;
  }
  /* corners of the vertices that sat in rewritten slots */
  const int vs[4] = {p, q, oA, oB};
  for (int i = 0; i < 4; i++) {
    const int v = vs[i];
    if (v < 0) {
      continue;
    }
    const int f = o->vcorner[v] / 3;
    if (f == A) {
      o->vcorner[v] = (v == p) ? 3 * A : (v == q) ? 3 * NA + 1 : 3 * A + 2;
    }
    else if (B >= 0 && f == B) {
      o->vcorner[v] = (v == q) ? 3 * B : (v == p) ? 3 * NB + 1 : 3 * B + 2;
    }
  }
  o->vcorner[m] = 3 * A + 1;
  new_faces[0] = NA;
  new_faces[1] = NB;
  opps[0] = oA;
  opps[1] = oB;
  return m;
}

/* Second phase of a split: each new outer corner k meets its neighbour, moved by its
 * own split this round (remap) or not (then the neighbour's corner is pointed back
 * at k here; no other split touches it). */
static void split_relink(OrcDyn *o, const int *outer, int stamp)
{
  for (int i = 0; i < 4; i++) {
    const int k = outer[i];
    if (k < 0) {
      continue;
    }
    const int y = o->topp[k];
    if (y < 0) {
      continue;
    }
    if (o->remap_ser[y] == stamp) {
      o->topp[k] = o->remap[y];
    }
    else {
      o->topp[y] = k;
    }
  }
}

/* Collapse the edge opposite corner cA (pbvh_bmesh_collapse_edge, pbvh_bmesh.c:
 * 1264-1429): the vertex with the lower mask goes (v1 = the lower label), the other
 * moves to the midpoint; the edge's faces go, their outer neighbours meet. */
static void collapse_edge(OrcDyn *o, int cA, int *touched)
{
  const int cB = o->topp[cA];
  const int A = cA / 3, B = cB >= 0 ? cB / 3 : -1;
  const int p = o->tv[c_next(cA)], q = o->tv[c_prev(cA)], oA = o->tv[cA];
  const int oB = cB >= 0 ? o->tv[cB] : -1;
  const int v1 = o->lab[p] < o->lab[q] ? p : q, v2 = v1 == p ? q : p;
  const float m1 = o->mask ? o->mask[v1] : 0.0f, m2 = o->mask ? o->mask[v2] : 0.0f;
  const int vdel = (m1 < m2) ? v1 : v2, vcon = vdel == v1 ? v2 : v1;
  int cs[DYN_RING], bnd;
  const int n = d_ring(o, vdel, cs, DYN_RING, &bnd);
  const int xA = o->topp[c_next(cA)], yA = o->topp[c_prev(cA)];
  const int xB = cB >= 0 ? o->topp[c_next(cB)] : -1, yB = cB >= 0 ? o->topp[c_prev(cB)] : -1;
  /* leaves touched: the removed faces' and those around both vertices */
  o->nodes[o->tnode[A]].flag |= TOUCH;
  if (B >= 0) {
    o->nodes[o->tnode[B]].flag |= TOUCH;
  }
  if (xA >= 0) {
    o->topp[xA] = yA;
  }
  if (yA >= 0) {
    o->topp[yA] = xA;
  }
  if (B >= 0) {
    if (xB >= 0) {
      o->topp[xB] = yB;
    }
    if (yB >= 0) {
      o->topp[yB] = xB;
    }
  }
  for (int i = 0; i < n; i++) {
    const int f = cs[i] / 3;
    if (f == A || f == B) {
      continue;
    }
    o->tv[cs[i]] = vcon;
  }
  /* corners that sat in the removed faces move to surviving faces */
  const int fA = o->vcorner[oA] / 3;
  if (fA == A || fA == B) {
    o->vcorner[oA] = xA >= 0 ? c_next(xA) : c_prev(yA);
  }
  if (B >= 0) {
    const int fB = o->vcorner[oB] / 3;
    if (fB == A || fB == B) {
      o->vcorner[oB] = xB >= 0 ? c_next(xB) : c_prev(yB);
    }
  }
  const int fc = o->vcorner[vcon] / 3;
  if (fc == A || fc == B) {
    o->vcorner[vcon] = xA >= 0 ? c_prev(xA) : c_next(yA);
  }
  float *pc = o->co + 3 * vcon, *pd = o->co + 3 * vdel;
  for (int i = 0; i < 3; i++) {
    pc[i] = 0.5f * (pc[i] + pd[i]);
  }
  float nn[3];
  for (int i = 0; i < 3; i++) {
    nn[i] = o->no[3 * vcon + i] + o->no[3 * vdel + i];
  }
  d_normalize(nn);
  memcpy(o->no + 3 * vcon, nn, sizeof(float) * 3);
  /* the removed faces leave their leaves */
  const int fs[2] = {A, B};
  for (int i = 0; i < 2; i++) {
    const int f = fs[i];
    if (f < 0) {
      continue;
    }
    o->slab[(size_t)o->tnode[f] * DYN_SLAB + o->tslot[f]] = -1;
    o->tnode[f] = -1;
  }
  o->vnode[vdel] = -1;
  o->vcorner[vdel] = -1;
  touched[0] = vcon;
  touched[1] = oA;
  touched[2] = oB;
}

/* After a round of collapses: a vertex whose owner leaf has none of its faces goes to
 * the smallest-numbered leaf among its faces (both owners marked for the refit); for
 * the kept vertex every leaf around it is marked (pbvh_bmesh.c:1409-1414). */
static void fix_owner(OrcDyn *o, int v, int mark_faces)
{
  if (v < 0 || o->vcorner[v] < 0) {
    return;
  }
  int cs[DYN_RING], bnd;
  const int n = d_ring(o, v, cs, DYN_RING, &bnd);
  int has = 0, best = -1;
  for (int i = 0; i < n; i++) {
    const int leaf = o->tnode[cs[i] / 3];
    if (mark_faces) {
      o->nodes[leaf].flag |= TOUCH;
    }
    has |= leaf == o->vnode[v];
    if (best < 0 || leaf < best) {
      best = leaf;
    }
  }
  if (!has) {
    if (o->vnode[v] >= 0) {
      o->nodes[o->vnode[v]].flag |= ORC_UpdateBB | ORC_UpdateDrawBuffers;
    }
    o->vnode[v] = best;
    if (best >= 0) {
      o->nodes[best].flag |= ORC_UpdateBB | ORC_UpdateDrawBuffers;
    }
  }
}

/* Whether the edge opposite corner c may collapse without making the surface
 * non-manifold (the faces around both vertices are *claimed). */
static int collapse_ok(const OrcDyn *o, int c)
{
  const int cB = o->topp[c];
  const int p = o->tv[c_next(c)], q = o->tv[c_prev(c)], oA = o->tv[c];
  const int oB = cB >= 0 ? o->tv[cB] : -1;
  if (o->vnode[p] < 0 || o->vnode[q] < 0) {
    return 0;
  }
  int np[DYN_RING + 1], nq[DYN_RING + 1];
  const int a = d_neighbors(o, p, np, DYN_RING + 1), b = d_neighbors(o, q, nq, DYN_RING + 1);
  if (a < 0 || b < 0) {
    return 0;
  }
  int cs[DYN_RING], bp, bq;
  d_ring(o, p, cs, DYN_RING, &bp);
  d_ring(o, q, cs, DYN_RING, &bq);
  if (cB >= 0 && bp && bq) {
    return 0; /* an inner edge joining two boundaries */
  }
  /* ears: a removed face with both other edges on the boundary */
  if (o->topp[c_next(c)] < 0 && o->topp[c_prev(c)] < 0) {
    return 0;
  }
  if (cB >= 0 && o->topp[c_next(cB)] < 0 && o->topp[c_prev(cB)] < 0) {
    return 0;
  }
  int common = 0;
  for (int i = 0; i < a; i++) {
    for (int j = 0; j < b; j++) {
      if (np[i] == nq[j]) {
        if (np[i] != oA && np[i] != oB) {
          return 0;
        }
        common++;
      }
    }
  }
  if (common != (cB >= 0 ? 2 : 1)) {
    return 0;
  }
  /* the opposite vertices keep enough edges */
  const int vA = d_valence(o, oA);
  int cs2[DYN_RING], bA;
  d_ring(o, oA, cs2, DYN_RING, &bA);
  if (vA - 1 < (bA ? 2 : 3)) {
    return 0;
  }
  if (oB >= 0) {
    int bB;
    d_ring(o, oB, cs2, DYN_RING, &bB);
    if (d_valence(o, oB) - 1 < (bB ? 2 : 3)) {
      return 0;
    }
  }
  return 1;
}

/* The faces an operation on the edge opposite c owns for its round, each once: a
 * split its two faces (it rewrites only those; its neighbours are joined in a second
 * phase); a collapse every face around both vertices.  Returns the count (-1: a ring
 * too large). */
static int claim_set(const OrcDyn *o, int c, int collapse, int *out, int cap)
{
  int n = 0;
  if (!collapse) {
    const int cB = o->topp[c];
    out[n++] = c / 3;
    if (cB >= 0) {
      out[n++] = cB / 3;
    }
    return n;
  }
  /* faces around p, then those around q without p (the edge's faces once) */
  int cs[DYN_RING], bnd;
  const int p = o->tv[c_next(c)], q = o->tv[c_prev(c)];
  const int k1 = d_ring(o, p, cs, DYN_RING, &bnd);
  if (k1 < 0 || k1 > cap) {
    return -1;
  }
  for (int j = 0; j < k1; j++) {
    out[n++] = cs[j] / 3;
  }
  const int k2 = d_ring(o, q, cs, DYN_RING, &bnd);
  if (k2 < 0 || n + k2 > cap) {
    return -1;
  }
  for (int j = 0; j < k2; j++) {
    const int f = cs[j] / 3;
    if (o->tv[3 * f] != p && o->tv[3 * f + 1] != p && o->tv[3 * f + 2] != p) {
      out[n++] = f;
    }
  }
  return n;
}

/* ------------------------------------------------------------------------- */
/* Update topology in rounds.                                                 */
/* ------------------------------------------------------------------------- */

/* Priority of an edge within a round: its squared length's band (exponent and two
 * mantissa bits: four bands per factor of two), longest first for splits, shortest
 * first for collapses, then a hash of its labels.  Bands keep the order close to the
 * reference's while lengths that vary smoothly across the surface do not chain the
 * rounds one edge after another. */
static uint64_t edge_key(const OrcDyn *o, int c, int collapse)
{
  const uint32_t band = f_bits(edge_len2(o, c)) >> 21;
  const uint64_t h = d_edge_hash33(o->lab[o->tv[c_next(c)]], o->lab[o->tv[c_prev(c)]]);
  return ((uint64_t)(collapse ? 0x3FFu - band : band) << 33) | h;
}

typedef struct {
  OrcDyn *o;
  int ncand, cand_cap;
  int eser; /* this round's stamp (corner registrations) */
} Reg;

static void reg_edge(Reg *r, int c)
{
  OrcDyn *o = r->o;
  if (!edge_insert_ok(o, c)) {
    return;
  }
  const int k = canon(o, c);
  if (o->estamp[k] == r->eser) {
    return;
  }
  o->estamp[k] = r->eser;
  if (r->ncand < r->cand_cap) {
    if (r->ncand >= o->cand_mem) {
      o->cand_mem = o->cand_mem ? 2 * o->cand_mem : 4096;
      o->cand = (int *)realloc(o->cand, sizeof(int) * (size_t)o->cand_mem);
    }
    o->cand[r->ncand] = k;
  }
  r->ncand++;
}

/* long_edge_queue_edge_add_recursive (pbvh_bmesh.c:849-903) from the edge opposite
 * corner c, seen from c's face; iterative, the same set of edges. */
static int reg_even(Reg *r, int c0, float len2_0, float limit0)
{
  OrcDyn *o = r->o;
  int sc[DYN_STACK];
  float sl[DYN_STACK], sm[DYN_STACK];
  int top = 0;
  sc[top] = c0;
  sl[top] = len2_0;
  sm[top] = limit0;
  top++;
  while (top) {
    top--;
    const int c = sc[top];
    const float len2 = sl[top];
    float limit = sm[top];
    const int g = o->topp[c];
    const int entered = g >= 0 ? g / 3 : c / 3;
    if (!front_ok(o, entered)) {
      continue;
    }
    reg_edge(r, c);
    if (g < 0) {
      continue;
    }
    const float cmp = len2 * 1.2f;
    limit = limit * 1.6f;
    const float limit2 = limit * limit;
    const int ks[2] = {c_next(g), c_prev(g)};
    for (int i = 0; i < 2; i++) {
      const float lo = edge_len2(o, ks[i]);
      if (lo > d_max(cmp, limit2)) {
        if (top >= DYN_STACK) {
          return -1;
        }
        sc[top] = ks[i];
        sl[top] = lo;
        sm[top] = limit;
        top++;
      }
    }
  }
  return 0;
}

typedef struct {
  int *w, wn, wcap;
  int *wstamp, wser;
} WSet;

static void w_add(WSet *w, int f)
{
  if (w->wstamp[f] == w->wser) {
    return;
  }
  w->wstamp[f] = w->wser;
  w->w[w->wn++] = f;
}

/* One phase (collapse or subdivide) of DUNE_pbvh_bmesh_update_topology. */
static int phase(OrcDyn *o, int collapse, const float center[3], float radius, WSet *ws)
{
  const float r2 = radius * radius;
  float cproj[3];
  if (o->use_projected) {
    d_project_plane(cproj, center, o->view_normal);
  }
  const float limit = collapse ? o->min_len : o->max_len;
  const float limit2 = limit * limit;
  Reg r;
  r.o = o;
  r.cand_cap = 2 * o->Tcap + 8;
  r.eser = 0;
  uint64_t *best = o->best, *tie = o->tie; /* largest key, largest tied key */
  int *vst = o->vst;
  uint64_t *key = NULL;
  uint8_t *valid = NULL;
  int *touched = NULL;
  int buf_mem = 0;
  int *hv = NULL, *hv2 = NULL, nhv = 0, nhv2 = 0, hv_mem = 0;
  int *set = (int *)malloc(sizeof(int) * 2 * DYN_RING);
  int modified = 0, rounds = 0;
  int64_t ops = 0, ncand_total = 0;
  for (int round = 0; round < DYN_MAX_ROUNDS; round++) {
    r.eser = ++o->stamp;
    r.ncand = 0;
    for (int i = 0; i < ws->wn; i++) {
      const int f = ws->w[i];
      if (o->tnode[f] < 0) {
        continue;
      }
      if (!front_ok(o, f) || !tri_in_range(o, f, center, cproj, r2, o->use_projected)) {
        continue;
      }
      for (int k = 0; k < 3; k++) {
        const int c = 3 * f + k;
        const float l2 = edge_len2(o, c);
        if (collapse) {
          if (l2 < limit2) {
            reg_edge(&r, c);
          }
        }
        else if (l2 > limit2) {
          if (reg_even(&r, c, l2, limit) < 0) {
            o->stats[8] |= 2;
          }
        }
      }
    }
    /* opposite vertices of the last round's splits with more than 8 edges: their
     * long edges (pbvh_bmesh.c:1191-1198) */
    for (int i = 0; i < nhv; i++) {
      const int v = hv[i];
      if (d_valence(o, v) <= 8) {
        continue;
      }
      int cs[DYN_RING], bnd;
      const int n = d_ring(o, v, cs, DYN_RING, &bnd);
      for (int j = 0; j < n; j++) {
        /* the edges (v, next) and, on a boundary's last face, (prev, v) */
        const int e1 = c_prev(cs[j]);
        if (edge_len2(o, e1) > limit2) {
          reg_edge(&r, e1);
        }
        if (bnd && j == 0) {
          const int e2 = c_next(cs[j]);
          if (edge_len2(o, e2) > limit2) {
            reg_edge(&r, e2);
          }
        }
      }
    }
    nhv = 0;
    ncand_total += r.ncand;
    if (r.ncand == 0) {
      break;
    }
    if (r.ncand > r.cand_cap) {
      o->stats[8] |= 1;
      break;
    }
    rounds++;
    if (r.ncand > buf_mem) {
      buf_mem = 2 * r.ncand;
      key = (uint64_t *)realloc(key, sizeof(uint64_t) * (size_t)buf_mem);
      valid = (uint8_t *)realloc(valid, (size_t)buf_mem);
      touched = (int *)realloc(touched, sizeof(int) * 3 * (size_t)buf_mem);
    }
    if (2 * r.ncand > hv_mem) {
      hv_mem = 4 * r.ncand;
      hv = (int *)realloc(hv, sizeof(int) * (size_t)hv_mem);
      hv2 = (int *)realloc(hv2, sizeof(int) * (size_t)hv_mem);
    }
    /* validity and keys */
    int nvalid = 0;
    for (int i = 0; i < r.ncand; i++) {
      const int c = o->cand[i];
      const int p = o->tv[c_next(c)], q = o->tv[c_prev(c)];
      int ok = o->vnode[p] >= 0 && o->vnode[q] >= 0;
      if (ok && collapse) {
        ok = collapse_ok(o, c);
      }
      if (ok && !collapse) {
        const int cB = o->topp[c];
        ok = o->scount[o->tnode[c / 3]] <= DYN_SLAB / 2 && (cB < 0 || o->scount[o->tnode[cB / 3]] <= DYN_SLAB / 2);
      }
      valid[i] = (uint8_t)ok;
      if (!ok) {
        continue;
      }
      nvalid++;
      key[i] = edge_key(o, c, collapse);
    }
    if (!collapse && (o->V + nvalid > o->Vcap || o->T + 2 * nvalid > o->Tcap)) {
      o->stats[8] |= 1;
      break;
    }
    /* claims */
    for (int i = 0; i < r.ncand; i++) {
      if (!valid[i]) {
        continue;
      }
      const int n = claim_set(o, o->cand[i], collapse, set, 2 * DYN_RING);
      for (int j = 0; j < n; j++) {
        best[set[j]] = 0;
        tie[set[j]] = 0;
      }
    }
    for (int i = 0; i < r.ncand; i++) {
      if (!valid[i]) {
        continue;
      }
      const int n = claim_set(o, o->cand[i], collapse, set, 2 * DYN_RING);
      if (n < 0) {
        valid[i] = 0;
        continue;
      }
      for (int j = 0; j < n; j++) {
        if (best[set[j]] == key[i]) {
          if (tie[set[j]] < key[i]) {
            tie[set[j]] = key[i];
          }
        }
        else if (best[set[j]] < key[i]) {
          best[set[j]] = key[i];
        }
      }
    }
    /* winners (independent: any order gives the same result) */
    int *win = (int *)malloc(sizeof(int) * (size_t)(r.ncand + 1));
    int *outer = (int *)malloc(sizeof(int) * 4 * (size_t)(r.ncand + 1));
    int nwin = 0;
    for (int i = 0; i < r.ncand; i++) {
      if (!valid[i]) {
        continue;
      }
      const int n = claim_set(o, o->cand[i], collapse, set, 2 * DYN_RING);
      int own = 1;
      for (int j = 0; j < n && own; j++) {
        own = best[set[j]] == key[i] && tie[set[j]] != key[i];
      }
      if (own) {
        win[nwin++] = o->cand[i];
      }
    }
    int ntouched = 0;
    for (int i = 0; i < nwin; i++) {
      if (collapse) {
        collapse_edge(o, win[i], touched + 3 * ntouched);
        ntouched++;
      }
      else {
        int nf[2], op[2];
        const int cB = o->topp[win[i]];
        split_edge(o, win[i], o->serial, round, r.eser, nf, op, outer + 4 * i);
        w_add(ws, win[i] / 3);
        if (cB >= 0) {
          w_add(ws, cB / 3);
        }
        for (int j = 0; j < 2; j++) {
          if (nf[j] >= 0) {
            w_add(ws, nf[j]);
          }
          if (op[j] >= 0 && vst[op[j]] != r.eser) {
            vst[op[j]] = r.eser;
            hv2[nhv2++] = op[j];
          }
        }
      }
    }
    if (!collapse) {
      for (int i = 0; i < nwin; i++) {
        split_relink(o, outer + 4 * i, r.eser);
      }
    }
    for (int i = 0; i < ntouched; i++) {
      for (int j = 0; j < 3; j++) {
        fix_owner(o, touched[3 * i + j], j == 0);
      }
    }
    free(win);
    free(outer);
    ops += nwin;
    if (nwin) {
      modified = 1;
    }
    else {
      break;
    }
    int *t = hv;
    hv = hv2;
    hv2 = t;
    nhv = nhv2;
    nhv2 = 0;
  }
  o->stats[collapse ? 3 : 2] += ops;
  o->stats[collapse ? 5 : 4] += rounds;
  o->stats[collapse ? 7 : 6] += ncand_total;
  free(key);
  free(valid);
  free(hv);
  free(hv2);
  free(set);
  free(touched);
  return modified;
}

/* DUNE_pbvh_bmesh_update_topology (pbvh_bmesh.c:1931-2005) over the leaves marked
 * ORC_UpdateTopology (not fully hidden). */
int orc_dyn_update_topology(OrcDyn *o, int mode, const float center[3], float radius)
{
  WSet ws;
  ws.wcap = o->Tcap;
  ws.w = o->wlist;
  ws.wstamp = o->wstamp;
  ws.wser = ++o->stamp;
  ws.wn = 0;
  int nl = 0;
  for (int ni = 0; ni < o->totnode; ni++) {
    const ONode *n = &o->nodes[ni];
    if ((n->flag & ORC_Leaf) && (n->flag & ORC_UpdateTopology) && !(n->flag & ORC_FullyHidden)) {
      nl++;
      const int *sl = o->slab + (size_t)ni * DYN_SLAB;
      for (int i = 0; i < o->scount[ni]; i++) {
        if (sl[i] >= 0) {
          w_add(&ws, sl[i]);
        }
      }
    }
  }
  o->stats[9] = nl;
  int modified = 0;
  if (mode & ORC_Collapse) {
    modified |= phase(o, 1, center, radius, &ws);
  }
  if (mode & ORC_Subdivide) {
    modified |= phase(o, 0, center, radius, &ws);
  }
  for (int ni = 0; ni < o->totnode; ni++) {
    o->nodes[ni].flag &= ~ORC_UpdateTopology;
  }
  o->serial++;
  return modified;
}

/* ------------------------------------------------------------------------- */
/* The dab.                                                                   */
/* ------------------------------------------------------------------------- */

static int cmp_u64(const void *a, const void *b)
{
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return (x > y) - (x < y);
}

/* One dab: the topology update of the leaves within 1.25 r (Draw, Inflate, Smooth;
 * mode from orc_dyn_topology_set), then the faces dab's spec over the unique vertices
 * of the leaves within r, the normals of flagged leaves and their bounds.  Affected
 * vertices are returned as labels, ascending. */
int orc_dyn_dab(OrcDyn *o, const OrcDab *dab, uint64_t *out_affected, int cap)
{
  memset(o->stats, 0, sizeof(o->stats));
  const int tool = dab->tool;
  const float r = dab->radius, r2 = r * r;
  const float *c = dab->center;
  const float bstr = orc_bstrength(dab);
  int *leaves = (int *)malloc(sizeof(int) * (size_t)(o->totnode + 1));
  if (tool != ORC_TOOL_GRAB && o->mode) {
    const float rt = r * 1.25f;
    const int nt = orc_dyn_search(o, c, rt, leaves, o->totnode + 1);
    for (int i = 0; i < nt; i++) {
      o->nodes[leaves[i]].flag |= ORC_UpdateTopology | ORC_UpdateNormals | ORC_UpdateBB | ORC_UpdateOriginalBB |
                                  ORC_UpdateDrawBuffers | ORC_UpdateRedraw;
    }
    if (nt) {
      orc_dyn_update_topology(o, o->mode, c, r);
    }
    free(leaves);
    leaves = (int *)malloc(sizeof(int) * (size_t)(o->totnode + 1));
  }
  const int nl = orc_dyn_search(o, c, r, leaves, o->totnode + 1);
  o->stats[0] = nl;
  /* the gathered leaves' unique vertices, each once */
  int *cand = (int *)malloc(sizeof(int) * ((size_t)o->V + 1));
  const int seen_stamp = ++o->stamp;
  int ncand = 0;
  for (int i = 0; i < nl; i++) {
    const int ni = leaves[i];
    const int *sl = o->slab + (size_t)ni * DYN_SLAB;
    for (int j = 0; j < o->scount[ni]; j++) {
      const int f = sl[j];
      if (f < 0) {
        continue;
      }
      for (int k = 0; k < 3; k++) {
        const int v = o->tv[3 * f + k];
        if (o->vnode[v] == ni && o->vst[v] != seen_stamp) {
          o->vst[v] = seen_stamp;
          cand[ncand++] = v;
        }
      }
    }
  }
  uint8_t *aff = (uint8_t *)calloc((size_t)ncand + 1, 1);
  float *fac = (float *)malloc(sizeof(float) * ((size_t)ncand + 1));
  float *newco = (float *)malloc(sizeof(float) * 3 * ((size_t)ncand + 1));
  int naff = 0;
  for (int i = 0; i < ncand; i++) {
    const int v = cand[i];
    if (o->respect_hide && o->vhide[v]) {
      continue;
    }
    const float *p = o->co + 3 * v;
    const float d[3] = {c[0] - p[0], c[1] - p[1], c[2] - p[2]};
    const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (d2 <= r2) {
      aff[i] = 1;
      naff++;
      fac[i] = orc_falloff(dab, d2) * (1.0f - (o->mask ? o->mask[v] : 0.0f));
    }
  }
  float offset[3] = {0.0f, 0.0f, 0.0f};
  if (tool == ORC_TOOL_DRAW) {
    const float nr = r * dab->normal_radius_factor, nr2 = nr * nr;
    int64_t s[3] = {0, 0, 0};
    for (int i = 0; i < ncand; i++) {
      const int v = cand[i];
      if (o->respect_hide && o->vhide[v]) {
        continue;
      }
      const float *p = o->co + 3 * v;
      const float d[3] = {c[0] - p[0], c[1] - p[1], c[2] - p[2]};
      if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= nr2) {
        for (int q = 0; q < 3; q++) {
          s[q] += (int64_t)llrint((double)o->no[3 * v + q] * 4294967296.0);
        }
      }
    }
    float an[3];
    for (int q = 0; q < 3; q++) {
      an[q] = (float)((double)s[q] * (1.0 / 4294967296.0));
    }
    d_normalize(an);
    for (int q = 0; q < 3; q++) {
      offset[q] = (an[q] * r) * bstr;
    }
  }
  if (tool == ORC_TOOL_SMOOTH) {
    /* spec S8: k Jacobi steps, neighbour mean as an exact fixed-point sum */
    const int iters = dab->smooth_iters > 0 ? dab->smooth_iters : 1;
    const int aff_stamp = ++o->stamp;
    float *prev = (float *)malloc(sizeof(float) * 3 * ((size_t)ncand + 1));
    for (int i = 0; i < ncand; i++) {
      if (aff[i]) {
        o->vst[cand[i]] = aff_stamp;
        o->vslot[cand[i]] = i;
      }
      memcpy(prev + 3 * i, o->co + 3 * cand[i], sizeof(float) * 3);
    }
    int nb[DYN_RING + 1];
    for (int it = 0; it < iters; it++) {
      for (int i = 0; i < ncand; i++) {
        if (!aff[i]) {
          continue;
        }
        const int nn = d_neighbors(o, cand[i], nb, DYN_RING + 1);
        int64_t s[3] = {0, 0, 0};
        for (int q = 0; q < nn; q++) {
          const int u = nb[q];
          const float *pu = o->vst[u] == aff_stamp ? prev + 3 * o->vslot[u] : o->co + 3 * u;
          for (int a = 0; a < 3; a++) {
            s[a] += (int64_t)llrint((double)pu[a] * 4294967296.0);
          }
        }
        const float *self = prev + 3 * i;
        float *dst = newco + 3 * i;
        if (nn > 0) {
          const float fade = bstr * fac[i];
          for (int a = 0; a < 3; a++) {
            const float avg = (float)(((double)s[a] * (1.0 / 4294967296.0)) / (double)nn);
            dst[a] = self[a] + (avg - self[a]) * fade;
          }
        }
        else {
          memcpy(dst, self, sizeof(float) * 3);
        }
      }
      for (int i = 0; i < ncand; i++) {
        if (aff[i]) {
          memcpy(prev + 3 * i, newco + 3 * i, sizeof(float) * 3);
        }
      }
    }
    free(prev);
  }
  else {
    for (int i = 0; i < ncand; i++) {
      if (!aff[i]) {
        continue;
      }
      const float *p = o->co + 3 * cand[i];
      float *dst = newco + 3 * i;
      const float f = fac[i];
      if (tool == ORC_TOOL_DRAW) {
        for (int q = 0; q < 3; q++) {
          dst[q] = p[q] + offset[q] * f;
        }
      }
      else if (tool == ORC_TOOL_INFLATE) {
        const float s = (bstr * f) * r;
        for (int q = 0; q < 3; q++) {
          dst[q] = p[q] + o->no[3 * cand[i] + q] * s;
        }
      }
      else {
        const float fade = bstr * f;
        for (int q = 0; q < 3; q++) {
          dst[q] = p[q] + dab->grab_delta[q] * fade;
        }
      }
    }
  }
  for (int i = 0; i < ncand; i++) {
    if (aff[i]) {
      memcpy(o->co + 3 * cand[i], newco + 3 * i, sizeof(float) * 3);
    }
  }
  for (int i = 0; i < nl; i++) {
    o->nodes[leaves[i]].flag |=
        ORC_UpdateNormals | ORC_UpdateBB | ORC_UpdateOriginalBB | ORC_UpdateDrawBuffers | ORC_UpdateRedraw;
  }
  d_update_normals(o);
  d_update_bounds(o);
  o->stats[1] = naff;
  if (out_affected) {
    int w = 0;
    for (int i = 0; i < ncand && w < cap; i++) {
      if (aff[i]) {
        out_affected[w++] = o->lab[cand[i]];
      }
    }
    qsort(out_affected, (size_t)w, sizeof(uint64_t), cmp_u64);
  }
  free(leaves);
  free(cand);
  free(aff);
  free(fac);
  free(newco);
  return naff;
}

/* ------------------------------------------------------------------------- */
/* After a stroke: leaf lists compacted, leaves above the limit split.        */
/* ------------------------------------------------------------------------- */

/* Faces of a leaf in a canonical order: by their labels, starting from the corner
 * with the smallest label (rotation kept). */
typedef struct {
  uint64_t k[3];
  int f;
} FaceKey;
static int cmp_fk(const void *a, const void *b)
{
  const FaceKey *x = (const FaceKey *)a, *y = (const FaceKey *)b;
  for (int i = 0; i < 3; i++) {
    if (x->k[i] != y->k[i]) {
      return x->k[i] < y->k[i] ? -1 : 1;
    }
  }
  return 0;
}
static void face_key(const OrcDyn *o, int f, FaceKey *fk)
{
  int s = 0;
  for (int k = 1; k < 3; k++) {
    if (o->lab[o->tv[3 * f + k]] < o->lab[o->tv[3 * f + s]]) {
      s = k;
    }
  }
  for (int k = 0; k < 3; k++) {
    fk->k[k] = o->lab[o->tv[3 * f + (s + k) % 3]];
  }
  fk->f = f;
}

/* pbvh_bmesh_node_finalize (pbvh_bmesh.c:179-235) */
static void node_finalize(OrcDyn *o, int ni)
{
  ONode *n = &o->nodes[ni];
  dbb_reset(&n->vb);
  int has_visible = 0;
  const int *sl = o->slab + (size_t)ni * DYN_SLAB;
  for (int i = 0; i < o->scount[ni]; i++) {
    const int f = sl[i];
    o->tnode[f] = ni;
    o->tslot[f] = i;
    for (int k = 0; k < 3; k++) {
      const int v = o->tv[3 * f + k];
      if (o->vnode[v] < 0) {
        o->vnode[v] = ni;
      }
      dbb_expand(&n->vb, o->co + 3 * v);
    }
    if (!o->fhide[f]) {
      has_visible = 1;
    }
  }
  n->orig_vb = n->vb;
  n->flag |= ORC_UpdateDrawBuffers | ORC_UpdateNormals;
  if (has_visible) {
    n->flag &= ~ORC_FullyHidden;
  }
  else {
    n->flag |= ORC_FullyHidden;
  }
}

/* pbvh_bmesh_node_split (pbvh_bmesh.c:237-361); faces taken in the canonical order */
static void node_split(OrcDyn *o, const OBBC *bbc, int ni)
{
  if (o->scount[ni] <= o->leaf_limit) {
    node_finalize(o, ni);
    return;
  }
  OBB cb;
  dbb_reset(&cb);
  const int cnt = o->scount[ni];
  int *fl = (int *)malloc(sizeof(int) * (size_t)cnt);
  memcpy(fl, o->slab + (size_t)ni * DYN_SLAB, sizeof(int) * (size_t)cnt);
  for (int i = 0; i < cnt; i++) {
    dbb_expand(&cb, bbc[fl[i]].bcentroid);
  }
  const int axis = dbb_widest_axis(&cb);
  const float mid = (cb.bmax[axis] + cb.bmin[axis]) * 0.5f;
  const int children = o->totnode;
  o->nodes[ni].children_offset = children;
  d_grow_nodes(o, o->totnode + 2);
  for (int k = 0; k < 2; k++) {
    memset(&o->nodes[children + k], 0, sizeof(ONode));
    o->nodes[children + k].flag = ORC_Leaf;
    o->nodes[children + k].children_offset = -1;
    o->scount[children + k] = 0;
  }
  int *s1 = o->slab + (size_t)children * DYN_SLAB, *s2 = o->slab + (size_t)(children + 1) * DYN_SLAB;
  for (int i = 0; i < cnt; i++) {
    const int f = fl[i];
    if (bbc[f].bcentroid[axis] < mid) {
      s1[o->scount[children]++] = f;
    }
    else {
      s2[o->scount[children + 1]++] = f;
    }
  }
  if (o->scount[children] == 0 || o->scount[children + 1] == 0) {
    /* move the first face of the other child (GSet iteration's first) */
    const int e = o->scount[children] == 0 ? children : children + 1, other = e == children ? children + 1 : children;
    int *se = o->slab + (size_t)e * DYN_SLAB, *so = o->slab + (size_t)other * DYN_SLAB;
    se[o->scount[e]++] = so[0];
    memmove(so, so + 1, sizeof(int) * (size_t)(o->scount[other] - 1));
    o->scount[other]--;
  }
  /* the node's unique vertices become unclaimed */
  for (int i = 0; i < cnt; i++) {
    for (int k = 0; k < 3; k++) {
      const int v = o->tv[3 * fl[i] + k];
      if (o->vnode[v] == ni) {
        o->vnode[v] = -1;
      }
    }
  }
  o->scount[ni] = 0;
  o->nodes[ni].flag &= ~ORC_Leaf;
  free(fl);
  node_split(o, bbc, children);
  node_split(o, bbc, children + 1);
  ONode *n = &o->nodes[ni];
  dbb_reset(&n->vb);
  dbb_expand_bb(&n->vb, &o->nodes[n->children_offset].vb);
  dbb_expand_bb(&n->vb, &o->nodes[n->children_offset + 1].vb);
  n->orig_vb = n->vb;
}

/* DUNE_pbvh_bmesh_after_stroke (pbvh_bmesh.c:2065-2078): every leaf's list compacted
 * (dead slots dropped, canonical order), leaves above the limit split. */
void orc_dyn_after_stroke(OrcDyn *o)
{
  OBBC *bbc = (OBBC *)malloc(sizeof(OBBC) * (size_t)(o->T + 1));
  FaceKey *fk = (FaceKey *)malloc(sizeof(FaceKey) * DYN_SLAB);
  for (int ni = 0; ni < o->totnode; ni++) {
    if (!(o->nodes[ni].flag & ORC_Leaf)) {
      continue;
    }
    int *sl = o->slab + (size_t)ni * DYN_SLAB;
    int n = 0;
    for (int i = 0; i < o->scount[ni]; i++) {
      if (sl[i] >= 0) {
        face_key(o, sl[i], &fk[n++]);
      }
    }
    qsort(fk, (size_t)n, sizeof(FaceKey), cmp_fk);
    for (int i = 0; i < n; i++) {
      sl[i] = fk[i].f;
      o->tslot[fk[i].f] = i;
    }
    o->scount[ni] = n;
  }
  free(fk);
  for (int ni = 0; ni < o->totnode; ni++) {
    if (!(o->nodes[ni].flag & ORC_Leaf) || o->scount[ni] <= o->leaf_limit) {
      continue;
    }
    const int *sl = o->slab + (size_t)ni * DYN_SLAB;
    for (int i = 0; i < o->scount[ni]; i++) {
      face_bbc(o, sl[i], &bbc[sl[i]]);
    }
    node_split(o, bbc, ni);
  }
  free(bbc);
}

/* ------------------------------------------------------------------------- */
/* The reference's sequential subdivision (one heap), for pinning the rounds. */
/* ------------------------------------------------------------------------- */

typedef struct {
  float pri;
  int64_t seq;
  uint64_t la, lb;
} HeapItem;

typedef struct {
  HeapItem *a;
  int n, cap;
  int64_t seq;
} Heap;

static int heap_less(const HeapItem *x, const HeapItem *y)
{
  if (x->pri != y->pri) {
    return x->pri < y->pri;
  }
  return x->seq < y->seq;
}
static void heap_push(Heap *h, float pri, uint64_t la, uint64_t lb)
{
  if (h->n == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 256;
    h->a = (HeapItem *)realloc(h->a, sizeof(HeapItem) * (size_t)h->cap);
  }
  HeapItem it = {pri, h->seq++, la, lb};
  int i = h->n++;
  while (i > 0) {
    const int pa = (i - 1) / 2;
    if (!heap_less(&it, &h->a[pa])) {
      break;
    }
    h->a[i] = h->a[pa];
    i = pa;
  }
  h->a[i] = it;
}
static HeapItem heap_pop(Heap *h)
{
  HeapItem top = h->a[0];
  HeapItem last = h->a[--h->n];
  int i = 0;
  for (;;) {
    int c = 2 * i + 1;
    if (c >= h->n) {
      break;
    }
    if (c + 1 < h->n && heap_less(&h->a[c + 1], &h->a[c])) {
      c++;
    }
    if (!heap_less(&h->a[c], &last)) {
      break;
    }
    h->a[i] = h->a[c];
    i = c;
  }
  if (h->n) {
    h->a[i] = last;
  }
  return top;
}

typedef struct {
  OrcDyn *o;
  Heap heap;
  uint64_t *qa, *qb; /* edges in the queue (EDGE_QUEUE tag), as label pairs */
  int nq, qcap;
  float limit2, limit, r2;
  const float *center;
  float cproj[3];
} Seq;

static int seq_tagged(const Seq *s, uint64_t a, uint64_t b)
{
  for (int i = 0; i < s->nq; i++) {
    if ((s->qa[i] == a && s->qb[i] == b) || (s->qa[i] == b && s->qb[i] == a)) {
      return 1;
    }
  }
  return 0;
}
static void seq_untag(Seq *s, uint64_t a, uint64_t b)
{
  for (int i = 0; i < s->nq; i++) {
    if ((s->qa[i] == a && s->qb[i] == b) || (s->qa[i] == b && s->qb[i] == a)) {
      s->qa[i] = s->qa[s->nq - 1];
      s->qb[i] = s->qb[s->nq - 1];
      s->nq--;
      return;
    }
  }
}
static void seq_insert(Seq *s, int c, float pri)
{
  OrcDyn *o = s->o;
  if (!edge_insert_ok(o, c)) {
    return;
  }
  const uint64_t a = o->lab[o->tv[c_next(c)]], b = o->lab[o->tv[c_prev(c)]];
  if (s->nq == s->qcap) {
    s->qcap = s->qcap ? 2 * s->qcap : 256;
    s->qa = (uint64_t *)realloc(s->qa, sizeof(uint64_t) * (size_t)s->qcap);
    s->qb = (uint64_t *)realloc(s->qb, sizeof(uint64_t) * (size_t)s->qcap);
  }
  s->qa[s->nq] = a;
  s->qb[s->nq] = b;
  s->nq++;
  heap_push(&s->heap, pri, a, b);
}
static void seq_recursive(Seq *s, int c, float len2, float limit)
{
  OrcDyn *o = s->o;
  const int g = o->topp[c];
  const int entered = g >= 0 ? g / 3 : c / 3;
  if (!front_ok(o, entered)) {
    return;
  }
  if (!seq_tagged(s, o->lab[o->tv[c_next(c)]], o->lab[o->tv[c_prev(c)]])) {
    seq_insert(s, c, -len2);
  }
  if (g < 0) {
    return;
  }
  const float cmp = len2 * 1.2f;
  limit = limit * 1.6f;
  const float limit2 = limit * limit;
  const int ks[2] = {c_next(g), c_prev(g)};
  for (int i = 0; i < 2; i++) {
    const float lo = edge_len2(o, ks[i]);
    if (lo > d_max(cmp, limit2)) {
      seq_recursive(s, ks[i], lo, limit);
    }
  }
}
static void seq_face_add(Seq *s, int f)
{
  OrcDyn *o = s->o;
  if (!front_ok(o, f) || !tri_in_range(o, f, s->center, s->cproj, s->r2, o->use_projected)) {
    return;
  }
  for (int k = 0; k < 3; k++) {
    const float l2 = edge_len2(o, 3 * f + k);
    if (l2 > s->limit2) {
      seq_recursive(s, 3 * f + k, l2, s->limit);
    }
  }
}
static int find_vert(const OrcDyn *o, uint64_t lab)
{
  for (int v = 0; v < o->V; v++) {
    if (o->vcorner[v] >= 0 && o->lab[v] == lab) {
      return v;
    }
  }
  return -1;
}

/* pbvh_bmesh_subdivide_long_edges with long_edge_queue_create (pbvh_bmesh.c:
 * 978-1039, :1211-1262) over the marked leaves: one heap, longest edge first (ties
 * in insertion order), faces in leaf-list order.  New labels take round = the split's
 * ordinal. */
int orc_dyn_subdivide_sequential(OrcDyn *o, const float center[3], float radius)
{
  Seq s;
  memset(&s, 0, sizeof(s));
  s.o = o;
  s.limit = o->max_len;
  s.limit2 = o->max_len * o->max_len;
  s.r2 = radius * radius;
  s.center = center;
  if (o->use_projected) {
    d_project_plane(s.cproj, center, o->view_normal);
  }
  for (int ni = 0; ni < o->totnode; ni++) {
    const ONode *n = &o->nodes[ni];
    if ((n->flag & ORC_Leaf) && (n->flag & ORC_UpdateTopology) && !(n->flag & ORC_FullyHidden)) {
      const int *sl = o->slab + (size_t)ni * DYN_SLAB;
      for (int i = 0; i < o->scount[ni]; i++) {
        if (sl[i] >= 0) {
          seq_face_add(&s, sl[i]);
        }
      }
    }
  }
  int nsplit = 0;
  while (s.heap.n) {
    const HeapItem it = heap_pop(&s.heap);
    seq_untag(&s, it.la, it.lb);
    const int v1 = find_vert(o, it.la), v2 = find_vert(o, it.lb);
    if (v1 < 0 || v2 < 0) {
      continue;
    }
    int cs[DYN_RING], bnd, ce = -1;
    const int n = d_ring(o, v1, cs, DYN_RING, &bnd);
    for (int i = 0; i < n && ce < 0; i++) {
      if (o->tv[c_next(cs[i])] == v2) {
        ce = c_prev(cs[i]);
      }
      else if (o->tv[c_prev(cs[i])] == v2) {
        ce = c_next(cs[i]);
      }
    }
    if (ce < 0 || o->vnode[v1] < 0 || o->vnode[v2] < 0) {
      continue;
    }
    int nf[2], op[2];
    const int cB = o->topp[ce];
    const int A = ce / 3, B = cB >= 0 ? cB / 3 : -1;
    {
      int outer[4];
      const int stamp = ++o->stamp;
      split_edge(o, ce, o->serial, 1000000 + nsplit, stamp, nf, op, outer);
      split_relink(o, outer, stamp);
    }
    nsplit++;
    const int fs[4] = {A, nf[0], B, nf[1]};
    for (int i = 0; i < 4; i++) {
      if (fs[i] >= 0) {
        seq_face_add(&s, fs[i]);
      }
    }
    for (int j = 0; j < 2; j++) {
      if (op[j] >= 0 && d_valence(o, op[j]) > 8) {
        int cs2[DYN_RING], b2;
        const int k = d_ring(o, op[j], cs2, DYN_RING, &b2);
        for (int i = 0; i < k; i++) {
          const int e1 = c_prev(cs2[i]);
          const float l2 = edge_len2(o, e1);
          if (l2 > s.limit2 && !seq_tagged(&s, o->lab[o->tv[c_next(e1)]], o->lab[o->tv[c_prev(e1)]])) {
            seq_insert(&s, e1, -l2);
          }
          if (b2 && i == 0) {
            const int e2 = c_next(cs2[i]);
            const float l3 = edge_len2(o, e2);
            if (l3 > s.limit2 && !seq_tagged(&s, o->lab[o->tv[c_next(e2)]], o->lab[o->tv[c_prev(e2)]])) {
              seq_insert(&s, e2, -l3);
            }
          }
        }
      }
    }
  }
  for (int ni = 0; ni < o->totnode; ni++) {
    o->nodes[ni].flag &= ~ORC_UpdateTopology;
  }
  o->serial++;
  free(s.heap.a);
  free(s.qa);
  free(s.qb);
  return nsplit;
}

/* ------------------------------------------------------------------------- */
/* Accessors.                                                                 */
/* ------------------------------------------------------------------------- */

void orc_dyn_counts(const OrcDyn *o, int out[6])
{
  int nv = 0, nt = 0;
  for (int v = 0; v < o->V; v++) {
    nv += o->vcorner[v] >= 0;
  }
  for (int f = 0; f < o->T; f++) {
    nt += o->tnode[f] >= 0;
  }
  out[0] = nv;
  out[1] = nt;
  out[2] = o->totnode;
  out[3] = o->V;
  out[4] = o->T;
  out[5] = DYN_SLAB;
}

/* Live vertices in slot order: label, position, normal, mask, owner node. */
void orc_dyn_export_verts(const OrcDyn *o, uint64_t *lab, float *co, float *no, float *mask, int *node)
{
  int w = 0;
  for (int v = 0; v < o->V; v++) {
    if (o->vcorner[v] < 0) {
      continue;
    }
    lab[w] = o->lab[v];
    memcpy(co + 3 * w, o->co + 3 * v, sizeof(float) * 3);
    memcpy(no + 3 * w, o->no + 3 * v, sizeof(float) * 3);
    if (mask) {
      mask[w] = o->mask ? o->mask[v] : 0.0f;
    }
    node[w] = o->vnode[v];
    w++;
  }
}

/* Live faces in slot order: corner labels, leaf, face normal, hidden. */
void orc_dyn_export_faces(const OrcDyn *o, uint64_t *labs, int *node, float *fno, uint8_t *hide)
{
  int w = 0;
  for (int f = 0; f < o->T; f++) {
    if (o->tnode[f] < 0) {
      continue;
    }
    for (int k = 0; k < 3; k++) {
      labs[3 * w + k] = o->lab[o->tv[3 * f + k]];
    }
    node[w] = o->tnode[f];
    memcpy(fno + 3 * w, o->fno + 3 * f, sizeof(float) * 3);
    hide[w] = o->fhide[f];
    w++;
  }
}

void orc_dyn_node(const OrcDyn *o, int ni, int out[3], float vb[6], float orig_vb[6])
{
  const ONode *n = &o->nodes[ni];
  out[0] = n->flag;
  out[1] = n->children_offset;
  int nf = 0;
  if (n->flag & ORC_Leaf) {
    for (int i = 0; i < o->scount[ni]; i++) {
      nf += o->slab[(size_t)ni * DYN_SLAB + i] >= 0;
    }
  }
  out[2] = nf;
  memcpy(vb, n->vb.bmin, sizeof(float) * 3);
  memcpy(vb + 3, n->vb.bmax, sizeof(float) * 3);
  memcpy(orig_vb, n->orig_vb.bmin, sizeof(float) * 3);
  memcpy(orig_vb + 3, n->orig_vb.bmax, sizeof(float) * 3);
}

void orc_dyn_last_stats(const OrcDyn *o, int64_t out[12]) { memcpy(out, o->stats, sizeof(o->stats)); }

/* Consistency of the corner table and the leaf lists: 0 when sound, else a code. */
int orc_dyn_check(const OrcDyn *o)
{
  for (int f = 0; f < o->T; f++) {
    if (o->tnode[f] < 0) {
      continue;
    }
    for (int k = 0; k < 3; k++) {
      const int c = 3 * f + k, x = o->topp[c];
      if (x >= 0) {
        if (o->topp[x] != c || o->tnode[x / 3] < 0) {
          return 1;
        }
        if (o->tv[c_next(c)] != o->tv[c_prev(x)] || o->tv[c_prev(c)] != o->tv[c_next(x)]) {
          return 2;
        }
      }
      if (o->vcorner[o->tv[c]] < 0) {
        return 3;
      }
    }
    if (o->slab[(size_t)o->tnode[f] * DYN_SLAB + o->tslot[f]] != f) {
      return 4;
    }
  }
  int *cnt = (int *)calloc((size_t)o->V + 1, sizeof(int));
  for (int f = 0; f < o->T; f++) {
    if (o->tnode[f] >= 0) {
      for (int k = 0; k < 3; k++) {
        cnt[o->tv[3 * f + k]]++;
      }
    }
  }
  int rc = 0;
  for (int v = 0; v < o->V && !rc; v++) {
    if (o->vcorner[v] < 0) {
      rc = cnt[v] ? 5 : 0;
      continue;
    }
    if (o->tv[o->vcorner[v]] != v || o->tnode[o->vcorner[v] / 3] < 0) {
      rc = 6;
      break;
    }
    int cs[DYN_RING], bnd;
    if (d_ring(o, v, cs, DYN_RING, &bnd) != cnt[v]) {
      rc = 7;
      break;
    }
    int owned = 0;
    for (int i = 0; i < cnt[v]; i++) {
      owned |= o->tnode[cs[i] / 3] == o->vnode[v];
    }
    if (!owned) {
      rc = 8;
    }
  }
  free(cnt);
  return rc;
}
