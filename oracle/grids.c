/*
 * Multires grids (PBVH_GRIDS) in the CPU oracle -- TEST INFRASTRUCTURE ONLY (see
 * oracle.h).  Plain C, FP32 without contraction, restating:
 *
 *   grids, coarse edge / vertex neighbourhoods   kernel/intern/subdiv_ccg.c:312-530
 *   neighbour coordinates                        subdiv_ccg.c:1365-1909
 *   grid normals, seam averaging                 subdiv_ccg.c:670-874, :876-1190
 *   tree over grids                              pbvh.c:152-214, :322-329, :331-361,
 *                                                :383-407, :409-440, :454-507, :598-643
 *   normals update on grids                      pbvh.c:4559-4587, :1605-1640
 *   bounds refit                                 pbvh.c:108-128, :1369-1421
 *   the dab                                      the faces dab's spec (oracle.c orc_dab)
 *                                                over grid elements; the averages of the
 *                                                normals update also move elements of
 *                                                leaves the dab did not gather, and those
 *                                                leaves are refit too
 *
 * Coarse topology conventions (OpenSubdiv's refiner in the reference, absent there):
 * edges numbered in order of first use by the polygons, each with the vertex order of
 * that first use; a vertex's edges by ascending edge number.  The paint_is_grid_face_hidden
 * rule (absent from the reference) is upstream Blender's: a quad with a hidden corner.
 *
 * Elements are held by grid id: element (g, x, y) is g * grid_area + y * grid_size + x.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_int.h"

typedef struct {
  int g, x, y;
} GCoord;

struct OrcGrids {
  int gs, ga, G, F, Vc, Ed;
  int has_normal, has_mask, respect_hide, leaf_limit;
  float *co, *no, *mask;
  uint8_t *hide;
  int *grid_mat;
  int *face_start, *face_n, *grid_face;
  int *corner_vert, *corner_edge; /* per grid (= per face corner) */
  int *ev0, *ev1;                 /* edge vertices */
  int *edge_foff;                 /* [Ed + 1] */
  GCoord *edge_bc;                /* 2 * gs per (edge, adjacent face) */
  int *vert_foff;                 /* [Vc + 1] */
  GCoord *vert_cc;
  int *vedge_off, *vedges;
  ONode *nodes;
  int totnode, node_mem;
  int *prim_indices;
  int64_t stats[4];
};

static inline float g_min(float a, float b) { return (a < b) ? a : b; }
static inline float g_max(float a, float b) { return (a > b) ? a : b; }

static inline int el(const OrcGrids *o, int g, int x, int y) { return g * o->ga + y * o->gs + x; }
static inline int el_c(const OrcGrids *o, GCoord c) { return el(o, c.g, c.x, c.y); }
static inline GCoord gc(int g, int x, int y)
{
  GCoord c = {g, x, y};
  return c;
}

static void gbb_reset(OBB *b)
{
  b->bmin[0] = b->bmin[1] = b->bmin[2] = FLT_MAX;
  b->bmax[0] = b->bmax[1] = b->bmax[2] = -FLT_MAX;
}
static void gbb_expand(OBB *b, const float *co)
{
  for (int i = 0; i < 3; i++) {
    b->bmin[i] = g_min(b->bmin[i], co[i]);
    b->bmax[i] = g_max(b->bmax[i], co[i]);
  }
}
static void gbb_expand_bb(OBB *b, const OBB *c)
{
  for (int i = 0; i < 3; i++) {
    b->bmin[i] = g_min(b->bmin[i], c->bmin[i]);
    b->bmax[i] = g_max(b->bmax[i], c->bmax[i]);
  }
}

/* normalize_v3, math_vector_inline.c:1165-1180 */
static float g_normalize(float n[3])
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

/* normal_quad_v3, math_geom.cc:48-67 */
static void g_quad_normal(float n[3], const float *v1, const float *v2, const float *v3, const float *v4)
{
  float n1[3], n2[3];
  n1[0] = v1[0] - v3[0];
  n1[1] = v1[1] - v3[1];
  n1[2] = v1[2] - v3[2];
  n2[0] = v2[0] - v4[0];
  n2[1] = v2[1] - v4[1];
  n2[2] = v2[2] - v4[2];
  n[0] = n1[1] * n2[2] - n1[2] * n2[1];
  n[1] = n1[2] * n2[0] - n1[0] * n2[2];
  n[2] = n1[0] * n2[1] - n1[1] * n2[0];
  g_normalize(n);
}

/* ------------------------------------------------------------------------- */
/* Topology: subdiv_ccg.c:312-530.                                           */
/* ------------------------------------------------------------------------- */

static int find_edge(const OrcGrids *o, const int *head, const int *next, int a, int b)
{
  for (int e = head[a < b ? a : b]; e >= 0; e = next[e]) {
    if ((o->ev0[e] == a && o->ev1[e] == b) || (o->ev0[e] == b && o->ev1[e] == a)) {
      return e;
    }
  }
  return -1;
}

static void build_topology(OrcGrids *o, int totvert, const int *poly_off, const int *poly_verts, int totpoly)
{
  const int G = o->G, gs = o->gs, n2 = 2 * gs;
  o->F = totpoly;
  o->Vc = totvert;
  o->face_start = (int *)malloc(sizeof(int) * (size_t)totpoly);
  o->face_n = (int *)malloc(sizeof(int) * (size_t)totpoly);
  o->grid_face = (int *)malloc(sizeof(int) * (size_t)G);
  o->corner_vert = (int *)malloc(sizeof(int) * (size_t)G);
  o->corner_edge = (int *)malloc(sizeof(int) * (size_t)G);
  o->ev0 = (int *)malloc(sizeof(int) * (size_t)G);
  o->ev1 = (int *)malloc(sizeof(int) * (size_t)G);
  /* edges by first use; a chain per smaller vertex id finds existing ones */
  int *head = (int *)malloc(sizeof(int) * (size_t)totvert);
  int *next = (int *)malloc(sizeof(int) * (size_t)G);
  for (int v = 0; v < totvert; v++) {
    head[v] = -1;
  }
  int ne = 0;
  for (int f = 0; f < totpoly; f++) {
    const int n = poly_off[f + 1] - poly_off[f], s = poly_off[f] - poly_off[0];
    o->face_start[f] = s;
    o->face_n[f] = n;
    for (int c = 0; c < n; c++) {
      const int a = poly_verts[poly_off[f] + c], b = poly_verts[poly_off[f] + (c + 1) % n];
      o->grid_face[s + c] = f;
      o->corner_vert[s + c] = a;
      int e = find_edge(o, head, next, a, b);
      if (e < 0) {
        e = ne++;
        o->ev0[e] = a;
        o->ev1[e] = b;
        const int k = a < b ? a : b;
        next[e] = head[k];
        head[k] = e;
      }
      o->corner_edge[s + c] = e;
    }
  }
  o->Ed = ne;
  free(head);
  free(next);
  /* vertex edges, ascending */
  o->vedge_off = (int *)calloc((size_t)totvert + 1, sizeof(int));
  for (int e = 0; e < ne; e++) {
    o->vedge_off[o->ev0[e] + 1]++;
    o->vedge_off[o->ev1[e] + 1]++;
  }
  for (int v = 0; v < totvert; v++) {
    o->vedge_off[v + 1] += o->vedge_off[v];
  }
  o->vedges = (int *)malloc(sizeof(int) * (size_t)(o->vedge_off[totvert] + 1));
  int *fill = (int *)malloc(sizeof(int) * ((size_t)(ne > totvert ? ne : totvert) + 1));
  memcpy(fill, o->vedge_off, sizeof(int) * (size_t)totvert);
  for (int e = 0; e < ne; e++) {
    o->vedges[fill[o->ev0[e]]++] = e;
    o->vedges[fill[o->ev1[e]]++] = e;
  }
  /* subdiv_ccg_init_faces_edge_neighborhood (:397-468) */
  o->edge_foff = (int *)calloc((size_t)ne + 1, sizeof(int));
  for (int g = 0; g < G; g++) {
    o->edge_foff[o->corner_edge[g] + 1]++;
  }
  for (int e = 0; e < ne; e++) {
    o->edge_foff[e + 1] += o->edge_foff[e];
  }
  o->edge_bc = (GCoord *)malloc(sizeof(GCoord) * (size_t)o->edge_foff[ne] * (size_t)n2);
  memcpy(fill, o->edge_foff, sizeof(int) * (size_t)ne);
  for (int f = 0; f < totpoly; f++) {
    const int n = o->face_n[f], s = o->face_start[f];
    for (int c = 0; c < n; c++) {
      const int v = o->corner_vert[s + c], e = o->corner_edge[s + c];
      const int flipped = o->ev0[e] != v;
      const int cur = s + c, nxt = s + (c + 1) % n;
      GCoord *bc = &o->edge_bc[(size_t)(fill[e]++) * n2];
      int k = 0;
      for (int i = 0; i < gs; i++) {
        bc[k++] = flipped ? gc(nxt, gs - i - 1, gs - 1) : gc(cur, gs - 1, gs - i - 1);
      }
      for (int i = 0; i < gs; i++) {
        bc[k++] = flipped ? gc(cur, gs - 1, i) : gc(nxt, i, gs - 1);
      }
    }
  }
  /* subdiv_ccg_init_faces_vertex_neighborhood (:492-530) */
  o->vert_foff = (int *)calloc((size_t)totvert + 1, sizeof(int));
  for (int g = 0; g < G; g++) {
    o->vert_foff[o->corner_vert[g] + 1]++;
  }
  for (int v = 0; v < totvert; v++) {
    o->vert_foff[v + 1] += o->vert_foff[v];
  }
  o->vert_cc = (GCoord *)malloc(sizeof(GCoord) * ((size_t)o->vert_foff[totvert] + 1));
  memcpy(fill, o->vert_foff, sizeof(int) * (size_t)totvert);
  for (int g = 0; g < G; g++) {
    o->vert_cc[fill[o->corner_vert[g]]++] = gc(g, gs - 1, gs - 1);
  }
  free(fill);
}

/* ------------------------------------------------------------------------- */
/* Neighbour coordinates: subdiv_ccg.c:1365-1909.                            */
/* ------------------------------------------------------------------------- */

static int prev_grid(const OrcGrids *o, int g)
{
  const int f = o->grid_face[g], s = o->face_start[f];
  int k = g - 1 - s;
  if (k < 0) {
    k = o->face_n[f] - 1;
  }
  return s + k;
}
static int next_grid(const OrcGrids *o, int g)
{
  const int f = o->grid_face[g], s = o->face_start[f];
  int k = g + 1 - s;
  if (k == o->face_n[f]) {
    k = 0;
  }
  return s + k;
}

static GCoord step_inside(const OrcGrids *o, GCoord c)
{
  const int g1 = o->gs - 1;
  if (c.x == g1) {
    c.x--;
  }
  else if (c.y == g1) {
    c.y--;
  }
  else if (c.x == 0) {
    c.x++;
  }
  else if (c.y == 0) {
    c.y++;
  }
  return c;
}

/* Neighbours of (g, x, y): writes up to cap coords, returns the count; *ndup = how many
 * of them (the last ones) are duplicates of the coordinate itself. */
static int neighbors(const OrcGrids *o, GCoord c, int dup, GCoord *out, int cap, int *ndup)
{
  const int gs = o->gs, g1 = gs - 1, n2 = 2 * gs;
  int n = 0, nd = 0;
#define PUSH(v)             \
  do {                      \
    if (n < cap) {          \
      out[n] = (v);         \
    }                       \
    n++;                    \
  } while (0)
  const int corner = (c.x == 0 && c.y == 0) || (c.x == 0 && c.y == g1) || (c.x == g1 && c.y == g1) ||
                     (c.x == g1 && c.y == 0);
  const int boundary = c.x == 0 || c.y == 0 || c.x == g1 || c.y == g1;
  if (corner && c.x == 0 && c.y == 0) {
    /* face centre (:1496-1521) */
    const int f = o->grid_face[c.g], s = o->face_start[f], nf = o->face_n[f];
    for (int k = 0; k < nf; k++) {
      PUSH(gc(s + k, 1, 0));
    }
    if (dup) {
      for (int k = 0; k < nf; k++) {
        if (s + k != c.g) {
          PUSH(gc(s + k, 0, 0));
          nd++;
        }
      }
    }
  }
  else if (corner && c.x == g1 && c.y == g1) {
    /* coarse vertex (:1523-1610) */
    const int v = o->corner_vert[c.g];
    for (int q = o->vedge_off[v]; q < o->vedge_off[v + 1]; q++) {
      const int e = o->vedges[q];
      const int point = (o->ev0[e] == v) ? 1 : n2 - 2;
      PUSH(o->edge_bc[(size_t)o->edge_foff[e] * n2 + point]);
    }
    if (dup) {
      for (int q = o->vert_foff[v]; q < o->vert_foff[v + 1]; q++) {
        if (o->vert_cc[q].g != c.g) {
          PUSH(o->vert_cc[q]);
          nd++;
        }
      }
    }
  }
  else if (corner || (boundary && !((c.x == 0 && c.y > 0 && c.y < g1) || (c.y == 0 && c.x > 0 && c.x < g1)))) {
    /* on a coarse edge (:1612-1774) */
    const int f = o->grid_face[c.g], s = o->face_start[f], nfg = o->face_n[f], k0 = c.g - s;
    const int e = (c.x == g1) ? o->corner_edge[c.g] : o->corner_edge[s + (k0 == 0 ? nfg - 1 : k0 - 1)];
    const int v = o->corner_vert[c.g];
    int point, dirv;
    if (c.x == g1) {
      point = gs - c.y - 1;
      dirv = o->ev0[e];
    }
    else {
      point = gs + c.x;
      dirv = o->ev1[e];
    }
    if (v != dirv) {
      point = n2 - point - 1;
    }
    const int point_dup = (point == gs) ? point - 1 : point + 1;
    const int next_point = (point == gs - 1) ? point + 2 : point + 1;
    const int prev_point = (point == gs) ? point - 2 : point - 1;
    const int nf = o->edge_foff[e + 1] - o->edge_foff[e];
    GCoord first2[2] = {{0, 0, 0}, {0, 0, 0}};
    GCoord inside[64], dups[128];
    int ndu = 0;
    for (int k = 0; k < nf && k < 64; k++) {
      const GCoord *bc = &o->edge_bc[(size_t)(o->edge_foff[e] + k) * n2];
      const GCoord p = bc[point];
      inside[k] = step_inside(o, p);
      if (p.g == c.g) {
        first2[0] = bc[prev_point];
        first2[1] = bc[next_point];
      }
      else if (dup) {
        dups[ndu++] = p;
      }
      if (dup && corner) {
        dups[ndu++] = bc[point_dup];
      }
    }
    PUSH(first2[0]);
    PUSH(first2[1]);
    for (int k = 0; k < nf && k < 64; k++) {
      PUSH(inside[k]);
    }
    for (int k = 0; k < ndu; k++) {
      PUSH(dups[k]);
    }
    nd = ndu;
  }
  else if (boundary) {
    /* seam between two grids of one face (:1807-1843) */
    if (c.x == 0) {
      PUSH(gc(c.g, c.x, c.y - 1));
      PUSH(gc(c.g, c.x, c.y + 1));
      PUSH(gc(c.g, c.x + 1, c.y));
      PUSH(gc(prev_grid(o, c.g), c.y, 1));
      if (dup) {
        PUSH(gc(prev_grid(o, c.g), c.y, 0));
        nd = 1;
      }
    }
    else {
      PUSH(gc(c.g, c.x - 1, c.y));
      PUSH(gc(c.g, c.x + 1, c.y));
      PUSH(gc(c.g, c.x, c.y + 1));
      PUSH(gc(next_grid(o, c.g), 1, c.x));
      if (dup) {
        PUSH(gc(next_grid(o, c.g), 0, c.x));
        nd = 1;
      }
    }
  }
  else {
    /* inside the grid (:1870-1880) */
    PUSH(gc(c.g, c.x, c.y - 1));
    PUSH(gc(c.g, c.x, c.y + 1));
    PUSH(gc(c.g, c.x - 1, c.y));
    PUSH(gc(c.g, c.x + 1, c.y));
  }
#undef PUSH
  if (ndup) {
    *ndup = nd;
  }
  return n;
}

int orc_grids_neighbors(const OrcGrids *o, int g, int x, int y, int include_duplicates, int *out, int cap,
                        int *r_ndup)
{
  GCoord tmp[512];
  const int n = neighbors(o, gc(g, x, y), include_duplicates, tmp, 512, r_ndup);
  for (int i = 0; i < n && i < 512 && i < cap; i++) {
    out[3 * i] = tmp[i].g;
    out[3 * i + 1] = tmp[i].x;
    out[3 * i + 2] = tmp[i].y;
  }
  return n;
}

/* ------------------------------------------------------------------------- */
/* Normals and seams: subdiv_ccg.c:670-874, :876-1190.                       */
/* ------------------------------------------------------------------------- */

static void grid_normals(OrcGrids *o, int g, float *fn)
{
  const int gs = o->gs, g1 = gs - 1;
  for (int y = 0; y < g1; y++) {
    for (int x = 0; x < g1; x++) {
      g_quad_normal(fn + 3 * (y * g1 + x), o->co + 3 * el(o, g, x, y + 1), o->co + 3 * el(o, g, x + 1, y + 1),
                    o->co + 3 * el(o, g, x + 1, y), o->co + 3 * el(o, g, x, y));
    }
  }
  for (int y = 0; y < gs; y++) {
    for (int x = 0; x < gs; x++) {
      float acc[3] = {0.0f, 0.0f, 0.0f};
      int counter = 0;
      int q[4], nq = 0;
      if (x < g1 && y < g1) {
        q[nq++] = y * g1 + x;
      }
      if (x >= 1) {
        if (y < g1) {
          q[nq++] = y * g1 + (x - 1);
        }
        if (y >= 1) {
          q[nq++] = (y - 1) * g1 + (x - 1);
        }
      }
      if (y >= 1 && x < g1) {
        q[nq++] = (y - 1) * g1 + x;
      }
      for (int i = 0; i < nq; i++) {
        acc[0] += fn[3 * q[i]];
        acc[1] += fn[3 * q[i] + 1];
        acc[2] += fn[3 * q[i] + 2];
        counter++;
      }
      const float f = 1.0f / (float)counter;
      float *no = o->no + 3 * el(o, g, x, y);
      no[0] = acc[0] * f;
      no[1] = acc[1] * f;
      no[2] = acc[2] * f;
    }
  }
}

/* During a dab's normals update: the grids whose elements an average moved (their
 * leaves are refit with the gathered ones; a leaf none of whose elements moved keeps
 * its box).  NULL otherwise. */
static uint8_t *s_moved;

static void note_moved(const OrcGrids *o, int v, const float old[3])
{
  if (s_moved && memcmp(old, o->co + 3 * v, sizeof(float) * 3) != 0) {
    s_moved[v / o->ga] = 1;
  }
}

static void pair_avg(OrcGrids *o, int a, int b)
{
  float *ca = o->co + 3 * a, *cb = o->co + 3 * b;
  float olda[3], oldb[3];
  memcpy(olda, ca, sizeof(olda));
  memcpy(oldb, cb, sizeof(oldb));
  for (int i = 0; i < 3; i++) {
    ca[i] = ca[i] + cb[i];
  }
  for (int i = 0; i < 3; i++) {
    ca[i] = ca[i] * 0.5f;
    cb[i] = ca[i];
  }
  if (o->has_normal) {
    float *na = o->no + 3 * a, *nb = o->no + 3 * b;
    for (int i = 0; i < 3; i++) {
      na[i] = na[i] + nb[i];
    }
    for (int i = 0; i < 3; i++) {
      na[i] = na[i] * 0.5f;
      nb[i] = na[i];
    }
  }
  if (o->has_mask) {
    const float m = (o->mask[a] + o->mask[b]) * 0.5f;
    o->mask[a] = m;
    o->mask[b] = m;
  }
  note_moved(o, a, olda);
  note_moved(o, b, oldb);
}

/* mean over a list of elements written back to each */
static void list_avg(OrcGrids *o, const int *els, int n)
{
  float c[3] = {0.0f, 0.0f, 0.0f}, nn[3] = {0.0f, 0.0f, 0.0f}, m = 0.0f;
  for (int k = 0; k < n; k++) {
    for (int i = 0; i < 3; i++) {
      c[i] += o->co[3 * els[k] + i];
      if (o->has_normal) {
        nn[i] += o->no[3 * els[k] + i];
      }
    }
    if (o->has_mask) {
      m += o->mask[els[k]];
    }
  }
  const float f = 1.0f / (float)n;
  for (int i = 0; i < 3; i++) {
    c[i] *= f;
    nn[i] *= f;
  }
  m *= f;
  for (int k = 0; k < n; k++) {
    float old[3];
    memcpy(old, o->co + 3 * els[k], sizeof(old));
    memcpy(o->co + 3 * els[k], c, sizeof(c));
    note_moved(o, els[k], old);
    if (o->has_normal) {
      memcpy(o->no + 3 * els[k], nn, sizeof(nn));
    }
    if (o->has_mask) {
      o->mask[els[k]] = m;
    }
  }
}

static void inner_face_avg(OrcGrids *o, int f)
{
  const int s = o->face_start[f], n = o->face_n[f], gs = o->gs;
  int prev = s + n - 1;
  for (int c = 0; c < n; c++) {
    for (int i = 1; i < gs; i++) {
      pair_avg(o, el(o, prev, i, 0), el(o, s + c, 0, i));
    }
    prev = s + c;
  }
  int *ctr = (int *)malloc(sizeof(int) * (size_t)n);
  for (int c = 0; c < n; c++) {
    ctr[c] = el(o, s + c, 0, 0);
  }
  list_avg(o, ctr, n);
  free(ctr);
}

static void edge_avg(OrcGrids *o, int e)
{
  const int nf = o->edge_foff[e + 1] - o->edge_foff[e], n2 = 2 * o->gs;
  if (nf == 1) {
    return;
  }
  int *els = (int *)malloc(sizeof(int) * (size_t)nf);
  for (int i = 1; i < n2 - 1; i++) {
    for (int k = 0; k < nf; k++) {
      els[k] = el_c(o, o->edge_bc[(size_t)(o->edge_foff[e] + k) * n2 + i]);
    }
    list_avg(o, els, nf);
  }
  free(els);
}

static void corner_avg(OrcGrids *o, int v)
{
  const int b = o->vert_foff[v], nf = o->vert_foff[v + 1] - b;
  if (nf <= 1) {
    return;
  }
  int *els = (int *)malloc(sizeof(int) * (size_t)nf);
  for (int k = 0; k < nf; k++) {
    els[k] = el_c(o, o->vert_cc[b + k]);
  }
  list_avg(o, els, nf);
  free(els);
}

void orc_grids_average(OrcGrids *o)
{
  for (int f = 0; f < o->F; f++) {
    inner_face_avg(o, f);
  }
  for (int e = 0; e < o->Ed; e++) {
    edge_avg(o, e);
  }
  for (int v = 0; v < o->Vc; v++) {
    corner_avg(o, v);
  }
}

void orc_grids_recalc_normals(OrcGrids *o)
{
  if (!o->has_normal) {
    return;
  }
  float *fn = (float *)malloc(sizeof(float) * 3 * (size_t)(o->gs - 1) * (size_t)(o->gs - 1));
  for (int g = 0; g < o->G; g++) {
    grid_normals(o, g, fn);
  }
  free(fn);
  orc_grids_average(o);
}

/* KERNEL_subdiv_ccg_update_normals over the given faces (each listed once); moved (may
 * be NULL) gets 1 for every grid with an element whose position an average changed. */
static void update_normals_faces(OrcGrids *o, const int *faces, int nfaces, uint8_t *moved)
{
  if (!o->has_normal || nfaces == 0) {
    return;
  }
  s_moved = moved;
  float *fn = (float *)malloc(sizeof(float) * 3 * (size_t)(o->gs - 1) * (size_t)(o->gs - 1));
  for (int i = 0; i < nfaces; i++) {
    const int f = faces[i];
    for (int c = 0; c < o->face_n[f]; c++) {
      grid_normals(o, o->face_start[f] + c, fn);
    }
    inner_face_avg(o, f);
  }
  free(fn);
  uint8_t *eseen = (uint8_t *)calloc((size_t)o->Ed + 1, 1), *vseen = (uint8_t *)calloc((size_t)o->Vc + 1, 1);
  for (int i = 0; i < nfaces; i++) {
    const int f = faces[i];
    for (int c = 0; c < o->face_n[f]; c++) {
      const int g = o->face_start[f] + c, e = o->corner_edge[g], v = o->corner_vert[g];
      if (!eseen[e]) {
        eseen[e] = 1;
        edge_avg(o, e);
      }
      if (!vseen[v]) {
        vseen[v] = 1;
        corner_avg(o, v);
      }
    }
  }
  free(eseen);
  free(vseen);
  s_moved = NULL;
}

void orc_grids_update_normals_faces(OrcGrids *o, const int *faces, int nfaces)
{
  update_normals_faces(o, faces, nfaces, NULL);
}

/* ------------------------------------------------------------------------- */
/* Tree over grids: pbvh.c:152-214, :322-329, :331-361, :383-507, :598-643.  */
/* ------------------------------------------------------------------------- */

static void g_grow(OrcGrids *o, int totnode)
{
  if (totnode > o->node_mem) {
    int mem = o->node_mem ? o->node_mem : 100;
    while (mem < totnode) {
      mem *= 2;
    }
    o->nodes = (ONode *)realloc(o->nodes, sizeof(ONode) * (size_t)mem);
    memset(o->nodes + o->node_mem, 0, sizeof(ONode) * (size_t)(mem - o->node_mem));
    o->node_mem = mem;
  }
  o->totnode = totnode;
}

static int g_partition(int *prim, int lo, int hi, int axis, float mid, const OBBC *bbc)
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
    const int t = prim[i];
    prim[i] = prim[j];
    prim[j] = t;
    i++;
  }
}

static int g_mat_match(const OrcGrids *o, int a, int b) { return !o->grid_mat || o->grid_mat[a] == o->grid_mat[b]; }

static int g_partition_material(OrcGrids *o, int lo, int hi)
{
  int *ind = o->prim_indices;
  const int first = ind[lo];
  int i = lo, j = hi;
  for (;;) {
    for (; g_mat_match(o, first, ind[i]); i++) {
    }
    for (; !g_mat_match(o, first, ind[j]); j--) {
    }
    if (!(i < j)) {
      return i;
    }
    const int t = ind[i];
    ind[i] = ind[j];
    ind[j] = t;
    i++;
  }
}

static int g_needs_split(const OrcGrids *o, int offset, int count)
{
  if (count <= 1 || !o->grid_mat) {
    return 0;
  }
  const int first = o->prim_indices[offset];
  for (int i = offset + count - 1; i > offset; i--) {
    if (!g_mat_match(o, first, o->prim_indices[i])) {
      return 1;
    }
  }
  return 0;
}

static void g_update_vb(OrcGrids *o, ONode *n, const OBBC *bbc, int offset, int count)
{
  gbb_reset(&n->vb);
  for (int i = offset + count - 1; i >= offset; i--) {
    gbb_expand_bb(&n->vb, (const OBB *)&bbc[o->prim_indices[i]]);
  }
  n->orig_vb = n->vb;
}

static int quad_hidden(const OrcGrids *o, int g, int x, int y)
{
  const uint8_t *h = o->hide + (size_t)g * o->ga;
  const int gs = o->gs;
  return h[y * gs + x] || h[y * gs + x + 1] || h[(y + 1) * gs + x + 1] || h[(y + 1) * gs + x];
}

int orc_grids_count_quads(const OrcGrids *o, const int *grids, int n)
{
  const int g1 = o->gs - 1;
  int tot = 0;
  for (int i = 0; i < n; i++) {
    if (!o->hide) {
      tot += g1 * g1;
      continue;
    }
    for (int y = 0; y < g1; y++) {
      for (int x = 0; x < g1; x++) {
        if (!quad_hidden(o, grids[i], x, y)) {
          tot++;
        }
      }
    }
  }
  return tot;
}

static void g_build_leaf(OrcGrids *o, int ni, const OBBC *bbc, int offset, int count)
{
  ONode *n = &o->nodes[ni];
  n->flag |= ORC_Leaf;
  n->prim_off = offset;
  n->totprim = count;
  g_update_vb(o, n, bbc, offset, count);
  if (orc_grids_count_quads(o, o->prim_indices + offset, count) == 0) {
    n->flag |= ORC_FullyHidden;
  }
  n->flag |= ORC_UpdateDrawBuffers | ORC_UpdateRedraw;
}

static void g_build_sub(OrcGrids *o, int ni, OBB *cb, const OBBC *bbc, int offset, int count)
{
  OBB cb_backing;
  const int below = count <= o->leaf_limit;
  if (below && !g_needs_split(o, offset, count)) {
    g_build_leaf(o, ni, bbc, offset, count);
    return;
  }
  o->nodes[ni].children_offset = o->totnode;
  g_grow(o, o->totnode + 2);
  g_update_vb(o, &o->nodes[ni], bbc, offset, count);
  int end;
  if (!below) {
    if (!cb) {
      cb = &cb_backing;
      gbb_reset(cb);
      for (int i = offset + count - 1; i >= offset; i--) {
        gbb_expand(cb, bbc[o->prim_indices[i]].bcentroid);
      }
    }
    float dim[3];
    for (int i = 0; i < 3; i++) {
      dim[i] = cb->bmax[i] - cb->bmin[i];
    }
    const int axis = (dim[0] > dim[1]) ? ((dim[0] > dim[2]) ? 0 : 2) : ((dim[1] > dim[2]) ? 1 : 2);
    end = g_partition(o->prim_indices, offset, offset + count - 1, axis, (cb->bmax[axis] + cb->bmin[axis]) * 0.5f,
                      bbc);
  }
  else {
    end = g_partition_material(o, offset, offset + count - 1);
  }
  const int c = o->nodes[ni].children_offset;
  g_build_sub(o, c, NULL, bbc, offset, end - offset);
  g_build_sub(o, c + 1, NULL, bbc, end, offset + count - end);
}

static void build_tree(OrcGrids *o)
{
  const int G = o->G, ga = o->ga;
  OBBC *bbc = (OBBC *)malloc(sizeof(OBBC) * (size_t)G);
  OBB cb;
  gbb_reset(&cb);
  for (int g = 0; g < G; g++) {
    gbb_reset((OBB *)&bbc[g]);
    for (int j = 0; j < ga; j++) {
      gbb_expand((OBB *)&bbc[g], o->co + 3 * ((size_t)g * ga + j));
    }
    for (int i = 0; i < 3; i++) {
      bbc[g].bcentroid[i] = (bbc[g].bmin[i] + bbc[g].bmax[i]) * 0.5f;
    }
    gbb_expand(&cb, bbc[g].bcentroid);
  }
  o->prim_indices = (int *)malloc(sizeof(int) * (size_t)G);
  for (int g = 0; g < G; g++) {
    o->prim_indices[g] = g;
  }
  g_grow(o, 1);
  g_build_sub(o, 0, &cb, bbc, 0, G);
  free(bbc);
}

/* ------------------------------------------------------------------------- */
/* Construction and accessors.                                               */
/* ------------------------------------------------------------------------- */

OrcGrids *orc_grids_new(int totvert, const int *poly_off, const int *poly_verts, int totpoly, int grid_size,
                        const float *co, const float *no, const float *mask, const uint8_t *hidden,
                        const int *grid_mat, int leaf_limit, int respect_hide)
{
  OrcGrids *o = (OrcGrids *)calloc(1, sizeof(OrcGrids));
  o->gs = grid_size;
  o->ga = grid_size * grid_size;
  o->G = poly_off[totpoly] - poly_off[0];
  const size_t E = (size_t)o->G * o->ga;
  o->co = (float *)malloc(sizeof(float) * 3 * E);
  memcpy(o->co, co, sizeof(float) * 3 * E);
  o->has_normal = no != NULL;
  o->no = (float *)calloc(3 * E, sizeof(float));
  if (no) {
    memcpy(o->no, no, sizeof(float) * 3 * E);
  }
  o->has_mask = mask != NULL;
  o->mask = (float *)calloc(E, sizeof(float));
  if (mask) {
    memcpy(o->mask, mask, sizeof(float) * E);
  }
  if (hidden) {
    o->hide = (uint8_t *)malloc(E);
    memcpy(o->hide, hidden, E);
  }
  if (grid_mat) {
    o->grid_mat = (int *)malloc(sizeof(int) * (size_t)o->G);
    memcpy(o->grid_mat, grid_mat, sizeof(int) * (size_t)o->G);
  }
  o->respect_hide = respect_hide;
  const int lim = leaf_limit > 0 ? leaf_limit : 10000;
  o->leaf_limit = lim / o->ga > 1 ? lim / o->ga : 1;
  build_topology(o, totvert, poly_off, poly_verts, totpoly);
  build_tree(o);
  return o;
}

void orc_grids_free(OrcGrids *o)
{
  if (!o) {
    return;
  }
  void *p[] = {o->co, o->no, o->mask, o->hide, o->grid_mat, o->face_start, o->face_n, o->grid_face,
               o->corner_vert, o->corner_edge, o->ev0, o->ev1, o->edge_foff, o->edge_bc, o->vert_foff,
               o->vert_cc, o->vedge_off, o->vedges, o->nodes, o->prim_indices};
  for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++) {
    free(p[i]);
  }
  free(o);
}

int orc_grids_totnode(const OrcGrids *o) { return o->totnode; }
void orc_grids_node_info(const OrcGrids *o, int node, int out[4])
{
  const ONode *n = &o->nodes[node];
  out[0] = n->flag;
  out[1] = (n->flag & ORC_Leaf) ? -1 : n->children_offset;
  out[2] = n->prim_off;
  out[3] = n->totprim;
}
void orc_grids_node_bb(const OrcGrids *o, int node, float vb[6], float orig_vb[6])
{
  memcpy(vb, &o->nodes[node].vb, sizeof(float) * 6);
  memcpy(orig_vb, &o->nodes[node].orig_vb, sizeof(float) * 6);
}
void orc_grids_prim_indices(const OrcGrids *o, int *out) { memcpy(out, o->prim_indices, sizeof(int) * (size_t)o->G); }
void orc_grids_get(const OrcGrids *o, float *co, float *no, float *mask)
{
  const size_t E = (size_t)o->G * o->ga;
  if (co) {
    memcpy(co, o->co, sizeof(float) * 3 * E);
  }
  if (no) {
    memcpy(no, o->no, sizeof(float) * 3 * E);
  }
  if (mask) {
    memcpy(mask, o->mask, sizeof(float) * E);
  }
}
void orc_grids_set_co(OrcGrids *o, const float *co) { memcpy(o->co, co, sizeof(float) * 3 * (size_t)o->G * o->ga); }
void orc_grids_last_stats(const OrcGrids *o, int64_t out[4]) { memcpy(out, o->stats, sizeof(o->stats)); }
int orc_grids_edge_count(const OrcGrids *o) { return o->Ed; }

/* ------------------------------------------------------------------------- */
/* Search, refit, dab.                                                       */
/* ------------------------------------------------------------------------- */

static int g_sphere_test(const OBB *bb, const float c[3], float r2)
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

/* pbvh.c:704-849 traversal: depth first, left child first, internal nodes pruned by
 * the same test; fully hidden leaves skipped (spec S1). */
int orc_grids_search(const OrcGrids *o, const float c[3], float radius, int use_orig, int *out, int cap)
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
    if (!g_sphere_test(use_orig ? &n->orig_vb : &n->vb, c, r2)) {
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

/* BKE_pbvh_update_bounds(UpdateBB | UpdateOriginalBB | UpdateRedraw): flagged leaves
 * get min/max over their elements (grids in prim order, elements row-major), orig_vb
 * = vb; then every ancestor of a refit leaf the union of its children. */
static void g_update_bounds(OrcGrids *o)
{
  for (int ni = 0; ni < o->totnode; ni++) {
    ONode *n = &o->nodes[ni];
    if (!(n->flag & ORC_Leaf) || !(n->flag & ORC_UpdateBB)) {
      continue;
    }
    OBB vb;
    gbb_reset(&vb);
    for (int i = 0; i < n->totprim; i++) {
      const int g = o->prim_indices[n->prim_off + i];
      for (int j = 0; j < o->ga; j++) {
        gbb_expand(&vb, o->co + 3 * ((size_t)g * o->ga + j));
      }
    }
    n->vb = vb;
    if (n->flag & ORC_UpdateOriginalBB) {
      n->orig_vb = n->vb;
    }
  }
  /* pbvh_flush_bb: children before parents */
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
      gbb_reset(&vb);
      gbb_expand_bb(&vb, &o->nodes[n->children_offset].vb);
      gbb_expand_bb(&vb, &o->nodes[n->children_offset + 1].vb);
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

static int cmp_i(const void *a, const void *b)
{
  const int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* One dab (the faces dab's spec over grid elements; see the file comment). */
int orc_grids_dab(OrcGrids *o, const OrcDab *dab, int *out_affected, int cap)
{
  memset(o->stats, 0, sizeof(o->stats));
  const int tool = dab->tool, ga = o->ga;
  const float r = dab->radius, r2 = r * r;
  const float *c = dab->center;
  const float bstr = orc_bstrength(dab);
  int *leaves = (int *)malloc(sizeof(int) * (size_t)(o->totnode + 1));
  const int nl = orc_grids_search(o, c, r, tool == ORC_TOOL_GRAB, leaves, o->totnode + 1);
  o->stats[0] = nl;
  /* candidates: the gathered leaves' elements in order */
  size_t ncand = 0;
  for (int i = 0; i < nl; i++) {
    ncand += (size_t)o->nodes[leaves[i]].totprim * ga;
  }
  int *cand = (int *)malloc(sizeof(int) * (ncand + 1));
  {
    size_t w = 0;
    for (int i = 0; i < nl; i++) {
      const ONode *n = &o->nodes[leaves[i]];
      for (int k = 0; k < n->totprim; k++) {
        const int g = o->prim_indices[n->prim_off + k];
        for (int j = 0; j < ga; j++) {
          cand[w++] = g * ga + j;
        }
      }
    }
  }
  uint8_t *aff = (uint8_t *)calloc(ncand + 1, 1);
  float *fac = (float *)malloc(sizeof(float) * (ncand + 1));
  float *newco = (float *)malloc(sizeof(float) * 3 * (ncand + 1));
  int naff = 0;
  for (size_t i = 0; i < ncand; i++) {
    const int v = cand[i];
    if (o->respect_hide && o->hide && o->hide[v]) {
      continue;
    }
    const float *p = o->co + 3 * v;
    float d[3] = {c[0] - p[0], c[1] - p[1], c[2] - p[2]};
    const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (d2 <= r2) {
      aff[i] = 1;
      naff++;
      fac[i] = orc_falloff(dab, d2) * (1.0f - (o->has_mask ? o->mask[v] : 0.0f));
    }
  }
  float offset[3] = {0.0f, 0.0f, 0.0f};
  if (tool == ORC_TOOL_DRAW) {
    const float nr = r * dab->normal_radius_factor, nr2 = nr * nr;
    int64_t s[3] = {0, 0, 0};
    for (size_t i = 0; i < ncand; i++) {
      const int v = cand[i];
      if (o->respect_hide && o->hide && o->hide[v]) {
        continue;
      }
      const float *p = o->co + 3 * v;
      float d[3] = {c[0] - p[0], c[1] - p[1], c[2] - p[2]};
      const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (d2 <= nr2) {
        for (int q = 0; q < 3; q++) {
          s[q] += (int64_t)llrint((double)o->no[3 * v + q] * 4294967296.0);
        }
      }
    }
    float an[3];
    for (int q = 0; q < 3; q++) {
      an[q] = (float)((double)s[q] * (1.0 / 4294967296.0));
    }
    g_normalize(an);
    for (int q = 0; q < 3; q++) {
      offset[q] = (an[q] * r) * bstr;
    }
  }
  if (tool == ORC_TOOL_SMOOTH) {
    /* spec S8: k Jacobi steps over the affected set, factors from the pre-dab state */
    const int iters = dab->smooth_iters > 0 ? dab->smooth_iters : 1;
    const size_t E = (size_t)o->G * ga;
    int *slot = (int *)malloc(sizeof(int) * E);
    for (size_t v = 0; v < E; v++) {
      slot[v] = -1;
    }
    float *prev = (float *)malloc(sizeof(float) * 3 * (ncand + 1));
    for (size_t i = 0; i < ncand; i++) {
      if (aff[i]) {
        slot[cand[i]] = (int)i;
      }
      memcpy(prev + 3 * i, o->co + 3 * cand[i], sizeof(float) * 3);
    }
    GCoord nb[512];
    for (int it = 0; it < iters; it++) {
      for (size_t i = 0; i < ncand; i++) {
        if (!aff[i]) {
          continue;
        }
        const int v = cand[i];
        const int g = v / ga, j = v % ga;
        int nd;
        const int nn = neighbors(o, gc(g, j % o->gs, j / o->gs), 0, nb, 512, &nd);
        float sum[3] = {0.0f, 0.0f, 0.0f};
        for (int q = 0; q < nn; q++) {
          const int u = el_c(o, nb[q]);
          const float *pu = slot[u] >= 0 ? prev + 3 * slot[u] : o->co + 3 * u;
          sum[0] = sum[0] + pu[0];
          sum[1] = sum[1] + pu[1];
          sum[2] = sum[2] + pu[2];
        }
        const float *self = prev + 3 * i;
        float *dst = newco + 3 * i;
        if (nn > 0) {
          const float fade = bstr * fac[i];
          const float inv = 1.0f / (float)nn;
          for (int q = 0; q < 3; q++) {
            dst[q] = self[q] + (sum[q] * inv - self[q]) * fade;
          }
        }
        else {
          memcpy(dst, self, sizeof(float) * 3);
        }
      }
      for (size_t i = 0; i < ncand; i++) {
        if (aff[i]) {
          memcpy(prev + 3 * i, newco + 3 * i, sizeof(float) * 3);
        }
      }
    }
    free(prev);
    free(slot);
  }
  else {
    for (size_t i = 0; i < ncand; i++) {
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
  for (size_t i = 0; i < ncand; i++) {
    if (aff[i]) {
      memcpy(o->co + 3 * cand[i], newco + 3 * i, sizeof(float) * 3);
    }
  }
  /* every gathered leaf: UpdateNormals | UpdateBB | UpdateOriginalBB | redraw */
  for (int i = 0; i < nl; i++) {
    o->nodes[leaves[i]].flag |=
        ORC_UpdateNormals | ORC_UpdateBB | ORC_UpdateOriginalBB | ORC_UpdateDrawBuffers | ORC_UpdateRedraw;
  }
  /* BKE_pbvh_update_normals: faces of the grids of UpdateNormals leaves, each once */
  if (o->has_normal) {
    uint8_t *fseen = (uint8_t *)calloc((size_t)o->F + 1, 1);
    int *faces = (int *)malloc(sizeof(int) * ((size_t)o->F + 1));
    int nf = 0;
    for (int ni = 0; ni < o->totnode; ni++) {
      ONode *n = &o->nodes[ni];
      if (!(n->flag & ORC_Leaf) || !(n->flag & ORC_UpdateNormals)) {
        continue;
      }
      for (int k = 0; k < n->totprim; k++) {
        const int f = o->grid_face[o->prim_indices[n->prim_off + k]];
        if (!fseen[f]) {
          fseen[f] = 1;
          faces[nf++] = f;
        }
      }
      n->flag &= ~ORC_UpdateNormals;
    }
    uint8_t *touch = (uint8_t *)calloc((size_t)o->G, 1);
    update_normals_faces(o, faces, nf, touch);
    /* the leaves of grids an average moved are refit with the gathered ones */
    for (int ni = 0; ni < o->totnode; ni++) {
      ONode *n = &o->nodes[ni];
      if (!(n->flag & ORC_Leaf)) {
        continue;
      }
      for (int k = 0; k < n->totprim; k++) {
        if (touch[o->prim_indices[n->prim_off + k]]) {
          n->flag |= ORC_UpdateBB | ORC_UpdateOriginalBB;
          break;
        }
      }
    }
    o->stats[2] = nf;
    free(touch);
    free(fseen);
    free(faces);
  }
  else {
    for (int i = 0; i < nl; i++) {
      o->nodes[leaves[i]].flag &= ~ORC_UpdateNormals;
    }
  }
  g_update_bounds(o);
  o->stats[1] = naff;
  if (out_affected) {
    int w = 0;
    for (size_t i = 0; i < ncand && w < cap; i++) {
      if (aff[i]) {
        out_affected[w++] = cand[i];
      }
    }
    qsort(out_affected, (size_t)w, sizeof(int), cmp_i);
  }
  free(leaves);
  free(cand);
  free(aff);
  free(fac);
  free(newco);
  return naff;
}

/* The drop-in update path: flag the listed leaves (node ids), then
 * BKE_pbvh_update_normals (faces of UpdateNormals leaves; no refit) or
 * BKE_pbvh_update_bounds (UpdateBB | UpdateOriginalBB leaves). */
void orc_grids_mark(OrcGrids *o, const int *nodes, int n, int flag)
{
  for (int i = 0; i < n; i++) {
    o->nodes[nodes[i]].flag |= flag;
  }
}
void orc_grids_update_normals(OrcGrids *o)
{
  uint8_t *fseen = (uint8_t *)calloc((size_t)o->F + 1, 1);
  int *faces = (int *)malloc(sizeof(int) * ((size_t)o->F + 1));
  int nf = 0;
  for (int ni = 0; ni < o->totnode; ni++) {
    ONode *n = &o->nodes[ni];
    if (!(n->flag & ORC_Leaf) || !(n->flag & ORC_UpdateNormals)) {
      continue;
    }
    for (int k = 0; k < n->totprim; k++) {
      const int f = o->grid_face[o->prim_indices[n->prim_off + k]];
      if (!fseen[f]) {
        fseen[f] = 1;
        faces[nf++] = f;
      }
    }
    n->flag &= ~ORC_UpdateNormals;
  }
  update_normals_faces(o, faces, nf, NULL);
  free(fseen);
  free(faces);
}
void orc_grids_update_bounds(OrcGrids *o) { g_update_bounds(o); }
