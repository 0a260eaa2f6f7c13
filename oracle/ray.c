/*
 * CPU oracle, ray cast -- TEST INFRASTRUCTURE ONLY (see oracle.h).  Never part of
 * the product: only tests/, smoke() and bench.py's cpu_baseline leg load it.
 *
 * Restates, in FP32 with no contraction, the reference's ray queries on a faces
 * PBVH and the absent sculpt callbacks that drive them:
 *
 *   ray/box       isect_ray_aabb_v3_precalc, isect_ray_aabb_v3
 *                 (lib/intern/math_geom.cc:3017-3080)
 *   ray/tri       isect_ray_tri_watertight_v3_precalc, isect_ray_tri_watertight_v3
 *                 (math_geom.cc:1755-1855), axis_dominant_v3_single
 *                 (math_geom_inline.c:76-82), xor_fl (math_bits_inline.c:159-162)
 *   ray/segment   isect_ray_line_v3 (math_geom.cc:1997-2021),
 *                 dist_squared_ray_to_seg_v3 (:574-606), interp_v3_v3v3
 *                 (math_vector.c:24-31)
 *   ray/box dist  dist_squared_ray_to_aabb_v3_precalc, dist_squared_ray_to_aabb_v3
 *                 (math_geom.cc:610-744)
 *   traversal     pbvh_iter_next_occluded (kernel/intern/pbvh.c:789-816),
 *                 node_tree_insert / traverse_tree (:871-912),
 *                 BKE_pbvh_search_callback_occluded (:934-970)
 *   raycast       ray_aabb_intersect + BKE_pbvh_raycast (:3895-3928),
 *                 ray_face_intersection_tri (:3951-3966),
 *                 pbvh_faces_node_raycast (:4041-4107), BKE_pbvh_node_raycast (:4203-4260)
 *   nearest       dist_squared_ray_to_tri_v3_fast (:3968-3990), ray_face_nearest_tri
 *                 (:4019-4037), nearest_to_ray_aabb_dist_sq + BKE_pbvh_find_nearest_to_ray
 *                 (:4316-4351), pbvh_faces_node_nearest_to_ray (:4353-4392)
 *   project       BKE_pbvh_raycast_project_ray_root (:4262-4307)
 *
 * The callbacks the sculpt operator passes (sculpt.c is absent) follow the project
 * spec: visit leaves in the traversal's tmin order; a leaf is tested only if its tmin
 * is below the running bound (FLT_MAX before the first hit, then the hit depth for the
 * ray cast, the squared distance for the nearest query); a hit sets the bound.  The
 * node raycast is called with no original-coordinate array (current positions).
 */
#include "oracle.h"
#include "oracle_int.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline float r_dot(const float a[3], const float b[3])
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* ---------------------------------------------------------------- ray / box */

typedef struct {
  float origin[3], inv[3];
  int sign[3];
} RayBox;

static void raybox_precalc(RayBox *d, const float origin[3], const float dir[3])
{
  for (int i = 0; i < 3; i++) {
    d->origin[i] = origin[i];
    d->inv[i] = 1.0f / dir[i];
    d->sign[i] = d->inv[i] < 0.0f;
  }
}

/* Slab test; writes *tmin only on a pass, like the reference. */
static int raybox_isect(const RayBox *d, const float bmin[3], const float bmax[3], float *tmin_out)
{
  const float *bb[2] = {bmin, bmax};
  float tmin = (bb[d->sign[0]][0] - d->origin[0]) * d->inv[0];
  float tmax = (bb[1 - d->sign[0]][0] - d->origin[0]) * d->inv[0];
  const float tymin = (bb[d->sign[1]][1] - d->origin[1]) * d->inv[1];
  const float tymax = (bb[1 - d->sign[1]][1] - d->origin[1]) * d->inv[1];
  if ((tmin > tymax) || (tymin > tmax)) {
    return 0;
  }
  if (tymin > tmin) {
    tmin = tymin;
  }
  if (tymax < tmax) {
    tmax = tymax;
  }
  const float tzmin = (bb[d->sign[2]][2] - d->origin[2]) * d->inv[2];
  const float tzmax = (bb[1 - d->sign[2]][2] - d->origin[2]) * d->inv[2];
  if ((tmin > tzmax) || (tzmin > tmax)) {
    return 0;
  }
  if (tzmin > tmin) {
    tmin = tzmin;
  }
  if (tmin_out) {
    *tmin_out = tmin;
  }
  return 1;
}

/* ---------------------------------------------------------------- ray / tri */

typedef struct {
  int kx, ky, kz;
  float sx, sy, sz;
} RayTri;

static void raytri_precalc(RayTri *p, const float dir[3])
{
  const float x = fabsf(dir[0]), y = fabsf(dir[1]), z = fabsf(dir[2]);
  const int kz = (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2);
  int kx = (kz != 2) ? (kz + 1) : 0;
  int ky = (kx != 2) ? (kx + 1) : 0;
  if (dir[kz] < 0.0f) { /* keep the winding */
    const int t = kx;
    kx = ky;
    ky = t;
  }
  const float inv_z = 1.0f / dir[kz];
  p->sx = dir[kx] * inv_z;
  p->sy = dir[ky] * inv_z;
  p->sz = inv_z;
  p->kx = kx;
  p->ky = ky;
  p->kz = kz;
}

static inline uint32_t f_bits(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static inline float bits_f(uint32_t u)
{
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static int raytri_isect(const float o[3], const RayTri *p, const float v0[3], const float v1[3],
                        const float v2[3], float *r_lambda)
{
  const int kx = p->kx, ky = p->ky, kz = p->kz;
  const float a[3] = {v0[0] - o[0], v0[1] - o[1], v0[2] - o[2]};
  const float b[3] = {v1[0] - o[0], v1[1] - o[1], v1[2] - o[2]};
  const float c[3] = {v2[0] - o[0], v2[1] - o[1], v2[2] - o[2]};
  const float ax = a[kx] - p->sx * a[kz];
  const float ay = a[ky] - p->sy * a[kz];
  const float bx = b[kx] - p->sx * b[kz];
  const float by = b[ky] - p->sy * b[kz];
  const float cx = c[kx] - p->sx * c[kz];
  const float cy = c[ky] - p->sy * c[kz];
  const float u = cx * by - cy * bx;
  const float v = ax * cy - ay * cx;
  const float w = bx * ay - by * ax;
  if ((u < 0.0f || v < 0.0f || w < 0.0f) && (u > 0.0f || v > 0.0f || w > 0.0f)) {
    return 0;
  }
  const float det = u + v + w;
  if (det == 0.0f || !isfinite(det)) {
    return 0;
  }
  const float t = (u * a[kz] + v * b[kz] + w * c[kz]) * p->sz;
  const float sign_t = bits_f(f_bits(t) ^ (f_bits(det) & 0x80000000u));
  if (sign_t < 0.0f) {
    return 0;
  }
  const float inv_det = 1.0f / det;
  *r_lambda = t * inv_det;
  return 1;
}

/* ------------------------------------------------------- ray / segment, box */

static int ray_line(const float o[3], const float dir[3], const float v0[3], const float v1[3],
                    float *r_lambda)
{
  float a[3], t[3], n[3], c[3], cray[3];
  for (int i = 0; i < 3; i++) {
    a[i] = v1[i] - v0[i];
    t[i] = v0[i] - o[i];
  }
  n[0] = a[1] * dir[2] - a[2] * dir[1];
  n[1] = a[2] * dir[0] - a[0] * dir[2];
  n[2] = a[0] * dir[1] - a[1] * dir[0];
  const float nlen = r_dot(n, n);
  if (nlen == 0.0f) {
    return 0; /* parallel */
  }
  for (int i = 0; i < 3; i++) {
    c[i] = n[i] - t[i];
  }
  cray[0] = c[1] * dir[2] - c[2] * dir[1];
  cray[1] = c[2] * dir[0] - c[0] * dir[2];
  cray[2] = c[0] * dir[1] - c[1] * dir[0];
  *r_lambda = r_dot(cray, n) / nlen;
  return 1;
}

static float ray_seg_dist_sq(const float o[3], const float dir[3], const float v0[3],
                             const float v1[3], float r_point[3], float *r_depth)
{
  float lambda;
  if (ray_line(o, dir, v0, v1, &lambda)) {
    if (lambda <= 0.0f) {
      memcpy(r_point, v0, sizeof(float) * 3);
    }
    else if (lambda >= 1.0f) {
      memcpy(r_point, v1, sizeof(float) * 3);
    }
    else {
      const float s = 1.0f - lambda;
      for (int i = 0; i < 3; i++) {
        r_point[i] = s * v0[i] + lambda * v1[i];
      }
    }
  }
  else {
    memcpy(r_point, v0, sizeof(float) * 3);
  }
  float dvec[3];
  for (int i = 0; i < 3; i++) {
    dvec[i] = r_point[i] - o[i];
  }
  const float depth = r_dot(dvec, dir);
  if (r_depth) {
    *r_depth = depth;
  }
  return r_dot(dvec, dvec) - depth * depth;
}

typedef struct {
  float origin[3], dir[3], inv[3];
} RayBoxDist;

static void rayboxdist_precalc(RayBoxDist *d, const float origin[3], const float dir[3])
{
  for (int i = 0; i < 3; i++) {
    d->origin[i] = origin[i];
    d->dir[i] = dir[i];
    d->inv[i] = (dir[i] != 0.0f) ? (1.0f / dir[i]) : FLT_MAX;
  }
}

static float rayboxdist(const RayBoxDist *d, const float bb_min[3], const float bb_max[3],
                        float r_point[3], float *r_depth)
{
  float lo[3], hi[3]; /* near / far corner along the ray */
  for (int i = 0; i < 3; i++) {
    if (d->dir[i] < 0.0f) {
      lo[i] = bb_max[i];
      hi[i] = bb_min[i];
    }
    else {
      lo[i] = bb_min[i];
      hi[i] = bb_max[i];
    }
  }
  const float tmin[3] = {(lo[0] - d->origin[0]) * d->inv[0], (lo[1] - d->origin[1]) * d->inv[1],
                         (lo[2] - d->origin[2]) * d->inv[2]};
  const float tmax[3] = {(hi[0] - d->origin[0]) * d->inv[0], (hi[1] - d->origin[1]) * d->inv[1],
                         (hi[2] - d->origin[2]) * d->inv[2]};
  float va[3], vb[3], rtmin, rtmax;
  int main_axis;
  if ((tmax[0] <= tmax[1]) && (tmax[0] <= tmax[2])) {
    rtmax = tmax[0];
    va[0] = vb[0] = hi[0];
    main_axis = 3;
  }
  else if ((tmax[1] <= tmax[0]) && (tmax[1] <= tmax[2])) {
    rtmax = tmax[1];
    va[1] = vb[1] = hi[1];
    main_axis = 2;
  }
  else {
    rtmax = tmax[2];
    va[2] = vb[2] = hi[2];
    main_axis = 1;
  }
  if ((tmin[0] >= tmin[1]) && (tmin[0] >= tmin[2])) {
    rtmin = tmin[0];
    va[0] = vb[0] = lo[0];
    main_axis -= 3;
  }
  else if ((tmin[1] >= tmin[0]) && (tmin[1] >= tmin[2])) {
    rtmin = tmin[1];
    va[1] = vb[1] = lo[1];
    main_axis -= 1;
  }
  else {
    rtmin = tmin[2];
    va[2] = vb[2] = lo[2];
    main_axis -= 2;
  }
  if (main_axis < 0) {
    main_axis += 3;
  }
  if (rtmin <= rtmax) { /* the ray meets the box */
    float dvec[3];
    memcpy(r_point, hi, sizeof(float) * 3);
    for (int i = 0; i < 3; i++) {
      dvec[i] = hi[i] - d->origin[i];
    }
    *r_depth = r_dot(dvec, d->dir);
    return 0.0f;
  }
  if (d->dir[main_axis] >= 0.0f) {
    va[main_axis] = lo[main_axis];
    vb[main_axis] = hi[main_axis];
  }
  else {
    va[main_axis] = hi[main_axis];
    vb[main_axis] = lo[main_axis];
  }
  return ray_seg_dist_sq(d->origin, d->dir, va, vb, r_point, r_depth);
}

/* dist_squared_ray_to_tri_v3_fast: the three edges (2,0), (0,1), (1,2); the first
 * edge always initialises the result. */
static float ray_tri_edges_dist_sq(const float o[3], const float dir[3], const float *tri[3],
                                   float *r_depth)
{
  float best = FLT_MAX;
  for (int i = 0, j = 2; i < 3; j = i++) {
    float pt[3], depth_test = FLT_MAX;
    const float dsq = ray_seg_dist_sq(o, dir, tri[i], tri[j], pt, &depth_test);
    if (dsq < best || i == 0) {
      *r_depth = depth_test;
      best = dsq;
    }
  }
  return best;
}

/* ------------------------------------------------------ occluded traversal */

typedef int (*occ_scb)(const OrcPBVH *o, int node, void *data, float *tmin);

/* pbvh_iter_next_occluded over the whole tree, then node_tree_insert's unbalanced
 * binary tree (left iff new.tmin < node.tmin, so ties and NaN go right), read back in
 * order.  Iterative (the reference recurses).  Returns the leaves in visiting order
 * with their tmin. */
static int occluded_order(const OrcPBVH *o, occ_scb scb, void *data, int **r_leaf, float **r_tmin,
                          int64_t *ntests)
{
  *r_leaf = NULL;
  *r_tmin = NULL;
  if (!o->nodes || o->totnode == 0) {
    return 0;
  }
  float *ntmin = (float *)malloc(sizeof(float) * (size_t)o->totnode);
  int cap = 64, tot = 0;
  int *leaf = (int *)malloc(sizeof(int) * (size_t)cap);
  int *left = (int *)malloc(sizeof(int) * (size_t)cap);
  int *right = (int *)malloc(sizeof(int) * (size_t)cap);
  int scap = 128, ssz = 0;
  int *st = (int *)malloc(sizeof(int) * (size_t)scap);
  st[ssz++] = 0;
  while (ssz) {
    const int ni = st[--ssz];
    const ONode *n = &o->nodes[ni];
    if (ntests) {
      (*ntests)++;
    }
    if (!scb(o, ni, data, &ntmin[ni])) {
      continue;
    }
    if (n->flag & ORC_Leaf) {
      if (tot == cap) {
        cap *= 2;
        leaf = (int *)realloc(leaf, sizeof(int) * (size_t)cap);
        left = (int *)realloc(left, sizeof(int) * (size_t)cap);
        right = (int *)realloc(right, sizeof(int) * (size_t)cap);
      }
      leaf[tot] = ni;
      left[tot] = right[tot] = -1;
      if (tot > 0) {
        int cur = 0;
        for (;;) {
          int *slot = (ntmin[ni] < ntmin[leaf[cur]]) ? &left[cur] : &right[cur];
          if (*slot < 0) {
            *slot = tot;
            break;
          }
          cur = *slot;
        }
      }
      tot++;
      continue;
    }
    if (ssz + 2 > scap) {
      scap *= 2;
      st = (int *)realloc(st, sizeof(int) * (size_t)scap);
    }
    st[ssz++] = n->children_offset + 1;
    st[ssz++] = n->children_offset;
  }
  int *out = (int *)malloc(sizeof(int) * (size_t)(tot ? tot : 1));
  float *out_t = (float *)malloc(sizeof(float) * (size_t)(tot ? tot : 1));
  int k = 0;
  if (tot) {
    /* in-order walk with an explicit stack */
    int *ws = (int *)malloc(sizeof(int) * (size_t)tot);
    int wsz = 0, cur = 0;
    while (cur >= 0 || wsz) {
      while (cur >= 0) {
        ws[wsz++] = cur;
        cur = left[cur];
      }
      cur = ws[--wsz];
      out[k] = leaf[cur];
      out_t[k] = ntmin[leaf[cur]];
      k++;
      cur = right[cur];
    }
    free(ws);
  }
  free(st);
  free(leaf);
  free(left);
  free(right);
  free(ntmin);
  *r_leaf = out;
  *r_tmin = out_t;
  return k;
}

static inline int tri_hidden(const OrcPBVH *o, int t)
{
  if (!o->respect_hide || !o->vhide) {
    return 0;
  }
  const int *c = &o->tris[3 * t];
  return o->vhide[c[0]] || o->vhide[c[1]] || o->vhide[c[2]];
}

/* ---------------------------------------------------------------- ray cast */

typedef struct {
  RayBox box;
  int original;
} CastSearch;

static int cast_scb(const OrcPBVH *o, int ni, void *data, float *tmin)
{
  const CastSearch *d = (const CastSearch *)data;
  const OBB *bb = d->original ? &o->nodes[ni].orig_vb : &o->nodes[ni].vb;
  return raybox_isect(&d->box, bb->bmin, bb->bmax, tmin);
}

/* pbvh_faces_node_raycast with origco = NULL. */
static int node_raycast(const OrcPBVH *o, const ONode *n, const float start[3], const float dir[3],
                        const RayTri *pre, float *depth, OrcRayHit *h, int64_t *ntris)
{
  if (n->flag & ORC_FullyHidden) {
    return 0;
  }
  int hit = 0;
  for (int i = 0; i < n->totprim; i++) {
    const int t = o->prim_indices[n->prim_off + i];
    if (tri_hidden(o, t)) {
      continue;
    }
    if (ntris) {
      (*ntris)++;
    }
    const int *c = &o->tris[3 * t];
    const float *co[3] = {&o->co[3 * c[0]], &o->co[3 * c[1]], &o->co[3 * c[2]]};
    float dt;
    if (raytri_isect(start, pre, co[0], co[1], co[2], &dt) && dt < *depth) {
      *depth = dt;
      hit = 1;
      orc_tri_normal(co[0], co[1], co[2], h->face_normal);
      float loc[3], best[3] = {0.0f, 0.0f, 0.0f};
      for (int k = 0; k < 3; k++) {
        loc[k] = start[k] + dir[k] * *depth;
      }
      for (int j = 0; j < 3; j++) {
        float dj[3], db[3];
        for (int k = 0; k < 3; k++) {
          dj[k] = co[j][k] - loc[k];
          db[k] = best[k] - loc[k];
        }
        if (j == 0 || r_dot(dj, dj) < r_dot(db, db)) {
          memcpy(best, co[j], sizeof(best));
          h->active_vertex = c[j];
          h->tri = t;
        }
      }
    }
  }
  return hit;
}

int orc_raycast(const OrcPBVH *o, const float start[3], const float dir[3], float depth0,
                int original, OrcRayHit *h, int64_t stats[4])
{
  CastSearch d;
  raybox_precalc(&d.box, start, dir);
  d.original = original;
  RayTri pre;
  raytri_precalc(&pre, dir);
  memset(h, 0, sizeof(*h));
  h->tri = h->active_vertex = h->leaf = -1;
  h->depth = depth0;
  int64_t ntests = 0, nleaves = 0, ntris = 0;
  int *leaf;
  float *ltmin;
  const int tot = occluded_order(o, cast_scb, &d, &leaf, &ltmin, &ntests);
  float bound = FLT_MAX; /* traverse_tree's tmin */
  for (int i = 0; i < tot; i++) {
    if (!(ltmin[i] < bound)) {
      continue;
    }
    nleaves++;
    if (node_raycast(o, &o->nodes[leaf[i]], start, dir, &pre, &h->depth, h, &ntris)) {
      h->hit = 1;
      h->leaf = leaf[i];
      bound = h->depth;
    }
  }
  free(leaf);
  free(ltmin);
  if (h->hit) {
    for (int k = 0; k < 3; k++) {
      h->location[k] = start[k] + dir[k] * h->depth;
    }
  }
  if (stats) {
    stats[0] = ntests;
    stats[1] = tot;
    stats[2] = nleaves;
    stats[3] = ntris;
  }
  return h->hit;
}

/* ------------------------------------------------------- nearest to a ray */

typedef struct {
  RayBoxDist box;
  int original;
} NearSearch;

static int near_scb(const OrcPBVH *o, int ni, void *data, float *tmin)
{
  const NearSearch *d = (const NearSearch *)data;
  const OBB *bb = d->original ? &o->nodes[ni].orig_vb : &o->nodes[ni].vb;
  float pt[3], depth;
  *tmin = rayboxdist(&d->box, bb->bmin, bb->bmax, pt, &depth);
  return depth > 0.0f;
}

/* pbvh_faces_node_nearest_to_ray with origco = NULL. */
static int node_nearest(const OrcPBVH *o, const ONode *n, const float start[3], const float dir[3],
                        float *depth, float *dist_sq, int *r_tri)
{
  if (n->flag & ORC_FullyHidden) {
    return 0;
  }
  int hit = 0;
  for (int i = 0; i < n->totprim; i++) {
    const int t = o->prim_indices[n->prim_off + i];
    if (tri_hidden(o, t)) {
      continue;
    }
    const int *c = &o->tris[3 * t];
    const float *co[3] = {&o->co[3 * c[0]], &o->co[3 * c[1]], &o->co[3 * c[2]]};
    float dt;
    const float dsq = ray_tri_edges_dist_sq(start, dir, co, &dt);
    if (dsq < *dist_sq) {
      *dist_sq = dsq;
      *depth = dt;
      *r_tri = t;
      hit = 1;
    }
  }
  return hit;
}

int orc_find_nearest_to_ray(const OrcPBVH *o, const float start[3], const float dir[3],
                            int original, OrcRayNearest *r)
{
  NearSearch d;
  rayboxdist_precalc(&d.box, start, dir);
  d.original = original;
  memset(r, 0, sizeof(*r));
  r->depth = FLT_MAX;
  r->dist_sq = FLT_MAX;
  r->tri = r->leaf = -1;
  int *leaf;
  float *ltmin;
  const int tot = occluded_order(o, near_scb, &d, &leaf, &ltmin, NULL);
  float bound = FLT_MAX;
  for (int i = 0; i < tot; i++) {
    if (!(ltmin[i] < bound)) {
      continue;
    }
    if (node_nearest(o, &o->nodes[leaf[i]], start, dir, &r->depth, &r->dist_sq, &r->tri)) {
      r->hit = 1;
      r->leaf = leaf[i];
      bound = r->dist_sq;
    }
  }
  free(leaf);
  free(ltmin);
  if (r->hit) {
    for (int k = 0; k < 3; k++) {
      r->location[k] = start[k] + dir[k] * r->depth;
    }
  }
  return r->hit;
}

/* --------------------------------------------------- visit order (tests) */

int orc_occluded_order(const OrcPBVH *o, const float start[3], const float dir[3], int original,
                       int mode, int *nodes, float *tmin, int cap)
{
  CastSearch cs;
  NearSearch ns;
  raybox_precalc(&cs.box, start, dir);
  cs.original = original;
  rayboxdist_precalc(&ns.box, start, dir);
  ns.original = original;
  int *leaf;
  float *lt;
  const int tot = mode == 0 ? occluded_order(o, cast_scb, &cs, &leaf, &lt, NULL)
                            : occluded_order(o, near_scb, &ns, &leaf, &lt, NULL);
  for (int i = 0; i < tot && i < cap; i++) {
    nodes[i] = leaf[i];
    tmin[i] = lt[i];
  }
  free(leaf);
  free(lt);
  return tot;
}

/* ------------------------------------------------ root projection, pieces */

int orc_project_ray_root(const OrcPBVH *o, int original, float start[3], float end[3], float dir[3])
{
  if (!o->nodes || o->totnode == 0) {
    return 0;
  }
  const OBB *root = original ? &o->nodes[0].orig_vb : &o->nodes[0].vb;
  const float offset = 1.0f + 1e-3f;
  float bmin[3], bmax[3], center[3], diff[3];
  for (int i = 0; i < 3; i++) {
    center[i] = 0.5f * (root->bmax[i] + root->bmin[i]);
    diff[i] = root->bmax[i] - center[i];
    diff[i] = diff[i] + 1e-3f;
    bmax[i] = center[i] + diff[i] * offset;
    bmin[i] = center[i] + diff[i] * -offset;
  }
  RayBox rb;
  float t_start, t_end, inv_dir[3];
  raybox_precalc(&rb, start, dir);
  if (!raybox_isect(&rb, bmin, bmax, &t_start)) {
    return 0;
  }
  for (int i = 0; i < 3; i++) {
    inv_dir[i] = dir[i] * -1.0f;
  }
  raybox_precalc(&rb, end, inv_dir);
  if (!raybox_isect(&rb, bmin, bmax, &t_end)) {
    return 0;
  }
  for (int i = 0; i < 3; i++) {
    start[i] = start[i] + dir[i] * t_start;
    end[i] = end[i] + inv_dir[i] * t_end;
  }
  return 1;
}

int orc_isect_ray_tri(const float start[3], const float dir[3], const float v0[3], const float v1[3],
                      const float v2[3], float *lambda)
{
  RayTri p;
  raytri_precalc(&p, dir);
  return raytri_isect(start, &p, v0, v1, v2, lambda);
}

int orc_isect_ray_box(const float start[3], const float dir[3], const float bmin[3],
                      const float bmax[3], float *tmin)
{
  RayBox d;
  raybox_precalc(&d, start, dir);
  return raybox_isect(&d, bmin, bmax, tmin);
}

float orc_dist_sq_ray_box(const float start[3], const float dir[3], const float bmin[3],
                          const float bmax[3], float *depth)
{
  RayBoxDist d;
  float pt[3];
  rayboxdist_precalc(&d, start, dir);
  return rayboxdist(&d, bmin, bmax, pt, depth);
}

float orc_dist_sq_ray_tri(const float start[3], const float dir[3], const float v0[3],
                          const float v1[3], const float v2[3], float *depth)
{
  const float *tri[3] = {v0, v1, v2};
  return ray_tri_edges_dist_sq(start, dir, tri, depth);
}
