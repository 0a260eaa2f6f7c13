/*
 * CPU oracle, draw buffers -- TEST INFRASTRUCTURE ONLY (see oracle.h).  Never part of
 * the product: only tests/, smoke() and bench.py's cpu_baseline leg load it.
 *
 * Restates what the draw code writes for one PBVH leaf of a faces mesh:
 *
 *   vertex format   gpu_pbvh_init (gpu/intern/gpu_buffers.c:85-100) packed by
 *                   VertexFormat_pack (gpu/intern/gpu_vertex_format.cc:60-70, :281-329):
 *                   pos F32x3 @0, nor I16x3 @16, msk U8 @22, ac U16x4 @24, fset U8x3 @32,
 *                   stride 36
 *   visible tris    gpu_pbvh_is_looptri_visible (:165-172): no hidden corner and a face
 *                   set above SCULPT_FACE_SET_NONE (0; the constant's header is absent)
 *   fill            gpu_pbvh_mesh_buffers_update (:174-306) with the update flags of
 *                   pbvh_get_buffers_update_flags (kernel/intern/pbvh.c:1244-1249):
 *                   positions, normals (vertex normals when the leaf's first face is
 *                   smooth, else the face normal, BKE_mesh_calc_poly_normal,
 *                   kernel/intern/mesh_evaluate.c:39-84), mask and face-set colour
 *   normal->short   normal_float_to_short_v3 (lib/intern/math_vector_inline.c:1260-1265)
 *   overlay colour  BKE_paint_face_set_overlay_color_get (kernel/intern/paint.c:2219-2235),
 *                   hsv_to_rgb (lib/intern/math_color.c:6-22), unit_float_to_uchar_clamp
 *                   (lib/intern/math_base_inline.c:678-682); BLI_hash_int_01 is absent
 *                   (BLI_hash.h is not in the tree): it is Bob Jenkins' lookup3 hashword
 *                   of (k, 0) with initial value 13, restated from lookup3's published
 *                   source and pinned by lookup3's published hash vectors
 *                   (tests/test_oracle_draw.py).
 *
 * The reference leaves the padding bytes and, without vertex colours, the `ac` bytes
 * uninitialised; here they are written as zero, as the GPU writes them.
 */
#include "oracle.h"
#include "oracle_int.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void orc_tri_normal(const float v1[3], const float v2[3], const float v3[3], float n[3]);

static inline float d_dot(const float a[3], const float b[3])
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* normalize_v3 (math_vector_inline.c:1165-1180 via :1242-1245) */
static inline float d_normalize(float n[3])
{
  float d = d_dot(n, n);
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

/* normal_quad_v3 (math_geom.cc:48-68): cross of the diagonals. */
static void quad_normal(const float v1[3], const float v2[3], const float v3[3], const float v4[3],
                        float n[3])
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
  d_normalize(n);
}

/* BKE_mesh_calc_poly_normal (mesh_evaluate.c:62-84), Newell for more than 4 corners
 * (:39-60, add_newell_cross_v3_v3v3 math_vector_inline.c). */
static void poly_normal(const float *co, const int *pv, int n, float r[3])
{
  if (n > 4) {
    r[0] = r[1] = r[2] = 0.0f;
    const float *prev = co + 3 * pv[n - 1];
    for (int i = 0; i < n; i++) {
      const float *cur = co + 3 * pv[i];
      r[0] += (prev[1] - cur[1]) * (prev[2] + cur[2]);
      r[1] += (prev[2] - cur[2]) * (prev[0] + cur[0]);
      r[2] += (prev[0] - cur[0]) * (prev[1] + cur[1]);
      prev = cur;
    }
    if (d_normalize(r) == 0.0f) {
      r[2] = 1.0f;
    }
  }
  else if (n == 3) {
    orc_tri_normal(co + 3 * pv[0], co + 3 * pv[1], co + 3 * pv[2], r);
  }
  else if (n == 4) {
    quad_normal(co + 3 * pv[0], co + 3 * pv[1], co + 3 * pv[2], co + 3 * pv[3], r);
  }
  else {
    r[0] = 0.0f;
    r[1] = 0.0f;
    r[2] = 1.0f;
  }
}

static inline void to_short3(short out[3], const float in[3])
{
  out[0] = (short)(in[0] * 32767.0f);
  out[1] = (short)(in[1] * 32767.0f);
  out[2] = (short)(in[2] * 32767.0f);
}

/* lookup3's final mix (Bob Jenkins, lookup3.c `final`), exported so the tests can run
 * lookup3's published hash vectors through it. */
static inline uint32_t h_rot(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
void orc_lookup3_final(uint32_t abc[3])
{
  uint32_t a = abc[0], b = abc[1], c = abc[2];
  c ^= b;
  c -= h_rot(b, 14);
  a ^= c;
  a -= h_rot(c, 11);
  b ^= a;
  b -= h_rot(a, 25);
  c ^= b;
  c -= h_rot(b, 16);
  a ^= c;
  a -= h_rot(c, 4);
  b ^= a;
  b -= h_rot(a, 14);
  c ^= b;
  c -= h_rot(b, 24);
  abc[0] = a;
  abc[1] = b;
  abc[2] = c;
}

/* BLI_hash_int_01(k) = BLI_hash_int_2d(k, 0) / 0xFFFFFFFF, where BLI_hash_int_2d is
 * lookup3's hashword of the two words (k, 0) with initial value 13: a = b = c =
 * 0xdeadbeef + (2 << 2) + 13, a += k, final.  (Published algorithm; the header that
 * holds it, BLI_hash.h, is absent from the reference.) */
float orc_hash_int_01(uint32_t k)
{
  uint32_t abc[3];
  abc[0] = abc[1] = abc[2] = 0xdeadbeefu + (2u << 2) + 13u;
  abc[0] += k;
  orc_lookup3_final(abc);
  return (float)abc[2] * (1.0f / (float)0xFFFFFFFFu);
}

static inline unsigned char unit_to_uchar(float v)
{
  return (unsigned char)((v <= 0.0f) ? 0.0f : ((v > (1.0f - 0.5f / 255.0f)) ? 255.0f : ((255.0f * v) + 0.5f)));
}

void orc_face_set_color(int face_set, int seed, unsigned char out[3])
{
  const int fs = abs(face_set);
  float hue = 0.618033988749895f * (float)(fs + (seed % 10));
  hue = hue - floorf(hue);
  const float sat = orc_hash_int_01((uint32_t)(fs + seed + 1));
  const float val = orc_hash_int_01((uint32_t)(fs + seed + 2));
  const float s = 0.6f + (sat * 0.25f), v = 1.0f - (val * 0.35f);
  float nr = fabsf(hue * 6.0f - 3.0f) - 1.0f;
  float ng = 2.0f - fabsf(hue * 6.0f - 2.0f);
  float nb = 2.0f - fabsf(hue * 6.0f - 4.0f);
  nr = nr < 0.0f ? 0.0f : (nr > 1.0f ? 1.0f : nr);
  nb = nb < 0.0f ? 0.0f : (nb > 1.0f ? 1.0f : nb);
  ng = ng < 0.0f ? 0.0f : (ng > 1.0f ? 1.0f : ng);
  out[0] = unit_to_uchar(((nr - 1.0f) * s + 1.0f) * v);
  out[1] = unit_to_uchar(((ng - 1.0f) * s + 1.0f) * v);
  out[2] = unit_to_uchar(((nb - 1.0f) * s + 1.0f) * v);
}

static inline int tri_hidden(const OrcPBVH *o, int t)
{
  if (!o->vhide) {
    return 0;
  }
  const int *c = &o->tris[3 * t];
  return o->vhide[c[0]] || o->vhide[c[1]] || o->vhide[c[2]];
}

int orc_draw_leaf(const OrcPBVH *o, int node, const OrcPolys *pl, const float *mask,
                  const int *face_sets, int seed, int fset_default, unsigned char *out, int cap_verts,
                  int *r_smooth, int *r_material)
{
  const ONode *n = &o->nodes[node];
  if (!(n->flag & ORC_Leaf) || n->totprim == 0) {
    return 0;
  }
  const int first_poly = pl->tri_poly[o->prim_indices[n->prim_off]];
  const int smooth = pl->poly_smooth[first_poly] != 0;
  if (r_smooth) {
    *r_smooth = smooth;
  }
  if (r_material) {
    *r_material = pl->poly_mat[first_poly];
  }
  int nv = 0;
  int prev_poly = -1;
  short fno[3] = {0, 0, 0};
  for (int i = 0; i < n->totprim; i++) {
    const int t = o->prim_indices[n->prim_off + i];
    const int poly = pl->tri_poly[t];
    if (tri_hidden(o, t) || !face_sets || !(face_sets[poly] > 0)) {
      continue; /* gpu_pbvh_is_looptri_visible */
    }
    const int *vt = &o->tris[3 * t];
    if (poly != prev_poly && !smooth) {
      float pn[3];
      const int b = pl->poly_off[poly], e = pl->poly_off[poly + 1];
      poly_normal(o->co, pl->poly_verts + b, e - b, pn);
      to_short3(fno, pn);
      prev_poly = poly;
    }
    unsigned char col[3] = {255, 255, 255};
    const int fs = abs(face_sets[poly]);
    if (fs != fset_default) {
      orc_face_set_color(fs, seed, col);
    }
    unsigned char cmask = 0;
    if (mask && !smooth) {
      const float fm = (mask[vt[0]] + mask[vt[1]] + mask[vt[2]]) / 3.0f;
      cmask = (unsigned char)(fm * 255);
    }
    for (int j = 0; j < 3; j++) {
      if (nv >= cap_verts) {
        return -1;
      }
      unsigned char *rec = out + 36 * (size_t)nv;
      memset(rec, 0, 36);
      memcpy(rec, o->co + 3 * vt[j], 12);
      short no[3];
      if (smooth) {
        to_short3(no, o->no + 3 * vt[j]);
      }
      else {
        memcpy(no, fno, sizeof(no));
      }
      memcpy(rec + 16, no, 6);
      if (mask && smooth) {
        cmask = (unsigned char)(mask[vt[j]] * 255);
      }
      rec[22] = cmask;
      memcpy(rec + 32, col, 3);
      nv++;
    }
  }
  return nv;
}
