/*
 * CPU oracle, elastic deform -- TEST INFRASTRUCTURE ONLY (see oracle.h).  Never part of
 * the product: only tests/, smoke() and bench.py's cpu_baseline leg load it.
 *
 * The Kelvinlet displacements follow kernel/intern/kelvinlet.c step by step:
 *   params        BKE_kelvinlet_init_params (:6-21)
 *   grab terms    init_kelvinlet_grab, Formula (6) (:23-45)
 *   grab modes    BKE_kelvinlet_grab / _grab_biscale / _grab_triscale (:47-109)
 *   scale, twist  kelvinlet_scale, Formula (16) (:141-155); kelvinlet_twist, Formula
 *                 (15) (:167-182), through sculpt_kelvinet_integrate (:114-138), whose
 *                 fourth stage is evaluated at vertex + k2 - location
 * with pow2f / pow3f / pow5f (lib/intern/math_base_inline.c:22-38), len_v3v3 and
 * cross_v3_v3v3 (math_vector_inline.c:1113-1119, :937-943).  The elastic deform brush
 * task itself is absent from the reference; it is restated from Blender's published
 * one: force = |grab_delta| * direction * bstrength, params with shear modulus 1 and
 * the brush's Poisson ratio, every node (the brush needs all of them), each visible
 * uniq vertex displaced from its original coordinate, the grab modes scaled by
 * bstrength * 20, the displacement scaled by 1 - mask, and the proxy combine from the
 * original coordinates (vertex = original + displacement).
 */
#include "oracle.h"
#include "oracle_int.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

float orc_bstrength(const OrcDab *dab);

typedef struct {
  float a, b, c, f, rs[3];
} KParams;

static inline float p2(float x) { return x * x; }
static inline float p3(float x) { return p2(x) * x; }
static inline float p5(float x) { return p2(p2(x)) * x; }

static void kparams(KParams *p, float radius, float force, float shear_modulus, float poisson_ratio)
{
  p->a = 1.0f / (4.0f * (float)M_PI * shear_modulus);
  p->b = p->a / (4.0f * (1.0f - poisson_ratio));
  p->c = 2 * (3.0f * p->a - 2.0f * p->b);
  p->f = force;
  const float re[3] = {1.0f, 2.0f, 2.0f};
  p->rs[0] = radius * re[0];
  p->rs[1] = p->rs[0] * re[1];
  p->rs[2] = p->rs[1] * re[2];
}

static void grab_terms(const KParams *p, float radius, int n, float kv[3])
{
  float re[3];
  for (int i = 0; i < n; i++) {
    re[i] = sqrtf(p2(radius) + p2(p->rs[i]));
  }
  for (int i = 0; i < n; i++) {
    kv[i] = ((p->a - p->b) / re[i]) + ((p->b * p2(radius)) / p3(re[i])) +
            ((p->a * p2(p->rs[i])) / (2.0f * p3(re[i])));
  }
}

static float dist3(const float a[3], const float b[3])
{
  const float d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  return sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

static void grab_disp(const KParams *p, int mode, const float co[3], const float loc[3],
                      const float delta[3], float out[3])
{
  const float radius = dist3(loc, co);
  float kv[3], fade;
  if (mode == 0) {
    grab_terms(p, radius, 1, kv);
    fade = kv[0] * p->c;
  }
  else if (mode == 1) {
    grab_terms(p, radius, 2, kv);
    const float u = kv[0] - kv[1];
    fade = u * p->c / ((1.0f / p->rs[0]) - (1.0f / p->rs[1]));
  }
  else {
    grab_terms(p, radius, 3, kv);
    float w[3];
    w[0] = 1.0f;
    w[1] = -((p2(p->rs[2]) - p2(p->rs[0])) / (p2(p->rs[2]) - p2(p->rs[1])));
    w[2] = ((p2(p->rs[1]) - p2(p->rs[0])) / (p2(p->rs[2]) - p2(p->rs[1])));
    const float u = w[0] * kv[0] + w[1] * kv[1] + w[2] * kv[2];
    fade = u * p->c / (w[0] / p->rs[0] + w[1] / p->rs[1] + w[2] / p->rs[2]);
  }
  out[0] = delta[0] * fade;
  out[1] = delta[1] * fade;
  out[2] = delta[2] * fade;
}

/* Formula (16) (twist = 0) or (15) (twist = 1) */
static void scale_twist(const KParams *p, int twist, const float co[3], const float loc[3],
                        const float nor[3], float out[3])
{
  const float rv[3] = {co[0] - loc[0], co[1] - loc[1], co[2] - loc[2]};
  const float radius = sqrtf(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
  const float re = sqrtf(p2(radius) + p2(p->rs[0]));
  float u;
  if (twist) {
    u = -p->a * ((1.0f / p3(re))) + ((3.0f * p2(p->rs[0])) / (2.0f * p5(re)));
  }
  else {
    u = (2.0f * p->b - p->a) * ((1.0f / p3(re))) + ((3.0f * p2(p->rs[0])) / (2.0f * p5(re)));
  }
  const float fade = u * p->c;
  const float s = fade * p->f;
  if (twist) {
    const float q[3] = {nor[1] * rv[2] - nor[2] * rv[1], nor[2] * rv[0] - nor[0] * rv[2],
                        nor[0] * rv[1] - nor[1] * rv[0]};
    out[0] = q[0] * s;
    out[1] = q[1] * s;
    out[2] = q[2] * s;
  }
  else {
    out[0] = rv[0] * s;
    out[1] = rv[1] * s;
    out[2] = rv[2] * s;
  }
}

static void integrate(const KParams *p, int twist, const float co[3], const float loc[3],
                      const float nor[3], float r[3])
{
  float k[4][3], it[3];
  scale_twist(p, twist, co, loc, nor, k[0]);
  for (int i = 0; i < 3; i++) {
    it[i] = co[i] + k[0][i] * 0.5f;
  }
  scale_twist(p, twist, it, loc, nor, k[1]);
  for (int i = 0; i < 3; i++) {
    it[i] = co[i] + k[1][i] * 0.5f;
  }
  scale_twist(p, twist, it, loc, nor, k[2]);
  for (int i = 0; i < 3; i++) {
    it[i] = (co[i] + k[2][i]) - loc[i];
  }
  scale_twist(p, twist, it, loc, nor, k[3]);
  for (int i = 0; i < 3; i++) {
    float v = k[0][i];
    v += k[1][i] * 2.0f;
    v += k[2][i] * 2.0f;
    v += k[3][i];
    r[i] = v * (1.0f / 6.0f);
  }
}

static void elastic_disp(const OrcDab *dab, float bstr, const float co[3], float out[3])
{
  const float *delta = dab->grab_delta;
  const float dir = (dab->flag & ORC_BRUSH_DIR_IN) ? -1.0f : 1.0f;
  const float glen = sqrtf(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
  const float force = glen * dir * bstr;
  KParams p;
  kparams(&p, dab->radius, force, 1.0f, dab->poisson_ratio);
  if (dab->elastic_type <= 2) {
    grab_disp(&p, dab->elastic_type, co, dab->center, delta, out);
    const float g = bstr * 20.0f;
    out[0] = out[0] * g;
    out[1] = out[1] * g;
    out[2] = out[2] * g;
  }
  else {
    integrate(&p, dab->elastic_type == 4, co, dab->center, dab->surface_normal, out);
  }
}

void orc_kelvinlet(const OrcDab *dab, const float *co, const float *bstr, float *out)
{
  elastic_disp(dab, bstr[0], co, out);
}

int orc_dab_elastic(OrcPBVH *o, const OrcDab *dab, const float *orco, int *out_affected, int cap)
{
  memset(o->stats, 0, sizeof(o->stats));
  if (!o->nodes || o->totnode == 0 || dab->tool != ORC_TOOL_ELASTIC) {
    return 0;
  }
  const float bstr = orc_bstrength(dab);
  int nleaf = 0, naff = 0, w = 0;
  for (int i = 0; i < o->totnode; i++) {
    ONode *n = &o->nodes[i];
    if (!(n->flag & ORC_Leaf)) {
      continue;
    }
    nleaf++;
    n->flag |= ORC_UpdateNormals | ORC_UpdateBB | ORC_UpdateOriginalBB | ORC_UpdateDrawBuffers |
               ORC_UpdateRedraw;
    for (int k = 0; k < n->uniq_verts; k++) {
      const int v = n->vert_indices[k];
      if (o->respect_hide && o->vhide && o->vhide[v]) {
        continue;
      }
      const float *oc = orco + 3 * (size_t)v;
      float disp[3];
      elastic_disp(dab, bstr, oc, disp);
      if (o->mask) {
        const float keep = 1.0f - o->mask[v];
        disp[0] = disp[0] * keep;
        disp[1] = disp[1] * keep;
        disp[2] = disp[2] * keep;
      }
      float *co = o->co + 3 * (size_t)v;
      co[0] = oc[0] + disp[0];
      co[1] = oc[1] + disp[1];
      co[2] = oc[2] + disp[2];
      if (o->co64) {
        for (int q = 0; q < 3; q++) {
          o->co64[3 * (size_t)v + q] = co[q];
        }
      }
      o->vert_bitmap[v] = 1;
      naff++;
      if (out_affected && w < cap) {
        out_affected[w++] = v;
      }
    }
  }
  o->stats[0] = nleaf;
  o->stats[5] = naff;
  o->stats[6] = orc_update_normals_bounds(o);
  if (out_affected) {
    for (int i = 1; i < w; i++) { /* insertion into ascending order (small cases) */
      int x = out_affected[i], j = i - 1;
      while (j >= 0 && out_affected[j] > x) {
        out_affected[j + 1] = out_affected[j];
        j--;
      }
      out_affected[j + 1] = x;
    }
  }
  return naff;
}
