// Device helpers shared by the dab kernels (kernels.hip) and the learned Smooth
// kernels (learned.hip).  FP32 expressions are written in the exact order the brush
// spec fixes; both files are compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "sg_kernels.h"

#define SG_WAVE 64

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sg_min(float a, float b) { return (a < b) ? a : b; }  // min_ff
__device__ __forceinline__ float sg_max(float a, float b) { return (a > b) ? a : b; }  // max_ff

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// prefix popcount of the lanes below this one in a 64-bit mask
__device__ __forceinline__ int mask_prefix(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// normalize_v3 (math_vector_inline.c:1165-1180): zero below 1e-35, else multiply by
// the rounded reciprocal of the correctly rounded sqrt.
__device__ __forceinline__ float sg_normalize(float &x, float &y, float &z) {
  float d = x * x + y * y + z * z;
  if (d > 1.0e-35f) {
    d = sqrtf(d);
    const float inv = 1.0f / d;
    x = x * inv;
    y = y * inv;
    z = z * inv;
  } else {
    x = y = z = 0.0f;
    d = 0.0f;
  }
  return d;
}

// normal_tri_v3 (math_geom.cc:31-46)
__device__ __forceinline__ void sg_tri_normal(float ax, float ay, float az, float bx, float by,
                                              float bz, float cx, float cy, float cz, float &nx,
                                              float &ny, float &nz) {
  const float n1x = ax - bx, n2x = bx - cx;
  const float n1y = ay - by, n2y = by - cy;
  const float n1z = az - bz, n2z = bz - cz;
  nx = n1y * n2z - n1z * n2y;
  ny = n1z * n2x - n1x * n2z;
  nz = n1x * n2y - n1y * n2x;
  sg_normalize(nx, ny, nz);
}

// BKE_mesh_calc_poly_normal (mesh_evaluate.c:62-84) of polygon q, corners read from
// the position arrays (x, y, z): more than 4 corners Newell's method from the last
// corner (mesh_calc_ngon_normal :39-60, add_newell_cross_v3_v3v3
// math_vector_inline.c:961-966; a zero result is +z), 3 normal_tri_v3, 4 normal_quad_v3
// (math_geom.cc:48-68: the cross product of the diagonals), fewer +z.
__device__ __forceinline__ void sg_poly_normal(const int *__restrict__ poly_off, const int *__restrict__ poly_v,
                                               const float *__restrict__ x, const float *__restrict__ y,
                                               const float *__restrict__ z, int q, float &nx, float &ny,
                                               float &nz) {
  const int b = poly_off[q], n = poly_off[q + 1] - b;
  if (n == 3 || n == 4) {
    int c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = k < n ? poly_v[b + k] : 0;
    float P[4][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (k < n) {
        P[k][0] = x[c[k]];
        P[k][1] = y[c[k]];
        P[k][2] = z[c[k]];
      }
    }
    if (n == 3) {
      sg_tri_normal(P[0][0], P[0][1], P[0][2], P[1][0], P[1][1], P[1][2], P[2][0], P[2][1], P[2][2], nx, ny, nz);
    } else {
      const float n10 = P[0][0] - P[2][0], n11 = P[0][1] - P[2][1], n12 = P[0][2] - P[2][2];
      const float n20 = P[1][0] - P[3][0], n21 = P[1][1] - P[3][1], n22 = P[1][2] - P[3][2];
      nx = n11 * n22 - n12 * n21;
      ny = n12 * n20 - n10 * n22;
      nz = n10 * n21 - n11 * n20;
      sg_normalize(nx, ny, nz);
    }
  } else if (n > 4) {
    nx = ny = nz = 0.0f;
    int u = poly_v[b + n - 1];
    float ax = x[u], ay = y[u], az = z[u];
    for (int k = 0; k < n; k++) {
      u = poly_v[b + k];
      const float cx = x[u], cy = y[u], cz = z[u];
      nx += (ay - cy) * (az + cz);
      ny += (az - cz) * (ax + cx);
      nz += (ax - cx) * (ay + cy);
      ax = cx;
      ay = cy;
      az = cz;
    }
    if (sg_normalize(nx, ny, nz) == 0.0f) nz = 1.0f;
  } else {
    nx = 0.0f;
    ny = 0.0f;
    nz = 1.0f;
  }
}

// Ordered bounding-box combine: `b` covers later elements than `a`.  Equal to
// BB_expand_with_bb(a, b) (pbvh.c:66-72), which keeps the later value on ties.
struct BB6 {
  float mn[3], mx[3];
};
__device__ __forceinline__ void bb_reset(BB6 &b) {
  b.mn[0] = b.mn[1] = b.mn[2] = FLT_MAX;
  b.mx[0] = b.mx[1] = b.mx[2] = -FLT_MAX;
}
__device__ __forceinline__ void bb_add_point(BB6 &b, float x, float y, float z) {
  b.mn[0] = sg_min(b.mn[0], x);
  b.mn[1] = sg_min(b.mn[1], y);
  b.mn[2] = sg_min(b.mn[2], z);
  b.mx[0] = sg_max(b.mx[0], x);
  b.mx[1] = sg_max(b.mx[1], y);
  b.mx[2] = sg_max(b.mx[2], z);
}
__device__ __forceinline__ void bb_add_bb(BB6 &a, const BB6 &b) {
  for (int i = 0; i < 3; i++) {
    a.mn[i] = sg_min(a.mn[i], b.mn[i]);
    a.mx[i] = sg_max(a.mx[i], b.mx[i]);
  }
}
__device__ __forceinline__ BB6 bb_load(const float *p) {
  const float4 a = *(const float4 *)p;
  const float4 b = *(const float4 *)(p + 4);
  BB6 r;
  r.mn[0] = a.x;
  r.mn[1] = a.y;
  r.mn[2] = a.z;
  r.mx[0] = b.x;
  r.mx[1] = b.y;
  r.mx[2] = b.z;
  return r;
}
__device__ __forceinline__ void bb_store(float *p, const BB6 &b) {
  *(float4 *)p = make_float4(b.mn[0], b.mn[1], b.mn[2], 0.0f);
  *(float4 *)(p + 4) = make_float4(b.mx[0], b.mx[1], b.mx[2], 0.0f);
}
// wave-ordered reduction: lane i's box covers elements before lane i+1's.
// The same ordered fold by an inclusive scan over DPP lane moves (VALU, no LDS round
// trips): each lane folds in the partial box of the lanes before it, row by row
// (row_shr 1, 2, 4, 8), then across rows (row_bcast 15, 31); lane 63 holds the fold of
// all 64 lanes in lane order.  Lanes whose source lies outside their row keep their
// own value (DPP returns `old`), and folding a box with itself changes nothing.
// Whole wave active.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_from(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROWS, 0xf, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ void bb_scan_step(BB6 &b) {
  BB6 e;
  for (int i = 0; i < 3; i++) {
    e.mn[i] = dpp_from<CTRL, ROWS>(b.mn[i]);
    e.mx[i] = dpp_from<CTRL, ROWS>(b.mx[i]);
  }
  bb_add_bb(e, b);  // earlier lanes first
  b = e;
}
__device__ __forceinline__ BB6 bb_wave_scan(BB6 b) {
  bb_scan_step<0x111, 0xf>(b);  // row_shr:1
  bb_scan_step<0x112, 0xf>(b);  // row_shr:2
  bb_scan_step<0x114, 0xf>(b);  // row_shr:4
  bb_scan_step<0x118, 0xf>(b);  // row_shr:8
  bb_scan_step<0x142, 0xa>(b);  // row_bcast:15 into rows 1 and 3
  bb_scan_step<0x143, 0xc>(b);  // row_bcast:31 into rows 2 and 3
  return b;                     // lane 63 holds the ordered fold
}

__device__ __forceinline__ BB6 bb_wave_reduce(BB6 b) {
  for (int off = 1; off < SG_WAVE; off <<= 1) {
    BB6 o;
    for (int i = 0; i < 3; i++) {
      o.mn[i] = __shfl_down(b.mn[i], off);
      o.mx[i] = __shfl_down(b.mx[i], off);
    }
    // lane l combines with lane l+off (later elements); lanes whose partner is out
    // of range keep their own value (their result is never used)
    if (lane_id() + off < SG_WAVE && (lane_id() & (2 * off - 1)) == 0) bb_add_bb(b, o);
  }
  return b;  // lane 0 holds the ordered combination
}

// A dab placed by a ray cast on the GPU reads its centre (and Grab translation) from
// the placement slot the ray cast wrote: cx cy cz gx gy gz hit 0.  A ray that missed
// leaves the dab empty (no box or vertex passes a negative squared radius).
__device__ __forceinline__ void sg_place(SgDabK &k, const float *place) {
  if (!place) return;
  const float4 a = *(const float4 *)place;
  const float4 b = *(const float4 *)(place + 4);
  k.cx = a.x;
  k.cy = a.y;
  k.cz = a.z;
  k.gx = a.w;
  k.gy = b.x;
  k.gz = b.y;
  if (b.z == 0.0f) {
    k.r2 = -1.0f;
    k.nr2 = -1.0f;
  }
}

// Affected-list slot t: its vertex when slot t holds a record, else -1 (gathered leaf
// t >> aff_shift owns the slots below its count).
__device__ __forceinline__ int aff_slot_vertex(const SgDevView &d, int t, int cnt) {
  return (t & ((1 << d.aff_shift) - 1)) < cnt ? d.aff_v[t] : -1;
}

// Undo recording: save a leaf the running dab is about to write, once per step (one
// wave; the leaf's uniq vertices are the contiguous range vr): positions, normals and
// the leaf's two boxes, as they were before the dab.
__device__ __forceinline__ bool undo_active(const SgDevView &d) { return d.undo && d.undo->active; }

// Caller has seen undo_active(d).  (sx, sy, sz): the position arrays that still hold
// the leaf's positions before the dab.  Returns the record's first vertex in the arena
// when this call saved the leaf, -1 when it was saved before or the arena is full.
__device__ __forceinline__ long long undo_save_leaf(const SgDevView &d, int leaf, int2 vr, const float *sx,
                                                    const float *sy, const float *sz) {
  SgUndoCtl *u = d.undo;
  const int lane = lane_id();
  const int n = vr.y - vr.x;
  // the leaf's first 128 vertices and its two boxes are loaded while lane 0 checks the
  // stamp and reserves the record, so the copy below only stores them
  float c0[2][6];
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int i = lane + j * SG_WAVE;
    if (i < n) {
      const int v = vr.x + i;
      c0[j][0] = sx[v];
      c0[j][1] = sy[v];
      c0[j][2] = sz[v];
      c0[j][3] = d.nx[v];
      c0[j][4] = d.ny[v];
      c0[j][5] = d.nz[v];
    }
  }
  const float bbv = lane < 16 ? (lane < 8 ? d.leaf_bb[8 * (size_t)leaf + lane] : d.leaf_obb[8 * (size_t)leaf + lane - 8])
                              : 0.0f;
  int e = -1;
  unsigned long long vb = 0;
  if (lane == 0 && u->stamp[leaf] != u->step) {
    e = atomicAdd(&u->n_e, 1);
    vb = atomicAdd(&u->n_v, (unsigned long long)n);
    if (e >= u->cap_e || vb + (unsigned long long)n > u->cap_v) {
      u->overflow = 1;
      e = -1;
    } else {
      u->tab[e] = make_int4(leaf, n, (int)(unsigned)vb, (int)(unsigned)(vb >> 32));
      u->slot[leaf] = e;
      u->stamp[leaf] = u->step;
    }
  }
  e = __shfl(e, 0);
  if (e < 0) return -1;
  vb = ((unsigned long long)(unsigned)__shfl((int)(unsigned)(vb >> 32), 0) << 32) |
       (unsigned)__shfl((int)(unsigned)vb, 0);
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int i = lane + j * SG_WAVE;
    if (i < n) {
      const size_t o = 3 * (size_t)(vb + i);
      u->pos[o] = c0[j][0];
      u->pos[o + 1] = c0[j][1];
      u->pos[o + 2] = c0[j][2];
      u->nor[o] = c0[j][3];
      u->nor[o + 1] = c0[j][4];
      u->nor[o + 2] = c0[j][5];
    }
  }
  for (int i = lane + 2 * SG_WAVE; i < n; i += SG_WAVE) {
    const int v = vr.x + i;
    const size_t o = 3 * (size_t)(vb + i);
    u->pos[o] = sx[v];
    u->pos[o + 1] = sy[v];
    u->pos[o + 2] = sz[v];
    u->nor[o] = d.nx[v];
    u->nor[o + 1] = d.ny[v];
    u->nor[o + 2] = d.nz[v];
  }
  if (lane < 16) u->bbs[16 * (size_t)e + lane] = bbv;
  return (long long)vb;
}

// Spec S1 sphere-vs-box test: nearest point of the box, strict '<' against r*r.
__device__ __forceinline__ bool sphere_box(const BB6 &b, float cx, float cy, float cz, float r2) {
  const float c[3] = {cx, cy, cz};
  float t[3];
  for (int i = 0; i < 3; i++) {
    float nearest;
    if (b.mn[i] > c[i])
      nearest = b.mn[i];
    else if (b.mx[i] < c[i])
      nearest = b.mx[i];
    else
      nearest = c[i];
    t[i] = c[i] - nearest;
  }
  return (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) < r2;
}

// Subtree box of a search cut from its leaf boxes (used by the search's refresh).
#define SEARCH_UNROLL 8

__device__ __forceinline__ void leaves_box(const float *bbs, const float *obbs, int lb, int le, BB6 &nb,
                                           BB6 &ob) {
  // ordered BB_expand_with_bb over leaf boxes [lb, le) (current and original): lane l
  // folds a contiguous chunk of `per` leaves (all loads issued together), then the
  // wave combines lanes in order
  const int lane = lane_id();
  const int per = (le - lb + SG_WAVE - 1) / SG_WAVE;
  bb_reset(nb);
  bb_reset(ob);
  for (int q0 = 0; q0 < per; q0 += SEARCH_UNROLL) {
    BB6 t[SEARCH_UNROLL], u[SEARCH_UNROLL];
#pragma unroll
    for (int j = 0; j < SEARCH_UNROLL; j++) {
      const int l = lb + lane * per + q0 + j;
      if (q0 + j < per && l < le) {
        t[j] = bb_load(bbs + 8 * (size_t)l);
        u[j] = bb_load(obbs + 8 * (size_t)l);
      } else {
        bb_reset(t[j]);
        bb_reset(u[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < SEARCH_UNROLL; j++) {
      bb_add_bb(nb, t[j]);
      bb_add_bb(ob, u[j]);
    }
  }
  nb = bb_wave_reduce(nb);
  ob = bb_wave_reduce(ob);
}


// Spec S3 falloff: hardness remap, then the preset curve, clamped to [0,1].
__device__ __forceinline__ float sg_falloff(const SgDabK &k, const float *curve, float dist) {
  const float r = k.r, h = k.hardness;
  {
    float pr = dist / r;
    if (pr < h) {
      dist = 0.0f;
    } else if (h == 1.0f) {
      dist = r;
    } else {
      pr = (pr - h) / (1.0f - h);
      dist = pr * r;
    }
  }
  if (dist >= r) return 0.0f;
  float p = dist / r;
  p = 1.0f - p;
  float s;
  switch (k.falloff) {
    case SG_CURVE_CUSTOM:
      if (curve) {
        const float value = 1.0f - p;
        float fi = value * (float)SG_CM_TABLE;
        const int i = (int)fi;
        if (fi < 0.0f || fi > (float)SG_CM_TABLE) {
          s = fi < 0.0f ? curve[0] : curve[SG_CM_TABLE];
        } else if (i < 0) {
          s = curve[0];
        } else if (i >= SG_CM_TABLE) {
          s = curve[SG_CM_TABLE];
        } else {
          fi = fi - (float)i;
          s = (1.0f - fi) * curve[i] + (fi)*curve[i + 1];
        }
      } else {
        s = 1.0f;
      }
      break;
    case SG_CURVE_SHARP:
      s = p * p;
      break;
    case SG_CURVE_SMOOTHER:
      s = ((p * p) * p) * (6.0f * p * p - 15.0f * p + 10.0f);
      break;
    case SG_CURVE_ROOT:
      s = sqrtf(p);
      break;
    case SG_CURVE_LIN:
      s = p;
      break;
    case SG_CURVE_CONSTANT:
      s = 1.0f;
      break;
    case SG_CURVE_SPHERE:
      s = sqrtf(2.0f * p - p * p);
      break;
    case SG_CURVE_POW4:
      s = p * p * p * p;
      break;
    case SG_CURVE_INVSQUARE:
      s = p * (2.0f - p);
      break;
    case SG_CURVE_SMOOTH:
    default:
      s = 3.0f * p * p - 2.0f * p * p * p;
      break;
  }
  if (s < 0.0f)
    s = 0.0f;
  else if (s > 1.0f)
    s = 1.0f;
  return s;
}

__device__ __forceinline__ unsigned long long as_u64(long long x) { return (unsigned long long)x; }

// wave-wide sums
__device__ __forceinline__ long long wave_sum_ll(long long v) {
  for (int off = SG_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
  for (int off = SG_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Sum of the (previous-iterate) positions of v's edge neighbours in CSR order, with
// the neighbour ids and then their positions fetched eight at a time.
__device__ __forceinline__ void nbr_sum(const int *__restrict__ vv_idx, int b, int e, const float *__restrict__ sx_,
                                        const float *__restrict__ sy_, const float *__restrict__ sz_, float &sx,
                                        float &sy, float &sz) {
  sx = 0.0f;
  sy = 0.0f;
  sz = 0.0f;
  for (int q0 = b; q0 < e; q0 += 8) {
    int u[8];
#pragma unroll
    for (int j = 0; j < 8; j++) u[j] = (q0 + j < e) ? vv_idx[q0 + j] : -1;
    float x[8], y[8], z[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (u[j] >= 0) {
        x[j] = sx_[u[j]];
        y[j] = sy_[u[j]];
        z[j] = sz_[u[j]];
      }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (u[j] >= 0) {
        sx = sx + x[j];
        sy = sy + y[j];
        sz = sz + z[j];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// K3: brush.  One wave per gathered leaf (grid-stride); lanes over the leaf's uniq
// vertices (one coalesced range).  Spec S2 vertex test, S3/S4 factor, S6-S9
// displacement.  Affected vertices are compacted with ballot + prefix popcount and
// one atomic per wave into records (vertex, its vertex->vertex and vertex->tri CSR
// ranges) so the later kernels skip those lookups.  Positions are written to both
// the position array and its shadow copy.  Smooth runs its first Jacobi step here,
// reading the untouched positions and writing the shadow copy.
// ---------------------------------------------------------------------------

// Normal of vertex v (spec S10): normalize(sum over incident tris, in CSR order, of
// normal_tri_v3).  Deterministic gather, no atomics.  The incident tris' corner ids
// are stored flattened in CSR order (vt4), so a batch of eight tris needs one load
// level for the corners and one for their positions.
__device__ __forceinline__ void vertex_normal(const SgDevView &d, int v, int b, int e, int stats, int stamp,
                                              long long &ntri, long long &nfa, long long &nvfa) {
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  if (d.poly_of_tri) {
    // a mesh with polygons: each incident looptri adds its polygon's normal
    // (pbvh.c:1041-1058: BKE_mesh_calc_poly_normal once per looptri)
    for (int q = b; q < e; q++) {
      const int t = d.vt4[q].w & 0x3fffffff;
      float fx, fy, fz;
      sg_poly_normal(d.poly_off, d.poly_v, d.px, d.py, d.pz, d.poly_of_tri[t], fx, fy, fz);
      ax = ax + fx;
      ay = ay + fy;
      az = az + fz;
    }
    ntri += e - b;
    sg_normalize(ax, ay, az);
    d.nx[v] = ax;
    d.ny[v] = ay;
    d.nz[v] = az;
    return;
  }
  for (int q0 = b; q0 < e; q0 += 8) {
    int4 c[8];
#pragma unroll
    for (int j = 0; j < 8; j++) c[j] = (q0 + j < e) ? d.vt4[q0 + j] : make_int4(-1, -1, -1, -1);
    float P[8][9];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (c[j].x >= 0) {
        P[j][0] = d.px[c[j].x];
        P[j][1] = d.py[c[j].x];
        P[j][2] = d.pz[c[j].x];
        P[j][3] = d.px[c[j].y];
        P[j][4] = d.py[c[j].y];
        P[j][5] = d.pz[c[j].y];
        P[j][6] = d.px[c[j].z];
        P[j][7] = d.py[c[j].z];
        P[j][8] = d.pz[c[j].z];
      }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (c[j].x >= 0) {
        float fx, fy, fz;
        sg_tri_normal(P[j][0], P[j][1], P[j][2], P[j][3], P[j][4], P[j][5], P[j][6], P[j][7], P[j][8], fx, fy,
                      fz);
        ax = ax + fx;
        ay = ay + fy;
        az = az + fz;
      }
    }
    if (stats) {
      for (int j = 0; j < 8; j++) {
        if (c[j].x < 0) continue;
        const int t = c[j].w & 0x3fffffff;
        if (atomicExch(&d.stamp_t[t], stamp) != stamp) {
          nfa++;
          if (atomicExch(&d.stamp_v2[c[j].x], stamp) != stamp) nvfa++;
          if (atomicExch(&d.stamp_v2[c[j].y], stamp) != stamp) nvfa++;
          if (atomicExch(&d.stamp_v2[c[j].z], stamp) != stamp) nvfa++;
        }
      }
    }
  }
  ntri += e - b;
  sg_normalize(ax, ay, az);
  d.nx[v] = ax;
  d.ny[v] = ay;
  d.nz[v] = az;
}

// Leaf bounds over ALL its vertices, uniq then shared (update_node_vb pbvh.c:108-128,
// BB_expand in vert_indices order).  Each lane folds a contiguous chunk, then the
// wave combines lanes in order.  Writes vb and orig_vb (UpdateOriginalBB).
__device__ __forceinline__ BB6 leaf_box(const SgDevView &d, int2 vr, int2 sr);
__device__ __forceinline__ void leaf_refit(const SgDevView &d, int leaf, int2 vr, int2 sr, long long &nsh) {
  const BB6 acc = leaf_box(d, vr, sr);
  if (lane_id() == 0) {
    bb_store(d.leaf_bb + 8 * (size_t)leaf, acc);
    bb_store(d.leaf_obb + 8 * (size_t)leaf, acc);
    nsh += sr.y - sr.x;
  }
}

// The box over a leaf's vertices (uniq range vr, then its shared vertices sr) in
// vert_indices order; lane 0 holds the result.
__device__ __forceinline__ BB6 leaf_box(const SgDevView &d, int2 vr, int2 sr) {
  const int nu = vr.y - vr.x, ns = sr.y - sr.x, tot = nu + ns;
  const int per = (tot + SG_WAVE - 1) / SG_WAVE;
  const int lane = lane_id();
  BB6 acc;
  bb_reset(acc);
  for (int q0 = 0; q0 < per; q0 += 8) {
    int vid[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int i = lane * per + q0 + j;
      vid[j] = -1;
      if (q0 + j < per && i < tot) vid[j] = (i < nu) ? vr.x + i : d.leaf_sh[sr.x + i - nu];
    }
    float x[8], y[8], z[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (vid[j] >= 0) {
        x[j] = d.px[vid[j]];
        y[j] = d.py[vid[j]];
        z[j] = d.pz[vid[j]];
      }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (vid[j] >= 0) bb_add_point(acc, x[j], y[j], z[j]);
    }
  }
  return bb_wave_reduce(acc);
}


// Per-vertex inputs of the brush, loaded together for two vertices per lane.
struct BrushIn {
  float x, y, z, d2, nx, ny, nz, mk;
  int2 vt, vv;
  bool aff;
};

__device__ __forceinline__ BrushIn brush_load(const SgDevView &d, const SgDabK &k, int v, int vend, bool smooth) {
  BrushIn in;
  in.aff = false;
  in.x = in.y = in.z = in.d2 = in.nx = in.ny = in.nz = in.mk = 0.0f;
  in.vt = make_int2(0, 0);
  in.vv = make_int2(0, 0);
  if (v < vend) {
    in.x = d.px[v];
    in.y = d.py[v];
    in.z = d.pz[v];
    in.vt = make_int2(d.vt_off[v], d.vt_off[v + 1]);
    if (smooth) in.vv = make_int2(d.vv_off[v], d.vv_off[v + 1]);
    if (k.tool == SG_TOOL_INFLATE) {
      in.nx = d.nx[v];
      in.ny = d.ny[v];
      in.nz = d.nz[v];
    }
    if (d.mask) in.mk = d.mask[v];
    const float dx = k.cx - in.x;
    const float dy = k.cy - in.y;
    const float dz = k.cz - in.z;
    in.d2 = dx * dx + dy * dy + dz * dz;
    in.aff = in.d2 <= k.r2;
    if (d.vhide && d.respect_hide && d.vhide[v]) in.aff = false;
  }
  return in;
}

// Spec S3/S4 factor and S6-S9 displacement of one affected vertex; returns the Smooth
// factor (0 for other tools).  Smooth step 1 reads the untouched positions and writes
// the shadow copy; the other tools write both copies.
__device__ __forceinline__ float brush_apply(const SgDevView &d, const SgDabK &k, const float *curve, int v,
                                             const BrushIn &in, float ox, float oy, float oz, long long &nbr) {
  const float f = sg_falloff(k, curve, sqrtf(in.d2)) * (1.0f - in.mk);
  float rx = in.x, ry = in.y, rz = in.z, fade = 0.0f;
  if (k.tool == SG_TOOL_DRAW) {
    rx = in.x + ox * f;
    ry = in.y + oy * f;
    rz = in.z + oz * f;
  } else if (k.tool == SG_TOOL_INFLATE) {
    const float s = (k.bstr * f) * k.r;
    rx = in.x + in.nx * s;
    ry = in.y + in.ny * s;
    rz = in.z + in.nz * s;
  } else if (k.tool == SG_TOOL_GRAB) {
    fade = k.bstr * f;
    rx = in.x + k.gx * fade;
    ry = in.y + k.gy * fade;
    rz = in.z + k.gz * fade;
  } else {
    fade = k.bstr * f;
    float sx, sy, sz;
    nbr_sum(d.vv_idx, in.vv.x, in.vv.y, d.px, d.py, d.pz, sx, sy, sz);
    nbr += in.vv.y - in.vv.x;
    if (in.vv.y > in.vv.x) {
      const float inv = 1.0f / (float)(in.vv.y - in.vv.x);
      const float avx = sx * inv, avy = sy * inv, avz = sz * inv;
      rx = in.x + (avx - in.x) * fade;
      ry = in.y + (avy - in.y) * fade;
      rz = in.z + (avz - in.z) * fade;
    }
  }
  if (k.tool != SG_TOOL_SMOOTH) {
    d.px[v] = rx;
    d.py[v] = ry;
    d.pz[v] = rz;
  }
  d.qx[v] = rx;
  d.qy[v] = ry;
  d.qz[v] = rz;
  return fade;
}

