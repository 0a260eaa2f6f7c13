// Ray cast on the GPU (gfx950, wave64): BKE_pbvh_raycast (kernel/intern/pbvh.c:3915-3928)
// over the occluded search (:934-970) with the sculpt operator's hit callback, and the
// leaf test BKE_pbvh_node_raycast -> pbvh_faces_node_raycast (:4041-4107, :4203-4260).
//
// What the reference computes, per ray: every leaf whose box and ancestors' boxes pass
// isect_ray_aabb_v3 (math_geom.cc:3032-3080) is visited in ascending tmin (ties in
// left->right leaf order: node_tree_insert :878-896 sends ties right); a leaf is
// ray-tested only while its tmin is below the running bound (FLT_MAX, then the hit
// depth), and a triangle replaces the hit only when strictly nearer (:3951-3966).
//
// How it runs here, one workgroup per ray (256 threads per ray of a batch, 1024 for the
// single ray that places a dab):
//   1. the search cut's subtree boxes (stale ones, touched by the previous dab and not
//      yet refreshed, are kept: their leaves are tested anyway)
//   2. the leaves of the passing subtrees, flattened over the workgroup
//   3. every passing leaf's triangles (a wave per leaf): the leaf's first nearest hit
//      as a 64-bit key (depth bits, position in leaf order), kept when below the ray's
//      depth bound
//   4. one wave orders those leaves by (tmin, leaf) and folds them exactly as the
//      reference's visit does (the skip rule included), then writes the result.
// Only leaves with a hit can change the result, so the fold over them equals the
// reference's visit over all passing leaves.
//
// Box nesting: with every 1/dir finite, (b - o) * inv is monotone in b, so a parent's
// slab interval contains its child's and a passing leaf implies passing ancestors; the
// leaf test alone is then exact.  When some 1/dir is infinite, (b - o) * inv can be
// NaN and the comparisons stop nesting, so each passing leaf also walks its ancestors
// (their boxes must be current: the host refits internal nodes first).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "sg_internal.h"
#include "sg_kernels.h"
#include "sg_device.h"

// Workgroup sizes: a batch of rays runs 256 threads per ray (several rays in flight per
// CU hide each ray's dependent load levels); a single placement ray runs 1024.
#define RAY_LEAF_UNROLL 2
#define RAY_HITS 512  // leaves with a hit, per ray (both sizes: the same limit)

struct SgRayView {
  const int4 *ltri;        // [T] tris in leaf order: new corner ids, .w = looptri | hidden << 31
  const int *toff;         // [L+1] each leaf's range in ltri
  const int *tri_poly;     // [T] looptri -> poly
  const int *node_parent;  // [N] (-1 at the root)
  const int *leaf_node;    // [L] DFS leaf -> node id
};

// isect_ray_aabb_v3_precalc + isect_ray_tri_watertight_v3_precalc (math_geom.cc:3017-3030,
// :1755-1780) of one ray.
struct RayPre {
  float o[3], dir[3], inv[3];
  int sgn[3];
  int kx, ky, kz;
  float sx, sy, sz;
  float depth0;
  bool slow;
};

__device__ __forceinline__ float comp(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

__device__ __forceinline__ RayPre ray_pre(const SgRay *r) {
  const float4 a = *(const float4 *)r;
  const float4 b = *(const float4 *)((const float *)r + 4);
  RayPre p;
  p.o[0] = a.x;
  p.o[1] = a.y;
  p.o[2] = a.z;
  p.dir[0] = a.w;
  p.dir[1] = b.x;
  p.dir[2] = b.y;
  p.depth0 = b.z;
  p.slow = false;
  for (int i = 0; i < 3; i++) {
    p.inv[i] = 1.0f / p.dir[i];
    p.sgn[i] = p.inv[i] < 0.0f;
    p.slow |= !isfinite(p.inv[i]);
  }
  // axis_dominant_v3_single (math_geom_inline.c:76-82)
  const float x = fabsf(p.dir[0]), y = fabsf(p.dir[1]), z = fabsf(p.dir[2]);
  const int kz = (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2);
  int kx = (kz != 2) ? (kz + 1) : 0;
  int ky = (kx != 2) ? (kx + 1) : 0;
  const float dz = comp(p.dir[0], p.dir[1], p.dir[2], kz);
  if (dz < 0.0f) {  // keep the winding
    const int t = kx;
    kx = ky;
    ky = t;
  }
  const float inv_z = 1.0f / dz;
  p.sx = comp(p.dir[0], p.dir[1], p.dir[2], kx) * inv_z;
  p.sy = comp(p.dir[0], p.dir[1], p.dir[2], ky) * inv_z;
  p.sz = inv_z;
  p.kx = kx;
  p.ky = ky;
  p.kz = kz;
  return p;
}

// isect_ray_aabb_v3 (math_geom.cc:3032-3080): writes tmin only on a pass.
__device__ __forceinline__ bool ray_box(const RayPre &r, const BB6 &b, float &tmin_out) {
  float tmin = ((r.sgn[0] ? b.mx[0] : b.mn[0]) - r.o[0]) * r.inv[0];
  float tmax = ((r.sgn[0] ? b.mn[0] : b.mx[0]) - r.o[0]) * r.inv[0];
  const float tymin = ((r.sgn[1] ? b.mx[1] : b.mn[1]) - r.o[1]) * r.inv[1];
  const float tymax = ((r.sgn[1] ? b.mn[1] : b.mx[1]) - r.o[1]) * r.inv[1];
  if ((tmin > tymax) || (tymin > tmax)) return false;
  if (tymin > tmin) tmin = tymin;
  if (tymax < tmax) tmax = tymax;
  const float tzmin = ((r.sgn[2] ? b.mx[2] : b.mn[2]) - r.o[2]) * r.inv[2];
  const float tzmax = ((r.sgn[2] ? b.mn[2] : b.mx[2]) - r.o[2]) * r.inv[2];
  if ((tmin > tzmax) || (tzmin > tmax)) return false;
  if (tzmin > tmin) tmin = tzmin;
  tmin_out = tmin;
  return true;
}

// isect_ray_tri_watertight_v3 (math_geom.cc:1782-1855).
__device__ __forceinline__ bool ray_tri(const RayPre &r, float v0x, float v0y, float v0z, float v1x, float v1y,
                                        float v1z, float v2x, float v2y, float v2z, float &lambda) {
  const float ax_ = v0x - r.o[0], ay_ = v0y - r.o[1], az_ = v0z - r.o[2];
  const float bx_ = v1x - r.o[0], by_ = v1y - r.o[1], bz_ = v1z - r.o[2];
  const float cx_ = v2x - r.o[0], cy_ = v2y - r.o[1], cz_ = v2z - r.o[2];
  const float a_kx = comp(ax_, ay_, az_, r.kx), a_ky = comp(ax_, ay_, az_, r.ky), a_kz = comp(ax_, ay_, az_, r.kz);
  const float b_kx = comp(bx_, by_, bz_, r.kx), b_ky = comp(bx_, by_, bz_, r.ky), b_kz = comp(bx_, by_, bz_, r.kz);
  const float c_kx = comp(cx_, cy_, cz_, r.kx), c_ky = comp(cx_, cy_, cz_, r.ky), c_kz = comp(cx_, cy_, cz_, r.kz);
  const float ax = a_kx - r.sx * a_kz;
  const float ay = a_ky - r.sy * a_kz;
  const float bx = b_kx - r.sx * b_kz;
  const float by = b_ky - r.sy * b_kz;
  const float cx = c_kx - r.sx * c_kz;
  const float cy = c_ky - r.sy * c_kz;
  const float u = cx * by - cy * bx;
  const float v = ax * cy - ay * cx;
  const float w = bx * ay - by * ax;
  if ((u < 0.0f || v < 0.0f || w < 0.0f) && (u > 0.0f || v > 0.0f || w > 0.0f)) return false;
  const float det = u + v + w;
  if (det == 0.0f || !isfinite(det)) return false;
  const float t = (u * a_kz + v * b_kz + w * c_kz) * r.sz;
  const float sign_t = __uint_as_float(__float_as_uint(t) ^ (__float_as_uint(det) & 0x80000000u));
  if (sign_t < 0.0f) return false;
  const float inv_det = 1.0f / det;
  lambda = t * inv_det;
  return true;
}

// Orderable key of a hit: depth bits (a hit's depth is +0, -0 or positive; -0 counts as
// +0, as '<' does), then the triangle's position in leaf order (first of equals wins).
__device__ __forceinline__ unsigned long long hit_key(float lambda, int pos) {
  unsigned u = __float_as_uint(lambda);
  if (u == 0x80000000u) u = 0u;
  return ((unsigned long long)u << 32) | (unsigned)pos;
}
__device__ __forceinline__ float key_depth(unsigned long long k) { return __uint_as_float((unsigned)(k >> 32)); }

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long k) {
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned lo = __shfl_xor((unsigned)k, off);
    const unsigned hi = __shfl_xor((unsigned)(k >> 32), off);
    const unsigned long long o = ((unsigned long long)hi << 32) | lo;
    k = o < k ? o : k;
  }
  return k;
}

template <int NT>
struct RayShared {
  static constexpr int kRound = NT * RAY_LEAF_UNROLL;  // leaves tested per round
  static constexpr int kCand = kRound + 256;            // passing leaves held before their scan
  int cut_lb[1024];                                     // passing subtrees: first leaf, leaf count
  int cut_n[1024];
  int pre[1025];
  int cand_leaf[kCand];  // passing leaves of the round: leaf, tmin, triangle range
  float cand_t[kCand];
  int cand_tb[kCand];
  int cand_te[kCand];
  int hit_leaf[RAY_HITS];  // leaves with a hit: leaf, tmin, key, exact depth of their hit
  float hit_t[RAY_HITS];
  unsigned long long hit_key[RAY_HITS];
  float hit_lam[RAY_HITS];
  int order[RAY_HITS];
  int nc, ncand, nhit, status;
  SgRayHit res[2];
};

// Scan the listed candidate leaves' tris (a wave per leaf); leaves with a hit below
// the depth bound go to the hit list with the exact depth of that hit.
template <int NT>
__device__ __forceinline__ void scan_candidates(const SgDevView &d, const SgRayView &rv, const RayPre &r,
                                                RayShared<NT> &S, int ncand, long long &ntris) {
  const int lane = lane_id();
  const int wave = threadIdx.x / SG_WAVE;
  for (int ci = wave; ci < ncand; ci += NT / SG_WAVE) {
    const int l = S.cand_leaf[ci];
    const int tb = S.cand_tb[ci], te = S.cand_te[ci];
    unsigned long long best = ~0ull;
    float best_lam = 0.0f;
    for (int t0 = tb; t0 < te; t0 += 2 * SG_WAVE) {
      int4 tr[2];
      bool ok[2];
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const int t = t0 + j * SG_WAVE + lane;
        ok[j] = t < te;
        if (ok[j]) {
          tr[j] = rv.ltri[t];
          if (d.respect_hide && tr[j].w < 0) ok[j] = false;
        }
      }
      float c[2][9];
#pragma unroll
      for (int j = 0; j < 2; j++) {
        if (ok[j]) {
          c[j][0] = d.px[tr[j].x];
          c[j][1] = d.py[tr[j].x];
          c[j][2] = d.pz[tr[j].x];
          c[j][3] = d.px[tr[j].y];
          c[j][4] = d.py[tr[j].y];
          c[j][5] = d.pz[tr[j].y];
          c[j][6] = d.px[tr[j].z];
          c[j][7] = d.py[tr[j].z];
          c[j][8] = d.pz[tr[j].z];
        }
      }
#pragma unroll
      for (int j = 0; j < 2; j++) {
        if (ok[j]) {
          ntris++;
          float lam;
          if (ray_tri(r, c[j][0], c[j][1], c[j][2], c[j][3], c[j][4], c[j][5], c[j][6], c[j][7], c[j][8], lam)) {
            const unsigned long long k = hit_key(lam, t0 + j * SG_WAVE + lane);
            if (k < best) {
              best = k;
              best_lam = lam;
            }
          }
        }
      }
    }
    const unsigned long long wb = wave_min_u64(best);
    if (wb != ~0ull && key_depth(wb) < r.depth0 && best == wb) {  // the one lane holding the minimum
      const int h = atomicAdd(&S.nhit, 1);
      if (h < RAY_HITS) {
        S.hit_leaf[h] = l;
        S.hit_t[h] = S.cand_t[ci];
        S.hit_key[h] = wb;
        S.hit_lam[h] = best_lam;
      } else {
        S.status = SG_RAY_OVERFLOW;
      }
    }
  }
}

// One ray, the whole workgroup; lane 0 of wave 0 leaves the result in S.res[slot].
// Global load levels: subtree boxes and ranges; leaf boxes and triangle ranges;
// triangle ids; corner positions; then, unless only the hit point is wanted, the
// winner's ids and positions again for the face normal, corner and poly.
template <int NT>
__device__ __forceinline__ void ray_one(const SgDevView &d, const SgRayView &rv, const SgRay *ray, int original,
                                        int stale, RayShared<NT> &S, int slot, long long *stats, bool point_only) {
  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid / SG_WAVE;
  if (tid == 0) {
    S.nc = 0;
    S.ncand = 0;
    S.nhit = 0;
    S.status = 0;
  }
  __syncthreads();
  const RayPre r = ray_pre(ray);
  const float *cut_bbs = original ? d.cut_obb : d.cut_bb;
  const float *leaf_bbs = original ? d.leaf_obb : d.leaf_bb;
  const float *node_bbs = original ? d.node_obb : d.node_bb;
  long long ntests = 0, ntris = 0, ncands = 0;

  // 1. search-cut subtree boxes, with their leaf ranges
  for (int c = tid; c < d.C; c += NT) {
    const int cfl = d.cut_flags[c];
    const int touch = d.cut_touch[c];
    const BB6 cb = bb_load(cut_bbs + 8 * (size_t)c);
    const int lb = d.cut_lbeg[c], le = d.cut_lend[c];
    if (cfl & 1) continue;  // a single fully hidden leaf: never hit
    float t;
    // a box not refreshed since the previous dab moved its leaves counts as passing
    const bool pass = touch == stale || ray_box(r, cb, t);
    ntests++;
    if (pass) {
      const int k = atomicAdd(&S.nc, 1);
      S.cut_lb[k] = lb;
      S.cut_n[k] = le - lb;
    }
  }
  __syncthreads();
  const int nc = S.nc;
  // leaf-count prefix over the passing subtrees (wave 0: chunk sums, then a scan)
  if (wave == 0) {
    const int per = (nc + SG_WAVE - 1) / SG_WAVE;
    int sum = 0;
    for (int q = 0; q < per; q++) {
      const int k = lane * per + q;
      if (k < nc) sum += S.cut_n[k];
    }
    int incl = sum;
    for (int off = 1; off < SG_WAVE; off <<= 1) {
      const int o = __shfl_up(incl, off);
      if (lane >= off) incl += o;
    }
    int run = incl - sum;
    for (int q = 0; q < per; q++) {
      const int k = lane * per + q;
      if (k < nc) {
        S.pre[k] = run;
        run += S.cut_n[k];
      }
    }
    if (lane == SG_WAVE - 1) S.pre[nc] = incl;
  }
  __syncthreads();
  const int total = S.pre[nc];

  // 2. leaves in rounds; 3. the passing leaves' triangles, once at the end (or
  // whenever the next round could overflow the list)
  for (int g0 = 0; g0 < total; g0 += RayShared<NT>::kRound) {
    int leaf[RAY_LEAF_UNROLL], fl[RAY_LEAF_UNROLL], tb[RAY_LEAF_UNROLL], te[RAY_LEAF_UNROLL];
    BB6 b[RAY_LEAF_UNROLL];
#pragma unroll
    for (int j = 0; j < RAY_LEAF_UNROLL; j++) {
      const int g = g0 + j * NT + tid;
      leaf[j] = -1;
      if (g < total) {
        int lo = 0, hi = nc - 1;  // last k with pre[k] <= g
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (S.pre[mid] <= g)
            lo = mid;
          else
            hi = mid - 1;
        }
        leaf[j] = S.cut_lb[lo] + (g - S.pre[lo]);
        b[j] = bb_load(leaf_bbs + 8 * (size_t)leaf[j]);
        fl[j] = d.leaf_flags[leaf[j]];
        tb[j] = rv.toff[leaf[j]];
        te[j] = rv.toff[leaf[j] + 1];
      }
    }
#pragma unroll
    for (int j = 0; j < RAY_LEAF_UNROLL; j++) {
      if (leaf[j] < 0) continue;
      ntests++;
      float t;
      // fully hidden leaves never hit (BKE_pbvh_node_raycast returns false for them);
      // tmin must be below FLT_MAX to be visited at all (NaN and +inf never are)
      bool pass = !(fl[j] & 1) && ray_box(r, b[j], t) && t < FLT_MAX;
      if (pass && r.slow) {
        for (int nd = rv.node_parent[rv.leaf_node[leaf[j]]]; nd >= 0 && pass; nd = rv.node_parent[nd]) {
          float ta;
          pass = ray_box(r, bb_load(node_bbs + 8 * (size_t)nd), ta);
          ntests++;
        }
      }
      if (pass) {
        const int q = atomicAdd(&S.ncand, 1);
        S.cand_leaf[q] = leaf[j];
        S.cand_t[q] = t;
        S.cand_tb[q] = tb[j];
        S.cand_te[q] = te[j];
      }
    }
    __syncthreads();
    const int ncand = S.ncand;
    if (g0 + RayShared<NT>::kRound < total && ncand + RayShared<NT>::kRound > RayShared<NT>::kCand) {
      ncands += (tid == 0) ? ncand : 0;
      scan_candidates(d, rv, r, S, ncand, ntris);
      __syncthreads();
      if (tid == 0) S.ncand = 0;
      __syncthreads();
    }
  }
  {
    const int ncand = S.ncand;
    ncands += (tid == 0) ? ncand : 0;
    scan_candidates(d, rv, r, S, ncand, ntris);
    __syncthreads();
  }

  // 4. ordered fold over the leaves with a hit
  if (wave == 0) {
    const int nh = min(S.nhit, RAY_HITS);
    for (int i = lane; i < nh; i += SG_WAVE) {
      const float ti = S.hit_t[i];
      const int li = S.hit_leaf[i];
      int rank = 0;
      for (int j = 0; j < nh; j++) {
        const float tj = S.hit_t[j];
        rank += (tj < ti || (tj == ti && S.hit_leaf[j] < li)) ? 1 : 0;
      }
      S.order[rank] = i;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    if (lane == 0) {
      SgRayHit &h = S.res[slot];
      h.hit = 0;
      h.location[0] = h.location[1] = h.location[2] = 0.0f;
      h.face_normal[0] = h.face_normal[1] = h.face_normal[2] = 0.0f;
      h.reserved = 0;
      h.depth = r.depth0;
      h.active_vertex = h.active_face = h.looptri = h.node = -1;
      h.status = S.status;
      float bound = FLT_MAX, depth = r.depth0;
      int win = -1;
      for (int k = 0; k < nh; k++) {
        const int i = S.order[k];
        if (S.hit_t[i] < bound) {
          const float ld = key_depth(S.hit_key[i]);
          if (ld < depth) {
            depth = ld;
            bound = depth;
            win = i;
          }
        }
      }
      if (win >= 0 && h.status == 0) {
        const float lam = S.hit_lam[win];  // the exact depth (sign of zero included)
        h.hit = 1;
        h.depth = lam;
        for (int k = 0; k < 3; k++) h.location[k] = r.o[k] + r.dir[k] * lam;
        if (!point_only) {
          // the winning triangle's normal, nearest corner (pbvh.c:4073-4101) and ids
          const int pos = (int)(unsigned)S.hit_key[win];
          const int4 tr = rv.ltri[pos];
          const int cv[3] = {tr.x, tr.y, tr.z};
          float co[3][3];
          for (int j = 0; j < 3; j++) {
            co[j][0] = d.px[cv[j]];
            co[j][1] = d.py[cv[j]];
            co[j][2] = d.pz[cv[j]];
          }
          sg_tri_normal(co[0][0], co[0][1], co[0][2], co[1][0], co[1][1], co[1][2], co[2][0], co[2][1], co[2][2],
                        h.face_normal[0], h.face_normal[1], h.face_normal[2]);
          float best[3] = {0.0f, 0.0f, 0.0f};
          int bv = cv[0];
          for (int j = 0; j < 3; j++) {
            const float dj0 = co[j][0] - h.location[0], dj1 = co[j][1] - h.location[1], dj2 = co[j][2] - h.location[2];
            const float db0 = best[0] - h.location[0], db1 = best[1] - h.location[1], db2 = best[2] - h.location[2];
            if (j == 0 || (dj0 * dj0 + dj1 * dj1 + dj2 * dj2) < (db0 * db0 + db1 * db1 + db2 * db2)) {
              best[0] = co[j][0];
              best[1] = co[j][1];
              best[2] = co[j][2];
              bv = cv[j];
            }
          }
          h.active_vertex = d.vperm_old[bv];
          h.looptri = tr.w & 0x7fffffff;
          h.active_face = rv.tri_poly[h.looptri];
          h.node = rv.leaf_node[S.hit_leaf[win]];
        }
      }
    }
  }
  if (stats) {
    // per-ray work: boxes tested, leaves whose tris were scanned, tris tested
    __shared__ long long s_acc[3];
    if (tid < 3) s_acc[tid] = 0;
    __syncthreads();
    atomicAdd((unsigned long long *)&s_acc[0], (unsigned long long)ntests);
    atomicAdd((unsigned long long *)&s_acc[2], (unsigned long long)ntris);
    if (tid == 0) s_acc[1] = ncands;
    __syncthreads();
    if (tid < 3) atomicAdd((unsigned long long *)&stats[tid], (unsigned long long)s_acc[tid]);
  }
  __syncthreads();
}

// Rays in a grid-stride loop (hits written per ray), or one dab placement: ray 0 gives
// the dab centre and, for Grab, ray 1's hit minus ray 0's gives the translation.
// place[8]: cx cy cz gx gy gz hit 0.
template <int NT>
__attribute__((amdgpu_waves_per_eu(4, 8))) __global__ __launch_bounds__(NT) void k_raycast(SgDevView d, SgRayView rv, const SgRay *rays, int n, int original,
                                                int stale, SgRayHit *out, float *place, int grab_pair,
                                                long long *stats) {
  __shared__ RayShared<NT> S;
  if (place && grab_pair) {
    // A Grab placement's two rays run in two workgroups side by side.  Each leaves its
    // hit point, hit flag and status in d.place_pair with agent-scope atomics and counts
    // itself in; the second to arrive writes the placement and resets the count.  No
    // block waits for the other.
    if (blockIdx.x > 1) return;
    const int b = blockIdx.x;
    ray_one(d, rv, rays + b, original, stale, S, 0, stats, true);
    __shared__ int s_last;
    if (threadIdx.x == 0) {
      const SgRayHit &h = S.res[0];
      int *pp = d.place_pair + 1 + 5 * b;
      for (int k = 0; k < 3; k++)
        __hip_atomic_store(pp + k, __float_as_int(h.location[k]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(pp + 3, h.hit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(pp + 4, h.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_last = __hip_atomic_fetch_add(d.place_pair, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 1;
    }
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
      float loc[2][3];
      int hit[2], status[2];
      for (int q = 0; q < 2; q++) {
        const int *pp = d.place_pair + 1 + 5 * q;
        for (int k = 0; k < 3; k++)
          loc[q][k] = __int_as_float(__hip_atomic_load(pp + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        hit[q] = __hip_atomic_load(pp + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        status[q] = __hip_atomic_load(pp + 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      float g[3] = {0.0f, 0.0f, 0.0f};
      if (hit[0] && hit[1]) {
        for (int k = 0; k < 3; k++) g[k] = loc[1][k] - loc[0][k];
      }
      const bool ok = hit[0] && status[0] == 0;
      *(float4 *)place = make_float4(loc[0][0], loc[0][1], loc[0][2], g[0]);
      *(float4 *)(place + 4) = make_float4(g[1], g[2], ok ? 1.0f : 0.0f, 0.0f);
      __hip_atomic_store(d.place_pair, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // for the next dab
    }
    return;
  }
  if (place) {
    if (blockIdx.x != 0) return;
    ray_one(d, rv, rays, original, stale, S, 0, stats, true);
    if (threadIdx.x == 0) {
      const SgRayHit &a = S.res[0];
      const bool hit = a.hit && a.status == 0;
      *(float4 *)place = make_float4(a.location[0], a.location[1], a.location[2], 0.0f);
      *(float4 *)(place + 4) = make_float4(0.0f, 0.0f, hit ? 1.0f : 0.0f, 0.0f);
    }
    return;
  }
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    ray_one(d, rv, rays + i, original, stale, S, 0, stats, false);
    if (threadIdx.x == 0) out[i] = S.res[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// The reference's callback API: which leaves the occluded search visits, and the leaf
// tests its callbacks run.
// ---------------------------------------------------------------------------

// dist_squared_ray_to_aabb_v3_precalc / dist_squared_ray_to_aabb_v3 (math_geom.cc:643-744)
// and dist_squared_ray_to_seg_v3 with isect_ray_line_v3 (:574-606, :1997-2021).
struct RayDistPre {
  float o[3], dir[3], inv[3];
};
__device__ __forceinline__ RayDistPre ray_dist_pre(const float o[3], const float dir[3]) {
  RayDistPre p;
  for (int i = 0; i < 3; i++) {
    p.o[i] = o[i];
    p.dir[i] = dir[i];
    p.inv[i] = (dir[i] != 0.0f) ? (1.0f / dir[i]) : FLT_MAX;
  }
  return p;
}

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ __forceinline__ float ray_seg_dist_sq(const float o[3], const float dir[3], const float v0[3],
                                                 const float v1[3], float &r_depth) {
  float a[3], t[3], n[3], pt[3];
  for (int i = 0; i < 3; i++) {
    a[i] = v1[i] - v0[i];
    t[i] = v0[i] - o[i];
  }
  n[0] = a[1] * dir[2] - a[2] * dir[1];
  n[1] = a[2] * dir[0] - a[0] * dir[2];
  n[2] = a[0] * dir[1] - a[1] * dir[0];
  const float nlen = dot3(n, n);
  bool on_v0 = true;
  float lambda = 0.0f;
  if (nlen != 0.0f) {
    float c[3], cray[3];
    for (int i = 0; i < 3; i++) c[i] = n[i] - t[i];
    cray[0] = c[1] * dir[2] - c[2] * dir[1];
    cray[1] = c[2] * dir[0] - c[0] * dir[2];
    cray[2] = c[0] * dir[1] - c[1] * dir[0];
    lambda = dot3(cray, n) / nlen;
    on_v0 = lambda <= 0.0f;
  }
  if (on_v0) {
    for (int i = 0; i < 3; i++) pt[i] = v0[i];
  } else if (lambda >= 1.0f) {
    for (int i = 0; i < 3; i++) pt[i] = v1[i];
  } else {
    const float s = 1.0f - lambda;
    for (int i = 0; i < 3; i++) pt[i] = s * v0[i] + lambda * v1[i];
  }
  float dv[3];
  for (int i = 0; i < 3; i++) dv[i] = pt[i] - o[i];
  const float depth = dot3(dv, dir);
  r_depth = depth;
  return dot3(dv, dv) - depth * depth;
}

__device__ __forceinline__ float ray_box_dist_sq(const RayDistPre &d, const BB6 &b, float &r_depth) {
  float lo[3], hi[3];
  for (int i = 0; i < 3; i++) {
    const bool neg = d.dir[i] < 0.0f;
    lo[i] = neg ? b.mx[i] : b.mn[i];
    hi[i] = neg ? b.mn[i] : b.mx[i];
  }
  float tmin[3], tmax[3];
  for (int i = 0; i < 3; i++) {
    tmin[i] = (lo[i] - d.o[i]) * d.inv[i];
    tmax[i] = (hi[i] - d.o[i]) * d.inv[i];
  }
  float va[3], vb[3], rtmin, rtmax;
  int main_axis;
  // the fixed-index writes below keep va/vb in registers
  va[0] = va[1] = va[2] = vb[0] = vb[1] = vb[2] = 0.0f;
  if ((tmax[0] <= tmax[1]) && (tmax[0] <= tmax[2])) {
    rtmax = tmax[0];
    va[0] = vb[0] = hi[0];
    main_axis = 3;
  } else if ((tmax[1] <= tmax[0]) && (tmax[1] <= tmax[2])) {
    rtmax = tmax[1];
    va[1] = vb[1] = hi[1];
    main_axis = 2;
  } else {
    rtmax = tmax[2];
    va[2] = vb[2] = hi[2];
    main_axis = 1;
  }
  if ((tmin[0] >= tmin[1]) && (tmin[0] >= tmin[2])) {
    rtmin = tmin[0];
    va[0] = vb[0] = lo[0];
    main_axis -= 3;
  } else if ((tmin[1] >= tmin[0]) && (tmin[1] >= tmin[2])) {
    rtmin = tmin[1];
    va[1] = vb[1] = lo[1];
    main_axis -= 1;
  } else {
    rtmin = tmin[2];
    va[2] = vb[2] = lo[2];
    main_axis -= 2;
  }
  if (main_axis < 0) main_axis += 3;
  if (rtmin <= rtmax) {
    float dv[3];
    for (int i = 0; i < 3; i++) dv[i] = hi[i] - d.o[i];
    r_depth = dot3(dv, d.dir);
    return 0.0f;
  }
  const float dm = comp(d.dir[0], d.dir[1], d.dir[2], main_axis);
  const float lm = comp(lo[0], lo[1], lo[2], main_axis), hm = comp(hi[0], hi[1], hi[2], main_axis);
  const float am = dm >= 0.0f ? lm : hm, bm = dm >= 0.0f ? hm : lm;
  va[0] = main_axis == 0 ? am : va[0];
  va[1] = main_axis == 1 ? am : va[1];
  va[2] = main_axis == 2 ? am : va[2];
  vb[0] = main_axis == 0 ? bm : vb[0];
  vb[1] = main_axis == 1 ? bm : vb[1];
  vb[2] = main_axis == 2 ? bm : vb[2];
  return ray_seg_dist_sq(d.o, d.dir, va, vb, r_depth);
}

// dist_squared_ray_to_tri_v3_fast (pbvh.c:3968-3990): edges (2,0), (0,1), (1,2); the
// first always initialises the result.
__device__ __forceinline__ float ray_tri_edges_dist_sq(const float o[3], const float dir[3], const float c[9],
                                                       float &r_depth) {
  float best = FLT_MAX;
  for (int i = 0, j = 2; i < 3; j = i++) {
    float dt = FLT_MAX;
    const float dsq = ray_seg_dist_sq(o, dir, c + 3 * i, c + 3 * j, dt);
    if (dsq < best || i == 0) {
      r_depth = dt;
      best = dsq;
    }
  }
  return best;
}

// Leaves the occluded search passes for one ray (pbvh_iter_next_occluded, pbvh.c:789-816):
// mode 0 isect_ray_aabb_v3 (ray_aabb_intersect, :3895-3913), mode 1 depth > 0 of
// dist_squared_ray_to_aabb_v3 (nearest_to_ray_aabb_dist_sq, :4316-4338).  A leaf is listed
// with its tmin when it and every ancestor pass; the ancestors are walked whenever the
// leaf test alone does not imply them (mode 1 always, mode 0 for a non-finite 1/dir).
// One workgroup; the list is unordered.
__global__ __launch_bounds__(1024) void k_ray_leaves(SgDevView d, SgRayView rv, const SgRay *ray, int mode,
                                                     int original, int2 *out, int *count) {
  __shared__ int s_cut[1024];
  __shared__ int s_nc;
  const int tid = threadIdx.x;
  if (tid == 0) s_nc = 0;
  __syncthreads();
  const RayPre r = ray_pre(ray);
  const RayDistPre dp = ray_dist_pre(r.o, r.dir);
  const float *cut_bbs = original ? d.cut_obb : d.cut_bb;
  const float *leaf_bbs = original ? d.leaf_obb : d.leaf_bb;
  const float *node_bbs = original ? d.node_obb : d.node_bb;
  auto test = [&](const BB6 &b, float &t) -> bool {
    if (mode == 0) return ray_box(r, b, t);
    float depth;
    t = ray_box_dist_sq(dp, b, depth);
    return depth > 0.0f;
  };
  const bool walk = mode == 1 || r.slow;
  for (int c = tid; c < d.C; c += 1024) {
    float t;
    if (test(bb_load(cut_bbs + 8 * (size_t)c), t)) s_cut[atomicAdd(&s_nc, 1)] = c;
  }
  __syncthreads();
  const int nc = s_nc;
  for (int k = 0; k < nc; k++) {
    const int c = s_cut[k];
    const int lb = d.cut_lbeg[c], le = d.cut_lend[c];
    for (int l = lb + tid; l < le; l += 1024) {
      float t;
      bool pass = test(bb_load(leaf_bbs + 8 * (size_t)l), t);
      if (pass && walk) {
        for (int nd = rv.node_parent[rv.leaf_node[l]]; nd >= 0 && pass; nd = rv.node_parent[nd]) {
          float ta;
          pass = test(bb_load(node_bbs + 8 * (size_t)nd), ta);
        }
      }
      if (pass) out[atomicAdd(count, 1)] = make_int2(l, __float_as_int(t));
    }
  }
}

// One leaf's triangles against one ray, as the callbacks' node tests do.  mode 0
// (pbvh_faces_node_raycast, pbvh.c:4041-4107): the first nearest watertight hit below
// *depth updates depth, face normal, active vertex and face.  mode 1
// (pbvh_faces_node_nearest_to_ray, :4353-4392): the first smallest edge distance below
// *dist_sq updates dist_sq and depth.  With oco (node-local original coordinates,
// AoS) and ofv (the node's face_vert_indices) the corners come from there.
// args[14]: start xyz, dir xyz, kx ky kz (int bits), sx sy sz, depth, dist_sq.
// res[0..8]: depth, dist_sq, hit, normal xyz, active vertex, poly, looptri (bits).
__global__ __launch_bounds__(64) void k_node_ray(SgDevView d, SgRayView rv, const float *args, int mode, int leaf,
                                                 const float *oco, const int *ofv, float *res) {
  const int lane = lane_id();
  RayPre r;
  for (int i = 0; i < 3; i++) {
    r.o[i] = args[i];
    r.dir[i] = args[3 + i];
  }
  r.kx = __float_as_int(args[6]);
  r.ky = __float_as_int(args[7]);
  r.kz = __float_as_int(args[8]);
  r.sx = args[9];
  r.sy = args[10];
  r.sz = args[11];
  const float depth_in = args[12], dsq_in = args[13];
  const int tb = rv.toff[leaf], te = rv.toff[leaf + 1];
  unsigned long long best = ~0ull;
  float best_depth = 0.0f;
  for (int t = tb + lane; t < te; t += SG_WAVE) {
    const int4 tr = rv.ltri[t];
    if (d.respect_hide && tr.w < 0) continue;
    float c[9];
    if (oco) {
      const int q = t - tb;
      for (int j = 0; j < 3; j++) {
        const int lv = ofv[3 * q + j];
        c[3 * j] = oco[3 * lv];
        c[3 * j + 1] = oco[3 * lv + 1];
        c[3 * j + 2] = oco[3 * lv + 2];
      }
    } else {
      const int cv[3] = {tr.x, tr.y, tr.z};
      for (int j = 0; j < 3; j++) {
        c[3 * j] = d.px[cv[j]];
        c[3 * j + 1] = d.py[cv[j]];
        c[3 * j + 2] = d.pz[cv[j]];
      }
    }
    if (mode == 0) {
      float lam;
      if (ray_tri(r, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], lam)) {
        const unsigned long long k = hit_key(lam, t);
        best = k < best ? k : best;
      }
    } else {
      float dep;
      const float dsq = ray_tri_edges_dist_sq(r.o, r.dir, c, dep);
      // squared distances can be negative by rounding: order them as signed floats
      unsigned u = __float_as_uint(dsq);
      u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
      const unsigned long long k = ((unsigned long long)u << 32) | (unsigned)t;
      if (k < best) {
        best = k;
        best_depth = dep;
      }
    }
  }
  const unsigned long long wb = wave_min_u64(best);
  const unsigned long long mine = __ballot(best == wb && wb != ~0ull);
  if (wb == ~0ull) {
    if (lane == 0) res[2] = 0.0f;
    return;
  }
  const int owner = __ffsll((long long)mine) - 1;
  const float dep_owner = __shfl(best_depth, owner);
  if (lane != 0) return;
  const int t = (int)(unsigned)wb;
  const int4 tr = rv.ltri[t];
  float c[9];
  int cv[3] = {tr.x, tr.y, tr.z};
  if (oco) {
    const int q = t - tb;
    for (int j = 0; j < 3; j++) {
      const int lv = ofv[3 * q + j];
      c[3 * j] = oco[3 * lv];
      c[3 * j + 1] = oco[3 * lv + 1];
      c[3 * j + 2] = oco[3 * lv + 2];
    }
  } else {
    for (int j = 0; j < 3; j++) {
      c[3 * j] = d.px[cv[j]];
      c[3 * j + 1] = d.py[cv[j]];
      c[3 * j + 2] = d.pz[cv[j]];
    }
  }
  if (mode == 1) {
    unsigned u = (unsigned)(wb >> 32);
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    const float dsq = __uint_as_float(u);
    if (!(dsq < dsq_in)) {
      res[2] = 0.0f;
      return;
    }
    res[0] = dep_owner;
    res[1] = dsq;
    res[2] = 1.0f;
    return;
  }
  float lam = 0.0f;
  ray_tri(r, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], lam);
  if (!(lam < depth_in)) {
    res[2] = 0.0f;
    return;
  }
  res[0] = lam;
  res[2] = 1.0f;
  sg_tri_normal(c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], res[3], res[4], res[5]);
  float loc[3];
  for (int k = 0; k < 3; k++) loc[k] = r.o[k] + r.dir[k] * lam;
  float bx = 0.0f, by = 0.0f, bz = 0.0f;
  int bv = cv[0];
  for (int j = 0; j < 3; j++) {
    const float dj0 = c[3 * j] - loc[0], dj1 = c[3 * j + 1] - loc[1], dj2 = c[3 * j + 2] - loc[2];
    const float db0 = bx - loc[0], db1 = by - loc[1], db2 = bz - loc[2];
    if (j == 0 || (dj0 * dj0 + dj1 * dj1 + dj2 * dj2) < (db0 * db0 + db1 * db1 + db2 * db2)) {
      bx = c[3 * j];
      by = c[3 * j + 1];
      bz = c[3 * j + 2];
      bv = cv[j];
    }
  }
  const int looptri = tr.w & 0x7fffffff;
  res[6] = __int_as_float(d.vperm_old[bv]);
  res[7] = __int_as_float(rv.tri_poly[looptri]);
  res[8] = __int_as_float(looptri);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static SgRayView ray_view(PBVH *p) {
  SgRayView v;
  v.ltri = p->d.ltri;
  v.toff = p->d.leaf_toff;
  v.tri_poly = p->d.tri_poly;
  v.node_parent = p->d.node_parent;
  v.leaf_node = p->d.leaf_node;
  return v;
}

template <class T>
static int upload_now(T **dst, const T *src, size_t count) {
  if (sg_check_hip(hipMalloc((void **)dst, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc")) return -1;
  return sg_check_hip(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice), "copy");
}

// Leaf-ordered triangle table and tree links, built on the first ray cast.
int sg_ray_ensure(PBVH *p) {
  if (p->no_device || !p->build_ok) return sg_set_error("no device mesh");
  SgDev &d = p->d;
  if (d.ltri) return 0;
  const int T = p->totprim, L = (int)p->leaf_node.size(), N = (int)p->nodes.size();
  std::vector<int4> lt((size_t)std::max(T, 1));
  std::vector<int> toff(L + 1, 0), poly((size_t)std::max(T, 1)), parent(std::max(N, 1), -1);
  int run = 0;
  for (int l = 0; l < L; l++) {
    const PBVHNode &n = p->nodes[p->leaf_node[l]];
    if (n.prim_off != run) return sg_set_error("leaf triangle ranges are not contiguous");
    toff[l] = run;
    run += n.totprim;
  }
  toff[L] = run;
  if (run != T) return sg_set_error("leaf triangle ranges do not cover the mesh");
  const bool hide = !p->vhide.empty();
  for (int pos = 0; pos < T; pos++) {
    const int t = p->prim_indices[pos];
    const int *c = &p->tris[3 * (size_t)t];
    int w = t;
    if (hide && (p->vhide[c[0]] || p->vhide[c[1]] || p->vhide[c[2]])) w |= (int)0x80000000u;
    lt[pos] = make_int4(p->perm_new_of_old[c[0]], p->perm_new_of_old[c[1]], p->perm_new_of_old[c[2]], w);
  }
  for (int t = 0; t < T; t++) poly[t] = p->looptri ? (int)p->looptri[t].poly : t;
  for (int i = 0; i < N; i++) {
    const PBVHNode &n = p->nodes[i];
    if (!(n.flag & PBVH_Leaf)) {
      parent[n.children_offset] = i;
      parent[n.children_offset + 1] = i;
    }
  }
  int rc = upload_now(&d.ltri, lt.data(), (size_t)T);
  if (!rc) rc = upload_now(&d.leaf_toff, toff.data(), (size_t)L + 1);
  if (!rc) rc = upload_now(&d.tri_poly, poly.data(), (size_t)T);
  if (!rc) rc = upload_now(&d.node_parent, parent.data(), (size_t)N);
  if (!rc) rc = upload_now(&d.leaf_node, p->leaf_node.data(), (size_t)L);
  if (rc) {
    void *q[] = {d.ltri, d.leaf_toff, d.tri_poly, d.node_parent, d.leaf_node};
    for (void *x : q)
      if (x) (void)hipFree(x);
    d.ltri = nullptr;
    d.leaf_toff = d.tri_poly = d.node_parent = d.leaf_node = nullptr;
  }
  return rc;
}

static bool ray_needs_ancestors(const SgRay &r) {
  for (int i = 0; i < 3; i++) {
    const float inv = 1.0f / r.normal[i];
    if (!std::isfinite(inv)) return true;
  }
  return false;
}

int sg_enqueue_raycast(PBVH *p, const SgRay *rays_dev, int n, int original, int stale, SgRayHit *out_dev,
                       float *place_dev, int grab_pair, long long *stats_dev) {
  if (int rc = sg_ray_ensure(p)) return rc;
  const SgDevView v = sg_view(p);
  const SgRayView rv = ray_view(p);
  if (place_dev)  // one workgroup per ray of the placement
    hipLaunchKernelGGL(k_raycast<1024>, dim3(grab_pair ? 2 : 1), dim3(1024), 0, p->stream, v, rv, rays_dev, n,
                       original, stale, out_dev, place_dev, grab_pair, stats_dev);
  else
    hipLaunchKernelGGL(k_raycast<256>, dim3(std::max(1, std::min(n, 65536))), dim3(256), 0, p->stream, v, rv,
                       rays_dev, n, original, stale, out_dev, place_dev, grab_pair, stats_dev);
  return sg_check_hip(hipGetLastError(), "raycast launch");
}

bool sg_ray_host_needs_refit(const SgRay *rays, int n) {
  for (int i = 0; i < n; i++)
    if (ray_needs_ancestors(rays[i])) return true;
  return false;
}

static int ensure_ray_buffers(PBVH *p, int n) {
  SgDev &d = p->d;
  if (n <= d.ray_cap) return 0;
  if (d.ray_in) (void)hipFree(d.ray_in);
  if (d.ray_out) (void)hipFree(d.ray_out);
  d.ray_in = nullptr;
  d.ray_out = nullptr;
  d.ray_cap = 0;
  if (sg_check_hip(hipMalloc(&d.ray_in, (size_t)n * sizeof(SgRay)), "hipMalloc")) return -1;
  if (sg_check_hip(hipMalloc(&d.ray_out, (size_t)n * sizeof(SgRayHit)), "hipMalloc")) return -1;
  d.ray_cap = n;
  return 0;
}

int sg_raycast_device(PBVH *p, const SgRay *rays_dev, int n, int original, SgRayHit *hits_dev) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: ray casts on PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p || n < 0 || (n > 0 && (!rays_dev || !hits_dev))) return sg_set_error("bad arguments");
  if (((uintptr_t)rays_dev | (uintptr_t)hits_dev) & 15) return sg_set_error("ray arrays must be 16-byte aligned");
  if (p->no_device) return sg_set_error("no HIP device");
  if (n == 0) return 0;
  // boxes of internal nodes are needed current only for rays along a coordinate plane;
  // they are refit here whenever a stroke left them pending
  if (p->nodes_dirty_deferred) {
    if (int rc = sg_enqueue_flush(p)) return rc;
    p->nodes_dirty_deferred = false;
  }
  return sg_enqueue_raycast(p, rays_dev, n, original ? 1 : 0, -1, hits_dev, nullptr, 0, nullptr);
}

int sg_raycast(PBVH *p, const SgRay *rays, int n, int original, SgRayHit *hits) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: ray casts on PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p || n < 0 || (n > 0 && (!rays || !hits))) return sg_set_error("bad arguments");
  if (p->no_device) return sg_set_error("no HIP device");
  if (n == 0) return 0;
  if (int rc = ensure_ray_buffers(p, n)) return rc;
  int rc = sg_check_hip(hipMemcpyAsync(p->d.ray_in, rays, (size_t)n * sizeof(SgRay), hipMemcpyHostToDevice, p->stream),
                        "copy");
  if (!rc) rc = sg_raycast_device(p, p->d.ray_in, n, original, p->d.ray_out);
  if (!rc)
    rc = sg_check_hip(
        hipMemcpyAsync(hits, p->d.ray_out, (size_t)n * sizeof(SgRayHit), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "raycast");
  return rc;
}

int sg_raycast_stats(PBVH *p, const SgRay *rays, int n, int original, long long out[3]) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: ray casts on PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p || n <= 0 || !rays || !out) return sg_set_error("bad arguments");
  if (p->no_device) return sg_set_error("no HIP device");
  if (int rc = ensure_ray_buffers(p, n)) return rc;
  long long *st = nullptr;
  int rc = sg_check_hip(hipMalloc(&st, 3 * sizeof(long long)), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMemsetAsync(st, 0, 3 * sizeof(long long), p->stream), "memset");
  if (!rc)
    rc = sg_check_hip(hipMemcpyAsync(p->d.ray_in, rays, (size_t)n * sizeof(SgRay), hipMemcpyHostToDevice, p->stream),
                      "copy");
  if (!rc && p->nodes_dirty_deferred) {
    rc = sg_enqueue_flush(p);
    p->nodes_dirty_deferred = false;
  }
  if (!rc) rc = sg_enqueue_raycast(p, p->d.ray_in, n, original ? 1 : 0, -1, p->d.ray_out, nullptr, 0, st);
  if (!rc) rc = sg_check_hip(hipMemcpyAsync(out, st, 3 * sizeof(long long), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "raycast");
  if (st) (void)hipFree(st);
  return rc;
}

// ---------------------------------------------------------------------------
// sg_raycast_stats without allocation.  Its scratch -- the three work counters, the
// staged rays and the hits of at most SG_NOALLOC_RAYS rays -- lives in one file-scope
// device array; each GPU the library runs on holds its own copy until the process
// ends.  Host code reaches it only through the address hipGetSymbolAddress returns for
// the current device (the array's name is not a device address on the host).
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocStats = 256;  // three counters, padded to the rays' alignment
constexpr size_t kNoAllocBytes =
    kNoAllocStats + (size_t)SG_NOALLOC_RAYS * sizeof(SgRay) + (size_t)SG_NOALLOC_RAYS * sizeof(SgRayHit);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch[kMaxDevices];

// The scratch array's address on the current device, looked up once per device.
static unsigned char *scratch_address() {
  int dev = 0;
  if (sg_check_hip(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
  if (dev < 0 || dev >= kMaxDevices) {
    sg_set_error("device %d beyond the %d the scratch table holds", dev, kMaxDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    if (sg_check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")) return nullptr;
    g_scratch[dev] = (unsigned char *)ptr;
  }
  return g_scratch[dev];
}

__global__ void k_zero_ll(long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// The reference's callback-driven ray API (host side of the drop-ins)
// ---------------------------------------------------------------------------

static inline float h_bits_f(int i) {
  float f;
  memcpy(&f, &i, 4);
  return f;
}
static inline int h_f_bits(float f) {
  int i;
  memcpy(&i, &f, 4);
  return i;
}

static int ensure_ray_scratch(PBVH *p) {
  SgDev &d = p->d;
  if (d.rl_list) return 0;
  if (sg_check_hip(hipMalloc(&d.rl_list, (size_t)std::max(d.L, 1) * sizeof(int2)), "hipMalloc")) return -1;
  if (sg_check_hip(hipMalloc(&d.rl_args, 64 * sizeof(float)), "hipMalloc")) return -1;
  return 0;
}

// The occluded search's visiting order over the listed leaves: they enter
// node_tree_insert (pbvh.c:878-896) in left->right order; a leaf goes left of a node
// iff its tmin is smaller, so ties and NaN go right.  The in-order result is the
// stable sort by tmin for the non-NaN leaves, and a NaN leaf sits right after every
// leaf that entered before it and every later leaf whose tmin is below the largest
// non-NaN tmin present when it entered: sorting by (tmin, or that largest tmin for a
// NaN, then entry order) gives exactly the tree's order without building it.
static void occluded_order(std::vector<int2> &v) {
  std::sort(v.begin(), v.end(), [](const int2 &a, const int2 &b) { return a.x < b.x; });
  std::vector<float> key(v.size());
  float run = -INFINITY;
  for (size_t i = 0; i < v.size(); i++) {
    const float t = h_bits_f(v[i].y);
    if (t != t) {
      key[i] = run;
    } else {
      key[i] = t;
      run = std::max(run, t);
    }
  }
  std::vector<int> idx(v.size());
  for (size_t i = 0; i < idx.size(); i++) idx[i] = (int)i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
  std::vector<int2> out(v.size());
  for (size_t i = 0; i < idx.size(); i++) out[i] = v[idx[i]];
  v.swap(out);
}

static int ray_leaves(PBVH *p, const float start[3], const float dir[3], int mode, bool original,
                      std::vector<int2> &out) {
  out.clear();
  if (int rc = sg_ray_ensure(p)) return rc;
  if (int rc = ensure_ray_buffers(p, 1)) return rc;
  if (int rc = ensure_ray_scratch(p)) return rc;
  if (p->nodes_dirty_deferred) {
    if (int rc = sg_enqueue_flush(p)) return rc;
    p->nodes_dirty_deferred = false;
  }
  SgRay r;
  memcpy(r.start, start, sizeof(r.start));
  memcpy(r.normal, dir, sizeof(r.normal));
  r.depth = FLT_MAX;
  r.reserved = 0;
  int *cnt = (int *)(p->d.rl_args + 32);
  int rc = sg_check_hip(hipMemcpyAsync(p->d.ray_in, &r, sizeof(r), hipMemcpyHostToDevice, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipMemsetAsync(cnt, 0, sizeof(int), p->stream), "memset");
  if (!rc) {
    hipLaunchKernelGGL(k_ray_leaves, dim3(1), dim3(1024), 0, p->stream, sg_view(p), ray_view(p), p->d.ray_in, mode,
                       original ? 1 : 0, p->d.rl_list, cnt);
    rc = sg_check_hip(hipGetLastError(), "ray leaves launch");
  }
  int n = 0;
  if (!rc) rc = sg_check_hip(hipMemcpyAsync(&n, cnt, sizeof(int), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "ray leaves");
  if (rc) return rc;
  out.resize(n);
  if (n) rc = sg_check_hip(hipMemcpy(out.data(), p->d.rl_list, (size_t)n * sizeof(int2), hipMemcpyDeviceToHost), "copy");
  if (!rc) occluded_order(out);
  return rc;
}

static void occluded_visit(PBVH *p, const float start[3], const float dir[3], int mode, bool original,
                           BKE_pbvh_HitOccludedCallback cb, void *data) {
  if (!p || !cb || p->nodes.empty() || p->no_device || !p->build_ok) return;
  std::vector<int2> v;
  if (ray_leaves(p, start, dir, mode, original, v)) return;
  for (const int2 &e : v) p->nodes[p->leaf_node[e.x]].tmin = h_bits_f(e.y);
  float tmin = FLT_MAX;  // traverse_tree's bound (pbvh.c:966-968)
  for (const int2 &e : v) cb(&p->nodes[p->leaf_node[e.x]], data, &tmin);
}

// One leaf test on the GPU (k_node_ray); res as documented there.
static bool node_ray(PBVH *p, PBVHNode *node, float (*origco)[3], const float start[3], const float dir[3],
                     const IsectRayPrecalc *pre, int mode, float depth, float dist_sq, float res[9]) {
  if (!p || !node || node->leaf < 0 || p->no_device || !p->build_ok) return false;
  if (node->flag & PBVH_FullyHidden) return false;
  if (sg_ray_ensure(p) || ensure_ray_scratch(p)) return false;
  float args[14];
  memcpy(args, start, 3 * sizeof(float));
  memcpy(args + 3, dir, 3 * sizeof(float));
  IsectRayPrecalc own;
  if (!pre) {
    sg_isect_ray_tri_precalc(&own, dir);
    pre = &own;
  }
  memcpy(args + 6, &pre->kx, 4);
  memcpy(args + 7, &pre->ky, 4);
  memcpy(args + 8, &pre->kz, 4);
  args[9] = pre->sx;
  args[10] = pre->sy;
  args[11] = pre->sz;
  args[12] = depth;
  args[13] = dist_sq;
  float *oco = nullptr;
  int *ofv = nullptr;
  int rc = sg_check_hip(hipMemcpyAsync(p->d.rl_args, args, sizeof(args), hipMemcpyHostToDevice, p->stream), "copy");
  if (!rc && origco) {
    const int nv = node->uniq_verts + node->face_verts;
    rc = sg_check_hip(hipMalloc(&oco, (size_t)std::max(nv, 1) * 3 * sizeof(float)), "hipMalloc");
    if (!rc) rc = sg_check_hip(hipMalloc(&ofv, (size_t)std::max(node->totprim, 1) * 3 * sizeof(int)), "hipMalloc");
    if (!rc)
      rc = sg_check_hip(hipMemcpyAsync(oco, origco, (size_t)nv * 3 * sizeof(float), hipMemcpyHostToDevice, p->stream),
                        "copy");
    if (!rc)
      rc = sg_check_hip(hipMemcpyAsync(ofv, node->face_vert_indices, (size_t)node->totprim * 3 * sizeof(int),
                                       hipMemcpyHostToDevice, p->stream),
                        "copy");
  }
  float *dres = p->d.rl_args + 16;
  if (!rc) {
    hipLaunchKernelGGL(k_node_ray, dim3(1), dim3(64), 0, p->stream, sg_view(p), ray_view(p), p->d.rl_args, mode,
                       node->leaf, oco, ofv, dres);
    rc = sg_check_hip(hipGetLastError(), "node ray launch");
  }
  if (!rc) rc = sg_check_hip(hipMemcpyAsync(res, dres, 9 * sizeof(float), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "node ray");
  if (oco) (void)hipFree(oco);
  if (ofv) (void)hipFree(ofv);
  return !rc && res[2] != 0.0f;
}

extern "C" {

int sg_raycast_stats_noalloc(PBVH *p, const SgRay *rays, int n, int original, long long out[3]) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: ray casts on PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p || n <= 0 || !rays || !out) return sg_set_error("bad arguments");
  if (p->no_device || !p->build_ok) return sg_set_error("no device mesh");
  if (n > SG_NOALLOC_RAYS)
    return sg_set_error("sg_raycast_stats_noalloc: %d rays exceed the static scratch of %d rays", n, SG_NOALLOC_RAYS);
  if (!p->d.ltri)
    return sg_set_error("sg_raycast_stats_noalloc: the ray layout is not built yet (one sg_raycast builds it)");
  unsigned char *base = scratch_address();
  if (!base) return -1;
  long long *st = (long long *)base;
  SgRay *rin = (SgRay *)(base + kNoAllocStats);
  SgRayHit *rout = (SgRayHit *)(base + kNoAllocStats + (size_t)SG_NOALLOC_RAYS * sizeof(SgRay));
  hipLaunchKernelGGL(k_zero_ll, dim3(1), dim3(64), 0, p->stream, st, 3);  // a kernel, not a memset: capturable
  int rc = sg_check_hip(hipGetLastError(), "zero launch");
  if (!rc)
    rc = sg_check_hip(hipMemcpyAsync(rin, rays, (size_t)n * sizeof(SgRay), hipMemcpyHostToDevice, p->stream), "copy");
  if (!rc && p->nodes_dirty_deferred) {
    rc = sg_enqueue_flush(p);
    p->nodes_dirty_deferred = false;
  }
  if (!rc) rc = sg_enqueue_raycast(p, rin, n, original ? 1 : 0, -1, rout, nullptr, 0, st);
  if (!rc) rc = sg_check_hip(hipMemcpyAsync(out, st, 3 * sizeof(long long), hipMemcpyDeviceToHost, p->stream), "copy");
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "raycast");
  return rc;
}

void sg_isect_ray_tri_precalc(IsectRayPrecalc *out, const float dir[3]) {
  if (!out || !dir) return;
  const float x = fabsf(dir[0]), y = fabsf(dir[1]), z = fabsf(dir[2]);
  const int kz = (x > y) ? ((x > z) ? 0 : 2) : ((y > z) ? 1 : 2);
  int kx = (kz != 2) ? (kz + 1) : 0;
  int ky = (kx != 2) ? (kx + 1) : 0;
  if (dir[kz] < 0.0f) std::swap(kx, ky);
  const float inv_z = 1.0f / dir[kz];
  out->sx = dir[kx] * inv_z;
  out->sy = dir[ky] * inv_z;
  out->sz = inv_z;
  out->kx = kx;
  out->ky = ky;
  out->kz = kz;
}

void BKE_pbvh_raycast(PBVH *p, BKE_pbvh_HitOccludedCallback cb, void *data, const float ray_start[3],
                      const float ray_normal[3], bool original) {
  if (p && p->type != PBVH_FACES) { sg_set_error("faces only: ray casts on PBVH_GRIDS / PBVH_BMESH are not built"); return; }
  occluded_visit(p, ray_start, ray_normal, 0, original, cb, data);
}

void BKE_pbvh_find_nearest_to_ray(PBVH *p, BKE_pbvh_SearchNearestCallback cb, void *data, const float ray_start[3],
                                  const float ray_normal[3], bool original) {
  if (p && p->type != PBVH_FACES) { sg_set_error("faces only: ray casts on PBVH_GRIDS / PBVH_BMESH are not built"); return; }
  occluded_visit(p, ray_start, ray_normal, 1, original, cb, data);
}

float BKE_pbvh_node_get_tmin(PBVHNode *node) { return node ? node->tmin : 0.0f; }

bool BKE_pbvh_node_raycast(PBVH *p, PBVHNode *node, float (*origco)[3], bool use_origco, const float ray_start[3],
                           const float ray_normal[3], IsectRayPrecalc *isect_precalc, float *depth,
                           int *active_vertex_index, int *active_face_grid_index, float *face_normal) {
  if (p && p->type != PBVH_FACES) { sg_set_error("faces only: ray casts on PBVH_GRIDS / PBVH_BMESH are not built"); return false; }
  (void)use_origco;  // dyntopo only
  if (!depth) return false;
  float res[9];
  if (!node_ray(p, node, origco, ray_start, ray_normal, isect_precalc, 0, *depth, FLT_MAX, res)) return false;
  *depth = res[0];
  if (face_normal) memcpy(face_normal, res + 3, 3 * sizeof(float));
  if (active_vertex_index) {
    *active_vertex_index = h_f_bits(res[6]);
    if (active_face_grid_index) *active_face_grid_index = h_f_bits(res[7]);
  }
  return true;
}

bool BKE_pbvh_node_find_nearest_to_ray(PBVH *p, PBVHNode *node, float (*origco)[3], bool use_origco,
                                       const float ray_start[3], const float ray_normal[3], float *depth,
                                       float *dist_sq) {
  if (p && p->type != PBVH_FACES) { sg_set_error("faces only: ray casts on PBVH_GRIDS / PBVH_BMESH are not built"); return false; }
  (void)use_origco;
  if (!depth || !dist_sq) return false;
  float res[9];
  if (!node_ray(p, node, origco, ray_start, ray_normal, nullptr, 1, *depth, *dist_sq, res)) return false;
  *depth = res[0];
  *dist_sq = res[1];
  return true;
}

// BKE_pbvh_raycast_project_ray_root (pbvh.c:4262-4307): clip the segment to the root box
// grown by 1e-3 (absolute) and 0.1% (relative); a few FP32 operations on one box.
void BKE_pbvh_raycast_project_ray_root(PBVH *p, bool original, float ray_start[3], float ray_end[3],
                                       float ray_normal[3]) {
  if (p && p->type != PBVH_FACES) { sg_set_error("faces only: ray casts on PBVH_GRIDS / PBVH_BMESH are not built"); return; }
  if (!p || p->nodes.empty()) return;
  sg_sync_host_bbs(p);
  const float *vb = original ? p->nodes[0].orig_vb : p->nodes[0].vb;
  const float offset = 1.0f + 1e-3f;
  float bmin[3], bmax[3];
  for (int i = 0; i < 3; i++) {
    const float center = 0.5f * (vb[3 + i] + vb[i]);
    float diff = vb[3 + i] - center;
    diff = diff + 1e-3f;
    bmax[i] = center + diff * offset;
    bmin[i] = center + diff * -offset;
  }
  auto slab = [&](const float o[3], const float dir[3], float *tmin_out) -> bool {
    float inv[3];
    int sg[3];
    for (int i = 0; i < 3; i++) {
      inv[i] = 1.0f / dir[i];
      sg[i] = inv[i] < 0.0f;
    }
    const float *bb[2] = {bmin, bmax};
    float tmin = (bb[sg[0]][0] - o[0]) * inv[0];
    float tmax = (bb[1 - sg[0]][0] - o[0]) * inv[0];
    const float tymin = (bb[sg[1]][1] - o[1]) * inv[1];
    const float tymax = (bb[1 - sg[1]][1] - o[1]) * inv[1];
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    const float tzmin = (bb[sg[2]][2] - o[2]) * inv[2];
    const float tzmax = (bb[1 - sg[2]][2] - o[2]) * inv[2];
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    *tmin_out = tmin;
    return true;
  };
  float t_start, t_end, inv_dir[3];
  if (!slab(ray_start, ray_normal, &t_start)) return;
  for (int i = 0; i < 3; i++) inv_dir[i] = ray_normal[i] * -1.0f;
  if (!slab(ray_end, inv_dir, &t_end)) return;
  for (int i = 0; i < 3; i++) {
    ray_start[i] = ray_start[i] + ray_normal[i] * t_start;
    ray_end[i] = ray_end[i] + inv_dir[i] * t_end;
  }
}

}  // extern "C"
