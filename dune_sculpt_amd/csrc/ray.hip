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
#include <vector>

#include "sg_internal.h"
#include "sg_kernels.h"
#include "sg_device.h"

// Workgroup sizes: a batch of rays runs 256 threads per ray (several rays in flight per
// CU hide each ray's dependent load levels); a single placement ray runs 1024.
#define RAY_LEAF_UNROLL 2
#define RAY_HITS 1024  // leaves with a hit, per ray (both sizes: the same limit)

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
  static constexpr int kCand = NT * RAY_LEAF_UNROLL;  // one round of leaves
  int cut[1024];
  int pre[1025];
  int cand_leaf[kCand];
  float cand_t[kCand];
  int hit_leaf[RAY_HITS];
  float hit_t[RAY_HITS];
  unsigned long long hit_key[RAY_HITS];
  int order[RAY_HITS];
  int nc, ncand, nhit, status;
  SgRayHit res[2];
};

// Scan the listed candidate leaves' tris (a wave per leaf); leaves with a hit below
// the depth bound go to the hit list.
template <int NT>
__device__ __forceinline__ void scan_candidates(const SgDevView &d, const SgRayView &rv, const RayPre &r,
                                                RayShared<NT> &S, int ncand, long long &ntris) {
  const int lane = lane_id();
  const int wave = threadIdx.x / SG_WAVE;
  for (int ci = wave; ci < ncand; ci += NT / SG_WAVE) {
    const int l = S.cand_leaf[ci];
    const int tb = rv.toff[l], te = rv.toff[l + 1];
    unsigned long long best = ~0ull;
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
            best = k < best ? k : best;
          }
        }
      }
    }
    best = wave_min_u64(best);
    if (lane == 0 && best != ~0ull && key_depth(best) < r.depth0) {
      const int h = atomicAdd(&S.nhit, 1);
      if (h < RAY_HITS) {
        S.hit_leaf[h] = l;
        S.hit_t[h] = S.cand_t[ci];
        S.hit_key[h] = best;
      } else {
        S.status = SG_RAY_OVERFLOW;
      }
    }
  }
}

// One ray, the whole workgroup; lane 0 of wave 0 leaves the result in S.res[slot].
template <int NT>
__device__ __forceinline__ void ray_one(const SgDevView &d, const SgRayView &rv, const SgRay *ray, int original,
                                        int stale, RayShared<NT> &S, int slot, long long *stats) {
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

  // 1. search-cut subtree boxes
  for (int c = tid; c < d.C; c += NT) {
    if (d.cut_flags[c] & 1) continue;  // a single fully hidden leaf: never hit
    bool pass = d.cut_touch[c] == stale;
    if (!pass) {
      float t;
      pass = ray_box(r, bb_load(cut_bbs + 8 * (size_t)c), t);
    }
    ntests++;
    if (pass) S.cut[atomicAdd(&S.nc, 1)] = c;
  }
  __syncthreads();
  const int nc = S.nc;
  // leaf-count prefix over the passing subtrees (wave 0: chunk sums, then a scan)
  if (wave == 0) {
    const int per = (nc + SG_WAVE - 1) / SG_WAVE;
    int sum = 0;
    for (int q = 0; q < per; q++) {
      const int k = lane * per + q;
      if (k < nc) sum += d.cut_lend[S.cut[k]] - d.cut_lbeg[S.cut[k]];
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
        run += d.cut_lend[S.cut[k]] - d.cut_lbeg[S.cut[k]];
      }
    }
    if (lane == SG_WAVE - 1) S.pre[nc] = incl;
  }
  __syncthreads();
  const int total = S.pre[nc];

  // 2. + 3. leaves in rounds, each round's passing leaves scanned at its end
  for (int g0 = 0; g0 < total; g0 += NT * RAY_LEAF_UNROLL) {
    int leaf[RAY_LEAF_UNROLL];
    BB6 b[RAY_LEAF_UNROLL];
    int fl[RAY_LEAF_UNROLL];
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
        leaf[j] = d.cut_lbeg[S.cut[lo]] + (g - S.pre[lo]);
        b[j] = bb_load(leaf_bbs + 8 * (size_t)leaf[j]);
        fl[j] = d.leaf_flags[leaf[j]];
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
      }
    }
    __syncthreads();
    const int ncand = S.ncand;
    ncands += (tid == 0) ? ncand : 0;
    scan_candidates(d, rv, r, S, ncand, ntris);
    __syncthreads();
    if (tid == 0) S.ncand = 0;
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
        // the winning triangle again: exact depth (sign of zero included), normal,
        // nearest corner (pbvh.c:4076-4101)
        const int pos = (int)(unsigned)S.hit_key[win];
        const int4 tr = rv.ltri[pos];
        const int cv[3] = {tr.x, tr.y, tr.z};
        float co[3][3];
        for (int j = 0; j < 3; j++) {
          co[j][0] = d.px[cv[j]];
          co[j][1] = d.py[cv[j]];
          co[j][2] = d.pz[cv[j]];
        }
        float lam = 0.0f;
        ray_tri(r, co[0][0], co[0][1], co[0][2], co[1][0], co[1][1], co[1][2], co[2][0], co[2][1], co[2][2], lam);
        h.hit = 1;
        h.depth = lam;
        for (int k = 0; k < 3; k++) h.location[k] = r.o[k] + r.dir[k] * lam;
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
__global__ __launch_bounds__(NT) void k_raycast(SgDevView d, SgRayView rv, const SgRay *rays, int n, int original,
                                                int stale, SgRayHit *out, float *place, int grab_pair,
                                                long long *stats) {
  __shared__ RayShared<NT> S;
  if (place) {
    if (blockIdx.x != 0) return;
    ray_one(d, rv, rays, original, stale, S, 0, stats);
    if (grab_pair) ray_one(d, rv, rays + 1, original, stale, S, 1, stats);
    if (threadIdx.x == 0) {
      const SgRayHit &a = S.res[0];
      float g[3] = {0.0f, 0.0f, 0.0f};
      if (grab_pair && a.hit && S.res[1].hit) {
        for (int k = 0; k < 3; k++) g[k] = S.res[1].location[k] - a.location[k];
      }
      const bool hit = a.hit && a.status == 0;
      *(float4 *)place = make_float4(a.location[0], a.location[1], a.location[2], g[0]);
      *(float4 *)(place + 4) = make_float4(g[1], g[2], hit ? 1.0f : 0.0f, 0.0f);
      if (out) {
        out[0] = a;
        if (grab_pair) out[1] = S.res[1];
      }
    }
    return;
  }
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    ray_one(d, rv, rays + i, original, stale, S, 0, stats);
    if (threadIdx.x == 0) out[i] = S.res[0];
    __syncthreads();
  }
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
      if (x) hipFree(x);
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
  if (place_dev)
    hipLaunchKernelGGL(k_raycast<1024>, dim3(1), dim3(1024), 0, p->stream, v, rv, rays_dev, n, original, stale,
                       out_dev, place_dev, grab_pair, stats_dev);
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
  if (d.ray_in) hipFree(d.ray_in);
  if (d.ray_out) hipFree(d.ray_out);
  d.ray_in = nullptr;
  d.ray_out = nullptr;
  d.ray_cap = 0;
  if (sg_check_hip(hipMalloc(&d.ray_in, (size_t)n * sizeof(SgRay)), "hipMalloc")) return -1;
  if (sg_check_hip(hipMalloc(&d.ray_out, (size_t)n * sizeof(SgRayHit)), "hipMalloc")) return -1;
  d.ray_cap = n;
  return 0;
}

int sg_raycast_device(PBVH *p, const SgRay *rays_dev, int n, int original, SgRayHit *hits_dev) {
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
  if (st) hipFree(st);
  return rc;
}
