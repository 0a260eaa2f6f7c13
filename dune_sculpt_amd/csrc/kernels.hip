// HIP kernels for gfx950 (CDNA4, wave64): one sculpt dab = search -> brush ->
// [smooth iterations] -> post (normals + leaf bounds), plus the deferred refit of
// internal nodes and the whole-mesh normal kernel.
//
// FP32 expressions are written in exactly the order the spec fixes and the file is
// compiled with -ffp-contract=off, so every rounding matches the CPU statement of
// the same spec; sqrtf and '/' are the correctly rounded HIP defaults.
//
// Layout (see pbvh_build.cpp): vertices SoA in leaf-major order, each leaf's uniq
// vertices one contiguous range; tris int4 in leaf order; vertex->tri and
// vertex->vertex CSR; bounding boxes as 8 floats (min xyz, pad, max xyz, pad).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "sg_internal.h"
#include "sg_kernels.h"

#define SG_WAVE 64

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
int sg_set_error(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}
const char *sg_last_error_str() { return g_err.c_str(); }
int sg_check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  sg_set_error("%s: %s", what, hipGetErrorString(e));
  return -(int)e - 1000;
}

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

// ---------------------------------------------------------------------------
// K1: search.  One workgroup of 1024 threads.
//  phase 0  refresh the boxes of cut subtrees the previous dab touched (ordered
//           reduction of their leaf boxes; these are the only boxes the search needs)
//  phase A  test every cut subtree box, ordered compaction of the passing ones
//  phase B  test the leaves of passing subtrees, ordered compaction -> gathered[]
// The gathered list is in ascending (DFS) leaf order, as BKE_pbvh_search_gather
// returns it.
// ---------------------------------------------------------------------------
#define SEARCH_THREADS 1024
#define SEARCH_WAVES (SEARCH_THREADS / SG_WAVE)
#define MAX_CUT 1024

__global__ __launch_bounds__(SEARCH_THREADS) void k_search(SgDevView d, SgDabK k, long long *log,
                                                           int log_index) {
  __shared__ int s_list[MAX_CUT];
  __shared__ int s_pre[MAX_CUT + 1];
  __shared__ int s_has[MAX_CUT];
  __shared__ int s_wave[SEARCH_WAVES + 1];
  __shared__ int s_total;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / SG_WAVE;

  if (tid == 0 && log && log_index > 0) {
    for (int i = 0; i < SG_CTR_COUNT; i++) log[(size_t)(log_index - 1) * SG_CTR_COUNT + i] = d.ctr[i];
  }

  // phase 0: one wave per previously touched cut
  const int nprev = (int)d.ctr[SG_CTR_NPREV];
  for (int pc = wave; pc < nprev; pc += SEARCH_WAVES) {
    const int c = d.prev_cuts[pc];
    const int lb = d.cut_lbeg[c], le = d.cut_lend[c];
    const int nl = le - lb;
    const int per = (nl + SG_WAVE - 1) / SG_WAVE;
    BB6 acc, oacc;
    bb_reset(acc);
    bb_reset(oacc);
    for (int q = 0; q < per; q++) {
      const int l = lb + lane * per + q;
      if (l < le) {
        bb_add_bb(acc, bb_load(d.leaf_bb + 8 * (size_t)l));
        bb_add_bb(oacc, bb_load(d.leaf_obb + 8 * (size_t)l));
      }
    }
    acc = bb_wave_reduce(acc);
    oacc = bb_wave_reduce(oacc);
    if (lane == 0) {
      bb_store(d.node_bb + 8 * (size_t)d.cut_node[c], acc);
      bb_store(d.node_obb + 8 * (size_t)d.cut_node[c], oacc);
    }
  }
  __syncthreads();

  // phase A
  const float *cut_bbs = k.use_orig ? d.node_obb : d.node_bb;
  const int C = d.C;
  bool pass = false;
  int c = tid;
  if (c < C) {
    const int nd = d.cut_node[c];
    pass = sphere_box(bb_load(cut_bbs + 8 * (size_t)nd), k.cx, k.cy, k.cz, k.r2);
    const int lb = d.cut_lbeg[c];
    if (pass && d.cut_lend[c] - lb == 1 && (d.leaf_flags[lb] & 1)) pass = false;
  }
  {
    const unsigned long long m = __ballot(pass);
    if (lane == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int w = 0; w < SEARCH_WAVES; w++) {
        const int t = s_wave[w];
        s_wave[w] = acc;
        acc += t;
      }
      s_total = acc;
    }
    __syncthreads();
    if (pass) {
      const int pos = s_wave[wave] + mask_prefix(m);
      s_list[pos] = c;
      s_pre[pos] = d.cut_lend[c] - d.cut_lbeg[c];
      s_has[pos] = 0;
    }
    __syncthreads();
  }
  const int npass = s_total;
  // exclusive scan of leaf counts (npass <= 1024): simple two-level scan
  {
    int v = (tid < npass) ? s_pre[tid] : 0;
    // inclusive wave scan
    for (int off = 1; off < SG_WAVE; off <<= 1) {
      const int o = __shfl_up(v, off);
      if (lane >= off) v += o;
    }
    if (lane == SG_WAVE - 1) s_wave[wave] = v;
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int w = 0; w < SEARCH_WAVES; w++) {
        const int t = s_wave[w];
        s_wave[w] = acc;
        acc += t;
      }
      s_wave[SEARCH_WAVES] = acc;
    }
    __syncthreads();
    const int incl = v + s_wave[wave];
    const int cnt = (tid < npass) ? s_pre[tid] : 0;
    __syncthreads();
    if (tid < npass) s_pre[tid] = incl - cnt;
    if (tid == 0) s_pre[npass] = s_wave[SEARCH_WAVES];
    __syncthreads();
  }
  const int total = s_pre[npass];
  const float *leaf_bbs = k.use_orig ? d.leaf_obb : d.leaf_bb;
  int gbase = 0;
  for (int base = 0; base < total; base += SEARCH_THREADS) {
    const int idx = base + tid;
    bool lp = false;
    int leaf = 0;
    if (idx < total) {
      int lo = 0, hi = npass - 1;  // last j with s_pre[j] <= idx
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[mid] <= idx)
          lo = mid;
        else
          hi = mid - 1;
      }
      leaf = d.cut_lbeg[s_list[lo]] + (idx - s_pre[lo]);
      lp = !(d.leaf_flags[leaf] & 1) &&
           sphere_box(bb_load(leaf_bbs + 8 * (size_t)leaf), k.cx, k.cy, k.cz, k.r2);
      if (lp) s_has[lo] = 1;
    }
    const unsigned long long m = __ballot(lp);
    if (lane == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int w = 0; w < SEARCH_WAVES; w++) {
        const int t = s_wave[w];
        s_wave[w] = acc;
        acc += t;
      }
      s_wave[SEARCH_WAVES] = acc;
    }
    __syncthreads();
    if (lp) d.gathered[gbase + s_wave[wave] + mask_prefix(m)] = leaf;
    gbase += s_wave[SEARCH_WAVES];
    __syncthreads();
  }
  // cuts with gathered leaves: dirty for the deferred refit, refreshed by the next dab
  if (tid == 0) {
    int n = 0;
    for (int j = 0; j < npass; j++) {
      if (s_has[j]) {
        d.prev_cuts[n++] = s_list[j];
        d.cut_dirty[s_list[j]] = 1;
      }
    }
    d.ctr[SG_CTR_NPREV] = n;
    d.ctr[SG_CTR_NGATHER] = gbase;
    d.ctr[SG_CTR_NAFF] = 0;
    d.ctr[SG_CTR_AREA_X] = 0;
    d.ctr[SG_CTR_AREA_Y] = 0;
    d.ctr[SG_CTR_AREA_Z] = 0;
    d.ctr[SG_CTR_TESTS] = C + total;
    d.ctr[SG_CTR_NCAND] = 0;
    d.ctr[SG_CTR_NAREA] = 0;
    d.ctr[SG_CTR_NBR] = 0;
    d.ctr[SG_CTR_TRIS] = 0;
    for (int i = 11; i < SG_CTR_COUNT; i++) d.ctr[i] = 0;
  }
}

// wave-wide sums
__device__ __forceinline__ long long wave_sum_ll(long long v) {
  for (int off = SG_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
  for (int off = SG_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// ---------------------------------------------------------------------------
// K2 (Draw only): area normal.  Sum of the normals of uniq verts of gathered leaves
// within normal_radius_factor * r, each component rounded to a multiple of 2^-32
// and accumulated in int64 -- exact, so the atomic order cannot change the bits.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_area(SgDevView d, SgDabK k) {
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  const int ng = (int)d.ctr[SG_CTR_NGATHER];
  long long sx = 0, sy = 0, sz = 0;
  int na = 0;
  for (int g = gw; g < ng; g += nw) {
    const int leaf = d.gathered[g];
    const int vb = d.leaf_vbeg[leaf], ve = d.leaf_vbeg[leaf + 1];
    for (int v = vb + lane; v < ve; v += SG_WAVE) {
      if (d.vhide && d.respect_hide && d.vhide[v]) continue;
      const float dx = k.cx - d.px[v];
      const float dy = k.cy - d.py[v];
      const float dz = k.cz - d.pz[v];
      const float d2 = dx * dx + dy * dy + dz * dz;
      if (d2 <= k.nr2) {
        sx += __double2ll_rn((double)d.nx[v] * 4294967296.0);
        sy += __double2ll_rn((double)d.ny[v] * 4294967296.0);
        sz += __double2ll_rn((double)d.nz[v] * 4294967296.0);
        na++;
      }
    }
  }
  sx = wave_sum_ll(sx);
  sy = wave_sum_ll(sy);
  sz = wave_sum_ll(sz);
  na = wave_sum_i(na);
  if (lane == 0 && na) {
    atomicAdd((unsigned long long *)&d.ctr[SG_CTR_AREA_X], as_u64(sx));
    atomicAdd((unsigned long long *)&d.ctr[SG_CTR_AREA_Y], as_u64(sy));
    atomicAdd((unsigned long long *)&d.ctr[SG_CTR_AREA_Z], as_u64(sz));
    atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NAREA], (unsigned long long)na);
  }
}

// ---------------------------------------------------------------------------
// K3: brush.  One wave per gathered leaf (grid-stride); lanes over the leaf's uniq
// vertices (one coalesced range).  Spec S2 vertex test, S3/S4 factor, S6-S9
// displacement.  Affected vertices are compacted with ballot + prefix popcount and
// one atomic per wave.  For Smooth it also records the factor and slot and runs the
// first Jacobi step (neighbours are read from the untouched positions).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_brush(SgDevView d, SgDabK k, const float *curve) {
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  const int ng = (int)d.ctr[SG_CTR_NGATHER];
  const int V = d.V;
  // Draw offset: (normalize(area) * r) * bstrength
  float ox = 0.0f, oy = 0.0f, oz = 0.0f;
  if (k.tool == SG_TOOL_DRAW) {
    float ax = (float)((double)d.ctr[SG_CTR_AREA_X] * (1.0 / 4294967296.0));
    float ay = (float)((double)d.ctr[SG_CTR_AREA_Y] * (1.0 / 4294967296.0));
    float az = (float)((double)d.ctr[SG_CTR_AREA_Z] * (1.0 / 4294967296.0));
    sg_normalize(ax, ay, az);
    ox = (ax * k.r) * k.bstr;
    oy = (ay * k.r) * k.bstr;
    oz = (az * k.r) * k.bstr;
  }
  int ncand = 0;
  long long nbr = 0;
  for (int g = gw; g < ng; g += nw) {
    const int leaf = d.gathered[g];
    const int vb = d.leaf_vbeg[leaf], ve = d.leaf_vbeg[leaf + 1];
    ncand += ve - vb;
    for (int base = vb; base < ve; base += SG_WAVE) {
      const int v = base + lane;
      bool aff = false;
      float x = 0.0f, y = 0.0f, z = 0.0f, d2 = 0.0f;
      if (v < ve) {
        x = d.px[v];
        y = d.py[v];
        z = d.pz[v];
        const float dx = k.cx - x;
        const float dy = k.cy - y;
        const float dz = k.cz - z;
        d2 = dx * dx + dy * dy + dz * dz;
        aff = d2 <= k.r2;
        if (d.vhide && d.respect_hide && d.vhide[v]) aff = false;
      }
      const unsigned long long m = __ballot(aff);
      if (m == 0ull) continue;
      int wbase = 0;
      if (lane == 0) wbase = (int)atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NAFF], (unsigned long long)__popcll(m));
      wbase = __shfl(wbase, 0);
      if (!aff) continue;
      const int pos = wbase + mask_prefix(m);
      d.affected[pos] = v;
      const float mk = d.mask ? d.mask[v] : 0.0f;
      const float f = sg_falloff(k, curve, sqrtf(d2)) * (1.0f - mk);
      if (k.tool == SG_TOOL_DRAW) {
        d.px[v] = x + ox * f;
        d.py[v] = y + oy * f;
        d.pz[v] = z + oz * f;
      } else if (k.tool == SG_TOOL_INFLATE) {
        const float s = (k.bstr * f) * k.r;
        d.px[v] = x + d.nx[v] * s;
        d.py[v] = y + d.ny[v] * s;
        d.pz[v] = z + d.nz[v] * s;
      } else if (k.tool == SG_TOOL_GRAB) {
        const float fade = k.bstr * f;
        d.px[v] = x + k.gx * fade;
        d.py[v] = y + k.gy * fade;
        d.pz[v] = z + k.gz * fade;
      } else {  // SMOOTH: factor, slot, first Jacobi step into sbuf[0]
        const float fade = k.bstr * f;
        d.aff_fade[pos] = fade;
        d.slot[v] = pos;
        const int b = d.vv_off[v], e = d.vv_off[v + 1];
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        for (int q = b; q < e; q++) {
          const int u = d.vv_idx[q];
          sx = sx + d.px[u];
          sy = sy + d.py[u];
          sz = sz + d.pz[u];
        }
        nbr += e - b;
        float nxv = x, nyv = y, nzv = z;
        if (e > b) {
          const float inv = 1.0f / (float)(e - b);
          const float avx = sx * inv, avy = sy * inv, avz = sz * inv;
          nxv = x + (avx - x) * fade;
          nyv = y + (avy - y) * fade;
          nzv = z + (avz - z) * fade;
        }
        float *s0 = d.sbuf0;
        s0[pos] = nxv;
        s0[(size_t)V + pos] = nyv;
        s0[2 * (size_t)V + pos] = nzv;
      }
    }
  }
  if (k.stats) {
    ncand = wave_sum_i(ncand);
    nbr = wave_sum_ll(nbr);
    if (lane == 0) {
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NCAND], (unsigned long long)ncand);
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NBR], as_u64(nbr));
    }
  }
}

// ---------------------------------------------------------------------------
// K4: Jacobi step it >= 2 of Smooth(k) (spec S8).  Thread per affected vertex; a
// neighbour's previous iterate comes from the compact buffer when it is affected
// (slot >= 0) and from the untouched positions otherwise.  The last step writes the
// positions directly (no reader needs the old value of an affected vertex).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_smooth_iter(SgDevView d, int src_buf, int last, int stats,
                                                     int stamp) {
  const int na = (int)d.ctr[SG_CTR_NAFF];
  const int V = d.V;
  const float *S = src_buf ? d.sbuf1 : d.sbuf0;
  float *D = src_buf ? d.sbuf0 : d.sbuf1;
  long long nbr = 0, r1 = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
    const int v = d.affected[i];
    const float x = S[i], y = S[(size_t)V + i], z = S[2 * (size_t)V + i];
    const int b = d.vv_off[v], e = d.vv_off[v + 1];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int q = b; q < e; q++) {
      const int u = d.vv_idx[q];
      const int su = d.slot[u];
      if (su >= 0) {
        sx = sx + S[su];
        sy = sy + S[(size_t)V + su];
        sz = sz + S[2 * (size_t)V + su];
      } else {
        sx = sx + d.px[u];
        sy = sy + d.py[u];
        sz = sz + d.pz[u];
        if (stats && atomicExch(&d.stamp_v[u], stamp) != stamp) r1++;
      }
    }
    nbr += e - b;
    const float fade = d.aff_fade[i];
    float nxv = x, nyv = y, nzv = z;
    if (e > b) {
      const float inv = 1.0f / (float)(e - b);
      const float avx = sx * inv, avy = sy * inv, avz = sz * inv;
      nxv = x + (avx - x) * fade;
      nyv = y + (avy - y) * fade;
      nzv = z + (avz - z) * fade;
    }
    if (last) {
      d.px[v] = nxv;
      d.py[v] = nyv;
      d.pz[v] = nzv;
    } else {
      D[i] = nxv;
      D[(size_t)V + i] = nyv;
      D[2 * (size_t)V + i] = nzv;
    }
  }
  if (stats) {
    nbr = wave_sum_ll(nbr);
    r1 = wave_sum_ll(r1);
    if (lane_id() == 0) {
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NBR], as_u64(nbr));
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_R1X], as_u64(r1));
    }
  }
}

// Smooth(1): copy the single iterate back.
__global__ __launch_bounds__(256) void k_smooth_copy(SgDevView d) {
  const int na = (int)d.ctr[SG_CTR_NAFF];
  const int V = d.V;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
    const int v = d.affected[i];
    d.px[v] = d.sbuf0[i];
    d.py[v] = d.sbuf0[(size_t)V + i];
    d.pz[v] = d.sbuf0[2 * (size_t)V + i];
  }
}

// Normal of vertex v (spec S10): normalize(sum over incident tris, in CSR order, of
// normal_tri_v3).  Deterministic gather, no atomics.
__device__ __forceinline__ void vertex_normal(const SgDevView &d, int v, int stats, int stamp,
                                              long long &ntri, long long &nfa, long long &nvfa) {
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  const int b = d.vt_off[v], e = d.vt_off[v + 1];
  for (int q = b; q < e; q++) {
    const int t = d.vt_idx[q] & 0x3fffffff;
    const int4 c = d.tri[t];
    float fx, fy, fz;
    sg_tri_normal(d.px[c.x], d.py[c.x], d.pz[c.x], d.px[c.y], d.py[c.y], d.pz[c.y], d.px[c.z],
                  d.py[c.z], d.pz[c.z], fx, fy, fz);
    ax = ax + fx;
    ay = ay + fy;
    az = az + fz;
    if (stats) {
      if (atomicExch(&d.stamp_t[t], stamp) != stamp) {
        nfa++;
        if (atomicExch(&d.stamp_v2[c.x], stamp) != stamp) nvfa++;
        if (atomicExch(&d.stamp_v2[c.y], stamp) != stamp) nvfa++;
        if (atomicExch(&d.stamp_v2[c.z], stamp) != stamp) nvfa++;
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
__device__ __forceinline__ void leaf_refit(const SgDevView &d, int leaf, long long &nsh) {
  const int vb = d.leaf_vbeg[leaf], ve = d.leaf_vbeg[leaf + 1];
  const int sb = d.leaf_sh_off[leaf], se = d.leaf_sh_off[leaf + 1];
  const int nu = ve - vb, ns = se - sb, tot = nu + ns;
  const int per = (tot + SG_WAVE - 1) / SG_WAVE;
  const int lane = lane_id();
  BB6 acc;
  bb_reset(acc);
  for (int q = 0; q < per; q++) {
    const int i = lane * per + q;
    if (i < tot) {
      const int v = (i < nu) ? vb + i : d.leaf_sh[sb + i - nu];
      bb_add_point(acc, d.px[v], d.py[v], d.pz[v]);
    }
  }
  acc = bb_wave_reduce(acc);
  if (lane == 0) {
    bb_store(d.leaf_bb + 8 * (size_t)leaf, acc);
    bb_store(d.leaf_obb + 8 * (size_t)leaf, acc);
  }
  nsh += ns;
}

// ---------------------------------------------------------------------------
// K5: post.  Normals of the affected vertices and bounds of the gathered leaves (all
// gathered leaves are marked, BKE_pbvh_node_mark_update).  Both only read positions.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_post(SgDevView d, int smooth, int stats, int stamp) {
  const int na = (int)d.ctr[SG_CTR_NAFF];
  const int ng = (int)d.ctr[SG_CTR_NGATHER];
  long long ntri = 0, nfa = 0, nvfa = 0, nsh = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
    const int v = d.affected[i];
    vertex_normal(d, v, stats, stamp, ntri, nfa, nvfa);
    if (smooth) d.slot[v] = -1;
  }
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  for (int g = gw; g < ng; g += nw) leaf_refit(d, d.gathered[g], nsh);
  if (stats) {
    ntri = wave_sum_ll(ntri);
    nfa = wave_sum_ll(nfa);
    nvfa = wave_sum_ll(nvfa);
    if (lane_id() == 0) {
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_TRIS], as_u64(ntri));
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_FA], as_u64(nfa));
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_VFA], as_u64(nvfa));
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_SHARED], as_u64(nsh / SG_WAVE));
    }
  }
}

// ---------------------------------------------------------------------------
// Deferred refit of internal nodes.  Sub: one workgroup per dirty cut subtree,
// leaf boxes staged in LDS, internal nodes computed height by height
// (parent = BB_expand_with_bb(child0, child1), pbvh.c:123-124).  Top: one workgroup
// over the tree above the cut, with the cut roots as its leaves.
// ---------------------------------------------------------------------------
#define FLUSH_THREADS 256
__global__ __launch_bounds__(FLUSH_THREADS) void k_flush_sub(SgDevView d, int maxn) {
  extern __shared__ float s_bb[];  // [maxn * 6]
  const int c = blockIdx.x;
  if (!d.cut_dirty[c]) return;
  const int lb = d.cut_lbeg[c], nl = d.cut_lend[c] - lb;
  for (int i = threadIdx.x; i < nl; i += blockDim.x) {
    const BB6 b = bb_load(d.leaf_bb + 8 * (size_t)(lb + i));
    for (int q = 0; q < 3; q++) {
      s_bb[6 * i + q] = b.mn[q];
      s_bb[6 * i + 3 + q] = b.mx[q];
    }
  }
  __syncthreads();
  const int eb = d.sub_off[c], ee = d.sub_off[c + 1];
  int h = 1;
  for (int done = eb; done < ee && h < 256; h++) {
    int nxt = done;
    // entries are sorted by height: [done, nxt) is this height
    while (nxt < ee && d.sub_tab[nxt].w == h) nxt++;
    for (int e = done + threadIdx.x; e < nxt; e += blockDim.x) {
      const int4 t = d.sub_tab[e];
      const int li = nl + (e - eb);
      for (int q = 0; q < 3; q++) {
        s_bb[6 * li + q] = sg_min(s_bb[6 * t.y + q], s_bb[6 * t.z + q]);
        s_bb[6 * li + 3 + q] = sg_max(s_bb[6 * t.y + 3 + q], s_bb[6 * t.z + 3 + q]);
      }
    }
    __syncthreads();
    done = nxt;
  }
  for (int e = eb + threadIdx.x; e < ee; e += blockDim.x) {
    const int li = nl + (e - eb);
    BB6 b;
    for (int q = 0; q < 3; q++) {
      b.mn[q] = s_bb[6 * li + q];
      b.mx[q] = s_bb[6 * li + 3 + q];
    }
    const int nd = d.sub_tab[e].x;
    bb_store(d.node_bb + 8 * (size_t)nd, b);
    bb_store(d.node_obb + 8 * (size_t)nd, b);
  }
  if (threadIdx.x == 0 && ee == eb) {  // cut root is a leaf: mirror its box
    const BB6 b = bb_load(d.leaf_bb + 8 * (size_t)lb);
    bb_store(d.node_bb + 8 * (size_t)d.cut_node[c], b);
    bb_store(d.node_obb + 8 * (size_t)d.cut_node[c], b);
  }
}

__global__ __launch_bounds__(1024) void k_flush_top(SgDevView d, int top_n, int clear) {
  extern __shared__ float s_bb[];  // [(C + top_n) * 6]
  const int C = d.C;
  for (int i = threadIdx.x; i < C; i += blockDim.x) {
    const BB6 b = bb_load(d.node_bb + 8 * (size_t)d.cut_node[i]);
    for (int q = 0; q < 3; q++) {
      s_bb[6 * i + q] = b.mn[q];
      s_bb[6 * i + 3 + q] = b.mx[q];
    }
  }
  __syncthreads();
  int h = 1;
  for (int done = 0; done < top_n && h < 256; h++) {
    int nxt = done;
    while (nxt < top_n && d.top_tab[nxt].w == h) nxt++;
    for (int e = done + threadIdx.x; e < nxt; e += blockDim.x) {
      const int4 t = d.top_tab[e];
      const int li = C + e;
      for (int q = 0; q < 3; q++) {
        s_bb[6 * li + q] = sg_min(s_bb[6 * t.y + q], s_bb[6 * t.z + q]);
        s_bb[6 * li + 3 + q] = sg_max(s_bb[6 * t.y + 3 + q], s_bb[6 * t.z + 3 + q]);
      }
    }
    __syncthreads();
    done = nxt;
  }
  for (int e = threadIdx.x; e < top_n; e += blockDim.x) {
    const int li = C + e;
    BB6 b;
    for (int q = 0; q < 3; q++) {
      b.mn[q] = s_bb[6 * li + q];
      b.mx[q] = s_bb[6 * li + 3 + q];
    }
    const int nd = d.top_tab[e].x;
    bb_store(d.node_bb + 8 * (size_t)nd, b);
    bb_store(d.node_obb + 8 * (size_t)nd, b);
  }
  if (clear) {
    for (int i = threadIdx.x; i < C; i += blockDim.x) d.cut_dirty[i] = 0;
  }
}

// Refresh cut-root boxes of all dirty cuts from their leaves (used before the top
// flush when the sub flush is skipped for leaf-root cuts) -- folded into k_flush_sub.

// ---------------------------------------------------------------------------
// Host-driven updates: refit a list of leaves; normals of marked verts of a list of
// leaves (BKE_pbvh_update_bounds / BKE_pbvh_update_normals with caller flags).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_refit_list(SgDevView d, const int *leaves, int n) {
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  long long nsh = 0;
  for (int g = gw; g < n; g += nw) {
    const int leaf = leaves[g];
    leaf_refit(d, leaf, nsh);
    if (lane_id() == 0) d.cut_dirty[d.leaf_cut[leaf]] = 1;
  }
}

__global__ __launch_bounds__(256) void k_normals_list(SgDevView d, const int *leaves, int n) {
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  long long a = 0, b = 0, c = 0;
  for (int g = gw; g < n; g += nw) {
    const int leaf = leaves[g];
    const int vb = d.leaf_vbeg[leaf], ve = d.leaf_vbeg[leaf + 1];
    for (int v = vb + lane_id(); v < ve; v += SG_WAVE) {
      if (!d.vmark[v]) continue;
      vertex_normal(d, v, 0, 0, a, b, c);
      d.vmark[v] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// Whole-mesh angle-weighted vertex normals (KERNEL_mesh_calc_normals_poly_and_vertex,
// mesh_normals.c:225-299) as a per-vertex CSR gather in ascending tri order: the poly
// normal by Newell's method, weighted by the corner angle saacos(-dot(e_prev,
// e_next)) with e_prev = normalize(c_prev - c_i), e_next = normalize(c_i - c_next);
// zero result falls back to the normalized position.  mode 1 uses normal_tri_v3 for
// the face normal (the looptri variant, mesh_normals.c:455-507).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sg_saacos(float f) {
  return (f <= -1.0f) ? 3.14159265358979323846f : ((f >= 1.0f) ? 0.0f : acosf(f));
}

__global__ __launch_bounds__(256) void k_full_normals(SgDevView d, int mode) {
  const int V = d.V;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (int q = d.vt_off[v]; q < d.vt_off[v + 1]; q++) {
      const int e = d.vt_idx[q];
      const int t = e & 0x3fffffff;
      const int ci = (int)((unsigned)e >> 30);
      const int4 c = d.tri[t];
      const int id[3] = {c.x, c.y, c.z};
      float P[3][3];
      for (int j = 0; j < 3; j++) {
        P[j][0] = d.px[id[j]];
        P[j][1] = d.py[id[j]];
        P[j][2] = d.pz[id[j]];
      }
      float fx, fy, fz;
      if (mode == 0) {
        fx = 0.0f;
        fy = 0.0f;
        fz = 0.0f;
        // Newell over corners (v_curr starts at the last corner), add_newell_cross
        int cur = 2;
        for (int nx = 0; nx <= 2; nx++) {
          const float *a = P[cur], *b = P[nx];
          fx += (a[1] - b[1]) * (a[2] + b[2]);
          fy += (a[2] - b[2]) * (a[0] + b[0]);
          fz += (a[0] - b[0]) * (a[1] + b[1]);
          cur = nx;
        }
        if (sg_normalize(fx, fy, fz) == 0.0f) fz = 1.0f;
      } else {
        sg_tri_normal(P[0][0], P[0][1], P[0][2], P[1][0], P[1][1], P[1][2], P[2][0], P[2][1],
                      P[2][2], fx, fy, fz);
      }
      const int ip = (ci + 2) % 3, in = (ci + 1) % 3;
      float epx = P[ip][0] - P[ci][0], epy = P[ip][1] - P[ci][1], epz = P[ip][2] - P[ci][2];
      float enx = P[ci][0] - P[in][0], eny = P[ci][1] - P[in][1], enz = P[ci][2] - P[in][2];
      sg_normalize(epx, epy, epz);
      sg_normalize(enx, eny, enz);
      const float fac = sg_saacos(-(epx * enx + epy * eny + epz * enz));
      ax += fx * fac;
      ay += fy * fac;
      az += fz * fac;
    }
    if (sg_normalize(ax, ay, az) == 0.0f) {
      ax = d.px[v];
      ay = d.py[v];
      az = d.pz[v];
      sg_normalize(ax, ay, az);
    }
    d.nx[v] = ax;
    d.ny[v] = ay;
    d.nz[v] = az;
  }
}

// Whole-mesh coordinate replace (BKE_pbvh_vert_coords_apply pbvh.c:4707-4749):
// bitwise compare, copy, mark.  Streams V*12 B in, V*12 B out.
__global__ __launch_bounds__(256) void k_coords_apply(SgDevView d, const float *src_old_order) {
  const int V = d.V;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const int o = d.vperm_old[v];
    const float x = src_old_order[3 * (size_t)o], y = src_old_order[3 * (size_t)o + 1],
                z = src_old_order[3 * (size_t)o + 2];
    const bool same = __float_as_uint(x) == __float_as_uint(d.px[v]) &&
                      __float_as_uint(y) == __float_as_uint(d.py[v]) &&
                      __float_as_uint(z) == __float_as_uint(d.pz[v]);
    if (!same) {
      d.px[v] = x;
      d.py[v] = y;
      d.pz[v] = z;
      d.vmark[v] = 1;
    }
  }
}

__global__ __launch_bounds__(256) void k_refit_all(SgDevView d) {
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  long long nsh = 0;
  for (int l = gw; l < d.L; l += nw) leaf_refit(d, l, nsh);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < d.C; c += gridDim.x * blockDim.x)
    d.cut_dirty[c] = 1;
}

__global__ void k_clear_marks(SgDevView d) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < d.V; v += gridDim.x * blockDim.x) d.vmark[v] = 0;
}

// gather a device array in original vertex order (AoS float3)
__global__ __launch_bounds__(256) void k_gather_aos(const float *x, const float *y, const float *z,
                                                    const int *perm_old, int V, float *out) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const int o = perm_old[v];
    out[3 * (size_t)o] = x[v];
    out[3 * (size_t)o + 1] = y[v];
    out[3 * (size_t)o + 2] = z[v];
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
SgDevView sg_view(PBVH *p) {
  SgDev &d = p->d;
  SgDevView v;
  v.V = d.V;
  v.T = d.T;
  v.L = d.L;
  v.N = d.N;
  v.C = d.C;
  v.respect_hide = p->respect_hide ? 1 : 0;
  v.px = d.px;
  v.py = d.py;
  v.pz = d.pz;
  v.nx = d.nx;
  v.ny = d.ny;
  v.nz = d.nz;
  v.mask = d.mask;
  v.tri = d.tri;
  v.vt_off = d.vt_off;
  v.vt_idx = d.vt_idx;
  v.vv_off = d.vv_off;
  v.vv_idx = d.vv_idx;
  v.leaf_vbeg = d.leaf_vbeg;
  v.leaf_sh_off = d.leaf_sh_off;
  v.leaf_sh = d.leaf_sh;
  v.leaf_flags = d.leaf_flags;
  v.leaf_cut = d.leaf_cut;
  v.leaf_bb = d.leaf_bb;
  v.leaf_obb = d.leaf_obb;
  v.node_bb = d.node_bb;
  v.node_obb = d.node_obb;
  v.cut_node = d.cut_node;
  v.cut_lbeg = d.cut_lbeg;
  v.cut_lend = d.cut_lend;
  v.sub_off = d.sub_off;
  v.sub_tab = d.sub_tab;
  v.top_tab = d.top_tab;
  v.cut_dirty = d.cut_dirty;
  v.vhide = d.vhide;
  v.vmark = d.vmark;
  v.gathered = d.gathered;
  v.affected = d.affected;
  v.aff_fade = d.aff_fade;
  v.slot = d.slot;
  v.sbuf0 = d.sbuf[0];
  v.sbuf1 = d.sbuf[1];
  v.prev_cuts = d.prev_cuts;
  v.ctr = d.ctr;
  v.vperm_old = d.vperm_old;
  v.stamp_v = d.stamp_v;
  v.stamp_v2 = d.stamp_v2;
  v.stamp_t = d.stamp_t;
  return v;
}

SgDabK sg_make_dab(const SgDabParams *dab, int stats) {
  SgDabK k;
  memset(&k, 0, sizeof(k));
  k.tool = dab->sculpt_tool;
  k.falloff = dab->curve_preset;
  k.cx = dab->location[0];
  k.cy = dab->location[1];
  k.cz = dab->location[2];
  k.r = dab->radius;
  k.r2 = k.r * k.r;
  const float nr = k.r * dab->normal_radius_factor;
  k.nr2 = nr * nr;
  k.hardness = dab->hardness;
  k.gx = dab->grab_delta[0];
  k.gy = dab->grab_delta[1];
  k.gz = dab->grab_delta[2];
  k.iters = dab->smooth_iterations > 0 ? dab->smooth_iterations : 1;
  k.use_orig = (k.tool == SG_TOOL_GRAB) ? 1 : 0;
  k.stats = stats;
  // Spec S5 bstrength
  const float s = dab->strength;
  const float flip = (dab->flag & SG_BRUSH_DIR_IN) ? -1.0f : 1.0f;
  const float alpha = s * s;
  switch (k.tool) {
    case SG_TOOL_DRAW: k.bstr = alpha * flip; break;
    case SG_TOOL_INFLATE: k.bstr = 0.25f * alpha * flip; break;
    case SG_TOOL_SMOOTH: k.bstr = s < 0.0f ? 0.0f : (s > 1.0f ? 1.0f : s); break;
    case SG_TOOL_GRAB: k.bstr = s; break;
    default: k.bstr = 0.0f; break;
  }
  return k;
}

int sg_enqueue_dab(PBVH *p, const SgDabK &k, const float *curve_dev, long long *log, int log_index,
                   int stamp) {
  const SgDevView v = sg_view(p);
  hipStream_t s = p->stream;
  const int G = p->grid_wg;
  hipLaunchKernelGGL(k_search, dim3(1), dim3(SEARCH_THREADS), 0, s, v, k, log, log_index);
  if (k.tool == SG_TOOL_DRAW) hipLaunchKernelGGL(k_area, dim3(G), dim3(256), 0, s, v, k);
  hipLaunchKernelGGL(k_brush, dim3(G), dim3(256), 0, s, v, k, curve_dev);
  if (k.tool == SG_TOOL_SMOOTH) {
    if (k.iters == 1) {
      hipLaunchKernelGGL(k_smooth_copy, dim3(G), dim3(256), 0, s, v);
    } else {
      int src = 0;
      for (int it = 2; it <= k.iters; it++) {
        hipLaunchKernelGGL(k_smooth_iter, dim3(G), dim3(256), 0, s, v, src, it == k.iters ? 1 : 0,
                           k.stats, stamp);
        src ^= 1;
      }
    }
  }
  hipLaunchKernelGGL(k_post, dim3(G), dim3(256), 0, s, v, k.tool == SG_TOOL_SMOOTH ? 1 : 0, k.stats,
                     stamp);
  return sg_check_hip(hipGetLastError(), "dab launch");
}

int sg_enqueue_flush(PBVH *p) {
  const SgDevView v = sg_view(p);
  hipStream_t s = p->stream;
  const size_t sub_lds = (size_t)p->d.sub_maxn * 6 * sizeof(float);
  hipLaunchKernelGGL(k_flush_sub, dim3(p->d.C), dim3(FLUSH_THREADS), sub_lds, s, v, p->d.sub_maxn);
  const size_t top_lds = (size_t)(p->d.C + p->d.top_n) * 6 * sizeof(float);
  hipLaunchKernelGGL(k_flush_top, dim3(1), dim3(1024), top_lds, s, v, p->d.top_n, 1);
  return sg_check_hip(hipGetLastError(), "flush launch");
}

int sg_enqueue_full_normals(PBVH *p, int mode) {
  const SgDevView v = sg_view(p);
  const int blocks = (p->d.V + 255) / 256;
  hipLaunchKernelGGL(k_full_normals, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, p->stream, v, mode);
  return sg_check_hip(hipGetLastError(), "normals launch");
}

int sg_enqueue_refit_list(PBVH *p, const int *leaves_dev, int n) {
  const SgDevView v = sg_view(p);
  if (n > 0) hipLaunchKernelGGL(k_refit_list, dim3(p->grid_wg), dim3(256), 0, p->stream, v, leaves_dev, n);
  return sg_check_hip(hipGetLastError(), "refit launch");
}

int sg_enqueue_normals_list(PBVH *p, const int *leaves_dev, int n) {
  const SgDevView v = sg_view(p);
  if (n > 0) hipLaunchKernelGGL(k_normals_list, dim3(p->grid_wg), dim3(256), 0, p->stream, v, leaves_dev, n);
  return sg_check_hip(hipGetLastError(), "normals launch");
}

int sg_enqueue_coords_apply(PBVH *p, const float *src_dev) {
  const SgDevView v = sg_view(p);
  const int blocks = (p->d.V + 255) / 256;
  hipLaunchKernelGGL(k_coords_apply, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, p->stream, v, src_dev);
  hipLaunchKernelGGL(k_refit_all, dim3(p->grid_wg), dim3(256), 0, p->stream, v);
  return sg_check_hip(hipGetLastError(), "coords launch");
}

int sg_enqueue_clear_marks(PBVH *p) {
  const SgDevView v = sg_view(p);
  hipLaunchKernelGGL(k_clear_marks, dim3(256), dim3(256), 0, p->stream, v);
  return sg_check_hip(hipGetLastError(), "clear launch");
}

int sg_enqueue_gather(PBVH *p, int which, float *out_dev) {
  const SgDev &d = p->d;
  const int blocks = (d.V + 255) / 256;
  if (which == 0)
    hipLaunchKernelGGL(k_gather_aos, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, p->stream, d.px, d.py,
                       d.pz, d.vperm_old, d.V, out_dev);
  else
    hipLaunchKernelGGL(k_gather_aos, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, p->stream, d.nx, d.ny,
                       d.nz, d.vperm_old, d.V, out_dev);
  return sg_check_hip(hipGetLastError(), "gather launch");
}
