// Learned Smooth: an MLP over 1-ring features that replaces k Jacobi steps of the
// plain Smooth brush, with BF16 MFMA inference (FP32 accumulate), FP32 training
// kernels (forward, backward, ordered gradient reduction, Adam), pair collection from
// the plain brush, and the per-dab fallback decision.
//
// Model: 28 -> 64 (SiLU) -> 64 (SiLU) -> 3, 6,211 parameters.
// Input (per affected vertex v with Smooth factor f and m edge neighbours u_j in CSR
// order): a local frame (t1, t2, n) from the vertex normal n and the first edge;
// h = mean edge length; offsets R (p(u_j) - p(v)) / h for j < 8 (zero padded);
// then m/8, f, f^2, 1 - (1 - f)^4.  Output: the displacement R (p_k(v) - p(v)) / h of
// plain Smooth(k), normalised per component.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "sg_batch.h"
#include "sg_device.h"
#include "sg_internal.h"
#include "sg_kernels.h"
#include "sg_learned.h"

typedef short sg_short4 __attribute__((ext_vector_type(4)));
typedef float sg_float4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN stays NaN or becomes Inf: both fail the
// fallback test, which is all that matters for a corrupted model)
__host__ __device__ inline uint16_t sg_f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float sg_bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

// SiLU with the hardware exp and reciprocal (about 1 ulp each; the network runs in
// BF16 for inference and its FP32 gradients are checked to 2e-4 relative)
__device__ inline float sg_sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float sg_silu(float z) { return z * sg_sigmoid(z); }
__device__ inline float sg_silu_grad(float z) {
  const float s = sg_sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// MFMA operand layout of v_mfma_f32_16x16x16_bf16 (checked on the GPU by
// tests/test_gpu_learned.py::test_mfma_layout): lane L holds A[L%16][4(L/16)+i],
// B[4(L/16)+i][L%16] for i < 4, and D[4(L/16)+i][L%16].
// ---------------------------------------------------------------------------
__global__ void k_mfma_probe(const uint16_t *A, const uint16_t *B, float *D) {
  // A, B: 16x16 bf16 row-major; D: 16x16 f32 row-major
  const int L = threadIdx.x;
  const int r = L & 15, kq = (L >> 4) * 4;
  sg_short4 a, b;
  for (int i = 0; i < 4; i++) {
    a[i] = (short)A[r * 16 + kq + i];
    b[i] = (short)B[(kq + i) * 16 + r];
  }
  sg_float4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  for (int i = 0; i < 4; i++) D[(kq + i) * 16 + r] = acc[i];
}

// Inference weights in LDS: W1t[64][32], W2t[64][64], W3t[16][64] (output-major, so a
// B fragment is one 8-byte read), biases b1[64], b2[64], b3[16].  Same layout as the
// packed device copy, so the block loads it with 16-byte vector loads.
struct __align__(16) MlpLds {
  uint16_t w1t[64 * 32];
  uint16_t w2t[64 * 64];
  uint16_t w3t[16 * 64];
  float b1[64], b2[64], b3[16];
};
#define MLP_W_VEC ((64 * 32 + 64 * 64 + 16 * 64) * 2 / 16)  // 896 uint4
#define MLP_B_VEC (144 / 4)                                 // 36 float4

// Every load is issued before the first LDS store (a fixed, unrolled count per thread),
// so the fill costs one memory round trip.
template <int NT>
__device__ inline void mlp_lds_load(MlpLds &s, const SgMlpView &m) {
  const int t = threadIdx.x;
  constexpr int NW = (MLP_W_VEC + NT - 1) / NT;
  const uint4 *src = (const uint4 *)m.wb;
  uint4 *dst = (uint4 *)s.w1t;
  uint4 r[NW];
#pragma unroll
  for (int j = 0; j < NW; j++) {
    const int i = t + j * NT;
    r[j] = i < MLP_W_VEC ? src[i] : make_uint4(0u, 0u, 0u, 0u);
  }
  const float4 b = t < MLP_B_VEC ? ((const float4 *)m.bias)[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int j = 0; j < NW; j++) {
    const int i = t + j * NT;
    if (i < MLP_W_VEC) dst[i] = r[j];
  }
  if (t < MLP_B_VEC) ((float4 *)s.b1)[t] = b;
}

// Activation staging rows are padded to 72 bf16 (144 B) so the 8-byte MFMA fragment
// reads of 16 consecutive rows fall in distinct LDS banks.
#define MLP_ST 72

// One wave evaluates the MLP for 16 rows: lane r < 16 supplies row r's normalised
// inputs x[28] and receives y[3]; all 64 lanes run the MFMAs (28 of them: 8 + 16 + 4),
// each lane 16 SiLUs per hidden layer.  bufA / bufB are this wave's 16 x MLP_ST bf16
// staging rows, sOut 16x4 floats.  A short chain per wave: the dab kernels are latency
// bound, so more waves with 16 rows each beat fewer waves with 64.
__device__ inline void mlp_wave16(const MlpLds &s, uint16_t *bufA, uint16_t *bufB, float *sOut, const float *x,
                                  float *y) {
  const int L = __lane_id();
  const int r = L & 15, kq = (L >> 4) * 4;
  if (L < 16) {
    uint32_t *rowA = (uint32_t *)&bufA[L * MLP_ST];
#pragma unroll
    for (int k = 0; k < 14; k++) rowA[k] = (uint32_t)sg_f2bf(x[2 * k]) | ((uint32_t)sg_f2bf(x[2 * k + 1]) << 16);
    rowA[14] = 0u;
    rowA[15] = 0u;
  }
  wave_lds_sync();
  // layer 1: [16 x 32] x [32 x 64]
  sg_float4 acc[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) acc[nt] = sg_float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kt = 0; kt < 2; kt++) {
    const sg_short4 a = *(const sg_short4 *)&bufA[r * MLP_ST + kt * 16 + kq];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const sg_short4 bb = *(const sg_short4 *)&s.w1t[(nt * 16 + r) * 32 + kt * 16 + kq];
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[nt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float z = acc[nt][i] + s.b1[nt * 16 + r];
      bufB[(kq + i) * MLP_ST + nt * 16 + r] = sg_f2bf(sg_silu(z));
    }
  }
  wave_lds_sync();
  // layer 2: [16 x 64] x [64 x 64]
#pragma unroll
  for (int nt = 0; nt < 4; nt++) acc[nt] = sg_float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kt = 0; kt < 4; kt++) {
    const sg_short4 a = *(const sg_short4 *)&bufB[r * MLP_ST + kt * 16 + kq];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const sg_short4 bb = *(const sg_short4 *)&s.w2t[(nt * 16 + r) * 64 + kt * 16 + kq];
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[nt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float z = acc[nt][i] + s.b2[nt * 16 + r];
      bufA[(kq + i) * MLP_ST + nt * 16 + r] = sg_f2bf(sg_silu(z));
    }
  }
  wave_lds_sync();
  // layer 3: [16 x 64] x [64 x 16] (3 used)
  sg_float4 o = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int kt = 0; kt < 4; kt++) {
    const sg_short4 a = *(const sg_short4 *)&bufA[r * MLP_ST + kt * 16 + kq];
    const sg_short4 bb = *(const sg_short4 *)&s.w3t[r * 64 + kt * 16 + kq];
    o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, o, 0, 0, 0);
  }
  if (r < 4) {
#pragma unroll
    for (int i = 0; i < 4; i++) sOut[(kq + i) * 4 + r] = o[i] + s.b3[r];
  }
  wave_lds_sync();
  if (L < 16) {
    y[0] = sOut[L * 4 + 0];
    y[1] = sOut[L * 4 + 1];
    y[2] = sOut[L * 4 + 2];
  }
  wave_lds_sync();
}

// Features of vertex v from the current positions/normals; returns false when the
// vertex is outside what the model was trained on (valence outside [3, 8] or a
// degenerate frame).
__device__ inline bool learned_features(const SgDevView &d, int2 vv, float x, float y, float z, float nx, float ny,
                                        float nz, float fade, float *feat, float *R, float &h) {
  const int m = vv.y - vv.x;
  int u[8];
#pragma unroll
  for (int j = 0; j < 8; j++) u[j] = (j < m) ? d.vv_idx[vv.x + j] : -1;
  float ex[8], ey[8], ez[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    if (u[j] >= 0) {
      ex[j] = d.px[u[j]] - x;
      ey[j] = d.py[u[j]] - y;
      ez[j] = d.pz[u[j]] - z;
    } else {
      ex[j] = ey[j] = ez[j] = 0.0f;
    }
  }
  const int mm = m < 8 ? m : 8;
  float hs = 0.0f;
  for (int j = 0; j < mm; j++) hs += sqrtf(ex[j] * ex[j] + ey[j] * ey[j] + ez[j] * ez[j]);
  h = mm > 0 ? hs / (float)mm : 0.0f;
  // frame: t1 = e0 projected on the tangent plane, t2 = n x t1
  const float en = ex[0] * nx + ey[0] * ny + ez[0] * nz;
  float t1x = ex[0] - en * nx, t1y = ey[0] - en * ny, t1z = ez[0] - en * nz;
  const float t1l = sqrtf(t1x * t1x + t1y * t1y + t1z * t1z);
  bool ok = (m >= 3 && m <= 8 && h > 0.0f && t1l > 1e-6f * h);
  const float it = ok ? 1.0f / t1l : 0.0f;
  t1x *= it;
  t1y *= it;
  t1z *= it;
  const float t2x = ny * t1z - nz * t1y, t2y = nz * t1x - nx * t1z, t2z = nx * t1y - ny * t1x;
  R[0] = t1x;
  R[1] = t1y;
  R[2] = t1z;
  R[3] = t2x;
  R[4] = t2y;
  R[5] = t2z;
  R[6] = nx;
  R[7] = ny;
  R[8] = nz;
  const float ih = ok ? 1.0f / h : 0.0f;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    feat[3 * j + 0] = (t1x * ex[j] + t1y * ey[j] + t1z * ez[j]) * ih;
    feat[3 * j + 1] = (t2x * ex[j] + t2y * ey[j] + t2z * ez[j]) * ih;
    feat[3 * j + 2] = (nx * ex[j] + ny * ey[j] + nz * ez[j]) * ih;
  }
  const float omf = 1.0f - fade;
  feat[24] = (float)m * 0.125f;
  feat[25] = fade;
  feat[26] = fade * fade;
  feat[27] = 1.0f - (omf * omf) * (omf * omf);
  return ok;
}

// ---------------------------------------------------------------------------
// Pair collection from the plain brush: features of the affected vertices before the
// dab (positions are untouched until the last Jacobi step) and, after it, the
// displacement in their local frames.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_learned_feat(SgDevView d, float *X, float *frame, int cap) {
  // one row per slot of the affected list; a slot without a record gives a row marked
  // not valid (vertex -1)
  const int na = (int)d.ctr[SG_CTR_NGATHER] << d.aff_shift;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < na && i < cap; i += gridDim.x * blockDim.x) {
    const int v = aff_slot_vertex(d, i, d.aff_cnt[i >> d.aff_shift]);
    if (v < 0) {
      for (int k = 0; k < 28; k++) X[(size_t)i * 28 + k] = 0.0f;
      float *fr = frame + (size_t)i * 16;
      for (int k = 0; k < 16; k++) fr[k] = 0.0f;
      fr[14] = __int_as_float(-1);
      continue;
    }
    const int2 vv = d.aff_vv[i];
    const float fade = d.aff_fade[i];
    const float x = d.px[v], y = d.py[v], z = d.pz[v];
    float feat[28], R[9], h;
    const bool ok = learned_features(d, vv, x, y, z, d.nx[v], d.ny[v], d.nz[v], fade, feat, R, h);
    for (int k = 0; k < 28; k++) X[(size_t)i * 28 + k] = feat[k];
    float *fr = frame + (size_t)i * 16;
    for (int k = 0; k < 9; k++) fr[k] = R[k];
    fr[9] = h;
    fr[10] = x;
    fr[11] = y;
    fr[12] = z;
    fr[13] = ok ? 1.0f : 0.0f;
    fr[14] = __int_as_float(v);  // index bits, exact for any size
  }
}

__global__ __launch_bounds__(256) void k_learned_target(SgDevView d, const float *frame, float *T, int cap) {
  const long long na = (long long)(int)d.ctr[SG_CTR_NGATHER] << d.aff_shift;
  const int n = (int)(na < cap ? na : cap);  // rows the feature kernel wrote
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float *fr = frame + (size_t)i * 16;
    const int v = __float_as_int(fr[14]);
    if (v < 0) {  // a slot without a record
      T[(size_t)i * 3 + 0] = T[(size_t)i * 3 + 1] = T[(size_t)i * 3 + 2] = 0.0f;
      continue;
    }
    const float dx = d.px[v] - fr[10], dy = d.py[v] - fr[11], dz = d.pz[v] - fr[12];
    const float ih = fr[9] > 0.0f ? 1.0f / fr[9] : 0.0f;
    T[(size_t)i * 3 + 0] = (fr[0] * dx + fr[1] * dy + fr[2] * dz) * ih;
    T[(size_t)i * 3 + 1] = (fr[3] * dx + fr[4] * dy + fr[5] * dz) * ih;
    T[(size_t)i * 3 + 2] = (fr[6] * dx + fr[7] * dy + fr[8] * dz) * ih;
  }
}

// ---------------------------------------------------------------------------
// Inference on a flat batch (validation, tests): the dab path's MLP, 16 rows per wave.
// ---------------------------------------------------------------------------
#define LRN_WAVES 4
#define LRN_THREADS (64 * LRN_WAVES)
#define LRN_ROWS 16
__global__ __launch_bounds__(LRN_THREADS) void k_mlp_predict(SgMlpView m, const float *X, int n, float *Y) {
  __shared__ MlpLds s;
  __shared__ __align__(16) uint16_t bufA[LRN_WAVES][LRN_ROWS * MLP_ST];
  __shared__ __align__(16) uint16_t bufB[LRN_WAVES][LRN_ROWS * MLP_ST];
  __shared__ __align__(16) float sOut[LRN_WAVES][LRN_ROWS * 4];
  mlp_lds_load<LRN_THREADS>(s, m);
  __syncthreads();
  const int w = threadIdx.x / 64, L = __lane_id();
  const int nw = gridDim.x * LRN_WAVES;
  for (int base = (blockIdx.x * LRN_WAVES + w) * LRN_ROWS; base < n; base += nw * LRN_ROWS) {
    const int row = base + L;
    float x[28], y[3] = {0.0f, 0.0f, 0.0f};
    for (int k = 0; k < 28; k++) {
      const float raw = (L < LRN_ROWS && row < n) ? X[(size_t)row * 28 + k] : 0.0f;
      x[k] = (raw - m.norm[k]) * m.norm[28 + k];
    }
    mlp_wave16(s, bufA[w], bufB[w], sOut[w], x, y);
    if (L < LRN_ROWS && row < n) {
      for (int o = 0; o < 3; o++) Y[(size_t)row * 3 + o] = y[o] * m.norm[56 + o];
    }
  }
}

// ---------------------------------------------------------------------------
// Learned dab.  k_brush_learned: work items (gathered leaf, slot), the wave of slot t
// takes the leaf's 16-vertex tiles t, t+4, ...: sphere test and factor as the plain
// Smooth brush, then features -> MLP (MFMA, 16 rows) -> predicted position written to
// the shadow copy; affected vertices get the dab's stamp.  The error estimate
// sum((|d|/r)^2) (2^-32 fixed point), the out-of-range count and the affected count
// are reduced per block and added once per block.
// k_post_learned: accept (error estimate and range checks within tolerance): normals
// and leaf bounds from the shadow copy, commit it to the positions; reject: undo the
// shadow copy and raise the halt flag so the rest of the stroke stops and the host
// runs this dab with the plain brush.  Both walk the gathered leaves' vertex ranges
// and test the stamp (no per-vertex list, so no slot reservations).
// ---------------------------------------------------------------------------
#define LRN_SLOTS 4   // waves per gathered leaf (batch)
#define LRN_DSLOTS 8  // waves per gathered leaf (single mesh): one 16-row tile each for leaves up to 128 uniq vertices

// Evaluate one 16-vertex tile of a learned dab (every lane of the wave calls it; lanes
// < 16 own a vertex).  Returns the Smooth factor of the lane's vertex.
__device__ inline float learned_eval(const SgDevView &d, const SgDabK &k, const float *curve, const SgMlpView &m,
                                     const MlpLds &s, uint16_t *bufA, uint16_t *bufB, float *sOut, int v,
                                     const BrushIn &in, float nx, float ny, float nz, long long &esum, int &nood) {
  float fade = 0.0f, feat[28], R[9], h = 0.0f;
  bool ok = true;
  if (in.aff) {
    fade = k.bstr * (sg_falloff(k, curve, sqrtf(in.d2)) * (1.0f - in.mk));
    ok = learned_features(d, in.vv, in.x, in.y, in.z, nx, ny, nz, fade, feat, R, h);
  } else {
    for (int q = 0; q < 28; q++) feat[q] = 0.0f;
  }
  float xn[28], y[3] = {0.0f, 0.0f, 0.0f};
  for (int q = 0; q < 28; q++) xn[q] = in.aff ? (feat[q] - m.norm[q]) * m.norm[28 + q] : 0.0f;
  mlp_wave16(s, bufA, bufB, sOut, xn, y);
  if (in.aff) {
    const float o0 = y[0] * m.norm[56] * h, o1 = y[1] * m.norm[57] * h, o2 = y[2] * m.norm[58] * h;
    const float dx = R[0] * o0 + R[3] * o1 + R[6] * o2;
    const float dy = R[1] * o0 + R[4] * o1 + R[7] * o2;
    const float dz = R[2] * o0 + R[5] * o1 + R[8] * o2;
    d.qx[v] = in.x + dx;
    d.qy[v] = in.y + dy;
    d.qz[v] = in.z + dz;
    const float rel2 = (dx * dx + dy * dy + dz * dz) / k.r2;
    if (!ok || !(rel2 < 1.0f)) {  // out of range, NaN or a displacement beyond r
      nood++;
    } else {
      esum += __double2ll_rn((double)rel2 * 4294967296.0);  // 2^32: rel2 < 1, no overflow below 2^31 vertices
    }
  }
  return fade;
}

// Loads of one tile: lanes < 16 take vertices base + lane (< end), the rest none.
__device__ __forceinline__ BrushIn learned_tile_load(const SgDevView &d, const SgDabK &k, int base, int end, int lane,
                                                     float &nx, float &ny, float &nz) {
  const int v = base + lane;
  const int vend = min(end, base + LRN_ROWS);
  const BrushIn in = brush_load(d, k, v, vend, true);
  nx = ny = nz = 0.0f;
  if (v < vend) {
    nx = d.nx[v];
    ny = d.ny[v];
    nz = d.nz[v];
  }
  return in;
}

__global__ __launch_bounds__(LRN_THREADS) void k_brush_learned(SgDevView d, SgDabK k, const float *curve,
                                                               SgMlpView m, const float *place) {
  __shared__ MlpLds s;
  __shared__ __align__(16) uint16_t bufA[LRN_WAVES][LRN_ROWS * MLP_ST];
  __shared__ __align__(16) uint16_t bufB[LRN_WAVES][LRN_ROWS * MLP_ST];
  __shared__ __align__(16) float sOut[LRN_WAVES][LRN_ROWS * 4];
  __shared__ unsigned long long s_acc[3];
  if (d.halt && d.halt[0]) return;
  sg_place(k, place);
  const int w = threadIdx.x / 64, lane = __lane_id();
  const int gw = blockIdx.x * LRN_WAVES + w;
  const int nw = gridDim.x * LRN_WAVES;
  // the wave's first leaf range, loaded beside the counter (speculatively: any index
  // below L is inside the gathered-leaf list; the loop uses it only when it is live)
  const int2 vr0 = gw / LRN_DSLOTS < d.L ? d.gl_v[gw / LRN_DSLOTS] : make_int2(0, 0);
  const long long ng_ctr = d.ctr[SG_CTR_NGATHER];
  const int nitems = LRN_DSLOTS * (int)ng_ctr;
  if (blockIdx.x * LRN_WAVES >= nitems) return;  // whole block idle: skip the weight fill
  // the first tile's loads go out before the weight fill, so both share one round trip
  const int base0 = vr0.x + (gw % LRN_DSLOTS) * LRN_ROWS;
  float nx0 = 0.0f, ny0 = 0.0f, nz0 = 0.0f;
  BrushIn in0{};
  if (gw < nitems && base0 < vr0.y) in0 = learned_tile_load(d, k, base0, vr0.y, lane, nx0, ny0, nz0);
  mlp_lds_load<LRN_THREADS>(s, m);
  if (threadIdx.x < 3) s_acc[threadIdx.x] = 0ull;
  __syncthreads();
  const int stamp = d.seq[0];
  long long esum = 0;
  int nood = 0, naff = 0;
  if (ng_ctr & SG_NGATHER_UNDO) {
    // an undo step is open: the learned brush writes the shadow copy and leaves the
    // positions alone (k_post_learned copies or redoes them), so each leaf is saved
    // from the positions by the wave of its first slot, whatever the other slots write
    for (int it = gw; it < nitems; it += nw)
      if (it % LRN_DSLOTS == 0)
        undo_save_leaf(d, d.gathered[it / LRN_DSLOTS], d.gl_v[it / LRN_DSLOTS], d.px, d.py, d.pz);
  }
  for (int it = gw; it < nitems; it += nw) {
    int2 vr = vr0;
    if (it != gw) vr = d.gl_v[it / LRN_DSLOTS];
    for (int base = vr.x + (it % LRN_DSLOTS) * LRN_ROWS; base < vr.y; base += LRN_DSLOTS * LRN_ROWS) {
      float nx = nx0, ny = ny0, nz = nz0;
      BrushIn in = in0;
      if (base != base0 || it != gw) in = learned_tile_load(d, k, base, vr.y, lane, nx, ny, nz);
      if (__ballot(in.aff) == 0ull) continue;
      const int v = base + lane;
      learned_eval(d, k, curve, m, s, bufA[w], bufB[w], sOut[w], v, in, nx, ny, nz, esum, nood);
      if (in.aff) {
        d.vstamp[v] = stamp;
        naff++;
      }
    }
  }
  esum = wave_sum_ll(esum);
  nood = wave_sum_i(nood);
  naff = wave_sum_i(naff);
  if (lane == 0 && naff) {
    atomicAdd(&s_acc[0], (unsigned long long)esum);
    atomicAdd(&s_acc[1], (unsigned long long)nood);
    atomicAdd(&s_acc[2], (unsigned long long)naff);
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_acc[2]) {
    atomicAdd((unsigned long long *)&d.ctr[SG_CTR_LERR], s_acc[0]);
    atomicAdd((unsigned long long *)&d.ctr[SG_CTR_LOOD], s_acc[1]);
    atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NAFF], s_acc[2]);
  }
}

// Batched learned brush: items (gathered leaf of a learned-Smooth mesh, slot) of every
// mesh; per-mesh error / range / affected counts feed the per-mesh decision.
__global__ __launch_bounds__(LRN_THREADS) void k_bbrush_learned(SgBatchView b, SgMlpView m) {
  __shared__ MlpLds s;
  __shared__ __align__(16) uint16_t bufA[LRN_WAVES][LRN_ROWS * MLP_ST];
  __shared__ __align__(16) uint16_t bufB[LRN_WAVES][LRN_ROWS * MLP_ST];
  __shared__ __align__(16) float sOut[LRN_WAVES][LRN_ROWS * 4];
  const SgDevView &d = b.d;
  const int nitems = LRN_SLOTS * (int)b.ctr[SB_NGL];  // leaves of learned meshes only
  if (blockIdx.x * LRN_WAVES >= nitems) return;
  mlp_lds_load<LRN_THREADS>(s, m);
  __syncthreads();
  const int w = threadIdx.x / 64, lane = __lane_id();
  const int gw = blockIdx.x * LRN_WAVES + w;
  const int nw = gridDim.x * LRN_WAVES;
  const int serial = b.step[0], row = b.step[1];
  for (int it = gw; it < nitems; it += nw) {
    const int g = b.gl_lrn[it / LRN_SLOTS];
    const int mesh = __builtin_amdgcn_readfirstlane(b.gl_mesh[g]);
    const SgDabK k = b.tab[(size_t)row * b.M + mesh];
    const int2 vr = d.gl_v[g];
    long long esum = 0, degt = 0;
    int nood = 0, naff = 0;
    for (int base = vr.x + (it % LRN_SLOTS) * LRN_ROWS; base < vr.y; base += LRN_SLOTS * LRN_ROWS) {
      float nx, ny, nz;
      const BrushIn in = learned_tile_load(d, k, base, vr.y, lane, nx, ny, nz);
      if (__ballot(in.aff) == 0ull) continue;
      const int v = base + lane;
      const float fade = learned_eval(d, k, nullptr, m, s, bufA[w], bufB[w], sOut[w], v, in, nx, ny, nz, esum, nood);
      if (in.aff) {
        b.vstamp[v] = serial;
        b.vfade[v] = fade;
        naff++;
        degt += in.vt.y - in.vt.x;
      }
    }
    esum = wave_sum_ll(esum);
    nood = wave_sum_i(nood);
    naff = wave_sum_i(naff);
    degt = wave_sum_ll(degt);
    if (lane == 0) {
      long long *mc = b.ctr + SB_COUNT + SB_PM * (size_t)mesh;
      if (it % LRN_SLOTS == 0) atomicAdd((unsigned long long *)&mc[SBM_CAND], (unsigned long long)(vr.y - vr.x));
      if (naff) {
        atomicAdd((unsigned long long *)&mc[SBM_LERR], (unsigned long long)esum);
        atomicAdd((unsigned long long *)&mc[SBM_LOOD], (unsigned long long)nood);
        atomicAdd((unsigned long long *)&mc[SBM_LNAFF], (unsigned long long)naff);
        atomicAdd((unsigned long long *)&mc[SBM_NAFF], (unsigned long long)naff);
        atomicAdd((unsigned long long *)&mc[SBM_DEGT], (unsigned long long)degt);
      }
    }
  }
}

// Batched decision, per stamped vertex of a learned-Smooth mesh (wave per gathered
// leaf): accepted -> the prediction is committed to the positions; rejected -> the
// shadow copy gets the plain brush's first Jacobi step from the untouched positions
// (steps 2..k follow in the plain kernels).
__global__ __launch_bounds__(256) void k_bdecide(SgBatchView b) {
  const SgDevView &d = b.d;
  const int lane = __lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  const int serial = b.step[0], row = b.step[1];
  const int ng = (int)b.ctr[SB_NGATHER];
  for (int g = gw; g < ng; g += nw) {
    const int mesh = __builtin_amdgcn_readfirstlane(b.gl_mesh[g]);
    if (!b.tab[(size_t)row * b.M + mesh].learned) continue;
    const bool rejected = sb_learned_rejected(b, mesh);
    const int2 vr = d.gl_v[g];
    for (int v = vr.x + lane; v < vr.y; v += SG_WAVE) {
      if (b.vstamp[v] != serial) continue;
      if (!rejected) {
        d.px[v] = d.qx[v];
        d.py[v] = d.qy[v];
        d.pz[v] = d.qz[v];
        continue;
      }
      const int2 vv = make_int2(d.vv_off[v], d.vv_off[v + 1]);
      const float fade = b.vfade[v];
      const float x = d.px[v], y = d.py[v], z = d.pz[v];
      float sx, sy, sz;
      nbr_sum(d.vv_idx, vv.x, vv.y, d.px, d.py, d.pz, sx, sy, sz);
      float rx = x, ry = y, rz = z;
      if (vv.y > vv.x) {
        const float inv = 1.0f / (float)(vv.y - vv.x);
        const float avx = sx * inv, avy = sy * inv, avz = sz * inv;
        rx = x + (avx - x) * fade;
        ry = y + (avy - y) * fade;
        rz = z + (avz - z) * fade;
      }
      d.qx[v] = rx;
      d.qy[v] = ry;
      d.qz[v] = rz;
    }
  }
}

// Items (gathered leaf, part): part 0 = the leaf's stamped vertices (accept: normals
// from the shadow copy, then commit; reject: restore the shadow copy), part 1 = the
// leaf's bounds from the shadow copy (accept only).
__global__ __launch_bounds__(256) void k_post_learned(SgDevView d, SgDevView dq, float tol, float alpha,
                                                      int serial) {
  if (d.halt && d.halt[0]) return;
  const int lane = __lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  // Work items: three per gathered leaf, vertices [0, 64) and [64, 128) of its uniq
  // range (then every 128 on), and its box refit, so a wave runs one 64-vertex chain.
  // The first item's leaf range, and then its vertices' stamps and incidence ranges,
  // are loaded beside the counters (speculatively: any index below L is inside the
  // gathered-leaf list and any vertex of a listed range inside the mesh).
  const int g0 = gw / 3, part0 = gw % 3;
  const int2 vr0 = g0 < d.L ? d.gl_v[g0] : make_int2(0, 0);
  const int stamp = d.seq[0];
  const int na = (int)d.ctr[SG_CTR_NAFF];
  const int ng = (int)d.ctr[SG_CTR_NGATHER];
  const long long es = d.ctr[SG_CTR_LERR];
  const long long nood = d.ctr[SG_CTR_LOOD];
  const int v0 = vr0.x + part0 * SG_WAVE + lane;
  const bool pre = part0 < 2 && v0 < vr0.y;
  const int st0 = pre ? d.vstamp[v0] : 0;
  int2 vt0 = make_int2(0, 0);
  if (pre) vt0 = make_int2(d.vt_off[v0], d.vt_off[v0 + 1]);
  // error estimate: alpha * RMS(|d| / r), alpha = held-out RMS error / RMS displacement
  const double rms = na > 0 ? sqrt((double)es / 4294967296.0 / (double)na) : 0.0;
  const bool accept = nood == 0 && (double)alpha * rms <= (double)tol;
  long long a = 0, b = 0, c = 0, nsh = 0;
  for (int it = gw; it < 3 * ng; it += nw) {
    const int g = it / 3, part = it % 3;
    if (part == 2) {
      if (accept) leaf_refit(dq, d.gathered[g], d.gl_v[g], d.gl_s[g], nsh);
      continue;
    }
    int2 vr = vr0;
    if (it != gw) vr = d.gl_v[g];
    for (int v = vr.x + part * SG_WAVE + lane; v < vr.y; v += 2 * SG_WAVE) {
      const bool first = it == gw && v == v0;
      int st = st0;
      if (!first) st = d.vstamp[v];
      if (st != stamp) continue;
      if (accept) {
        int2 vt = vt0;
        if (!first) vt = make_int2(d.vt_off[v], d.vt_off[v + 1]);
        vertex_normal(dq, v, vt.x, vt.y, 0, 0, a, b, c);  // corners from the shadow copy
        d.px[v] = d.qx[v];
        d.py[v] = d.qy[v];
        d.pz[v] = d.qz[v];
      } else {
        d.qx[v] = d.px[v];
        d.qy[v] = d.py[v];
        d.qz[v] = d.pz[v];
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (accept)
      d.ctr[SG_CTR_LACC] = 1;
    else if (d.halt)
      d.halt[0] = serial + 1;
  }
}

// ---------------------------------------------------------------------------
// Training (FP32).  One workgroup = TB samples: forward, loss, backward with LDS
// activations; per-workgroup gradient written to its own slot, then reduced in a
// fixed order and applied with Adam (deterministic).
// ---------------------------------------------------------------------------
#define TB 32
__global__ __launch_bounds__(256) void k_mlp_train(const float *w, const float *X, const float *T, const int *idx,
                                                  int n_valid, int base, float inv_count, const float *norm,
                                                  float *grad_part, float *loss_part) {
  __shared__ float sX[TB][28], sZ1[TB][64], sA1[TB][64], sZ2[TB][64], sA2[TB][64], sD[TB][4];
  __shared__ float sLoss[256];
  const int tid = threadIdx.x;
  const int s0 = base + blockIdx.x * TB;
  for (int q = tid; q < TB * 28; q += 256) {
    const int s = q / 28, i = q % 28;
    const int gs = s0 + s;
    const float raw = gs < n_valid ? X[(size_t)idx[gs] * 28 + i] : 0.0f;
    sX[s][i] = (raw - norm[i]) * norm[28 + i];
  }
  __syncthreads();
  const float *W1 = w + SG_OFF_W1, *B1 = w + SG_OFF_B1, *W2 = w + SG_OFF_W2, *B2 = w + SG_OFF_B2;
  const float *W3 = w + SG_OFF_W3, *B3 = w + SG_OFF_B3;
  for (int o = tid; o < TB * 64; o += 256) {
    const int s = o / 64, j = o % 64;
    float z = B1[j];
    for (int i = 0; i < 28; i++) z += sX[s][i] * W1[i * 64 + j];
    sZ1[s][j] = z;
    sA1[s][j] = sg_silu(z);
  }
  __syncthreads();
  for (int o = tid; o < TB * 64; o += 256) {
    const int s = o / 64, j = o % 64;
    float z = B2[j];
    for (int i = 0; i < 64; i++) z += sA1[s][i] * W2[i * 64 + j];
    sZ2[s][j] = z;
    sA2[s][j] = sg_silu(z);
  }
  __syncthreads();
  float lsum = 0.0f;
  if (tid < TB * 3) {
    const int s = tid / 3, o = tid % 3;
    const int gs = s0 + s;
    float yv = B3[o];
    for (int j = 0; j < 64; j++) yv += sA2[s][j] * W3[j * 3 + o];
    float dv = 0.0f;
    if (gs < n_valid) {
      const float t = T[(size_t)idx[gs] * 3 + o] / norm[56 + o];
      const float e = yv - t;
      lsum = e * e;
      dv = 2.0f * e * inv_count;
    }
    sD[s][o] = dv;
  }
  sLoss[tid] = lsum;
  __syncthreads();
  float *G = grad_part + (size_t)blockIdx.x * SG_MLP_NPARAM;
  // dW3, db3
  for (int q = tid; q < 64 * 3; q += 256) {
    const int j = q / 3, o = q % 3;
    float g = 0.0f;
    for (int s = 0; s < TB; s++) g += sA2[s][j] * sD[s][o];
    G[SG_OFF_W3 + q] = g;
  }
  if (tid < 3) {
    float g = 0.0f;
    for (int s = 0; s < TB; s++) g += sD[s][tid];
    G[SG_OFF_B3 + tid] = g;
  }
  __syncthreads();
  // dZ2 in place
  for (int q = tid; q < TB * 64; q += 256) {
    const int s = q / 64, j = q % 64;
    const float da = sD[s][0] * W3[j * 3 + 0] + sD[s][1] * W3[j * 3 + 1] + sD[s][2] * W3[j * 3 + 2];
    sZ2[s][j] = da * sg_silu_grad(sZ2[s][j]);
  }
  __syncthreads();
  for (int q = tid; q < 64 * 64; q += 256) {
    const int i = q / 64, j = q % 64;
    float g = 0.0f;
    for (int s = 0; s < TB; s++) g += sA1[s][i] * sZ2[s][j];
    G[SG_OFF_W2 + q] = g;
  }
  if (tid < 64) {
    float g = 0.0f;
    for (int s = 0; s < TB; s++) g += sZ2[s][tid];
    G[SG_OFF_B2 + tid] = g;
  }
  __syncthreads();  // dW2 above still reads A1
  // dZ1 in place
  for (int q = tid; q < TB * 64; q += 256) {
    const int s = q / 64, i = q % 64;
    float da = 0.0f;
    for (int j = 0; j < 64; j++) da += sZ2[s][j] * W2[i * 64 + j];
    sA1[s][i] = da * sg_silu_grad(sZ1[s][i]);  // (A1 no longer needed)
  }
  __syncthreads();
  for (int q = tid; q < 28 * 64; q += 256) {
    const int i = q / 64, j = q % 64;
    float g = 0.0f;
    for (int s = 0; s < TB; s++) g += sX[s][i] * sA1[s][j];
    G[SG_OFF_W1 + q] = g;
  }
  if (tid < 64) {
    float g = 0.0f;
    for (int s = 0; s < TB; s++) g += sA1[s][tid];
    G[SG_OFF_B1 + tid] = g;
  }
  // per-workgroup loss
  for (int off = 128; off > 0; off >>= 1) {
    __syncthreads();
    if (tid < off) sLoss[tid] += sLoss[tid + off];
  }
  if (tid == 0) loss_part[blockIdx.x] = sLoss[0];
}

__global__ __launch_bounds__(256) void k_mlp_adam(float *w, float *mo, float *ve, const float *grad_part, int nwg,
                                                 float lr, float b1, float b2, float eps, float c1, float c2,
                                                 float *grad_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= SG_MLP_NPARAM) return;
  float g = 0.0f;
  for (int i = 0; i < nwg; i++) g += grad_part[(size_t)i * SG_MLP_NPARAM + p];
  if (grad_out) {
    grad_out[p] = g;
    return;  // gradient only (checks)
  }
  const float m = b1 * mo[p] + (1.0f - b1) * g;
  const float v = b2 * ve[p] + (1.0f - b2) * g * g;
  mo[p] = m;
  ve[p] = v;
  w[p] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
SgMlpView sg_mlp_view(SgMlp *m) {
  SgMlpView v;
  v.wb = m->wb;
  v.bias = m->bias;
  v.norm = m->norm;
  return v;
}

static int mlp_export(SgMlp *m) {
  std::vector<float> w(SG_MLP_NPARAM);
  if (sg_check_hip(hipMemcpy(w.data(), m->w, SG_MLP_NPARAM * 4, hipMemcpyDeviceToHost), "copy")) return -1;
  std::vector<uint16_t> wb(64 * 32 + 64 * 64 + 16 * 64, 0);
  for (int n = 0; n < 64; n++)
    for (int k = 0; k < 28; k++) wb[n * 32 + k] = sg_f2bf(w[SG_OFF_W1 + k * 64 + n]);
  for (int n = 0; n < 64; n++)
    for (int k = 0; k < 64; k++) wb[64 * 32 + n * 64 + k] = sg_f2bf(w[SG_OFF_W2 + k * 64 + n]);
  for (int n = 0; n < 3; n++)
    for (int k = 0; k < 64; k++) wb[64 * 32 + 64 * 64 + n * 64 + k] = sg_f2bf(w[SG_OFF_W3 + k * 3 + n]);
  std::vector<float> bias(144, 0.0f);
  for (int j = 0; j < 64; j++) {
    bias[j] = w[SG_OFF_B1 + j];
    bias[64 + j] = w[SG_OFF_B2 + j];
  }
  for (int o = 0; o < 3; o++) bias[128 + o] = w[SG_OFF_B3 + o];
  if (sg_check_hip(hipMemcpy(m->wb, wb.data(), wb.size() * 2, hipMemcpyHostToDevice), "copy")) return -1;
  if (sg_check_hip(hipMemcpy(m->bias, bias.data(), bias.size() * 4, hipMemcpyHostToDevice), "copy")) return -1;
  return sg_check_hip(hipMemcpy(m->norm, m->host_norm, sizeof(m->host_norm), hipMemcpyHostToDevice), "copy");
}

extern "C" {

SgMlp *sg_mlp_new(uint64_t seed) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    sg_set_error("no HIP device");
    return nullptr;
  }
  SgMlp *m = new SgMlp();
  int rc = sg_check_hip(hipMalloc(&m->w, SG_MLP_NPARAM * 4), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&m->mo, SG_MLP_NPARAM * 4), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&m->ve, SG_MLP_NPARAM * 4), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&m->wb, (64 * 32 + 64 * 64 + 16 * 64) * 2), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&m->bias, 144 * 4), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&m->norm, 64 * 4), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream");
  if (rc) {
    sg_mlp_free(m);
    return nullptr;
  }
  hipMemset(m->mo, 0, SG_MLP_NPARAM * 4);
  hipMemset(m->ve, 0, SG_MLP_NPARAM * 4);
  // He-style init from a seeded generator
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  std::vector<float> w(SG_MLP_NPARAM, 0.0f);
  for (int i = 0; i < 28 * 64; i++) w[SG_OFF_W1 + i] = nd(rng) * sqrtf(2.0f / 28.0f);
  for (int i = 0; i < 64 * 64; i++) w[SG_OFF_W2 + i] = nd(rng) * sqrtf(2.0f / 64.0f);
  for (int i = 0; i < 64 * 3; i++) w[SG_OFF_W3 + i] = nd(rng) * sqrtf(1.0f / 64.0f);
  for (int i = 0; i < 64; i++) m->host_norm[i] = 0.0f;
  for (int i = 0; i < 28; i++) m->host_norm[28 + i] = 1.0f;
  for (int o = 0; o < 3; o++) m->host_norm[56 + o] = 1.0f;
  if (sg_check_hip(hipMemcpy(m->w, w.data(), SG_MLP_NPARAM * 4, hipMemcpyHostToDevice), "copy") || mlp_export(m)) {
    sg_mlp_free(m);
    return nullptr;
  }
  return m;
}

void sg_mlp_free(SgMlp *m) {
  if (!m) return;
  if (m->stream) hipStreamSynchronize(m->stream);
  void *ptrs[] = {m->w, m->mo, m->ve, m->wb, m->bias, m->norm, m->grad_part, m->loss_part, m->dX, m->dT, m->perm};
  for (void *q : ptrs)
    if (q) hipFree(q);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

int sg_mlp_get_params(SgMlp *m, float *out) {
  return sg_check_hip(hipMemcpy(out, m->w, SG_MLP_NPARAM * 4, hipMemcpyDeviceToHost), "copy");
}

int sg_mlp_set_params(SgMlp *m, const float *w) {
  if (sg_check_hip(hipMemcpy(m->w, w, SG_MLP_NPARAM * 4, hipMemcpyHostToDevice), "copy")) return -1;
  return mlp_export(m);
}

// in_mu[28], in_inv_sigma[28], out_scale[3]
int sg_mlp_set_norm(SgMlp *m, const float *in_mu, const float *in_isig, const float *out_scale) {
  for (int i = 0; i < 28; i++) {
    m->host_norm[i] = in_mu[i];
    m->host_norm[28 + i] = in_isig[i];
  }
  for (int o = 0; o < 3; o++) m->host_norm[56 + o] = out_scale[o];
  return mlp_export(m);
}

int sg_mlp_get_norm(SgMlp *m, float *out64) {
  memcpy(out64, m->host_norm, sizeof(m->host_norm));
  return 0;
}

int sg_mlp_set_decision(SgMlp *m, float alpha, float tol) {
  m->alpha = alpha;
  m->tol = tol;
  return 0;
}

int sg_mlp_get_decision(SgMlp *m, float *out2) {
  out2[0] = m->alpha;
  out2[1] = m->tol;
  return 0;
}

// Training data lives on the device for the whole run.
int sg_mlp_set_data(SgMlp *m, const float *X, const float *T, int n) {
  if (!m || n <= 0) return sg_set_error("bad arguments");
  if (m->dX) hipFree(m->dX);
  if (m->dT) hipFree(m->dT);
  if (m->perm) hipFree(m->perm);
  if (sg_check_hip(hipMalloc(&m->dX, (size_t)n * 28 * 4), "hipMalloc")) return -1;
  if (sg_check_hip(hipMalloc(&m->dT, (size_t)n * 3 * 4), "hipMalloc")) return -1;
  if (sg_check_hip(hipMalloc(&m->perm, (size_t)n * 4), "hipMalloc")) return -1;
  if (sg_check_hip(hipMemcpy(m->dX, X, (size_t)n * 28 * 4, hipMemcpyHostToDevice), "copy")) return -1;
  if (sg_check_hip(hipMemcpy(m->dT, T, (size_t)n * 3 * 4, hipMemcpyHostToDevice), "copy")) return -1;
  m->n = n;
  return 0;
}

static int ensure_parts(SgMlp *m, int nwg) {
  if (nwg <= m->part_cap) return 0;
  if (m->grad_part) hipFree(m->grad_part);
  if (m->loss_part) hipFree(m->loss_part);
  if (sg_check_hip(hipMalloc(&m->grad_part, (size_t)nwg * SG_MLP_NPARAM * 4), "hipMalloc")) return -1;
  if (sg_check_hip(hipMalloc(&m->loss_part, (size_t)nwg * 4), "hipMalloc")) return -1;
  m->part_cap = nwg;
  return 0;
}

// Gradient of the mean squared (normalised) error over rows [0, n) of the data in
// index order (for checks against the CPU oracle).  Writes 6,211 floats.
int sg_mlp_gradient(SgMlp *m, int n, float *grad_out, float *loss_out) {
  if (!m || !m->dX || n <= 0 || n > m->n) return sg_set_error("bad arguments");
  const int nwg = (n + TB - 1) / TB;
  if (ensure_parts(m, nwg)) return -1;
  std::vector<int> id(m->n);
  for (int i = 0; i < m->n; i++) id[i] = i;
  if (sg_check_hip(hipMemcpy(m->perm, id.data(), (size_t)m->n * 4, hipMemcpyHostToDevice), "copy")) return -1;
  float *dg = nullptr;
  if (sg_check_hip(hipMalloc(&dg, SG_MLP_NPARAM * 4), "hipMalloc")) return -1;
  hipLaunchKernelGGL(k_mlp_train, dim3(nwg), dim3(256), 0, m->stream, m->w, m->dX, m->dT, m->perm, n, 0,
                     1.0f / (float)(n * 3), m->norm, m->grad_part, m->loss_part);
  hipLaunchKernelGGL(k_mlp_adam, dim3((SG_MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, m->w, m->mo, m->ve,
                     m->grad_part, nwg, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, dg);
  hipMemcpyAsync(grad_out, dg, SG_MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->stream);
  std::vector<float> lp(nwg);
  hipMemcpyAsync(lp.data(), m->loss_part, (size_t)nwg * 4, hipMemcpyDeviceToHost, m->stream);
  int rc = sg_check_hip(hipStreamSynchronize(m->stream), "gradient");
  hipFree(dg);
  if (loss_out) {
    double s = 0.0;
    for (float x : lp) s += x;
    *loss_out = (float)(s / (double)(n * 3));
  }
  return rc;
}

// Adam over shuffled minibatches of batch_wg*32 rows for `epochs` passes; writes the
// mean training loss of each epoch to loss_curve (may be NULL).
int sg_mlp_train(SgMlp *m, int epochs, int batch_wg, float lr, uint64_t seed, float *loss_curve) {
  if (!m || !m->dX || epochs <= 0 || batch_wg <= 0) return sg_set_error("bad arguments");
  if (ensure_parts(m, batch_wg)) return -1;
  std::mt19937_64 rng(seed);
  std::vector<int> id(m->n);
  std::vector<float> lp(batch_wg);
  const int batch = batch_wg * TB;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int e = 0; e < epochs; e++) {
    for (int i = 0; i < m->n; i++) id[i] = i;
    std::shuffle(id.begin(), id.end(), rng);
    hipMemcpyAsync(m->perm, id.data(), (size_t)m->n * 4, hipMemcpyHostToDevice, m->stream);
    double lsum = 0.0;
    // cosine-free schedule: constant, then /10 for the last quarter of the epochs
    const float lr_e = (e >= (3 * epochs) / 4) ? lr * 0.1f : lr;
    for (int base = 0; base + batch <= m->n; base += batch) {
      m->step++;
      const float c1 = 1.0f / (1.0f - powf(b1, (float)m->step));
      const float c2 = 1.0f / (1.0f - powf(b2, (float)m->step));
      hipLaunchKernelGGL(k_mlp_train, dim3(batch_wg), dim3(256), 0, m->stream, m->w, m->dX, m->dT, m->perm,
                         m->n, base, 1.0f / (float)(batch * 3), m->norm, m->grad_part, m->loss_part);
      hipLaunchKernelGGL(k_mlp_adam, dim3((SG_MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, m->w, m->mo,
                         m->ve, m->grad_part, batch_wg, lr_e, b1, b2, eps, c1, c2, (float *)nullptr);
      if (loss_curve && base == 0) {
        hipMemcpyAsync(lp.data(), m->loss_part, (size_t)batch_wg * 4, hipMemcpyDeviceToHost, m->stream);
        hipStreamSynchronize(m->stream);
        double s = 0.0;
        for (float x : lp) s += x;
        lsum = s / (double)(batch * 3);
      }
    }
    if (loss_curve) loss_curve[e] = (float)lsum;
  }
  if (sg_check_hip(hipStreamSynchronize(m->stream), "train")) return -1;
  return mlp_export(m);
}

// Inference through the MFMA path on raw features X[n][28] -> displacement Y[n][3]
// (local frame / h).
int sg_mlp_predict(SgMlp *m, const float *X, int n, float *Y) {
  if (!m || n <= 0) return sg_set_error("bad arguments");
  float *dX = nullptr, *dY = nullptr;
  if (sg_check_hip(hipMalloc(&dX, (size_t)n * 28 * 4), "hipMalloc")) return -1;
  if (sg_check_hip(hipMalloc(&dY, (size_t)n * 3 * 4), "hipMalloc")) {
    hipFree(dX);
    return -1;
  }
  hipMemcpyAsync(dX, X, (size_t)n * 28 * 4, hipMemcpyHostToDevice, m->stream);
  const int blocks = (n + LRN_ROWS * LRN_WAVES - 1) / (LRN_ROWS * LRN_WAVES);
  hipLaunchKernelGGL(k_mlp_predict, dim3(blocks < 4096 ? blocks : 4096), dim3(LRN_THREADS), 0, m->stream,
                     sg_mlp_view(m), dX, n, dY);
  hipMemcpyAsync(Y, dY, (size_t)n * 3 * 4, hipMemcpyDeviceToHost, m->stream);
  int rc = sg_check_hip(hipStreamSynchronize(m->stream), "predict");
  hipFree(dX);
  hipFree(dY);
  return rc;
}

int sg_mfma_probe(const uint16_t *A, const uint16_t *B, float *D) {
  uint16_t *dA = nullptr, *dB = nullptr;
  float *dD = nullptr;
  int rc = sg_check_hip(hipMalloc(&dA, 512), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&dB, 512), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&dD, 1024), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMemcpy(dA, A, 512, hipMemcpyHostToDevice), "copy");
  if (!rc) rc = sg_check_hip(hipMemcpy(dB, B, 512, hipMemcpyHostToDevice), "copy");
  if (!rc) {
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, dA, dB, dD);
    rc = sg_check_hip(hipMemcpy(D, dD, 1024, hipMemcpyDeviceToHost), "probe");
  }
  hipFree(dA);
  hipFree(dB);
  hipFree(dD);
  return rc;
}

}  // extern "C"

// ---- launch helpers used by the dab path (api.cpp / kernels.hip) ----
int sg_enqueue_learned_brush(PBVH *p, const SgDevView &v, const SgDabK &k, const float *curve, SgMlp *m,
                             const float *place) {
  // at least 512 workgroups (2048 waves): a wave per (leaf, slot) for up to 256 gathered
  // leaves, so no wave runs two tiles' load chains back to back
  const int G = std::max(512, p->grid_brush > 0 ? p->grid_brush : p->grid_wg);
  hipLaunchKernelGGL(k_brush_learned, dim3(G), dim3(LRN_THREADS), 0, p->stream, v, k, curve, sg_mlp_view(m),
                     place);
  return sg_check_hip(hipGetLastError(), "learned brush launch");
}

int sg_enqueue_learned_post(PBVH *p, const SgDevView &v, SgMlp *m, int serial) {
  SgDevView q = v;
  q.px = v.qx;
  q.py = v.qy;
  q.pz = v.qz;
  const int G = p->grid_post > 0 ? p->grid_post : p->grid_wg;
  hipLaunchKernelGGL(k_post_learned, dim3(G), dim3(256), 0, p->stream, v, q, m->tol, m->alpha, serial);
  return sg_check_hip(hipGetLastError(), "learned post launch");
}

// Squared distances on the affected vertices of the last dab: learned result vs the
// plain result (acc[0]) and the plain result vs the positions before the dab (acc[1]).
__global__ __launch_bounds__(256) void k_judge(SgDevView d, int na, const float *px, const float *py,
                                              const float *pz, const float *lrn, int V, const float *bx,
                                              const float *by, const float *bz, double *acc) {
  double e = 0.0, m = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
    const int v = aff_slot_vertex(d, i, d.aff_cnt[i >> d.aff_shift]);
    if (v < 0) continue;
    const double x = px[v], y = py[v], z = pz[v];
    const double ex = lrn[v] - x, ey = lrn[V + v] - y, ez = lrn[2 * V + v] - z;
    const double mx = x - bx[v], my = y - by[v], mz = z - bz[v];
    e += ex * ex + ey * ey + ez * ez;
    m += mx * mx + my * my + mz * mz;
  }
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    m += __shfl_xor(m, off);
  }
  if (__lane_id() == 0) {
    atomicAdd(&acc[0], e);
    atomicAdd(&acc[1], m);
  }
}

int sg_enqueue_judge(PBVH *p, int na, const float *lrn, const float *bx, const float *by, const float *bz,
                     double *acc) {
  hipMemsetAsync(acc, 0, 16, p->stream);
  const int blocks = std::min((na + 255) / 256, 1024);
  hipLaunchKernelGGL(k_judge, dim3(blocks), dim3(256), 0, p->stream, sg_view(p), na, p->d.px, p->d.py, p->d.pz, lrn,
                     p->d.V, bx, by, bz, acc);
  return sg_check_hip(hipGetLastError(), "judge launch");
}

int sg_batch_enqueue_learned_brush(SgBatch *b, const SgBatchView &v) {
  hipLaunchKernelGGL(k_bbrush_learned, dim3(b->g_learned), dim3(LRN_THREADS), 0, b->stream, v,
                     sg_mlp_view(b->mlp));
  return sg_check_hip(hipGetLastError(), "batch learned brush launch");
}

int sg_batch_enqueue_decide(SgBatch *b, const SgBatchView &v) {
  hipLaunchKernelGGL(k_bdecide, dim3(b->g_smooth), dim3(256), 0, b->stream, v);
  return sg_check_hip(hipGetLastError(), "batch decide launch");
}

int sg_enqueue_collect(PBVH *p, const SgDevView &v, float *X, float *frame, int cap) {
  hipLaunchKernelGGL(k_learned_feat, dim3(p->grid_wg), dim3(256), 0, p->stream, v, X, frame, cap);
  return sg_check_hip(hipGetLastError(), "collect launch");
}

int sg_enqueue_target(PBVH *p, const SgDevView &v, const float *frame, float *T, int n) {
  hipLaunchKernelGGL(k_learned_target, dim3(p->grid_wg), dim3(256), 0, p->stream, v, frame, T, n);
  return sg_check_hip(hipGetLastError(), "target launch");
}
