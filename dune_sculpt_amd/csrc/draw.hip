// Draw buffers on the GPU (gfx950, wave64): the vertex data the reference's draw code
// builds per PBVH leaf (GPU_pbvh_mesh_buffers_build / gpu_pbvh_mesh_buffers_update,
// gpu/intern/gpu_buffers.c:174-306, :308-350, driven by pbvh_update_draw_buffers,
// kernel/intern/pbvh.c:1251-1367), written straight into one device arena that a
// renderer on the same GPU can read in place instead of receiving a host upload.
//
// Layout per vertex (VertexFormat_pack of gpu_pbvh_init's format,
// gpu/intern/gpu_vertex_format.cc:60-70, :281-329): pos F32x3 @0, nor I16x3 @16,
// msk U8 @22, ac U16x4 @24, fset U8x3 @32, stride 36 (padding and `ac` written as
// zero; the reference leaves them uninitialised without vertex colours).  Three
// vertices per visible tri, leaf after leaf (DFS order), tris in prim order.
//
// Visible tri (gpu_pbvh_is_looptri_visible, :165-172): no hidden corner and a face
// set above 0.  A leaf draws smooth (vertex normals) when its first face is ME_SMOOTH,
// else with face normals (BKE_mesh_calc_poly_normal, kernel/intern/mesh_evaluate.c:
// 39-84).  Mask bytes follow :246-265; face-set colours BKE_paint_face_set_overlay_color_get
// (kernel/intern/paint.c:2219-2235).
//
// Which leaves are refilled: those of every search subtree a dab gathered leaves from
// (the brush marks its gathered nodes UpdateDrawBuffers; the search marks their subtree,
// one store per subtree, in an array that lives as long as the PBVH), plus those of
// subtrees holding host-marked nodes.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "sg_internal.h"
#include "sg_kernels.h"
#include "sg_device.h"

#define DRAW_STRIDE 36

struct SgDrawView {
  const int4 *ltri;
  const int *toff, *tri_poly;
  int *dtri;   // visible tris' ltri positions
  int *dleaf;  // bit0: the leaf draws smooth
  const int *doff, *fset, *poly_off, *poly_v;
  const int *leaf_cut;  // leaf -> search subtree
  int *dirty;           // per subtree
  uint8_t *vbo;
  int seed, fset_default;
};

__device__ __forceinline__ bool draw_visible(const SgDrawView &w, int4 tr) {
  if (tr.w < 0) return false;  // a hidden corner (paint_is_face_hidden)
  return w.fset[w.tri_poly[tr.w & 0x7fffffff]] > 0;
}

// per leaf: visible tri count; smooth flag of the first face
__global__ __launch_bounds__(256) void k_draw_count(SgDrawView w, int L, const uint8_t *poly_smooth, int *count) {
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  for (int l = gw; l < L; l += nw) {
    const int tb = w.toff[l], te = w.toff[l + 1];
    int n = 0;
    for (int t = tb + lane; t < te; t += SG_WAVE) n += draw_visible(w, w.ltri[t]) ? 1 : 0;
    n = wave_sum_i(n);
    if (lane == 0) {
      count[l] = n;
      w.dleaf[l] = (te > tb) ? (int)poly_smooth[w.tri_poly[w.ltri[tb].w & 0x7fffffff]] : 0;
    }
  }
}

// exclusive scan of L counts into off[0..L] (one workgroup)
__global__ __launch_bounds__(1024) void k_draw_scan(const int *count, int L, int *off) {
  __shared__ int s_part[1024];
  const int tid = threadIdx.x;
  const int per = (L + 1023) / 1024;
  const int b = tid * per, e = min(L, b + per);
  int sum = 0;
  for (int i = b; i < e; i++) sum += count[i];
  s_part[tid] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = tid >= o ? s_part[tid - o] : 0;
    __syncthreads();
    s_part[tid] += v;
    __syncthreads();
  }
  int run = s_part[tid] - sum;
  for (int i = b; i < e; i++) {
    off[i] = run;
    run += count[i];
  }
  if (tid == 1023) off[L] = s_part[1023];
}

// per leaf: the visible tris' ltri positions, in prim order; bit 31 when the tri's
// polygon is the same triangle in the same corner order (tri_is_poly[looptri])
__global__ __launch_bounds__(256) void k_draw_compact(SgDrawView w, int L, const uint8_t *tri_is_poly) {
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  for (int l = gw; l < L; l += nw) {
    const int tb = w.toff[l], te = w.toff[l + 1];
    int base = w.doff[l];
    for (int t0 = tb; t0 < te; t0 += SG_WAVE) {
      const int t = t0 + lane;
      int4 tr = make_int4(0, 0, 0, -1);
      if (t < te) tr = w.ltri[t];
      const bool vis = t < te && draw_visible(w, tr);
      const unsigned long long m = __ballot(vis);
      if (vis) w.dtri[base + mask_prefix(m)] = t | (tri_is_poly[tr.w & 0x7fffffff] ? (int)0x80000000u : 0);
      base += __popcll(m);
    }
  }
}

__device__ __forceinline__ short to_short(float f) { return (short)(f * 32767.0f); }

__device__ __forceinline__ float d_normalize(float &x, float &y, float &z) { return sg_normalize(x, y, z); }

// BKE_mesh_calc_poly_normal on the current positions (corners in the new numbering)
__device__ void poly_normal(const SgDevView &d, const SgDrawView &w, int poly, float &nx, float &ny, float &nz) {
  const int b = w.poly_off[poly], n = w.poly_off[poly + 1] - b;
  auto P = [&](int k, float &x, float &y, float &z) {
    const int v = w.poly_v[b + k];
    x = d.px[v];
    y = d.py[v];
    z = d.pz[v];
  };
  if (n > 4) {  // Newell (add_newell_cross_v3_v3v3), zero -> +z
    nx = ny = nz = 0.0f;
    float ax, ay, az;
    P(n - 1, ax, ay, az);
    for (int k = 0; k < n; k++) {
      float cx, cy, cz;
      P(k, cx, cy, cz);
      nx += (ay - cy) * (az + cz);
      ny += (az - cz) * (ax + cx);
      nz += (ax - cx) * (ay + cy);
      ax = cx;
      ay = cy;
      az = cz;
    }
    if (d_normalize(nx, ny, nz) == 0.0f) nz = 1.0f;
  } else if (n == 3) {
    float a0, a1, a2, b0, b1, b2, c0, c1, c2;
    P(0, a0, a1, a2);
    P(1, b0, b1, b2);
    P(2, c0, c1, c2);
    sg_tri_normal(a0, a1, a2, b0, b1, b2, c0, c1, c2, nx, ny, nz);
  } else if (n == 4) {  // normal_quad_v3: cross of the diagonals
    float v1[3], v2[3], v3[3], v4[3];
    P(0, v1[0], v1[1], v1[2]);
    P(1, v2[0], v2[1], v2[2]);
    P(2, v3[0], v3[1], v3[2]);
    P(3, v4[0], v4[1], v4[2]);
    const float n10 = v1[0] - v3[0], n11 = v1[1] - v3[1], n12 = v1[2] - v3[2];
    const float n20 = v2[0] - v4[0], n21 = v2[1] - v4[1], n22 = v2[2] - v4[2];
    nx = n11 * n22 - n12 * n21;
    ny = n12 * n20 - n10 * n22;
    nz = n10 * n21 - n11 * n20;
    d_normalize(nx, ny, nz);
  } else {
    nx = 0.0f;
    ny = 0.0f;
    nz = 1.0f;
  }
}

// BLI_hash_int_2d(k, 0) / 0xFFFFFFFF: lookup3's final mix (the header holding it is
// absent from the reference; this is its published algorithm)
__device__ __forceinline__ unsigned h_rot(unsigned x, int k) { return (x << k) | (x >> (32 - k)); }
__device__ float hash_int_01(unsigned k) {
  unsigned a, b, c;
  a = b = c = 0xdeadbeefu + (2u << 2) + 13u;
  a += k;
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
  return (float)c * (1.0f / (float)0xFFFFFFFFu);
}

__device__ __forceinline__ unsigned char unit_to_uchar(float v) {
  return (unsigned char)((v <= 0.0f) ? 0.0f : ((v > (1.0f - 0.5f / 255.0f)) ? 255.0f : ((255.0f * v) + 0.5f)));
}

// BKE_paint_face_set_overlay_color_get -> rgb; white for the default face set
__device__ unsigned face_set_rgb(int fs, int seed, int fdefault) {
  if (fs == fdefault) return 0xffffffu;
  float hue = 0.618033988749895f * (float)(fs + (seed % 10));
  hue = hue - floorf(hue);
  const float sat = hash_int_01((unsigned)(fs + seed + 1));
  const float val = hash_int_01((unsigned)(fs + seed + 2));
  const float s = 0.6f + (sat * 0.25f), v = 1.0f - (val * 0.35f);
  float nr = fabsf(hue * 6.0f - 3.0f) - 1.0f;
  float ng = 2.0f - fabsf(hue * 6.0f - 2.0f);
  float nb = 2.0f - fabsf(hue * 6.0f - 4.0f);
  nr = nr < 0.0f ? 0.0f : (nr > 1.0f ? 1.0f : nr);
  nb = nb < 0.0f ? 0.0f : (nb > 1.0f ? 1.0f : nb);
  ng = ng < 0.0f ? 0.0f : (ng > 1.0f ? 1.0f : ng);
  const unsigned r = unit_to_uchar(((nr - 1.0f) * s + 1.0f) * v);
  const unsigned g = unit_to_uchar(((ng - 1.0f) * s + 1.0f) * v);
  const unsigned bl = unit_to_uchar(((nb - 1.0f) * s + 1.0f) * v);
  return r | (g << 8) | (bl << 16);
}

// Fill the records of stale leaves (all = every leaf): wave per leaf, lane per tri in
// chunks of 64 tris; each lane builds its tri's 27 dwords in LDS (stride 27, odd: no
// bank conflicts), then the wave stores the chunk's 64 x 108 bytes contiguously.
// A draw_tri entry with bit 31 set is a tri whose polygon is that same triangle in the
// same corner order: its face normal comes from the corners already loaded.
#define DRAW_WAVES 4
__global__ __launch_bounds__(64 * DRAW_WAVES) void k_draw_fill(SgDevView d, SgDrawView w, int L, int all) {
  __shared__ uint32_t s_rec[DRAW_WAVES][64 * 27];
  const int lane = lane_id();
  const int wid = threadIdx.x / SG_WAVE;
  uint32_t *rec = s_rec[wid];
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  for (int l = gw; l < L; l += nw) {
    if (!all && !w.dirty[w.leaf_cut[l]]) continue;
    const int sb = w.doff[l], se = w.doff[l + 1];
    const bool smooth = (w.dleaf[l] & 1) != 0;
    for (int s0 = sb; s0 < se; s0 += SG_WAVE) {
      const int s = s0 + lane;
      if (s < se) {
        const int e = w.dtri[s];
        const int4 tr = w.ltri[e & 0x7fffffff];
        const int poly = w.tri_poly[tr.w & 0x7fffffff];
        const int cv[3] = {tr.x, tr.y, tr.z};
        float co[3][3];
        for (int j = 0; j < 3; j++) {
          co[j][0] = d.px[cv[j]];
          co[j][1] = d.py[cv[j]];
          co[j][2] = d.pz[cv[j]];
        }
        short no[3][3];
        if (smooth) {
          for (int j = 0; j < 3; j++) {
            no[j][0] = to_short(d.nx[cv[j]]);
            no[j][1] = to_short(d.ny[cv[j]]);
            no[j][2] = to_short(d.nz[cv[j]]);
          }
        } else {
          float fx, fy, fz;
          if (e < 0)
            sg_tri_normal(co[0][0], co[0][1], co[0][2], co[1][0], co[1][1], co[1][2], co[2][0], co[2][1], co[2][2],
                          fx, fy, fz);
          else
            poly_normal(d, w, poly, fx, fy, fz);
          for (int j = 0; j < 3; j++) {
            no[j][0] = to_short(fx);
            no[j][1] = to_short(fy);
            no[j][2] = to_short(fz);
          }
        }
        unsigned char msk[3] = {0, 0, 0};
        if (d.mask) {
          if (smooth) {
            for (int j = 0; j < 3; j++) msk[j] = (unsigned char)(d.mask[cv[j]] * 255);
          } else {
            const float fm = (d.mask[cv[0]] + d.mask[cv[1]] + d.mask[cv[2]]) / 3.0f;
            msk[0] = msk[1] = msk[2] = (unsigned char)(fm * 255);
          }
        }
        const unsigned rgb = face_set_rgb(abs(w.fset[poly]), w.seed, w.fset_default);
        uint32_t *r = rec + 27 * lane;
        for (int j = 0; j < 3; j++, r += 9) {
          r[0] = __float_as_uint(co[j][0]);
          r[1] = __float_as_uint(co[j][1]);
          r[2] = __float_as_uint(co[j][2]);
          r[3] = 0u;
          r[4] = (uint32_t)(uint16_t)no[j][0] | ((uint32_t)(uint16_t)no[j][1] << 16);
          r[5] = (uint32_t)(uint16_t)no[j][2] | ((uint32_t)msk[j] << 16);
          r[6] = 0u;
          r[7] = 0u;
          r[8] = rgb;
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
      const int nd = min(SG_WAVE, se - s0) * 27;
      uint32_t *out = (uint32_t *)(w.vbo + (size_t)s0 * 3 * DRAW_STRIDE);
      for (int k = lane; k < nd; k += SG_WAVE) out[k] = rec[k];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    }
  }
}

__global__ void k_draw_mark(int *dirty, const int *leaf_cut, const int *leaves, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dirty[leaf_cut[leaves[i]]] = 1;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static SgDrawView draw_view(PBVH *p) {
  SgDrawView w;
  const SgDev &d = p->d;
  w.ltri = d.ltri;
  w.toff = d.leaf_toff;
  w.tri_poly = d.tri_poly;
  w.dtri = d.draw_tri;
  w.doff = d.draw_off;
  w.dleaf = d.draw_leaf;
  w.fset = d.draw_fset;
  w.poly_off = d.draw_poly_off;
  w.poly_v = d.draw_poly_v;
  w.dirty = d.draw_dirty;
  w.leaf_cut = d.leaf_cut;
  w.vbo = d.draw_vbo;
  w.seed = d.draw_seed;
  w.fset_default = d.draw_fset_default;
  return w;
}

extern "C" void sg_pbvh_draw_buffers_free(PBVH *p);

template <class T>
static int dev_upload(T **dst, const T *src, size_t n) {
  if (sg_check_hip(hipMalloc((void **)dst, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc")) return -1;
  return n ? sg_check_hip(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice), "copy") : 0;
}

extern "C" {

int sg_pbvh_draw_buffers_build(PBVH *p, const int *face_sets, int npoly, int seed, int fset_default) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: draw buffers of PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p || npoly < 0 || (npoly > 0 && !face_sets)) return sg_set_error("bad arguments");
  if (p->no_device || !p->build_ok) return sg_set_error("no device mesh");
  if (npoly != (int)p->poly_smooth.size()) return sg_set_error("face sets: %d polygons, the mesh has %zu", npoly,
                                                               p->poly_smooth.size());
  if (int rc = sg_ray_ensure(p)) return rc;  // leaf-ordered tri table, looptri -> poly
  sg_pbvh_draw_buffers_free(p);
  SgDev &d = p->d;
  const int L = d.L, P = npoly;
  // polygon corners in the leaf-major numbering; per-poly smooth flag and material
  std::vector<int> pv(p->poly_verts.size());
  for (size_t k = 0; k < pv.size(); k++) pv[k] = p->perm_new_of_old[p->poly_verts[k]];
  std::vector<uint8_t> sm(std::max(P, 1), 0);
  for (int q = 0; q < P; q++) sm[q] = p->poly_smooth[q];
  // a looptri whose polygon is that triangle, corners in the same order
  std::vector<uint8_t> tip((size_t)std::max(p->totprim, 1), 0);
  for (int t = 0; t < p->totprim; t++) {
    const int q = p->looptri ? (int)p->looptri[t].poly : t;
    const int b = p->poly_off[q];
    tip[t] = (p->poly_off[q + 1] - b == 3 && p->poly_verts[b] == p->tris[3 * (size_t)t] &&
              p->poly_verts[b + 1] == p->tris[3 * (size_t)t + 1] && p->poly_verts[b + 2] == p->tris[3 * (size_t)t + 2])
                 ? 1
                 : 0;
  }
  uint8_t *dtp = nullptr;
  uint8_t *dsm = nullptr;
  int *count = nullptr;
  int rc = dev_upload(&d.draw_fset, face_sets, (size_t)P);
  if (!rc) rc = dev_upload(&d.draw_poly_off, p->poly_off.data(), p->poly_off.size());
  if (!rc) rc = dev_upload(&d.draw_poly_v, pv.data(), pv.size());
  if (!rc) rc = dev_upload(&dsm, sm.data(), sm.size());
  if (!rc) rc = dev_upload(&dtp, tip.data(), tip.size());
  if (!rc) rc = sg_check_hip(hipMalloc(&d.draw_off, ((size_t)L + 1) * sizeof(int)), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&d.draw_leaf, (size_t)std::max(L, 1) * sizeof(int)), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMalloc(&count, (size_t)std::max(L, 1) * sizeof(int)), "hipMalloc");
  if (!rc) rc = sg_check_hip(hipMemset(d.draw_dirty, 0, (size_t)std::max(d.C, 1) * sizeof(int)), "memset");
  d.draw_seed = seed;
  d.draw_fset_default = fset_default;
  int ntri = 0;
  if (!rc) {
    SgDrawView w = draw_view(p);
    hipLaunchKernelGGL(k_draw_count, dim3(std::max(1, std::min((L + 3) / 4, 4096))), dim3(256), 0, p->stream, w, L,
                       dsm, count);
    hipLaunchKernelGGL(k_draw_scan, dim3(1), dim3(1024), 0, p->stream, count, L, d.draw_off);
    rc = sg_check_hip(hipGetLastError(), "draw build launch");
    if (!rc)
      rc = sg_check_hip(hipMemcpyAsync(&ntri, d.draw_off + L, sizeof(int), hipMemcpyDeviceToHost, p->stream), "copy");
    if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "draw build");
  }
  if (!rc) rc = sg_check_hip(hipMalloc(&d.draw_tri, (size_t)std::max(ntri, 1) * sizeof(int)), "hipMalloc");
  d.draw_ntri = ntri;
  d.draw_bytes = (long long)ntri * 3 * DRAW_STRIDE;
  if (!rc) rc = sg_check_hip(hipMalloc(&d.draw_vbo, (size_t)std::max(d.draw_bytes, 1LL)), "hipMalloc");
  if (!rc) {
    SgDrawView w = draw_view(p);
    hipLaunchKernelGGL(k_draw_compact, dim3(std::max(1, std::min((L + 3) / 4, 4096))), dim3(256), 0, p->stream, w,
                       L, dtp);
    hipLaunchKernelGGL(k_draw_fill, dim3(std::max(1, std::min((L + DRAW_WAVES - 1) / DRAW_WAVES, 8192))),
                       dim3(64 * DRAW_WAVES), 0, p->stream, sg_view(p), w, L, 1);
    rc = sg_check_hip(hipGetLastError(), "draw fill launch");
    if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "draw build");
  }
  if (dsm) (void)hipFree(dsm);
  if (dtp) (void)hipFree(dtp);
  if (!rc) {  // the full fill consumed every node's draw marks
    for (PBVHNode &n : p->nodes) n.flag &= ~(PBVH_UpdateDrawBuffers | PBVH_RebuildDrawBuffers);
    p->draw_marks_pending = false;
  }
  if (count) (void)hipFree(count);
  if (rc) sg_pbvh_draw_buffers_free(p);
  return rc;
}

// Refill the stale leaves (all != 0: every leaf): those the dabs since the last
// update gathered and those marked on the host (PBVH_UpdateDrawBuffers /
// PBVH_RebuildDrawBuffers, cleared here as pbvh_update_draw_buffers clears them).
int sg_pbvh_draw_buffers_update(PBVH *p, int all) {
  if (p && p->type != PBVH_FACES) return sg_set_error("faces only: draw buffers of PBVH_GRIDS / PBVH_BMESH are not built");
  if (!p) return sg_set_error("bad arguments");
  if (!p->d.draw_vbo) return sg_set_error("draw buffers not built");
  std::vector<int> marked;
  if (p->draw_marks_pending) {
    for (PBVHNode &n : p->nodes) {
      if (n.leaf >= 0 && (n.flag & (PBVH_UpdateDrawBuffers | PBVH_RebuildDrawBuffers))) {
        marked.push_back(n.leaf);
        n.flag &= ~(PBVH_UpdateDrawBuffers | PBVH_RebuildDrawBuffers);
      }
    }
    p->draw_marks_pending = false;
  }
  int rc = 0;
  int *dl = nullptr;
  if (!marked.empty() && !all) {  // host marks join the search's marks on the device
    rc = sg_check_hip(hipMalloc(&dl, marked.size() * sizeof(int)), "hipMalloc");
    if (!rc)
      rc = sg_check_hip(
          hipMemcpyAsync(dl, marked.data(), marked.size() * sizeof(int), hipMemcpyHostToDevice, p->stream), "copy");
    if (!rc) {
      const int n = (int)marked.size();
      hipLaunchKernelGGL(k_draw_mark, dim3(std::max(1, std::min((n + 255) / 256, 1024))), dim3(256), 0, p->stream,
                         p->d.draw_dirty, p->d.leaf_cut, dl, n);
      rc = sg_check_hip(hipGetLastError(), "draw mark launch");
    }
    if (rc) {
      if (dl) (void)hipFree(dl);
      return rc;
    }
  }
  const int L = p->d.L;
  hipLaunchKernelGGL(k_draw_fill, dim3(std::max(1, std::min((L + DRAW_WAVES - 1) / DRAW_WAVES, 8192))),
                     dim3(64 * DRAW_WAVES), 0, p->stream, sg_view(p), draw_view(p), L, all ? 1 : 0);
  rc = sg_check_hip(hipGetLastError(), "draw fill launch");
  if (!rc) rc = sg_check_hip(hipMemsetAsync(p->d.draw_dirty, 0, (size_t)p->d.C * sizeof(int), p->stream), "memset");
  if (!rc) rc = sg_check_hip(hipStreamSynchronize(p->stream), "draw update");
  if (dl) (void)hipFree(dl);
  return rc;
}

int sg_pbvh_draw_buffers_info(PBVH *p, void **vbo_dev, long long *bytes, int *visible_tris) {
  if (!p) return sg_set_error("bad arguments");
  if (vbo_dev) *vbo_dev = p->d.draw_vbo;
  if (bytes) *bytes = p->d.draw_bytes;
  if (visible_tris) *visible_tris = p->d.draw_ntri;
  return p->d.draw_vbo ? 0 : sg_set_error("draw buffers not built");
}

int sg_pbvh_draw_buffer_node(PBVH *p, PBVHNode *node, long long *byte_offset, int *vert_count, int *smooth,
                             int *material) {
  if (!p || !node || node->leaf < 0) return sg_set_error("not a leaf");
  if (!p->d.draw_vbo) return sg_set_error("draw buffers not built");
  int off[2];
  if (int rc = sg_check_hip(hipMemcpy(off, p->d.draw_off + node->leaf, 2 * sizeof(int), hipMemcpyDeviceToHost), "copy"))
    return rc;
  if (byte_offset) *byte_offset = (long long)off[0] * 3 * DRAW_STRIDE;
  if (vert_count) *vert_count = (off[1] - off[0]) * 3;
  const int first = p->prim_indices[node->prim_off];
  const int poly = p->looptri ? (int)p->looptri[first].poly : first;
  if (smooth) *smooth = p->poly_smooth.empty() ? 0 : p->poly_smooth[poly];
  if (material) *material = p->poly_mat.empty() ? 0 : p->poly_mat[poly];
  return 0;
}

// Host copy of one leaf's records (tests, or a renderer in another process).
long long sg_pbvh_draw_buffer_read(PBVH *p, PBVHNode *node, void *out, long long cap) {
  long long off = 0;
  int nv = 0;
  if (int rc = sg_pbvh_draw_buffer_node(p, node, &off, &nv, nullptr, nullptr)) return rc;
  const long long bytes = (long long)nv * DRAW_STRIDE;
  if (bytes > cap) return sg_set_error("buffer too small");
  if (bytes)
    if (int rc = sg_check_hip(hipMemcpy(out, p->d.draw_vbo + off, (size_t)bytes, hipMemcpyDeviceToHost), "copy"))
      return rc;
  return bytes;
}

void sg_pbvh_draw_buffers_free(PBVH *p) {
  if (!p) return;
  SgDev &d = p->d;
  // draw_dirty stays: captured strokes keep marking it for the PBVH's lifetime
  void *q[] = {d.draw_tri, d.draw_off, d.draw_leaf, d.draw_fset, d.draw_poly_off, d.draw_poly_v, d.draw_vbo};
  for (void *x : q)
    if (x) (void)hipFree(x);
  d.draw_tri = d.draw_off = d.draw_leaf = d.draw_fset = d.draw_poly_off = d.draw_poly_v = nullptr;
  d.draw_vbo = nullptr;
  d.draw_bytes = 0;
  d.draw_ntri = 0;
}

}  // extern "C"
