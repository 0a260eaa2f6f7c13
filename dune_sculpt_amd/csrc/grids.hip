// PBVH_GRIDS on the GPU: the device layout of a tree over multires grids, one dab
// (search -> brush -> Smooth -> grid normals and seam averaging -> refit), the
// drop-in normals / bounds updates, and the host entry points of the grids API.
//
// A dab follows the faces dab's written spec with grid elements in place of mesh
// vertices: leaves gathered by the sphere test on their box (orig box for Grab),
// every visible element of a gathered leaf tested (grids in the leaf's prim order,
// elements row-major, the order of the reference's vertex iterator), falloff and
// displacement as for faces, Smooth(k) as k Jacobi steps over the element's
// neighbours (KERNEL_subdiv_ccg_neighbor_coords_get without duplicates, subdiv_ccg.c:
// 1882).  Every gathered leaf is marked UpdateNormals | UpdateBB | UpdateOriginalBB
// as on faces; BKE_pbvh_update_normals for grids (pbvh.c:4559-4587 -> :1605
// BKE_pbvh_get_grid_updates -> KERNEL_subdiv_ccg_update_normals, subdiv_ccg.c:847)
// then recomputes the grids of every face with a grid in a marked leaf and averages
// their seams, edges and corners.  Those averages also move elements of leaves the
// dab did not gather; their leaves are refit with the marked ones, so every box keeps
// holding its elements.
//
// Elements are SoA in HBM, numbered grid by grid in leaf order, each grid row-major;
// per-element work is a thread per element over chunks of 1024 elements of the
// gathered leaves (a leaf holds up to leaf_limit elements, 10000 by default).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "sg_device.h"
#include "sg_grids.h"

#define GR_THREADS 256
#define GR_CHUNK 256  // elements per work item (1 per thread)

namespace {

struct GView {  // kernel view of SgGridDev (by value)
  int G, gs, ga, L, N, H, nchunk;
  float *px, *py, *pz, *qx, *qy, *qz, *nx, *ny, *nz, *mask;
  const unsigned char *hide;
  int *stamp;
  float *fade;
  const int *grid_pos, *grid_old, *grid_leaf, *leaf_gbeg, *leaf_flags;
  int *leaf_mark;
  float *leaf_bb, *leaf_obb;
  const int *node_child, *node_leaf;
  int *node_mark;
  float *node_bb, *node_obb;
  const int *lvl_off, *lvl_node;
  const long long *nb_off;
  const int *nb_idx;
  int F, Ed, Vc;
  const int *face_gbeg, *grid_face, *face_edge, *face_vert, *edge_foff, *edge_el, *vert_foff, *vert_el;
  int *face_stamp, *edge_stamp, *vert_stamp;
  int *gathered, *dfaces, *ditems;
  long long *ctr, *ctr_last;
  int *dser;  // [0] the running operation's serial, [1] the last dab's
  int *leaf_dirty, *node_dirty;  // refit since the last internal-node flush
};

GView view_of(const SgGridDev &d) {
  GView v;
  memset(&v, 0, sizeof(v));
  v.G = d.G;
  v.gs = d.gs;
  v.ga = d.ga;
  v.L = d.L;
  v.N = d.N;
  v.H = d.H;
  v.nchunk = (int)(((long long)d.maxleafg * d.ga + GR_CHUNK - 1) / GR_CHUNK);
  v.px = d.px;
  v.py = d.py;
  v.pz = d.pz;
  v.qx = d.qx;
  v.qy = d.qy;
  v.qz = d.qz;
  v.nx = d.nx;
  v.ny = d.ny;
  v.nz = d.nz;
  v.mask = d.mask;
  v.hide = d.hide;
  v.stamp = d.stamp;
  v.fade = d.fade;
  v.grid_pos = d.grid_pos;
  v.grid_old = d.grid_old;
  v.grid_leaf = d.grid_leaf;
  v.leaf_gbeg = d.leaf_gbeg;
  v.leaf_flags = d.leaf_flags;
  v.leaf_mark = d.leaf_mark;
  v.leaf_bb = d.leaf_bb;
  v.leaf_obb = d.leaf_obb;
  v.node_child = d.node_child;
  v.node_leaf = d.node_leaf;
  v.node_mark = d.node_mark;
  v.node_bb = d.node_bb;
  v.node_obb = d.node_obb;
  v.lvl_off = d.lvl_off;
  v.lvl_node = d.lvl_node;
  v.nb_off = d.nb_off;
  v.nb_idx = d.nb_idx;
  v.F = d.F;
  v.Ed = d.Ed;
  v.Vc = d.Vc;
  v.face_gbeg = d.face_gbeg;
  v.grid_face = d.grid_face;
  v.face_edge = d.face_edge;
  v.face_vert = d.face_vert;
  v.edge_foff = d.edge_foff;
  v.edge_el = d.edge_el;
  v.vert_foff = d.vert_foff;
  v.vert_el = d.vert_el;
  v.face_stamp = d.face_stamp;
  v.edge_stamp = d.edge_stamp;
  v.vert_stamp = d.vert_stamp;
  v.gathered = d.gathered;
  v.dfaces = d.dfaces;
  v.ditems = d.ditems;
  v.ctr = d.ctr;
  v.ctr_last = d.ctr_last;
  v.dser = d.dser;
  v.leaf_dirty = d.leaf_dirty;
  v.node_dirty = d.node_dirty;
  return v;
}

__device__ __forceinline__ BB6 gbb_load(const float *p) { return bb_load(p); }

// The work item `it` of the element chunks of gathered leaves: element range [b, e).
__device__ __forceinline__ bool chunk_range(const GView &g, int it, int ng, int &b, int &e) {
  const int gi = it / g.nchunk, ch = it - gi * g.nchunk;
  if (gi >= ng) return false;
  const int l = g.gathered[gi];
  const int lb = g.leaf_gbeg[l] * g.ga, le = g.leaf_gbeg[l + 1] * g.ga;
  b = lb + ch * GR_CHUNK;
  e = min(le, b + GR_CHUNK);
  return b < e;
}

// ---------------------------------------------------------------------------
// search: a thread per leaf (spec S1 on vb, orig_vb for Grab; fully hidden leaves
// skipped); every gathered leaf is marked for the refit
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GR_THREADS) void kg_search(GView g, SgDabK k) {
  const int serial = g.dser[0];
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  const int lc = min(l, g.L - 1);
  const BB6 b = gbb_load((k.use_orig ? g.leaf_obb : g.leaf_bb) + 8 * (size_t)lc);
  const bool pass = l < g.L && !((g.leaf_flags[lc] & 1) && k.hide_skip) && sphere_box(b, k.cx, k.cy, k.cz, k.r2);
  const unsigned long long m = __ballot(pass);
  if (m == 0) return;
  int base = 0;
  if (lane_id() == 0) base = (int)atomicAdd((unsigned long long *)&g.ctr[SGG_NGATHER], (unsigned long long)__popcll(m));
  base = __shfl(base, 0);
  if (pass) {
    g.gathered[base + mask_prefix(m)] = l;
    g.leaf_mark[l] = serial;
  }
}

// ---------------------------------------------------------------------------
// Draw area normal: normals of the visible elements within normal_radius_factor * r,
// summed exactly in 2^-32 fixed point (any order gives the same bits)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GR_THREADS) void kg_area(GView g, SgDabK k, int respect_hide) {
  const int ng = (int)g.ctr[SGG_NGATHER];
  long long sx = 0, sy = 0, sz = 0;
  int na = 0;
  for (int it = blockIdx.x; it < ng * g.nchunk; it += gridDim.x) {
    int b, e;
    if (!chunk_range(g, it, ng, b, e)) continue;
    for (int v = b + threadIdx.x; v < e; v += GR_THREADS) {
      if (g.hide && respect_hide && g.hide[v]) continue;
      const float dx = k.cx - g.px[v], dy = k.cy - g.py[v], dz = k.cz - g.pz[v];
      const float d2 = dx * dx + dy * dy + dz * dz;
      if (d2 <= k.nr2) {
        sx += __double2ll_rn((double)g.nx[v] * 4294967296.0);
        sy += __double2ll_rn((double)g.ny[v] * 4294967296.0);
        sz += __double2ll_rn((double)g.nz[v] * 4294967296.0);
        na++;
      }
    }
  }
  sx = wave_sum_ll(sx);
  sy = wave_sum_ll(sy);
  sz = wave_sum_ll(sz);
  na = wave_sum_i(na);
  if (lane_id() == 0 && na) {
    atomicAdd((unsigned long long *)&g.ctr[SGG_AREA_X], as_u64(sx));
    atomicAdd((unsigned long long *)&g.ctr[SGG_AREA_Y], as_u64(sy));
    atomicAdd((unsigned long long *)&g.ctr[SGG_AREA_Z], as_u64(sz));
    atomicAdd((unsigned long long *)&g.ctr[SGG_NAREA], (unsigned long long)na);
  }
}

// ---------------------------------------------------------------------------
// brush: spec S2 test, S3/S4 factor, S6 Draw / S7 Inflate / S9 Grab written to both
// position copies; Smooth keeps the factor for the Jacobi steps
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GR_THREADS) void kg_brush(GView g, SgDabK k, const float *curve, int respect_hide,
                                                       int claim_faces) {
  const int serial = g.dser[0];
  const int ng = (int)g.ctr[SGG_NGATHER];
  float ox = 0.0f, oy = 0.0f, oz = 0.0f;
  if (k.tool == SG_TOOL_DRAW) {
    float ax = (float)((double)g.ctr[SGG_AREA_X] * (1.0 / 4294967296.0));
    float ay = (float)((double)g.ctr[SGG_AREA_Y] * (1.0 / 4294967296.0));
    float az = (float)((double)g.ctr[SGG_AREA_Z] * (1.0 / 4294967296.0));
    sg_normalize(ax, ay, az);
    ox = (ax * k.r) * k.bstr;
    oy = (ay * k.r) * k.bstr;
    oz = (az * k.r) * k.bstr;
  }
  int naff = 0;
  for (int it = blockIdx.x; it < ng * g.nchunk; it += gridDim.x) {
    int b, e;
    if (!chunk_range(g, it, ng, b, e)) continue;
    if (claim_faces && it % g.nchunk == 0) {
      // the faces of this leaf's grids, each once (BKE_pbvh_get_grid_updates,
      // pbvh.c:1605-1640): their normals and seams are redone after the brush
      const int l = g.gathered[it / g.nchunk];
      for (int p = g.leaf_gbeg[l] + threadIdx.x; p < g.leaf_gbeg[l + 1]; p += GR_THREADS) {
        const int f = g.grid_face[g.grid_old[p]];
        if (atomicExch(&g.face_stamp[f], serial) != serial) {
          const int slot = (int)atomicAdd((unsigned long long *)&g.ctr[SGG_NFACES], 1ull);
          g.dfaces[slot] = f;
        }
      }
    }
    for (int v = b + threadIdx.x; v < e; v += GR_THREADS) {
      if (g.hide && respect_hide && g.hide[v]) continue;
      const float x = g.px[v], y = g.py[v], z = g.pz[v];
      const float dx = k.cx - x, dy = k.cy - y, dz = k.cz - z;
      const float d2 = dx * dx + dy * dy + dz * dz;
      if (!(d2 <= k.r2)) continue;
      naff++;
      g.stamp[v] = serial;
      const float mk = g.mask ? g.mask[v] : 0.0f;
      const float f = sg_falloff(k, curve, sqrtf(d2)) * (1.0f - mk);
      if (k.tool == SG_TOOL_SMOOTH) {
        g.fade[v] = k.bstr * f;
        continue;
      }
      float rx, ry, rz;
      if (k.tool == SG_TOOL_DRAW) {
        rx = x + ox * f;
        ry = y + oy * f;
        rz = z + oz * f;
      } else if (k.tool == SG_TOOL_INFLATE) {
        const float s = (k.bstr * f) * k.r;
        rx = x + g.nx[v] * s;
        ry = y + g.ny[v] * s;
        rz = z + g.nz[v] * s;
      } else {
        const float fade = k.bstr * f;
        rx = x + k.gx * fade;
        ry = y + k.gy * fade;
        rz = z + k.gz * fade;
      }
      g.px[v] = rx;
      g.py[v] = ry;
      g.pz[v] = rz;
      g.qx[v] = rx;
      g.qy[v] = ry;
      g.qz[v] = rz;
    }
  }
  naff = wave_sum_i(naff);
  if (lane_id() == 0 && naff) atomicAdd((unsigned long long *)&g.ctr[SGG_NAFF], (unsigned long long)naff);
}

// One Jacobi step of spec S8: src -> dst for this dab's affected elements; the others
// hold the same values in both copies.
__global__ __launch_bounds__(GR_THREADS) void kg_smooth(GView g, const float *sx, const float *sy, const float *sz,
                                                        float *dx, float *dy, float *dz) {
  const int serial = g.dser[0];
  const int ng = (int)g.ctr[SGG_NGATHER];
  for (int it = blockIdx.x; it < ng * g.nchunk; it += gridDim.x) {
    int b, e;
    if (!chunk_range(g, it, ng, b, e)) continue;
    for (int v = b + threadIdx.x; v < e; v += GR_THREADS) {
      if (g.stamp[v] != serial) continue;
      const int local = v % g.ga, y = local / g.gs, x = local - y * g.gs;
      const float px = sx[v], py = sy[v], pz = sz[v];
      float ax = 0.0f, ay = 0.0f, az = 0.0f;
      int cnt;
      if (x > 0 && y > 0 && x < g.gs - 1 && y < g.gs - 1) {
        // inside the grid: previous row, next row, previous column, next column
        // (neighbor_coords_inner_get, subdiv_ccg.c:1870-1880)
        const int u[4] = {v - g.gs, v + g.gs, v - 1, v + 1};
        for (int q = 0; q < 4; q++) {
          ax = ax + sx[u[q]];
          ay = ay + sy[u[q]];
          az = az + sz[u[q]];
        }
        cnt = 4;
      } else {
        const long long nb = g.nb_off[v], ne = g.nb_off[v + 1];
        for (long long q = nb; q < ne; q++) {
          const int u = g.nb_idx[q];
          ax = ax + sx[u];
          ay = ay + sy[u];
          az = az + sz[u];
        }
        cnt = (int)(ne - nb);
      }
      float rx = px, ry = py, rz = pz;
      if (cnt > 0) {
        const float fade = g.fade[v];
        const float inv = 1.0f / (float)cnt;
        rx = px + (ax * inv - px) * fade;
        ry = py + (ay * inv - py) * fade;
        rz = pz + (az * inv - pz) * fade;
      }
      dx[v] = rx;
      dy[v] = ry;
      dz[v] = rz;
    }
  }
}

__global__ __launch_bounds__(GR_THREADS) void kg_copy_aff(GView g, const float *sx, const float *sy, const float *sz,
                                                          float *dx, float *dy, float *dz) {
  const int serial = g.dser[0];
  const int ng = (int)g.ctr[SGG_NGATHER];
  for (int it = blockIdx.x; it < ng * g.nchunk; it += gridDim.x) {
    int b, e;
    if (!chunk_range(g, it, ng, b, e)) continue;
    for (int v = b + threadIdx.x; v < e; v += GR_THREADS) {
      if (g.stamp[v] != serial) continue;
      dx[v] = sx[v];
      dy[v] = sy[v];
      dz[v] = sz[v];
    }
  }
}

// Faces of the grids of the listed leaves (the gathered ones when leaves == null),
// each once (BKE_pbvh_get_grid_updates, pbvh.c:1605-1640).
__global__ __launch_bounds__(GR_THREADS) void kg_mark_faces(GView g, const int *leaves, int nleaves) {
  const int serial = g.dser[0];
  const int n = leaves ? nleaves : (int)g.ctr[SGG_NGATHER];
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int l = leaves ? leaves[i] : g.gathered[i];
    const int gb = g.leaf_gbeg[l], ge = g.leaf_gbeg[l + 1];
    for (int p = gb + threadIdx.x; p < ge; p += GR_THREADS) {
      const int f = g.grid_face[g.grid_old[p]];
      if (atomicExch(&g.face_stamp[f], serial) != serial) {
        const int slot = (int)atomicAdd((unsigned long long *)&g.ctr[SGG_NFACES], 1ull);
        g.dfaces[slot] = f;
      }
    }
  }
}

__device__ __forceinline__ int elem_id(const GView &g, int grid, int x, int y) {
  return g.grid_pos[grid] * g.ga + y * g.gs + x;
}

// normal_quad_v3 (math_geom.cc:48-67) of grid quad (qx, qy); P is the grid's
// positions, xyz interleaved, row-major
__device__ __forceinline__ void quad_normal(const float *P, int gs, int qx, int qy, float &nx, float &ny,
                                            float &nz) {
  const float *v1 = P + 3 * ((qy + 1) * gs + qx), *v2 = P + 3 * ((qy + 1) * gs + qx + 1);
  const float *v3 = P + 3 * (qy * gs + qx + 1), *v4 = P + 3 * (qy * gs + qx);
  const float a0 = v1[0] - v3[0], a1 = v1[1] - v3[1], a2 = v1[2] - v3[2];
  const float b0 = v2[0] - v4[0], b1 = v2[1] - v4[1], b2 = v2[2] - v4[2];
  nx = a1 * b2 - a2 * b1;
  ny = a2 * b0 - a0 * b2;
  nz = a0 * b1 - a1 * b0;
  sg_normalize(nx, ny, nz);
}

// the leaf of element v joins this operation's refit (its position changed): listed
// once, after the gathered leaves (which the search marked)
__device__ __forceinline__ void mark_moved(const GView &g, int v, int serial) {
  const int l = g.grid_leaf[v / g.ga];
  if (g.leaf_mark[l] != serial && atomicExch(&g.leaf_mark[l], serial) != serial) {
    const int ng = (int)g.ctr[SGG_NGATHER];
    const int slot = (int)atomicAdd((unsigned long long *)&g.ctr[SGG_NMOVED], 1ull);
    g.gathered[ng + slot] = l;
  }
}

__device__ __forceinline__ void avg_pair(const GView &g, int a, int b, int serial) {  // subdiv_ccg.c:876-900
  const float ax0 = g.px[a], ay0 = g.py[a], az0 = g.pz[a], bx0 = g.px[b], by0 = g.py[b], bz0 = g.pz[b];
  float x = ax0 + bx0, y = ay0 + by0, z = az0 + bz0;
  x *= 0.5f;
  y *= 0.5f;
  z *= 0.5f;
  g.px[a] = g.px[b] = g.qx[a] = g.qx[b] = x;
  g.py[a] = g.py[b] = g.qy[a] = g.qy[b] = y;
  g.pz[a] = g.pz[b] = g.qz[a] = g.qz[b] = z;
  const unsigned xb = __float_as_uint(x), yb = __float_as_uint(y), zb = __float_as_uint(z);
  if (xb != __float_as_uint(ax0) || yb != __float_as_uint(ay0) || zb != __float_as_uint(az0)) mark_moved(g, a, serial);
  if (xb != __float_as_uint(bx0) || yb != __float_as_uint(by0) || zb != __float_as_uint(bz0)) mark_moved(g, b, serial);
  float u = g.nx[a] + g.nx[b], w = g.ny[a] + g.ny[b], t = g.nz[a] + g.nz[b];
  u *= 0.5f;
  w *= 0.5f;
  t *= 0.5f;
  g.nx[a] = g.nx[b] = u;
  g.ny[a] = g.ny[b] = w;
  g.nz[a] = g.nz[b] = t;
  if (g.mask) {
    const float m = (g.mask[a] + g.mask[b]) * 0.5f;
    g.mask[a] = g.mask[b] = m;
  }
}

// the mean of the listed elements written back to each (element_accumulator_*,
// subdiv_ccg.c:902-949); leaves of moved elements marked
__device__ __forceinline__ void avg_list(const GView &g, const int *els, int stride, int n, int serial) {
  float cx = 0.0f, cy = 0.0f, cz = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f, m = 0.0f;
  for (int k = 0; k < n; k++) {
    const int v = els[(size_t)k * stride];
    cx += g.px[v];
    cy += g.py[v];
    cz += g.pz[v];
    nx += g.nx[v];
    ny += g.ny[v];
    nz += g.nz[v];
    if (g.mask) m += g.mask[v];
  }
  const float inv = 1.0f / n;
  cx *= inv;
  cy *= inv;
  cz *= inv;
  nx *= inv;
  ny *= inv;
  nz *= inv;
  m *= inv;
  for (int k = 0; k < n; k++) {
    const int v = els[(size_t)k * stride];
    if (__float_as_uint(g.px[v]) != __float_as_uint(cx) || __float_as_uint(g.py[v]) != __float_as_uint(cy) ||
        __float_as_uint(g.pz[v]) != __float_as_uint(cz))
      mark_moved(g, v, serial);
    g.px[v] = g.qx[v] = cx;
    g.py[v] = g.qy[v] = cy;
    g.pz[v] = g.qz[v] = cz;
    g.nx[v] = nx;
    g.ny[v] = ny;
    g.nz[v] = nz;
    if (g.mask) g.mask[v] = m;
  }
}

// A workgroup per dirty face: its grids' element normals (subdiv_ccg.c:670-740), then
// the seams between its grids and the shared centre (:951-984); its edges and coarse
// vertices are queued once each for the boundary and corner averages.  Leaves of
// elements the averages move are marked for the refit.
#define FN_THREADS 512  // a face's grids (4 x 81 elements) in one pass

__global__ __launch_bounds__(FN_THREADS) void kg_face_normals(GView g, int batch) {
  extern __shared__ float s_pos[];  // `batch` grids' positions, xyz interleaved
  const int serial = g.dser[0];
  const int nf = (int)g.ctr[SGG_NFACES];
  const int gs = g.gs, g1 = gs - 1, ga = g.ga;
  for (int i = blockIdx.x; i < nf; i += gridDim.x) {
    const int f = g.dfaces[i];
    const int g0 = g.face_gbeg[f], n = g.face_gbeg[f + 1] - g0;
    // the face's grids' positions staged in LDS, as many grids at a time as `batch`
    // (one coalesced pass), then their quad normals read them there
    for (int c0 = 0; c0 < n; c0 += batch) {
      const int nc = min(batch, n - c0);
      for (int t = threadIdx.x; t < nc * ga; t += FN_THREADS) {
        const int c = t / ga, local = t - c * ga;
        const int v = g.grid_pos[g0 + c0 + c] * ga + local;
        s_pos[3 * t] = g.px[v];
        s_pos[3 * t + 1] = g.py[v];
        s_pos[3 * t + 2] = g.pz[v];
      }
      __syncthreads();
      for (int t = threadIdx.x; t < nc * ga; t += FN_THREADS) {
        const int c = t / ga, local = t - c * ga;
        const int y = local / gs, x = local - y * gs;
        const float *P = s_pos + 3 * c * ga;
        float ax = 0.0f, ay = 0.0f, az = 0.0f, qx, qy, qz;
        int counter = 0;
        if (x < g1 && y < g1) {
          quad_normal(P, gs, x, y, qx, qy, qz);
          ax += qx;
          ay += qy;
          az += qz;
          counter++;
        }
        if (x >= 1) {
          if (y < g1) {
            quad_normal(P, gs, x - 1, y, qx, qy, qz);
            ax += qx;
            ay += qy;
            az += qz;
            counter++;
          }
          if (y >= 1) {
            quad_normal(P, gs, x - 1, y - 1, qx, qy, qz);
            ax += qx;
            ay += qy;
            az += qz;
            counter++;
          }
        }
        if (y >= 1 && x < g1) {
          quad_normal(P, gs, x, y - 1, qx, qy, qz);
          ax += qx;
          ay += qy;
          az += qz;
          counter++;
        }
        const float inv = 1.0f / (float)counter;
        const int v = g.grid_pos[g0 + c0 + c] * ga + local;
        g.nx[v] = ax * inv;
        g.ny[v] = ay * inv;
        g.nz[v] = az * inv;
      }
      __syncthreads();
    }
    for (int t = threadIdx.x; t < n * g1; t += FN_THREADS) {
      const int c = t / g1, ii = 1 + (t - c * g1);
      const int prev = g0 + (c + n - 1) % n;
      avg_pair(g, elem_id(g, prev, ii, 0), elem_id(g, g0 + c, 0, ii), serial);
    }
    // the shared centre: the grids' centre values gathered in parallel, summed in grid
    // order by one thread, written back in parallel
    __shared__ float s_c[7][FN_THREADS];
    __shared__ float s_mean[7];
    const bool par = n <= FN_THREADS;
    if (par && threadIdx.x < n) {
      const int v = g.grid_pos[g0 + threadIdx.x] * ga;
      s_c[0][threadIdx.x] = g.px[v];
      s_c[1][threadIdx.x] = g.py[v];
      s_c[2][threadIdx.x] = g.pz[v];
      s_c[3][threadIdx.x] = g.nx[v];
      s_c[4][threadIdx.x] = g.ny[v];
      s_c[5][threadIdx.x] = g.nz[v];
      s_c[6][threadIdx.x] = g.mask ? g.mask[v] : 0.0f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      float a[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      for (int c = 0; c < n; c++) {
        const int v = par ? 0 : g.grid_pos[g0 + c] * ga;
        const float val[7] = {par ? s_c[0][c] : g.px[v], par ? s_c[1][c] : g.py[v], par ? s_c[2][c] : g.pz[v],
                              par ? s_c[3][c] : g.nx[v], par ? s_c[4][c] : g.ny[v], par ? s_c[5][c] : g.nz[v],
                              par ? s_c[6][c] : (g.mask ? g.mask[v] : 0.0f)};
        for (int q = 0; q < 7; q++) a[q] += val[q];
      }
      const float inv = 1.0f / n;
      for (int q = 0; q < 7; q++) s_mean[q] = a[q] * inv;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < n; c += FN_THREADS) {
      const int v = g.grid_pos[g0 + c] * ga;
      const float cx = s_mean[0], cy = s_mean[1], cz = s_mean[2];
      if (__float_as_uint(g.px[v]) != __float_as_uint(cx) || __float_as_uint(g.py[v]) != __float_as_uint(cy) ||
          __float_as_uint(g.pz[v]) != __float_as_uint(cz))
        mark_moved(g, v, serial);
      g.px[v] = g.qx[v] = cx;
      g.py[v] = g.qy[v] = cy;
      g.pz[v] = g.qz[v] = cz;
      g.nx[v] = s_mean[3];
      g.ny[v] = s_mean[4];
      g.nz[v] = s_mean[5];
      if (g.mask) g.mask[v] = s_mean[6];
    }
    for (int c = threadIdx.x; c < n; c += FN_THREADS) {
      const int e = g.face_edge[g0 + c], vv = g.face_vert[g0 + c];
      if (atomicExch(&g.edge_stamp[e], serial) != serial) {
        const int slot = (int)atomicAdd((unsigned long long *)&g.ctr[SGG_NITEMS], 1ull);
        g.ditems[slot] = e;
      }
      if (atomicExch(&g.vert_stamp[vv], serial) != serial) {
        const int slot = (int)atomicAdd((unsigned long long *)&g.ctr[SGG_NITEMS], 1ull);
        g.ditems[slot] = g.Ed + vv;
      }
    }
    __syncthreads();
  }
}

// A wave per queued edge or coarse vertex: the mean over its adjacent faces written
// back to each (subdiv_ccg.c:1010-1048 boundaries, :1081-1104 corners); leaves of the
// elements it moves are marked for the refit.
__global__ __launch_bounds__(GR_THREADS) void kg_edges_verts(GView g) {
  const int serial = g.dser[0];
  const int ni = (int)g.ctr[SGG_NITEMS];
  const int lane = lane_id();
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  const int n2 = 2 * g.gs;
  for (int i = w; i < ni; i += nw) {
    const int item = g.ditems[i];
    if (item < g.Ed) {
      const int fb = g.edge_foff[item], nf = g.edge_foff[item + 1] - fb;
      if (nf == 1) continue;
      for (int p = 1 + lane; p < n2 - 1; p += SG_WAVE) avg_list(g, g.edge_el + (size_t)fb * n2 + p, n2, nf, serial);
    } else {
      const int vv = item - g.Ed;
      const int fb = g.vert_foff[vv], nf = g.vert_foff[vv + 1] - fb;
      if (nf > 1 && lane == 0) avg_list(g, g.vert_el + fb, 1, nf, serial);
    }
  }
}


// Leaf refit (update_node_vb, pbvh.c:108-128): vb = min/max over the leaf's elements
// in iteration order.  min_ff / max_ff keep the later of equal values, so any fold that
// combines contiguous pieces left to right gives the sequential result.  The listed
// leaves (the host's list, or the dab's gathered leaves and those its averages moved)
// are cut into GR_SPLIT contiguous parts, a workgroup per part: each wave folds a
// contiguous quarter, 64 consecutive elements per round (one coalesced load per
// coordinate, an ordered DPP scan over the lanes), 4 rounds in flight; the waves
// combine in order into the part's box.
#define GR_SPLIT 8
__global__ __launch_bounds__(GR_THREADS) void kg_refit_parts(GView g, const int *list, int list_n, float *parts) {
  constexpr int NW = GR_THREADS / SG_WAVE;
  __shared__ float s_bb[NW][6];
  const int lane = lane_id(), w = threadIdx.x / SG_WAVE;
  const int nl = list ? list_n : (int)(g.ctr[SGG_NGATHER] + g.ctr[SGG_NMOVED]);
  for (int item = blockIdx.x; item < nl * GR_SPLIT; item += gridDim.x) {
    const int li = item / GR_SPLIT, part = item - li * GR_SPLIT;
    const int l = list ? list[li] : g.gathered[li];
    const int b = g.leaf_gbeg[l] * g.ga, e = g.leaf_gbeg[l + 1] * g.ga;
    const int seg = (e - b + GR_SPLIT - 1) / GR_SPLIT;
    const int pb = min(e, b + part * seg), pe = min(e, pb + seg);
    const int q = (pe - pb + NW - 1) / NW;
    const int wb = min(pe, pb + w * q), we = min(pe, wb + q);
    BB6 acc;
    bb_reset(acc);
    for (int base = wb; base < we; base += SG_WAVE * 4) {
      float x[4], y[4], z[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {  // every load issued (clamped index) before any is folded
        const int v = min(base + r * SG_WAVE + lane, we - 1);
        x[r] = g.px[v];
        y[r] = g.py[v];
        z[r] = g.pz[v];
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        BB6 t;
        bb_reset(t);
        if (base + r * SG_WAVE + lane < we) bb_add_point(t, x[r], y[r], z[r]);
        t = bb_wave_scan(t);
        BB6 tot;
        for (int i = 0; i < 3; i++) {
          tot.mn[i] = __shfl(t.mn[i], SG_WAVE - 1);
          tot.mx[i] = __shfl(t.mx[i], SG_WAVE - 1);
        }
        bb_add_bb(acc, tot);
      }
    }
    if (lane == 0)
      for (int i = 0; i < 3; i++) {
        s_bb[w][i] = acc.mn[i];
        s_bb[w][3 + i] = acc.mx[i];
      }
    __syncthreads();
    if (threadIdx.x == 0) {
      BB6 r;
      bb_reset(r);
      for (int k = 0; k < NW; k++) {
        BB6 xk;
        for (int i = 0; i < 3; i++) {
          xk.mn[i] = s_bb[k][i];
          xk.mx[i] = s_bb[k][3 + i];
        }
        bb_add_bb(r, xk);
      }
      bb_store(parts + 8 * (size_t)item, r);
    }
    __syncthreads();
  }
}

// The listed leaves' boxes from their parts, in order; orig_vb = vb when set_orig; the
// leaves stay dirty for the next internal-node flush.  With `finish` this also closes
// the dab: its counters kept (and logged in the faces counters' layout) and cleared,
// the next serial taken (a captured stroke replays with fresh serials).
__global__ __launch_bounds__(GR_THREADS) void kg_refit_combine(GView g, const int *list, int list_n,
                                                               const float *parts, int set_orig, int finish,
                                                               long long *log) {
  const int t = threadIdx.x;
  const int nl = list ? list_n : (int)(g.ctr[SGG_NGATHER] + g.ctr[SGG_NMOVED]);
  for (int li = t; li < nl; li += blockDim.x) {
    const int l = list ? list[li] : g.gathered[li];
    BB6 r;
    bb_reset(r);
    for (int p = 0; p < GR_SPLIT; p++) bb_add_bb(r, gbb_load(parts + 8 * ((size_t)li * GR_SPLIT + p)));
    bb_store(g.leaf_bb + 8 * (size_t)l, r);
    if (set_orig) bb_store(g.leaf_obb + 8 * (size_t)l, r);
    g.leaf_dirty[l] = 1;
  }
  if (!finish) return;
  __shared__ long long s_c[SGG_CTR_COUNT];
  if (t < SGG_CTR_COUNT) s_c[t] = g.ctr[t];
  __syncthreads();
  if (t < SGG_CTR_COUNT) {
    const long long c = t == SGG_NLEAVES ? nl : s_c[t];
    g.ctr_last[t] = c;
    g.ctr[t] = 0;
  }
  if (log && t < SG_CTR_COUNT) {
    const int src = t == SG_CTR_NGATHER ? SGG_NGATHER : t == SG_CTR_NAFF ? SGG_NAFF : t == SG_CTR_AREA_X ? SGG_AREA_X
                  : t == SG_CTR_AREA_Y ? SGG_AREA_Y : t == SG_CTR_AREA_Z ? SGG_AREA_Z : t == SG_CTR_NAREA ? SGG_NAREA
                  : -1;
    log[t] = src >= 0 ? s_c[src] : 0;
  }
  if (t == 0) {
    g.dser[1] = g.dser[0];
    g.dser[0] = g.dser[0] + 1;
  }
}

__device__ __forceinline__ bool child_dirty(const GView &g, int c) {
  const int lf = g.node_leaf[c];
  return lf >= 0 ? g.leaf_dirty[lf] != 0 : g.node_dirty[c] != 0;
}
__device__ __forceinline__ BB6 child_box(const GView &g, int c) {
  const int lf = g.node_leaf[c];
  return lf >= 0 ? gbb_load(g.leaf_bb + 8 * (size_t)lf) : gbb_load(g.node_bb + 8 * (size_t)c);
}

// One workgroup: internal nodes height by height; a node with a refit child since the
// last flush becomes the union of its two children (pbvh_flush_bb, pbvh.c:1369-1399),
// orig_vb following when set_orig; then every dirty flag is cleared.  The per-dab
// sphere search reads leaf boxes only, so this runs once per stroke or update.
__global__ __launch_bounds__(1024) void kg_flush_nodes(GView g, int set_orig) {
  for (int h = 0; h < g.H; h++) {
    for (int i = g.lvl_off[h] + threadIdx.x; i < g.lvl_off[h + 1]; i += blockDim.x) {
      const int n = g.lvl_node[i];
      const int c0 = g.node_child[n], c1 = c0 + 1;
      if (!child_dirty(g, c0) && !child_dirty(g, c1)) continue;
      BB6 b;
      bb_reset(b);
      bb_add_bb(b, child_box(g, c0));
      bb_add_bb(b, child_box(g, c1));
      bb_store(g.node_bb + 8 * (size_t)n, b);
      if (set_orig) bb_store(g.node_obb + 8 * (size_t)n, b);
      g.node_dirty[n] = 1;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < g.L; i += blockDim.x) g.leaf_dirty[i] = 0;
  for (int i = threadIdx.x; i < g.N; i += blockDim.x) g.node_dirty[i] = 0;
}

// Around a drop-in update (and before a dab after a failed launch): zero the counters
// and take the next serial on the device (a dab ends the same way in kg_refit_combine).
__global__ void kg_clear_ctr(GView g) {
  if (threadIdx.x < SGG_CTR_COUNT) g.ctr[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    g.dser[0] = g.dser[0] + 1;
  }
}

template <class T>
int gupload(T **dst, const T *src, size_t n) {
  if (sg_check_hip(hipMalloc(dst, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc")) return -1;
  if (n && sg_check_hip(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice), "copy")) return -1;
  return 0;
}
template <class T>
int gzeros(T **dst, size_t n) {
  if (sg_check_hip(hipMalloc(dst, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc")) return -1;
  return sg_check_hip(hipMemset(*dst, 0, std::max<size_t>(n, 1) * sizeof(T)), "memset");
}
template <class T>
void gfree(T *&p) {
  if (p) hipFree((void *)p);
  p = nullptr;
}

int elem_blocks(const SgGridDev &d) {
  const long long items = (long long)d.L * (((long long)d.maxleafg * d.ga + GR_CHUNK - 1) / GR_CHUNK);
  return (int)std::max(1ll, std::min(items, 512ll));  // a dab gathers a few leaves: blocks stride the items
}

void free_topology(SgGridDev &d) {
  gfree(d.nb_off);
  gfree(d.nb_idx);
  gfree(d.face_gbeg);
  gfree(d.grid_face);
  gfree(d.face_edge);
  gfree(d.face_vert);
  gfree(d.edge_foff);
  gfree(d.edge_el);
  gfree(d.vert_foff);
  gfree(d.vert_el);
  gfree(d.face_stamp);
  gfree(d.edge_stamp);
  gfree(d.vert_stamp);
  gfree(d.dfaces);
  gfree(d.ditems);
  d.F = d.Ed = d.Vc = 0;
  d.topo_of = nullptr;
  d.topo_version = -1;
}

}  // namespace

// ---------------------------------------------------------------------------
// device layout
// ---------------------------------------------------------------------------
void sg_grids_free_device(PBVH *p) {
  if (!p->gd) return;
  SgGridDev &d = *p->gd;
  free_topology(d);
  float **fl[] = {&d.px, &d.py, &d.pz, &d.qx, &d.qy, &d.qz, &d.nx, &d.ny, &d.nz, &d.mask, &d.fade,
                  &d.leaf_bb, &d.leaf_obb, &d.node_bb, &d.node_obb, &d.curve};
  for (float **q : fl) gfree(*q);
  int **il[] = {&d.stamp, &d.grid_pos, &d.grid_old, &d.grid_leaf, &d.leaf_gbeg, &d.leaf_flags, &d.leaf_mark,
                &d.node_child, &d.node_leaf, &d.node_mark, &d.lvl_off, &d.lvl_node, &d.gathered};
  for (int **q : il) gfree(*q);
  gfree(d.hide);
  gfree(d.ctr);
  gfree(d.ctr_last);
  gfree(d.hlist);
  gfree(d.dser);
  gfree(d.parts);
  gfree(d.leaf_dirty);
  gfree(d.node_dirty);
  if (d.snap) hipFree(d.snap);
  delete p->gd;
  p->gd = nullptr;
}

int sg_grids_build_device(PBVH *p) {
  sg_grids_free_device(p);
  p->gd = new SgGridDev();
  SgGridDev &d = *p->gd;
  const CCGKey &key = p->gridkey;
  d.G = p->totgrid;
  d.gs = key.grid_size;
  d.ga = key.grid_area;
  d.L = (int)p->leaf_node.size();
  d.N = (int)p->nodes.size();
  d.E = (long long)d.G * d.ga;
  if (d.E >= (1ll << 31) - 1) return sg_set_error("grids: %lld elements exceed the 2^31 element ids", d.E);
  // grid order: leaf by leaf (DFS), each leaf's prims in prim_indices order
  std::vector<int> grid_pos(d.G, -1), grid_old, grid_leaf, leaf_gbeg{0}, leaf_flags;
  for (int l = 0; l < d.L; l++) {
    const PBVHNode &n = p->nodes[p->leaf_node[l]];
    for (int i = 0; i < n.totprim; i++) {
      const int g = p->prim_indices[n.prim_off + i];
      grid_pos[g] = (int)grid_old.size();
      grid_old.push_back(g);
      grid_leaf.push_back(l);
    }
    leaf_gbeg.push_back((int)grid_old.size());
    leaf_flags.push_back((n.flag & PBVH_FullyHidden) ? 1 : 0);
    d.maxleafg = std::max(d.maxleafg, n.totprim);
  }
  if ((int)grid_old.size() != d.G) return sg_set_error("grids: the leaves hold %zu of %d grids", grid_old.size(), d.G);
  // elements
  const size_t E = (size_t)d.E;
  std::vector<float> x(E), y(E), z(E), nx(E, 0.0f), ny(E, 0.0f), nz(E, 0.0f), mk;
  std::vector<unsigned char> hide;
  bool any_hidden = false;
  if (key.has_mask) mk.assign(E, 0.0f);
  for (int pos = 0; pos < d.G; pos++) {
    const int g = grid_old[pos];
    const unsigned char *base = (const unsigned char *)p->grids[g];
    const BLI_bitmap *gh = p->grid_hidden ? p->grid_hidden[g] : nullptr;
    for (int j = 0; j < d.ga; j++) {
      const float *el = (const float *)(base + (size_t)j * key.elem_size);
      const size_t v = (size_t)pos * d.ga + j;
      x[v] = el[0];
      y[v] = el[1];
      z[v] = el[2];
      if (key.has_normals) {
        const float *no = (const float *)((const unsigned char *)el + key.normal_offset);
        nx[v] = no[0];
        ny[v] = no[1];
        nz[v] = no[2];
      }
      if (key.has_mask) mk[v] = *(const float *)((const unsigned char *)el + key.mask_offset);
      if (gh && (gh[j >> 5] & (1u << (j & 31)))) {
        if (!any_hidden) hide.assign(E, 0);
        any_hidden = true;
        hide[v] = 1;
      }
    }
  }
  // tree tables: leaf boxes, internal nodes by height
  std::vector<float> lbb((size_t)d.L * 8, 0.0f), lobb((size_t)d.L * 8, 0.0f), nbb((size_t)d.N * 8, 0.0f),
      nobb((size_t)d.N * 8, 0.0f);
  std::vector<int> child(d.N, -1), nleaf(d.N, -1), height(d.N, 0);
  for (int i = 0; i < d.N; i++) {
    const PBVHNode &n = p->nodes[i];
    float *b = (n.leaf >= 0) ? &lbb[8 * (size_t)n.leaf] : &nbb[8 * (size_t)i];
    float *ob = (n.leaf >= 0) ? &lobb[8 * (size_t)n.leaf] : &nobb[8 * (size_t)i];
    for (int q = 0; q < 3; q++) {
      b[q] = n.vb[q];
      b[4 + q] = n.vb[3 + q];
      ob[q] = n.orig_vb[q];
      ob[4 + q] = n.orig_vb[3 + q];
    }
    child[i] = (n.flag & PBVH_Leaf) ? -1 : n.children_offset;
    nleaf[i] = n.leaf;
  }
  for (int i = d.N - 1; i >= 0; i--)  // children have larger indices than parents
    if (child[i] >= 0) height[i] = 1 + std::max(height[child[i]], height[child[i] + 1]);
  int H = 0;
  for (int i = 0; i < d.N; i++) H = std::max(H, height[i]);
  std::vector<int> lvl_off(H + 1, 0), lvl_node;
  for (int h = 1; h <= H; h++) {
    for (int i = 0; i < d.N; i++)
      if (child[i] >= 0 && height[i] == h) lvl_node.push_back(i);
    lvl_off[h] = (int)lvl_node.size();
  }
  d.H = H;
  int rc = 0;
  rc |= gupload(&d.px, x.data(), E);
  rc |= gupload(&d.py, y.data(), E);
  rc |= gupload(&d.pz, z.data(), E);
  rc |= gupload(&d.qx, x.data(), E);
  rc |= gupload(&d.qy, y.data(), E);
  rc |= gupload(&d.qz, z.data(), E);
  rc |= gupload(&d.nx, nx.data(), E);
  rc |= gupload(&d.ny, ny.data(), E);
  rc |= gupload(&d.nz, nz.data(), E);
  if (key.has_mask) rc |= gupload(&d.mask, mk.data(), E);
  if (any_hidden) rc |= gupload(&d.hide, hide.data(), E);
  rc |= gzeros(&d.stamp, E);
  rc |= gzeros(&d.fade, E);
  rc |= gupload(&d.grid_pos, grid_pos.data(), grid_pos.size());
  rc |= gupload(&d.grid_old, grid_old.data(), grid_old.size());
  rc |= gupload(&d.grid_leaf, grid_leaf.data(), grid_leaf.size());
  rc |= gupload(&d.leaf_gbeg, leaf_gbeg.data(), leaf_gbeg.size());
  rc |= gupload(&d.leaf_flags, leaf_flags.data(), leaf_flags.size());
  rc |= gzeros(&d.leaf_mark, (size_t)d.L);
  rc |= gupload(&d.leaf_bb, lbb.data(), lbb.size());
  rc |= gupload(&d.leaf_obb, lobb.data(), lobb.size());
  rc |= gupload(&d.node_child, child.data(), child.size());
  rc |= gupload(&d.node_leaf, nleaf.data(), nleaf.size());
  rc |= gzeros(&d.node_mark, (size_t)d.N);
  rc |= gupload(&d.node_bb, nbb.data(), nbb.size());
  rc |= gupload(&d.node_obb, nobb.data(), nobb.size());
  rc |= gupload(&d.lvl_off, lvl_off.data(), lvl_off.size());
  rc |= gupload(&d.lvl_node, lvl_node.data(), lvl_node.size());
  rc |= gzeros(&d.gathered, (size_t)d.L);
  rc |= gzeros(&d.ctr, (size_t)SGG_CTR_COUNT);
  rc |= gzeros(&d.ctr_last, (size_t)SGG_CTR_COUNT);
  rc |= gzeros(&d.curve, (size_t)(SG_CM_TABLE + 1));
  rc |= gzeros(&d.hlist, (size_t)d.L);
  rc |= gzeros(&d.dser, (size_t)4);
  rc |= gzeros(&d.leaf_dirty, (size_t)d.L);
  rc |= gzeros(&d.parts, (size_t)d.L * GR_SPLIT * 8);
  rc |= gzeros(&d.node_dirty, (size_t)d.N);
  {
    const int ser0[4] = {1, 0, 0, 0};  // the first operation's serial: above every stamp's 0
    rc |= sg_check_hip(hipMemcpy(d.dser, ser0, sizeof(ser0), hipMemcpyHostToDevice), "copy");
  }
  return rc ? -1 : 0;
}

int sg_grids_attach_topology(PBVH *p, const SubdivCCG *s) {
  if (!p->gd) return sg_set_error("grids: no device layout (BKE_pbvh_build_grids first)");
  SgGridDev &d = *p->gd;
  if (d.topo_of == s && d.topo_version == s->topo_version) return 0;
  if (s->num_grids != d.G || s->grid_size != d.gs)
    return sg_set_error("grids: the SubdivCCG has %d grids of size %d, the PBVH %d of size %d", s->num_grids,
                        s->grid_size, d.G, d.gs);
  free_topology(d);
  std::vector<int> grid_pos(d.G);
  if (sg_check_hip(hipMemcpy(grid_pos.data(), d.grid_pos, d.G * sizeof(int), hipMemcpyDeviceToHost), "copy"))
    return -1;
  const int ga = d.ga, gs = d.gs;
  auto el = [&](const SubdivCCGCoord &c) { return grid_pos[c.grid_index] * ga + c.y * gs + c.x; };
  // neighbour CSR in element ids of the device order
  std::vector<long long> off_old;
  std::vector<int> idx_old;
  sg_ccg_neighbor_csr(s, off_old, idx_old);
  const size_t E = (size_t)d.E;
  std::vector<long long> off(E + 1, 0);
  std::vector<int> idx(idx_old.size());
  {
    std::vector<int> old_of(E);
    for (int g = 0; g < d.G; g++)
      for (int j = 0; j < ga; j++) old_of[(size_t)grid_pos[g] * ga + j] = g * ga + j;
    long long w = 0;
    for (size_t v = 0; v < E; v++) {
      const int o = old_of[v];
      for (long long q = off_old[o]; q < off_old[o + 1]; q++) {
        const int u = idx_old[q];
        idx[w++] = grid_pos[u / ga] * ga + (u % ga);
      }
      off[v + 1] = w;
    }
  }
  std::vector<int> face_gbeg(s->num_faces + 1), grid_face(d.G);
  for (int f = 0; f < s->num_faces; f++) face_gbeg[f] = s->faces[f].start_grid_index;
  face_gbeg[s->num_faces] = s->num_grids;
  for (int g = 0; g < d.G; g++) grid_face[g] = (int)(s->grid_faces[g] - s->faces.data());
  std::vector<int> edge_el(s->edge_coords.size()), vert_el(s->vert_coords.size());
  for (size_t i = 0; i < edge_el.size(); i++) edge_el[i] = el(s->edge_coords[i]);
  for (size_t i = 0; i < vert_el.size(); i++) vert_el[i] = el(s->vert_coords[i]);
  d.F = s->num_faces;
  d.Ed = s->num_edges;
  d.Vc = s->num_coarse_verts;
  if ((size_t)d.ga * 12 > 64 * 1024)
    return sg_set_error("grids: a grid of %d elements exceeds the 64 KiB of LDS its normals use", d.ga);
  int rc = 0;
  rc |= gupload(&d.nb_off, off.data(), off.size());
  rc |= gupload(&d.nb_idx, idx.data(), idx.size());
  rc |= gupload(&d.face_gbeg, face_gbeg.data(), face_gbeg.size());
  rc |= gupload(&d.grid_face, grid_face.data(), grid_face.size());
  rc |= gupload(&d.face_edge, s->face_edges.data(), s->face_edges.size());
  rc |= gupload(&d.face_vert, s->face_verts.data(), s->face_verts.size());
  rc |= gupload(&d.edge_foff, s->edge_face_off.data(), s->edge_face_off.size());
  rc |= gupload(&d.edge_el, edge_el.data(), edge_el.size());
  rc |= gupload(&d.vert_foff, s->vert_face_off.data(), s->vert_face_off.size());
  rc |= gupload(&d.vert_el, vert_el.data(), vert_el.size());
  rc |= gzeros(&d.face_stamp, (size_t)d.F);
  rc |= gzeros(&d.edge_stamp, (size_t)d.Ed);
  rc |= gzeros(&d.vert_stamp, (size_t)d.Vc);
  rc |= gzeros(&d.dfaces, (size_t)d.F);
  rc |= gzeros(&d.ditems, (size_t)d.Ed + d.Vc);
  if (rc) return -1;
  d.topo_of = s;
  d.topo_version = s->topo_version;
  return 0;
}

// ---------------------------------------------------------------------------
// enqueue
// ---------------------------------------------------------------------------
static int launch_check(const char *what) { return sg_check_hip(hipGetLastError(), what); }

// normals and seams of the faces of the listed leaves (the gathered ones when null),
// then nothing else: the refit comes after
// normals and seams of the faces of the listed leaves (faces already claimed when
// leaves is null: the brush claims those of the gathered leaves)
static int enqueue_normals(PBVH *p, const GView &g, const int *leaves, int n) {
  const SgGridDev &d = *p->gd;
  if (leaves) kg_mark_faces<<<std::max(1, std::min(n, 1024)), GR_THREADS, 0, p->stream>>>(g, leaves, n);
  const int batch = std::max(1, (32 * 1024) / (d.ga * 12));  // grids staged together: up to 32 KiB of LDS
  kg_face_normals<<<std::max(1, std::min(d.F, 512)), FN_THREADS, (size_t)batch * d.ga * 12, p->stream>>>(g, batch);
  kg_edges_verts<<<std::max(1, std::min((d.Ed + d.Vc + 3) / 4, 256)), GR_THREADS, 0, p->stream>>>(g);
  return launch_check("grid normals");
}

int sg_grids_enqueue_dab(PBVH *p, const SgDabK &k, const float *curve_dev, long long *log, int log_index) {
  if (!p->gd) return sg_set_error("grids: no device layout");
  SgGridDev &d = *p->gd;
  if (!d.topo_of) return sg_set_error("grids: no SubdivCCG attached (DUNE_pbvh_subdiv_cgg_set)");
  if (k.tool != SG_TOOL_DRAW && k.tool != SG_TOOL_INFLATE && k.tool != SG_TOOL_SMOOTH && k.tool != SG_TOOL_GRAB)
    return sg_set_error("grids: tool %d is not supported on PBVH_GRIDS (Draw, Inflate, Smooth, Grab)", k.tool);
  if (k.learned || k.orig_co) return sg_set_error("grids: learned Smooth and original-coordinate Grab are faces only");
  const GView g = view_of(d);
  const int rh = p->respect_hide ? 1 : 0;
  const int eb = elem_blocks(d);
  const bool normals = p->gridkey.has_normals != 0;
  if (d.ctr_unclean) kg_clear_ctr<<<1, 64, 0, p->stream>>>(g);  // a failed launch left the counters
  d.ctr_unclean = true;
  kg_search<<<(d.L + GR_THREADS - 1) / GR_THREADS, GR_THREADS, 0, p->stream>>>(g, k);
  if (k.tool == SG_TOOL_DRAW) kg_area<<<eb, GR_THREADS, 0, p->stream>>>(g, k, rh);
  kg_brush<<<eb, GR_THREADS, 0, p->stream>>>(g, k, curve_dev, rh, normals ? 1 : 0);
  if (k.tool == SG_TOOL_SMOOTH) {
    for (int it = 0; it < k.iters; it++) {
      if (it & 1)
        kg_smooth<<<eb, GR_THREADS, 0, p->stream>>>(g, d.qx, d.qy, d.qz, d.px, d.py, d.pz);
      else
        kg_smooth<<<eb, GR_THREADS, 0, p->stream>>>(g, d.px, d.py, d.pz, d.qx, d.qy, d.qz);
    }
    if (k.iters & 1)
      kg_copy_aff<<<eb, GR_THREADS, 0, p->stream>>>(g, d.qx, d.qy, d.qz, d.px, d.py, d.pz);
    else
      kg_copy_aff<<<eb, GR_THREADS, 0, p->stream>>>(g, d.px, d.py, d.pz, d.qx, d.qy, d.qz);
  }
  if (normals)
    if (int rc = enqueue_normals(p, g, nullptr, 0)) return rc;
  kg_refit_parts<<<std::max(1, std::min(d.L * GR_SPLIT, 512)), GR_THREADS, 0, p->stream>>>(g, nullptr, 0, d.parts);
  kg_refit_combine<<<1, GR_THREADS, 0, p->stream>>>(g, nullptr, 0, d.parts, 1, 1,
                                                    log ? log + (size_t)log_index * SG_CTR_COUNT : nullptr);
  if (int rc = launch_check("grid dab")) return rc;
  d.ctr_unclean = false;
  return 0;
}

int sg_grids_enqueue_flush(PBVH *p) {
  if (!p->gd) return 0;
  kg_flush_nodes<<<1, 1024, 0, p->stream>>>(view_of(*p->gd), 1);
  return launch_check("grid flush");
}

int sg_grids_enqueue_update_normals(PBVH *p, const int *leaves_dev, int n) {
  if (!p->gd || !p->gd->topo_of) return sg_set_error("grids: no SubdivCCG attached");
  SgGridDev &d = *p->gd;
  if (!p->gridkey.has_normals || n <= 0) return 0;
  const GView g = view_of(d);
  kg_clear_ctr<<<1, 64, 0, p->stream>>>(g);
  if (int rc = enqueue_normals(p, g, leaves_dev, n)) return rc;
  kg_clear_ctr<<<1, 64, 0, p->stream>>>(g);
  return launch_check("grid normals");
}

int sg_grids_enqueue_refit(PBVH *p, const int *leaves_dev, int n, int set_orig) {
  if (!p->gd) return sg_set_error("grids: no device layout");
  SgGridDev &d = *p->gd;
  if (n <= 0) return 0;
  const GView g = view_of(d);
  kg_clear_ctr<<<1, 64, 0, p->stream>>>(g);
  kg_refit_parts<<<std::max(1, std::min(n * GR_SPLIT, 512)), GR_THREADS, 0, p->stream>>>(g, leaves_dev, n, d.parts);
  kg_refit_combine<<<1, GR_THREADS, 0, p->stream>>>(g, leaves_dev, n, d.parts, set_orig, 0, nullptr);
  kg_flush_nodes<<<1, 1024, 0, p->stream>>>(g, set_orig);
  kg_clear_ctr<<<1, 64, 0, p->stream>>>(g);
  return launch_check("grid refit");
}

int sg_grids_download(PBVH *p, float *co, float *no, float *mask) {
  const SgGridDev &d = *p->gd;
  const size_t E = (size_t)d.E;
  std::vector<float> a(E), b(E), c(E);
  std::vector<int> grid_old(d.G);
  if (sg_check_hip(hipStreamSynchronize(p->stream), "sync")) return -1;
  if (sg_check_hip(hipMemcpy(grid_old.data(), d.grid_old, d.G * sizeof(int), hipMemcpyDeviceToHost), "copy"))
    return -1;
  auto scatter3 = [&](const float *x, const float *y, const float *z, float *out) -> int {
    if (sg_check_hip(hipMemcpy(a.data(), x, E * 4, hipMemcpyDeviceToHost), "copy")) return -1;
    if (sg_check_hip(hipMemcpy(b.data(), y, E * 4, hipMemcpyDeviceToHost), "copy")) return -1;
    if (sg_check_hip(hipMemcpy(c.data(), z, E * 4, hipMemcpyDeviceToHost), "copy")) return -1;
    for (int pos = 0; pos < d.G; pos++) {
      const size_t src = (size_t)pos * d.ga, dst = (size_t)grid_old[pos] * d.ga;
      for (int j = 0; j < d.ga; j++) {
        out[3 * (dst + j)] = a[src + j];
        out[3 * (dst + j) + 1] = b[src + j];
        out[3 * (dst + j) + 2] = c[src + j];
      }
    }
    return 0;
  };
  if (co && scatter3(d.px, d.py, d.pz, co)) return -1;
  if (no && scatter3(d.nx, d.ny, d.nz, no)) return -1;
  if (mask && d.mask) {
    if (sg_check_hip(hipMemcpy(a.data(), d.mask, E * 4, hipMemcpyDeviceToHost), "copy")) return -1;
    for (int pos = 0; pos < d.G; pos++)
      memcpy(mask + (size_t)grid_old[pos] * d.ga, &a[(size_t)pos * d.ga], d.ga * sizeof(float));
  }
  return 0;
}

int sg_grids_upload(PBVH *p, const float *co, const float *no, const float *mask) {
  const SgGridDev &d = *p->gd;
  const size_t E = (size_t)d.E;
  std::vector<float> a(E), b(E), c(E);
  std::vector<int> grid_old(d.G);
  if (sg_check_hip(hipStreamSynchronize(p->stream), "sync")) return -1;
  if (sg_check_hip(hipMemcpy(grid_old.data(), d.grid_old, d.G * sizeof(int), hipMemcpyDeviceToHost), "copy"))
    return -1;
  auto gather3 = [&](const float *in, float *x, float *y, float *z, float *x2, float *y2, float *z2) -> int {
    for (int pos = 0; pos < d.G; pos++) {
      const size_t dst = (size_t)pos * d.ga, src = (size_t)grid_old[pos] * d.ga;
      for (int j = 0; j < d.ga; j++) {
        a[dst + j] = in[3 * (src + j)];
        b[dst + j] = in[3 * (src + j) + 1];
        c[dst + j] = in[3 * (src + j) + 2];
      }
    }
    float *dsts[6] = {x, y, z, x2, y2, z2};
    const float *srcs[3] = {a.data(), b.data(), c.data()};
    for (int i = 0; i < 6; i++)
      if (dsts[i] && sg_check_hip(hipMemcpy(dsts[i], srcs[i % 3], E * 4, hipMemcpyHostToDevice), "copy")) return -1;
    return 0;
  };
  if (co && gather3(co, d.px, d.py, d.pz, d.qx, d.qy, d.qz)) return -1;
  if (no && gather3(no, d.nx, d.ny, d.nz, nullptr, nullptr, nullptr)) return -1;
  if (mask && d.mask) {
    for (int pos = 0; pos < d.G; pos++)
      memcpy(&a[(size_t)pos * d.ga], mask + (size_t)grid_old[pos] * d.ga, d.ga * sizeof(float));
    if (sg_check_hip(hipMemcpy(d.mask, a.data(), E * 4, hipMemcpyHostToDevice), "copy")) return -1;
  }
  return 0;
}

int sg_grids_read_boxes(PBVH *p, float *leaf_bb, float *node_bb) {
  const SgGridDev &d = *p->gd;
  if (sg_check_hip(hipStreamSynchronize(p->stream), "sync")) return -1;
  if (leaf_bb) {
    if (sg_check_hip(hipMemcpy(leaf_bb, d.leaf_bb, (size_t)d.L * 32, hipMemcpyDeviceToHost), "copy")) return -1;
    if (sg_check_hip(hipMemcpy(leaf_bb + (size_t)d.L * 8, d.leaf_obb, (size_t)d.L * 32, hipMemcpyDeviceToHost),
                     "copy"))
      return -1;
  }
  if (node_bb) {
    if (sg_check_hip(hipMemcpy(node_bb, d.node_bb, (size_t)d.N * 32, hipMemcpyDeviceToHost), "copy")) return -1;
    if (sg_check_hip(hipMemcpy(node_bb + (size_t)d.N * 8, d.node_obb, (size_t)d.N * 32, hipMemcpyDeviceToHost),
                     "copy"))
      return -1;
  }
  return 0;
}

int sg_grids_read_stamps(PBVH *p, int *stamp) {
  const SgGridDev &d = *p->gd;
  if (sg_check_hip(hipStreamSynchronize(p->stream), "sync")) return -1;
  return sg_check_hip(hipMemcpy(stamp, d.stamp, (size_t)d.E * 4, hipMemcpyDeviceToHost), "copy");
}

// Device checkpoint of the mutable state (elements, every box); slot per PBVH.
int sg_grids_snapshot(PBVH *p, int restore) {
  SgGridDev &d = *p->gd;
  const size_t E = (size_t)d.E;
  struct Part {
    void *dev;
    size_t bytes;
  } parts[] = {{d.px, E * 4}, {d.py, E * 4}, {d.pz, E * 4}, {d.nx, E * 4}, {d.ny, E * 4}, {d.nz, E * 4},
               {d.mask, d.mask ? E * 4 : 0}, {d.leaf_bb, (size_t)d.L * 32}, {d.leaf_obb, (size_t)d.L * 32},
               {d.node_bb, (size_t)d.N * 32}, {d.node_obb, (size_t)d.N * 32}};
  const int np = sizeof(parts) / sizeof(parts[0]);
  if (!restore && !d.snap) {
    size_t tot = 0;
    for (const Part &q : parts) tot += q.bytes;
    if (sg_check_hip(hipMalloc(&d.snap, tot), "hipMalloc")) return -1;
  }
  if (!d.snap) return sg_set_error("no grids snapshot saved");
  size_t off = 0;
  for (int i = 0; i < np; i++) {
    if (!parts[i].bytes) continue;
    unsigned char *s = (unsigned char *)d.snap + off;
    if (restore) {
      if (sg_check_hip(hipMemcpyAsync(parts[i].dev, s, parts[i].bytes, hipMemcpyDeviceToDevice, p->stream), "copy"))
        return -1;
    } else {
      if (sg_check_hip(hipMemcpyAsync(s, parts[i].dev, parts[i].bytes, hipMemcpyDeviceToDevice, p->stream), "copy"))
        return -1;
    }
    off += parts[i].bytes;
  }
  if (restore) {  // the shadow copy equals the positions between dabs
    const float *src[3] = {d.px, d.py, d.pz};
    float *dst[3] = {d.qx, d.qy, d.qz};
    for (int i = 0; i < 3; i++)
      if (sg_check_hip(hipMemcpyAsync(dst[i], src[i], E * 4, hipMemcpyDeviceToDevice, p->stream), "copy")) return -1;
  }
  return sg_check_hip(hipStreamSynchronize(p->stream), "snapshot");
}
