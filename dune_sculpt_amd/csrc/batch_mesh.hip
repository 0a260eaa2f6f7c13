// Persistent batch step: one workgroup owns one mesh of the batch and runs every row of
// the loaded [step][mesh] table on it in a single launch.  The meshes are independent,
// so a mesh's dabs need no grid-wide ordering: the phases of a dab are separated by
// workgroup barriers only, and a mesh never waits for another mesh.
//
// Per dab (tool != 0), all over the mesh's own ranges of the arena:
//   1 search, subtree level   a thread per search cut of the mesh; passing cuts listed
//                             in LDS with their leaf ranges
//   2 search, leaf level      a thread per leaf of the passing cuts (flat index); a
//                             passing leaf takes (list slot, first candidate) from one
//                             64-bit LDS counter, so slots and candidate offsets rise
//                             together and a candidate index maps back to its leaf by
//                             binary search
//   3 area normal (Draw)      a thread per candidate (uniq vertex of a gathered leaf)
//   4 brush                   a thread per candidate; affected vertices appended to the
//                             mesh's own list (vertex, Smooth factor)
//   5 Smooth steps 2..k       a thread per affected vertex (held in registers across
//                             the steps when the list fits), then the final iterate
//                             copied into the other array
//   6 normals + leaf refits   a thread per affected vertex; a wave per gathered leaf
//   7 subtree boxes           a wave per passing cut, rebuilt from its leaf boxes
// Each phase's loads issue together (PM_U items per thread per round).  The arithmetic
// is the single-mesh engine's (the same device helpers in the same FP32 order), so each
// mesh ends bit-identical to the same dabs applied to it alone.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sg_batch.h"
#include "sg_device.h"

#define PM_THREADS 512
#define PM_WAVES (PM_THREADS / SG_WAVE)
#define PM_U 4            // items per thread per round
#define PM_CUT_CAP 1024   // passing cuts of one mesh listed in LDS (a mesh has at most this many cuts)
#define PM_G_CAP 1024     // gathered leaves whose records stay in LDS (more: read from memory)
#define PM_R 4            // leaves refit per wave per round
#define PM_PER 4          // points per lane of a leaf refit in the unrolled path

struct PmArgs {
  const int *mesh_cut;   // [M+1] first search cut of each mesh
  const int *mesh_leaf;  // [M+1] first leaf
  const int *mesh_vert;  // [M+1] first vertex
  int *aff_v;            // [V] affected vertices of the running dab, in the mesh's vertex range
  float *aff_f;          // [V] their Smooth factors
  int *g_off;            // [L] first candidate of each gathered leaf, in the mesh's leaf range
  int *err;              // [1] set when a mesh has more cuts than PM_CUT_CAP
  long long *tot;        // [SBT_COUNT] running totals
};

// largest g in [0, ng) with off[g] <= t (off[0] = 0 <= t < off[ng])
__device__ __forceinline__ int pm_find(const int *s_off, const int *g_off, int ng, int t) {
  int lo = 0, hi = ng;
  if (ng <= PM_G_CAP) {
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (s_off[mid] <= t) lo = mid; else hi = mid;
    }
  } else {
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (g_off[mid] <= t) lo = mid; else hi = mid;
    }
  }
  return lo;
}

// leaves_box with the ordered DPP scan: lane 63 holds the folds
__device__ __forceinline__ void pm_leaves_box(const float *bbs, const float *obbs, int lb, int le, BB6 &nb, BB6 &ob) {
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
  nb = bb_wave_scan(nb);
  ob = bb_wave_scan(ob);
}

__global__ __launch_bounds__(PM_THREADS) void k_bmesh(SgBatchView b, PmArgs a, int row0, int steps) {
  __shared__ int s_cut[PM_CUT_CAP], s_clb[PM_CUT_CAP], s_cpre[PM_CUT_CAP + 1];
  __shared__ int s_off[PM_G_CAP + 1], s_vb[PM_G_CAP], s_ve[PM_G_CAP], s_sb[PM_G_CAP], s_se[PM_G_CAP],
      s_leaf[PM_G_CAP];
  __shared__ unsigned long long s_pack;  // gathered leaves << 32 | candidates
  __shared__ int s_ncut, s_naff;
  __shared__ long long s_area[4];
  __shared__ long long s_sum[PM_WAVES][4];
  const SgDevView &d = b.d;
  const int mesh = blockIdx.x, tid = threadIdx.x, lane = lane_id(), w = tid / SG_WAVE;
  const int cb = a.mesh_cut[mesh], ce = a.mesh_cut[mesh + 1];
  const int lbase = a.mesh_leaf[mesh], vbase = a.mesh_vert[mesh];
  int *gathered = d.gathered + lbase;
  int2 *gl_v = d.gl_v + lbase, *gl_s = d.gl_s + lbase;
  int *g_off = a.g_off + lbase;
  int *aff_v = a.aff_v + vbase;
  float *aff_f = a.aff_f + vbase;
  if (ce - cb > PM_CUT_CAP) {
    if (tid == 0) a.err[0] = 1;
    return;
  }
  // per-thread totals: gathered, affected, tests, candidates, vertex->tri degrees,
  // Smooth neighbour reads x k, shared vertices refit, area vertices, Smooth affected x k
  long long t_gather = 0, t_aff = 0, t_tests = 0, t_cand = 0, t_degt = 0, t_nbr = 0, t_sh = 0, t_area = 0, t_ak = 0;
  for (int s = 0; s < steps; s++) {
    const SgDabK k = b.tab[(size_t)(row0 + s) * b.M + mesh];
    if (k.tool == 0) continue;  // uniform: every thread skips
    const bool smooth = k.tool == SG_TOOL_SMOOTH, orig = k.use_orig & 1;
    if (tid == 0) {
      s_ncut = 0;
      s_naff = 0;
      s_pack = 0ull;
      s_area[0] = s_area[1] = s_area[2] = s_area[3] = 0;
    }
    __syncthreads();
    // 1. subtree boxes
    for (int c = cb + tid; c < ce; c += PM_THREADS) {
      const BB6 bx = bb_load((orig ? d.cut_obb : d.cut_bb) + 8 * (size_t)c);
      const int cfl = d.cut_flags[c], lb = d.cut_lbeg[c], le = d.cut_lend[c];
      if (!(cfl & 1) && sphere_box(bx, k.cx, k.cy, k.cz, k.r2)) {
        const int j = atomicAdd(&s_ncut, 1);
        s_cut[j] = c;
        s_clb[j] = lb;
        s_cpre[j] = le - lb;  // leaf count, turned into offsets below
      }
    }
    if (tid == 0) t_tests += ce - cb;
    __syncthreads();
    const int ncut = s_ncut;
    if (w == 0) {  // exclusive offsets of the listed cuts' leaf counts
      int run = 0;
      for (int j0 = 0; j0 < ncut; j0 += SG_WAVE) {
        const int j = j0 + lane;
        const int n = j < ncut ? s_cpre[j] : 0;
        int incl = n;
        for (int o = 1; o < SG_WAVE; o <<= 1) {
          const int up = __shfl_up(incl, o);
          if (lane >= o) incl += up;
        }
        if (j < ncut) s_cpre[j] = run + incl - n;
        run += __shfl(incl, SG_WAVE - 1);
      }
      if (lane == 0) s_cpre[ncut] = run;
    }
    __syncthreads();
    // 2. leaves of the listed cuts
    const int nleaf = s_cpre[ncut];
    const float *leaf_bbs = orig ? d.leaf_obb : d.leaf_bb;
    for (int t0 = tid; t0 < nleaf; t0 += PM_THREADS * PM_U) {
      int l[PM_U];
      BB6 bx[PM_U];
      int fl[PM_U];
      int2 vr[PM_U], sr[PM_U];
#pragma unroll
      for (int u = 0; u < PM_U; u++) {
        const int t = min(t0 + u * PM_THREADS, nleaf - 1);
        int lo = 0, hi = ncut;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_cpre[mid] <= t) lo = mid; else hi = mid;
        }
        l[u] = s_clb[lo] + (t - s_cpre[lo]);
        bx[u] = bb_load(leaf_bbs + 8 * (size_t)l[u]);
        fl[u] = d.leaf_flags[l[u]];
        vr[u] = b.leaf_vr[l[u]];
        sr[u] = b.leaf_shr[l[u]];
      }
#pragma unroll
      for (int u = 0; u < PM_U; u++) {
        if (t0 + u * PM_THREADS < nleaf && !(fl[u] & 1) && sphere_box(bx[u], k.cx, k.cy, k.cz, k.r2)) {
          const int nu = vr[u].y - vr[u].x;
          const unsigned long long old = atomicAdd(&s_pack, (1ull << 32) | (unsigned long long)nu);
          const int g = (int)(old >> 32), off = (int)(old & 0xffffffffull);
          gathered[g] = l[u];
          gl_v[g] = vr[u];
          gl_s[g] = sr[u];
          g_off[g] = off;
          if (g < PM_G_CAP) {
            s_off[g] = off;
            s_vb[g] = vr[u].x;
            s_ve[g] = vr[u].y;
            s_sb[g] = sr[u].x;
            s_se[g] = sr[u].y;
            s_leaf[g] = l[u];
          }
        }
      }
    }
    if (tid == 0) t_tests += nleaf;
    __syncthreads();
    const int ng = (int)(s_pack >> 32), ncand = (int)(s_pack & 0xffffffffull);
    if (tid == 0) {
      if (ng <= PM_G_CAP) s_off[ng] = ncand;
      t_gather += ng;
      t_cand += ncand;
    }
    __syncthreads();
    auto cand_vertex = [&](int t) {
      const int g = pm_find(s_off, g_off, ng, t);
      const int vb = ng <= PM_G_CAP ? s_vb[g] : gl_v[g].x;
      const int o = ng <= PM_G_CAP ? s_off[g] : g_off[g];
      return vb + (t - o);
    };
    // 3. Draw: area normal (2^-32 fixed point, exact in any order)
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    if (k.tool == SG_TOOL_DRAW) {
      long long sx = 0, sy = 0, sz = 0;
      int na = 0;
      for (int t0 = tid; t0 < ncand; t0 += PM_THREADS * PM_U) {
        float px[PM_U], py[PM_U], pz[PM_U], nx[PM_U], ny[PM_U], nz[PM_U];
        bool hid[PM_U];
#pragma unroll
        for (int u = 0; u < PM_U; u++) {
          const int v = cand_vertex(min(t0 + u * PM_THREADS, ncand - 1));
          px[u] = d.px[v];
          py[u] = d.py[v];
          pz[u] = d.pz[v];
          nx[u] = d.nx[v];
          ny[u] = d.ny[v];
          nz[u] = d.nz[v];
          hid[u] = d.vhide && d.respect_hide && d.vhide[v];
        }
#pragma unroll
        for (int u = 0; u < PM_U; u++) {
          if (t0 + u * PM_THREADS >= ncand || hid[u]) continue;
          const float dx = k.cx - px[u];
          const float dy = k.cy - py[u];
          const float dz = k.cz - pz[u];
          const float d2 = dx * dx + dy * dy + dz * dz;
          if (d2 <= k.nr2) {
            sx += __double2ll_rn((double)nx[u] * 4294967296.0);
            sy += __double2ll_rn((double)ny[u] * 4294967296.0);
            sz += __double2ll_rn((double)nz[u] * 4294967296.0);
            na++;
          }
        }
      }
      sx = wave_sum_ll(sx);
      sy = wave_sum_ll(sy);
      sz = wave_sum_ll(sz);
      na = wave_sum_i(na);
      if (lane == 0 && na) {
        atomicAdd((unsigned long long *)&s_area[0], as_u64(sx));
        atomicAdd((unsigned long long *)&s_area[1], as_u64(sy));
        atomicAdd((unsigned long long *)&s_area[2], as_u64(sz));
        atomicAdd((unsigned long long *)&s_area[3], (unsigned long long)na);
      }
      __syncthreads();
      float ax = (float)((double)s_area[0] * (1.0 / 4294967296.0));
      float ay = (float)((double)s_area[1] * (1.0 / 4294967296.0));
      float az = (float)((double)s_area[2] * (1.0 / 4294967296.0));
      sg_normalize(ax, ay, az);
      ox = (ax * k.r) * k.bstr;
      oy = (ay * k.r) * k.bstr;
      oz = (az * k.r) * k.bstr;
      if (tid == 0) t_area += s_area[3];
    }
    // 4. brush
    long long nbr = 0;
    for (int t0 = tid; t0 < ncand; t0 += PM_THREADS * PM_U) {
      BrushIn in[PM_U];
      int vv[PM_U];
#pragma unroll
      for (int u = 0; u < PM_U; u++) {
        const int t = t0 + u * PM_THREADS;
        vv[u] = cand_vertex(min(t, ncand - 1));
        in[u] = brush_load(d, k, vv[u], t < ncand ? vv[u] + 1 : vv[u], smooth);
      }
#pragma unroll
      for (int u = 0; u < PM_U; u++) {
        float fade = 0.0f;
        if (in[u].aff) {
          fade = brush_apply(d, k, nullptr, vv[u], in[u], ox, oy, oz, nbr);
          t_aff++;
          t_degt += in[u].vt.y - in[u].vt.x;
        }
        // list the affected vertices: one LDS reservation per wave
        const unsigned long long m = __ballot(in[u].aff);
        if (m) {
          int base = 0;
          if (lane == 0) base = atomicAdd(&s_naff, __popcll(m));
          base = __shfl(base, 0);
          if (in[u].aff) {
            const int i = base + mask_prefix(m);
            aff_v[i] = vv[u];
            aff_f[i] = fade;
          }
        }
      }
    }
    __syncthreads();
    const int naff = s_naff;
    // 5. Smooth steps 2..k over the affected list (even steps read the shadow copy and
    // write the positions, odd steps the reverse; each step completes before the next),
    // then the final iterate is copied into the other array: odd k ends in the shadow
    // copy, even k in the positions (the single-mesh engine's smooth-sync and post)
    if (smooth) {
      t_nbr += nbr * k.iters;
      if (tid == 0) t_ak += (long long)naff * k.iters;
      // one Jacobi step of list items [i0, i0 + PM_THREADS * PM_U)
      auto jacobi = [&](int it, const int *v, const float *fade, const int2 *vv, int i0) {
        const bool from_shadow = (it % 2) == 0;
        const float *sx_ = from_shadow ? d.qx : d.px;
        const float *sy_ = from_shadow ? d.qy : d.py;
        const float *sz_ = from_shadow ? d.qz : d.pz;
        float *dx_ = from_shadow ? d.px : d.qx;
        float *dy_ = from_shadow ? d.py : d.qy;
        float *dz_ = from_shadow ? d.pz : d.qz;
#pragma unroll
        for (int u = 0; u < PM_U; u++) {
          if (i0 + tid + u * PM_THREADS >= naff) continue;
          const float x = sx_[v[u]], y = sy_[v[u]], z = sz_[v[u]];
          float sx, sy, sz;
          nbr_sum(d.vv_idx, vv[u].x, vv[u].y, sx_, sy_, sz_, sx, sy, sz);
          float rx = x, ry = y, rz = z;
          if (vv[u].y > vv[u].x) {
            const float inv = 1.0f / (float)(vv[u].y - vv[u].x);
            const float avx = sx * inv, avy = sy * inv, avz = sz * inv;
            rx = x + (avx - x) * fade[u];
            ry = y + (avy - y) * fade[u];
            rz = z + (avz - z) * fade[u];
          }
          dx_[v[u]] = rx;
          dy_[v[u]] = ry;
          dz_[v[u]] = rz;
        }
      };
      auto load_items = [&](int i0, int *v, float *fade, int2 *vv) {
#pragma unroll
        for (int u = 0; u < PM_U; u++) {
          const int i = min(i0 + tid + u * PM_THREADS, naff - 1);
          v[u] = aff_v[i];
          fade[u] = aff_f[i];
        }
#pragma unroll
        for (int u = 0; u < PM_U; u++) vv[u] = make_int2(d.vv_off[v[u]], d.vv_off[v[u] + 1]);
      };
      if (naff > 0 && naff <= PM_THREADS * PM_U) {
        // the whole list in registers across the steps, with the first 8 neighbour ids
        // of each vertex: a step is one level of position loads
        int v[PM_U];
        float fade[PM_U];
        int2 vv[PM_U];
        load_items(0, v, fade, vv);
        int nb[PM_U][8];
#pragma unroll
        for (int u = 0; u < PM_U; u++)
#pragma unroll
          for (int j = 0; j < 8; j++) nb[u][j] = d.vv_idx[vv[u].x + j < vv[u].y ? vv[u].x + j : vv[u].x];
        for (int it = 2; it <= k.iters; it++) {
          const bool from_shadow = (it % 2) == 0;
          const float *sx_ = from_shadow ? d.qx : d.px;
          const float *sy_ = from_shadow ? d.qy : d.py;
          const float *sz_ = from_shadow ? d.qz : d.pz;
          float *dx_ = from_shadow ? d.px : d.qx;
          float *dy_ = from_shadow ? d.py : d.qy;
          float *dz_ = from_shadow ? d.pz : d.qz;
#pragma unroll
          for (int u = 0; u < PM_U; u++) {
            if (tid + u * PM_THREADS >= naff) continue;
            const int cnt = vv[u].y - vv[u].x;
            float nx[8], ny[8], nz[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
              nx[j] = sx_[nb[u][j]];
              ny[j] = sy_[nb[u][j]];
              nz[j] = sz_[nb[u][j]];
            }
            const float x = sx_[v[u]], y = sy_[v[u]], z = sz_[v[u]];
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;  // CSR order, as nbr_sum
#pragma unroll
            for (int j = 0; j < 8; j++) {
              if (j < cnt) {
                sx = sx + nx[j];
                sy = sy + ny[j];
                sz = sz + nz[j];
              }
            }
            for (int q = vv[u].x + 8; q < vv[u].y; q++) {
              const int n = d.vv_idx[q];
              sx = sx + sx_[n];
              sy = sy + sy_[n];
              sz = sz + sz_[n];
            }
            float rx = x, ry = y, rz = z;
            if (cnt > 0) {
              const float inv = 1.0f / (float)cnt;
              const float avx = sx * inv, avy = sy * inv, avz = sz * inv;
              rx = x + (avx - x) * fade[u];
              ry = y + (avy - y) * fade[u];
              rz = z + (avz - z) * fade[u];
            }
            dx_[v[u]] = rx;
            dy_[v[u]] = ry;
            dz_[v[u]] = rz;
          }
          __syncthreads();
        }
      } else if (naff > 0) {
        for (int it = 2; it <= k.iters; it++) {
          for (int i0 = 0; i0 < naff; i0 += PM_THREADS * PM_U) {
            int v[PM_U];
            float fade[PM_U];
            int2 vv[PM_U];
            load_items(i0, v, fade, vv);
            jacobi(it, v, fade, vv, i0);
          }
          __syncthreads();
        }
      }
      const bool odd = (k.iters & 1) != 0;
      for (int i = tid; i < naff; i += PM_THREADS) {
        const int v = aff_v[i];
        if (odd) {
          d.px[v] = d.qx[v];
          d.py[v] = d.qy[v];
          d.pz[v] = d.qz[v];
        } else {
          d.qx[v] = d.px[v];
          d.qy[v] = d.py[v];
          d.qz[v] = d.pz[v];
        }
      }
      __syncthreads();
    }
    // 6. normals of the affected vertices; leaf refits (a wave per gathered leaf)
    {
      long long a0 = 0, a1 = 0, a2 = 0;
      for (int i0 = tid; i0 < naff; i0 += PM_THREADS * PM_U) {
        int v[PM_U];
        int2 vt[PM_U];
#pragma unroll
        for (int u = 0; u < PM_U; u++) {
          v[u] = aff_v[min(i0 + u * PM_THREADS, naff - 1)];
          vt[u] = make_int2(d.vt_off[v[u]], d.vt_off[v[u] + 1]);
        }
#pragma unroll
        for (int u = 0; u < PM_U; u++)
          if (i0 + u * PM_THREADS < naff) vertex_normal(d, v[u], vt[u].x, vt[u].y, 0, 0, a0, a1, a2);
      }
      // leaf refits: PM_R leaves per wave per round, every load of the round issued
      // together (the leaf's shared-vertex ids, then the positions), each leaf folded in
      // vert_indices order and combined across lanes in order (leaf_box's fold)
      long long nsh = 0;
      for (int g0 = w; g0 < ng; g0 += PM_WAVES * PM_R) {
        int leaf[PM_R], per[PM_R];
        int2 vr[PM_R], sr[PM_R];
        int vid[PM_R][PM_PER];
#pragma unroll
        for (int r = 0; r < PM_R; r++) {
          const int g = min(g0 + r * PM_WAVES, ng - 1);
          const bool lds = g < PM_G_CAP;
          leaf[r] = lds ? s_leaf[g] : gathered[g];
          vr[r] = lds ? make_int2(s_vb[g], s_ve[g]) : gl_v[g];
          sr[r] = lds ? make_int2(s_sb[g], s_se[g]) : gl_s[g];
          const int nu = vr[r].y - vr[r].x, tot = nu + sr[r].y - sr[r].x;
          per[r] = (tot + SG_WAVE - 1) / SG_WAVE;
#pragma unroll
          for (int c = 0; c < PM_PER; c++) {
            const int i = lane * per[r] + c;
            vid[r][c] = -1;
            if (c < per[r] && i < tot) vid[r][c] = i < nu ? vr[r].x + i : d.leaf_sh[sr[r].x + i - nu];
          }
        }
        float px[PM_R][PM_PER], py[PM_R][PM_PER], pz[PM_R][PM_PER];
#pragma unroll
        for (int r = 0; r < PM_R; r++)
#pragma unroll
          for (int c = 0; c < PM_PER; c++) {
            const int v = vid[r][c] < 0 ? vr[r].x : vid[r][c];
            px[r][c] = d.px[v];
            py[r][c] = d.py[v];
            pz[r][c] = d.pz[v];
          }
#pragma unroll
        for (int r = 0; r < PM_R; r++) {
          if (g0 + r * PM_WAVES >= ng) continue;  // wave-uniform
          if (per[r] > PM_PER) {  // a leaf of more than 64 * PM_PER points: the general fold
            leaf_refit(d, leaf[r], vr[r], sr[r], nsh);
            continue;
          }
          BB6 acc;
          bb_reset(acc);
#pragma unroll
          for (int c = 0; c < PM_PER; c++)
            if (vid[r][c] >= 0) bb_add_point(acc, px[r][c], py[r][c], pz[r][c]);
          acc = bb_wave_scan(acc);
          if (lane == SG_WAVE - 1) {
            bb_store(d.leaf_bb + 8 * (size_t)leaf[r], acc);
            bb_store(d.leaf_obb + 8 * (size_t)leaf[r], acc);
            nsh += sr[r].y - sr[r].x;
          }
        }
      }
      t_sh += nsh;
    }
    __syncthreads();
    // 7. the listed cuts' boxes from their (refit) leaf boxes
    for (int j = w; j < ncut; j += PM_WAVES) {
      const int c = s_cut[j];
      BB6 nb, ob;
      pm_leaves_box(d.leaf_bb, d.leaf_obb, d.cut_lbeg[c], d.cut_lend[c], nb, ob);
      if (lane == SG_WAVE - 1) {
        bb_store(d.cut_bb + 8 * (size_t)c, nb);
        bb_store(d.cut_obb + 8 * (size_t)c, ob);
      }
    }
    __syncthreads();
  }
  // totals of this mesh: one set of atomics per workgroup
  long long v4[4] = {t_aff, t_degt, t_nbr, t_sh};
  for (int i = 0; i < 4; i++) {
    const long long x = wave_sum_ll(v4[i]);
    if (lane == 0) s_sum[w][i] = x;
  }
  __syncthreads();
  if (tid == 0) {
    long long acc[4] = {0, 0, 0, 0};
    for (int q = 0; q < PM_WAVES; q++)
      for (int i = 0; i < 4; i++) acc[i] += s_sum[q][i];
    atomicAdd((unsigned long long *)&a.tot[SBT_AFF], (unsigned long long)acc[0]);
    atomicAdd((unsigned long long *)&a.tot[SBT_DEGT], (unsigned long long)acc[1]);
    atomicAdd((unsigned long long *)&a.tot[SBT_NBR], (unsigned long long)acc[2]);
    atomicAdd((unsigned long long *)&a.tot[SBT_SHARED], (unsigned long long)acc[3]);
    atomicAdd((unsigned long long *)&a.tot[SBT_GATHER], (unsigned long long)t_gather);
    atomicAdd((unsigned long long *)&a.tot[SBT_TESTS], (unsigned long long)t_tests);
    atomicAdd((unsigned long long *)&a.tot[SBT_CAND], (unsigned long long)t_cand);
    atomicAdd((unsigned long long *)&a.tot[SBT_NAREA], (unsigned long long)t_area);
    atomicAdd((unsigned long long *)&a.tot[SBT_AK], (unsigned long long)t_ak);
    if (mesh == 0) atomicAdd((unsigned long long *)&a.tot[SBT_STEPS], (unsigned long long)steps);
  }
}

// the step counters advance as the step graph would have advanced them
__global__ void k_bmesh_advance(int *step, int steps) {
  step[0] += steps;
  step[1] += steps;
}

int sg_batch_enqueue_persistent(SgBatch *b, int row0, int steps) {
  SgBatchView v = sg_batch_view(b);
  PmArgs a;
  a.mesh_cut = b->mesh_cut;
  a.mesh_leaf = b->mesh_leaf;
  a.mesh_vert = b->mesh_vert;
  a.aff_v = b->pm_aff_v;
  a.aff_f = b->pm_aff_f;
  a.g_off = b->pm_g_off;
  a.err = b->pm_err;
  a.tot = b->tot;
  hipLaunchKernelGGL(k_bmesh, dim3(b->M), dim3(PM_THREADS), 0, b->stream, v, a, row0, steps);
  hipLaunchKernelGGL(k_bmesh_advance, dim3(1), dim3(1), 0, b->stream, b->step, steps);
  return sg_check_hip(hipGetLastError(), "persistent batch launch");
}
