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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "sg_internal.h"
#include "sg_kernels.h"
#include "sg_device.h"
#include "sg_learned.h"

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
// K1: search.  A thread per leaf tests the sphere against the leaf's own box: a subtree
// box only prunes (it contains its leaf boxes, so a leaf it rejects fails its own
// test, and a fully hidden subtree has only hidden leaves), so the gathered set is the
// pruned search's, and a passing leaf waits on two levels: its box, then its record
// loaded beside one slot reservation per wave.  Reading every leaf box (32 B each)
// costs bandwidth the pruned search would not spend, but no dependent level.
// The gathered list is unordered (consumers never depend on its order; the API sorts
// it into the reference's left->right order when asked).  A wave after the subtree
// waves logs the previous dab's counters and clears the next dab's counter set.
// ---------------------------------------------------------------------------
// The last `copy_blocks` workgroups of the search bring in the results of the previous
// dab when it was a fused dab (fused kernel below): it wrote its affected vertices'
// positions into the shadow copy and their normals into (mx, my, mz), because its own
// workgroups read those arrays' old values; the copy restores positions == shadow copy.
#define COUNT_UNROLL 8
// A wave's sum of the per-leaf affected counts aff_cnt[0, ng): the first 64 *
// COUNT_UNROLL counts load unconditionally (clamped indices), so they are issued before
// ng (itself a load) is known and the sum waits on one round trip.
struct CountPrefetch {
  int c[COUNT_UNROLL];
};
__device__ __forceinline__ CountPrefetch aff_count_prefetch(const SgDevView &d, int lane) {
  CountPrefetch p;
#pragma unroll
  for (int u = 0; u < COUNT_UNROLL; u++) p.c[u] = d.aff_cnt[min(lane + u * SG_WAVE, d.L - 1)];
  return p;
}
__device__ __forceinline__ int aff_count_sum(const SgDevView &d, const CountPrefetch &p, int lane, int ng) {
  int s = 0;
#pragma unroll
  for (int u = 0; u < COUNT_UNROLL; u++) s += lane + u * SG_WAVE < ng ? p.c[u] : 0;
  for (int g = lane + COUNT_UNROLL * SG_WAVE; g < ng; g += SG_WAVE) s += d.aff_cnt[g];
  return wave_sum_i(s);
}

#define COPY_UNROLL 4
// Caller knows a fused dab's results are pending (its record: pend[0] = its gathered
// count).
__device__ __forceinline__ void fused_copy_back(const SgDevView &d, int tid, int nthreads) {
  // The first COPY_UNROLL slots of each thread load beside the pending record, and all
  // their positions and normals load before any store: unconditional loads from clamped
  // indices (slots below L << aff_shift are allocated; vertex 0 stands in for an unused
  // slot), so the copies wait on two levels.
  const int nslot = d.L << d.aff_shift, msk = (1 << d.aff_shift) - 1;
  int cnt[COPY_UNROLL], vs[COPY_UNROLL];
#pragma unroll
  for (int u = 0; u < COPY_UNROLL; u++) {
    const int i = min(tid + u * nthreads, nslot - 1);
    cnt[u] = d.aff_cnt[i >> d.aff_shift];
    vs[u] = d.aff_v[i];
  }
  const int na = d.pend[0] << d.aff_shift;
  bool ok[COPY_UNROLL];
  float x[COPY_UNROLL], y[COPY_UNROLL], z[COPY_UNROLL], a[COPY_UNROLL], b[COPY_UNROLL], c[COPY_UNROLL];
#pragma unroll
  for (int u = 0; u < COPY_UNROLL; u++) {
    const int i = tid + u * nthreads;
    ok[u] = i < na && (i & msk) < cnt[u];
    const int v = ok[u] ? vs[u] : 0;
    x[u] = d.qx[v];
    y[u] = d.qy[v];
    z[u] = d.qz[v];
    a[u] = d.mx[v];
    b[u] = d.my[v];
    c[u] = d.mz[v];
  }
#pragma unroll
  for (int u = 0; u < COPY_UNROLL; u++) {
    if (!ok[u]) continue;
    const int v = vs[u];
    d.px[v] = x[u];
    d.py[v] = y[u];
    d.pz[v] = z[u];
    d.nx[v] = a[u];
    d.ny[v] = b[u];
    d.nz[v] = c[u];
  }
  for (int i = tid + COPY_UNROLL * nthreads; i < na; i += nthreads) {
    const int v = aff_slot_vertex(d, i, d.aff_cnt[i >> d.aff_shift]);
    if (v < 0) continue;
    const float px = d.qx[v], py = d.qy[v], pz = d.qz[v];
    const float nx = d.mx[v], ny = d.my[v], nz = d.mz[v];
    d.px[v] = px;
    d.py[v] = py;
    d.pz[v] = pz;
    d.nx[v] = nx;
    d.ny[v] = ny;
    d.nz[v] = nz;
  }
}

// Block roles, in grid order:
//   [0, leaf_blocks)          a thread per leaf: box test, then for passing leaves the
//                             records load beside one slot reservation per wave
//   next cut_blocks           a wave per subtree: refresh the box of a subtree the
//                             previous dab changed (for ray casts and the statistics);
//                             the wave after the last subtree does the bookkeeping
//   last copy_blocks          copy-back of a fused previous dab (copy_pending: known
//                             when the launch is recorded)
// No test waits on a subtree box: a leaf's own box decides, so a passing leaf's chain
// is its box, then the reservation.
__global__ __launch_bounds__(256) void k_search(SgDevView d, SgDabK k, long long *log, int log_index,
                                                int serial, const float *place, int leaf_blocks, int cut_blocks,
                                                int copy_blocks, int copy_pending) {
  if (d.halt && d.halt[0]) return;
  const int lane = lane_id();
  // Block roles in launch order: subtree waves, copy blocks, then leaves.  The
  // dispatcher starts blocks in index order, and the subtree refreshes and the copy are
  // the longest blocks, so they start first.  vb: the block's number in the order
  // leaves, subtrees, copy.
  const int bx = (int)blockIdx.x;
  const int vb = bx < cut_blocks ? leaf_blocks + bx
                                 : (bx < cut_blocks + copy_blocks ? leaf_blocks + bx : bx - cut_blocks - copy_blocks);
  if (vb >= leaf_blocks + cut_blocks) {
    const int cb = vb - (leaf_blocks + cut_blocks);
    if (copy_pending) fused_copy_back(d, cb * blockDim.x + threadIdx.x, copy_blocks * blockDim.x);
    return;
  }
  sg_place(k, place);
  const float *leaf_bbs = k.use_orig ? d.leaf_obb : d.leaf_bb;
  if (vb < leaf_blocks) {
    const int l = vb * blockDim.x + threadIdx.x;
    const int lc = min(l, d.L - 1);
    const BB6 b = bb_load(leaf_bbs + 8 * (size_t)lc);
    const int fl = k.hide_skip ? d.leaf_flags[lc] : 0;
    const bool pass = l < d.L && !(fl & k.hide_skip) && sphere_box(b, k.cx, k.cy, k.cz, k.r2);
    const unsigned long long m = __ballot(pass);
    if (m == 0) return;
    // the passing leaves' records and subtrees load beside the reservation (one round
    // trip for both); every lane loads, the passing ones use them
    int4 r0, r1;
    if (d.hl_rec) {
      r0 = d.hl_rec[2 * lc];
      r1 = d.hl_rec[2 * lc + 1];
    } else {
      r0 = make_int4(d.leaf_vbeg[lc], d.leaf_sh_off[lc], 0, 0);
      r1 = make_int4((d.leaf_vbeg[lc + 1] - d.leaf_vbeg[lc]) | ((d.leaf_sh_off[lc + 1] - d.leaf_sh_off[lc]) << 16), 0,
                     0, 0);
    }
    const int cut = d.leaf_cut[lc];
    const int stamp = d.dstamp ? d.dstamp[serial & 1] : 0;
    int wbase = 0;
    if (lane == 0) wbase = (int)atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NGATHER], (unsigned long long)__popcll(m));
    wbase = __shfl(wbase, 0);
    if (pass) {
      const int pos = wbase + mask_prefix(m);
      d.gathered[pos] = l;
      d.gl_v[pos] = make_int2(r0.x, r0.x + (r1.x & 0xffff));
      d.gl_s[pos] = make_int2(r0.y, r0.y + (r1.x >> 16));
      if (k.fused) {
        d.gl_rec[2 * pos] = r0;
        d.gl_rec[2 * pos + 1] = r1;
        d.leaf_gstamp[l] = stamp;
      }
      // the subtree's flags (every passing leaf of it writes the same values)
      d.cut_touch[cut] = serial;
      d.cut_dirty[cut] = 1;
      d.draw_dirty[cut] = 1;  // the subtree's leaves need new draw records (UpdateDrawBuffers)
    }
    return;
  }
  // subtree waves
  const int c = (vb - leaf_blocks) * (blockDim.x / SG_WAVE) + threadIdx.x / SG_WAVE;
  if (c == d.C) {  // bookkeeping
    const int stamp = d.dstamp ? d.dstamp[serial & 1] : 0;
    if (log && log_index > 0) {
      // a fused previous dab left its affected count as per-leaf counts: sum them
      long long naff = -1;
      if (copy_pending) {
        const CountPrefetch cp = aff_count_prefetch(d, lane);
        naff = aff_count_sum(d, cp, lane, d.pend[0]);
      }
      if (lane < SG_CTR_COUNT) {
        long long prev = d.ctr_next[lane];  // the set the previous dab used
        if (lane == SG_CTR_NAFF && naff >= 0) prev = naff;
        log[(size_t)(log_index - 1) * SG_CTR_COUNT + lane] = lane == SG_CTR_NGATHER ? (prev & 0xffffffffll) : prev;
      }
    }
    if (lane < SG_CTR_COUNT) d.ctr_next[lane] = 0;
    if (lane == 0 && d.seq) d.seq[0] += 1;                                 // this dab's stamp for the later kernels
    if (lane == 0 && d.dstamp) d.dstamp[(serial & 1) ^ 1] = stamp + 1;   // the next dab's
    if (lane == 0 && undo_active(d)) atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NGATHER], SG_NGATHER_UNDO);
    return;
  }
  if (c > d.C) return;
  // The previous dab changed this subtree (its leaves' boxes moved), so its box is
  // rebuilt here from its leaf boxes, both the current and the original one.  With the
  // statistics on, the wave also counts the tests a pruned search makes: the subtree
  // box, and every leaf of a passing or refreshed subtree.
  const int touch = d.cut_touch[c];
  const int lb = d.cut_lbeg[c], le = d.cut_lend[c];
  const bool refresh = touch == serial - 1;
  long long ntests = 0;
  if (k.stats) {
    const BB6 cb = bb_load((k.use_orig ? d.cut_obb : d.cut_bb) + 8 * (size_t)c);
    const int cfl = d.cut_flags[c];
    const bool hit = !(cfl & k.hide_skip) && sphere_box(cb, k.cx, k.cy, k.cz, k.r2);
    ntests = 1 + ((refresh || hit) ? le - lb : 0);
  }
  if (refresh) {
    BB6 tb, ob;
    bb_reset(tb);
    bb_reset(ob);
    for (int q0 = lb; q0 < le; q0 += SG_WAVE * SEARCH_UNROLL) {
      BB6 nb[SEARCH_UNROLL], xb[SEARCH_UNROLL];
#pragma unroll
      for (int j = 0; j < SEARCH_UNROLL; j++) {
        const int l = min(q0 + j * SG_WAVE + lane, le - 1);
        nb[j] = bb_load(d.leaf_bb + 8 * (size_t)l);
        xb[j] = bb_load(d.leaf_obb + 8 * (size_t)l);
      }
#pragma unroll
      for (int j = 0; j < SEARCH_UNROLL; j++) {
        if (q0 + j * SG_WAVE + lane < le) {
          bb_add_bb(tb, nb[j]);
          bb_add_bb(ob, xb[j]);
        }
      }
    }
    tb = bb_wave_scan(tb);
    ob = bb_wave_scan(ob);
    if (lane == SG_WAVE - 1) {
      bb_store(d.cut_bb + 8 * (size_t)c, tb);
      bb_store(d.cut_obb + 8 * (size_t)c, ob);
    }
  }
  if (k.stats && lane == 0) atomicAdd((unsigned long long *)&d.ctr[SG_CTR_TESTS], (unsigned long long)ntests);
}

// zero both counter sets and the halt flag at the start of a stroke; the dab stamps
// only grow (the first dab's stamp is above every earlier one)
__global__ void k_clear_ctr(long long *ctr, int *halt, int *dstamp) {
  if (threadIdx.x < 2 * SG_CTR_COUNT) ctr[threadIdx.x] = 0;
  if (threadIdx.x == 0) halt[0] = 0;
  if (threadIdx.x == 0 && dstamp) dstamp[0] = max(dstamp[0], dstamp[1]) + 1;
}

// ---------------------------------------------------------------------------
// K2 (Draw only): area normal.  Sum of the normals of uniq verts of gathered leaves
// within normal_radius_factor * r, each component rounded to a multiple of 2^-32
// and accumulated in int64 -- exact, so the atomic order cannot change the bits.
// Positions and normals are fetched in the same load level.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_area(SgDevView d, SgDabK k, const float *place) {
  if (d.halt && d.halt[0]) return;
  sg_place(k, place);
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  const int2 vr0 = gw < d.L ? d.gl_v[gw] : make_int2(0, 0);
  const int ng = (int)d.ctr[SG_CTR_NGATHER];
  long long sx = 0, sy = 0, sz = 0;
  int na = 0;
  for (int g = gw; g < ng; g += nw) {
    const int2 vr = (g == gw) ? vr0 : d.gl_v[g];
    for (int v = vr.x + lane; v < vr.y; v += SG_WAVE) {
      const float px = d.px[v], py = d.py[v], pz = d.pz[v];
      const float nx = d.nx[v], ny = d.ny[v], nz = d.nz[v];
      if (d.vhide && d.respect_hide && d.vhide[v]) continue;
      const float dx = k.cx - px;
      const float dy = k.cy - py;
      const float dz = k.cz - pz;
      const float d2 = dx * dx + dy * dy + dz * dz;
      if (d2 <= k.nr2) {
        sx += __double2ll_rn((double)nx * 4294967296.0);
        sy += __double2ll_rn((double)ny * 4294967296.0);
        sz += __double2ll_rn((double)nz * 4294967296.0);
        na++;
      }
    }
  }
  sx = wave_sum_ll(sx);
  sy = wave_sum_ll(sy);
  sz = wave_sum_ll(sz);
  na = wave_sum_i(na);
  // one set of atomics per workgroup (the four sums of its waves added in LDS first):
  // every wave's atomics go to the same four counters
  __shared__ long long s_part[4][4];
  const int w = threadIdx.x / SG_WAVE;
  if (lane == 0) {
    s_part[w][0] = sx;
    s_part[w][1] = sy;
    s_part[w][2] = sz;
    s_part[w][3] = na;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t[4] = {0, 0, 0, 0};
    for (int i = 0; i < (int)(blockDim.x / SG_WAVE); i++)
      for (int q = 0; q < 4; q++) t[q] += s_part[i][q];
    if (t[3]) {
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_AREA_X], as_u64(t[0]));
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_AREA_Y], as_u64(t[1]));
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_AREA_Z], as_u64(t[2]));
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NAREA], (unsigned long long)t[3]);
    }
  }
}

// ---------------------------------------------------------------------------
// K3: brush.  One wave per gathered leaf (grid-stride); two vertices per lane, so a
// leaf of up to 128 uniq vertices is one load level and one ballot pair.  Spec S2
// vertex test, S3/S4 factor, S6-S9 displacement.  Affected vertices are compacted with
// ballot + prefix popcount into records (vertex, its vertex->tri and vertex->vertex CSR
// ranges) in the gathered leaf's own slots of the list, with the leaf's count beside
// them, so the later kernels skip those lookups and no wave waits on a shared counter.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_brush(SgDevView d, SgDabK k, const float *curve, int stamp,
                                               const float *place) {
  if (d.halt && d.halt[0]) return;
  sg_place(k, place);
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  const int2 vr0 = gw < d.L ? d.gl_v[gw] : make_int2(0, 0);
  const long long ng_ctr = d.ctr[SG_CTR_NGATHER];
  const int ng = (int)ng_ctr;
  const bool smooth = k.tool == SG_TOOL_SMOOTH;
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
  if (ng_ctr & SG_NGATHER_UNDO) {
    // an undo step is open: save this wave's leaves before it writes them (a leaf's
    // positions and shadow copy are written only by the wave that owns it, so both
    // still hold the values before the dab); a loop of its own keeps the brush loop's
    // registers as they are without it
    for (int g = gw; g < ng; g += nw) undo_save_leaf(d, d.gathered[g], (g == gw) ? vr0 : d.gl_v[g], d.qx, d.qy, d.qz);
  }
  for (int g = gw; g < ng; g += nw) {
    const int2 vr = (g == gw) ? vr0 : d.gl_v[g];
    ncand += vr.y - vr.x;
    // the leaf's affected records go to its own slots: no reservation to wait for
    int used = 0;
    for (int base = vr.x; base < vr.y; base += 2 * SG_WAVE) {
      const int v0 = base + lane, v1 = base + SG_WAVE + lane;
      const BrushIn in0 = brush_load(d, k, v0, vr.y, smooth);
      const BrushIn in1 = brush_load(d, k, v1, vr.y, smooth);
      const unsigned long long m0 = __ballot(in0.aff);
      const unsigned long long m1 = __ballot(in1.aff);
      const int c0 = __popcll(m0), cnt = c0 + __popcll(m1);
      if (cnt == 0) continue;
      const int wbase = (g << d.aff_shift) + used;
      used += cnt;
      float fade0 = 0.0f, fade1 = 0.0f;
      if (in0.aff) fade0 = brush_apply(d, k, curve, v0, in0, ox, oy, oz, nbr);
      if (in1.aff) fade1 = brush_apply(d, k, curve, v1, in1, ox, oy, oz, nbr);
      if (k.stats) {
        if (in0.aff) d.stamp_a[v0] = stamp;
        if (in1.aff) d.stamp_a[v1] = stamp;
      }
      if (in0.aff) {
        const int pos = wbase + mask_prefix(m0);
        d.aff_v[pos] = v0;
        d.aff_vt[pos] = in0.vt;
        if (smooth) {
          d.aff_vv[pos] = in0.vv;
          d.aff_fade[pos] = fade0;
        }
      }
      if (in1.aff) {
        const int pos = wbase + c0 + mask_prefix(m1);
        d.aff_v[pos] = v1;
        d.aff_vt[pos] = in1.vt;
        if (smooth) {
          d.aff_vv[pos] = in1.vv;
          d.aff_fade[pos] = fade1;
        }
      }
    }
    if (lane == 0) d.aff_cnt[g] = used;
  }
  if (k.stats) {
    ncand = (lane == 0) ? ncand : 0;
    ncand = wave_sum_i(ncand);
    nbr = wave_sum_ll(nbr);
    if (lane == 0) {
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NCAND], (unsigned long long)ncand);
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_NBR], as_u64(nbr));
    }
  }
}

// Grab from the original coordinates (SG_DAB_ORIGINAL).  The reference's Grab tests and
// displaces each vertex from its original coordinate (SculptOrigVertData, the undo
// node's copy) into a proxy, and the proxy combine then writes original + proxy to
// every visible uniq vertex of the node (use_orco), so vertices outside the sphere go
// back to their original positions.  A wave per gathered leaf; the original
// coordinates are the open undo step's record of the leaf (saved here first when this
// dab is the step's first to touch it), or the current positions when no step is
// open.  Every written vertex joins the affected list, so its normal is recomputed.
__global__ __launch_bounds__(256) void k_brush_orig(SgDevView d, SgDabK k, const float *curve, int stamp,
                                                    const float *place) {
  if (d.halt && d.halt[0]) return;
  sg_place(k, place);
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  const long long ng_ctr = d.ctr[SG_CTR_NGATHER];
  const int ng = (int)ng_ctr;
  const bool rec = (ng_ctr & SG_NGATHER_UNDO) != 0;
  for (int g = gw; g < ng; g += nw) {
    const int leaf = d.gathered[g];
    const int2 vr = d.gl_v[g];
    const float *orig = nullptr;
    if (rec) {
      const SgUndoCtl *u = d.undo;
      long long vb;
      if (u->stamp[leaf] == u->step) {  // saved by an earlier kernel of the step
        const int4 r = u->tab[u->slot[leaf]];
        vb = (long long)(((unsigned long long)(unsigned)r.w << 32) | (unsigned)r.z);
      } else {
        vb = undo_save_leaf(d, leaf, vr, d.px, d.py, d.pz);  // -1: arena full
      }
      if (vb >= 0) orig = u->pos + 3 * (size_t)vb;  // read back lane by lane: same lane wrote it
    }
    int used = 0;
    for (int base = vr.x; base < vr.y; base += SG_WAVE) {
      const int v = base + lane;
      const bool vis = v < vr.y && !(d.vhide && d.respect_hide && d.vhide[v]);
      const unsigned long long m = __ballot(vis);
      if (m == 0ull) continue;
      const int wbase = (g << d.aff_shift) + used;
      used += __popcll(m);
      if (!vis) continue;
      float ox, oy, oz;
      if (orig) {
        const size_t i = 3 * (size_t)(v - vr.x);
        ox = orig[i];
        oy = orig[i + 1];
        oz = orig[i + 2];
      } else {
        ox = d.px[v];
        oy = d.py[v];
        oz = d.pz[v];
      }
      const float dx = k.cx - ox, dy = k.cy - oy, dz = k.cz - oz;
      const float d2 = dx * dx + dy * dy + dz * dz;
      float rx = ox, ry = oy, rz = oz;
      if (d2 <= k.r2) {
        const float f = sg_falloff(k, curve, sqrtf(d2)) * (1.0f - (d.mask ? d.mask[v] : 0.0f));
        const float fade = k.bstr * f;
        rx = ox + k.gx * fade;
        ry = oy + k.gy * fade;
        rz = oz + k.gz * fade;
      }
      d.px[v] = rx;
      d.py[v] = ry;
      d.pz[v] = rz;
      d.qx[v] = rx;
      d.qy[v] = ry;
      d.qz[v] = rz;
      const int pos = wbase + mask_prefix(m);
      d.aff_v[pos] = v;
      d.aff_vt[pos] = make_int2(d.vt_off[v], d.vt_off[v + 1]);
      if (k.stats) d.stamp_a[v] = stamp;
    }
    if (lane == 0) d.aff_cnt[g] = used;
  }
}

// ---------------------------------------------------------------------------
// Elastic deform (SCULPT_TOOL_ELASTIC_DEFORM).  The Kelvinlet displacements restate
// kernel/intern/kelvinlet.c (BKE_kelvinlet_init_params :6-21, init_kelvinlet_grab
// :23-45, grab / bi-scale / tri-scale :47-109, the integration :114-138, scale :141-155
// and twist :167-182) operation for operation.  The brush task is absent from the
// reference; it follows Blender's published one: every node, each visible uniq vertex
// displaced from its original coordinate, the grab modes scaled by bstrength * 20, the
// result scaled by 1 - mask and combined as original + displacement.
// ---------------------------------------------------------------------------
struct Kelvinlet {
  float a, b, c, f, rs[3];
};

__device__ __forceinline__ Kelvinlet kelvinlet_params(float radius, float force, float shear_modulus,
                                                     float poisson_ratio) {
  Kelvinlet p;
  p.a = 1.0f / (4.0f * (float)M_PI * shear_modulus);
  p.b = p.a / (4.0f * (1.0f - poisson_ratio));
  p.c = 2.0f * (3.0f * p.a - 2.0f * p.b);
  p.f = force;
  p.rs[0] = radius * 1.0f;
  p.rs[1] = p.rs[0] * 2.0f;
  p.rs[2] = p.rs[1] * 2.0f;
  return p;
}

__device__ __forceinline__ float kl_p2(float x) { return x * x; }
__device__ __forceinline__ float kl_p3(float x) { return kl_p2(x) * x; }
__device__ __forceinline__ float kl_p5(float x) { return kl_p2(kl_p2(x)) * x; }

// Formula (6) for the first n radii
__device__ __forceinline__ void kelvinlet_grab_terms(const Kelvinlet &p, float radius, int n, float kv[3]) {
  for (int i = 0; i < n; i++) {
    const float re = sqrtf(kl_p2(radius) + kl_p2(p.rs[i]));
    kv[i] = ((p.a - p.b) / re) + ((p.b * kl_p2(radius)) / kl_p3(re)) + ((p.a * kl_p2(p.rs[i])) / (2.0f * kl_p3(re)));
  }
}

// grab, bi-scale or tri-scale: the displacement factor along brush_delta
__device__ __forceinline__ float kelvinlet_grab_fade(const Kelvinlet &p, int mode, float radius) {
  float kv[3];
  if (mode == SG_ELASTIC_GRAB) {
    kelvinlet_grab_terms(p, radius, 1, kv);
    return kv[0] * p.c;
  }
  if (mode == SG_ELASTIC_GRAB_BISCALE) {
    kelvinlet_grab_terms(p, radius, 2, kv);
    const float u = kv[0] - kv[1];
    return u * p.c / ((1.0f / p.rs[0]) - (1.0f / p.rs[1]));
  }
  kelvinlet_grab_terms(p, radius, 3, kv);
  const float w0 = 1.0f;
  const float w1 = -((kl_p2(p.rs[2]) - kl_p2(p.rs[0])) / (kl_p2(p.rs[2]) - kl_p2(p.rs[1])));
  const float w2 = ((kl_p2(p.rs[1]) - kl_p2(p.rs[0])) / (kl_p2(p.rs[2]) - kl_p2(p.rs[1])));
  const float u = w0 * kv[0] + w1 * kv[1] + w2 * kv[2];
  return u * p.c / (w0 / p.rs[0] + w1 / p.rs[1] + w2 / p.rs[2]);
}

// Formula (16) (scale) or (15) (twist) at vertex co
__device__ __forceinline__ void kelvinlet_scale_twist(const Kelvinlet &p, bool twist, const float co[3],
                                                      const float loc[3], const float nor[3], float disp[3]) {
  const float rv[3] = {co[0] - loc[0], co[1] - loc[1], co[2] - loc[2]};
  const float radius = sqrtf(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
  const float re = sqrtf(kl_p2(radius) + kl_p2(p.rs[0]));
  const float lead = twist ? -p.a : (2.0f * p.b - p.a);
  const float u = lead * ((1.0f / kl_p3(re))) + ((3.0f * kl_p2(p.rs[0])) / (2.0f * kl_p5(re)));
  const float fade = u * p.c;
  const float s = fade * p.f;
  if (twist) {
    const float q[3] = {nor[1] * rv[2] - nor[2] * rv[1], nor[2] * rv[0] - nor[0] * rv[2],
                        nor[0] * rv[1] - nor[1] * rv[0]};
    disp[0] = q[0] * s;
    disp[1] = q[1] * s;
    disp[2] = q[2] * s;
  } else {
    disp[0] = rv[0] * s;
    disp[1] = rv[1] * s;
    disp[2] = rv[2] * s;
  }
}

// the four-stage integration of scale / twist (the fourth stage offsets by -location
// as the reference does)
__device__ __forceinline__ void kelvinlet_integrate(const Kelvinlet &p, bool twist, const float co[3],
                                                    const float loc[3], const float nor[3], float r[3]) {
  float k0[3], k1[3], k2[3], k3[3], it[3];
  kelvinlet_scale_twist(p, twist, co, loc, nor, k0);
  for (int i = 0; i < 3; i++) it[i] = co[i] + k0[i] * 0.5f;
  kelvinlet_scale_twist(p, twist, it, loc, nor, k1);
  for (int i = 0; i < 3; i++) it[i] = co[i] + k1[i] * 0.5f;
  kelvinlet_scale_twist(p, twist, it, loc, nor, k2);
  for (int i = 0; i < 3; i++) it[i] = (co[i] + k2[i]) - loc[i];
  kelvinlet_scale_twist(p, twist, it, loc, nor, k3);
  const float sixth = 1.0f / 6.0f;
  for (int i = 0; i < 3; i++) {
    float v = k0[i];
    v += k1[i] * 2.0f;
    v += k2[i] * 2.0f;
    v += k3[i];
    r[i] = v * sixth;
  }
}

// A wave per leaf over every leaf.  Original coordinates as k_brush_orig: the open undo
// step's record (saved here first), else the current positions.
__global__ __launch_bounds__(256) void k_elastic(SgDevView d, SgDabK k, const float *place, long long nwritten) {
  if (d.halt && d.halt[0]) return;
  sg_place(k, place);
  const int lane = lane_id();
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  const bool rec = undo_active(d);
  const Kelvinlet prm = kelvinlet_params(k.r, k.eforce, 1.0f, k.epois);
  const float loc[3] = {k.cx, k.cy, k.cz}, nor[3] = {k.enx, k.eny, k.enz};
  const float delta[3] = {k.gx, k.gy, k.gz};
  const bool grab = k.etype <= SG_ELASTIC_GRAB_TRISCALE;
  const float gscale = k.bstr * 20.0f;
  // the written count is known on the host (every visible uniq vertex)
  if (blockIdx.x == 0 && threadIdx.x == 0) d.ctr[SG_CTR_NAFF] = nwritten;
  for (int leaf = gw; leaf < d.L; leaf += nw) {
    const int2 vr = make_int2(d.leaf_vbeg[leaf], d.leaf_vbeg[leaf + 1]);
    const float *orig = nullptr;
    if (rec) {
      const SgUndoCtl *u = d.undo;
      long long vb;
      if (u->stamp[leaf] == u->step) {
        const int4 r = u->tab[u->slot[leaf]];
        vb = (long long)(((unsigned long long)(unsigned)r.w << 32) | (unsigned)r.z);
      } else {
        vb = undo_save_leaf(d, leaf, vr, d.px, d.py, d.pz);
      }
      if (vb >= 0) orig = u->pos + 3 * (size_t)vb;  // same lane wrote each record it reads
    }
    for (int v = vr.x + lane; v < vr.y; v += SG_WAVE) {
      if (d.vhide && d.respect_hide && d.vhide[v]) continue;
      float co[3];
      if (orig) {
        const size_t i = 3 * (size_t)(v - vr.x);
        co[0] = orig[i];
        co[1] = orig[i + 1];
        co[2] = orig[i + 2];
      } else {
        co[0] = d.px[v];
        co[1] = d.py[v];
        co[2] = d.pz[v];
      }
      float disp[3];
      if (grab) {
        const float e[3] = {co[0] - loc[0], co[1] - loc[1], co[2] - loc[2]};
        const float radius = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        const float fade = kelvinlet_grab_fade(prm, k.etype, radius);
        for (int i = 0; i < 3; i++) disp[i] = delta[i] * fade;
        for (int i = 0; i < 3; i++) disp[i] = disp[i] * gscale;
      } else {
        kelvinlet_integrate(prm, k.etype == SG_ELASTIC_TWIST, co, loc, nor, disp);
      }
      if (d.mask) {
        const float keep = 1.0f - d.mask[v];
        for (int i = 0; i < 3; i++) disp[i] = disp[i] * keep;
      }
      const float rx = co[0] + disp[0], ry = co[1] + disp[1], rz = co[2] + disp[2];
      d.px[v] = rx;
      d.py[v] = ry;
      d.pz[v] = rz;
      d.qx[v] = rx;
      d.qy[v] = ry;
      d.qz[v] = rz;
    }
  }
}

// Normals of every visible uniq vertex (the combine flags them all), spec S10.
__global__ __launch_bounds__(256) void k_elastic_normals(SgDevView d) {
  const int nu = d.leaf_vbeg[d.L];
  long long a = 0, b = 0, c = 0;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nu; v += gridDim.x * blockDim.x) {
    if (d.vhide && d.respect_hide && d.vhide[v]) continue;
    vertex_normal(d, v, d.vt_off[v], d.vt_off[v + 1], 0, 0, a, b, c);
  }
}

// ---------------------------------------------------------------------------
// K4: Jacobi step it >= 2 of Smooth(k) (spec S8).  Thread per affected vertex.  Odd
// steps read the positions and write the shadow copy, even steps the reverse; a
// vertex outside the affected set has the same value in both, so neighbours are read
// straight from the source array with no lookup.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_smooth_iter(SgDevView d, int from_shadow, int stats, int stamp_dab,
                                                     int stamp) {
  if (d.halt && d.halt[0]) return;
  const int nthreads = gridDim.x * blockDim.x;
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const float *sx_ = from_shadow ? d.qx : d.px;
  const float *sy_ = from_shadow ? d.qy : d.py;
  const float *sz_ = from_shadow ? d.qz : d.pz;
  float *dx_ = from_shadow ? d.px : d.qx;
  float *dy_ = from_shadow ? d.py : d.qy;
  float *dz_ = from_shadow ? d.pz : d.qz;
  const int na = (int)d.ctr[SG_CTR_NGATHER] << d.aff_shift;  // slots of the gathered leaves
  long long nbr = 0, r1 = 0;
  for (int i = i0; i < na; i += nthreads) {
    const int v = aff_slot_vertex(d, i, d.aff_cnt[i >> d.aff_shift]);
    if (v < 0) continue;
    const int2 vv = d.aff_vv[i];
    const float fade = d.aff_fade[i];
    const float x = sx_[v], y = sy_[v], z = sz_[v];
    float sx, sy, sz;
    nbr_sum(d.vv_idx, vv.x, vv.y, sx_, sy_, sz_, sx, sy, sz);
    nbr += vv.y - vv.x;
    if (stats) {
      for (int q = vv.x; q < vv.y; q++) {
        const int u = d.vv_idx[q];
        if (d.stamp_a[u] != stamp_dab && atomicExch(&d.stamp_v[u], stamp) != stamp) r1++;
      }
    }
    float rx = x, ry = y, rz = z;
    if (vv.y > vv.x) {
      const float inv = 1.0f / (float)(vv.y - vv.x);
      const float avx = sx * inv, avy = sy * inv, avz = sz * inv;
      rx = x + (avx - x) * fade;
      ry = y + (avy - y) * fade;
      rz = z + (avz - z) * fade;
    }
    dx_[v] = rx;
    dy_[v] = ry;
    dz_[v] = rz;
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

// Copy the final Jacobi iterate from the shadow copy (odd k) or to it (even k), so
// both arrays agree again.
__global__ __launch_bounds__(256) void k_smooth_sync(SgDevView d, int from_shadow) {
  if (d.halt && d.halt[0]) return;
  const int na = (int)d.ctr[SG_CTR_NGATHER] << d.aff_shift;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
    const int v = aff_slot_vertex(d, i, d.aff_cnt[i >> d.aff_shift]);
    if (v < 0) continue;
    if (from_shadow) {
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

// ---------------------------------------------------------------------------
// K5: post.  Normals of the affected vertices and bounds of the gathered leaves (all
// gathered leaves are marked, BKE_pbvh_node_mark_update).  Both only read positions.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_post(SgDevView d, int stats, int stamp, int sync_q) {
  if (d.halt && d.halt[0]) return;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int nthreads = gridDim.x * blockDim.x;
  const int gw = tid / SG_WAVE;
  const int nw = nthreads / SG_WAVE;
  // The normals take the affected records from the first thread on, the leaf refits
  // the gathered leaves from the last wave back, so the two independent chains run in
  // different waves side by side instead of one after the other in the first waves.
  const int rw = nw - 1 - gw;
  // speculative first-record loads alongside the counters
  const int nslot = d.L << d.aff_shift;
  const int v0 = tid < nslot ? d.aff_v[tid] : 0;
  const int2 vt0 = tid < nslot ? d.aff_vt[tid] : make_int2(0, 0);
  const int c0 = (tid >> d.aff_shift) < d.L ? d.aff_cnt[tid >> d.aff_shift] : 0;
  const int leaf0 = rw < d.L ? d.gathered[rw] : 0;
  const int2 vr0 = rw < d.L ? d.gl_v[rw] : make_int2(0, 0);
  const int2 sr0 = rw < d.L ? d.gl_s[rw] : make_int2(0, 0);
  CountPrefetch cp;
  if (gw == nw / 2) cp = aff_count_prefetch(d, lane_id());
  const int ng = (int)d.ctr[SG_CTR_NGATHER];
  const int na = ng << d.aff_shift;  // slots of the gathered leaves
  long long ntri = 0, nfa = 0, nvfa = 0, nsh = 0;
  // the dab's affected count for the host and the logs: a wave halfway through the
  // grid (between the normals' waves and the refits') sums the leaves' counts
  if (gw == nw / 2) {
    const int s = aff_count_sum(d, cp, lane_id(), ng);
    if (lane_id() == 0) d.ctr[SG_CTR_NAFF] = s;
  }
  for (int i = tid; i < na; i += nthreads) {
    const int v = (i == tid) ? ((i & ((1 << d.aff_shift) - 1)) < c0 ? v0 : -1)
                             : aff_slot_vertex(d, i, d.aff_cnt[i >> d.aff_shift]);
    if (v < 0) continue;
    const int2 vt = (i == tid) ? vt0 : d.aff_vt[i];
    if (sync_q) {
      d.qx[v] = d.px[v];
      d.qy[v] = d.py[v];
      d.qz[v] = d.pz[v];
    }
    vertex_normal(d, v, vt.x, vt.y, stats, stamp, ntri, nfa, nvfa);
  }
  for (int g = rw; g < ng; g += nw) {
    if (g == rw)
      leaf_refit(d, leaf0, vr0, sr0, nsh);
    else
      leaf_refit(d, d.gathered[g], d.gl_v[g], d.gl_s[g], nsh);
  }
  if (stats) {
    ntri = wave_sum_ll(ntri);
    nfa = wave_sum_ll(nfa);
    nvfa = wave_sum_ll(nvfa);
    if (lane_id() == 0) {
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_TRIS], as_u64(ntri));
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_FA], as_u64(nfa));
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_VFA], as_u64(nvfa));
      atomicAdd((unsigned long long *)&d.ctr[SG_CTR_SHARED], as_u64(nsh));
    }
  }
}

// ---------------------------------------------------------------------------
// Fused dab: one workgroup per gathered leaf runs the brush, Smooth's Jacobi steps, the
// normals of the leaf's affected vertices and the leaf's box in one pass, on the leaf's
// halo staged in LDS (its uniq vertices, ring 0 = its shared vertices and their edge
// neighbours' ring, and for Smooth(k) k more rings; tables from the build).  Every
// halo vertex's displacement is computed here from the same inputs and in the same
// FP32 order as the vertex's own workgroup computes it (a vertex of ring j is exact
// after k - j Jacobi steps), so the results equal the separate kernels' bit for bit.
// Workgroups read other leaves' vertices as they were before the dab, so the results
// go to the shadow copy (positions) and to (mx, my, mz) (normals); the next dab's
// search, or k_dab_finish, copies them in.  A halo vertex counts as affected only when
// its own leaf was gathered (the search stamps gathered leaves), as the search and the
// per-leaf vertex test decide it for that leaf.
// ---------------------------------------------------------------------------
#define FUSED_THREADS 256
#define FUSED_ROW 8  // neighbour / incidence entries per local held in registers

// One workgroup per gathered leaf; each thread holds LPT locals (local i = tid + j * 256).
// Global loads come in three dependent levels, each issued for every local at once:
//   1  the counters and the workgroup's first gathered leaf (ranges, halo header)
//   2  halo vertex ids, neighbour-row and incidence ranges, shared-vertex local ids
//   3  positions, normals, masks, owner stamps, neighbour rows (first 8 entries) and
//      incidence corners (first 8), into registers
// after which the brush, the Jacobi steps, the normals and the box run from LDS and
// registers.  Rows longer than 8 entries read the rest from memory where they are used.
#define FUSED_FPT 2              // faces per face thread whose corners load up front
#define FUSED_FACE_THREADS 192   // the first three waves compute the face normals, the last the leaf box

template <int LPT>
__global__ __launch_bounds__(FUSED_THREADS) void k_dab_fused(SgDevView d, SgDabK k, const float *curve, int serial,
                                                            int cap, int fcap) {
  if (d.halt && d.halt[0]) return;
  extern __shared__ __align__(16) unsigned char fused_smem[];
  __shared__ int s_cnt;
  // positions of the locals, two buffers of `cap` (x, y, z, 0), the normals of the
  // leaf's faces (`fcap` of them), the shared vertices' local ids
  float4 *const S = (float4 *)fused_smem;
  float4 *const Sfn = S + 2 * cap;
  unsigned short *Sshl = (unsigned short *)(Sfn + fcap);
  const int tid = threadIdx.x;
  // level 1: counters, stamp and the first gathered leaf's records together.  Loads
  // below are unconditional from clamped (always valid) indices, with the selection
  // after them: a load under a branch makes the compiler wait for it at the join.
  const int g0 = blockIdx.x;
  const int gc = min(g0, d.L - 1);
  const int leaf0 = d.gathered[gc];
  const int4 ra0 = d.gl_rec[2 * gc], rb0 = d.gl_rec[2 * gc + 1];
  const long long ng_ctr = d.ctr[SG_CTR_NGATHER];
  const int ng = (int)ng_ctr;
  const int stamp = d.dstamp[serial & 1];
  const bool smooth = k.tool == SG_TOOL_SMOOTH;
  const int K = smooth ? k.iters : 0;
  if (blockIdx.x == 0 && tid == 0) {  // this dab's results wait for the copy-back
    d.pend[0] = ng;
    d.pend[1] = stamp;
  }
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
  const bool hide = d.vhide && d.respect_hide;
  for (int g = g0; g < ng; g += gridDim.x) {
    const bool first = g == g0;
    const int leaf = first ? leaf0 : d.gathered[g];
    const int4 ra = first ? ra0 : d.gl_rec[2 * g], rb = first ? rb0 : d.gl_rec[2 * g + 1];
    const int nu = rb.x & 0xffff, ns = rb.x >> 16;
    const int2 vr = make_int2(ra.x, ra.x + nu), sr = make_int2(ra.y, ra.y + ns);
    const int ext_off = ra.z, nb_blk = ra.w;
    const int nlt = rb.w >> 16;  // faces incident to the uniq vertices
    // ext count at the end of ring j
    auto ring_end = [&](int j) {
      return j <= 0 ? rb.y & 0xffff : j == 1 ? rb.y >> 16 : j == 2 ? rb.z & 0xffff : j == 3 ? rb.z >> 16 : rb.w & 0xffff;
    };
    const int nloc = nu + ring_end(K);
    const int ninner = K > 0 ? nu + ring_end(K - 1) : 0;  // locals the first Jacobi step updates
    if (tid == 0) s_cnt = 0;
    if ((ng_ctr & SG_NGATHER_UNDO) && tid < SG_WAVE) undo_save_leaf(d, leaf, vr, d.px, d.py, d.pz);
    // level 2
    int vid[LPT], owner[LPT], nbb[LPT], nbe[LPT];
#pragma unroll
    for (int j = 0; j < LPT; j++) {
      const int i = tid + j * FUSED_THREADS;
      const bool ext = i >= nu && i < nloc, inner = i < ninner;
      // (id, owner) as one 64-bit load, so both halves stay where they are issued
      const long long e = *(const long long *)(d.hl_v + (ext ? ext_off + i - nu : 0));
      const int b = d.hl_nboff[inner ? nb_blk + i : 0], en = d.hl_nboff[inner ? nb_blk + i + 1 : 0];
      int id = (int)e, ow = (int)(e >> 32);
      if (!ext) {
        id = -1;
        ow = -1;
      }
      if (i < nu) id = vr.x + i;
      vid[j] = id;
      owner[j] = ow;
      nbb[j] = inner ? b : 0;
      nbe[j] = inner ? en : 0;
    }
    const bool uq0 = tid < nu;
    const int v0 = uq0 ? vr.x + tid : 0;
    const int qb0r = d.vt_off[v0], qe0r = d.vt_off[v0 + 1];
    const int shv0 = d.hl_sh[tid < ns ? sr.x + tid : 0];
    const int qb0 = uq0 ? qb0r : 0, qe0 = uq0 ? qe0r : 0;
    // the leaf's faces: their corners follow the leaf's neighbour offsets
    const int lt_base = nb_blk + nu + ring_end(SG_HALO_RINGS - 1) + 1;
    unsigned int fc0[FUSED_FPT], fc1[FUSED_FPT];
#pragma unroll
    for (int f = 0; f < FUSED_FPT; f++) {
      const int j = tid < FUSED_FACE_THREADS ? tid + f * FUSED_FACE_THREADS : nlt;
      fc0[f] = (unsigned)d.hl_nboff[j < nlt ? lt_base + 2 * j : 0];
      fc1[f] = (unsigned)d.hl_nboff[j < nlt ? lt_base + 2 * j + 1 : 0];
    }
    // level 3
    // stand-ins keep the loads unconditional without new traffic: a tool that needs no
    // normals, or a mesh with no mask or hidden vertices, reads the position lines again
    const bool inflate = k.tool == SG_TOOL_INFLATE;
    const float *mp = d.mask ? d.mask : d.px;
    const int *hp = hide ? d.vhide : (const int *)d.px;
    const float *nxp = inflate ? d.nx : d.px, *nyp = inflate ? d.ny : d.py, *nzp = inflate ? d.nz : d.pz;
    float x[LPT], y[LPT], z[LPT], nxv[LPT], nyv[LPT], nzv[LPT], mk[LPT];
    bool own[LPT], hid[LPT];
    unsigned short nb[LPT][FUSED_ROW];
#pragma unroll
    for (int j = 0; j < LPT; j++) {
      const int v = vid[j] < 0 ? 0 : vid[j];
      x[j] = d.px[v];
      y[j] = d.py[v];
      z[j] = d.pz[v];
      nxv[j] = nxp[v];
      nyv[j] = nyp[v];
      nzv[j] = nzp[v];
      const float m = mp[v];
      const int hv = hp[v];
      const int gs = d.leaf_gstamp[owner[j] < 0 ? 0 : owner[j]];
      mk[j] = d.mask ? m : 0.0f;
      hid[j] = hide && hv != 0;
      own[j] = owner[j] < 0 || gs == stamp;  // its leaf was gathered by this search
#pragma unroll
      for (int q = 0; q < FUSED_ROW; q++) nb[j][q] = d.hl_nbl[nbb[j] + q < nbe[j] ? nbb[j] + q : 0];
    }
    unsigned short ti0[FUSED_ROW];  // the first incidences' local faces
#pragma unroll
    for (int q = 0; q < FUSED_ROW; q++) ti0[q] = d.hl_ti[qb0 + q < qe0 ? qb0 + q : 0];
    // brush (spec S2-S9) on every local
    float fade[LPT];
    bool aff[LPT];
#pragma unroll
    for (int j = 0; j < LPT; j++) {
      const int i = tid + j * FUSED_THREADS;
      const float dx = k.cx - x[j];
      const float dy = k.cy - y[j];
      const float dz = k.cz - z[j];
      const float d2 = dx * dx + dy * dy + dz * dz;
      aff[j] = vid[j] >= 0 && own[j] && d2 <= k.r2 && !hid[j];
      fade[j] = 0.0f;
      if (aff[j]) {
        const float f = sg_falloff(k, curve, sqrtf(d2)) * (1.0f - mk[j]);
        if (k.tool == SG_TOOL_DRAW) {
          x[j] = x[j] + ox * f;
          y[j] = y[j] + oy * f;
          z[j] = z[j] + oz * f;
        } else if (k.tool == SG_TOOL_INFLATE) {
          const float sc = (k.bstr * f) * k.r;
          x[j] = x[j] + nxv[j] * sc;
          y[j] = y[j] + nyv[j] * sc;
          z[j] = z[j] + nzv[j] * sc;
        } else if (k.tool == SG_TOOL_GRAB) {
          fade[j] = k.bstr * f;
          x[j] = x[j] + k.gx * fade[j];
          y[j] = y[j] + k.gy * fade[j];
          z[j] = z[j] + k.gz * fade[j];
        } else {
          fade[j] = k.bstr * f;
        }
      }
      if (i < nloc) S[i] = make_float4(x[j], y[j], z[j], 0.0f);
    }
    if (tid < ns) Sshl[tid] = (unsigned short)shv0;
    for (int i = tid + FUSED_THREADS; i < ns; i += FUSED_THREADS) Sshl[i] = d.hl_sh[sr.x + i];
    __syncthreads();
    // Smooth: Jacobi step t updates the locals of rings <= K - t from step t - 1
    float inv[LPT];  // 1 / neighbour count, the same value each step
#pragma unroll
    for (int j = 0; j < LPT; j++) inv[j] = K > 0 && nbe[j] > nbb[j] ? 1.0f / (float)(nbe[j] - nbb[j]) : 0.0f;
    int cur = 0;
    for (int t = 1; t <= K; t++) {
      const int nact = nu + ring_end(K - t);
      const float4 *X = S + cur * cap;
      float4 *W = S + (cur ^ 1) * cap;
#pragma unroll
      for (int j = 0; j < LPT; j++) {
        const int i = tid + j * FUSED_THREADS;
        if (i < nact) {
          if (aff[j]) {
            const int cnt = nbe[j] - nbb[j];
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
            for (int q = 0; q < FUSED_ROW; q++) {
              if (q < cnt) {
                const float4 pu = X[nb[j][q]];
                sx = sx + pu.x;
                sy = sy + pu.y;
                sz = sz + pu.z;
              }
            }
            for (int q = FUSED_ROW; q < cnt; q++) {
              const float4 pu = X[d.hl_nbl[nbb[j] + q]];
              sx = sx + pu.x;
              sy = sy + pu.y;
              sz = sz + pu.z;
            }
            if (cnt > 0) {
              const float avx = sx * inv[j], avy = sy * inv[j], avz = sz * inv[j];
              x[j] = x[j] + (avx - x[j]) * fade[j];
              y[j] = y[j] + (avy - y[j]) * fade[j];
              z[j] = z[j] + (avz - z[j]) * fade[j];
            }
          }
          W[i] = make_float4(x[j], y[j], z[j], 0.0f);
        }
      }
      __syncthreads();
      cur ^= 1;
    }
    // the affected uniq vertices: normal (spec S10, incident tris in vertex->tri order),
    // position, list record
    const float4 *F = S + cur * cap;  // final positions of the locals
    if (tid < FUSED_FACE_THREADS) {
      // each face normal once (spec S10's face normal; faces in the tables' order)
      auto face_normal = [&](unsigned c0, unsigned c1, int j) {
        const float4 a = F[c0 & 0xffff], b = F[c0 >> 16], e = F[c1];
        float fx, fy, fz;
        sg_tri_normal(a.x, a.y, a.z, b.x, b.y, b.z, e.x, e.y, e.z, fx, fy, fz);
        Sfn[j] = make_float4(fx, fy, fz, 0.0f);
      };
#pragma unroll
      for (int f = 0; f < FUSED_FPT; f++) {
        const int j = tid + f * FUSED_FACE_THREADS;
        if (j < nlt) face_normal(fc0[f], fc1[f], j);
      }
      for (int j = tid + FUSED_FPT * FUSED_FACE_THREADS; j < nlt; j += FUSED_FACE_THREADS)
        face_normal((unsigned)d.hl_nboff[lt_base + 2 * j], (unsigned)d.hl_nboff[lt_base + 2 * j + 1], j);
    } else {
      // meanwhile the last wave folds the leaf's box over its uniq then shared vertices
      // (contiguous runs per lane, lanes in order, as leaf_box does)
      const int lane = tid - FUSED_FACE_THREADS;
      const int tot = nu + ns, per = (tot + SG_WAVE - 1) / SG_WAVE;
      BB6 acc;
      bb_reset(acc);
      for (int c = 0; c < per; c++) {
        const int i = lane * per + c;
        if (i < tot) {
          const float4 pu = F[i < nu ? i : Sshl[i - nu]];
          bb_add_point(acc, pu.x, pu.y, pu.z);
        }
      }
      acc = bb_wave_scan(acc);
      if (lane == SG_WAVE - 1) {
        bb_store(d.leaf_bb + 8 * (size_t)leaf, acc);
        bb_store(d.leaf_obb + 8 * (size_t)leaf, acc);
      }
    }
    __syncthreads();
    // the affected uniq vertices: normal (the incident faces' normals summed in
    // vertex->tri order), position, list record
#pragma unroll
    for (int j = 0; j < LPT; j++) {
      const int i = tid + j * FUSED_THREADS;
      if (i >= nu || !aff[j]) continue;
      const int v = vr.x + i;
      float ax = 0.0f, ay = 0.0f, az = 0.0f;
      auto add_face = [&](int f) {
        const float4 n = Sfn[f];
        ax = ax + n.x;
        ay = ay + n.y;
        az = az + n.z;
      };
      if (j == 0) {
        const int cnt = qe0 - qb0;
#pragma unroll
        for (int q = 0; q < FUSED_ROW; q++)
          if (q < cnt) add_face(ti0[q]);
        for (int q = qb0 + FUSED_ROW; q < qe0; q++) add_face(d.hl_ti[q]);
      } else {  // leaves of more than FUSED_THREADS uniq vertices
        const int qb = d.vt_off[v], qe = d.vt_off[v + 1];
        for (int q = qb; q < qe; q++) add_face(d.hl_ti[q]);
      }
      sg_normalize(ax, ay, az);
      d.mx[v] = ax;
      d.my[v] = ay;
      d.mz[v] = az;
      d.qx[v] = x[j];
      d.qy[v] = y[j];
      d.qz[v] = z[j];
      const int pos = atomicAdd(&s_cnt, 1);
      d.aff_v[(g << d.aff_shift) + pos] = v;
    }
    __syncthreads();
    if (tid == 0) d.aff_cnt[g] = s_cnt;
    __syncthreads();  // LDS is reused by the next leaf
  }
}

// After the last dab of a stroke (or a single dab): when that dab was fused, copy its
// results in, and give it its affected count.
__global__ __launch_bounds__(256) void k_dab_finish(SgDevView d, int serial) {
  const int stamp = d.dstamp[serial & 1];
  if (d.pend[1] != stamp) return;
  fused_copy_back(d, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
  if (blockIdx.x == 0 && threadIdx.x < SG_WAVE) {
    const CountPrefetch cp = aff_count_prefetch(d, threadIdx.x);
    const int s = aff_count_sum(d, cp, threadIdx.x, d.pend[0]);
    if (threadIdx.x == 0) d.ctr[SG_CTR_NAFF] = s;
  }
}

// ---------------------------------------------------------------------------
// Deferred refit of internal nodes.  Sub: one workgroup per dirty cut subtree,
// leaf boxes staged in LDS, internal nodes computed height by height
// (parent = BB_expand_with_bb(child0, child1), pbvh.c:123-124).  Top: one workgroup
// over the tree above the cut, with the cut roots as its leaves.
// ---------------------------------------------------------------------------
#define FLUSH_THREADS 256
#define FLUSH_MAXH 256

// Where each height's run starts in a table sorted by height (.w): hs[h] = first entry
// of height >= h for h = 1..FLUSH_MAXH, n past the end.  Every entry compares its
// height with its predecessor's, so the runs are found in one parallel pass instead of
// a serial scan per height.  Block-wide; ends with a barrier.
__device__ __forceinline__ void height_starts(const int4 *tab, int n, int *hs) {
  for (int h = threadIdx.x; h <= FLUSH_MAXH; h += blockDim.x) hs[h] = n;
  __syncthreads();
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int w = min(tab[e].w, FLUSH_MAXH), pw = e > 0 ? min(tab[e - 1].w, FLUSH_MAXH) : 0;
    for (int h = pw + 1; h <= w; h++) hs[h] = e;  // one writer per h: the run boundary
  }
  __syncthreads();
}

__global__ __launch_bounds__(FLUSH_THREADS) void k_flush_sub(SgDevView d, int maxn) {
  extern __shared__ float s_bb[];  // [maxn * 6]
  __shared__ int s_hs[FLUSH_MAXH + 1];
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
  height_starts(d.sub_tab + eb, ee - eb, s_hs);  // entries are sorted by height
  for (int h = 1; h < FLUSH_MAXH && s_hs[h] < ee - eb; h++) {
    for (int e = eb + s_hs[h] + threadIdx.x; e < eb + s_hs[h + 1]; e += blockDim.x) {
      const int4 t = d.sub_tab[e];
      const int li = nl + (e - eb);
      for (int q = 0; q < 3; q++) {
        s_bb[6 * li + q] = sg_min(s_bb[6 * t.y + q], s_bb[6 * t.z + q]);
        s_bb[6 * li + 3 + q] = sg_max(s_bb[6 * t.y + 3 + q], s_bb[6 * t.z + 3 + q]);
      }
    }
    __syncthreads();
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
    if (e == ee - 1) {  // the last entry is the subtree root
      bb_store(d.cut_bb + 8 * (size_t)c, b);
      bb_store(d.cut_obb + 8 * (size_t)c, b);
    }
  }
  if (threadIdx.x == 0 && ee == eb) {  // subtree root is a leaf: mirror its box
    const BB6 b = bb_load(d.leaf_bb + 8 * (size_t)lb);
    bb_store(d.node_bb + 8 * (size_t)d.cut_node[c], b);
    bb_store(d.node_obb + 8 * (size_t)d.cut_node[c], b);
    bb_store(d.cut_bb + 8 * (size_t)c, b);
    bb_store(d.cut_obb + 8 * (size_t)c, b);
  }
}

__global__ __launch_bounds__(1024) void k_flush_top(SgDevView d, int top_n, int clear) {
  extern __shared__ float s_bb[];  // [(C + top_n) * 6]
  __shared__ int s_hs[FLUSH_MAXH + 1];
  const int C = d.C;
  for (int i = threadIdx.x; i < C; i += blockDim.x) {
    const BB6 b = bb_load(d.cut_bb + 8 * (size_t)i);
    for (int q = 0; q < 3; q++) {
      s_bb[6 * i + q] = b.mn[q];
      s_bb[6 * i + 3 + q] = b.mx[q];
    }
  }
  height_starts(d.top_tab, top_n, s_hs);  // also the barrier after the staging
  for (int h = 1; h < FLUSH_MAXH && s_hs[h] < top_n; h++) {
    for (int e = s_hs[h] + threadIdx.x; e < s_hs[h + 1]; e += blockDim.x) {
      const int4 t = d.top_tab[e];
      const int li = C + e;
      for (int q = 0; q < 3; q++) {
        s_bb[6 * li + q] = sg_min(s_bb[6 * t.y + q], s_bb[6 * t.z + q]);
        s_bb[6 * li + 3 + q] = sg_max(s_bb[6 * t.y + 3 + q], s_bb[6 * t.z + 3 + q]);
      }
    }
    __syncthreads();
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
__device__ __forceinline__ void leaf_refit_id(const SgDevView &d, int leaf, long long &nsh) {
  leaf_refit(d, leaf, make_int2(d.leaf_vbeg[leaf], d.leaf_vbeg[leaf + 1]),
             make_int2(d.leaf_sh_off[leaf], d.leaf_sh_off[leaf + 1]), nsh);
}

// cut subtree box from its leaves (ordered), for the host-driven refits
__device__ __forceinline__ void cut_refresh(const SgDevView &d, int c) {
  const int lb = d.cut_lbeg[c], le = d.cut_lend[c];
  const int per = (le - lb + SG_WAVE - 1) / SG_WAVE;
  BB6 acc, oacc;
  bb_reset(acc);
  bb_reset(oacc);
  for (int q = 0; q < per; q++) {
    const int l = lb + lane_id() * per + q;
    if (l < le) {
      bb_add_bb(acc, bb_load(d.leaf_bb + 8 * (size_t)l));
      bb_add_bb(oacc, bb_load(d.leaf_obb + 8 * (size_t)l));
    }
  }
  acc = bb_wave_reduce(acc);
  oacc = bb_wave_reduce(oacc);
  if (lane_id() == 0) {
    bb_store(d.cut_bb + 8 * (size_t)c, acc);
    bb_store(d.cut_obb + 8 * (size_t)c, oacc);
  }
}

__global__ __launch_bounds__(256) void k_refit_list(SgDevView d, const int *leaves, int n) {
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  long long nsh = 0;
  for (int g = gw; g < n; g += nw) {
    const int leaf = leaves[g];
    leaf_refit_id(d, leaf, nsh);
    if (lane_id() == 0) d.cut_dirty[d.leaf_cut[leaf]] = 1;
  }
}

// after a host-driven leaf refit: every dirty subtree's search box from its leaves
__global__ __launch_bounds__(256) void k_cut_refresh(SgDevView d) {
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  for (int c = gw; c < d.C; c += nw) {
    if (d.cut_dirty[c]) cut_refresh(d, c);
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
      vertex_normal(d, v, d.vt_off[v], d.vt_off[v + 1], 0, 0, a, b, c);
      d.vmark[v] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// Whole-mesh angle-weighted vertex normals (KERNEL_mesh_calc_normals_poly_and_vertex,
// mesh_normals.c:225-299) as a per-vertex gather in ascending polygon order: each
// polygon's normal by Newell's method, weighted by the angle at the vertex's corner,
// saacos(-dot(e_prev, e_next)) with e_prev = normalize(c_prev - c_i), e_next =
// normalize(c_i - c_next); a zero sum falls back to the normalized position.  A
// triangle mesh walks its vertex->tri incidences (each tri its own polygon); a mesh
// with polygons its vertex->(polygon, corner) list.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sg_saacos(float f) {
  return (f <= -1.0f) ? 3.14159265358979323846f : ((f >= 1.0f) ? 0.0f : acosf(f));
}

__global__ __launch_bounds__(256) void k_full_normals(SgDevView d) {
  const int V = d.V;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    if (d.vc_off) {
      for (int q = d.vc_off[v]; q < d.vc_off[v + 1]; q++) {
        const int2 pc = d.vc[q];
        const int b = d.poly_off[pc.x], n = d.poly_off[pc.x + 1] - b;
        // Newell normal of the polygon, from its last corner
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        int u = d.poly_v[b + n - 1];
        float cx = d.px[u], cy = d.py[u], cz = d.pz[u];
        for (int k = 0; k < n; k++) {
          u = d.poly_v[b + k];
          const float nx = d.px[u], ny = d.py[u], nz = d.pz[u];
          fx += (cy - ny) * (cz + nz);
          fy += (cz - nz) * (cx + nx);
          fz += (cx - nx) * (cy + ny);
          cx = nx;
          cy = ny;
          cz = nz;
        }
        if (sg_normalize(fx, fy, fz) == 0.0f) fz = 1.0f;
        // the angle at corner pc.y between its two polygon edges
        const int up = d.poly_v[b + (pc.y + n - 1) % n], uc = d.poly_v[b + pc.y], un = d.poly_v[b + (pc.y + 1) % n];
        float epx = d.px[up] - d.px[uc], epy = d.py[up] - d.py[uc], epz = d.pz[up] - d.pz[uc];
        float enx = d.px[uc] - d.px[un], eny = d.py[uc] - d.py[un], enz = d.pz[uc] - d.pz[un];
        sg_normalize(epx, epy, epz);
        sg_normalize(enx, eny, enz);
        const float fac = sg_saacos(-(epx * enx + epy * eny + epz * enz));
        ax += fx * fac;
        ay += fy * fac;
        az += fz * fac;
      }
    } else {
      for (int q = d.vt_off[v]; q < d.vt_off[v + 1]; q++) {
        const int4 c = d.vt4[q];
        const int ci = (int)((unsigned)c.w >> 30);
        const int id[3] = {c.x, c.y, c.z};
        float P[3][3];
        for (int j = 0; j < 3; j++) {
          P[j][0] = d.px[id[j]];
          P[j][1] = d.py[id[j]];
          P[j][2] = d.pz[id[j]];
        }
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
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
      d.qx[v] = x;
      d.qy[v] = y;
      d.qz[v] = z;
      d.vmark[v] = 1;
    }
  }
}

__global__ __launch_bounds__(256) void k_refit_all(SgDevView d) {
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  long long nsh = 0;
  for (int l = gw; l < d.L; l += nw) leaf_refit_id(d, l, nsh);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < d.C; c += gridDim.x * blockDim.x)
    d.cut_dirty[c] = 1;
}

// Vertex boxes of listed leaves into out[8 g] (not into the tree): the host flush of
// BKE_pbvh_update_bounds with partial flags decides which box each becomes.
__global__ __launch_bounds__(256) void k_leaf_boxes(SgDevView d, const int *leaves, int n, float *out) {
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) / SG_WAVE;
  const int nw = gridDim.x * blockDim.x / SG_WAVE;
  for (int g = gw; g < n; g += nw) {
    const int leaf = leaves[g];
    const BB6 b = leaf_box(d, make_int2(d.leaf_vbeg[leaf], d.leaf_vbeg[leaf + 1]),
                           make_int2(d.leaf_sh_off[leaf], d.leaf_sh_off[leaf + 1]));
    if (lane_id() == 0) bb_store(out + 8 * (size_t)g, b);
  }
}

// Box records written by the host flush: .x kind (0 leaf, 1 internal node, 2 search
// subtree), .y index; then the current and the original box (8 floats each).
struct SgBoxRec {
  int kind, idx, pad0, pad1;
  float vb[8], ob[8];
};

__global__ __launch_bounds__(256) void k_scatter_boxes(SgDevView d, const SgBoxRec *recs, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 16 * n; i += gridDim.x * blockDim.x) {
    const SgBoxRec &r = recs[i >> 4];
    const int k = i & 15;
    float *vb = r.kind == 0 ? d.leaf_bb : (r.kind == 1 ? d.node_bb : d.cut_bb);
    float *ob = r.kind == 0 ? d.leaf_obb : (r.kind == 1 ? d.node_obb : d.cut_obb);
    if (k < 8)
      vb[8 * (size_t)r.idx + k] = r.vb[k];
    else
      ob[8 * (size_t)r.idx + k - 8] = r.ob[k - 8];
  }
}

// Host marks (BKE_pbvh_vert_mark_update) join the device's marks: list of new ids.
__global__ __launch_bounds__(256) void k_set_marks(SgDevView d, const int *verts, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d.vmark[verts[i]] = 1;
}

// copy the counters of the last dab of a stroke into its log row
__global__ void k_log_row(const long long *ctr, long long *log, int row) {
  const int i = threadIdx.x;
  if (i < SG_CTR_COUNT) log[(size_t)row * SG_CTR_COUNT + i] = i == SG_CTR_NGATHER ? (ctr[i] & 0xffffffffll) : ctr[i];
}

int sg_enqueue_log_row(PBVH *p, long long *log, int row) {
  hipLaunchKernelGGL(k_log_row, dim3(1), dim3(64), 0, p->stream, p->d.ctr + (size_t)(row & 1) * SG_CTR_COUNT, log,
                     row);
  return sg_check_hip(hipGetLastError(), "log launch");
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
SgDevView sg_view(PBVH *p) { return sg_view_of(p->d, p->respect_hide, p->halt_live); }

SgDevView sg_view_of(SgDev &d, bool respect_hide, bool halt_live) {
  SgDevView v;
  v.V = d.V;
  v.T = d.T;
  v.L = d.L;
  v.N = d.N;
  v.C = d.C;
  v.respect_hide = respect_hide ? 1 : 0;
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
  v.vt4 = d.vt4;
  v.vt_off = d.vt_off;
  v.vv_off = d.vv_off;
  v.vv_idx = d.vv_idx;
  v.poly_off = d.poly_off;
  v.poly_v = d.poly_v;
  v.poly_of_tri = d.poly_of_tri;
  v.vc_off = d.vc_off;
  v.vc = d.vc;
  v.leaf_vbeg = d.leaf_vbeg;
  v.leaf_sh_off = d.leaf_sh_off;
  v.leaf_sh = d.leaf_sh;
  v.leaf_flags = d.leaf_flags;
  v.leaf_cut = d.leaf_cut;
  v.leaf_bb = d.leaf_bb;
  v.leaf_obb = d.leaf_obb;
  v.node_bb = d.node_bb;
  v.node_obb = d.node_obb;
  v.cut_bb = d.cut_bb;
  v.cut_obb = d.cut_obb;
  v.cut_node = d.cut_node;
  v.cut_lbeg = d.cut_lbeg;
  v.cut_lend = d.cut_lend;
  v.cut_flags = d.cut_flags;
  v.sub_off = d.sub_off;
  v.sub_tab = d.sub_tab;
  v.top_tab = d.top_tab;
  v.cut_dirty = d.cut_dirty;
  v.vhide = d.vhide;
  v.vmark = d.vmark;
  v.gathered = d.gathered;
  v.draw_dirty = d.draw_dirty;
  v.undo = d.undo;
  v.place_pair = d.place_pair;
  v.gl_v = d.gl_v;
  v.gl_s = d.gl_s;
  v.aff_v = d.aff_v;
  v.aff_vt = d.aff_vt;
  v.aff_vv = d.aff_vv;
  v.aff_fade = d.aff_fade;
  v.aff_cnt = d.aff_cnt;
  v.aff_shift = d.aff_shift;
  v.prev_cuts = d.prev_cuts;
  v.ctr = d.ctr;
  v.ctr_next = d.ctr + SG_CTR_COUNT;
  v.cut_touch = d.cut_touch;
  v.hl_rec = d.hl_rec;
  v.hl_v = d.hl_v;
  v.hl_nboff = d.hl_nboff;
  v.hl_nbl = d.hl_nbl;
  v.hl_sh = d.hl_sh;
  v.hl_ti = d.hl_ti;
  v.gl_rec = d.gl_rec;
  v.leaf_gstamp = d.leaf_gstamp;
  v.dstamp = d.dstamp;
  v.pend = d.pend;
  v.mx = d.mx;
  v.my = d.my;
  v.mz = d.mz;
  v.halt = halt_live ? d.halt : nullptr;  // no flag test in strokes without learned dabs
  v.seq = d.seq;
  v.vstamp = d.vstamp;
  v.vperm_old = d.vperm_old;
  v.stamp_v = d.stamp_v;
  v.stamp_v2 = d.stamp_v2;
  v.stamp_t = d.stamp_t;
  v.stamp_a = d.stamp_a;
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
  k.hide_skip = 1;
  k.etype = dab->elastic_deform_type;
  k.epois = dab->elastic_deform_volume_preservation;
  k.enx = dab->surface_normal[0];
  k.eny = dab->surface_normal[1];
  k.enz = dab->surface_normal[2];
  k.eforce = 0.0f;
  k.orig_co = (k.tool == SG_TOOL_GRAB && (dab->flag & SG_DAB_ORIGINAL)) ? 1 : 0;
  k.learned = (k.tool == SG_TOOL_SMOOTH && (dab->flag & SG_SMOOTH_LEARNED)) ? 1 : 0;
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
    case SG_TOOL_ELASTIC_DEFORM: k.bstr = s; break;
    default: k.bstr = 0.0f; break;
  }
  if (k.tool == SG_TOOL_ELASTIC_DEFORM) {
    // Every leaf (the brush needs all nodes: a Kelvinlet has no finite support), the
    // fully hidden ones too (the node search without a callback returns them).  The
    // force: |grab_delta| * direction * bstrength, the direction -1 with DIR_IN (the
    // sign the stroke's mouse motion gives).
    k.r2 = __builtin_inff();
    k.hide_skip = 0;
    const float gl = sqrtf(k.gx * k.gx + k.gy * k.gy + k.gz * k.gz);
    k.eforce = gl * flip * k.bstr;
  }
  return k;
}

// Event pair around one launch when per-kernel profiling is on.
// Per-kernel timing: with profiling on, every launch goes through
// hipExtLaunchKernelGGL with a start/stop event pair that the dispatch packet itself
// timestamps, so the measured span is the kernel's execution on this stream.
// Launches captured into a graph are not profiled.
struct ProfScope {
  PBVH *p;
  int kid;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ProfScope(PBVH *pp, int k) : p(pp), kid(k) {
    if (!p->profile || p->capturing) return;
    while (p->ev_pool.size() < p->ev_used + 2) {
      hipEvent_t e;
      // no system-scope fence: the events only time kernels on this stream
      hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
      p->ev_pool.push_back(e);
    }
    e0 = p->ev_pool[p->ev_used];
    e1 = p->ev_pool[p->ev_used + 1];
    p->ev_used += 2;
    p->ev_kid.push_back(kid);
  }
};

#define SG_LAUNCH(ps, kern, grid, block, shm, stream, ...)                                   \
  do {                                                                                      \
    if ((ps).e0)                                                                            \
      hipExtLaunchKernelGGL(kern, grid, block, shm, stream, (ps).e0, (ps).e1, 0, __VA_ARGS__); \
    else                                                                                    \
      hipLaunchKernelGGL(kern, grid, block, shm, stream, __VA_ARGS__);                      \
  } while (0)

int sg_profile_collect(PBVH *p) {
  if (p->ev_kid.empty()) return 0;
  if (sg_check_hip(hipStreamSynchronize(p->stream), "profile sync")) return -1;
  for (size_t i = 0; i < p->ev_kid.size(); i++) {
    float ms = 0.0f;
    hipEventElapsedTime(&ms, p->ev_pool[2 * i], p->ev_pool[2 * i + 1]);
    p->prof_ms[p->ev_kid[i]] += ms;
    p->prof_n[p->ev_kid[i]] += 1;
  }
  p->ev_kid.clear();
  p->ev_used = 0;
  return 0;
}

int sg_enqueue_clear_ctr(PBVH *p) {
  hipLaunchKernelGGL(k_clear_ctr, dim3(1), dim3(64), 0, p->stream, p->d.ctr, p->d.halt, p->d.dstamp);
  return sg_check_hip(hipGetLastError(), "clear launch");
}

// One dab; `serial` is the dab's position in its stroke (selects the counter set and
// the "touched by the previous dab" test of the search).
int sg_enqueue_dab(PBVH *p, const SgDabK &k, const float *curve_dev, long long *log, int log_index,
                   int stamp, int serial, const SgCollect *col, const float *place) {
  SgDevView v = sg_view(p);
  v.ctr = p->d.ctr + (size_t)(serial & 1) * SG_CTR_COUNT;
  v.ctr_next = p->d.ctr + (size_t)((serial + 1) & 1) * SG_CTR_COUNT;
  hipStream_t s = p->stream;
  const int G = p->grid_wg;
  const int GB = p->grid_brush > 0 ? p->grid_brush : G;
  const int GP = p->grid_post > 0 ? p->grid_post : G;
  const int GS = p->grid_smooth > 0 ? p->grid_smooth : G;
  // the fused dab (one kernel after the search) when the mesh has halo tables and the
  // dab is a plain brush dab it covers
  SgDabK kk = k;
  // no leaf is fully hidden: the hidden test passes every leaf, and the search skips
  // loading the flags
  if (p->d.nhidden_leaves == 0) kk.hide_skip = 0;
  kk.fused = 0;
  const bool fused_tool = k.tool == SG_TOOL_DRAW || k.tool == SG_TOOL_INFLATE || k.tool == SG_TOOL_GRAB ||
                          (k.tool == SG_TOOL_SMOOTH && k.iters >= 1 && k.iters <= SG_HALO_RINGS);
  int fused_cap = 0, fused_fcap = 1, lpt = 0;
  size_t lds = 0;
  if (p->d.halo_ok && p->fused_enabled && fused_tool && !k.stats && !k.orig_co && !(k.learned && p->mlp) && !col &&
      !place) {
    fused_cap = p->d.halo_maxlocr[k.tool == SG_TOOL_SMOOTH ? k.iters : 0];
    lpt = (fused_cap + FUSED_THREADS - 1) / FUSED_THREADS;
    fused_fcap = std::max(1, p->d.halo_maxlt);
    lds = (size_t)fused_cap * 2 * sizeof(float4) + (size_t)fused_fcap * sizeof(float4) +
          (size_t)(p->d.leaf_maxsh + 2) * 2 + 16;
    // larger halos (very large leaves) take the separate kernels
    kk.fused = lpt >= 1 && lpt <= 4 && lds <= 65536 ? 1 : 0;
  }
  const int copy_blocks = p->d.halo_ok ? 64 : 0;
  {
    ProfScope ps(p, SG_K_SEARCH);
    const int leaf_blocks = std::max(1, (p->d.L + 255) / 256);
    const int cut_blocks = (p->d.C + 1 + 3) / 4;  // a wave per subtree and one for the bookkeeping
    SG_LAUNCH(ps, k_search, dim3(leaf_blocks + cut_blocks + copy_blocks), dim3(256), 0, s, v, kk, log, log_index,
              serial, place, leaf_blocks, cut_blocks, copy_blocks, p->fused_pending ? 1 : 0);
    p->fused_pending = kk.fused != 0;
  }
  if (kk.fused) {
    if (k.tool == SG_TOOL_DRAW) {
      ProfScope ps(p, SG_K_AREA);
      SG_LAUNCH(ps, k_area, dim3(GB), dim3(256), 0, s, v, k, place);
    }
    ProfScope ps(p, SG_K_BRUSH);
    const dim3 gf(p->grid_fused), bf(FUSED_THREADS);
    if (lpt == 1)
      SG_LAUNCH(ps, k_dab_fused<1>, gf, bf, lds, s, v, kk, curve_dev, serial, fused_cap, fused_fcap);
    else if (lpt == 2)
      SG_LAUNCH(ps, k_dab_fused<2>, gf, bf, lds, s, v, kk, curve_dev, serial, fused_cap, fused_fcap);
    else if (lpt == 3)
      SG_LAUNCH(ps, k_dab_fused<3>, gf, bf, lds, s, v, kk, curve_dev, serial, fused_cap, fused_fcap);
    else
      SG_LAUNCH(ps, k_dab_fused<4>, gf, bf, lds, s, v, kk, curve_dev, serial, fused_cap, fused_fcap);
    return sg_check_hip(hipGetLastError(), "dab launch");
  }
  if (k.tool == SG_TOOL_ELASTIC_DEFORM) {  // every leaf; then all normals, leaf boxes and the tree
    // whole-mesh passes: a wave per leaf, a thread per vertex (enough waves in flight
    // to cover the memory latency, rather than the per-dab kernels' fixed grid)
    const int GL = std::max(1, (p->d.L + 3) / 4);
    {
      ProfScope ps(p, SG_K_BRUSH);
      const long long nw = p->respect_hide ? p->d.nuniq_visible : p->d.nuniq;
      SG_LAUNCH(ps, k_elastic, dim3(GL), dim3(256), 0, s, v, k, place, nw);
    }
    {
      ProfScope ps(p, SG_K_POST);
      const int nb = std::max(1, (p->d.V + 255) / 256);
      SG_LAUNCH(ps, k_elastic_normals, dim3(nb), dim3(256), 0, s, v);
    }
    {
      ProfScope ps(p, SG_K_POST);
      SG_LAUNCH(ps, k_refit_all, dim3(GL), dim3(256), 0, s, v);
    }
    {
      ProfScope ps(p, SG_K_FLUSH_SUB);
      const size_t sub_lds = (size_t)p->d.sub_maxn * 6 * sizeof(float);
      SG_LAUNCH(ps, k_flush_sub, dim3(p->d.C), dim3(FLUSH_THREADS), sub_lds, s, v, p->d.sub_maxn);
    }
    {
      ProfScope ps(p, SG_K_FLUSH_TOP);
      const size_t top_lds = (size_t)(p->d.C + p->d.top_n) * 6 * sizeof(float);
      SG_LAUNCH(ps, k_flush_top, dim3(1), dim3(1024), top_lds, s, v, p->d.top_n, 1);
    }
    return sg_check_hip(hipGetLastError(), "dab launch");
  }
  if (k.learned && p->mlp) {  // learned Smooth: brush + MLP, then decide / commit
    {
      ProfScope ps(p, SG_K_BRUSH);
      sg_enqueue_learned_brush(p, v, k, curve_dev, p->mlp, place);
    }
    {
      ProfScope ps(p, SG_K_POST);
      sg_enqueue_learned_post(p, v, p->mlp, serial);
    }
    return sg_check_hip(hipGetLastError(), "dab launch");
  }
  if (k.tool == SG_TOOL_DRAW) {
    ProfScope ps(p, SG_K_AREA);
    SG_LAUNCH(ps, k_area, dim3(GB), dim3(256), 0, s, v, k, place);
  }
  if (k.orig_co) {
    ProfScope ps(p, SG_K_BRUSH);
    SG_LAUNCH(ps, k_brush_orig, dim3(GB), dim3(256), 0, s, v, k, curve_dev, stamp, place);
  } else {
    ProfScope ps(p, SG_K_BRUSH);
    SG_LAUNCH(ps, k_brush, dim3(GB), dim3(256), 0, s, v, k, curve_dev, stamp, place);
  }
  if (col) sg_enqueue_collect(p, v, col->X, col->frame, col->cap);
  if (k.tool == SG_TOOL_SMOOTH) {
    // step 1 (in k_brush) wrote the shadow copy; step it reads the shadow copy when it
    // is even and writes the positions, and the reverse when it is odd
    for (int it = 2; it <= k.iters; it++) {
      ProfScope ps(p, SG_K_SMOOTH);
      SG_LAUNCH(ps, k_smooth_iter, dim3(GS), dim3(256), 0, s, v, (it % 2 == 0) ? 1 : 0, k.stats, stamp, stamp + it);
    }
    // odd k ends in the shadow copy: copy it back before the normals read positions;
    // even k ends in the positions and k_post refreshes the shadow copy
    if (k.iters % 2 == 1) {
      ProfScope ps(p, SG_K_SMOOTH);
      SG_LAUNCH(ps, k_smooth_sync, dim3(GS), dim3(256), 0, s, v, 1);
    }
  }
  {
    ProfScope ps(p, SG_K_POST);
    SG_LAUNCH(ps, k_post, dim3(GP), dim3(256), 0, s, v, k.stats, stamp,
              (k.tool == SG_TOOL_SMOOTH && k.iters % 2 == 0) ? 1 : 0);
  }
  if (col) sg_enqueue_target(p, v, col->frame, col->T, col->cap);
  if (p->profile && !p->capturing && p->ev_kid.size() > 4096) sg_profile_collect(p);
  return sg_check_hip(hipGetLastError(), "dab launch");
}

int sg_enqueue_flush(PBVH *p) {
  const SgDevView v = sg_view(p);
  hipStream_t s = p->stream;
  const size_t sub_lds = (size_t)p->d.sub_maxn * 6 * sizeof(float);
  {
    ProfScope ps(p, SG_K_FLUSH_SUB);
    SG_LAUNCH(ps, k_flush_sub, dim3(p->d.C), dim3(FLUSH_THREADS), sub_lds, s, v, p->d.sub_maxn);
  }
  const size_t top_lds = (size_t)(p->d.C + p->d.top_n) * 6 * sizeof(float);
  {
    ProfScope ps(p, SG_K_FLUSH_TOP);
    SG_LAUNCH(ps, k_flush_top, dim3(1), dim3(1024), top_lds, s, v, p->d.top_n, 1);
  }
  return sg_check_hip(hipGetLastError(), "flush launch");
}

int sg_enqueue_full_normals(PBVH *p) {
  const SgDevView v = sg_view(p);
  const int blocks = (p->d.V + 255) / 256;
  hipLaunchKernelGGL(k_full_normals, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, p->stream, v);
  return sg_check_hip(hipGetLastError(), "normals launch");
}

int sg_enqueue_full_normals_view(const SgDevView &v, hipStream_t s) {
  const int blocks = (v.V + 255) / 256;
  hipLaunchKernelGGL(k_full_normals, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, s, v);
  return sg_check_hip(hipGetLastError(), "normals launch");
}

int sg_enqueue_gather_raw(hipStream_t s, const float *x, const float *y, const float *z, const int *perm_old, int V,
                          float *out_dev) {
  const int blocks = (V + 255) / 256;
  hipLaunchKernelGGL(k_gather_aos, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, s, x, y, z, perm_old, V, out_dev);
  return sg_check_hip(hipGetLastError(), "gather launch");
}

int sg_enqueue_refit_list(PBVH *p, const int *leaves_dev, int n) {
  const SgDevView v = sg_view(p);
  if (n > 0) hipLaunchKernelGGL(k_refit_list, dim3(p->grid_wg), dim3(256), 0, p->stream, v, leaves_dev, n);
  hipLaunchKernelGGL(k_cut_refresh, dim3(p->grid_wg), dim3(256), 0, p->stream, v);
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
  hipLaunchKernelGGL(k_refit_all, dim3(std::max(1, (p->d.L + 3) / 4)), dim3(256), 0, p->stream, v);  // a wave per leaf
  hipLaunchKernelGGL(k_cut_refresh, dim3(p->grid_wg), dim3(256), 0, p->stream, v);
  return sg_check_hip(hipGetLastError(), "coords launch");
}

int sg_enqueue_finish(PBVH *p, int serial_last) {
  p->fused_pending = false;
  if (!p->d.halo_ok) return 0;
  SgDevView v = sg_view(p);
  v.ctr = p->d.ctr + (size_t)(serial_last & 1) * SG_CTR_COUNT;  // the last dab's counter set
  hipLaunchKernelGGL(k_dab_finish, dim3(64), dim3(256), 0, p->stream, v, serial_last);
  return sg_check_hip(hipGetLastError(), "finish launch");
}

int sg_enqueue_leaf_boxes(PBVH *p, const int *leaves_dev, int n, float *out_dev) {
  const SgDevView v = sg_view(p);
  if (n > 0) hipLaunchKernelGGL(k_leaf_boxes, dim3(std::max(1, std::min(p->grid_wg, (n + 3) / 4))), dim3(256), 0,
                                p->stream, v, leaves_dev, n, out_dev);
  return sg_check_hip(hipGetLastError(), "leaf boxes launch");
}

int sg_enqueue_scatter_boxes(PBVH *p, const void *recs_dev, int n) {
  const SgDevView v = sg_view(p);
  if (n > 0) hipLaunchKernelGGL(k_scatter_boxes, dim3(std::max(1, std::min(1024, (16 * n + 255) / 256))), dim3(256), 0,
                                p->stream, v, (const SgBoxRec *)recs_dev, n);
  return sg_check_hip(hipGetLastError(), "scatter launch");
}

int sg_enqueue_set_marks(PBVH *p, const int *verts_dev, int n) {
  const SgDevView v = sg_view(p);
  if (n > 0) hipLaunchKernelGGL(k_set_marks, dim3(std::max(1, std::min(1024, (n + 255) / 256))), dim3(256), 0,
                                p->stream, v, verts_dev, n);
  return sg_check_hip(hipGetLastError(), "marks launch");
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
