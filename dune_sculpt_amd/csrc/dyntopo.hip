// Dynamic topology (PBVH_BMESH) on the GPU: the corner table, leaf face slabs and
// work lists in HBM; a topology update runs in rounds (pbvh_bmesh.c:1931-2005 with the
// queue rules of :773-1090, the split of :1093-1209 and the collapse of :1264-1429),
// then the dab's brush over the gathered leaves' own vertices, the normals of flagged
// leaves (pbvh_bmesh.c:1635-1656) and their bounds.
//
// One round, each step a kernel over a list:
//   begin    counters of the round
//   cand     faces of the working set (and the last round's high-valence opposite
//            vertices): the queue's edges, each registered once (corner stamp); the
//            registrant clears the claims of the faces the edge would own
//   claim    validity, the edge's key (length band, then a hash of its labels),
//            atomicMax on every owned face; equal maxima recorded per face
//   verify   an edge that holds every owned face's unshared maximum wins
//   apply    winners rewrite only faces they own (a split's two faces, a collapse's
//            faces around both vertices); a split records where its outer corners moved
//   relink / owner   splits join their outer edges; collapses fix vertex owners
// The host reads the round's counters (one wait per round) and stops when no edge
// qualifies, none wins, or a capacity limit is reached.  Slots are allocated with
// atomics in any order: results depend on labels only, and match the oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "sg_device.h"
#include "sg_dyn.h"
#include "sg_kernels.h"

#define DY_THREADS 256
#define DY_BLOCKS 160
#define LB_THREADS 1024  // a leaf's face slots (hundreds between after_stroke calls) in one pass

namespace {

struct DV {
  int Vcap, Tcap, N, L, has_mask, respect_hide, use_frontface, use_projected, ccap;
  float vnx, vny, vnz;
  float *px, *py, *pz, *nx, *ny, *nz, *mask;
  const uint8_t *vhide;
  unsigned long long *lab;
  int *vnode, *vcorner, *vstamp, *astamp, *aslot, *hvst;
  int *tv, *topp, *tnode, *tslot, *wstamp;
  unsigned long long *tie;
  float *fnx, *fny, *fnz;
  uint8_t *fhide;
  unsigned long long *claim;
  int *estamp, *remap, *remap_ser;
  int *node_flag, *node_child, *node_dirty, *leaf_node, *lvl_off, *lvl_node;
  int *node_parent, *node_height, *anc_stamp, *anc_pos, *anc, *lvl_list, *lvl_cnt;
  float *node_bb, *node_obb;
  int *slab, *scount;
  int *wlist, *cand, *win, *outer, *hv_in, *hv_out, *touched, *gathered, *vlist, *aff, *nlist, *nlist2;
  int *llist, *lstamp, *seed0, *seed1, *sstamp;
  uint8_t *fin;
  unsigned long long *key;
  uint8_t *valid;
  float *fade, *buf0, *buf1;
  long long *ctl;
};

const int TOUCH = PBVH_UpdateNormals | PBVH_UpdateBB | PBVH_UpdateDrawBuffers;

__device__ __forceinline__ int dn(int c) { return (c % 3 == 2) ? c - 2 : c + 1; }
__device__ __forceinline__ int dp(int c) { return (c % 3 == 0) ? c + 2 : c - 1; }
__device__ __forceinline__ unsigned long long ull(long long x) { return (unsigned long long)x; }

// Loads of words other threads update with atomics (performed at L2): atomic loads, so
// the topology workgroup never reads a stale cached copy across its barriers.
__device__ __forceinline__ long long aload(const long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long aload(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int aload(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void azero(long long *p) { atomicExch((unsigned long long *)p, 0ull); }

__device__ __forceinline__ float len2(const DV &g, int a, int b) {
  const float x = g.px[a] - g.px[b], y = g.py[a] - g.py[b], z = g.pz[a] - g.pz[b];
  return x * x + y * y + z * z;
}
__device__ __forceinline__ float edge_len2(const DV &g, int c) { return len2(g, g.tv[dn(c)], g.tv[dp(c)]); }

// splitmix64's finalizer (Steele, Lea, Flood, OOPSLA 2014)
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}
__device__ __forceinline__ unsigned long long new_label(unsigned long long a, unsigned long long b, int serial,
                                                        int round) {
  const unsigned long long lo = a < b ? a : b, hi = a < b ? b : a;
  return mix64(lo ^ mix64(hi ^ mix64(((unsigned long long)(unsigned)serial << 20) ^ (unsigned long long)(unsigned)round))) |
         (1ULL << 63);
}
__device__ __forceinline__ unsigned long long edge_hash33(unsigned long long a, unsigned long long b) {
  const unsigned long long lo = a < b ? a : b, hi = a < b ? b : a;
  return mix64(lo * 0x9E3779B97F4A7C15ULL ^ hi) >> 31;
}

// Priority within a round: the squared length's band (exponent and two mantissa
// bits), longest first for splits, shortest first for collapses, then the label hash.
__device__ __forceinline__ unsigned long long edge_key(const DV &g, int c, int collapse) {
  const unsigned band = __float_as_uint(edge_len2(g, c)) >> 21;
  const unsigned long long h = edge_hash33(g.lab[g.tv[dn(c)]], g.lab[g.tv[dp(c)]]);
  return ((unsigned long long)(collapse ? 0x3FFu - band : band) << 33) | h;
}

// The first corner of v's fan (rewound over its boundary), -1 for a dead vertex, -2
// past the walk bound.
__device__ __forceinline__ int ring_start(const DV &g, int v, int &bnd) {
  bnd = 0;
  const int c0 = g.vcorner[v];
  if (c0 < 0) return -1;
  int c = c0, steps = 0;
  for (;;) {
    const int x = g.topp[dn(c)];
    if (x < 0) {
      bnd = 1;
      break;
    }
    c = dn(x);
    if (c == c0) break;
    if (++steps > SG_DYN_RING) return -2;
  }
  return c;
}

// f(corner, i) for the corners around v in fan order; the count, or -1 past the bound
// (f may then have seen part of the ring).  Two laps: rewind, then walk (the second
// over lines the first brought into the cache).
template <class F>
__device__ __forceinline__ int ring_for(const DV &g, int v, int &bnd, F f) {
  const int start = ring_start(g, v, bnd);
  if (start == -1) return 0;
  if (start == -2) return -1;
  int c = start, n = 0;
  for (;;) {
    if (n >= SG_DYN_RING) return -1;
    f(c, n);
    n++;
    const int x = g.topp[dp(c)];
    if (x < 0) break;
    c = dp(x);
    if (c == start) break;
  }
  return n;
}

#define RB 16  // corners of a closed fan held in registers

// One lap from vcorner[v]: true when v's fan closes within RB corners, which are then
// b[0..n) in fan order (a closed fan starts at vcorner[v], as ring_start finds); false
// for an open, longer or dead fan.  Unused entries hold vcorner[v].  Callers index b
// only with unrolled (constant) indices, so it stays in registers.
__device__ __forceinline__ bool ring_closed(const DV &g, int v, int (&b)[RB], int &n) {
  const int c0 = g.vcorner[v];
  n = 0;
#pragma unroll
  for (int j = 0; j < RB; j++) b[j] = c0;
  if (c0 < 0) return false;
  int c = c0, state = 0;
#pragma unroll
  for (int j = 0; j < RB; j++) {
    if (state == 0) {
      b[j] = c;
      n = j + 1;
      const int x = g.topp[dp(c)];
      if (x < 0) {
        state = 2;
      } else {
        c = dp(x);
        if (c == c0) state = 1;
      }
    }
  }
  return state == 1;
}


// ring_for's count and boundary flag: one lap for a closed fan of at most RB corners
__device__ __forceinline__ int ring_count(const DV &g, int v, int &bnd) {
  const int c0 = g.vcorner[v];
  if (c0 >= 0) {
    int c = c0;
    for (int j = 1; j <= RB; j++) {
      const int x = g.topp[dp(c)];
      if (x < 0) break;
      c = dp(x);
      if (c == c0) {
        bnd = 0;
        return j;
      }
    }
  }
  return ring_for(g, v, bnd, [](int, int) {});
}

// closest_on_tri_to_point_v3 (math_geom.cc:966-1044)
__device__ void closest_on_tri(float r[3], const float p[3], const float v1[3], const float v2[3], const float v3[3]) {
  float ab[3], ac[3], ap[3], bp[3], cp[3];
  for (int i = 0; i < 3; i++) {
    ab[i] = v2[i] - v1[i];
    ac[i] = v3[i] - v1[i];
    ap[i] = p[i] - v1[i];
  }
  const float d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
  const float d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
  if (d1 <= 0.0f && d2 <= 0.0f) {
    for (int i = 0; i < 3; i++) r[i] = v1[i];
    return;
  }
  for (int i = 0; i < 3; i++) bp[i] = p[i] - v2[i];
  const float d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
  const float d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
  if (d3 >= 0.0f && d4 <= d3) {
    for (int i = 0; i < 3; i++) r[i] = v2[i];
    return;
  }
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float v = d1 / (d1 - d3);
    for (int i = 0; i < 3; i++) r[i] = v1[i] + ab[i] * v;
    return;
  }
  for (int i = 0; i < 3; i++) cp[i] = p[i] - v3[i];
  const float d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
  const float d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
  if (d6 >= 0.0f && d5 <= d6) {
    for (int i = 0; i < 3; i++) r[i] = v3[i];
    return;
  }
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = d2 / (d2 - d6);
    for (int i = 0; i < 3; i++) r[i] = v1[i] + ac[i] * w;
    return;
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int i = 0; i < 3; i++) r[i] = (v3[i] - v2[i]) * w + v2[i];
    return;
  }
  const float denom = 1.0f / (va + vb + vc);
  const float v = vb * denom;
  const float w = vc * denom;
  for (int i = 0; i < 3; i++) r[i] = (v1[i] + ab[i] * v) + ac[i] * w;
}

// project_plane_normalized_v3_v3v3 (math_vector.c:635-641)
__device__ __forceinline__ void project_plane(const DV &g, float out[3], const float p[3]) {
  const float n[3] = {g.vnx, g.vny, g.vnz};
  const float mul = p[0] * n[0] + p[1] * n[1] + p[2] * n[2];
  for (int i = 0; i < 3; i++) out[i] = p[i] + n[i] * -mul;
}

struct Range {
  float c[3], cp[3], r2;
};

// edge_queue_tri_in_sphere / _in_circle (pbvh_bmesh.c:773-805), the corners' positions
// in face order
__device__ bool tri_in_range_v(const DV &g, const float v0[3], const float v1[3], const float v2[3],
                               const Range &R) {
  float c[3];
  if (g.use_projected) {
    float p[3][3];
    project_plane(g, p[0], v0);
    project_plane(g, p[1], v1);
    project_plane(g, p[2], v2);
    closest_on_tri(c, R.cp, p[0], p[1], p[2]);
    const float x = R.cp[0] - c[0], y = R.cp[1] - c[1], z = R.cp[2] - c[2];
    return x * x + y * y + z * z <= R.r2;
  }
  closest_on_tri(c, R.c, v0, v1, v2);
  const float x = R.c[0] - c[0], y = R.c[1] - c[1], z = R.c[2] - c[2];
  return x * x + y * y + z * z <= R.r2;
}

__device__ bool tri_in_range(const DV &g, int f, const Range &R) {
  float v[3][3];
  for (int k = 0; k < 3; k++) {
    const int u = g.tv[3 * f + k];
    v[k][0] = g.px[u];
    v[k][1] = g.py[u];
    v[k][2] = g.pz[u];
  }
  return tri_in_range_v(g, v[0], v[1], v[2], R);
}

__device__ __forceinline__ bool front_ok_v(const DV &g, float x, float y, float z) {
  if (!g.use_frontface) return true;
  return !(x * g.vnx + y * g.vny + z * g.vnz < 0.0f);
}

__device__ __forceinline__ bool front_ok(const DV &g, int f) { return front_ok_v(g, g.fnx[f], g.fny[f], g.fnz[f]); }

__device__ __forceinline__ float len2_v(const float a[3], const float b[3]) {
  const float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
  return x * x + y * y + z * z;
}

// edge_queue_insert's filter (pbvh_bmesh.c:812-833)
__device__ __forceinline__ bool edge_insert_ok(const DV &g, int c) {
  const int a = g.tv[dn(c)], b = g.tv[dp(c)];
  if (g.has_mask && !(g.mask[a] < 1.0f || g.mask[b] < 1.0f)) return false;
  return !(g.vhide[a] || g.vhide[b]);
}

__device__ __forceinline__ int canon(const DV &g, int c) {
  const int x = g.topp[c];
  return (x >= 0 && x < c) ? x : c;
}

// The faces an edge's operation owns: a split its two faces; a collapse the faces
// around both vertices (the edge's faces once).  f(face) for each; false when a walk
// fails.
template <class F>
__device__ bool claim_set(const DV &g, int c, int collapse, F f) {
  if (!collapse) {
    f(c / 3);
    const int cB = g.topp[c];
    if (cB >= 0) f(cB / 3);
    return true;
  }
  const int p = g.tv[dn(c)], q = g.tv[dp(c)];
  int bnd;
  if (ring_for(g, p, bnd, [&](int k, int) { f(k / 3); }) < 0) return false;
  const int n = ring_for(g, q, bnd, [&](int k, int) {
    const int t = k / 3;
    if (g.tv[3 * t] != p && g.tv[3 * t + 1] != p && g.tv[3 * t + 2] != p) f(t);
  });
  return n >= 0;
}

// The round's registered edges: an LDS set of canonical corners (open addressing, a
// bounded probe); a key whose probe finds every slot taken by other keys uses the global
// stamp instead, and does so on every attempt (slots only fill during a round), so each
// edge registers exactly once.
#define ESET_BITS 13
#define ESET_N (1 << ESET_BITS)
#define ESET_PROBES 8
#define ESET_EMPTY 0xFFFFFFFFu

__device__ __forceinline__ bool edge_first(const DV &g, unsigned *eset, int k, int stamp) {
  const unsigned h = ((unsigned)k * 0x9E3779B1u) >> (32 - ESET_BITS);
  for (int i = 0; i < ESET_PROBES; i++) {
    const unsigned old = atomicCAS(&eset[(h + i) & (ESET_N - 1)], ESET_EMPTY, (unsigned)k);
    if (old == ESET_EMPTY) return true;
    if (old == (unsigned)k) return false;
  }
  return atomicExch(&g.estamp[k], stamp) != stamp;
}

__device__ void reg_edge(const DV &g, unsigned *eset, int c, int collapse, int stamp) {
  if (!edge_insert_ok(g, c)) return;
  const int k = canon(g, c);
  if (!edge_first(g, eset, k, stamp)) return;
  const unsigned long long slot = atomicAdd((unsigned long long *)&g.ctl[SGD_NCAND], 1ull);
  if (slot < (unsigned long long)g.ccap) g.cand[slot] = k;
}

// long_edge_queue_edge_add_recursive (pbvh_bmesh.c:849-903), iterative
__device__ int reg_even(const DV &g, unsigned *eset, int c0, float l0, float lim0, int stamp) {
  int sc[SG_DYN_STACK];
  float sl[SG_DYN_STACK], sm[SG_DYN_STACK];
  int top = 0;
  sc[0] = c0;
  sl[0] = l0;
  sm[0] = lim0;
  top = 1;
  while (top) {
    top--;
    const int c = sc[top];
    const float l2 = sl[top];
    float limit = sm[top];
    const int gc = g.topp[c];
    const int entered = gc >= 0 ? gc / 3 : c / 3;
    if (!front_ok(g, entered)) continue;
    // the entered face's other edges, read before the registration's stores
    const int ks[2] = {gc >= 0 ? dn(gc) : 0, gc >= 0 ? dp(gc) : 0};
    const float lo[2] = {gc >= 0 ? edge_len2(g, ks[0]) : 0.0f, gc >= 0 ? edge_len2(g, ks[1]) : 0.0f};
    reg_edge(g, eset, c, 0, stamp);
    if (gc < 0) continue;
    const float cmp = l2 * 1.2f;
    limit = limit * 1.6f;
    const float limit2 = limit * limit;
    for (int i = 0; i < 2; i++) {
      if (lo[i] > sg_max(cmp, limit2)) {
        if (top >= SG_DYN_STACK) return -1;
        sc[top] = ks[i];
        sl[top] = lo[i];
        sm[top] = limit;
        top++;
      }
    }
  }
  return 0;
}

// Whether collapsing the edge opposite c keeps the surface manifold (link condition,
// no inner edge between two boundaries, no ears, opposite vertices keep enough edges).
__device__ bool collapse_ok_walk(const DV &g, int c) {
  const int cB = g.topp[c];
  const int p = g.tv[dn(c)], q = g.tv[dp(c)], oA = g.tv[c];
  const int oB = cB >= 0 ? g.tv[cB] : -1;
  if (g.vnode[p] < 0 || g.vnode[q] < 0) return false;
  int bp, bq;
  if (ring_for(g, p, bp, [](int, int) {}) < 0 || ring_for(g, q, bq, [](int, int) {}) < 0) return false;
  if (cB >= 0 && bp && bq) return false;
  if (g.topp[dn(c)] < 0 && g.topp[dp(c)] < 0) return false;
  if (cB >= 0 && g.topp[dn(cB)] < 0 && g.topp[dp(cB)] < 0) return false;
  // common neighbours: each neighbour x of q, tested against p's neighbours
  int common = 0;
  bool bad = false;
  auto nbr_of_p = [&](int x) {
    bool hit = false;
    int b2, first = -1;
    ring_for(g, p, b2, [&](int k, int i) {
      if (i == 0) first = k;
      if (g.tv[dn(k)] == x) hit = true;
    });
    if (b2 && first >= 0 && g.tv[dp(first)] == x) hit = true;
    return hit;
  };
  int b3, first = -1;
  ring_for(g, q, b3, [&](int k, int i) {
    if (i == 0) first = k;
    const int x = g.tv[dn(k)];
    if (nbr_of_p(x)) {
      if (x != oA && x != oB) bad = true;
      common++;
    }
  });
  if (b3 && first >= 0) {
    const int x = g.tv[dp(first)];
    if (nbr_of_p(x)) {
      if (x != oA && x != oB) bad = true;
      common++;
    }
  }
  if (bad || common != (cB >= 0 ? 2 : 1)) return false;
  int bA;
  const int nA = ring_for(g, oA, bA, [](int, int) {});
  const int vA = nA < 0 ? SG_DYN_RING + 1 : nA + bA;
  if (vA - 1 < (bA ? 2 : 3)) return false;
  if (oB >= 0) {
    int bB;
    const int nB = ring_for(g, oB, bB, [](int, int) {});
    const int vB = nB < 0 ? SG_DYN_RING + 1 : nB + bB;
    if (vB - 1 < (bB ? 2 : 3)) return false;
  }
  return true;
}

// collapse_ok_walk's result; when both fans close within RB corners, p's neighbours are
// held in registers and each fan is walked once.
__device__ bool collapse_ok(const DV &g, int c) {
  const int p = g.tv[dn(c)], q = g.tv[dp(c)];
  if (g.vnode[p] < 0 || g.vnode[q] < 0) return false;
  int bP[RB], bQ[RB], nP, nQ;
  if (!ring_closed(g, p, bP, nP) || !ring_closed(g, q, bQ, nQ)) return collapse_ok_walk(g, c);
  const int cB = g.topp[c];
  const int oA = g.tv[c], oB = cB >= 0 ? g.tv[cB] : -1;
  if (g.topp[dn(c)] < 0 && g.topp[dp(c)] < 0) return false;
  if (cB >= 0 && g.topp[dn(cB)] < 0 && g.topp[dp(cB)] < 0) return false;
  int xP[RB], xQ[RB];
#pragma unroll
  for (int j = 0; j < RB; j++) {
    xP[j] = g.tv[dn(bP[j])];
    xQ[j] = g.tv[dn(bQ[j])];
  }
  // common neighbours (both fans closed: no boundary neighbour)
  int common = 0;
  bool bad = false;
#pragma unroll
  for (int i = 0; i < RB; i++) {
    if (i >= nQ) break;
    const int x = xQ[i];
    bool hit = false;
#pragma unroll
    for (int j = 0; j < RB; j++) hit = hit || (j < nP && xP[j] == x);
    if (hit) {
      if (x != oA && x != oB) bad = true;
      common++;
    }
  }
  if (bad || common != (cB >= 0 ? 2 : 1)) return false;
  int bA;
  const int nA = ring_for(g, oA, bA, [](int, int) {});
  const int vA = nA < 0 ? SG_DYN_RING + 1 : nA + bA;
  if (vA - 1 < (bA ? 2 : 3)) return false;
  if (oB >= 0) {
    int bB;
    const int nB = ring_for(g, oB, bB, [](int, int) {});
    const int vB = nB < 0 ? SG_DYN_RING + 1 : nB + bB;
    if (vB - 1 < (bB ? 2 : 3)) return false;
  }
  return true;
}

// ---------------------------------------------------------------------------------
// search: a thread per leaf (spec S1 on the leaf box, fully hidden leaves skipped);
// gathered leaves get `flags`
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(DY_THREADS) void kd_search(DV g, float cx, float cy, float cz, float r2, int flags) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  bool pass = false;
  int node = 0;
  if (l < g.L) {
    node = g.leaf_node[l];
    const int fl = g.node_flag[node];
    const BB6 b = bb_load(g.node_bb + 8 * (size_t)node);
    pass = !(fl & PBVH_FullyHidden) && sphere_box(b, cx, cy, cz, r2);
  }
  const unsigned long long m = __ballot(pass);
  if (m == 0) return;
  int base = 0;
  if (lane_id() == 0) base = (int)atomicAdd((unsigned long long *)&g.ctl[SGD_NGATHER], (unsigned long long)__popcll(m));
  base = __shfl(base, 0);
  if (pass) {
    g.gathered[base + mask_prefix(m)] = node;
    atomicOr(&g.node_flag[node], flags);
  }
}

// working set: the live faces of leaves flagged UpdateTopology (not fully hidden)
__global__ __launch_bounds__(DY_THREADS) void kd_wbuild(DV g, int wser, int gate) {
  if (gate && g.ctl[SGD_NGATHER] == 0) return;
  for (int l = blockIdx.x; l < g.L; l += gridDim.x) {
    const int node = g.leaf_node[l];
    const int fl = g.node_flag[node];
    if (!(fl & PBVH_UpdateTopology) || (fl & PBVH_FullyHidden)) continue;
    if (threadIdx.x == 0) atomicAdd((unsigned long long *)&g.ctl[SGD_NLIST], 1ull);
    const int n = g.scount[node];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const int f = g.slab[(size_t)node * SG_DYN_SLAB + i];
      if (f < 0) continue;
      if (atomicExch(&g.wstamp[f], wser) != wser) {
        const unsigned long long s = atomicAdd((unsigned long long *)&g.ctl[SGD_WN], 1ull);
        g.wlist[s] = f;
      }
    }
  }
}

__device__ __forceinline__ void slab_append(const DV &g, int node, int f) {
  const int pos = atomicAdd(&g.scount[node], 1);
  g.slab[(size_t)node * SG_DYN_SLAB + pos] = f;
  g.tnode[f] = node;
  g.tslot[f] = pos;
}

// a leaf touched by a split: TOUCH set and FullyHidden cleared (face create /
// remove, pbvh_bmesh.c:495-519, :634-672); the atomics only when a bit changes
__device__ __forceinline__ void touch_leaf(const DV &g, int leaf) {
  const int f = g.node_flag[leaf];
  if ((f & TOUCH) != TOUCH) atomicOr(&g.node_flag[leaf], TOUCH);
  if (f & PBVH_FullyHidden) atomicAnd(&g.node_flag[leaf], ~PBVH_FullyHidden);
}

__device__ __forceinline__ void hv_add(const DV &g, int v, int stamp, int *hv_out) {
  if (v < 0) return;
  if (atomicExch(&g.hvst[v], stamp) != stamp) {
    const unsigned long long s = atomicAdd((unsigned long long *)&g.ctl[SGD_HV2], 1ull);
    hv_out[s] = v;
  }
}

// a working-set face in range: alive, front facing, meeting the sphere / circle
__device__ __forceinline__ bool face_in(const DV &g, int f, const Range &R) {
  return g.tnode[f] >= 0 && front_ok(g, f) && tri_in_range(g, f, R);
}

// A face the round changed (rewritten, created, or moved by a collapse) joins the
// working set when a split made it, is tested again, and joins the scan list if in range.
__device__ __forceinline__ int scan_update(const DV &g, int f, const Range &R, int lser, int *out) {
  const bool in = face_in(g, f, R);
  g.fin[f] = in ? 1 : 0;
  if (in && g.lstamp[f] != lser) {
    g.lstamp[f] = lser;
    *out = f;
    return 1;
  }
  return 0;
}

// A seed face of a phase: an edge shorter than the limit (collapse) or longer (split).
// Positions move only where faces are rewritten or a collapse moves a vertex, so a face
// seeds the next round iff it seeded this one untouched, or a change left it seeding.
__device__ __forceinline__ bool face_seeds(const DV &g, int f, float limit2, int collapse) {
  bool any = false;
  for (int k = 0; k < 3; k++) {
    const float l2 = edge_len2(g, 3 * f + k);
    any = any || (collapse ? l2 < limit2 : l2 > limit2);
  }
  return any;
}

__device__ __forceinline__ void seed_push(const DV &g, int f, int stamp, int *next) {
  if (atomicExch(&g.sstamp[f], stamp) != stamp) {
    const unsigned long long s = atomicAdd((unsigned long long *)&g.ctl[SGD_NS2], 1ull);
    next[s] = f;
  }
}

// Split the edge opposite cA at a new vertex m (pbvh_bmesh.c:1093-1209): A (p, q, oA)
// in its winding becomes A' = (p, m, oA) in its slot and NA = (m, q, oA); B likewise.
// Outer corners keep their old neighbour ids and record where they moved (remap).
// Everything is read before the first store (a store may alias any array), and the four
// faces' range and seed tests run on the positions held in registers.
__device__ void split_apply(const DV &g, int wi, int cA, int serial, int round, int stamp, int wser, int lser,
                            const Range &R, int *hv_out, float limit2, int *seed_next) {
  const int cB = g.topp[cA];
  const int A = cA / 3, B = cB >= 0 ? cB / 3 : -1;
  const int p = g.tv[dn(cA)], q = g.tv[dp(cA)], oA = g.tv[cA];
  const int oB = cB >= 0 ? g.tv[cB] : -1;
  const int yA = g.topp[dp(cA)], xA = g.topp[dn(cA)];
  const int yB = cB >= 0 ? g.topp[dp(cB)] : -1, xB = cB >= 0 ? g.topp[dn(cB)] : -1;
  const int oldAy = dp(cA), oldAx = dn(cA), oldBy = cB >= 0 ? dp(cB) : -1, oldBx = cB >= 0 ? dn(cB) : -1;
  const float P[3] = {g.px[p], g.py[p], g.pz[p]}, Q[3] = {g.px[q], g.py[q], g.pz[q]};
  const float OA[3] = {g.px[oA], g.py[oA], g.pz[oA]};
  float OB[3] = {0.0f, 0.0f, 0.0f};
  if (oB >= 0) {
    OB[0] = g.px[oB];
    OB[1] = g.py[oB];
    OB[2] = g.pz[oB];
  }
  const float npx = g.nx[p], npy = g.ny[p], npz = g.nz[p], nqx = g.nx[q], nqy = g.ny[q], nqz = g.nz[q];
  const float mp = g.has_mask ? g.mask[p] : 0.0f, mq = g.has_mask ? g.mask[q] : 0.0f;
  const unsigned long long lp = g.lab[p], lq = g.lab[q], loA = g.lab[oA], loB = oB >= 0 ? g.lab[oB] : 0ull;
  const int leafA = g.tnode[A], leafB = B >= 0 ? g.tnode[B] : -1;
  const float fAx = g.fnx[A], fAy = g.fny[A], fAz = g.fnz[A];
  const float fBx = B >= 0 ? g.fnx[B] : 0.0f, fBy = B >= 0 ? g.fny[B] : 0.0f, fBz = B >= 0 ? g.fnz[B] : 0.0f;
  const uint8_t hA = g.fhide[A], hB = B >= 0 ? g.fhide[B] : 0;
  const int vs[4] = {p, q, oA, oB};
  int vc[4];
  for (int i = 0; i < 4; i++) vc[i] = vs[i] >= 0 ? g.vcorner[vs[i]] : -1;
  const int m = (int)atomicAdd((unsigned long long *)&g.ctl[SGD_V], 1ull);
  const int NA = (int)atomicAdd((unsigned long long *)&g.ctl[SGD_T], B >= 0 ? 2ull : 1ull);
  const int NB = B >= 0 ? NA + 1 : -1;
  const int own[4] = {A, NA, B, NB};
  int ls[4];
  for (int i = 0; i < 4; i++) ls[i] = own[i] >= 0 ? g.lstamp[own[i]] : lser;
  // the new vertex
  const float M[3] = {0.5f * (P[0] + Q[0]), 0.5f * (P[1] + Q[1]), 0.5f * (P[2] + Q[2])};
  float nx = 0.5f * (npx + nqx), ny = 0.5f * (npy + nqy), nz = 0.5f * (npz + nqz);
  sg_normalize(nx, ny, nz);
  // the four faces: in range (alive, front facing, meeting the sphere / circle) and seeds
  bool in[4], sd[4];
  {
    const bool frA = front_ok_v(g, fAx, fAy, fAz), frB = front_ok_v(g, fBx, fBy, fBz);
    in[0] = frA && tri_in_range_v(g, P, M, OA, R);
    in[1] = frA && tri_in_range_v(g, M, Q, OA, R);
    in[2] = B >= 0 && frB && tri_in_range_v(g, Q, M, OB, R);
    in[3] = B >= 0 && frB && tri_in_range_v(g, M, P, OB, R);
    const float lpm = len2_v(P, M), lmq = len2_v(M, Q), lmA = len2_v(M, OA);
    sd[0] = lpm > limit2 || lmA > limit2 || len2_v(OA, P) > limit2;
    sd[1] = lmq > limit2 || len2_v(Q, OA) > limit2 || lmA > limit2;
    const float lmB = len2_v(M, OB);
    sd[2] = lmq > limit2 || lmB > limit2 || len2_v(OB, Q) > limit2;
    sd[3] = lpm > limit2 || len2_v(P, OB) > limit2 || lmB > limit2;
  }
  g.px[m] = M[0];
  g.py[m] = M[1];
  g.pz[m] = M[2];
  g.nx[m] = nx;
  g.ny[m] = ny;
  g.nz[m] = nz;
  if (g.has_mask) g.mask[m] = 0.5f * (mp + mq);
  ((uint8_t *)g.vhide)[m] = 0;
  g.lab[m] = new_label(lp, lq, serial, round);
  g.vnode[m] = (B >= 0 && loB < loA) ? leafB : leafA;
  int *tv = g.tv, *tp = g.topp;
  tv[3 * A] = p;
  tv[3 * A + 1] = m;
  tv[3 * A + 2] = oA;
  tv[3 * NA] = m;
  tv[3 * NA + 1] = q;
  tv[3 * NA + 2] = oA;
  tp[3 * A] = 3 * NA + 1;
  tp[3 * NA + 1] = 3 * A;
  tp[3 * A + 1] = yA;
  tp[3 * NA] = xA;
  g.remap[oldAy] = 3 * A + 1;
  g.remap_ser[oldAy] = stamp;
  g.remap[oldAx] = 3 * NA;
  g.remap_ser[oldAx] = stamp;
  int4 out = make_int4(3 * A + 1, 3 * NA, -1, -1);
  if (B >= 0) {
    tv[3 * B] = q;
    tv[3 * B + 1] = m;
    tv[3 * B + 2] = oB;
    tv[3 * NB] = m;
    tv[3 * NB + 1] = p;
    tv[3 * NB + 2] = oB;
    tp[3 * B] = 3 * NB + 1;
    tp[3 * NB + 1] = 3 * B;
    tp[3 * B + 1] = yB;
    tp[3 * NB] = xB;
    g.remap[oldBy] = 3 * B + 1;
    g.remap_ser[oldBy] = stamp;
    g.remap[oldBx] = 3 * NB;
    g.remap_ser[oldBx] = stamp;
    out.z = 3 * B + 1;
    out.w = 3 * NB;
    tp[3 * A + 2] = 3 * NB + 2;
    tp[3 * NB + 2] = 3 * A + 2;
    tp[3 * NA + 2] = 3 * B + 2;
    tp[3 * B + 2] = 3 * NA + 2;
  } else {
    tp[3 * A + 2] = -1;
    tp[3 * NA + 2] = -1;
  }
  ((int4 *)g.outer)[wi] = out;
  g.fnx[NA] = fAx;
  g.fny[NA] = fAy;
  g.fnz[NA] = fAz;
  g.fhide[NA] = hA;
  slab_append(g, leafA, NA);
  touch_leaf(g, leafA);
  if (B >= 0) {
    g.fnx[NB] = fBx;
    g.fny[NB] = fBy;
    g.fnz[NB] = fBz;
    g.fhide[NB] = hB;
    slab_append(g, leafB, NB);
    touch_leaf(g, leafB);
  }
  for (int i = 0; i < 4; i++) {
    const int v = vs[i];
    if (v < 0) continue;
    const int f = vc[i] / 3;
    if (f == A)
      g.vcorner[v] = (v == p) ? 3 * A : (v == q) ? 3 * NA + 1 : 3 * A + 2;
    else if (B >= 0 && f == B)
      g.vcorner[v] = (v == q) ? 3 * B : (v == p) ? 3 * NB + 1 : 3 * B + 2;
  }
  g.vcorner[m] = 3 * A + 1;
  // this split's faces (owned here) join the working set; in-range ones the scan list,
  // seeding ones the next round's seeds
  int fs[4], nf = 0;
  for (int i = 0; i < 4; i++) {
    const int f = own[i];
    if (f < 0) continue;
    g.wstamp[f] = wser;
    g.fin[f] = in[i] ? 1 : 0;
    if (in[i] && ls[i] != lser) {
      g.lstamp[f] = lser;
      fs[nf++] = f;
    }
    if (in[i] && sd[i]) seed_push(g, f, stamp, seed_next);
  }
  if (nf) {
    const int base = (int)atomicAdd((unsigned long long *)&g.ctl[SGD_LN], (unsigned long long)nf);
    for (int i = 0; i < nf; i++) g.llist[base + i] = fs[i];
  }
  hv_add(g, oA, stamp, hv_out);
  hv_add(g, oB, stamp, hv_out);
}

// Collapse the edge opposite cA (pbvh_bmesh.c:1264-1429): the lower-mask vertex goes
// (v1 = the lower label), the other moves to the midpoint.
__device__ void collapse_apply(const DV &g, int cA) {
  const int cB = g.topp[cA];
  const int A = cA / 3, B = cB >= 0 ? cB / 3 : -1;
  const int p = g.tv[dn(cA)], q = g.tv[dp(cA)], oA = g.tv[cA];
  const int oB = cB >= 0 ? g.tv[cB] : -1;
  const int v1 = g.lab[p] < g.lab[q] ? p : q, v2 = v1 == p ? q : p;
  const float m1 = g.has_mask ? g.mask[v1] : 0.0f, m2 = g.has_mask ? g.mask[v2] : 0.0f;
  const int vdel = (m1 < m2) ? v1 : v2, vcon = vdel == v1 ? v2 : v1;
  const int xA = g.topp[dn(cA)], yA = g.topp[dp(cA)];
  const int xB = cB >= 0 ? g.topp[dn(cB)] : -1, yB = cB >= 0 ? g.topp[dp(cB)] : -1;
  atomicOr(&g.node_flag[g.tnode[A]], TOUCH);
  if (B >= 0) atomicOr(&g.node_flag[g.tnode[B]], TOUCH);
  // re-point vdel's other corners (walked before any link changes)
  int bnd;
  ring_for(g, vdel, bnd, [&](int k, int) {
    const int f = k / 3;
    if (f != A && f != B) g.tv[k] = vcon;
  });
  if (xA >= 0) g.topp[xA] = yA;
  if (yA >= 0) g.topp[yA] = xA;
  if (B >= 0) {
    if (xB >= 0) g.topp[xB] = yB;
    if (yB >= 0) g.topp[yB] = xB;
  }
  const int fA = g.vcorner[oA] / 3;
  if (fA == A || fA == B) g.vcorner[oA] = xA >= 0 ? dn(xA) : dp(yA);
  if (B >= 0) {
    const int fB = g.vcorner[oB] / 3;
    if (fB == A || fB == B) g.vcorner[oB] = xB >= 0 ? dn(xB) : dp(yB);
  }
  const int fc = g.vcorner[vcon] / 3;
  if (fc == A || fc == B) g.vcorner[vcon] = xA >= 0 ? dp(xA) : dn(yA);
  g.px[vcon] = 0.5f * (g.px[vcon] + g.px[vdel]);
  g.py[vcon] = 0.5f * (g.py[vcon] + g.py[vdel]);
  g.pz[vcon] = 0.5f * (g.pz[vcon] + g.pz[vdel]);
  float nx = g.nx[vcon] + g.nx[vdel], ny = g.ny[vcon] + g.ny[vdel], nz = g.nz[vcon] + g.nz[vdel];
  sg_normalize(nx, ny, nz);
  g.nx[vcon] = nx;
  g.ny[vcon] = ny;
  g.nz[vcon] = nz;
  const int fs[2] = {A, B};
  for (int i = 0; i < 2; i++) {
    const int f = fs[i];
    if (f < 0) continue;
    g.slab[(size_t)g.tnode[f] * SG_DYN_SLAB + g.tslot[f]] = -1;
    g.tnode[f] = -1;
  }
  g.vnode[vdel] = -1;
  g.vcorner[vdel] = -1;
  const int s = (int)atomicAdd((unsigned long long *)&g.ctl[SGD_NTOUCH], 1ull);
  g.touched[3 * s] = vcon;
  g.touched[3 * s + 1] = oA;
  g.touched[3 * s + 2] = oB;
}

// ---------------------------------------------------------------------------------
// the topology update: the working set's in-range faces (a thread per face), then one
// workgroup runs every round of both phases, its stages separated by barriers
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(DY_THREADS) void kd_wrange(DV g, Range R, int lser, int gate) {
  if (gate && g.ctl[SGD_NGATHER] == 0) return;
  const int wn = (int)g.ctl[SGD_WN];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < wn; i += gridDim.x * blockDim.x) {
    const int f = g.wlist[i];
    const bool in = face_in(g, f, R);
    g.fin[f] = in ? 1 : 0;
    if (in) {
      g.lstamp[f] = lser;
      const unsigned long long s = atomicAdd((unsigned long long *)&g.ctl[SGD_LN], 1ull);
      g.llist[s] = f;
    }
  }
}

#define TW_THREADS 768

__global__ __launch_bounds__(TW_THREADS) void kd_topology(DV g0, int mode, Range R, float min_len, float max_len,
                                                           int gate, int stamp0, int wser, int lser, int tag_base) {
  const int tid = threadIdx.x, nt = blockDim.x;
  // a dab's update runs only when its range search gathered a leaf
  if (gate && g0.ctl[SGD_NGATHER] == 0) return;
  // the counters live in LDS while the workgroup runs (no other wave touches them):
  // its many same-word atomics stay out of the L2
  __shared__ long long sctl[SGD_COUNT];
  __shared__ unsigned eset[ESET_N];
  for (int i = tid; i < SGD_COUNT; i += nt) sctl[i] = g0.ctl[i];
  __syncthreads();
  DV g = g0;
  g.ctl = sctl;
  const int serial = (int)sctl[SGD_SERIAL];
  long long st_ops[2] = {0, 0}, st_rounds[2] = {0, 0}, st_cand[2] = {0, 0};
  int stamp = stamp0;
  for (int ph = 0; ph < 2; ph++) {
    const int collapse = ph == 0;
    if (!(mode & (collapse ? PBVH_Collapse : PBVH_Subdivide))) continue;
    const float limit = collapse ? min_len : max_len;
    const float limit2 = limit * limit;
    int *hv_in = g.hv_in, *hv_out = g.hv_out;  // the two lists, swapped per round
    int *seed_cur = g.seed0, *seed_next = g.seed1;
    if (tid == 0) {
      azero(&g.ctl[SGD_HV]);
      azero(&g.ctl[SGD_HV2]);
      azero(&g.ctl[SGD_NS]);
      azero(&g.ctl[SGD_NS2]);
    }
    __syncthreads();
    for (int round = 0; round < SG_DYN_MAX_ROUNDS; round++) {
      stamp++;
      for (int i = tid; i < ESET_N; i += nt) eset[i] = ESET_EMPTY;
      if (tid == 0) {
        azero(&g.ctl[SGD_NCAND]);
        azero(&g.ctl[SGD_NVALID]);
        azero(&g.ctl[SGD_NWIN]);
        azero(&g.ctl[SGD_NTOUCH]);
      }
      __syncthreads();
      // candidates: the queue's edges of in-range working-set faces, and the long
      // edges of the last round's high-valence opposite vertices (pbvh_bmesh.c:1191-1198).
      // The first round of a phase scans every in-range face; later rounds only the
      // seed faces (with an edge past the limit) the previous round kept or changed.
      const int ln = round == 0 ? (int)aload(&g.ctl[SGD_LN]) : (int)aload(&g.ctl[SGD_NS]);
      const int *scan = round == 0 ? g.llist : seed_cur;
      const int hn = collapse ? 0 : (int)aload(&g.ctl[SGD_HV]);
      for (int i = tid; i < ln + hn; i += nt) {
        if (i < ln) {
          const int f = scan[i];
          if (g.tnode[f] < 0 || !g.fin[f]) continue;
          bool seeds = false;
          float l2s[3];
          for (int k = 0; k < 3; k++) l2s[k] = edge_len2(g, 3 * f + k);  // before any registration's stores
          for (int k = 0; k < 3; k++) {
            const int c = 3 * f + k;
            const float l2 = l2s[k];
            if (collapse) {
              if (l2 < limit2) {
                seeds = true;
                reg_edge(g, eset, c, 1, stamp);
              }
            } else if (l2 > limit2) {
              seeds = true;
              if (reg_even(g, eset, c, l2, limit, stamp) < 0) atomicOr((unsigned long long *)&g.ctl[SGD_FLAGS], 2ull);
            }
          }
          if (seeds) seed_push(g, f, stamp, seed_next);
        } else {
          const int v = hv_in[i - ln];
          int b0;
          const int n0 = ring_count(g, v, b0);
          if (n0 < 0 || n0 + b0 <= 8) continue;
          int bnd;
          ring_for(g, v, bnd, [&](int k, int j) {
            const int e1 = dp(k);
            if (edge_len2(g, e1) > limit2) reg_edge(g, eset, e1, 0, stamp);
            if (bnd && j == 0) {
              const int e2 = dn(k);
              if (edge_len2(g, e2) > limit2) reg_edge(g, eset, e2, 0, stamp);
            }
          });
        }
      }
      __syncthreads();
      const long long nc = aload(&g.ctl[SGD_NCAND]);
      st_cand[ph] += nc;
      if (nc == 0) break;
      if (nc > g.ccap) {
        if (tid == 0) atomicOr((unsigned long long *)&g.ctl[SGD_FLAGS], 1ull);
        break;
      }
      st_rounds[ph]++;
      const int n = (int)nc;
      // validity, keys, claims
      for (int i = tid; i < n; i += nt) {
        const int c = g.cand[i];
        const int p = g.tv[dn(c)], q = g.tv[dp(c)];
        bool ok = g.vnode[p] >= 0 && g.vnode[q] >= 0 && g.tnode[c / 3] >= 0;
        if (ok && collapse) ok = collapse_ok(g, c);
        if (ok && !collapse) {
          const int cB = g.topp[c];
          ok = aload(&g.scount[g.tnode[c / 3]]) <= SG_DYN_SLAB / 2 &&
               (cB < 0 || (g.tnode[cB / 3] >= 0 && aload(&g.scount[g.tnode[cB / 3]]) <= SG_DYN_SLAB / 2));
        }
        g.valid[i] = ok ? 1 : 0;
        if (!ok) continue;
        atomicAdd((unsigned long long *)&g.ctl[SGD_NVALID], 1ull);
        // the round's tag above the key: claims of earlier rounds are smaller, so no
        // face's claim needs clearing
        const unsigned long long kk = ((unsigned long long)(stamp - tag_base) << 43) | edge_key(g, c, collapse);
        g.key[i] = kk;
        claim_set(g, c, collapse, [&](int t) {
          const unsigned long long old = atomicMax(&g.claim[t], kk);
          if (old == kk) atomicMax(&g.tie[t], kk);
        });
      }
      __syncthreads();
      if (!collapse) {
        // room for the round's new vertices and faces
        const long long nv = aload(&g.ctl[SGD_NVALID]);
        if (aload(&g.ctl[SGD_V]) + nv > g.Vcap || aload(&g.ctl[SGD_T]) + 2 * nv > g.Tcap) {
          if (tid == 0) atomicOr((unsigned long long *)&g.ctl[SGD_FLAGS], 1ull);
          break;
        }
      }
      // winners: the unshared largest key on every owned face
      for (int i = tid; i < n; i += nt) {
        if (!g.valid[i]) continue;
        const unsigned long long kk = g.key[i];
        bool own = true;
        claim_set(g, g.cand[i], collapse, [&](int t) { own = own && aload(&g.claim[t]) == kk && aload(&g.tie[t]) != kk; });
        if (own) {
          const unsigned long long s = atomicAdd((unsigned long long *)&g.ctl[SGD_NWIN], 1ull);
          g.win[s] = g.cand[i];
        }
      }
      __syncthreads();
      const int nw = (int)aload(&g.ctl[SGD_NWIN]);
      for (int i = tid; i < nw; i += nt) {
        if (collapse)
          collapse_apply(g, g.win[i]);
        else
          split_apply(g, i, g.win[i], serial, round, stamp, wser, lser, R, hv_out, limit2, seed_next);
      }
      __syncthreads();
      if (!collapse) {
        // outer edges meet their neighbours
        for (int i = tid; i < nw; i += nt) {
          const int4 o = ((const int4 *)g.outer)[i];
          const int ks[4] = {o.x, o.y, o.z, o.w};
          for (int j = 0; j < 4; j++) {
            const int k = ks[j];
            if (k < 0) continue;
            const int y = g.topp[k];
            if (y < 0) continue;
            if (g.remap_ser[y] == stamp)
              g.topp[k] = g.remap[y];
            else
              g.topp[y] = k;
          }
        }
      } else {
        // owners after the collapses: a vertex whose owner leaf lost its last face with
        // it goes to the smallest-numbered leaf among its faces; around the kept vertex
        // every leaf is marked and every working-set face tested again (it moved)
        const int nto = 3 * (int)aload(&g.ctl[SGD_NTOUCH]);
        for (int i = tid; i < nto; i += nt) {
          const int v = g.touched[i];
          if (v < 0 || g.vcorner[v] < 0) continue;
          const bool kept = (i % 3) == 0;
          const int own = g.vnode[v];
          bool has = false;
          int best = -1, bnd;
          ring_for(g, v, bnd, [&](int k, int) {
            const int f = k / 3;
            const int leaf = g.tnode[f];
            if (leaf < 0) return;
            if (kept) {
              atomicOr(&g.node_flag[leaf], TOUCH);
              if (g.wstamp[f] == wser) {
                int out;
                if (scan_update(g, f, R, lser, &out)) {
                  const unsigned long long s = atomicAdd((unsigned long long *)&g.ctl[SGD_LN], 1ull);
                  g.llist[s] = out;
                }
                if (g.fin[f] && face_seeds(g, f, limit2, 1)) seed_push(g, f, stamp, seed_next);
              }
            }
            has = has || leaf == own;
            if (best < 0 || leaf < best) best = leaf;
          });
          if (!has) {
            if (own >= 0) atomicOr(&g.node_flag[own], PBVH_UpdateBB | PBVH_UpdateDrawBuffers);
            g.vnode[v] = best;
            if (best >= 0) atomicOr(&g.node_flag[best], PBVH_UpdateBB | PBVH_UpdateDrawBuffers);
          }
        }
      }
      __syncthreads();
      st_ops[ph] += nw;
      if (nw == 0) break;
      int *t = hv_in;
      hv_in = hv_out;
      hv_out = t;
      t = seed_cur;
      seed_cur = seed_next;
      seed_next = t;
      if (tid == 0) {
        atomicExch((unsigned long long *)&g.ctl[SGD_HV], (unsigned long long)aload(&g.ctl[SGD_HV2]));
        azero(&g.ctl[SGD_HV2]);
        atomicExch((unsigned long long *)&g.ctl[SGD_NS], (unsigned long long)aload(&g.ctl[SGD_NS2]));
        azero(&g.ctl[SGD_NS2]);
      }
      __syncthreads();
    }
    __syncthreads();
  }
  if (tid == 0) {
    sctl[SGD_ST + 0] = st_ops[1];
    sctl[SGD_ST + 1] = st_ops[0];
    sctl[SGD_ST + 2] = st_rounds[1];
    sctl[SGD_ST + 3] = st_rounds[0];
    sctl[SGD_ST + 4] = st_cand[1];
    sctl[SGD_ST + 5] = st_cand[0];
    sctl[SGD_ST + 6] = sctl[SGD_NLIST];
    sctl[SGD_SERIAL] = serial + 1;
  }
  __syncthreads();
  for (int i = tid; i < SGD_COUNT; i += nt) g0.ctl[i] = sctl[i];
}

__global__ void kd_clear_topology_flags(DV g, int gate) {
  if (gate && g.ctl[SGD_NGATHER] == 0) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.N; i += gridDim.x * blockDim.x)
    g.node_flag[i] &= ~PBVH_UpdateTopology;
}

// ---------------------------------------------------------------------------------
// the brush
// ---------------------------------------------------------------------------------
// the gathered leaves' own vertices, each once
__global__ __launch_bounds__(LB_THREADS) void kd_unique(DV g, int stamp) {
  const int ng = (int)g.ctl[SGD_NGATHER];
  for (int gi = blockIdx.x; gi < ng; gi += gridDim.x) {
    const int node = g.gathered[gi];
    const int n = g.scount[node];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const int f = g.slab[(size_t)node * SG_DYN_SLAB + i];
      if (f < 0) continue;
      for (int k = 0; k < 3; k++) {
        const int v = g.tv[3 * f + k];
        const bool add = g.vnode[v] == node && atomicExch(&g.vstamp[v], stamp) != stamp;
        // one list reservation per wave (its first lane that adds)
        const unsigned long long m = __ballot(add);
        if (m == 0) continue;
        const int leader = __ffsll((long long)m) - 1;
        int base = 0;
        if (lane_id() == leader) base = (int)atomicAdd((unsigned long long *)&g.ctl[SGD_NVL], (unsigned long long)__popcll(m));
        base = __shfl(base, leader);
        if (add) g.vlist[base + mask_prefix(m)] = v;
      }
    }
  }
}

__global__ __launch_bounds__(DY_THREADS) void kd_area(DV g, SgDabK k) {
  const int n = (int)g.ctl[SGD_NVL];
  long long sx = 0, sy = 0, sz = 0;
  int na = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = g.vlist[i];
    if (g.respect_hide && g.vhide[v]) continue;
    const float dx = k.cx - g.px[v], dy = k.cy - g.py[v], dz = k.cz - g.pz[v];
    if (dx * dx + dy * dy + dz * dz <= k.nr2) {
      sx += __double2ll_rn((double)g.nx[v] * 4294967296.0);
      sy += __double2ll_rn((double)g.ny[v] * 4294967296.0);
      sz += __double2ll_rn((double)g.nz[v] * 4294967296.0);
      na++;
    }
  }
  sx = wave_sum_ll(sx);
  sy = wave_sum_ll(sy);
  sz = wave_sum_ll(sz);
  na = wave_sum_i(na);
  if (lane_id() == 0 && na) {
    atomicAdd((unsigned long long *)&g.ctl[SGD_AREA_X], ull(sx));
    atomicAdd((unsigned long long *)&g.ctl[SGD_AREA_Y], ull(sy));
    atomicAdd((unsigned long long *)&g.ctl[SGD_AREA_Z], ull(sz));
    atomicAdd((unsigned long long *)&g.ctl[SGD_NAREA], (unsigned long long)na);
  }
}

// spec S2 test, S3/S4 factor, S6 Draw / S7 Inflate / S9 Grab; Smooth keeps its factor
__global__ __launch_bounds__(DY_THREADS) void kd_brush(DV g, SgDabK k, const float *curve, int stamp) {
  const int n = (int)g.ctl[SGD_NVL];
  float ox = 0.0f, oy = 0.0f, oz = 0.0f;
  if (k.tool == SG_TOOL_DRAW) {
    float ax = (float)((double)g.ctl[SGD_AREA_X] * (1.0 / 4294967296.0));
    float ay = (float)((double)g.ctl[SGD_AREA_Y] * (1.0 / 4294967296.0));
    float az = (float)((double)g.ctl[SGD_AREA_Z] * (1.0 / 4294967296.0));
    sg_normalize(ax, ay, az);
    ox = (ax * k.r) * k.bstr;
    oy = (ay * k.r) * k.bstr;
    oz = (az * k.r) * k.bstr;
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = g.vlist[i];
    if (g.respect_hide && g.vhide[v]) continue;
    const float x = g.px[v], y = g.py[v], z = g.pz[v];
    const float dx = k.cx - x, dy = k.cy - y, dz = k.cz - z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 <= k.r2)) continue;
    const int s = (int)atomicAdd((unsigned long long *)&g.ctl[SGD_NAFF], 1ull);
    g.aff[s] = v;
    const float mk = g.has_mask ? g.mask[v] : 0.0f;
    const float f = sg_falloff(k, curve, sqrtf(d2)) * (1.0f - mk);
    if (k.tool == SG_TOOL_SMOOTH) {
      g.fade[s] = k.bstr * f;
      g.astamp[v] = stamp;
      g.aslot[v] = s;
      continue;
    }
    float rx, ry, rz;
    if (k.tool == SG_TOOL_DRAW) {
      rx = x + ox * f;
      ry = y + oy * f;
      rz = z + oz * f;
    } else if (k.tool == SG_TOOL_INFLATE) {
      const float sc = (k.bstr * f) * k.r;
      rx = x + g.nx[v] * sc;
      ry = y + g.ny[v] * sc;
      rz = z + g.nz[v] * sc;
    } else {
      const float fade = k.bstr * f;
      rx = x + k.gx * fade;
      ry = y + k.gy * fade;
      rz = z + k.gz * fade;
    }
    g.px[v] = rx;
    g.py[v] = ry;
    g.pz[v] = rz;
  }
}

__global__ __launch_bounds__(DY_THREADS) void kd_smooth_init(DV g) {
  const int n = (int)g.ctl[SGD_NAFF];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = g.aff[i];
    g.buf0[3 * (size_t)i] = g.px[v];
    g.buf0[3 * (size_t)i + 1] = g.py[v];
    g.buf0[3 * (size_t)i + 2] = g.pz[v];
  }
}

// one Jacobi step of spec S8: the neighbour mean as an exact 2^-32 fixed-point sum
__global__ __launch_bounds__(DY_THREADS) void kd_smooth(DV g, const float *src, float *dst, int stamp) {
  const int n = (int)g.ctl[SGD_NAFF];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = g.aff[i];
    long long s0 = 0, s1 = 0, s2 = 0;
    auto acc = [&](int u) {
      float x, y, z;
      if (g.astamp[u] == stamp) {
        const int j = g.aslot[u];
        x = src[3 * (size_t)j];
        y = src[3 * (size_t)j + 1];
        z = src[3 * (size_t)j + 2];
      } else {
        x = g.px[u];
        y = g.py[u];
        z = g.pz[u];
      }
      s0 += __double2ll_rn((double)x * 4294967296.0);
      s1 += __double2ll_rn((double)y * 4294967296.0);
      s2 += __double2ll_rn((double)z * 4294967296.0);
    };
    int bnd, first = -1;
    int cnt = ring_for(g, v, bnd, [&](int k, int j) {
      if (j == 0) first = k;
      acc(g.tv[dn(k)]);
    });
    if (cnt > 0 && bnd) {
      acc(g.tv[dp(first)]);
      cnt++;
    }
    const float sx = src[3 * (size_t)i], sy = src[3 * (size_t)i + 1], sz = src[3 * (size_t)i + 2];
    float rx = sx, ry = sy, rz = sz;
    if (cnt > 0) {
      const float fade = g.fade[i];
      const float ax = (float)(((double)s0 * (1.0 / 4294967296.0)) / (double)cnt);
      const float ay = (float)(((double)s1 * (1.0 / 4294967296.0)) / (double)cnt);
      const float az = (float)(((double)s2 * (1.0 / 4294967296.0)) / (double)cnt);
      rx = sx + (ax - sx) * fade;
      ry = sy + (ay - sy) * fade;
      rz = sz + (az - sz) * fade;
    }
    dst[3 * (size_t)i] = rx;
    dst[3 * (size_t)i + 1] = ry;
    dst[3 * (size_t)i + 2] = rz;
  }
}

__global__ __launch_bounds__(DY_THREADS) void kd_smooth_store(DV g, const float *src) {
  const int n = (int)g.ctl[SGD_NAFF];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = g.aff[i];
    g.px[v] = src[3 * (size_t)i];
    g.py[v] = src[3 * (size_t)i + 1];
    g.pz[v] = src[3 * (size_t)i + 2];
  }
}

// ---------------------------------------------------------------------------------
// normals and bounds of flagged leaves
// ---------------------------------------------------------------------------------
// both lists in one pass: leaves with a bit of mask_n (normals: nlist) and of mask_b
// (bounds: nlist2); the normals step clears only UpdateNormals, so the bounds list is
// the one a pass after it would make
__global__ __launch_bounds__(DY_THREADS) void kd_flagged(DV g, int mask_n, int mask_b) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  int fl = 0, node = 0;
  if (l < g.L) {
    node = g.leaf_node[l];
    fl = g.node_flag[node];
  }
  const bool pn = l < g.L && (fl & mask_n) != 0, pb = l < g.L && (fl & mask_b) != 0;
  const unsigned long long mn = __ballot(pn), mb = __ballot(pb);
  int bn = 0, bb = 0;
  if (lane_id() == 0) {
    if (mn) bn = (int)atomicAdd((unsigned long long *)&g.ctl[SGD_NLIST], (unsigned long long)__popcll(mn));
    if (mb) bb = (int)atomicAdd((unsigned long long *)&g.ctl[SGD_NLIST2], (unsigned long long)__popcll(mb));
  }
  bn = __shfl(bn, 0);
  bb = __shfl(bb, 0);
  if (pn) g.nlist[bn + mask_prefix(mn)] = node;
  if (pb) g.nlist2[bb + mask_prefix(mb)] = node;
}

// face normals (normal_tri_v3) of the listed leaves' faces; their vertices, each once
__global__ __launch_bounds__(LB_THREADS) void kd_face_normals(DV g, int stamp) {
  const int nl = (int)g.ctl[SGD_NLIST];
  for (int li = blockIdx.x; li < nl; li += gridDim.x) {
    const int node = g.nlist[li];
    const int n = g.scount[node];
    int nf = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const int f = g.slab[(size_t)node * SG_DYN_SLAB + i];
      if (f < 0) continue;
      nf++;
      const int a = g.tv[3 * f], b = g.tv[3 * f + 1], c = g.tv[3 * f + 2];
      float x, y, z;
      sg_tri_normal(g.px[a], g.py[a], g.pz[a], g.px[b], g.py[b], g.pz[b], g.px[c], g.py[c], g.pz[c], x, y, z);
      g.fnx[f] = x;
      g.fny[f] = y;
      g.fnz[f] = z;
      const int vs[3] = {a, b, c};
      for (int k = 0; k < 3; k++) {
        if (atomicExch(&g.vstamp[vs[k]], stamp) != stamp) {
          const unsigned long long s = atomicAdd((unsigned long long *)&g.ctl[SGD_NVL], 1ull);
          g.vlist[s] = vs[k];
        }
      }
    }
    nf = wave_sum_i(nf);
    if (lane_id() == 0 && nf) atomicAdd((unsigned long long *)&g.ctl[SGD_NFN], (unsigned long long)nf);
    if (threadIdx.x == 0) atomicAnd(&g.node_flag[node], ~PBVH_UpdateNormals);
  }
}

// mesh_vert_calc_normal (mesh_polygon.c:591-634): corner angle x face normal,
// normalized; exact fixed-point sum; saacos in double
__global__ __launch_bounds__(DY_THREADS) void kd_vert_normals(DV g) {
  const int n = (int)g.ctl[SGD_NVL];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int v = g.vlist[i];
    long long s0 = 0, s1 = 0, s2 = 0;
    const float vx = g.px[v], vy = g.py[v], vz = g.pz[v];
    int bnd;
    const int cnt = ring_for(g, v, bnd, [&](int c, int) {
      const int a = g.tv[dp(c)], b = g.tv[dn(c)];
      float e1x = vx - g.px[a], e1y = vy - g.py[a], e1z = vz - g.pz[a];
      float e2x = g.px[b] - vx, e2y = g.py[b] - vy, e2z = g.pz[b] - vz;
      sg_normalize(e1x, e1y, e1z);
      sg_normalize(e2x, e2y, e2z);
      const float dot = -(e1x * e2x + e1y * e2y + e1z * e2z);
      const float fac = (dot <= -1.0f) ? 3.14159265358979323846f : ((dot >= 1.0f) ? 0.0f : (float)acos((double)dot));
      const int f = c / 3;
      s0 += __double2ll_rn((double)(g.fnx[f] * fac) * 4294967296.0);
      s1 += __double2ll_rn((double)(g.fny[f] * fac) * 4294967296.0);
      s2 += __double2ll_rn((double)(g.fnz[f] * fac) * 4294967296.0);
    });
    if (cnt <= 0) continue;
    float x = (float)((double)s0 * (1.0 / 4294967296.0));
    float y = (float)((double)s1 * (1.0 / 4294967296.0));
    float z = (float)((double)s2 * (1.0 / 4294967296.0));
    sg_normalize(x, y, z);
    g.nx[v] = x;
    g.ny[v] = y;
    g.nz[v] = z;
  }
}

// leaf boxes over the leaf faces' vertices (zeros as +0): a workgroup per listed leaf
__global__ __launch_bounds__(LB_THREADS) void kd_leaf_bounds(DV g, int stamp) {
  __shared__ float s_mn[3][LB_THREADS / 64], s_mx[3][LB_THREADS / 64];
  const int nl = (int)g.ctl[SGD_NLIST2];
  for (int li = blockIdx.x; li < nl; li += gridDim.x) {
    const int node = g.nlist2[li];
    const int fl = g.node_flag[node];
    if (fl & PBVH_UpdateBB) {
      float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
      const int n = g.scount[node];
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int f = g.slab[(size_t)node * SG_DYN_SLAB + i];
        if (f < 0) continue;
        for (int k = 0; k < 3; k++) {
          const int v = g.tv[3 * f + k];
          const float q[3] = {g.px[v] + 0.0f, g.py[v] + 0.0f, g.pz[v] + 0.0f};
          for (int a = 0; a < 3; a++) {
            mn[a] = sg_min(mn[a], q[a]);
            mx[a] = sg_max(mx[a], q[a]);
          }
        }
      }
      for (int a = 0; a < 3; a++) {
        for (int o = 32; o > 0; o >>= 1) {
          mn[a] = sg_min(mn[a], __shfl_xor(mn[a], o));
          mx[a] = sg_max(mx[a], __shfl_xor(mx[a], o));
        }
      }
      const int w = threadIdx.x / 64;
      if (lane_id() == 0)
        for (int a = 0; a < 3; a++) {
          s_mn[a][w] = mn[a];
          s_mx[a][w] = mx[a];
        }
      __syncthreads();
      if (threadIdx.x == 0) {
        float b[8];
        for (int a = 0; a < 3; a++) {
          float lo = s_mn[a][0], hi = s_mx[a][0];
          for (int j = 1; j < LB_THREADS / 64; j++) {
            lo = sg_min(lo, s_mn[a][j]);
            hi = sg_max(hi, s_mx[a][j]);
          }
          b[a] = lo;
          b[4 + a] = hi;
        }
        b[3] = b[7] = 0.0f;
        for (int j = 0; j < 8; j++) g.node_bb[8 * (size_t)node + j] = b[j];
        atomicAdd((unsigned long long *)&g.ctl[SGD_NREFIT], 1ull);
      }
      __syncthreads();
    }
    const int dirty = fl & (PBVH_UpdateBB | PBVH_UpdateOriginalBB);
    if (threadIdx.x == 0) {
      if (fl & PBVH_UpdateOriginalBB)
        for (int j = 0; j < 8; j++) g.node_obb[8 * (size_t)node + j] = g.node_bb[8 * (size_t)node + j];
      g.node_dirty[node] = dirty;
      atomicAnd(&g.node_flag[node], ~(PBVH_UpdateBB | PBVH_UpdateOriginalBB | PBVH_UpdateRedraw));
    }
    // the ancestors the flush visits, each listed once by height: lane i of the first
    // wave takes the i-th ancestor from the table (the stamp lists each once, whichever
    // leaf's lane gets there first); a deeper chain goes on from the table's last entry
    if (dirty && threadIdx.x < SG_WAVE) {
      const int lane = threadIdx.x;
      auto list = [&](int a) {
        if (atomicExch(&g.anc_stamp[a], stamp) == stamp) return false;
        const int h = g.node_height[a];
        const int s = atomicAdd(&g.lvl_cnt[h], 1);
        g.lvl_list[g.lvl_off[h - 1] + s] = a;
        g.anc_pos[a] = s;
        return true;
      };
      for (int j = lane; j < SG_DYN_ANC; j += SG_WAVE) {
        const int a = g.anc[(size_t)node * SG_DYN_ANC + j];
        if (a >= 0) list(a);
      }
      if (lane == 0) {
        const int last = g.anc[(size_t)node * SG_DYN_ANC + SG_DYN_ANC - 1];
        if (last >= 0)
          for (int a = g.node_parent[last]; a >= 0; a = g.node_parent[a])
            if (!list(a)) break;
      }
    }
  }
}

// internal nodes with an updated child: union of the children (child 1, then 2),
// height by height over the listed ancestors of refit leaves; one workgroup.  Nodes
// off those paths have no updated child (pbvh_flush_bb leaves them as they are).
// Up to FL_CAP listed ancestors run from LDS: every input (their boxes, their children's
// slots or boxes and flags) is read in one parallel pass, the heights then take one
// barrier each, and the results go back in one pass.  More ancestors run over HBM.
#define FL_CAP 512
#define FL_MAXH 64

__global__ __launch_bounds__(1024) void kd_flush(DV g, int H, int stamp) {
  __shared__ int s_base[FL_MAXH + 2];  // first slot per height
  __shared__ int s_node[FL_CAP], s_k[2][FL_CAP], s_d[FL_CAP], s_cd[2][FL_CAP];
  __shared__ float s_bb[FL_CAP][8], s_cb[2][FL_CAP][6];
  const int tid = threadIdx.x, nt = blockDim.x;
  const bool lds = H + 1 < FL_MAXH;
  if (lds) {
    for (int h = tid; h <= H; h += nt) s_base[h + 1] = h >= 1 ? g.lvl_cnt[h] : 0;
    __syncthreads();
    if (tid == 0) {
      s_base[0] = 0;
      s_base[1] = 0;
      for (int h = 1; h <= H; h++) s_base[h + 1] += s_base[h];
    }
    __syncthreads();
  }
  const int total = lds ? s_base[H + 1] : 0;
  if (lds && total <= FL_CAP) {
    for (int i = tid; i < total; i += nt) {
      int h = 1;
      while (s_base[h + 1] <= i) h++;
      const int node = g.lvl_list[g.lvl_off[h - 1] + (i - s_base[h])];
      const int c = g.node_child[node];
      s_node[i] = node;
      for (int j = 0; j < 8; j++) s_bb[i][j] = g.node_bb[8 * (size_t)node + j];
      for (int e = 0; e < 2; e++) {
        const int cc = c + e;
        int k = -1;
        if (g.anc_stamp[cc] == stamp) {
          const int hc = g.node_height[cc];
          k = s_base[hc] + g.anc_pos[cc];
        } else {
          for (int a = 0; a < 3; a++) {
            s_cb[e][i][a] = g.node_bb[8 * (size_t)cc + a];
            s_cb[e][i][3 + a] = g.node_bb[8 * (size_t)cc + 4 + a];
          }
          s_cd[e][i] = g.node_dirty[cc];
        }
        s_k[e][i] = k;
      }
    }
    __syncthreads();
    for (int h = 1; h <= H; h++) {
      for (int i = s_base[h] + tid; i < s_base[h + 1]; i += nt) {
        const float *cb[2];
        int u = 0;
        for (int e = 0; e < 2; e++) {
          const int k = s_k[e][i];
          if (k >= 0) {
            u |= s_d[k];
            cb[e] = nullptr;
          } else {
            u |= s_cd[e][i];
            cb[e] = s_cb[e][i];
          }
        }
        if (u & PBVH_UpdateBB) {
          for (int a = 0; a < 3; a++) {
            float lo = FLT_MAX, hi = -FLT_MAX;
            for (int e = 0; e < 2; e++) {
              const int k = s_k[e][i];
              lo = sg_min(lo, k >= 0 ? s_bb[k][a] : cb[e][a]);
              hi = sg_max(hi, k >= 0 ? s_bb[k][4 + a] : cb[e][3 + a]);
            }
            s_bb[i][a] = lo;
            s_bb[i][4 + a] = hi;
          }
          s_bb[i][3] = s_bb[i][7] = 0.0f;
        }
        s_d[i] = u;
      }
      __syncthreads();
    }
    for (int i = tid; i < total; i += nt) {
      const int node = s_node[i], u = s_d[i];
      if (u & PBVH_UpdateBB)
        for (int j = 0; j < 8; j++) g.node_bb[8 * (size_t)node + j] = s_bb[i][j];
      if (u & PBVH_UpdateOriginalBB)
        for (int j = 0; j < 8; j++) g.node_obb[8 * (size_t)node + j] = s_bb[i][j];
      g.node_dirty[node] = 0;
    }
  } else {
    for (int h = 1; h <= H; h++) {
      const int b = g.lvl_off[h - 1], n = g.lvl_cnt[h];
      for (int i = tid; i < n; i += nt) {
        const int node = g.lvl_list[b + i];
        const int c = g.node_child[node];
        const int u = g.node_dirty[c] | g.node_dirty[c + 1];
        if (u & PBVH_UpdateBB) {
          float bb[8];
          for (int a = 0; a < 3; a++) {
            float lo = FLT_MAX, hi = -FLT_MAX;
            lo = sg_min(lo, g.node_bb[8 * (size_t)c + a]);
            hi = sg_max(hi, g.node_bb[8 * (size_t)c + 4 + a]);
            lo = sg_min(lo, g.node_bb[8 * (size_t)(c + 1) + a]);
            hi = sg_max(hi, g.node_bb[8 * (size_t)(c + 1) + 4 + a]);
            bb[a] = lo;
            bb[4 + a] = hi;
          }
          bb[3] = bb[7] = 0.0f;
          for (int j = 0; j < 8; j++) g.node_bb[8 * (size_t)node + j] = bb[j];
        }
        if (u & PBVH_UpdateOriginalBB)
          for (int j = 0; j < 8; j++) g.node_obb[8 * (size_t)node + j] = g.node_bb[8 * (size_t)node + j];
        g.node_dirty[node] = u;
      }
      __syncthreads();
    }
    // clear the marks of the listed ancestors
    for (int h = 1; h <= H; h++) {
      const int b = g.lvl_off[h - 1], n = g.lvl_cnt[h];
      for (int i = tid; i < n; i += nt) g.node_dirty[g.lvl_list[b + i]] = 0;
    }
  }
  // clear the marks of the refit leaves and the level counts
  const int nl = (int)g.ctl[SGD_NLIST2];
  for (int i = tid; i < nl; i += nt) g.node_dirty[g.nlist2[i]] = 0;
  __syncthreads();
  for (int h = tid; h <= H; h += nt) g.lvl_cnt[h] = 0;
}

__global__ void kd_ctl_zero(DV g, unsigned mask) {
  const int i = threadIdx.x;
  if (i < SGD_COUNT && ((mask >> i) & 1u)) g.ctl[i] = 0;
}

// ---------------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------------
template <class T>
int dalloc(T *&p, size_t n) {
  if (p) hipFree(p);
  p = nullptr;
  if (sg_check_hip(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc")) return -1;
  return sg_check_hip(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)), "hipMemset");
}
template <class T>
void dfree(T *&p) {
  if (p) hipFree(p);
  p = nullptr;
}

DV view(PBVH *p) {
  SgDynState *s = p->dyn;
  SgDynDev &d = s->d;
  DV g;
  memset(&g, 0, sizeof(g));
  g.Vcap = d.Vcap;
  g.Tcap = d.Tcap;
  g.N = d.N;
  g.L = d.L;
  g.has_mask = d.has_mask;
  g.respect_hide = p->respect_hide ? 1 : 0;
  g.use_frontface = s->use_frontface;
  g.use_projected = s->use_projected;
  g.ccap = d.ccap;
  g.vnx = s->view_normal[0];
  g.vny = s->view_normal[1];
  g.vnz = s->view_normal[2];
  g.px = d.px;
  g.py = d.py;
  g.pz = d.pz;
  g.nx = d.nx;
  g.ny = d.ny;
  g.nz = d.nz;
  g.mask = d.mask;
  g.vhide = d.vhide;
  g.lab = d.lab;
  g.vnode = d.vnode;
  g.vcorner = d.vcorner;
  g.vstamp = d.vstamp;
  g.astamp = d.astamp;
  g.aslot = d.aslot;
  g.hvst = d.hvst;
  g.tv = d.tv;
  g.topp = d.topp;
  g.tnode = d.tnode;
  g.tslot = d.tslot;
  g.wstamp = d.wstamp;
  g.tie = d.tie;
  g.fnx = d.fnx;
  g.fny = d.fny;
  g.fnz = d.fnz;
  g.fhide = d.fhide;
  g.claim = d.claim;
  g.estamp = d.estamp;
  g.remap = d.remap;
  g.remap_ser = d.remap_ser;
  g.node_flag = d.node_flag;
  g.node_child = d.node_child;
  g.node_dirty = d.node_dirty;
  g.leaf_node = d.leaf_node;
  g.lvl_off = d.lvl_off;
  g.lvl_node = d.lvl_node;
  g.node_parent = d.node_parent;
  g.node_height = d.node_height;
  g.anc_stamp = d.anc_stamp;
  g.anc_pos = d.anc_pos;
  g.anc = d.anc;
  g.lvl_list = d.lvl_list;
  g.lvl_cnt = d.lvl_cnt;
  g.node_bb = d.node_bb;
  g.node_obb = d.node_obb;
  g.slab = d.slab;
  g.scount = d.scount;
  g.wlist = d.wlist;
  g.cand = d.cand;
  g.win = d.win;
  g.outer = d.outer;
  g.hv_in = d.hv0;
  g.hv_out = d.hv1;
  g.touched = d.touched;
  g.gathered = d.gathered;
  g.vlist = d.vlist;
  g.aff = d.aff;
  g.nlist = d.nlist;
  g.nlist2 = d.nlist2;
  g.llist = d.llist;
  g.lstamp = d.lstamp;
  g.seed0 = d.seed0;
  g.seed1 = d.seed1;
  g.sstamp = d.sstamp;
  g.fin = d.fin;
  g.key = d.key;
  g.valid = d.valid;
  g.fade = d.fade;
  g.buf0 = d.buf0;
  g.buf1 = d.buf1;
  g.ctl = d.ctl;
  return g;
}

int read_ctl(PBVH *p) {
  SgDynDev &d = p->dyn->d;
  if (sg_check_hip(hipMemcpyAsync(d.ctl_host, d.ctl, SGD_COUNT * sizeof(long long), hipMemcpyDeviceToHost,
                                  p->stream),
                   "copy"))
    return -1;
  return sg_check_hip(hipStreamSynchronize(p->stream), "sync");
}

// slots [b, e) as a mask for zero_ctl
unsigned slots(int b, int e) {
  unsigned m = 0;
  for (int i = b; i < e; i++) m |= 1u << i;
  return m;
}
void zero_ctl(PBVH *p, const DV &g, unsigned mask) { kd_ctl_zero<<<1, 64, 0, p->stream>>>(g, mask); }
// the control words a topology update starts from zero (the range search's count aside)
unsigned topology_slots() { return slots(SGD_NCAND, SGD_ST + 7) & ~(1u << SGD_NGATHER); }

int launch_check(const char *what) { return sg_check_hip(hipGetLastError(), what); }

}  // namespace

// ---------------------------------------------------------------------------------
// upload / download
// ---------------------------------------------------------------------------------
void sg_dyn_free_device(PBVH *p) {
  if (!p || !p->dyn) return;
  SgDynDev &d = p->dyn->d;
  float **fl[] = {&d.px, &d.py, &d.pz, &d.nx, &d.ny, &d.nz, &d.mask, &d.fnx, &d.fny, &d.fnz, &d.node_bb,
                  &d.node_obb, &d.fade, &d.buf0, &d.buf1, &d.curve};
  for (float **q : fl) dfree(*q);
  int **il[] = {&d.vnode, &d.vcorner, &d.vstamp, &d.astamp, &d.aslot, &d.hvst, &d.tv, &d.topp, &d.tnode, &d.tslot,
                &d.wstamp, &d.estamp, &d.remap, &d.remap_ser, &d.node_flag, &d.node_child, &d.node_dirty,
                &d.leaf_node, &d.lvl_off, &d.lvl_node, &d.node_parent, &d.node_height, &d.anc_stamp, &d.anc_pos, &d.anc, &d.lvl_list,
                &d.lvl_cnt, &d.slab, &d.scount, &d.wlist, &d.cand, &d.win, &d.outer,
                &d.hv0, &d.hv1, &d.touched, &d.gathered, &d.vlist, &d.aff, &d.nlist, &d.nlist2, &d.llist, &d.lstamp,
                &d.seed0, &d.seed1, &d.sstamp};
  for (int **q : il) dfree(*q);
  dfree(d.vhide);
  dfree(d.fhide);
  dfree(d.valid);
  dfree(d.fin);
  dfree(d.lab);
  dfree(d.claim);
  dfree(d.tie);
  dfree(d.key);
  dfree(d.ctl);
  if (d.ctl_host) hipHostFree(d.ctl_host);
  d.ctl_host = nullptr;
  d.Vcap = d.Tcap = d.Ncap = 0;
}

int sg_dyn_upload(PBVH *p) {
  SgDynState *s = p->dyn;
  SgDynMesh &m = *s->bm;
  SgDynTree &t = s->t;
  SgDynDev &d = s->d;
  const int N = (int)t.child.size();
  hipStream_t st = p->stream;
  if (d.Vcap != m.Vcap || d.Tcap != m.Tcap) {
    sg_dyn_free_device(p);
    const size_t V = m.Vcap, T = m.Tcap, C = 3 * (size_t)m.Tcap;
    d.ccap = (int)std::min<long long>(2LL * m.Tcap + 8, INT32_MAX);
    int rc = 0;
    rc |= dalloc(d.px, V) | dalloc(d.py, V) | dalloc(d.pz, V) | dalloc(d.nx, V) | dalloc(d.ny, V) | dalloc(d.nz, V);
    rc |= dalloc(d.mask, m.has_mask ? V : 1) | dalloc(d.vhide, V) | dalloc(d.lab, V);
    rc |= dalloc(d.vnode, V) | dalloc(d.vcorner, V) | dalloc(d.vstamp, V) | dalloc(d.astamp, V) | dalloc(d.aslot, V);
    rc |= dalloc(d.hvst, V) | dalloc(d.hv0, V) | dalloc(d.hv1, V) | dalloc(d.vlist, V) | dalloc(d.aff, V);
    rc |= dalloc(d.fade, V) | dalloc(d.buf0, 3 * V) | dalloc(d.buf1, 3 * V);
    rc |= dalloc(d.tv, C) | dalloc(d.topp, C) | dalloc(d.estamp, C) | dalloc(d.remap, C) | dalloc(d.remap_ser, C);
    rc |= dalloc(d.tnode, T) | dalloc(d.tslot, T) | dalloc(d.wstamp, T) | dalloc(d.fnx, T) | dalloc(d.fny, T);
    rc |= dalloc(d.llist, T) | dalloc(d.lstamp, T) | dalloc(d.fin, T);
    rc |= dalloc(d.seed0, T) | dalloc(d.seed1, T) | dalloc(d.sstamp, T);
    rc |= dalloc(d.fnz, T) | dalloc(d.fhide, T) | dalloc(d.claim, T) | dalloc(d.wlist, T);
    rc |= dalloc(d.tie, T);
    rc |= dalloc(d.cand, (size_t)d.ccap) | dalloc(d.key, (size_t)d.ccap) | dalloc(d.valid, (size_t)d.ccap);
    rc |= dalloc(d.win, (size_t)d.ccap) | dalloc(d.outer, 4 * (size_t)d.ccap) | dalloc(d.touched, 3 * (size_t)d.ccap);
    rc |= dalloc(d.ctl, SGD_COUNT) | dalloc(d.curve, SG_CM_TABLE + 1);
    if (!rc) rc = sg_check_hip(hipHostMalloc((void **)&d.ctl_host, SGD_COUNT * sizeof(long long)), "hipHostMalloc");
    if (rc) return -1;
    d.Vcap = m.Vcap;
    d.Tcap = m.Tcap;
  }
  if (N > d.Ncap) {
    const size_t Nc = (size_t)std::max(N + N / 2, 64);
    int rc = 0;
    rc |= dalloc(d.node_flag, Nc) | dalloc(d.node_child, Nc) | dalloc(d.node_dirty, Nc) | dalloc(d.leaf_node, Nc);
    rc |= dalloc(d.lvl_off, Nc + 1) | dalloc(d.lvl_node, Nc) | dalloc(d.node_bb, 8 * Nc) | dalloc(d.node_obb, 8 * Nc);
    rc |= dalloc(d.slab, Nc * SG_DYN_SLAB) | dalloc(d.scount, Nc) | dalloc(d.gathered, Nc) | dalloc(d.nlist, Nc);
    rc |= dalloc(d.node_parent, Nc) | dalloc(d.node_height, Nc) | dalloc(d.anc_stamp, Nc) | dalloc(d.lvl_list, Nc);
    rc |= dalloc(d.anc_pos, Nc) | dalloc(d.anc, Nc * SG_DYN_ANC);
    rc |= dalloc(d.lvl_cnt, Nc + 1) | dalloc(d.nlist2, Nc);
    if (rc) return -1;
    d.Ncap = (int)Nc;
  }
  d.N = N;
  d.has_mask = m.has_mask ? 1 : 0;
  // mesh
  std::vector<float> x(m.V), y(m.V), z(m.V);
  auto put3 = [&](const std::vector<float> &src, float *dx, float *dy, float *dz, int n) -> int {
    for (int i = 0; i < n; i++) {
      x[i] = src[3 * (size_t)i];
      y[i] = src[3 * (size_t)i + 1];
      z[i] = src[3 * (size_t)i + 2];
    }
    int rc = sg_check_hip(hipMemcpy(dx, x.data(), n * 4, hipMemcpyHostToDevice), "copy");
    rc |= sg_check_hip(hipMemcpy(dy, y.data(), n * 4, hipMemcpyHostToDevice), "copy");
    rc |= sg_check_hip(hipMemcpy(dz, z.data(), n * 4, hipMemcpyHostToDevice), "copy");
    return rc;
  };
  int rc = put3(m.co, d.px, d.py, d.pz, m.V) | put3(m.no, d.nx, d.ny, d.nz, m.V);
  x.resize(std::max(m.V, m.T));
  y.resize(x.size());
  z.resize(x.size());
  rc |= put3(m.fno, d.fnx, d.fny, d.fnz, m.T);
  if (m.has_mask) rc |= sg_check_hip(hipMemcpy(d.mask, m.mask.data(), m.V * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.vhide, m.vhide.data(), m.V, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.lab, m.lab.data(), m.V * 8, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.vnode, m.vnode.data(), m.V * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.vcorner, m.vcorner.data(), m.V * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.tv, m.tv.data(), 3 * (size_t)m.T * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.topp, m.topp.data(), 3 * (size_t)m.T * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.tnode, m.tnode.data(), m.T * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.tslot, m.tslot.data(), m.T * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.fhide, m.fhide.data(), m.T, hipMemcpyHostToDevice), "copy");
  // tree
  std::vector<int> leaves, height(N, 0);
  {
    std::vector<int> st{0};
    while (!st.empty()) {
      const int i = st.back();
      st.pop_back();
      if (t.flag[i] & PBVH_Leaf) {
        leaves.push_back(i);
        continue;
      }
      st.push_back(t.child[i] + 1);
      st.push_back(t.child[i]);
    }
    // heights, children before parents (children always have larger ids)
    for (int i = N - 1; i >= 0; i--)
      if (!(t.flag[i] & PBVH_Leaf)) height[i] = 1 + std::max(height[t.child[i]], height[t.child[i] + 1]);
  }
  int H = 0;
  for (int i = 0; i < N; i++) H = std::max(H, height[i]);
  std::vector<int> lvl_off(H + 1, 0), lvl_node;
  for (int h = 1; h <= H; h++) {
    lvl_off[h - 1] = (int)lvl_node.size();
    for (int i = 0; i < N; i++)
      if (height[i] == h) lvl_node.push_back(i);
  }
  lvl_off[H] = (int)lvl_node.size();
  d.L = (int)leaves.size();
  d.H = H;
  rc |= sg_check_hip(hipMemcpy(d.node_flag, t.flag.data(), N * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.node_child, t.child.data(), N * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.node_bb, t.bb.data(), (size_t)N * 32, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.node_obb, t.obb.data(), (size_t)N * 32, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.scount, t.scount.data(), N * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.slab, t.slab.data(), (size_t)N * SG_DYN_SLAB * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.leaf_node, leaves.data(), leaves.size() * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemcpy(d.lvl_off, lvl_off.data(), lvl_off.size() * 4, hipMemcpyHostToDevice), "copy");
  if (!lvl_node.empty())
    rc |= sg_check_hip(hipMemcpy(d.lvl_node, lvl_node.data(), lvl_node.size() * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemset(d.node_dirty, 0, N * 4), "memset");
  std::vector<int> parent(N, -1);
  for (int i = 0; i < N; i++)
    if (!(t.flag[i] & PBVH_Leaf)) parent[t.child[i]] = parent[t.child[i] + 1] = i;
  rc |= sg_check_hip(hipMemcpy(d.node_parent, parent.data(), N * 4, hipMemcpyHostToDevice), "copy");
  {
    std::vector<int> anc((size_t)N * SG_DYN_ANC, -1);
    for (int i = 0; i < N; i++) {
      int a = parent[i];
      for (int j = 0; j < SG_DYN_ANC && a >= 0; j++, a = parent[a]) anc[(size_t)i * SG_DYN_ANC + j] = a;
    }
    rc |= sg_check_hip(hipMemcpy(d.anc, anc.data(), anc.size() * 4, hipMemcpyHostToDevice), "copy");
  }
  rc |= sg_check_hip(hipMemcpy(d.node_height, height.data(), N * 4, hipMemcpyHostToDevice), "copy");
  rc |= sg_check_hip(hipMemset(d.lvl_cnt, 0, (H + 1) * 4), "memset");
  std::vector<long long> ctl(SGD_COUNT, 0);
  ctl[SGD_V] = m.V;
  ctl[SGD_T] = m.T;
  ctl[SGD_SERIAL] = s->serial;
  rc |= sg_check_hip(hipMemcpy(d.ctl, ctl.data(), SGD_COUNT * 8, hipMemcpyHostToDevice), "copy");
  (void)st;
  s->pend.assign(N, 0);
  return rc ? -1 : 0;
}

int sg_dyn_download(PBVH *p, bool mesh) {
  SgDynState *s = p->dyn;
  SgDynDev &d = s->d;
  SgDynTree &t = s->t;
  if (sg_check_hip(hipStreamSynchronize(p->stream), "sync")) return -1;
  const int N = d.N;
  int rc = 0;
  long long ser = 0;
  rc |= sg_check_hip(hipMemcpy(&ser, d.ctl + SGD_SERIAL, 8, hipMemcpyDeviceToHost), "copy");
  s->serial = (int)ser;
  rc |= sg_check_hip(hipMemcpy(t.flag.data(), d.node_flag, N * 4, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(t.bb.data(), d.node_bb, (size_t)N * 32, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(t.obb.data(), d.node_obb, (size_t)N * 32, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(t.scount.data(), d.scount, N * 4, hipMemcpyDeviceToHost), "copy");
  for (int i = 0; i < N; i++) {
    if ((size_t)i < s->pend.size()) t.flag[i] |= s->pend[i];
  }
  if (!mesh || !s->bm) return rc ? -1 : 0;
  rc |= sg_check_hip(hipMemcpy(t.slab.data(), d.slab, (size_t)N * SG_DYN_SLAB * 4, hipMemcpyDeviceToHost), "copy");
  SgDynMesh &m = *s->bm;
  long long ctl[2];
  rc |= sg_check_hip(hipMemcpy(ctl, d.ctl, 16, hipMemcpyDeviceToHost), "copy");
  m.V = (int)ctl[SGD_V];
  m.T = (int)ctl[SGD_T];
  std::vector<float> x(std::max(m.V, m.T)), y(x.size()), z(x.size());
  auto get3 = [&](std::vector<float> &dst, const float *dx, const float *dy, const float *dz, int n) -> int {
    int r = sg_check_hip(hipMemcpy(x.data(), dx, n * 4, hipMemcpyDeviceToHost), "copy");
    r |= sg_check_hip(hipMemcpy(y.data(), dy, n * 4, hipMemcpyDeviceToHost), "copy");
    r |= sg_check_hip(hipMemcpy(z.data(), dz, n * 4, hipMemcpyDeviceToHost), "copy");
    for (int i = 0; i < n; i++) {
      dst[3 * (size_t)i] = x[i];
      dst[3 * (size_t)i + 1] = y[i];
      dst[3 * (size_t)i + 2] = z[i];
    }
    return r;
  };
  rc |= get3(m.co, d.px, d.py, d.pz, m.V) | get3(m.no, d.nx, d.ny, d.nz, m.V) | get3(m.fno, d.fnx, d.fny, d.fnz, m.T);
  if (m.has_mask) rc |= sg_check_hip(hipMemcpy(m.mask.data(), d.mask, m.V * 4, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(m.vhide.data(), d.vhide, m.V, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(m.lab.data(), d.lab, m.V * 8, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(m.vnode.data(), d.vnode, m.V * 4, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(m.vcorner.data(), d.vcorner, m.V * 4, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(m.tv.data(), d.tv, 3 * (size_t)m.T * 4, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(m.topp.data(), d.topp, 3 * (size_t)m.T * 4, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(m.tnode.data(), d.tnode, m.T * 4, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(m.tslot.data(), d.tslot, m.T * 4, hipMemcpyDeviceToHost), "copy");
  rc |= sg_check_hip(hipMemcpy(m.fhide.data(), d.fhide, m.T, hipMemcpyDeviceToHost), "copy");
  return rc ? -1 : 0;
}

// Host-set node flags (pend) go down: device flags |= pend.
int sg_dyn_sync_flags_down(PBVH *p) {
  SgDynState *s = p->dyn;
  SgDynDev &d = s->d;
  bool any = false;
  for (int v : s->pend) any |= v != 0;
  if (!any) return 0;
  std::vector<int> f(d.N);
  int rc = sg_check_hip(hipMemcpy(f.data(), d.node_flag, d.N * 4, hipMemcpyDeviceToHost), "copy");
  for (int i = 0; i < d.N && i < (int)s->pend.size(); i++) f[i] |= s->pend[i];
  rc |= sg_check_hip(hipMemcpy(d.node_flag, f.data(), d.N * 4, hipMemcpyHostToDevice), "copy");
  std::fill(s->pend.begin(), s->pend.end(), 0);
  return rc ? -1 : 0;
}

// ---------------------------------------------------------------------------------
// topology update (rounds)
// ---------------------------------------------------------------------------------
// The topology update's launches (no wait).  gate: run only if the range search that
// precedes it on the stream gathered a leaf (a dab); the counters stay in HBM.
// clear_ctl: false when the caller already cleared the topology's control words
int topology_launch(PBVH *p, int mode, const float center[3], float radius, int gate, bool clear_ctl = true) {
  SgDynState *s = p->dyn;
  SgDynDev &d = s->d;
  hipStream_t st = p->stream;
  DV g = view(p);
  Range R;
  for (int i = 0; i < 3; i++) R.c[i] = center[i];
  R.r2 = radius * radius;
  if (s->use_projected) {
    const float *n = s->view_normal;
    const float mul = center[0] * n[0] + center[1] * n[1] + center[2] * n[2];
    for (int i = 0; i < 3; i++) R.cp[i] = center[i] + n[i] * -mul;
  } else {
    R.cp[0] = R.cp[1] = R.cp[2] = 0.0f;
  }
  const int wser = ++s->stamp, lser = ++s->stamp;
  const int stamp0 = s->stamp;
  s->stamp += 2 * SG_DYN_MAX_ROUNDS + 2;  // the workgroup's round stamps
  if (clear_ctl) zero_ctl(p, g, topology_slots());
  kd_wbuild<<<std::max(1, std::min(d.L, 4096)), DY_THREADS, 0, st>>>(g, wser, gate);
  kd_wrange<<<DY_BLOCKS, DY_THREADS, 0, st>>>(g, R, lser, gate);
  // round tags stay below 2^21 (keys take bits 0..42): past that, claims restart from zero
  if (stamp0 + 2 * SG_DYN_MAX_ROUNDS + 2 - s->tag_base >= (1 << 21)) {
    if (sg_check_hip(hipMemsetAsync(d.claim, 0, (size_t)d.Tcap * 8, st), "memset") ||
        sg_check_hip(hipMemsetAsync(d.tie, 0, (size_t)d.Tcap * 8, st), "memset"))
      return -1;
    s->tag_base = stamp0;
  }
  kd_topology<<<1, TW_THREADS, 0, st>>>(g, mode, R, s->min_len, s->max_len, gate, stamp0, wser, lser, s->tag_base);
  kd_clear_topology_flags<<<(d.N + 255) / 256, 256, 0, st>>>(g, gate);
  return launch_check("dyntopo topology");
}

// the topology counters of the last update, from the host copy of the control words
void topology_stats(SgDynState *s, const long long *c) {
  s->stats[2] += c[SGD_ST + 0];
  s->stats[3] += c[SGD_ST + 1];
  s->stats[4] += c[SGD_ST + 2];
  s->stats[5] += c[SGD_ST + 3];
  s->stats[6] += c[SGD_ST + 4];
  s->stats[7] += c[SGD_ST + 5];
  s->stats[8] |= c[SGD_FLAGS] & 3;
  s->stats[9] = c[SGD_ST + 6];
  s->serial = (int)c[SGD_SERIAL];
}

int sg_dyn_update_topology(PBVH *p, int mode, const float center[3], float radius) {
  SgDynState *s = p->dyn;
  if (topology_launch(p, mode, center, radius, 0) || read_ctl(p)) return -1;
  const long long *c = s->d.ctl_host;
  topology_stats(s, c);
  s->host_stale = true;
  return (c[SGD_ST + 0] + c[SGD_ST + 1]) > 0 ? 1 : 0;
}

// ---------------------------------------------------------------------------------
// normals and bounds of flagged leaves
// ---------------------------------------------------------------------------------
int normals_bounds_launch(PBVH *p, bool normals, bool bounds) {
  SgDynState *s = p->dyn;
  SgDynDev &d = s->d;
  hipStream_t st = p->stream;
  DV g = view(p);
  const int lb = std::max(1, (d.L + DY_THREADS - 1) / DY_THREADS);
  zero_ctl(p, g, (1u << SGD_NVL) | slots(SGD_NLIST, SGD_NFN + 1) | (1u << SGD_NLIST2) | (1u << SGD_NREFIT));
  kd_flagged<<<lb, DY_THREADS, 0, st>>>(g, normals ? PBVH_UpdateNormals : 0,
                                        bounds ? (PBVH_UpdateBB | PBVH_UpdateOriginalBB | PBVH_UpdateRedraw) : 0);
  if (normals) {
    kd_face_normals<<<DY_BLOCKS, LB_THREADS, 0, st>>>(g, ++s->stamp);
    kd_vert_normals<<<DY_BLOCKS, DY_THREADS, 0, st>>>(g);
  }
  if (bounds) {
    const int bstamp = ++s->stamp;
    kd_leaf_bounds<<<DY_BLOCKS, LB_THREADS, 0, st>>>(g, bstamp);
    kd_flush<<<1, 1024, 0, st>>>(g, d.H, bstamp);
  }
  return launch_check("dyntopo normals/bounds");
}

int sg_dyn_update_normals_bounds(PBVH *p, bool normals, bool bounds) {
  SgDynState *s = p->dyn;
  SgDynDev &d = s->d;
  if (normals_bounds_launch(p, normals, bounds) || read_ctl(p)) return -1;
  if (normals) s->stats[10] = d.ctl_host[SGD_NFN];
  if (bounds) s->stats[11] = d.ctl_host[SGD_NREFIT];
  s->host_stale = true;
  return 0;
}

// ---------------------------------------------------------------------------------
// the dab
// ---------------------------------------------------------------------------------
int sg_dyn_dab_apply(PBVH *p, const SgDabParams *dab) {
  SgDynState *s = p->dyn;
  SgDynDev &d = s->d;
  hipStream_t st = p->stream;
  const int tool = dab->sculpt_tool;
  if (tool != SG_TOOL_DRAW && tool != SG_TOOL_INFLATE && tool != SG_TOOL_SMOOTH && tool != SG_TOOL_GRAB)
    return sg_set_error("PBVH_BMESH: Draw, Inflate, Smooth and Grab only");
  if (dab->flag & (SG_DAB_ORIGINAL | SG_SMOOTH_LEARNED))
    return sg_set_error("PBVH_BMESH: no original data or learned Smooth");
  if (sg_dyn_sync_flags_down(p)) return -1;
  memset(s->stats, 0, sizeof(s->stats));
  const SgDabK k = sg_make_dab(dab, 0);
  const float *curve = nullptr;
  if (k.falloff == SG_CURVE_CUSTOM && dab->curve_table) {
    if (sg_check_hip(hipMemcpyAsync(d.curve, dab->curve_table, (SG_CM_TABLE + 1) * 4, hipMemcpyHostToDevice, st),
                     "copy"))
      return -1;
    curve = d.curve;
  }
  DV g = view(p);
  const int lb = std::max(1, (d.L + DY_THREADS - 1) / DY_THREADS);
  const int update = PBVH_UpdateNormals | PBVH_UpdateBB | PBVH_UpdateOriginalBB | PBVH_UpdateDrawBuffers |
                     PBVH_UpdateRedraw;
  // one wait per dab: every step is queued, the counters read once at the end
  const bool topo = tool != SG_TOOL_GRAB && s->mode;
  if (topo) {
    const float rt = k.r * 1.25f;
    // the range search's count and the topology's words in one clear (the search
    // touches only its count)
    zero_ctl(p, g, (1u << SGD_NGATHER) | topology_slots());
    kd_search<<<lb, DY_THREADS, 0, st>>>(g, k.cx, k.cy, k.cz, rt * rt, update | PBVH_UpdateTopology);
    const float c[3] = {k.cx, k.cy, k.cz};
    if (launch_check("dyntopo search") || topology_launch(p, s->mode, c, k.r, 1, false)) return -1;
  }
  zero_ctl(p, g, slots(SGD_NGATHER, SGD_NREFIT + 1));
  kd_search<<<lb, DY_THREADS, 0, st>>>(g, k.cx, k.cy, k.cz, k.r2, update);
  const int stamp = ++s->stamp;
  kd_unique<<<DY_BLOCKS, LB_THREADS, 0, st>>>(g, stamp);
  if (tool == SG_TOOL_DRAW) kd_area<<<DY_BLOCKS, DY_THREADS, 0, st>>>(g, k);
  kd_brush<<<DY_BLOCKS, DY_THREADS, 0, st>>>(g, k, curve, stamp);
  if (tool == SG_TOOL_SMOOTH) {
    kd_smooth_init<<<DY_BLOCKS, DY_THREADS, 0, st>>>(g);
    float *a = d.buf0, *b = d.buf1;
    for (int it = 0; it < k.iters; it++) {
      kd_smooth<<<DY_BLOCKS, DY_THREADS, 0, st>>>(g, a, b, stamp);
      std::swap(a, b);
    }
    kd_smooth_store<<<DY_BLOCKS, DY_THREADS, 0, st>>>(g, a);
  }
  if (launch_check("dyntopo brush") || normals_bounds_launch(p, true, true) || read_ctl(p)) return -1;
  // the gathered and affected counts survive the normals and bounds (other words)
  const long long *c = d.ctl_host;
  s->stats[0] = c[SGD_NGATHER];
  s->stats[1] = c[SGD_NAFF];
  s->stats[10] = c[SGD_NFN];
  s->stats[11] = c[SGD_NREFIT];
  if (topo) topology_stats(s, c);
  s->host_stale = true;
  p->dab_count++;
  return 0;
}

int sg_dyn_last_affected(PBVH *p, uint64_t *out, int cap) {
  SgDynState *s = p->dyn;
  SgDynDev &d = s->d;
  const int n = (int)s->stats[1];
  long long nv = 0;
  if (sg_check_hip(hipMemcpy(&nv, d.ctl + SGD_V, 8, hipMemcpyDeviceToHost), "copy")) return -1;
  std::vector<int> aff(n);
  std::vector<uint64_t> lab((size_t)nv);
  if (n && sg_check_hip(hipMemcpy(aff.data(), d.aff, n * 4, hipMemcpyDeviceToHost), "copy")) return -1;
  if (nv && sg_check_hip(hipMemcpy(lab.data(), d.lab, (size_t)nv * 8, hipMemcpyDeviceToHost), "copy")) return -1;
  std::vector<uint64_t> r(n);
  for (int i = 0; i < n; i++) r[i] = lab[aff[i]];
  std::sort(r.begin(), r.end());
  for (int i = 0; i < n && i < cap; i++) out[i] = r[i];
  return n;
}
