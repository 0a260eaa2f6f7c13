// Deterministic synthetic meshes for tests and benchmarks: icosphere level k
// (V = 10*4^k + 2, T = 20*4^k), torus, and seeded 3D value-noise displacement.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/sculpt_gpu.h"

namespace {

// open-addressing edge -> midpoint map
struct EdgeMap {
  std::vector<uint64_t> key;
  std::vector<int> val;
  uint64_t mask;
  explicit EdgeMap(size_t n) {
    size_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    key.assign(cap, ~0ull);
    val.assign(cap, -1);
    mask = cap - 1;
  }
  // returns existing value or inserts `fresh`
  int get_or_insert(uint64_t k, int fresh, bool *inserted) {
    uint64_t h = k * 0x9E3779B97F4A7C15ull;
    size_t i = (size_t)(h >> 20) & mask;
    for (;;) {
      if (key[i] == k) {
        *inserted = false;
        return val[i];
      }
      if (key[i] == ~0ull) {
        key[i] = k;
        val[i] = fresh;
        *inserted = true;
        return fresh;
      }
      i = (i + 1) & mask;
    }
  }
};

uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

double lattice(int ix, int iy, int iz, uint32_t seed) {
  uint32_t h = mix32((uint32_t)ix * 0x8da6b343u ^ mix32((uint32_t)iy * 0xd8163841u ^
                     mix32((uint32_t)iz * 0xcb1ab31fu ^ seed)));
  return (double)h / 4294967295.0 * 2.0 - 1.0;
}

double value_noise(double x, double y, double z, uint32_t seed) {
  const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  double tx = x - fx, ty = y - fy, tz = z - fz;
  tx = tx * tx * (3.0 - 2.0 * tx);
  ty = ty * ty * (3.0 - 2.0 * ty);
  tz = tz * tz * (3.0 - 2.0 * tz);
  double c[2][2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < 2; e++) c[a][b][e] = lattice(ix + a, iy + b, iz + e, seed);
  const double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
  const double x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
  const double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
  const double x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
  const double y0 = x00 + (x10 - x00) * ty;
  const double y1 = x01 + (x11 - x01) * ty;
  return y0 + (y1 - y0) * tz;
}

}  // namespace

extern "C" {

int sg_mesh_icosphere_size(int level, int *r_totvert, int *r_tottri) {
  if (level < 0 || level > 11) return -1;
  long long p = 1;
  for (int i = 0; i < level; i++) p *= 4;
  *r_totvert = (int)(10 * p + 2);
  *r_tottri = (int)(20 * p);
  return 0;
}

// Icosahedron (12 vertices, 20 faces, outward winding), each level splits every
// face into 4 with midpoints pushed to the sphere.  Vertex order: the previous
// level's vertices, then new midpoints in order of first use while faces are
// visited in order; face (a,b,c) -> (a,ab,ca) (ab,b,bc) (ca,bc,c) (ab,bc,ca).
int sg_mesh_icosphere(int level, float radius, float *co, int *tris) {
  int V, T;
  if (sg_mesh_icosphere_size(level, &V, &T)) return -1;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<double> P = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                           0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
  for (int i = 0; i < 12; i++) {
    double *q = &P[3 * i];
    const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    q[0] /= n;
    q[1] /= n;
    q[2] /= n;
  }
  std::vector<int> F = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                        11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                        3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  P.reserve(3 * (size_t)V);
  for (int l = 0; l < level; l++) {
    const size_t nf = F.size() / 3;
    EdgeMap em(nf * 3 / 2 + 16);
    std::vector<int> NF(nf * 12);
    auto mid = [&](int a, int b) {
      const uint64_t k = a < b ? ((uint64_t)a << 32) | (uint32_t)b : ((uint64_t)b << 32) | (uint32_t)a;
      bool ins;
      const int fresh = (int)(P.size() / 3);
      const int m = em.get_or_insert(k, fresh, &ins);
      if (ins) {
        double q[3] = {(P[3 * a] + P[3 * b]) * 0.5, (P[3 * a + 1] + P[3 * b + 1]) * 0.5,
                       (P[3 * a + 2] + P[3 * b + 2]) * 0.5};
        const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        P.push_back(q[0] / n);
        P.push_back(q[1] / n);
        P.push_back(q[2] / n);
      }
      return m;
    };
    for (size_t f = 0; f < nf; f++) {
      const int a = F[3 * f], b = F[3 * f + 1], c = F[3 * f + 2];
      const int ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
      const int q[12] = {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca};
      memcpy(&NF[12 * f], q, sizeof(q));
    }
    F.swap(NF);
  }
  if ((int)(P.size() / 3) != V || (int)(F.size() / 3) != T) return -2;
  for (size_t i = 0; i < (size_t)V * 3; i++) co[i] = (float)(P[i] * (double)radius);
  memcpy(tris, F.data(), sizeof(int) * F.size());
  return 0;
}

// Torus around +z: vertex (i, j) = i * nv + j at u = 2*pi*i/nu, v = 2*pi*j/nv;
// quad (i,j),(i+1,j),(i+1,j+1),(i,j+1) split along its first diagonal, outward.
int sg_mesh_torus(int nu, int nv, float major_r, float minor_r, float *co, int *tris) {
  if (nu < 3 || nv < 3) return -1;
  const double two_pi = 6.283185307179586476925286766559;
  for (int i = 0; i < nu; i++) {
    const double u = two_pi * i / nu;
    for (int j = 0; j < nv; j++) {
      const double v = two_pi * j / nv;
      const double rr = (double)major_r + (double)minor_r * std::cos(v);
      float *p = co + 3 * ((size_t)i * nv + j);
      p[0] = (float)(rr * std::cos(u));
      p[1] = (float)(rr * std::sin(u));
      p[2] = (float)((double)minor_r * std::sin(v));
    }
  }
  size_t k = 0;
  for (int i = 0; i < nu; i++) {
    const int i1 = (i + 1) % nu;
    for (int j = 0; j < nv; j++) {
      const int j1 = (j + 1) % nv;
      const int a = i * nv + j, b = i1 * nv + j, c = i1 * nv + j1, d = i * nv + j1;
      tris[k++] = a;
      tris[k++] = b;
      tris[k++] = c;
      tris[k++] = a;
      tris[k++] = c;
      tris[k++] = d;
    }
  }
  return 0;
}

int sg_mesh_noise_displace(float *co, const float *dir, int totvert, uint32_t seed, float amplitude,
                           float frequency, int octaves) {
  if (totvert < 0 || octaves < 1 || octaves > 16) return -1;
  double norm = 0.0, g = 1.0;
  for (int o = 0; o < octaves; o++, g *= 0.5) norm += g;
#pragma omp parallel for schedule(static)
  for (int v = 0; v < totvert; v++) {
    float *p = co + 3 * (size_t)v;
    double f = frequency, amp = 1.0, n = 0.0;
    for (int o = 0; o < octaves; o++) {
      n += amp * value_noise(p[0] * f, p[1] * f, p[2] * f, seed + 0x9e37u * (uint32_t)o);
      f *= 2.0;
      amp *= 0.5;
    }
    const double s = (double)amplitude * (n / norm);
    const float *dv = dir + 3 * (size_t)v;
    p[0] = (float)(p[0] + dv[0] * s);
    p[1] = (float)(p[1] + dv[1] * s);
    p[2] = (float)(p[2] + dv[2] * s);
  }
  return 0;
}

}  // extern "C"
