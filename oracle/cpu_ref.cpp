// oracle/cpu_ref.cpp -- CPU restatement of the reference render / d_render path.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
// and bench.py's CPU-baseline leg load this library; the shipped path never does.
//
// Built with -ffp-contract=off and no fast-math: every expression below keeps the
// reference's evaluation order (LiteMath, common_src/LiteMath.h), so results are
// bit-identical to the reference's BF path (checked against fixtures written by
// oracle/_ref, the reference compiled from its own sources).
#include "cpu_ref.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

inline f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 mul(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 mul(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline f3 mulv(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline f3 divs(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// LiteMath dot: a.x*b.x + a.y*b.y + a.z*b.z, left to right (LiteMath.h:792)
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(f3 a) { return std::sqrt(dot(a, a)); }
// normalize = a * (1/length) (LiteMath.h:795)
inline f3 normalize(f3 a) { float li = 1.0f / length(a); return mul(a, li); }
// cross via the yzx-shuffle form (LiteMath.h:816-821); same component formulas
inline f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
inline float length2(f2 a) { return std::sqrt(dot2(a, a)); }
// LiteMath sign (LiteMath.h:109-114)
inline float sgn(float x) { return x == 0.0f ? 0.0f : (x < 0.0f ? -1.0f : 1.0f); }
// std::max / std::min semantics
inline float fmax_std(float a, float b) { return (a < b) ? b : a; }
inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
inline int imax(int a, int b) { return (a < b) ? b : a; }
inline int imin(int a, int b) { return (b < a) ? b : a; }
// LiteMath clamp(float) = min(max(a, u), b) (LiteMath.h:84)
inline float clampf(float u, float a, float b) { return fmin_std(fmax_std(a, u), b); }
inline int clampi(int u, int a, int b) { return imin(imax(a, u), b); }

// float4x4, column-major memory: m[c*4 + r]
struct M44 { float m[16]; float operator()(int r, int c) const { return m[c * 4 + r]; } };
inline M44 identity() { M44 r; std::memset(r.m, 0, sizeof r.m); r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f; return r; }
// mat4_colmajor_mul_vec4 (LiteMath.h:1342-1348)
inline f4 mv(const M44& B, f4 V) {
  f4 r;
  r.x = V.x * B.m[0] + V.y * B.m[4] + V.z * B.m[8] + V.w * B.m[12];
  r.y = V.x * B.m[1] + V.y * B.m[5] + V.z * B.m[9] + V.w * B.m[13];
  r.z = V.x * B.m[2] + V.y * B.m[6] + V.z * B.m[10] + V.w * B.m[14];
  r.w = V.x * B.m[3] + V.y * B.m[7] + V.z * B.m[11] + V.w * B.m[15];
  return r;
}
// float4x4 * float3 uses w = 1 (LiteMath.h:1451-1457)
inline f3 mp(const M44& B, f3 v) { f4 r = mv(B, {v.x, v.y, v.z, 1.0f}); return {r.x, r.y, r.z}; }
inline f3 mdir(const M44& B, f3 v) { f4 r = mv(B, {v.x, v.y, v.z, 0.0f}); return {r.x, r.y, r.z}; }
// m1*m2: column j = m1 * m2.col(j) (LiteMath.h:1536-1549)
inline M44 mm(const M44& a, const M44& b) {
  M44 r;
  for (int c = 0; c < 4; c++) {
    f4 col = mv(a, {b.m[c * 4 + 0], b.m[c * 4 + 1], b.m[c * 4 + 2], b.m[c * 4 + 3]});
    r.m[c * 4 + 0] = col.x; r.m[c * 4 + 1] = col.y; r.m[c * 4 + 2] = col.z; r.m[c * 4 + 3] = col.w;
  }
  return r;
}
inline M44 transpose(const M44& a) { M44 r; for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) r.m[j * 4 + i] = a.m[i * 4 + j]; return r; }

// Cofactor-pair 4x4 inverse with the reference's exact term order
// (inverse4x4, LiteMath.h:1551-1634).
M44 inverse4x4(const M44& a) {
  float t[12];
  float o[16];
  auto m1 = [&](int r, int c) { return a(r, c); };
  auto M = [&](int r, int c) -> float& { return o[c * 4 + r]; };
  t[0] = m1(2,2) * m1(3,3); t[1] = m1(2,3) * m1(3,2); t[2] = m1(2,1) * m1(3,3);
  t[3] = m1(2,3) * m1(3,1); t[4] = m1(2,1) * m1(3,2); t[5] = m1(2,2) * m1(3,1);
  t[6] = m1(2,0) * m1(3,3); t[7] = m1(2,3) * m1(3,0); t[8] = m1(2,0) * m1(3,2);
  t[9] = m1(2,2) * m1(3,0); t[10] = m1(2,0) * m1(3,1); t[11] = m1(2,1) * m1(3,0);
  M(0,0) = t[0] * m1(1,1) + t[3] * m1(1,2) + t[4] * m1(1,3);
  M(0,0) -= t[1] * m1(1,1) + t[2] * m1(1,2) + t[5] * m1(1,3);
  M(1,0) = t[1] * m1(1,0) + t[6] * m1(1,2) + t[9] * m1(1,3);
  M(1,0) -= t[0] * m1(1,0) + t[7] * m1(1,2) + t[8] * m1(1,3);
  M(2,0) = t[2] * m1(1,0) + t[7] * m1(1,1) + t[10] * m1(1,3);
  M(2,0) -= t[3] * m1(1,0) + t[6] * m1(1,1) + t[11] * m1(1,3);
  M(3,0) = t[5] * m1(1,0) + t[8] * m1(1,1) + t[11] * m1(1,2);
  M(3,0) -= t[4] * m1(1,0) + t[9] * m1(1,1) + t[10] * m1(1,2);
  M(0,1) = t[1] * m1(0,1) + t[2] * m1(0,2) + t[5] * m1(0,3);
  M(0,1) -= t[0] * m1(0,1) + t[3] * m1(0,2) + t[4] * m1(0,3);
  M(1,1) = t[0] * m1(0,0) + t[7] * m1(0,2) + t[8] * m1(0,3);
  M(1,1) -= t[1] * m1(0,0) + t[6] * m1(0,2) + t[9] * m1(0,3);
  M(2,1) = t[3] * m1(0,0) + t[6] * m1(0,1) + t[11] * m1(0,3);
  M(2,1) -= t[2] * m1(0,0) + t[7] * m1(0,1) + t[10] * m1(0,3);
  M(3,1) = t[4] * m1(0,0) + t[9] * m1(0,1) + t[10] * m1(0,2);
  M(3,1) -= t[5] * m1(0,0) + t[8] * m1(0,1) + t[11] * m1(0,2);
  t[0] = m1(0,2) * m1(1,3); t[1] = m1(0,3) * m1(1,2); t[2] = m1(0,1) * m1(1,3);
  t[3] = m1(0,3) * m1(1,1); t[4] = m1(0,1) * m1(1,2); t[5] = m1(0,2) * m1(1,1);
  t[6] = m1(0,0) * m1(1,3); t[7] = m1(0,3) * m1(1,0); t[8] = m1(0,0) * m1(1,2);
  t[9] = m1(0,2) * m1(1,0); t[10] = m1(0,0) * m1(1,1); t[11] = m1(0,1) * m1(1,0);
  M(0,2) = t[0] * m1(3,1) + t[3] * m1(3,2) + t[4] * m1(3,3);
  M(0,2) -= t[1] * m1(3,1) + t[2] * m1(3,2) + t[5] * m1(3,3);
  M(1,2) = t[1] * m1(3,0) + t[6] * m1(3,2) + t[9] * m1(3,3);
  M(1,2) -= t[0] * m1(3,0) + t[7] * m1(3,2) + t[8] * m1(3,3);
  M(2,2) = t[2] * m1(3,0) + t[7] * m1(3,1) + t[10] * m1(3,3);
  M(2,2) -= t[3] * m1(3,0) + t[6] * m1(3,1) + t[11] * m1(3,3);
  M(3,2) = t[5] * m1(3,0) + t[8] * m1(3,1) + t[11] * m1(3,2);
  M(3,2) -= t[4] * m1(3,0) + t[9] * m1(3,1) + t[10] * m1(3,2);
  M(0,3) = t[2] * m1(2,2) + t[5] * m1(2,3) + t[1] * m1(2,1);
  M(0,3) -= t[4] * m1(2,3) + t[0] * m1(2,1) + t[3] * m1(2,2);
  M(1,3) = t[8] * m1(2,3) + t[0] * m1(2,0) + t[7] * m1(2,2);
  M(1,3) -= t[6] * m1(2,2) + t[9] * m1(2,3) + t[1] * m1(2,0);
  M(2,3) = t[6] * m1(2,1) + t[11] * m1(2,3) + t[3] * m1(2,0);
  M(2,3) -= t[10] * m1(2,3) + t[2] * m1(2,0) + t[7] * m1(2,1);
  M(3,3) = t[10] * m1(2,2) + t[4] * m1(2,0) + t[9] * m1(2,1);
  M(3,3) -= t[8] * m1(2,1) + t[11] * m1(2,2) + t[5] * m1(2,0);
  const float k = 1.0f / (m1(0,0) * M(0,0) + m1(0,1) * M(1,0) + m1(0,2) * M(2,0) + m1(0,3) * M(3,0));
  M44 r;
  for (int i = 0; i < 16; i++) r.m[i] = o[i] * k;
  return r;
}

M44 load44(const float* p) { M44 r; std::memcpy(r.m, p, sizeof r.m); return r; }

// ----------------------------------------------------------------------------
// Sample sequences
// ----------------------------------------------------------------------------
// prng::NextState / rndFloat / RandomGenInit (common_src/qmc.h:81-87, qmc.cpp:250-268).
struct Prng { uint32_t x, y; };
inline uint32_t prng_next(Prng& g) {
  const uint32_t x = g.x * 17u + g.y * 13123u;
  g.x = (x << 13) ^ x;
  g.y ^= (x << 7);
  return x;
}
inline float prng_float(Prng& g) {
  const uint32_t x = prng_next(g);
  const uint32_t tmp = x * (x * x * 15731u + 74323u) + 871483u;
  return (float)tmp * (1.0f / 4294967296.0f);
}
Prng prng_init(int seed) {
  const uint32_t s = (uint32_t)seed;  // the reference's int arithmetic wraps mod 2^32
  Prng g;
  g.x = s * (s * s * 15731u + 74323u) + 871483u;
  g.y = s * (s * s * 13734u + 37828u) + 234234u;
  for (int i = 0; i < seed % 7; i++) prng_next(g);
  return g;
}

// The reference draws edge samples in `#pragma omp parallel for num_threads(14)`
// with the default static schedule (src/drender.h:324-331): thread t owns the
// contiguous range [t*q + min(t,r), ...) with q = N/14, r = N%14, and draws
// (rnd0, rnd1) from stream RandomGenInit(7777 + t*t + 1) in order.  Both libomp
// and libgomp partition static loops this way.
void prng_table(int n, float* out) {
  const int T = 14;
  const int q = n / T, r = n % T;
  for (int t = 0; t < T; t++) {
    const int start = t * q + (t < r ? t : r);
    const int cnt = q + (t < r ? 1 : 0);
    Prng g = prng_init(7777 + t * t + 1);
    for (int k = 0; k < cnt; k++) {
      const float a = prng_float(g);
      const float b = prng_float(g);
      out[2 * (start + k) + 0] = a;
      out[2 * (start + k) + 1] = b;
    }
  }
}

// Niederreiter base-2 direction numbers after Bratley, Fox & Niederreiter,
// "Implementation and tests of low-discrepancy sequences", ACM TOMACS 2(3), 1992
// (the construction behind qmc::init, common_src/qmc.cpp:14-187).  Dimension d
// uses the d-th irreducible polynomial over GF(2) in increasing numeric order;
// 63 output bits per generator column, of which bits 32..62 form the table.
void qmc_table(uint32_t* table) {
  const int DIMS = 11, RES = 31, NB = 63;
  // irreducible polynomials, found by trial division in increasing order
  std::vector<int> polys;
  polys.push_back(2);
  for (int p = 3; (int)polys.size() < DIMS; p++) {
    int deg = 31 - __builtin_clz((unsigned)p);
    bool irreducible = true;
    for (int q : polys) {
      int dq = 31 - __builtin_clz((unsigned)q);
      if (dq > deg) break;
      int rem = p;
      for (int dr = deg; dr >= dq; dr--)
        if (rem & (1 << dr)) rem ^= q << (dr - dq);
      if (rem == 0) { irreducible = false; break; }
    }
    if (irreducible) polys.push_back(p);
  }
  std::vector<uint64_t> C((size_t)NB * DIMS, 0);  // C[bit*DIMS + d]
  for (int d = 0; d < DIMS; d++) {
    const int P = polys[d];
    const int e = 31 - __builtin_clz((unsigned)P);
    // coefficient arrays are stored highest degree first
    std::vector<int> pc(e + 1);
    for (int i = 0; i <= e; i++) pc[i] = (P >> (e - i)) & 1;
    std::vector<int> b(1, 1);  // b(x) = 1
    std::vector<int> v(NB + e, 0);
    for (int r = 0; r < NB; r++) {
      const int j = NB - 1 - r;  // output bit position
      const int u = r % e;
      if (u == 0) {
        // b <- b * p  (GF(2) convolution)
        const int m1 = (int)b.size() - 1;
        const int m = m1 + e;
        std::vector<int> nb(m + 1, 0);
        for (int i = 0; i <= m1; i++)
          if (b[i])
            for (int k = 0; k <= e; k++) nb[i + k] ^= pc[k];
        b.swap(nb);
        // v: m1 zeros, then ones up to degree m, then the linear recurrence with
        // characteristic polynomial b
        for (int i = 0; i < m1; i++) v[i] = 0;
        for (int i = m1; i < m; i++) v[i] = 1;
        for (int i = m; i < NB + e - 1; i++) {
          int s = 0;
          for (int k = 1; k <= m; k++) s ^= v[i - k] & b[k];
          v[i] = s;
        }
      }
      for (int i = 0; i < NB; i++) C[(size_t)i * DIMS + d] |= (uint64_t)v[i + u] << j;
    }
  }
  for (int d = 0; d < DIMS; d++)
    for (int bit = 0; bit < RES; bit++)
      table[d * RES + bit] = (uint32_t)((C[(size_t)bit * DIMS + d] >> 32) & 0x7FFFFFFFu);
}

// qmc::rndFloat (common_src/qmc.cpp:190-197)
inline float qmc_rnd(uint32_t pos, int dim, const uint32_t* table) {
  uint32_t res = 0, data = pos;
  for (int bit = 0; bit < 31; bit++, data >>= 1)
    if (data & 1) res ^= table[bit + dim * 31];
  const float INT_SCALE = 1.0f / (float)0x80000001U;
  return (float)(res + 1) * INT_SCALE;
}

// qmc::planeHammersley (common_src/qmc.cpp:227-243)
void hammersley(int n, float* out) {
  for (int k = 0; k < n; k++) {
    float u = 0;
    int kk = k;
    for (float p = 0.5f; kk; p *= 0.5f, kk >>= 1)
      if (kk & 1) u += p;
    out[2 * k + 0] = u;
    out[2 * k + 1] = (k + 0.5f) / n;
  }
}

// ----------------------------------------------------------------------------
// Scene (prepare_for_render, common_src/scene_utils.cpp:98-173)
// ----------------------------------------------------------------------------
struct Hit { uint32_t prim, inst, geom; float t, u, v; uint32_t gtri; };

struct BvhNode { float lo[3], hi[3]; int left, right, first, count; };

struct Scene {
  int nm = 0, V = 0, F = 0;
  std::vector<int> vo, to;                  // vertex / triangle offsets per mesh (size nm+1)
  std::vector<M44> tr, tr_inv;
  std::vector<f3> pos, pos_orig, col, nrm, nrm2d;
  std::vector<f2> tc;
  std::vector<uint32_t> idx_local;          // 3F mesh-local indices
  std::vector<uint32_t> idx;                // 3F global indices (get_index)
  std::vector<int> tri_mesh;                // mesh of each global triangle
  std::vector<f3> ta, tb, tcv;              // BF's per-triangle positions (transform applied again)
  std::vector<int> tex_w, tex_h, tex_off;
  std::vector<float> texels;
  std::vector<int> diff;
  int tracer = ORC_TRACER_BF;
  // BVH (CPU stand-in for Embree)
  std::vector<BvhNode> nodes;
  std::vector<int> order;
  bool bvh_built = false;
  // triangles BF may accept with a clamped determinant (|det| < 1e-6), tested by brute force
  // after BVH traversal: for camera rays of the current view (shared origin) and for shadow
  // rays (shared direction).  Set before the parallel loops.
  std::vector<int> sus_cam, sus_light;
  // jittered pixel samples (orc_set_jitter)
  int jit_on = 0;
  uint32_t jit_seed = 0;
};

// Pixel-sample positions.  The reference: sample (dx, dy) at the centre of its cell of the
// sq x sq grid (src/drender.h:124-128).  Jittered (the build's own option, no counterpart in the
// reference): a uniform point of that cell, offsets ux, uy = (k >> 8) * 2^-24 from a chain of
// Wellons' "lowbias32" integer hash (hash-prospector, 2020) over (seed, view, x, y, sample).
static uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}
static void pixel_sample(const Scene& s, int view, int x, int y, int si, int sq, float& xo, float& yo) {
  float ux = 0.5f, uy = 0.5f;
  if (s.jit_on) {
    uint32_t k = lowbias32(s.jit_seed ^ lowbias32((uint32_t)view + 0x9e3779b9U));
    k = lowbias32(k ^ (uint32_t)x);
    k = lowbias32(k ^ (uint32_t)y);
    k = lowbias32(k ^ (uint32_t)si);
    ux = (float)(k >> 8) * (1.0f / 16777216.0f);
    uy = (float)(lowbias32(k) >> 8) * (1.0f / 16777216.0f);
  }
  xo = ((si % sq) + ux) / float(sq);
  yo = ((si / sq) + uy) / float(sq);
}

// Clamped-determinant suspects.  BF's invDet = sign/max(|det|, 1e-6) lets a nearly parallel
// ray accept a triangle it never touches.  With d = a*e1 + b*e2 + c*(o-A), acceptance needs
// |a|, |b|, |c| < 1e-6/|V|, V = (o-A).N, so |V| < 1e-6 (|e1| + |e2| + |o-A|) for a shared
// origin o; for a shared direction d it needs |d.N| < 1e-6.  Bounds checked in double with a
// 2x margin.
void suspects_origin(Scene& s, f3 o) {
  s.sus_cam.clear();
  for (int g = 0; g < s.F; g++) {
    const double ax = s.ta[g].x, ay = s.ta[g].y, az = s.ta[g].z;
    const double e1x = (double)s.tb[g].x - ax, e1y = (double)s.tb[g].y - ay, e1z = (double)s.tb[g].z - az;
    const double e2x = (double)s.tcv[g].x - ax, e2y = (double)s.tcv[g].y - ay, e2z = (double)s.tcv[g].z - az;
    const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const double tx = o.x - ax, ty = o.y - ay, tz = o.z - az;
    const double V = std::fabs(tx * nx + ty * ny + tz * nz);
    const double b = std::sqrt(e1x * e1x + e1y * e1y + e1z * e1z) + std::sqrt(e2x * e2x + e2y * e2y + e2z * e2z) +
                     std::sqrt(tx * tx + ty * ty + tz * tz);
    if (V <= 2e-6 * b + 1e-12 || V != V) s.sus_cam.push_back(g);
  }
}
void suspects_direction(Scene& s, f3 d) {
  s.sus_light.clear();
  for (int g = 0; g < s.F; g++) {
    const double ax = s.ta[g].x, ay = s.ta[g].y, az = s.ta[g].z;
    const double e1x = (double)s.tb[g].x - ax, e1y = (double)s.tb[g].y - ay, e1z = (double)s.tb[g].z - az;
    const double e2x = (double)s.tcv[g].x - ax, e2y = (double)s.tcv[g].y - ay, e2z = (double)s.tcv[g].z - az;
    const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const double dn = std::fabs(d.x * nx + d.y * ny + d.z * nz);
    if (dn <= 2e-6 || dn != dn) s.sus_light.push_back(g);
  }
}

int grad_layout(const Scene& s, int mode, std::vector<int>* mesh_off) {
  int off = 0;
  if (mesh_off) mesh_off->assign(s.nm, -1);
  for (int i = 0; i < s.nm; i++) {
    if (!s.diff[i]) continue;
    if (mesh_off) (*mesh_off)[i] = off;
    const int nv = s.vo[i + 1] - s.vo[i];
    off += 3 * nv;                                   // positions
    if (mode == ORC_VERTEX_COLOR || mode == ORC_LAMBERT_VERTEX_COLOR) off += 3 * nv;    // colours
    if (mode == ORC_TEXTURE_COLOR || mode == ORC_LAMBERT) off += 3 * s.tex_w[i] * s.tex_h[i];
    off += 12;                                       // one instance transform (DMesh::TRANSFORM_SIZE)
  }
  return off;
}

// ----------------------------------------------------------------------------
// Tracing
// ----------------------------------------------------------------------------
// Moller-Trumbore exactly as BruteForce3D::GetNearestHit (raytrace3d_bf.cpp:123-150)
inline bool mt_test(const Scene& s, int g, f3 o, f3 d, float tNear, float& t, float& u, float& v) {
  const f3 A = s.ta[g], B = s.tb[g], C = s.tcv[g];
  const f3 e1 = sub(B, A);
  const f3 e2 = sub(C, A);
  const f3 pvec = cross(d, e2);
  const f3 tvec = sub(o, A);
  const f3 qvec = cross(tvec, e1);
  const float det = dot(e1, pvec);
  const float sv = sgn(det);
  const float invDet = sv / fmax_std(sv * det, 1e-6f);
  v = dot(tvec, pvec) * invDet;
  u = dot(qvec, d) * invDet;
  t = dot(e2, qvec) * invDet;
  return v > 0.0f && u > 0.0f && (u + v < 1.0f) && t > tNear;
}

Hit miss_hit(float tFar) { Hit h; h.prim = h.inst = h.geom = 0xFFFFFFFFu; h.t = tFar; h.u = h.v = 0.0f; h.gtri = 0xFFFFFFFFu; return h; }

Hit trace_bf(const Scene& s, f3 o, f3 d, float tNear, float tFar) {
  Hit h = miss_hit(tFar);
  for (int g = 0; g < s.F; g++) {
    float t, u, v;
    if (mt_test(s, g, o, d, tNear, t, u, v) && t < h.t) {
      h.t = t; h.u = u; h.v = v; h.gtri = g;
    }
  }
  if (h.gtri != 0xFFFFFFFFu) { h.geom = s.tri_mesh[h.gtri]; h.inst = 0; h.prim = h.gtri - s.to[h.geom]; }
  return h;
}

void build_bvh(Scene& s) {
  s.nodes.clear();
  s.order.resize(s.F);
  for (int i = 0; i < s.F; i++) s.order[i] = i;
  std::vector<f3> cen(s.F);
  std::vector<f3> lo(s.F), hi(s.F);
  for (int i = 0; i < s.F; i++) {
    f3 a = s.ta[i], b = s.tb[i], c = s.tcv[i];
    lo[i] = {std::min(a.x, std::min(b.x, c.x)), std::min(a.y, std::min(b.y, c.y)), std::min(a.z, std::min(b.z, c.z))};
    hi[i] = {std::max(a.x, std::max(b.x, c.x)), std::max(a.y, std::max(b.y, c.y)), std::max(a.z, std::max(b.z, c.z))};
    cen[i] = {0.5f * (lo[i].x + hi[i].x), 0.5f * (lo[i].y + hi[i].y), 0.5f * (lo[i].z + hi[i].z)};
  }
  struct Item { int node, first, count; };
  std::vector<Item> stack;
  s.nodes.push_back({});
  stack.push_back({0, 0, s.F});
  while (!stack.empty()) {
    Item it = stack.back(); stack.pop_back();
    BvhNode nd;
    float bl[3] = {INFINITY, INFINITY, INFINITY}, bh[3] = {-INFINITY, -INFINITY, -INFINITY};
    float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = it.first; k < it.first + it.count; k++) {
      int i = s.order[k];
      const float l[3] = {lo[i].x, lo[i].y, lo[i].z}, h[3] = {hi[i].x, hi[i].y, hi[i].z};
      const float c[3] = {cen[i].x, cen[i].y, cen[i].z};
      for (int a = 0; a < 3; a++) { bl[a] = std::min(bl[a], l[a]); bh[a] = std::max(bh[a], h[a]); cl[a] = std::min(cl[a], c[a]); ch[a] = std::max(ch[a], c[a]); }
    }
    for (int a = 0; a < 3; a++) {
      // conservative padding: the MT test may accept points a few ulps outside the triangle
      const float pad = 1e-5f * (std::fabs(bl[a]) + std::fabs(bh[a]) + (bh[a] - bl[a])) + 1e-30f;
      nd.lo[a] = bl[a] - pad; nd.hi[a] = bh[a] + pad;
    }
    if (it.count <= 4) {
      nd.left = nd.right = -1; nd.first = it.first; nd.count = it.count;
      s.nodes[it.node] = nd;
      continue;
    }
    int axis = 0;
    for (int a = 1; a < 3; a++) if (ch[a] - cl[a] > ch[axis] - cl[axis]) axis = a;
    const int mid = it.first + it.count / 2;
    std::nth_element(s.order.begin() + it.first, s.order.begin() + mid, s.order.begin() + it.first + it.count,
                     [&](int x, int y) {
                       const float cx = axis == 0 ? cen[x].x : axis == 1 ? cen[x].y : cen[x].z;
                       const float cy = axis == 0 ? cen[y].x : axis == 1 ? cen[y].y : cen[y].z;
                       return cx < cy || (cx == cy && x < y);
                     });
    nd.left = (int)s.nodes.size();
    nd.right = nd.left + 1;
    nd.first = nd.count = 0;
    s.nodes[it.node] = nd;
    s.nodes.push_back({});
    s.nodes.push_back({});
    stack.push_back({nd.left, it.first, mid - it.first});
    stack.push_back({nd.right, mid, it.first + it.count - mid});
  }
  s.bvh_built = true;
}

inline bool box_hit(const BvhNode& n, const float o[3], const float id[3], float tmax, float& tnear) {
  float t0 = 0.0f, t1 = tmax;
  for (int a = 0; a < 3; a++) {
    float ta = (n.lo[a] - o[a]) * id[a];
    float tb = (n.hi[a] - o[a]) * id[a];
    if (ta != ta) ta = -INFINITY;  // 0 * inf: ray parallel to the slab and on its plane
    if (tb != tb) tb = INFINITY;
    const float tl = std::min(ta, tb), th = std::max(ta, tb);
    t0 = std::max(t0, tl);
    t1 = std::min(t1, th);
  }
  tnear = t0;
  return t0 <= t1;
}

// Closest hit through the BVH; the same MT test and the BF tie rule (on equal t the
// triangle that comes first in (mesh, triangle) order wins).
Hit trace_bvh(const Scene& s, f3 o, f3 d, float tNear, float tFar) {
  Hit h = miss_hit(tFar);
  if (s.F == 0) return h;
  const float oo[3] = {o.x, o.y, o.z};
  const float id[3] = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
  if (o.x != o.x || o.y != o.y || o.z != o.z || d.x != d.x || d.y != d.y || d.z != d.z) return h;
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BvhNode& n = s.nodes[stack[--sp]];
    float tn;
    const float lim = h.t * 1.0001f + 1e-6f;
    if (!box_hit(n, oo, id, lim, tn)) continue;
    if (n.left < 0) {
      for (int k = n.first; k < n.first + n.count; k++) {
        const int g = s.order[k];
        float t, u, v;
        if (mt_test(s, g, o, d, tNear, t, u, v) && (t < h.t || (t == h.t && (uint32_t)g < h.gtri))) {
          h.t = t; h.u = u; h.v = v; h.gtri = g;
        }
      }
    } else {
      stack[sp++] = n.left;
      stack[sp++] = n.right;
    }
  }
  if (h.gtri != 0xFFFFFFFFu) { h.geom = s.tri_mesh[h.gtri]; h.inst = 0; h.prim = h.gtri - s.to[h.geom]; }
  return h;
}

// Closest hit (camera rays) or, with shadow = true, a shadow ray's hit.  The BVH path adds
// the suspect triangles of the ray family, so both tracers return the BF answer.
inline Hit trace(const Scene& s, f3 o, f3 d, float tNear = 0.0f, float tFar = 1e9f, bool shadow = false) {
  if (s.tracer != ORC_TRACER_BVH) return trace_bf(s, o, d, tNear, tFar);
  Hit h = trace_bvh(s, o, d, tNear, tFar);
  for (int g : (shadow ? s.sus_light : s.sus_cam)) {
    float t, u, v;
    if (mt_test(s, g, o, d, tNear, t, u, v) && (t < h.t || (t == h.t && h.gtri != 0xFFFFFFFFu && (uint32_t)g < h.gtri))) {
      h.t = t; h.u = u; h.v = v; h.gtri = g;
    }
  }
  if (h.gtri != 0xFFFFFFFFu) { h.geom = s.tri_mesh[h.gtri]; h.inst = 0; h.prim = h.gtri - s.to[h.geom]; }
  return h;
}

struct Cam {
  M44 proj_inv, wv_inv;
  float W, H;
  float wvp[16];  // CamInfo::mWVP = mProj*mWorldView in memory order (camera.h:39-43)
};
Cam make_cam(const orc_camera& c) {
  Cam r;
  const M44 wv = load44(c.world_view), pr = load44(c.proj);
  r.proj_inv = inverse4x4(pr);   // BruteForce3D::SetCamera (raytrace3d_bf.cpp:93-100)
  r.wv_inv = inverse4x4(wv);
  r.W = c.width; r.H = c.height;
  M44 wvp = mm(pr, wv);
  std::memcpy(r.wvp, wvp.m, sizeof r.wvp);
  return r;
}

// The camera origin every CastSingleRay of this view starts from: WorldViewInv*(0,0,0,1).
inline f3 cam_origin(const Cam& c) { return mp(c.wv_inv, f3{0, 0, 0}); }

// CastSingleRay (raytrace3d_bf.cpp:158-174) + EyeRayDirNormalized / transform_ray3f (raytrace.h:49-76)
inline void eye_ray(const Cam& c, float x, float y, f3& o, f3& d) {
  const float xn = x / c.W, yn = y / c.H;
  f4 p = {2.0f * xn - 1.0f, -2.0f * yn + 1.0f, 0.0f, 1.0f};
  p = mv(c.proj_inv, p);
  p = {p.x / p.w, p.y / p.w, p.z / p.w, p.w / p.w};
  f3 dir = normalize(f3{p.x, p.y, p.z});
  const f3 o0 = {0, 0, 0};
  const f3 pos = mp(c.wv_inv, o0);
  const f3 pos2 = mp(c.wv_inv, add(o0, mul(100.0f, dir)));
  o = pos;
  d = normalize(sub(pos2, pos));
}

// ----------------------------------------------------------------------------
// Shading (src/shade_simple.cpp, src/shade_lambert.cpp, src/shade_common.cpp)
// ----------------------------------------------------------------------------
struct Lam { f3 L, Lc, amb; };

inline f3 tex_fetch(const Scene& s, int mesh, int x, int y) {
  const float* p = s.texels.data() + s.tex_off[mesh] + 3 * (y * s.tex_w[mesh] + x);
  return {p[0], p[1], p[2]};
}
// sample_bilinear_clamp_3f (shade_common.cpp:3-21)
f3 bilinear(const Scene& s, int mesh, f2 tc) {
  const int w = s.tex_w[mesh], h = s.tex_h[mesh];
  tc.x = clampf(tc.x, 0.0f, 1.0f); tc.y = clampf(tc.y, 0.0f, 1.0f);
  tc.x *= (float)w; tc.y *= (float)h;
  const int ix = (int)tc.x, iy = (int)tc.y;
  const int x0 = clampi(ix, 0, w - 1), y0 = clampi(iy, 0, h - 1);
  const int x1 = clampi(ix + 1, 0, w - 1), y1 = clampi(iy + 1, 0, h - 1);
  const float dx = tc.x - (float)x0, dy = tc.y - (float)y0;
  const f3 p00 = tex_fetch(s, mesh, x0, y0), p01 = tex_fetch(s, mesh, x0, y1);
  const f3 p10 = tex_fetch(s, mesh, x1, y0), p11 = tex_fetch(s, mesh, x1, y1);
  const float P00[3] = {p00.x, p00.y, p00.z}, P01[3] = {p01.x, p01.y, p01.z};
  const float P10[3] = {p10.x, p10.y, p10.z}, P11[3] = {p11.x, p11.y, p11.z};
  float r[3];
  for (int i = 0; i < 3; i++)
    r[i] = (1 - dx) * (1 - dy) * P00[i] + (1 - dx) * dy * P01[i] + dx * (1 - dy) * P10[i] + dx * dy * P11[i];
  return {r[0], r[1], r[2]};
}
// sample_bilinear_clamp_3f_grad (shade_common.cpp:23-45): tc is NOT clamped here.  The terms
// are float products as in the reference; Acc is the type they are summed in.
template <class Acc>
void bilinear_grad(const Scene& s, int mesh, f3 val, f2 tc, Acc* g) {
  const int w = s.tex_w[mesh], h = s.tex_h[mesh];
  const float px = tc.x * (float)w, py = tc.y * (float)h;
  const int ix = (int)px, iy = (int)py;
  const int x0 = clampi(ix, 0, w - 1), y0 = clampi(iy, 0, h - 1);
  const int x1 = clampi(ix + 1, 0, w - 1), y1 = clampi(iy + 1, 0, h - 1);
  const float dx = px - (float)x0, dy = py - (float)y0;
  Acc* d00 = g + 3 * (y0 * w + x0);
  Acc* d01 = g + 3 * (y1 * w + x0);
  Acc* d10 = g + 3 * (y0 * w + x1);
  Acc* d11 = g + 3 * (y1 * w + x1);
  d00[0] += (1 - dx) * (1 - dy) * val.x; d00[1] += (1 - dx) * (1 - dy) * val.y; d00[2] += (1 - dx) * (1 - dy) * val.z;
  d01[0] += (1 - dx) * dy * val.x;       d01[1] += (1 - dx) * dy * val.y;       d01[2] += (1 - dx) * dy * val.z;
  d10[0] += dx * (1 - dy) * val.x;       d10[1] += dx * (1 - dy) * val.y;       d10[2] += dx * (1 - dy) * val.z;
  d11[0] += dx * dy * val.x;             d11[1] += dx * dy * val.y;             d11[2] += dx * dy * val.z;
}

// barycentric blend a*(1-u-v) + b*v + u*c (shade_simple.cpp:36)
inline f3 blend3(f3 a, f3 b, f3 c, float u, float v) {
  const float w = 1.0f - u - v;
  return add(add(mul(a, w), mul(b, v)), mul(u, c));
}
inline f2 blend2(f2 a, f2 b, f2 c, float u, float v) {
  const float w = 1.0f - u - v;
  return {a.x * w + b.x * v + u * c.x, a.y * w + b.y * v + u * c.y};
}

// LAMBERT's light factor at a closest hit (shade_lambert.cpp:26-34):
// (shadow * Lc) * max(0, n.(-L)) + ambient with the interpolated, un-normalised normal n and
// shadow = 1 unless the ray from pos + (t - 1e-5) dir toward -L hits anything.
f3 lambert_light(const Scene& s, const Lam& lam, f3 o, f3 d, const Hit& h, const std::vector<f3>& nrm) {
  const uint32_t g = h.gtri;
  const uint32_t A = s.idx[3 * g + 0], B = s.idx[3 * g + 1], C = s.idx[3 * g + 2];
  const f3 n = blend3(nrm[A], nrm[B], nrm[C], h.u, h.v);
  const float BIAS = 1e-5;
  const f3 surf = add(o, mul(h.t - BIAS, d));
  const f3 negL = mul(-1.0f, lam.L);
  const float shadow = trace(s, surf, negL, 0.0f, 1e9f, true).prim == 0xFFFFFFFFu ? 1.0f : 0.0f;
  const f3 lit = mul(mul(shadow, lam.Lc), fmax_std(0.0f, dot(n, negL)));
  return add(lit, lam.amb);
}

// shade<m> (shade_simple.cpp:6-37,113-130; shade_lambert.cpp:6-35). `nrm` selects the
// normal array: the 3D scene's for render, the projected scene's for edge samples.
// ORC_LAMBERT_VERTEX_COLOR: LAMBERT with the interpolated vertex colour as albedo.
f3 shade(const Scene& s, int mode, const Lam& lam, const Cam& c, f2 sp, const std::vector<f3>& nrm) {
  f3 o, d;
  eye_ray(c, sp.x, sp.y, o, d);
  const Hit h = trace(s, o, d);
  if (h.prim == 0xFFFFFFFFu) return {0, 0, 0};
  if (mode == ORC_SILHOUETTE) return {1, 1, 1};
  const uint32_t g = h.gtri;
  const uint32_t A = s.idx[3 * g + 0], B = s.idx[3 * g + 1], C = s.idx[3 * g + 2];
  if (mode == ORC_VERTEX_COLOR) return blend3(s.col[A], s.col[B], s.col[C], h.u, h.v);
  f3 diffuse;
  if (mode == ORC_LAMBERT_VERTEX_COLOR) {
    diffuse = blend3(s.col[A], s.col[B], s.col[C], h.u, h.v);   // the vertex colour as albedo
  } else {
    const f2 tc = blend2(s.tc[A], s.tc[B], s.tc[C], h.u, h.v);
    diffuse = bilinear(s, h.geom, tc);
    if (mode == ORC_TEXTURE_COLOR) return diffuse;
  }
  return mulv(lambert_light(s, lam, o, d, h, nrm), diffuse);
}

// ----------------------------------------------------------------------------
// Projection and its adjoint (src/functions.h:5-13, src/functions.cpp:21-163)
// ----------------------------------------------------------------------------
inline void vertex_shader(const Cam& c, float vx, float vy, float vz, float out[2]) {
  const float* M = c.wvp;
  const float W = vx * M[3] + vy * M[7] + vz * M[11] + M[15];
  const float xNDC = (vx * M[0] + vy * M[4] + vz * M[8] + M[12]) / W;
  const float yNDC = -(vx * M[1] + vy * M[5] + vz * M[9] + M[13]) / W;
  out[0] = (xNDC * 0.5f + 0.5f) * c.W;
  out[1] = (yNDC * 0.5f + 0.5f) * c.H;
}
// d(screen x)/dV and d(screen y)/dV of vertex_shader; the reverse-mode chain
// (d/dNDC = size/2, then the quotient rule) with the reference's operation order.
void vs_grad(const Cam& c, f3 V, int axis, float dV[3]) {
  const float* M = c.wvp;
  const float W = V.x * M[3] + V.y * M[7] + V.z * M[11] + M[15];
  const int r = axis == 0 ? 0 : 1;
  float num = V.x * M[r] + V.y * M[4 + r] + V.z * M[8 + r] + M[12 + r];
  if (axis == 1) num = -num;
  const float size = axis == 0 ? c.W : c.H;
  const float dN = (1 * size) * 0.5F;      // d out / d NDC
  const float q = dN / W;                   // d out / d numerator
  const float dW = dN * -num / (W * W);     // d out / d W
  if (axis == 0) {
    dV[0] = 0 + q * M[0]; dV[1] = 0 + q * M[4]; dV[2] = 0 + q * M[8];
  } else {
    dV[0] = 0 + -q * M[1]; dV[1] = 0 + -q * M[5]; dV[2] = 0 + -q * M[9];
  }
  dV[0] += dW * M[3]; dV[1] += dW * M[7]; dV[2] += dW * M[11];
}

struct EdgeRec { int v0, v1, vn0, vn1, mesh, inst; };

struct EdgeSet {
  std::vector<EdgeRec> edges;
  std::vector<float> pmf, cdf;
  std::vector<f3> proj;  // projected positions of every vertex (sx, sy, z_world)
};

// edge_derivatives part 1 (drender.h:241-302) + build_edge_sampler (drender.cpp:5-22)
void build_edges(const Scene& s, const Cam& c, EdgeSet& E) {
  E.proj.resize(s.V);
  for (int i = 0; i < s.nm; i++)
    for (int k = s.vo[i]; k < s.vo[i + 1]; k++) {
      // transform(m, tr): mul(tr, v) (scene_utils.cpp:87-96), then VertexShader in place
      f3 v = mp(s.tr[i], s.pos_orig[k]);
      float o2[2];
      vertex_shader(c, v.x, v.y, v.z, o2);
      // the projected scene is prepared with an identity transform: I*(x,y,z,1)
      f3 p = {o2[0], o2[1], v.z};
      E.proj[k] = mp(identity(), p);
    }
  // std::set<Edge> ordered by (v0, v1), first insertion kept (drender.h:26-52,274-285)
  std::vector<std::pair<uint64_t, EdgeRec>> keyed;
  for (int i = 0; i < s.nm; i++) {
    if (!s.diff[i]) continue;
    for (int f = s.to[i]; f < s.to[i + 1]; f++) {
      for (int e = 0; e < 3; e++) {
        const int a = (int)s.idx[3 * f + e], b = (int)s.idx[3 * f + (e + 1) % 3];
        const int an = (int)s.idx_local[3 * f + e], bn = (int)s.idx_local[3 * f + (e + 1) % 3];
        EdgeRec r;
        r.mesh = i; r.inst = 0;
        if (a < b) { r.v0 = a; r.vn0 = an; r.v1 = b; r.vn1 = bn; }
        else { r.v0 = b; r.vn0 = bn; r.v1 = a; r.vn1 = an; }
        keyed.push_back({((uint64_t)(uint32_t)r.v0 << 32) | (uint32_t)r.v1, r});
      }
    }
  }
  std::stable_sort(keyed.begin(), keyed.end(), [](const std::pair<uint64_t, EdgeRec>& x, const std::pair<uint64_t, EdgeRec>& y) { return x.first < y.first; });
  E.edges.clear();
  for (size_t i = 0; i < keyed.size(); i++)
    if (i == 0 || keyed[i].first != keyed[i - 1].first) E.edges.push_back(keyed[i].second);
  const int n = (int)E.edges.size();
  E.pmf.resize(n);
  E.cdf.resize(n + 1);
  E.cdf[0] = 0;
  for (int i = 0; i < n; i++) {
    const f3 p0 = E.proj[E.edges[i].v0], p1 = E.proj[E.edges[i].v1];
    E.pmf[i] = length(sub(p1, p0));
    E.cdf[i + 1] = E.pmf[i] + E.cdf[i];
  }
  const float total = E.cdf[n];
  for (auto& p : E.pmf) p /= total;
  for (auto& p : E.cdf) p /= total;
}

// sample() (drender.cpp:25-29): clamp(upper_bound(cdf, u) - 1, 0, E - 1)
inline int sample_edge(const std::vector<float>& cdf, float u) {
  const int ub = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
  return clampi(ub - 1, 0, (int)cdf.size() - 2);
}

}  // namespace

// ============================================================================
// C API
// ============================================================================
extern "C" {

void* orc_scene_create(int n_meshes, const int* vcount, const int* fcount, const float* transforms,
                       const float* pos, const float* col, const float* tc, const float* nrm,
                       const uint32_t* idx, const int* tex_w, const int* tex_h, const float* texels,
                       const int* differentiable) {
  Scene* s = new Scene();
  s->nm = n_meshes;
  s->vo.assign(n_meshes + 1, 0);
  s->to.assign(n_meshes + 1, 0);
  for (int i = 0; i < n_meshes; i++) { s->vo[i + 1] = s->vo[i] + vcount[i]; s->to[i + 1] = s->to[i] + fcount[i]; }
  s->V = s->vo[n_meshes]; s->F = s->to[n_meshes];
  s->tex_w.resize(n_meshes); s->tex_h.resize(n_meshes); s->tex_off.resize(n_meshes); s->diff.resize(n_meshes);
  int toff = 0;
  for (int i = 0; i < n_meshes; i++) {
    s->tex_w[i] = tex_w ? tex_w[i] : 0; s->tex_h[i] = tex_h ? tex_h[i] : 0;
    s->tex_off[i] = toff; toff += 3 * s->tex_w[i] * s->tex_h[i];
    s->diff[i] = differentiable ? differentiable[i] : 1;
  }
  s->texels.assign(texels ? texels : (const float*)nullptr, texels ? texels + toff : (const float*)nullptr);
  s->pos.resize(s->V); s->pos_orig.resize(s->V); s->col.assign(s->V, {0, 0, 0});
  s->tc.assign(s->V, {0, 0}); s->nrm.assign(s->V, {1, 0, 0}); s->nrm2d.assign(s->V, {1, 0, 0});
  s->idx.resize(3 * s->F); s->idx_local.resize(3 * s->F); s->tri_mesh.resize(s->F);
  s->ta.resize(s->F); s->tb.resize(s->F); s->tcv.resize(s->F);
  for (int i = 0; i < n_meshes; i++) {
    const M44 T = load44(transforms + 16 * i);
    s->tr.push_back(T);
    s->tr_inv.push_back(inverse4x4(T));
    const M44 NT = transpose(inverse4x4(T));
    for (int k = s->vo[i]; k < s->vo[i + 1]; k++) {
      const f3 v = {pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]};
      s->pos_orig[k] = v;
      s->pos[k] = mp(T, v);                                   // scene_utils.cpp:150
      if (col) s->col[k] = {col[3 * k], col[3 * k + 1], col[3 * k + 2]};
      if (tc) s->tc[k] = {tc[2 * k], tc[2 * k + 1]};
      if (nrm) {
        const f3 n = {nrm[3 * k], nrm[3 * k + 1], nrm[3 * k + 2]};
        s->nrm[k] = mp(NT, n);                                 // scene_utils.cpp:162 (w = 1)
        s->nrm2d[k] = mp(identity(), mdir(T, n));              // transform() then identity prepare
      }
    }
    for (int f = s->to[i]; f < s->to[i + 1]; f++) {
      s->tri_mesh[f] = i;
      for (int e = 0; e < 3; e++) {
        s->idx_local[3 * f + e] = idx[3 * f + e];
        s->idx[3 * f + e] = (uint32_t)s->vo[i] + idx[3 * f + e];
      }
      // BF applies the instance transform to the already transformed positions
      // (raytrace3d_bf.cpp:123-125)
      s->ta[f] = mp(T, s->pos[s->idx[3 * f + 0]]);
      s->tb[f] = mp(T, s->pos[s->idx[3 * f + 1]]);
      s->tcv[f] = mp(T, s->pos[s->idx[3 * f + 2]]);
    }
  }
  return s;
}

void orc_scene_destroy(void* s) { delete (Scene*)s; }

void orc_set_jitter(void* sp, int on, uint32_t seed) {
  Scene* s = (Scene*)sp;
  s->jit_on = on != 0;
  s->jit_seed = seed;
}

void orc_pixel_sample(int on, uint32_t seed, int view, int x, int y, int si, int sq, float* xo, float* yo) {
  Scene s;
  s.jit_on = on != 0;
  s.jit_seed = seed;
  pixel_sample(s, view, x, y, si, sq, *xo, *yo);
}

void orc_scene_set_tracer(void* sp, int kind) {
  Scene* s = (Scene*)sp;
  s->tracer = kind;
  if (kind == ORC_TRACER_BVH && !s->bvh_built) build_bvh(*s);
}

int orc_grad_size(void* s, int mode) { return grad_layout(*(Scene*)s, mode, nullptr); }

void orc_hits(void* sp, const orc_camera* cam, uint32_t* prim, uint32_t* geom, float* t, float* u, float* v) {
  Scene& s = *(Scene*)sp;
  const Cam c = make_cam(*cam);
  if (s.tracer == ORC_TRACER_BVH) suspects_origin(s, cam_origin(c));
  const int W = (int)cam->width, H = (int)cam->height;
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      f3 o, d;
      eye_ray(c, x + 0.5f, y + 0.5f, o, d);
      const Hit h = trace(s, o, d);
      const int i = y * W + x;
      prim[i] = h.prim; geom[i] = h.geom; t[i] = h.t; u[i] = h.u; v[i] = h.v;
    }
}

static Lam make_lam(const orc_lambert* l) {
  Lam r;
  if (l) {
    r.L = {l->light_dir[0], l->light_dir[1], l->light_dir[2]};
    r.Lc = {l->light_color[0], l->light_color[1], l->light_color[2]};
    r.amb = {l->ambient[0], l->ambient[1], l->ambient[2]};
  } else {  // shade_lambert.cpp:10-12
    r.L = normalize(f3{-1, -1, 0});
    r.Lc = {1, 0.5, 0.2};
    r.amb = mul(0.1f, f3{1, 1, 1});
  }
  return r;
}

void orc_render(void* sp, int mode, int spp, const orc_lambert* lam, const orc_camera* cams, int views, float* images) {
  Scene& s = *(Scene*)sp;
  const Lam L = make_lam(lam);
  if (s.tracer == ORC_TRACER_BVH) suspects_direction(s, mul(-1.0f, L.L));
  const int sq = (int)std::sqrt((float)spp);
  const int nsamp = sq * sq;
  size_t img_off = 0;
  for (int vi = 0; vi < views; vi++) {
    const Cam c = make_cam(cams[vi]);
    if (s.tracer == ORC_TRACER_BVH) suspects_origin(s, cam_origin(c));
    const int W = (int)cams[vi].width, H = (int)cams[vi].height;
    float* img = images + img_off;
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) {
        f3 acc = {0, 0, 0};
        for (int dy = 0; dy < sq; dy++)
          for (int dx = 0; dx < sq; dx++) {
            float xo, yo;
            pixel_sample(s, vi, x, y, dy * sq + dx, sq, xo, yo);
            const f3 col = shade(s, mode, L, c, f2{x + xo, y + yo}, s.nrm);
            acc = add(acc, divs(col, (float)nsamp));
          }
        float* p = img + 3 * ((size_t)y * W + x);
        p[0] = acc.x; p[1] = acc.y; p[2] = acc.z;
      }
    img_off += (size_t)3 * W * H;
  }
}

float orc_loss(const float* image, const float* target, int npix, float* adjoint) {
  double acc = 0.0;
  for (int i = 0; i < npix; i++) {
    const f3 b = {image[3 * i], image[3 * i + 1], image[3 * i + 2]};
    const f3 a = {target[3 * i], target[3 * i + 1], target[3 * i + 2]};
    const f3 dv = sub(b, a);
    const f3 ad = mul(2.0f, dv);
    adjoint[3 * i] = ad.x; adjoint[3 * i + 1] = ad.y; adjoint[3 * i + 2] = ad.z;
    acc += double(dot(dv, dv));
  }
  return float(acc);
}

void orc_d_render(void* sp, int mode, int spp, const orc_lambert* lam, const orc_camera* cams, int views,
                  const float* adjoints, int N, int rng_mode, float* grad) {
  orc_d_render_ex(sp, mode, spp, lam, cams, views, adjoints, N, rng_mode, 0, grad);
}

void orc_d_render_ex(void* sp, int mode, int spp, const orc_lambert* lam, const orc_camera* cams, int views,
                     const float* adjoints, int N, int rng_mode, int flags, float* grad) {
  Scene& s = *(Scene*)sp;
  const Lam L = make_lam(lam);
  const bool aligned = (flags & ORC_INTERIOR_ALIGNED) != 0;
  if (s.tracer == ORC_TRACER_BVH) suspects_direction(s, mul(-1.0f, L.L));
  std::vector<int> moff;
  const int G = grad_layout(s, mode, &moff);
  const int sq = (int)std::sqrt((float)spp);
  const int nsamp = sq * sq;
  std::vector<float> hamm(2 * std::max(spp, 1));
  hammersley(spp, hamm.data());
  std::vector<float> rnd(2 * (size_t)std::max(N, 0));
  if (rng_mode == ORC_RNG_REF_PRNG14) {
    prng_table(N, rnd.data());
  } else {
    uint32_t tab[11 * 31];
    qmc_table(tab);
    for (int i = 0; i < N; i++) { rnd[2 * i] = qmc_rnd((uint32_t)i, 0, tab); rnd[2 * i + 1] = qmc_rnd((uint32_t)i, 1, tab); }
  }
  int nthreads = 1;
#ifdef _OPENMP
  nthreads = omp_get_max_threads();
#endif
  size_t img_off = 0;
  for (int vi = 0; vi < views; vi++) {
    const Cam c = make_cam(cams[vi]);
    if (s.tracer == ORC_TRACER_BVH) suspects_origin(s, cam_origin(c));
    const int W = (int)cams[vi].width, H = (int)cams[vi].height;
    const float* adj = adjoints + img_off;
    auto adj_at = [&](int x, int y) { const float* p = adj + 3 * ((size_t)y * W + x); return f3{p[0], p[1], p[2]}; };
    // Per-thread partial sums in double: the result then does not depend on the thread count
    // or the OpenMP schedule (the reference sums float partials per thread, so its order
    // varies with both; what remains between the two is its float rounding).
    std::vector<std::vector<double>> part(nthreads, std::vector<double>(G, 0.0));
    // ---- interior_derivatives (drender.h:197-233) ----
    if (mode != ORC_SILHOUETTE) {
#pragma omp parallel for schedule(dynamic, 2)
      for (int y = 0; y < H; y++) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* g = part[tid].data();
        for (int x = 0; x < W; x++)
          for (int si = 0; si < nsamp; si++) {
            // Hammersley offsets (drender.h:217-218), or with `aligned` the render's sample
            // positions (drender.h:124-128, jittered when jitter is on)
            float xo = hamm[2 * si], yo = hamm[2 * si + 1];
            if (aligned) pixel_sample(s, vi, x, y, si, sq, xo, yo);
            const f3 val = divs(adj_at(x, y), (float)nsamp);
            f3 o, d;
            eye_ray(c, x + xo, y + yo, o, d);
            const Hit h = trace(s, o, d);
            if (h.prim == 0xFFFFFFFFu) continue;
            const int m = (int)h.geom;
            if (moff[m] < 0) continue;
            const uint32_t gt = h.gtri;
            const int nv = s.vo[m + 1] - s.vo[m];
            if (mode == ORC_VERTEX_COLOR || mode == ORC_LAMBERT_VERTEX_COLOR) {
              // shade_grad<VERTEX_COLOR> (shade_simple.cpp:39-74); with the light factor of the
              // hit for LAMBERT_VERTEX_COLOR (the derivative of its shade in the colours)
              const f3 vs = mode == ORC_VERTEX_COLOR ? val : mulv(val, lambert_light(s, L, o, d, h, s.nrm));
              const float w = 1.0f - h.u - h.v;
              const f3 cA = mul(w, vs), cB = mul(h.v, vs), cC = mul(h.u, vs);
              double* gc = g + moff[m] + 3 * nv;
              const uint32_t A = s.idx_local[3 * gt], B = s.idx_local[3 * gt + 1], C = s.idx_local[3 * gt + 2];
              gc[3 * A] += cA.x; gc[3 * A + 1] += cA.y; gc[3 * A + 2] += cA.z;
              gc[3 * B] += cB.x; gc[3 * B + 1] += cB.y; gc[3 * B + 2] += cB.z;
              gc[3 * C] += cC.x; gc[3 * C + 1] += cC.y; gc[3 * C + 2] += cC.z;
            } else {
              // shade_grad<TEXTURE_COLOR> (shade_simple.cpp:132-162), used by LAMBERT too
              if (s.tex_w[m] * s.tex_h[m] == 0) continue;
              const uint32_t A = s.idx[3 * gt], B = s.idx[3 * gt + 1], C = s.idx[3 * gt + 2];
              const f2 tcv = blend2(s.tc[A], s.tc[B], s.tc[C], h.u, h.v);
              bilinear_grad(s, m, val, tcv, g + moff[m] + 3 * nv);
            }
          }
      }
    }
    // ---- edge_derivatives (drender.h:235-425) ----
    EdgeSet E;
    build_edges(s, c, E);
    const int ne = (int)E.edges.size();
    if (ne > 0) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < N; i++) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* g = part[tid].data();
        const float rnd0 = rnd[2 * i], rnd1 = rnd[2 * i + 1];
        const int eid = sample_edge(E.cdf, rnd0);
        const EdgeRec& e = E.edges[eid];
        const float pmf = E.pmf[eid];
        const f2 v0 = {E.proj[e.v0].x, E.proj[e.v0].y}, v1 = {E.proj[e.v1].x, E.proj[e.v1].y};
        const float t = rnd1;
        const f2 p = {v0.x + t * (v1.x - v0.x), v0.y + t * (v1.y - v0.y)};
        // int() truncation; NaN counts as off-screen (x86 cvttss2si yields INT_MIN)
        if (p.x != p.x || p.y != p.y) continue;
        const int xi = (p.x >= 2147483648.0f || p.x < -2147483648.0f) ? INT32_MIN : (int)p.x;
        const int yi = (p.y >= 2147483648.0f || p.y < -2147483648.0f) ? INT32_MIN : (int)p.y;
        if (xi < 0 || yi < 0 || xi >= W || yi >= H) continue;
        const f3 a = adj_at(xi, yi);
        if ((double)length(a) < 1e-3) continue;
        const f2 dv = {v1.x - v0.x, v1.y - v0.y};
        const float len = length2(dv);
        const f2 nd = {dv.x / len, dv.y / len};
        const f2 n = {-nd.y, nd.x};  // normal2D (utils.h:29)
        const f2 cin = {p.x - 1e-3f * n.x, p.y - 1e-3f * n.y};
        const f2 cout = {p.x + 1e-3f * n.x, p.y + 1e-3f * n.y};
        const f3 col_in = shade(s, mode, L, c, cin, s.nrm2d);
        const f3 col_out = shade(s, mode, L, c, cout, s.nrm2d);
        const float pdf = pmf / len;
        const float weight = 1.0f / (pdf * float(N));
        const float adjv = dot(sub(col_in, col_out), a);
        if (!(adjv * adjv > 0.0f)) continue;
        const f2 d0 = {((1 - t) * n.x) * adjv * weight, ((1 - t) * n.y) * adjv * weight};
        const f2 d1 = {(t * n.x) * adjv * weight, (t * n.y) * adjv * weight};
        // edge_grad (drender.cpp:31-87), diagonal screen Jacobian as in the reference
        float g0x[3], g0y[3], g1x[3], g1y[3];
        vs_grad(c, s.pos[e.v0], 0, g0x); vs_grad(c, s.pos[e.v0], 1, g0y);
        vs_grad(c, s.pos[e.v1], 0, g1x); vs_grad(c, s.pos[e.v1], 1, g1y);
        const float dv0_dx = g0x[0] * d0.x;
        const float dv0_dy = g0y[1] * d0.y;
        const float dv0_dz = (g0x[2] * d0.x + g0y[2] * d0.y);
        const float dv1_dx = g1x[0] * d1.x;
        const float dv1_dy = g1y[1] * d1.y;
        const float dv1_dz = (g1x[2] * d1.x + g1y[2] * d1.y);
        const M44& T = s.tr_inv[0];  // mesh 0's inverse transform for every mesh (drender.cpp:60)
        if (moff[e.mesh] < 0) continue;
        double* gp = g + moff[e.mesh];
        gp[e.vn0 * 3 + 0] += T(0,0) * dv0_dx + T(0,1) * dv0_dy + T(0,2) * dv0_dz;
        gp[e.vn0 * 3 + 1] += T(1,0) * dv0_dx + T(1,1) * dv0_dy + T(1,2) * dv0_dz;
        gp[e.vn0 * 3 + 2] += T(2,0) * dv0_dx + T(2,1) * dv0_dy + T(2,2) * dv0_dz;
        gp[e.vn1 * 3 + 0] += T(0,0) * dv1_dx + T(0,1) * dv1_dy + T(0,2) * dv1_dz;
        gp[e.vn1 * 3 + 1] += T(1,0) * dv1_dx + T(1,1) * dv1_dy + T(1,2) * dv1_dz;
        gp[e.vn1 * 3 + 2] += T(2,0) * dv1_dx + T(2,1) * dv1_dy + T(2,2) * dv1_dz;
        const int nvm = s.vo[e.mesh + 1] - s.vo[e.mesh];
        int trs = 3 * nvm;
        if (mode == ORC_VERTEX_COLOR || mode == ORC_LAMBERT_VERTEX_COLOR) trs += 3 * nvm;
        if (mode == ORC_TEXTURE_COLOR || mode == ORC_LAMBERT) trs += 3 * s.tex_w[e.mesh] * s.tex_h[e.mesh];
        double* gt = gp + trs + 12 * e.inst;
        const f3 o0 = s.pos_orig[e.v0], o1 = s.pos_orig[e.v1];
        gt[0] += o0.x * dv0_dx + o1.x * dv1_dx; gt[1] += o0.y * dv0_dx + o1.y * dv1_dx;
        gt[2] += o0.z * dv0_dx + o1.z * dv1_dx; gt[3] += 1 * dv0_dx + 1 * dv1_dx;
        gt[4] += o0.x * dv0_dy + o1.x * dv1_dy; gt[5] += o0.y * dv0_dy + o1.y * dv1_dy;
        gt[6] += o0.z * dv0_dy + o1.z * dv1_dy; gt[7] += 1 * dv0_dy + 1 * dv1_dy;
        gt[8] += o0.x * dv0_dz + o1.x * dv1_dz; gt[9] += o0.y * dv0_dz + o1.y * dv1_dz;
        gt[10] += o0.z * dv0_dz + o1.z * dv1_dz; gt[11] += 1 * dv0_dz + 1 * dv1_dz;
      }
    }
    for (int j = 0; j < G; j++) {
      double sum = 0.0;
      for (int t = 0; t < nthreads; t++) sum += part[t][j];
      grad[j] += (float)sum;
    }
    img_off += (size_t)3 * W * H;
  }
}

int orc_edge_sampler(void* sp, const orc_camera* cam, int* v0, int* v1, float* pmf, float* cdf, float* proj) {
  const Scene& s = *(Scene*)sp;
  const Cam c = make_cam(*cam);
  EdgeSet E;
  build_edges(s, c, E);
  const int n = (int)E.edges.size();
  if (v0) {
    for (int i = 0; i < n; i++) { v0[i] = E.edges[i].v0; v1[i] = E.edges[i].v1; pmf[i] = E.pmf[i]; }
    for (int i = 0; i <= n; i++) cdf[i] = E.cdf[i];
  }
  if (proj)
    for (int k = 0; k < s.V; k++) { proj[3 * k] = E.proj[k].x; proj[3 * k + 1] = E.proj[k].y; proj[3 * k + 2] = E.proj[k].z; }
  return n;
}

void orc_qmc_table(uint32_t* table) { qmc_table(table); }
float orc_qmc_rnd(uint32_t pos, int dim, const uint32_t* table) { return qmc_rnd(pos, dim, table); }
void orc_prng_table(int n, float* out) { prng_table(n, out); }
void orc_hammersley(int n, float* out) { hammersley(n, out); }

// perspectiveMatrix (LiteMath.h:1706-1727)
void orc_perspective(float fovy, float aspect, float zNear, float zFar, float* out) {
  const float ymax = zNear * tanf(fovy * 3.14159265358979323846f / 360.0f);
  const float xmax = ymax * aspect;
  const float left = -xmax, right = +xmax, bottom = -ymax, top = +ymax;
  const float temp = 2.0f * zNear, temp2 = right - left, temp3 = top - bottom, temp4 = zFar - zNear;
  float m[16] = {temp / temp2, 0, 0, 0,
                 0, temp / temp3, 0, 0,
                 (right + left) / temp2, (top + bottom) / temp3, (-zFar - zNear) / temp4, -1.0f,
                 0, 0, (-temp * zFar) / temp4, 0};
  std::memcpy(out, m, sizeof m);
}

// lookAt (LiteMath.h:1679-1704)
void orc_look_at(const float* e, const float* ce, const float* u, float* out) {
  f3 z = {e[0] - ce[0], e[1] - ce[1], e[2] - ce[2]};
  z = normalize(z);
  f3 y = {u[0], u[1], u[2]};
  f3 x = cross(y, z);
  y = cross(z, x);
  x = normalize(x);
  y = normalize(y);
  float m[16] = {x.x, y.x, z.x, 0.0f,
                 x.y, y.y, z.y, 0.0f,
                 x.z, y.z, z.z, 0.0f,
                 -x.x * e[0] - x.y * e[1] - x.z * e[2],
                 -y.x * e[0] - y.y * e[1] - y.z * e[2],
                 -z.x * e[0] - z.y * e[1] - z.z * e[2], 1.0f};
  std::memcpy(out, m, sizeof m);
}

void orc_inverse4x4(const float* m, float* out) { M44 r = inverse4x4(load44(m)); std::memcpy(out, r.m, sizeof r.m); }
void orc_mat_mul(const float* a, const float* b, float* out) { M44 r = mm(load44(a), load44(b)); std::memcpy(out, r.m, sizeof r.m); }

// OptSimple::OptStep, alg == GD_Adam (optimizer_simple.cpp:188-236) for one mesh block.
// m_GSquare is indexed without the mesh offset in the normalisation line (:224);
// with one differentiable mesh that is the same element.
void orc_optsimple_step(int n, float* grad, float* m, float* G, int nblocks, const int* bs, const float* blr) {
  const float alpha = 0.5f, beta = 0.25f;
  for (int i = 0; i < n; i++) m[i] = m[i] * beta + grad[i] * (1.0f - beta);
  for (int i = 0; i < n; i++) G[i] = 2.0f * (G[i] * alpha + (grad[i] * grad[i]) * (1.0f - alpha));
  for (int i = 0; i < n; i++) grad[i] = m[i];
  for (int i = 0; i < n; i++) grad[i] = grad[i] / (std::sqrt(G[i] + float(1e-8f)));
  for (int b = 0; b < nblocks; b++) {
    const int end = (b == nblocks - 1) ? n : bs[b + 1];
    for (int j = bs[b]; j < end; j++) grad[j] *= blr[b];
  }
}

void orc_adam_step(int n, float* p, const float* g, float* m, float* v, int step, float lr, float b1, float b2, float eps) {
  const double bc1 = 1.0 - std::pow((double)b1, step), bc2 = 1.0 - std::pow((double)b2, step);
  for (int i = 0; i < n; i++) {
    m[i] = b1 * m[i] + (1.0f - b1) * g[i];
    v[i] = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
    const double mh = m[i] / bc1, vh = v[i] / bc2;
    p[i] = (float)(p[i] - lr * mh / (std::sqrt(vh) + eps));
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// FP64 finite-difference pin of the edge-sampled position gradient (the check of
// src/fin_diff.cpp / Tester::test_fin_diff, src/tests.cpp:491-676, made exact): in SILHOUETTE
// mode the image, box-filtered over each pixel, is the area of the pixel the mesh covers.
// For a closed convex mesh seen from outside the front-facing triangles tile the silhouette
// without overlap, so L(x) = sum_p a_p * covered_area_p(x), a_p = sum of the pixel's adjoint
// channels, is a sum over front-facing triangles of sum_p a_p * area(triangle ∩ pixel p),
// each area computed exactly (Sutherland-Hodgman clip to the pixel square, shoelace) in
// double.  Edge sampling estimates dL/dx (its colour jump is 1 across the silhouette and 0
// elsewhere), so central differences of L in double pin it.  The projection is the
// reference's VertexShader (src/functions.h:5-13) in double with mProj*mWorldView.
// ---------------------------------------------------------------------------
namespace {
struct P2 { double x, y; };
double clip_area(const P2* tri, double x0, double y0) {
  P2 a[8], b[8];
  int n = 3;
  for (int i = 0; i < 3; i++) a[i] = tri[i];
  const double lo[2] = {x0, y0}, hi[2] = {x0 + 1.0, y0 + 1.0};
  for (int e = 0; e < 4 && n > 0; e++) {
    const int ax = e & 1;          // 0: x, 1: y
    const bool upper = e >= 2;
    const double c = upper ? hi[ax] : lo[ax];
    auto inside = [&](const P2& p) { const double v = ax ? p.y : p.x; return upper ? v <= c : v >= c; };
    int m = 0;
    for (int i = 0; i < n; i++) {
      const P2 &P = a[i], &Q = a[(i + 1) % n];
      const bool ip = inside(P), iq = inside(Q);
      if (ip) b[m++] = P;
      if (ip != iq) {
        const double pv = ax ? P.y : P.x, qv = ax ? Q.y : Q.x;
        const double t = (c - pv) / (qv - pv);
        b[m++] = P2{P.x + t * (Q.x - P.x), P.y + t * (Q.y - P.y)};
      }
    }
    n = m;
    for (int i = 0; i < n; i++) a[i] = b[i];
  }
  double ar = 0.0;
  for (int i = 0; i < n; i++) { const P2 &P = a[i], &Q = a[(i + 1) % n]; ar += P.x * Q.y - Q.x * P.y; }
  return std::fabs(0.5 * ar);
}
struct FdCam { double m[16]; double W, H; int w, h; };
P2 project_d(const FdCam& c, const double* v) {
  const double* M = c.m;  // column-major mWVP
  const double W = v[0] * M[3] + v[1] * M[7] + v[2] * M[11] + M[15];
  const double xN = (v[0] * M[0] + v[1] * M[4] + v[2] * M[8] + M[12]) / W;
  const double yN = -(v[0] * M[1] + v[1] * M[5] + v[2] * M[9] + M[13]) / W;
  return P2{(xN * 0.5 + 0.5) * c.W, (yN * 0.5 + 0.5) * c.H};
}
double tri_functional(const FdCam& c, const double* A, const double* B, const double* C, const double* a) {
  const P2 t[3] = {project_d(c, A), project_d(c, B), project_d(c, C)};
  const double xmin = std::min(t[0].x, std::min(t[1].x, t[2].x)), xmax = std::max(t[0].x, std::max(t[1].x, t[2].x));
  const double ymin = std::min(t[0].y, std::min(t[1].y, t[2].y)), ymax = std::max(t[0].y, std::max(t[1].y, t[2].y));
  const int x0 = std::max(0, (int)std::floor(xmin)), x1 = std::min(c.w - 1, (int)std::floor(xmax));
  const int y0 = std::max(0, (int)std::floor(ymin)), y1 = std::min(c.h - 1, (int)std::floor(ymax));
  double L = 0.0;
  for (int y = y0; y <= y1; y++)
    for (int x = x0; x <= x1; x++) {
      const double ap = a[(size_t)y * c.w + x];
      if (ap != 0.0) L += ap * clip_area(t, x, y);
    }
  return L;
}
}  // namespace

extern "C" {
double orc_silhouette_functional(void* sp, const orc_camera* cam, const float* adjoint) {
  Scene& s = *(Scene*)sp;
  FdCam c;
  M44 wvp = mm(load44(cam->proj), load44(cam->world_view));
  for (int i = 0; i < 16; i++) c.m[i] = wvp.m[i];
  c.W = cam->width; c.H = cam->height; c.w = (int)cam->width; c.h = (int)cam->height;
  std::vector<double> a((size_t)c.w * c.h);
  for (size_t p = 0; p < a.size(); p++) a[p] = (double)adjoint[3 * p] + adjoint[3 * p + 1] + adjoint[3 * p + 2];
  const M44 wi = inverse4x4(load44(cam->world_view));
  const double eye[3] = {wi.m[12], wi.m[13], wi.m[14]};
  double L = 0.0;
  for (int g = 0; g < s.F; g++) {
    double v[3][3];
    for (int k = 0; k < 3; k++) { const f3 q = s.pos[s.idx[3 * g + k]]; v[k][0] = q.x; v[k][1] = q.y; v[k][2] = q.z; }
    const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
    const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
    const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double vd = n[0] * (v[0][0] - eye[0]) + n[1] * (v[0][1] - eye[1]) + n[2] * (v[0][2] - eye[2]);
    if (vd < 0.0) L += tri_functional(c, v[0], v[1], v[2], a.data());
  }
  return L;
}

// Central differences (L(x + h) - L(x - h)) / 2h of the functional above for `n` coordinates
// (3 * global vertex + axis); only the triangles around the vertex change, and a triangle
// keeps the facing it has at x (h is far below any facing change of a non-degenerate one).
void orc_fd_silhouette(void* sp, const orc_camera* cam, const float* adjoint, int n, const int* coords, double h,
                       double* out) {
  Scene& s = *(Scene*)sp;
  FdCam c;
  M44 wvp = mm(load44(cam->proj), load44(cam->world_view));
  for (int i = 0; i < 16; i++) c.m[i] = wvp.m[i];
  c.W = cam->width; c.H = cam->height; c.w = (int)cam->width; c.h = (int)cam->height;
  std::vector<double> a((size_t)c.w * c.h);
  for (size_t p = 0; p < a.size(); p++) a[p] = (double)adjoint[3 * p] + adjoint[3 * p + 1] + adjoint[3 * p + 2];
  const M44 wi = inverse4x4(load44(cam->world_view));
  const double eye[3] = {wi.m[12], wi.m[13], wi.m[14]};
  std::vector<std::vector<int>> adj_tri(s.V);
  for (int g = 0; g < s.F; g++)
    for (int k = 0; k < 3; k++) adj_tri[s.idx[3 * g + k]].push_back(g);
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < n; i++) {
    const int vtx = coords[i] / 3, ax = coords[i] % 3;
    double acc[2] = {0.0, 0.0};
    for (int g : adj_tri[vtx]) {
      double v[3][3];
      for (int k = 0; k < 3; k++) { const f3 q = s.pos[s.idx[3 * g + k]]; v[k][0] = q.x; v[k][1] = q.y; v[k][2] = q.z; }
      const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
      const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
      const double nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      const double vd = nn[0] * (v[0][0] - eye[0]) + nn[1] * (v[0][1] - eye[1]) + nn[2] * (v[0][2] - eye[2]);
      if (!(vd < 0.0)) continue;
      for (int sgn = 0; sgn < 2; sgn++) {
        double w[3][3];
        std::memcpy(w, v, sizeof w);
        for (int k = 0; k < 3; k++)
          if ((int)s.idx[3 * g + k] == vtx) w[k][ax] += sgn == 0 ? h : -h;
        acc[sgn] += tri_functional(c, w[0], w[1], w[2], a.data());
      }
    }
    out[i] = (acc[0] - acc[1]) / (2.0 * h);
  }
}
}  // extern "C"

// ---------------------------------------------------------------------------
// FP64 pixel-area functional of the LAMBERT / LAMBERT_VERTEX_COLOR image, the finite-difference
// pin of the position gradient with its interior term (src/fin_diff.cpp's check, made exact up
// to quadrature): L(x) = sum_p adj_p . (integral over pixel p of the shade), pixels of area 1.
// For a closed convex mesh seen from outside the front-facing triangles tile the silhouette
// and nothing else is visible (background 0).  On triangle T the shade at screen point q is
//   albedo(u, v) * ((lit_T Lc) max(0, n(u, v).(-L)) + ambient)
// (shade<LAMBERT>, src/shade_lambert.cpp:6-35), (u, v) the barycentrics of the camera ray
// through q (eye_ray in double), n the interpolated prepared normal, albedo the bilinear texel
// of the interpolated texture coordinate (LAMBERT) or the interpolated vertex colour; lit_T:
// the reference's shadow ray from a convex mesh's surface meets the mesh iff the face turns
// away from the light.  Each triangle-pixel polygon (Sutherland-Hodgman) is fanned into
// triangles, each integrated with an n x n Gauss-Legendre rule on the collapsed square
// (Duffy's transform); the nodes come from Newton's method on P_n.
// ---------------------------------------------------------------------------
namespace {
struct D3 { double x, y, z; };
D3 d3(f3 a) { return D3{a.x, a.y, a.z}; }
D3 dadd3(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 dsub3(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 dmul3(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
double ddot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 dcross3(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// 4x4 inverse by Gauss-Jordan with partial pivoting, column-major in and out
bool inv44d(const double* m, double* out) {
  double a[4][8];
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) { a[r][c] = m[c * 4 + r]; a[r][c + 4] = r == c ? 1.0 : 0.0; }
  for (int c = 0; c < 4; c++) {
    int piv = c;
    for (int r = c + 1; r < 4; r++) if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
    if (a[piv][c] == 0.0) return false;
    if (piv != c) for (int k = 0; k < 8; k++) std::swap(a[c][k], a[piv][k]);
    const double d = a[c][c];
    for (int k = 0; k < 8; k++) a[c][k] /= d;
    for (int r = 0; r < 4; r++) {
      if (r == c) continue;
      const double f = a[r][c];
      for (int k = 0; k < 8; k++) a[r][k] -= f * a[c][k];
    }
  }
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) out[c * 4 + r] = a[r][c + 4];
  return true;
}

// Sutherland-Hodgman of triangle t against pixel [x0, x0+1] x [y0, y0+1]; returns the count
int clip_poly(const P2* tri, double x0, double y0, P2* out) {
  P2 a[8], b[8];
  int n = 3;
  for (int i = 0; i < 3; i++) a[i] = tri[i];
  const double lo[2] = {x0, y0}, hi[2] = {x0 + 1.0, y0 + 1.0};
  for (int e = 0; e < 4 && n > 0; e++) {
    const int ax = e & 1;
    const bool upper = e >= 2;
    const double c = upper ? hi[ax] : lo[ax];
    auto inside = [&](const P2& p) { const double v = ax ? p.y : p.x; return upper ? v <= c : v >= c; };
    int m = 0;
    for (int i = 0; i < n; i++) {
      const P2 &P = a[i], &Q = a[(i + 1) % n];
      const bool ip = inside(P), iq = inside(Q);
      if (ip) b[m++] = P;
      if (ip != iq) {
        const double pv = ax ? P.y : P.x, qv = ax ? Q.y : Q.x;
        const double t = (c - pv) / (qv - pv);
        b[m++] = P2{P.x + t * (Q.x - P.x), P.y + t * (Q.y - P.y)};
      }
    }
    n = m;
    for (int i = 0; i < n; i++) a[i] = b[i];
  }
  for (int i = 0; i < n; i++) out[i] = a[i];
  return n;
}

// Gauss-Legendre nodes and weights of order n on [0, 1] (Newton on P_n from Chebyshev guesses)
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < n; i++) {
    double z = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; k++) { const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
      const double pn = n == 1 ? z : p1, pm = n == 1 ? 1.0 : p0;
      dp = n * (z * pn - pm) / (z * z - 1.0);
      const double dz = pn / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * dp * dp);   // 2/((1-z^2) P'^2) on [-1, 1], halved for [0, 1]
  }
}

struct LambertFd {
  const Scene* s;
  int mode;
  D3 negL, Lc, amb;
  FdCam fc;
  double proj_inv[16], wv_inv[16];
  D3 eye;
  const float* adj;   // H x W x 3
  std::vector<double> gx, gw;   // Gauss-Legendre nodes on [0, 1] and weights

  void ray(double x, double y, D3& o, D3& d) const {
    const double xn = x / fc.W, yn = y / fc.H;
    const double q[4] = {2.0 * xn - 1.0, -2.0 * yn + 1.0, 0.0, 1.0};
    double p[4];
    for (int r = 0; r < 4; r++) p[r] = proj_inv[r] * q[0] + proj_inv[4 + r] * q[1] + proj_inv[8 + r] * q[2] + proj_inv[12 + r] * q[3];
    D3 dir{p[0] / p[3], p[1] / p[3], p[2] / p[3]};
    dir = dmul3(dir, 1.0 / std::sqrt(ddot3(dir, dir)));
    const D3 far = dmul3(dir, 100.0);
    D3 b{wv_inv[0] * far.x + wv_inv[4] * far.y + wv_inv[8] * far.z + wv_inv[12],
         wv_inv[1] * far.x + wv_inv[5] * far.y + wv_inv[9] * far.z + wv_inv[13],
         wv_inv[2] * far.x + wv_inv[6] * far.y + wv_inv[10] * far.z + wv_inv[14]};
    o = eye;
    d = dsub3(b, o);
    d = dmul3(d, 1.0 / std::sqrt(ddot3(d, d)));
  }
  // shade of global triangle g (positions V[3] in world space) at screen point (x, y)
  D3 shade(int g, const D3* V, bool lit, double x, double y) const {
    D3 o, d;
    ray(x, y, o, d);
    const D3 e1 = dsub3(V[1], V[0]), e2 = dsub3(V[2], V[0]);
    const D3 pv = dcross3(d, e2);
    const double det = ddot3(e1, pv);
    if (det == 0.0) return D3{0, 0, 0};
    const D3 tv = dsub3(o, V[0]);
    const double bb = ddot3(tv, pv) / det;          // weight of B
    const D3 qv = dcross3(tv, e1);
    const double cc = ddot3(d, qv) / det;           // weight of C
    const double aa = 1.0 - bb - cc;
    const uint32_t A = s->idx[3 * g], B = s->idx[3 * g + 1], C = s->idx[3 * g + 2];
    const D3 n = dadd3(dadd3(dmul3(d3(s->nrm[A]), aa), dmul3(d3(s->nrm[B]), bb)), dmul3(d3(s->nrm[C]), cc));
    const double ndl = std::max(0.0, ddot3(n, negL));
    D3 alb;
    if (mode == ORC_LAMBERT_VERTEX_COLOR) {
      alb = dadd3(dadd3(dmul3(d3(s->col[A]), aa), dmul3(d3(s->col[B]), bb)), dmul3(d3(s->col[C]), cc));
    } else {
      const int m = s->tri_mesh[g];
      const int w = s->tex_w[m], h = s->tex_h[m];
      if (w * h == 0) return D3{0, 0, 0};
      double tx = s->tc[A].x * aa + s->tc[B].x * bb + s->tc[C].x * cc;
      double ty = s->tc[A].y * aa + s->tc[B].y * bb + s->tc[C].y * cc;
      tx = std::min(std::max(tx, 0.0), 1.0) * w;
      ty = std::min(std::max(ty, 0.0), 1.0) * h;
      const int ix = (int)tx, iy = (int)ty;
      const int x0 = std::min(std::max(ix, 0), w - 1), y0 = std::min(std::max(iy, 0), h - 1);
      const int x1 = std::min(std::max(ix + 1, 0), w - 1), y1 = std::min(std::max(iy + 1, 0), h - 1);
      const double fx = tx - x0, fy = ty - y0;
      auto T = [&](int xx, int yy) { return d3(tex_fetch(*s, m, xx, yy)); };
      alb = dadd3(dadd3(dmul3(T(x0, y0), (1 - fx) * (1 - fy)), dmul3(T(x0, y1), (1 - fx) * fy)),
                  dadd3(dmul3(T(x1, y0), fx * (1 - fy)), dmul3(T(x1, y1), fx * fy)));
    }
    const double sh = lit ? 1.0 : 0.0;
    return D3{alb.x * (sh * Lc.x * ndl + amb.x), alb.y * (sh * Lc.y * ndl + amb.y), alb.z * (sh * Lc.z * ndl + amb.z)};
  }
  // sum over pixels of adj . integral of the shade over triangle g's projection
  double tri(int g, const D3* V, bool lit) const {
    double w3[3][3] = {{V[0].x, V[0].y, V[0].z}, {V[1].x, V[1].y, V[1].z}, {V[2].x, V[2].y, V[2].z}};
    const P2 t[3] = {project_d(fc, w3[0]), project_d(fc, w3[1]), project_d(fc, w3[2])};
    const double xmin = std::min(t[0].x, std::min(t[1].x, t[2].x)), xmax = std::max(t[0].x, std::max(t[1].x, t[2].x));
    const double ymin = std::min(t[0].y, std::min(t[1].y, t[2].y)), ymax = std::max(t[0].y, std::max(t[1].y, t[2].y));
    const int x0 = std::max(0, (int)std::floor(xmin)), x1 = std::min(fc.w - 1, (int)std::floor(xmax));
    const int y0 = std::max(0, (int)std::floor(ymin)), y1 = std::min(fc.h - 1, (int)std::floor(ymax));
    double L = 0.0;
    const int ng = (int)gx.size();
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++) {
        const float* ap = adj + 3 * ((size_t)y * fc.w + x);
        if (ap[0] == 0.0f && ap[1] == 0.0f && ap[2] == 0.0f) continue;
        P2 poly[8];
        const int n = clip_poly(t, x, y, poly);
        if (n < 3) continue;
        D3 acc{0, 0, 0};
        for (int k = 1; k + 1 < n; k++) {
          const P2 a = poly[0], b = poly[k], c = poly[k + 1];
          const double J = std::fabs((b.x - a.x) * (c.y - b.y) - (b.y - a.y) * (c.x - b.x));
          if (J == 0.0) continue;
          // Duffy: (s, r) in [0, 1]^2 -> a + s (b - a) + s r (c - b), area element s J
          for (int i = 0; i < ng; i++)
            for (int j = 0; j < ng; j++) {
              const double ss = gx[i], rr = gx[j];
              const double qx = a.x + ss * (b.x - a.x) + ss * rr * (c.x - b.x);
              const double qy = a.y + ss * (b.y - a.y) + ss * rr * (c.y - b.y);
              const double wgt = gw[i] * gw[j] * ss * J;
              acc = dadd3(acc, dmul3(shade(g, V, lit, qx, qy), wgt));
            }
        }
        L += acc.x * ap[0] + acc.y * ap[1] + acc.z * ap[2];
      }
    return L;
  }
};

void lambert_fd_setup(LambertFd& F, const Scene& s, int mode, const orc_lambert* lam, const orc_camera* cam,
                      const float* adj, int order) {
  F.s = &s;
  F.mode = mode;
  const Lam L = make_lam(lam);
  F.negL = D3{-(double)L.L.x, -(double)L.L.y, -(double)L.L.z};
  F.Lc = d3(L.Lc);
  F.amb = d3(L.amb);
  const M44 wvp = mm(load44(cam->proj), load44(cam->world_view));
  for (int i = 0; i < 16; i++) F.fc.m[i] = wvp.m[i];
  F.fc.W = cam->width; F.fc.H = cam->height; F.fc.w = (int)cam->width; F.fc.h = (int)cam->height;
  double pr[16], wv[16];
  for (int i = 0; i < 16; i++) { pr[i] = cam->proj[i]; wv[i] = cam->world_view[i]; }
  inv44d(pr, F.proj_inv);
  inv44d(wv, F.wv_inv);
  F.eye = D3{F.wv_inv[12], F.wv_inv[13], F.wv_inv[14]};
  F.adj = adj;
  gauss_legendre(order, F.gx, F.gw);
}

// facing (front to the camera, lit by the light) of triangle g with world positions V
void facing(const LambertFd& F, const D3* V, bool& front, bool& lit) {
  const D3 n = dcross3(dsub3(V[1], V[0]), dsub3(V[2], V[0]));
  front = ddot3(n, dsub3(V[0], F.eye)) < 0.0;
  lit = ddot3(n, F.negL) > 0.0;
}
}  // namespace

extern "C" {
// The functional itself, every front-facing triangle (for checks on small scenes).
double orc_lambert_functional(void* sp, int mode, const orc_lambert* lam, const orc_camera* cam, const float* adjoint,
                              int order) {
  Scene& s = *(Scene*)sp;
  LambertFd F;
  lambert_fd_setup(F, s, mode, lam, cam, adjoint, order);
  double L = 0.0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : L)
  for (int g = 0; g < s.F; g++) {
    D3 V[3];
    for (int k = 0; k < 3; k++) V[k] = d3(s.pos[s.idx[3 * g + k]]);
    bool front, lit;
    facing(F, V, front, lit);
    if (front) L += F.tri(g, V, lit);
  }
  return L;
}

// Central differences (L(x + h) - L(x - h)) / 2h of the Lambert functional for `n` coordinates
// (3 * global vertex + axis, world space); only the triangles around the vertex change, each
// keeping its facings at x.
void orc_fd_lambert(void* sp, int mode, const orc_lambert* lam, const orc_camera* cam, const float* adjoint, int n,
                    const int* coords, double h, int order, double* out) {
  Scene& s = *(Scene*)sp;
  LambertFd F;
  lambert_fd_setup(F, s, mode, lam, cam, adjoint, order);
  std::vector<std::vector<int>> ring(s.V);
  for (int g = 0; g < s.F; g++)
    for (int k = 0; k < 3; k++) ring[s.idx[3 * g + k]].push_back(g);
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < n; i++) {
    const int vtx = coords[i] / 3, ax = coords[i] % 3;
    double acc[2] = {0.0, 0.0};
    for (int g : ring[vtx]) {
      D3 V[3];
      for (int k = 0; k < 3; k++) V[k] = d3(s.pos[s.idx[3 * g + k]]);
      bool front, lit;
      facing(F, V, front, lit);
      if (!front) continue;
      for (int sgn = 0; sgn < 2; sgn++) {
        D3 W[3] = {V[0], V[1], V[2]};
        for (int k = 0; k < 3; k++)
          if ((int)s.idx[3 * g + k] == vtx) {
            double* c = ax == 0 ? &W[k].x : ax == 1 ? &W[k].y : &W[k].z;
            *c += sgn == 0 ? h : -h;
          }
        acc[sgn] += F.tri(g, W, lit);
      }
    }
    out[i] = (acc[0] - acc[1]) / (2.0 * h);
  }
}
}  // extern "C"


