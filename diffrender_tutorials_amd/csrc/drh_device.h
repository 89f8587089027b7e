// drh_device.h -- device math, ray generation, LBVH / brute-force closest hit and the
// shading models of the render / d_render path, for gfx950.
//
// Bit-exactness: every translation unit is built with -ffp-contract=off and HIP's
// default correctly rounded f32 divide and sqrt, and each expression keeps the
// reference's evaluation order (LiteMath, common_src/LiteMath.h), so a ray, its
// Moller-Trumbore test and the shaded colour are the same IEEE operations the
// reference's brute-force tracer performs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace drh {

constexpr uint32_t NO_HIT = 0xFFFFFFFFu;

struct F2 { float x, y; };
struct F3 { float x, y, z; };

__device__ __forceinline__ F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 mul(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ F3 mulv(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 divs(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// a.x*b.x + a.y*b.y + a.z*b.z evaluated left to right (LiteMath.h:792)
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(F3 a) { return sqrtf(dot(a, a)); }
// a * (1/|a|) (LiteMath.h:795)
__device__ __forceinline__ F3 normalize(F3 a) { const float li = 1.0f / length(a); return mul(a, li); }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// LiteMath sign (LiteMath.h:109-114)
__device__ __forceinline__ float sgn(float x) { return x == 0.0f ? 0.0f : (x < 0.0f ? -1.0f : 1.0f); }
// std::max / std::min (NaN in the first argument propagates differently from fmaxf)
__device__ __forceinline__ float max_std(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float min_std(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ int clampi(int u, int a, int b) { int t = (a < u) ? u : a; return (b < t) ? b : t; }
__device__ __forceinline__ float clampf(float u, float a, float b) { return min_std(max_std(a, u), b); }

__device__ __forceinline__ F3 ld3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ F3 f4xyz(float4 v) { return {v.x, v.y, v.z}; }

// float4x4 (column-major memory) times (x, y, z, w) with the reference's term order
// (mat4_colmajor_mul_vec4, LiteMath.h:1342-1348)
__device__ __forceinline__ void mv4(const float* B, float x, float y, float z, float w, float r[4]) {
  r[0] = x * B[0] + y * B[4] + z * B[8] + w * B[12];
  r[1] = x * B[1] + y * B[5] + z * B[9] + w * B[13];
  r[2] = x * B[2] + y * B[6] + z * B[10] + w * B[14];
  r[3] = x * B[3] + y * B[7] + z * B[11] + w * B[15];
}
// The same product with B read through the constant address space: a matrix the whole wave
// shares (a camera's) then arrives by scalar loads, one copy per wave instead of one per lane.
__device__ __forceinline__ float cload(const float* p) { return *(const __attribute__((address_space(4))) float*)p; }
__device__ __forceinline__ void mv4c(const float* B, float x, float y, float z, float w, float r[4]) {
  r[0] = x * cload(B) + y * cload(B + 4) + z * cload(B + 8) + w * cload(B + 12);
  r[1] = x * cload(B + 1) + y * cload(B + 5) + z * cload(B + 9) + w * cload(B + 13);
  r[2] = x * cload(B + 2) + y * cload(B + 6) + z * cload(B + 10) + w * cload(B + 14);
  r[3] = x * cload(B + 3) + y * cload(B + 7) + z * cload(B + 11) + w * cload(B + 15);
}
__device__ __forceinline__ F3 mp(const float* B, F3 v) { float r[4]; mv4(B, v.x, v.y, v.z, 1.0f, r); return {r[0], r[1], r[2]}; }

// Per-view camera constants (host computes inverse4x4 and mProj*mWorldView exactly as
// BruteForce3D::SetCamera and CamInfo::commit do, raytrace3d_bf.cpp:93-100, camera.h:39-43).
struct DevCam {
  float proj_inv[16];
  float wv_inv[16];
  float wvp[16];
  float W, H;          // CamInfo::width / height
  int   w, h;          // integer image size
  long long img_off;   // float offset of this view's image in a views x H x W x 3 buffer
};

// CastSingleRay: EyeRayDirNormalized + transform_ray3f (raytrace.h:49-76, raytrace3d_bf.cpp:158-174)
// q / s.  When s is a power of two (2^-126 < 1/s < 2^127) its reciprocal is exact and q * (1/s)
// is the same correctly rounded quotient without the divide sequence.  The callers' s is the
// same in every lane (an image size, a sample count), so this is a branch, not a select.
__device__ __forceinline__ float div_pow2(float q, float s) {
  const uint32_t b = __float_as_uint(s), e = b >> 23;
  if ((b & 0x007FFFFFu) == 0u && e >= 1u && e <= 253u) return q * __uint_as_float((254u - e) << 23);
  return q / s;
}
__device__ __forceinline__ F3 divs_pow2(F3 a, float s) { return {div_pow2(a.x, s), div_pow2(a.y, s), div_pow2(a.z, s)}; }

__device__ __forceinline__ void eye_ray(const DevCam& c, float x, float y, F3& o, F3& d) {
  const float xn = div_pow2(x, cload(&c.W)), yn = div_pow2(y, cload(&c.H));
  float p[4];
  mv4c(c.proj_inv, 2.0f * xn - 1.0f, -2.0f * yn + 1.0f, 0.0f, 1.0f, p);
  const float w = p[3];
  const F3 dir = normalize(F3{p[0] / w, p[1] / w, p[2] / w});
  float a[4], b[4];
  mv4c(c.wv_inv, 0.0f, 0.0f, 0.0f, 1.0f, a);
  mv4c(c.wv_inv, 0.0f + 100.0f * dir.x, 0.0f + 100.0f * dir.y, 0.0f + 100.0f * dir.z, 1.0f, b);
  o = {a[0], a[1], a[2]};
  d = normalize(F3{b[0] - a[0], b[1] - a[1], b[2] - a[2]});
}

// Pixel-sample positions.  The reference puts sample (dx, dy) of a pixel at the centre of its
// cell of the sqrt(spp) x sqrt(spp) grid, ((dx + 0.5)/sq, (dy + 0.5)/sq) (src/drender.h:124-128).
// With jitter on, the sample moves to a uniform point of that cell, ((dx + ux)/sq, (dy + uy)/sq)
// with ux, uy in [0, 1) multiples of 2^-24 drawn from an integer hash of (seed + step, view,
// x, y, sample): stratified jitter, a fresh pattern per seed and per optimiser step (step is
// read from the device when the caller passes a counter).  Off, ux = uy = 0.5 exactly.
// The hash is Wellons' "lowbias32" (hash-prospector, 2020) chained over the keys.
struct Jitter { uint32_t seed; int on; const uint32_t* step; };
__host__ __device__ __forceinline__ uint32_t jitter_hash(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}
__device__ __forceinline__ void sample_offset(const Jitter& j, int view, int x, int y, int si, int sq,
                                              float& xo, float& yo) {
  float ux = 0.5f, uy = 0.5f;
  if (j.on) {
    const uint32_t base = j.seed + (j.step ? *j.step : 0u);
    uint32_t k = jitter_hash(base ^ jitter_hash((uint32_t)view + 0x9e3779b9U));
    k = jitter_hash(k ^ (uint32_t)x);
    k = jitter_hash(k ^ (uint32_t)y);
    k = jitter_hash(k ^ (uint32_t)si);
    ux = (float)(k >> 8) * (1.0f / 16777216.0f);
    uy = (float)(jitter_hash(k) >> 8) * (1.0f / 16777216.0f);
  }
  xo = div_pow2((si % sq) + ux, float(sq));
  yo = div_pow2((si / sq) + uy, float(sq));
}

// ---------------------------------------------------------------------------
// Scene data on the device
// ---------------------------------------------------------------------------
// Triangles in LBVH leaf order, SoA: A (w = original global triangle id as bits),
// e1 = B - A, e2 = C - A -- the same subtractions BruteForce3D::GetNearestHit does
// (raytrace3d_bf.cpp:123-129), done once per commit.
// LBVH node: both child boxes in the parent (n0..n2) and child links in n3.x / n3.y
// (>= 0 internal node, < 0 leaf: ~slot in the triangle arrays).
struct DevScene {
  const float4* __restrict__ tri_a;
  const float4* __restrict__ tri_e1;
  const float4* __restrict__ tri_e2;
  const float4* __restrict__ nodes;   // 4 float4 per internal node
  int F;                              // triangles
  int V;                              // vertices
  const uint32_t* __restrict__ idx;   // 3F global vertex indices (get_index)
  const int* __restrict__ tri_mesh;   // mesh of each global triangle
  const int* __restrict__ mesh_vo;    // vertex offset per mesh (n_meshes+1)
  const int* __restrict__ mesh_to;    // triangle offset per mesh (n_meshes+1)
  const float* __restrict__ pos;      // 3V prepared positions (transform applied)
  const float* __restrict__ pos_orig; // 3V mesh-local positions
  const float* __restrict__ col;      // 3V
  const float* __restrict__ tc;       // 2V
  const float* __restrict__ nrm;      // 3V normals as the 3D scene prepares them
  const float* __restrict__ nrm2d;    // 3V normals as the projected (edge) scene prepares them
  const float* __restrict__ texels;
  const int* __restrict__ tex_w;
  const int* __restrict__ tex_h;
  const int* __restrict__ tex_off;
  int n_meshes;
  // 1: every lane walks the LBVH on its own (trace_bvh_lane) -- for incoherent rays, where a
  // wave's rays open many more nodes together than each of them alone; 0: the wave walks it
  // as one packet (trace_bvh).  Chosen by the host per launch; the launchers pick the
  // kernels' TF_LANE instantiation by it (the walk is a compile-time choice on the device, so
  // the packet walk's code is what it is without the other).
  int lane_walk;
  // Device-side fault word (bit 0: a traversal stack overflowed and a subtree was skipped);
  // read back by the host at its next synchronisation and reported as DRH_E_DEVICE.
  int* fault;
};
constexpr int FAULT_STACK_OVERFLOW = 1;
// Trace flags, a template parameter of trace / shade and the kernels that cast rays:
// TF_WT the watertight triangle test, TF_LANE the per-lane LBVH walk.
constexpr int TF_WT = 1, TF_LANE = 2;

struct Hit { uint32_t gtri; float t, u, v; };

// Moller-Trumbore exactly as BruteForce3D::GetNearestHit (raytrace3d_bf.cpp:123-150).
__device__ __forceinline__ bool mt_test(float4 A4, float4 E14, float4 E24, F3 o, F3 d, float tNear,
                                        float& t, float& u, float& v) {
  const F3 A = f4xyz(A4), e1 = f4xyz(E14), e2 = f4xyz(E24);
  const F3 pvec = cross(d, e2);
  const F3 tvec = sub(o, A);
  const F3 qvec = cross(tvec, e1);
  const float det = dot(e1, pvec);
  const float sv = sgn(det);
  const float nv = dot(tvec, pvec), nu = dot(qvec, d), nt = dot(e2, qvec);
  // invDet has det's sign and is 0 only when det is, so v > 0, u > 0 and t > tNear >= 0 need
  // numerators of det's sign: a ray that fails there fails the test below too, and skips the
  // divide (t, u, v are then left unset, as for any rejected ray)
  if (!(nv * sv > 0.0f && nu * sv > 0.0f && (tNear < 0.0f || nt * sv > 0.0f))) return false;
  const float invDet = sv / max_std(sv * det, 1e-6f);
  v = nv * invDet;
  u = nu * invDet;
  t = nt * invDet;
  return v > 0.0f && u > 0.0f && (u + v < 1.0f) && t > tNear;
}

// Watertight ray/triangle test (S. Woop, C. Benthin, I. Wald, "Watertight Ray/Triangle
// Intersection", JCGT 2(1), 2013), the opt-in test of the edge pass's side rays: the ray's
// dominant axis becomes z, the triangle is sheared into the ray's frame and the signs of the
// 2D edge functions U, V, W decide (recomputed in double when one is exactly 0).  Unlike
// Moller-Trumbore's divide by a tiny determinant, a nearly edge-on triangle cannot accept a
// ray that misses it by more than the edge functions' rounding, and no ray slips between two
// triangles that share an edge.  Barycentrics as the reference's: u = weight of C, v of B.
struct WtRay { int kx, ky, kz; float Sx, Sy, Sz; };
__device__ __forceinline__ float comp(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ WtRay wt_ray(F3 d) {
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  WtRay r;
  r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  r.kx = r.kz == 2 ? 0 : r.kz + 1;
  r.ky = r.kx == 2 ? 0 : r.kx + 1;
  const float dz = comp(d, r.kz);
  if (dz < 0.0f) { const int t = r.kx; r.kx = r.ky; r.ky = t; }   // keep the winding
  r.Sx = comp(d, r.kx) / dz;
  r.Sy = comp(d, r.ky) / dz;
  r.Sz = 1.0f / dz;
  return r;
}
__device__ __forceinline__ bool wt_test(const WtRay& r, F3 o, F3 A, F3 B, F3 C, float tNear, float tFar,
                                        float& t, float& u, float& v) {
  const F3 a = sub(A, o), b = sub(B, o), c = sub(C, o);
  const float az = comp(a, r.kz), bz = comp(b, r.kz), cz = comp(c, r.kz);
  const float Ax = comp(a, r.kx) - r.Sx * az, Ay = comp(a, r.ky) - r.Sy * az;
  const float Bx = comp(b, r.kx) - r.Sx * bz, By = comp(b, r.ky) - r.Sy * bz;
  const float Cx = comp(c, r.kx) - r.Sx * cz, Cy = comp(c, r.ky) - r.Sy * cz;
  float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float T = U * (r.Sz * az) + V * (r.Sz * bz) + W * (r.Sz * cz);
  const float tn = tNear * det, tf = tFar * det;
  if (det > 0.0f ? (T <= tn || T >= tf) : (T >= tn || T <= tf)) return false;   // t in (tNear, tFar)
  const float rd = 1.0f / det;
  t = T * rd;
  v = V * rd;
  u = W * rd;
  return true;
}
__device__ __forceinline__ F3 vertex_of(const DevScene& s, uint32_t g, int k) {
  const uint32_t i = s.idx[3 * g + k];
  return F3{s.pos[3 * i], s.pos[3 * i + 1], s.pos[3 * i + 2]};
}

// Read-only scene data through the constant address space: a wave-uniform index then
// compiles to a scalar load into SGPRs, shared by the 64 rays of the wavefront.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) v4f cv4f;
typedef const __attribute__((address_space(4))) int cint;
__device__ __forceinline__ float4 uload(const float4* p, int i) {
  const v4f v = ((cv4f*)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int uload(const int* p) { return *(cint*)p; }
// The same at a 32-bit byte offset (element index times 16, no wider than 32 bits): the
// address is then the base register pair plus an offset register, one scalar shift instead of
// a sign-extended 64-bit multiply-add per load.
__device__ __forceinline__ float4 uload_at(const float4* p, uint32_t byte_off) {
  const v4f v = *(cv4f*)((const __attribute__((address_space(4))) char*)p + byte_off);
  return make_float4(v.x, v.y, v.z, v.w);
}
// Per-lane load at a 32-bit byte offset: base in scalar registers plus a 32-bit lane offset.
__device__ __forceinline__ float4 vload_at(const float4* p, uint32_t byte_off) {
  return *(const float4*)((const char*)p + byte_off);
}

// 1/d for the slab test; a zero (or denormal) component becomes +-1e30 so the fused form
// below stays finite: that axis' slab is then all-pass or all-fail, as for an infinite 1/d.
__device__ __forceinline__ float slab_rcp(float d) { return fabsf(d) < 1e-30f ? copysignf(1e30f, d) : 1.0f / d; }

// Slab test of a padded child box, fused form t = fma(plane, 1/d, -o/d).  Its rounding differs
// from (plane - o) * (1/d) by about 2^-24 (|o| + |plane - o|) in world units, which the boxes'
// padding (1e-5 relative plus an absolute term of the scene's scale) absorbs; the triangle
// test itself stays exact, so the padding only decides how many boxes are opened.  (Picking
// the entry / exit planes by a wave-uniform direction octant saves the min / max pairs on the
// vector unit but costs more on the scalar unit, which issues at the same rate.)
__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz,
                                     F3 id, F3 nod, float tmax, float& tn) {
  const float t0x = fmaf(lx, id.x, nod.x), t1x = fmaf(hx, id.x, nod.x);
  const float t0y = fmaf(ly, id.y, nod.y), t1y = fmaf(hy, id.y, nod.y);
  const float t0z = fmaf(lz, id.z, nod.z), t1z = fmaf(hz, id.z, nod.z);
  const float tmin = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), 0.0f));
  const float tmx = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
  tn = tmin;
  return tmin <= tmx;
}

// LBVH walk, one ray per lane: each lane keeps its own stack (private memory, 64 entries for
// the same depth bound as below) and opens only the boxes its own ray passes, nearer child
// first; a popped subtree whose entry distance is beyond the lane's closest hit so far is
// skipped.  Same hits as the wavefront walk below: the (t, triangle id) rule picks the same
// triangle from any superset of the triangles the lane's own boxes open.
template <bool ANY, bool WT>
__device__ __forceinline__ Hit trace_bvh_lane(const DevScene& s, F3 o, F3 d, float tNear, float tFar, bool retired) {
  Hit h{NO_HIT, tFar, 0.0f, 0.0f};
  if (retired) return h;
  const F3 id = {slab_rcp(d.x), slab_rcp(d.y), slab_rcp(d.z)};
  const F3 nod = {-(o.x * id.x), -(o.y * id.y), -(o.z * id.z)};
  const WtRay wr = WT ? wt_ray(d) : WtRay{};
  float lim = tFar * 1.00002f + 1e-6f;
  auto leaf = [&](int slot) -> bool {   // true: an any-hit ray is done
    const uint32_t off = (uint32_t)slot << 4;
    const float4 A = vload_at(s.tri_a, off);
    float t, u, v;
    bool hit;
    if (WT) {
      const uint32_t gg = __float_as_uint(A.w);
      hit = wt_test(wr, o, vertex_of(s, gg, 0), vertex_of(s, gg, 1), vertex_of(s, gg, 2), tNear, tFar, t, u, v);
    } else {
      hit = mt_test(A, vload_at(s.tri_e1, off), vload_at(s.tri_e2, off), o, d, tNear, t, u, v);
    }
    if (!hit) return false;
    const uint32_t g = __float_as_uint(A.w);
    if (ANY) {
      if (t < tFar) { h.gtri = g; h.t = t; h.u = u; h.v = v; return true; }
    } else if (t < h.t || (t == h.t && h.gtri != NO_HIT && g < h.gtri)) {
      h.gtri = g; h.t = t; h.u = u; h.v = v;
      lim = t * 1.00002f + 1e-6f;
    }
    return false;
  };
  if (s.F == 1) { leaf(0); return h; }
  int stk[64];
  float stt[64];
  int sp = 0;
  int node = 0;
  for (int visits = 0; visits < s.F; visits++) {
    const uint32_t off = (uint32_t)node << 6;
    const float4 n0 = vload_at(s.nodes, off), n1 = vload_at(s.nodes, off + 16);
    const float4 n2 = vload_at(s.nodes, off + 32), n3 = vload_at(s.nodes, off + 48);
    const int cl = __float_as_int(n3.x), cr = __float_as_int(n3.y);
    float tl, tr;
    bool hl = slab(n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, id, nod, lim, tl);
    bool hr = slab(n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, id, nod, lim, tr);
    if (hl && cl < 0) { if (leaf(~cl)) break; hl = false; }
    if (hr && cr < 0) { if (leaf(~cr)) break; hr = false; }
    if (hl && hr) {
      const bool lfirst = tl <= tr;
      if (sp < 64) {
        stk[sp] = lfirst ? cr : cl;
        stt[sp] = lfirst ? tr : tl;
        sp++;
      } else if (s.fault) {
        atomicOr(s.fault, FAULT_STACK_OVERFLOW);
      }
      node = lfirst ? cl : cr;
    } else if (hl) {
      node = cl;
    } else if (hr) {
      node = cr;
    } else {
      // pop the next subtree still in front of the closest hit
      bool found = false;
      while (sp > 0) {
        sp--;
        if (stt[sp] <= lim) { node = stk[sp]; found = true; break; }
      }
      if (!found) break;
    }
  }
  return h;
}

// LBVH walk, wavefront-coherent: the active lanes of a wave traverse together.  The node
// index is wave-uniform, so a node's two child boxes and links arrive by one scalar load
// for all 64 rays; each lane tests its own ray (per-lane closest-hit culling through `lim`),
// ballots pick the children any lane needs (nearer child first by majority), and the
// traversal stack is wave-uniform too: 64 ints of LDS per wave (`wst`), written by every
// active lane with the same value (no exec-mask juggling for a single writer), read as a
// broadcast.  The stack holds at most one far child per level of the path to the current
// node; a Karras tree over unique keys (30-bit Morton code, then the 32-bit triangle index
// as tie-break) splits at a strictly deeper key bit on every level, so it is at most 63
// levels deep and 64 entries always suffice.  Should a push ever find the stack full, the
// far child is dropped and the scene's fault word records it (the host then fails the call
// with DRH_E_DEVICE rather than return a hit that may be wrong).
// A lane whose ray is NaN, or an any-hit lane that has its hit, gets
// lim = -1 and no box passes for it any more.
// A lane tests every leaf its own slab test opens and possibly more (leaves opened for
// other lanes): the closest hit with the (t, triangle id) tie rule is the same over any
// superset of the triangles a lane's own walk would test, so the result is the brute-force
// tracer's answer for every lane.
// WT: the watertight test on the triangles' shared vertex positions instead of Moller-Trumbore.
template <bool ANY, bool WT = false>
__device__ __forceinline__ Hit trace_bvh(const DevScene& s, F3 o, F3 d, float tNear, float tFar, int* wst) {
  if (s.F == 0) return Hit{NO_HIT, tFar, 0.0f, 0.0f};
  const bool retired = o.x != o.x || o.y != o.y || o.z != o.z || d.x != d.x || d.y != d.y || d.z != d.z;
  Hit h{NO_HIT, tFar, 0.0f, 0.0f};
  const F3 id = {slab_rcp(d.x), slab_rcp(d.y), slab_rcp(d.z)};
  const F3 nod = {-(o.x * id.x), -(o.y * id.y), -(o.z * id.z)};
  const WtRay wr = WT ? wt_ray(d) : WtRay{};
  float lim = retired ? -1.0f : tFar * 1.00002f + 1e-6f;
  bool live = !retired;   // any-hit: still looking
  auto leaf = [&](int slot, bool mine) {
    const uint32_t off = (uint32_t)slot << 4;
    const float4 A = uload_at(s.tri_a, off), E1 = uload_at(s.tri_e1, off), E2 = uload_at(s.tri_e2, off);
    float t, u, v;
    bool hit = false;
    if (mine) {
      if (WT) {
        const uint32_t gg = __float_as_uint(A.w);
        hit = wt_test(wr, o, vertex_of(s, gg, 0), vertex_of(s, gg, 1), vertex_of(s, gg, 2), tNear, tFar, t, u, v);
      } else {
        hit = mt_test(A, E1, E2, o, d, tNear, t, u, v);
      }
    }
    if (hit) {
      const uint32_t g = __float_as_uint(A.w);
      if (ANY) { if (t < tFar) { h.gtri = g; h.t = t; h.u = u; h.v = v; lim = -1.0f; live = false; } }
      else if (t < h.t || (t == h.t && h.gtri != NO_HIT && g < h.gtri)) {
        h.gtri = g; h.t = t; h.u = u; h.v = v;
        lim = t * 1.00002f + 1e-6f;
      }
    }
  };
  if (s.F == 1) { leaf(0, !retired); return h; }
  int sp = 0;
  int node = 0;
  // a tree over F leaves has F-1 internal nodes, each visited at most once per wave
  for (int visits = 0; visits < s.F; visits++) {
    node = __builtin_amdgcn_readfirstlane(node);   // wave-uniform by construction; keep it scalar
    const uint32_t off = (uint32_t)node << 6;
    const float4 n0 = uload_at(s.nodes, off), n1 = uload_at(s.nodes, off + 16);
    const float4 n2 = uload_at(s.nodes, off + 32), n3 = uload_at(s.nodes, off + 48);
    const int cl = __float_as_int(n3.x), cr = __float_as_int(n3.y);
    float tl, tr;
    // any-hit lanes that are done skip the box tests (exec-masked); closest-hit lanes are
    // culled by lim alone
    const bool hl = (!ANY || live) && slab(n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, id, nod, lim, tl);
    const bool hr = (!ANY || live) && slab(n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, id, nod, lim, tr);
    const uint64_t bl = __ballot(hl), br = __ballot(hr);
    bool goL = false, goR = false;
    if (bl) { if (cl < 0) leaf(~cl, hl); else goL = true; }
    if (br) { if (cr < 0) leaf(~cr, hr); else goR = true; }
    if (ANY && (bl | br) && !__ballot(lim >= 0.0f)) break;
    if (goL && goR) {
      const uint64_t both = bl & br;
      const uint64_t lf = __ballot(hl && hr && tl <= tr);
      const bool lfirst = both ? 2 * __popcll(lf) >= __popcll(both) : __popcll(bl) >= __popcll(br);
      const int nearc = lfirst ? cl : cr, farc = lfirst ? cr : cl;
      if (sp < 64) { wst[sp] = farc; sp++; }   // every active lane writes the same value
      else if (s.fault) atomicOr(s.fault, FAULT_STACK_OVERFLOW);
      node = nearc;
    } else if (goL) {
      node = cl;
    } else if (goR) {
      node = cr;
    } else {
      if (sp == 0) break;
      sp--;
      node = __builtin_amdgcn_readfirstlane(wst[sp]);
    }
  }
  return h;
}

// Brute force over every triangle in the reference's order (raytrace3d_bf.cpp:113-153).
template <bool ANY>
__device__ __forceinline__ Hit trace_bf(const DevScene& s, const float4* tri_a_orig, const float4* tri_e1_orig,
                                        const float4* tri_e2_orig, F3 o, F3 d, float tNear, float tFar) {
  Hit h{NO_HIT, tFar, 0.0f, 0.0f};
  if (o.x != o.x || o.y != o.y || o.z != o.z || d.x != d.x || d.y != d.y || d.z != d.z) return h;
  for (int g = 0; g < s.F; g++) {
    float t, u, v;
    if (mt_test(uload(tri_a_orig, g), uload(tri_e1_orig, g), uload(tri_e2_orig, g), o, d, tNear, t, u, v) && t < h.t) {
      h.gtri = (uint32_t)g; h.t = t; h.u = u; h.v = v;
      if (ANY && t < tFar) return h;
    }
  }
  return h;
}

// Tracer selector: BVH or BF.  BF reads triangles in original order (bf_* arrays).
//
// Clamped determinants: BF's invDet = sign(det)/max(|det|, 1e-6) scales u, v, t by
// |det|/1e-6 when |det| < 1e-6, so BF can accept a triangle a nearly parallel ray never
// touches (reported t shrinks by the same factor).  Writing d = a*e1 + b*e2 + c*(o-A),
// acceptance needs |d.(e2 x (o-A))|, |d.((o-A) x e1)|, |d.N| < 1e-6, i.e. |a|, |b|, |c| <
// 1e-6/|V| with V = (o-A).N, hence |V| < 1e-6 (|e1| + |e2| + |o-A|).  Camera rays share o,
// so only the triangles meeting that bound for the view's origin can be accepted this way;
// shadow rays share d = -L, so only triangles with |L.N| < 1e-6 can.  Those "suspect"
// triangles are found per view / per light (k_suspects) and tested by brute force after the
// LBVH traversal, which makes the LBVH answer the BF answer.
// View entries are one float4 (N.xyz, triangle id); rays with |d.N| >= 1.5e-6 skip an entry
// (their det is not clamped).  A shadow ray's determinant depends on the triangle alone, so
// k_suspects computes it as mt_test would and lists a triangle only when 0 < |det| < 1e-6,
// in two float4: (A.xyz, id) and (pvec.xyz, invDet) of that test; a ray then skips an entry
// unless its v = (o - A).pvec * invDet -- mt_test's own v -- lies in (0, 1), which
// acceptance needs (v > 0, u > 0 and u + v < 1).
// Both lists are also binned (k_sus_bin, k_light_bin): per cell, the records of the suspects
// that some ray of the cell could be accepted by, in a pool (cell[c] = (offset, n), two float4
// per record; n < 0: the cell overflowed, its rays use the whole list).  View lists use screen
// cells of SUS_CELL_PX pixels and records (N.xyz, id), (qvec.xyz, ub): qvec = (o - A) x e1 is
// mt_test's own qvec (every camera ray of a view starts at the same o) and acceptance needs
// |qvec.d| <= ub.  Light lists use a grid over the plane across the light, located by
// grid = (a1.xyz, u0), (a2.xyz, v0), (1/du, 1/dv, -, -), and the light records.  A ray outside
// its grid uses the whole list.
// Records are read by scalar loads (the index is wave-uniform) and carry what the per-lane
// filters need, so only a record some lane passes loads its triangle.
struct Suspects {
  const float4* list;   // view: (N.x, N.y, N.z, id bits); light: (A, id bits), (pvec, invDet); any order
  const int* count;
  const float4* pool;   // binned records, or NULL
  const int2* cell;     // (offset, n) per cell
  int cells_x, cells_y;
  const float4* grid;   // light lists: the grid's placement (3 float4)
};
constexpr int SUS_CELL_PX = 8;
constexpr int SUS_BIN_MIN_F = 32768;   // lists of smaller scenes are not binned (short lists cost less than binning)
// LDS per wave for the walks: the traversal stack.
constexpr int WST_INTS = 64;

struct Tracer {
  DevScene s;
  const float4* bf_a;
  const float4* bf_e1;
  const float4* bf_e2;
  int use_bf;
  Suspects cam;     // for closest-hit (camera) rays of the current view
  Suspects light;   // for any-hit (shadow) rays
  // Ray counters of the counting instantiations (COUNT = true) only: [0] closest-hit rays,
  // [1] any-hit rays traced by the kernel this tracer is passed to.
  unsigned long long* rays;
};

// Counting instantiations: add the wave's active lanes whose ray is cast (one atomic per wave).
template <bool COUNT>
__device__ __forceinline__ void count_rays(unsigned long long* c, bool cast) {
  if (!COUNT || !c) return;
  const uint64_t b = __ballot(cast);
  const int leader = __ffsll((long long)__ballot(1)) - 1;
  if ((int)(threadIdx.x & 63) == leader && b) atomicAdd(c, (unsigned long long)__popcll(b));
}

// Calls f(r0, r1) for the records [0, n) of `rec` (two float4 each, scalar loads: the index
// is wave-uniform); true as soon as f does (the lane is done; the others go on).
template <class Fn>
__device__ __forceinline__ bool walk_records(const float4* rec, int n, Fn f) {
  for (int k = 0; k < n; k++)
    if (f(uload(rec, 2 * k), uload(rec, 2 * k + 1))) return true;
  return false;
}
__device__ __forceinline__ int2 uload2(const int2* p, int i) {
  const int* q = (const int*)(p + i);
  return make_int2(uload(q), uload(q + 1));
}
// Walks the pooled records of every distinct cell of the active lanes (a lane testing another
// lane's cell as well only tests a superset); true when this lane is done.
template <class Fn>
__device__ __forceinline__ bool walk_cells(const Suspects& S, int cell, Fn f) {
  bool todo = true;
  for (;;) {
    const uint64_t p = __ballot(todo);
    if (!p) break;
    const int c = __builtin_amdgcn_readlane(cell, __ffsll((long long)p) - 1);
    const int2 cc = uload2(S.cell, c);
    if (walk_records(S.pool + 2ll * cc.x, cc.y, f)) return true;
    if (cell == c) todo = false;
  }
  return false;
}

// The view's suspects for camera rays (closest hit, or any hit for `covered`); (sx, sy) is
// the ray's screen position (its cell), negative when unknown.
template <bool ANY>
__device__ __forceinline__ void test_cam_suspects(const Tracer& T, F3 o, F3 d, float tNear, float tFar, Hit& h,
                                                  int* wst, float sx, float sy) {
  const Suspects& S = T.cam;
  if (!S.count) return;
  // one record: true when an any-hit lane has its hit
  auto test = [&](float4 e) -> bool {
    const int g = __builtin_amdgcn_readfirstlane(__float_as_int(e.w));   // the same in every lane
    float t, u, v;
    if (mt_test(uload(T.bf_a, g), uload(T.bf_e1, g), uload(T.bf_e2, g), o, d, tNear, t, u, v)) {
      if (ANY) { if (t < tFar) { h.gtri = (uint32_t)g; h.t = t; h.u = u; h.v = v; return true; } }
      else if (t < h.t || (t == h.t && h.gtri != NO_HIT && (uint32_t)g < h.gtri)) {
        h.gtri = (uint32_t)g; h.t = t; h.u = u; h.v = v;
      }
    }
    return false;
  };
  if (S.pool) {
    const float fx = floorf(sx * (1.0f / SUS_CELL_PX)), fy = floorf(sy * (1.0f / SUS_CELL_PX));
    int cell = -1;
    if (fx >= 0.0f && fy >= 0.0f && fx < (float)S.cells_x && fy < (float)S.cells_y) cell = (int)fy * S.cells_x + (int)fx;
    const int n = cell >= 0 ? S.cell[cell].y : -1;
    if (!__ballot(n < 0)) {
      walk_cells(S, cell, [&](float4 e, float4 q) -> bool {
        const float dn = fabsf(fmaf(d.x, e.x, fmaf(d.y, e.y, d.z * e.z)));
        if (!(dn < 1.5e-6f) || !(fabsf(dot(f4xyz(q), d)) <= q.w)) return false;
        return test(e);
      });
      return;
    }
  }
  const int n = uload(S.count);
  for (int k = 0; k < n; k++) {
    const float4 e = uload(S.list, k);
    const float dn = fabsf(fmaf(d.x, e.x, fmaf(d.y, e.y, d.z * e.z)));
    if (!(dn < 1.5e-6f)) continue;
    if (test(e)) return;
  }
}

// The light's suspects for shadow rays (any hit, d = -L exactly as k_suspects used it).
__device__ __forceinline__ void test_light_suspects(const Tracer& T, F3 o, F3 d, float tNear, float tFar, Hit& h,
                                                    int* wst) {
  const Suspects& S = T.light;
  if (!S.count) return;
  auto rec = [&](float4 r0, float4 r1) -> bool {
    const float v0 = dot(sub(o, f4xyz(r0)), f4xyz(r1)) * r1.w;
    if (!(v0 > 0.0f && v0 < 1.0f)) return false;
    const int g = __builtin_amdgcn_readfirstlane(__float_as_int(r0.w));
    float t, u, v;
    if (mt_test(uload(T.bf_a, g), uload(T.bf_e1, g), uload(T.bf_e2, g), o, d, tNear, t, u, v) && t < tFar) {
      h.gtri = (uint32_t)g; h.t = t; h.u = u; h.v = v;
      return true;
    }
    return false;
  };
  if (S.pool) {
    const float4 g0 = uload(S.grid, 0), g1 = uload(S.grid, 1), g2 = uload(S.grid, 2);
    const float fu = floorf((dot(o, f4xyz(g0)) - g0.w) * g2.x), fv = floorf((dot(o, f4xyz(g1)) - g1.w) * g2.y);
    int cell = -1;
    if (fu >= 0.0f && fv >= 0.0f && fu < (float)S.cells_x && fv < (float)S.cells_y) cell = (int)fv * S.cells_x + (int)fu;
    const int n = cell >= 0 ? S.cell[cell].y : -1;
    if (!__ballot(n < 0)) {
      walk_cells(S, cell, rec);
      return;
    }
  }
  walk_records(S.list, uload(S.count), rec);
}

// Brute force with the watertight test (the WT form of trace_bf, original triangle order).
template <bool ANY>
__device__ __forceinline__ Hit trace_bf_wt(const DevScene& s, F3 o, F3 d, float tNear, float tFar) {
  Hit h{NO_HIT, tFar, 0.0f, 0.0f};
  if (o.x != o.x || o.y != o.y || o.z != o.z || d.x != d.x || d.y != d.y || d.z != d.z) return h;
  const WtRay wr = wt_ray(d);
  for (int g = 0; g < s.F; g++) {
    float t, u, v;
    if (wt_test(wr, o, vertex_of(s, g, 0), vertex_of(s, g, 1), vertex_of(s, g, 2), tNear, tFar, t, u, v) && t < h.t) {
      h.gtri = (uint32_t)g; h.t = t; h.u = u; h.v = v;
      if (ANY) return h;
    }
  }
  return h;
}

// LBVH walk of trace flags TF (packet or per lane, MT or watertight test)
template <bool ANY, int TF>
__device__ __forceinline__ Hit walk(const DevScene& s, F3 o, F3 d, float tNear, float tFar, int* wst) {
  constexpr bool WT = (TF & TF_WT) != 0;
  if (TF & TF_LANE) {
    if (s.F == 0) return Hit{NO_HIT, tFar, 0.0f, 0.0f};
    const bool retired = o.x != o.x || o.y != o.y || o.z != o.z || d.x != d.x || d.y != d.y || d.z != d.z;
    return trace_bvh_lane<ANY, WT>(s, o, d, tNear, tFar, retired);
  }
  return trace_bvh<ANY, WT>(s, o, d, tNear, tFar, wst);
}

// (sx, sy): a camera ray's screen position (selects its suspect cell); negative if unknown.
template <bool ANY, bool COUNT = false, int TF = 0>
__device__ __forceinline__ Hit trace(const Tracer& T, F3 o, F3 d, float tNear, float tFar, int* wst,
                                     float sx = -1.0f, float sy = -1.0f) {
  count_rays<COUNT>(T.rays + (ANY ? 1 : 0), !(o.x != o.x || o.y != o.y || o.z != o.z || d.x != d.x || d.y != d.y || d.z != d.z));
  if (TF & TF_WT) {
    // no clamped determinants to emulate: no suspect lists
    if (T.use_bf) return trace_bf_wt<ANY>(T.s, o, d, tNear, tFar);
    return walk<ANY, TF>(T.s, o, d, tNear, tFar, wst);
  }
  if (T.use_bf) return trace_bf<ANY>(T.s, T.bf_a, T.bf_e1, T.bf_e2, o, d, tNear, tFar);
  Hit h = walk<ANY, TF>(T.s, o, d, tNear, tFar, wst);
  if (ANY && h.gtri != NO_HIT) return h;
  if (o.x != o.x || o.y != o.y || o.z != o.z || d.x != d.x || d.y != d.y || d.z != d.z) return h;
  // any-hit rays here are the shadow rays (direction -L); camera any-hit queries use covered()
  if (ANY) test_light_suspects(T, o, d, tNear, tFar, h, wst);
  else test_cam_suspects<false>(T, o, d, tNear, tFar, h, wst, sx, sy);
  return h;
}

// Does a camera ray hit anything?  Any-hit traversal with the view's suspect list: the same
// yes / no as the closest-hit query (both accept the first triangle whose MT test passes with
// 0 < t < tFar; the closest-hit query only keeps looking for a nearer one).
template <bool COUNT = false, int TF = 0>
__device__ __forceinline__ bool covered(const Tracer& T, F3 o, F3 d, int* wst, float sx = -1.0f, float sy = -1.0f) {
  count_rays<COUNT>(T.rays + 1, !(o.x != o.x || o.y != o.y || o.z != o.z || d.x != d.x || d.y != d.y || d.z != d.z));
  if (T.use_bf) return trace_bf<true>(T.s, T.bf_a, T.bf_e1, T.bf_e2, o, d, 0.0f, 1e9f).gtri != NO_HIT;
  Hit h = walk<true, TF & TF_LANE>(T.s, o, d, 0.0f, 1e9f, wst);
  if (h.gtri != NO_HIT) return true;
  if (o.x != o.x || o.y != o.y || o.z != o.z || d.x != d.x || d.y != d.y || d.z != d.z) return false;
  test_cam_suspects<true>(T, o, d, 0.0f, 1e9f, h, wst, sx, sy);
  return h.gtri != NO_HIT;
}

// Derivatives of the Moller-Trumbore barycentrics u (weight of C) and v (weight of B) of the
// ray (o, d) against triangle (A, A + e1, A + e2) with respect to A, B and C, closed form:
// with tvec = o - A, pvec = d x e2, det = e1.pvec, u = d.(tvec x e1) / det and
// v = tvec.pvec / det, so du/dB = (d x tvec - u pvec) / det, du/dC = -u (e1 x d) / det,
// dv/dB = -v pvec / det, dv/dC = (tvec x d - v (e1 x d)) / det, and the A terms are minus
// the sum of the tvec, B and C ones (what BarU_grad / BarV_grad compute, src/functions.cpp:
// 226-470).  False (and no output) where det is clamped (|det| < 1e-6).
__device__ __forceinline__ bool bary_grad(F3 o, F3 d, F3 A, F3 e1, F3 e2, float u, float v, F3 du[3], F3 dv[3]) {
  const F3 tvec = sub(o, A), pvec = cross(d, e2), e1xd = cross(e1, d);
  const float det = dot(e1, pvec);
  if (!(fabsf(det) >= 1e-6f)) return false;
  const float id = 1.0f / det;
  du[1] = mul(id, sub(cross(d, tvec), mul(u, pvec)));
  du[2] = mul(-u * id, e1xd);
  du[0] = mul(-1.0f, add(add(mul(id, e1xd), du[1]), du[2]));
  dv[1] = mul(-v * id, pvec);
  dv[2] = mul(id, sub(cross(tvec, d), mul(v, e1xd)));
  dv[0] = mul(-1.0f, add(add(mul(id, pvec), dv[1]), dv[2]));
  return true;
}

// ---------------------------------------------------------------------------
// Shading (src/shade_simple.cpp, src/shade_lambert.cpp, src/shade_common.cpp)
// ---------------------------------------------------------------------------
struct Light { F3 L, Lc, amb; };

__device__ __forceinline__ F3 tex_fetch(const DevScene& s, int m, int x, int y) {
  return ld3(s.texels + s.tex_off[m] + 3 * (y * s.tex_w[m] + x));
}
// sample_bilinear_clamp_3f (shade_common.cpp:3-21)
__device__ __forceinline__ F3 bilinear(const DevScene& s, int m, F2 tc) {
  const int w = s.tex_w[m], h = s.tex_h[m];
  if (w == 0 || h == 0) return {0, 0, 0};
  tc.x = clampf(tc.x, 0.0f, 1.0f); tc.y = clampf(tc.y, 0.0f, 1.0f);
  tc.x *= (float)w; tc.y *= (float)h;
  const int ix = (int)tc.x, iy = (int)tc.y;
  const int x0 = clampi(ix, 0, w - 1), y0 = clampi(iy, 0, h - 1);
  const int x1 = clampi(ix + 1, 0, w - 1), y1 = clampi(iy + 1, 0, h - 1);
  const float dx = tc.x - (float)x0, dy = tc.y - (float)y0;
  const F3 p00 = tex_fetch(s, m, x0, y0), p01 = tex_fetch(s, m, x0, y1);
  const F3 p10 = tex_fetch(s, m, x1, y0), p11 = tex_fetch(s, m, x1, y1);
  const float a = (1 - dx) * (1 - dy), b = (1 - dx) * dy, c = dx * (1 - dy), e = dx * dy;
  return {a * p00.x + b * p01.x + c * p10.x + e * p11.x,
          a * p00.y + b * p01.y + c * p10.y + e * p11.y,
          a * p00.z + b * p01.z + c * p10.z + e * p11.z};
}

// a*(1-u-v) + b*v + u*c (shade_simple.cpp:36)
__device__ __forceinline__ F3 blend3(F3 a, F3 b, F3 c, float u, float v) {
  const float w = 1.0f - u - v;
  return add(add(mul(a, w), mul(b, v)), mul(u, c));
}
__device__ __forceinline__ F2 blend2(const float* tc, uint32_t A, uint32_t B, uint32_t C, float u, float v) {
  const float w = 1.0f - u - v;
  return {tc[2 * A] * w + tc[2 * B] * v + u * tc[2 * C], tc[2 * A + 1] * w + tc[2 * B + 1] * v + u * tc[2 * C + 1]};
}

// Shading modes: the reference's SHADING_MODEL numbers (src/dmesh.h:16-23) and one of this
// build's: LAMBERT with the interpolated vertex colour as albedo instead of a texture (the
// multi-view fitting workload's shading; no counterpart in the reference).
constexpr int MODE_LAMBERT_VCOL = 8;

// LAMBERT's light terms at a closest hit (src/shade_lambert.cpp:26-34): ndl = max(0, n.(-L))
// with the interpolated, un-normalised normal n, and shadow = 1 unless the ray from
// pos + (t - 1e-5) dir toward -L hits anything.  The lit colour is (shadow * Lc) * ndl.  With
// ndl == +0 and a finite light colour the shadow ray cannot change that product (bit for bit),
// so it is only cast where ndl > 0 (shadow stays 1 otherwise).
template <bool COUNT = false, int TF = 0>
__device__ __forceinline__ void lambert_terms(const Tracer& T, const Light& L, F3 o, F3 d, const Hit& h,
                                              uint32_t A, uint32_t B, uint32_t C, const float* nrm, int* wst,
                                              float& shadow, float& ndl) {
  const F3 n = blend3(ld3(nrm + 3 * A), ld3(nrm + 3 * B), ld3(nrm + 3 * C), h.u, h.v);
  const float BIAS = 1e-5f;
  const F3 negL = mul(-1.0f, L.L);
  ndl = max_std(0.0f, dot(n, negL));   // +0 or positive
  shadow = 1.0f;
  if (ndl != 0.0f || !(fabsf(L.Lc.x) <= 3.4e38f && fabsf(L.Lc.y) <= 3.4e38f && fabsf(L.Lc.z) <= 3.4e38f)) {
    const F3 surf = add(o, mul(h.t - BIAS, d));
    const Hit sh = trace<true, COUNT, TF>(T, surf, negL, 0.0f, 1e9f, wst);
    shadow = sh.gtri == NO_HIT ? 1.0f : 0.0f;
  }
}
// (shadow * Lc) * ndl + ambient: the factor LAMBERT multiplies the albedo by
__device__ __forceinline__ F3 lambert_factor(const Light& L, float shadow, float ndl) {
  return add(mul(mul(shadow, L.Lc), ndl), L.amb);
}
// The pair (shadow, ndl) in one float, for the stored hits of MODE_LAMBERT_VCOL: ndl when the
// point is lit, -ndl in shadow.  ndl == +0 always comes with shadow = 1 (no ray cast), so the
// sign decides the shadow and the magnitude is ndl.
__device__ __forceinline__ float lit_encode(float shadow, float ndl) { return shadow != 0.0f ? ndl : -ndl; }
__device__ __forceinline__ void lit_decode(float w, float& shadow, float& ndl) {
  shadow = w < 0.0f ? 0.0f : 1.0f;
  ndl = fabsf(w);
}

// shade<m>(screen_pos) for SIL / VCOL / TEXTURE / LAMBERT / LAMBERT_VCOL.  `nrm` is the 3D
// scene's normals for render and the projected scene's for edge samples (src/drender.h:353-354).
// lit_out (LAMBERT_VCOL, may be NULL) receives the hit's light terms, lit_encode'd.
template <int MODE, bool COUNT = false, int TF = 0>
__device__ __forceinline__ F3 shade(const Tracer& T, const Light& L, const DevCam& c, float sx, float sy,
                                    const float* nrm, int* wst, Hit* hit_out = nullptr, float* lit_out = nullptr) {
  F3 o, d;
  eye_ray(c, sx, sy, o, d);
  const Hit h = trace<false, COUNT, TF>(T, o, d, 0.0f, 1e9f, wst, sx, sy);
  if (hit_out) *hit_out = h;
  if (h.gtri == NO_HIT) return {0, 0, 0};
  if (MODE == 1) return {1, 1, 1};
  const DevScene& s = T.s;
  const uint32_t A = s.idx[3 * h.gtri], B = s.idx[3 * h.gtri + 1], C = s.idx[3 * h.gtri + 2];
  if (MODE == 2) return blend3(ld3(s.col + 3 * A), ld3(s.col + 3 * B), ld3(s.col + 3 * C), h.u, h.v);
  F3 diffuse;
  if (MODE == MODE_LAMBERT_VCOL) {
    diffuse = blend3(ld3(s.col + 3 * A), ld3(s.col + 3 * B), ld3(s.col + 3 * C), h.u, h.v);
  } else {
    const int m = s.tri_mesh[h.gtri];
    const F2 tcv = blend2(s.tc, A, B, C, h.u, h.v);
    diffuse = bilinear(s, m, tcv);
    if (MODE == 3) return diffuse;
  }
  float shadow, ndl;
  lambert_terms<COUNT, TF>(T, L, o, d, h, A, B, C, nrm, wst, shadow, ndl);
  if (MODE == MODE_LAMBERT_VCOL && lit_out) *lit_out = lit_encode(shadow, ndl);
  return mulv(lambert_factor(L, shadow, ndl), diffuse);
}

// ---------------------------------------------------------------------------
// Projection and its adjoint (src/functions.h:5-13, src/functions.cpp:21-163)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void vertex_shader(const DevCam& c, float vx, float vy, float vz, float& sx, float& sy) {
  const float* M = c.wvp;
  const float W = vx * M[3] + vy * M[7] + vz * M[11] + M[15];
  const float xN = (vx * M[0] + vy * M[4] + vz * M[8] + M[12]) / W;
  const float yN = -(vx * M[1] + vy * M[5] + vz * M[9] + M[13]) / W;
  sx = (xN * 0.5f + 0.5f) * c.W;
  sy = (yN * 0.5f + 0.5f) * c.H;
}
// d(screen x)/dV (axis 0) or d(screen y)/dV (axis 1): the reverse-mode chain with the
// reference's term order.
__device__ __forceinline__ F3 vs_grad(const DevCam& c, F3 V, int axis) {
  const float* M = c.wvp;
  const float W = V.x * M[3] + V.y * M[7] + V.z * M[11] + M[15];
  const int r = axis;
  float num = V.x * M[r] + V.y * M[4 + r] + V.z * M[8 + r] + M[12 + r];
  if (axis == 1) num = -num;
  const float size = axis == 0 ? c.W : c.H;
  const float dN = (1 * size) * 0.5f;
  const float q = dN / W;
  const float dW = dN * -num / (W * W);
  F3 g;
  if (axis == 0) { g = {0 + q * M[0], 0 + q * M[4], 0 + q * M[8]}; }
  else { g = {0 + -q * M[1], 0 + -q * M[5], 0 + -q * M[9]}; }
  g.x += dW * M[3]; g.y += dW * M[7]; g.z += dW * M[11];
  return g;
}

}  // namespace drh
