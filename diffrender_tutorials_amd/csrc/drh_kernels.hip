// drh_kernels.hip -- HIP kernels of the render / d_render path for gfx950 (MI355X).
//
//   commit   k_prepare, k_tri_setup, k_morton, k_radix_{hist,scan,scatter},
//            k_karras, k_refit_local (+ k_refit_top above one chunk), k_pack_nodes (with the
//            leaf reorder), k_suspects, k_sus_bin / k_light_bin + k_bin_offsets
//                                                                     (replaces IRayTracer::Init)
//   render   k_render<MODE> (+ fused L2 loss / adjoint), k_loss_final
//   interior k_interior<MODE>, k_coverage, k_interior_pos              (interior_derivatives)
//   edges    k_project (with edge lengths), k_cdf (one wave, small edge sets) or
//            k_cdf_{sums,scan,spec,walk,final}, k_edge_bin, k_bin_scan, k_edge_scatter
//            (pixel-bin order), k_edge_trace<MODE> (side colours), k_edge_grad
//                                                                     (edge_derivatives)
//   optimise k_optsimple_{moments,apply}, k_adam, k_transform_update (k_step_inc only for a
//            scene without vertices; otherwise k_prepare advances the step counter)
//
// Wave64 everywhere.  Rays are traced a wavefront at a time (trace_bvh in drh_device.h): a
// pixel tile or 32 edge samples (both sides) per wave, one wave-uniform LBVH walk per packet.
#include <hip/hip_runtime.h>

#include <mutex>

#include "drh_device.h"
#include "drh_kernels.h"
#include "drh_mat4.h"

namespace drh {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t f2key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}
__device__ __forceinline__ float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void k_fill_f32(float* p, long long n, float v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_copy_f32(float4* __restrict__ d, const float4* __restrict__ s, long long n4) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) d[i] = s[i];
}
// One upload of many arrays: segment blockIdx.y of the staged block `src` (4-byte words at
// word offset off[y], words[y] of them) to its destination buffer.
__global__ void k_scatter(ScatterArgs a) {
  const int g = blockIdx.y;
  const uint32_t* __restrict__ src = a.src + a.off[g];
  uint32_t* __restrict__ dst = a.dst[g];
  const long long n = a.words[g];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
__global__ void k_fill_u32(uint32_t* p, long long n, uint32_t v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// commit: prepare_for_render (common_src/scene_utils.cpp:98-173)
// ---------------------------------------------------------------------------
// pos = T*v, nrm = transpose(inverse(T))*(n,1) for the 3D scene; the projected scene's
// normals are I*(T*(n,0), 1) (transform() then an identity prepare, src/drender.h:260-268).
__global__ void k_prepare(PrepareArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) {
    // state for the kernels that follow (one launch fewer each)
    if (a.bounds) {
      a.bounds[0] = a.bounds[1] = a.bounds[2] = 0xFFFFFFFFu;
      a.bounds[3] = a.bounds[4] = a.bounds[5] = 0u;
    }
    if (a.zero) *a.zero = 0u;
    if (a.step) *a.step += 1;
  }
  if (k >= a.V) return;
  int m = 0;
  while (m + 1 < a.n_meshes && a.mesh_vo[m + 1] <= k) m++;
  const float* T = a.mats + 48 * m;        // T, transpose(inverse(T)), identity
  const float* NT = T + 16;
  const float* I = T + 32;
  const F3 v = ld3(a.pos_orig + 3 * k);
  const F3 p = mp(T, v);
  a.pos[3 * k] = p.x; a.pos[3 * k + 1] = p.y; a.pos[3 * k + 2] = p.z;
  if (a.has_nrm) {
    const F3 n = ld3(a.nrm_in + 3 * k);
    const F3 n3 = mp(NT, n);
    a.nrm[3 * k] = n3.x; a.nrm[3 * k + 1] = n3.y; a.nrm[3 * k + 2] = n3.z;
    float r[4];
    mv4(T, n.x, n.y, n.z, 0.0f, r);
    const F3 n2 = mp(I, F3{r[0], r[1], r[2]});
    a.nrm2d[3 * k] = n2.x; a.nrm2d[3 * k + 1] = n2.y; a.nrm2d[3 * k + 2] = n2.z;
  }
}

// Per triangle: BF's transformed corners T*pos (raytrace3d_bf.cpp:123-125), its edges, AABB
// and centroid; scene centroid bounds by ordered-key atomics.
// A triangle of the light's suspect list (see k_suspects): clamped shadow-ray determinant.
__device__ __forceinline__ void light_suspect(const SuspectArgs& a, int g, float4 A, float4 E1, float4 E2) {
  const F3 negL = mul(-1.0f, F3{a.ox, a.oy, a.oz});
  const F3 pvec = cross(negL, f4xyz(E2));
  const float det = dot(f4xyz(E1), pvec);
  const float sv = sgn(det);
  if (!(sv != 0.0f && sv * det < 1e-6f)) return;
  const float invDet = sv / max_std(sv * det, 1e-6f);
  const int k = atomicAdd(a.count, 1);
  a.list[2 * k] = make_float4(A.x, A.y, A.z, __uint_as_float((uint32_t)g));
  a.list[2 * k + 1] = make_float4(pvec.x, pvec.y, pvec.z, invDet);
}
__global__ void k_tri_setup(TriSetupArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ uint32_t smin[3], smax[3];
  if (threadIdx.x < 3) { smin[threadIdx.x] = 0xFFFFFFFFu; smax[threadIdx.x] = 0u; }
  if (a.counter && g <= a.F) a.counter[g] = 0u;
  __syncthreads();
  uint32_t kmin[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, kmax[3] = {0u, 0u, 0u};
  if (g < a.F) {
    const int m = a.tri_mesh[g];
    const float* T = a.mats + 48 * m;
    const uint32_t ia = a.idx[3 * g], ib = a.idx[3 * g + 1], ic = a.idx[3 * g + 2];
    const F3 A = mp(T, ld3(a.pos + 3 * ia));
    const F3 B = mp(T, ld3(a.pos + 3 * ib));
    const F3 C = mp(T, ld3(a.pos + 3 * ic));
    const F3 e1 = sub(B, A), e2 = sub(C, A);
    const float4 A4 = make_float4(A.x, A.y, A.z, __uint_as_float((uint32_t)g));
    const float4 E14 = make_float4(e1.x, e1.y, e1.z, 0.0f), E24 = make_float4(e2.x, e2.y, e2.z, 0.0f);
    a.bf_a[g] = A4;
    a.bf_e1[g] = E14;
    a.bf_e2[g] = E24;
    if (a.light.list) light_suspect(a.light, g, A4, E14, E24);   // the light's suspect list, k_suspects' light form
    const float lx = fminf(A.x, fminf(B.x, C.x)), ly = fminf(A.y, fminf(B.y, C.y)), lz = fminf(A.z, fminf(B.z, C.z));
    const float hx = fmaxf(A.x, fmaxf(B.x, C.x)), hy = fmaxf(A.y, fmaxf(B.y, C.y)), hz = fmaxf(A.z, fmaxf(B.z, C.z));
    a.box_lo[g] = make_float4(lx, ly, lz, 0.0f);
    a.box_hi[g] = make_float4(hx, hy, hz, 0.0f);
    const float cx = (lx + hx) * 0.5f, cy = (ly + hy) * 0.5f, cz = (lz + hz) * 0.5f;
    a.cen[g] = make_float4(cx, cy, cz, 0.0f);
    if (cx == cx && cy == cy && cz == cz) {
      kmin[0] = f2key(cx); kmin[1] = f2key(cy); kmin[2] = f2key(cz);
      kmax[0] = kmin[0]; kmax[1] = kmin[1]; kmax[2] = kmin[2];
    }
  }
  // the centroid key bounds over the wave first (min / max are order-free), then one LDS
  // atomic per wave and key instead of one per lane on the same six words
#pragma unroll
  for (int k = 0; k < 3; k++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      kmin[k] = min(kmin[k], (uint32_t)__shfl_xor((int)kmin[k], off, 64));
      kmax[k] = max(kmax[k], (uint32_t)__shfl_xor((int)kmax[k], off, 64));
    }
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) { atomicMin(&smin[k], kmin[k]); atomicMax(&smax[k], kmax[k]); }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    atomicMin(&a.bounds[threadIdx.x], smin[threadIdx.x]);
    atomicMax(&a.bounds[3 + threadIdx.x], smax[threadIdx.x]);
  }
}

// 30-bit Morton code of the centroid in the scene's centroid box.
__device__ __forceinline__ uint32_t spread10(uint32_t x) {
  x &= 0x3FFu;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
__device__ __forceinline__ uint32_t quant10(float c, float lo, float hi) {
  const float ext = hi - lo;
  float q = ext > 0.0f ? (c - lo) / ext : 0.0f;
  q = q * 1024.0f;
  if (!(q > 0.0f)) q = 0.0f;   // also NaN
  if (q > 1023.0f) q = 1023.0f;
  return (uint32_t)q;
}
__global__ void k_morton(MortonArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.F) return;
  const float lx = key2f(a.bounds[0]), ly = key2f(a.bounds[1]), lz = key2f(a.bounds[2]);
  const float hx = key2f(a.bounds[3]), hy = key2f(a.bounds[4]), hz = key2f(a.bounds[5]);
  const float4 c = a.cen[g];
  const uint32_t code = (spread10(quant10(c.x, lx, hx)) << 2) | (spread10(quant10(c.y, ly, hy)) << 1) |
                        spread10(quant10(c.z, lz, hz));
  a.keys[g] = code;
  a.vals[g] = (uint32_t)g;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits, RADIX_TILE keys per block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* keys, int n, int shift, uint32_t* hist, int nblocks) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * RADIX_TILE;
  for (int i = threadIdx.x; i < RADIX_TILE; i += 256) {
    const int k = base + i;
    if (k < n) atomicAdd(&cnt[(keys[k] >> shift) & 0xFFu], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of hist[256*nblocks] (digit-major) in one block.
__global__ __launch_bounds__(1024) void k_radix_scan(uint32_t* hist, int total) {
  __shared__ uint32_t part[1024];
  const int per = (total + 1023) / 1024;
  const int b = threadIdx.x * per;
  uint32_t s = 0;
  for (int i = 0; i < per; i++) if (b + i < total) s += hist[b + i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
  for (int i = 0; i < per; i++)
    if (b + i < total) { const uint32_t h = hist[b + i]; hist[b + i] = run; run += h; }
}

// Stable scatter: the tile is taken in rounds of 256 keys in index order; the rank of a key
// among equal digits of its round comes from wave ballots over the 8 digit bits.
__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                       uint32_t* vout, int n, int shift, const uint32_t* offs,
                                                       int nblocks) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcnt[4][256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  base[threadIdx.x] = offs[threadIdx.x * nblocks + blockIdx.x];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < RADIX_TILE / 256; r++) {
    const int k = blockIdx.x * RADIX_TILE + r * 256 + threadIdx.x;
    const bool valid = k < n;
    uint32_t key = 0, val = 0, dig = 0;
    if (valid) { key = kin[k]; val = vin[k]; dig = (key >> shift) & 0xFFu; }
    for (int i = threadIdx.x; i < 4 * 256; i += 256) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < 8; b++) {
      const uint64_t bb = __ballot((dig >> b) & 1u);
      peers &= ((dig >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wcnt[wave][dig] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t before = 0;
      for (int w = 0; w < wave; w++) before += wcnt[w][dig];
      const uint32_t dst = base[dig] + before + rank;
      kout[dst] = key;
      vout[dst] = val;
    }
    __syncthreads();
    {  // advance the running base of each digit by this round's count
      const int d = threadIdx.x;
      base[d] += wcnt[0][d] + wcnt[1][d] + wcnt[2][d] + wcnt[3][d];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Karras (2012) hierarchy over sorted (key, index) and bottom-up refit
// ---------------------------------------------------------------------------
__device__ __forceinline__ int kdelta(const uint32_t* k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = k[i], b = k[j];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}
__global__ void k_karras(KarrasArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = a.F;
  if (i >= n - 1) return;
  const uint32_t* k = a.keys;
  const int d = (kdelta(k, n, i, i + 1) - kdelta(k, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = kdelta(k, n, i, i - d);
  int lmax = 2;
  while (kdelta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (kdelta(k, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = kdelta(k, n, i, j);
  int s = 0, t = l;
  while (true) {
    t = (t + 1) >> 1;
    if (kdelta(k, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  const int gamma = i + s * d + (d < 0 ? d : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const int left = (lo == gamma) ? ~gamma : gamma;
  const int right = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
  a.child[2 * i] = left;
  a.child[2 * i + 1] = right;
  a.range[i] = make_int2(lo, hi);
  if (left < 0) a.leaf_parent[~left] = i; else a.node_parent[left] = i;
  if (right < 0) a.leaf_parent[~right] = i; else a.node_parent[right] = i;
  if (i == 0) a.node_parent[0] = -1;
}

// Bottom-up refit in two passes without device-scope fences.  k_refit_local: one workgroup per
// chunk of REFIT_CHUNK sorted leaves climbs from its leaves with LDS arrival counters and LDS
// boxes while the parent's leaf range stays inside the chunk (the second arrival at a node
// merges both child boxes and goes on); a node whose parent spans chunks is handed to the top
// pass through `frontier`.  k_refit_top: one workgroup finishes the few spanning nodes the
// same way with global counters; its boxes move between its own waves through L2 (volatile
// loads), so workgroup-scope fences order them.
__device__ __forceinline__ void box_union(float4& L, float4& H, float4 l, float4 h) {
  L = make_float4(fminf(L.x, l.x), fminf(L.y, l.y), fminf(L.z, l.z), 0.0f);
  H = make_float4(fmaxf(H.x, h.x), fmaxf(H.y, h.y), fmaxf(H.z, h.z), 0.0f);
}
__global__ __launch_bounds__(REFIT_CHUNK) void k_refit_local(RefitArgs a) {
  __shared__ int cnt[REFIT_CHUNK];
  __shared__ float lbox[REFIT_CHUNK][6];   // leaf boxes of the chunk's slots
  __shared__ float nbox[REFIT_CHUNK][6];   // boxes of the internal nodes numbered in the chunk
  __shared__ int2 lrange[REFIT_CHUNK], lchild[REFIT_CHUNK];
  __shared__ int lpar[REFIT_CHUNK];
  const int base = blockIdx.x * REFIT_CHUNK, end = min(base + REFIT_CHUNK, a.F);
  const int k = threadIdx.x, sl = base + k;
  // stage the chunk's leaves and the internal nodes numbered base..end-1 (coalesced)
  cnt[k] = 0;
  float4 L = {0, 0, 0, 0}, H = {0, 0, 0, 0};
  int p = -1;
  if (sl < end) {
    const int g = (int)a.order[sl];
    L = a.box_lo[g];
    H = a.box_hi[g];
    a.leaf_lo[sl] = L;
    a.leaf_hi[sl] = H;
    lbox[k][0] = L.x; lbox[k][1] = L.y; lbox[k][2] = L.z; lbox[k][3] = H.x; lbox[k][4] = H.y; lbox[k][5] = H.z;
    p = a.leaf_parent[sl];
  }
  if (sl < end && sl < a.F - 1) {
    lrange[k] = a.range[sl];
    lchild[k] = make_int2(a.child[2 * sl], a.child[2 * sl + 1]);
    lpar[k] = a.node_parent[sl];
  }
  __syncthreads();
  if (sl >= end) return;
  int q = ~sl;
  // an internal node's number lies in its own leaf range, so a parent numbered outside the
  // chunk spans chunks; one numbered inside spans iff its staged range leaves the chunk
  for (int level = 0; p >= 0 && level < 64; level++) {
    const int pk = p - base;
    if (p < base || p >= end || lrange[pk].x < base || lrange[pk].y >= end) {
      a.frontier[atomicAdd(a.counter + a.F, 1u)] = q;
      return;
    }
    __threadfence_block();
    if (atomicAdd(&cnt[pk], 1) == 0) return;  // the other child is still on its way
    __threadfence_block();
    const int2 c = lchild[pk];
    const int sib = c.x == q ? c.y : c.x;
    const float* sb = sib < 0 ? lbox[~sib - base] : nbox[sib - base];
    box_union(L, H, make_float4(sb[0], sb[1], sb[2], 0.0f), make_float4(sb[3], sb[4], sb[5], 0.0f));
    nbox[pk][0] = L.x; nbox[pk][1] = L.y; nbox[pk][2] = L.z; nbox[pk][3] = H.x; nbox[pk][4] = H.y; nbox[pk][5] = H.z;
    a.node_lo[p] = L;
    a.node_hi[p] = H;
    q = p;
    p = lpar[pk];
  }
}
// The nodes whose leaf range crosses a refit chunk: listed once per full build (the topology
// stays until the next one), with each one's parent and children as list indices, so that
// k_refit_top can climb them in LDS.
__device__ __forceinline__ bool spans_chunks(int2 r) { return r.x / REFIT_CHUNK != r.y / REFIT_CHUNK; }
__global__ void k_span_mark(SpanArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.F - 1 || !spans_chunks(a.range[i])) return;
  const int k = (int)atomicAdd(a.span_n, 1u);
  a.span_of[i] = k;
  if (k < SPAN_MAX) a.span_p[k] = i;
}
__global__ void k_span_link(SpanArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = min((int)*a.span_n, SPAN_MAX);
  if (k >= n) return;
  const int p = a.span_p[k];
  const int par = a.node_parent[p];   // the parent's range holds this one's: it spans too
  const int c0 = a.child[2 * p], c1 = a.child[2 * p + 1];
  const int s0 = (c0 >= 0 && spans_chunks(a.range[c0])) ? a.span_of[c0] : -1;
  const int s1 = (c1 >= 0 && spans_chunks(a.range[c1])) ? a.span_of[c1] : -1;
  a.span[k] = make_int4(p, par >= 0 ? a.span_of[par] : -1, s0, s1);
  a.span_cid[k] = make_int2(c0, c1);
}

// Refit of the spanning nodes in LDS (one workgroup, at most SPAN_MAX of them): each node's
// children are spanning nodes of the list or chunk-local nodes / leaves whose boxes
// k_refit_local has written; the latter are staged first and counted as arrived, then the
// nodes with both children staged start climbs, and a climb continues into the parent when
// it brings the parent's second child (LDS arrival counters).
__device__ __forceinline__ void refit_top_lds(const RefitArgs& a, int ns) {
  __shared__ float sbox[SPAN_MAX][6];      // the spanning nodes' boxes
  __shared__ float cbox[SPAN_MAX][2][6];   // their chunk-local children's boxes
  __shared__ int4 sinfo[SPAN_MAX];
  __shared__ int scnt[SPAN_MAX];
  // one entry per thread (SPAN_MAX <= blockDim); the climbs' start nodes are fixed here,
  // before any climb can raise a counter to 2
  static_assert(SPAN_MAX <= 1024, "one spanning node per thread of the 1024-thread workgroup");
  int start = -1;
  for (int k = threadIdx.x; k < ns; k += blockDim.x) {
    const int4 e = a.span[k];
    const int2 cid = a.span_cid[k];
    sinfo[k] = e;
    int staged = 0;
    for (int j = 0; j < 2; j++) {
      if ((j ? e.w : e.z) >= 0) continue;
      const int c = j ? cid.y : cid.x;
      const float4 L = c < 0 ? a.leaf_lo[~c] : a.node_lo[c];
      const float4 H = c < 0 ? a.leaf_hi[~c] : a.node_hi[c];
      cbox[k][j][0] = L.x; cbox[k][j][1] = L.y; cbox[k][j][2] = L.z;
      cbox[k][j][3] = H.x; cbox[k][j][4] = H.y; cbox[k][j][5] = H.z;
      staged++;
    }
    scnt[k] = staged;
    if (staged == 2) start = k;
  }
  __syncthreads();
  if (start >= 0) {
    int k = start;
    for (int level = 0; level < 64; level++) {
      const int4 e = sinfo[k];
      const float* b0 = e.z >= 0 ? sbox[e.z] : cbox[k][0];
      const float* b1 = e.w >= 0 ? sbox[e.w] : cbox[k][1];
      float4 L = make_float4(b0[0], b0[1], b0[2], 0.0f), H = make_float4(b0[3], b0[4], b0[5], 0.0f);
      box_union(L, H, make_float4(b1[0], b1[1], b1[2], 0.0f), make_float4(b1[3], b1[4], b1[5], 0.0f));
      sbox[k][0] = L.x; sbox[k][1] = L.y; sbox[k][2] = L.z; sbox[k][3] = H.x; sbox[k][4] = H.y; sbox[k][5] = H.z;
      a.node_lo[e.x] = L;
      a.node_hi[e.x] = H;
      if (e.y < 0) break;   // the root
      __threadfence_block();
      if (atomicAdd(&scnt[e.y], 1) != 1) break;   // the parent's other child is still on its way
      __threadfence_block();
      k = e.y;
    }
  }
}

__global__ __launch_bounds__(1024) void k_refit_top(RefitArgs a) {
  const int ns = a.span_n ? (int)*a.span_n : SPAN_MAX + 1;
  if (ns <= SPAN_MAX) {
    refit_top_lds(a, ns);
    return;
  }
  const int n = (int)a.counter[a.F];
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    int q = a.frontier[k];
    float4 L = q < 0 ? a.leaf_lo[~q] : a.node_lo[q];
    float4 H = q < 0 ? a.leaf_hi[~q] : a.node_hi[q];
    int p = q < 0 ? a.leaf_parent[~q] : a.node_parent[q];
    for (int level = 0; p >= 0 && level < 64; level++) {
      const int c0 = a.child[2 * p], c1 = a.child[2 * p + 1];   // in flight during the atomic
      const int pp = a.node_parent[p];
      __threadfence_block();
      if (atomicAdd(&a.counter[p], 1u) == 0u) break;
      __threadfence_block();
      const int sib = c0 == q ? c1 : c0;
      // a spanning sibling was written by another wave of this workgroup: read it from L2
      const volatile float* sl = (const volatile float*)(sib < 0 ? a.leaf_lo + ~sib : a.node_lo + sib);
      const volatile float* sh = (const volatile float*)(sib < 0 ? a.leaf_hi + ~sib : a.node_hi + sib);
      box_union(L, H, make_float4(sl[0], sl[1], sl[2], 0.0f), make_float4(sh[0], sh[1], sh[2], 0.0f));
      a.node_lo[p] = L;
      a.node_hi[p] = H;
      q = p;
      p = pp;
    }
  }
}

// Box padding: the MT test can accept a point a few ulps outside its triangle, and the slab
// test rounds; widen every box so traversal never culls a triangle BF would accept.
// Child boxes are padded by 1e-5 of their own magnitude plus 1e-6 of the scene's (the largest
// |coordinate| of the triangle centroids): the slab test's rounding scales with |origin|.
__device__ __forceinline__ void pad_box(float4& lo, float4& hi, float pa) {
  const float px = 1e-5f * (fabsf(lo.x) + fabsf(hi.x) + (hi.x - lo.x)) + pa;
  const float py = 1e-5f * (fabsf(lo.y) + fabsf(hi.y) + (hi.y - lo.y)) + pa;
  const float pz = 1e-5f * (fabsf(lo.z) + fabsf(hi.z) + (hi.z - lo.z)) + pa;
  lo.x -= px; lo.y -= py; lo.z -= pz;
  hi.x += px; hi.y += py; hi.z += pz;
}
__device__ __forceinline__ void pack_node(const PackArgs& a, int i) {
  const int c0 = a.child[2 * i], c1 = a.child[2 * i + 1];
  float4 l0 = c0 < 0 ? a.leaf_lo[~c0] : a.node_lo[c0];
  float4 h0 = c0 < 0 ? a.leaf_hi[~c0] : a.node_hi[c0];
  float4 l1 = c1 < 0 ? a.leaf_lo[~c1] : a.node_lo[c1];
  float4 h1 = c1 < 0 ? a.leaf_hi[~c1] : a.node_hi[c1];
  float S = 0.0f;
  for (int k = 0; k < 6; k++) {
    const float b = key2f(a.bounds[k]);
    if (fabsf(b) <= 3.4e38f) S = fmaxf(S, fabsf(b));
  }
  const float pa = 1e-6f * S + 1e-30f;
  pad_box(l0, h0, pa);
  pad_box(l1, h1, pa);
  float4* n = a.nodes + 4 * i;
  n[0] = make_float4(l0.x, l0.y, l0.z, h0.x);
  n[1] = make_float4(h0.y, h0.z, l1.x, l1.y);
  n[2] = make_float4(l1.z, h1.x, h1.y, h1.z);
  n[3] = make_float4(__int_as_float(c0), __int_as_float(c1), 0.0f, 0.0f);
}

// Suspect triangles (see Tracer in drh_device.h): light == 0 tests the view origin bound
// |(o-A).N| <= 2e-6 (|e1| + |e2| + |o-A|) in double with a 2x margin over the exact bound;
// light == 1 lists the triangles whose shadow-ray determinant, computed as mt_test computes it
// for d = -L, is clamped (0 < |det| < 1e-6), with that test's pvec and invDet.  det = 0, NaN
// or infinite gives an invDet that rejects every ray; |det| >= 1e-6 is the unclamped test the
// LBVH walk makes.
__global__ void k_suspects(SuspectArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.F) return;
  const float4 A = a.bf_a[g], E1 = a.bf_e1[g], E2 = a.bf_e2[g];
  if (a.light) {
    light_suspect(a, g, A, E1, E2);
    return;
  }
  const double e1x = E1.x, e1y = E1.y, e1z = E1.z, e2x = E2.x, e2y = E2.y, e2z = E2.z;
  const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  const double tx = (double)a.ox - A.x, ty = (double)a.oy - A.y, tz = (double)a.oz - A.z;
  const double V = fabs(tx * nx + ty * ny + tz * nz);
  const double bound = sqrt(e1x * e1x + e1y * e1y + e1z * e1z) + sqrt(e2x * e2x + e2y * e2y + e2z * e2z) +
                       sqrt(tx * tx + ty * ty + tz * tz);
  const bool hit = V <= 2e-6 * bound + 1e-12;
  if (hit || !(nx == nx && ny == ny && nz == nz)) {
    const int k = atomicAdd(a.count, 1);
    a.list[k] = make_float4((float)nx, (float)ny, (float)nz, __uint_as_float((uint32_t)g));
  }
}

// ---------------------------------------------------------------------------
// Suspect binning.  Every pass runs one wavefront per suspect: its lanes walk the suspect's
// candidate cells (a box that bounds where it can be accepted) and keep a cell when an exact
// test says some ray of the cell could be accepted; the count pass adds 1 per kept cell, the
// scan turns the counts into pool offsets (a cell past the pool's capacity gets n = -1) and
// the fill pass writes each kept pair's record at its cell's next slot.  Count and fill compute
// the same cells (the same arithmetic).
// ---------------------------------------------------------------------------

// View binning.  A record is kept for a screen cell unless a necessary condition of its
// clamped acceptance fails for every direction d through the cell:
//   C: |d.N| < 1.5e-6 (the per-lane filter the rays apply, N from the list);
//   U: |qvec.d| <= ub, with qvec = (o - A) x e1 as mt_test computes it for this view's rays
//      and ub >= (1 + 2^-23) max(|det|, 1e-6) for every det the C filter lets through (u in
//      (0, 1) needs |qvec.d| |invDet| < 1);
//   V: |w.d| <= ub + ev, w = e2 x (o - A): mt_test's v needs |(o - A).(d x e2)| < ub as well;
//   K: |(a qvec - b w).d| <= |a| (ub + eu) + |b| (ub + ev), U and V combined with the
//      coordinates (a, b) of o - A along (e1, e2) in the basis (e1, e2, N/|N|).  qvec and w
//      are nearly parallel to N (the triangle's plane passes close to o), so the U and V
//      bands are two almost coincident lines across the image; in a qvec - b w the N parts
//      cancel, and its band crosses them, which confines the record to a short stretch.
// eu, ev bound the float dot products' rounding and the ray direction's error (2e-5 + 1e-7 |o|
// relative), scaled by |qvec| and |o - A| |e2|.  The exact direction through screen point
// (x, y) is +-D / |D| with D = R p (R: the rotation of the view's inverse world-view matrix,
// p: proj_inv (ndc, 0, 1), both affine in x and y; the sign that of p.w), so n.D is affine
// over a cell and |D| convex: if n.D has one sign at the cell's four corners and
// min |n.D| > tau max |D|, then |n.d| > tau everywhere in the cell.  Candidate cells: the box
// shared by the parallelograms where two of the conditions hold with |D| at its largest over
// the grid, over every pair of non-parallel bands.  Double precision.
struct CamBin {
  double D0[3], Dx[3], Dy[3];   // D(x, y) = D0 + x Dx + y Dy
  double dmax;                  // max |D| over the grid
  double n[4][3], tau[4];       // conditions C, U, V, K (K: tau < 0 when absent)
  double f[4][3];               // n_i.D(x, y) = f[i][0] + x f[i][1] + y f[i][2]
  double t2[4];                 // tau_i^2
  double g[6];                  // |D(x, y)|^2 = g0 + x (g1 + x g3 + y g4) + y (g2 + y g5)
  float4 e, q;                  // the record
};
__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void cam_bin_setup(const SusBinArgs& a, int view, int s, int cells_x, int cells_y, CamBin& b) {
  const DevCam& c = a.cams[view];
  const float* B = c.proj_inv;
  const float* M = c.wv_inv;
  double P0[3], Px[3], Py[3];
  for (int r = 0; r < 3; r++) {
    P0[r] = -(double)B[r] + (double)B[4 + r] + (double)B[12 + r];
    Px[r] = 2.0 / (double)c.W * (double)B[r];
    Py[r] = -2.0 / (double)c.H * (double)B[4 + r];
  }
  for (int r = 0; r < 3; r++) {
    b.D0[r] = P0[0] * M[r] + P0[1] * M[4 + r] + P0[2] * M[8 + r];
    b.Dx[r] = Px[0] * M[r] + Px[1] * M[4 + r] + Px[2] * M[8 + r];
    b.Dy[r] = Py[0] * M[r] + Py[1] * M[4 + r] + Py[2] * M[8 + r];
  }
  const double gw = (double)(cells_x * SUS_CELL_PX), gh = (double)(cells_y * SUS_CELL_PX);
  b.dmax = 0.0;
  for (int k = 0; k < 4; k++) {
    const double x = (k & 1) ? gw : 0.0, y = (k & 2) ? gh : 0.0;
    double D[3];
    for (int r = 0; r < 3; r++) D[r] = b.D0[r] + x * b.Dx[r] + y * b.Dy[r];
    b.dmax = fmax(b.dmax, sqrt(dot3(D, D)));
  }
  float oa[4];
  mv4(c.wv_inv, 0.0f, 0.0f, 0.0f, 1.0f, oa);   // eye_ray's origin, bit for bit
  const F3 o = {oa[0], oa[1], oa[2]};
  const double delta = 2e-5 + 1e-7 * sqrt((double)o.x * o.x + (double)o.y * o.y + (double)o.z * o.z);
  const float4 e = a.list[(long long)view * a.F + s];
  const int g = __float_as_int(e.w);
  const float4 A = a.bf_a[g], E1 = a.bf_e1[g], E2 = a.bf_e2[g];
  const F3 tv = sub(o, f4xyz(A));   // mt_test's tvec
  const F3 qvec = cross(tv, f4xyz(E1));
  const double l1 = sqrt((double)E1.x * E1.x + (double)E1.y * E1.y + (double)E1.z * E1.z);
  const double l2 = sqrt((double)E2.x * E2.x + (double)E2.y * E2.y + (double)E2.z * E2.z);
  const float ub = (float)(1.6e-6 + 1e-6 * l1 * l2);
  b.e = e;
  b.q = make_float4(qvec.x, qvec.y, qvec.z, ub);
  b.n[0][0] = e.x; b.n[0][1] = e.y; b.n[0][2] = e.z;
  b.tau[0] = 1.5e-6 + (delta + 1e-6) * sqrt(dot3(b.n[0], b.n[0]));
  b.n[1][0] = qvec.x; b.n[1][1] = qvec.y; b.n[1][2] = qvec.z;
  const double eu = (delta + 1e-6) * sqrt(dot3(b.n[1], b.n[1]));
  b.tau[1] = (double)ub + eu;
  const double tx = tv.x, ty = tv.y, tz = tv.z, ax = E1.x, ay = E1.y, az = E1.z, bx = E2.x, by = E2.y, bz = E2.z;
  b.n[2][0] = by * tz - bz * ty; b.n[2][1] = bz * tx - bx * tz; b.n[2][2] = bx * ty - by * tx;   // e2 x tvec
  const double ev = (delta + 1e-6) * sqrt(tx * tx + ty * ty + tz * tz) * l2;
  b.tau[2] = (double)ub + ev;
  // (ka, kb): tvec = ka e1 + kb e2 + c N/|N|, by Cramer's rule with det = |N|
  const double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
  const double nn = sqrt(nx * nx + ny * ny + nz * nz);
  b.tau[3] = -1.0;
  if (nn > 0.0) {
    const double hx = nx / nn, hy = ny / nn, hz = nz / nn;
    const double ka = (tx * (by * hz - bz * hy) + ty * (bz * hx - bx * hz) + tz * (bx * hy - by * hx)) / nn;
    const double kb = (tx * (hy * az - hz * ay) + ty * (hz * ax - hx * az) + tz * (hx * ay - hy * ax)) / nn;
    for (int r = 0; r < 3; r++) b.n[3][r] = ka * b.n[1][r] - kb * b.n[2][r];
    b.tau[3] = (fabs(ka) * ((double)ub + eu) + fabs(kb) * ((double)ub + ev)) * (1.0 + 1e-9);
  }
  for (int i = 0; i < 4; i++) {
    b.f[i][0] = dot3(b.D0, b.n[i]); b.f[i][1] = dot3(b.Dx, b.n[i]); b.f[i][2] = dot3(b.Dy, b.n[i]);
    b.t2[i] = b.tau[i] * b.tau[i];
  }
  b.g[0] = dot3(b.D0, b.D0); b.g[1] = 2.0 * dot3(b.D0, b.Dx); b.g[2] = 2.0 * dot3(b.D0, b.Dy);
  b.g[3] = dot3(b.Dx, b.Dx); b.g[4] = 2.0 * dot3(b.Dx, b.Dy); b.g[5] = dot3(b.Dy, b.Dy);
}
// false when |n_i.d| > tau_i for every direction through the cell, for some condition i
__device__ __forceinline__ bool cam_cell_kept(const CamBin& b, int cx, int cy) {
  double f[4][4], dm2 = 0.0;   // dm2: max |D|^2 over the corners
  for (int k = 0; k < 4; k++) {
    const double x = (double)((cx + (k & 1)) * SUS_CELL_PX), y = (double)((cy + (k >> 1)) * SUS_CELL_PX);
    dm2 = fmax(dm2, b.g[0] + x * (b.g[1] + x * b.g[3] + y * b.g[4]) + y * (b.g[2] + y * b.g[5]));
    for (int i = 0; i < 4; i++) f[i][k] = b.f[i][0] + x * b.f[i][1] + y * b.f[i][2];
  }
  for (int i = 0; i < 4; i++) {
    if (!(b.tau[i] >= 0.0)) continue;
    const bool pos = f[i][0] > 0.0 && f[i][1] > 0.0 && f[i][2] > 0.0 && f[i][3] > 0.0;
    const bool neg = f[i][0] < 0.0 && f[i][1] < 0.0 && f[i][2] < 0.0 && f[i][3] < 0.0;
    if (!pos && !neg) continue;
    const double lo = fmin(fmin(fabs(f[i][0]), fabs(f[i][1])), fmin(fabs(f[i][2]), fabs(f[i][3])));
    if (lo * lo > b.t2[i] * dm2) return false;   // min |n.D| > tau max |D|, squared
  }
  return true;
}
// candidate cells: [cx0, cx1] x [cy0, cy1], the box shared by the parallelograms of every pair
// of non-parallel conditions (|f_i| <= tau_i max |D| for both)
__device__ __forceinline__ void cam_candidates(const CamBin& b, int cells_x, int cells_y, int& cx0, int& cx1, int& cy0,
                                               int& cy1) {
  double x0 = 0.0, x1 = (double)(cells_x * SUS_CELL_PX), y0 = 0.0, y1 = (double)(cells_y * SUS_CELL_PX);
  for (int i = 0; i < 4; i++)
    for (int j = i + 1; j < 4; j++) {
      if (!(b.tau[i] >= 0.0) || !(b.tau[j] >= 0.0)) continue;
      const double ti = b.tau[i] * b.dmax, tj = b.tau[j] * b.dmax;
      const double det = b.f[i][1] * b.f[j][2] - b.f[i][2] * b.f[j][1];
      const double scale = (fabs(b.f[i][1]) + fabs(b.f[i][2])) * (fabs(b.f[j][1]) + fabs(b.f[j][2]));
      if (!(fabs(det) > 1e-9 * scale) || !(scale < 1e300)) continue;
      double px0 = 1e300, px1 = -1e300, py0 = 1e300, py1 = -1e300;
      for (int k = 0; k < 4; k++) {
        const double ri = ((k & 1) ? ti : -ti) - b.f[i][0], rj = ((k & 2) ? tj : -tj) - b.f[j][0];
        const double x = (ri * b.f[j][2] - rj * b.f[i][2]) / det, y = (b.f[i][1] * rj - b.f[j][1] * ri) / det;
        px0 = fmin(px0, x); px1 = fmax(px1, x); py0 = fmin(py0, y); py1 = fmax(py1, y);
      }
      if (!(px0 == px0 && px1 == px1 && py0 == py0 && py1 == py1)) continue;
      x0 = fmax(x0, px0); x1 = fmin(x1, px1); y0 = fmax(y0, py0); y1 = fmin(y1, py1);
    }
  const double g0 = floor(x0 / SUS_CELL_PX) - 1.0, g1 = floor(x1 / SUS_CELL_PX) + 1.0;
  const double h0 = floor(y0 / SUS_CELL_PX) - 1.0, h1 = floor(y1 / SUS_CELL_PX) + 1.0;
  cx0 = (int)fmax(0.0, fmin(g0, (double)cells_x)); cx1 = (int)fmin((double)(cells_x - 1), fmax(g1, -1.0));
  cy0 = (int)fmax(0.0, fmin(h0, (double)cells_y)); cy1 = (int)fmin((double)(cells_y - 1), fmax(h1, -1.0));
}

// Light binning over a grid across the light.  Shadow rays share d = -L, so a record's
// determinant and invDet are fixed and its acceptance confines the ray's origin o: with
// s = sign(invDet), l1 = s (o - A).pvec (mt_test's v before the scale |invDet| = 1e6) and
// l2 = s (o - A).(e1 x d) (its u), acceptance needs l1 > 0, l2 > 0 and
// l1 + l2 < 1e-6 (1 + 2.4e-7).  l1, l2 are affine in the origin's coordinates (u, v) =
// (o.a1, o.a2) across the light (a1, a2 orthonormal, perpendicular to L; the component along
// L enters only through pvec's rounding), so the origins lie in the triangle of the (u, v)
// plane where l1 >= -m1, l2 >= -m2, l1 + l2 <= 1e-6 (1 + 2.4e-7) + m1 + m2, with m1, m2
// bounds on the float tvec / dot rounding (3e-7 and 5e-7 of |o - A| |pvec| and |o - A| |e1|,
// |o - A| at most the scene box's diagonal) and on the component along L.  Candidate cells: the
// triangle's box; a cell is kept when no triangle edge has all four (widened) corners outside
// (separating axes).  A cell is widened by 2% of its size plus the rounding of a ray's own
// float (u, v).  The grid covers the scene box (the LBVH root's child boxes) with a margin.
struct LightGrid {
  double b1[3], b2[3], L[3];
  float u0, v0, iu, iv;
  double wmax, diag, slack;
  bool ok;
};
__device__ __forceinline__ void light_grid(const LightBinArgs& a, LightGrid& G) {
  const double ll = sqrt((double)a.Lx * a.Lx + (double)a.Ly * a.Ly + (double)a.Lz * a.Lz);
  const double Lx = a.Lx / ll, Ly = a.Ly / ll, Lz = a.Lz / ll;
  G.L[0] = Lx; G.L[1] = Ly; G.L[2] = Lz;
  const double ax = fabs(Lx), ay = fabs(Ly), az = fabs(Lz);
  // a1 = L x (the axis least aligned with L), normalised; a2 = L x a1
  double ex = 0, ey = 0, ez = 0;
  if (ax <= ay && ax <= az) ex = 1; else if (ay <= az) ey = 1; else ez = 1;
  double b1x = Ly * ez - Lz * ey, b1y = Lz * ex - Lx * ez, b1z = Lx * ey - Ly * ex;
  const double bl = sqrt(b1x * b1x + b1y * b1y + b1z * b1z);
  b1x /= bl; b1y /= bl; b1z /= bl;
  G.b1[0] = b1x; G.b1[1] = b1y; G.b1[2] = b1z;
  G.b2[0] = Ly * b1z - Lz * b1y; G.b2[1] = Lz * b1x - Lx * b1z; G.b2[2] = Lx * b1y - Ly * b1x;
  const float4 n0 = a.nodes[0], n1 = a.nodes[1], n2 = a.nodes[2];
  const double lo[3] = {fmin((double)n0.x, (double)n1.z), fmin((double)n0.y, (double)n1.w), fmin((double)n0.z, (double)n2.x)};
  const double hi[3] = {fmax((double)n0.w, (double)n2.y), fmax((double)n1.x, (double)n2.z), fmax((double)n1.y, (double)n2.w)};
  double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300, cmax = 0.0;
  G.wmax = 0.0;
  for (int k = 0; k < 8; k++) {
    const double p[3] = {(k & 1) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1], (k & 4) ? hi[2] : lo[2]};
    const double u = dot3(p, G.b1), v = dot3(p, G.b2), w = dot3(p, G.L);
    umin = fmin(umin, u); umax = fmax(umax, u); vmin = fmin(vmin, v); vmax = fmax(vmax, v);
    G.wmax = fmax(G.wmax, fabs(w));
    cmax = fmax(cmax, sqrt(dot3(p, p)));
  }
  G.diag = sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) + (hi[2] - lo[2]) * (hi[2] - lo[2]));
  const double pad = 0.01 * (umax - umin + vmax - vmin) + 1e-3 * (1.0 + fmax(fmax(fabs(umin), fabs(umax)), fmax(fabs(vmin), fabs(vmax))));
  G.u0 = (float)(umin - pad);
  G.v0 = (float)(vmin - pad);
  G.iu = (float)(a.bp.cells_x / (umax - umin + 2 * pad));
  G.iv = (float)(a.bp.cells_y / (vmax - vmin + 2 * pad));
  G.slack = 0.02 / fmin((double)G.iu, (double)G.iv) + 1e-5 * (1.0 + cmax);
  G.ok = G.diag == G.diag && G.diag < 1e30 && G.iu > 0.0f && G.iv > 0.0f && G.iu < 3.0e38f && G.iv < 3.0e38f;
}
struct LightRec {
  double j11, j12, j21, j22, c1, c2, m1, m2, T;
  float4 r0, r1;
};
__device__ __forceinline__ void light_rec(const LightBinArgs& a, const LightGrid& G, int s, LightRec& R) {
  R.r0 = a.list[2 * s];
  R.r1 = a.list[2 * s + 1];
  const float4 E1 = a.bf_e1[__float_as_int(R.r0.w)];
  const F3 d = mul(-1.0f, F3{a.Lx, a.Ly, a.Lz});   // the shadow rays' direction, as lambert_terms forms it
  const double sg = R.r1.w > 0.0f ? 1.0 : -1.0;
  const double p[3] = {R.r1.x, R.r1.y, R.r1.z};                                               // pvec
  const double q[3] = {(double)E1.y * d.z - (double)E1.z * d.y, (double)E1.z * d.x - (double)E1.x * d.z,
                       (double)E1.x * d.y - (double)E1.y * d.x};                                 // e1 x d
  const double A[3] = {R.r0.x, R.r0.y, R.r0.z};
  const double lp = sqrt(dot3(p, p)), lq = sqrt(dot3(q, q));
  const double le1 = sqrt((double)E1.x * E1.x + (double)E1.y * E1.y + (double)E1.z * E1.z);
  R.j11 = sg * dot3(G.b1, p); R.j12 = sg * dot3(G.b2, p);
  R.j21 = sg * dot3(G.b1, q); R.j22 = sg * dot3(G.b2, q);
  R.c1 = -sg * dot3(A, p); R.c2 = -sg * dot3(A, q);
  const double w1 = fabs(dot3(G.L, p)) * G.wmax + 1e-7 * lp * G.wmax;
  const double w2 = fabs(dot3(G.L, q)) * G.wmax + 1e-7 * lq * G.wmax;
  R.m1 = 2.0 * (3e-7 * G.diag * lp + w1) + 1e-12;
  R.m2 = 2.0 * (5e-7 * G.diag * le1 + w2) + 1e-12;
  R.T = 1e-6 * (1.0 + 2.4e-7) + R.m1 + R.m2;
}
__device__ __forceinline__ void light_candidates(const LightGrid& G, const LightRec& R, int cells_x, int cells_y,
                                                 int& cx0, int& cx1, int& cy0, int& cy1) {
  cx0 = 0; cx1 = cells_x - 1; cy0 = 0; cy1 = cells_y - 1;
  const double det = R.j11 * R.j22 - R.j12 * R.j21;
  if (!(det != 0.0) || det != det) return;
  const double L1[3] = {-R.m1, R.T + R.m2, -R.m1}, L2[3] = {-R.m2, -R.m2, R.T + R.m1};
  double ulo = 1e300, uhi = -1e300, vlo = 1e300, vhi = -1e300;
  for (int k = 0; k < 3; k++) {
    const double r1 = L1[k] - R.c1, r2 = L2[k] - R.c2;
    const double u = (r1 * R.j22 - r2 * R.j12) / det, v = (R.j11 * r2 - R.j21 * r1) / det;
    ulo = fmin(ulo, u); uhi = fmax(uhi, u); vlo = fmin(vlo, v); vhi = fmax(vhi, v);
  }
  if (!(ulo == ulo && uhi == uhi && vlo == vlo && vhi == vhi)) return;
  const double g0 = floor((ulo - G.slack - G.u0) * G.iu), g1 = floor((uhi + G.slack - G.u0) * G.iu);
  const double h0 = floor((vlo - G.slack - G.v0) * G.iv), h1 = floor((vhi + G.slack - G.v0) * G.iv);
  cx0 = (int)fmax(0.0, fmin(g0, (double)cells_x)); cx1 = (int)fmin((double)(cells_x - 1), fmax(g1, -1.0));
  cy0 = (int)fmax(0.0, fmin(h0, (double)cells_y)); cy1 = (int)fmin((double)(cells_y - 1), fmax(h1, -1.0));
}
__device__ __forceinline__ bool light_cell_kept(const LightGrid& G, const LightRec& R, int gx, int gy) {
  const double ua = G.u0 + gx / (double)G.iu - G.slack, ub = G.u0 + (gx + 1) / (double)G.iu + G.slack;
  const double va = G.v0 + gy / (double)G.iv - G.slack, vb = G.v0 + (gy + 1) / (double)G.iv + G.slack;
  bool out1 = true, out2 = true, out3 = true;
  for (int k = 0; k < 4; k++) {
    const double u = (k & 1) ? ub : ua, v = (k & 2) ? vb : va;
    const double l1 = R.j11 * u + R.j12 * v + R.c1, l2 = R.j21 * u + R.j22 * v + R.c2;
    out1 = out1 && l1 < -R.m1;
    out2 = out2 && l2 < -R.m2;
    out3 = out3 && l1 + l2 > R.T;
  }
  return !(out1 || out2 || out3);
}

// kept pairs into the pool: PASS 0 counts, PASS 1 fills (after the scan)
template <int PASS>
__device__ __forceinline__ void bin_emit(const BinPool& bp, int cell, float4 r0, float4 r1) {
  if (PASS == 0) {
    atomicAdd(bp.cnt + cell, 1);
  } else {
    const int2 ci = bp.cell[cell];
    if (ci.y < 0) return;
    const int k = atomicAdd(bp.cnt + cell, 1);
    bp.pool[2ll * (ci.x + k)] = r0;
    bp.pool[2ll * (ci.x + k) + 1] = r1;
  }
}

// Calls f(X, Y) for the cells of [cx0, cx1] x [cy0, cy1] (in cell units: cell (X, Y) is the unit
// square at (X, Y)) that the band |A + BX x + BY y| <= t meets, widened by one cell; the lanes
// take the rows (or columns, for a band closer to horizontal) in turn.  Without a usable
// direction every cell of the box is visited.
template <class Fn>
__device__ __forceinline__ void band_walk(double A, double BX, double BY, double t, int cx0, int cx1, int cy0, int cy1,
                                          int lane, Fn f) {
  const int wx = cx1 - cx0 + 1, wy = cy1 - cy0 + 1;
  if (wx <= 0 || wy <= 0) return;   // an empty range has no cell
  const bool dir = fabs(BX) + fabs(BY) > 0.0 && fabs(A) + fabs(BX) + fabs(BY) + t < 1e300 && t == t;
  if (!dir) {
    for (long long k = lane; k < (long long)wx * wy; k += 64) f(cx0 + (int)(k % wx), cy0 + (int)(k / wx));
    return;
  }
  const bool rows = fabs(BX) >= fabs(BY);
  const double P = rows ? BX : BY, Q = rows ? BY : BX;   // |A + P a + Q b| <= t along a row (b fixed) or column
  const int o0 = rows ? cy0 : cx0, o1 = rows ? cy1 : cx1, i0 = rows ? cx0 : cy0, i1 = rows ? cx1 : cy1;
  for (int r = o0 + lane; r <= o1; r += 64) {
    double lo = 1e300, hi = -1e300;
    for (int k = 0; k < 4; k++) {
      const double b = (double)(r + (k & 1)), sg = (k & 2) ? t : -t;
      const double a = (sg - A - Q * b) / P;
      lo = fmin(lo, a); hi = fmax(hi, a);
    }
    int a0 = i0, a1 = i1;
    if (lo == lo && hi == hi) {
      a0 = (int)fmax((double)i0, fmin(floor(lo) - 1.0, (double)i1 + 1.0));
      a1 = (int)fmin((double)i1, fmax(floor(hi) + 1.0, (double)i0 - 1.0));
    }
    for (int c = a0; c <= a1; c++) {
      if (rows) f(c, r); else f(r, c);
    }
  }
}

template <int PASS>
__global__ __launch_bounds__(256) void k_sus_bin(SusBinArgs a) {
  const int view = blockIdx.y;
  BinPool bp = a.bp;
  bp.pool += 2 * a.pool_stride * view;
  bp.cell += a.cell_stride * view;
  bp.cnt += a.cell_stride * view;
  bp.cells_x = (a.cams[view].w + SUS_CELL_PX - 1) / SUS_CELL_PX;
  bp.cells_y = (a.cams[view].h + SUS_CELL_PX - 1) / SUS_CELL_PX;
  const int lane = threadIdx.x & 63;
  const int n = a.count[view];
  const int waves = gridDim.x * (blockDim.x >> 6);
  for (int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < n; s += waves) {
    CamBin b;
    cam_bin_setup(a, view, s, bp.cells_x, bp.cells_y, b);
    int cx0, cx1, cy0, cy1;
    cam_candidates(b, bp.cells_x, bp.cells_y, cx0, cx1, cy0, cy1);
    // walk along the thinnest band (in cell units: x = SUS_CELL_PX X)
    int best = -1;
    double bw = 1e300;
    for (int i = 0; i < 4; i++) {
      if (!(b.tau[i] >= 0.0)) continue;
      const double g = sqrt(b.f[i][1] * b.f[i][1] + b.f[i][2] * b.f[i][2]);
      const double w = b.tau[i] * b.dmax / g;
      if (w < bw) { bw = w; best = i; }
    }
    const double A = best >= 0 ? b.f[best][0] : 0.0;
    const double BX = best >= 0 ? SUS_CELL_PX * b.f[best][1] : 0.0, BY = best >= 0 ? SUS_CELL_PX * b.f[best][2] : 0.0;
    const double t = best >= 0 ? b.tau[best] * b.dmax : 0.0;
    band_walk(A, BX, BY, t, cx0, cx1, cy0, cy1, lane, [&](int cx, int cy) {
      if (cam_cell_kept(b, cx, cy)) bin_emit<PASS>(bp, cy * bp.cells_x + cx, b.e, b.q);
    });
  }
}

template <int PASS>
__global__ __launch_bounds__(256) void k_light_bin(LightBinArgs a) {
  LightGrid G;
  light_grid(a, G);
  if (PASS == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    a.grid[0] = make_float4((float)G.b1[0], (float)G.b1[1], (float)G.b1[2], G.u0);
    a.grid[1] = make_float4((float)G.b2[0], (float)G.b2[1], (float)G.b2[2], G.v0);
    a.grid[2] = make_float4(G.iu, G.iv, 0.0f, 0.0f);
  }
  const int lane = threadIdx.x & 63;
  const int n = *a.count;
  const int waves = gridDim.x * (blockDim.x >> 6);
  for (int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < n; s += waves) {
    LightRec R;
    light_rec(a, G, s, R);
    int cx0 = 0, cx1 = a.bp.cells_x - 1, cy0 = 0, cy1 = a.bp.cells_y - 1;
    if (!G.ok) {
      band_walk(0.0, 0.0, 0.0, 0.0, cx0, cx1, cy0, cy1, lane, [&](int gx, int gy) {
        bin_emit<PASS>(a.bp, gy * a.bp.cells_x + gx, R.r0, R.r1);
      });
      continue;
    }
    light_candidates(G, R, a.bp.cells_x, a.bp.cells_y, cx0, cx1, cy0, cy1);
    // the thinner of the bands -m1 <= l1 <= T + m2 and -m2 <= l2 <= T + m1, in cell units
    // (u = u0 + X / iu, v = v0 + Y / iv)
    const double h1 = 0.5 * (R.T + R.m1 + R.m2), mid1 = 0.5 * (R.T + R.m2 - R.m1);
    const double h2 = 0.5 * (R.T + R.m1 + R.m2), mid2 = 0.5 * (R.T + R.m1 - R.m2);
    const double X1 = R.j11 / G.iu, Y1 = R.j12 / G.iv, X2 = R.j21 / G.iu, Y2 = R.j22 / G.iv;
    const double A1 = R.c1 - mid1 + R.j11 * G.u0 + R.j12 * G.v0, A2 = R.c2 - mid2 + R.j21 * G.u0 + R.j22 * G.v0;
    const bool first = h1 / sqrt(X1 * X1 + Y1 * Y1) <= h2 / sqrt(X2 * X2 + Y2 * Y2);
    // the cell widening (slack) is far below a cell: the walk's one-cell margin covers it
    band_walk(first ? A1 : A2, first ? X1 : X2, first ? Y1 : Y2, first ? h1 : h2, cx0, cx1, cy0, cy1, lane,
              [&](int gx, int gy) {
                if (light_cell_kept(G, R, gx, gy)) bin_emit<PASS>(a.bp, gy * a.bp.cells_x + gx, R.r0, R.r1);
              });
  }
}

// Pool offsets from the counts of n cells (one workgroup per list, blockIdx.x = list): cell =
// (offset, count), count = -1 past the capacity; the counts become the fill cursors (zero).
__global__ __launch_bounds__(1024) void k_bin_offsets(BinPool bp, long long cell_stride, int n) {
  __shared__ int part[1024];
  __shared__ int carry;
  int* cnt = bp.cnt + cell_stride * blockIdx.x;
  int2* cell = bp.cell + cell_stride * blockIdx.x;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 1024) {
    const int i = base + threadIdx.x;
    const int c = i < n ? cnt[i] : 0;
    part[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {   // inclusive scan
      const int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    const int start = carry + part[threadIdx.x] - c;
    if (i < n) {
      cell[i] = (start + c <= bp.cap) ? make_int2(start, c) : make_int2(0, -1);
      cnt[i] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
}

// the triangles in leaf order, beside the node packing (one launch)
__global__ void k_pack_nodes(PackArgs pk, ReorderArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (pk.zero && s < pk.nzero) pk.zero[s] = 0u;
  if (s < pk.F - 1) pack_node(pk, s);
  if (s >= a.F) return;
  const int g = (int)a.order[s];
  a.tri_a[s] = a.bf_a[g];
  a.tri_e1[s] = a.bf_e1[g];
  a.tri_e2[s] = a.bf_e2[g];
}

// ---------------------------------------------------------------------------
// render (src/drender.h:98-141) with the L2 loss fused (src/utils.cpp:15-27)
// ---------------------------------------------------------------------------
// Work of one wavefront.  With sq*sq in {1, 4, 16, 64} samples per pixel a wave takes a
// (8/sq) x (8/sq) pixel tile with one sample per lane, so its 64 rays cover a small footprint
// and walk the LBVH as one coherent packet; otherwise an 8x8 tile, one pixel per lane, whose
// samples are traced one after another.
struct TileMap { int x, y, S, s; };   // pixel, lanes per pixel, this lane's sample (0 if S == 1)
__device__ __forceinline__ TileMap tile_map_xy(int sq, int tx, int ty, int lane) {
  const bool inter = sq == 1 || sq == 2 || sq == 4 || sq == 8;
  const int S = inter ? sq * sq : 1;
  const int ts = tile_side(sq);               // tile side in pixels
  const int p = lane / S;                     // pixel within the tile
  return TileMap{tx * ts + p % ts, ty * ts + p / ts, S, lane % S};
}
// The tile of workgroup b of a grid of n (one per tile).  The dispatcher deals workgroups to
// the 8 XCDs round robin (b goes to XCD b % 8); this hands each XCD runs of 256 consecutive
// tiles instead (run j to XCD j % 8), so neighbouring tiles, which open the same LBVH nodes
// and triangles, share one XCD's L2 and scalar caches, while every XCD still gets runs from
// all over the image (one contiguous band per XCD left some XCDs with mostly empty
// background: config-2 render 18% slower; 64-tile runs 12% slower; 256-tile runs as fast as
// round robin there and 10% faster on the 16-view 200k-triangle scene; the edge pass's sample
// chunks dealt the same way ran 24% slower at config 2, so it keeps round robin).  A bijection
// on [0, n); the last n mod 2048 workgroups keep their own tile.
constexpr int XCD_RUN = 256;
__device__ __forceinline__ int xcd_tile(int b, int n) {
  if (b >= (n / (8 * XCD_RUN)) * (8 * XCD_RUN)) return b;
  const int x = b & 7, k = b >> 3;
  return ((k / XCD_RUN) * 8 + x) * XCD_RUN + k % XCD_RUN;
}
__device__ __forceinline__ TileMap tile_map(int sq, int w, int tile, int lane) {
  const int tiles_x = (w + tile_side(sq) - 1) / tile_side(sq);
  return tile_map_xy(sq, tile % tiles_x, tile / tiles_x, lane);
}

template <int MODE, bool COUNT = false, int TF = 0>
__global__ __launch_bounds__(64) void k_render(RenderArgs a) {
  __shared__ int wst[WST_INTS];   // the wave's traversal stack
  const Tracer& T = a.tr;
  const DevCam& c = a.cams[a.view];
  const int lane = threadIdx.x;
  const int sq = a.sq;
  const float ns = (float)(sq * sq);
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const TileMap tm = tile_map(sq, c.w, tile, lane);
  const int x = tm.x, y = tm.y;
  const bool inside = x < c.w && y < c.h;
  F3 acc = {0, 0, 0};
  if (tm.S > 1) {
    // one subsample per lane; the pixel's S lanes are adjacent, sample k = dy*sq + dx in lane
    // base + k, so the gather below adds col/spp in the reference's dy-major order
    F3 col = {0, 0, 0};
    if (inside) {
      float xo, yo;
      sample_offset(a.jit, a.view, x, y, tm.s, sq, xo, yo);
      Hit h;
      float lit = 0.0f;
      col = shade<MODE, COUNT, TF>(T, a.light, c, x + xo, y + yo, a.tr.s.nrm, wst, &h, &lit);
      if (a.hits)
        a.hits[(c.img_off / 3 + (long long)y * c.w + x) * tm.S + tm.s] =
            make_float4(__uint_as_float(h.gtri), h.u, h.v, MODE == MODE_LAMBERT_VCOL ? lit : h.t);
    }
    const int base = lane - tm.s;
    for (int k = 0; k < tm.S; k++) {
      const F3 ck = {__shfl(col.x, base + k, 64), __shfl(col.y, base + k, 64), __shfl(col.z, base + k, 64)};
      acc = add(acc, divs_pow2(ck, ns));
    }
  } else if (inside) {
    for (int dy = 0; dy < sq; dy++)
      for (int dx = 0; dx < sq; dx++) {
        float xo, yo;
        sample_offset(a.jit, a.view, x, y, dy * sq + dx, sq, xo, yo);
        Hit h;
        float lit = 0.0f;
        const F3 col = shade<MODE, COUNT, TF>(T, a.light, c, x + xo, y + yo, a.tr.s.nrm, wst, &h, &lit);
        if (a.hits)
          a.hits[(c.img_off / 3 + (long long)y * c.w + x) * (sq * sq) + dy * sq + dx] =
              make_float4(__uint_as_float(h.gtri), h.u, h.v, MODE == MODE_LAMBERT_VCOL ? lit : h.t);
        acc = add(acc, divs_pow2(col, ns));
      }
  }
  double l2 = 0.0;
  if (inside && tm.s == 0) {
    const long long p = c.img_off + 3ll * ((long long)y * c.w + x);
    if (a.img) { a.img[p] = acc.x; a.img[p + 1] = acc.y; a.img[p + 2] = acc.z; }
    if (a.target) {
      const F3 t = ld3(a.target + p);
      const F3 dv = sub(acc, t);
      const F3 ad = mul(2.0f, dv);
      a.adj[p] = ad.x; a.adj[p + 1] = ad.y; a.adj[p + 2] = ad.z;
      l2 = double(dot(dv, dv));
    }
  }
  if (a.target) {
    l2 = wave_sum_d(l2);
    if (lane == 0) a.loss_part[a.part_off + tile] = l2;
  }
}

// Deterministic final sum of the per-tile loss partials: mse / (views*W*H) as a float.
__global__ __launch_bounds__(1024) void k_loss_final(const double* part, int n, float* out, double denom) {
  __shared__ double s[1024];
  double v = 0.0;
  int i = threadIdx.x;
  // eight loads in flight per thread, added in the same order as one at a time
  for (; i + 7 * 1024 < n; i += 8 * 1024) {
    double x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = part[i + k * 1024];
#pragma unroll
    for (int k = 0; k < 8; k++) v += x[k];
  }
  for (; i < n; i += 1024) v += part[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = float(float(s[0]) / float(denom));
}

// Sum `v` over each run of equal `key` in the wavefront (all 64 lanes active; lanes that
// hold no contribution carry a key of their own); afterwards the last lane of a run holds
// the run's total.  The doubling steps stop once no run is longer than the step.
template <int K, typename Key>
__device__ __forceinline__ bool seg_reduce(Key key, float* v) {
  const int lane = threadIdx.x & 63;
  const Key prev = __shfl_up(key, 1, 64);
  const Key next = __shfl_down(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const uint64_t heads = __ballot(head);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int start = 63 - __clzll((long long)(heads & upto));
  for (int off = 1; off < 64; off <<= 1) {
    if (!__ballot(lane - off >= start)) break;
#pragma unroll
    for (int q = 0; q < K; q++) {
      const float o = __shfl_up(v[q], off, 64);
      if (lane - off >= start) v[q] += o;
    }
  }
  return lane == 63 || next != key;
}

// Add a float3 to gradient target `key` (its offset in grad) through the wave's LDS hash
// table: linear probing, slots claimed by compare-and-swap; a target that finds no slot
// within INTERIOR_HASH_PROBES goes straight to grad.
__device__ __forceinline__ void hash_add(int* hkey, float* hval, int nh, float* grad, int key, float x, float y, float z) {
  unsigned h = ((unsigned)key * 0x9E3779B1u) >> 16;
  for (int p = 0; p < INTERIOR_HASH_PROBES; p++) {
    h &= (unsigned)(nh - 1);
    const int k = atomicCAS(&hkey[h], -1, key);
    if (k == -1 || k == key) {
      atomicAdd(&hval[3 * h], x); atomicAdd(&hval[3 * h + 1], y); atomicAdd(&hval[3 * h + 2], z);
      return;
    }
    h++;
  }
  atomicAdd(grad + key, x); atomicAdd(grad + key + 1, y); atomicAdd(grad + key + 2, z);
}

// Deterministic mode: a gradient term as a 32.32 fixed-point integer added with a 64-bit
// integer atomic.  Integer addition is associative, so the sums do not depend on the order
// the waves arrive in and repeated runs agree bit for bit.  v * 2^32 is exact in float and its
// conversion exact for |v| >= 2^-9; smaller terms are rounded to a multiple of 2^-32.  Every
// partial sum must stay below 2^31 in magnitude.
__device__ __forceinline__ void dadd(unsigned long long* acc, int off, float v) {
  atomicAdd(acc + off, (unsigned long long)__float2ll_rn(v * 4294967296.0f));
}

__global__ void k_det_finish(const unsigned long long* acc, int G, float* grad, float* trscr, int ntr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G + ntr) return;
  const float v = (float)((double)(long long)acc[i] * (1.0 / 4294967296.0));
  if (i < G) grad[i] += v;
  else trscr[i - G] = v;
}
void launch_det_finish(const unsigned long long* dacc, int G, float* grad, float* trscr, int ntr, hipStream_t s) {
  if (G + ntr > 0) hipLaunchKernelGGL(k_det_finish, dim3((G + ntr + 255) / 256), dim3(256), 0, s, dacc, G, grad, trscr, ntr);
}

// ---------------------------------------------------------------------------
// interior_derivatives (src/drender.h:197-233): Hammersley subpixel offsets,
// shade_grad<VERTEX_COLOR> (shade_simple.cpp:39-74) / <TEXTURE_COLOR> (:132-162, LAMBERT too)
// ---------------------------------------------------------------------------
// POS (VERTEX_COLOR only): also the position gradient of the interior term, dF/du du/dV +
// dF/dv dv/dV through the barycentrics of each sample's hit (bary_grad), the term the
// reference's shade_grad<VERTEX_COLOR> has written out but disabled (src/shade_simple.cpp:75-108).
// LAMBERT_VCOL: the colour gradient of VERTEX_COLOR (shade_simple.cpp:39-74) times the light
// factor (shadow * Lc) * ndl + ambient of the hit (the derivative of its shade), which needs the
// hit's normal and shadow ray (traced again, or decoded from the render's stored hit).
template <int MODE, bool POS = false, bool COUNT = false, int TF = 0>
__global__ __launch_bounds__(64) void k_interior(InteriorArgs a) {
  constexpr bool VCOL_LIKE = MODE == 2 || MODE == MODE_LAMBERT_VCOL;
  extern __shared__ float lgrad[];  // a.lds_n floats (dynamic)
  __shared__ int wst[WST_INTS];
  const Tracer& T = a.tr;
  const DevCam& c = a.cams[a.view];
  const int lane = threadIdx.x;
  const bool use_lds = a.lds_n > 0;
  const int nh = a.hash;                       // hash table: nh keys, then 3*nh sums
  int* hkey = reinterpret_cast<int*>(lgrad);
  float* hval = lgrad + nh;
  for (int i = lane; i < a.lds_n; i += 64) lgrad[i] = 0.0f;
  for (int i = lane; i < nh; i += 64) { hkey[i] = -1; hval[3 * i] = 0.0f; hval[3 * i + 1] = 0.0f; hval[3 * i + 2] = 0.0f; }
  __syncthreads();
  const int sq = a.sq;
  const int rt = a.region;
  const int ts = tile_side(sq);
  const int tiles_x = (c.w + ts - 1) / ts, tiles_y = (c.h + ts - 1) / ts;
  const int units_x = (tiles_x + rt - 1) / rt;
  const int nunits = units_x * ((tiles_y + rt - 1) / rt);
  F3 uni = {0, 0, 0};  // uniform-target accumulator (single 1x1 texture)
  // A wave's samples cover one compact pixel tile, so neighbouring lanes mostly hit the same
  // triangle (vertex colours) or texel quad (texture): the per-sample terms are summed over
  // runs of equal target in the wave and each run issues one set of atomics.
  const bool runs = !a.coverage && (VCOL_LIKE || !a.uniform);
  for (int unit0 = blockIdx.x; unit0 < nunits; unit0 += gridDim.x) {
    const int unit = xcd_tile(unit0, nunits);   // runs of neighbouring units per XCD
    for (int j = 0; j < rt * rt; j++) {
      const int tx = (unit % units_x) * rt + j % rt, ty = (unit / units_x) * rt + j / rt;
      if (tx >= tiles_x || ty >= tiles_y) continue;   // whole tile outside the image
      const TileMap tm = tile_map_xy(sq, tx, ty, lane);
      const int x = tm.x, y = tm.y;
      const bool inside = x < c.w && y < c.h;
      const F3 adj = inside ? ld3(a.adj + c.img_off + 3ll * ((long long)y * c.w + x)) : F3{0, 0, 0};
      // one Hammersley sample per lane in the interleaved layout, all of them otherwise (the
      // trip count is the same in every lane)
      const int s_begin = tm.S > 1 ? tm.s : 0, s_end = tm.S > 1 ? tm.s + 1 : a.nsamp;
      for (int si = s_begin; si < s_end; si++) {
        const F3 val = divs_pow2(adj, (float)a.nsamp);
        long long key = -1 - lane;   // lanes without a contribution form runs of their own
        constexpr int NG = POS ? 18 : 12;
        float g[NG];
#pragma unroll
        for (int q = 0; q < NG; q++) g[q] = 0.0f;
        int o0 = 0, o1 = 0, o2 = 0, o3 = 0;   // offsets (in G) the four float3 terms go to
        int p0 = 0, p1 = 0, p2 = 0;           // POS: offsets of the three vertex positions
        F3 ro = {0, 0, 0}, rd = {0, 0, 1};   // this sample's camera ray
        float lit_stored = 0.0f;             // LAMBERT_VCOL with stored hits: lit_encode'd terms
        do {
          if (!inside) break;
          Hit h;
          if (a.hits) {
            // the render's stored closest hit of this (grid) sample
            const float4 r = a.hits[(c.img_off / 3 + (long long)y * c.w + x) * a.nsamp + si];
            h = Hit{__float_as_uint(r.x), r.w, r.y, r.z};
            lit_stored = r.w;
            if (a.coverage) { if (h.gtri != NO_HIT) uni = add(uni, val); break; }
          } else {
            float xo = a.hamm[2 * si], yo = a.hamm[2 * si + 1];
            if (a.aligned) sample_offset(a.jit, a.view, x, y, si, sq, xo, yo);
            eye_ray(c, x + xo, y + yo, ro, rd);
            const F3 o = ro, d = rd;
            if (a.coverage) {
              // 1x1 texture and tc == 0 everywhere: the bilinear weights are (1, 0, 0, 0) at any
              // hit, so the sample adds exactly `val` if the ray hits anything
              if (covered<COUNT, TF>(T, o, d, wst, x + xo, y + yo)) uni = add(uni, val);
              break;
            }
            h = trace<false, COUNT, TF>(T, o, d, 0.0f, 1e9f, wst, x + xo, y + yo);
          }
          if (h.gtri == NO_HIT) break;
          const int m = a.tr.s.tri_mesh[h.gtri];
          const int goff = a.layout[4 * m + 0];
          if (goff < 0) break;
          const int vo = a.tr.s.mesh_vo[m];
          if (VCOL_LIKE) {
            F3 vs = val;
            if (MODE == MODE_LAMBERT_VCOL) {
              const uint32_t* I3 = a.tr.s.idx + 3 * h.gtri;
              float shadow, ndl;
              if (a.hits) lit_decode(lit_stored, shadow, ndl);
              else lambert_terms<COUNT, TF>(T, a.light, ro, rd, h, I3[0], I3[1], I3[2], a.tr.s.nrm, wst, shadow, ndl);
              vs = mulv(val, lambert_factor(a.light, shadow, ndl));
            }
            const float w = 1.0f - h.u - h.v;
            const F3 cA = mul(w, vs), cB = mul(h.v, vs), cC = mul(h.u, vs);
            g[0] = cA.x; g[1] = cA.y; g[2] = cA.z;
            g[3] = cB.x; g[4] = cB.y; g[5] = cB.z;
            g[6] = cC.x; g[7] = cC.y; g[8] = cC.z;
            const int base = goff + a.layout[4 * m + 1];
            const uint32_t* I = a.tr.s.idx + 3 * h.gtri;
            const int gb = use_lds ? base - a.lds_base : base;
            o0 = gb + 3 * ((int)I[0] - vo); o1 = gb + 3 * ((int)I[1] - vo); o2 = gb + 3 * ((int)I[2] - vo);
            key = h.gtri;
            if (POS) {
              // dF/du = (c_C - c_A).val, dF/dv = (c_B - c_A).val; world-space du/dV, dv/dV,
              // then mesh 0's inverse transform as in edge_grad (src/drender.cpp:60-71)
              const F3 cAv = ld3(a.tr.s.col + 3 * I[0]), cBv = ld3(a.tr.s.col + 3 * I[1]), cCv = ld3(a.tr.s.col + 3 * I[2]);
              const float dFu = dot(sub(cCv, cAv), val), dFv = dot(sub(cBv, cAv), val);
              F3 du[3], dv[3];
              const bool ok = bary_grad(ro, rd, f4xyz(T.bf_a[h.gtri]), f4xyz(T.bf_e1[h.gtri]), f4xyz(T.bf_e2[h.gtri]),
                                        h.u, h.v, du, dv);
              const float* Ti = a.tr_inv0;
#pragma unroll
              for (int k = 0; k < 3; k++) {
                const F3 wgt = ok ? add(mul(dFu, du[k]), mul(dFv, dv[k])) : F3{0, 0, 0};
                g[9 + 3 * k + 0] = Ti[0] * wgt.x + Ti[4] * wgt.y + Ti[8] * wgt.z;
                g[9 + 3 * k + 1] = Ti[1] * wgt.x + Ti[5] * wgt.y + Ti[9] * wgt.z;
                g[9 + 3 * k + 2] = Ti[2] * wgt.x + Ti[6] * wgt.y + Ti[10] * wgt.z;
              }
              const int pb = use_lds ? goff - a.lds_base : goff;
              p0 = pb + 3 * ((int)I[0] - vo); p1 = pb + 3 * ((int)I[1] - vo); p2 = pb + 3 * ((int)I[2] - vo);
            }
          } else {
            const int w = a.tr.s.tex_w[m], hh = a.tr.s.tex_h[m];
            if (w == 0 || hh == 0) break;
            const uint32_t* I = a.tr.s.idx + 3 * h.gtri;
            const F2 tcv = blend2(a.tr.s.tc, I[0], I[1], I[2], h.u, h.v);
            // sample_bilinear_clamp_3f_grad (shade_common.cpp:23-45): tc not clamped
            const float px = tcv.x * (float)w, py = tcv.y * (float)hh;
            const int ix = (int)px, iy = (int)py;
            const int x0 = clampi(ix, 0, w - 1), y0 = clampi(iy, 0, hh - 1);
            const int x1 = clampi(ix + 1, 0, w - 1), y1 = clampi(iy + 1, 0, hh - 1);
            const float dx = px - (float)x0, dy = py - (float)y0;
            const float w00 = (1 - dx) * (1 - dy), w01 = (1 - dx) * dy, w10 = dx * (1 - dy), w11 = dx * dy;
            if (a.uniform) {
              uni.x += w00 * val.x + w01 * val.x + w10 * val.x + w11 * val.x;
              uni.y += w00 * val.y + w01 * val.y + w10 * val.y + w11 * val.y;
              uni.z += w00 * val.z + w01 * val.z + w10 * val.z + w11 * val.z;
              break;
            }
            const int base = goff + a.layout[4 * m + 2];
            const int gb = use_lds ? base - a.lds_base : base;
            o0 = gb + 3 * (y0 * w + x0); o1 = gb + 3 * (y1 * w + x0);
            o2 = gb + 3 * (y0 * w + x1); o3 = gb + 3 * (y1 * w + x1);
            g[0] = w00 * val.x; g[1] = w00 * val.y; g[2] = w00 * val.z;
            g[3] = w01 * val.x; g[4] = w01 * val.y; g[5] = w01 * val.z;
            g[6] = w10 * val.x; g[7] = w10 * val.y; g[8] = w10 * val.z;
            g[9] = w11 * val.x; g[10] = w11 * val.y; g[11] = w11 * val.z;
            // the quad's corner (o0) and whether its far corners clamp onto it fix all four
            key = 4ll * o0 + (x1 != x0 ? 2 : 0) + (y1 != y0 ? 1 : 0);
          }
        } while (false);
        if (runs) {
          const bool tail = seg_reduce<VCOL_LIKE ? (POS ? 18 : 9) : 12>(key, g);
          if (POS && tail && key >= 0) {
            const int op[3] = {p0, p1, p2};
            for (int q = 0; q < 3; q++) {
              const float* gv = g + 9 + 3 * q;
              if (nh > 0) hash_add(hkey, hval, nh, a.grad, op[q], gv[0], gv[1], gv[2]);
              else if (a.dacc) { for (int c3 = 0; c3 < 3; c3++) dadd(a.dacc, op[q] + c3, gv[c3]); }
              else {
                float* G = use_lds ? lgrad : a.grad;
                atomicAdd(G + op[q], gv[0]); atomicAdd(G + op[q] + 1, gv[1]); atomicAdd(G + op[q] + 2, gv[2]);
              }
            }
          }
          if (tail && key >= 0) {
            if (nh > 0) {
              hash_add(hkey, hval, nh, a.grad, o0, g[0], g[1], g[2]);
              hash_add(hkey, hval, nh, a.grad, o1, g[3], g[4], g[5]);
              hash_add(hkey, hval, nh, a.grad, o2, g[6], g[7], g[8]);
              if (!VCOL_LIKE) hash_add(hkey, hval, nh, a.grad, o3, g[9], g[10], g[11]);
            } else if (a.dacc) {
              const int o[4] = {o0, o1, o2, o3};
              for (int q = 0; q < (VCOL_LIKE ? 3 : 4); q++)
                for (int c3 = 0; c3 < 3; c3++) dadd(a.dacc, o[q] + c3, g[3 * q + c3]);
            } else {
              float* G = use_lds ? lgrad : a.grad;
              atomicAdd(G + o0, g[0]); atomicAdd(G + o0 + 1, g[1]); atomicAdd(G + o0 + 2, g[2]);
              atomicAdd(G + o1, g[3]); atomicAdd(G + o1 + 1, g[4]); atomicAdd(G + o1 + 2, g[5]);
              atomicAdd(G + o2, g[6]); atomicAdd(G + o2 + 1, g[7]); atomicAdd(G + o2 + 2, g[8]);
              if (!VCOL_LIKE) { atomicAdd(G + o3, g[9]); atomicAdd(G + o3 + 1, g[10]); atomicAdd(G + o3 + 2, g[11]); }
            }
          }
        }
      }
    }
    if (nh > 0) {
      // the region's sums: one set of atomics per target, then an empty table
      __syncthreads();
      for (int i = lane; i < nh; i += 64) {
        const int k = hkey[i];
        if (k >= 0) {
          atomicAdd(a.grad + k, hval[3 * i]); atomicAdd(a.grad + k + 1, hval[3 * i + 1]); atomicAdd(a.grad + k + 2, hval[3 * i + 2]);
          hkey[i] = -1; hval[3 * i] = 0.0f; hval[3 * i + 1] = 0.0f; hval[3 * i + 2] = 0.0f;
        }
      }
      __syncthreads();
    }
  }
  if (a.uniform) {
    uni.x = wave_sum(uni.x); uni.y = wave_sum(uni.y); uni.z = wave_sum(uni.z);
    if (lane == 0 && a.dacc) {
      dadd(a.dacc, a.uniform_off, uni.x); dadd(a.dacc, a.uniform_off + 1, uni.y); dadd(a.dacc, a.uniform_off + 2, uni.z);
    } else if (lane == 0) {
      atomicAdd(a.grad + a.uniform_off, uni.x);
      atomicAdd(a.grad + a.uniform_off + 1, uni.y);
      atomicAdd(a.grad + a.uniform_off + 2, uni.z);
    }
  }
  if (use_lds) {
    __syncthreads();
    for (int i = lane; i < a.lds_n; i += 64)
      if (lgrad[i] != 0.0f) atomicAdd(a.grad + a.lds_base + i, lgrad[i]);
  }
}

// The interior pass when its only output is a 1x1 albedo reached through texture
// coordinates that are all 0 (one differentiable mesh; TEXTURE_COLOR / LAMBERT): the bilinear
// weights are (1, 0, 0, 0) at any hit, so a sample adds exactly adj / nsamp to the albedo's
// gradient iff its ray hits anything (an any-hit ray, shade_grad<TEXTURE_COLOR>,
// src/shade_simple.cpp:132-162 with src/shade_common.cpp:23-45).  One wave per pixel tile as
// the render (one sample per lane), the wave's sum of its covered samples into part[tile];
// k_coverage_sum adds the partial sums in a fixed order, so the result does not depend on
// scheduling.  Small in registers, it keeps as many waves resident as the render.
template <bool COUNT = false, int TF = 0>
__global__ __launch_bounds__(64) void k_coverage(InteriorArgs a, float4* __restrict__ part) {
  __shared__ int wst[WST_INTS];
  const Tracer& T = a.tr;
  const DevCam& c = a.cams[a.view];
  const int lane = threadIdx.x;
  const int sq = a.sq;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const TileMap tm = tile_map(sq, c.w, tile, lane);
  const bool inside = tm.x < c.w && tm.y < c.h;
  F3 acc = {0, 0, 0};
  if (inside) {
    const F3 val = divs_pow2(ld3(a.adj + c.img_off + 3ll * ((long long)tm.y * c.w + tm.x)), (float)a.nsamp);
    const int s_begin = tm.S > 1 ? tm.s : 0, s_end = tm.S > 1 ? tm.s + 1 : a.nsamp;
    for (int si = s_begin; si < s_end; si++) {
      bool hit;
      if (a.hits) {
        hit = __float_as_uint(a.hits[(c.img_off / 3 + (long long)tm.y * c.w + tm.x) * a.nsamp + si].x) != NO_HIT;
      } else {
        float xo = a.hamm[2 * si], yo = a.hamm[2 * si + 1];
        if (a.aligned) sample_offset(a.jit, a.view, tm.x, tm.y, si, sq, xo, yo);
        F3 o, d;
        eye_ray(c, tm.x + xo, tm.y + yo, o, d);
        hit = covered<COUNT, TF>(T, o, d, wst, tm.x + xo, tm.y + yo);
      }
      if (hit) acc = add(acc, val);
    }
  }
  acc.x = wave_sum(acc.x); acc.y = wave_sum(acc.y); acc.z = wave_sum(acc.z);
  if (lane == 0) part[tile] = make_float4(acc.x, acc.y, acc.z, 0.0f);
}

// grad[0..2] += the sum of n partial float3 (part[i].xyz), in a fixed order, in double.
// Interior position term of LAMBERT / LAMBERT_VCOL (the opt-in counterpart of
// k_interior<VERTEX_COLOR, true>; the reference has no interior position gradient): a sample
// whose ray hits triangle (A, B, C) at barycentrics (u, v) sees the shade
//   F(u, v) = val . (albedo(u, v) * ((shadow * Lc) * max(0, n(u, v).(-L)) + ambient)),
// n = nA (1-u-v) + nB v + nC u the interpolated (prepared) normal and albedo the interpolated
// vertex colour (LAMBERT_VCOL) or the bilinear texel of the interpolated texture coordinate
// (LAMBERT, sample_bilinear_clamp_3f, src/shade_common.cpp).  Moving a vertex moves (u, v):
// dF/dV = dF/du du/dV + dF/dv dv/dV, with the closed-form du/dV, dv/dV of the ray's
// barycentrics (BarU_grad / BarV_grad, src/functions.cpp:226-470) and mesh 0's inverse
// transform as in edge_grad (src/drender.cpp:60-71).  The shadow is held fixed (it changes
// only where a triangle's facing to the light flips).  One wave per pixel tile as the render.
template <int MODE, bool COUNT = false, int TF = 0>
__global__ __launch_bounds__(64) void k_interior_pos(InteriorArgs a) {
  __shared__ int wst[WST_INTS];
  const Tracer& T = a.tr;
  const DevScene& S = a.tr.s;
  const DevCam& c = a.cams[a.view];
  const int lane = threadIdx.x;
  const int sq = a.sq;
  const TileMap tm = tile_map(sq, c.w, xcd_tile(blockIdx.x, gridDim.x), lane);
  if (!(tm.x < c.w && tm.y < c.h)) return;
  const F3 val = divs_pow2(ld3(a.adj + c.img_off + 3ll * ((long long)tm.y * c.w + tm.x)), (float)a.nsamp);
  const F3 negL = mul(-1.0f, a.light.L);
  const int s_begin = tm.S > 1 ? tm.s : 0, s_end = tm.S > 1 ? tm.s + 1 : a.nsamp;
  for (int si = s_begin; si < s_end; si++) {
    float xo = a.hamm[2 * si], yo = a.hamm[2 * si + 1];
    if (a.aligned) sample_offset(a.jit, a.view, tm.x, tm.y, si, sq, xo, yo);
    F3 ro, rd;
    eye_ray(c, tm.x + xo, tm.y + yo, ro, rd);
    Hit h;
    if (a.hits) {
      const float4 r = a.hits[(c.img_off / 3 + (long long)tm.y * c.w + tm.x) * a.nsamp + si];
      h = Hit{__float_as_uint(r.x), r.w, r.y, r.z};
    } else {
      h = trace<false, COUNT, TF>(T, ro, rd, 0.0f, 1e9f, wst, tm.x + xo, tm.y + yo);
    }
    if (h.gtri == NO_HIT) continue;
    const int m = S.tri_mesh[h.gtri];
    const int goff = a.layout[4 * m + 0];
    if (goff < 0) continue;
    const uint32_t* I = S.idx + 3 * h.gtri;
    float shadow, ndl;
    if (a.hits && MODE == MODE_LAMBERT_VCOL) lit_decode(h.t, shadow, ndl);   // stored hits keep the light terms
    else lambert_terms<COUNT, TF>(T, a.light, ro, rd, h, I[0], I[1], I[2], S.nrm, wst, shadow, ndl);
    const F3 light = lambert_factor(a.light, shadow, ndl);
    // d(light)/du, d(light)/dv: zero where max(0, n.(-L)) is clamped or the point is in shadow
    const F3 nA = ld3(S.nrm + 3 * I[0]), nB = ld3(S.nrm + 3 * I[1]), nC = ld3(S.nrm + 3 * I[2]);
    const bool on = shadow != 0.0f && ndl > 0.0f;
    const F3 dLu = on ? mul(dot(sub(nC, nA), negL) * shadow, a.light.Lc) : F3{0, 0, 0};
    const F3 dLv = on ? mul(dot(sub(nB, nA), negL) * shadow, a.light.Lc) : F3{0, 0, 0};
    F3 alb, dAu, dAv;
    if (MODE == MODE_LAMBERT_VCOL) {
      const F3 cA = ld3(S.col + 3 * I[0]), cB = ld3(S.col + 3 * I[1]), cC = ld3(S.col + 3 * I[2]);
      alb = blend3(cA, cB, cC, h.u, h.v);
      dAu = sub(cC, cA);
      dAv = sub(cB, cA);
    } else {
      const int w = S.tex_w[m], hh = S.tex_h[m];
      const F2 t0 = blend2(S.tc, I[0], I[1], I[2], h.u, h.v);
      alb = bilinear(S, m, t0);
      // gradient of the clamped bilinear lookup in tc, then tc's in (u, v)
      F3 gx = {0, 0, 0}, gy = {0, 0, 0};
      if (w > 0 && hh > 0) {
        const float tx = clampf(t0.x, 0.0f, 1.0f) * (float)w, ty = clampf(t0.y, 0.0f, 1.0f) * (float)hh;
        const int ix = (int)tx, iy = (int)ty;
        const int x0 = clampi(ix, 0, w - 1), y0 = clampi(iy, 0, hh - 1);
        const int x1 = clampi(ix + 1, 0, w - 1), y1 = clampi(iy + 1, 0, hh - 1);
        const float dx = tx - (float)x0, dy = ty - (float)y0;
        const F3 p00 = tex_fetch(S, m, x0, y0), p01 = tex_fetch(S, m, x0, y1);
        const F3 p10 = tex_fetch(S, m, x1, y0), p11 = tex_fetch(S, m, x1, y1);
        const bool inx = t0.x > 0.0f && t0.x < 1.0f, iny = t0.y > 0.0f && t0.y < 1.0f;
        if (inx) gx = mul((float)w, add(mul(1.0f - dy, sub(p10, p00)), mul(dy, sub(p11, p01))));
        if (iny) gy = mul((float)hh, add(mul(1.0f - dx, sub(p01, p00)), mul(dx, sub(p11, p10))));
      }
      const float tAx = S.tc[2 * I[0]], tAy = S.tc[2 * I[0] + 1];
      const float tBx = S.tc[2 * I[1]], tBy = S.tc[2 * I[1] + 1];
      const float tCx = S.tc[2 * I[2]], tCy = S.tc[2 * I[2] + 1];
      dAu = add(mul(tCx - tAx, gx), mul(tCy - tAy, gy));
      dAv = add(mul(tBx - tAx, gx), mul(tBy - tAy, gy));
    }
    const float dFu = dot(val, add(mulv(dAu, light), mulv(alb, dLu)));
    const float dFv = dot(val, add(mulv(dAv, light), mulv(alb, dLv)));
    F3 du[3], dv[3];
    if (!bary_grad(ro, rd, f4xyz(T.bf_a[h.gtri]), f4xyz(T.bf_e1[h.gtri]), f4xyz(T.bf_e2[h.gtri]), h.u, h.v, du, dv))
      continue;
    const float* Ti = a.tr_inv0;
    const int vo = S.mesh_vo[m];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const F3 wgt = add(mul(dFu, du[k]), mul(dFv, dv[k]));
      const float gxk = Ti[0] * wgt.x + Ti[4] * wgt.y + Ti[8] * wgt.z;
      const float gyk = Ti[1] * wgt.x + Ti[5] * wgt.y + Ti[9] * wgt.z;
      const float gzk = Ti[2] * wgt.x + Ti[6] * wgt.y + Ti[10] * wgt.z;
      const int o = goff + 3 * ((int)I[k] - vo);
      if (a.dacc) { dadd(a.dacc, o, gxk); dadd(a.dacc, o + 1, gyk); dadd(a.dacc, o + 2, gzk); }
      else { atomicAdd(a.grad + o, gxk); atomicAdd(a.grad + o + 1, gyk); atomicAdd(a.grad + o + 2, gzk); }
    }
  }
}

__global__ __launch_bounds__(1024) void k_coverage_sum(const float4* __restrict__ part, int n, float* grad) {
  __shared__ double s[3][1024];
  double x = 0.0, y = 0.0, z = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    const float4 p = part[i];
    x += p.x; y += p.y; z += p.z;
  }
  s[0][threadIdx.x] = x; s[1][threadIdx.x] = y; s[2][threadIdx.x] = z;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 3; k++) s[k][threadIdx.x] += s[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) grad[threadIdx.x] += (float)s[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// edge_derivatives (src/drender.h:235-425)
// ---------------------------------------------------------------------------
// VertexShader of every vertex (transform() already applied: pos is T*v), then the
// projected scene's identity prepare: I*(sx, sy, z, 1).
__device__ __forceinline__ float4 project_vertex(const DevCam& c, const float* pos, int k) {
  const F3 v = ld3(pos + 3 * k);
  float sx, sy;
  vertex_shader(c, v.x, v.y, v.z, sx, sy);
  const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  const F3 p = mp(I, F3{sx, sy, v.z});
  return make_float4(p.x, p.y, p.z, 0.0f);
}
// Projected vertices and, in the same launch, the edges' lengths -- the pmf before
// normalisation: float3 length in (screen x, screen y, world z) (build_edge_sampler,
// src/drender.cpp:13-15).  An edge projects its two endpoints itself (the same arithmetic as
// the vertex threads, so the same values).
__global__ void k_project(ProjectArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const DevCam& c = a.cams[a.view];
  if (k < a.V) a.proj[k] = project_vertex(c, a.pos, k);
  if (k < a.E) {
    const int2 v = a.edges[k];
    const float4 p0 = project_vertex(c, a.pos, v.x), p1 = project_vertex(c, a.pos, v.y);
    a.pmf_raw[k] = length(F3{p1.x - p0.x, p1.y - p0.y, p1.z - p0.z});
  }
}

// The reference's CDF is a sequential float running sum (src/drender.cpp:12-17); its
// rounding decides which edge a sample lands on, so it is reproduced exactly: chunks of
// pmf are staged in LDS by the whole block and one lane carries the running sum.  Then
// pmf and cdf are divided by the total (:18-20).
__global__ __launch_bounds__(1024) void k_cdf(const float* pmf_raw, int E, float* pmf, float* cdf) {
  __shared__ float buf[CDF_CHUNK];
  __shared__ float total_s;
  float run = 0.0f;
  for (int start = 0; start < E; start += CDF_CHUNK) {
    const int n = min(CDF_CHUNK, E - start);
    for (int i = threadIdx.x; i < n; i += 1024) buf[i] = pmf_raw[start + i];
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 0; i < n; i++) { run = buf[i] + run; buf[i] = run; }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 1024) cdf[start + 1 + i] = buf[i];
    __syncthreads();
  }
  if (threadIdx.x == 0) total_s = run;
  __syncthreads();
  const float total = total_s;
  for (int i = threadIdx.x; i < E; i += 1024) pmf[i] = pmf_raw[i] / total;
  for (int i = threadIdx.x; i <= E; i += 1024) cdf[i] = (i == 0 ? 0.0f : cdf[i]) / total;
}

// The same sequential float running sum, computed in parallel and exactly.
// Chunk j (CDF_CH edges) starts from a speculated value s_j = float(double prefix) and runs
// the float recurrence c <- p + c over its range.  If the true start T_j differs, the true
// sequence is the speculated one shifted by d = T_j - s_j whenever every speculated value
// stays in one binade (one grid spacing u), no addition was a rounding tie (TwoSum error
// exactly u/2; ties-to-even depends on parity, which an odd shift flips), and the shifted
// end stays in the binade: round-to-nearest commutes with adding a grid multiple.  One
// wavefront walks the chunks applying shifts, 64 chunks per round (replaying a chunk where
// the conditions fail, from LDS); then every chunk is recomputed from its true start and
// written out normalised by the total.
//   k_cdf_sums   double sum per chunk            k_cdf_scan   speculated starts
//   k_cdf_spec   speculated ends, shift test     k_cdf_walk   true starts (64 per round)
//   k_cdf_final  cdf / pmf normalised
__device__ __forceinline__ int fexp(float x) { return (int)((__float_as_uint(x) >> 23) & 0xFFu); }

struct CdfWork {
  double* csum;   // NCH
  float* s0;      // NCH speculated starts
  float* se;      // NCH speculated ends
  int* ok;        // NCH shift-invariant
  float* ts;      // NCH + 1 true starts, ts[NCH] = total
};
__host__ __device__ inline int cdf_chunks(int E) { return (E + CDF_CH - 1) / CDF_CH; }
size_t cdf_work_bytes(int E) {
  const size_t n = (size_t)cdf_chunks(E);
  return 8 * n + 4 * n * 3 + 4 * (n + 1) + 64;
}
static CdfWork cdf_work(void* w, int E) {
  const size_t n = (size_t)cdf_chunks(E);
  char* p = (char*)w;
  CdfWork c;
  c.csum = (double*)p; p += 8 * n;
  c.s0 = (float*)p; p += 4 * n;
  c.se = (float*)p; p += 4 * n;
  c.ok = (int*)p; p += 4 * n;
  c.ts = (float*)p;
  return c;
}

// CDF_CPB chunks of src into t[chunk][*] by the workgroup's CDF_STAGE_T threads, coalesced,
// every load of a thread issued before its LDS stores; lane L of wave 0 then walks row L
// (stride CDF_CH + 1 floats: conflict-free)
constexpr int CDF_STAGE_T = 256;
__device__ __forceinline__ void cdf_stage(const float* __restrict__ src, int E, int chunk0,
                                          float (*t)[CDF_CH + 1]) {
  constexpr int PER = CDF_CPB * CDF_CH / CDF_STAGE_T;
  const int base = chunk0 * CDF_CH;
  float v[PER];
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int k = base + (int)threadIdx.x + i * CDF_STAGE_T;
    v[i] = k < E ? src[k] : 0.0f;
  }
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int q = (int)threadIdx.x + i * CDF_STAGE_T;
    t[q / CDF_CH][q % CDF_CH] = v[i];
  }
}

__global__ __launch_bounds__(CDF_STAGE_T) void k_cdf_sums(const float* __restrict__ raw, int E, CdfWork w) {
  __shared__ float t[CDF_CPB][CDF_CH + 1];
  cdf_stage(raw, E, blockIdx.x * CDF_CPB, t);
  __syncthreads();
  const int j = blockIdx.x * CDF_CPB + threadIdx.x;
  if (threadIdx.x >= CDF_CPB || j >= cdf_chunks(E)) return;
  const int n = min(CDF_CH, E - j * CDF_CH);
  double s = 0.0;
  for (int q = 0; q < n; q++) s += (double)t[threadIdx.x][q];
  w.csum[j] = s;
}

__global__ __launch_bounds__(1024) void k_cdf_scan(int E, CdfWork w) {
  __shared__ double part[1024];
  const int n = cdf_chunks(E);
  const int per = (n + 1023) / 1024;
  const int b = min((int)threadIdx.x * per, n), e = min(b + per, n);
  double s = 0.0;
  for (int j = b; j < e; j++) s += w.csum[j];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const double v = (int)threadIdx.x >= o ? part[threadIdx.x - o] : 0.0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  double run = threadIdx.x == 0 ? 0.0 : part[threadIdx.x - 1];
  for (int j = b; j < e; j++) {
    w.s0[j] = j == 0 ? 0.0f : (float)run;
    run += w.csum[j];
  }
}

__global__ __launch_bounds__(CDF_STAGE_T) void k_cdf_spec(const float* __restrict__ raw, int E, CdfWork w) {
  __shared__ float t[CDF_CPB][CDF_CH + 1];
  cdf_stage(raw, E, blockIdx.x * CDF_CPB, t);
  __syncthreads();
  const int j = blockIdx.x * CDF_CPB + threadIdx.x;
  if (threadIdx.x >= CDF_CPB || j >= cdf_chunks(E)) return;
  const int n = min(CDF_CH, E - j * CDF_CH);
  const float s0 = w.s0[j];
  const int ex = fexp(s0);
  const float u = __uint_as_float((uint32_t)max(ex - 23, 1) << 23);  // grid spacing of the binade
  float c = s0;
  int good = 1;
  for (int q = 0; q < n; q++) {
    const float p = t[threadIdx.x][q];
    const float sum = p + c;
    // TwoSum: err = (p + c) - sum exactly
    const float bb = sum - c;
    const float aa = sum - bb;
    const float err = (p - bb) + (c - aa);
    if (fabsf(err) * 2.0f == u && err != 0.0f) good = 0;
    c = sum;
  }
  // one binade for the whole chunk; tiny (grid below 2^-126) and non-finite values replay
  if (fexp(c) != ex || ex <= 24 || ex == 255) good = 0;
  w.se[j] = c;
  w.ok[j] = good;
}

// Serial float sum T = x + T over the len (<= CDF_CH = 128) values lanes hold in v0 (first
// 64) and v1 (next 64), in order; T and the result are the same in every lane.
static_assert(CDF_CH == 128, "chunk_sum reads two values per lane");
__device__ __forceinline__ float chunk_sum(float T, float v0, float v1, int len) {
  const int a0 = __float_as_int(v0), a1 = __float_as_int(v1);
  if (len == CDF_CH) {
#pragma unroll
    for (int k = 0; k < 64; k++) T = __int_as_float(__builtin_amdgcn_readlane(a0, k)) + T;
#pragma unroll
    for (int k = 0; k < 64; k++) T = __int_as_float(__builtin_amdgcn_readlane(a1, k)) + T;
  } else {
    for (int k = 0; k < len; k++) T = __int_as_float(__builtin_amdgcn_readlane(k < 64 ? a0 : a1, k & 63)) + T;
  }
  return T;
}

// Inclusive prefix sum over the 64 lanes of a fully active wavefront (mod 2^32): row shifts
// within each 16-lane row, then row broadcasts of lanes 15 and 31.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// True chunk starts.  The chunks that are not shift-invariant are staged in LDS first (one
// wavefront per chunk).  Wave 0 then resolves up to 64 chunks per round.  With
// D_k = bits(T_k) - bits(s0_k) (true and speculated start of chunk k as integers), a chunk
// that takes a fast path -- T_k == s0_k, or a shift-invariant chunk whose true start lies in
// s0_k's binade and whose shifted end stays there -- ends at bits(se_k) + D_k, so
//   D_{k+1} = D_k + bits(se_k) - bits(s0_{k+1}),
// and a wave prefix sum of those differences gives every start up to the round's first chunk
// that fails both tests; that chunk is summed serially from its true start and the next
// round begins after it.  The starts are bit-identical to the chunk-by-chunk walk.
__global__ __launch_bounds__(1024) void k_cdf_walk(const float* __restrict__ raw, int E, CdfWork w) {
  constexpr int MAXST = CDF_WALK_STAGE / CDF_CH;   // staged chunks
  __shared__ int scnt[1024];
  __shared__ int slot_chunk[MAXST];
  __shared__ float stage[CDF_WALK_STAGE];
  __shared__ float ls0[CDF_WALK_CHUNKS], lse[CDF_WALK_CHUNKS];
  __shared__ int lok[CDF_WALK_CHUNKS];
  const int n = cdf_chunks(E);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool lds_spec = n <= CDF_WALK_CHUNKS;
  if (lds_spec)
    for (int j = tid; j < n; j += 1024) { ls0[j] = w.s0[j]; lse[j] = w.se[j]; lok[j] = w.ok[j]; }
  const float* S0 = lds_spec ? ls0 : w.s0;
  const float* SE = lds_spec ? lse : w.se;
  const int* OK = lds_spec ? lok : w.ok;
  __syncthreads();
  // rank of every chunk that is not shift-invariant, in chunk order
  const int per = (n + 1023) / 1024;
  const int b = min(tid * per, n), e = min(b + per, n);
  int cnt = 0;
  for (int j = b; j < e; j++) cnt += OK[j] ? 0 : 1;
  scnt[tid] = cnt;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = tid >= o ? scnt[tid - o] : 0;
    __syncthreads();
    scnt[tid] += v;
    __syncthreads();
  }
  int r = scnt[tid] - cnt;
  for (int j = b; j < e; j++)
    if (!OK[j]) { if (r < MAXST) slot_chunk[r] = j; r++; }
  const int nst = min(scnt[1023], MAXST);
  __syncthreads();
  for (int q = wave; q < nst; q += 16) {
    const int j = slot_chunk[q];
    const int len = min(CDF_CH, E - j * CDF_CH);
    for (int k = lane; k < len; k += 64) stage[q * CDF_CH + k] = raw[(size_t)j * CDF_CH + k];
  }
  __syncthreads();
  if (wave != 0) return;
  float T = 0.0f;
  int rank = 0;   // chunks before j0 that are not shift-invariant (= LDS slot of the next one)
  for (int j0 = 0; j0 < n;) {
    const int j = j0 + lane;
    const bool in = j < n;
    const uint32_t uss = in ? __float_as_uint(S0[j]) : 0u, use = in ? __float_as_uint(SE[j]) : 0u;
    const bool okj = in && OK[j] != 0;
    // exclusive prefix over lanes of bits(se_k) - bits(s0_{k+1}), mod 2^32.  Up to the first
    // slow chunk every start and speculated start is a float with the sign bit clear, so the
    // true D_k lies in (-2^31, 2^31) and the wrapped sum read as signed is exact.
    const uint32_t dl = use - (uint32_t)__shfl_down((int)uss, 1, 64);
    const uint32_t pre = wave_incl_scan(dl) - dl;
    const uint32_t tbits = __float_as_uint(T);
    const long long D = (long long)(int32_t)(tbits - (uint32_t)__builtin_amdgcn_readfirstlane((int)uss) + pre);
    const int bx = (int)((uss >> 23) & 0xFF);
    const bool nan = bx == 255 && (uss & 0x7FFFFFu) != 0;
    const long long tb = (long long)uss + D, teb = (long long)use + D;   // true start / shifted end
    const bool sane = (tbits >> 31) == 0 && (uss >> 31) == 0 && (use >> 31) == 0;
    const bool fast = in && sane && ((D == 0 && !nan) || (okj && (tb >> 23) == bx && (teb >> 23) == bx));
    const uint64_t inm = __ballot(in), slow = inm & ~__ballot(fast), notok = inm & ~__ballot(okj);
    const int L = __popcll(inm);
    const int k = slow ? __ffsll((long long)slow) - 1 : L;
    const float start = lane == 0 ? T : __uint_as_float((uint32_t)tb);
    if (in && lane <= k) w.ts[j] = start;
    if (k == L) {
      T = __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)teb, L - 1));
      rank += __popcll(notok);
      j0 += L;
      continue;
    }
    // chunk j0 + k: summed from its true start
    const int jr = j0 + k;
    T = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(start), k));
    const uint64_t before = (1ull << k) - 1ull;      // k <= 63
    const int slot = rank + __popcll(notok & before);
    const bool staged = ((notok >> k) & 1ull) && slot < MAXST;
    rank = slot + (int)((notok >> k) & 1ull);
    const int len = min(CDF_CH, E - jr * CDF_CH);
    const float* src = staged ? stage + slot * CDF_CH : raw + (size_t)jr * CDF_CH;
    const float v0 = lane < len ? src[lane] : 0.0f, v1 = 64 + lane < len ? src[64 + lane] : 0.0f;
    T = chunk_sum(T, v0, v1, len);
    j0 = jr + 1;
  }
  if (lane == 0) w.ts[n] = T;
}

__global__ __launch_bounds__(CDF_STAGE_T) void k_cdf_final(const float* __restrict__ raw, int E, CdfWork w,
                                                       float* __restrict__ pmf, float* __restrict__ cdf) {
  __shared__ float t[CDF_CPB][CDF_CH + 1];
  const int chunk0 = blockIdx.x * CDF_CPB;
  cdf_stage(raw, E, chunk0, t);
  __syncthreads();
  const int j = chunk0 + threadIdx.x;
  const float total = w.ts[cdf_chunks(E)];
  if (threadIdx.x < CDF_CPB && j < cdf_chunks(E)) {
    const int n = min(CDF_CH, E - j * CDF_CH);
    float c = w.ts[j];
    for (int q = 0; q < n; q++) { c = t[threadIdx.x][q] + c; t[threadIdx.x][q] = c; }
  }
  __syncthreads();
  const int base = chunk0 * CDF_CH;
  for (int q = threadIdx.x; q < CDF_CPB * CDF_CH; q += CDF_STAGE_T) {
    const int k = base + q;
    if (k < E) {
      cdf[k + 1] = t[q / CDF_CH][q % CDF_CH] / total;
      pmf[k] = raw[k] / total;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) cdf[0] = 0.0f / total;
}

// sample(): clamp(upper_bound(cdf, u) - 1, 0, E - 1) (src/drender.cpp:25-29)
__device__ __forceinline__ int sample_edge(const float* cdf, int E, float u) {
  int first = 0, count = E + 1;
  while (count > 0) {
    const int step = count >> 1;
    const int it = first + step;
    if (!(u < cdf[it])) { first = it + 1; count -= step + 1; }
    else count = step;
  }
  return clampi(first - 1, 0, E - 1);
}

// Sample point of sample (rnd0 -> edge, rnd1 = t) and the reference's rejection tests
// (src/drender.h:338-345): off-screen pixel (int() truncation; NaN and out-of-range values
// are off-screen as x86 cvttss2si's INT_MIN) and |adjoint| < 1e-3.
struct EdgePoint { int2 ev; float4 P0, P1; float t, px, py; int xi, yi; };
__device__ __forceinline__ void edge_point(const int2* edges, const float4* proj, int eid, float t, EdgePoint& p) {
  p.ev = edges[eid];
  p.P0 = proj[p.ev.x]; p.P1 = proj[p.ev.y];
  p.t = t;
  p.px = p.P0.x + t * (p.P1.x - p.P0.x); p.py = p.P0.y + t * (p.P1.y - p.P0.y);
  p.xi = 0; p.yi = 0;
}
__device__ __forceinline__ bool edge_keep(const DevCam& c, const float* adj, EdgePoint& p) {
  if (p.px > -1.0f && p.py > -1.0f && p.px < (float)c.w && p.py < (float)c.h) {
    p.xi = (int)p.px; p.yi = (int)p.py;
    if (p.xi >= 0 && p.yi >= 0 && p.xi < c.w && p.yi < c.h) {
      const F3 aj = ld3(adj + c.img_off + 3ll * ((long long)p.yi * c.w + p.xi));
      return !((double)length(aj) < 1e-3);
    }
  }
  return false;
}

// Bin every kept sample: its edge, bin and rank within the bin.  Slots in rnd0 order sample
// a few consecutive edges per wavefront, so lanes form runs of equal bin that take their
// ranks with one counter increment per run (rnd0 order kept within a run).
__global__ __launch_bounds__(256) void k_edge_bin(EdgeBinArgs a) {
  const DevCam& c = a.cams[a.view];
  const int lane = threadIdx.x & 63;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long wbase = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); wbase < a.N; wbase += stride) {
    const long long slot = wbase + lane;
    int bin = -1 - lane, eid = 0;   // rejected and past-the-end lanes: runs of their own
    if (slot < a.N) {
      const float2 r = a.rnd[slot];
      eid = sample_edge(a.cdf, a.E, r.x);
      EdgePoint p;
      edge_point(a.edges, a.proj, eid, r.y, p);
      if (edge_keep(c, a.adj, p)) bin = (p.yi / EDGE_BIN) * a.bins_x + p.xi / EDGE_BIN;
    }
    const int prev = __shfl_up(bin, 1, 64);
    const bool head = lane == 0 || prev != bin;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __clzll((long long)(heads & upto));
    const uint64_t later = lane == 63 ? 0ull : (heads >> (lane + 1));
    const int end = later ? lane + __ffsll((long long)later) : 64;   // one past the run
    int base = 0;
    if (head && bin >= 0) base = atomicAdd(&a.bin_count[bin], end - lane);
    base = __shfl(base, start, 64);
    if (slot < a.N) a.rec[slot] = make_int4(eid, bin, base + lane - start, 0);
  }
}

// Bin counts -> exclusive starts in place; cnt[n] = total.
__global__ __launch_bounds__(1024) void k_bin_scan(int* cnt, int n) {
  __shared__ int part[1024];
  const int per = (n + 1023) / 1024;
  const int b = min((int)threadIdx.x * per, n), e = min(b + per, n);
  int s = 0;
  for (int j = b; j < e; j++) s += cnt[j];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = (int)threadIdx.x >= o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = part[threadIdx.x] - s;
  for (int j = b; j < e; j++) { const int x = cnt[j]; cnt[j] = run; run += x; }
  if (threadIdx.x == 1023) cnt[n] = part[1023];
}

__global__ void k_edge_scatter(EdgeBinArgs a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const int4 r = a.rec[i];
  if (r.y >= 0) a.order[a.bin_count[r.y] + r.z] = make_int2((int)i, r.x);
}

// Edge pass, part 1: the colour on either side of every kept sample (shade at the sample
// point offset by -/+1e-3 px along the edge normal, src/drender.h:347-360).  A wavefront takes
// 32 samples with both sides in adjacent lanes, one ray (and shadow ray) per lane, so its
// packet covers 32 nearby points and the kernel stays small enough for 8 waves per SIMD.
// ecol[i] = (colour inside - colour outside, edge id bits; -1 for a rejected sample).
template <int MODE, bool COUNT = false, int TF = 0>
__global__ __launch_bounds__(256) void k_edge_trace(EdgeArgs a) {
  __shared__ int wstack[4][WST_INTS];
  const Tracer& T = a.tr;
  const DevCam& c = a.cams[a.view];
  const int lane = threadIdx.x & 63;
  int* wst = wstack[threadIdx.x >> 6];
  const long long items = a.order ? (long long)__builtin_amdgcn_readfirstlane(*a.nkept) : a.N;
  const long long stride = (long long)gridDim.x * (blockDim.x / 2);
  for (long long wbase = ((long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) / 2; wbase < items; wbase += stride) {
    const long long item = wbase + (lane >> 1);
    const int sd = lane & 1;
    if (item >= items) continue;
    EdgePoint p;
    int eid;
    bool live;
    if (a.order) {
      const int2 o = a.order[item];   // binned samples passed the rejection tests
      eid = o.y;
      edge_point(a.edges, a.proj, eid, a.rnd[o.x].y, p);
      live = true;
    } else {
      const float2 r = a.rnd[item];
      eid = sample_edge(a.cdf, a.E, r.x);
      edge_point(a.edges, a.proj, eid, r.y, p);
      live = edge_keep(c, a.adj, p);
    }
    F3 col = {0.0f, 0.0f, 0.0f};
    if (live) {
      const float dvx = p.P1.x - p.P0.x, dvy = p.P1.y - p.P0.y;
      const float len = sqrtf(dvx * dvx + dvy * dvy);
      const float ndx = dvx / len, ndy = dvy / len;
      const float nx = -ndy, ny = ndx;  // normal2D (src/utils.h:29)
      const float qx = sd == 0 ? p.px - 1e-3f * nx : p.px + 1e-3f * nx;
      const float qy = sd == 0 ? p.py - 1e-3f * ny : p.py + 1e-3f * ny;
      col = shade<MODE, COUNT, TF>(T, a.light, c, qx, qy, a.tr.s.nrm2d, wst);
    }
    // both lanes of a sample are here together (same item, same `live`)
    const float ox = __shfl_xor(col.x, 1, 64), oy = __shfl_xor(col.y, 1, 64), oz = __shfl_xor(col.z, 1, 64);
    if (sd == 0) a.ecol[item] = make_float4(col.x - ox, col.y - oy, col.z - oz, __int_as_float(live ? eid : -1));
  }
}

// Edge pass, part 2: each kept sample's position / transform gradient from its two side
// colours (src/drender.h:361-425, edge_grad src/drender.cpp:31-87).
__global__ __launch_bounds__(256) void k_edge_grad(EdgeArgs a) {
  extern __shared__ float lpos[];  // a.lds_n floats (dynamic)
  __shared__ float ltr[4][12];
  const DevCam& c = a.cams[a.view];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool use_lds = a.lds_n > 0;
  for (int i = threadIdx.x; i < a.lds_n; i += 256) lpos[i] = 0.0f;
  if (threadIdx.x < 48) ltr[threadIdx.x / 12][threadIdx.x % 12] = 0.0f;
  __syncthreads();
  const float* Ti = a.tr_inv0;  // mesh 0's inverse transform for every mesh (drender.cpp:60)
  const long long stride = (long long)gridDim.x * blockDim.x;
  // Items are the slots in rnd0 order (a wavefront's 64 slots sample a few consecutive edges)
  // or the kept samples in pixel-bin order.  Whole wavefronts step together (the run
  // reduction needs all 64 lanes).
  const long long items = a.order ? (long long)__builtin_amdgcn_readfirstlane(*a.nkept) : a.N;
  for (long long wbase = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); wbase < items; wbase += stride) {
    const long long item = wbase + lane;
    // position gradient of v0, v1 (after mesh 0's inverse transform), then for a single
    // mesh the raw screen-chain terms of v0, v1 that its transform gradient is built from
    float gsum[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int key = -1 - lane;   // lanes without a contribution form runs of their own
    int vn0 = 0, vn1 = 0, goff = -1;
    bool live = false;
    int eid = 0, xi = 0, yi = 0;
    float t = 0.0f;
    float4 P0 = {0, 0, 0, 0}, P1 = {0, 0, 0, 0};
    int2 ev = {0, 0};
    F3 dcol = {0, 0, 0};
    if (item < items) {
      const float4 c0 = a.ecol[item];
      eid = __float_as_int(c0.w);
      if (eid >= 0) {
        // the sample point exactly as the trace pass computed it
        EdgePoint p;
        edge_point(a.edges, a.proj, eid, a.rnd[a.order ? a.order[item].x : item].y, p);
        p.xi = (int)p.px; p.yi = (int)p.py;
        live = true;
        ev = p.ev; P0 = p.P0; P1 = p.P1; t = p.t; xi = p.xi; yi = p.yi;
        dcol = F3{c0.x, c0.y, c0.z};
      }
    }
    if (live) {
    const float pmf = a.pmf[eid];
    const F3 adj = ld3(a.adj + c.img_off + 3ll * ((long long)yi * c.w + xi));
    const float dvx = P1.x - P0.x, dvy = P1.y - P0.y;
    const float len = sqrtf(dvx * dvx + dvy * dvy);
    const float ndx = dvx / len, ndy = dvy / len;
    const float nx = -ndy, ny = ndx;  // normal2D (src/utils.h:29)
    const float pdf = pmf / len;
    const float weight = 1.0f / (pdf * float(a.N));
    const float adjv = dot(dcol, adj);   // dot(cin - cout, adj)
    if (adjv * adjv > 0.0f) {
    const float d0x = ((1 - t) * nx) * adjv * weight, d0y = ((1 - t) * ny) * adjv * weight;
    const float d1x = (t * nx) * adjv * weight, d1y = (t * ny) * adjv * weight;
    // edge_grad (src/drender.cpp:31-87): the reference keeps the diagonal of the screen
    // Jacobian (dSx/dVx, dSy/dVy, and both for z), exact only for views whose screen axes are
    // world x and y; the full-Jacobian mode adds the cross terms dSy/dVx and dSx/dVy
    const F3 V0 = ld3(a.tr.s.pos + 3 * ev.x), V1 = ld3(a.tr.s.pos + 3 * ev.y);
    const F3 g0x = vs_grad(c, V0, 0), g0y = vs_grad(c, V0, 1);
    const F3 g1x = vs_grad(c, V1, 0), g1y = vs_grad(c, V1, 1);
    float dv0_dx = g0x.x * d0x, dv0_dy = g0y.y * d0y;
    float dv1_dx = g1x.x * d1x, dv1_dy = g1y.y * d1y;
    const float dv0_dz = (g0x.z * d0x + g0y.z * d0y);
    const float dv1_dz = (g1x.z * d1x + g1y.z * d1y);
    if (a.full_jacobian) {
      dv0_dx = dv0_dx + g0y.x * d0y; dv0_dy = g0x.y * d0x + dv0_dy;
      dv1_dx = dv1_dx + g1y.x * d1y; dv1_dy = g1x.y * d1x + dv1_dy;
    }
    const int m = a.edge_mesh[eid];
    goff = a.layout[4 * m + 0];
    if (goff >= 0) {
    const int vo = a.tr.s.mesh_vo[m];
    vn0 = ev.x - vo; vn1 = ev.y - vo;
    key = eid;
    gsum[0] = Ti[0] * dv0_dx + Ti[4] * dv0_dy + Ti[8] * dv0_dz;
    gsum[1] = Ti[1] * dv0_dx + Ti[5] * dv0_dy + Ti[9] * dv0_dz;
    gsum[2] = Ti[2] * dv0_dx + Ti[6] * dv0_dy + Ti[10] * dv0_dz;
    gsum[3] = Ti[0] * dv1_dx + Ti[4] * dv1_dy + Ti[8] * dv1_dz;
    gsum[4] = Ti[1] * dv1_dx + Ti[5] * dv1_dy + Ti[9] * dv1_dz;
    gsum[5] = Ti[2] * dv1_dx + Ti[6] * dv1_dy + Ti[10] * dv1_dz;
    if (a.single_mesh) {
      gsum[6] = dv0_dx; gsum[7] = dv0_dy; gsum[8] = dv0_dz;
      gsum[9] = dv1_dx; gsum[10] = dv1_dy; gsum[11] = dv1_dz;
    } else {
      // d_tr (src/drender.cpp:73-86): orig_vertices times the screen-chain terms
      const F3 o0 = ld3(a.tr.s.pos_orig + 3 * ev.x), o1 = ld3(a.tr.s.pos_orig + 3 * ev.y);
      const float tv[12] = {o0.x * dv0_dx + o1.x * dv1_dx, o0.y * dv0_dx + o1.y * dv1_dx,
                            o0.z * dv0_dx + o1.z * dv1_dx, 1 * dv0_dx + 1 * dv1_dx,
                            o0.x * dv0_dy + o1.x * dv1_dy, o0.y * dv0_dy + o1.y * dv1_dy,
                            o0.z * dv0_dy + o1.z * dv1_dy, 1 * dv0_dy + 1 * dv1_dy,
                            o0.x * dv0_dz + o1.x * dv1_dz, o0.y * dv0_dz + o1.y * dv1_dz,
                            o0.z * dv0_dz + o1.z * dv1_dz, 1 * dv0_dz + 1 * dv1_dz};
      if (a.dacc) {
        const int to = a.trscr ? a.dG + 12 * m : goff + a.layout[4 * m + 3];
        for (int q = 0; q < 12; q++) dadd(a.dacc, to + q, tv[q]);
      } else {
        float* gt = a.trscr ? a.trscr + 12 * m : a.grad + goff + a.layout[4 * m + 3];
#pragma unroll
        for (int q = 0; q < 12; q++) atomicAdd(gt + q, tv[q]);
      }
    }
    }  // differentiable mesh
    }  // adj != 0
    }  // accepted sample
    // one atomic per run of equal edges for the position gradient of its two vertices
    bool tail;
    if (a.single_mesh) tail = seg_reduce<12>(key, gsum);
    else tail = seg_reduce<6>(key, gsum);
    if (tail && key >= 0) {
      if (a.dacc) {
        for (int c3 = 0; c3 < 3; c3++) {
          dadd(a.dacc, goff + 3 * vn0 + c3, gsum[c3]);
          dadd(a.dacc, goff + 3 * vn1 + c3, gsum[3 + c3]);
        }
      } else {
      float* G = use_lds ? lpos : a.grad;
      const int gb = use_lds ? goff - a.lds_base : goff;
      atomicAdd(G + gb + 3 * vn0, gsum[0]); atomicAdd(G + gb + 3 * vn0 + 1, gsum[1]); atomicAdd(G + gb + 3 * vn0 + 2, gsum[2]);
      atomicAdd(G + gb + 3 * vn1, gsum[3]); atomicAdd(G + gb + 3 * vn1 + 1, gsum[4]); atomicAdd(G + gb + 3 * vn1 + 2, gsum[5]);
      }
      if (a.single_mesh) {
        // the run's transform gradient: orig_vertices of its two (shared) vertices times the
        // run's summed screen-chain terms (src/drender.cpp:73-86), into the wave's LDS row
        const F3 o0 = ld3(a.tr.s.pos_orig + 3 * ev.x), o1 = ld3(a.tr.s.pos_orig + 3 * ev.y);
#pragma unroll
        for (int r = 0; r < 3; r++) {
          const float d0 = gsum[6 + r], d1 = gsum[9 + r];
          const float t4[4] = {o0.x * d0 + o1.x * d1, o0.y * d0 + o1.y * d1, o0.z * d0 + o1.z * d1, 1 * d0 + 1 * d1};
          if (a.dacc) {
            const int to = a.trscr ? a.dG + 12 * a.single_id : a.single_tr_off;
            for (int c4 = 0; c4 < 4; c4++) dadd(a.dacc, to + 4 * r + c4, t4[c4]);
          } else {
            for (int c4 = 0; c4 < 4; c4++) atomicAdd(&ltr[wave][4 * r + c4], t4[c4]);
          }
        }
      }
    }
  }
  __syncthreads();
  if (a.single_mesh && threadIdx.x < 12) {
    const float v = ltr[0][threadIdx.x] + ltr[1][threadIdx.x] + ltr[2][threadIdx.x] + ltr[3][threadIdx.x];
    if (v != 0.0f) atomicAdd((a.trscr ? a.trscr + 12 * a.single_id : a.grad + a.single_tr_off) + threadIdx.x, v);
  }
  if (use_lds)
    for (int i = threadIdx.x; i < a.lds_n; i += 256)
      if (lpos[i] != 0.0f) atomicAdd(a.grad + a.lds_base + i, lpos[i]);
}

// ---------------------------------------------------------------------------
// optimisers
// ---------------------------------------------------------------------------
// OptSimple::OptStep, GD_Adam (src/optimizer_simple.cpp:208-216) on one mesh block.
__global__ void k_optsimple_moments(const float* grad, float* m, float* G, int off, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = grad[off + i];
  m[off + i] = m[off + i] * 0.25f + g * (1.0f - 0.25f);
  G[off + i] = 2.0f * (G[off + i] * 0.5f + (g * g) * (1.0f - 0.5f));
}
// Normalisation (:223-224; the reference indexes m_GSquare without the mesh offset),
// per-block learning rate (:227-232), then OptUpdateScene's p -= step (:92-186).
__global__ void k_optsimple_apply(OptApplyArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float step = a.m[a.off + i] / sqrtf(a.G[i] + float(1e-8f));
  int seg = 0;
  for (int s = 1; s < a.nseg; s++) if (i >= a.seg_start[s]) seg = s;
  step = step * a.seg_lr[seg];
  float* p = a.seg_param[seg];
  if (p) p[i - a.seg_start[seg]] -= step;
}

// Standard Adam (Kingma & Ba) with bias correction.
__global__ void k_adam(AdamArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  int seg = 0;
  for (int s = 1; s < a.nseg; s++) if (i >= a.seg_start[s]) seg = s;
  float* p = a.seg_param[seg];
  if (!p) return;
  const float g = a.grad[a.off + i];
  const float m = a.b1 * a.m[a.off + i] + (1.0f - a.b1) * g;
  const float v = a.b2 * a.v[a.off + i] + (1.0f - a.b2) * g * g;
  a.m[a.off + i] = m;
  a.v[a.off + i] = v;
  const int t = *a.step;
  const float bc1 = (float)(1.0 - pow((double)a.b1, t)), bc2 = (float)(1.0 - pow((double)a.b2, t));
  const float mh = m / bc1, vh = v / bc2;
  p[i - a.seg_start[seg]] -= a.seg_lr[seg] * mh / (sqrtf(vh) + a.eps);
}

__global__ void k_transform_update(TransformArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_meshes || a.layout[4 * i] < 0) return;
  float* mt = a.mats + 48 * i;
  M4 T = m4_load(mt);
  if (a.rflag && a.rflag[i]) {
    T = m4_restricted(a.rparams + 7 * i);
  } else {
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 4; c++) T.at(r, c) = a.rows[12 * i + 4 * r + c];
  }
  const M4 NT = m4_transpose(m4_inverse(T));
  for (int k = 0; k < 16; k++) { mt[k] = T.m[k]; mt[16 + k] = NT.m[k]; }
}
__global__ void k_restricted_grad(RestrictedArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_meshes || a.layout[4 * i] < 0) return;
  const float* d = a.scr + 12 * i;
  float* g = a.grad + a.layout[4 * i] + a.layout[4 * i + 3];
  if (a.rflag[i]) {
    float J[7][12];
    restricted_jacobian(a.rparams + 7 * i, J);
    for (int k = 0; k < 7; k++) {
      float acc = 0.0f;
      for (int l = 0; l < 12; l++) acc += J[k][l] * d[l];
      g[k] += acc;
    }
  } else {
    for (int l = 0; l < 12; l++) g[l] += d[l];
  }
}
void launch_restricted_grad(const RestrictedArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_restricted_grad, dim3((a.n_meshes + 63) / 64), dim3(64), 0, s, a);
}
void launch_transform_update(const TransformArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_transform_update, dim3((a.n_meshes + 63) / 64), dim3(64), 0, s, a);
}

__global__ void k_step_inc(int* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) step[0] += 1;
}
void launch_step_inc(int* step, hipStream_t s) { hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, s, step); }

// ---------------------------------------------------------------------------
// debug: closest hit at pixel centres
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_debug_hits(HitsArgs a) {
  __shared__ int wst[WST_INTS];
  const Tracer& T = a.tr;
  const DevCam& c = a.cams[0];
  const int lane = threadIdx.x;
  const int tiles_x = (c.w + 7) >> 3;
  const int x = (blockIdx.x % tiles_x) * 8 + (lane & 7);
  const int y = (blockIdx.x / tiles_x) * 8 + (lane >> 3);
  if (x >= c.w || y >= c.h) return;
  F3 o, d;
  eye_ray(c, x + 0.5f, y + 0.5f, o, d);
  const Hit h = trace<false>(T, o, d, 0.0f, 1e9f, wst, x + 0.5f, y + 0.5f);
  const int p = y * c.w + x;
  uint32_t prim = NO_HIT, geom = NO_HIT;
  if (h.gtri != NO_HIT) { geom = (uint32_t)a.tr.s.tri_mesh[h.gtri]; prim = h.gtri - (uint32_t)a.tr.s.mesh_to[geom]; }
  a.prim[p] = prim; a.geom[p] = geom; a.t[p] = h.t; a.u[p] = h.u; a.v[p] = h.v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define DRH_DISPATCH_MODE(MODE_VAR, BODY)          \
  switch (MODE_VAR) {                              \
    case 1: { constexpr int M = 1; BODY; } break;  \
    case 2: { constexpr int M = 2; BODY; } break;  \
    case 3: { constexpr int M = 3; BODY; } break;  \
    case MODE_LAMBERT_VCOL: { constexpr int M = MODE_LAMBERT_VCOL; BODY; } break; \
    default: { constexpr int M = 4; BODY; } break; \
  }
void launch_fill_f32(float* p, long long n, float v, hipStream_t s) {
  if (n <= 0) return;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_f32, dim3(blocks), dim3(256), 0, s, p, n, v);
}
void launch_scatter(const ScatterArgs& a, hipStream_t s) {
  if (a.n <= 0) return;
  long long most = 0;
  for (int g = 0; g < a.n; g++) most = std::max(most, a.words[g]);
  const int blocks = (int)std::max<long long>(1, std::min<long long>((most + 255) / 256, 1024));
  hipLaunchKernelGGL(k_scatter, dim3(blocks, a.n), dim3(256), 0, s, a);
}
void launch_copy_f32(float* d, const float* src, long long n, hipStream_t s) {
  // n is a multiple of 4 floats (16 B) for the calibration buffers
  const long long n4 = n / 4;
  if (n4 <= 0) return;
  const int blocks = (int)std::min<long long>((n4 + 255) / 256, 8192);
  hipLaunchKernelGGL(k_copy_f32, dim3(blocks), dim3(256), 0, s, (float4*)d, (const float4*)src, n4);
}
void launch_fill_u32(uint32_t* p, long long n, uint32_t v, hipStream_t s) {
  if (n <= 0) return;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, n, v);
}
void launch_prepare(const PrepareArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_prepare, dim3((a.V + 255) / 256), dim3(256), 0, s, a);
}
void launch_tri_setup(const TriSetupArgs& a, hipStream_t s) {
  // F + 1 threads at least: the last counter is cleared too
  hipLaunchKernelGGL(k_tri_setup, dim3((a.F + 1 + 255) / 256), dim3(256), 0, s, a);
}
void launch_morton(const MortonArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_morton, dim3((a.F + 255) / 256), dim3(256), 0, s, a);
}
void launch_radix_pass(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout, int n, int shift,
                       uint32_t* hist, hipStream_t s) {
  const int nb = (n + RADIX_TILE - 1) / RADIX_TILE;
  hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(256), 0, s, kin, n, shift, hist, nb);
  hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, s, hist, 256 * nb);
  hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(256), 0, s, kin, vin, kout, vout, n, shift, hist, nb);
}
void launch_karras(const KarrasArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_karras, dim3((a.F + 254) / 256), dim3(256), 0, s, a);
}
void launch_refit(const RefitArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_refit_local, dim3((a.F + REFIT_CHUNK - 1) / REFIT_CHUNK), dim3(REFIT_CHUNK), 0, s, a);
  // one chunk holds the whole tree up to its root: no node spans chunks, the frontier is empty
  if (a.F > REFIT_CHUNK) hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(1024), 0, s, a);
}
void launch_span(const SpanArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, a.span_n, 1ll, 0u);
  hipLaunchKernelGGL(k_span_mark, dim3((a.F + 254) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_span_link, dim3((SPAN_MAX + 255) / 256), dim3(256), 0, s, a);
}
void launch_pack_reorder(const PackArgs& pk, const ReorderArgs& ro, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_nodes, dim3((ro.F + 255) / 256), dim3(256), 0, s, pk, ro);
}
void launch_suspects(const SuspectArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_suspects, dim3((a.F + 255) / 256), dim3(256), 0, s, a);
}
void launch_sus_bin(const SusBinArgs& a, int views, hipStream_t s) {
  const dim3 grid(std::max(1, std::min(1024, (a.F + 1023) / 1024)), views);   // waves loop over the suspects
  hipLaunchKernelGGL(k_sus_bin<0>, grid, dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_bin_offsets, dim3(views), dim3(1024), 0, s, a.bp, a.cell_stride, (int)a.cell_stride);
  hipLaunchKernelGGL(k_sus_bin<1>, grid, dim3(256), 0, s, a);
}
void launch_light_bin(const LightBinArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_light_bin<0>, dim3(256), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_bin_offsets, dim3(1), dim3(1024), 0, s, a.bp, 0ll, a.bp.cells_x * a.bp.cells_y);
  hipLaunchKernelGGL(k_light_bin<1>, dim3(256), dim3(256), 0, s, a);
}
// The ray-counting (COUNT) and watertight instantiations walk the LBVH as packets only: a
// walk changes no hit and no ray count.
void launch_render(int mode, const RenderArgs& a, int tiles, hipStream_t s) {
  if (a.tr.rays) DRH_DISPATCH_MODE(mode, hipLaunchKernelGGL((k_render<M, true>), dim3(tiles), dim3(64), 0, s, a))
  else if (a.tr.s.lane_walk) DRH_DISPATCH_MODE(mode, hipLaunchKernelGGL((k_render<M, false, TF_LANE>), dim3(tiles), dim3(64), 0, s, a))
  else DRH_DISPATCH_MODE(mode, hipLaunchKernelGGL((k_render<M>), dim3(tiles), dim3(64), 0, s, a))
}
void launch_loss_final(const double* part, int n, float* out, double denom, hipStream_t s) {
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(1024), 0, s, part, n, out, denom);
}
void launch_interior(int mode, const InteriorArgs& a, int blocks, hipStream_t s) {
  const size_t lds = sizeof(float) * (size_t)std::max(a.lds_n, 4 * a.hash);
  const bool cnt = a.tr.rays != nullptr;   // the ray-counting instantiations
  const bool lane = !cnt && a.tr.s.lane_walk;
  if (mode == 2 && a.pos_grad) {
    if (cnt) hipLaunchKernelGGL((k_interior<2, true, true>), dim3(blocks), dim3(64), lds, s, a);
    else if (lane) hipLaunchKernelGGL((k_interior<2, true, false, TF_LANE>), dim3(blocks), dim3(64), lds, s, a);
    else hipLaunchKernelGGL((k_interior<2, true>), dim3(blocks), dim3(64), lds, s, a);
  } else if (mode == 2) {
    if (cnt) hipLaunchKernelGGL((k_interior<2, false, true>), dim3(blocks), dim3(64), lds, s, a);
    else if (lane) hipLaunchKernelGGL((k_interior<2, false, false, TF_LANE>), dim3(blocks), dim3(64), lds, s, a);
    else hipLaunchKernelGGL((k_interior<2>), dim3(blocks), dim3(64), lds, s, a);
  } else if (mode == MODE_LAMBERT_VCOL) {
    if (cnt) hipLaunchKernelGGL((k_interior<MODE_LAMBERT_VCOL, false, true>), dim3(blocks), dim3(64), lds, s, a);
    else if (lane) hipLaunchKernelGGL((k_interior<MODE_LAMBERT_VCOL, false, false, TF_LANE>), dim3(blocks), dim3(64), lds, s, a);
    else hipLaunchKernelGGL((k_interior<MODE_LAMBERT_VCOL>), dim3(blocks), dim3(64), lds, s, a);
  } else {
    if (cnt) hipLaunchKernelGGL((k_interior<3, false, true>), dim3(blocks), dim3(64), lds, s, a);
    else if (lane) hipLaunchKernelGGL((k_interior<3, false, false, TF_LANE>), dim3(blocks), dim3(64), lds, s, a);
    else hipLaunchKernelGGL((k_interior<3>), dim3(blocks), dim3(64), lds, s, a);
  }
}
void launch_coverage(const InteriorArgs& a, int tiles, float4* part, hipStream_t s) {
  if (tiles <= 0) return;
  if (a.tr.rays) hipLaunchKernelGGL((k_coverage<true>), dim3(tiles), dim3(64), 0, s, a, part);
  else if (a.tr.s.lane_walk) hipLaunchKernelGGL((k_coverage<false, TF_LANE>), dim3(tiles), dim3(64), 0, s, a, part);
  else hipLaunchKernelGGL((k_coverage<false>), dim3(tiles), dim3(64), 0, s, a, part);
  hipLaunchKernelGGL(k_coverage_sum, dim3(1), dim3(1024), 0, s, (const float4*)part, tiles, a.grad + a.uniform_off);
}
void launch_interior_pos(int mode, const InteriorArgs& a, int tiles, hipStream_t s) {
  if (tiles <= 0) return;
  if (mode == MODE_LAMBERT_VCOL) {
    if (a.tr.rays) hipLaunchKernelGGL((k_interior_pos<MODE_LAMBERT_VCOL, true>), dim3(tiles), dim3(64), 0, s, a);
    else if (a.tr.s.lane_walk) hipLaunchKernelGGL((k_interior_pos<MODE_LAMBERT_VCOL, false, TF_LANE>), dim3(tiles), dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_interior_pos<MODE_LAMBERT_VCOL, false>), dim3(tiles), dim3(64), 0, s, a);
  } else {
    if (a.tr.rays) hipLaunchKernelGGL((k_interior_pos<4, true>), dim3(tiles), dim3(64), 0, s, a);
    else if (a.tr.s.lane_walk) hipLaunchKernelGGL((k_interior_pos<4, false, TF_LANE>), dim3(tiles), dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_interior_pos<4, false>), dim3(tiles), dim3(64), 0, s, a);
  }
}
void launch_project(const ProjectArgs& a, hipStream_t s) {
  const int n = std::max(a.V, a.E);
  if (n > 0) hipLaunchKernelGGL(k_project, dim3((n + 255) / 256), dim3(256), 0, s, a);
}
void launch_cdf(const float* pmf_raw, int E, float* pmf, float* cdf, void* work, hipStream_t s, bool parallel) {
  if (E <= 0) return;
  if (!parallel || E <= CDF_SERIAL_MAX) {   // a short sum: one workgroup, one launch
    hipLaunchKernelGGL(k_cdf, dim3(1), dim3(1024), 0, s, pmf_raw, E, pmf, cdf);
    return;
  }
  const CdfWork w = cdf_work(work, E);
  const int blocks = (cdf_chunks(E) + CDF_CPB - 1) / CDF_CPB;
  hipLaunchKernelGGL(k_cdf_sums, dim3(blocks), dim3(CDF_STAGE_T), 0, s, pmf_raw, E, w);
  hipLaunchKernelGGL(k_cdf_scan, dim3(1), dim3(1024), 0, s, E, w);
  hipLaunchKernelGGL(k_cdf_spec, dim3(blocks), dim3(CDF_STAGE_T), 0, s, pmf_raw, E, w);
  hipLaunchKernelGGL(k_cdf_walk, dim3(1), dim3(1024), 0, s, pmf_raw, E, w);
  hipLaunchKernelGGL(k_cdf_final, dim3(blocks), dim3(CDF_STAGE_T), 0, s, pmf_raw, E, w, pmf, cdf);
}

// The exact CDF's work buffer for callers that cannot allocate (drh_build_edge_sampler_noalloc):
// a file-scope device array, one copy per GPU, held until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in for the symbol): the
// device address comes from hipGetSymbolAddress, once per device with that device current.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static_assert(8 * ((NOALLOC_MAX_EDGES + CDF_CH - 1) / CDF_CH) + 12 * ((NOALLOC_MAX_EDGES + CDF_CH - 1) / CDF_CH) +
                      4 * ((NOALLOC_MAX_EDGES + CDF_CH - 1) / CDF_CH + 1) + 64 <= NOALLOC_SCRATCH_BYTES,
              "the CDF work buffer of the largest edge count fits the scratch array");

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch_ptr[device] = p;
  if (cur != device) (void)hipSetDevice(cur);
  return e;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

void launch_edge_bin(const EdgeBinArgs& a, hipStream_t s) {
  if (a.N <= 0) return;
  const int blocks = (int)std::max<long long>(1, std::min<long long>((a.N + 255) / 256, 4096));
  hipLaunchKernelGGL(k_edge_bin, dim3(blocks), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, a.bin_count, a.nbins);
  hipLaunchKernelGGL(k_edge_scatter, dim3((unsigned)((a.N + 255) / 256)), dim3(256), 0, s, a);
}
void launch_edge_trace(int mode, const EdgeArgs& a, hipStream_t s) {
  const int blocks = (int)std::max<long long>(1, std::min<long long>((2 * a.N + 255) / 256, 8192));
  if (a.watertight) {
    if (a.tr.rays) DRH_DISPATCH_MODE(mode, hipLaunchKernelGGL((k_edge_trace<M, true, TF_WT>), dim3(blocks), dim3(256), 0, s, a))
    else DRH_DISPATCH_MODE(mode, hipLaunchKernelGGL((k_edge_trace<M, false, TF_WT>), dim3(blocks), dim3(256), 0, s, a))
  } else if (a.tr.rays) {
    DRH_DISPATCH_MODE(mode, hipLaunchKernelGGL((k_edge_trace<M, true>), dim3(blocks), dim3(256), 0, s, a))
  } else if (a.tr.s.lane_walk) {
    DRH_DISPATCH_MODE(mode, hipLaunchKernelGGL((k_edge_trace<M, false, TF_LANE>), dim3(blocks), dim3(256), 0, s, a))
  } else {
    DRH_DISPATCH_MODE(mode, hipLaunchKernelGGL((k_edge_trace<M>), dim3(blocks), dim3(256), 0, s, a))
  }
}
void launch_edge_grad(const EdgeArgs& a, hipStream_t s) {
  const int blocks = (int)std::max<long long>(1, std::min<long long>((a.N + 255) / 256, 2048));
  hipLaunchKernelGGL(k_edge_grad, dim3(blocks), dim3(256), a.lds_n * sizeof(float), s, a);
}
void launch_optsimple(const float* grad, float* m, float* G, int off, const OptApplyArgs& a, hipStream_t s) {
  const int b = (a.n + 255) / 256;
  hipLaunchKernelGGL(k_optsimple_moments, dim3(b), dim3(256), 0, s, grad, m, G, off, a.n);
  hipLaunchKernelGGL(k_optsimple_apply, dim3(b), dim3(256), 0, s, a);
}
void launch_adam(const AdamArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
}
void launch_debug_hits(const HitsArgs& a, int tiles, hipStream_t s) {
  hipLaunchKernelGGL(k_debug_hits, dim3(tiles), dim3(64), 0, s, a);
}

}  // namespace drh
