// drh_api.cpp -- host runtime behind the C ABI (include/drh.h).
//
// Owns one HIP stream and the device buffers of a committed scene; orders the commit
// (prepare + LBVH build), render (+ fused loss), interior and edge kernels the way
// DiffRender<SHADING_MODEL> orders its loops (src/drender.h:74-434), and the optimiser
// step of OptSimple (src/optimizer_simple.cpp:188-309).
//
// Host float math (camera inverses, transforms) is built with -ffp-contract=off and keeps
// LiteMath's term order, so the matrices the kernels receive are the ones the reference
// computes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/drh.h"
#include "drh_internal.h"
#include "drh_kernels.h"
#include "drh_mat4.h"

using namespace drh;

namespace {

// ---------------------------------------------------------------------------
// sample sequences on the host
// ---------------------------------------------------------------------------
// The reference's edge-sample streams: `omp parallel for num_threads(14)` with a static
// schedule gives thread t the contiguous range [t*q + min(t, r), +q + (t < r)), q = N/14,
// r = N%14; each thread draws (rnd0, rnd1) per sample from RandomGenInit(7777 + t*t + 1)
// (src/drender.h:306-331; common_src/qmc.h:81-87, common_src/qmc.cpp:250-268).
void prng14_table(long long N, std::vector<float>& out) {
  out.resize(2 * (size_t)N);
  const long long T = 14, q = N / T, r = N % T;
  for (long long t = 0; t < T; t++) {
    const uint32_t seed = (uint32_t)(7777 + t * t + 1);
    uint32_t sx = seed * (seed * seed * 15731u + 74323u) + 871483u;
    uint32_t sy = seed * (seed * seed * 13734u + 37828u) + 234234u;
    auto next = [&]() { const uint32_t x = sx * 17u + sy * 13123u; sx = (x << 13) ^ x; sy ^= (x << 7); return x; };
    for (uint32_t i = 0; i < seed % 7; i++) next();
    auto rnd = [&]() { const uint32_t x = next(); const uint32_t v = x * (x * x * 15731u + 74323u) + 871483u; return (float)v * (1.0f / 4294967296.0f); };
    const long long start = t * q + std::min(t, r), cnt = q + (t < r ? 1 : 0);
    for (long long k = 0; k < cnt; k++) {
      const float a = rnd();
      const float b = rnd();
      out[2 * (start + k)] = a;
      out[2 * (start + k) + 1] = b;
    }
  }
}

// Niederreiter base-2 generator (Bratley, Fox & Niederreiter, ACM TOMACS 2(3), 1992), the
// construction of qmc::init (common_src/qmc.cpp:14-187): dimension d uses the d-th
// irreducible GF(2) polynomial in numeric order; each block of deg(p) output bits comes
// from the linear recurrence with characteristic polynomial p^k.  table[d][bit] keeps
// bits 32..62 of the 63-bit generator column.
void niederreiter_table(uint32_t table[11 * 31]) {
  std::vector<unsigned> irr;
  for (unsigned p = 2; irr.size() < 11; p++) {
    const int dp = 31 - __builtin_clz(p);
    bool ok = true;
    for (unsigned q : irr) {
      const int dq = 31 - __builtin_clz(q);
      if (2 * dq > dp) break;
      unsigned rem = p;
      for (int k = dp; k >= dq; k--) if (rem >> k & 1u) rem ^= q << (k - dq);
      if (!rem) { ok = false; break; }
    }
    if (ok) irr.push_back(p);
  }
  for (int d = 0; d < 11; d++) {
    const unsigned P = irr[d];
    const int e = 31 - __builtin_clz(P);
    std::vector<uint8_t> b{1};             // b(x) = p(x)^k, coefficients highest degree first
    std::vector<uint8_t> v(63 + e, 0);
    uint64_t col[63] = {0};
    for (int r = 0; r < 63; r++) {
      const int u = r % e;
      if (u == 0) {
        const size_t m1 = b.size() - 1;
        std::vector<uint8_t> nb(m1 + e + 1, 0);
        for (size_t i = 0; i <= m1; i++)
          if (b[i]) for (int k = 0; k <= e; k++) nb[i + k] ^= (uint8_t)((P >> (e - k)) & 1u);
        b.swap(nb);
        const size_t m = b.size() - 1;
        for (size_t i = 0; i < 63 + (size_t)e - 1; i++) {
          if (i < m1) v[i] = 0;
          else if (i < m) v[i] = 1;
          else { uint8_t s = 0; for (size_t k = 1; k <= m; k++) s ^= v[i - k] & b[k]; v[i] = s; }
        }
      }
      for (int i = 0; i < 63; i++) col[i] |= (uint64_t)v[i + u] << (62 - r);
    }
    for (int bit = 0; bit < 31; bit++) table[d * 31 + bit] = (uint32_t)((col[bit] >> 32) & 0x7FFFFFFFu);
  }
}

// qmc::rndFloat (common_src/qmc.cpp:190-197)
float qmc_rnd_host(uint32_t pos, int dim, const uint32_t* table) {
  uint32_t r = 0, d = pos;
  for (int bit = 0; bit < 31; bit++, d >>= 1)
    if (d & 1) r ^= table[bit + dim * 31];
  return (float)(r + 1) * (1.0f / 2147483648.0f);
}

// The (rnd0, rnd1) pairs of sample indices 0..N-1 in either mode, ordered by rnd0 (ties by
// index).  The sampled edge is a non-decreasing function of rnd0 (upper_bound in the CDF), so
// in this order consecutive slots sample consecutive edges: the edge kernel's wavefronts trace
// nearby rays and reduce each run of equal edges before one atomic, with no per-step sort.
// A sample's contribution depends on its (rnd0, rnd1) only, so the order changes nothing but
// the summation order of the gradient.  Tables are cached per (mode, N) for the process.
std::shared_ptr<const std::vector<float>> sorted_sample_table(int mode, long long N) {
  static std::mutex mu;
  static std::map<std::pair<int, long long>, std::shared_ptr<const std::vector<float>>> cache;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({mode, N});
    if (it != cache.end()) return it->second;
  }
  std::vector<float> raw;
  if (mode == DRH_RNG_REF_PRNG14) {
    prng14_table(N, raw);
  } else {
    uint32_t tab[11 * 31];
    niederreiter_table(tab);
    raw.resize(2 * (size_t)N);
    for (long long i = 0; i < N; i++) {
      raw[2 * i] = qmc_rnd_host((uint32_t)i, 0, tab);
      raw[2 * i + 1] = qmc_rnd_host((uint32_t)i, 1, tab);
    }
  }
  // rnd0 is in [0, 1]: its bit pattern orders like its value
  std::vector<uint64_t> key((size_t)N);
  for (long long i = 0; i < N; i++) {
    uint32_t b;
    std::memcpy(&b, &raw[2 * i], 4);
    key[i] = ((uint64_t)b << 32) | (uint64_t)i;
  }
  std::sort(key.begin(), key.end());
  auto out = std::make_shared<std::vector<float>>(2 * (size_t)N);
  for (long long k = 0; k < N; k++) {
    const uint64_t i = key[k] & 0xFFFFFFFFull;
    (*out)[2 * k] = raw[2 * i];
    (*out)[2 * k + 1] = raw[2 * i + 1];
  }
  std::lock_guard<std::mutex> lk(mu);
  return cache.emplace(std::make_pair(mode, N), out).first->second;
}

// planeHammersley (common_src/qmc.cpp:227-243)
void hammersley(int n, std::vector<float>& out) {
  out.resize(2 * (size_t)std::max(n, 1));
  for (int k = 0; k < n; k++) {
    float u = 0;
    int kk = k;
    for (float p = 0.5f; kk; p *= 0.5f, kk >>= 1) if (kk & 1) u += p;
    out[2 * k] = u;
    out[2 * k + 1] = (k + 0.5f) / n;
  }
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
// Buffers built as members of a handle belong to it (its BufOwner base is constructed before
// the members and names itself in t_buf_owner):
//  - a buffer that moves bumps its owner's reallocation epoch: a captured graph holds raw
//    pointers into the handle's buffers, so it is replayed only while the epoch it was
//    captured in still holds;
//  - once the handle's stream exists, its buffers come from the device's stream-ordered pool
//    (hipMallocAsync / hipFreeAsync on that stream): freeing them does not wait for the whole
//    device, which matters when a server tears jobs down while others run.
// Other buffers use hipMalloc / hipFree and count into a process-wide epoch.
std::atomic<unsigned long long> g_realloc_epoch{0};

struct BufOwner;
thread_local BufOwner* t_buf_owner = nullptr;

struct BufOwner {
  std::atomic<unsigned long long> realloc_epoch{0};
  hipStream_t pool_stream = nullptr;
  BufOwner() { t_buf_owner = this; }
};

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  BufOwner* owner = t_buf_owner;
  bool pooled = false;
  ~DBuf() { release(); }
  void release() {
    if (!p) return;
    if (pooled) (void)hipFreeAsync(p, owner->pool_stream);   // after the stream's work
    else (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t reserve(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (owner) owner->realloc_epoch++; else g_realloc_epoch++;
    release();
    if (n == 0) return hipSuccess;
    hipStream_t st = owner ? owner->pool_stream : nullptr;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (st) (void)hipStreamIsCapturing(st, &cs);
    hipError_t e;
    if (st && cs == hipStreamCaptureStatusNone) {
      // ordered on the handle's stream: its copies and kernels follow on that stream, and the
      // second stream starts from an event recorded there later
      e = hipMallocAsync(&p, n, st);
      pooled = true;
    } else {
      e = hipMalloc(&p, n);
      pooled = false;
    }
    if (e == hipSuccess) bytes = n; else p = nullptr;
    return e;
  }
  template <class T> T* as() const { return (T*)p; }
};

// Pooled memory stays with the process once freed (reused by the next allocation) instead
// of going back to the driver at every synchronisation.
void keep_pool_memory(int device) {
  static std::mutex m;
  static std::vector<int> done;
  std::lock_guard<std::mutex> lk(m);
  if (std::find(done.begin(), done.end(), device) != done.end()) return;
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
    uint64_t thr = ~0ull;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
  }
  done.push_back(device);
}

// Pinned host blocks (the handles' staging buffers).  hipHostFree waits for the whole
// device, every stream's queued work included, which stalls a process serving many handles
// for milliseconds per free; freed blocks are therefore cached by size (powers of two from
// 64 KiB) for the next handle that needs one, up to PINNED_CACHE_MAX bytes in all.
constexpr size_t PINNED_CACHE_MAX = 512ull << 20;
std::mutex g_pinned_mu;
std::multimap<size_t, char*> g_pinned_free;
size_t g_pinned_cached = 0;

size_t pinned_class(size_t need) {
  size_t cap = 64 << 10;
  while (cap < need) cap *= 2;
  return cap;
}
char* pinned_take(size_t need, size_t* cap) {
  *cap = pinned_class(need);
  {
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    auto it = g_pinned_free.find(*cap);
    if (it != g_pinned_free.end()) {
      char* p = it->second;
      g_pinned_free.erase(it);
      g_pinned_cached -= *cap;
      return p;
    }
  }
  char* p = nullptr;
  if (hipHostMalloc((void**)&p, *cap, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}
// p must have no transfer in flight
void pinned_give(char* p, size_t cap) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    if (g_pinned_cached + cap <= PINNED_CACHE_MAX) {
      g_pinned_free.emplace(cap, p);
      g_pinned_cached += cap;
      return;
    }
  }
  (void)hipHostFree(p);
}

struct ProfEvent { std::string name; hipEvent_t a, b; };

}  // namespace

struct drh_renderer : BufOwner {
  drh_settings st{};
  int sq = 1, nsamp = 1;
  hipStream_t stream = nullptr;
  // second stream (default priority) for the edge sampler's preparation of view 0, which
  // depends on the geometry only and so runs beside the render; fork / join by events
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool prep0_forked = false;
  // the interior pass runs on the second stream beside the edge pass (fork after the
  // gradient is cleared, join before its final conversions)
  hipEvent_t ev_ifork = nullptr, ev_ijoin = nullptr;
  // with view 0's edge preparation forked, the fused step also clears the gradient on the
  // second stream before the render (grad_zeroed) and reduces the loss there after it
  // (loss_on_side: fork after the render, join at the end of d_render), so the render is
  // followed directly by the edge pass on the main stream
  hipEvent_t ev_lfork = nullptr, ev_ljoin = nullptr;
  bool grad_zeroed = false, loss_on_side = false, loss_side_pending = false;
  int sharing = 1;          // handles sharing the device (drh_set_device_sharing)
  // interior samples: Hammersley (reference) or the render's grid (aligned); store: the render
  // keeps each sample's closest hit and the interior pass reads it back
  int interior_aligned = 0, interior_store = 0;   // interior_store: DRH_HITS_*
  // pixel-sample jitter (drh_set_pixel_jitter): the pattern of seed + step, where step is 0
  // outside the fused step, the host's step number for OptSimple steps and the device's Adam
  // step counter for Adam steps (so a captured step draws a fresh pattern on every replay)
  int jit_on = 0;
  uint32_t jit_seed = 0, jit_host_step = 0;
  const uint32_t* jit_dev_step = nullptr;
  Jitter jitter() const { return Jitter{jit_seed + jit_host_step, jit_on, jit_on ? jit_dev_step : nullptr}; }
  size_t hit_limit = ~(size_t)0;                  // DRH_HITS_AUTO stores only within this many bytes
  // Free a buffer unless the handle's stream is being captured (the free would become a node
  // of the captured step); the eager step before any capture has freed it already.
  void release_uncaptured(DBuf& b) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(stream, &cs);
    if (cs == hipStreamCaptureStatusNone) b.release();
  }
  // bytes of stored hits a d_render_and_compare of `views` views needs
  size_t hit_bytes(int views) const {
    size_t px = 0;
    for (int v = 0; v < views; v++) px += (size_t)cam_cache[v].w * cam_cache[v].h;
    return 16 * px * (size_t)nsamp;
  }
  int edge_order = DRH_EDGE_ORDER_AUTO;   // sample order of the edge pass
  int edge_grad_mode = DRH_EDGE_GRAD_REFERENCE_DIAGONAL;
  int side_test = DRH_SIDE_TEST_REFERENCE;   // ray/triangle test of the edge samples' side rays
  DBuf bin_count, bin_rec, bin_order;     // pixel-bin order of the kept edge samples
  DBuf ecol;                              // edge samples' side colour differences (float4 each)
  DBuf cov_part;                          // coverage interior pass: per-tile partial sums
  DBuf hitbuf;
  bool hits_stored = false;
  std::string err;
  Light light{};
  bool committed = false;

  // scene (host view)
  int nm = 0, V = 0, F = 0, G = 0, n_diff = 0;
  std::vector<int> vo, to, tex_w, tex_h, tex_off, diff, layout;  // layout: 4 per mesh
  std::vector<float> mats;      // 48 per mesh
  std::vector<float> tr_inv0;   // 16
  std::vector<uint32_t> idx_local_cache;
  std::vector<int> diff_cache;
  int E = 0;
  bool has_nrm = false, has_col = false;

  // device scene
  DBuf pos_orig, pos, col, tc, nrm_in, nrm, nrm2d, texels, idx, tri_mesh, mesh_vo, mesh_to, dtex_w, dtex_h,
      dtex_off, dmats, dlayout, dtr_inv0, tr_rows, drflag, drparams, trscr, dacc;
  bool deterministic = false;   // fixed-point gradient sums (drh_set_deterministic)
  bool interior_pos = false;    // interior position term in VERTEX_COLOR (drh_set_interior_pos_grad)
  // RESTRICTED_TRANSFORM meshes: flag and the 7 parameters per mesh
  std::vector<int> rflag;
  std::vector<float> rparams_h;
  bool any_restricted = false;
  int tr_block(int i) const { return rflag[i] ? 7 : 12; }
  // LBVH
  DBuf bf_a, bf_e1, bf_e2, box_lo, box_hi, cen, bounds, keys0, vals0, keys1, vals1, hist, child, node_parent,
      leaf_parent, counter, nrange, frontier, leaf_lo, leaf_hi, node_lo, node_hi, nodes, tri_a, tri_e1, tri_e2;
  // the spanning nodes of the current topology (k_span_mark / k_span_link at each full build)
  DBuf span, span_cid, span_of, span_p, span_n;
  // clamped-determinant suspects (per view for camera rays, per light for shadow rays)
  DBuf sus_view, sus_view_cnt, sus_light, sus_light_cnt;
  int view_cnt_zeroed = 0;   // view suspect counters the last BVH build cleared and no list build used
  // the suspect lists binned per cell (k_sus_bin, k_light_bin) when F >= SUS_BIN_MIN_F:
  // records in a pool, (offset, count) and a counter per cell
  DBuf sus_pool, sus_cell, sus_cell_cnt, sus_light_pool, sus_light_cell, sus_light_cell_cnt, sus_light_grid;
  long long sus_cells_stride = 0, sus_pool_cap = 0;   // cells and pool records per view
  static constexpr int LIGHT_GRID = 64;                // light grid: LIGHT_GRID^2 cells
  static constexpr int LIGHT_POOL = 16 * LIGHT_GRID * LIGHT_GRID;
  bool light_sus = false;
  bool tc_zero = false;     // every texture coordinate of the committed scene is 0
  // edges
  DBuf edges, edge_mesh, pmf_raw, pmf, cdf, cdfw, proj;
  // sampling
  DBuf rnd, hamm;   // sorted edge-sample table, Hammersley offsets
  long long rnd_n = -1;
  // per call
  DBuf cams, img, adj, tgt, loss_part, loss, grad;
  DBuf fault;   // the kernels' fault word (DevScene::fault), zero while all is well
  DBuf job_tgt, job_loss;   // a served job's target images and loss word (drh_job_upload)
  // Host staging: what a commit, a camera change or a handle's own edge-sample table sends to
  // the device is first copied into a pinned buffer of the handle and sent from there by DMA
  // on the stream; the call returns without waiting for the transfer (the caller's arrays
  // are free at once).  The buffers form a small ring: a staging takes one whose transfers
  // have passed (asked without waiting), else adds one, and waits (an event) only when
  // STAGE_MAX are in flight -- in a busy process a handle's transfers queue behind the
  // kernels that share its stream's hardware queue, so waiting on them would stall the caller.
  struct Stage { char* p = nullptr; size_t cap = 0; hipEvent_t done = nullptr; bool pending = false; unsigned long long seq = 0; };
  static constexpr size_t STAGE_MAX = 6;
  std::vector<Stage> stg;
  int stg_cur = 0;
  size_t stg_used = 0;
  unsigned long long stg_seq = 0;
  drh_status stage_begin(size_t need) {
    int pick = -1, oldest = -1;
    for (size_t i = 0; i < stg.size(); i++) {
      Stage& b = stg[i];
      if (b.pending) {
        if (hipEventQuery(b.done) != hipSuccess) {
          (void)hipGetLastError();   // not ready is no error of the next call
          if (oldest < 0 || b.seq < stg[oldest].seq) oldest = (int)i;
          continue;
        }
        b.pending = false;
      }
      if (pick < 0 || (b.cap >= need && stg[pick].cap < need)) pick = (int)i;
    }
    if (pick < 0 && stg.size() < STAGE_MAX) {
      stg.emplace_back();
      pick = (int)stg.size() - 1;
    }
    if (pick < 0) {
      pick = oldest;
      if (hipEventSynchronize(stg[pick].done) != hipSuccess) return hip(hipGetLastError(), "staging");
      stg[pick].pending = false;
    }
    stg_cur = pick;
    Stage& b = stg[stg_cur];
    if (need > b.cap) {
      pinned_give(b.p, b.cap);   // its transfers are done
      b.p = nullptr;
      b.cap = 0;
      size_t cap = 0;
      if (!(b.p = pinned_take(need, &cap))) return fail(DRH_E_OOM, "pinned staging buffer of %zu bytes", pinned_class(need));
      b.cap = cap;
    }
    if (!b.done && hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess) return hip(hipGetLastError(), "staging");
    stg_used = 0;
    segs.clear();
    return DRH_OK;
  }
  static size_t stage_round(size_t bytes) { return (bytes + 255) & ~(size_t)255; }
  // copy src into the current staging buffer (stage_begin reserved room for it) and send it
  drh_status stage_send(void* dst, const void* src, size_t bytes) {
    if (!bytes) return DRH_OK;
    if (stg_used + bytes > stg[stg_cur].cap) return fail(DRH_E_INVALID, "staging: %zu bytes beyond the reserved room", bytes);
    char* q = stg[stg_cur].p + stg_used;
    std::memcpy(q, src, bytes);
    stg_used += stage_round(bytes);
    return hip(hipMemcpyAsync(dst, q, bytes, hipMemcpyHostToDevice, stream), "upload");
  }
  // Many arrays at once: stage_put collects them (copied into the staging buffer now),
  // stage_flush sends the staged block with one copy into `arena` and scatters it to the
  // arrays' buffers with one kernel (instead of a copy call per array).
  struct Seg { void* dst; size_t off, bytes; };
  std::vector<Seg> segs;
  DBuf arena;
  drh_status stage_put(void* dst, const void* src, size_t bytes) {
    if (!bytes) return DRH_OK;
    if (stg_used + bytes > stg[stg_cur].cap || bytes % 4) {
      segs.clear();
      return fail(DRH_E_INVALID, "staging: %zu bytes beyond the reserved room", bytes);
    }
    char* q = stg[stg_cur].p + stg_used;
    std::memcpy(q, src, bytes);
    segs.push_back(Seg{dst, stg_used, bytes});
    stg_used += stage_round(bytes);
    return DRH_OK;
  }
  drh_status stage_flush() {
    drh_status s;
    if (segs.empty()) return DRH_OK;
    if ((s = hip(arena.reserve(stg_used), "alloc")) ||
        (s = hip(hipMemcpyAsync(arena.p, stg[stg_cur].p, stg_used, hipMemcpyHostToDevice, stream), "upload")))
      return s;
    for (size_t b = 0; b < segs.size(); b += SCATTER_MAX) {
      ScatterArgs a{};
      a.src = arena.as<uint32_t>();
      for (size_t k = b; k < segs.size() && a.n < SCATTER_MAX; k++, a.n++) {
        a.dst[a.n] = (uint32_t*)segs[k].dst;
        a.off[a.n] = (long long)(segs[k].off / 4);
        a.words[a.n] = (long long)(segs[k].bytes / 4);
      }
      launch_scatter(a, stream);
    }
    segs.clear();
    return hip(hipGetLastError(), "upload");
  }
  drh_status stage_end() {
    Stage& b = stg[stg_cur];
    if (hipEventRecord(b.done, stream) != hipSuccess) return hip(hipGetLastError(), "staging");
    b.pending = true;
    b.seq = ++stg_seq;
    return DRH_OK;
  }
  void stage_free() {
    for (Stage& b : stg) {
      if (b.pending) (void)hipEventSynchronize(b.done);
      if (b.done) (void)hipEventDestroy(b.done);
      pinned_give(b.p, b.cap);
    }
    stg.clear();
  }
  const float2* rnd_ptr = nullptr;   // the edge-sample table (shared, or this handle's rnd)
  // Ray counting (drh_count_rays): the traversal kernels' counting instantiations add their
  // rays to rays_buf: [0, 1] render closest / any-hit, [2, 3] interior, [4, 5] edge trace.
  bool count_on = false;
  DBuf rays_buf;
  unsigned long long* ray_slot(int kernel) const {
    return count_on ? rays_buf.as<unsigned long long>() + 2 * kernel : nullptr;
  }
  // After a synchronisation: a set fault word fails the call with DRH_E_DEVICE (and is
  // cleared, so the handle stays usable).
  drh_status check_fault() {
    uint32_t f = 0;
    drh_status s;
    if ((s = hip(hipMemcpyAsync(&f, fault.p, 4, hipMemcpyDeviceToHost, stream), "fault word")) ||
        (s = hip(hipStreamSynchronize(stream), "fault word")))
      return s;
    if (!f) return DRH_OK;
    (void)hipMemsetAsync(fault.p, 0, 4, stream);
    return fail(DRH_E_DEVICE, "device check failed (fault word 0x%x: %s)", f,
                (f & FAULT_STACK_OVERFLOW) ? "a traversal stack overflowed" : "unknown");
  }
  std::vector<DevCam> cam_cache;
  // optimiser
  int opt_kind = -1;
  float pos_lr = 0.1f, col_lr = 0.1f, tex_lr = 0.0f, tr_lr = 0.0f;
  bool skip_tex_interior = false;   // set by the fused step: the interior pass feeds nothing
  int opt_iter = 0;
  DBuf opt_m, opt_v, opt_t;
  // captured optimisation step
  bool use_graph = true;
  // two captured forms: [1] with the full LBVH build, [0] with the refit only
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<char> graph_sig[2];
  unsigned long long graph_epoch[2] = {0, 0};
  // k-step graphs (drh_opt_steps): k consecutive fused steps captured as one graph, keyed by k,
  // the steps' LBVH forms (bit s: full build at step s), the arguments and the epoch.
  struct MultiGraph {
    int k;
    unsigned forms;
    std::vector<char> sig;
    unsigned long long epoch;
    hipGraph_t g;
    hipGraphExec_t e;
  };
  std::vector<MultiGraph> multi;
  static constexpr size_t MULTI_KEEP = 6;
  // Dropping graphs does not wait for the stream: the executables and graphs wait, with an
  // event recorded behind their last launch, and are destroyed once that event has passed
  // (checked without waiting at every later drop; beyond GRAVES_MAX waiting entries the
  // oldest is waited for).  A server thread that steps many handles then never blocks on a
  // handle whose kernels queue behind the other handles' work.
  struct Grave { hipEvent_t ev; std::vector<hipGraphExec_t> execs; std::vector<hipGraph_t> graphs; };
  std::deque<Grave> graves;
  std::vector<hipEvent_t> grave_events;   // recycled
  static constexpr size_t GRAVES_MAX = 8;
  void bury(std::vector<hipGraphExec_t>&& execs, std::vector<hipGraph_t>&& graphs) {
    if (execs.empty() && graphs.empty()) return;
    hipEvent_t ev = nullptr;
    if (!grave_events.empty()) {
      ev = grave_events.back();
      grave_events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      ev = nullptr;
    }
    if (!ev || hipEventRecord(ev, stream) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipStreamSynchronize(stream);   // no event: wait here instead
      for (auto e : execs) (void)hipGraphExecDestroy(e);
      for (auto g : graphs) (void)hipGraphDestroy(g);
      if (ev) grave_events.push_back(ev);
      return;
    }
    graves.push_back(Grave{ev, std::move(execs), std::move(graphs)});
    reap(false);
  }
  // destroy the waiting entries whose event has passed (with wait: all, waiting for them)
  void reap(bool wait) {
    while (!graves.empty()) {
      Grave& g = graves.front();
      if (wait || graves.size() > GRAVES_MAX) {
        (void)hipEventSynchronize(g.ev);
      } else if (hipEventQuery(g.ev) != hipSuccess) {
        (void)hipGetLastError();   // not ready is no error of the next call
        break;
      }
      for (auto e : g.execs) (void)hipGraphExecDestroy(e);
      for (auto x : g.graphs) (void)hipGraphDestroy(x);
      grave_events.push_back(g.ev);
      graves.pop_front();
    }
  }
  void drop_form(int k) {
    graph_sig[k].clear();
    std::vector<hipGraphExec_t> e;
    std::vector<hipGraph_t> g;
    if (gexec[k]) e.push_back(gexec[k]);
    if (graph[k]) g.push_back(graph[k]);
    gexec[k] = nullptr;
    graph[k] = nullptr;
    bury(std::move(e), std::move(g));
  }
  // drop the oldest `n` k-step graphs (all with n < 0)
  void drop_multi(long n = -1) {
    const size_t cnt = n < 0 ? multi.size() : std::min(multi.size(), (size_t)n);
    std::vector<hipGraphExec_t> e;
    std::vector<hipGraph_t> g;
    for (size_t i = 0; i < cnt; i++) {
      e.push_back(multi[i].e);
      g.push_back(multi[i].g);
    }
    multi.erase(multi.begin(), multi.begin() + cnt);
    bury(std::move(e), std::move(g));
  }
  void drop_graph() { drop_form(0); drop_form(1); drop_multi(); }
  void destroy_graphs() {
    drop_graph();
    reap(true);
    for (auto ev : grave_events) (void)hipEventDestroy(ev);
    grave_events.clear();
  }
  // executable for captured graph g of form k
  bool instantiate(int k, hipGraph_t g) {
    reap(false);
    return hipGraphInstantiate(&gexec[k], g, nullptr, nullptr, 0) == hipSuccess;
  }
  // profiling
  bool prof = false;
  std::vector<ProfEvent> pending;
  std::map<std::string, std::pair<double, int>> prof_acc;
  std::vector<hipEvent_t> event_pool;


  drh_status fail(drh_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return s;
  }
  drh_status hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return DRH_OK;
    return fail(e == hipErrorOutOfMemory ? DRH_E_OOM : DRH_E_HIP, "%s: %s", what, hipGetErrorString(e));
  }

  hipEvent_t ev() {
    if (!event_pool.empty()) { hipEvent_t e = event_pool.back(); event_pool.pop_back(); return e; }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
  template <class F> void timed(const char* name, F&& launch) {
    if (!prof) { launch(); return; }
    ProfEvent pe{name, ev(), ev()};
    (void)hipEventRecord(pe.a, stream);
    launch();
    (void)hipEventRecord(pe.b, stream);
    pending.push_back(pe);
  }
  void drain_prof() {
    if (pending.empty()) return;
    (void)hipStreamSynchronize(stream);
    for (auto& p : pending) {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, p.a, p.b);
      auto& acc = prof_acc[p.name];
      acc.first += ms;
      acc.second += 1;
      event_pool.push_back(p.a);
      event_pool.push_back(p.b);
    }
    pending.clear();
  }

  Tracer tracer() const {
    Tracer t{};
    DevScene& s = t.s;
    s.tri_a = tri_a.as<float4>(); s.tri_e1 = tri_e1.as<float4>(); s.tri_e2 = tri_e2.as<float4>();
    s.nodes = nodes.as<float4>();
    s.F = F; s.V = V;
    s.idx = idx.as<uint32_t>(); s.tri_mesh = tri_mesh.as<int>(); s.mesh_vo = mesh_vo.as<int>(); s.mesh_to = mesh_to.as<int>();
    s.pos = pos.as<float>(); s.pos_orig = pos_orig.as<float>(); s.col = col.as<float>(); s.tc = tc.as<float>();
    s.nrm = nrm.as<float>(); s.nrm2d = nrm2d.as<float>(); s.texels = texels.as<float>();
    s.tex_w = dtex_w.as<int>(); s.tex_h = dtex_h.as<int>(); s.tex_off = dtex_off.as<int>();
    s.n_meshes = nm;
    s.fault = fault.as<int>();
    t.bf_a = bf_a.as<float4>(); t.bf_e1 = bf_e1.as<float4>(); t.bf_e2 = bf_e2.as<float4>();
    t.use_bf = st.tracer == DRH_TRACER_BF;
    t.cam = {nullptr, nullptr};
    t.light = {light_sus ? sus_light.as<float4>() : nullptr, light_sus ? sus_light_cnt.as<int>() : nullptr};
    if (light_sus && F >= SUS_BIN_MIN_F)
      t.light = {sus_light.as<float4>(), sus_light_cnt.as<int>(), sus_light_pool.as<float4>(), sus_light_cell.as<int2>(),
                 LIGHT_GRID, LIGHT_GRID, sus_light_grid.as<float4>()};
    return t;
  }
  // Tracer for the camera rays of view v (its suspect list must have been built).
  Tracer tracer_view(int v) const {
    Tracer t = tracer();
    if (!t.use_bf && F > 0) {
      const DevCam& c = cam_cache[v];
      t.cam = {sus_view.as<float4>() + (size_t)v * F, sus_view_cnt.as<int>() + v, nullptr, nullptr, 0, 0, nullptr};
      if (F >= SUS_BIN_MIN_F) {
        t.cam.pool = sus_pool.as<float4>() + 2 * sus_pool_cap * v;
        t.cam.cell = sus_cell.as<int2>() + sus_cells_stride * v;
        t.cam.cells_x = (c.w + SUS_CELL_PX - 1) / SUS_CELL_PX;
        t.cam.cells_y = (c.h + SUS_CELL_PX - 1) / SUS_CELL_PX;
      }
    }
    t.s.lane_walk = lane_walk_cam(v);
    return t;
  }
  // LBVH walk per launch (drh_set_traversal).  A wave's camera rays cover a tile of 64
  // pixels: with many triangles per pixel the nodes the packet opens together are many times
  // those of one ray, and each lane walking alone is faster.  The edge pass's side rays are
  // as coherent as its samples are dense on the edges.
  int traversal = DRH_TRAVERSAL_AUTO;
  static constexpr double LANE_TRIS_PER_PIXEL = 0.5, LANE_EDGES_PER_SAMPLE = 0.5;
  int lane_walk_cam(int v) const {
    if (traversal != DRH_TRAVERSAL_AUTO) return traversal == DRH_TRAVERSAL_LANE;
    return (double)F > LANE_TRIS_PER_PIXEL * (double)cam_cache[v].w * (double)cam_cache[v].h;
  }
  int lane_walk_edge(long long n) const {
    if (traversal != DRH_TRAVERSAL_AUTO) return traversal == DRH_TRAVERSAL_LANE;
    return (double)E > LANE_EDGES_PER_SAMPLE * (double)n;
  }
  // Suspect triangles of every view's camera origin (CastSingleRay's origin,
  // WorldViewInv * (0,0,0,1), raytrace.h:67-76).
  int views_built = -1;   // views whose suspect lists match the current geometry and cameras
  drh_status build_view_suspects(int views) {
    if (st.tracer == DRH_TRACER_BF || F == 0) return DRH_OK;
    if (views_built == views) return DRH_OK;
    drh_status s;
    long long cells = 1;
    for (int v = 0; v < views; v++)
      cells = std::max<long long>(cells, (long long)((cam_cache[v].w + SUS_CELL_PX - 1) / SUS_CELL_PX) *
                                             ((cam_cache[v].h + SUS_CELL_PX - 1) / SUS_CELL_PX));
    const bool bin = F >= SUS_BIN_MIN_F;
    const long long cap = std::max<long long>(4096, 8 * cells);   // pool records per view
    if ((s = hip(sus_view.reserve(16 * (size_t)F * views), "alloc")) ||
        (s = hip(sus_view_cnt.reserve(4 * (size_t)views), "alloc")))
      return s;
    if (bin && ((s = hip(sus_pool.reserve(32 * (size_t)cap * views), "alloc")) ||
                (s = hip(sus_cell.reserve(8 * (size_t)cells * views), "alloc")) ||
                (s = hip(sus_cell_cnt.reserve(4 * (size_t)cells * views), "alloc"))))
      return s;
    sus_cells_stride = cells;
    sus_pool_cap = cap;
    if (view_cnt_zeroed < views) timed("k_fill_u32", [&] { launch_fill_u32(sus_view_cnt.as<uint32_t>(), views, 0u, stream); });
    view_cnt_zeroed = 0;
    if (bin) timed("k_fill_u32", [&] { launch_fill_u32(sus_cell_cnt.as<uint32_t>(), cells * views, 0u, stream); });
    for (int v = 0; v < views; v++) {
      const M4 wi = m4_load(cam_cache[v].wv_inv);
      const float z4[4] = {0.0f, 0.0f, 0.0f, 1.0f};
      float o[4];
      m4_vec(wi, z4, o);
      SuspectArgs a{F, 0, o[0], o[1], o[2], bf_a.as<float4>(), bf_e1.as<float4>(), bf_e2.as<float4>(),
                    sus_view.as<float4>() + (size_t)v * F, sus_view_cnt.as<int>() + v};
      timed("k_suspects", [&] { launch_suspects(a, stream); });
    }
    if (bin) {
      SusBinArgs b{cams.as<DevCam>(), sus_view.as<float4>(), sus_view_cnt.as<int>(), F, bf_a.as<float4>(),
                   bf_e1.as<float4>(), bf_e2.as<float4>(),
                   BinPool{sus_pool.as<float4>(), (int)cap, sus_cell.as<int2>(), sus_cell_cnt.as<int>(), 0, 0}, cap, cells};
      timed("k_sus_bin", [&] { launch_sus_bin(b, views, stream); });
    }
    views_built = views;
    return hip(hipGetLastError(), "suspects");
  }

  // full: Morton sort + Karras topology + refit; otherwise (same triangle count, a topology
  // from an earlier full build) only the boxes are refitted over the moved geometry
  drh_status build_bvh(bool full = true);
  int* prep_step = nullptr;   // set by a fused step: k_prepare advances this Adam step counter
  bool topo_valid = false;   // child / parent / leaf-order buffers hold a topology for F
  int rebuild_every = 8;     // fused steps: a full build every this many steps, refits between
  drh_status upload_scene(const drh_scene_desc* d);
  drh_status set_cams(const drh_camera* cams, int views, long long* total_floats);
  drh_status run_render(int views, float* d_img, const float* d_target, float* d_adj, float* d_loss,
                        bool store_hits = false);
  drh_status run_d_render(int views, const float* d_adj, long long N, float* d_grad, bool accumulate);
  drh_status ensure_rnd(long long N);
  // projection, edge lengths and the exact CDF of view v on stream q
  void edge_prep(int v, hipStream_t q) {
    ProjectArgs pa{V, cams.as<DevCam>(), v, pos.as<float>(), proj.as<float4>(), edges.as<int2>(), E, pmf_raw.as<float>()};
    timed("k_project", [&] { launch_project(pa, q); });   // and the edge lengths
    timed("k_cdf_par", [&] { launch_cdf(pmf_raw.as<float>(), E, pmf.as<float>(), cdf.as<float>(), cdfw.p, q, true); });
  }
  // fork view 0's edge preparation onto the side stream (not while profiling: per-kernel
  // events time the main stream)
  // (zero_grad, n: the gradient the step's d_render will overwrite, cleared there as well)
  drh_status fork_prep0(long long N, float* zero_grad = nullptr, long long n = 0) {
    prep0_forked = false;
    grad_zeroed = false;
    if (prof || sharing > 1 || E <= 0 || N <= 0) return DRH_OK;
    drh_status s;
    if ((s = hip(hipEventRecord(ev_fork, stream), "fork")) || (s = hip(hipStreamWaitEvent(side, ev_fork, 0), "fork")))
      return s;
    edge_prep(0, side);
    if (zero_grad && n > 0) {
      launch_fill_f32(zero_grad, n, 0.0f, side);
      grad_zeroed = true;
    }
    if ((s = hip(hipEventRecord(ev_join, side), "fork"))) return s;
    prep0_forked = true;
    return hip(hipGetLastError(), "edge prep");
  }
};

// ---------------------------------------------------------------------------
// commit
// ---------------------------------------------------------------------------
drh_status drh_renderer::upload_scene(const drh_scene_desc* d) {
  if (!d || d->n_meshes <= 0 || !d->vcount || !d->fcount || !d->transforms || !d->pos || !d->idx)
    return fail(DRH_E_INVALID, "commit: scene arrays missing");
  const int n = d->n_meshes;
  std::vector<int> nvo(n + 1, 0), nto(n + 1, 0);
  long long tv = 0, tf = 0;
  for (int i = 0; i < n; i++) {
    if (d->vcount[i] < 0 || d->fcount[i] < 0) return fail(DRH_E_INVALID, "commit: negative size in mesh %d", i);
    tv += d->vcount[i];
    tf += d->fcount[i];
    // LBVH nodes and leaves are addressed by 32-bit byte offsets (node * 64, triangle * 16)
    if (tf > MAX_TRIANGLES || tv > MAX_TRIANGLES * 3)
      return fail(DRH_E_INVALID, "commit: more than %lld triangles or %lld vertices", MAX_TRIANGLES, 3 * MAX_TRIANGLES);
    nvo[i + 1] = nvo[i] + d->vcount[i];
    nto[i + 1] = nto[i] + d->fcount[i];
  }
  const int nV = nvo[n], nF = nto[n];
  for (int i = 0; i < n; i++)
    for (int f = nto[i]; f < nto[i + 1]; f++)
      for (int k = 0; k < 3; k++)
        if (d->idx[3 * f + k] >= (uint32_t)d->vcount[i])
          return fail(DRH_E_INVALID, "commit: index %u out of range in mesh %d", d->idx[3 * f + k], i);
  std::vector<int> tw(n), th(n), toff(n), df(n);
  int tsum = 0;
  for (int i = 0; i < n; i++) {
    tw[i] = d->tex_w ? d->tex_w[i] : 0;
    th[i] = d->tex_h ? d->tex_h[i] : 0;
    toff[i] = tsum;
    tsum += 3 * tw[i] * th[i];
    df[i] = d->differentiable ? (d->differentiable[i] != 0) : 1;
    if ((st.mode == DRH_TEXTURE_COLOR || st.mode == DRH_LAMBERT) && tw[i] * th[i] == 0)
      return fail(DRH_E_UNSUPPORTED, "commit: mesh %d has no texture (TEXTURE_COLOR / LAMBERT read texture 0)", i);
  }
  if ((st.mode == DRH_TEXTURE_COLOR || st.mode == DRH_LAMBERT) && !d->tc)
    return fail(DRH_E_UNSUPPORTED, "commit: texture coordinates required for this shading mode");
  // RESTRICTED_TRANSFORM meshes: the transform is TransformR::to_mat of their 7 parameters
  // (Scene::prepare_for_render, common_src/scene_utils.cpp:113-118)
  std::vector<int> rf(n, 0);
  std::vector<float> rp(7 * (size_t)n, 0.0f);
  std::vector<float> eff(d->transforms, d->transforms + 16 * (size_t)n);
  for (int i = 0; i < n; i++) {
    rf[i] = d->restricted ? (d->restricted[i] != 0) : 0;
    if (!rf[i]) continue;
    if (!d->restricted_params) return fail(DRH_E_INVALID, "commit: restricted mesh %d without restricted_params", i);
    std::memcpy(&rp[7 * (size_t)i], d->restricted_params + 7 * (size_t)i, 28);
    const M4 T = m4_restricted(&rp[7 * (size_t)i]);
    std::memcpy(&eff[16 * (size_t)i], T.m, 64);
  }
  nm = n; V = nV; F = nF; vo = nvo; to = nto; tex_w = tw; tex_h = th; tex_off = toff; diff = df;
  rflag = rf; rparams_h = rp;
  any_restricted = false;
  for (int i = 0; i < n; i++) any_restricted = any_restricted || (rf[i] && df[i]);
  has_nrm = d->nrm != nullptr;
  has_col = d->col != nullptr;
  // gradient layout (DMesh, src/dmesh.h:35-85)
  layout.assign(4 * n, -1);
  G = 0; n_diff = 0;
  for (int i = 0; i < n; i++) {
    if (!df[i]) continue;
    n_diff++;
    const int nv = vo[i + 1] - vo[i];
    int off = 3 * nv;
    int colo = -1, texo = -1;
    if (st.mode == DRH_VERTEX_COLOR || st.mode == DRH_LAMBERT_VERTEX_COLOR) { colo = off; off += 3 * nv; }
    if (st.mode == DRH_TEXTURE_COLOR || st.mode == DRH_LAMBERT) { texo = off; off += 3 * tw[i] * th[i]; }
    layout[4 * i] = G; layout[4 * i + 1] = colo; layout[4 * i + 2] = texo; layout[4 * i + 3] = off;
    G += off + (rf[i] ? 7 : 12);   // RESTRICTED_TRANSFORM_SIZE or TRANSFORM_SIZE
  }
  // matrices: T, transpose(inverse(T)), identity
  mats.assign(48 * (size_t)n, 0.0f);
  for (int i = 0; i < n; i++) {
    const M4 T = m4_load(&eff[16 * (size_t)i]);
    const M4 NT = m4_transpose(m4_inverse(T));
    const M4 I = m4_identity();
    std::memcpy(&mats[48 * i], T.m, 64);
    std::memcpy(&mats[48 * i + 16], NT.m, 64);
    std::memcpy(&mats[48 * i + 32], I.m, 64);
  }
  {
    const M4 T0inv = m4_inverse(m4_load(eff.data()));
    tr_inv0.assign(T0inv.m, T0inv.m + 16);
  }
  // rows 0-2 of each transform, row-major: the optimiser's copy (the order of d_tr)
  std::vector<float> rows(12 * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 4; c++) rows[12 * i + 4 * r + c] = eff[16 * i + 4 * c + r];
  // edge set: std::set<Edge> of differentiable meshes ordered by (v0, v1), first kept
  // (src/drender.h:26-52,274-285); depends on topology only, so cached.
  std::vector<uint32_t> idx_local(d->idx, d->idx + 3 * (size_t)nF);
  const bool topo_same = idx_local == idx_local_cache && df == diff_cache && E >= 0 && edges.p;
  std::vector<int2> hedges;
  std::vector<int> hemesh;
  if (!topo_same) {
    // (v0, v1) with v0 < v1 in ascending order, duplicates once: a counting sort by v0, then
    // each vertex's few v1 in order.  Meshes own disjoint vertex ranges, so an edge occurs
    // in one mesh only and whichever copy is kept carries the same mesh id.
    std::vector<int> start(nV + 1, 0);
    auto each_edge = [&](auto&& fn) {
      for (int i = 0; i < n; i++) {
        if (!df[i]) continue;
        for (int f = nto[i]; f < nto[i + 1]; f++)
          for (int k = 0; k < 3; k++) {
            const uint32_t a = nvo[i] + idx_local[3 * f + k], b = nvo[i] + idx_local[3 * f + (k + 1) % 3];
            fn(std::min(a, b), std::max(a, b), i);
          }
      }
    };
    each_edge([&](uint32_t v0, uint32_t, int) { start[v0 + 1]++; });   // indices checked above
    for (int v = 0; v < nV; v++) start[v + 1] += start[v];
    std::vector<uint32_t> other(start[nV]);
    std::vector<int> omesh(start[nV]);
    std::vector<int> fill(start.begin(), start.end() - 1);
    each_edge([&](uint32_t v0, uint32_t v1, int i) { other[fill[v0]] = v1; omesh[fill[v0]] = i; fill[v0]++; });
    for (int v = 0; v < nV; v++) {
      const int b = start[v], e = start[v + 1];
      for (int x = b + 1; x < e; x++)   // insertion sort: a vertex has a handful of neighbours
        for (int y = x; y > b && other[y] < other[y - 1]; y--) { std::swap(other[y], other[y - 1]); std::swap(omesh[y], omesh[y - 1]); }
      for (int x = b; x < e; x++)
        if (x == b || other[x] != other[x - 1]) {
          hedges.push_back(make_int2(v, (int)other[x]));
          hemesh.push_back(omesh[x]);
        }
    }
  }
  std::vector<uint32_t> gidx(3 * (size_t)nF);
  std::vector<int> tmesh(nF);
  for (int i = 0; i < n; i++)
    for (int f = nto[i]; f < nto[i + 1]; f++) {
      tmesh[f] = i;
      for (int k = 0; k < 3; k++) gidx[3 * f + k] = (uint32_t)nvo[i] + idx_local[3 * f + k];
    }

  auto R = [&](DBuf& b, size_t bytes, const char* what) -> drh_status { return hip(b.reserve(std::max<size_t>(bytes, 16)), what); };
  auto up = [&](DBuf& b, const void* src, size_t bytes) -> drh_status { return stage_put(b.p, src, bytes); };
  drh_status s;
  const size_t v3 = 12 * (size_t)V;
  {
    auto q = [](size_t b) { return stage_round(b); };
    size_t need = q(v3) + q(v3) + q(8 * (size_t)V) + 2 * q(v3) + q(4 * (size_t)tsum) + q(12 * (size_t)F) +
                  q(4 * (size_t)F) + 2 * q(4 * (size_t)(n + 1)) + 3 * q(4 * (size_t)n) + q(4 * mats.size()) +
                  q(4 * layout.size()) + q(64) + q(48 * (size_t)n) + q(4 * (size_t)n) + q(28 * (size_t)n);
    if (!topo_same) need += q(8 * hedges.size()) + q(4 * hemesh.size());
    if ((s = stage_begin(need))) return s;
  }
  if ((s = R(pos_orig, v3, "alloc")) || (s = R(pos, v3, "alloc")) || (s = R(col, v3, "alloc")) ||
      (s = R(tc, 8 * (size_t)V, "alloc")) || (s = R(nrm_in, v3, "alloc")) || (s = R(nrm, v3, "alloc")) ||
      (s = R(nrm2d, v3, "alloc")) || (s = R(texels, 4 * (size_t)tsum, "alloc")) || (s = R(idx, 12 * (size_t)F, "alloc")) ||
      (s = R(tri_mesh, 4 * (size_t)F, "alloc")) || (s = R(mesh_vo, 4 * (size_t)(n + 1), "alloc")) ||
      (s = R(mesh_to, 4 * (size_t)(n + 1), "alloc")) || (s = R(dtex_w, 4 * (size_t)n, "alloc")) ||
      (s = R(dtex_h, 4 * (size_t)n, "alloc")) || (s = R(dtex_off, 4 * (size_t)n, "alloc")) ||
      (s = R(dmats, 4 * mats.size(), "alloc")) || (s = R(dlayout, 4 * layout.size(), "alloc")) ||
      (s = R(dtr_inv0, 64, "alloc")) || (s = R(tr_rows, 48 * (size_t)n, "alloc")) ||
      (s = R(drflag, 4 * (size_t)n, "alloc")) || (s = R(drparams, 28 * (size_t)n, "alloc")) ||
      (s = R(trscr, 48 * (size_t)n, "alloc")))
    return s;
  if ((s = up(pos_orig, d->pos, v3))) return s;
  if (d->col) { if ((s = up(col, d->col, v3))) return s; }
  else { launch_fill_f32(col.as<float>(), 3ll * V, 0.0f, stream); }
  tc_zero = true;
  if (d->tc) {
    if ((s = up(tc, d->tc, 8 * (size_t)V))) return s;
    for (size_t k = 0; k < 2 * (size_t)V && tc_zero; k++) tc_zero = d->tc[k] == 0.0f;
  } else {
    launch_fill_f32(tc.as<float>(), 2ll * V, 0.0f, stream);
  }
  if (d->nrm) { if ((s = up(nrm_in, d->nrm, v3))) return s; }
  else {
    std::vector<float> def(3 * (size_t)V, 0.0f);
    for (int k = 0; k < V; k++) def[3 * k] = 1.0f;  // prepared default normal (1,0,0)
    if ((s = up(nrm, def.data(), v3)) || (s = up(nrm2d, def.data(), v3))) return s;
  }
  if (tsum && d->texels) { if ((s = up(texels, d->texels, 4 * (size_t)tsum))) return s; }
  if ((s = up(idx, gidx.data(), 12 * (size_t)F)) || (s = up(tri_mesh, tmesh.data(), 4 * (size_t)F)) ||
      (s = up(mesh_vo, vo.data(), 4 * (size_t)(n + 1))) || (s = up(mesh_to, to.data(), 4 * (size_t)(n + 1))) ||
      (s = up(dtex_w, tw.data(), 4 * (size_t)n)) || (s = up(dtex_h, th.data(), 4 * (size_t)n)) ||
      (s = up(dtex_off, toff.data(), 4 * (size_t)n)) || (s = up(dmats, mats.data(), 4 * mats.size())) ||
      (s = up(dlayout, layout.data(), 4 * layout.size())) || (s = up(dtr_inv0, tr_inv0.data(), 64)) ||
      (s = up(tr_rows, rows.data(), 48 * (size_t)n)) || (s = up(drflag, rf.data(), 4 * (size_t)n)) ||
      (s = up(drparams, rp.data(), 28 * (size_t)n)))
    return s;
  if (!topo_same) {
    E = (int)hedges.size();
    if ((s = R(edges, 8 * (size_t)E, "alloc")) || (s = R(edge_mesh, 4 * (size_t)E, "alloc")) ||
        (s = R(pmf_raw, 4 * (size_t)E, "alloc")) || (s = R(pmf, 4 * (size_t)E, "alloc")) ||
        (s = R(cdf, 4 * (size_t)(E + 1), "alloc")) || (s = R(cdfw, cdf_work_bytes(E), "alloc")))
      return s;
    if ((s = up(edges, hedges.data(), 8 * (size_t)E)) || (s = up(edge_mesh, hemesh.data(), 4 * (size_t)E))) return s;
    idx_local_cache = idx_local;
    diff_cache = df;
  }
  if ((s = R(proj, 16 * (size_t)V, "alloc"))) return s;
  if ((s = stage_flush())) return s;
  return stage_end();   // the caller's arrays were copied into the staging buffer
}

drh_status drh_renderer::build_bvh(bool full) {
  views_built = -1;   // the views' suspect lists belong to the old geometry
  PrepareArgs pa{V, nm, has_nrm ? 1 : 0, mesh_vo.as<int>(), dmats.as<float>(), pos_orig.as<float>(),
                 nrm_in.as<float>(), pos.as<float>(), nrm.as<float>(), nrm2d.as<float>(), nullptr, nullptr, prep_step};
  if (F == 0) {
    if (V > 0) timed("k_prepare", [&] { launch_prepare(pa, stream); });
    else if (prep_step) timed("k_step_inc", [&] { launch_step_inc(prep_step, stream); });
    return hip(hipGetLastError(), "commit kernels");
  }
  auto R = [&](DBuf& b, size_t bytes) -> drh_status { return hip(b.reserve(std::max<size_t>(bytes, 16)), "alloc"); };
  const size_t f16 = 16 * (size_t)F, f4 = 4 * (size_t)F;
  const int nb = (F + RADIX_TILE - 1) / RADIX_TILE;
  drh_status s;
  if ((s = R(bf_a, f16)) || (s = R(bf_e1, f16)) || (s = R(bf_e2, f16)) || (s = R(box_lo, f16)) || (s = R(box_hi, f16)) ||
      (s = R(cen, f16)) || (s = R(bounds, 32)) || (s = R(keys0, f4)) || (s = R(vals0, f4)) || (s = R(keys1, f4)) ||
      (s = R(vals1, f4)) || (s = R(hist, 4 * 256 * (size_t)nb)) || (s = R(child, 8 * (size_t)F)) ||
      (s = R(node_parent, f4)) || (s = R(leaf_parent, f4)) || (s = R(counter, f4 + 4)) ||
      (s = R(nrange, 2 * f4)) || (s = R(frontier, f4)) || (s = R(leaf_lo, f16)) ||
      (s = R(leaf_hi, f16)) || (s = R(node_lo, f16)) || (s = R(node_hi, f16)) || (s = R(nodes, 4 * f16)) ||
      (s = R(tri_a, f16)) || (s = R(tri_e1, f16)) || (s = R(tri_e2, f16)))
    return s;
  if (F > REFIT_CHUNK &&
      ((s = R(span, 16 * (size_t)SPAN_MAX)) || (s = R(span_cid, 8 * (size_t)SPAN_MAX)) || (s = R(span_p, 4 * (size_t)SPAN_MAX)) ||
       (s = R(span_of, f4)) || (s = R(span_n, 16))))
    return s;
  light_sus = (st.mode == DRH_LAMBERT || st.mode == DRH_LAMBERT_VERTEX_COLOR) && st.tracer != DRH_TRACER_BF;
  if (light_sus && ((s = R(sus_light, 2 * f16)) || (s = R(sus_light_cnt, 16))))   // two float4 per entry
    return s;
  if (light_sus && F >= SUS_BIN_MIN_F &&
      ((s = R(sus_light_pool, 32 * (size_t)LIGHT_POOL)) || (s = R(sus_light_cell, 8 * (size_t)LIGHT_GRID * LIGHT_GRID)) ||
       (s = R(sus_light_cell_cnt, 4 * (size_t)LIGHT_GRID * LIGHT_GRID)) || (s = R(sus_light_grid, 48))))
    return s;
  // k_prepare also empties the centroid bounds and the light's suspect count, and
  // k_tri_setup clears the refit counters
  pa.bounds = bounds.as<uint32_t>();
  pa.zero = light_sus ? sus_light_cnt.as<uint32_t>() : nullptr;
  timed("k_prepare", [&] { launch_prepare(pa, stream); });
  // with the light's suspect list (k_prepare zeroed its count)
  SuspectArgs la{F, 1, light.L.x, light.L.y, light.L.z, bf_a.as<float4>(), bf_e1.as<float4>(), bf_e2.as<float4>(),
                 light_sus ? sus_light.as<float4>() : nullptr, light_sus ? sus_light_cnt.as<int>() : nullptr};
  TriSetupArgs ta{F, tri_mesh.as<int>(), dmats.as<float>(), idx.as<uint32_t>(), pos.as<float>(), bf_a.as<float4>(),
                  bf_e1.as<float4>(), bf_e2.as<float4>(), box_lo.as<float4>(), box_hi.as<float4>(), cen.as<float4>(),
                  bounds.as<uint32_t>(), F > 1 ? counter.as<uint32_t>() : nullptr, la};
  timed("k_tri_setup", [&] { launch_tri_setup(ta, stream); });
  if (!topo_valid) full = true;
  if (full) {
  MortonArgs ma{F, cen.as<float4>(), bounds.as<uint32_t>(), keys0.as<uint32_t>(), vals0.as<uint32_t>()};
  timed("k_morton", [&] { launch_morton(ma, stream); });
  for (int pass = 0; pass < 4; pass++) {
    const bool even = (pass & 1) == 0;
    const uint32_t* kin = even ? keys0.as<uint32_t>() : keys1.as<uint32_t>();
    const uint32_t* vin = even ? vals0.as<uint32_t>() : vals1.as<uint32_t>();
    uint32_t* kout = even ? keys1.as<uint32_t>() : keys0.as<uint32_t>();
    uint32_t* vout = even ? vals1.as<uint32_t>() : vals0.as<uint32_t>();
    timed("k_radix_sort", [&] { launch_radix_pass(kin, vin, kout, vout, F, 8 * pass, hist.as<uint32_t>(), stream); });
  }
  // after 4 passes the sorted keys / triangle order are back in keys0 / vals0
  if (F > 1) {
    KarrasArgs ka{F, keys0.as<uint32_t>(), nrange.as<int2>(), child.as<int>(), node_parent.as<int>(), leaf_parent.as<int>()};
    timed("k_karras", [&] { launch_karras(ka, stream); });
    if (F > REFIT_CHUNK) {
      SpanArgs sa{F, nrange.as<int2>(), child.as<int>(), node_parent.as<int>(), span_of.as<int>(), span_p.as<int>(),
                  span.as<int4>(), span_cid.as<int2>(), span_n.as<uint32_t>()};
      timed("k_span", [&] { launch_span(sa, stream); });
    }
  }
  }  // full
  topo_valid = true;
  if (F > 1) {
    RefitArgs ra{F, vals0.as<uint32_t>(), box_lo.as<float4>(), box_hi.as<float4>(), child.as<int>(),
                 node_parent.as<int>(), leaf_parent.as<int>(), nrange.as<int2>(), counter.as<uint32_t>(),
                 frontier.as<int>(), leaf_lo.as<float4>(), leaf_hi.as<float4>(), node_lo.as<float4>(),
                 node_hi.as<float4>(), span.as<int4>(), span_cid.as<int2>(),
                 F > REFIT_CHUNK ? span_n.as<uint32_t>() : nullptr};
    timed("k_refit", [&] { launch_refit(ra, stream); });
  }
  // node packing (F > 1) and the triangles in leaf order, one launch
  // the packing launch also clears the views' suspect counters for the next list build
  // (k_pack_nodes has at least 256 threads)
  const int nzero = (int)std::min<size_t>(sus_view_cnt.bytes / 4, 256);
  PackArgs pk{F, bounds.as<uint32_t>(), child.as<int>(), leaf_lo.as<float4>(), leaf_hi.as<float4>(), node_lo.as<float4>(),
              node_hi.as<float4>(), nodes.as<float4>(), nzero > 0 ? sus_view_cnt.as<uint32_t>() : nullptr, nzero};
  view_cnt_zeroed = nzero;
  ReorderArgs ro{F, vals0.as<uint32_t>(), bf_a.as<float4>(), bf_e1.as<float4>(), bf_e2.as<float4>(),
                 tri_a.as<float4>(), tri_e1.as<float4>(), tri_e2.as<float4>()};
  timed("k_pack_nodes", [&] { launch_pack_reorder(pk, ro, stream); });
  if (light_sus && F >= SUS_BIN_MIN_F) {
    // the light's suspects over a grid across the light (needs the tree's root boxes)
    timed("k_fill_u32", [&] { launch_fill_u32(sus_light_cell_cnt.as<uint32_t>(), LIGHT_GRID * LIGHT_GRID, 0u, stream); });
    LightBinArgs lb{sus_light.as<float4>(), sus_light_cnt.as<int>(), bf_e1.as<float4>(), nodes.as<float4>(),
                    light.L.x, light.L.y, light.L.z, sus_light_grid.as<float4>(),
                    BinPool{sus_light_pool.as<float4>(), LIGHT_POOL, sus_light_cell.as<int2>(), sus_light_cell_cnt.as<int>(),
                            LIGHT_GRID, LIGHT_GRID}};
    timed("k_light_bin", [&] { launch_light_bin(lb, stream); });
  }
  return hip(hipGetLastError(), "commit kernels");
}

drh_status drh_renderer::set_cams(const drh_camera* c, int views, long long* total) {
  if (!c || views <= 0) return fail(DRH_E_INVALID, "no cameras");
  std::vector<DevCam> dc(views);
  long long off = 0;
  for (int v = 0; v < views; v++) {
    if (!(c[v].width >= 1.0f) || !(c[v].height >= 1.0f) || c[v].width > 65536.0f || c[v].height > 65536.0f)
      return fail(DRH_E_INVALID, "camera %d: bad size %g x %g", v, c[v].width, c[v].height);
    const M4 wv = m4_load(c[v].world_view), pr = m4_load(c[v].proj);
    const M4 pi = m4_inverse(pr), wi = m4_inverse(wv), wvp = m4_mul(pr, wv);
    std::memcpy(dc[v].proj_inv, pi.m, 64);
    std::memcpy(dc[v].wv_inv, wi.m, 64);
    std::memcpy(dc[v].wvp, wvp.m, 64);
    dc[v].W = c[v].width; dc[v].H = c[v].height;
    dc[v].w = (int)c[v].width; dc[v].h = (int)c[v].height;
    dc[v].img_off = off;
    off += 3ll * dc[v].w * dc[v].h;
  }
  *total = off;
  const bool same = cam_cache.size() == dc.size() && std::memcmp(cam_cache.data(), dc.data(), sizeof(DevCam) * dc.size()) == 0;
  if (!same) {
    drh_status s;
    if ((s = hip(cams.reserve(sizeof(DevCam) * views), "alloc"))) return s;
    if ((s = stage_begin(stage_round(sizeof(DevCam) * views))) ||
        (s = stage_send(cams.p, dc.data(), sizeof(DevCam) * views)) || (s = stage_end()))
      return s;
    cam_cache = dc;
    views_built = -1;
  }
  return DRH_OK;
}

drh_status drh_renderer::run_render(int views, float* d_img, const float* d_target, float* d_adj, float* d_loss,
                                    bool store_hits) {
  drh_status s0;
  if ((s0 = build_view_suspects(views))) return s0;
  hits_stored = false;
  if (store_hits) {
    const hipError_t e = hitbuf.reserve(hit_bytes(views));
    if (e != hipSuccess) {
      // the automatic policy recomputes when the memory is not there; a forced store fails
      if (interior_store != DRH_HITS_AUTO || e != hipErrorOutOfMemory) return hip(e, "alloc");
      (void)hipGetLastError();
      release_uncaptured(hitbuf);
      store_hits = false;
    }
  }
  int part = 0;
  std::vector<int> tiles(views);
  for (int v = 0; v < views; v++) { tiles[v] = num_tiles(sq, cam_cache[v].w, cam_cache[v].h); part += tiles[v]; }
  if (d_target) {
    drh_status s;
    if ((s = hip(loss_part.reserve(8 * (size_t)std::max(part, 1)), "alloc"))) return s;
  }
  int part_off = 0;
  for (int v = 0; v < views; v++) {
    RenderArgs ra{tracer_view(v), light, cams.as<DevCam>(), v, sq, d_img, d_target, d_adj, loss_part.as<double>(), part_off,
                  store_hits ? hitbuf.as<float4>() : nullptr, jitter()};
    ra.tr.rays = ray_slot(0);
    timed("k_render", [&] { launch_render(st.mode, ra, tiles[v], stream); });
    part_off += tiles[v];
  }
  if (d_target && d_loss) {
    double denom = 0;
    for (int v = 0; v < views; v++) denom += (double)cam_cache[0].w * cam_cache[0].h;
    if (loss_on_side) {   // nothing on the main stream waits for the loss before d_render ends
      drh_status s;
      if ((s = hip(hipEventRecord(ev_lfork, stream), "fork")) || (s = hip(hipStreamWaitEvent(side, ev_lfork, 0), "fork")))
        return s;
      launch_loss_final(loss_part.as<double>(), part, d_loss, denom, side);
      loss_side_pending = true;
    } else {
      timed("k_loss_final", [&] { launch_loss_final(loss_part.as<double>(), part, d_loss, denom, stream); });
    }
  }
  hits_stored = store_hits;
  return hip(hipGetLastError(), "render kernels");
}

// Edge-sample tables on the device, shared read-only by every handle of the device (a table
// depends only on the random-number mode and N): uploaded once, kept until the process ends,
// SHARED_RND_BYTES of them at most per device (beyond that a handle uploads its own).
constexpr size_t SHARED_RND_BYTES = 512ull << 20;
const float2* shared_sample_table(int device, int mode, long long N) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, long long>, float2*> tables;
  static std::map<int, size_t> used;
  std::lock_guard<std::mutex> lk(mu);
  const auto key = std::make_tuple(device, mode, N);
  auto it = tables.find(key);
  if (it != tables.end()) return it->second;
  const size_t bytes = 8 * (size_t)N;
  if (used[device] + bytes > SHARED_RND_BYTES) return nullptr;
  const auto t = sorted_sample_table(mode, N);
  float2* p = nullptr;
  if (hipMalloc((void**)&p, bytes) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
  if (hipMemcpy(p, t->data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(p);
    return nullptr;
  }
  used[device] += bytes;
  tables.emplace(key, p);
  return p;
}

drh_status drh_renderer::ensure_rnd(long long N) {
  if (N == rnd_n) return DRH_OK;
  if (N > 0x7FFFFFFFll) return fail(DRH_E_INVALID, "edge_samples too large");
  drh_status s;
  rnd_ptr = nullptr;
  if (N > 0 && (rnd_ptr = shared_sample_table(st.device, st.rng_mode, N)) == nullptr) {
    if ((s = hip(rnd.reserve(8 * (size_t)N), "alloc"))) return s;
    const auto t = sorted_sample_table(st.rng_mode, N);
    if ((s = stage_begin(stage_round(8 * (size_t)N))) || (s = stage_send(rnd.p, t->data(), 8 * (size_t)N)) ||
        (s = stage_end()))
      return s;
    rnd_ptr = rnd.as<float2>();
  }
  rnd_n = N;
  return DRH_OK;
}

drh_status drh_renderer::run_d_render(int views, const float* d_adj, long long N, float* d_grad, bool accumulate) {
  drh_status s;
  if ((s = ensure_rnd(N))) return s;
  if ((s = build_view_suspects(views))) return s;
  const bool zeroed_on_side = grad_zeroed && !accumulate;   // cleared beside the render (fork_prep0)
  grad_zeroed = false;
  if (!accumulate && !zeroed_on_side) timed("k_fill_f32", [&] { launch_fill_f32(d_grad, G, 0.0f, stream); });
  // with restricted meshes the edge pass gathers every mesh's 12 transform entries apart,
  // then k_restricted_grad maps them into the gradient blocks
  const bool tr_apart = any_restricted && E > 0 && N > 0;
  if (tr_apart) timed("k_fill_f32", [&] { launch_fill_f32(trscr.as<float>(), 12ll * nm, 0.0f, stream); });
  // deterministic mode: every gradient term into 32.32 fixed-point integer sums (G floats,
  // then the 12 transform entries per mesh gathered apart), converted once at the end
  unsigned long long* dacc_p = nullptr;
  if (deterministic && G > 0) {
    const size_t nacc = (size_t)G + 12 * (size_t)nm;
    if ((s = hip(dacc.reserve(8 * nacc), "alloc"))) return s;
    dacc_p = dacc.as<unsigned long long>();
    timed("k_fill_u32", [&] { launch_fill_u32(dacc.as<uint32_t>(), 2ll * (long long)nacc, 0u, stream); });
  }
  // single differentiable mesh: position / transform gradient offsets
  int single = -1;
  for (int i = 0; i < nm; i++) if (diff[i]) { single = (single == -1) ? i : -2; }
  // Edge samples in pixel-bin order when there are few samples per edge (rnd0 order then
  // scatters a wavefront's rays over distant edges); buffers sized before any launch.
  // Deterministic mode keeps the rnd0 slot order whatever was asked: a sample's rank inside
  // its bin comes from an atomic counter (wave arrival order), which would regroup the float
  // run sums of k_edge_grad from run to run.
  const bool binned = E > 0 && N > 0 && !deterministic &&
                      (edge_order == DRH_EDGE_ORDER_BINNED ||
                       (edge_order == DRH_EDGE_ORDER_AUTO && N < (long long)EDGE_BIN_AUTO_PER_EDGE * E));
  if (E > 0 && N > 0 && (s = hip(ecol.reserve(16 * (size_t)N), "alloc"))) return s;
  {
    int max_tiles = 1;
    for (int v = 0; v < views; v++) max_tiles = std::max(max_tiles, num_tiles(sq, cam_cache[v].w, cam_cache[v].h));
    if ((s = hip(cov_part.reserve(16 * (size_t)max_tiles), "alloc"))) return s;
  }
  if (binned) {
    int max_bins = 1;
    for (int v = 0; v < views; v++) {
      const DevCam& c = cam_cache[v];
      max_bins = std::max(max_bins, ((c.w + EDGE_BIN - 1) / EDGE_BIN) * ((c.h + EDGE_BIN - 1) / EDGE_BIN));
    }
    if ((s = hip(bin_count.reserve(4 * ((size_t)max_bins + 1)), "alloc")) ||
        (s = hip(bin_rec.reserve(16 * (size_t)N), "alloc")) || (s = hip(bin_order.reserve(8 * (size_t)N), "alloc")))
      return s;
  }
  // The interior pass reads the adjoint and the scene and adds to other gradient entries than
  // the edge pass (or to the same ones by atomics): with the device to itself the handle runs
  // it on the second stream beside the edge pass.  Not while profiling (per-kernel events time
  // the main stream) or when other handles share the device (their work overlaps instead).
  const bool fork_int = !prof && sharing == 1 && st.mode != DRH_SILHOUETTE && G > 0 &&
                        (!skip_tex_interior || (interior_pos && (st.mode == DRH_LAMBERT || st.mode == DRH_LAMBERT_VERTEX_COLOR)));
  hipStream_t iq = stream;
  if (fork_int) {   // the side stream cleared the gradient before this point in its order
    if ((s = hip(hipEventRecord(ev_ifork, stream), "fork")) || (s = hip(hipStreamWaitEvent(side, ev_ifork, 0), "fork")))
      return s;
    iq = side;
  } else if (zeroed_on_side && (s = hip(hipStreamWaitEvent(stream, ev_join, 0), "join"))) {
    return s;   // an interior pass on the main stream writes the gradient: the clear comes first
  }
  for (int v = 0; v < views; v++) {
    const DevCam& c = cam_cache[v];
    const int tiles = num_tiles(sq, c.w, c.h);   // interior samples: nsamp = sq*sq Hammersley points
    if (st.mode != DRH_SILHOUETTE && G > 0 && !skip_tex_interior) {
      InteriorArgs ia{};
      ia.tr = tracer_view(v); ia.tr.rays = ray_slot(1); ia.cams = cams.as<DevCam>(); ia.view = v; ia.nsamp = nsamp; ia.sq = sq; ia.hamm = hamm.as<float>();
      ia.adj = d_adj; ia.grad = d_grad; ia.layout = dlayout.as<int>();
      ia.lds_base = 0; ia.lds_n = 0; ia.uniform = 0; ia.uniform_off = 0;
      ia.aligned = interior_aligned;
      ia.jit = jitter();
      ia.light = light;
      ia.hits = (interior_store && hits_stored) ? hitbuf.as<float4>() : nullptr;
      if (interior_pos && st.mode == DRH_VERTEX_COLOR) {
        ia.pos_grad = 1; ia.tr_inv0 = dtr_inv0.as<float>();
        ia.hits = nullptr;   // the position term needs each sample's ray: traced again
      }
      int blocks = tiles;
      if ((st.mode == DRH_TEXTURE_COLOR || st.mode == DRH_LAMBERT) && nm == 1 && diff[0] && tex_w[0] * tex_h[0] == 1) {
        ia.uniform = 1; ia.uniform_off = layout[2];
        ia.coverage = tc_zero ? 1 : 0;
        blocks = std::min(tiles, 4096);
      } else if (G <= INTERIOR_LDS_FLOATS) {
        ia.lds_base = 0; ia.lds_n = G;
        blocks = std::min(tiles, 1024);
      } else {
        // regions of several tiles only when the samples come from the stored hits: a wave
        // that traces walks its region's tiles one after another, and one tile per wave
        // keeps as many packets in flight as the render
        ia.hash = INTERIOR_HASH;
        ia.region = (ia.hits && (sq == 1 || sq == 2 || sq == 4 || sq == 8)) ? INTERIOR_REGION : 1;
        blocks = num_regions(sq, c.w, c.h, ia.region);
      }
      if (ia.region == 0) ia.region = 1;
      if (dacc_p) {
        // no LDS sums or hash tables (their float atomics would make the order matter)
        ia.dacc = dacc_p; ia.lds_base = 0; ia.lds_n = 0; ia.hash = 0; ia.region = 1;
        if (!ia.uniform) blocks = tiles;
      }
      if (ia.uniform && ia.coverage)
        timed("k_interior", [&] { launch_coverage(ia, tiles, cov_part.as<float4>(), iq); });
      else
        timed("k_interior", [&] { launch_interior(st.mode, ia, blocks, iq); });
    }
    if (interior_pos && (st.mode == DRH_LAMBERT || st.mode == DRH_LAMBERT_VERTEX_COLOR) && G > 0) {
      // the Lambert interior position term: a pass of its own over the same samples
      InteriorArgs pa{};
      pa.tr = tracer_view(v); pa.tr.rays = ray_slot(1); pa.cams = cams.as<DevCam>(); pa.view = v;
      pa.nsamp = nsamp; pa.sq = sq; pa.hamm = hamm.as<float>(); pa.adj = d_adj; pa.grad = d_grad;
      pa.layout = dlayout.as<int>(); pa.aligned = interior_aligned; pa.jit = jitter(); pa.light = light;
      pa.hits = (interior_store && hits_stored) ? hitbuf.as<float4>() : nullptr;
      pa.tr_inv0 = dtr_inv0.as<float>(); pa.dacc = dacc_p;
      timed("k_interior_pos", [&] { launch_interior_pos(st.mode, pa, tiles, iq); });
    }
    if (E > 0 && N > 0) {
      if (v == 0 && prep0_forked) {
        if ((s = hip(hipStreamWaitEvent(stream, ev_join, 0), "join"))) return s;   // prepared beside the render
      } else {
        edge_prep(v, stream);
      }
      EdgeArgs ea{};
      ea.tr = tracer_view(v); ea.tr.rays = ray_slot(2); ea.light = light; ea.cams = cams.as<DevCam>(); ea.view = v; ea.N = N; ea.E = E;
      ea.tr.s.lane_walk = lane_walk_edge(N);
      ea.edges = edges.as<int2>(); ea.edge_mesh = edge_mesh.as<int>(); ea.pmf = pmf.as<float>(); ea.cdf = cdf.as<float>();
      ea.proj = proj.as<float4>();
      ea.rnd = rnd_ptr;
      ea.adj = d_adj; ea.grad = d_grad; ea.layout = dlayout.as<int>();
      ea.tr_inv0 = dtr_inv0.as<float>();
      ea.lds_base = 0; ea.lds_n = 0;
      if (single >= 0 && 3 * (vo[single + 1] - vo[single]) <= EDGE_LDS_FLOATS) {
        ea.lds_base = layout[4 * single]; ea.lds_n = 3 * (vo[single + 1] - vo[single]);
      }
      if (binned) {
        EdgeBinArgs ba{};
        ba.cams = cams.as<DevCam>(); ba.view = v; ba.N = N; ba.E = E;
        ba.edges = edges.as<int2>(); ba.cdf = cdf.as<float>(); ba.proj = proj.as<float4>();
        ba.rnd = rnd_ptr; ba.adj = d_adj;
        ba.bins_x = (c.w + EDGE_BIN - 1) / EDGE_BIN;
        ba.nbins = ba.bins_x * ((c.h + EDGE_BIN - 1) / EDGE_BIN);
        ba.bin_count = bin_count.as<int>(); ba.rec = bin_rec.as<int4>(); ba.order = bin_order.as<int2>();
        timed("k_fill_u32", [&] { launch_fill_u32(bin_count.as<uint32_t>(), ba.nbins + 1, 0u, stream); });
        timed("k_edge_bin", [&] { launch_edge_bin(ba, stream); });
        ea.order = ba.order; ea.nkept = ba.bin_count + ba.nbins;
      }
      ea.single_mesh = single >= 0 ? 1 : 0;
      ea.full_jacobian = edge_grad_mode == DRH_EDGE_GRAD_FULL_JACOBIAN ? 1 : 0;
      ea.watertight = side_test == DRH_SIDE_TEST_WATERTIGHT ? 1 : 0;
      ea.single_tr_off = single >= 0 ? layout[4 * single] + layout[4 * single + 3] : 0;
      ea.single_id = single >= 0 ? single : 0;
      ea.trscr = tr_apart ? trscr.as<float>() : nullptr;
      ea.dacc = dacc_p; ea.dG = G;
      if (dacc_p) { ea.lds_base = 0; ea.lds_n = 0; }
      ea.ecol = ecol.as<float4>();
      timed("k_edge_trace", [&] { launch_edge_trace(st.mode, ea, stream); });
      timed("k_edge_grad", [&] { launch_edge_grad(ea, stream); });
    }
  }
  if (fork_int &&
      ((s = hip(hipEventRecord(ev_ijoin, side), "join")) || (s = hip(hipStreamWaitEvent(stream, ev_ijoin, 0), "join"))))
    return s;
  if (loss_side_pending) {
    loss_side_pending = false;
    if ((s = hip(hipEventRecord(ev_ljoin, side), "join")) || (s = hip(hipStreamWaitEvent(stream, ev_ljoin, 0), "join")))
      return s;
  }
  if (dacc_p)
    timed("k_det_finish", [&] {
      launch_det_finish(dacc_p, G, d_grad, tr_apart ? trscr.as<float>() : nullptr, tr_apart ? 12 * nm : 0, stream);
    });
  if (tr_apart) {
    RestrictedArgs ra{nm, dlayout.as<int>(), drflag.as<int>(), drparams.as<float>(), trscr.as<float>(), d_grad};
    timed("k_restricted_grad", [&] { launch_restricted_grad(ra, stream); });
  }
  return hip(hipGetLastError(), "d_render kernels");
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

static bool settings_valid(const drh_settings* s) {
  return s && s->spp >= 1 && (s->mode >= DRH_SILHOUETTE && (s->mode <= DRH_LAMBERT || s->mode == DRH_LAMBERT_VERTEX_COLOR)) &&
         (s->rng_mode == DRH_RNG_REF_PRNG14 || s->rng_mode == DRH_RNG_QMC) &&
         (s->tracer == DRH_TRACER_BVH || s->tracer == DRH_TRACER_BF);
}

// settings, sample grid and light (src/shade_lambert.cpp:10-12 unless a custom one is given)
static void apply_settings(drh_renderer& r, const drh_settings& settings) {
  r.st = settings;
  r.sq = (int)std::sqrt((float)settings.spp);
  r.nsamp = r.sq * r.sq;
  if (settings.light_custom) {
    r.light.L = {settings.light_dir[0], settings.light_dir[1], settings.light_dir[2]};
    r.light.Lc = {settings.light_color[0], settings.light_color[1], settings.light_color[2]};
    r.light.amb = {settings.ambient[0], settings.ambient[1], settings.ambient[2]};
  } else {
    const float x = -1, y = -1, z = 0;
    const float li = 1.0f / std::sqrt(x * x + y * y + z * z);
    r.light.L = {x * li, y * li, z * li};
    r.light.Lc = {1.0f, 0.5f, 0.2f};
    r.light.amb = {0.1f * 1.0f, 0.1f * 1.0f, 0.1f * 1.0f};
  }
}

// A handle's streams, events and buffers kept, its settings replaced: what drh_create would
// give for `settings` on the same device, once the caller commits a scene (the server hands
// an idle handle to a job with other settings this way).
drh_status drh_reset_settings(drh_handle h, const drh_settings* settings) {
  if (!h || !settings_valid(settings) || settings->device != h->st.device) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  const bool spp_changed = settings->spp != h->st.spp;
  const bool rng_changed = settings->rng_mode != h->st.rng_mode;
  h->drop_graph();
  h->committed = false;
  h->hits_stored = false;
  // every option back to drh_create's value
  h->opt_kind = -1;
  h->opt_iter = 0;
  h->pos_lr = 0.1f, h->col_lr = 0.1f, h->tex_lr = 0.0f, h->tr_lr = 0.0f;
  h->skip_tex_interior = false;
  h->sharing = 1;
  h->interior_aligned = 0, h->interior_store = 0;
  h->hit_limit = ~(size_t)0;
  h->jit_on = 0, h->jit_seed = 0, h->jit_host_step = 0;
  h->edge_order = DRH_EDGE_ORDER_AUTO;
  h->edge_grad_mode = DRH_EDGE_GRAD_REFERENCE_DIAGONAL;
  h->side_test = DRH_SIDE_TEST_REFERENCE;
  h->traversal = DRH_TRAVERSAL_AUTO;
  h->deterministic = false;
  h->interior_pos = false;
  h->use_graph = true;
  h->prof = false;
  h->count_on = false;
  h->rebuild_every = 8;
  apply_settings(*h, *settings);
  if (rng_changed) h->rnd_n = -1;
  if (spp_changed) {
    std::vector<float> hm;
    hammersley(settings->spp, hm);
    drh_status s;
    if ((s = h->hip(h->hamm.reserve(4 * hm.size()), "alloc")) || (s = h->stage_begin(drh_renderer::stage_round(4 * hm.size()))) ||
        (s = h->stage_send(h->hamm.p, hm.data(), 4 * hm.size())) || (s = h->stage_end()))
      return s;
  }
  return DRH_OK;
}

drh_status drh_create(const drh_settings* settings, drh_handle* out) {
  if (!settings || !out) return DRH_E_INVALID;
  *out = nullptr;
  if (!settings_valid(settings)) return DRH_E_INVALID;
  std::unique_ptr<drh_renderer> r(new drh_renderer());
  t_buf_owner = nullptr;   // later buffers of this thread belong to no handle
  r->st = *settings;
  if (hipSetDevice(settings->device) != hipSuccess) return DRH_E_HIP;
  if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) return DRH_E_HIP;
  // the allocation-free CDF's scratch address on this device (looked up once per device)
  if (noalloc_scratch_init(settings->device) != hipSuccess) return DRH_E_HIP;
  keep_pool_memory(settings->device);
  r->pool_stream = r->stream;
  if (hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&r->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&r->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&r->ev_ifork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&r->ev_ijoin, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&r->ev_lfork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&r->ev_ljoin, hipEventDisableTiming) != hipSuccess)
    return DRH_E_HIP;
  apply_settings(*r, *settings);
  // queued on the new stream, no wait: in a busy process a synchronisation here waits for
  // whatever shares the stream's hardware queue
  std::vector<float> hm;
  hammersley(settings->spp, hm);
  if (r->hamm.reserve(4 * hm.size()) != hipSuccess || r->loss.reserve(16) != hipSuccess ||
      r->fault.reserve(16) != hipSuccess)
    return DRH_E_OOM;
  if (r->stage_begin(drh_renderer::stage_round(4 * hm.size())) || r->stage_send(r->hamm.p, hm.data(), 4 * hm.size()) ||
      r->stage_end() || hipMemsetAsync(r->fault.p, 0, 16, r->stream) != hipSuccess)
    return DRH_E_HIP;
  *out = r.release();
  return DRH_OK;
}

void drh_destroy(drh_handle h) {
  if (!h) return;
  (void)hipSetDevice(h->st.device);
  (void)hipStreamSynchronize(h->stream);
  h->destroy_graphs();
  h->stage_free();
  for (auto& p : h->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto e : h->event_pool) (void)hipEventDestroy(e);
  (void)hipStreamSynchronize(h->side);
  (void)hipEventDestroy(h->ev_fork);
  (void)hipEventDestroy(h->ev_join);
  (void)hipEventDestroy(h->ev_ifork);
  (void)hipEventDestroy(h->ev_ijoin);
  (void)hipEventDestroy(h->ev_lfork);
  (void)hipEventDestroy(h->ev_ljoin);
  hipStream_t side = h->side, main = h->stream;
  delete h;   // pooled buffers are freed on the main stream, which still exists
  (void)hipStreamDestroy(side);
  (void)hipStreamDestroy(main);
}

const char* drh_last_error(drh_handle h) { return h ? h->err.c_str() : "null handle"; }

drh_status drh_commit(drh_handle h, const drh_scene_desc* scene) {
  if (!h) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  h->committed = false;
  h->topo_valid = false;
  drh_status s = h->upload_scene(scene);
  if (s) return s;
  if ((s = h->build_bvh())) return s;
  h->committed = true;
  h->opt_kind = -1;  // optimiser state belongs to the previous scene
  return DRH_OK;
}

int drh_grad_size(drh_handle h) { return h ? h->G : -1; }

drh_status drh_render_dev(drh_handle h, const drh_camera* cams, int views, float* d_images) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "render: renderer was not prepared for this scene (commit first)");
  if (!d_images) return h->fail(DRH_E_INVALID, "render: images NULL");
  (void)hipSetDevice(h->st.device);
  long long total;
  drh_status s;
  if ((s = h->set_cams(cams, views, &total))) return s;
  return h->run_render(views, d_images, nullptr, nullptr, nullptr);
}

drh_status drh_render(drh_handle h, const drh_camera* cams, int views, float* images) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "render: renderer was not prepared for this scene (commit first)");
  if (!images) return h->fail(DRH_E_INVALID, "render: images NULL");
  (void)hipSetDevice(h->st.device);
  long long total;
  drh_status s;
  if ((s = h->set_cams(cams, views, &total))) return s;
  if ((s = h->hip(h->img.reserve(4 * (size_t)total), "alloc"))) return s;
  if ((s = h->run_render(views, h->img.as<float>(), nullptr, nullptr, nullptr))) return s;
  if ((s = h->hip(hipMemcpyAsync(images, h->img.p, 4 * (size_t)total, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  return h->check_fault();
}

drh_status drh_d_render_dev(drh_handle h, const drh_camera* cams, int views, const float* d_adj, int edge_samples,
                            float* d_grad, int accumulate) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "d_render: renderer was not prepared for this scene (commit first)");
  if (!d_adj || !d_grad || edge_samples < 0) return h->fail(DRH_E_INVALID, "d_render: bad arguments");
  (void)hipSetDevice(h->st.device);
  long long total;
  drh_status s;
  if ((s = h->set_cams(cams, views, &total))) return s;
  return h->run_d_render(views, d_adj, edge_samples, d_grad, accumulate != 0);
}

drh_status drh_d_render(drh_handle h, const drh_camera* cams, int views, const float* adjoints, int edge_samples,
                        float* grad, int accumulate) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "d_render: renderer was not prepared for this scene (commit first)");
  if (!adjoints || !grad || edge_samples < 0) return h->fail(DRH_E_INVALID, "d_render: bad arguments");
  (void)hipSetDevice(h->st.device);
  long long total;
  drh_status s;
  if ((s = h->set_cams(cams, views, &total))) return s;
  if ((s = h->hip(h->adj.reserve(4 * (size_t)total), "alloc")) || (s = h->hip(h->grad.reserve(4 * (size_t)std::max(h->G, 1)), "alloc")))
    return s;
  if ((s = h->hip(hipMemcpyAsync(h->adj.p, adjoints, 4 * (size_t)total, hipMemcpyHostToDevice, h->stream), "upload"))) return s;
  if (accumulate && h->G > 0 &&
      (s = h->hip(hipMemcpyAsync(h->grad.p, grad, 4 * (size_t)h->G, hipMemcpyHostToDevice, h->stream), "upload")))
    return s;
  if ((s = h->run_d_render(views, h->adj.as<float>(), edge_samples, h->grad.as<float>(), accumulate != 0))) return s;
  if (h->G > 0 && (s = h->hip(hipMemcpyAsync(grad, h->grad.p, 4 * (size_t)h->G, hipMemcpyDeviceToHost, h->stream), "download")))
    return s;
  return h->check_fault();
}

drh_status drh_d_render_and_compare_dev(drh_handle h, const drh_camera* cams, int views, const float* d_targets,
                                        int edge_samples, float* d_grad, float* d_images, float* d_loss) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "d_render_and_compare: commit first");
  if (!d_targets || !d_grad || edge_samples < 0) return h->fail(DRH_E_INVALID, "d_render_and_compare: bad arguments");
  (void)hipSetDevice(h->st.device);
  long long total;
  drh_status s;
  if ((s = h->set_cams(cams, views, &total))) return s;
  for (int v = 1; v < views; v++)
    if (h->cam_cache[v].w != h->cam_cache[0].w || h->cam_cache[v].h != h->cam_cache[0].h)
      return h->fail(DRH_E_INVALID, "d_render_and_compare: views must share one resolution");
  if ((s = h->hip(h->adj.reserve(4 * (size_t)total), "alloc"))) return s;
  float* img = d_images;
  if (!img) {
    if ((s = h->hip(h->img.reserve(4 * (size_t)total), "alloc"))) return s;
    img = h->img.as<float>();
  }
  float* loss = d_loss ? d_loss : h->loss.as<float>();
  // the views' suspect lists (the render's first kernels) before the fork, so that the edge
  // preparation's branch starts beside the render itself
  if ((s = h->ensure_rnd(edge_samples)) || (s = h->build_view_suspects(views)) ||
      (s = h->fork_prep0(edge_samples, d_grad, h->G)))
    return s;
  bool store = h->interior_store && h->interior_aligned && !h->skip_tex_interior && h->st.mode != DRH_SILHOUETTE;
  if (store && h->interior_store == DRH_HITS_AUTO && h->hit_bytes(views) > h->hit_limit) {
    store = false;
    h->release_uncaptured(h->hitbuf);   // recompute: hold no hit buffer beyond the limit
  }
  h->loss_on_side = h->prep0_forked;
  s = h->run_render(views, img, d_targets, h->adj.as<float>(), loss, store);
  h->loss_on_side = false;
  if (s) {
    h->prep0_forked = h->grad_zeroed = h->loss_side_pending = false;
    return s;
  }
  s = h->run_d_render(views, h->adj.as<float>(), edge_samples, d_grad, false);
  h->prep0_forked = false;
  h->grad_zeroed = false;
  h->hits_stored = false;
  return s;
}

drh_status drh_d_render_and_compare(drh_handle h, const drh_camera* cams, int views, const float* targets,
                                    int edge_samples, float* grad, float* out_images, float* loss) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "d_render_and_compare: commit first");
  if (!targets || !grad || edge_samples < 0) return h->fail(DRH_E_INVALID, "d_render_and_compare: bad arguments");
  (void)hipSetDevice(h->st.device);
  long long total;
  drh_status s;
  if ((s = h->set_cams(cams, views, &total))) return s;
  if ((s = h->hip(h->tgt.reserve(4 * (size_t)total), "alloc")) || (s = h->hip(h->img.reserve(4 * (size_t)total), "alloc")) ||
      (s = h->hip(h->grad.reserve(4 * (size_t)std::max(h->G, 1)), "alloc")))
    return s;
  if ((s = h->hip(hipMemcpyAsync(h->tgt.p, targets, 4 * (size_t)total, hipMemcpyHostToDevice, h->stream), "upload"))) return s;
  if ((s = drh_d_render_and_compare_dev(h, cams, views, h->tgt.as<float>(), edge_samples, h->grad.as<float>(),
                                        h->img.as<float>(), h->loss.as<float>())))
    return s;
  if (h->G > 0 && (s = h->hip(hipMemcpyAsync(grad, h->grad.p, 4 * (size_t)h->G, hipMemcpyDeviceToHost, h->stream), "download")))
    return s;
  if (out_images && (s = h->hip(hipMemcpyAsync(out_images, h->img.p, 4 * (size_t)total, hipMemcpyDeviceToHost, h->stream), "download")))
    return s;
  float l = 0.0f;
  if ((s = h->hip(hipMemcpyAsync(&l, h->loss.p, 4, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  if ((s = h->check_fault())) return s;
  if (loss) *loss = l;
  return DRH_OK;
}

drh_status drh_opt_init(drh_handle h, int kind, float pos_lr, float col_lr) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "opt_init: commit first");
  if (kind != DRH_OPT_SIMPLE && kind != DRH_OPT_ADAM) return h->fail(DRH_E_INVALID, "opt_init: kind");
  (void)hipSetDevice(h->st.device);
  drh_status s;
  const size_t n = 4 * (size_t)std::max(h->G, 1);
  if ((s = h->hip(h->opt_m.reserve(n), "alloc")) || (s = h->hip(h->opt_v.reserve(n), "alloc")) ||
      (s = h->hip(h->grad.reserve(n), "alloc")))
    return s;
  if ((s = h->hip(h->opt_t.reserve(16), "alloc"))) return s;
  launch_fill_f32(h->opt_m.as<float>(), h->G, 0.0f, h->stream);
  launch_fill_f32(h->opt_v.as<float>(), h->G, 0.0f, h->stream);
  launch_fill_u32(h->opt_t.as<uint32_t>(), 1, 1u, h->stream);  // first Adam step is t = 1
  h->drop_graph();
  h->opt_kind = kind;
  h->pos_lr = pos_lr;
  h->col_lr = col_lr;
  h->opt_iter = 0;
  return h->hip(hipGetLastError(), "opt_init");
}

// One optimisation step, every launch on the handle's stream (capturable: no allocation,
// upload or synchronisation once the first step has sized every buffer).
static drh_status opt_step_body(drh_handle h, const drh_camera* cams, int views, const float* d_targets,
                                int edge_samples, float* d_loss) {
  drh_status s;
  // re-commit from the device parameters (the optimiser moved them)
  // (k_prepare then also advances Adam's device step counter: step t, from 1, is in place for
  // this step's jitter pattern and Adam update; the first step uses the commit's t = 1)
  if (h->opt_iter > 0) {
    h->prep_step = h->opt_kind == DRH_OPT_ADAM ? h->opt_t.as<int>() : nullptr;
    s = h->build_bvh(h->opt_iter % h->rebuild_every == 0);   // a fresh commit's tree is current
    h->prep_step = nullptr;
    if (s) return s;
  }
  // TEXTURE / LAMBERT interior adjoints reach only the texture block (shade_grad,
  // src/shade_lambert.cpp:37-42): with the texture frozen they feed nothing
  h->skip_tex_interior = (h->st.mode == DRH_TEXTURE_COLOR || h->st.mode == DRH_LAMBERT) && h->tex_lr == 0.0f;
  // step t (from 1) samples the jitter pattern of seed + t
  if (h->opt_kind == DRH_OPT_ADAM) h->jit_dev_step = h->opt_t.as<uint32_t>();
  else h->jit_host_step = (uint32_t)h->opt_iter + 1u;
  s = drh_d_render_and_compare_dev(h, cams, views, d_targets, edge_samples, h->grad.as<float>(), nullptr, d_loss);
  h->jit_dev_step = nullptr;
  h->jit_host_step = 0;
  h->skip_tex_interior = false;
  if (s) return s;
  // StepDecay (src/optimizer_simple.cpp:82-90): decayPeriod 100, decay_mult 0.75 for GD_Adam,
  // applied after iterations 100, 200, ...
  // scales every block's rate: positions, colours, texture and transform
  float plr = h->pos_lr, clr = h->col_lr, tlr = h->tex_lr, trlr = h->tr_lr;
  if (h->opt_kind == DRH_OPT_SIMPLE)
    for (int k = 1; k < h->opt_iter; k++)
      if (k % 100 == 0) { plr *= 0.75f; clr *= 0.75f; tlr *= 0.75f; trlr *= 0.75f; }
  for (int i = 0; i < h->nm; i++) {
    if (!h->diff[i]) continue;
    const int off = h->layout[4 * i];
    const int nv = h->vo[i + 1] - h->vo[i];
    const int colo = h->layout[4 * i + 1];
    const int n = h->layout[4 * i + 3] + h->tr_block(i);
    int seg_start[MAX_SEG] = {0};
    float seg_lr[MAX_SEG] = {0};
    float* seg_param[MAX_SEG] = {nullptr};
    int nseg = 0;
    // OptSimple::GetLR (src/optimizer_simple.cpp:61-80) blocks.  OptUpdateScene
    // (src/optimizer_simple.cpp:92-186) moves positions and colours, the texture when its rate
    // is set, and rows 0-2 of the transform when its rate is set (all of OPT_ALL)
    seg_start[nseg] = 0; seg_lr[nseg] = plr; seg_param[nseg] = h->pos_orig.as<float>() + 3ll * h->vo[i]; nseg++;
    if (colo >= 0) {
      seg_start[nseg] = colo; seg_lr[nseg] = clr; seg_param[nseg] = h->col.as<float>() + 3ll * h->vo[i]; nseg++;
    }
    const int texo = h->layout[4 * i + 2];
    if (texo >= 0) {
      seg_start[nseg] = texo;
      seg_lr[nseg] = tlr;
      seg_param[nseg] = h->tex_lr != 0.0f ? h->texels.as<float>() + h->tex_off[i] : nullptr;
      nseg++;
    }
    seg_start[nseg] = h->layout[4 * i + 3];
    seg_lr[nseg] = trlr;
    seg_param[nseg] = h->tr_lr == 0.0f ? nullptr
                      : h->rflag[i] ? h->drparams.as<float>() + 7ll * i   // TransformR parameters
                                    : h->tr_rows.as<float>() + 12ll * i;
    nseg++;
    (void)nv;
    if (h->opt_kind == DRH_OPT_SIMPLE) {
      OptApplyArgs a{};
      a.n = n; a.off = off; a.nseg = nseg; a.m = h->opt_m.as<float>(); a.G = h->opt_v.as<float>();
      for (int k = 0; k < nseg; k++) { a.seg_start[k] = seg_start[k]; a.seg_lr[k] = seg_lr[k]; a.seg_param[k] = seg_param[k]; }
      h->timed("k_optsimple", [&] { launch_optsimple(h->grad.as<float>(), h->opt_m.as<float>(), h->opt_v.as<float>(), off, a, h->stream); });
    } else {
      AdamArgs a{};
      a.n = n; a.off = off; a.nseg = nseg; a.grad = h->grad.as<float>(); a.m = h->opt_m.as<float>(); a.v = h->opt_v.as<float>();
      a.b1 = 0.9f; a.b2 = 0.999f; a.eps = 1e-8f;
      a.step = h->opt_t.as<int>();
      for (int k = 0; k < nseg; k++) { a.seg_start[k] = seg_start[k]; a.seg_lr[k] = seg_lr[k]; a.seg_param[k] = seg_param[k]; }
      h->timed("k_adam", [&] { launch_adam(a, h->stream); });
    }
  }
  if (h->tr_lr != 0.0f) {
    // the moved transforms into the matrices the next commit applies; the inverse of mesh 0's
    // transform used by the edge gradient stays the one of the first commit, as the
    // reference's Scene::transforms_inv is appended to and never cleared
    // (common_src/scene_utils.cpp:110-122)
    TransformArgs ta{h->nm, h->dlayout.as<int>(), h->tr_rows.as<float>(), h->dmats.as<float>(), h->drflag.as<int>(),
                     h->drparams.as<float>()};
    h->timed("k_transform_update", [&] { launch_transform_update(ta, h->stream); });
  }
  return h->hip(hipGetLastError(), "opt_step");
}

drh_status drh_opt_step(drh_handle h, const drh_camera* cams, int views, const float* d_targets, int edge_samples,
                        float* d_loss) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "opt_step: commit first");
  if (h->opt_kind < 0) return h->fail(DRH_E_INVALID, "opt_step: call drh_opt_init first");
  if (!cams || views <= 0 || !d_targets || edge_samples < 0) return h->fail(DRH_E_INVALID, "opt_step: bad arguments");
  (void)hipSetDevice(h->st.device);
  drh_status s;
  // Adam steps are replayed from a hipGraph once the first (eager) step has sized every
  // buffer; the graph is re-captured when the arguments change.  OptSimple stays eager (its
  // step-decayed learning rates are host values).
  std::vector<char> sig(sizeof(drh_camera) * views + 4 * sizeof(void*));
  std::memcpy(sig.data(), cams, sizeof(drh_camera) * views);
  const void* ptrs[4] = {d_targets, d_loss, (const void*)(intptr_t)edge_samples, (const void*)(intptr_t)views};
  std::memcpy(sig.data() + sizeof(drh_camera) * views, ptrs, sizeof ptrs);
  const bool graphable = h->use_graph && h->opt_kind == DRH_OPT_ADAM && h->opt_iter > 0 && !h->prof && !h->count_on;
  if (graphable) {
    // the graph reads the camera block and the sample table: make sure they hold this call's
    // (uploads only if another call changed them; a reallocation shows in the epoch)
    long long total;
    if ((s = h->set_cams(cams, views, &total)) || (s = h->ensure_rnd(edge_samples))) return s;
  }
  // this step's form: full LBVH build or refit only (opt_step_body makes the same choice)
  const int form = (h->opt_iter % h->rebuild_every == 0 || !h->topo_valid) ? 1 : 0;
  const unsigned long long epoch = h->realloc_epoch.load();
  auto valid = [&](int k) { return h->gexec[k] && sig == h->graph_sig[k] && h->graph_epoch[k] == epoch; };
  if (graphable && valid(form)) {
    if ((s = h->hip(hipGraphLaunch(h->gexec[form], h->stream), "graph launch"))) return s;
    h->opt_iter++;
    return DRH_OK;
  }
  if (graphable) {
    // this form is missing or stale; the other one is kept while it still holds
    if (valid(1 - form)) h->drop_form(form); else h->drop_graph();
    if ((s = h->hip(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "begin capture"))) return s;
    const drh_status body = opt_step_body(h, cams, views, d_targets, edge_samples, d_loss);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(h->stream, &g);
    if (body == DRH_OK && ec == hipSuccess && g && h->instantiate(form, g)) {
      h->graph[form] = g;
      h->graph_sig[form] = sig;
      h->graph_epoch[form] = h->realloc_epoch.load();
      if ((s = h->hip(hipGraphLaunch(h->gexec[form], h->stream), "graph launch"))) return s;
      h->opt_iter++;
      return DRH_OK;
    }
    // capture failed: forget it and run the step eagerly
    if (g) (void)hipGraphDestroy(g);
    h->gexec[form] = nullptr;
    (void)hipGetLastError();
    h->use_graph = false;
  }
  if ((s = opt_step_body(h, cams, views, d_targets, edge_samples, d_loss))) return s;
  h->opt_iter++;
  return DRH_OK;
}

drh_status drh_set_pixel_jitter(drh_handle h, int on, unsigned seed) {
  if (!h) return DRH_E_INVALID;
  if ((on != 0) != (h->jit_on != 0) || seed != h->jit_seed) h->drop_graph();   // baked into the captured step
  h->jit_on = on != 0;
  h->jit_seed = seed;
  return DRH_OK;
}

drh_status drh_set_traversal(drh_handle h, int mode) {
  if (!h) return DRH_E_INVALID;
  if (mode != DRH_TRAVERSAL_AUTO && mode != DRH_TRAVERSAL_WAVE && mode != DRH_TRAVERSAL_LANE)
    return h->fail(DRH_E_INVALID, "set_traversal: unknown mode %d", mode);
  if (mode != h->traversal) h->drop_graph();   // baked into the captured step's arguments
  h->traversal = mode;
  return DRH_OK;
}

drh_status drh_set_edge_side_test(drh_handle h, int test) {
  if (!h || (test != DRH_SIDE_TEST_REFERENCE && test != DRH_SIDE_TEST_WATERTIGHT)) return DRH_E_INVALID;
  if (test != h->side_test) h->drop_graph();   // the captured step bakes the choice in
  h->side_test = test;
  return DRH_OK;
}

drh_status drh_opt_steps(drh_handle h, const drh_camera* cams, int views, const float* d_targets, int edge_samples,
                         float* d_loss, int steps) {
  if (!h) return DRH_E_INVALID;
  if (steps < 1) return h->fail(DRH_E_INVALID, "opt_steps: steps must be at least 1");
  if (steps == 1) return drh_opt_step(h, cams, views, d_targets, edge_samples, d_loss);
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "opt_steps: commit first");
  if (h->opt_kind < 0) return h->fail(DRH_E_INVALID, "opt_steps: call drh_opt_init first");
  if (!cams || views <= 0 || !d_targets || edge_samples < 0) return h->fail(DRH_E_INVALID, "opt_steps: bad arguments");
  (void)hipSetDevice(h->st.device);
  drh_status s;
  // the first step of a fresh optimisation sizes every buffer eagerly; without graphs, step by step
  if (h->opt_iter == 0) {
    if ((s = drh_opt_step(h, cams, views, d_targets, edge_samples, d_loss))) return s;
    return drh_opt_steps(h, cams, views, d_targets, edge_samples, d_loss, steps - 1);
  }
  const bool graphable = h->use_graph && h->opt_kind == DRH_OPT_ADAM && !h->prof && !h->count_on;
  if (!graphable) {
    for (int k = 0; k < steps; k++)
      if ((s = drh_opt_step(h, cams, views, d_targets, edge_samples, d_loss))) return s;
    return DRH_OK;
  }
  std::vector<char> sig(sizeof(drh_camera) * views + 4 * sizeof(void*));
  std::memcpy(sig.data(), cams, sizeof(drh_camera) * views);
  const void* ptrs[4] = {d_targets, d_loss, (const void*)(intptr_t)edge_samples, (const void*)(intptr_t)views};
  std::memcpy(sig.data() + sizeof(drh_camera) * views, ptrs, sizeof ptrs);
  long long total;
  if ((s = h->set_cams(cams, views, &total)) || (s = h->ensure_rnd(edge_samples))) return s;
  const int it0 = h->opt_iter;
  unsigned forms = 0;
  for (int k = 0; k < steps && k < 32; k++)
    if ((it0 + k) % h->rebuild_every == 0 || (k == 0 && !h->topo_valid)) forms |= 1u << k;
  const unsigned long long epoch = h->realloc_epoch.load();
  for (auto& m : h->multi)
    if (m.k == steps && m.forms == forms && m.epoch == epoch && m.sig == sig) {
      if ((s = h->hip(hipGraphLaunch(m.e, h->stream), "graph launch"))) return s;
      h->opt_iter = it0 + steps;
      return DRH_OK;
    }
  if (h->multi.size() >= drh_renderer::MULTI_KEEP) h->drop_multi(1);
  if ((s = h->hip(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "begin capture"))) return s;
  drh_status body = DRH_OK;
  for (int k = 0; k < steps && body == DRH_OK; k++) {
    h->opt_iter = it0 + k;
    body = opt_step_body(h, cams, views, d_targets, edge_samples, d_loss);
  }
  h->opt_iter = it0;
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  const hipError_t ec = hipStreamEndCapture(h->stream, &g);
  if (body == DRH_OK && ec == hipSuccess && g && hipGraphInstantiate(&e, g, nullptr, nullptr, 0) == hipSuccess) {
    h->multi.push_back(drh_renderer::MultiGraph{steps, forms, sig, h->realloc_epoch.load(), g, e});
    if ((s = h->hip(hipGraphLaunch(e, h->stream), "graph launch"))) return s;
    h->opt_iter = it0 + steps;
    return DRH_OK;
  }
  // capture failed: the host state of the captured steps never ran; forget it and step eagerly
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  h->topo_valid = false;
  for (int k = 0; k < steps; k++)
    if ((s = drh_opt_step(h, cams, views, d_targets, edge_samples, d_loss))) return s;
  return DRH_OK;
}

drh_status drh_opt_use_graph(drh_handle h, int on) {
  if (!h) return DRH_E_INVALID;
  h->use_graph = on != 0;
  if (!on) h->drop_graph();
  return DRH_OK;
}

drh_status drh_export_params(drh_handle h, float* pos, float* col, const void* d_loss, float* loss, uint32_t* fault) {
  if (!h || !pos || !col || !loss || !d_loss || !fault) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "export_params: commit first");
  (void)hipSetDevice(h->st.device);
  const size_t b = 12 * (size_t)h->V;
  drh_status s;
  if ((s = h->hip(hipMemcpyAsync(pos, h->pos_orig.p, b, hipMemcpyDeviceToHost, h->stream), "export")) ||
      (s = h->hip(hipMemcpyAsync(col, h->col.p, b, hipMemcpyDeviceToHost, h->stream), "export")) ||
      (s = h->hip(hipMemcpyAsync(loss, d_loss, 4, hipMemcpyDeviceToHost, h->stream), "export")) ||
      (s = h->hip(hipMemcpyAsync(fault, h->fault.p, 4, hipMemcpyDeviceToHost, h->stream), "export")))
    return s;
  return DRH_OK;
}

drh_status drh_job_upload(drh_handle h, const void* targets, size_t bytes, void** d_targets, void** d_loss) {
  if (!h || (bytes && !targets) || !d_targets || !d_loss) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  drh_status s;
  if ((s = h->hip(h->job_tgt.reserve(std::max<size_t>(bytes, 4)), "alloc")) || (s = h->hip(h->job_loss.reserve(16), "alloc")) ||
      (s = h->stage_begin(drh_renderer::stage_round(bytes))) || (s = h->stage_send(h->job_tgt.p, targets, bytes)) ||
      (s = h->stage_end()))
    return s;
  *d_targets = h->job_tgt.p;
  *d_loss = h->job_loss.p;
  return DRH_OK;
}

drh_status drh_get_params(drh_handle h, float* pos, float* col) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "get_params: commit first");
  (void)hipSetDevice(h->st.device);
  drh_status s;
  if (pos && (s = h->hip(hipMemcpyAsync(pos, h->pos_orig.p, 12 * (size_t)h->V, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  if (col && (s = h->hip(hipMemcpyAsync(col, h->col.p, 12 * (size_t)h->V, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  return h->hip(hipStreamSynchronize(h->stream), "get_params");
}

drh_status drh_get_transforms(drh_handle h, float* transforms) {
  if (!h || !transforms) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "get_transforms: commit first");
  (void)hipSetDevice(h->st.device);
  std::vector<float> m(48 * (size_t)h->nm);
  drh_status s;
  if ((s = h->hip(hipMemcpyAsync(m.data(), h->dmats.p, 4 * m.size(), hipMemcpyDeviceToHost, h->stream), "download")) ||
      (s = h->hip(hipStreamSynchronize(h->stream), "get_transforms")))
    return s;
  for (int i = 0; i < h->nm; i++) std::memcpy(transforms + 16 * (size_t)i, &m[48 * (size_t)i], 64);
  return DRH_OK;
}

drh_status drh_prepare_steps(drh_handle h, const drh_camera* cams, int views, int edge_samples) {
  if (!h || !cams || views <= 0 || edge_samples < 0) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  long long total;
  drh_status s;
  if ((s = h->set_cams(cams, views, &total)) || (s = h->ensure_rnd(edge_samples))) return s;
  return DRH_OK;
}

drh_status drh_set_interior_pos_grad(drh_handle h, int on) {
  if (!h || (on != 0 && on != 1)) return DRH_E_INVALID;
  if (on && h->st.mode != DRH_VERTEX_COLOR && h->st.mode != DRH_LAMBERT && h->st.mode != DRH_LAMBERT_VERTEX_COLOR)
    return h->fail(DRH_E_UNSUPPORTED, "set_interior_pos_grad: VERTEX_COLOR, LAMBERT and LAMBERT_VERTEX_COLOR only");
  if ((on != 0) != h->interior_pos) h->drop_graph();
  h->interior_pos = on != 0;
  return DRH_OK;
}

drh_status drh_set_deterministic(drh_handle h, int on) {
  if (!h || (on != 0 && on != 1)) return DRH_E_INVALID;
  if ((on != 0) != h->deterministic) h->drop_graph();
  h->deterministic = on != 0;
  return DRH_OK;
}

drh_status drh_get_restricted(drh_handle h, float* params) {
  if (!h || !params) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "get_restricted: commit first");
  (void)hipSetDevice(h->st.device);
  std::vector<float> p(7 * (size_t)h->nm);
  drh_status s;
  if ((s = h->hip(hipMemcpyAsync(p.data(), h->drparams.p, 4 * p.size(), hipMemcpyDeviceToHost, h->stream), "download")) ||
      (s = h->hip(hipStreamSynchronize(h->stream), "get_restricted")))
    return s;
  for (int i = 0; i < h->nm; i++)
    for (int k = 0; k < 7; k++) params[7 * (size_t)i + k] = h->rflag[i] ? p[7 * (size_t)i + k] : 0.0f;
  return DRH_OK;
}

drh_status drh_opt_set_transform_lr(drh_handle h, float lr) {
  if (!h) return DRH_E_INVALID;
  if (!(lr >= 0.0f) || lr > 1e30f) return h->fail(DRH_E_INVALID, "opt_set_transform_lr: bad rate");
  if (lr != h->tr_lr) h->drop_graph();   // the captured step bakes the segment table in
  h->tr_lr = lr;
  return DRH_OK;
}

drh_status drh_opt_set_texture_lr(drh_handle h, float tex_lr) {
  if (!h) return DRH_E_INVALID;
  if (!(tex_lr >= 0.0f) || tex_lr > 1e30f) return h->fail(DRH_E_INVALID, "opt_set_texture_lr: bad rate");
  if (tex_lr != h->tex_lr) h->drop_graph();   // the captured step bakes the segment table in
  h->tex_lr = tex_lr;
  return DRH_OK;
}

drh_status drh_set_interior_samples(drh_handle h, int aligned, int store) {
  if (!h || (aligned != 0 && aligned != 1) || store < DRH_HITS_RECOMPUTE || store > DRH_HITS_AUTO) return DRH_E_INVALID;
  if (store && !aligned) return h->fail(DRH_E_INVALID, "set_interior_samples: storing hits needs aligned samples");
  if (aligned != h->interior_aligned || store != h->interior_store) h->drop_graph();
  h->interior_aligned = aligned;
  h->interior_store = store;
  return DRH_OK;
}

drh_status drh_set_hit_buffer_limit(drh_handle h, size_t bytes) {
  if (!h) return DRH_E_INVALID;
  if (bytes != h->hit_limit) h->drop_graph();   // the captured step bakes the choice in
  h->hit_limit = bytes;
  return DRH_OK;
}

drh_status drh_count_rays(drh_handle h, int on) {
  if (!h || (on != 0 && on != 1)) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  h->drop_graph();   // counting runs other kernel instantiations, eagerly
  h->count_on = on != 0;
  if (!on) return DRH_OK;
  drh_status s;
  if ((s = h->hip(h->rays_buf.reserve(64), "alloc")) || (s = h->hip(hipMemsetAsync(h->rays_buf.p, 0, 64, h->stream), "zero")))
    return s;
  return h->hip(hipStreamSynchronize(h->stream), "count_rays");
}

drh_status drh_read_ray_counts(drh_handle h, unsigned long long counts[6]) {
  if (!h || !counts) return DRH_E_INVALID;
  if (!h->rays_buf.p) { std::memset(counts, 0, 6 * sizeof(unsigned long long)); return DRH_OK; }
  (void)hipSetDevice(h->st.device);
  drh_status s;
  if ((s = h->hip(hipMemcpyAsync(counts, h->rays_buf.p, 48, hipMemcpyDeviceToHost, h->stream), "download")))
    return s;
  return h->check_fault();
}

drh_status drh_get_hit_buffer_bytes(drh_handle h, size_t* bytes) {
  if (!h || !bytes) return DRH_E_INVALID;
  *bytes = h->hitbuf.bytes;
  return DRH_OK;
}

drh_status drh_set_edge_grad_mode(drh_handle h, int mode) {
  if (!h || (mode != DRH_EDGE_GRAD_REFERENCE_DIAGONAL && mode != DRH_EDGE_GRAD_FULL_JACOBIAN)) return DRH_E_INVALID;
  if (mode != h->edge_grad_mode) h->drop_graph();
  h->edge_grad_mode = mode;
  return DRH_OK;
}

drh_status drh_set_bvh_rebuild(drh_handle h, int every) {
  if (!h || every < 1) return DRH_E_INVALID;
  h->rebuild_every = every;
  return DRH_OK;
}

drh_status drh_set_edge_order(drh_handle h, int order) {
  if (!h || order < DRH_EDGE_ORDER_AUTO || order > DRH_EDGE_ORDER_BINNED) return DRH_E_INVALID;
  if (order != h->edge_order) h->drop_graph();
  h->edge_order = order;
  return DRH_OK;
}

drh_status drh_set_device_sharing(drh_handle h, int handles) {
  if (!h || handles < 1) return DRH_E_INVALID;
  if ((handles > 1) != (h->sharing > 1)) h->drop_graph();   // the captured step bakes the fork in
  h->sharing = handles;
  return DRH_OK;
}

drh_status drh_get_texels(drh_handle h, float* out) {
  if (!h || !out) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "get_texels: commit first");
  (void)hipSetDevice(h->st.device);
  size_t n = 0;
  for (int i = 0; i < h->nm; i++) n += 3 * (size_t)h->tex_w[i] * h->tex_h[i];
  drh_status s;
  if (n && (s = h->hip(hipMemcpyAsync(out, h->texels.p, 4 * n, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  return h->hip(hipStreamSynchronize(h->stream), "get_texels");
}

void* drh_stream(drh_handle h) { return h ? (void*)h->stream : nullptr; }

drh_status drh_sync(drh_handle h) {
  if (!h) return DRH_E_INVALID;
  return h->check_fault();
}

drh_status drh_profile_enable(drh_handle h, int on) {
  if (!h) return DRH_E_INVALID;
  h->drain_prof();
  h->prof_acc.clear();
  h->prof = on != 0;
  return DRH_OK;
}

drh_status drh_profile_read(drh_handle h, const char* name, double* total_ms, int* launches) {
  if (!h || !name) return DRH_E_INVALID;
  h->drain_prof();
  auto it = h->prof_acc.find(name);
  if (total_ms) *total_ms = it == h->prof_acc.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == h->prof_acc.end() ? 0 : it->second.second;
  return DRH_OK;
}

drh_status drh_debug_hits(drh_handle h, const drh_camera* cam, uint32_t* prim, uint32_t* geom, float* t, float* u,
                          float* v) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "debug_hits: commit first");
  if (!prim || !geom || !t || !u || !v) return h->fail(DRH_E_INVALID, "debug_hits: NULL output");
  (void)hipSetDevice(h->st.device);
  long long total;
  drh_status s;
  if ((s = h->set_cams(cam, 1, &total))) return s;
  const int n = h->cam_cache[0].w * h->cam_cache[0].h;
  DBuf b;
  if ((s = h->hip(b.reserve(20 * (size_t)n), "alloc"))) return s;
  if ((s = h->build_view_suspects(1))) return s;
  HitsArgs a{h->tracer_view(0), h->cams.as<DevCam>(), b.as<uint32_t>(), b.as<uint32_t>() + n, (float*)(b.as<uint32_t>() + 2 * n),
             (float*)(b.as<uint32_t>() + 3 * n), (float*)(b.as<uint32_t>() + 4 * n)};
  const int tiles = ((h->cam_cache[0].w + 7) / 8) * ((h->cam_cache[0].h + 7) / 8);
  launch_debug_hits(a, tiles, h->stream);
  if ((s = h->hip(hipGetLastError(), "debug_hits"))) return s;
  std::vector<uint32_t> hb(5 * (size_t)n);
  if ((s = h->hip(hipMemcpyAsync(hb.data(), b.p, 20 * (size_t)n, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  if ((s = h->check_fault())) return s;
  std::memcpy(prim, hb.data(), 4 * (size_t)n);
  std::memcpy(geom, hb.data() + n, 4 * (size_t)n);
  std::memcpy(t, hb.data() + 2 * n, 4 * (size_t)n);
  std::memcpy(u, hb.data() + 3 * n, 4 * (size_t)n);
  std::memcpy(v, hb.data() + 4 * n, 4 * (size_t)n);
  return DRH_OK;
}

drh_status drh_debug_bvh(drh_handle h, uint32_t* morton_sorted, uint32_t* tri_order, float* scene_bounds) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "debug_bvh: commit first");
  (void)hipSetDevice(h->st.device);
  drh_status s;
  if (h->F == 0) return DRH_OK;
  if (morton_sorted && (s = h->hip(hipMemcpyAsync(morton_sorted, h->keys0.p, 4 * (size_t)h->F, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  if (tri_order && (s = h->hip(hipMemcpyAsync(tri_order, h->vals0.p, 4 * (size_t)h->F, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  uint32_t b[6];
  if ((s = h->hip(hipMemcpyAsync(b, h->bounds.p, 24, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  if ((s = h->hip(hipStreamSynchronize(h->stream), "debug_bvh"))) return s;
  if (scene_bounds)
    for (int i = 0; i < 6; i++) {
      const uint32_t k = b[i];
      const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
      std::memcpy(&scene_bounds[i], &u, 4);
    }
  return DRH_OK;
}

drh_status drh_debug_edges(drh_handle h, int* n_edges, int* v0, int* v1, float* pmf, float* cdf) {
  if (!h) return DRH_E_INVALID;
  if (!h->committed) return h->fail(DRH_E_NOT_COMMITTED, "debug_edges: commit first");
  (void)hipSetDevice(h->st.device);
  if (n_edges) *n_edges = h->E;
  if (!v0 || h->E == 0) return DRH_OK;
  std::vector<int2> e(h->E);
  drh_status s;
  if ((s = h->hip(hipMemcpyAsync(e.data(), h->edges.p, 8 * (size_t)h->E, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  if (pmf && (s = h->hip(hipMemcpyAsync(pmf, h->pmf.p, 4 * (size_t)h->E, hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  if (cdf && (s = h->hip(hipMemcpyAsync(cdf, h->cdf.p, 4 * (size_t)(h->E + 1), hipMemcpyDeviceToHost, h->stream), "download"))) return s;
  if ((s = h->hip(hipStreamSynchronize(h->stream), "debug_edges"))) return s;
  for (int i = 0; i < h->E; i++) { v0[i] = e[i].x; if (v1) v1[i] = e[i].y; }
  return DRH_OK;
}

}  // extern "C"

extern "C" {

drh_status drh_malloc(drh_handle h, size_t bytes, void** d_ptr) {
  if (!h || !d_ptr) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  *d_ptr = nullptr;
  drh_status s = h->hip(hipMallocAsync(d_ptr, std::max<size_t>(bytes, 4), h->stream), "drh_malloc");
  if (s) return s;
  return h->hip(hipStreamSynchronize(h->stream), "drh_malloc");
}

drh_status drh_free(drh_handle h, void* d_ptr) {
  if (!h) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  drh_status s = h->hip(hipStreamSynchronize(h->stream), "drh_free");
  if (s) return s;
  return d_ptr ? h->hip(hipFreeAsync(d_ptr, h->stream), "drh_free") : DRH_OK;
}

drh_status drh_memcpy_htod(drh_handle h, void* d_dst, const void* src, size_t bytes) {
  if (!h || (bytes && (!d_dst || !src))) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  drh_status s = h->hip(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, h->stream), "drh_memcpy_htod");
  if (s) return s;
  return h->hip(hipStreamSynchronize(h->stream), "drh_memcpy_htod");
}

drh_status drh_memcpy_dtoh(drh_handle h, void* dst, const void* d_src, size_t bytes) {
  if (!h || (bytes && (!dst || !d_src))) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  drh_status s = h->hip(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, h->stream), "drh_memcpy_dtoh");
  if (s) return s;
  return h->hip(hipStreamSynchronize(h->stream), "drh_memcpy_dtoh");
}

}  // extern "C"

extern "C" drh_status drh_debug_copy(drh_handle h, float* d_dst, const float* d_src, long long n) {
  if (!h || n < 0 || (n && (!d_dst || !d_src))) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  h->timed("k_copy_f32", [&] { launch_copy_f32(d_dst, d_src, n, h->stream); });
  return h->hip(hipStreamSynchronize(h->stream), "debug_copy");
}

extern "C" drh_status drh_debug_cdf(drh_handle h, const float* pmf_raw, int E, float* pmf_out, float* cdf_out,
                                    int parallel) {
  if (!h || E <= 0 || !pmf_raw || !pmf_out || !cdf_out) return DRH_E_INVALID;
  (void)hipSetDevice(h->st.device);
  DBuf a, p, c, w;
  drh_status s;
  if ((s = h->hip(a.reserve(4 * (size_t)E), "alloc")) || (s = h->hip(p.reserve(4 * (size_t)E), "alloc")) ||
      (s = h->hip(c.reserve(4 * (size_t)(E + 1)), "alloc")) || (s = h->hip(w.reserve(cdf_work_bytes(E)), "alloc")))
    return s;
  if ((s = h->hip(hipMemcpyAsync(a.p, pmf_raw, 4 * (size_t)E, hipMemcpyHostToDevice, h->stream), "upload"))) return s;
  h->timed(parallel ? "k_cdf_par" : "k_cdf", [&] {
    launch_cdf(a.as<float>(), E, p.as<float>(), c.as<float>(), w.p, h->stream, parallel != 0);
  });
  if ((s = h->hip(hipGetLastError(), "debug_cdf"))) return s;
  if ((s = h->hip(hipMemcpyAsync(pmf_out, p.p, 4 * (size_t)E, hipMemcpyDeviceToHost, h->stream), "download")) ||
      (s = h->hip(hipMemcpyAsync(cdf_out, c.p, 4 * (size_t)(E + 1), hipMemcpyDeviceToHost, h->stream), "download")))
    return s;
  return h->hip(hipStreamSynchronize(h->stream), "debug_cdf");
}

// drh_debug_cdf's parallel exact CDF with device inputs / outputs and the work buffer in the
// device's static scratch array: no allocation, no synchronisation (capturable).
extern "C" drh_status drh_build_edge_sampler_noalloc(drh_handle h, const float* d_pmf_raw, int E, float* d_pmf,
                                                     float* d_cdf) {
  if (!h) return DRH_E_INVALID;
  if (E <= 0 || !d_pmf_raw || !d_pmf || !d_cdf) return h->fail(DRH_E_INVALID, "build_edge_sampler_noalloc: bad arguments");
  if (E > NOALLOC_MAX_EDGES)
    return h->fail(DRH_E_INVALID, "build_edge_sampler_noalloc: %d edges exceed the maximum of %d (scratch of %zu bytes)",
                   E, NOALLOC_MAX_EDGES, NOALLOC_SCRATCH_BYTES);
  (void)hipSetDevice(h->st.device);
  void* scr = noalloc_scratch(h->st.device);
  if (!scr) return h->fail(DRH_E_HIP, "build_edge_sampler_noalloc: scratch address not set up on device %d", h->st.device);
  h->timed("k_cdf_par", [&] { launch_cdf(d_pmf_raw, E, d_pmf, d_cdf, scr, h->stream, true); });
  return h->hip(hipGetLastError(), "build_edge_sampler_noalloc");
}
