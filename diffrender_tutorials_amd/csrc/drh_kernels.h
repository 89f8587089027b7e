// drh_kernels.h -- kernel argument blocks and launchers shared by the host runtime
// (drh_api.cpp) and the kernels (drh_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "drh_device.h"

namespace drh {

constexpr int RADIX_TILE = 2048;           // keys per radix-sort block (8 rounds of 256)
constexpr int REFIT_CHUNK = 1024;          // sorted leaves refit per workgroup in LDS (one per thread)
constexpr int CDF_CHUNK = 8192;            // pmf floats staged in LDS per CDF round (sequential CDF)
constexpr int CDF_SERIAL_MAX = 2048;       // edge sets up to this size take the sequential CDF (one launch)
constexpr int CDF_CH = 128;                // edges per chunk of the parallel exact CDF
constexpr int CDF_CPB = 64;                // CDF chunks per workgroup (one per lane)
constexpr int CDF_WALK_STAGE = 8192;       // floats of LDS for the chunks replayed serially
constexpr int CDF_WALK_CHUNKS = 8192;      // chunk records the serial walk reads from LDS
constexpr int INTERIOR_LDS_FLOATS = 4096;  // largest gradient target accumulated in LDS (interior)
constexpr int EDGE_LDS_FLOATS = 4096;      // largest position gradient accumulated in LDS (edges)

// pixel tiles per view, one wavefront each (matches tile_map in the kernels)
__host__ __device__ inline int tile_side(int sq) { return (sq == 1 || sq == 2 || sq == 4 || sq == 8) ? 8 / sq : 8; }
// Largest scene: the LBVH walks address nodes (64 B) and leaf triangles (16 B per array) by
// 32-bit byte offsets.
constexpr long long MAX_TRIANGLES = 1ll << 26;
__host__ __device__ inline int num_tiles(int sq, int w, int h) {
  const int ts = tile_side(sq);
  return ((w + ts - 1) / ts) * ((h + ts - 1) / ts);
}
// regions of rt x rt tiles (one wavefront each in the interior pass)
__host__ __device__ inline int num_regions(int sq, int w, int h, int rt) {
  const int ts = tile_side(sq) * rt;
  return ((w + ts - 1) / ts) * ((h + ts - 1) / ts);
}
// Interior gradient targets (float3 each: a vertex colour or a texel) are summed in a
// per-wavefront LDS hash table of this many entries before the global atomics; a region of
// INTERIOR_REGION x INTERIOR_REGION interleaved tiles (256 samples) shares one table.
constexpr int INTERIOR_HASH = 256;
constexpr int INTERIOR_HASH_PROBES = 16;
constexpr int INTERIOR_REGION = 2;

struct PrepareArgs {
  int V, n_meshes, has_nrm;
  const int* mesh_vo;
  const float* mats;      // per mesh: T, transpose(inverse(T)), identity (48 floats)
  const float* pos_orig;
  const float* nrm_in;
  float* pos;
  float* nrm;
  float* nrm2d;
  uint32_t* bounds;       // optional: set to the empty key range (6 values) by thread 0
  uint32_t* zero;         // optional: one counter cleared by thread 0
  int* step;              // optional: the Adam step counter, advanced by thread 0 (a fused step's first kernel)
};

struct SuspectArgs {
  int F;
  int light;              // 0: view origin (ox, oy, oz); 1: light direction (ox, oy, oz)
  float ox, oy, oz;
  const float4* bf_a;
  const float4* bf_e1;
  const float4* bf_e2;
  float4* list;           // capacity F entries (view: one float4 each; light: two)
  int* count;             // zeroed by the caller
};

struct TriSetupArgs {
  int F;
  const int* tri_mesh;
  const float* mats;
  const uint32_t* idx;
  const float* pos;
  float4* bf_a;
  float4* bf_e1;
  float4* bf_e2;
  float4* box_lo;
  float4* box_hi;
  float4* cen;
  uint32_t* bounds;       // 6 ordered float keys: min xyz, max xyz
  uint32_t* counter;      // optional: F + 1 refit arrival counters cleared here
  SuspectArgs light;      // the light's suspect list built here too when light.list is set
};

struct MortonArgs {
  int F;
  const float4* cen;
  const uint32_t* bounds;
  uint32_t* keys;
  uint32_t* vals;
};

struct KarrasArgs {
  int F;
  const uint32_t* keys;   // sorted
  int2* range;            // per internal node: first and last sorted leaf it covers
  int* child;             // 2 per internal node
  int* node_parent;
  int* leaf_parent;
};

struct RefitArgs {
  int F;
  const uint32_t* order;  // sorted slot -> original triangle
  const float4* box_lo;
  const float4* box_hi;
  const int* child;
  const int* node_parent;
  const int* leaf_parent;
  const int2* range;
  uint32_t* counter;      // F + 1, zeroed; counter[F] counts the frontier
  int* frontier;          // F: nodes (or ~leaves) whose parent spans refit chunks
  float4* leaf_lo;
  float4* leaf_hi;
  float4* node_lo;
  float4* node_hi;
  // the spanning nodes listed at the last full build (k_span_mark / k_span_link); k_refit_top
  // climbs them in LDS when there are at most SPAN_MAX, and through `frontier` otherwise
  const int4* span;        // (node, parent's span index or -1, child 0 / 1: span index or -1)
  const int2* span_cid;    // the two children's ids (~leaf for a leaf)
  const uint32_t* span_n;  // spanning nodes counted (may exceed SPAN_MAX)
};
constexpr int SPAN_MAX = 1024;

// Spanning-node list (one per full build): node i spans when its leaf range crosses a
// REFIT_CHUNK boundary.
struct SpanArgs {
  int F;
  const int2* range;
  const int* child;
  const int* node_parent;
  int* span_of;            // F - 1: a spanning node's index in the list
  int* span_p;             // SPAN_MAX: the list's nodes
  int4* span;              // SPAN_MAX
  int2* span_cid;          // SPAN_MAX
  uint32_t* span_n;        // zeroed before k_span_mark
};

struct PackArgs {
  int F;
  const uint32_t* bounds;  // scene centroid bounds (ordered keys), for the absolute padding
  const int* child;
  const float4* leaf_lo;
  const float4* leaf_hi;
  const float4* node_lo;
  const float4* node_hi;
  float4* nodes;
  uint32_t* zero;          // optional: nzero (<= 256) counters cleared for the next list build
  int nzero;
};

struct ReorderArgs {
  int F;
  const uint32_t* order;
  const float4* bf_a;
  const float4* bf_e1;
  const float4* bf_e2;
  float4* tri_a;
  float4* tri_e1;
  float4* tri_e2;
};


// Binning of suspect lists into per-cell records in a pool (see Suspects in drh_device.h), in
// three passes over the suspects -- count the kept (suspect, cell) pairs per cell, scan the
// counts into pool offsets, fill -- one wavefront per suspect walking its candidate cells.
struct BinPool {
  float4* pool;           // two float4 per record
  int cap;                // records
  int2* cell;             // (offset, n) per cell (n = -1: overflowed, the rays use the list)
  int* cnt;               // per cell: kept pairs, then (after the scan) the fill cursor; zeroed
  int cells_x, cells_y;
};
struct SusBinArgs {        // the views' lists by screen cell (blockIdx.y = view)
  const DevCam* cams;
  const float4* list;     // view v's suspects (N.xyz, id) at list + v * F
  const int* count;       // per view
  int F;
  const float4* bf_a;
  const float4* bf_e1;
  const float4* bf_e2;
  BinPool bp;             // view v's at pool + 2 v pool_stride, cell / cnt + v cell_stride (its own cells_x / y)
  long long pool_stride, cell_stride;
};
struct LightBinArgs {      // the light's list over a grid across the light
  const float4* list;     // the light's suspects (A, id), (pvec, invDet)
  const int* count;
  const float4* bf_e1;
  const float4* nodes;    // LBVH (the root's child boxes give the scene's extent)
  float Lx, Ly, Lz;       // the light direction L (shadow rays go along -L)
  float4* grid;           // out: the grid's placement (3 float4)
  BinPool bp;
};

struct RenderArgs {
  Tracer tr;
  Light light;
  const DevCam* cams;
  int view;
  int sq;                 // (int)sqrt(spp)
  float* img;
  const float* target;    // NULL: no loss
  float* adj;
  double* loss_part;
  int part_off;
  float4* hits;           // non-NULL: store each sample's closest hit (id bits, u, v, t) for the
                          // interior pass, at ((pixel) * sq*sq + sample) of the views' pixels;
                          // LAMBERT_VCOL stores lit_encode(shadow, ndl) in place of t
  Jitter jit;             // sample positions in the pixel (off: the grid's centres)
};

struct InteriorArgs {
  Tracer tr;
  const DevCam* cams;
  int view;
  int nsamp, sq;          // nsamp = sq*sq
  const float* hamm;      // 2*nsamp Hammersley offsets
  const float* adj;
  float* grad;
  const int* layout;      // per mesh: grad offset (-1: not differentiable), colour, texture, transform offsets
  int lds_base, lds_n;    // gradient range accumulated in LDS (lds_n = 0: global atomics)
  int uniform, uniform_off;  // one 1x1 texture: wave-reduce into grad[uniform_off..+3)
  int coverage;           // uniform and tc == 0 everywhere: only whether a sample hits matters
  int aligned;            // 1: samples at the render's grid offsets instead of Hammersley points
  const float4* hits;     // non-NULL (aligned only): the render's stored hits instead of tracing
  int hash;               // entries of the per-wave LDS hash table (0: atomics straight to grad)
  int region;             // tiles per side of a wave's region (1: one tile)
  unsigned long long* dacc;  // deterministic mode: 32.32 fixed-point sums per gradient float (else NULL)
  int pos_grad;           // VERTEX_COLOR: add the interior position term (k_interior<2, true>)
  const float* tr_inv0;   // mesh 0's inverse transform (for pos_grad)
  Light light;            // LAMBERT_VCOL: the light of the colour adjoint's factor
  Jitter jit;             // aligned: the render's sample positions
};

struct ProjectArgs {
  int V;
  const DevCam* cams;
  int view;
  const float* pos;
  float4* proj;
  // edge lengths (pmf before normalisation) in the same launch
  const int2* edges;
  int E;
  float* pmf_raw;
};

struct EdgeArgs {
  Tracer tr;
  Light light;
  const DevCam* cams;
  int view;
  long long N;
  int E;
  const int2* edges;
  const int* edge_mesh;
  const float* pmf;
  const float* cdf;
  const float4* proj;
  const float2* rnd;      // N (rnd0, rnd1) pairs ordered by rnd0
  const int2* order;      // non-NULL: the kept samples as (slot, edge) in pixel-bin order ...
  const int* nkept;       // ... and their count (device); NULL: every slot in rnd0 order
  float4* ecol;           // per item: inside - outside colour (w: edge id, -1 rejected)
  const float* adj;
  float* grad;
  const int* layout;
  const float* tr_inv0;   // 16 floats
  int lds_base, lds_n;
  int single_mesh, single_tr_off;
  int single_id;          // the single differentiable mesh
  float* trscr;           // non-NULL: transform gradients go here, 12 per mesh (restricted meshes present)
  unsigned long long* dacc;  // deterministic mode: 32.32 fixed-point sums, G for grad then 12 per mesh for trscr
  int dG;                 // G: where trscr's entries start in dacc
  int full_jacobian;      // 1: edge_grad with the full screen Jacobian (cross terms kept)
  int watertight;         // 1: side rays traced with the watertight ray/triangle test
};

// Edge samples binned by the EDGE_BIN x EDGE_BIN pixel tile of their sample point.  Only the
// samples the reference keeps (on screen, adjoint not ~0) are binned, so the sampling pass
// traces rays that share a small screen area whatever the number of samples per edge.
constexpr int EDGE_BIN = 8;
constexpr int EDGE_BIN_AUTO_PER_EDGE = 1;   // automatic binning below this many samples per edge
struct EdgeBinArgs {
  const DevCam* cams;
  int view;
  long long N;
  int E;
  const int2* edges;
  const float* cdf;
  const float4* proj;
  const float2* rnd;
  const float* adj;
  int bins_x, nbins;
  int* bin_count;         // nbins + 1, zero on entry; then bin starts and [nbins] = kept samples
  int4* rec;              // per slot: edge, bin (negative: rejected), rank within the bin
  int2* order;            // kept samples (slot, edge) in bin order
};

constexpr int MAX_SEG = 8;
struct OptApplyArgs {
  int n, off, nseg;
  const float* m;
  const float* G;
  int seg_start[MAX_SEG];
  float seg_lr[MAX_SEG];
  float* seg_param[MAX_SEG];
};

struct AdamArgs {
  int n, off, nseg;
  const float* grad;
  float* m;
  float* v;
  float b1, b2, eps;
  const int* step;        // 1-based Adam step, kept on the device so a captured graph replays correctly
  int seg_start[MAX_SEG];
  float seg_lr[MAX_SEG];
  float* seg_param[MAX_SEG];
};

// The optimiser's transform update: rows 0-2 of each differentiable mesh's transform
// (row-major 3x4, the order of its 12 gradient entries) back into T, then the normal matrix
// transpose(inverse(T)) (Scene::prepare_for_render, common_src/scene_utils.cpp:140-143).
struct TransformArgs {
  int n_meshes;
  const int* layout;   // 4 per mesh; layout[4i] < 0: mesh i is not differentiable
  const float* rows;   // 12 per mesh
  float* mats;         // 48 per mesh: T, transpose(inverse(T)), identity (column-major)
  const int* rflag;    // per mesh 1: T = TransformR::to_mat of rparams (NULL: none)
  const float* rparams;  // 7 per mesh
};

// Transform gradients of restricted meshes: the 12 matrix entries gathered by the edge pass
// (scr) through the Jacobian of TransformR's matrix into the mesh's 7-entry block
// (src/drender.h:398-422); other differentiable meshes get their 12 entries added as they are.
struct RestrictedArgs {
  int n_meshes;
  const int* layout;
  const int* rflag;
  const float* rparams;
  const float* scr;    // 12 per mesh
  float* grad;
};

struct HitsArgs {
  Tracer tr;
  const DevCam* cams;
  uint32_t* prim;
  uint32_t* geom;
  float* t;
  float* u;
  float* v;
};

void launch_fill_f32(float* p, long long n, float v, hipStream_t s);
void launch_fill_u32(uint32_t* p, long long n, uint32_t v, hipStream_t s);
void launch_copy_f32(float* d, const float* src, long long n, hipStream_t s);
// up to SCATTER_MAX arrays copied out of one staged block in one launch
constexpr int SCATTER_MAX = 24;
struct ScatterArgs {
  int n;
  const uint32_t* src;
  uint32_t* dst[SCATTER_MAX];
  long long off[SCATTER_MAX], words[SCATTER_MAX];
};
void launch_scatter(const ScatterArgs& a, hipStream_t s);
void launch_prepare(const PrepareArgs& a, hipStream_t s);
void launch_tri_setup(const TriSetupArgs& a, hipStream_t s);
void launch_morton(const MortonArgs& a, hipStream_t s);
void launch_radix_pass(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout, int n, int shift,
                       uint32_t* hist, hipStream_t s);
void launch_karras(const KarrasArgs& a, hipStream_t s);
void launch_refit(const RefitArgs& a, hipStream_t s);
void launch_span(const SpanArgs& a, hipStream_t s);   // F > REFIT_CHUNK, after k_karras
void launch_pack(const PackArgs& a, hipStream_t s);
void launch_pack_reorder(const PackArgs& pk, const ReorderArgs& ro, hipStream_t s);   // F > 1; one launch
void launch_suspects(const SuspectArgs& a, hipStream_t s);
void launch_sus_bin(const SusBinArgs& a, int views, hipStream_t s);   // count, scan and fill
void launch_light_bin(const LightBinArgs& a, hipStream_t s);
void launch_render(int mode, const RenderArgs& a, int tiles, hipStream_t s);
void launch_loss_final(const double* part, int n, float* out, double denom, hipStream_t s);
void launch_interior(int mode, const InteriorArgs& a, int blocks, hipStream_t s);
// The interior pass of a single 1x1 albedo with all texture coordinates 0 (a.uniform and
// a.coverage): one wave per pixel tile, partial sums into part (16 B per tile), then a fixed-
// order sum added to grad[uniform_off .. +3).
// LAMBERT (mode 4) / LAMBERT_VCOL interior position term (drh_set_interior_pos_grad)
void launch_interior_pos(int mode, const InteriorArgs& a, int tiles, hipStream_t s);
void launch_coverage(const InteriorArgs& a, int tiles, float4* part, hipStream_t s);
void launch_project(const ProjectArgs& a, hipStream_t s);
// exact CDF; `work` holds cdf_work_bytes(E) of device memory (parallel form only)
size_t cdf_work_bytes(int E);
void launch_cdf(const float* pmf_raw, int E, float* pmf, float* cdf, void* work, hipStream_t s, bool parallel = true);
// The allocation-free CDF's work buffer: a static device array of NOALLOC_SCRATCH_BYTES per
// GPU, large enough for NOALLOC_MAX_EDGES edges.  noalloc_scratch_init(device) looks its
// address up once per device (at handle creation); noalloc_scratch(device) returns it (NULL
// before that).
constexpr int NOALLOC_MAX_EDGES = 1 << 24;
constexpr size_t NOALLOC_SCRATCH_BYTES = 4u << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
hipError_t noalloc_scratch_init(int device);
void* noalloc_scratch(int device);
void launch_edge_trace(int mode, const EdgeArgs& a, hipStream_t s);   // side colour differences into ecol
void launch_edge_grad(const EdgeArgs& a, hipStream_t s);              // gradients from ecol
void launch_edge_bin(const EdgeBinArgs& a, hipStream_t s);   // bin_count must be zero
void launch_optsimple(const float* grad, float* m, float* G, int off, const OptApplyArgs& a, hipStream_t s);
void launch_adam(const AdamArgs& a, hipStream_t s);
void launch_step_inc(int* step, hipStream_t s);
void launch_transform_update(const TransformArgs& a, hipStream_t s);
void launch_restricted_grad(const RestrictedArgs& a, hipStream_t s);
// Deterministic mode: grad[i] += dacc[i] / 2^32 (i < G); trscr[j] = dacc[G + j] / 2^32.
void launch_det_finish(const unsigned long long* dacc, int G, float* grad, float* trscr, int ntr, hipStream_t s);
void launch_debug_hits(const HitsArgs& a, int tiles, hipStream_t s);

}  // namespace drh
