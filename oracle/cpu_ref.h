/*
 * oracle/cpu_ref.h -- CPU restatement of the reference differentiable-rendering
 * path (FROL256/diffrender_tutorials: DiffRender<SHADING_MODEL>, src/drender.h:74-434).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.  The
 * product (diffrender_tutorials_amd/, libdrh.so) never links or calls it.
 *
 * Every entry point follows a reference function; the .cpp cites file:line for
 * each.  Floating point follows the reference's semantics exactly: no FP
 * contraction (built with -ffp-contract=off), IEEE division and sqrt, the
 * reference's operation order (LiteMath, common_src/LiteMath.h).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Shading models: same numbering as SHADING_MODEL (src/dmesh.h:16-23). */
enum { ORC_SILHOUETTE = 1, ORC_VERTEX_COLOR = 2, ORC_TEXTURE_COLOR = 3, ORC_LAMBERT = 4,
       /* LAMBERT with the interpolated vertex colour as albedo (no counterpart in the reference) */
       ORC_LAMBERT_VERTEX_COLOR = 8 };
/* Edge-sample random numbers: the reference's 14 OpenMP PRNG streams
 * (src/drender.h:306-331) or index-addressed Niederreiter QMC (common_src/qmc.cpp:190-197). */
enum { ORC_RNG_REF_PRNG14 = 0, ORC_RNG_QMC = 1 };
/* Ray tracer: brute force (raytrace_src/raytrace3d_bf.cpp:102-156) or a CPU BVH
 * that returns the same closest hit (stand-in for the Embree path). */
enum { ORC_TRACER_BF = 0, ORC_TRACER_BVH = 1 };

typedef struct {
  float world_view[16]; /* CamInfo::mWorldView, float4x4 memory order (column-major) */
  float proj[16];       /* CamInfo::mProj */
  float width, height;  /* CamInfo::width/height */
} orc_camera;

typedef struct {
  float light_dir[3];   /* direction the light travels (reference: normalize(-1,-1,0)) */
  float light_color[3]; /* reference: (1, 0.5, 0.2) */
  float ambient[3];     /* reference: 0.1*(1,1,1) */
} orc_lambert;

/* Scene: meshes concatenated.  Per mesh i: vcount[i] vertices, fcount[i] triangles,
 * one instance with transform transforms[16*i].  col/tc/nrm may be NULL (reference
 * defaults (0,0,0), (0,0), (1,0,0), common_src/scene_utils.cpp:135-140).
 * Textures: tex_w[i] x tex_h[i] RGB float texels, concatenated (0x0 = none). */
void* orc_scene_create(int n_meshes, const int* vcount, const int* fcount, const float* transforms,
                       const float* pos, const float* col, const float* tc, const float* nrm,
                       const uint32_t* idx, const int* tex_w, const int* tex_h, const float* texels,
                       const int* differentiable);
void  orc_scene_destroy(void* s);
void  orc_scene_set_tracer(void* s, int kind);
/* Jittered pixel samples for orc_render and the aligned interior samples of orc_d_render_ex
 * (the device's drh_set_pixel_jitter pattern of `seed`); off: the reference's cell centres. */
void  orc_set_jitter(void* s, int on, uint32_t seed);
/* The sample position of sample si of pixel (x, y) in view `view` (offsets in the pixel). */
void  orc_pixel_sample(int on, uint32_t seed, int view, int x, int y, int si, int sq, float* xo, float* yo);
/* Size of the flat gradient (DScene full_data of every differentiable mesh, in mesh order). */
int   orc_grad_size(void* s, int mode);

/* Per-pixel closest hit at pixel centres, spp=1 (CastSingleRay(x+0.5, y+0.5)). */
void  orc_hits(void* s, const orc_camera* cam, uint32_t* prim, uint32_t* geom, float* t, float* u, float* v);
/* DiffRender::render (src/drender.h:98-141); images: views x H x W x 3. */
void  orc_render(void* s, int mode, int spp, const orc_lambert* lam, const orc_camera* cams, int views,
                 float* images);
/* DiffRender::d_render (src/drender.h:144-167); accumulates into grad. */
void  orc_d_render(void* s, int mode, int spp, const orc_lambert* lam, const orc_camera* cams, int views,
                   const float* adjoints, int edge_samples, int rng_mode, float* grad);
/* The same with flags: ORC_INTERIOR_ALIGNED puts the interior samples at the render's grid
 * offsets (d + 0.5)/sqrt(spp) (src/drender.h:124-128) instead of the reference's Hammersley
 * offsets (src/drender.h:217-218). */
enum { ORC_INTERIOR_ALIGNED = 1 };
void  orc_d_render_ex(void* s, int mode, int spp, const orc_lambert* lam, const orc_camera* cams, int views,
                      const float* adjoints, int edge_samples, int rng_mode, int flags, float* grad);
/* LossAndDiffLoss (src/utils.cpp:15-27). Returns the float-cast double sum. */
float orc_loss(const float* image, const float* target, int npixels, float* adjoint);

/* Edge set + sampler for one view (src/drender.h:241-302, src/drender.cpp:5-22).
 * Call with v0 == NULL to get the edge count. */
int   orc_edge_sampler(void* s, const orc_camera* cam, int* v0, int* v1, float* pmf, float* cdf,
                       float* proj_pos /* 3*V projected (sx, sy, z_world) or NULL */);

/* Sample sequences. */
void  orc_qmc_table(uint32_t* table /* 11*31 */);
float orc_qmc_rnd(uint32_t pos, int dim, const uint32_t* table);
void  orc_prng_table(int n, float* out /* 2n: (rnd0, rnd1) per sample index */);
void  orc_hammersley(int n, float* out /* 2n */);

/* LiteMath camera helpers (common_src/LiteMath.h:1551-1727). */
void  orc_perspective(float fovy_deg, float aspect, float znear, float zfar, float* out16);
void  orc_look_at(const float* eye, const float* center, const float* up, float* out16);
void  orc_inverse4x4(const float* m, float* out16);
void  orc_mat_mul(const float* a, const float* b, float* out16);

/* OptSimple::OptStep with GD_Adam (src/optimizer_simple.cpp:188-236) for one
 * differentiable mesh: updates m, G in place, turns grad into the step,
 * scaled per block: [0,col_off) pos_lr, [col_off, tex_off) colour lr, ... */
void  orc_optsimple_step(int n, float* grad, float* m, float* G, int nblocks, const int* block_start,
                         const float* block_lr);

/* Standard Adam (Kingma & Ba) on a flat vector; step is 1-based. */
/* FP64 box-filtered silhouette functional L = sum_p (sum of adjoint channels)_p * covered area
 * of pixel p (front-facing triangles of a closed convex mesh), and its central differences for
 * coordinates 3*vertex+axis of the committed scene (pins the edge-sampled position gradient). */
double orc_silhouette_functional(void* s, const orc_camera* cam, const float* adjoint);
void  orc_fd_silhouette(void* s, const orc_camera* cam, const float* adjoint, int n, const int* coords, double h,
                        double* out);
/* FP64 pixel-area functional of the LAMBERT (4) / LAMBERT_VERTEX_COLOR (8) image
 * L = sum_p adj_p . integral over pixel p of the shade (front-facing triangles of a closed
 * convex mesh, the shadow of a face = it turns away from the light; n x n Gauss-Legendre on
 * each triangle-pixel polygon), and its central differences for coordinates 3*vertex+axis
 * (world space).  Pins the position gradient with its interior term. */
double orc_lambert_functional(void* s, int mode, const orc_lambert* lam, const orc_camera* cam, const float* adjoint,
                              int order);
void  orc_fd_lambert(void* s, int mode, const orc_lambert* lam, const orc_camera* cam, const float* adjoint, int n,
                     const int* coords, double h, int order, double* out);
void  orc_adam_step(int n, float* param, const float* grad, float* m, float* v, int step, float lr,
                    float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif
