/*
 * drh.h -- C ABI of the MI355X differentiable-rendering path (libdrh.so).
 *
 * Drop-in for the reference's renderer plugin interface IDiffRender
 * (FROL256/diffrender_tutorials, src/virtual_drender.h:14-26) and its factory
 * MakeDifferentialRenderer (src/drender.cpp:89-147).  Plain pointers and sizes only;
 * host arrays are read or written only during a call.  Every call returns a
 * drh_status (the reference prints and returns, src/drender.h:103-107; here the
 * message is kept in drh_last_error()).  One handle per host thread: a handle owns
 * its own HIP stream and device buffers and is not re-entrant (neither is the
 * reference's DiffRender, whose m_aux / tracer camera change on every call).
 *
 * Layouts follow the reference:
 *   images   views x H x W x 3 float32, row-major, y = 0 is the top row
 *            (Img = LiteImage::Image2D<float3>, src/utils.h:20)
 *   gradient DScene::full_data of every differentiable mesh, in mesh order:
 *            per mesh [pos 3V][colour 3V if VERTEX_COLOR][texture w*h*3 if TEXTURE_COLOR
 *            or LAMBERT][transform 12] (DMesh, src/dmesh.h:35-85)
 *   cameras  CamInfo::mWorldView and mProj as float4x4 memory (column-major),
 *            width, height (common_src/camera.h:6-44)
 */
#ifndef DRH_H
#define DRH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct drh_renderer* drh_handle;

typedef enum {
  DRH_OK = 0,
  DRH_E_INVALID = -1,        /* bad argument (sizes, mode, NULL pointers) */
  DRH_E_NOT_COMMITTED = -2,  /* render/d_render before commit (reference: "renderer was not prepared") */
  DRH_E_HIP = -3,            /* HIP runtime error; see drh_last_error */
  DRH_E_OOM = -4,            /* device allocation failed */
  DRH_E_UNSUPPORTED = -5,    /* a scene feature outside this path (see drh_last_error) */
  DRH_E_DEVICE = -6,         /* a kernel's own check failed (a traversal stack overflowed); the
                                outputs of the calls since the last synchronisation are not valid */
} drh_status;

/* SHADING_MODEL numbering (src/dmesh.h:16-23); PHONG / GGX / PATH_TEST are not on this path.
 * DRH_LAMBERT_VERTEX_COLOR has no counterpart in the reference: shade<LAMBERT>
 * (src/shade_lambert.cpp:6-35) with the interpolated vertex colour as the albedo instead of
 * texture 0, the shading of multi-view mesh fitting with per-vertex colours.  Its gradient
 * block is VERTEX_COLOR's ([pos 3V][colour 3V][transform 12]); the interior term's colour
 * adjoint is shade_grad<VERTEX_COLOR>'s (src/shade_simple.cpp:39-74) times the hit's light
 * factor (shadow * Lc) * max(0, n.(-L)) + ambient, the derivative of the shaded colour. */
enum { DRH_SILHOUETTE = 1, DRH_VERTEX_COLOR = 2, DRH_TEXTURE_COLOR = 3, DRH_LAMBERT = 4,
       DRH_LAMBERT_VERTEX_COLOR = 8 };
/* Edge-sample random numbers: the reference's 14 OpenMP PRNG streams (src/drender.h:306-331),
 * or index-addressed Niederreiter QMC (common_src/qmc.cpp:190-197). */
enum { DRH_RNG_REF_PRNG14 = 0, DRH_RNG_QMC = 1 };
/* Ray tracer: GPU LBVH (default), or brute force in the reference's triangle order. */
enum { DRH_TRACER_BVH = 0, DRH_TRACER_BF = 1 };
/* Optimiser for drh_opt_step: the reference OptSimple "Adam" (src/optimizer_simple.cpp:188-236)
 * or standard Adam. */
enum { DRH_OPT_SIMPLE = 0, DRH_OPT_ADAM = 1 };

typedef struct {
  int mode;             /* DRH_SILHOUETTE ... DRH_LAMBERT           (DiffRenderSettings::mode) */
  int spp;              /* samples per pixel; sqrt(spp)^2 are used   (DiffRenderSettings::spp)  */
  int rng_mode;         /* DRH_RNG_*                                                             */
  int tracer;           /* DRH_TRACER_*                                                          */
  int device;           /* HIP device ordinal                                                    */
  int light_custom;     /* 0: the reference's hard-coded light (src/shade_lambert.cpp:10-12)     */
  float light_dir[3];   /* direction the light travels (used when light_custom != 0)             */
  float light_color[3];
  float ambient[3];
} drh_settings;

typedef struct {
  float world_view[16];
  float proj[16];
  float width, height;
} drh_camera;

/* A scene: meshes concatenated (Scene::add_mesh, common_src/scene.h:130-138), one instance per
 * mesh with a 4x4 transform.  col / tc / nrm may be NULL (reference defaults (0,0,0), (0,0),
 * (1,0,0)); textures are tex_w[i] x tex_h[i] RGB float32 texels, concatenated (0 x 0 = none). */
typedef struct {
  int n_meshes;
  const int* vcount;
  const int* fcount;
  const float* transforms;   /* 16 per mesh, column-major */
  const float* pos;          /* 3 per vertex */
  const float* col;          /* 3 per vertex or NULL */
  const float* tc;           /* 2 per vertex or NULL */
  const float* nrm;          /* 3 per vertex or NULL */
  const uint32_t* idx;       /* 3 per triangle, mesh-local vertex indices */
  const int* tex_w;
  const int* tex_h;
  const float* texels;
  const int* differentiable; /* per mesh 0/1 (the DScene's differentiable mesh ids) */
  /* Optional (NULL: none).  restricted[i] != 0 makes mesh i a RESTRICTED_TRANSFORM mesh
   * (Scene::add_mesh with TransformR, common_src/scene.h:91-145): its transform is
   * TransformR::to_mat of the 7 values restricted_params[7i..7i+6] = translate xyz, rotate xyz
   * (radians), scale (its transforms[] entry is ignored), and its gradient block holds those
   * 7 parameters' derivatives instead of the 12 matrix entries (DMesh, src/dmesh.h:35-85;
   * src/drender.h:398-422). */
  const int* restricted;
  const float* restricted_params;
} drh_scene_desc;

/* MakeDifferentialRenderer + IDiffRender::init (src/drender.cpp:89-147, src/drender.h:79-88). */
drh_status drh_create(const drh_settings* settings, drh_handle* out);
void       drh_destroy(drh_handle h);
const char* drh_last_error(drh_handle h);

/* IDiffRender::commit (src/drender.h:91-96): flattens the scene (Scene::prepare_for_render,
 * common_src/scene_utils.cpp:98-173) into device buffers and builds the LBVH.  At most 2^26
 * triangles and 3 * 2^26 vertices in all (DRH_E_INVALID beyond). */
drh_status drh_commit(drh_handle h, const drh_scene_desc* scene);
/* Number of floats of the flat gradient (DScene::full_size, src/dmesh.h:231-237). */
int        drh_grad_size(drh_handle h);

/* IDiffRender::render (src/drender.h:98-141).  images: host, views x H x W x 3. */
drh_status drh_render(drh_handle h, const drh_camera* cams, int views, float* images);
/* IDiffRender::d_render (src/drender.h:144-167): ACCUMULATES into grad when accumulate != 0
 * (the reference always accumulates; the caller clears), else overwrites. */
drh_status drh_d_render(drh_handle h, const drh_camera* cams, int views, const float* adjoints,
                        int edge_samples, float* grad, int accumulate);
/* IDiffRender::d_render_and_compare (src/drender.h:169-191): commit must have been called
 * (the reference calls commit itself; a caller does drh_commit then this).  Renders, writes
 * out_images (may be NULL), computes the L2 loss and its adjoint 2*(I - T)
 * (LossAndDiffLoss, src/utils.cpp:15-27), clears grad and runs d_render.  *loss =
 * sum of squared errors / (views*W*H).  Views must share one resolution (as in the reference,
 * src/drender.h:180-182). */
drh_status drh_d_render_and_compare(drh_handle h, const drh_camera* cams, int views, const float* targets,
                                    int edge_samples, float* grad, float* out_images, float* loss);

/* Device-resident forms: every array argument is a device pointer (hipMalloc'd, on the
 * handle's device); work is ordered on the handle's stream; no host synchronisation except
 * where a host output (loss) is requested. */
drh_status drh_render_dev(drh_handle h, const drh_camera* cams, int views, float* d_images);
drh_status drh_d_render_dev(drh_handle h, const drh_camera* cams, int views, const float* d_adjoints,
                            int edge_samples, float* d_grad, int accumulate);
drh_status drh_d_render_and_compare_dev(drh_handle h, const drh_camera* cams, int views,
                                        const float* d_targets, int edge_samples, float* d_grad,
                                        float* d_images, float* d_loss);

/* One optimisation step on the device-resident scene (OptSimple::Run's loop body,
 * src/optimizer_simple.cpp:293-305): re-commit from the device parameters (LBVH rebuild),
 * d_render_and_compare against d_targets, then the optimiser update of positions
 * (and colours for DRH_VERTEX_COLOR).  kind DRH_OPT_SIMPLE follows OptStep + OptUpdateScene
 * + StepDecay with OptimizerParameters(GD_Adam) learning rates (pos_lr, col_lr, decay every
 * 100 steps by 0.75); DRH_OPT_ADAM is standard Adam(lr, 0.9, 0.999, 1e-8) on the same
 * parameters.  d_loss (device float, may be NULL) receives the loss of this step. */
drh_status drh_opt_init(drh_handle h, int kind, float pos_lr, float col_lr);
drh_status drh_opt_step(drh_handle h, const drh_camera* cams, int views, const float* d_targets,
                        int edge_samples, float* d_loss);
/* DRH_OPT_ADAM steps after the first are captured once as a hipGraph and replayed (on by
 * default); the graph is re-captured when the arguments or any device buffer change. */
drh_status drh_opt_use_graph(drh_handle h, int on);
/* `steps` consecutive drh_opt_step calls with the same arguments, captured (Adam, graphs on)
 * as one hipGraph and replayed with one launch: the same kernels in the same order, fewer host
 * calls per step.  Graphs are kept per step count and per pattern of LBVH builds and refits
 * (drh_set_bvh_rebuild); the first step after drh_opt_init always runs eagerly.  d_loss
 * receives the loss of the last step.  No counterpart in the reference. */
drh_status drh_opt_steps(drh_handle h, const drh_camera* cams, int views, const float* d_targets,
                         int edge_samples, float* d_loss, int steps);
/* Download the current (optimised) vertex positions (3V) and colours (3V, may be NULL). */
drh_status drh_get_params(drh_handle h, float* pos, float* col);
/* Learning rate of the texture block (OptimizerParameters::textures_lr with _OPT_TEX,
 * src/optimizer.h:26-45, src/optimizer_simple.cpp:168-183); 0 (the default) keeps textures
 * fixed.  With 0 in TEXTURE_COLOR / LAMBERT mode the fused step leaves out the interior pass,
 * whose only output there is the texture gradient: the step's results are unchanged. */
drh_status drh_opt_set_texture_lr(drh_handle h, float tex_lr);
/* Learning rate of the transform block (OptimizerParameters::transforms_lr with _OPT_TFM,
 * src/optimizer.h:26-45; OptUpdateScene, src/optimizer_simple.cpp:113-135): each step moves
 * rows 0-2 of every differentiable mesh's transform by its scaled d_tr entries (row-major,
 * 4 per row), then derives the normal matrix again.  0 (the default) keeps the transforms
 * fixed.  As in the reference, the inverse of mesh 0's transform that the edge gradient
 * multiplies by stays the one of the last drh_commit (Scene::transforms_inv is appended to
 * and never cleared, common_src/scene_utils.cpp:110-122).  The reference's default flags
 * (OPT_ALL) move the transforms with transforms_lr = 0.1. */
drh_status drh_opt_set_transform_lr(drh_handle h, float lr);
/* Download the current (optimised) transforms: 16 floats per mesh, column-major. */
drh_status drh_get_transforms(drh_handle h, float* transforms);
/* Deterministic gradients (off by default): d_render, d_render_and_compare and the fused step
 * add every gradient term as a 32.32 fixed-point integer with 64-bit integer atomics, whose
 * sums do not depend on the order the GPU's waves finish in, and convert once at the end, so
 * repeated runs give the same gradient bit for bit.  Terms of magnitude >= 2^-9 enter
 * exactly; smaller ones are rounded to a multiple of 2^-32; every partial sum must stay below
 * 2^31 in magnitude.  Slower (no LDS sums, more atomics); the edge pass keeps the rnd0 slot
 * order whatever drh_set_edge_order asks (pixel-bin ranks come from atomic counters).  The
 * build counterpart of the
 * reference's serial sum of its per-thread gradient copies (src/drender.h:229-232, 380-425),
 * which is ordered; no counterpart in the reference's interface. */
drh_status drh_set_deterministic(drh_handle h, int on);
/* The interior term's vertex position gradient (off by default, as in the reference, whose
 * shade_grad<VERTEX_COLOR> has it written out but disabled, src/shade_simple.cpp:75-108, and
 * whose shade_grad<LAMBERT> has none): each interior sample adds dF/du du/dV + dF/dv dv/dV
 * for the three vertices of its hit, with the closed-form derivatives of the ray's
 * barycentrics (BarU_grad / BarV_grad, src/functions.cpp:226-470), through mesh 0's inverse
 * transform like the edge term.  VERTEX_COLOR: dF/du = (c_C - c_A).val, dF/dv = (c_B - c_A).val.
 * LAMBERT / LAMBERT_VERTEX_COLOR: F = val.(albedo * ((shadow Lc) max(0, n.(-L)) + ambient))
 * with the interpolated normal n and albedo (texel of the interpolated texture coordinate, or
 * the interpolated vertex colour) differentiated in u and v, the shadow held fixed; a pass of
 * its own (k_interior_pos).  With it the position gradient is the derivative of the image away
 * from silhouettes too.  Other modes: DRH_E_UNSUPPORTED. */
drh_status drh_set_interior_pos_grad(drh_handle h, int on);
/* Download the current (optimised) restricted-transform parameters: 7 floats per mesh
 * (translate, rotate, scale); zeros for meshes that are not RESTRICTED_TRANSFORM meshes. */
drh_status drh_get_restricted(drh_handle h, float* params);
/* Interior samples (interior_derivatives, src/drender.h:197-233).  aligned = 0 (default): the
 * reference's Hammersley offsets, so every interior sample is a fresh ray (recompute).
 * aligned = 1: the render's grid offsets (d + 0.5)/sqrt(spp) instead -- another unbiased
 * estimate of the same pixel integral, not the reference's samples.  store (aligned only):
 *   DRH_HITS_RECOMPUTE (0)  the interior pass traces every sample again;
 *   DRH_HITS_STORE (1)      d_render_and_compare and the fused step keep each render sample's
 *                           closest hit (16 B per sample in HBM) and the interior pass reads it
 *                           back instead of tracing (d_render alone recomputes);
 *   DRH_HITS_AUTO (2)       store when the buffer (16 B x pixels x samples of all views) fits
 *                           the handle's hit-buffer limit (drh_set_hit_buffer_limit) and can be
 *                           allocated, recompute otherwise (holding no buffer).
 * Stored and recomputed hits give the same gradient up to float summation order.  No
 * counterpart in the reference. */
enum { DRH_HITS_RECOMPUTE = 0, DRH_HITS_STORE = 1, DRH_HITS_AUTO = 2 };
drh_status drh_set_interior_samples(drh_handle h, int aligned, int store);
/* Jittered pixel samples (off by default; no counterpart in the reference, whose render puts
 * sample (dx, dy) at the centre of its cell of the sqrt(spp) x sqrt(spp) grid,
 * src/drender.h:124-128).  On: the sample moves to a uniform point of its cell (stratified
 * jitter) drawn from an integer hash of (seed + step, view, x, y, sample), step = 0 for
 * render / d_render / d_render_and_compare and t = 1, 2, ... for the fused step's t-th step, so
 * every optimiser step draws a fresh pattern.  The interior pass with aligned samples uses
 * the same points; the Hammersley interior samples are unchanged. */
drh_status drh_set_pixel_jitter(drh_handle h, int on, unsigned seed);
/* The most HBM DRH_HITS_AUTO may use for stored hits (default: no limit but the device's free
 * memory); the server sets it from its HBM budget.  No counterpart in the reference. */
drh_status drh_set_hit_buffer_limit(drh_handle h, size_t bytes);
/* Bytes the handle holds for stored hits now (0 while it recomputes). */
drh_status drh_get_hit_buffer_bytes(drh_handle h, size_t* bytes);
/* Order in which the edge pass (edge_derivatives, src/drender.h:235-425) visits its samples.
 * The samples, their edges and their contributions are the reference's in every order; only
 * the order of the gradient sums (float rounding) differs.
 *   DRH_EDGE_ORDER_SAMPLE  slots in rnd0 order: a wavefront samples a few consecutive edges.
 *   DRH_EDGE_ORDER_BINNED  the kept samples (on screen, adjoint not ~0) binned by the 8x8
 *                          pixel tile of their sample point first: a wavefront's rays share a
 *                          small screen area whatever the number of samples per edge.
 *   DRH_EDGE_ORDER_AUTO    (default) binned when there are fewer samples than edges, else
 *                          sample order.
 * No counterpart in the reference. */
enum { DRH_EDGE_ORDER_AUTO = 0, DRH_EDGE_ORDER_SAMPLE = 1, DRH_EDGE_ORDER_BINNED = 2 };
drh_status drh_set_edge_order(drh_handle h, int order);
/* LBVH upkeep in drh_opt_step.  The reference re-commits the scene every iteration (a new
 * accelerator over the moved vertices).  every = 1 does that: Morton sort, Karras topology
 * and refit on each step.  every = k > 1 (default 8): step i (counted from drh_opt_init;
 * step 0 uses the commit's tree) runs the full build when i is a multiple of k and otherwise
 * refits the last topology's boxes over the moved geometry.  Traversal is exact over any
 * valid tree, so hits and images are bit-identical for every k and gradients differ only by
 * float summation order, as between any two runs.  No counterpart in the reference. */
drh_status drh_set_bvh_rebuild(drh_handle h, int every);
/* Screen Jacobian in the edge gradient (edge_grad, src/drender.cpp:31-87).
 *   DRH_EDGE_GRAD_REFERENCE_DIAGONAL (default): the reference's terms, dSx/dVx, dSy/dVy and
 *     both z terms -- its parity mode, exact when the view's screen axes are world x and y.
 *   DRH_EDGE_GRAD_FULL_JACOBIAN: adds the cross terms dSy/dVx and dSx/dVy, the derivative of
 *     the rendered silhouette for any view (checked against finite differences).
 * For an axis-aligned view both modes give the same gradient.  Replaces no reference entry
 * point (the reference has only the diagonal form). */
enum { DRH_EDGE_GRAD_REFERENCE_DIAGONAL = 0, DRH_EDGE_GRAD_FULL_JACOBIAN = 1 };
drh_status drh_set_edge_grad_mode(drh_handle h, int mode);
/* Ray/triangle test of the edge samples' two side rays (and their shadow rays), which shade
 * the points 1e-3 px either side of a sampled edge (src/drender.h:345-355).
 *   DRH_SIDE_TEST_REFERENCE (default): the reference's Moller-Trumbore test
 *     (raytrace3d_bf.cpp:123-150), the parity mode.  Its divide by the determinant lets a
 *     nearly edge-on triangle accept a side ray that misses it by far more than the 1e-3 px
 *     offset, which drops that sample's colour difference.
 *   DRH_SIDE_TEST_WATERTIGHT: the watertight test of Woop, Benthin and Wald (JCGT 2013) on the
 *     triangles' shared vertices: such rays are not accepted, and none slips between two
 *     triangles that share an edge.
 * Camera, interior and render rays keep the reference's test in both modes.  No counterpart
 * in the reference's interface. */
enum { DRH_SIDE_TEST_REFERENCE = 0, DRH_SIDE_TEST_WATERTIGHT = 1 };
drh_status drh_set_edge_side_test(drh_handle h, int test);
/* How rays walk the LBVH.
 *   DRH_TRAVERSAL_AUTO (default): per launch, from the rays' expected coherence -- camera
 *     rays walk per lane when the scene has more than 0.5 triangles per pixel of the view,
 *     the edge pass's side rays when there are more than 0.5 edges per edge sample;
 *   DRH_TRAVERSAL_WAVE: a wave's 64 rays walk the tree as one packet (shared node loads and
 *     stack), fastest for coherent rays;
 *   DRH_TRAVERSAL_LANE: each ray walks alone (own stack), fastest when a packet's rays open
 *     many different subtrees.
 * Every mode returns the same hits bit for bit (the closest hit with the (t, triangle id)
 * rule does not depend on the order the triangles are tested in).  No counterpart in the
 * reference's interface (its tracer is brute force, raytrace3d_bf.cpp). */
enum { DRH_TRAVERSAL_AUTO = 0, DRH_TRAVERSAL_WAVE = 1, DRH_TRAVERSAL_LANE = 2 };
drh_status drh_set_traversal(drh_handle h, int mode);
/* How many handles share the device (1 by default).  A handle alone runs view 0's edge
 * preparation on a second stream beside the render; when several share the device (the
 * server sets its slot count) that overlap comes from the other handles instead and the
 * second stream is left idle (it would contend for hardware queues).  No counterpart in the
 * reference. */
drh_status drh_set_device_sharing(drh_handle h, int handles);
/* Download the committed scene's textures (all meshes, in commit order; the floats of
 * drh_scene_desc.texels), as the optimiser has left them. */
drh_status drh_get_texels(drh_handle h, float* texels);

/* Device memory plumbing (no counterpart in the reference): for callers that hold no HIP
 * allocator of their own, e.g. Python tests and the benchmark.  Allocations come from the
 * device's stream-ordered pool on the handle's stream and are usable at return; drh_free
 * waits for the handle's stream, then returns the memory to the pool without waiting for the
 * rest of the device (free it only once no other stream still uses it).  Copies are ordered
 * on the handle's stream and complete before return. */
drh_status drh_malloc(drh_handle h, size_t bytes, void** d_ptr);
drh_status drh_free(drh_handle h, void* d_ptr);
drh_status drh_memcpy_htod(drh_handle h, void* d_dst, const void* src, size_t bytes);
drh_status drh_memcpy_dtoh(drh_handle h, void* dst, const void* d_src, size_t bytes);

/* ---------------------------------------------------------------------------------------------
 * Serving (no counterpart in the reference, whose optimiser fits one scene at a time,
 * src/optimizer_simple.cpp:272-309): a continuous-batching server for mesh-fitting jobs.
 * Jobs are queued by drh_server_submit (every input copied), admitted by a worker thread while
 * fewer than max_active run and the HBM budget holds their estimated footprint, and stepped
 * concurrently (one fused commit + render + loss + d_render + Adam step per iteration, each job
 * on its own stream, replayed as a hipGraph).  A job larger than the whole budget is refused
 * with DRH_E_OOM.
 * ------------------------------------------------------------------------------------------- */
typedef struct drh_server_s* drh_server;

typedef struct {
  drh_settings settings;       /* mode, spp, light, rng (tracer / device are the server's) */
  drh_scene_desc scene;        /* initial mesh */
  const drh_camera* cams;
  int views;
  const float* targets;        /* views x H x W x 3 target images */
  int edge_samples;            /* per view per iteration */
  int iterations;
  float lr_pos, lr_col;        /* Adam learning rates */
  int flags;                   /* DRH_JOB_* (0: the defaults of a fresh handle) */
} drh_job_desc;

/* drh_job_desc.flags:
 *   DRH_JOB_DETERMINISTIC     drh_set_deterministic(1): a job's results do not depend on what
 *                             else the server runs beside it, bit for bit;
 *   DRH_JOB_ALIGNED_INTERIOR  drh_set_interior_samples(1, DRH_HITS_AUTO): the render's hits are
 *                             stored for the interior pass when they fit what is left of the
 *                             server's HBM budget at admission (charged to the job), and
 *                             recomputed otherwise;
 *   DRH_JOB_FULL_JACOBIAN     drh_set_edge_grad_mode(DRH_EDGE_GRAD_FULL_JACOBIAN). */
enum { DRH_JOB_DETERMINISTIC = 1, DRH_JOB_ALIGNED_INTERIOR = 2, DRH_JOB_FULL_JACOBIAN = 4 };

typedef struct {
  drh_status status;
  int iterations;              /* completed */
  float final_loss;            /* loss of the last iteration */
  double queue_ms, run_ms;     /* submit -> admit, admit -> done */
  float* pos;                  /* out (may be NULL): 3V fitted positions */
  float* col;                  /* out (may be NULL): 3V colours */
  float* latencies_ms;         /* out (may be NULL): per-iteration latency */
  int latencies_cap;
  size_t hbm_bytes;            /* HBM charged against the budget (footprint + stored hits) */
  size_t hit_bytes;            /* of which the stored hits (0: the interior pass recomputed) */
} drh_job_result;

typedef struct {
  long long iterations;        /* completed over all jobs */
  int jobs_done, jobs_failed, jobs_queued;
  double p50_ms, p99_ms;       /* per-iteration latency */
  int peak_active;
  double mean_active;
  size_t hbm_used;
  double elapsed_s;
  double admit_s, step_s, finish_s;  /* worker-thread time: admitting jobs (create, commit,
                                        uploads), issuing steps, finishing jobs (downloads, free) */
} drh_server_stats_t;

drh_status drh_server_create(int device, size_t hbm_budget_bytes, int max_active, drh_server* out);
drh_status drh_server_submit(drh_server s, const drh_job_desc* job, int* job_id);
drh_status drh_server_wait(drh_server s, int job_id, double timeout_s, drh_job_result* result);
drh_status drh_server_stats(drh_server s, drh_server_stats_t* out);
/* The HBM footprint the server charges a job against its budget (bytes). */
size_t     drh_server_job_bytes(const drh_job_desc* job);
void       drh_server_destroy(drh_server s);

/* The handle's HIP stream (a hipStream_t) and device synchronisation of it. */
void*      drh_stream(drh_handle h);
drh_status drh_sync(drh_handle h);

/* Ray counts (no counterpart in the reference): with counting on, render / d_render /
 * d_render_and_compare / the fused step run counting copies of the traversal kernels (eagerly,
 * not from a graph) that add up the rays they cast; drh_count_rays(h, 1) also zeroes the
 * counts.  drh_read_ray_counts returns [0] render camera rays, [1] render shadow rays,
 * [2] interior closest-hit rays, [3] interior any-hit rays (coverage or shadow rays),
 * [4] edge-sample side rays, [5] edge-sample shadow rays, summed since counting was turned on. */
drh_status drh_count_rays(drh_handle h, int on);
drh_status drh_read_ray_counts(drh_handle h, unsigned long long counts[6]);

/* Per-kernel timing with HIP events recorded on the handle's stream around each launch.
 * drh_profile_read returns the summed milliseconds and launch count of kernel `name`
 * (names as in the rocprof trace) since the last drh_profile_enable. */
drh_status drh_profile_enable(drh_handle h, int on);
drh_status drh_profile_read(drh_handle h, const char* name, double* total_ms, int* launches);

/* Diagnostics for parity tests: per-pixel closest hit at pixel centres (spp 1) for one camera,
 * and the LBVH's sorted Morton keys / triangle order and the edge sampler of the last
 * d_render view. */
drh_status drh_debug_hits(drh_handle h, const drh_camera* cam, uint32_t* prim, uint32_t* geom,
                          float* t, float* u, float* v);
drh_status drh_debug_bvh(drh_handle h, uint32_t* morton_sorted, uint32_t* tri_order, float* scene_bounds);
drh_status drh_debug_edges(drh_handle h, int* n_edges, int* v0, int* v1, float* pmf, float* cdf);
/* The edge sampler's CDF over an arbitrary pmf (host arrays): pmf_out = pmf_raw / total,
 * cdf_out[0..E] = sequential float running sum / total, by the parallel exact kernel
 * (parallel != 0) or the single-lane chain. */
drh_status drh_debug_cdf(drh_handle h, const float* pmf_raw, int E, float* pmf_out, float* cdf_out, int parallel);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  The edge sampler's exact CDF (build_edge_sampler,
 * src/drender.cpp:5-22; drh_debug_cdf with parallel != 0) over device arrays: d_pmf = d_pmf_raw /
 * total, d_cdf[0..E] = the sequential float running sum / total.  It allocates nothing: its
 * work buffer is a static 4 MiB device array per GPU, so E is limited to 2^24 edges (larger
 * E: DRH_E_INVALID).  Work is queued on the handle's stream and not waited for.  Every call on
 * one GPU shares that GPU's array (it keeps the last call's contents; every word a call reads
 * it writes first): calls that overlap on one GPU (other handles' streams) race. */
drh_status drh_build_edge_sampler_noalloc(drh_handle h, const float* d_pmf_raw, int E, float* d_pmf, float* d_cdf);
/* A plain device-to-device copy kernel (k_copy_f32) of n floats: the known-size reference
 * that hardware traffic counters (FETCH_SIZE / WRITE_SIZE) are calibrated against. */
drh_status drh_debug_copy(drh_handle h, float* d_dst, const float* d_src, long long n);

#ifdef __cplusplus
}
#endif
#endif /* DRH_H */
