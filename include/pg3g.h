/*
 * pg3g.h -- C ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * The library replaces the inner work of PG3Render's CPU renderers behind the
 * reference's own renderer boundary, `AbstractRenderer` (src/renderer.hxx:232-278):
 *
 *   AbstractRenderer::RunIteration(Algorithm, uint32_t)  -> pg3g_render()
 *       (src/path_tracer_base.hxx:57-90, one sample per pixel per iteration)
 *   AbstractRenderer::GetFramebuffer(Framebuffer&)         -> pg3g_read_framebuffer()
 *       (src/renderer.hxx:255-261; FP64 sums, src/frame_buffer.hxx:30-77)
 *   Scene (geometry, materials, lights, camera, EM)        -> pg3g_upload_scene()
 *       (src/scene.hxx:42-113 queries, :176-956 builder output)
 *
 * Every call returns an int status (0 = PG3G_OK) and throws nothing across the
 * ABI; pg3g_last_error() returns the message of the last failure on the calling
 * host thread.  A context is used from one host thread at a time.  All pointer
 * arguments are host pointers unless the name ends in `_dev`.
 */
#ifndef PG3G_H
#define PG3G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------------------- */
#define PG3G_OK               0
#define PG3G_ERR_INVALID_ARG  1
#define PG3G_ERR_HIP          2
#define PG3G_ERR_UNSUPPORTED  3
#define PG3G_ERR_OUT_OF_MEMORY 4
#define PG3G_ERR_NO_SCENE     5

/* ---- algorithms: same numbering as `enum Algorithm` (src/config.hxx:23-34) ------------- */
#define PG3G_ALG_EYE_LIGHT            0
#define PG3G_ALG_NORMAL_VISUALISER    1
#define PG3G_ALG_DIRECT_BSDF          2  /* dbs  */
#define PG3G_ALG_DIRECT_LIGHT_ALL     3  /* dlsa */
#define PG3G_ALG_DIRECT_LIGHT_SINGLE  4  /* dlss */
#define PG3G_ALG_DIRECT_MIS           5  /* dmis */
#define PG3G_ALG_PATH_TRACING_NAIVE   6  /* ptn  */
#define PG3G_ALG_PATH_TRACING         7  /* pt   */

/* ---- materials (src/materials.hxx) ---------------------------------------------------- */
#define PG3G_MAT_LAMBERT           0  /* LambertMaterial                 :253-350  */
#define PG3G_MAT_PHONG             1  /* PhongMaterial                   :353-601  */
#define PG3G_MAT_GGX_CONDUCTOR     2  /* MicrofacetGGXConductorMaterial  :788-989  */
#define PG3G_MAT_GGX_DIELECTRIC    3  /* MicrofacetGGXDielectricMaterial :992-1298 */
#define PG3G_MAT_SMOOTH_CONDUCTOR  4  /* SmoothConductorMaterial         :657-710  */
#define PG3G_MAT_SMOOTH_DIELECTRIC 5  /* SmoothDielectricMaterial        :713-785  */

typedef struct pg3g_material {
    int32_t type;              /* PG3G_MAT_* */
    int32_t light_id;          /* emitter this material belongs to, -1 if none
                                  (the material->light map, src/scene.hxx:51-55) */
    float   diffuse[3];        /* Lambert reflectance; Phong diffuse reflectance as passed to the
                                  PhongMaterial ctor (halved by the library when glossy is on) */
    float   glossy[3];         /* Phong glossy reflectance; GGX specular tint (1,1,1 = reference) */
    float   phong_exponent;    /* Phong exponent */
    int32_t phong_diffuse_on;  /* PhongMaterial ctor flags aDiffuse / aGlossy */
    int32_t phong_glossy_on;
    float   alpha;             /* GGX roughness, clamped to [0.001, 1] as in the ctors */
    float   inner_ior;         /* n inside */
    float   outer_ior;         /* n outside */
    float   absorbance;        /* conductor k */
    int32_t back_side_light_sampling; /* GGX dielectric ctor flag (materials.hxx:1000-1007) */
} pg3g_material;

/* ---- lights (src/lights.hxx) ---------------------------------------------------------- */
#define PG3G_LIGHT_AREA        0  /* AreaLight: one-sided triangle p0,p1,p2 */
#define PG3G_LIGHT_POINT       1  /* PointLight at p0 */
#define PG3G_LIGHT_BACKGROUND  2  /* BackgroundLight: constant radiance or the EM of the scene */

typedef struct pg3g_light {
    int32_t type;
    float   p0[3], p1[3], p2[3];
    float   power[3];          /* flux as given to SetPower(); for the background: constant
                                  radiance (SetConstantRadiance) when the scene has no EM image */
} pg3g_light;

/* ---- environment map ------------------------------------------------------------------ */
#define PG3G_EM_NONE        0
#define PG3G_EM_COSINE      1  /* EnvironmentMapCosineSampler      src/em_cosine_sampler.hxx  */
#define PG3G_EM_SIMPLE      2  /* EnvironmentMapSimpleSphericalSampler (2-D luminance CDF)    */
#define PG3G_EM_STEERABLE   3  /* EnvironmentMapSteerableSampler   src/em_steerable_sampler.hxx */

typedef struct pg3g_em_desc {
    uint32_t     width, height;  /* lat-long RGB image; 0 = no image */
    const float *texels;         /* width*height*3 FP32, row-major (em_image.hxx:185-199) */
    int32_t      bilinear;       /* tent filtering (default on, lights.hxx:372) */
    int32_t      sampler;        /* PG3G_EM_* */
    /* steerable BuildParameters (em_steerable_sampler.hxx:27-57); negative = reference default */
    float        st_max_approx_error;   /* <= 0: default 5.0 */
    int32_t      st_min_subdiv_level;   /* < 0: default 4 */
    int32_t      st_max_subdiv_level;   /* < 0: default 9 */
} pg3g_em_desc;

/* ---- camera (src/camera.hxx:13-59) ---------------------------------------------------- */
typedef struct pg3g_camera_desc {
    float    position[3];
    float    forward[3];
    float    up[3];
    float    hfov_deg;
    uint32_t res_x, res_y;
} pg3g_camera_desc;

/* ---- render parameters (Config, src/config.hxx:556-578) ------------------------------- */
typedef struct pg3g_render_params {
    int32_t  algorithm;          /* PG3G_ALG_* */
    uint32_t min_path_length;    /* -minpl (default 1) */
    uint32_t max_path_length;    /* -maxpl (0 = Russian roulette) */
    float    indirect_clipping;  /* -iic (0 = off) */
    float    splitting_budget;   /* -sb (1..64; 1 with -iic 0 is the plain path tracer) */
    float    splitting_level;    /* -sl (default 1) */
    float    light_to_bsdf_ratio;/* -slbr (default 1) */
    uint64_t seed;               /* counter-based RNG key */
    uint32_t scene_id;           /* RNG stream id of the scene (batched renders) */
} pg3g_render_params;

/* ---- scene ------------------------------------------------------------------------------ */
typedef struct pg3g_scene_desc {
    uint32_t             num_triangles;
    const float         *triangles;       /* num_triangles * 9: p0, p1, p2 (scene_graph.hxx:88-101) */
    const int32_t       *triangle_mat;    /* material index per triangle */
    uint32_t             num_spheres;
    const float         *spheres;         /* num_spheres * 4: centre, radius (scene_graph.hxx:164-174) */
    const int32_t       *sphere_mat;
    uint32_t             num_materials;
    const pg3g_material *materials;
    uint32_t             num_lights;
    const pg3g_light    *lights;          /* Scene::mLights order; at most one background light */
    pg3g_em_desc         em;              /* image of the background light, if any */
    pg3g_camera_desc     camera;
    pg3g_render_params   params;
} pg3g_scene_desc;

/* ---- context -------------------------------------------------------------------------- */
typedef struct pg3g_context pg3g_context;

/* Creates a context on HIP device `device`.  Replaces the per-thread renderer construction in
 * CreateRenderer (src/pg3render.cxx:24-50). */
int pg3g_create(pg3g_context **ctx, int device);
int pg3g_destroy(pg3g_context *ctx);
/* Message of the last failed call on this host thread ("" if none). */
const char *pg3g_last_error(void);

/* Uploads a scene into slot 0 and builds its EM sampling tables (FP64 luminance CDF prefix sums
 * on the GPU; steerable triangulation + 6-ary tree on the host).  Replaces Scene::LoadCornellBox's
 * output (src/scene.hxx:176-956) as seen by the renderers and EnvironmentMap's sampler Init
 * (src/em.hxx:25-69).  Clears the framebuffer of the slot. */
int pg3g_upload_scene(pg3g_context *ctx, const pg3g_scene_desc *desc);
/* Same, into a numbered slot (0..PG3G_MAX_SCENES-1) for batched renders. */
#define PG3G_MAX_SCENES 256
int pg3g_upload_scene_slot(pg3g_context *ctx, uint32_t slot, const pg3g_scene_desc *desc);
/* Replaces the render parameters of a slot without rebuilding tables. */
int pg3g_set_params(pg3g_context *ctx, uint32_t slot, const pg3g_render_params *params);

/* Renders iterations [first_iter, first_iter + n_iter) of slot 0: each iteration is one sample per
 * pixel, exactly PathTracerBase::RunIteration (src/path_tracer_base.hxx:57-90), accumulated into the
 * slot's FP64 framebuffer.  Asynchronous on the context stream. */
int pg3g_render(pg3g_context *ctx, uint32_t first_iter, uint32_t n_iter);
/* PathTracerBase::RunIteration of the drop-in renderer (src/path_tracer_base.hxx:57-90) as
 * AbstractRenderer::RunIteration calls it (src/renderer.hxx:232-278): queues iteration `iteration`
 * of a slot.  Consecutive queued iterations are launched together, up to PG3G_ITERATION_RUN per
 * launch (one launch per iteration would pay the launch's ramp-up and drain every time; 256
 * iterations of a 512x512 frame are about 40 ms of GPU work);
 * launch_now != 0 launches what is queued at once (the last iteration of a fixed count, or every
 * iteration of a time-limited render).  An iteration that does not follow the queued ones launches
 * them first.  Reading the framebuffer (pg3g_read_framebuffer*, pg3g_framebuffer_device_ptr),
 * pg3g_synchronize, pg3g_render* of the slot and pg3g_set_params launch what is queued before they
 * act; pg3g_clear_framebuffer and an upload to the slot drop it.  Asynchronous. */
#define PG3G_ITERATION_RUN 256
int pg3g_run_iteration(pg3g_context *ctx, uint32_t slot, uint32_t iteration, int launch_now);
int pg3g_render_slot(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter);

/* Same as pg3g_render_slot, bracketed by HIP events recorded on the context stream (the stream
 * the kernels run on); waits for completion and returns the elapsed GPU time in milliseconds. */
int pg3g_render_timed(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter,
                      float *elapsed_ms);

/* Kernel organisation of pg3g_render*: one fused kernel (default) or wavefront queues. */
#define PG3G_KERNEL_WAVEFRONT 0
#define PG3G_KERNEL_MEGAKERNEL 1
int pg3g_set_kernel_mode(pg3g_context *ctx, int mode);

/* Copies the FP64 per-pixel RGB sums (res_x*res_y*3) and the iteration count of a slot.
 * The image the reference writes is sum / iterations (renderer.hxx:255-261). */
int pg3g_read_framebuffer(pg3g_context *ctx, double *sum_rgb, uint64_t *iterations);
int pg3g_read_framebuffer_slot(pg3g_context *ctx, uint32_t slot, double *sum_rgb, uint64_t *iterations);
int pg3g_clear_framebuffer(pg3g_context *ctx, uint32_t slot);
/* Device address of a slot's framebuffer (res_x*res_y*3 doubles) for zero-copy consumers. */
int pg3g_framebuffer_device_ptr(pg3g_context *ctx, uint32_t slot, double **fb_dev);
int pg3g_synchronize(pg3g_context *ctx);

/* Number of path segments (shaded path vertices: an extension ray and at most one shadow ray each)
 * and camera samples traced since the last reset. */
int pg3g_read_counters(pg3g_context *ctx, uint64_t *segments, uint64_t *camera_samples);
/* Rays the path-tracing kernels actually traced since the last reset: extension rays (camera rays
 * included; Scene::Intersect, src/scene.hxx:42-59) and shadow rays (Scene::Occluded, :61-74).
 * The differentiable kernels (pg3g_render_diff, pg3g_adjoint) count segments and samples only. */
int pg3g_read_ray_counters(pg3g_context *ctx, uint64_t *extension_rays, uint64_t *shadow_rays);
int pg3g_reset_counters(pg3g_context *ctx);
/* Primitives of an uploaded slot as the intersection kernels see them: axis-aligned rectangles
 * (pairs of the scene's triangles merged at upload: the reference's box walls, src/scene.hxx:
 * 459-488, tested as one primitive with the triangles' own hit distance, normal and material),
 * the remaining triangles and the spheres (GeometryList, src/scene_graph.hxx:33-82).  Any output
 * pointer may be null.  Introspection of the upload; no counterpart in the reference. */
int pg3g_scene_primitives(pg3g_context *ctx, uint32_t slot, uint32_t *rectangles, uint32_t *triangles,
                          uint32_t *spheres);

/* ---- function-level entry points (tier-1 parity; host arrays, n records) ----------------- */
/* AbstractMaterial::EvalBsdf with kOptSamplingProbs (materials.hxx:228-229): wil/wol local
 * directions (n*3) -> f (n*3), pdfW, compProb. */
int pg3g_bsdf_eval_batch(pg3g_context *ctx, const pg3g_material *mat, uint32_t n,
                         const float *wil, const float *wol,
                         float *f, float *pdf, float *comp_prob);
/* AbstractMaterial::SampleBsdf (materials.hxx:231-234) with explicit uniforms u (n*3):
 * u[0],u[1] = direction sample, u[2] = component / reflect-refract choice.
 * Outputs wil (n*3), attenuation (n*3), pdfW, compProb. */
int pg3g_bsdf_sample_batch(pg3g_context *ctx, const pg3g_material *mat, uint32_t n,
                           const float *wol, const float *u,
                           float *wil, float *f, float *pdf, float *comp_prob);
/* EnvironmentMapSampler::Sample / PdfW and EnvironmentMapImage::Evaluate of slot 0's EM
 * (em_sampler.hxx:49-64, em_image.hxx:95-150).  normals: n*3 world surface normals (frame from
 * Frame::SetFromZ), side_flags: bit0 front, bit1 back; u: n*3 uniforms.
 * Sample outputs: dir (n*3), pdf, radiance*cos (n*3), ok. */
int pg3g_em_sample_batch(pg3g_context *ctx, uint32_t n, const float *normals,
                         const int32_t *side_flags, const float *u,
                         float *dir, float *pdf, float *radiance_cos, int32_t *ok);
int pg3g_em_pdf_batch(pg3g_context *ctx, uint32_t n, const float *dirs, const float *normals,
                      const int32_t *side_flags, float *pdf);
int pg3g_em_eval_batch(pg3g_context *ctx, uint32_t n, const float *dirs, float *rgb);
/* Scene::Intersect (src/scene.hxx:42-59; GeometryList / Triangle / Sphere::Intersect,
 * src/scene_graph.hxx:33-245) of slot 0's scene: the closest hit beyond tmin of n rays
 * (org, dir: n*3; tmin: n).  Outputs dist (n; -1 for a miss), the geometric normal (n*3) and the
 * material index (n; -1 for a miss). */
int pg3g_intersect_batch(pg3g_context *ctx, uint32_t n, const float *org, const float *dir,
                         const float *tmin, float *dist, float *normal, int32_t *mat);
/* Scene::Occluded (src/scene.hxx:61-74): whether anything lies on the segment from
 * p + dir * 1e-3 to distance tmax - 2e-3 along dir (p, dir: n*3; tmax: n) -> occluded (n, 0/1). */
int pg3g_occluded_batch(pg3g_context *ctx, uint32_t n, const float *p, const float *dir,
                        const float *tmax, int32_t *occluded);

/* Introspection of the steerable tables of slot 0 (for parity with the reference build). */
int pg3g_steerable_info(pg3g_context *ctx, uint32_t *num_leaves, uint32_t *num_vertices,
                        uint32_t *num_levels);
/* Copies leaf triangles in tree order: vertex directions (num_leaves*9). */
int pg3g_steerable_leaves(pg3g_context *ctx, float *leaf_dirs);
/* Root weight (9 SH floats) = sum of all leaf weights grouped as the 6-ary tree does. */
int pg3g_steerable_root_weight(pg3g_context *ctx, float *weight9);

/* Host-only build of the steerable sampler tables of an EM (no context, no GPU): the
 * triangulation refined by the approximation-error test and the 6-ary tree.  Replaces
 * EnvironmentMapSteerableSampler::Init's build (TriangulateEm / RefineEmTriangulation /
 * BuildTriangleTree, src/em_steerable_sampler.hxx:2963-3472, 3755-3793), which refines one
 * triangle at a time; here each level's subdivision tests run on num_threads host threads
 * (0 = PG3G_BUILD_THREADS, else OMP_NUM_THREADS, else the hardware threads), with the same
 * tables for any thread count.  The same build runs inside pg3g_upload_scene.  Outputs may be
 * null: leaf_dirs needs num_leaves*9 floats (leaf triangles in tree order, as
 * pg3g_steerable_leaves), root_weight9 9 floats; build_ms receives the wall time of the build. */
typedef struct pg3g_steerable_build_info {
    uint32_t num_leaves, num_vertices, num_levels, threads;
    uint32_t from_cache;         /* 1: loaded from the steerable cache (pg3g_steerable_cache_dir) */
    double   build_ms;
} pg3g_steerable_build_info;
int pg3g_steerable_build(const pg3g_em_desc *em, uint32_t num_threads, pg3g_steerable_build_info *info,
                         float *leaf_dirs, uint32_t leaf_capacity, float *root_weight9);

/* Directory of the on-disk cache of steerable tables (null or "": none; default: the
 * PG3G_STEERABLE_CACHE_DIR environment variable).  Replaces the reference's .emssd files
 * (EnvironmentMapSteerableSampler::SaveToDisk / LoadFromDisk, src/em_steerable_sampler.hxx:
 * 1596-2030), which are keyed by the EM file name: entries here are keyed by a hash of the
 * texels, the image size and filter, and every build parameter.  Used by pg3g_upload_scene,
 * pg3g_diff_set_params (refreshed sampling) and pg3g_steerable_build.  Process-wide. */
int pg3g_steerable_cache_dir(const char *dir);

/* ---- image and environment-map files (host only) ---------------------------------------- */
/* Reads a Radiance RGBE (.hdr: flat or run-length scanlines, -Y h +X w), PFM or raw float image
 * (uint32 width, uint32 height, width*height RGB float32 top row first) into row-major RGB
 * floats, top row first.  rgb may be null to query the size; capacity counts floats. */
int pg3g_load_image(const char *path, uint32_t *width, uint32_t *height, float *rgb, uint64_t capacity);
/* EnvironmentMapImage::LoadImage (src/em_image.hxx:51-93) without OpenEXR: an image as
 * pg3g_load_image, rotated in azimuth by azimuth_rotation (fraction of a turn, taken modulo 1 as
 * Math::FmodX) and multiplied by scale. */
int pg3g_load_em(const char *path, float azimuth_rotation, float scale, uint32_t *width, uint32_t *height,
                 float *rgb, uint64_t capacity);
/* Framebuffer::SaveHDR / SaveBMP (src/frame_buffer.hxx:114-200) of an RGB FP64 image (row-major,
 * top row first) times scale (1 / iterations for pg3g_read_framebuffer sums).  The format follows
 * the extension: .hdr (RGBE, flat scanlines) or .bmp (24-bit, value^(1/gamma), clamped). */
int pg3g_save_image(const char *path, const double *rgb, uint32_t width, uint32_t height, double scale,
                    float gamma);
/* Per-pixel comparison of two RGB images on luminance, the check of the reference's regression
 * script (testing/full_test.sh:97, an external diff tool in relative-RMSE mode):
 * pixel_rel_rmse = sqrt(mean(((L - Lref) / max(|Lref|, 1e-3 mean|Lref|))^2)),
 * image_rel_rmse = sqrt(sum (L - Lref)^2 / sum Lref^2).  diff_rgb (may be null) receives
 * |image - reference| per channel. */
typedef struct pg3g_image_diff {
    double pixel_rel_rmse, image_rel_rmse, max_pixel_rel;
    double mean_luminance, mean_luminance_reference;
} pg3g_image_diff;
int pg3g_image_compare(const float *image, const float *reference, uint32_t width, uint32_t height,
                       pg3g_image_diff *out, float *diff_rgb);

/* ---- batched multi-scene rendering ------------------------------------------------------ */
/* No counterpart in the reference, which renders one scene per process (src/pg3render.cxx:
 * 54-174).  A batch renders many uploaded slots in ONE persistent launch per pass: every scene
 * is cut into work items (an 8x8 pixel tile x a run of iterations) that resident wavefronts take
 * from a device counter.  Pass p renders iterations [p*spp[i], (p+1)*spp[i]) of scene i, traced
 * exactly like pg3g_render_slot (same samples), into the slot's framebuffer; a pass is a
 * captured hipGraph (reset counter, persistent kernel, advance pass) replayed per call.
 * Scenes must use the path tracer (`pt`).  A batch keeps copies of its slots' scene tables:
 * uploading a scene or setting parameters on the context invalidates every batch of it
 * (pg3g_batch_render then fails with PG3G_ERR_NO_SCENE). */
typedef struct pg3g_batch pg3g_batch;
typedef struct pg3g_batch_config {
    uint32_t iters_per_item;   /* iterations of one tile per work item (>= 1) */
    uint32_t waves_per_cu;     /* persistent wavefronts per compute unit (multiple of 4, >= 4) */
    uint32_t use_graph;        /* 1: replay the captured hipGraph; 0: launch the three kernels */
} pg3g_batch_config;

int pg3g_batch_create(pg3g_context *ctx, uint32_t num_scenes, const uint32_t *slots,
                      const uint32_t *spp, pg3g_batch **batch);
/* A batch refers to its context: destroy every batch before pg3g_destroy of its context. */
int pg3g_batch_destroy(pg3g_batch *batch);
int pg3g_batch_set_config(pg3g_batch *batch, const pg3g_batch_config *config);
int pg3g_batch_get_config(const pg3g_batch *batch, pg3g_batch_config *config);
/* Renders `passes` passes (asynchronous). */
int pg3g_batch_render(pg3g_batch *batch, uint32_t passes);
/* Same, bracketed by HIP events on the context stream; waits and returns the GPU time. */
int pg3g_batch_render_timed(pg3g_batch *batch, uint32_t passes, float *elapsed_ms);
/* Autotuner: times one pass for each candidate configuration (iterations per item x
 * persistent wavefronts per CU, at most `max_candidates`, 0 = all), keeps the fastest, then
 * clears the batch's framebuffers and restarts at pass 0.  Returns the best pass time. */
int pg3g_batch_autotune(pg3g_batch *batch, uint32_t max_candidates, float *best_pass_ms);

/* ---- differentiable rendering ------------------------------------------------------------ */
/* No counterpart in the reference.  Path-replay backpropagation of an image loss for the path
 * tracer (`pt`, -maxpl > 0) with frozen sampling: directions, pdfs and MIS weights come from the
 * sampling parameters, integrand values (BSDF values, EM radiance) from the evaluation
 * parameters; the two are equal after a refresh.  Differentiated: Lambert albedo (3 scalars),
 * GGX conductor tint + alpha (4 scalars, in that order), and the EM texels (FP64 gradients). */
/* Selects the differentiated materials (their order fixes the scalar layout). */
int pg3g_diff_setup(pg3g_context *ctx, uint32_t slot, uint32_t num_materials,
                    const uint32_t *material_ids, uint32_t *num_scalars);
/* Current evaluation parameters: scalars (num_scalars) and EM texels (w*h*3), either may be NULL. */
int pg3g_diff_get_params(pg3g_context *ctx, uint32_t slot, float *scalars, float *em_texels);
/* Sets evaluation parameters (NULL keeps them); refresh_sampling != 0 also makes them the sampling
 * parameters (materials and EM sampling tables rebuilt). */
int pg3g_diff_set_params(pg3g_context *ctx, uint32_t slot, const float *scalars, const float *em_texels,
                         int refresh_sampling);
/* Differentiable forward render of iterations [first_iter, first_iter + n_iter) into the slot's
 * framebuffer (like pg3g_render_slot).  elapsed_ms != NULL: waits and returns the GPU time. */
int pg3g_render_diff(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter,
                     float *elapsed_ms);
/* Adjoint of the same iterations: dloss_dimage (res_x*res_y*3) is dLoss/dImage with image =
 * sum / n_iter.  Writes grad_scalars (num_scalars) and grad_em (w*h*3, FP64 accumulated on the
 * GPU); either may be NULL.  Waits; elapsed_ms (may be NULL) receives the GPU time. */
int pg3g_adjoint(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter,
                 const float *dloss_dimage, double *grad_scalars, double *grad_em, float *elapsed_ms);

/* Multi-view differentiable rendering (no counterpart in the reference; the inverse-rendering
 * workload renders one scene from several cameras at every optimisation step): the slot's scene,
 * parameters and sampling tables seen from num_views cameras (each with the slot's resolution;
 * num_views = 0 removes them).  View v's samples use the RNG streams of scene id scene_ids[v],
 * so they equal the samples of the same scene uploaded with that camera and scene id.  One launch
 * renders every view; pg3g_render_diff_views returns the FP64 sums of iterations [first_iter,
 * first_iter + n_iter) of each view (num_views x res_y x res_x x 3, view-major; the image is sum /
 * n_iter), and pg3g_adjoint_views takes dLoss/dImage of every view (the same layout, float) and
 * returns the gradients summed over the views, with pg3g_adjoint's conventions (it reuses the
 * per-sample radiance of a pg3g_render_diff_views of the same iterations). */
#define PG3G_MAX_DIFF_VIEWS 64
int pg3g_diff_set_views(pg3g_context *ctx, uint32_t slot, uint32_t num_views, const pg3g_camera_desc *cameras,
                        const uint32_t *scene_ids);
int pg3g_render_diff_views(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter, double *sums,
                           float *elapsed_ms);
int pg3g_adjoint_views(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter,
                       const float *dloss_dimages, double *grad_scalars, double *grad_em, float *elapsed_ms);

/* Allocation-free variant of pg3g_adjoint.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight optimisation loops,
 * harnesses that forbid hipMalloc.  It calls no hipMalloc / hipFree and does not synchronise:
 * asynchronous on the context stream (pg3g_synchronize waits for it).  Inputs and outputs are
 * DEVICE pointers: dloss_dimage_dev (res_x*res_y*3 floats, dLoss/dImage with image = sum / n_iter),
 * grad_scalars_dev (num_scalars doubles, may be null) and grad_em_dev (w*h*3 doubles, may be null).
 * The gradients accumulate (FP64) in one fixed device-side scratch array of the library, 769 KiB per
 * GPU held until the process ends, which limits the EM to PG3G_NOALLOC_MAX_EM_CHANNELS FP64
 * channels (256x128 RGB); a larger EM is rejected with PG3G_ERR_UNSUPPORTED.  Every call on one GPU
 * shares that GPU's scratch array: calls that overlap on one GPU race (serialise them, e.g. on one
 * stream).  It reuses the per-sample radiance of the last pg3g_render_diff of the same iterations
 * when there is one, else replays each path twice.  The gradients equal pg3g_adjoint's. */
#define PG3G_NOALLOC_MAX_EM_CHANNELS (256 * 128 * 3)
int pg3g_adjoint_noalloc(pg3g_context *ctx, uint32_t slot, uint32_t first_iter, uint32_t n_iter,
                         const float *dloss_dimage_dev, double *grad_scalars_dev, double *grad_em_dev);

#ifdef __cplusplus
}
#endif

#endif /* PG3G_H */
