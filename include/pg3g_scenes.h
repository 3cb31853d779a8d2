/*
 * pg3g_scenes.h -- host-side scene construction for libpg3g.
 *
 * The reference builds its scenes in Scene::LoadCornellBox (src/scene.hxx:176-956) and hands the
 * renderers a fully populated Scene.  These builders produce the same kind of flat description
 * (pg3g_scene_desc) for the benchmark workloads and for the reference's own Cornell-box scene IDs,
 * so both the GPU path and CPU checkers consume identical inputs.  Builders allocate host memory
 * owned by the returned handle.
 */
#ifndef PG3G_SCENES_H
#define PG3G_SCENES_H

#include "pg3g.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pg3g_scene_build pg3g_scene_build;

/* Workload scenes:
 *   1: unit Cornell box, 5 Lambert walls, 0.25x0.25 ceiling area light of radiance 25, a Lambert
 *      (0.6) and a GGX conductor (alpha 0.1, F0 0.95) sphere of radius 0.25; 256x256, maxpl 8.
 *   2: the box lit by a seeded 1024x512 sky-gradient + sun EM, steerable sampler (levels <= 4);
 *      second sphere a GGX dielectric (alpha 0.3, IOR 1.5); 512x512, maxpl 10.
 *   3: one of 64 seeded variants (variant = 0..63) of workload 1 with a 256x128 EM.
 *   4: workload-1 geometry lit only by a 128x64 EM, two GGX-conductor spheres with RGB tints;
 *      256x256, maxpl 8 (differentiation).
 *   5: a box holding 4 GGX-conductor spheres, 128x64 EM of 3 Gaussian lobes on a sky gradient,
 *      view `variant` of 8 (the same scene for every variant: spheres and EM depend on the seed
 *      only); 256x256, maxpl 6 (inverse rendering).
 * res_override > 0 replaces both resolution axes (tests render smaller images).
 * Render parameters are set to the workload's (pt, -sb 1, -iic 0, minpl 1). */
int pg3g_build_workload_scene(int workload, uint64_t seed, uint32_t variant, uint32_t res_override,
                              pg3g_scene_build **out);

/* The reference's Cornell-box scene IDs 0..37 (src/config.hxx:45-115, scene.hxx:176-956) with
 * environment map em_type (scene.hxx:158-170, 830-943): the constant kEMConstBluish (0) and
 * kEMConstSRGBWhite (1), and the constant white images 8x4 (2) and 1024x512 (3), which are built in
 * memory.  Image EMs use the reference's default simple spherical sampler (hard_config.hxx:17),
 * tent filtering on.  resolution 0 = 512x512 (config.hxx:568).  em_type 4..12 (photographs)
 * needs pg3g_build_reference_scene_em. */
int pg3g_build_reference_scene(uint32_t scene_id, uint32_t em_type, uint32_t resolution,
                               pg3g_scene_build **out);
/* As pg3g_build_reference_scene, with the EM of em_type 2..12 read from em_path (Radiance .hdr,
 * PFM or raw float; see pg3g_load_em) and rotated and scaled as the reference loads that type
 * (LoadEnvironmentMap arguments, scene.hxx:850-943), in place of its OpenEXR file. */
int pg3g_build_reference_scene_em(uint32_t scene_id, uint32_t em_type, uint32_t resolution, const char *em_path,
                                  pg3g_scene_build **out);
/* The short name of a reference scene ID with environment map em_type, as the reference builds
 * it for default output file names (Scene::GetSceneName's acronym, src/scene.hxx:1035-1273:
 * geometry, lights with the EM type number, materials, separated by '_').  Writes at most len
 * bytes including the terminating zero. */
int pg3g_reference_scene_acronym(uint32_t scene_id, uint32_t em_type, char *buf, size_t len);

/* Description view of a built scene (valid until the build is freed; callers may edit params and
 * the material/light arrays in place through pg3g_scene_build_mutable). */
const pg3g_scene_desc *pg3g_scene_build_desc(const pg3g_scene_build *b);
pg3g_scene_desc *pg3g_scene_build_mutable(pg3g_scene_build *b);
/* Mutable EM texels (width*height*3) of a built scene, or NULL. */
float *pg3g_scene_build_em_texels(pg3g_scene_build *b);
/* Samples per pixel the workload renders (iterations of RunIteration). */
uint32_t pg3g_scene_build_spp(const pg3g_scene_build *b);
void pg3g_scene_build_free(pg3g_scene_build *b);

#ifdef __cplusplus
}
#endif

#endif /* PG3G_SCENES_H */
