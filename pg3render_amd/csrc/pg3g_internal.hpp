// Internal declarations shared by the host side of libpg3g (C ABI, table builders) and the
// kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "dev_scene.hpp"
#include "image_io.hpp"

namespace pg3 {

// S: host copy of the scene view (launch geometry, generic kernel argument); Sdev: the same view
// in device memory (read by the specialised path-tracer kernels).
// work: one device word for the persistent kernels' item counter (reset by the launch).
hipError_t LaunchMegakernel(const SceneView &S, const SceneView *Sdev, uint32_t firstIter, uint32_t nIter, double *fb,
                            unsigned long long *counters, uint32_t *work, int numCUs, hipStream_t stream);
struct WavefrontContext;   // queue storage of the wavefront path tracer (one per pg3g_context)
WavefrontContext *CreateWavefrontContext();
void DestroyWavefrontContext(WavefrontContext *w);
hipError_t LaunchWavefront(const SceneView &S, const SceneView *Sdev, uint32_t firstIter, uint32_t nIter, double *fb,
                           unsigned long long *counters, uint32_t *work, int numCUs, WavefrontContext *queues,
                           hipStream_t stream);
// Batched multi-scene pass (batch.hip).  Scene s owns global work items
// [itemBegin, itemBegin + numTiles * ceil(spp / itersPerItem)); item = chunk * numTiles + tile.
struct BatchScene {
    double *fb;                 // the scene's FP64 framebuffer
    uint32_t itemBegin;
    uint32_t tilesX, numTiles;  // 8x8 pixel tiles
    uint32_t spp;               // iterations per pass
    uint32_t itersPerItem;      // iterations of one tile per work item
    uint32_t pad;
};
constexpr uint32_t kBatchParts = 8;   // MI355X: 8 XCDs, each with its own L2
struct BatchLaunch {
    const SceneView *views;     // device, numScenes
    const BatchScene *info;     // device, numScenes
    uint32_t numScenes, numItems;
    uint32_t *work;             // device: [p] next item of part p, [kBatchParts] pass index,
                                // [kBatchParts + 1 + p] first item of part p (kBatchParts + 1 entries)
    unsigned long long *counters;
    uint32_t gridBlocks;        // persistent workgroups
    uint32_t blockLanes;        // lanes per workgroup (multiple of 64, at most 1024)
    int emKind;                 // EM sampler shared by every scene, or ignored when oneLight < 0
    int oneLight;               // 1 / 0 when shared by every scene, -1 = read per scene
};
hipError_t LaunchBatchPass(const BatchLaunch &L, hipStream_t stream);

// Differentiable render (diff.hip).  adjImg == nullptr: forward with frozen sampling into fb.
// Otherwise the adjoint of iterations [firstIter, firstIter + nIter): adjImg = dLoss/dImage
// (res_x*res_y*3, image = sum / nIter), gradients added to gradEm (w*h*3, may be null when the
// scene has no EM) and gradScal (numScal <= kMaxDiffScalars).  sampleL (may be null): per-sample
// radiance, nIter * res_x * res_y * 3 floats, written by the forward and read by the adjoint in
// place of its primal replay.  The adjoint is a persistent kernel (numCUs workgroups) taking work
// items from work[0 .. numViews), which must be zero at launch.  views (device, numViews cameras; null with
// numViews = 1: the scene's own camera): every view renders the same iterations; fb, adjImg and
// sampleL then hold numViews images one after another (sampleL: iteration-major, numViews x
// res_x x res_y samples per iteration), and the gradients are the sum over the views.
constexpr uint32_t kMaxDiffScalars = 64;
hipError_t LaunchDiff(const SceneView &S, const SceneView *Sdev, uint32_t firstIter, uint32_t nIter, double *fb,
                      const float *adjImg, double *gradEm, double *gradScal, uint32_t numScal, float *sampleL,
                      uint32_t *work, int numCUs, unsigned long long *counters, hipStream_t stream,
                      const DevView *views = nullptr, uint32_t numViews = 1);

// Allocation-free adjoint (noalloc.hip).  kNoAllocMaxEmChannels: FP64 EM-gradient channels the
// fixed scratch array holds (an RGB EM of up to 256 x 128 texels).  InitNoAllocScratch looks up
// the scratch array's address on `device` (the current device) once; LaunchAdjointNoAlloc clears
// the scratch, runs the adjoint of LaunchDiff into it and copies the gradients to outScal (numScal
// doubles, zeroMask bits set to 0) and outEm (S.em.w * S.em.h * 3 doubles); either may be null.
// sampleL (may be null): per-sample radiance of a differentiable forward of the same iterations.
constexpr uint32_t kNoAllocMaxEmChannels = 256u * 128u * 3u;
hipError_t InitNoAllocScratch(int device);
hipError_t LaunchAdjointNoAlloc(int device, const SceneView &S, const SceneView *Sdev, uint32_t firstIter,
                                uint32_t nIter, const float *adjImg, uint32_t numScal, uint64_t zeroMask,
                                const float *sampleL, double *outScal, double *outEm, int numCUs,
                                unsigned long long *counters, hipStream_t stream);

hipError_t LaunchBsdfEval(const DevMaterial &m, uint32_t n, const float *wil, const float *wol, float *f, float *pdf,
                          float *comp, hipStream_t s);
hipError_t LaunchBsdfSample(const DevMaterial &m, uint32_t n, const float *wol, const float *u, float *wil, float *f,
                            float *pdf, float *comp, hipStream_t s);
hipError_t LaunchEmSample(const SceneView &S, uint32_t n, const float *nrm, const int32_t *sides, const float *u,
                          float *dir, float *pdf, float *rc, int32_t *ok, hipStream_t s);
hipError_t LaunchEmPdf(const SceneView &S, uint32_t n, const float *dirs, const float *nrm, const int32_t *sides,
                       float *pdf, hipStream_t s);
hipError_t LaunchEmEval(const SceneView &S, uint32_t n, const float *dirs, float *rgb, hipStream_t s);
hipError_t LaunchIntersect(const SceneView &S, uint32_t n, const float *org, const float *dir, const float *tmin,
                           float *dist, float *nrm, int32_t *mat, hipStream_t s);
hipError_t LaunchOccluded(const SceneView &S, uint32_t n, const float *p, const float *dir, const float *tmax,
                          int32_t *occ, hipStream_t s);
// Texel luminances of an EM (the steerable pdf's luminance map).
hipError_t LaunchEmLum(const float4 *tex, float *lum, uint32_t n, hipStream_t s);
hipError_t LaunchCdfBuild(const DevEm &E, float *condCdf, float *margCdf, float *rowIntegral, hipStream_t s);

// Host-built steerable sampler tables (em_steerable_sampler.hxx build path), flattened:
struct SteerableTables {
    uint32_t numLeaves = 0, numVertices = 0, numLevels = 0;
    std::vector<uint32_t> levelCount, levelBase;   // levels[0] = leaves, last = root
    std::vector<float> nodeW;                      // 12 floats per node, all levels
    std::vector<float> leafVerts;                  // 36 floats per leaf (3 x (dir, SH weight))
    float rootW[9];
};
// Builds the tables from an RGB image (w*h*3 floats) with tent or box filtering on `threads`
// host threads (0: SteerableBuildThreads()); the tables do not depend on the thread count.
bool BuildSteerable(const float *texels, uint32_t w, uint32_t h, bool bilinear, float maxErr, uint32_t minLevel,
                    uint32_t maxLevel, SteerableTables &out, std::string &err, uint32_t threads = 0);
// PG3G_BUILD_THREADS, else OMP_NUM_THREADS, else the hardware threads (at most 64).
uint32_t SteerableBuildThreads();
// BuildSteerable through the on-disk cache (steerable_cache.cpp) when a cache directory is set
// (SetSteerableCacheDir, else PG3G_STEERABLE_CACHE_DIR); *fromCache tells whether it was loaded.
bool BuildSteerableCached(const float *texels, uint32_t w, uint32_t h, bool bilinear, float maxErr, uint32_t minLevel,
                          uint32_t maxLevel, SteerableTables &out, std::string &err, uint32_t threads = 0,
                          bool *fromCache = nullptr);
void SetSteerableCacheDir(const char *dir);   // null or "": no cache


}  // namespace pg3
