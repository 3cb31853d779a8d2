// gfx950 kernels of the path-tracing hot path and their launchers.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pg3g_internal.hpp"
#include "pt_kernels.hpp"

namespace pg3 {

// ---------------------------------------------------------------------------------------------
// Fused path-tracing kernel: one thread traces `itersPerThread` consecutive iterations of one pixel
// (PathTracerBase::RunIteration, path_tracer_base.hxx:57-90, over a range of iterations) and adds
// its FP64 sum to the framebuffer (frame_buffer.hxx:30-46).  Consecutive lanes are consecutive
// pixels of one 8x8 tile so a wavefront traces spatially coherent rays.
__global__ void __launch_bounds__(256) RenderMegakernel(SceneView S, uint32_t firstIter, uint32_t nIter,
                                                         uint32_t itersPerThread, uint32_t numChunks,
                                                         double *__restrict__ fb,
                                                         unsigned long long *__restrict__ counters) {
    const uint32_t npix = S.resX * S.resY;
    const uint32_t tilesX = (S.resX + 7u) / 8u, tilesY = (S.resY + 7u) / 8u;
    const uint32_t numTiles = tilesX * tilesY;
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = (uint32_t)(gid & 63u);
    const uint64_t waveId = gid >> 6;
    const uint32_t tile = (uint32_t)(waveId % numTiles);
    const uint32_t chunk = (uint32_t)(waveId / numTiles);
    const uint32_t x = (tile % tilesX) * 8u + (lane & 7u);
    const uint32_t y = (tile / tilesX) * 8u + (lane >> 3);
    RayTally tally;
    if (chunk < numChunks && x < S.resX && y < S.resY) {
        const uint32_t pixel = y * S.resX + x;
        const uint32_t itBegin = firstIter + chunk * itersPerThread;
        const uint32_t itEnd = min(firstIter + nIter, itBegin + itersPerThread);
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        for (uint32_t it = itBegin; it < itEnd; ++it) {
            const F3 L = RadianceSample(S, pixel, it, tally);
            acc0 += L.x;
            acc1 += L.y;
            acc2 += L.z;
            ++tally.samples;
        }
        if (itEnd > itBegin) {
            atomicAdd(fb + (size_t)pixel * 3 + 0, acc0);
            atomicAdd(fb + (size_t)pixel * 3 + 1, acc1);
            atomicAdd(fb + (size_t)pixel * 3 + 2, acc2);
        }
    }
    (void)npix;
    FlushTally(tally, counters);
}

// Path tracer (-sb 1) specialised on the EM sampler and the light count, with path
// regeneration: a lane whose path ends starts its pixel's next iteration at once instead of
// idling until the longest path of its wavefront ends.  The samples, their RNG streams and the
// per-pixel sums are those of RenderMegakernel; only the order of the vertex loop changes.
//   kEmKind: kEmAbsent = no EM image (constant or no background), else the sampler of the EM.
//   kOneLight: the scene has exactly one light (light picking is trivial, no contribution cache).
// Register budget of the path-tracing kernel: 4 waves per SIMD (128 VGPRs) measured fastest on
// the box + steerable-EM workload (1, 3, 4 and 5 waves tried); the excess spills to scratch.
#ifndef PG3_PT_WAVES_PER_EU
#define PG3_PT_WAVES_PER_EU 4
#endif
// Persistent: each CU holds 4 x PTWavesPerEU wavefronts for the whole launch, and every wavefront takes work
// items (8x8 tile x iteration chunk) from a device counter until none is left (TraceTasks: a lane
// whose task ends adds its radiance to the framebuffer and takes the next task at once).  With the
// steerable sampler one workgroup per CU holds all its resident wavefronts (1024 lanes at 4 waves
// per SIMD) and stages the inner levels of the EM's triangle tree in LDS once (kTreeLdsNodes
// nodes, 60 KiB): the tree descent of every light sample then gathers from LDS instead of from the
// L2.  With the simple sampler the marginal CDF
// is staged instead (up to kMargLdsMax entries, an EM of up to 4095 rows).
// The split path keeps more state per lane.  Its generic kernel (light samples in a loop with their
// shadow rays) runs at 3 waves per SIMD (168 VGPRs; the box + area light frame at the reference's
// defaults: 2.6 ms against 2.85 at 4 waves).  The kernel for a steerable EM as the only light
// (SplitPath<true>: shared EM lookups, the shadow ray after the BSDF sample) needs 143 VGPRs and
// runs faster at 4 waves per SIMD (128 VGPRs) despite the spill: the box + steerable EM frame
// at the reference's defaults takes 45.5 against 56.0 ms.
#ifndef PG3_SPLIT_WAVES_PER_EU
#define PG3_SPLIT_WAVES_PER_EU 3
#endif
#ifndef PG3_SPLIT_EM_WAVES_PER_EU
#define PG3_SPLIT_EM_WAVES_PER_EU 4
#endif
// Waves per SIMD of the chain path, of the generic split path and of the split path with shared
// EM lookups (kMerge)
template <bool kSplit, bool kMerge>
constexpr uint32_t PTWavesPerEU() {
    return !kSplit ? PG3_PT_WAVES_PER_EU : (kMerge ? PG3_SPLIT_EM_WAVES_PER_EU : PG3_SPLIT_WAVES_PER_EU);
}
template <bool kSplit, bool kMerge>
constexpr uint32_t PTWavesPerCU() { return 4u * PTWavesPerEU<kSplit, kMerge>(); }
template <int kEmKind, bool kOneLight, bool kSplit>
constexpr uint32_t PTBlock() {
    return kEmKind == kEmSteerable ? 64u * PTWavesPerCU<kSplit, kOneLight && kEmKind == kEmSteerable>() : 256u;
}
static_assert(64u * PTWavesPerCU<false, false>() <= 1024u && 64u * PTWavesPerCU<false, true>() <= 1024u &&
                  64u * PTWavesPerCU<true, false>() <= 1024u && 64u * PTWavesPerCU<true, true>() <= 1024u,
              "one steerable workgroup per CU holds every resident wavefront");

// kSplit: the path tracer with splitting at the first vertex and indirect clipping (SplitPath),
// else the -sb 1 -iic 0 chain (ChainPath).
template <int kEmKind, bool kOneLight, bool kSplit>
__global__ void __launch_bounds__((PTBlock<kEmKind, kOneLight, kSplit>()))
__attribute__((amdgpu_waves_per_eu(PTWavesPerEU<kSplit, kOneLight && kEmKind == kEmSteerable>())))
RenderPTKernel(const SceneView *__restrict__ Sp, uint32_t firstIter, uint32_t nIter, uint32_t itersPerThread,
               uint32_t numChunks, double *__restrict__ fb, uint32_t *__restrict__ work,
               unsigned long long *__restrict__ counters) {
    // The scene is read from a device copy rather than from kernel arguments: held in SGPRs,
    // its ~450 bytes of fields spill into VGPR lanes of this register-heavy kernel.
    const SceneView &S = *Sp;
    LightSpec spec{kEmKind, kOneLight ? 1u : S.numLights};
    if constexpr (kEmKind == kEmSteerable) {
        __shared__ F4v treeLds[kTreeLdsNodes * 3];
        spec.tree = StageTree(S.em, treeLds);
        __syncthreads();
    }
    if constexpr (kEmKind == kEmSimple) {
        __shared__ float margLds[kMargLdsMax];
        StageMarginal(S.em, margLds, spec.tree);
        __syncthreads();
    }
    const uint32_t tilesX = (S.resX + 7u) / 8u, numTiles = tilesX * ((S.resY + 7u) / 8u);
    RayTally tally;
    SceneItems<kOneLight && kEmKind == kEmSteerable> src{Sp,        spec,     work,     fb,      firstIter,
                                                         nIter,     itersPerThread, numTiles * numChunks, numTiles,
                                                         tilesX,    0u,       0u};
    src.SetTail(numChunks);
    if constexpr (kSplit) TraceTasks<SplitPath<kOneLight && kEmKind == kEmSteerable>, false>(src, tally);
    else TraceTasks<ChainPath<kOneLight && kEmKind == kEmSteerable, kOneLight && kEmKind == kEmSteerable>, false>(src, tally);
    FlushTally(tally, counters);
}

// Path tracer with splitting beyond the first vertex: one thread per (pixel, iteration chunk),
// each sample traced by RadiancePTTree (its vertex stack and chain records live in scratch).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) RenderTreeKernel(SceneView S, uint32_t firstIter, uint32_t nIter,
                                                         uint32_t itersPerThread, uint32_t numChunks,
                                                         double *__restrict__ fb,
                                                         unsigned long long *__restrict__ counters) {
    const LightSpec spec = SceneLightSpec(S);
    const uint32_t tilesX = (S.resX + 7u) / 8u, tilesY = (S.resY + 7u) / 8u;
    const uint32_t numTiles = tilesX * tilesY;
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = (uint32_t)(gid & 63u);
    const uint64_t waveId = gid >> 6;
    const uint32_t tile = (uint32_t)(waveId % numTiles);
    const uint32_t chunk = (uint32_t)(waveId / numTiles);
    const uint32_t x = (tile % tilesX) * 8u + (lane & 7u);
    const uint32_t y = (tile / tilesX) * 8u + (lane >> 3);
    RayTally tally;
    if (chunk < numChunks && x < S.resX && y < S.resY) {
        const uint32_t pixel = y * S.resX + x;
        const uint32_t itBegin = firstIter + chunk * itersPerThread;
        const uint32_t itEnd = min(firstIter + nIter, itBegin + itersPerThread);
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        for (uint32_t it = itBegin; it < itEnd; ++it) {
            RngKey key;
            F3 org, dir;
            CameraSample(S, pixel, it, key, org, dir);
            const F3 L = RadiancePTTree(S, spec, key, org, dir, tally);
            acc0 += L.x;
            acc1 += L.y;
            acc2 += L.z;
            ++tally.samples;
        }
        if (itEnd > itBegin) {
            atomicAdd(fb + (size_t)pixel * 3 + 0, acc0);
            atomicAdd(fb + (size_t)pixel * 3 + 1, acc1);
            atomicAdd(fb + (size_t)pixel * 3 + 2, acc2);
        }
    }
    (void)lane;
    FlushTally(tally, counters);
}

// ---------------------------------------------------------------------------------------------
// Function-level kernels (tier-1 parity of BSDF and EM functions)
__global__ void BsdfEvalKernel(DevMaterial m, uint32_t n, const float *__restrict__ wil, const float *__restrict__ wol,
                               float *__restrict__ f, float *__restrict__ pdf, float *__restrict__ comp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    BsdfRec r;
    EvalBsdf(m, mk(wil[i * 3], wil[i * 3 + 1], wil[i * 3 + 2]), mk(wol[i * 3], wol[i * 3 + 1], wol[i * 3 + 2]), r);
    f[i * 3] = r.att.x; f[i * 3 + 1] = r.att.y; f[i * 3 + 2] = r.att.z;
    pdf[i] = r.pdfW; comp[i] = r.compProb;
}
__global__ void BsdfSampleKernel(DevMaterial m, uint32_t n, const float *__restrict__ wol, const float *__restrict__ u,
                                 float *__restrict__ wil, float *__restrict__ f, float *__restrict__ pdf,
                                 float *__restrict__ comp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    BsdfRec r;
    SampleBsdf(m, mk(wol[i * 3], wol[i * 3 + 1], wol[i * 3 + 2]), u[i * 3], u[i * 3 + 1], u[i * 3 + 2], r);
    wil[i * 3] = r.wil.x; wil[i * 3 + 1] = r.wil.y; wil[i * 3 + 2] = r.wil.z;
    f[i * 3] = r.att.x; f[i * 3 + 1] = r.att.y; f[i * 3 + 2] = r.att.z;
    pdf[i] = r.pdfW; comp[i] = r.compProb;
}
__global__ void EmSampleKernel(SceneView S, uint32_t n, const float *__restrict__ nrm, const int32_t *__restrict__ sides,
                               const float *__restrict__ u, float *__restrict__ dir, float *__restrict__ pdf,
                               float *__restrict__ rc, int32_t *__restrict__ ok) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Frame fr = FrameFromZ(mk(nrm[i * 3], nrm[i * 3 + 1], nrm[i * 3 + 2]));
    F3 d = mk(0.f), radCos = mk(0.f);
    float p = 0.f;
    const bool good = EmSample(S.em, S.em.sampler, fr, (uint32_t)sides[i], u[i * 3], u[i * 3 + 1], u[i * 3 + 2], d, p, radCos);
    ok[i] = good ? 1 : 0;
    dir[i * 3] = d.x; dir[i * 3 + 1] = d.y; dir[i * 3 + 2] = d.z;
    pdf[i] = p;
    rc[i * 3] = radCos.x; rc[i * 3 + 1] = radCos.y; rc[i * 3 + 2] = radCos.z;
}
__global__ void EmPdfKernel(SceneView S, uint32_t n, const float *__restrict__ dirs, const float *__restrict__ nrm,
                            const int32_t *__restrict__ sides, float *__restrict__ pdf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Frame fr = FrameFromZ(mk(nrm[i * 3], nrm[i * 3 + 1], nrm[i * 3 + 2]));
    pdf[i] = EmPdf(S.em, S.em.sampler, mk(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]), fr, (uint32_t)sides[i]);
}
// Scene::Intersect / Occluded of n rays, one thread per ray (function-level parity).
__global__ void IntersectKernel(SceneView S, uint32_t n, const float *__restrict__ org, const float *__restrict__ dir,
                                const float *__restrict__ tmin, float *__restrict__ dist, float *__restrict__ nrm,
                                int32_t *__restrict__ mat) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Hit h;
    h.dist = 1e36f;
    h.mat = -1;
    h.n = mk(0.f);
    const bool hit = Intersect(S, mk(org[i * 3], org[i * 3 + 1], org[i * 3 + 2]),
                               mk(dir[i * 3], dir[i * 3 + 1], dir[i * 3 + 2]), tmin[i], h);
    dist[i] = hit ? h.dist : -1.f;
    mat[i] = hit ? h.mat : -1;
    nrm[i * 3] = hit ? h.n.x : 0.f;
    nrm[i * 3 + 1] = hit ? h.n.y : 0.f;
    nrm[i * 3 + 2] = hit ? h.n.z : 0.f;
}
__global__ void OccludedKernel(SceneView S, uint32_t n, const float *__restrict__ p, const float *__restrict__ dir,
                               const float *__restrict__ tmax, int32_t *__restrict__ occ) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    occ[i] = Occluded(S, mk(p[i * 3], p[i * 3 + 1], p[i * 3 + 2]), mk(dir[i * 3], dir[i * 3 + 1], dir[i * 3 + 2]),
                      tmax[i]) ? 1 : 0;
}
__global__ void EmEvalKernel(SceneView S, uint32_t n, const float *__restrict__ dirs, float *__restrict__ rgb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const F3 v = EmEval(S.em, mk(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]));
    rgb[i * 3] = v.x; rgb[i * 3 + 1] = v.y; rgb[i * 3 + 2] = v.z;
}

// ---------------------------------------------------------------------------------------------
// EM luminance CDF build (em_simple_spherical_sampler.hxx:117-152, distribution.hxx:45-74,
// 484-506): f(x,y) = sin(theta_mid(y)) * AveragePixelLuminance(x,y), per-row FP64 prefix sums of
// f/w, normalised rows stored as FP32; the marginal is built from the FP32 row integrals.
// One 256-thread workgroup per row; the row is scanned in 256-element FP64 tiles.
__device__ __forceinline__ float AvgPixelLum(const DevEm &E, uint32_t x, uint32_t y) {   // em_image.hxx:152-183
    if (!E.bilinear) return Luminance(Texel(E, x, y));
    uint32_t x0, y0, x2, y2;
    EmNorm(E, (int32_t)x - 1, (int32_t)y - 1, x0, y0);
    EmNorm(E, (int32_t)x + 1, (int32_t)y + 1, x2, y2);
    const float e = 1 / 24.f, c = 2 / 3.f;
    float s = Luminance(Texel(E, x0, y0)) * e;
    s = s + Luminance(Texel(E, x0, y)) * e;
    s = s + Luminance(Texel(E, x0, y2)) * e;
    s = s + Luminance(Texel(E, x, y0)) * e;
    s = s + Luminance(Texel(E, x, y)) * c;
    s = s + Luminance(Texel(E, x, y2)) * e;
    s = s + Luminance(Texel(E, x2, y0)) * e;
    s = s + Luminance(Texel(E, x2, y)) * e;
    s = s + Luminance(Texel(E, x2, y2)) * e;
    return s;
}

__device__ __forceinline__ double BlockScanInclusive(double v, double *lds) {
    // Hillis-Steele over 64 lanes with shuffles, then across the 4 waves through LDS.
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double t = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += t;
    }
    if (lane == 63) lds[wave] = v;
    __syncthreads();
    double off = 0.0;
    for (uint32_t w = 0; w < wave; ++w) off += lds[w];
    __syncthreads();
    return v + off;
}

__global__ void __launch_bounds__(256) CdfRowsKernel(DevEm E, float *__restrict__ condCdf, float *__restrict__ rowIntegral) {
    __shared__ double lds[4];
    __shared__ double carry;
    const uint32_t row = blockIdx.x;
    const uint32_t w = E.w;
    float *cdf = condCdf + (size_t)row * (w + 1);
    const float sinMid = SinMidTheta(E.h, row);
    if (threadIdx.x == 0) carry = 0.0;
    __syncthreads();
    // pass 1: total (FP64)
    for (uint32_t base = 0; base < w; base += 256) {
        const uint32_t x = base + threadIdx.x;
        const double v = (x < w) ? (double)(sinMid * AvgPixelLum(E, x, row)) / (double)w : 0.0;
        const double incl = BlockScanInclusive(v, lds);
        const double c = carry;
        __syncthreads();
        if (threadIdx.x == 255) carry = c + incl;
        __syncthreads();
    }
    const double total = carry;
    __syncthreads();
    if (threadIdx.x == 0) carry = 0.0;
    __syncthreads();
    // pass 2: normalised prefix sums
    for (uint32_t base = 0; base < w; base += 256) {
        const uint32_t x = base + threadIdx.x;
        const double v = (x < w) ? (double)(sinMid * AvgPixelLum(E, x, row)) / (double)w : 0.0;
        const double incl = BlockScanInclusive(v, lds);
        const double c = carry;
        if (x < w) cdf[x + 1] = (total == 0.0) ? (float)(x + 1) / w : (float)((c + incl) / total);
        __syncthreads();
        if (threadIdx.x == 255) carry = c + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cdf[0] = 0.f;
        rowIntegral[row] = (float)total;
    }
}
__global__ void __launch_bounds__(256) CdfMarginalKernel(uint32_t h, const float *__restrict__ rowIntegral,
                                                         float *__restrict__ margCdf) {
    __shared__ double lds[4];
    __shared__ double carry;
    if (threadIdx.x == 0) carry = 0.0;
    __syncthreads();
    for (uint32_t base = 0; base < h; base += 256) {
        const uint32_t y = base + threadIdx.x;
        const double v = (y < h) ? (double)rowIntegral[y] / (double)h : 0.0;
        const double incl = BlockScanInclusive(v, lds);
        const double c = carry;
        __syncthreads();
        if (threadIdx.x == 255) carry = c + incl;
        __syncthreads();
    }
    const double total = carry;
    __syncthreads();
    if (threadIdx.x == 0) carry = 0.0;
    __syncthreads();
    for (uint32_t base = 0; base < h; base += 256) {
        const uint32_t y = base + threadIdx.x;
        const double v = (y < h) ? (double)rowIntegral[y] / (double)h : 0.0;
        const double incl = BlockScanInclusive(v, lds);
        const double c = carry;
        if (y < h) margCdf[y + 1] = (total == 0.0) ? (float)(y + 1) / h : (float)((c + incl) / total);
        __syncthreads();
        if (threadIdx.x == 255) carry = c + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) margCdf[0] = 0.f;
}

// ---------------------------------------------------------------------------------------------
// Launchers (host)
using PTKernelFn = void (*)(const SceneView *, uint32_t, uint32_t, uint32_t, uint32_t, double *, uint32_t *,
                           unsigned long long *);

__global__ void ResetWorkCounter(uint32_t *work) {
    if (threadIdx.x == 0) work[0] = 0u;
}

// Whether every lane of the split path takes at most one light sample per step: nL <= nB light
// samples at the first vertex (ComputeSplittingCounts, path_tracer.hxx:440-450) and one at the others.
static bool OneLightSamplePerStep(const SceneView &S) {
    const float nB = (float)lroundf((1.f - S.splitLevel) * 1.f + S.splitLevel * S.splitBudget);
    const long nL0 = lroundf(std::max(S.lightRatio * nB, 1.f)), nL1 = lroundf(std::max(S.lightRatio, 1.f));
    return nL0 <= (long)nB && nL1 == 1;
}

// Specialised path-tracer kernel for a scene (`pt` only), its workgroup size and its resident
// wavefronts per CU.  The split path's one-light steerable kernel (SplitPath<true>) takes one light
// sample per lane and step; other light-sample counts run the generic split kernel.
struct PTLaunch {
    PTKernelFn fn;
    uint32_t block, wavesPerCU;
};
template <int kEmKind, bool kOneLight, bool kSplit>
static PTLaunch PTLaunchOf() {
    return {RenderPTKernel<kEmKind, kOneLight, kSplit>, PTBlock<kEmKind, kOneLight, kSplit>(),
            PTWavesPerCU<kSplit, kOneLight && kEmKind == kEmSteerable>()};
}
static PTLaunch PickPTKernel(const SceneView &S, bool split) {
    const int em = SceneLightSpec(S).em;
    const bool one = (S.numLights == 1) && !(split && em == kEmSteerable && !OneLightSamplePerStep(S));
#define PG3_PT_CASE(K)                                                                                    \
    if (em == (K)) {                                                                                      \
        if (split) return one ? PTLaunchOf<(K), true, true>() : PTLaunchOf<(K), false, true>();           \
        return one ? PTLaunchOf<(K), true, false>() : PTLaunchOf<(K), false, false>();                    \
    }
    PG3_PT_CASE(kEmAbsent)
    PG3_PT_CASE(kEmCosine)
    PG3_PT_CASE(kEmSimple)
    PG3_PT_CASE(kEmSteerable)
#undef PG3_PT_CASE
    return {nullptr, 0u, 0u};
}

hipError_t LaunchMegakernel(const SceneView &S, const SceneView *Sdev, uint32_t firstIter, uint32_t nIter, double *fb,
                            unsigned long long *counters, uint32_t *work, int numCUs, hipStream_t stream) {
    if (nIter == 0) return hipSuccess;
    // work-item and thread indices are 32-bit: a range of more than 2^30 (tile, iteration)
    // pairs is rendered as consecutive launches (iterations add independently)
    const uint64_t tiles64 = (uint64_t)((S.resX + 7u) / 8u) * ((S.resY + 7u) / 8u);
    const uint32_t maxIter = (uint32_t)std::max<uint64_t>(1u, (1ull << 30) / tiles64);
    if (nIter > maxIter) {
        for (uint32_t f = 0; f < nIter; f += maxIter) {
            const hipError_t e = LaunchMegakernel(S, Sdev, firstIter + f, std::min(maxIter, nIter - f), fb, counters, work,
                                                  numCUs, stream);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const uint32_t npix = S.resX * S.resY;
    // Enough lanes to fill 256 CUs several times over; each thread keeps >= 1 iteration.
    const uint64_t targetThreads = 1u << 21;
    uint32_t numChunks = (uint32_t)((targetThreads + npix - 1) / npix);
    numChunks = max(1u, min(numChunks, nIter));
    const uint32_t ipt = (nIter + numChunks - 1) / numChunks;
    numChunks = (nIter + ipt - 1) / ipt;
    const uint32_t tiles = ((S.resX + 7u) / 8u) * ((S.resY + 7u) / 8u);
    const uint64_t threads = (uint64_t)tiles * 64u * numChunks;
    const uint32_t blocks = (uint32_t)((threads + 255) / 256);
    const bool chain = (S.algorithm == 7) && S.splitBudget == 1.f && !(S.iic > 0.f);   // -sb 1 -iic 0
    // ComputeSplittingCounts at the first vertex: when it leaves budget 1 to the next vertex,
    // every later vertex has one BSDF sample (SplitPath); other budgets take the general tree
    const uint32_t nB0 = (uint32_t)lroundf((1.f - S.splitLevel) * 1.f + S.splitLevel * S.splitBudget);
    const bool rootSplit = (S.algorithm == 7) && !chain && nB0 >= 1u && (1 + ((S.splitBudget - nB0) / (float)nB0)) == 1.f;
    const PTLaunch pl = (chain || rootSplit) ? PickPTKernel(S, rootSplit) : PTLaunch{nullptr, 0u, 0u};
    const PTKernelFn pt = pl.fn;
    const uint32_t ptBlock = pl.block, wavesPerCU = pl.wavesPerCU;
    if (pt) {
        if (!work) return hipErrorInvalidValue;
        // items of a few iterations: the wave overlaps items, so small items cost no tail, and
        // they even out the waves' shares at the end of the launch (the box + steerable EM frame,
        // 64 spp, with items of 1, 2, 3, 4, 8 and 16 iterations: 11.94, 11.84, 11.93, 11.78, 11.75
        // and 11.80 ms; the last chunk is split into single iterations in any case)
        const uint32_t waves = (uint32_t)max(1, numCUs) * wavesPerCU;
        const uint64_t perWave = (uint64_t)nIter * tiles / (16u * waves);
        uint32_t ipi = perWave >= 4u ? 4u : (perWave >= 2u ? 2u : 1u);
        if (const char *e = getenv("PG3G_PT_ITEM_ITERS")) ipi = (uint32_t)max(1, atoi(e));
        const uint32_t ptIpi = min(ipi, nIter), ptChunks = (nIter + ptIpi - 1) / ptIpi;
        // resident wavefronts per CU: the kernel's register budget per SIMD x 4 SIMDs
        const uint32_t grid = (uint32_t)max(1, numCUs) * (64u * wavesPerCU / ptBlock);
        hipLaunchKernelGGL(ResetWorkCounter, dim3(1), dim3(64), 0, stream, work);
        hipLaunchKernelGGL(pt, dim3(grid), dim3(ptBlock), 0, stream, Sdev, firstIter, nIter, ptIpi, ptChunks, fb, work,
                           counters);
    }
    else if (S.algorithm == 7)
        hipLaunchKernelGGL(RenderTreeKernel, dim3(blocks), dim3(256), 0, stream, S, firstIter, nIter, ipt, numChunks, fb,
                           counters);
    else hipLaunchKernelGGL(RenderMegakernel, dim3(blocks), dim3(256), 0, stream, S, firstIter, nIter, ipt, numChunks, fb,
                            counters);
    return hipGetLastError();
}

hipError_t LaunchBsdfEval(const DevMaterial &m, uint32_t n, const float *wil, const float *wol, float *f, float *pdf,
                          float *comp, hipStream_t s) {
    hipLaunchKernelGGL(BsdfEvalKernel, dim3((n + 255) / 256), dim3(256), 0, s, m, n, wil, wol, f, pdf, comp);
    return hipGetLastError();
}
hipError_t LaunchBsdfSample(const DevMaterial &m, uint32_t n, const float *wol, const float *u, float *wil, float *f,
                            float *pdf, float *comp, hipStream_t s) {
    hipLaunchKernelGGL(BsdfSampleKernel, dim3((n + 255) / 256), dim3(256), 0, s, m, n, wol, u, wil, f, pdf, comp);
    return hipGetLastError();
}
hipError_t LaunchEmSample(const SceneView &S, uint32_t n, const float *nrm, const int32_t *sides, const float *u,
                          float *dir, float *pdf, float *rc, int32_t *ok, hipStream_t s) {
    hipLaunchKernelGGL(EmSampleKernel, dim3((n + 255) / 256), dim3(256), 0, s, S, n, nrm, sides, u, dir, pdf, rc, ok);
    return hipGetLastError();
}
hipError_t LaunchEmPdf(const SceneView &S, uint32_t n, const float *dirs, const float *nrm, const int32_t *sides,
                       float *pdf, hipStream_t s) {
    hipLaunchKernelGGL(EmPdfKernel, dim3((n + 255) / 256), dim3(256), 0, s, S, n, dirs, nrm, sides, pdf);
    return hipGetLastError();
}
hipError_t LaunchIntersect(const SceneView &S, uint32_t n, const float *org, const float *dir, const float *tmin,
                           float *dist, float *nrm, int32_t *mat, hipStream_t s) {
    hipLaunchKernelGGL(IntersectKernel, dim3((n + 255) / 256), dim3(256), 0, s, S, n, org, dir, tmin, dist, nrm, mat);
    return hipGetLastError();
}
hipError_t LaunchOccluded(const SceneView &S, uint32_t n, const float *p, const float *dir, const float *tmax,
                          int32_t *occ, hipStream_t s) {
    hipLaunchKernelGGL(OccludedKernel, dim3((n + 255) / 256), dim3(256), 0, s, S, n, p, dir, tmax, occ);
    return hipGetLastError();
}
hipError_t LaunchEmEval(const SceneView &S, uint32_t n, const float *dirs, float *rgb, hipStream_t s) {
    hipLaunchKernelGGL(EmEvalKernel, dim3((n + 255) / 256), dim3(256), 0, s, S, n, dirs, rgb);
    return hipGetLastError();
}
__global__ void EmLumKernel(const float4 *__restrict__ tex, float *__restrict__ lum, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float4 t = tex[i];
        lum[i] = Luminance(mk(t.x, t.y, t.z));
    }
}
hipError_t LaunchEmLum(const float4 *tex, float *lum, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(EmLumKernel, dim3((n + 255) / 256), dim3(256), 0, s, tex, lum, n);
    return hipGetLastError();
}
hipError_t LaunchCdfBuild(const DevEm &E, float *condCdf, float *margCdf, float *rowIntegral, hipStream_t s) {
    hipLaunchKernelGGL(CdfRowsKernel, dim3(E.h), dim3(256), 0, s, E, condCdf, rowIntegral);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(CdfMarginalKernel, dim3(1), dim3(256), 0, s, E.h, rowIntegral, margCdf);
    return hipGetLastError();
}

}  // namespace pg3
