// Differentiable forward render (frozen sampling) and its path-replay adjoint (diff_kernels.hpp).
#include <hip/hip_runtime.h>

#include "diff_kernels.hpp"
#include "pg3g_internal.hpp"

namespace pg3 {

// Differentiable forward: same work decomposition as the plain path-tracing kernel.
// The forward runs at 3 waves per SIMD (168 VGPRs, 40 B of spill): at 4 (128 VGPRs, 208 B of
// spill) the 256^2 x 32 spp forward of the box with two GGX conductors took 1.71-1.75 ms against
// 1.44-1.51 ms.
#ifndef PG3_DIFF_WAVES_PER_EU
#define PG3_DIFF_WAVES_PER_EU 3
#endif
template <int kEmKind, bool kOneLight>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PG3_DIFF_WAVES_PER_EU)))
DiffKernel(const SceneView *__restrict__ Sp, uint32_t firstIter, uint32_t nIter, uint32_t itersPerThread,
           uint32_t numChunks, double *__restrict__ fb, float *__restrict__ sampleL,
           unsigned long long *__restrict__ counters, const DevView *__restrict__ views, uint32_t numViews) {
    const SceneView &S = *Sp;
    const LightSpec spec = (kEmKind == kEmDynamic) ? SceneLightSpec(S)
                                                   : LightSpec{kEmKind, kOneLight ? 1u : S.numLights};
    __shared__ double accLds[4][64 * 3];
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    const uint32_t tilesX = (S.resX + 7u) / 8u, tilesY = (S.resY + 7u) / 8u;
    const uint32_t numTiles = tilesX * tilesY;
    const uint32_t waveId = (uint32_t)(((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    // wavefront -> (view, chunk, tile); one view unless the scene has several cameras
    const uint32_t perView = numTiles * numChunks;
    const uint32_t view = waveId / perView, local = waveId % perView;
    const uint32_t tile = local % numTiles, chunk = local / numTiles;
    uint32_t segments = 0, samples = 0;
    const uint32_t it = firstIter + chunk * itersPerThread;
    const uint32_t itEnd = min(firstIter + nIter, it + itersPerThread);
    const size_t npx = (size_t)S.resX * S.resY;
    if (view < numViews && it < itEnd)
        TraceTileDiff<false>(S, spec, (tile % tilesX) * 8u, (tile / tilesX) * 8u, it, itEnd, accLds[w],
                             fb ? fb + view * npx * 3 : nullptr, nullptr, 0.f, nullptr, nullptr,
                             sampleL ? sampleL + view * npx * 3 : nullptr, firstIter, nullptr, nullptr, segments,
                             samples, views ? views + view : nullptr, numViews * npx);
    const uint32_t segSum = WaveSum(segments), smpSum = WaveSum(samples);
    if (lane == 0 && counters) {
        atomicAdd(counters + 0, (unsigned long long)segSum);
        atomicAdd(counters + 1, (unsigned long long)smpSum);
    }
}

// Adjoint as a persistent kernel: one 1024-lane workgroup per CU (the LDS texel cache takes most
// of the CU's LDS), each wavefront taking work items (tile, iteration chunk) from a device counter
// until none is left; at the end the workgroup flushes its texel cache and its scalar gradients.
template <int kEmKind, bool kOneLight>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4)))
AdjointKernel(const SceneView *__restrict__ Sp, uint32_t firstIter, uint32_t nIter, uint32_t itersPerItem,
              uint32_t numItems, const float *__restrict__ adjImg, float invN, double *__restrict__ gradEm,
              double *__restrict__ gradScal, uint32_t numScal, float *__restrict__ sampleL,
              uint32_t *__restrict__ work, unsigned long long *__restrict__ counters, const DevView *__restrict__ views,
              uint32_t numViews) {
    const SceneView &S = *Sp;
    const LightSpec spec = (kEmKind == kEmDynamic) ? SceneLightSpec(S)
                                                   : LightSpec{kEmKind, kOneLight ? 1u : S.numLights};
    __shared__ int32_t cacheKeys[kEmCacheSlots];
    __shared__ double cacheVals[kEmCacheSlots * 3];
    __shared__ double scalLds[16][kMaxDiffScalars];
    const uint32_t tid = threadIdx.x, w = tid >> 6, lane = __lane_id();
    for (uint32_t i = tid; i < kEmCacheSlots; i += 1024u) cacheKeys[i] = -1;
    for (uint32_t i = tid; i < kEmCacheSlots * 3; i += 1024u) cacheVals[i] = 0.0;
    for (uint32_t i = tid; i < 16u * kMaxDiffScalars; i += 1024u) scalLds[i / kMaxDiffScalars][i % kMaxDiffScalars] = 0.0;
    __syncthreads();
    const bool useCache = (gradEm != nullptr);
    const uint32_t tilesX = (S.resX + 7u) / 8u, tilesY = (S.resY + 7u) / 8u;
    const uint32_t numTiles = tilesX * tilesY;
    uint32_t segments = 0, samples = 0;
    // Items of view v (chunk x tiles) come from counter work[v].  Workgroup w starts on view
    // w % numViews and moves to the next view when that one's items are used up: a workgroup's
    // texel cache then holds the texels one camera's paths reach, as with one launch per view.
    const uint32_t perView = numItems / numViews;
    uint32_t view = blockIdx.x % numViews, exhausted = 0;
    for (;;) {
        uint32_t local = 0;
        if (lane == 0) local = atomicAdd(work + view, 1u);
        local = __builtin_amdgcn_readfirstlane(__shfl(local, 0, 64));
        if (local >= perView) {
            // a view's counter never refills: every wave leaves after finding each view used up once
            if (++exhausted >= numViews) break;
            view = (view + 1u == numViews) ? 0u : view + 1u;
            continue;
        }
        const uint32_t tile = local % numTiles, chunk = local / numTiles;
        const uint32_t it = firstIter + chunk * itersPerItem;
        const uint32_t itEnd = min(firstIter + nIter, it + itersPerItem);
        const size_t npx = (size_t)S.resX * S.resY;
        TraceTileDiff<true>(S, spec, (tile % tilesX) * 8u, (tile / tilesX) * 8u, it, itEnd, nullptr, nullptr,
                            adjImg + view * npx * 3, invN, gradEm, scalLds[w], sampleL ? sampleL + view * npx * 3 : nullptr,
                            firstIter, useCache ? cacheKeys : nullptr, useCache ? cacheVals : nullptr, segments,
                            samples, views ? views + view : nullptr, numViews * npx);
    }
    __syncthreads();
    if (useCache) {
        for (uint32_t i = tid; i < kEmCacheSlots; i += 1024u) {
            const int32_t t = cacheKeys[i];
            if (t < 0) continue;
            double *p = gradEm + (size_t)t * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double v = cacheVals[i * 3 + c];
                if (v != 0.0) atomicAdd(p + c, v);
            }
        }
    }
    if (tid < numScal) {
        double v = 0.0;
#pragma unroll 1
        for (uint32_t k = 0; k < 16u; ++k) v += scalLds[k][tid];
        if (v != 0.0) atomicAdd(gradScal + tid, v);
    }
    const uint32_t segSum = WaveSum(segments), smpSum = WaveSum(samples);
    if (lane == 0 && counters) {
        atomicAdd(counters + 0, (unsigned long long)segSum);
        atomicAdd(counters + 1, (unsigned long long)smpSum);
    }
}

using AdjFn = void (*)(const SceneView *, uint32_t, uint32_t, uint32_t, uint32_t, const float *, float, double *,
                       double *, uint32_t, float *, uint32_t *, unsigned long long *, const DevView *, uint32_t);

using DiffFn = void (*)(const SceneView *, uint32_t, uint32_t, uint32_t, uint32_t, double *, float *,
                        unsigned long long *, const DevView *, uint32_t);

hipError_t LaunchDiff(const SceneView &S, const SceneView *Sdev, uint32_t firstIter, uint32_t nIter, double *fb,
                      const float *adjImg, double *gradEm, double *gradScal, uint32_t numScal, float *sampleL,
                      uint32_t *work, int numCUs, unsigned long long *counters, hipStream_t stream,
                      const DevView *views, uint32_t numViews) {
    if (nIter == 0) return hipSuccess;
    if (numScal > kMaxDiffScalars || numViews == 0 || (numViews > 1 && !views)) return hipErrorInvalidValue;
    const LightSpec ls = SceneLightSpec(S);
    const bool fast = (ls.em == kEmSimple && ls.numLights == 1);
    const uint32_t tiles = ((S.resX + 7u) / 8u) * ((S.resY + 7u) / 8u);
    if (adjImg) {
        if (!work) return hipErrorInvalidValue;
        const uint32_t ipi = min(nIter, 4u);
        const uint64_t items = (uint64_t)numViews * tiles * ((nIter + ipi - 1) / ipi);
        if (items > 0xFFFFFFF0ull) return hipErrorInvalidValue;
        const AdjFn fn = fast ? AdjointKernel<kEmSimple, true> : AdjointKernel<kEmDynamic, false>;
        hipLaunchKernelGGL(fn, dim3((uint32_t)max(1, numCUs)), dim3(1024), 0, stream, Sdev, firstIter, nIter, ipi,
                           (uint32_t)items, adjImg, 1.0f / (float)nIter, gradEm, gradScal, numScal, sampleL, work,
                           counters, views, numViews);
        return hipGetLastError();
    }
    const DiffFn fn = fast ? DiffKernel<kEmSimple, true> : DiffKernel<kEmDynamic, false>;
    const uint32_t npix = S.resX * S.resY;
    const uint64_t targetThreads = 1u << 21;
    uint32_t numChunks = (uint32_t)((targetThreads + npix - 1) / npix);
    numChunks = max(1u, min(numChunks, nIter));
    const uint32_t ipt = (nIter + numChunks - 1) / numChunks;
    numChunks = (nIter + ipt - 1) / ipt;
    const uint64_t threads = (uint64_t)numViews * tiles * 64u * numChunks;
    if (threads > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const uint32_t blocks = (uint32_t)((threads + 255) / 256);
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(256), 0, stream, Sdev, firstIter, nIter, ipt, numChunks, fb, sampleL,
                       counters, views, numViews);
    return hipGetLastError();
}

}  // namespace pg3
