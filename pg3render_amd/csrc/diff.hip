// Differentiable forward render (frozen sampling) and its path-replay adjoint (diff_kernels.hpp).
// Same work decomposition as the plain path-tracing kernel: one wavefront per (8x8 tile,
// iteration chunk), samples handed to lanes from the item's pool.
#include <hip/hip_runtime.h>

#include "diff_kernels.hpp"
#include "pg3g_internal.hpp"

namespace pg3 {

template <int kEmKind, bool kOneLight, bool kAdjoint>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
DiffKernel(const SceneView *__restrict__ Sp, uint32_t firstIter, uint32_t nIter, uint32_t itersPerThread,
           uint32_t numChunks, double *__restrict__ fb, const float *__restrict__ adjImg, float invN,
           double *__restrict__ gradEm, double *__restrict__ gradScal, uint32_t numScal,
           unsigned long long *__restrict__ counters) {
    const SceneView &S = *Sp;
    const LightSpec spec = (kEmKind == kEmDynamic) ? SceneLightSpec(S)
                                                   : LightSpec{kEmKind, kOneLight ? 1u : S.numLights};
    __shared__ double accLds[4][64 * 3];
    __shared__ double scalLds[4][kMaxDiffScalars];
    const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
    if (kAdjoint) {
        for (uint32_t i = lane; i < kMaxDiffScalars; i += 64u) scalLds[w][i] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const uint32_t tilesX = (S.resX + 7u) / 8u, tilesY = (S.resY + 7u) / 8u;
    const uint32_t numTiles = tilesX * tilesY;
    const uint32_t waveId = (uint32_t)(((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t tile = waveId % numTiles, chunk = waveId / numTiles;
    uint32_t segments = 0, samples = 0;
    const uint32_t it = firstIter + chunk * itersPerThread;
    const uint32_t itEnd = min(firstIter + nIter, it + itersPerThread);
    if (chunk < numChunks && it < itEnd)
        TraceTileDiff<kAdjoint>(S, spec, (tile % tilesX) * 8u, (tile / tilesX) * 8u, it, itEnd, accLds[w], fb, adjImg,
                                invN, gradEm, scalLds[w], segments, samples);
    if (kAdjoint) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t i = lane; i < numScal; i += 64u) {
            const double v = scalLds[w][i];
            if (v != 0.0) atomicAdd(gradScal + i, v);
        }
    }
    const uint32_t segSum = WaveSum(segments), smpSum = WaveSum(samples);
    if (lane == 0 && counters) {
        atomicAdd(counters + 0, (unsigned long long)segSum);
        atomicAdd(counters + 1, (unsigned long long)smpSum);
    }
}

using DiffFn = void (*)(const SceneView *, uint32_t, uint32_t, uint32_t, uint32_t, double *, const float *, float,
                        double *, double *, uint32_t, unsigned long long *);

hipError_t LaunchDiff(const SceneView &S, const SceneView *Sdev, uint32_t firstIter, uint32_t nIter, double *fb,
                      const float *adjImg, double *gradEm, double *gradScal, uint32_t numScal,
                      unsigned long long *counters, hipStream_t stream) {
    if (nIter == 0) return hipSuccess;
    if (numScal > kMaxDiffScalars) return hipErrorInvalidValue;
    const bool adjoint = (adjImg != nullptr);
    const LightSpec ls = SceneLightSpec(S);
    const bool fast = (ls.em == kEmSimple && ls.numLights == 1);
    DiffFn fn;
    if (adjoint) fn = fast ? DiffKernel<kEmSimple, true, true> : DiffKernel<kEmDynamic, false, true>;
    else fn = fast ? DiffKernel<kEmSimple, true, false> : DiffKernel<kEmDynamic, false, false>;
    const uint32_t npix = S.resX * S.resY;
    const uint64_t targetThreads = 1u << 21;
    uint32_t numChunks = (uint32_t)((targetThreads + npix - 1) / npix);
    numChunks = max(1u, min(numChunks, nIter));
    const uint32_t ipt = (nIter + numChunks - 1) / numChunks;
    numChunks = (nIter + ipt - 1) / ipt;
    const uint32_t tiles = ((S.resX + 7u) / 8u) * ((S.resY + 7u) / 8u);
    const uint64_t threads = (uint64_t)tiles * 64u * numChunks;
    const uint32_t blocks = (uint32_t)((threads + 255) / 256);
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(256), 0, stream, Sdev, firstIter, nIter, ipt, numChunks, fb, adjImg,
                       1.0f / (float)nIter, gradEm, gradScal, numScal, counters);
    return hipGetLastError();
}

}  // namespace pg3
