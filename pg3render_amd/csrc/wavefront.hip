// Wavefront organisation of the path tracer for gfx950.
//
// A batch of up to `capacity` camera samples is traced vertex by vertex with four kernels per
// bounce: Intersect (closest hit of every queued ray), Shade (emission + MIS, light sample, BSDF
// sample), Shadow (occlusion of the light samples) and, after the last bounce, Accumulate (FP64
// framebuffer).  Path state, hit records and both queues live in HBM as 16-byte-per-lane
// records, so every kernel reads and writes fully coalesced cache lines.  Queue appends are
// compacted per wavefront: a 64-lane ballot, one atomic per wave and a popcount prefix give each
// lane its slot.  Every launch is a grid-stride loop over a device-side count, so the host never
// waits for the GPU between bounces and the whole batch is a fixed launch sequence (capturable
// in a HIP graph).  Per-vertex shading is the same ShadeVertex the fused kernel uses.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>

#include "pg3g_internal.hpp"
#include "pt_kernels.hpp"

namespace pg3 {

struct WfQueues {
    uint32_t capacity;
    // path state (one entry per sample of the batch)
    float4 *pA;    // org.xyz, tmin
    float4 *pB;    // dir.xyz, prevBsdfPdf
    float4 *pC;    // beta.xyz, flags
    float4 *pD;    // eThr.xyz, iteration (the path length is the bounce index)
    float4 *pE;    // L.xyz, pixel
    float4 *pF;    // prevN.xyz, light-cache sum
    // hit records
    float4 *hA;    // n.xyz, dist (dist < 0: miss)
    int32_t *hMat;
    // ray queues (ping-pong) and the shadow queue
    uint32_t *rayQ[2];
    float4 *sA;    // p.xyz, tmax
    float4 *sB;    // dir.xyz, path index
    float4 *sC;    // contrib.xyz
    uint32_t *counts;   // [0], [1]: ray queues, [2]: shadow queue, [3]: segments (u32 per batch)
};

__device__ __forceinline__ float AsF(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t AsU(float f) { return __float_as_uint(f); }

// Appends the calling lane to a queue if `pred`; one atomic per wavefront.
__device__ __forceinline__ uint32_t WaveAppend(bool pred, uint32_t *counter) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return 0xFFFFFFFFu;
    const uint32_t lane = __lane_id();
    const uint64_t active = __ballot(1);
    const uint32_t leader = __ffsll((unsigned long long)active) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    const uint64_t below = mask & ((1ull << lane) - 1ull);
    return pred ? base + (uint32_t)__popcll(below) : 0xFFFFFFFFu;
}

__device__ __forceinline__ void StorePath(const WfQueues &Q, uint32_t i, const PathState &ps, uint32_t pixel,
                                          uint32_t iteration) {
    Q.pA[i] = make_float4(ps.org.x, ps.org.y, ps.org.z, ps.tmin);
    Q.pB[i] = make_float4(ps.dir.x, ps.dir.y, ps.dir.z, ps.prevBsdfPdf);
    Q.pC[i] = make_float4(ps.beta.x, ps.beta.y, ps.beta.z, AsF(ps.flags));
    Q.pD[i] = make_float4(ps.eThr.x, ps.eThr.y, ps.eThr.z, AsF(iteration));
    Q.pE[i] = make_float4(ps.L.x, ps.L.y, ps.L.z, AsF(pixel));
    Q.pF[i] = make_float4(ps.prevN.x, ps.prevN.y, ps.prevN.z, ps.lcSum);
}
__device__ __forceinline__ void LoadPath(const WfQueues &Q, uint32_t i, PathState &ps, uint32_t &pixel,
                                         uint32_t &iteration) {
    const float4 a = Q.pA[i], b = Q.pB[i], c = Q.pC[i], d = Q.pD[i], e = Q.pE[i], f = Q.pF[i];
    ps.org = mk(a.x, a.y, a.z);
    ps.tmin = a.w;
    ps.dir = mk(b.x, b.y, b.z);
    ps.prevBsdfPdf = b.w;
    ps.beta = mk(c.x, c.y, c.z);
    ps.flags = AsU(c.w);
    ps.eThr = mk(d.x, d.y, d.z);
    iteration = AsU(d.w);
    ps.L = mk(e.x, e.y, e.z);
    pixel = AsU(e.w);
    ps.prevN = mk(f.x, f.y, f.z);
    ps.lcSum = f.w;
}

// Sample s of the batch -> (pixel, iteration): iterations are the slow index, and pixels of one
// iteration are enumerated in 8x8 tiles so a wavefront starts with 64 neighbouring camera rays.
__device__ __forceinline__ void SampleToPixel(const SceneView &S, uint64_t s, uint32_t firstIter, uint32_t &pixel,
                                              uint32_t &iteration) {
    const uint32_t tilesX = (S.resX + 7u) / 8u, tilesY = (S.resY + 7u) / 8u;
    const uint64_t slots = (uint64_t)tilesX * tilesY * 64u;
    iteration = firstIter + (uint32_t)(s / slots);
    const uint32_t r = (uint32_t)(s % slots);
    const uint32_t tile = r >> 6, lane = r & 63u;
    const uint32_t x = (tile % tilesX) * 8u + (lane & 7u);
    const uint32_t y = (tile / tilesX) * 8u + (lane >> 3);
    pixel = (x < S.resX && y < S.resY) ? y * S.resX + x : 0xFFFFFFFFu;
}

__global__ void __launch_bounds__(256) WfGenerate(SceneView S, WfQueues Q, uint64_t batchBase, uint32_t count,
                                                  uint32_t firstIter) {
    uint32_t n = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        uint32_t pixel, it;
        SampleToPixel(S, batchBase + i, firstIter, pixel, it);
        const bool valid = pixel != 0xFFFFFFFFu;
        PathState ps;
        if (valid) {
            RngKey key{S.k0, S.k1, pixel, it, S.sceneId << 16};
            const float jx = RngGet(key, 0, 0, 0), jy = RngGet(key, 0, 0, 1);   // path_tracer_base.hxx:74-77
            F3 org, dir;
            GenerateRay(S, float(pixel % S.resX) + jx, float(pixel / S.resX) + jy, org, dir);
            InitPath(ps, org, dir);
        } else {
            InitPath(ps, mk(0.f), mk(0.f, 0.f, 1.f));
        }
        StorePath(Q, i, ps, valid ? pixel : 0xFFFFFFFFu, it);
        const uint32_t slot = WaveAppend(valid, &Q.counts[0]);
        if (valid) Q.rayQ[0][slot] = i;
        n += valid;
    }
    (void)n;
}

__global__ void __launch_bounds__(256) WfIntersect(SceneView S, WfQueues Q, uint32_t qi) {
    const uint32_t count = Q.counts[qi];
    const uint32_t *q = Q.rayQ[qi];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += gridDim.x * blockDim.x) {
        const uint32_t i = q[j];
        const float4 a = Q.pA[i], b = Q.pB[i];
        Hit h;
        h.dist = 1e36f;
        const bool hit = Intersect(S, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, h);
        Q.hA[i] = make_float4(h.n.x, h.n.y, h.n.z, hit ? h.dist : -1.f);
        Q.hMat[i] = hit ? h.mat : -1;
    }
}

// Shading stage, specialised like the fused kernel (kEmKind: EM sampler or kEmAbsent; kOneLight)
// and kept at 4 waves per SIMD.  The scene is read through a device copy (held by value in
// kernel arguments its fields would occupy SGPRs and spill), and a scene whose only light is a
// steerable EM takes the vertex that shares the EM lookups of escaped paths and light samples.
template <int kEmKind, bool kOneLight>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
WfShade(const SceneView *__restrict__ Sp, WfQueues Q, uint32_t qi, uint32_t qo, uint32_t bounce) {
    const SceneView &S = *Sp;
    const LightSpec spec{kEmKind, kOneLight ? 1u : S.numLights};
    const uint32_t count = Q.counts[qi];
    const uint32_t *q = Q.rayQ[qi];
    uint32_t segments = 0;
    for (uint32_t j0 = blockIdx.x * blockDim.x; j0 < count; j0 += gridDim.x * blockDim.x) {
        const uint32_t j = j0 + threadIdx.x;
        const bool active = j < count;
        bool cont = false, shadow = false;
        uint32_t i = 0;
        ShadowReq sr;
        sr.valid = false;
        if (active) {
            i = q[j];
            PathState ps;
            uint32_t pixel, it;
            LoadPath(Q, i, ps, pixel, it);
            ps.len = bounce;
            const float4 ha = Q.hA[i];
            Hit h;
            h.dist = ha.w;
            h.n = mk(ha.x, ha.y, ha.z);
            h.mat = Q.hMat[i];
            RngKey key{S.k0, S.k1, pixel, it, S.sceneId << 16};
            cont = ShadeVertex<kOneLight && kEmKind == kEmSteerable>(S, spec, key, ps, h.mat >= 0, h, sr, segments);
            shadow = sr.valid;
            StorePath(Q, i, ps, pixel, it);
        }
        // queue appends are wave-uniform calls (inactive tail lanes pass false)
        const uint32_t rs = WaveAppend(cont, &Q.counts[qo]);
        if (cont) Q.rayQ[qo][rs] = i;
        const uint32_t ss = WaveAppend(shadow, &Q.counts[2]);
        if (shadow) {
            Q.sA[ss] = make_float4(sr.p.x, sr.p.y, sr.p.z, sr.tmax);
            Q.sB[ss] = make_float4(sr.dir.x, sr.dir.y, sr.dir.z, AsF(i));
            Q.sC[ss] = make_float4(sr.contrib.x, sr.contrib.y, sr.contrib.z, 0.f);
        }
    }
    uint32_t w = segments;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
    if (__lane_id() == 0 && w) atomicAdd(&Q.counts[3], w);
}

__global__ void __launch_bounds__(256) WfShadow(SceneView S, WfQueues Q) {
    const uint32_t count = Q.counts[2];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += gridDim.x * blockDim.x) {
        const float4 a = Q.sA[j], b = Q.sB[j];
        if (!Occluded(S, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w)) {
            const uint32_t i = AsU(b.w);
            const float4 c = Q.sC[j];
            float4 L = Q.pE[i];
            L.x += c.x;
            L.y += c.y;
            L.z += c.z;
            Q.pE[i] = L;
        }
    }
}

__global__ void __launch_bounds__(256) WfAccumulate(WfQueues Q, uint32_t count, double *__restrict__ fb,
                                                    unsigned long long *__restrict__ counters) {
    uint32_t samples = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const float4 L = Q.pE[i];
        const uint32_t pixel = AsU(L.w);
        if (pixel == 0xFFFFFFFFu) continue;
        atomicAdd(fb + (size_t)pixel * 3 + 0, (double)L.x);
        atomicAdd(fb + (size_t)pixel * 3 + 1, (double)L.y);
        atomicAdd(fb + (size_t)pixel * 3 + 2, (double)L.z);
        ++samples;
    }
    uint32_t w = samples;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
    if (__lane_id() == 0 && w) atomicAdd(counters + 1, (unsigned long long)w);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(counters + 0, (unsigned long long)Q.counts[3]);
    }
}

__global__ void WfSetCounts(uint32_t *counts, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t mask) {
    if (threadIdx.x == 0) {
        if (mask & 1) counts[0] = a;
        if (mask & 2) counts[1] = b;
        if (mask & 4) counts[2] = c;
        if (mask & 8) counts[3] = d;
    }
}

// ---------------------------------------------------------------------------------------------
// Host side: queue storage and the launch sequence of one render call.
namespace {

struct WfStorage {
    int device = -1;
    uint32_t capacity = 0;
    void *mem = nullptr;
    WfQueues q{};
    ~WfStorage() {
        if (mem) (void)hipFree(mem);
    }
    hipError_t Ensure(uint32_t cap) {
        if (mem && cap <= capacity) return hipSuccess;
        if (mem) (void)hipFree(mem);
        mem = nullptr;
        const size_t f4 = sizeof(float4) * (size_t)cap;
        const size_t bytes = 6 * f4 + f4 + sizeof(int32_t) * (size_t)cap +
                             2 * sizeof(uint32_t) * (size_t)cap + 3 * f4 + 64 + 17 * 256;  // + alignment
        hipError_t e = hipMalloc(&mem, bytes);
        if (e != hipSuccess) return e;
        capacity = cap;
        char *p = static_cast<char *>(mem);
        auto take = [&](size_t n) { char *r = p; p += (n + 255) & ~size_t(255); return r; };
        q.capacity = cap;
        q.pA = (float4 *)take(f4); q.pB = (float4 *)take(f4); q.pC = (float4 *)take(f4);
        q.pD = (float4 *)take(f4); q.pE = (float4 *)take(f4); q.pF = (float4 *)take(f4);
        q.hA = (float4 *)take(f4);
        q.hMat = (int32_t *)take(sizeof(int32_t) * (size_t)cap);
        q.rayQ[0] = (uint32_t *)take(sizeof(uint32_t) * (size_t)cap);
        q.rayQ[1] = (uint32_t *)take(sizeof(uint32_t) * (size_t)cap);
        q.sA = (float4 *)take(f4); q.sB = (float4 *)take(f4); q.sC = (float4 *)take(f4);
        q.counts = (uint32_t *)take(64);
        return hipSuccess;
    }
};

}  // namespace

using ShadeFn = void (*)(const SceneView *, WfQueues, uint32_t, uint32_t, uint32_t);
static ShadeFn PickShadeKernel(const SceneView &S) {
    const bool one = (S.numLights == 1);
    const int em = SceneLightSpec(S).em;
#define PG3_WF_CASE(K) \
    if (em == (K)) return one ? WfShade<(K), true> : WfShade<(K), false>;
    PG3_WF_CASE(kEmAbsent)
    PG3_WF_CASE(kEmCosine)
    PG3_WF_CASE(kEmSimple)
    PG3_WF_CASE(kEmSteerable)
#undef PG3_WF_CASE
    return nullptr;
}

struct WavefrontContext {
    WfStorage st;
    uint32_t capacity = 1u << 21;   // samples per batch (~2M paths, ~300 MB of queues)
    uint32_t gridBlocks = 0;
};

WavefrontContext *CreateWavefrontContext() { return new (std::nothrow) WavefrontContext(); }
void DestroyWavefrontContext(WavefrontContext *w) { delete w; }

hipError_t LaunchWavefront(const SceneView &S, const SceneView *Sdev, uint32_t firstIter, uint32_t nIter, double *fb,
                           unsigned long long *counters, uint32_t *work, int numCUs, WavefrontContext *W,
                           hipStream_t stream) {
    if (!W) return hipErrorInvalidValue;
    if (S.algorithm != 7 || S.maxPL == 0 || S.splitBudget != 1.f || S.iic > 0.f) return LaunchMegakernel(S, Sdev, firstIter, nIter, fb, counters, work, numCUs, stream);
    if (nIter == 0) return hipSuccess;
    if (W->gridBlocks == 0) {
        int dev = 0, cus = 256;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        W->gridBlocks = (uint32_t)cus * 8u;
    }
    const uint32_t tilesX = (S.resX + 7u) / 8u, tilesY = (S.resY + 7u) / 8u;
    const uint64_t slotsPerIter = (uint64_t)tilesX * tilesY * 64u;
    const uint64_t total = slotsPerIter * nIter;
    const uint32_t cap = (uint32_t)std::min<uint64_t>(W->capacity, total);
    hipError_t e = W->st.Ensure(cap);
    if (e != hipSuccess) return e;
    const WfQueues &Q = W->st.q;
    const ShadeFn shade = PickShadeKernel(S);
    if (!shade) return LaunchMegakernel(S, Sdev, firstIter, nIter, fb, counters, work, numCUs, stream);
    const dim3 grid(W->gridBlocks), block(256);
    for (uint64_t base = 0; base < total; base += cap) {
        const uint32_t count = (uint32_t)std::min<uint64_t>(cap, total - base);
        hipLaunchKernelGGL(WfSetCounts, dim3(1), dim3(64), 0, stream, Q.counts, 0u, 0u, 0u, 0u, 15u);
        hipLaunchKernelGGL(WfGenerate, grid, block, 0, stream, S, Q, base, count, firstIter);
        for (uint32_t b = 1; b <= S.maxPL; ++b) {
            const uint32_t qi = (b - 1) & 1u, qo = qi ^ 1u;
            hipLaunchKernelGGL(WfIntersect, grid, block, 0, stream, S, Q, qi);
            hipLaunchKernelGGL(WfSetCounts, dim3(1), dim3(64), 0, stream, Q.counts, 0u, 0u, 0u, 0u,
                               (qo == 0 ? 1u : 2u) | 4u);
            hipLaunchKernelGGL(shade, grid, block, 0, stream, Sdev, Q, qi, qo, b);
            hipLaunchKernelGGL(WfShadow, grid, block, 0, stream, S, Q);
        }
        hipLaunchKernelGGL(WfAccumulate, grid, block, 0, stream, Q, count, fb, counters);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace pg3
