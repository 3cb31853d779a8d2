// Wavefront organisation of the path tracer for gfx950.
//
// A batch of up to `capacity` camera samples is traced vertex by vertex with four kernels per
// bounce: Intersect (closest hit of every live path), Shade (emission + MIS, light sample, BSDF
// sample), Shadow (occlusion of the light samples) and, after the last bounce, Accumulate (FP64
// framebuffer).  The queues hold the path state itself, not indices into it: the live paths of a
// bounce are six SoA arrays of 16-byte records in queue order, ping-ponged between two sets, so
// Intersect and Shade read lane j's path at position j and every load is a fully coalesced cache
// line.  Shade writes each continuing path to its slot in the other set, a finished path's radiance
// to the finished array (indexed by the sample's position in the batch) and a light sample to the
// shadow queue; Shadow adds an unoccluded sample's contribution to the path's radiance wherever
// Shade put it, and Accumulate sums each pixel's samples of the batch without atomics.  Queue appends
// are compacted per workgroup: a 64-lane ballot per wave, an LDS scan of the wave counts, one
// atomic per queue per workgroup and a popcount prefix give each lane its slot, so the lanes of a
// wave write consecutive records.  Every launch is a grid-stride loop over a device-side count, so
// the host never waits for the GPU between bounces and the whole batch is a fixed launch sequence
// (capturable in a HIP graph); under Russian roulette the host reads the live count back every
// 8 bounces instead.  Per-vertex shading is the same ShadeVertex the fused kernel uses.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>

#include "pg3g_internal.hpp"
#include "pt_kernels.hpp"

namespace pg3 {

// Path state of the live paths of one bounce, one record per path and field, in queue order.
struct WfPaths {
    float4 *A;   // org.xyz, tmin
    float4 *B;   // dir.xyz, prevBsdfPdf
    float4 *C;   // beta.xyz, flags
    float4 *D;   // eThr.xyz, iteration (the path length is the bounce index)
    float4 *E;   // L.xyz, pixel
    float4 *F;   // prevN.xyz, light-cache sum
};

struct WfQueues {
    uint32_t capacity;
    WfPaths P[2];    // live paths entering the current bounce and the next one (ping-pong)
    float4 *hA;      // hit record of live path j: n.xyz, dist (dist < 0: miss)
    int32_t *hMat;
    float4 *sA;      // shadow queue: p.xyz, tmax
    float4 *sB;      //   dir.xyz, destination of the contribution (below)
    float4 *sC;      //   contrib.xyz
    float4 *done;    // finished paths by sample position in the batch: L.xyz
    uint32_t *counts;   // [0], [1]: live paths in P[0], P[1]; [2], [4]: shadow queue of odd / even
                        // bounces (ShadowCounter); [3]: segments; [5], [6]: extension and shadow
                        // rays traced (the intersected live paths, the shadow requests)
};

// Destination of a shadow contribution: slot of the continuing path in the next live set, or
// the sample's position in the finished array with this bit set.
constexpr uint32_t kDoneBit = 0x80000000u;
constexpr int kWfBlock = 256;

__device__ __forceinline__ float AsF(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t AsU(float f) { return __float_as_uint(f); }

// Appends the calling lane to queue q if p[q], for kN queues at once: a ballot per wavefront, the
// wave counts scanned in LDS and one atomic per queue per workgroup; s[q] receives the lane's slot
// (0xFFFFFFFF where p[q] is false).  Lanes of one wave get consecutive slots.  A wave-level atomic
// would make every wavefront of the grid contend for the same few counters.  Every lane of the
// workgroup calls it (the caller's loop is workgroup-uniform) and kWaves is the workgroup's size
// in wavefronts.
template <int kWaves, int kN>
__device__ __forceinline__ void BlockAppend(const bool (&p)[kN], uint32_t *const (&c)[kN], uint32_t (&s)[kN]) {
    __shared__ uint32_t cnt[kN][kWaves + 1];
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    uint64_t m[kN];
#pragma unroll
    for (int q = 0; q < kN; ++q) m[q] = __ballot(p[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < kN; ++q) cnt[q][wave] = (uint32_t)__popcll(m[q]);
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)kN) {
        uint32_t *k = cnt[threadIdx.x];
        uint32_t total = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t n = k[w];
            k[w] = total;
            total += n;
        }
        // the counters are in global memory (a global atomic: the pointer array hides it)
        k[kWaves] = total ? __hip_atomic_fetch_add((__attribute__((address_space(1))) uint32_t *)c[threadIdx.x], total,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                          : 0u;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int q = 0; q < kN; ++q)
        s[q] = p[q] ? cnt[q][kWaves] + cnt[q][wave] + (uint32_t)__popcll(m[q] & below) : 0xFFFFFFFFu;
    __syncthreads();   // the next call rewrites the counts
}

// Orders a workgroup's chunk of live paths by material class (the hit's material, kClsMiss for an
// escaped path; inactive lanes last) so that its wavefronts shade one material each, except where
// a class's run ends inside a wavefront: returns the live-path position this lane takes after
// the reordering (0xFFFFFFFF: none).  A counting sort over the workgroup's lanes: a ballot per
// class and wavefront, class offsets scanned in LDS.  Every lane of the workgroup calls it.
template <int kWaves>
__device__ __forceinline__ uint32_t SortChunkByClass(uint32_t j, uint32_t cls) {
    __shared__ uint32_t cnt[kNumCls + 1][kWaves];
    __shared__ uint32_t perm[kWaves * 64];
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t c = 0; c <= (uint32_t)kNumCls; ++c) {
        const uint64_t m = __ballot(cls == c);
        if (cls == c) mine = m;
        if (lane == 0) cnt[c][wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t off = 0;
        for (uint32_t c = 0; c <= (uint32_t)kNumCls; ++c)
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t n = cnt[c][w];
                cnt[c][w] = off;
                off += n;
            }
    }
    __syncthreads();
    perm[cnt[cls][wave] + (uint32_t)__popcll(mine & ((1ull << lane) - 1ull))] = j;
    __syncthreads();
    const uint32_t r = perm[threadIdx.x];
    __syncthreads();   // the next call rewrites the table
    return r;
}

__device__ __forceinline__ void StorePath(const WfPaths &P, uint32_t i, const PathState &ps, uint32_t pixel,
                                          uint32_t iteration) {
    P.A[i] = make_float4(ps.org.x, ps.org.y, ps.org.z, ps.tmin);
    P.B[i] = make_float4(ps.dir.x, ps.dir.y, ps.dir.z, ps.prevBsdfPdf);
    P.C[i] = make_float4(ps.beta.x, ps.beta.y, ps.beta.z, AsF(ps.flags));
    P.D[i] = make_float4(ps.eThr.x, ps.eThr.y, ps.eThr.z, AsF(iteration));
    P.E[i] = make_float4(ps.L.x, ps.L.y, ps.L.z, AsF(pixel));
    P.F[i] = make_float4(ps.prevN.x, ps.prevN.y, ps.prevN.z, ps.lcSum);
}
__device__ __forceinline__ void LoadPath(const WfPaths &P, uint32_t i, PathState &ps, uint32_t &pixel,
                                         uint32_t &iteration) {
    const float4 a = P.A[i], b = P.B[i], c = P.C[i], d = P.D[i], e = P.E[i], f = P.F[i];
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

// Sample s of the render call -> (pixel, iteration): iterations are the slow index, and pixels of one
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
// The inverse: sample of the render call of an (in-image) pixel and iteration.
__device__ __forceinline__ uint64_t PixelToSample(const SceneView &S, uint32_t pixel, uint32_t iteration,
                                                  uint32_t firstIter) {
    const uint32_t tilesX = (S.resX + 7u) / 8u, tilesY = (S.resY + 7u) / 8u;
    const uint64_t slots = (uint64_t)tilesX * tilesY * 64u;
    const uint32_t x = pixel % S.resX, y = pixel / S.resX;
    const uint32_t r = ((y >> 3) * tilesX + (x >> 3)) * 64u + (y & 7u) * 8u + (x & 7u);
    return (uint64_t)(iteration - firstIter) * slots + r;
}

// Camera rays of samples [batchBase, batchBase + count) into live set 0; samples of the tile
// padding outside the image are not enqueued.  Without padding (both image sides multiples of
// 8, `dense`) every sample is valid and sample i takes slot i: no append, and the first lane
// sets the count.
__global__ void __launch_bounds__(kWfBlock) WfGenerate(SceneView S, WfQueues Q, uint64_t batchBase, uint32_t count,
                                                       uint32_t firstIter, bool dense) {
    if (dense && blockIdx.x == 0 && threadIdx.x == 0) Q.counts[0] = count;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < count; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t pixel = 0xFFFFFFFFu, it = 0;
        if (i < count) SampleToPixel(S, batchBase + i, firstIter, pixel, it);
        const bool valid = pixel != 0xFFFFFFFFu;
        uint32_t s[1] = {i};
        if (!dense) {
            const bool p[1] = {valid};
            uint32_t *const c[1] = {&Q.counts[0]};
            BlockAppend<kWfBlock / 64, 1>(p, c, s);
        }
        if (valid) {
            RngKey key{S.k0, S.k1, pixel, it, S.sceneId << 16};
            const float jx = RngGet(key, 0, 0, 0), jy = RngGet(key, 0, 0, 1);   // path_tracer_base.hxx:74-77
            F3 org, dir;
            GenerateRay(S, float(pixel % S.resX) + jx, float(pixel / S.resX) + jy, org, dir);
            PathState ps;
            InitPath(ps, org, dir);
            StorePath(Q.P[0], s[0], ps, pixel, it);
        }
    }
}

// Also zeroes the counters selected by clearMask (bit k: counts[k]; never qi): the shading stage
// that follows appends to them, and every earlier reader has finished (stream order), so no
// launch of its own is needed for the clear.
// The body runs as workgroup `blk` of `nblk` (the merged launch below gives it part of its grid).
__device__ __forceinline__ void IntersectStage(const SceneView &S, const WfQueues &Q, uint32_t qi, uint32_t clearMask,
                                               uint32_t blk, uint32_t nblk) {
    if (blk == 0 && threadIdx.x < 8u && (clearMask >> threadIdx.x & 1u)) Q.counts[threadIdx.x] = 0u;
    const uint32_t count = Q.counts[qi];
    if (blk == 0 && threadIdx.x == 0 && count) atomicAdd(&Q.counts[5], count);   // extension rays
    const WfPaths P = Q.P[qi];
    for (uint32_t j = blk * blockDim.x + threadIdx.x; j < count; j += nblk * blockDim.x) {
        const float4 a = P.A[j], b = P.B[j];
        Hit h;
        h.dist = 1e36f;
        const bool hit = Intersect(S, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, h);
        Q.hA[j] = make_float4(h.n.x, h.n.y, h.n.z, hit ? h.dist : -1.f);
        Q.hMat[j] = hit ? h.mat : -1;
    }
}
__global__ void __launch_bounds__(kWfBlock) WfIntersect(SceneView S, WfQueues Q, uint32_t qi, uint32_t clearMask) {
    IntersectStage(S, Q, qi, clearMask, blockIdx.x, gridDim.x);
}

// Shadow queue counter of bounce b: odd and even bounces alternate between counts[2] and
// counts[4], so that one launch can test bounce b's shadow rays while it intersects (and clears
// the counters for) bounce b + 1.
__host__ __device__ constexpr uint32_t ShadowCounter(uint32_t b) { return (b & 1u) ? 2u : 4u; }

// Shading stage, specialised like the fused kernel (kEmKind: EM sampler or kEmAbsent; kOneLight)
// and kept at 4 waves per SIMD.  The scene is read through a device copy (held by value in
// kernel arguments its fields would occupy SGPRs and spill), and a scene whose only light is a
// steerable EM takes the vertex that shares the EM lookups of escaped paths and light samples.
// Each 256-lane workgroup stages the top levels of the steerable tree that fit in
// PG3_WF_TREE_NODES nodes in LDS (the fused kernel's one 1024-lane workgroup per CU would stage
// the whole 1280-node budget once, but its queue-append barriers then hold all 16 wavefronts of
// the CU in step, which measured slower).  The simple sampler's marginal CDF is staged.
// Register budget of the shading stage in waves per SIMD (its grid is the resident workgroups).
// C2 frame: 2, 3, 4, 5, 6, 8 waves: 22.5, 21.3, 20.9, 24.9, 29.9, 40.7 ms (fewer waves hide less
// load latency; more spill).
#ifndef PG3_WF_SHADE_WAVES
#define PG3_WF_SHADE_WAVES 4
#endif
// Steerable tree nodes staged in LDS per shading workgroup (48 B each; the 4 resident workgroups
// of a CU share its 160 KiB); 0 reads the tree from the caches.  C2 frame: 0, 400, 800 nodes:
// 20.4, 19.4, 19.4 ms.
#ifndef PG3_WF_TREE_NODES
#define PG3_WF_TREE_NODES 400
#endif
template <int kEmKind, bool kOneLight>
__global__ void __launch_bounds__(kWfBlock) __attribute__((amdgpu_waves_per_eu(PG3_WF_SHADE_WAVES)))
WfShade(const SceneView *__restrict__ Sp, WfQueues Q, uint32_t qi, uint32_t qo, uint32_t bounce, uint64_t batchBase,
        uint32_t firstIter) {
    const SceneView &S = *Sp;
    LightSpec spec{kEmKind, kOneLight ? 1u : S.numLights};
    if constexpr (kEmKind == kEmSimple) {
        __shared__ float margLds[kMargLdsMax];
        StageMarginal(S.em, margLds, spec.tree);
        __syncthreads();
    }
#if PG3_WF_TREE_NODES > 0
    if constexpr (kEmKind == kEmSteerable) {
        __shared__ F4v treeLds[PG3_WF_TREE_NODES * 3];
        spec.tree = StageTree<PG3_WF_TREE_NODES>(S.em, treeLds);
        __syncthreads();
    }
#endif
    const uint32_t count = Q.counts[qi];
    const WfPaths Pi = Q.P[qi], Po = Q.P[qo];
    uint32_t segments = 0;
    for (uint32_t j0 = blockIdx.x * blockDim.x; j0 < count; j0 += gridDim.x * blockDim.x) {
        // material-coherent wavefronts: the chunk's paths reordered by material class
        uint32_t jl = j0 + threadIdx.x, cls = kNumCls;
        if (jl < count) {
            const int32_t m = Q.hMat[jl];
            cls = (m < 0) ? (uint32_t)kClsMiss : ClassOfType(AsGlobal(&S.mats[m].type)[0]);
        } else {
            jl = 0xFFFFFFFFu;
        }
        const uint32_t j = SortChunkByClass<kWfBlock / 64>(jl, cls);
        const bool active = j != 0xFFFFFFFFu;
        bool cont = false;
        uint32_t pixel = 0xFFFFFFFFu, it = 0;
        PathState ps;
        ShadowReq sr;
        sr.valid = false;
        if (active) {
            LoadPath(Pi, j, ps, pixel, it);
            ps.len = bounce;
            const float4 ha = Q.hA[j];
            Hit h;
            h.dist = ha.w;
            h.n = mk(ha.x, ha.y, ha.z);
            h.mat = Q.hMat[j];
            RngKey key{S.k0, S.k1, pixel, it, S.sceneId << 16};
            cont = ShadeVertex<kOneLight && kEmKind == kEmSteerable>(S, spec, key, ps, h.mat >= 0, h, sr, segments);
        }
        // workgroup-uniform appends (inactive tail lanes pass false)
        const bool p[2] = {cont, sr.valid};
        uint32_t *const c[2] = {&Q.counts[qo], &Q.counts[ShadowCounter(bounce)]};
        uint32_t s[2];
        BlockAppend<kWfBlock / 64, 2>(p, c, s);
        uint32_t fin = 0;
        if (cont) {
            StorePath(Po, s[0], ps, pixel, it);
        } else if (active) {
            fin = (uint32_t)(PixelToSample(S, pixel, it, firstIter) - batchBase);
            Q.done[fin] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.f);
        }
        if (sr.valid) {
            const uint32_t dst = cont ? s[0] : (fin | kDoneBit);
            Q.sA[s[1]] = make_float4(sr.p.x, sr.p.y, sr.p.z, sr.tmax);
            Q.sB[s[1]] = make_float4(sr.dir.x, sr.dir.y, sr.dir.z, AsF(dst));
            Q.sC[s[1]] = make_float4(sr.contrib.x, sr.contrib.y, sr.contrib.z, 0.f);
        }
    }
    const uint32_t w = WaveSum(segments);
    if (__lane_id() == 0 && w) atomicAdd(&Q.counts[3], w);
}

// Occlusion of the light samples of bounce b; qo is the live set the continuing paths were
// written to.  Runs as workgroup `blk` of `nblk`.
__device__ __forceinline__ void ShadowStage(const SceneView &S, const WfQueues &Q, uint32_t qo, uint32_t b,
                                            uint32_t blk, uint32_t nblk) {
    const uint32_t count = Q.counts[ShadowCounter(b)];
    if (blk == 0 && threadIdx.x == 0 && count) atomicAdd(&Q.counts[6], count);   // shadow rays
    float4 *const live = Q.P[qo].E;
    for (uint32_t j = blk * blockDim.x + threadIdx.x; j < count; j += nblk * blockDim.x) {
        const float4 a = Q.sA[j], b = Q.sB[j];
        if (!Occluded(S, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w)) {
            const uint32_t dst = AsU(b.w);
            float4 *const L = (dst & kDoneBit) ? Q.done + (dst & ~kDoneBit) : live + dst;
            const float4 c = Q.sC[j];
            float4 v = *L;
            v.x += c.x;
            v.y += c.y;
            v.z += c.z;
            *L = v;
        }
    }
}
__global__ void __launch_bounds__(kWfBlock) WfShadow(SceneView S, WfQueues Q, uint32_t qo, uint32_t b) {
    ShadowStage(S, Q, qo, b, blockIdx.x, gridDim.x);
}
// Bounce b's shadow rays (the first shadowBlocks workgroups) and bounce b + 1's intersections
// (the rest) in one launch: the two touch disjoint data (the radiance of the live paths and the
// shadow queue; the rays and hit records of the live paths) and disjoint counters.
__global__ void __launch_bounds__(kWfBlock) WfShadowIntersect(SceneView S, WfQueues Q, uint32_t qo, uint32_t b,
                                                              uint32_t clearMask, uint32_t shadowBlocks) {
    if (blockIdx.x < shadowBlocks) ShadowStage(S, Q, qo, b, blockIdx.x, shadowBlocks);
    else IntersectStage(S, Q, qo, clearMask, blockIdx.x - shadowBlocks, gridDim.x - shadowBlocks);
}

// Adds the finished paths of samples [batchBase, batchBase + count) to the FP64 framebuffer.  Lane
// r' owns the pixel of sample batchBase + r' and sums, in iteration order, every sample of the
// batch on that pixel (positions r', r' + slots, ...), so each pixel is read and written once, by
// one lane.  Every path has finished: the last bounce stops at the path-length limit.
__global__ void __launch_bounds__(kWfBlock) WfAccumulate(SceneView S, WfQueues Q, uint64_t batchBase, uint32_t count,
                                                         uint32_t firstIter, double *__restrict__ fb,
                                                         unsigned long long *__restrict__ counters) {
    const uint32_t tilesX = (S.resX + 7u) / 8u, tilesY = (S.resY + 7u) / 8u;
    const uint32_t slots = tilesX * tilesY * 64u;
    const uint32_t n = std::min(slots, count);
    uint32_t samples = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        uint32_t pixel, it;
        SampleToPixel(S, batchBase + r, firstIter, pixel, it);
        if (pixel == 0xFFFFFFFFu) continue;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (uint32_t i = r; i < count; i += slots) {
            const float4 L = Q.done[i];
            sx += (double)L.x;
            sy += (double)L.y;
            sz += (double)L.z;
            ++samples;
        }
        double *const f = fb + (size_t)pixel * 3;
        f[0] += sx;
        f[1] += sy;
        f[2] += sz;
    }
    uint32_t w = samples;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
    if (__lane_id() == 0 && w) atomicAdd(counters + 1, (unsigned long long)w);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(counters + 0, (unsigned long long)Q.counts[3]);
        atomicAdd(counters + 2, (unsigned long long)Q.counts[5]);
        atomicAdd(counters + 3, (unsigned long long)Q.counts[6]);
    }
}

// Zeroes the counters selected by `mask` (bit k: counts[k]).
__global__ void WfClearCounts(uint32_t *counts, uint32_t mask) {
    if (threadIdx.x < 8 && (mask >> threadIdx.x & 1u)) counts[threadIdx.x] = 0u;
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
        // two live sets of 6 records, the hit record (16 + 4 B), 3 shadow records, 1 finished record
        const size_t bytes = 12 * f4 + f4 + sizeof(int32_t) * (size_t)cap + 3 * f4 + f4 + 64 + 20 * 256;
        hipError_t e = hipMalloc(&mem, bytes);
        if (e != hipSuccess) return e;
        capacity = cap;
        char *p = static_cast<char *>(mem);
        auto take = [&](size_t n) { char *r = p; p += (n + 255) & ~size_t(255); return r; };
        q.capacity = cap;
        for (WfPaths &P : q.P) {
            P.A = (float4 *)take(f4); P.B = (float4 *)take(f4); P.C = (float4 *)take(f4);
            P.D = (float4 *)take(f4); P.E = (float4 *)take(f4); P.F = (float4 *)take(f4);
        }
        q.hA = (float4 *)take(f4);
        q.hMat = (int32_t *)take(sizeof(int32_t) * (size_t)cap);
        q.sA = (float4 *)take(f4); q.sB = (float4 *)take(f4); q.sC = (float4 *)take(f4);
        q.done = (float4 *)take(f4);
        q.counts = (uint32_t *)take(64);
        return hipSuccess;
    }
};

}  // namespace

using ShadeFn = void (*)(const SceneView *, WfQueues, uint32_t, uint32_t, uint32_t, uint64_t, uint32_t);
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
    uint32_t capacity = 1u << 21;   // samples per batch (~2M paths, ~580 MB of queues)
    uint32_t gridBlocks = 0;    // grid-stride stages: 8 workgroups per CU
    // shading stage: exactly the resident workgroups (16 wavefronts per CU at its 4 waves per
    // SIMD), each looping over the live paths; more workgroups measured slower (launch and
    // set-up per workgroup), e.g. 25.0 ms per C2 frame at 8 per CU against 23.2 at 4
    uint32_t shadeBlocks = 0;
};

WavefrontContext *CreateWavefrontContext() { return new (std::nothrow) WavefrontContext(); }
void DestroyWavefrontContext(WavefrontContext *w) { delete w; }

hipError_t LaunchWavefront(const SceneView &S, const SceneView *Sdev, uint32_t firstIter, uint32_t nIter, double *fb,
                           unsigned long long *counters, uint32_t *work, int numCUs, WavefrontContext *W,
                           hipStream_t stream) {
    if (!W) return hipErrorInvalidValue;
    if (S.algorithm != 7 || S.splitBudget != 1.f || S.iic > 0.f)
        return LaunchMegakernel(S, Sdev, firstIter, nIter, fb, counters, work, numCUs, stream);
    if (nIter == 0) return hipSuccess;
    if (W->gridBlocks == 0) {
        int dev = 0, cus = 256;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        W->gridBlocks = (uint32_t)cus * 8u;
        W->shadeBlocks = (uint32_t)cus * 4u * PG3_WF_SHADE_WAVES / (kWfBlock / 64u);
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
    const dim3 grid(W->gridBlocks), block(kWfBlock);
    // With a path-length limit the batch is a fixed launch sequence.  Under Russian roulette
    // (-maxpl 0) the number of bounces is open: bounces run in groups of 8 and the live-path
    // count is read back after each group; every path stops at 1200 vertices at the latest.
    const uint32_t maxBounces = S.maxPL > 0 ? S.maxPL : 1200u;
    const bool dense = (S.resX % 8u) == 0u && (S.resY % 8u) == 0u;
    for (uint64_t base = 0; base < total; base += cap) {
        const uint32_t count = (uint32_t)std::min<uint64_t>(cap, total - base);
        hipLaunchKernelGGL(WfClearCounts, dim3(1), dim3(64), 0, stream, Q.counts, 0xFFu);
        hipLaunchKernelGGL(WfGenerate, grid, block, 0, stream, S, Q, base, count, firstIter, dense);
        hipLaunchKernelGGL(WfIntersect, grid, block, 0, stream, S, Q, 0u, (1u << 1) | (1u << ShadowCounter(1)));
        for (uint32_t b = 1; b <= maxBounces; ++b) {
            // bounce b: live set qi (already intersected) -> qo; the clear mask of bounce b + 1
            // names the counters its shading appends to (qi's and its shadow queue's), whose
            // readers (bounce b's shading, bounce b - 1's shadow rays) have finished
            const uint32_t qi = (b - 1) & 1u, qo = qi ^ 1u;
            hipLaunchKernelGGL(shade, dim3(W->shadeBlocks), block, 0, stream, Sdev, Q, qi, qo, b, base,
                               firstIter);
            if (b < maxBounces)
                hipLaunchKernelGGL(WfShadowIntersect, dim3(2u * grid.x), block, 0, stream, S, Q, qo, b,
                                   (1u << qi) | (1u << ShadowCounter(b + 1u)), grid.x);
            else
                hipLaunchKernelGGL(WfShadow, grid, block, 0, stream, S, Q, qo, b);
            if (S.maxPL == 0 && (b % 8u) == 0u && b < maxBounces) {
                uint32_t live = 0;
                e = hipMemcpyAsync(&live, Q.counts + qo, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
                if (e == hipSuccess) e = hipStreamSynchronize(stream);
                if (e != hipSuccess) return e;
                if (live == 0) break;
            }
        }
        hipLaunchKernelGGL(WfAccumulate, grid, block, 0, stream, S, Q, base, count, firstIter, fb, counters);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace pg3
