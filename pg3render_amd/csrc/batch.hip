// Batched multi-scene path tracing: one persistent launch renders one pass (each scene's spp
// iterations) of many scenes of uneven resolution and sample count.  The reference renders one
// scene per process (pg3render.cxx:54-174, one AbstractRenderer per thread); here every scene is
// a set of work items (one 8x8 pixel tile x a run of iterations) that resident wavefronts take
// from a device counter until none is left.  Each item is traced exactly as the per-scene kernel
// traces it (TraceTasks: same samples, same RNG streams), so a batched pass and per-scene renders
// of the same iterations produce the same per-pixel sums.
#include <hip/hip_runtime.h>

#include "pg3g_internal.hpp"
#include "pt_kernels.hpp"

namespace pg3 {

// Wave-uniform scene of a global item index: the last scene whose first item is <= item.
__device__ __forceinline__ uint32_t SceneOfItem(const BatchScene *__restrict__ info, uint32_t numScenes,
                                                uint32_t item) {
    uint32_t lo = 0, hi = numScenes;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (info[mid].itemBegin <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Item source of a batch pass: global item i belongs to the last scene whose first item is <= i;
// within the scene, item = chunk * numTiles + tile as in the per-scene kernel.
template <int kEmKind, bool kOneLight>
struct BatchItems {
    static constexpr bool kMergeEm = kOneLight && kEmKind == kEmSteerable;
    const SceneView *views;
    const BatchScene *info;
    uint32_t numScenes, numItems, pass;
    uint32_t *work;
    PG3_DEV bool Next(TileItem &t) {
        uint32_t item = 0;
        if (__lane_id() == 0) item = atomicAdd(&work[0], 1u);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
        if (item >= numItems) return false;   // every wave leaves once the counter passes the last item
        const uint32_t s = SceneOfItem(info, numScenes, item);
        const BatchScene bi = info[s];
        const uint32_t local = item - bi.itemBegin;
        const uint32_t tile = local % bi.numTiles, chunk = local / bi.numTiles;
        const uint32_t base = pass * bi.spp;
        t.scene = s;
        t.x0 = (tile % bi.tilesX) * 8u;
        t.y0 = (tile / bi.tilesX) * 8u;
        t.it = base + chunk * bi.itersPerItem;
        t.total = 64u * (min(bi.spp, (chunk + 1u) * bi.itersPerItem) - chunk * bi.itersPerItem);
        t.fb = bi.fb;
        return true;
    }
    PG3_DEV const SceneView *View(uint32_t s) const { return views + s; }
    PG3_DEV LightSpec Spec(uint32_t s) const {
        const SceneView &S = views[s];
        return (kEmKind == kEmDynamic) ? SceneLightSpec(S) : LightSpec{kEmKind, kOneLight ? 1u : S.numLights};
    }
};

template <int kEmKind, bool kOneLight>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
BatchKernel(const SceneView *__restrict__ views, const BatchScene *__restrict__ info, uint32_t numScenes,
            uint32_t numItems, uint32_t *__restrict__ work, unsigned long long *__restrict__ counters) {
    RayTally tally;
    BatchItems<kEmKind, kOneLight> src{views, info, numScenes, numItems, work[1], work};
    TraceTasks<ChainPath<BatchItems<kEmKind, kOneLight>::kMergeEm>, true>(src, tally);
    FlushTally(tally, counters);
}

// Pass bookkeeping nodes of the captured graph (plain kernels: no memset nodes in graphs).
__global__ void BatchBeginPass(uint32_t *work) {
    if (threadIdx.x == 0) work[0] = 0u;
}
__global__ void BatchEndPass(uint32_t *work) {
    if (threadIdx.x == 0) work[1] += 1u;
}

using BatchFn = void (*)(const SceneView *, const BatchScene *, uint32_t, uint32_t, uint32_t *, unsigned long long *);

static BatchFn PickBatchKernel(int emKind, int oneLight) {
    if (oneLight < 0) return BatchKernel<kEmDynamic, false>;
#define PG3_BATCH_CASE(K)                                                                     \
    if (emKind == (K)) return oneLight ? BatchKernel<(K), true> : BatchKernel<(K), false>;
    PG3_BATCH_CASE(kEmAbsent)
    PG3_BATCH_CASE(kEmCosine)
    PG3_BATCH_CASE(kEmSimple)
    PG3_BATCH_CASE(kEmSteerable)
#undef PG3_BATCH_CASE
    return BatchKernel<kEmDynamic, false>;
}

hipError_t LaunchBatchPass(const BatchLaunch &L, hipStream_t stream) {
    if (L.numItems == 0 || L.numScenes == 0) return hipSuccess;
    hipLaunchKernelGGL(BatchBeginPass, dim3(1), dim3(64), 0, stream, L.work);
    hipLaunchKernelGGL(PickBatchKernel(L.emKind, L.oneLight), dim3(L.gridBlocks), dim3(256), 0, stream, L.views,
                       L.info, L.numScenes, L.numItems, L.work, L.counters);
    hipLaunchKernelGGL(BatchEndPass, dim3(1), dim3(64), 0, stream, L.work);
    return hipGetLastError();
}

}  // namespace pg3
