// Batched multi-scene path tracing: one persistent launch renders one pass (each scene's spp
// iterations) of many scenes of uneven resolution and sample count.  The reference renders one
// scene per process (pg3render.cxx:54-174, one AbstractRenderer per thread); here every scene is
// a set of work items (one 8x8 pixel tile x a run of iterations) that resident wavefronts take
// from a device counter until none is left.  Each item is traced exactly as the per-scene kernel
// traces it (TraceTasks: same samples, same RNG streams), so a batched pass and per-scene renders
// of the same iterations produce the same per-pixel sums.
#include <hip/hip_runtime.h>

#include "pg3g.h"
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

// LDS cache of a batch workgroup's sampling tables: two slots, each holding the steerable tree's
// inner levels (or the simple sampler's marginal CDF) of one scene.  Slot state word: bit 31 =
// being staged, bits 16-30 = scene + 1 (0: empty), bits 0-15 = wavefronts using it.  A wavefront
// entering a scene takes a slot holding that scene; if a slot is being staged with that scene it
// takes none for now; otherwise it claims a slot no wavefront uses and stages the scene's tables
// into it with its own lanes.  Each takes one compare-and-swap, and a wavefront without a slot
// reads the scene's tables from memory and, at each step of its task loop, joins a slot that has
// finished staging its scene (the copy equals the tables in memory, so paths in flight may switch
// to it).  No wavefront ever waits for another.
constexpr uint32_t kBatchBlockWaves = 16;   // wavefronts of the largest batch workgroup (1024 lanes)
constexpr uint32_t kSlotStaging = 1u << 31;
constexpr uint32_t kSlotSlots = 2;
static_assert(PG3G_MAX_SCENES + 1 < (1u << 15), "scene + 1 fits the slot word's 15 scene bits");
static_assert(kBatchBlockWaves < (1u << 16), "the slot word's user count holds every wavefront of a workgroup");
struct BatchSlots {
    PG3_DEV static F4v *Tables(uint32_t k) {
        __shared__ F4v tables[kSlotSlots][kTreeLdsNodes * 3];
        static_assert(sizeof(tables[0]) >= sizeof(float) * kMargLdsMax, "the marginal CDF fits a slot");
        return tables[k];
    }
    PG3_DEV static uint32_t *State() {
        __shared__ uint32_t state[kSlotSlots];
        return state;
    }
};

// Item source of a batch pass: global item i belongs to the last scene whose first item is <= i;
// within the scene, item = chunk * numTiles + tile as in the per-scene kernel.  The items are cut
// into kBatchParts contiguous parts of about equal sample counts, one per XCD (workgroups are
// dealt to the XCDs in turn, so workgroup w starts on part w % kBatchParts), each handed out in
// order from its own counter: the wavefronts of an XCD work through their part's scenes together
// (one scene's tables and framebuffer at a time in that XCD's L2), and a wavefront meets about
// kBatchParts times fewer scenes -- it drains its lanes at every change of scene -- than with one
// counter for the whole GPU.  A wavefront whose part is used up takes items of the next parts.
template <int kEmKind, bool kOneLight>
struct BatchItems {
    static constexpr bool kMergeEm = kOneLight && kEmKind == kEmSteerable;
    const SceneView *views;
    const BatchScene *info;
    uint32_t numScenes, pass;
    uint32_t *work;                    // [p]: items handed out of part p; [kBatchParts + 1 + p]: its first item
    uint32_t exhausted;                // parts found used up: the wave takes from part w + exhausted
    uint32_t cur;                      // scene entered (~0u: none yet)
    // Its tables, one word (the kernel's scalar registers are all in use): bits 0-1 the LDS slot + 1
    // (0: tables read from memory), bit 2 the slot holds the marginal CDF instead of tree levels,
    // bit 3 the scene has tables a slot can hold, bits 4-8 the staged tree's lowest level and
    // bits 9-31 its first node (TreeCache of StageTree; a tree refined to level 9 has 6.3 M nodes).
    uint32_t tab;
    PG3_DEV int32_t slot() const { return (int32_t)(tab & 3u) - 1; }
    PG3_DEV void setSlot(int32_t k) { tab = (tab & ~3u) | (uint32_t)(k + 1); }
    PG3_DEV bool marg() const { return (tab >> 2) & 1u; }
    PG3_DEV bool want() const { return (tab >> 3) & 1u; }
    PG3_DEV uint32_t minLevel() const { return (tab >> 4) & 31u; }
    PG3_DEV uint32_t baseNode() const { return tab >> 9; }
    // Without a slot: join one that holds the entered scene and is no longer being staged.
    PG3_DEV void Join() {
        if (slot() >= 0 || !want()) return;
        int32_t got = -1;
        if (__lane_id() == 0) {
            uint32_t *state = BatchSlots::State();
            const uint32_t tag = (cur + 1u) << 16;
            for (uint32_t k = 0; k < kSlotSlots && got < 0; ++k) {
                const uint32_t o = __hip_atomic_load(&state[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if ((o & ~0xFFFFu) == tag && (o & 0xFFFFu) < 0xFFFFu && atomicCAS(&state[k], o, o + 1u) == o)
                    got = (int32_t)k;
            }
        }
        setSlot(__builtin_amdgcn_readfirstlane(__shfl(got, 0, 64)));
    }
    PG3_DEV bool Next(TileItem &t) {
        uint32_t item;
        for (;;) {
            const uint32_t part = (blockIdx.x + exhausted) % kBatchParts;
            const uint32_t *parts = work + kBatchParts + 1u;
            uint32_t k = 0;
            if (__lane_id() == 0) k = atomicAdd(&work[part], 1u);
            item = parts[part] + __builtin_amdgcn_readfirstlane(__shfl(k, 0, 64));
            if (item < parts[part + 1]) break;
            // a part never refills: every wavefront leaves after finding each part used up once
            if (++exhausted >= kBatchParts) return false;
        }
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
    PG3_DEV bool Entered(uint32_t s) const { return cur == s; }
    // The wavefront's lanes are all idle: leave the slot of the previous scene, take one for s.
    PG3_DEV void Enter(uint32_t s) {
        uint32_t *state = BatchSlots::State();
        const bool lead = __lane_id() == 0;
        if (slot() >= 0 && lead) atomicSub(&state[slot()], 1u);
        tab = 0u;
        cur = s;
        const DevEm &E = views[s].em;
        const int kind = (kEmKind == kEmDynamic) ? (E.w > 0 ? E.sampler : kEmAbsent) : kEmKind;
        uint32_t count = 0;
        if (kind == kEmSteerable) {
            const uint32_t minLevel = TreeLevels(E, count);
            if (minLevel >= E.numLevels || E.levelBase[minLevel] >= (1u << 23)) return;   // memory tables
            tab = (E.levelBase[minLevel] << 9) | (minLevel << 4);
        } else if (kind == kEmSimple) {
            if (E.h + 1u > kMargLdsMax) return;
            tab = 4u;
        } else {
            return;
        }
        tab |= 8u;
        int32_t got = -1;
        uint32_t stage = 0;
        if (lead) {
            const uint32_t tag = (s + 1u) << 16;
            bool staging = false;          // a slot is being staged with s: join it later
            for (uint32_t k = 0; k < kSlotSlots && got < 0; ++k) {
                const uint32_t o = __hip_atomic_load(&state[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if ((o & ~(kSlotStaging | 0xFFFFu)) != tag) continue;
                if (o & kSlotStaging) staging = true;
                else if ((o & 0xFFFFu) < 0xFFFFu && atomicCAS(&state[k], o, o + 1u) == o) got = (int32_t)k;
            }
            for (uint32_t k = 0; k < kSlotSlots && got < 0 && !staging; ++k) {
                const uint32_t o = __hip_atomic_load(&state[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (!(o & kSlotStaging) && (o & 0xFFFFu) == 0u && atomicCAS(&state[k], o, kSlotStaging | tag | 1u) == o) {
                    got = (int32_t)k;
                    stage = 1;
                }
            }
        }
        got = __builtin_amdgcn_readfirstlane(__shfl(got, 0, 64));
        stage = __builtin_amdgcn_readfirstlane(__shfl(stage, 0, 64));
        if (got < 0) return;
        if (stage) {
            TreeCache tc;
            if (marg()) StageMarginal(E, reinterpret_cast<float *>(BatchSlots::Tables((uint32_t)got)), tc, __lane_id(), 64u);
            else StageTree(E, BatchSlots::Tables((uint32_t)got), __lane_id(), 64u);
            __threadfence_block();   // the tables are in LDS before the slot is published
            if (lead) atomicAnd(&state[got], ~kSlotStaging);
        }
        setSlot(got);
    }
    PG3_DEV const SceneView *View(uint32_t s) const { return views + s; }
    PG3_DEV LightSpec Spec(uint32_t s) const {
        const SceneView &S = views[s];
        LightSpec l = (kEmKind == kEmDynamic) ? SceneLightSpec(S) : LightSpec{kEmKind, kOneLight ? 1u : S.numLights};
        if (slot() >= 0 && s == cur) {
            const float *t = reinterpret_cast<const float *>(BatchSlots::Tables((uint32_t)slot()));
            if (marg()) {
                l.tree.marg = t;
            } else {
                l.tree.nodes = t;
                l.tree.minLevel = minLevel();
                l.tree.baseNode = baseNode();
            }
        }
        return l;
    }
};

// One workgroup of 1024 lanes per compute unit at 4 waves per SIMD.
constexpr uint32_t kBatchBlock = 1024;
template <int kEmKind, bool kOneLight>
__global__ void __launch_bounds__(kBatchBlock) __attribute__((amdgpu_waves_per_eu(4)))
BatchKernel(const SceneView *__restrict__ views, const BatchScene *__restrict__ info, uint32_t numScenes,
            uint32_t *__restrict__ work, unsigned long long *__restrict__ counters) {
    if (threadIdx.x < kSlotSlots) BatchSlots::State()[threadIdx.x] = 0u;
    __syncthreads();
    BatchItems<kEmKind, kOneLight> src{views, info, numScenes, work[kBatchParts], work, 0u, ~0u, 0u};
    RayTally tally;
    TraceTasks<ChainPath<BatchItems<kEmKind, kOneLight>::kMergeEm>, true>(src, tally);
    FlushTally(tally, counters);
}

// Pass bookkeeping nodes of the captured graph (plain kernels: no memset nodes in graphs).
__global__ void BatchBeginPass(uint32_t *work) {
    if (threadIdx.x < kBatchParts) work[threadIdx.x] = 0u;
}
__global__ void BatchEndPass(uint32_t *work) {
    if (threadIdx.x == 0) work[kBatchParts] += 1u;
}

using BatchFn = void (*)(const SceneView *, const BatchScene *, uint32_t, uint32_t *, unsigned long long *);

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
    if (L.blockLanes == 0 || L.blockLanes > kBatchBlock || L.blockLanes % 64u != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(BatchBeginPass, dim3(1), dim3(64), 0, stream, L.work);
    hipLaunchKernelGGL(PickBatchKernel(L.emKind, L.oneLight), dim3(L.gridBlocks), dim3(L.blockLanes), 0, stream,
                       L.views, L.info, L.numScenes, L.work, L.counters);
    hipLaunchKernelGGL(BatchEndPass, dim3(1), dim3(64), 0, stream, L.work);
    return hipGetLastError();
}

}  // namespace pg3
