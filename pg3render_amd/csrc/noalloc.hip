// Allocation-free adjoint (pg3g_adjoint_noalloc): the path-replay adjoint of diff.hip with its
// scratch memory -- the persistent kernel's item counter and the FP64 gradient accumulators -- in
// one fixed, file-scope device array instead of buffers the library allocates.  For callers that
// cannot allocate device memory (HIP graph capture, tight optimisation loops, harnesses that forbid
// hipMalloc).  Every call on one GPU shares that GPU's copy of the array: calls that overlap on
// one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pg3g_internal.hpp"

namespace pg3 {

// Layout of `scratch` (bytes): the adjoint's work-item counter, the scalar gradients, the EM-texel
// gradients (kNoAllocMaxEmChannels FP64 channels: an RGB EM of up to 256 x 128 texels).
constexpr size_t kNaWorkOff = 0;
constexpr size_t kNaScalOff = 256;
constexpr size_t kNaEmOff = kNaScalOff + sizeof(double) * kMaxDiffScalars;
constexpr size_t kNaBytes = kNaEmOff + sizeof(double) * kNoAllocMaxEmChannels;
static_assert(kNaEmOff % 256 == 0, "the EM-gradient block stays 256-byte aligned");

// One copy per GPU, held until the process ends (769 KiB).  Host code never uses this name as an
// address: the device address comes from hipGetSymbolAddress, per device, at context creation.
static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_naMutex;
void *g_naPtr[kMaxDevices] = {};   // device address of `scratch` on each device (null: not set up)
}  // namespace

hipError_t InitNoAllocScratch(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_naMutex);
    if (g_naPtr[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_naPtr[device] = p;
    return hipSuccess;
}

static void *NoAllocScratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_naMutex);
    return g_naPtr[device];
}

// Clears what the adjoint accumulates into: the item counter, the scalar gradients and the first
// emChannels EM-gradient channels (the rest keeps an earlier call's values and is never read).
__global__ void __launch_bounds__(256) NoAllocClear(unsigned char *base, uint32_t emChannels) {
    uint32_t *work = reinterpret_cast<uint32_t *>(base + kNaWorkOff);
    double *scal = reinterpret_cast<double *>(base + kNaScalOff);
    double *em = reinterpret_cast<double *>(base + kNaEmOff);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) work[0] = 0u;
    if (i < kMaxDiffScalars) scal[i] = 0.0;
    for (uint32_t j = i; j < emChannels; j += gridDim.x * blockDim.x) em[j] = 0.0;
}

// Copies the gradients to the caller's outputs; scalars whose bit is set in zeroMask get 0 (a GGX
// alpha outside the material's clamp range [0.001, 1], materials.hxx:805, has no gradient).
__global__ void __launch_bounds__(256) NoAllocFinish(const unsigned char *base, uint32_t numScal, uint64_t zeroMask,
                                                     double *outScal, uint32_t emChannels, double *outEm) {
    const double *scal = reinterpret_cast<const double *>(base + kNaScalOff);
    const double *em = reinterpret_cast<const double *>(base + kNaEmOff);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (outScal && i < numScal) outScal[i] = ((zeroMask >> i) & 1ull) ? 0.0 : scal[i];
    if (outEm)
        for (uint32_t j = i; j < emChannels; j += gridDim.x * blockDim.x) outEm[j] = em[j];
}

hipError_t LaunchAdjointNoAlloc(int device, const SceneView &S, const SceneView *Sdev, uint32_t firstIter,
                                uint32_t nIter, const float *adjImg, uint32_t numScal, uint64_t zeroMask,
                                const float *sampleL, double *outScal, double *outEm, int numCUs,
                                unsigned long long *counters, hipStream_t stream) {
    const uint64_t emChannels = (uint64_t)S.em.w * S.em.h * 3u;
    if (emChannels > kNoAllocMaxEmChannels || numScal > kMaxDiffScalars) return hipErrorInvalidValue;
    unsigned char *base = static_cast<unsigned char *>(NoAllocScratch(device));
    if (!base) return hipErrorNotInitialized;
    uint32_t *work = reinterpret_cast<uint32_t *>(base + kNaWorkOff);
    double *gScal = reinterpret_cast<double *>(base + kNaScalOff);
    double *gEm = emChannels ? reinterpret_cast<double *>(base + kNaEmOff) : nullptr;
    const uint32_t blocks = 96;   // 24,576 lanes: a few strides over the largest EM block
    hipLaunchKernelGGL(NoAllocClear, dim3(blocks), dim3(256), 0, stream, base, (uint32_t)emChannels);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the adjoint only reads the per-sample radiance
    e = LaunchDiff(S, Sdev, firstIter, nIter, nullptr, adjImg, gEm, gScal, numScal, const_cast<float *>(sampleL), work,
                   numCUs, counters, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(NoAllocFinish, dim3(blocks), dim3(256), 0, stream, base, numScal, zeroMask, outScal,
                       (uint32_t)emChannels, outEm);
    return hipGetLastError();
}

}  // namespace pg3
