// Closest-hit ray casting that allocates nothing (yrh_trace_closest_noalloc),
// for callers that cannot call hipMalloc: code under HIP graph capture, tight
// serving loops. The traversal's scratch memory -- the ray-claim counters and
// the per-lane stack spill region -- lives in one file-scope device array per
// GPU, sized once at load; the launch shape is fitted to it.
#include "yrh_internal.h"
#include "yrh_traverse.h"

#include <mutex>

namespace yrh {
namespace {

// Scratch layout: [0, kCounterBytes) the claim counters (trav::kParts ranges,
// trav::kPartStride words apart), then the stack spill region.
constexpr size_t kScratchBytes = size_t{64} << 20;
constexpr size_t kCounterBytes = 4096;
static_assert(trav::kParts * trav::kPartStride * sizeof(uint32_t) <= kCounterBytes, "claim counters fit");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` per GPU (each GPU holds its own copy of the
// array). The host-side name `scratch` is only a stand-in for the symbol, so
// kernels and copies always get the address looked up here.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

__global__ void ZeroWordsKernel(uint32_t* words, int count) {
    for (int i = static_cast<int>(threadIdx.x); i < count; i += static_cast<int>(blockDim.x)) words[i] = 0u;
}

}  // namespace

int InitNoAllocScratch(yrh_context* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return Fail(ctx, YRH_E_UNSUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    void*& ptr = g_scratch_ptr[ctx->device];
    if (ptr == nullptr) {
        // The caller (yrh_create) has made ctx->device current.
        void* p = nullptr;
        const int rc = CheckHip(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress scratch");
        if (rc != YRH_OK) return rc;
        ptr = p;
    }
    ctx->noalloc_scratch = ptr;
    return YRH_OK;
}

}  // namespace yrh

using namespace yrh;

extern "C" {

int yrh_trace_closest_noalloc(yrh_context* ctx, const yrh_ray_soa* rays, int64_t n, const yrh_hit_soa* hits) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (rays == nullptr || hits == nullptr || n < 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null rays/hits or negative count");
    if (n > 0 && (rays->org_tmin == nullptr || rays->dir_tmax == nullptr || hits->hit == nullptr)) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "ray or hit array missing");
    }
    if (n > INT32_MAX) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "at most 2^31-1 rays per call");
    if (!ctx->geo.loaded) return Fail(ctx, YRH_E_NO_SCENE, "no geometry uploaded");
    if (ctx->noalloc_scratch == nullptr) return Fail(ctx, YRH_E_HIP, "no scratch array on this device");
    if (n == 0) return YRH_OK;
    (void)hipSetDevice(ctx->device);
    auto* base = static_cast<unsigned char*>(ctx->noalloc_scratch);
    auto* counters = reinterpret_cast<uint32_t*>(base);
    // The array keeps the previous call's counters: clear them with a kernel
    // (no memset node, so the call can be captured into a graph).
    hipLaunchKernelGGL(ZeroWordsKernel, dim3(1), dim3(256), 0, ctx->stream, counters,
                       trav::kParts * trav::kPartStride);
    int rc = CheckHip(ctx, hipGetLastError(), "scratch clear launch");
    if (rc != YRH_OK) return rc;
    TraceRequest r;
    r.any = false;
    r.org = reinterpret_cast<const float4*>(rays->org_tmin);
    r.dir = reinterpret_cast<const float4*>(rays->dir_tmax);
    r.n_host = n;
    r.hit = reinterpret_cast<float4*>(hits->hit);
    r.instance = hits->instance;
    r.node_tests = hits->node_tests;
    r.triangle_tests = hits->triangle_tests;
    r.sphere_tests = hits->sphere_tests;
    r.work_counter = counters;
    r.parts = trav::kParts;
    r.spill = reinterpret_cast<int2*>(base + kCounterBytes);
    r.spill_bytes = kScratchBytes - kCounterBytes;
    return LaunchTrace(ctx, r);
}

}  // extern "C"
