// Internal declarations shared by the libyrh translation units.
//
// Device layout of the acceleration structure (MI355X / gfx950):
//   * A binary BVH is the reference's own tree (include/yaoray/accel/bvh.hpp:34-47),
//     repacked so that an interior node carries both child boxes: one 64-byte
//     record (four 16-byte loads issued together) instead of two dependent
//     32-byte node reads per interior step. Visit order, box arithmetic and the
//     node_tests counter are unchanged.
//   * A child is named by a 32-bit "code": >= 0 is an interior node index,
//     < 0 a leaf: bit 31 set, bits 27..30 = count-1 (15 = escape, count read from
//     leaf_count[first]), bits 0..26 = first primitive slot.
//   * Primitive slots follow the leaf order of primitive_indices; each slot
//     holds the triangle as (p0, e1 = p1 - p0, e2 = p2 - p0) plus its flat
//     triangle index: 3 x float4 = 48 bytes, one coalesced fetch per test in
//     place of the reference's 5-level index chase (bvh_traversal.cpp:59-79).
//     e1/e2 are the same IEEE subtractions the reference performs per test,
//     so every intersection is bit-identical.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/yrh.h"

namespace yrh {

// A roctx range over a host-side stage (entry points, pass stages): shown by
// `rocprofv3 --marker-trace` beside the kernels the stage enqueues; a
// push / pop pair costs nanoseconds without a profiler.
struct StageRange {
    explicit StageRange(const char* name) { roctxRangePushA(name); }
    ~StageRange() { roctxRangePop(); }
    StageRange(const StageRange&) = delete;
    StageRange& operator=(const StageRange&) = delete;
};

constexpr int kLeafCountEscape = 15;
// Context work-counter words (yrh_context::work_counter).
constexpr int kOverflowWord = 2;  // traversal stack overflow flag
constexpr int kCounterSetBase = 256;
constexpr int kCounterSetWords = 256;  // 8 ranges x 32 words
constexpr size_t kWorkCounterBytes = (kCounterSetBase + 2 * kCounterSetWords) * sizeof(uint32_t);
constexpr int32_t kMaxSlot = (1 << 27) - 1;

struct alignas(16) DevNode {
    float4 a;  // L.min.x L.min.y L.min.z L.max.x
    float4 b;  // L.max.y L.max.z R.min.x R.min.y
    float4 c;  // R.min.z R.max.x R.max.y R.max.z
    int4 d;    // left code, right code, 0, 0
};
static_assert(sizeof(DevNode) == 64, "node record is 64 bytes");


struct alignas(16) DevBlasRoot {
    float4 lo;  // root box min, w = root code bits
    float4 hi;  // root box max
};

// One TLAS leaf slot: the instance's world-to-object rows and its BLAS root,
// so a traversal step into an instance is one 80-byte fetch.
struct alignas(16) DevInstanceSlot {
    float4 r0;  // world_to_object m[0], m[4], m[8],  m[12]
    float4 r1;  //                 m[1], m[5], m[9],  m[13]
    float4 r2;  //                 m[2], m[6], m[10], m[14]
    float4 lo;  // BLAS root box min, w = root code bits
    float4 hi;  // BLAS root box max, w = instance index bits
};
static_assert(sizeof(DevInstanceSlot) == 80, "instance slot record is 80 bytes");

inline int32_t PackLeaf(int32_t first, int32_t count) {
    const int32_t nib = count <= kLeafCountEscape ? count - 1 : kLeafCountEscape;
    return static_cast<int32_t>(0x80000000u | (static_cast<uint32_t>(nib) << 27) |
                                static_cast<uint32_t>(first));
}

// Device buffer owned by a context.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& other) noexcept : ptr(other.ptr), bytes(other.bytes) {
        other.ptr = nullptr;
        other.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& other) noexcept {
        if (this != &other) {
            reset();
            ptr = other.ptr;
            bytes = other.bytes;
            other.ptr = nullptr;
            other.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { reset(); }
    void reset() {
        if (ptr != nullptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

struct DevGeometry {
    bool loaded = false;
    int accel_kind = YRH_ACCEL_FLAT;
    // flat BVH, or every BLAS concatenated (node and slot indices rebased)
    DevBuf nodes;
    DevBuf slots;       // float4[3 * slot_count]
    DevBuf leaf_count;  // int32[slot_count], read only for escaped leaves
    int64_t slot_count = 0;
    int32_t root_code = 0;
    float root_lo[3] = {0, 0, 0};
    float root_hi[3] = {0, 0, 0};
    bool empty = true;
    bool has_spheres = false;  // sphere slots exist: launch the traversal variant that tests them
    // two-level
    DevBuf tlas_nodes;
    DevBuf tlas_leaf_count;
    DevBuf instance_slots;      // DevInstanceSlot per TLAS slot
    int32_t instance_count = 0;
    int32_t blas_count = 0;
    int32_t stack_capacity = 0;  // entries needed by the deepest traversal
};

struct RenderState;  // yrh_render.hip
struct GuideState;   // yrh_guide.hip

}  // namespace yrh

namespace yrg {
struct GuideView;  // yrh_guide.h
}

struct yrh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int compute_units = 0;
    char device_name[256] = {0};
    std::string error;
    yrh::DevGeometry geo;
    // persistent traversal scratch
    // uint32 words: [2] stack overflow flag; from kCounterSetBase two sets of
    // per-range ray counters (kCounterSetWords each) used by alternate
    // launches, each launch zeroing the other set for its successor
    yrh::DevBuf work_counter;
    int launch_parity = 0;
    yrh::DevBuf stack_spill;    // int2 per lane per spilled stack level
    // Render passes run each bounce's shadow rays on a second stream, beside
    // the next bounce's path-segment traversal (yrh_render.hip); it has its
    // own spill region. Default priority, like the main stream.
    hipStream_t aux_stream = nullptr;
    yrh::DevBuf stack_spill_aux;
    hipEvent_t bounce_events[2] = {nullptr, nullptr};  // [0] shaded, [1] shadows resolved
    bool render_overlap = true;
    // Launch shape of the persistent traversal kernels; 0 = chosen per launch
    // (yrh_trace.hip PrepareTraceArgs: 8 LDS stack levels and 8 workgroups per
    // CU for flat trees, 16 levels and 5 workgroups for two-level ones).
    int blocks_per_cu = 0;
    int block_size = 256;
    int lds_levels = 0;  // traversal stack levels kept in LDS per lane
    // Ray-cast launches of at least kOrderMinRays rays claim their rays in
    // origin-cell order (yrh_order.hip) unless switched off.
    int trace_order = YRH_TRACE_ORDER_ORIGIN;
    yrh::DevBuf ray_order;          // int32 per ray: claim position -> ray index
    yrh::DevBuf ray_order_scratch;  // cell histograms and scan tile sums
    // Phase timing of ray-cast calls: events before the ordering pass, before
    // and after the traversal kernel of the last call.
    bool trace_timing = false;
    bool last_call_ordered = false;
    hipEvent_t phase_events[3] = {nullptr, nullptr, nullptr};
    // Device address of this GPU's scratch array for yrh_trace_closest_noalloc
    // (yrh_noalloc.hip), looked up once per device at context creation.
    void* noalloc_scratch = nullptr;
    yrh::RenderState* render = nullptr;  // surface/light data, film, path state (yrh_render.hip)
    yrh::GuideState* guide = nullptr;    // learned guiding proposal (yrh_guide.hip)
};

namespace yrh {

int Fail(yrh_context* ctx, int code, const std::string& message);
int CheckHip(yrh_context* ctx, hipError_t status, const char* what);
// Reads (and clears) the traversal stack overflow flag after synchronising the
// context stream: YRH_E_OVERFLOW when a traversal since the last check dropped
// stack entries.
int CheckStackOverflow(yrh_context* ctx);

// One launch of a traversal kernel over a ray queue (yrh_trace.hip).
struct TraceRequest {
    bool any = false;
    const float4* org = nullptr;
    const float4* dir = nullptr;
    int64_t n_host = 0;                  // ray count when n_device is null
    const uint32_t* n_device = nullptr;  // device-resident ray count (read by the kernel)
    int64_t capacity = 0;                // upper bound of *n_device
    int64_t expected = 0;                // typical live count (sizes the grid; 0 = capacity)
    uint32_t* work_counter = nullptr;    // zeroed ray counter(s) owned by the caller; null = context's
    int parts = 1;                       // caller's counters: 1, or trav::kParts ranges kPartStride words apart
    float4* hit = nullptr;
    int32_t* instance = nullptr;
    uint32_t* node_tests = nullptr;
    uint32_t* triangle_tests = nullptr;
    uint32_t* sphere_tests = nullptr;
    uint32_t* occluded = nullptr;
    unsigned long long* totals = nullptr;  // [0] node tests, [1] triangle tests, [2] sphere tests
    const int32_t* order = nullptr;        // claim position -> ray index (null = identity)
    bool aux = false;                      // on the context's second stream, with its spill region
    // Caller-provided stack spill region (never reallocated): the grid is
    // reduced until the spill of every lane fits in spill_bytes. Null: the
    // context's region, grown as needed.
    int2* spill = nullptr;
    size_t spill_bytes = 0;
};
int LaunchTrace(yrh_context* ctx, const TraceRequest& r);

// Looks up this device's scratch array for yrh_trace_closest_noalloc
// (yrh_noalloc.hip); called by yrh_create with the device current.
int InitNoAllocScratch(yrh_context* ctx);

// Order of n rays by the Morton code of their origin's cell in the scene box
// (yrh_order.hip), input order kept within a cell except among one wave's
// rays of mixed cells; the order lives in a context buffer that stays valid
// until the next call.
constexpr int64_t kOrderMinRays = 1 << 16;
int OrderRaysByOrigin(yrh_context* ctx, const float4* org, int64_t n, const int32_t** order);
// Durations of the last timed ray-cast call's phases (yrh_trace.hip).
int TracePhaseMs(yrh_context* ctx, float* order_ms, float* trace_ms);

// Render-side state (yrh_render.hip): shading copy of the geometry, taken at
// geometry upload; released with the context.
int UploadSurfaceGeometry(yrh_context* ctx, const yrh_geometry_desc* g);
void DestroyRenderState(yrh_context* ctx);

// Guiding state for one render pass (yrh_guide.hip): fills `view` (mode off
// when no proposal is configured), sizing the per-row buffers for `paths`
// rows and, when recording, `paths * max_depth` vertex records.
int GuidePrepareView(yrh_context* ctx, int64_t paths, int max_depth, yrg::GuideView* view);
void DestroyGuideState(yrh_context* ctx);

// Launches of the traversal kernels (yrh_trace.hip).
int LaunchTraceClosest(yrh_context* ctx, const yrh_ray_soa* rays, int64_t n, const yrh_hit_soa* hits);
int LaunchTraceAny(yrh_context* ctx, const yrh_ray_soa* rays, int64_t n, uint32_t* occluded,
                   uint32_t* node_tests, uint32_t* triangle_tests);

}  // namespace yrh
