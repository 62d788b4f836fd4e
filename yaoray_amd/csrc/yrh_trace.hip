// Closest-hit and any-hit ray-cast kernels for MI355X (gfx950) over a ray
// queue: the traversal loops of yrh_traverse.h with the ray-array policy.
#include "yrh_traverse.h"

namespace yrh {
namespace {

using trav::kMaxBlock;
using trav::TraceArgs;

// 8 waves per SIMD (at most 64 VGPRs and 100 SGPRs): with 8 workgroups of 256
// per CU every workgroup of the persistent grid is resident at once.
template <bool kAny, bool kSpheres>
__global__ __launch_bounds__(kMaxBlock, 8) void TraceFlatKernel(TraceArgs a) {
    extern __shared__ int2 lds_stack[];
    const trav::RayArrayPolicy<kAny> pol{&a};
    trav::TraceFlatLoop<kAny, kSpheres>(a, pol, lds_stack);
}

// At most 96 VGPRs (5 waves per SIMD, the most that 16 LDS stack levels allow).
template <bool kAny, bool kSpheres>
__global__ __launch_bounds__(kMaxBlock, 5) void TraceTwoLevelKernel(TraceArgs a) {
    extern __shared__ int2 lds_stack[];
    const trav::RayArrayPolicy<kAny> pol{&a};
    trav::TraceTwoLevelLoop<kAny, kSpheres>(a, pol, lds_stack);
}

template <bool kAny, bool kSpheres>
void LaunchTraceKernel(bool two, int grid, int block, size_t lds, hipStream_t st, const TraceArgs& a) {
    if (two) {
        hipLaunchKernelGGL((TraceTwoLevelKernel<kAny, kSpheres>), dim3(grid), dim3(block), lds, st, a);
    } else {
        hipLaunchKernelGGL((TraceFlatKernel<kAny, kSpheres>), dim3(grid), dim3(block), lds, st, a);
    }
}

bool Aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace

size_t LdsBytes(const trav::TraceArgs& a, int block) {
    return static_cast<size_t>(a.lds_levels) * static_cast<size_t>(block) * sizeof(int2);
}

int PrepareTraceArgs(yrh_context* ctx, const TraceRequest& r, trav::TraceArgs& a, int& grid, int& block) {
    const DevGeometry& g = ctx->geo;
    block = ctx->block_size;
    // Launch shape: LDS stack levels per lane (flat trees are shallow enough
    // for 8; two-level stacks hold a TLAS and a BLAS path, so 16) and as many
    // workgroups per CU as that LDS allows (at most 8), unless set explicitly.
    const int lds_levels = ctx->lds_levels > 0 ? ctx->lds_levels : (g.accel_kind == YRH_ACCEL_TWO_LEVEL ? 16 : 8);
    const int per_cu = ctx->blocks_per_cu > 0
                           ? ctx->blocks_per_cu
                           : std::max(1, std::min(8, (160 * 1024) / (lds_levels * block * static_cast<int>(sizeof(int2)))));
    grid = ctx->compute_units * per_cu;
    const int64_t lanes = static_cast<int64_t>(grid) * block;
    // Small queues: keep at least 4 rays per lane (the persistent lanes refill
    // from the counter, so fewer, fuller workgroups drain faster), and never
    // more lanes than rays.
    const int64_t bound = r.n_device != nullptr ? (r.expected > 0 ? std::min(r.expected, r.capacity) : r.capacity)
                                                : r.n_host;
    const int64_t want = std::max<int64_t>(std::min<int64_t>(bound, static_cast<int64_t>(ctx->compute_units) * block),
                                           (bound + 3) / 4);
    if (lanes > want) grid = static_cast<int>(std::max<int64_t>(1, (want + block - 1) / block));
    const int32_t capacity = g.stack_capacity;
    const int64_t spill_levels = std::max<int64_t>(0, capacity - lds_levels);
    int2* spill_ptr = nullptr;
    if (r.spill != nullptr) {
        // A fixed region: fewer workgroups when every lane's spill would not fit
        // (persistent lanes claim rays until none are left, so any grid is correct).
        const int64_t per_group = spill_levels * block * static_cast<int64_t>(sizeof(int2));
        if (per_group > 0) {
            const int64_t fit = static_cast<int64_t>(r.spill_bytes) / per_group;
            if (fit < 1) {
                return Fail(ctx, YRH_E_UNSUPPORTED,
                            "BVH too deep for the scratch stack region (" + std::to_string(spill_levels) +
                                " spilled levels x " + std::to_string(block) + " lanes > " +
                                std::to_string(r.spill_bytes) + " bytes)");
            }
            grid = static_cast<int>(std::min<int64_t>(grid, fit));
        }
        spill_ptr = r.spill;
    }
    const int64_t total = static_cast<int64_t>(grid) * block;
    const size_t spill_bytes = static_cast<size_t>(std::max<int64_t>(1, spill_levels * total)) * sizeof(int2);
    DevBuf& spill = r.aux ? ctx->stack_spill_aux : ctx->stack_spill;
    if (spill_ptr == nullptr && spill.bytes < spill_bytes) {
        // Earlier launches on either stream may still use the old region.
        int rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        if (rc == YRH_OK && ctx->aux_stream != nullptr) {
            rc = CheckHip(ctx, hipStreamSynchronize(ctx->aux_stream), "hipStreamSynchronize");
        }
        if (rc != YRH_OK) return rc;
        spill.reset();
        rc = CheckHip(ctx, hipMalloc(&spill.ptr, spill_bytes), "hipMalloc stack spill");
        if (rc != YRH_OK) return rc;
        spill.bytes = spill_bytes;
    }
    a = TraceArgs{};
    a.org = r.org;
    a.dir = r.dir;
    a.n = static_cast<int32_t>(r.n_host);
    a.n_device = r.n_device;
    a.totals = r.totals;
    a.hit = r.hit;
    a.inst_out = r.instance;
    a.node_tests = r.node_tests;
    a.triangle_tests = r.triangle_tests;
    a.sphere_tests = r.sphere_tests;
    a.occluded = r.occluded;
    a.nodes = g.nodes.as<DevNode>();
    a.slots = g.slots.as<float4>();
    a.leaf_count = g.leaf_count.as<int32_t>();
    a.root_code = g.root_code;
    a.root_lo_x = g.root_lo[0];
    a.root_lo_y = g.root_lo[1];
    a.root_lo_z = g.root_lo[2];
    a.root_hi_x = g.root_hi[0];
    a.root_hi_y = g.root_hi[1];
    a.root_hi_z = g.root_hi[2];
    a.empty = g.empty ? 1 : 0;
    a.tnodes = g.tlas_nodes.as<DevNode>();
    a.tleaf_count = g.tlas_leaf_count.as<int32_t>();
    a.instance_slots = g.instance_slots.as<DevInstanceSlot>();
    uint32_t* words = ctx->work_counter.as<uint32_t>();
    if (r.work_counter != nullptr) {
        a.counter = r.work_counter;
        a.next_counter = nullptr;
        a.parts = r.parts;
    } else {
        // Two sets of per-range counters; each launch zeroes the other set.
        uint32_t* sets = words + kCounterSetBase;
        a.counter = sets + ctx->launch_parity * kCounterSetWords;
        a.next_counter = sets + (1 - ctx->launch_parity) * kCounterSetWords;
        a.parts = trav::kParts;
    }
    a.order = r.order;
    a.overflow = words + kOverflowWord;
    a.spill = spill_ptr != nullptr ? spill_ptr : spill.as<int2>();
    a.spill_stride = total;
    a.capacity = capacity;
    a.lds_levels = lds_levels;
    return YRH_OK;
}

int LaunchTrace(yrh_context* ctx, const TraceRequest& r) {
    if (!Aligned16(r.org) || !Aligned16(r.dir) || (r.hit != nullptr && !Aligned16(r.hit))) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "ray and hit arrays must be 16-byte aligned");
    }
    if (r.any ? r.occluded == nullptr : r.hit == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "missing output array");
    (void)hipSetDevice(ctx->device);
    TraceArgs a;
    int grid = 0, block = 0;
    int rc = PrepareTraceArgs(ctx, r, a, grid, block);
    if (rc != YRH_OK) return rc;
    const bool two = ctx->geo.accel_kind == YRH_ACCEL_TWO_LEVEL;
    const bool spheres = ctx->geo.has_spheres;
    hipStream_t st = r.aux ? ctx->aux_stream : ctx->stream;
    const size_t lds = LdsBytes(a, block);
    if (r.any) {
        if (spheres) {
            LaunchTraceKernel<true, true>(two, grid, block, lds, st, a);
        } else {
            LaunchTraceKernel<true, false>(two, grid, block, lds, st, a);
        }
    } else {
        if (spheres) {
            LaunchTraceKernel<false, true>(two, grid, block, lds, st, a);
        } else {
            LaunchTraceKernel<false, false>(two, grid, block, lds, st, a);
        }
    }
    rc = CheckHip(ctx, hipGetLastError(), r.any ? "trace any launch" : "trace closest launch");
    if (rc == YRH_OK && r.work_counter == nullptr) ctx->launch_parity = 1 - ctx->launch_parity;
    return rc;
}

// Ray-cast launches claim large batches in origin-cell order; with phase
// timing on, events bracket the ordering pass and the traversal kernel.
int LaunchRayCast(yrh_context* ctx, TraceRequest& r) {
    const bool order = ctx->trace_order == YRH_TRACE_ORDER_ORIGIN && r.n_host >= kOrderMinRays && Aligned16(r.org);
    int rc = YRH_OK;
    if (ctx->trace_timing) {
        for (hipEvent_t& e : ctx->phase_events) {
            if (e == nullptr && (rc = CheckHip(ctx, hipEventCreate(&e), "hipEventCreate")) != YRH_OK) return rc;
        }
        if ((rc = CheckHip(ctx, hipEventRecord(ctx->phase_events[0], ctx->stream), "hipEventRecord")) != YRH_OK) return rc;
    }
    if (order && (rc = OrderRaysByOrigin(ctx, r.org, r.n_host, &r.order)) != YRH_OK) return rc;
    if (ctx->trace_timing) {
        if ((rc = CheckHip(ctx, hipEventRecord(ctx->phase_events[1], ctx->stream), "hipEventRecord")) != YRH_OK) return rc;
    }
    if ((rc = LaunchTrace(ctx, r)) != YRH_OK) return rc;
    if (ctx->trace_timing) {
        if ((rc = CheckHip(ctx, hipEventRecord(ctx->phase_events[2], ctx->stream), "hipEventRecord")) != YRH_OK) return rc;
        ctx->last_call_ordered = order;
    }
    return YRH_OK;
}

int LaunchTraceClosest(yrh_context* ctx, const yrh_ray_soa* rays, int64_t n, const yrh_hit_soa* hits) {
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
    return LaunchRayCast(ctx, r);
}

int LaunchTraceAny(yrh_context* ctx, const yrh_ray_soa* rays, int64_t n, uint32_t* occluded, uint32_t* node_tests,
                   uint32_t* triangle_tests) {
    TraceRequest r;
    r.any = true;
    r.org = reinterpret_cast<const float4*>(rays->org_tmin);
    r.dir = reinterpret_cast<const float4*>(rays->dir_tmax);
    r.n_host = n;
    r.occluded = occluded;
    r.node_tests = node_tests;
    r.triangle_tests = triangle_tests;
    return LaunchRayCast(ctx, r);
}

int TracePhaseMs(yrh_context* ctx, float* order_ms, float* trace_ms) {
    if (ctx->phase_events[2] == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "no timed ray-cast call yet");
    int rc = CheckHip(ctx, hipEventSynchronize(ctx->phase_events[2]), "hipEventSynchronize");
    if (rc != YRH_OK) return rc;
    float a = 0.0f, b = 0.0f;
    if ((rc = CheckHip(ctx, hipEventElapsedTime(&a, ctx->phase_events[0], ctx->phase_events[1]), "hipEventElapsedTime")) != YRH_OK) return rc;
    if ((rc = CheckHip(ctx, hipEventElapsedTime(&b, ctx->phase_events[1], ctx->phase_events[2]), "hipEventElapsedTime")) != YRH_OK) return rc;
    *order_ms = ctx->last_call_ordered ? a : 0.0f;
    *trace_ms = b;
    return YRH_OK;
}

}  // namespace yrh
