// Context lifetime, device memory helpers and acceleration-structure upload
// for libyrh (MI355X / gfx950).
#include "yrh_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>

namespace yrh {

int Fail(yrh_context* ctx, int code, const std::string& message) {
    if (ctx != nullptr) ctx->error = message;
    return code;
}

int CheckHip(yrh_context* ctx, hipError_t status, const char* what) {
    if (status == hipSuccess) return YRH_OK;
    return Fail(ctx, YRH_E_HIP, std::string(what) + ": " + hipGetErrorString(status));
}

int CheckStackOverflow(yrh_context* ctx) {
    if (ctx->work_counter.ptr == nullptr) return YRH_OK;
    uint32_t flag = 0;
    const uint32_t* word = ctx->work_counter.as<uint32_t>() + kOverflowWord;
    int rc = CheckHip(ctx, hipMemcpyAsync(&flag, word, sizeof(flag), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    if (rc == YRH_OK) rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (rc != YRH_OK || flag == 0) return rc;
    (void)hipMemsetAsync(const_cast<uint32_t*>(word), 0, sizeof(flag), ctx->stream);
    return Fail(ctx, YRH_E_OVERFLOW,
                "traversal stack overflow: a BVH is deeper than the max_depth it was uploaded with; results of the "
                "traversals since the last check are incomplete");
}

namespace {

template <typename T>
int Upload(yrh_context* ctx, DevBuf& buf, const std::vector<T>& host) {
    buf.reset();
    const size_t bytes = std::max<size_t>(host.size() * sizeof(T), 16);
    int rc = CheckHip(ctx, hipMalloc(&buf.ptr, bytes), "hipMalloc");
    if (rc != YRH_OK) return rc;
    buf.bytes = bytes;
    if (!host.empty()) {
        rc = yrh_copy_to_device(ctx, buf.ptr, host.data(), host.size() * sizeof(T));
    }
    return rc;
}

struct PackedBvh {
    std::vector<DevNode> nodes;
    std::vector<float4> slots;
    std::vector<int32_t> leaf_count;
    int32_t root_code = 0;
    float root_lo[3] = {0, 0, 0};
    float root_hi[3] = {0, 0, 0};
    bool empty = true;
};

// Checks one reference BVH's arrays (child and leaf ranges) and that its
// declared max_depth covers the tree: the traversal stack is sized from the
// declared depth, so an under-declared one would drop entries. Depth counts
// the root as 1, as the reference's builder does (src/accel/bvh_builder.cpp:226-244,
// bvh_serial_builder.cpp:44-55). Children follow their parent in the node
// array (left = i + 1, right > i), so one pass in index order finds the
// deepest node.
bool ValidateBvh(const yrh_bvh_desc& bvh, std::string& error) {
    if (bvh.node_count < 0 || bvh.primitive_index_count < 0 || bvh.primitive_count < 0) {
        error = "negative BVH array size";
        return false;
    }
    if (bvh.node_count > 0 && (bvh.nodes == nullptr || bvh.primitive_indices == nullptr)) {
        error = "BVH arrays missing";
        return false;
    }
    std::vector<int32_t> depth(static_cast<size_t>(bvh.node_count), 0);
    int32_t deepest = 0;
    if (bvh.node_count > 0) depth[0] = 1;
    for (int32_t i = 0; i < bvh.node_count; ++i) {
        const yrh_bvh_node& n = bvh.nodes[i];
        if (n.primitive_count > 0) {
            if (n.payload_offset < 0 ||
                static_cast<int64_t>(n.payload_offset) + n.primitive_count > bvh.primitive_index_count) {
                error = "BVH leaf primitive range out of bounds";
                return false;
            }
        } else {
            const int32_t left = i + 1;
            const int32_t right = n.payload_offset;
            if (left >= bvh.node_count || right <= i || right >= bvh.node_count) {
                error = "BVH interior node child out of bounds";
                return false;
            }
            const int32_t d = depth[static_cast<size_t>(i)] + 1;
            depth[static_cast<size_t>(left)] = std::max(depth[static_cast<size_t>(left)], d);
            depth[static_cast<size_t>(right)] = std::max(depth[static_cast<size_t>(right)], d);
        }
        deepest = std::max(deepest, depth[static_cast<size_t>(i)]);
    }
    if (bvh.max_depth < deepest) {
        error = "BVH max_depth " + std::to_string(bvh.max_depth) + " is below the tree's depth " +
                std::to_string(deepest) + " (the traversal stack is sized from it)";
        return false;
    }
    return true;
}

// Descriptor checks of yrh_upload_geometry that need no device: array sizes
// and presence, no spheres, and every BVH (flat, each BLAS, the TLAS) valid
// with a declared max_depth that covers it.
int ValidateGeometry(const yrh_geometry_desc* g, std::string& error) {
    if (g == nullptr) {
        error = "null argument";
        return YRH_E_INVALID_ARGUMENT;
    }
    if (g->sphere_count < 0 || (g->sphere_count > 0 && g->spheres == nullptr)) {
        error = "sphere array missing";
        return YRH_E_INVALID_ARGUMENT;
    }
    if (g->vertex_count < 0 || g->index_count < 0 || g->primitive_count < 0) {
        error = "negative geometry array size";
        return YRH_E_INVALID_ARGUMENT;
    }
    if ((g->vertex_count > 0 && g->vertices == nullptr) || (g->index_count > 0 && g->indices == nullptr) ||
        (g->primitive_count > 0 && g->primitives == nullptr)) {
        error = "geometry arrays missing";
        return YRH_E_INVALID_ARGUMENT;
    }
    if (g->accel_kind == YRH_ACCEL_FLAT) {
        if (!ValidateBvh(g->flat, error)) return YRH_E_INVALID_ARGUMENT;
    } else if (g->accel_kind == YRH_ACCEL_TWO_LEVEL) {
        if (g->blas_count < 0 || g->instance_count < 0 || (g->blas_count > 0 && g->blases == nullptr) ||
            (g->instance_count > 0 && g->instances == nullptr)) {
            error = "two-level arrays missing";
            return YRH_E_INVALID_ARGUMENT;
        }
        for (int32_t b = 0; b < g->blas_count; ++b) {
            if (!ValidateBvh(g->blases[b], error)) return YRH_E_INVALID_ARGUMENT;
        }
        if (!ValidateBvh(g->tlas, error)) return YRH_E_INVALID_ARGUMENT;
    } else {
        error = "unknown acceleration kind";
        return YRH_E_INVALID_ARGUMENT;
    }
    return YRH_OK;
}

// Validates one reference BVH and appends its repacked form to `out`, rebasing
// node indices by out.nodes.size() and slots by out.slots.size() / 3.
// `slot_payload(slot, prim)` writes the three float4 of a slot.
template <typename SlotFn>
bool PackBvh(const yrh_bvh_desc& bvh, PackedBvh& out, std::string& error, SlotFn&& slot_payload) {
    const int32_t node_base = static_cast<int32_t>(out.nodes.size());
    const int64_t slot_base = static_cast<int64_t>(out.leaf_count.size());
    if (!ValidateBvh(bvh, error)) return false;
    if (slot_base + bvh.primitive_index_count > kMaxSlot) {
        error = "too many BVH primitive slots for the device layout";
        return false;
    }
    const auto code_of = [&](int32_t node) -> int32_t {
        const yrh_bvh_node& n = bvh.nodes[node];
        if (n.primitive_count > 0) {
            return PackLeaf(static_cast<int32_t>(slot_base) + n.payload_offset, n.primitive_count);
        }
        return node_base + node;
    };
    out.nodes.resize(out.nodes.size() + static_cast<size_t>(bvh.node_count));
    for (int32_t i = 0; i < bvh.node_count; ++i) {
        const yrh_bvh_node& n = bvh.nodes[i];
        DevNode& rec = out.nodes[static_cast<size_t>(node_base + i)];
        if (n.primitive_count > 0) {
            rec.a = make_float4(0, 0, 0, 0);
            rec.b = rec.a;
            rec.c = rec.a;
            rec.d = make_int4(code_of(i), 0, 0, 0);
            continue;
        }
        const yrh_bvh_node& l = bvh.nodes[i + 1];
        const yrh_bvh_node& r = bvh.nodes[n.payload_offset];
        rec.a = make_float4(l.bounds_min[0], l.bounds_min[1], l.bounds_min[2], l.bounds_max[0]);
        rec.b = make_float4(l.bounds_max[1], l.bounds_max[2], r.bounds_min[0], r.bounds_min[1]);
        rec.c = make_float4(r.bounds_min[2], r.bounds_max[0], r.bounds_max[1], r.bounds_max[2]);
        rec.d = make_int4(code_of(i + 1), code_of(n.payload_offset), 0, 0);
    }
    out.slots.resize(out.slots.size() + 3 * static_cast<size_t>(bvh.primitive_index_count));
    out.leaf_count.resize(out.leaf_count.size() + static_cast<size_t>(bvh.primitive_index_count), 0);
    for (int32_t s = 0; s < bvh.primitive_index_count; ++s) {
        const int32_t p = bvh.primitive_indices[s];
        if (p < 0 || p >= bvh.primitive_count) {
            error = "BVH primitive index out of bounds";
            return false;
        }
        float4* dst = &out.slots[3 * static_cast<size_t>(slot_base + s)];
        if (!slot_payload(bvh.primitives[p], dst, error)) return false;
    }
    for (int32_t i = 0; i < bvh.node_count; ++i) {
        const yrh_bvh_node& n = bvh.nodes[i];
        if (n.primitive_count > kLeafCountEscape) {
            out.leaf_count[static_cast<size_t>(slot_base + n.payload_offset)] = n.primitive_count;
        }
    }
    if (bvh.node_count > 0) {
        out.empty = false;
        out.root_code = code_of(0);
        for (int k = 0; k < 3; ++k) {
            out.root_lo[k] = bvh.nodes[0].bounds_min[k];
            out.root_hi[k] = bvh.nodes[0].bounds_max[k];
        }
    }
    return true;
}

// Triangle slot: (p0, flat id bits) (e1, mesh primitive bits) (e2, local triangle bits).
// A reference triangle that its traversal would skip (bad handle or index,
// bvh_traversal.cpp:59-76) becomes a zero-area slot, which the |det| < 1e-8
// test rejects, so it is counted but never hit, as in the reference.
struct TriangleSlotWriter {
    const yrh_geometry_desc* g;
    bool operator()(const yrh_bvh_primitive& prim, float4* dst, std::string& error) const {
        if (prim.kind == 1) {
            // Sphere slot: (centre, -2 - sphere) (radius, 0, 0, -1) (0, 0, 0, -1).
            // A bad handle (skipped by the reference, bvh_traversal.cpp:130-133)
            // gets a NaN centre: counted, never hit.
            const bool ok = prim.sphere >= 0 && prim.sphere < g->sphere_count;
            const float nan = std::numeric_limits<float>::quiet_NaN();
            const yrh_sphere* sp = ok ? &g->spheres[prim.sphere] : nullptr;
            const int32_t id = -2 - (ok ? prim.sphere : 0);
            float fid, fneg;
            const int32_t neg = -1;
            std::memcpy(&fid, &id, 4);
            std::memcpy(&fneg, &neg, 4);
            dst[0] = ok ? make_float4(sp->center[0], sp->center[1], sp->center[2], fid) : make_float4(nan, nan, nan, fid);
            dst[1] = make_float4(ok ? sp->radius : nan, 0.0f, 0.0f, fneg);
            dst[2] = make_float4(0.0f, 0.0f, 0.0f, fneg);
            return true;
        }
        if (prim.kind != 0) {
            error = "unknown BVH primitive kind";
            return false;
        }
        float p[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        bool valid = prim.mesh_primitive >= 0 && prim.mesh_primitive < g->primitive_count &&
                     prim.local_triangle >= 0;
        if (valid) {
            const yrh_primitive& mp = g->primitives[prim.mesh_primitive];
            const uint64_t base = static_cast<uint64_t>(mp.first_index) +
                                  static_cast<uint64_t>(prim.local_triangle) * 3u;
            valid = base + 2 < static_cast<uint64_t>(g->index_count);
            if (valid) {
                for (int k = 0; k < 3; ++k) {
                    const uint32_t vi = g->indices[base + static_cast<uint64_t>(k)];
                    if (static_cast<int64_t>(vi) >= g->vertex_count) {
                        valid = false;
                        break;
                    }
                    for (int c = 0; c < 3; ++c) p[k][c] = g->vertices[vi].position[c];
                }
            }
        }
        float e1[3] = {0, 0, 0};
        float e2[3] = {0, 0, 0};
        if (valid) {
            for (int c = 0; c < 3; ++c) {
                e1[c] = p[1][c] - p[0][c];
                e2[c] = p[2][c] - p[0][c];
            }
        }
        int32_t id = prim.flat_triangle_index;
        int32_t mesh = prim.mesh_primitive;
        int32_t local = prim.local_triangle;
        float fid, fmesh, flocal;
        std::memcpy(&fid, &id, 4);
        std::memcpy(&fmesh, &mesh, 4);
        std::memcpy(&flocal, &local, 4);
        dst[0] = make_float4(p[0][0], p[0][1], p[0][2], fid);
        dst[1] = make_float4(e1[0], e1[1], e1[2], fmesh);
        dst[2] = make_float4(e2[0], e2[1], e2[2], flocal);
        return true;
    }
};

}  // namespace
}  // namespace yrh

using namespace yrh;

extern "C" {

int yrh_create(int device, yrh_context** out) {
    if (out == nullptr) return YRH_E_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return YRH_E_HIP;
    if (device < 0 || device >= count) return YRH_E_INVALID_ARGUMENT;
    auto* ctx = new (std::nothrow) yrh_context();
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return YRH_E_HIP;
    }
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->compute_units = prop.multiProcessorCount;
        std::strncpy(ctx->device_name, prop.gcnArchName, sizeof(ctx->device_name) - 1);
    }
    if (ctx->compute_units <= 0) ctx->compute_units = 256;
    if (hipMalloc(&ctx->work_counter.ptr, kWorkCounterBytes) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return YRH_E_HIP;
    }
    ctx->work_counter.bytes = kWorkCounterBytes;
    if (hipMemset(ctx->work_counter.ptr, 0, kWorkCounterBytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        InitNoAllocScratch(ctx) != YRH_OK) {
        ctx->work_counter.reset();
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return YRH_E_HIP;
    }
    *out = ctx;
    return YRH_OK;
}

void yrh_destroy(yrh_context* ctx) {
    if (ctx == nullptr) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->aux_stream != nullptr) (void)hipStreamSynchronize(ctx->aux_stream);
    DestroyRenderState(ctx);
    DestroyGuideState(ctx);
    ctx->geo = DevGeometry{};
    ctx->work_counter.reset();
    ctx->stack_spill.reset();
    ctx->stack_spill_aux.reset();
    for (hipEvent_t e : ctx->phase_events) {
        if (e != nullptr) (void)hipEventDestroy(e);
    }
    for (hipEvent_t e : ctx->bounce_events) {
        if (e != nullptr) (void)hipEventDestroy(e);
    }
    if (ctx->aux_stream != nullptr) (void)hipStreamDestroy(ctx->aux_stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* yrh_last_error(const yrh_context* ctx) {
    return ctx == nullptr ? "null context" : ctx->error.c_str();
}

int yrh_synchronize(yrh_context* ctx) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    const int rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return rc != YRH_OK ? rc : CheckStackOverflow(ctx);
}

void* yrh_stream(yrh_context* ctx) { return ctx == nullptr ? nullptr : static_cast<void*>(ctx->stream); }

int yrh_device_info(yrh_context* ctx, char* name, int name_capacity, int* compute_units) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (name != nullptr && name_capacity > 0) {
        std::strncpy(name, ctx->device_name, static_cast<size_t>(name_capacity - 1));
        name[name_capacity - 1] = '\0';
    }
    if (compute_units != nullptr) *compute_units = ctx->compute_units;
    return YRH_OK;
}

int yrh_device_alloc(yrh_context* ctx, size_t bytes, void** out) {
    if (ctx == nullptr || out == nullptr) return YRH_E_INVALID_ARGUMENT;
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    return CheckHip(ctx, hipMalloc(out, std::max<size_t>(bytes, 16)), "hipMalloc");
}

int yrh_device_free(yrh_context* ctx, void* ptr) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (ptr == nullptr) return YRH_OK;
    return CheckHip(ctx, hipFree(ptr), "hipFree");
}

int yrh_copy_to_device(yrh_context* ctx, void* dst, const void* src, size_t bytes) {
    if (ctx == nullptr || (bytes > 0 && (dst == nullptr || src == nullptr))) return YRH_E_INVALID_ARGUMENT;
    if (bytes == 0) return YRH_OK;
    int rc = CheckHip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpyAsync H2D");
    if (rc != YRH_OK) return rc;
    return CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int yrh_copy_to_host(yrh_context* ctx, void* dst, const void* src, size_t bytes) {
    if (ctx == nullptr || (bytes > 0 && (dst == nullptr || src == nullptr))) return YRH_E_INVALID_ARGUMENT;
    if (bytes == 0) return YRH_OK;
    int rc = CheckHip(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync D2H");
    if (rc != YRH_OK) return rc;
    return CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int yrh_set_trace_config(yrh_context* ctx, int blocks_per_cu, int block_size) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (blocks_per_cu < -1 || blocks_per_cu > 64) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "blocks_per_cu must be in [-1, 64]");
    if (block_size != 0 && block_size != 64 && block_size != 128 && block_size != 256) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "block_size must be 64, 128 or 256");
    }
    if (blocks_per_cu != 0) ctx->blocks_per_cu = blocks_per_cu < 0 ? 0 : blocks_per_cu;
    if (block_size > 0) ctx->block_size = block_size;
    return YRH_OK;
}

int yrh_set_trace_lds_levels(yrh_context* ctx, int levels) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (levels < 0 || levels > 16) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "lds levels must be in [0, 16]");
    ctx->lds_levels = levels;
    return YRH_OK;
}

int yrh_set_trace_order(yrh_context* ctx, int order) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (order != YRH_TRACE_ORDER_INPUT && order != YRH_TRACE_ORDER_ORIGIN) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "trace order must be YRH_TRACE_ORDER_INPUT or YRH_TRACE_ORDER_ORIGIN");
    }
    ctx->trace_order = order;
    return YRH_OK;
}

int yrh_set_trace_timing(yrh_context* ctx, int enable) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    ctx->trace_timing = enable != 0;
    return YRH_OK;
}

int yrh_set_render_overlap(yrh_context* ctx, int enable) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    ctx->render_overlap = enable != 0;
    return YRH_OK;
}

int yrh_trace_phase_ms(yrh_context* ctx, float* order_ms, float* trace_ms) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (order_ms == nullptr || trace_ms == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null output");
    (void)hipSetDevice(ctx->device);
    return TracePhaseMs(ctx, order_ms, trace_ms);
}

int yrh_event_create(yrh_context* ctx, void** event) {
    if (ctx == nullptr || event == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null argument");
    hipEvent_t e = nullptr;
    int rc = CheckHip(ctx, hipEventCreate(&e), "hipEventCreate");
    *event = e;
    return rc;
}

int yrh_event_destroy(yrh_context* ctx, void* event) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (event == nullptr) return YRH_OK;
    return CheckHip(ctx, hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
}

int yrh_event_record(yrh_context* ctx, void* event) {
    if (ctx == nullptr || event == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null argument");
    return CheckHip(ctx, hipEventRecord(static_cast<hipEvent_t>(event), ctx->stream), "hipEventRecord");
}

int yrh_event_elapsed_ms(yrh_context* ctx, void* start, void* stop, float* ms) {
    if (ctx == nullptr || start == nullptr || stop == nullptr || ms == nullptr) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null argument");
    }
    int rc = CheckHip(ctx, hipEventSynchronize(static_cast<hipEvent_t>(stop)), "hipEventSynchronize");
    if (rc != YRH_OK) return rc;
    return CheckHip(ctx, hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)),
                    "hipEventElapsedTime");
}

int yrh_validate_geometry(const yrh_geometry_desc* g, char* error, int error_capacity) {
    std::string msg;
    const int rc = ValidateGeometry(g, msg);
    if (rc != YRH_OK && error != nullptr && error_capacity > 0) {
        std::strncpy(error, msg.c_str(), static_cast<size_t>(error_capacity - 1));
        error[error_capacity - 1] = '\0';
    }
    return rc;
}

int yrh_upload_geometry(yrh_context* ctx, const yrh_geometry_desc* g) {
    const StageRange stage_range("yrh_upload_geometry");
    if (ctx == nullptr || g == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null argument");
    {
        std::string msg;
        const int vrc = ValidateGeometry(g, msg);
        if (vrc != YRH_OK) return Fail(ctx, vrc, msg);
    }
    (void)hipSetDevice(ctx->device);
    int rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (rc != YRH_OK) return rc;
    ctx->geo = DevGeometry{};
    DevGeometry& geo = ctx->geo;
    std::string error;
    TriangleSlotWriter writer{g};
    PackedBvh packed;

    if (g->accel_kind == YRH_ACCEL_FLAT) {
        if (!PackBvh(g->flat, packed, error, writer)) return Fail(ctx, YRH_E_INVALID_ARGUMENT, error);
        geo.stack_capacity = std::max(1, g->flat.max_depth) + 2;
    } else if (g->accel_kind == YRH_ACCEL_TWO_LEVEL) {
        if (g->blas_count < 0 || g->instance_count < 0 || (g->blas_count > 0 && g->blases == nullptr) ||
            (g->instance_count > 0 && g->instances == nullptr)) {
            return Fail(ctx, YRH_E_INVALID_ARGUMENT, "two-level arrays missing");
        }
        std::vector<DevBlasRoot> roots(static_cast<size_t>(g->blas_count));
        int32_t blas_depth = 1;
        for (int32_t b = 0; b < g->blas_count; ++b) {
            PackedBvh one;
            const size_t node_base = packed.nodes.size();
            const size_t slot_base = packed.leaf_count.size();
            // Pack into the shared arrays by pre-seeding sizes.
            one.nodes.swap(packed.nodes);
            one.slots.swap(packed.slots);
            one.leaf_count.swap(packed.leaf_count);
            const bool ok = PackBvh(g->blases[b], one, error, writer);
            packed.nodes.swap(one.nodes);
            packed.slots.swap(one.slots);
            packed.leaf_count.swap(one.leaf_count);
            if (!ok) return Fail(ctx, YRH_E_INVALID_ARGUMENT, error);
            if (one.empty) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "empty BLAS");
            (void)node_base;
            (void)slot_base;
            int32_t code = one.root_code;
            float fcode;
            std::memcpy(&fcode, &code, 4);
            roots[static_cast<size_t>(b)].lo = make_float4(one.root_lo[0], one.root_lo[1], one.root_lo[2], fcode);
            roots[static_cast<size_t>(b)].hi = make_float4(one.root_hi[0], one.root_hi[1], one.root_hi[2], 0.0f);
            blas_depth = std::max(blas_depth, g->blases[b].max_depth);
        }
        // TLAS: slots resolve primitive_indices -> tlas primitive -> instance.
        const yrh_bvh_desc& t = g->tlas;
        if (t.primitive_index_count > 0 && (t.primitive_indices == nullptr || g->tlas_primitives == nullptr)) {
            return Fail(ctx, YRH_E_INVALID_ARGUMENT, "TLAS arrays missing");
        }
        PackedBvh tl;
        std::vector<int32_t> slot_instance;
        yrh_bvh_desc tdesc = t;
        tdesc.primitives = nullptr;
        tdesc.primitive_count = g->tlas_primitive_count;
        // Reuse PackBvh for nodes/codes with a slot writer that records instances.
        // Per TLAS slot: an instance index (>= 0) or -2 - sphere index.
        std::vector<int32_t> resolved(static_cast<size_t>(std::max(0, t.primitive_index_count)), -1);
        for (int32_t s = 0; s < t.primitive_index_count; ++s) {
            const int32_t p = t.primitive_indices[s];
            if (p < 0 || p >= g->tlas_primitive_count) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "TLAS primitive index out of bounds");
            const yrh_tlas_primitive& tp = g->tlas_primitives[p];
            if (tp.kind == 1) {
                if (tp.sphere < 0 || tp.sphere >= g->sphere_count) {
                    return Fail(ctx, YRH_E_INVALID_ARGUMENT, "TLAS sphere out of bounds");
                }
                resolved[static_cast<size_t>(s)] = -2 - tp.sphere;
                continue;
            }
            if (tp.kind != 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "unknown TLAS primitive kind");
            if (tp.instance < 0 || tp.instance >= g->instance_count) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "TLAS instance out of bounds");
            const int32_t bi = g->instances[tp.instance].blas_index;
            if (bi < 0 || bi >= g->blas_count) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "instance BLAS out of bounds");
            resolved[static_cast<size_t>(s)] = tp.instance;
        }
        std::vector<yrh_bvh_primitive> dummy(static_cast<size_t>(std::max(1, g->tlas_primitive_count)));
        tdesc.primitives = dummy.data();
        tdesc.primitive_count = static_cast<int32_t>(dummy.size());
        const auto no_slot = [](const yrh_bvh_primitive&, float4* dst, std::string&) {
            dst[0] = dst[1] = dst[2] = make_float4(0, 0, 0, 0);
            return true;
        };
        if (!PackBvh(tdesc, tl, error, no_slot)) return Fail(ctx, YRH_E_INVALID_ARGUMENT, error);
        std::vector<DevInstanceSlot> slots(resolved.size());
        for (size_t s = 0; s < resolved.size(); ++s) {
            const int32_t i = resolved[s];
            if (i <= -2) {
                // TLAS sphere slot: r0 = (centre, radius), hi.w = -2 - sphere.
                const yrh_sphere& sp = g->spheres[-2 - i];
                float fi;
                std::memcpy(&fi, &i, 4);
                DevInstanceSlot& d = slots[s];
                d.r0 = make_float4(sp.center[0], sp.center[1], sp.center[2], sp.radius);
                d.r1 = d.r2 = d.lo = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                d.hi = make_float4(0.0f, 0.0f, 0.0f, fi);
                continue;
            }
            const float* m = g->instances[i].world_to_object;
            const DevBlasRoot& root = roots[static_cast<size_t>(g->instances[i].blas_index)];
            DevInstanceSlot& d = slots[s];
            d.r0 = make_float4(m[0], m[4], m[8], m[12]);
            d.r1 = make_float4(m[1], m[5], m[9], m[13]);
            d.r2 = make_float4(m[2], m[6], m[10], m[14]);
            float fi;
            std::memcpy(&fi, &i, 4);
            d.lo = root.lo;
            d.hi = make_float4(root.hi.x, root.hi.y, root.hi.z, fi);
        }
        geo.instance_count = g->instance_count;
        geo.blas_count = g->blas_count;
        geo.root_code = tl.root_code;
        std::memcpy(geo.root_lo, tl.root_lo, sizeof(geo.root_lo));
        std::memcpy(geo.root_hi, tl.root_hi, sizeof(geo.root_hi));
        geo.empty = tl.empty;
        geo.stack_capacity = std::max(1, t.max_depth) + blas_depth + 3;
        if ((rc = Upload(ctx, geo.tlas_nodes, tl.nodes)) != YRH_OK) return rc;
        if ((rc = Upload(ctx, geo.tlas_leaf_count, tl.leaf_count)) != YRH_OK) return rc;
        if ((rc = Upload(ctx, geo.instance_slots, slots)) != YRH_OK) return rc;
    } else {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "unknown acceleration kind");
    }

    if (g->accel_kind == YRH_ACCEL_FLAT) {
        geo.root_code = packed.root_code;
        std::memcpy(geo.root_lo, packed.root_lo, sizeof(geo.root_lo));
        std::memcpy(geo.root_hi, packed.root_hi, sizeof(geo.root_hi));
        geo.empty = packed.empty;
    }
    geo.accel_kind = g->accel_kind;
    // Sphere slots exist if any flat-BVH or TLAS primitive is a sphere (an
    // invalid sphere handle still counts a sphere test, as in the reference).
    geo.has_spheres = g->sphere_count > 0;
    if (g->accel_kind == YRH_ACCEL_FLAT) {
        for (int32_t k = 0; !geo.has_spheres && k < g->flat.primitive_count; ++k) {
            geo.has_spheres = g->flat.primitives[k].kind != 0;
        }
    } else {
        for (int32_t k = 0; !geo.has_spheres && k < g->tlas_primitive_count; ++k) {
            geo.has_spheres = g->tlas_primitives[k].kind != 0;
        }
    }
    geo.slot_count = static_cast<int64_t>(packed.leaf_count.size());
    if ((rc = Upload(ctx, geo.nodes, packed.nodes)) != YRH_OK) return rc;
    if ((rc = Upload(ctx, geo.slots, packed.slots)) != YRH_OK) return rc;
    if ((rc = Upload(ctx, geo.leaf_count, packed.leaf_count)) != YRH_OK) return rc;
    if ((rc = UploadSurfaceGeometry(ctx, g)) != YRH_OK) return rc;
    geo.loaded = true;
    return YRH_OK;
}

int yrh_trace_closest(yrh_context* ctx, const yrh_ray_soa* rays, int64_t n, const yrh_hit_soa* hits) {
    const StageRange stage_range("yrh_trace_closest");
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (rays == nullptr || hits == nullptr || n < 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null rays/hits or negative count");
    if (n > 0 && (rays->org_tmin == nullptr || rays->dir_tmax == nullptr || hits->hit == nullptr)) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "ray or hit array missing");
    }
    if (n > INT32_MAX) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "at most 2^31-1 rays per call");
    if (!ctx->geo.loaded) return Fail(ctx, YRH_E_NO_SCENE, "no geometry uploaded");
    if (n == 0) return YRH_OK;
    return LaunchTraceClosest(ctx, rays, n, hits);
}

int yrh_trace_any(yrh_context* ctx, const yrh_ray_soa* rays, int64_t n, uint32_t* occluded,
                  uint32_t* node_tests, uint32_t* triangle_tests) {
    const StageRange stage_range("yrh_trace_any");
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (rays == nullptr || n < 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null rays or negative count");
    if (n > 0 && (rays->org_tmin == nullptr || rays->dir_tmax == nullptr || occluded == nullptr)) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "ray or output array missing");
    }
    if (n > INT32_MAX) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "at most 2^31-1 rays per call");
    if (!ctx->geo.loaded) return Fail(ctx, YRH_E_NO_SCENE, "no geometry uploaded");
    if (n == 0) return YRH_OK;
    return LaunchTraceAny(ctx, rays, n, occluded, node_tests, triangle_tests);
}

}  // extern "C"
