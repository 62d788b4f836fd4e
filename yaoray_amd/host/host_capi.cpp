// C entry points of the host library (scene IR, builder, workloads, upload).
#include "yaoray_host.h"

#include <chrono>
#include <cstddef>
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "scenes.hpp"
#include "yr_ir.hpp"

// The reference IR arrays are handed to libyrh as-is: the C structs in yrh.h
// must be byte-compatible with the yr:: types.
static_assert(sizeof(yr::RenderVertex) == sizeof(yrh_vertex));
static_assert(sizeof(yr::RenderPrimitive) == sizeof(yrh_primitive));
static_assert(offsetof(yr::RenderPrimitive, material_index) == offsetof(yrh_primitive, material_index));
static_assert(offsetof(yr::RenderPrimitive, has_normals) == offsetof(yrh_primitive, has_normals));
static_assert(sizeof(yr::RenderBvhNode) == sizeof(yrh_bvh_node));
static_assert(offsetof(yr::RenderBvhNode, payload_offset) == offsetof(yrh_bvh_node, payload_offset));
static_assert(sizeof(yr::RenderBvhPrimitive) == sizeof(yrh_bvh_primitive));
static_assert(sizeof(yr::PreparedMeshInstance) == sizeof(yrh_instance));
static_assert(offsetof(yr::PreparedMeshInstance, world_to_object) == offsetof(yrh_instance, world_to_object));
static_assert(offsetof(yr::PreparedMeshInstance, world_bounds) == offsetof(yrh_instance, world_bounds_min));
static_assert(sizeof(yr::RenderTlasPrimitive) == sizeof(yrh_tlas_primitive));

struct yrhost_scene {
    yr::RenderSceneIR ir;
    yr::RenderAcceleration accel;
    bool accel_built = false;
    std::string error;
    std::vector<yrh_bvh_desc> blas_desc;
};

namespace {

int SetError(yrhost_scene* s, const std::string& msg) {
    if (s != nullptr) s->error = msg;
    return YRH_E_INVALID_ARGUMENT;
}

yrh_bvh_desc Describe(const yr::RenderBvh& b) {
    yrh_bvh_desc d{};
    d.nodes = reinterpret_cast<const yrh_bvh_node*>(b.nodes.data());
    d.node_count = static_cast<int32_t>(b.nodes.size());
    d.primitive_indices = b.primitive_indices.data();
    d.primitive_index_count = static_cast<int32_t>(b.primitive_indices.size());
    d.primitives = reinterpret_cast<const yrh_bvh_primitive*>(b.primitives.data());
    d.primitive_count = static_cast<int32_t>(b.primitives.size());
    d.max_depth = b.max_depth;
    return d;
}

}  // namespace

extern "C" {

int yrhost_scene_create(int kind, const int64_t* params, int param_count, uint64_t seed, yrhost_scene** out) {
    if (out == nullptr) return YRH_E_INVALID_ARGUMENT;
    *out = nullptr;
    auto s = std::unique_ptr<yrhost_scene>(new (std::nothrow) yrhost_scene());
    if (!s) return YRH_E_INVALID_ARGUMENT;
    auto param = [&](int i, int64_t def) { return (params != nullptr && i < param_count) ? params[i] : def; };
    switch (kind) {
        case YRHOST_SCENE_CORNELL:
            s->ir = yrhip::MakeCornellScene(seed);
            break;
        case YRHOST_SCENE_TRIANGLES: {
            const int64_t n = param(0, 1000000);
            if (n < 0 || n > 50000000) return YRH_E_INVALID_ARGUMENT;
            s->ir = yrhip::MakeRandomTriangleScene(static_cast<int>(n), seed);
            break;
        }
        case YRHOST_SCENE_INSTANCED: {
            const int64_t inst = param(0, 10000), sub = param(1, 4), var = param(2, 8), ew = param(3, 2048),
                          eh = param(4, 1024);
            if (inst < 0 || sub < 0 || sub > 7 || var < 1 || ew < 0 || eh < 0) return YRH_E_INVALID_ARGUMENT;
            s->ir = yrhip::MakeInstancedScene(static_cast<int>(inst), static_cast<int>(sub), static_cast<int>(var),
                                              seed, static_cast<int>(ew), static_cast<int>(eh));
            break;
        }
        default:
            return YRH_E_INVALID_ARGUMENT;
    }
    *out = s.release();
    return YRH_OK;
}

int yrhost_scene_from_arrays(const yrh_vertex* vertices, int64_t vertex_count, const uint32_t* indices,
                             int64_t index_count, const yrh_primitive* primitives, int32_t primitive_count,
                             const int32_t* instance_primitive, const float* instance_matrices,
                             int32_t instance_count, yrhost_scene** out) {
    if (out == nullptr || vertex_count < 0 || index_count < 0 || primitive_count < 0 || instance_count < 0) {
        return YRH_E_INVALID_ARGUMENT;
    }
    if ((vertex_count > 0 && vertices == nullptr) || (index_count > 0 && indices == nullptr) ||
        (primitive_count > 0 && primitives == nullptr) ||
        (instance_count > 0 && (instance_primitive == nullptr || instance_matrices == nullptr))) {
        return YRH_E_INVALID_ARGUMENT;
    }
    *out = nullptr;
    auto s = std::unique_ptr<yrhost_scene>(new (std::nothrow) yrhost_scene());
    if (!s) return YRH_E_INVALID_ARGUMENT;
    s->ir.vertices.resize(static_cast<size_t>(vertex_count));
    if (vertex_count > 0) std::memcpy(s->ir.vertices.data(), vertices, static_cast<size_t>(vertex_count) * sizeof(yrh_vertex));
    s->ir.indices.assign(indices, indices + index_count);
    s->ir.primitives.resize(static_cast<size_t>(primitive_count));
    for (int32_t i = 0; i < primitive_count; ++i) {
        yr::RenderPrimitive& p = s->ir.primitives[static_cast<size_t>(i)];
        p.first_index = primitives[i].first_index;
        p.index_count = primitives[i].index_count;
        p.material_index = primitives[i].material_index;
        p.has_normals = primitives[i].has_normals != 0;
        p.has_uvs = primitives[i].has_uvs != 0;
        p.has_tangents = primitives[i].has_tangents != 0;
    }
    for (int32_t i = 0; i < instance_count; ++i) {
        yr::RenderInstance inst;
        inst.primitive = instance_primitive[i];
        std::memcpy(inst.object_to_world.m.data(), instance_matrices + 16 * static_cast<size_t>(i), 16 * sizeof(float));
        s->ir.instances.push_back(inst);
    }
    int32_t max_material = -1;
    for (const yr::RenderPrimitive& p : s->ir.primitives) max_material = std::max(max_material, p.material_index);
    s->ir.materials.resize(static_cast<size_t>(std::max(1, max_material + 1)));
    *out = s.release();
    return YRH_OK;
}

void yrhost_scene_destroy(yrhost_scene* scene) { delete scene; }

const char* yrhost_scene_error(const yrhost_scene* scene) { return scene == nullptr ? "null scene" : scene->error.c_str(); }

int yrhost_scene_info_get(const yrhost_scene* s, yrhost_scene_info* info) {
    if (s == nullptr || info == nullptr) return YRH_E_INVALID_ARGUMENT;
    *info = yrhost_scene_info{};
    info->vertex_count = static_cast<int64_t>(s->ir.vertices.size());
    info->index_count = static_cast<int64_t>(s->ir.indices.size());
    info->primitive_count = static_cast<int32_t>(s->ir.primitives.size());
    info->instance_count = static_cast<int32_t>(s->ir.instances.size());
    info->material_count = static_cast<int32_t>(s->ir.materials.size());
    info->emissive_count = static_cast<int32_t>(s->ir.emissive_primitives.size());
    info->accel_built = s->accel_built ? 1 : 0;
    info->environment_active = s->ir.environment.active ? 1 : 0;
    if (s->accel_built) {
        info->accel_kind = s->accel.kind == yr::RenderAccelerationKind::TwoLevel ? YRH_ACCEL_TWO_LEVEL : YRH_ACCEL_FLAT;
        info->node_count = s->accel.NodeCount();
        info->max_depth = s->accel.MaxDepth();
        info->blas_count = static_cast<int32_t>(s->accel.two_level.blases.size());
    }
    const yr::RenderCamera& c = s->ir.camera;
    const float cam[13] = {c.origin.x,  c.origin.y,  c.origin.z,  c.forward.x, c.forward.y, c.forward.z, c.right.x,
                           c.right.y,   c.right.z,   c.up.x,      c.up.y,      c.up.z,      c.fov_y_radians};
    std::memcpy(info->camera, cam, sizeof(cam));
    return YRH_OK;
}

int yrhost_scene_geometry(const yrhost_scene* s, const yrh_vertex** vertices, const uint32_t** indices,
                          const yrh_primitive** primitives) {
    if (s == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (vertices != nullptr) *vertices = reinterpret_cast<const yrh_vertex*>(s->ir.vertices.data());
    if (indices != nullptr) *indices = s->ir.indices.data();
    if (primitives != nullptr) *primitives = reinterpret_cast<const yrh_primitive*>(s->ir.primitives.data());
    return YRH_OK;
}

int yrhost_build_accel(yrhost_scene* s, int thread_count, double* seconds) {
    if (s == nullptr || thread_count < 0) return YRH_E_INVALID_ARGUMENT;
    yr::BvhBuildOptions opt;
    opt.thread_count = thread_count;
    const auto t0 = std::chrono::steady_clock::now();
    yr::RenderAccelerationBuildResult r = yr::BuildRenderAcceleration(s->ir.Geometry(), opt);
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds != nullptr) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (!r.errors.empty()) {
        s->accel_built = false;
        return SetError(s, r.errors.front());
    }
    s->accel = std::move(r.acceleration);
    s->accel_built = true;
    s->blas_desc.clear();
    for (const yr::RenderBlas& b : s->accel.two_level.blases) s->blas_desc.push_back(Describe(b.bvh));
    return YRH_OK;
}

int yrhost_accel_flat(const yrhost_scene* s, yrh_bvh_desc* out) {
    if (s == nullptr || out == nullptr || !s->accel_built) return YRH_E_INVALID_ARGUMENT;
    *out = Describe(s->accel.flat);
    return YRH_OK;
}

int yrhost_accel_blas(const yrhost_scene* s, int32_t index, yrh_bvh_desc* out) {
    if (s == nullptr || out == nullptr || !s->accel_built) return YRH_E_INVALID_ARGUMENT;
    if (index < 0 || index >= static_cast<int32_t>(s->blas_desc.size())) return YRH_E_INVALID_ARGUMENT;
    *out = s->blas_desc[static_cast<size_t>(index)];
    return YRH_OK;
}

int yrhost_accel_tlas(const yrhost_scene* s, yrh_bvh_desc* tlas, const yrh_tlas_primitive** primitives,
                      int32_t* primitive_count, const yrh_instance** instances, int32_t* instance_count) {
    if (s == nullptr || !s->accel_built) return YRH_E_INVALID_ARGUMENT;
    const yr::TwoLevelBvh& tl = s->accel.two_level;
    if (tlas != nullptr) {
        *tlas = yrh_bvh_desc{};
        tlas->nodes = reinterpret_cast<const yrh_bvh_node*>(tl.tlas.nodes.data());
        tlas->node_count = static_cast<int32_t>(tl.tlas.nodes.size());
        tlas->primitive_indices = tl.tlas.primitive_indices.data();
        tlas->primitive_index_count = static_cast<int32_t>(tl.tlas.primitive_indices.size());
        tlas->max_depth = tl.tlas.max_depth;
    }
    if (primitives != nullptr) *primitives = reinterpret_cast<const yrh_tlas_primitive*>(tl.tlas.primitives.data());
    if (primitive_count != nullptr) *primitive_count = static_cast<int32_t>(tl.tlas.primitives.size());
    if (instances != nullptr) *instances = reinterpret_cast<const yrh_instance*>(tl.instances.data());
    if (instance_count != nullptr) *instance_count = static_cast<int32_t>(tl.instances.size());
    return YRH_OK;
}

int yrhost_upload_geometry(const yrhost_scene* s, yrh_context* ctx) {
    if (s == nullptr || ctx == nullptr || !s->accel_built) return YRH_E_INVALID_ARGUMENT;
    yrh_geometry_desc g{};
    g.vertices = reinterpret_cast<const yrh_vertex*>(s->ir.vertices.data());
    g.vertex_count = static_cast<int64_t>(s->ir.vertices.size());
    g.indices = s->ir.indices.data();
    g.index_count = static_cast<int64_t>(s->ir.indices.size());
    g.primitives = reinterpret_cast<const yrh_primitive*>(s->ir.primitives.data());
    g.primitive_count = static_cast<int32_t>(s->ir.primitives.size());
    if (s->accel.kind == yr::RenderAccelerationKind::FlatReference) {
        g.accel_kind = YRH_ACCEL_FLAT;
        g.flat = Describe(s->accel.flat);
    } else {
        g.accel_kind = YRH_ACCEL_TWO_LEVEL;
        g.blases = s->blas_desc.data();
        g.blas_count = static_cast<int32_t>(s->blas_desc.size());
        yrhost_accel_tlas(s, &g.tlas, &g.tlas_primitives, &g.tlas_primitive_count, &g.instances, &g.instance_count);
    }
    return yrh_upload_geometry(ctx, &g);
}

int yrhost_camera_rays(const yrhost_scene* s, int width, int height, float t_min, float t_max, float* org_tmin,
                       float* dir_tmax) {
    if (s == nullptr || width <= 0 || height <= 0 || org_tmin == nullptr || dir_tmax == nullptr) return YRH_E_INVALID_ARGUMENT;
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            yr::Point3f o;
            yr::Vec3f d;
            yrhip::CameraRay(s->ir.camera, width, height, x, y, 0.5f, 0.5f, o, d);
            const size_t i = 4 * (static_cast<size_t>(y) * static_cast<size_t>(width) + static_cast<size_t>(x));
            org_tmin[i + 0] = o.x;
            org_tmin[i + 1] = o.y;
            org_tmin[i + 2] = o.z;
            org_tmin[i + 3] = t_min;
            dir_tmax[i + 0] = d.x;
            dir_tmax[i + 1] = d.y;
            dir_tmax[i + 2] = d.z;
            dir_tmax[i + 3] = t_max;
        }
    }
    return YRH_OK;
}

int yrhost_random_rays(int64_t count, uint64_t seed, float t_min, float t_max, float* org_tmin, float* dir_tmax) {
    if (count < 0 || (count > 0 && (org_tmin == nullptr || dir_tmax == nullptr))) return YRH_E_INVALID_ARGUMENT;
    yrhip::Pcg32 rng(seed);
    for (int64_t i = 0; i < count; ++i) {
        const float ox = rng.NextFloat(), oy = rng.NextFloat(), oz = rng.NextFloat();
        const yr::Vec3f d = yrhip::UniformSphere(rng.NextFloat(), rng.NextFloat());
        float* o = org_tmin + 4 * i;
        float* v = dir_tmax + 4 * i;
        o[0] = ox;
        o[1] = oy;
        o[2] = oz;
        o[3] = t_min;
        v[0] = d.x;
        v[1] = d.y;
        v[2] = d.z;
        v[3] = t_max;
    }
    return YRH_OK;
}

}  // extern "C"
