// C entry points of the host library (scene IR, builder, workloads, upload).
#include "yaoray_host.h"

#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <memory>
#include <new>
#include <optional>
#include <stop_token>
#include <string>

#include "abi_marshal.hpp"
#include "hip_backend.hpp"
#include "scenes.hpp"
#include "yr_ir.hpp"

struct yrhost_scene {
    yr::RenderSceneIR ir;
    yr::RenderAcceleration accel;
    bool accel_built = false;
    std::string error;
    // Converted tables the C-ABI descriptors point into.
    mutable yrhip::Marshal marshal;
};

namespace {

int SetError(yrhost_scene* s, const std::string& msg) {
    if (s != nullptr) s->error = msg;
    return YRH_E_INVALID_ARGUMENT;
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
        case YRHOST_SCENE_ROOM: {
            const int64_t tris = param(0, 200000), ew = param(1, 1024), eh = param(2, 512);
            if (tris < 0 || tris > 20000000 || ew < 0 || eh < 0) return YRH_E_INVALID_ARGUMENT;
            s->ir = yrhip::MakeGuidingRoomScene(static_cast<int>(tris), seed, static_cast<int>(ew), static_cast<int>(eh));
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

int yrhost_scene_set_camera(yrhost_scene* s, const float* c) {
    if (s == nullptr || c == nullptr) return YRH_E_INVALID_ARGUMENT;
    yr::RenderCamera& cam = s->ir.camera;
    cam.origin = yr::Point3f{c[0], c[1], c[2]};
    cam.forward = yr::Vec3f{c[3], c[4], c[5]};
    cam.right = yr::Vec3f{c[6], c[7], c[8]};
    cam.up = yr::Vec3f{c[9], c[10], c[11]};
    cam.fov_y_radians = c[12];
    return YRH_OK;
}

int yrhost_scene_set_materials(yrhost_scene* s, const yrh_material* m, int32_t count) {
    if (s == nullptr || m == nullptr || count <= 0) return YRH_E_INVALID_ARGUMENT;
    s->ir.materials.clear();
    for (int32_t k = 0; k < count; ++k) {
        yr::RenderMaterial r;
        switch (m[k].kind) {
            case YRH_MATERIAL_DIFFUSE: r.kind = yr::RenderMaterialKind::Diffuse; break;
            case YRH_MATERIAL_CONDUCTOR: r.kind = yr::RenderMaterialKind::Conductor; break;
            case YRH_MATERIAL_DIELECTRIC: r.kind = yr::RenderMaterialKind::Dielectric; break;
            case YRH_MATERIAL_THIN_DIELECTRIC: r.kind = yr::RenderMaterialKind::ThinDielectric; break;
            default: return SetError(s, "unknown material kind");
        }
        r.reflectance.value = yr::Color3f{m[k].reflectance[0], m[k].reflectance[1], m[k].reflectance[2]};
        r.ior = m[k].ior;
        r.uroughness.value = m[k].uroughness;
        r.vroughness.value = m[k].vroughness;
        r.alpha.value = m[k].alpha;
        r.emission = yr::Color3f{m[k].emission[0], m[k].emission[1], m[k].emission[2]};
        r.absorption_distance = m[k].absorption_distance;
        r.absorption_color = yr::Color3f{m[k].absorption_color[0], m[k].absorption_color[1], m[k].absorption_color[2]};
        r.reflectance.texture = m[k].reflectance_texture;
        r.alpha.texture = m[k].alpha_texture;
        r.normal_map = m[k].normal_map;
        r.normal_scale = m[k].normal_scale;
        s->ir.materials.push_back(r);
    }
    return YRH_OK;
}

int yrhost_scene_set_emissives(yrhost_scene* s, const yrh_emissive* e, int32_t count) {
    if (s == nullptr || count < 0 || (count > 0 && e == nullptr)) return YRH_E_INVALID_ARGUMENT;
    s->ir.emissive_primitives.clear();
    for (int32_t k = 0; k < count; ++k) {
        s->ir.emissive_primitives.push_back(
            yr::EmissivePrimitive{e[k].primitive_index, yr::Color3f{e[k].radiance[0], e[k].radiance[1], e[k].radiance[2]},
                                  e[k].area});
    }
    yr::PrepareLightSampling(s->ir);
    return YRH_OK;
}

int yrhost_scene_set_analytic_lights(yrhost_scene* s, const yrh_analytic_light* l, int32_t count) {
    if (s == nullptr || count < 0 || (count > 0 && l == nullptr)) return YRH_E_INVALID_ARGUMENT;
    s->ir.analytic_lights.clear();
    for (int32_t k = 0; k < count; ++k) {
        if (l[k].kind < YRH_LIGHT_POINT || l[k].kind > YRH_LIGHT_DISTANT) return YRH_E_INVALID_ARGUMENT;
        yr::AnalyticLight a;
        a.kind = static_cast<yr::AnalyticLightKind>(l[k].kind);
        a.position = yr::Point3f{l[k].position[0], l[k].position[1], l[k].position[2]};
        a.direction = yr::Vec3f{l[k].direction[0], l[k].direction[1], l[k].direction[2]};
        a.intensity = yr::Color3f{l[k].intensity[0], l[k].intensity[1], l[k].intensity[2]};
        a.cone_angle = l[k].cone_angle;
        a.cone_cos_inner = l[k].cone_cos_inner;
        s->ir.analytic_lights.push_back(a);
    }
    yr::PrepareLightSampling(s->ir);
    return YRH_OK;
}

int yrhost_scene_set_constant_environment(yrhost_scene* s, const float* radiance3, float strength) {
    if (s == nullptr || radiance3 == nullptr) return YRH_E_INVALID_ARGUMENT;
    s->ir.environment = yr::RenderEnvironment{};
    s->ir.environment.active = false;
    s->ir.environment.radiance = yr::Color3f{radiance3[0], radiance3[1], radiance3[2]};
    s->ir.environment.strength = strength;
    s->ir.environment_distributions.clear();
    return YRH_OK;
}

int yrhost_scene_set_environment_map(yrhost_scene* s, int32_t texture_index, const float* radiance3, float strength,
                                     float rotation_radians) {
    if (s == nullptr || radiance3 == nullptr || texture_index < 0 ||
        texture_index >= static_cast<int32_t>(s->ir.textures.size())) {
        return YRH_E_INVALID_ARGUMENT;
    }
    const yr::RenderTexture& t = s->ir.textures[static_cast<size_t>(texture_index)];
    if (t.width <= 0 || t.height <= 0 || t.texels.empty()) return SetError(s, "environment map has no texels");
    s->ir.environment = yr::RenderEnvironment{};
    s->ir.environment.active = true;
    s->ir.environment.radiance = yr::Color3f{radiance3[0], radiance3[1], radiance3[2]};
    s->ir.environment.strength = strength;
    s->ir.environment.rotation_radians = rotation_radians;
    s->ir.environment.texture_index = texture_index;
    s->ir.environment_distributions.clear();
    s->ir.environment_distributions.push_back(yr::BuildEnvironmentDistribution(t));
    s->ir.environment.distribution_index = 0;
    return YRH_OK;
}

int yrhost_scene_add_texture(yrhost_scene* s, const yrh_texture* t) {
    if (s == nullptr || t == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (t->filter < YRH_FILTER_NEAREST || t->filter > YRH_FILTER_EWA || t->wrap_s < YRH_WRAP_REPEAT ||
        t->wrap_s > YRH_WRAP_MIRRORED || t->wrap_t < YRH_WRAP_REPEAT || t->wrap_t > YRH_WRAP_MIRRORED) {
        return SetError(s, "texture filter or wrap mode out of range");
    }
    yr::RenderTexture r;
    r.filter = static_cast<yr::TextureFilter>(t->filter);
    r.wrap_s = static_cast<yr::TextureWrap>(t->wrap_s);
    r.wrap_t = static_cast<yr::TextureWrap>(t->wrap_t);
    if (t->width > 0 && t->height > 0 && t->texels != nullptr) {
        r.width = t->width;
        r.height = t->height;
        r.texels.resize(static_cast<size_t>(t->width) * static_cast<size_t>(t->height));
        std::memcpy(r.texels.data(), t->texels, r.texels.size() * sizeof(yr::Color4f));
    }
    s->ir.textures.push_back(std::move(r));
    return static_cast<int>(s->ir.textures.size()) - 1;
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
    info->sphere_count = static_cast<int32_t>(s->ir.spheres.size());
    info->analytic_light_count = static_cast<int32_t>(s->ir.analytic_lights.size());
    info->texture_count = static_cast<int32_t>(s->ir.textures.size());
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

int yrhost_scene_set_spheres(yrhost_scene* s, const yrh_sphere* spheres, int32_t count) {
    if (s == nullptr || count < 0 || (count > 0 && spheres == nullptr)) return YRH_E_INVALID_ARGUMENT;
    static_assert(sizeof(yrh_sphere) == sizeof(yr::RenderSphere));
    s->ir.spheres.resize(static_cast<size_t>(count));
    if (count > 0) std::memcpy(s->ir.spheres.data(), spheres, sizeof(yrh_sphere) * static_cast<size_t>(count));
    s->accel_built = false;
    return YRH_OK;
}

int yrhost_scene_spheres(const yrhost_scene* s, const yrh_sphere** spheres, int32_t* count) {
    if (s == nullptr || spheres == nullptr || count == nullptr) return YRH_E_INVALID_ARGUMENT;
    *spheres = reinterpret_cast<const yrh_sphere*>(s->ir.spheres.data());
    *count = static_cast<int32_t>(s->ir.spheres.size());
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
    s->marshal.blas.clear();
    for (const yr::RenderBlas& b : s->accel.two_level.blases) s->marshal.blas.push_back(yrhip::DescribeBvh(b.bvh));
    return YRH_OK;
}

int yrhost_accel_flat(const yrhost_scene* s, yrh_bvh_desc* out) {
    if (s == nullptr || out == nullptr || !s->accel_built) return YRH_E_INVALID_ARGUMENT;
    *out = yrhip::DescribeBvh(s->accel.flat);
    return YRH_OK;
}

int yrhost_accel_blas(const yrhost_scene* s, int32_t index, yrh_bvh_desc* out) {
    if (s == nullptr || out == nullptr || !s->accel_built) return YRH_E_INVALID_ARGUMENT;
    if (index < 0 || index >= static_cast<int32_t>(s->marshal.blas.size())) return YRH_E_INVALID_ARGUMENT;
    *out = s->marshal.blas[static_cast<size_t>(index)];
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

int yrhost_geometry_desc(const yrhost_scene* s, yrh_geometry_desc* out) {
    if (s == nullptr || out == nullptr || !s->accel_built) return YRH_E_INVALID_ARGUMENT;
    yrhip::DescribeGeometry(s->ir, s->accel, s->marshal, *out);
    return YRH_OK;
}

int yrhost_upload_geometry(const yrhost_scene* s, yrh_context* ctx) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    yrh_geometry_desc g{};
    const int rc = yrhost_geometry_desc(s, &g);
    if (rc != YRH_OK) return rc;
    return yrh_upload_geometry(ctx, &g);
}

int yrhost_shading_desc(yrhost_scene* s, yrh_shading_desc* out) {
    if (s == nullptr || out == nullptr) return YRH_E_INVALID_ARGUMENT;
    const std::string err = yrhip::DescribeShading(s->ir, s->marshal, *out);
    return err.empty() ? YRH_OK : SetError(s, err);
}

int yrhost_upload_shading(yrhost_scene* s, yrh_context* ctx) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    yrh_shading_desc d{};
    const int rc = yrhost_shading_desc(s, &d);
    if (rc != YRH_OK) return rc;
    return yrh_upload_shading(ctx, &d);
}

int yrhost_render_desc(const yrhost_scene* s, int32_t width, int32_t height, int32_t sampler, int32_t spp,
                       int32_t max_depth, uint64_t seed, float radiance_clamp, int32_t sample_index,
                       yrh_render_desc* out) {
    if (s == nullptr || out == nullptr) return YRH_E_INVALID_ARGUMENT;
    const yr::RenderCamera& c = s->ir.camera;
    yrh_render_desc& d = *out;
    d = yrh_render_desc{};
    const float cam[12] = {c.origin.x, c.origin.y, c.origin.z, c.forward.x, c.forward.y, c.forward.z,
                           c.right.x,  c.right.y,  c.right.z,  c.up.x,      c.up.y,      c.up.z};
    std::memcpy(d.camera_origin, cam, 12);
    std::memcpy(d.camera_forward, cam + 3, 12);
    std::memcpy(d.camera_right, cam + 6, 12);
    std::memcpy(d.camera_up, cam + 9, 12);
    // The reference evaluates std::tan(fov_y_radians * 0.5f) per ray (camera.cpp:14).
    d.tan_half_fov = std::tan(c.fov_y_radians * 0.5f);
    d.width = width;
    d.height = height;
    d.sampler = sampler;
    d.spp = spp;
    d.max_depth = max_depth;
    d.seed = seed;
    d.radiance_clamp = radiance_clamp;
    d.sample_index = sample_index;
    return YRH_OK;
}

int yrhost_backend_render(const yrhost_scene* s, const yrhost_render_settings* settings,
                          const yrhost_render_request* request, yrhost_film_pixel* film_out, int64_t film_capacity,
                          yrhost_render_report* report) {
    static_assert(sizeof(yrhost_film_pixel) == sizeof(yr::FilmPixel));
    if (s == nullptr || settings == nullptr || report == nullptr) return YRH_E_INVALID_ARGUMENT;
    *report = yrhost_render_report{};
    auto set_error = [&](const std::string& e) {
        std::strncpy(report->error, e.c_str(), sizeof(report->error) - 1);
        report->error[sizeof(report->error) - 1] = '\0';
    };
    yr::RenderJob job;
    job.scene = s->ir;
    job.settings.requested_backend = settings->requested_backend == 0 ? yr::RenderBackendKind::Hip
                                                                      : yr::RenderBackendKind::Cpu;
    switch (settings->sampler) {
        case YRH_SAMPLER_INDEPENDENT: job.settings.sampler = yr::RenderSamplerKind::Independent; break;
        case YRH_SAMPLER_STRATIFIED: job.settings.sampler = yr::RenderSamplerKind::Stratified; break;
        case YRH_SAMPLER_OWEN_SOBOL: job.settings.sampler = yr::RenderSamplerKind::OwenSobol; break;
        case YRH_SAMPLER_ZSOBOL: job.settings.sampler = yr::RenderSamplerKind::ZSobol; break;
        default: return YRH_E_INVALID_ARGUMENT;
    }
    job.settings.width = settings->width;
    job.settings.height = settings->height;
    job.settings.spp = settings->spp;
    job.settings.max_depth = settings->max_depth;
    job.settings.seed = settings->seed;
    job.settings.radiance_clamp = settings->radiance_clamp;
    job.settings.adaptive_sampling.enabled = settings->adaptive_enabled != 0;
    job.settings.adaptive_sampling.min_spp = settings->adaptive_min_spp;
    job.settings.adaptive_sampling.relative_error = settings->adaptive_relative_error;
    job.settings.adaptive_sampling.absolute_error = settings->adaptive_absolute_error;
    job.settings.adaptive_sampling.confidence = settings->adaptive_confidence;

    yr::HipBackend backend(settings->device);
    yr::BackendPrepareResult prepared = backend.Prepare(std::move(job));
    report->prepare_seconds = prepared.elapsed_seconds;
    report->acceleration_build_seconds = prepared.acceleration_build_seconds;
    if (!prepared.ok) {
        set_error(prepared.error);
        return YRH_OK;
    }
    report->prepare_ok = 1;

    std::optional<yr::Film> resume;
    yr::RenderRequest req;
    std::stop_source stop;
    req.stop_token = stop.get_token();
    std::string cancel_error;
    if (request != nullptr) {
        if (request->resume_film != nullptr) {
            try {
                resume.emplace(request->resume_width, request->resume_height);
            } catch (const std::exception& e) {
                set_error(e.what());
                return YRH_OK;
            }
            for (int y = 0; y < request->resume_height; ++y) {
                for (int x = 0; x < request->resume_width; ++x) {
                    yr::FilmPixel p;
                    std::memcpy(&p, &request->resume_film[static_cast<size_t>(y) * request->resume_width + x], sizeof(p));
                    resume->SetPixelForCheckpoint(x, y, p);
                }
            }
            req.resume_film = &*resume;
        }
        req.resume_completed_spp = request->resume_completed_spp;
        if (request->stop_before_start != 0) stop.request_stop();
        if (request->cancel_error != nullptr) cancel_error = request->cancel_error;
    }
    const int cancel_after = request != nullptr ? request->cancel_after_passes : 0;
    if (cancel_after >= 0) req.progress_callback = [&](const yr::RenderProgress& p, const yr::Film&) {
        ++report->progress_calls;
        report->last_progress_spp = p.completed_spp;
        report->last_progress_samples = p.completed_samples;
        report->last_progress_active = p.active_pixels;
        yr::RenderProgressDecision d;
        if (cancel_after > 0 && report->progress_calls >= cancel_after) {
            d.cancel = true;
            d.error = cancel_error;
        }
        return d;
    };
    const yr::RenderResult r = backend.Render(*prepared.scene, req);
    report->ok = r.ok ? 1 : 0;
    set_error(r.error);
    const yr::RenderStats& st = r.stats;
    report->threads = st.threads;
    report->bvh_nodes = st.bvh_nodes;
    report->bvh_max_depth = st.bvh_max_depth;
    report->samples_rendered = st.samples_rendered;
    report->rays_traced = st.rays_traced;
    report->shadow_rays = st.shadow_rays;
    report->occluded_shadow_rays = st.occluded_shadow_rays;
    report->triangle_tests = st.triangle_tests;
    report->sphere_tests = st.sphere_tests;
    report->bvh_node_tests = st.bvh_node_tests;
    report->hits = st.hits;
    report->misses = st.misses;
    report->converged_pixels = st.converged_pixels;
    report->elapsed_seconds = st.elapsed_seconds;
    if (r.film.has_value()) {
        report->film_valid = 1;
        report->film_width = r.film->Width();
        report->film_height = r.film->Height();
        report->bad_samples = r.film->BadSampleCount();
        const std::vector<yr::FilmPixel>& px = r.film->Pixels();
        if (film_out != nullptr && film_capacity >= static_cast<int64_t>(px.size())) {
            std::memcpy(film_out, px.data(), px.size() * sizeof(yr::FilmPixel));
        }
    }
    return YRH_OK;
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
