// ORACLE — test infrastructure only (see oracle/oracle_trace.cpp).
//
// C entry points over the reference itself, compiled from its own sources
// where they lie by oracle/Makefile into oracle/_ref/libyrref.so. Used
// to pin the restated oracle and the product's BVH builder (byte-identical
// node arrays), and as the reference CPU baseline. Builds against the
// reference's headers; contains no reference code.
#include <yaoray/accel/acceleration.hpp>
#include <yaoray/backends/cpu/cpu_path_tracer.hpp>
#include <yaoray/backends/cpu/cpu_prepared_scene.hpp>
#include <yaoray/core/transform.hpp>
#include <yaoray/lighting/environment.hpp>
#include <yaoray/lighting/light_sampling.hpp>
#include <yaoray/core/rng.hpp>
#include <yaoray/sampling/sampler.hpp>
#include <yaoray/shading/bsdf.hpp>
#include <yaoray/shading/texture.hpp>
#include <yaoray/scene/render_scene.hpp>

#include "../include/yrh.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct RefScene {
    yr::RenderSceneIR scene;
    yr::RenderAcceleration accel;
};

// The reference material a yrh_material describes; texture indices outside
// [0, ntex) become -1, which the reference reads as "no texture" as well
// (the render appends its environment texture after the ntex material ones).
yr::RenderMaterial ToRefMaterial(const yrh_material& m, int32_t ntex = 0) {
    yr::RenderMaterial rm;
    rm.kind = m.kind == YRH_MATERIAL_DIFFUSE ? yr::RenderMaterialKind::Diffuse
            : m.kind == YRH_MATERIAL_CONDUCTOR ? yr::RenderMaterialKind::Conductor
            : m.kind == YRH_MATERIAL_DIELECTRIC ? yr::RenderMaterialKind::Dielectric
                                                : yr::RenderMaterialKind::ThinDielectric;
    rm.reflectance.value = yr::Color3f{m.reflectance[0], m.reflectance[1], m.reflectance[2]};
    rm.ior = m.ior;
    rm.uroughness.value = m.uroughness;
    rm.vroughness.value = m.vroughness;
    rm.alpha.value = m.alpha;
    rm.emission = yr::Color3f{m.emission[0], m.emission[1], m.emission[2]};
    rm.absorption_color = yr::Color3f{m.absorption_color[0], m.absorption_color[1], m.absorption_color[2]};
    rm.absorption_distance = m.absorption_distance;
    auto tex = [&](int32_t k) { return k >= 0 && k < ntex ? k : -1; };
    rm.reflectance.texture = tex(m.reflectance_texture);
    rm.alpha.texture = tex(m.alpha_texture);
    rm.normal_map = tex(m.normal_map);
    rm.normal_scale = m.normal_scale;
    return rm;
}

yr::RenderTexture ToRefTexture(const yrh_texture& t) {
    yr::RenderTexture r;
    r.kind = yr::RenderTextureKind::Image;
    r.filter = t.filter == YRH_FILTER_NEAREST     ? yr::TextureFilter::Nearest
             : t.filter == YRH_FILTER_BILINEAR    ? yr::TextureFilter::Bilinear
             : t.filter == YRH_FILTER_TRILINEAR   ? yr::TextureFilter::Trilinear
                                                  : yr::TextureFilter::Ewa;
    auto wrap = [](int32_t w) {
        return w == YRH_WRAP_CLAMP ? yr::TextureWrap::ClampToEdge
             : w == YRH_WRAP_MIRRORED ? yr::TextureWrap::MirroredRepeat
                                      : yr::TextureWrap::Repeat;
    };
    r.wrap_s = wrap(t.wrap_s);
    r.wrap_t = wrap(t.wrap_t);
    if (t.width > 0 && t.height > 0 && t.texels != nullptr) {
        r.width = t.width;
        r.height = t.height;
        r.texels.resize(static_cast<size_t>(t.width) * static_cast<size_t>(t.height));
        std::memcpy(r.texels.data(), t.texels, r.texels.size() * 16);
    }
    return r;
}

}  // namespace

extern "C" {
#pragma GCC visibility push(default)

// Spheres the next ref_build / ref_render calls add to their scene (yrh_sphere
// records, converted field by field into yr::RenderSphere); n = 0 clears them.
std::vector<yr::RenderSphere> g_spheres;

void ref_set_spheres(const yrh_sphere* sp, int32_t n) {
    g_spheres.clear();
    for (int32_t k = 0; k < n; ++k) {
        yr::RenderSphere r;
        r.center = yr::Point3f{sp[k].center[0], sp[k].center[1], sp[k].center[2]};
        r.radius = sp[k].radius;
        r.material_index = sp[k].material_index;
        r.area_light_index = sp[k].area_light_index;
        r.flip_normals = sp[k].flip_normals != 0;
        g_spheres.push_back(r);
    }
}

// Analytic lights the next ref_render calls add to their scene (yrh_analytic_light
// records, converted field by field); n = 0 clears them.
std::vector<yr::AnalyticLight> g_analytic;

void ref_set_analytic_lights(const yrh_analytic_light* l, int32_t n) {
    g_analytic.clear();
    for (int32_t k = 0; k < n; ++k) {
        yr::AnalyticLight a;
        a.kind = static_cast<yr::AnalyticLightKind>(l[k].kind);
        a.position = yr::Point3f{l[k].position[0], l[k].position[1], l[k].position[2]};
        a.direction = yr::Vec3f{l[k].direction[0], l[k].direction[1], l[k].direction[2]};
        a.intensity = yr::Color3f{l[k].intensity[0], l[k].intensity[1], l[k].intensity[2]};
        a.cone_angle = l[k].cone_angle;
        a.cone_cos_inner = l[k].cone_cos_inner;
        g_analytic.push_back(a);
    }
}

// yr::PrepareLightSampling's analytic tables for `n` lights: the power alias
// table (n entries each) and the light tree (at most 2n - 1 nodes, 40-byte
// LightTreeNode records); returns the node count.
int32_t ref_analytic_tables(const yrh_analytic_light* l, int32_t n, float* prob, int32_t* alias, float* pmf,
                            void* tree, int32_t* root, int32_t* use_tree) {
    ref_set_analytic_lights(l, n);
    yr::RenderSceneIR sc;
    sc.analytic_lights = g_analytic;
    g_analytic.clear();
    yr::PrepareLightSampling(sc);
    const yr::LightSamplingCache& c = sc.light_sampling;
    for (size_t k = 0; k < c.analytic_power.pmf.size(); ++k) {
        prob[k] = c.analytic_power.probability[k];
        alias[k] = c.analytic_power.alias[k];
        pmf[k] = c.analytic_power.pmf[k];
    }
    static_assert(sizeof(yr::LightTreeNode) == 40);
    if (!c.analytic_tree.empty()) std::memcpy(tree, c.analytic_tree.data(), c.analytic_tree.size() * 40);
    *root = c.analytic_tree_root;
    *use_tree = c.use_analytic_tree ? 1 : 0;
    return static_cast<int32_t>(c.analytic_tree.size());
}

// vertices: 48-byte RenderVertex records; primitives: 16-byte RenderPrimitive.
int ref_build(const void* vertices, int64_t nv, const uint32_t* indices, int64_t ni, const void* primitives,
              int32_t np, const int32_t* inst_prim, const float* inst_mat, int32_t ninst, int32_t threads,
              void** out, double* seconds) {
    static_assert(sizeof(yr::RenderVertex) == 48);
    static_assert(sizeof(yr::RenderPrimitive) == 16);
    auto* s = new RefScene();
    s->scene.vertices.resize(static_cast<size_t>(nv));
    if (nv > 0) std::memcpy(s->scene.vertices.data(), vertices, static_cast<size_t>(nv) * 48);
    s->scene.indices.assign(indices, indices + ni);
    s->scene.primitives.resize(static_cast<size_t>(np));
    if (np > 0) std::memcpy(s->scene.primitives.data(), primitives, static_cast<size_t>(np) * 16);
    for (int32_t i = 0; i < ninst; ++i) {
        yr::RenderInstance inst;
        inst.primitive = yr::MeshPrimitiveHandle{inst_prim[i]};
        std::memcpy(inst.object_to_world.m.data(), inst_mat + 16 * static_cast<size_t>(i), 64);
        s->scene.instances.push_back(inst);
    }
    s->scene.spheres = g_spheres;
    s->scene.materials.push_back(yr::RenderMaterial{});
    yr::BvhBuildOptions opt;
    opt.thread_count = threads;
    const auto t0 = std::chrono::steady_clock::now();
    auto r = yr::BuildRenderAcceleration(s->scene.Geometry(), opt);
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds != nullptr) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (!r.errors.empty()) {
        delete s;
        return -1;
    }
    s->accel = std::move(r.acceleration);
    *out = s;
    return 0;
}

void ref_free(void* h) { delete static_cast<RefScene*>(h); }

// kind: 0 flat, 1 two-level; counts of the flat BVH or the TLAS.
int ref_accel_info(void* h, int32_t* kind, int32_t* nodes, int32_t* prim_indices, int32_t* prims, int32_t* max_depth,
                   int32_t* blas_count, int32_t* instance_count) {
    const auto* s = static_cast<RefScene*>(h);
    const bool two = s->accel.kind == yr::RenderAccelerationKind::TwoLevel;
    *kind = two ? 1 : 0;
    if (!two) {
        *nodes = static_cast<int32_t>(s->accel.flat.nodes.size());
        *prim_indices = static_cast<int32_t>(s->accel.flat.primitive_indices.size());
        *prims = static_cast<int32_t>(s->accel.flat.primitives.size());
        *max_depth = s->accel.flat.max_depth;
        *blas_count = 0;
        *instance_count = 0;
    } else {
        const auto& t = s->accel.two_level.tlas;
        *nodes = static_cast<int32_t>(t.nodes.size());
        *prim_indices = static_cast<int32_t>(t.primitive_indices.size());
        *prims = static_cast<int32_t>(t.primitives.size());
        *max_depth = t.max_depth;
        *blas_count = static_cast<int32_t>(s->accel.two_level.blases.size());
        *instance_count = static_cast<int32_t>(s->accel.two_level.instances.size());
    }
    return 0;
}

// Copies one binary BVH (blas = -1: the flat BVH) as raw reference bytes:
// nodes 32 B each, primitive_indices int32, primitives 20 B each.
int ref_bvh_sizes(void* h, int32_t blas, int32_t* nodes, int32_t* prim_indices, int32_t* prims, int32_t* max_depth) {
    const auto* s = static_cast<RefScene*>(h);
    const yr::RenderBvh& b = blas < 0 ? s->accel.flat : s->accel.two_level.blases[static_cast<size_t>(blas)].bvh;
    *nodes = static_cast<int32_t>(b.nodes.size());
    *prim_indices = static_cast<int32_t>(b.primitive_indices.size());
    *prims = static_cast<int32_t>(b.primitives.size());
    *max_depth = b.max_depth;
    return 0;
}

int ref_bvh_copy(void* h, int32_t blas, void* nodes, int32_t* prim_indices, void* prims) {
    const auto* s = static_cast<RefScene*>(h);
    const yr::RenderBvh& b = blas < 0 ? s->accel.flat : s->accel.two_level.blases[static_cast<size_t>(blas)].bvh;
    static_assert(sizeof(yr::RenderBvhNode) == 32);
    std::memcpy(nodes, b.nodes.data(), b.nodes.size() * 32);
    std::memcpy(prim_indices, b.primitive_indices.data(), b.primitive_indices.size() * 4);
    auto* out = static_cast<int32_t*>(prims);
    for (size_t i = 0; i < b.primitives.size(); ++i) {
        const yr::RenderBvhPrimitive& p = b.primitives[i];
        out[5 * i + 0] = p.kind == yr::BvhPrimitiveKind::Triangle ? 0 : 1;
        out[5 * i + 1] = p.mesh_primitive.Value();
        out[5 * i + 2] = p.sphere.Value();
        out[5 * i + 3] = p.local_triangle;
        out[5 * i + 4] = p.flat_triangle_index;
    }
    return 0;
}

// TLAS nodes (32 B each) + primitive indices + instances as
// (primitive, blas, object_to_world[16], world_to_object[16], bounds[6]) = 40 floats/ints.
int ref_tlas_copy(void* h, void* nodes, int32_t* prim_indices, int32_t* prim_instance, float* instances) {
    const auto* s = static_cast<RefScene*>(h);
    const auto& tl = s->accel.two_level;
    std::memcpy(nodes, tl.tlas.nodes.data(), tl.tlas.nodes.size() * 32);
    std::memcpy(prim_indices, tl.tlas.primitive_indices.data(), tl.tlas.primitive_indices.size() * 4);
    for (size_t i = 0; i < tl.tlas.primitives.size(); ++i) prim_instance[i] = tl.tlas.primitives[i].instance.Value();
    for (size_t i = 0; i < tl.instances.size(); ++i) {
        const yr::PreparedMeshInstance& in = tl.instances[i];
        float* o = instances + 40 * i;
        const int32_t ids[2] = {in.primitive.Value(), in.blas_index};
        std::memcpy(o, ids, 8);
        std::memcpy(o + 2, in.object_to_world.m.data(), 64);
        std::memcpy(o + 18, in.world_to_object.m.data(), 64);
        const float b[6] = {in.world_bounds.min.x, in.world_bounds.min.y, in.world_bounds.min.z,
                            in.world_bounds.max.x, in.world_bounds.max.y, in.world_bounds.max.z};
        std::memcpy(o + 34, b, 24);
    }
    return 0;
}

// yr::IntersectAcceleration per ray over `threads` threads.
// hit: float4 (t, u, v, id bits) with id = triangle_index, or -2 - sphere for a
// sphere hit, or -1 for a miss; inst / counters may be null.
int ref_trace(void* h, const float* org_tmin, const float* dir_tmax, int64_t n, int32_t threads, float* hit,
              int32_t* inst, uint32_t* node_tests, uint32_t* tri_tests, uint32_t* sphere_tests) {
    const auto* s = static_cast<RefScene*>(h);
    const yr::GeometryView g = s->scene.Geometry();
    std::atomic_int64_t next{0};
    auto work = [&] {
        while (true) {
            const int64_t b = next.fetch_add(256);
            if (b >= n) break;
            const int64_t e = std::min(n, b + 256);
            for (int64_t i = b; i < e; ++i) {
                const yr::Ray3f ray{yr::Point3f{org_tmin[4 * i], org_tmin[4 * i + 1], org_tmin[4 * i + 2]},
                                    yr::Vec3f{dir_tmax[4 * i], dir_tmax[4 * i + 1], dir_tmax[4 * i + 2]}};
                yr::BvhTraceStats st;
                const yr::BvhHit hh = yr::IntersectAcceleration(g, s->accel, ray, st, org_tmin[4 * i + 3],
                                                                dir_tmax[4 * i + 3]);
                if (hit != nullptr) {
                    const int32_t id =
                        !hh.hit ? -1 : (hh.sphere.IsValid() ? -2 - hh.sphere.Value() : hh.triangle_index);
                    float fid;
                    std::memcpy(&fid, &id, 4);
                    hit[4 * i] = hh.t;
                    hit[4 * i + 1] = hh.bary_u;
                    hit[4 * i + 2] = hh.bary_v;
                    hit[4 * i + 3] = fid;
                }
                if (inst != nullptr) inst[i] = hh.hit ? hh.instance.Value() : -1;
                if (node_tests != nullptr) node_tests[i] = static_cast<uint32_t>(st.node_tests);
                if (tri_tests != nullptr) tri_tests[i] = static_cast<uint32_t>(st.triangle_tests);
                if (sphere_tests != nullptr) sphere_tests[i] = static_cast<uint32_t>(st.sphere_tests);
            }
        }
    };
    if (threads <= 1) {
        work();
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t) pool.emplace_back(work);
        for (auto& th : pool) th.join();
    }
    return 0;
}

// The reference's own render: RenderSceneIR built from the same inputs the GPU
// backend receives, PrepareCpuScene (light tables, BVH) and RenderCpuPathTrace
// for settings.spp passes. film: width*height FilmPixel (56 B) records.
// stats: rays, shadow, occluded, triangle tests, sphere tests, node tests,
// hits, misses, bad samples, threads; seconds = render seconds (excludes prepare).
// Adaptive-sampling settings the next ref_render calls use
// (yr::AdaptiveSamplingSettings, include/yaoray/scene/render_options.hpp:30-36),
// and the converged-pixel count of the last ref_render.
static yr::AdaptiveSamplingSettings g_adaptive{};
static uint64_t g_last_converged = 0;

void ref_set_adaptive(int32_t enabled, int32_t min_spp, float relative_error, float absolute_error, float confidence) {
    g_adaptive.enabled = enabled != 0;
    g_adaptive.min_spp = min_spp;
    g_adaptive.relative_error = relative_error;
    g_adaptive.absolute_error = absolute_error;
    g_adaptive.confidence = confidence;
}

uint64_t ref_last_converged() { return g_last_converged; }

int ref_render(const yrh_vertex* vertices, int64_t nv, const uint32_t* indices, int64_t ni,
               const yrh_primitive* primitives, int32_t np, const int32_t* inst_prim, const float* inst_mat,
               int32_t ninst, const yrh_shading_desc* sd, const float* camera /* origin, forward, right, up, fov */,
               const yrh_render_desc* rd, int32_t threads, void* film_out, uint64_t* stats, double* seconds) {
    yr::RenderJob job;
    yr::RenderSceneIR& sc = job.scene;
    sc.vertices.resize(static_cast<size_t>(nv));
    if (nv > 0) std::memcpy(sc.vertices.data(), vertices, static_cast<size_t>(nv) * 48);
    sc.indices.assign(indices, indices + ni);
    sc.primitives.resize(static_cast<size_t>(np));
    if (np > 0) std::memcpy(sc.primitives.data(), primitives, static_cast<size_t>(np) * 16);
    for (int32_t i = 0; i < ninst; ++i) {
        yr::RenderInstance inst;
        inst.primitive = yr::MeshPrimitiveHandle{inst_prim[i]};
        std::memcpy(inst.object_to_world.m.data(), inst_mat + 16 * static_cast<size_t>(i), 64);
        sc.instances.push_back(inst);
    }
    sc.spheres = g_spheres;
    sc.analytic_lights = g_analytic;
    const int32_t ntex = sd->textures != nullptr ? sd->texture_count : 0;
    for (int32_t k = 0; k < ntex; ++k) sc.textures.push_back(ToRefTexture(sd->textures[k]));
    for (int32_t k = 0; k < sd->material_count; ++k) sc.materials.push_back(ToRefMaterial(sd->materials[k], ntex));
    for (int32_t k = 0; k < sd->emissive_count; ++k) {
        const yrh_emissive& e = sd->emissives[k];
        sc.emissive_primitives.push_back(
            yr::EmissivePrimitive{e.primitive_index, yr::Color3f{e.radiance[0], e.radiance[1], e.radiance[2]}, e.area});
    }
    const yrh_environment_desc& env = sd->environment;
    sc.environment.active = env.active != 0;
    sc.environment.radiance = yr::Color3f{env.radiance[0], env.radiance[1], env.radiance[2]};
    sc.environment.strength = env.strength;
    sc.environment.rotation_radians = env.rotation_radians;
    if (env.texels != nullptr && env.width > 0 && env.height > 0) {
        yr::RenderTexture tex;
        tex.width = env.width;
        tex.height = env.height;
        tex.filter = env.nearest ? yr::TextureFilter::Nearest : yr::TextureFilter::Bilinear;
        tex.wrap_s = env.wrap_s == YRH_WRAP_CLAMP ? yr::TextureWrap::ClampToEdge
                   : env.wrap_s == YRH_WRAP_MIRRORED ? yr::TextureWrap::MirroredRepeat : yr::TextureWrap::Repeat;
        tex.wrap_t = env.wrap_t == YRH_WRAP_CLAMP ? yr::TextureWrap::ClampToEdge
                   : env.wrap_t == YRH_WRAP_MIRRORED ? yr::TextureWrap::MirroredRepeat : yr::TextureWrap::Repeat;
        tex.texels.resize(static_cast<size_t>(env.width) * static_cast<size_t>(env.height));
        std::memcpy(tex.texels.data(), env.texels, tex.texels.size() * 16);
        sc.textures.push_back(std::move(tex));
        sc.environment.texture_index = ntex;
        if (env.has_distribution) {
            sc.environment_distributions.push_back(yr::BuildEnvironmentDistribution(sc.textures[static_cast<size_t>(ntex)]));
            sc.environment.distribution_index = 0;
        }
    }
    sc.camera.origin = yr::Point3f{camera[0], camera[1], camera[2]};
    sc.camera.forward = yr::Vec3f{camera[3], camera[4], camera[5]};
    sc.camera.right = yr::Vec3f{camera[6], camera[7], camera[8]};
    sc.camera.up = yr::Vec3f{camera[9], camera[10], camera[11]};
    sc.camera.fov_y_radians = camera[12];
    yr::RenderSettings& st = job.settings;
    st.width = rd->width;
    st.height = rd->height;
    st.spp = rd->spp;
    st.max_depth = rd->max_depth;
    st.seed = rd->seed;
    st.threads = threads;
    st.radiance_clamp = rd->radiance_clamp;
    st.adaptive_sampling = g_adaptive;
    st.sampler = rd->sampler == YRH_SAMPLER_INDEPENDENT ? yr::RenderSamplerKind::Independent
               : rd->sampler == YRH_SAMPLER_STRATIFIED ? yr::RenderSamplerKind::Stratified
               : rd->sampler == YRH_SAMPLER_OWEN_SOBOL ? yr::RenderSamplerKind::OwenSobol
                                                       : yr::RenderSamplerKind::ZSobol;
    yr::CpuPrepareResult prep = yr::PrepareCpuScene(std::move(job));
    if (!prep.ok) return -1;
    const yr::CpuPathTraceResult r = yr::RenderCpuPathTrace(*prep.scene);
    if (!r.ok) return -2;
    g_last_converged = r.stats.converged_pixels;
    static_assert(sizeof(yr::FilmPixel) == 56);
    std::memcpy(film_out, r.film.Pixels().data(), r.film.Pixels().size() * 56);
    const uint64_t vals[10] = {r.stats.rays_traced, r.stats.shadow_rays, r.stats.occluded_shadow_rays,
                               r.stats.triangle_tests, r.stats.sphere_tests, r.stats.bvh_node_tests, r.stats.hits,
                               r.stats.misses, r.film.BadSampleCount(), static_cast<uint64_t>(r.stats.threads)};
    std::memcpy(stats, vals, sizeof(vals));
    if (seconds != nullptr) *seconds = r.stats.elapsed_seconds;
    return 0;
}

// yr::SampleTexture4 (src/shading/texture.cpp:312-322) of texture `t` after
// BuildTextureSamplingCaches (its mip chain), at n points: uv (n x 2),
// footprints (n x 4: dudx, dvdx, dudy, dvdy; NULL = none); out n x 4.
int ref_sample_texture(const yrh_texture* t, int64_t n, const float* uv, const float* footprint, float* out) {
    std::vector<yr::RenderTexture> texs{ToRefTexture(*t)};
    yr::BuildTextureSamplingCaches(texs);
    for (int64_t i = 0; i < n; ++i) {
        yr::TextureFootprint f;
        if (footprint != nullptr) {
            f.dudx = footprint[4 * i];
            f.dvdx = footprint[4 * i + 1];
            f.dudy = footprint[4 * i + 2];
            f.dvdy = footprint[4 * i + 3];
        }
        const yr::Color4f c = yr::SampleTexture4(texs[0], yr::Vec2f{uv[2 * i], uv[2 * i + 1]}, f);
        out[4 * i] = c.x;
        out[4 * i + 1] = c.y;
        out[4 * i + 2] = c.z;
        out[4 * i + 3] = c.w;
    }
    return 0;
}

// The mip chain yr::BuildTextureMipChain (texture.cpp:241-272) makes for `t`:
// level count (with level 0) and all texels, as oracle_texture_mips.
int ref_texture_mips(const yrh_texture* t, int32_t* levels, int64_t* texels, float* out) {
    yr::RenderTexture r = ToRefTexture(*t);
    yr::BuildTextureMipChain(r);
    *levels = r.texels.empty() ? 0 : 1 + static_cast<int32_t>(r.mip_levels.size());
    int64_t k = 0;
    auto put = [&](const std::vector<yr::Color4f>& v) {
        for (const yr::Color4f& c : v) {
            if (out != nullptr) {
                out[4 * k] = c.x;
                out[4 * k + 1] = c.y;
                out[4 * k + 2] = c.z;
                out[4 * k + 3] = c.w;
            }
            ++k;
        }
    };
    put(r.texels);
    for (const yr::TextureMipLevel& l : r.mip_levels) put(l.texels);
    *texels = k;
    return 0;
}

void ref_inverse(const float* m, float* out) {
    yr::Mat4f a;
    std::memcpy(a.m.data(), m, 64);
    const yr::Mat4f r = yr::Inverse(a);
    std::memcpy(out, r.m.data(), 64);
}

// yr::Sampler values at bounce 0 (include/yaoray/sampling/sampler.hpp:25-50):
// dims == 1 -> Sample1D(dim[i]), dims == 2 -> Sample2D(dim[i]) into out[2i..].
int ref_sample_values(int32_t kind, uint64_t seed, int32_t spp, int64_t n, const int32_t* x, const int32_t* y,
                      const int32_t* sample, const int32_t* dim, int32_t dims, float* out) {
    const yr::RenderSamplerKind k = kind == YRH_SAMPLER_INDEPENDENT ? yr::RenderSamplerKind::Independent
                                    : kind == YRH_SAMPLER_STRATIFIED ? yr::RenderSamplerKind::Stratified
                                    : kind == YRH_SAMPLER_OWEN_SOBOL ? yr::RenderSamplerKind::OwenSobol
                                                                     : yr::RenderSamplerKind::ZSobol;
    for (int64_t i = 0; i < n; ++i) {
        const yr::Sampler s(k, seed, x[i], y[i], sample[i], spp, 1);
        const auto d = static_cast<yr::SampleDimension>(static_cast<uint32_t>(dim[i]));
        if (dims == 1) {
            out[i] = s.Sample1D(d);
        } else {
            const yr::Vec2f v = s.Sample2D(d);
            out[2 * i] = v.x;
            out[2 * i + 1] = v.y;
        }
    }
    return 0;
}

// The reference's EvaluateBsdf / PdfBsdf / SampleBsdf / IsDeltaBsdf
// (src/shading/bsdf.cpp:17-329) on the query layout of oracle_bsdf_query:
// q[12i..] = wo, wi, normal, u (2), unused; out[12i..] = f (3), pdf, sampled
// wi (3), weight (3), pdf, flags (valid + 2 specular + 4 delta). A fresh
// Rng{1} per query (only layered BSDFs read it).
int ref_bsdf_query(const yrh_material* mats, int32_t nmat, int64_t n, const int32_t* mat, const float* q,
                   float* out) {
    std::vector<yr::RenderMaterial> rm;
    for (int32_t k = 0; k < nmat; ++k) rm.push_back(ToRefMaterial(mats[k]));
    for (int64_t i = 0; i < n; ++i) {
        if (mat[i] < 0 || mat[i] >= nmat) return -1;
        const yr::RenderMaterial& m = rm[static_cast<size_t>(mat[i])];
        const float* a = q + 12 * i;
        const yr::Vec3f wo{a[0], a[1], a[2]}, wi{a[3], a[4], a[5]}, nn{a[6], a[7], a[8]};
        yr::Rng rng{1u};
        const yr::Color3f f = yr::EvaluateBsdf(m, wo, wi, nn, rng);
        const float pdf = yr::PdfBsdf(m, wo, wi, nn, rng);
        const yr::BsdfSample s = yr::SampleBsdf(m, wo, nn, yr::Vec2f{a[9], a[10]}, rng);
        const float vals[12] = {f.x, f.y, f.z, pdf, s.wi.x, s.wi.y, s.wi.z, s.weight.x, s.weight.y, s.weight.z, s.pdf,
                                static_cast<float>((s.valid ? 1 : 0) + (s.specular ? 2 : 0) +
                                                   (yr::IsDeltaBsdf(m) ? 4 : 0))};
        std::memcpy(out + 12 * i, vals, sizeof(vals));
    }
    return 0;
}

#pragma GCC visibility pop
}  // extern "C"
