#include "abi_marshal.hpp"

#include <cstddef>
#include <string>

namespace yrhip {

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
static_assert(sizeof(yr::Color4f) == 16);
static_assert(sizeof(yr::RenderSphere) == sizeof(yrh_sphere));
static_assert(offsetof(yr::RenderSphere, radius) == offsetof(yrh_sphere, radius));
static_assert(offsetof(yr::RenderSphere, material_index) == offsetof(yrh_sphere, material_index));
static_assert(offsetof(yr::RenderSphere, flip_normals) == offsetof(yrh_sphere, flip_normals));
static_assert(sizeof(yr::AnalyticLight) == sizeof(yrh_analytic_light));
static_assert(offsetof(yr::AnalyticLight, position) == offsetof(yrh_analytic_light, position));
static_assert(offsetof(yr::AnalyticLight, direction) == offsetof(yrh_analytic_light, direction));
static_assert(offsetof(yr::AnalyticLight, intensity) == offsetof(yrh_analytic_light, intensity));
static_assert(offsetof(yr::AnalyticLight, cone_cos_inner) == offsetof(yrh_analytic_light, cone_cos_inner));
static_assert(static_cast<int>(yr::AnalyticLightKind::Point) == YRH_LIGHT_POINT);
static_assert(static_cast<int>(yr::AnalyticLightKind::Spot) == YRH_LIGHT_SPOT);
static_assert(static_cast<int>(yr::AnalyticLightKind::Distant) == YRH_LIGHT_DISTANT);
static_assert(sizeof(yr::LightTreeNode) == sizeof(yrh_light_tree_node));
static_assert(offsetof(yr::LightTreeNode, power) == offsetof(yrh_light_tree_node, power));
static_assert(offsetof(yr::LightTreeNode, light_index) == offsetof(yrh_light_tree_node, light_index));

static_assert(static_cast<int>(yr::TextureFilter::Nearest) == YRH_FILTER_NEAREST);
static_assert(static_cast<int>(yr::TextureFilter::Bilinear) == YRH_FILTER_BILINEAR);
static_assert(static_cast<int>(yr::TextureFilter::Trilinear) == YRH_FILTER_TRILINEAR);
static_assert(static_cast<int>(yr::TextureFilter::Ewa) == YRH_FILTER_EWA);
static_assert(static_cast<int>(yr::TextureWrap::Repeat) == YRH_WRAP_REPEAT);
static_assert(static_cast<int>(yr::TextureWrap::ClampToEdge) == YRH_WRAP_CLAMP);
static_assert(static_cast<int>(yr::TextureWrap::MirroredRepeat) == YRH_WRAP_MIRRORED);

namespace {

// Level-0 texels: the reference moves them into a shared sampling cache once
// BuildTextureSamplingCaches ran (src/shading/texture.cpp:274-298).
const std::vector<yr::Color4f>& Texels(const yr::RenderTexture& t) {
#if defined(YAORAY_HIP_DROPIN)
    if (t.sampling_cache) return t.sampling_cache->texels;
#endif
    return t.texels;
}

}  // namespace

yrh_bvh_desc DescribeBvh(const yr::RenderBvh& b) {
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

void DescribeGeometry(const yr::RenderSceneIR& ir, const yr::RenderAcceleration& accel, Marshal& m,
                      yrh_geometry_desc& g) {
    g = yrh_geometry_desc{};
    g.vertices = reinterpret_cast<const yrh_vertex*>(ir.vertices.data());
    g.vertex_count = static_cast<int64_t>(ir.vertices.size());
    g.indices = ir.indices.data();
    g.index_count = static_cast<int64_t>(ir.indices.size());
    g.primitives = reinterpret_cast<const yrh_primitive*>(ir.primitives.data());
    g.primitive_count = static_cast<int32_t>(ir.primitives.size());
    g.spheres = reinterpret_cast<const yrh_sphere*>(ir.spheres.data());
    g.sphere_count = static_cast<int32_t>(ir.spheres.size());
    if (accel.kind == yr::RenderAccelerationKind::FlatReference) {
        g.accel_kind = YRH_ACCEL_FLAT;
        g.flat = DescribeBvh(accel.flat);
        return;
    }
    const yr::TwoLevelBvh& tl = accel.two_level;
    m.blas.clear();
    for (const yr::RenderBlas& b : tl.blases) m.blas.push_back(DescribeBvh(b.bvh));
    g.accel_kind = YRH_ACCEL_TWO_LEVEL;
    g.blases = m.blas.data();
    g.blas_count = static_cast<int32_t>(m.blas.size());
    g.instances = reinterpret_cast<const yrh_instance*>(tl.instances.data());
    g.instance_count = static_cast<int32_t>(tl.instances.size());
    g.tlas = yrh_bvh_desc{};
    g.tlas.nodes = reinterpret_cast<const yrh_bvh_node*>(tl.tlas.nodes.data());
    g.tlas.node_count = static_cast<int32_t>(tl.tlas.nodes.size());
    g.tlas.primitive_indices = tl.tlas.primitive_indices.data();
    g.tlas.primitive_index_count = static_cast<int32_t>(tl.tlas.primitive_indices.size());
    g.tlas.max_depth = tl.tlas.max_depth;
    g.tlas_primitives = reinterpret_cast<const yrh_tlas_primitive*>(tl.tlas.primitives.data());
    g.tlas_primitive_count = static_cast<int32_t>(tl.tlas.primitives.size());
}

std::string DescribeShading(const yr::RenderSceneIR& ir, Marshal& m, yrh_shading_desc& d) {
    m.materials.clear();
    const int ntex = static_cast<int>(ir.textures.size());
    std::vector<char> used(static_cast<size_t>(ntex), 0);
    auto tex = [&](int k) {
        if (k < 0 || k >= ntex) return -1;  // TextureIndexValid (material_evaluator.cpp:19-21)
        used[static_cast<size_t>(k)] = 1;
        return k;
    };
    for (const yr::RenderMaterial& mat : ir.materials) {
        yrh_material a{};
        switch (mat.kind) {
            case yr::RenderMaterialKind::Diffuse: a.kind = YRH_MATERIAL_DIFFUSE; break;
            case yr::RenderMaterialKind::Conductor: a.kind = YRH_MATERIAL_CONDUCTOR; break;
            case yr::RenderMaterialKind::Dielectric: a.kind = YRH_MATERIAL_DIELECTRIC; break;
            case yr::RenderMaterialKind::ThinDielectric: a.kind = YRH_MATERIAL_THIN_DIELECTRIC; break;
            default: return "material kind outside the HIP backend's set";
        }
        // Only reflectance, alpha and the normal map are textured at a hit
        // (EvaluateMaterialAtSurface, material_evaluator.cpp:137-177); the
        // other parameters' texture slots are not read there either.
        a.reflectance_texture = tex(mat.reflectance.texture);
        a.alpha_texture = tex(mat.alpha.texture);
        a.normal_map = tex(mat.normal_map);
        a.normal_scale = mat.normal_scale;
        a.reflectance[0] = mat.reflectance.value.x;
        a.reflectance[1] = mat.reflectance.value.y;
        a.reflectance[2] = mat.reflectance.value.z;
        a.ior = mat.ior;
        a.uroughness = mat.uroughness.value;
        a.vroughness = mat.vroughness.value;
        a.alpha = mat.alpha.value;
        a.emission[0] = mat.emission.x;
        a.emission[1] = mat.emission.y;
        a.emission[2] = mat.emission.z;
        a.absorption_distance = mat.absorption_distance;
        a.absorption_color[0] = mat.absorption_color.x;
        a.absorption_color[1] = mat.absorption_color.y;
        a.absorption_color[2] = mat.absorption_color.z;
        m.materials.push_back(a);
    }
    m.emissives.clear();
    for (const yr::EmissivePrimitive& e : ir.emissive_primitives) {
        m.emissives.push_back(yrh_emissive{e.primitive_index, {e.radiance.x, e.radiance.y, e.radiance.z}, e.area});
    }
    d = yrh_shading_desc{};
    d.materials = m.materials.data();
    d.material_count = static_cast<int32_t>(m.materials.size());
    d.emissives = m.emissives.data();
    d.emissive_count = static_cast<int32_t>(m.emissives.size());
    const yr::AliasTable& power = ir.light_sampling.emissive_power;
    d.alias_probability = power.probability.data();
    d.alias_index = power.alias.data();
    d.alias_pmf = power.pmf.data();
    const yr::LightSamplingCache& cache = ir.light_sampling;
    d.analytic_lights = reinterpret_cast<const yrh_analytic_light*>(ir.analytic_lights.data());
    d.analytic_light_count = static_cast<int32_t>(ir.analytic_lights.size());
    if (cache.analytic_power.pmf.size() == ir.analytic_lights.size()) {
        d.analytic_alias_probability = cache.analytic_power.probability.data();
        d.analytic_alias_index = cache.analytic_power.alias.data();
        d.analytic_alias_pmf = cache.analytic_power.pmf.data();
    }
    d.analytic_tree = reinterpret_cast<const yrh_light_tree_node*>(cache.analytic_tree.data());
    d.analytic_tree_count = static_cast<int32_t>(cache.analytic_tree.size());
    d.analytic_tree_root = cache.analytic_tree_root;
    d.use_analytic_tree = cache.use_analytic_tree ? 1 : 0;
    // Every texture keeps its index; ones no material reads (the environment
    // map) go without texels, which nothing samples.
    m.textures.assign(static_cast<size_t>(ntex), yrh_texture{});
    for (int k = 0; k < ntex; ++k) {
        const yr::RenderTexture& t = ir.textures[static_cast<size_t>(k)];
        yrh_texture& o = m.textures[static_cast<size_t>(k)];
        o.width = t.width;
        o.height = t.height;
        o.filter = static_cast<int32_t>(t.filter);
        o.wrap_s = static_cast<int32_t>(t.wrap_s);
        o.wrap_t = static_cast<int32_t>(t.wrap_t);
        const std::vector<yr::Color4f>& texels = Texels(t);
        if (used[static_cast<size_t>(k)] && !texels.empty()) {
            if (static_cast<int64_t>(texels.size()) < static_cast<int64_t>(t.width) * static_cast<int64_t>(t.height)) {
                return "texture has fewer texels than width * height";
            }
            o.texels = reinterpret_cast<const float*>(texels.data());
        }
    }
    d.textures = m.textures.data();
    d.texture_count = ntex;
    yrh_environment_desc& e = d.environment;
    const yr::RenderEnvironment& env = ir.environment;
    e.active = env.active ? 1 : 0;
    e.radiance[0] = env.radiance.x;
    e.radiance[1] = env.radiance.y;
    e.radiance[2] = env.radiance.z;
    e.strength = env.strength;
    e.rotation_radians = env.rotation_radians;
    if (env.texture_index >= 0 && env.texture_index < static_cast<int>(ir.textures.size())) {
        const yr::RenderTexture& t = ir.textures[static_cast<size_t>(env.texture_index)];
        // EvaluateEnvironment samples without a footprint (environment.cpp:136):
        // bilinear unless the map is nearest-filtered, with its wrap modes.
        e.nearest = t.filter == yr::TextureFilter::Nearest ? 1 : 0;
        e.wrap_s = static_cast<int32_t>(t.wrap_s);
        e.wrap_t = static_cast<int32_t>(t.wrap_t);
        e.width = t.width;
        e.height = t.height;
        e.texels = reinterpret_cast<const float*>(Texels(t).data());
    }
    if (env.distribution_index >= 0 && env.distribution_index < static_cast<int>(ir.environment_distributions.size())) {
        const yr::RenderEnvironmentDistribution& dist = ir.environment_distributions[static_cast<size_t>(env.distribution_index)];
        e.has_distribution = 1;
        e.dist_width = dist.width;
        e.dist_height = dist.height;
        e.texel_weights = dist.texel_weights.data();
        e.row_weights = dist.row_weights.data();
        e.row_cdf = dist.row_cdf.data();
        e.conditional_cdfs = dist.conditional_cdfs.data();
        e.total_weight = dist.total_weight;
    }
    return {};
}

}  // namespace yrhip
