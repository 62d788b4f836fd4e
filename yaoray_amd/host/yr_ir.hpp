// Host-side data contracts of the wavefront path-tracing backend.
//
// These types carry the reference's names, field order and byte layout for the
// part of its scene IR the hot path consumes, so HipBackend reads like a
// reference backend and the arrays it uploads are the reference's own:
//   Vec/Bounds/Mat4f     include/yaoray/core/{vec,bounds,transform}.hpp
//   RenderVertex etc.    include/yaoray/scene/geometry.hpp:12-71
//   RenderMaterial       include/yaoray/scene/material.hpp:7-65 (in-scope fields)
//   lights / env         include/yaoray/scene/light.hpp:10-67
//   camera / settings    include/yaoray/scene/{camera,render_options}.hpp
//   BVH / two-level      include/yaoray/accel/{bvh,two_level_bvh,acceleration}.hpp
// Inside the reference tree these would be the reference headers themselves;
// this repository carries its own copy of the contracts so it builds standalone.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <span>
#include <string>
#include <vector>

namespace yr {

struct Vec2f {
    float x = 0.0f;
    float y = 0.0f;
};

struct Vec3f {
    float x = 0.0f;
    float y = 0.0f;
    float z = 0.0f;
    constexpr Vec3f operator-() const { return Vec3f{-x, -y, -z}; }
};
using Point3f = Vec3f;
using Color3f = Vec3f;

struct Color4f {
    float x = 0.0f;
    float y = 0.0f;
    float z = 0.0f;
    float w = 1.0f;
};

constexpr Vec3f operator+(Vec3f a, Vec3f b) { return Vec3f{a.x + b.x, a.y + b.y, a.z + b.z}; }
constexpr Vec3f operator-(Vec3f a, Vec3f b) { return Vec3f{a.x - b.x, a.y - b.y, a.z - b.z}; }
constexpr Vec3f operator*(Vec3f v, float s) { return Vec3f{v.x * s, v.y * s, v.z * s}; }
constexpr Vec3f operator*(float s, Vec3f v) { return v * s; }
constexpr Vec3f operator/(Vec3f v, float s) { return Vec3f{v.x / s, v.y / s, v.z / s}; }
constexpr float Dot(Vec3f a, Vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
constexpr Vec3f Cross(Vec3f a, Vec3f b) {
    return Vec3f{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float LengthSquared(Vec3f v) { return Dot(v, v); }
inline float Length(Vec3f v) { return std::sqrt(LengthSquared(v)); }
inline Vec3f Normalize(Vec3f v) {
    const float len = Length(v);
    return len > 0.0f ? v / len : Vec3f{};
}

struct Bounds3f {
    Point3f min{std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                std::numeric_limits<float>::infinity()};
    Point3f max{-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                -std::numeric_limits<float>::infinity()};
};

struct Mat4f {
    std::array<float, 16> m{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
};

Mat4f Multiply(Mat4f a, Mat4f b);
Mat4f Inverse(Mat4f m);
Point3f TransformPoint(Mat4f t, Point3f p);
Vec3f TransformVector(Mat4f t, Vec3f v);
Mat4f TranslationMatrix(Vec3f t);
Mat4f ScaleMatrix(Vec3f s);
Mat4f RotationAxisMatrix(float angle_degrees, Vec3f axis);

// ---- scene ----------------------------------------------------------------

struct RenderVertex {
    Point3f position;
    Vec3f normal;
    Vec2f uv;
    Vec3f tangent;
    float tangent_handedness = 1.0f;
};
static_assert(sizeof(RenderVertex) == 48);

struct RenderPrimitive {
    std::uint32_t first_index = 0;
    std::uint32_t index_count = 0;
    int material_index = 0;
    bool has_normals = false;
    bool has_uvs = false;
    bool has_tangents = false;
    // The reference's padding byte, named so that copies of the array (and
    // digests of its bytes) are deterministic.
    std::uint8_t reserved = 0;
};
static_assert(sizeof(RenderPrimitive) == 16);

// Analytic sphere (geometry.hpp:29-35): world-space centre and radius;
// flip_normals turns the outward normal inward.
struct RenderSphere {
    Point3f center{0.0f, 0.0f, 0.0f};
    float radius = 1.0f;
    int material_index = 0;
    int area_light_index = -1;
    bool flip_normals = false;
};

struct RenderInstance {
    int primitive = -1;  // MeshPrimitiveHandle value
    Mat4f object_to_world;
};

struct GeometryView {
    std::span<const RenderVertex> vertices;
    std::span<const std::uint32_t> indices;
    std::span<const RenderPrimitive> primitives;
    std::span<const RenderSphere> spheres;
    std::span<const RenderInstance> instances{};
};

enum class RenderMaterialKind {
    Diffuse,
    Conductor,
    Dielectric,
    ThinDielectric,
    CoatedDiffuse,
    CoatedConductor,
    DiffuseTransmission,
    Mix,
    Measured,
    Subsurface,
    SubsurfaceExit,
};

// A material parameter with an optional texture index into
// RenderSceneIR::textures (material.hpp:21-30; -1 = none).
struct TexParam1f {
    float value = 0.0f;
    int texture = -1;
};

struct TexParam3f {
    Color3f value{0.0f, 0.0f, 0.0f};
    int texture = -1;
};

// The fields of RenderMaterial (material.hpp:32-65) the in-scope kinds read,
// with the reference's names and types: diffuse / conductor / dielectric /
// thin dielectric with emission, alpha and interior absorption.
struct RenderMaterial {
    RenderMaterialKind kind = RenderMaterialKind::Diffuse;
    TexParam3f reflectance{{0.5f, 0.5f, 0.5f}};
    float ior = 1.5f;
    TexParam1f uroughness{0.0f};
    TexParam1f vroughness{0.0f};
    int normal_map = -1;
    float normal_scale = 1.0f;
    Color3f emission{0.0f, 0.0f, 0.0f};
    TexParam1f alpha{1.0f};
    Color3f absorption_color{1.0f, 1.0f, 1.0f};
    float absorption_distance = 1.0f;
};

struct EmissivePrimitive {
    int primitive_index = 0;
    Color3f radiance{0.0f, 0.0f, 0.0f};
    float area = 0.0f;
};

struct AliasTable {
    std::vector<float> probability;
    std::vector<int> alias;
    std::vector<float> pmf;
};

// AnalyticLight (light.hpp:36-45): point, spot or distant light.
enum class AnalyticLightKind { Point, Spot, Distant };

struct AnalyticLight {
    AnalyticLightKind kind = AnalyticLightKind::Point;
    Point3f position;
    Vec3f direction;
    Color3f intensity;
    float cone_angle = 0.0f;      // cosine of the outer cone
    float cone_cos_inner = 0.0f;  // cosine of the full-intensity cone
};

// LightTreeNode (light.hpp:53-59).
struct LightTreeNode {
    Bounds3f bounds;
    float power = 0.0f;
    int left = -1;
    int right = -1;
    int light_index = -1;
};

// LightSamplingCache (light.hpp:61-67): the emissive-power alias table, the
// analytic-power alias table and the analytic light tree.
struct LightSamplingCache {
    AliasTable emissive_power;
    AliasTable analytic_power;
    std::vector<LightTreeNode> analytic_tree;
    int analytic_tree_root = -1;
    bool use_analytic_tree = false;
};

// TextureFilter / TextureWrap (include/yaoray/scene/texture_options.hpp:7-18).
enum class TextureFilter { Nearest, Bilinear, Trilinear, Ewa };
enum class TextureWrap { Repeat, ClampToEdge, MirroredRepeat };

// RenderTexture (texture.hpp:38-51): linear RGBA texels of level 0 with the
// filter and wrap modes; mip chains are built where the texture is sampled.
struct RenderTexture {
    int width = 0;
    int height = 0;
    std::vector<Color4f> texels;
    TextureFilter filter = TextureFilter::Bilinear;
    TextureWrap wrap_s = TextureWrap::Repeat;
    TextureWrap wrap_t = TextureWrap::Repeat;
};

struct RenderEnvironment {
    bool active = false;
    Color3f radiance{0.0f, 0.0f, 0.0f};
    float strength = 1.0f;
    float rotation_radians = 0.0f;
    int texture_index = -1;
    int distribution_index = -1;
};

struct RenderEnvironmentDistribution {
    int width = 0;
    int height = 0;
    std::vector<float> texel_weights;
    std::vector<float> row_weights;
    std::vector<float> row_cdf;
    std::vector<float> conditional_cdfs;
    float total_weight = 0.0f;
    bool uniform = false;
};

struct RenderCamera {
    Point3f origin;
    Vec3f forward{0.0f, 0.0f, -1.0f};
    Vec3f right{1.0f, 0.0f, 0.0f};
    Vec3f up{0.0f, 1.0f, 0.0f};
    float fov_y_radians = 0.785398185f;
};

enum class RenderBackendKind { Cpu, Cuda, Hip };
enum class RenderSamplerKind { Independent, Stratified, OwenSobol, ZSobol };

struct AdaptiveSamplingSettings {
    bool enabled = false;
    int min_spp = 16;
    float relative_error = 0.05f;
    float absolute_error = 0.001f;
    float confidence = 1.96f;
};

struct RenderSettings {
    RenderBackendKind requested_backend = RenderBackendKind::Hip;
    RenderSamplerKind sampler = RenderSamplerKind::Independent;
    int width = 0;
    int height = 0;
    int spp = 1;
    int max_depth = 5;
    std::uint64_t seed = 0;
    int threads = 0;
    float radiance_clamp = 0.0f;
    AdaptiveSamplingSettings adaptive_sampling;
};

struct RenderSceneIR {
    RenderCamera camera;
    std::vector<RenderVertex> vertices;
    std::vector<std::uint32_t> indices;
    std::vector<RenderPrimitive> primitives;
    std::vector<RenderSphere> spheres;
    std::vector<RenderInstance> instances;
    std::vector<RenderMaterial> materials;
    std::vector<RenderTexture> textures;
    std::vector<EmissivePrimitive> emissive_primitives;
    std::vector<AnalyticLight> analytic_lights;
    RenderEnvironment environment;
    std::vector<RenderEnvironmentDistribution> environment_distributions;
    LightSamplingCache light_sampling;

    GeometryView Geometry() const { return GeometryView{vertices, indices, primitives, spheres, instances}; }
};

// ---- acceleration ------------------------------------------------------------

enum class BvhSplitMethod { LongestAxisMedian, SahBucketBinning };

struct BvhBuildOptions {
    BvhSplitMethod split_method = BvhSplitMethod::SahBucketBinning;
    int max_leaf_triangles = 4;
    int thread_count = 0;
    int parallel_min_subtree_size = 1024;
};

struct alignas(32) RenderBvhNode {
    Bounds3f bounds;
    int payload_offset = -1;
    int primitive_count = 0;
    bool IsLeaf() const { return primitive_count > 0; }
};
static_assert(sizeof(RenderBvhNode) == 32);

enum class BvhPrimitiveKind : std::int32_t { Triangle, Sphere };

struct RenderBvhPrimitive {
    BvhPrimitiveKind kind = BvhPrimitiveKind::Triangle;
    int mesh_primitive = -1;
    int sphere = -1;
    int local_triangle = -1;
    int flat_triangle_index = -1;
};
static_assert(sizeof(RenderBvhPrimitive) == 20);

struct RenderBvh {
    std::vector<RenderBvhNode> nodes;
    std::vector<int> primitive_indices;
    std::vector<RenderBvhPrimitive> primitives;
    int max_depth = 0;
    int total_primitives = 0;
    int total_triangles = 0;
    int total_spheres = 0;
};

struct RenderBlas {
    int primitive = -1;
    Bounds3f object_bounds;
    RenderBvh bvh;
};

struct PreparedMeshInstance {
    int primitive = -1;
    int blas_index = -1;
    Mat4f object_to_world;
    Mat4f world_to_object;
    Bounds3f world_bounds;
};

enum class TlasPrimitiveKind : std::int32_t { Instance, Sphere };

struct RenderTlasPrimitive {
    TlasPrimitiveKind kind = TlasPrimitiveKind::Instance;
    int instance = -1;
    int sphere = -1;
    Bounds3f world_bounds;
};

struct RenderTlas {
    std::vector<RenderBvhNode> nodes;
    std::vector<int> primitive_indices;
    std::vector<RenderTlasPrimitive> primitives;
    int max_depth = 0;
};

struct TwoLevelBvh {
    std::vector<RenderBlas> blases;
    std::vector<PreparedMeshInstance> instances;
    RenderTlas tlas;
};

enum class RenderAccelerationKind { FlatReference, TwoLevel };

struct RenderAcceleration {
    RenderAccelerationKind kind = RenderAccelerationKind::FlatReference;
    RenderBvh flat;
    TwoLevelBvh two_level;
    int NodeCount() const;
    int MaxDepth() const;
};

struct BvhBuildResult {
    RenderBvh bvh;
    std::vector<std::string> errors;
};

struct RenderAccelerationBuildResult {
    RenderAcceleration acceleration;
    std::vector<std::string> errors;
};

BvhBuildResult BuildBvh(GeometryView geometry, const BvhBuildOptions& options = {});
RenderAccelerationBuildResult BuildRenderAcceleration(GeometryView geometry, const BvhBuildOptions& options = {});

// ---- light tables ---------------------------------------------------------

AliasTable BuildAliasTable(std::span<const float> weights);
void PrepareLightSampling(RenderSceneIR& scene);
RenderEnvironmentDistribution BuildEnvironmentDistribution(const RenderTexture& texture);

}  // namespace yr
