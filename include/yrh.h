/*
 * yrh.h — C ABI of the MI355X (gfx950) wavefront ray-casting / path-tracing
 * library `libyrh.so`.
 *
 * This is the only surface through which host code (the C++20 HipBackend in
 * yaoray_amd/host, the Python bindings, tests and benchmarks) reaches HIP.
 * Signatures use plain C types, pointers and sizes only.
 *
 * Conventions
 *   - Every function returns an int status: YRH_OK (0) or a negative YRH_E_*
 *     code; yrh_last_error(ctx) returns a human-readable message for the last
 *     failure on that context (the reference reports failures as `ok = false`
 *     plus a string: include/yaoray/runtime/backend.hpp:57-79).
 *   - A context owns one device, one HIP stream and all device copies of the
 *     scene. Contexts are externally synchronised: one host thread per context,
 *     matching "no concurrent Render on one prepared scene"
 *     (src/backends/cpu/cpu_worker_pool.cpp:40-42).
 *   - Arrays handed to yrh_upload_* are HOST arrays, copied during the call.
 *     Ray / hit arrays handed to yrh_trace_* are DEVICE arrays (allocate them
 *     with yrh_device_alloc or any HIP allocator on the context's device).
 */
#ifndef YRH_H
#define YRH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

#define YRH_OK 0
#define YRH_E_INVALID_ARGUMENT (-1)
#define YRH_E_HIP (-2)
#define YRH_E_UNSUPPORTED (-3)
#define YRH_E_NO_SCENE (-4)
#define YRH_E_CANCELLED (-5)
/* A traversal dropped stack entries: a BVH is deeper than the max_depth it was
 * uploaded with (uploads reject such BVHs; reported at the next sync point). */
#define YRH_E_OVERFLOW (-6)

typedef struct yrh_context yrh_context;

/* ---- Scene data, byte-compatible with the reference's IR ---------------- */

/* == yr::RenderBvhNode (include/yaoray/accel/bvh.hpp:34-47), 32 bytes.
 * Leaf: payload_offset = first primitive slot, primitive_count > 0.
 * Interior: payload_offset = right child, left child = own index + 1. */
typedef struct yrh_bvh_node {
    float bounds_min[3];
    float bounds_max[3];
    int32_t payload_offset;
    int32_t primitive_count;
} yrh_bvh_node;

/* == yr::RenderBvhPrimitive (include/yaoray/accel/bvh.hpp:49-55), 20 bytes.
 * kind 0 = triangle, 1 = sphere. Handles are -1 when invalid. */
typedef struct yrh_bvh_primitive {
    int32_t kind;
    int32_t mesh_primitive;
    int32_t sphere;
    int32_t local_triangle;
    int32_t flat_triangle_index;
} yrh_bvh_primitive;

/* == yr::RenderVertex (include/yaoray/scene/geometry.hpp:12-18), 48 bytes. */
typedef struct yrh_vertex {
    float position[3];
    float normal[3];
    float uv[2];
    float tangent[3];
    float tangent_handedness;
} yrh_vertex;

/* == yr::RenderPrimitive (include/yaoray/scene/geometry.hpp:20-27), 16 bytes. */
typedef struct yrh_primitive {
    uint32_t first_index;
    uint32_t index_count;
    int32_t material_index;
    uint8_t has_normals;
    uint8_t has_uvs;
    uint8_t has_tangents;
    uint8_t reserved;
} yrh_primitive;

/* == yr::RenderSphere (include/yaoray/scene/geometry.hpp:29-35), 28 bytes:
 * world-space centre and radius; flip_normals turns the outward normal
 * inward. Analytic spheres carry no emission in the reference
 * (src/integrators/path_integrator.cpp:83-86). */
typedef struct yrh_sphere {
    float center[3];
    float radius;
    int32_t material_index;
    int32_t area_light_index;
    uint8_t flip_normals;
    uint8_t reserved[3];
} yrh_sphere;

/* One binary BVH: yr::RenderBvh (include/yaoray/accel/bvh.hpp:72-81). */
typedef struct yrh_bvh_desc {
    const yrh_bvh_node* nodes;
    int32_t node_count;
    const int32_t* primitive_indices;
    int32_t primitive_index_count;
    const yrh_bvh_primitive* primitives;
    int32_t primitive_count;
    int32_t max_depth;
} yrh_bvh_desc;

/* == yr::PreparedMeshInstance (include/yaoray/accel/two_level_bvh.hpp:20-26).
 * Matrices are column-major yr::Mat4f (translation in m[12..14]). The inverse
 * is the host-computed one (src/core/transform.cpp:119) so object rays match. */
typedef struct yrh_instance {
    int32_t primitive;
    int32_t blas_index;
    float object_to_world[16];
    float world_to_object[16];
    float world_bounds_min[3];
    float world_bounds_max[3];
} yrh_instance;

/* == yr::RenderTlasPrimitive (include/yaoray/accel/two_level_bvh.hpp:28-33).
 * kind 0 = instance, 1 = sphere. */
typedef struct yrh_tlas_primitive {
    int32_t kind;
    int32_t instance;
    int32_t sphere;
    float world_bounds_min[3];
    float world_bounds_max[3];
} yrh_tlas_primitive;

#define YRH_ACCEL_FLAT 0
#define YRH_ACCEL_TWO_LEVEL 1

/* GeometryView + RenderAcceleration (include/yaoray/accel/acceleration.hpp:12-19). */
typedef struct yrh_geometry_desc {
    const yrh_vertex* vertices;
    int64_t vertex_count;
    const uint32_t* indices;
    int64_t index_count;
    const yrh_primitive* primitives;
    int32_t primitive_count;
    int32_t sphere_count;
    const yrh_sphere* spheres; /* flat BVH primitives of kind 1, or TLAS primitives of kind 1 */

    int32_t accel_kind; /* YRH_ACCEL_FLAT or YRH_ACCEL_TWO_LEVEL */
    yrh_bvh_desc flat;  /* used when accel_kind == YRH_ACCEL_FLAT */

    /* used when accel_kind == YRH_ACCEL_TWO_LEVEL */
    const yrh_bvh_desc* blases; /* one per mesh primitive, in primitive order */
    int32_t blas_count;
    const yrh_instance* instances;
    int32_t instance_count;
    yrh_bvh_desc tlas; /* tlas.primitives is unused; see tlas_primitives */
    const yrh_tlas_primitive* tlas_primitives;
    int32_t tlas_primitive_count;
} yrh_geometry_desc;

/* ---- Ray / hit queues (device memory, structure of arrays) -------------- */

/* Rays: org_tmin[i] = (origin.xyz, t_min), dir_tmax[i] = (direction.xyz, t_max),
 * each an array of n float4 (16-byte aligned). */
typedef struct yrh_ray_soa {
    const float* org_tmin;
    const float* dir_tmax;
} yrh_ray_soa;

/* Closest-hit output: hit[i] = (t, bary_u, bary_v, id bits) as float4 with
 * id = the hit's flat triangle index (>= 0), -1 for a miss (t is then the
 * ray's t_max, as in yr::BvhHit with hit = false), or -2 - s for a hit on
 * sphere s, whose (bary_u, bary_v) are then SphereUv of the hit normal
 * (src/accel/bvh_traversal.cpp:123-150). instance may be NULL; otherwise it
 * receives the instance index of two-level triangle hits (-1 for flat scenes,
 * spheres and misses). node_tests / triangle_tests / sphere_tests may be NULL;
 * otherwise they receive the per-ray counters with yr::BvhTraceStats
 * semantics (src/accel/bvh_traversal.cpp:272, 302, 305, 318, 324). */
typedef struct yrh_hit_soa {
    float* hit;
    int32_t* instance;
    uint32_t* node_tests;
    uint32_t* triangle_tests;
    uint32_t* sphere_tests;
} yrh_hit_soa;

/* ---- Context -------------------------------------------------------------- */

int yrh_create(int device, yrh_context** out);
void yrh_destroy(yrh_context* ctx);
const char* yrh_last_error(const yrh_context* ctx);
int yrh_synchronize(yrh_context* ctx);
/* The context's HIP stream (a hipStream_t), for callers that order their own
 * work with the library's. */
void* yrh_stream(yrh_context* ctx);
/* Device name and compute-unit count of the context's device. */
int yrh_device_info(yrh_context* ctx, char* name, int name_capacity, int* compute_units);

/* Device memory helpers so callers need no HIP headers. */
int yrh_device_alloc(yrh_context* ctx, size_t bytes, void** out);
int yrh_device_free(yrh_context* ctx, void* ptr);
int yrh_copy_to_device(yrh_context* ctx, void* dst, const void* src, size_t bytes);
int yrh_copy_to_host(yrh_context* ctx, void* dst, const void* src, size_t bytes);

/* ---- Acceleration structure --------------------------------------------- */

/* The checks yrh_upload_geometry applies before touching the device: array
 * sizes and presence, sphere records, child and leaf ranges of every BVH, and each
 * BVH's declared max_depth at least its real depth (root = 1, as
 * yr::RenderBvh::max_depth counts it, src/accel/bvh_builder.cpp:226-244): the
 * traversal stack is sized from it. Needs no device or context. Writes a
 * message into error (may be NULL) on failure. */
int yrh_validate_geometry(const yrh_geometry_desc* geometry, char* error, int error_capacity);

/* Replaces the context's geometry. Consumes the reference's own BVH arrays
 * (flat: yr::RenderBvh; two-level: yr::TwoLevelBvh) exactly as built by
 * yr::BuildRenderAcceleration (src/accel/acceleration.cpp:28-46) and repacks
 * them on upload into the device layout (same tree, same visit order). */
int yrh_upload_geometry(yrh_context* ctx, const yrh_geometry_desc* geometry);

/* Batched closest hit over n rays: the GPU counterpart of calling
 * yr::IntersectAcceleration (src/accel/acceleration.cpp:48-61) once per ray.
 * Launched on the context's stream; returns without synchronising. A
 * traversal stack overflow (impossible for geometry that passed
 * yrh_validate_geometry) is reported as YRH_E_OVERFLOW by the next
 * yrh_synchronize, yrh_film_read, yrh_render_stats_read or yrh_batch_render. */
int yrh_trace_closest(yrh_context* ctx, const yrh_ray_soa* rays, int64_t n,
                      const yrh_hit_soa* hits);

/* yrh_trace_closest without any allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). Same rays, hits
 * and counters as yrh_trace_closest, bit for bit; rays are claimed in input
 * order. Its scratch memory (ray-claim counters and the traversal stack spill)
 * is one 64 MiB device array per GPU, held from library load to process end;
 * the launch uses fewer workgroups when the tree's spilled stack levels would
 * not fit, and fails with YRH_E_UNSUPPORTED if one workgroup's would not (a
 * tree deeper than 32768 + LDS levels). Every call on a GPU shares its array:
 * calls that overlap on one GPU (two streams or two contexts) race. */
int yrh_trace_closest_noalloc(yrh_context* ctx, const yrh_ray_soa* rays, int64_t n, const yrh_hit_soa* hits);

/* Batched any-hit (occlusion) over n rays: occluded[i] = 1 if any triangle
 * or sphere is hit with t_min < t <= t_max, else 0. Equal to (closest hit found) on the
 * same ray; the reference has no any-hit API (its shadow rays are closest-hit
 * queries, src/integrators/path_direct_lighting.cpp:23-113).
 * node_tests / triangle_tests may be NULL. */
int yrh_trace_any(yrh_context* ctx, const yrh_ray_soa* rays, int64_t n,
                  uint32_t* occluded, uint32_t* node_tests, uint32_t* triangle_tests);

/* Launch geometry of the persistent traversal kernels: workgroups per compute
 * unit and threads per workgroup (64, 128 or 256); 0 keeps the current value;
 * blocks_per_cu = -1 returns to the automatic choice (as many workgroups as
 * the LDS stack allows, at most 8). */
int yrh_set_trace_config(yrh_context* ctx, int blocks_per_cu, int block_size);
/* Traversal stack levels kept in LDS per lane (1..16; 0 = automatic: 8 for a
 * flat tree, 16 for two-level); deeper levels spill to a per-lane global
 * region. Fewer levels use less LDS per workgroup, so more workgroups fit on
 * a CU (by default as many as the LDS allows, at most 8). */
int yrh_set_trace_lds_levels(yrh_context* ctx, int levels);

/* Order in which yrh_trace_closest / yrh_trace_any claim their rays.
 * YRH_TRACE_ORDER_ORIGIN (the default): launches of at least 65536 rays first
 * sort the ray indices by the Morton code of the origin's cell in the scene
 * box (16 cells per axis; rays of one cell keep their input order, e.g. a
 * camera's pixel order), so rays starting close together share a wave and
 * its cached nodes; results still go to each ray's own index and are
 * identical. YRH_TRACE_ORDER_INPUT: rays are claimed in input order. */
#define YRH_TRACE_ORDER_INPUT 0
#define YRH_TRACE_ORDER_ORIGIN 1
int yrh_set_trace_order(yrh_context* ctx, int order);

/* ---- Shading data ----------------------------------------------------------- */

#define YRH_MATERIAL_DIFFUSE 0
#define YRH_MATERIAL_CONDUCTOR 1
#define YRH_MATERIAL_DIELECTRIC 2
#define YRH_MATERIAL_THIN_DIELECTRIC 3

/* The subset of yr::RenderMaterial (include/yaoray/scene/material.hpp:26-65)
 * that the in-scope BSDFs and EvaluateMaterialAtSurface read
 * (src/shading/bsdf.cpp:17-329; src/shading/material_evaluator.cpp:112-180).
 * Texture fields index yrh_shading_desc::textures; an index outside
 * [0, texture_count) means "no texture" (TextureIndexValid,
 * material_evaluator.cpp:19-21). Textures apply only to triangles of
 * primitives with has_uvs: reflectance is multiplied by the texture's RGB,
 * alpha by its first channel; a hit whose alpha is below 0.5 is skipped
 * (IsAlphaVisible, :182-184); normal_map (also needs has_tangents) perturbs
 * the shading normal in the interpolated tangent frame, tangent-space x and
 * y scaled by normal_scale (:35-92). Spheres are never textured. */
typedef struct yrh_material {
    int32_t kind; /* YRH_MATERIAL_* */
    float reflectance[3];
    float ior;
    float uroughness;
    float vroughness;
    float alpha;
    float emission[3];
    float absorption_distance;
    float absorption_color[3];
    float reserved;
    int32_t reflectance_texture;
    int32_t alpha_texture;
    int32_t normal_map;
    float normal_scale;
} yrh_material;

#define YRH_FILTER_NEAREST 0
#define YRH_FILTER_BILINEAR 1
#define YRH_FILTER_TRILINEAR 2
#define YRH_FILTER_EWA 3
#define YRH_WRAP_REPEAT 0
#define YRH_WRAP_CLAMP 1
#define YRH_WRAP_MIRRORED 2

/* == yr::RenderTexture (include/yaoray/scene/texture.hpp:38-51): linear RGBA
 * texels of level 0 (row-major, width * height * 4 floats), the filter
 * (YRH_FILTER_*, include/yaoray/scene/texture_options.hpp:7-12) and the wrap
 * modes (YRH_WRAP_*). The library builds the mip chain at upload
 * (BuildTextureMipChain, src/shading/texture.cpp:241-272). A texture without
 * texels (width or height <= 0, or texels NULL) samples as zero, as
 * SampleTexture4 returns for one (texture.cpp:312-313). Trilinear and EWA
 * filtering apply where the reference has a texture footprint: hits of camera
 * rays (ray differentials one pixel apart, src/scene/camera.cpp:20-39;
 * ComputeTextureFootprint, src/integrators/surface_query.cpp:35-97); every
 * other lookup is bilinear. */
typedef struct yrh_texture {
    int32_t width;
    int32_t height;
    int32_t filter;
    int32_t wrap_s;
    int32_t wrap_t;
    int32_t reserved;
    const float* texels;
} yrh_texture;

/* == yr::EmissivePrimitive (include/yaoray/scene/light.hpp:30-34). */
typedef struct yrh_emissive {
    int32_t primitive_index;
    float radiance[3];
    float area;
} yrh_emissive;

/* yr::RenderEnvironment + its RenderTexture (RGBA float) + RenderEnvironmentDistribution
 * (include/yaoray/scene/light.hpp:10-28). The map is read as SampleTexture
 * reads it without a footprint (src/lighting/environment.cpp:136): bilinear,
 * or nearest when `nearest` is set, with the texture's wrap modes (YRH_WRAP_*;
 * the reference's HDR / EXR loaders set repeat in u, clamp in v). */
typedef struct yrh_environment_desc {
    int32_t active;
    float radiance[3];
    float strength;
    float rotation_radians;
    int32_t width;
    int32_t height;
    const float* texels; /* width * height * 4 floats, or NULL */
    int32_t has_distribution;
    int32_t dist_width;
    int32_t dist_height;
    const float* texel_weights;    /* dist_width * dist_height */
    const float* row_weights;      /* dist_height */
    const float* row_cdf;          /* dist_height */
    const float* conditional_cdfs; /* dist_width * dist_height */
    float total_weight;
    int32_t nearest;
    int32_t wrap_s;
    int32_t wrap_t;
} yrh_environment_desc;

#define YRH_LIGHT_POINT 0
#define YRH_LIGHT_SPOT 1
#define YRH_LIGHT_DISTANT 2

/* == yr::AnalyticLight (include/yaoray/scene/light.hpp:36-45): a point, spot
 * or distant light (kind YRH_LIGHT_*). Spot lights: direction is the cone
 * axis, cone_angle the cosine of the outer cone, cone_cos_inner the cosine of
 * the inner (full-intensity) cone; distant lights shine along direction. */
typedef struct yrh_analytic_light {
    int32_t kind;
    float position[3];
    float direction[3];
    float intensity[3];
    float cone_angle;
    float cone_cos_inner;
} yrh_analytic_light;

/* == yr::LightTreeNode (include/yaoray/scene/light.hpp:53-59): bounds of the
 * light positions below, their summed power, children (-1 for a leaf) and the
 * leaf's light index (-1 for an interior node). */
typedef struct yrh_light_tree_node {
    float bounds_min[3];
    float bounds_max[3];
    float power;
    int32_t left;
    int32_t right;
    int32_t light_index;
} yrh_light_tree_node;

typedef struct yrh_shading_desc {
    const yrh_material* materials;
    int32_t material_count;
    const yrh_emissive* emissives;
    int32_t emissive_count;
    /* LightSamplingCache::emissive_power (src/lighting/light_sampling.cpp:148-216) */
    const float* alias_probability;
    const int32_t* alias_index;
    const float* alias_pmf;
    yrh_environment_desc environment;
    /* RenderSceneIR::analytic_lights and their LightSamplingCache tables
     * (src/lighting/light_sampling.cpp:206-269): the power alias table
     * (analytic_light_count entries) and the light tree, which light
     * selection descends instead when use_analytic_tree is set (more than 8
     * lights, none distant). NULL / 0 when the scene has none. */
    const yrh_analytic_light* analytic_lights;
    int32_t analytic_light_count;
    const float* analytic_alias_probability;
    const int32_t* analytic_alias_index;
    const float* analytic_alias_pmf;
    const yrh_light_tree_node* analytic_tree;
    int32_t analytic_tree_count;
    int32_t analytic_tree_root;
    int32_t use_analytic_tree;
    /* RenderSceneIR::textures that materials index (NULL / 0 when none). */
    const yrh_texture* textures;
    int32_t texture_count;
} yrh_shading_desc;

/* Uploads materials, emitters and the environment for the current geometry
 * (call after yrh_upload_geometry). */
int yrh_upload_shading(yrh_context* ctx, const yrh_shading_desc* shading);

/* ---- Wavefront path tracing ----------------------------------------------- */

#define YRH_SAMPLER_INDEPENDENT 0
#define YRH_SAMPLER_STRATIFIED 1
#define YRH_SAMPLER_OWEN_SOBOL 2
#define YRH_SAMPLER_ZSOBOL 3

/* One pass = one sample for every pixel of the frame, the unit the reference
 * commits per pass (src/backends/cpu/cpu_path_tracer.cpp:109-180); a pass may
 * carry several consecutive sample indices (sample_count) to fill the GPU. */
typedef struct yrh_render_desc {
    float camera_origin[3];
    float camera_forward[3];
    float camera_right[3];
    float camera_up[3];
    float tan_half_fov; /* std::tan(fov_y_radians * 0.5f), computed on the host */
    int32_t width;
    int32_t height;
    int32_t sampler; /* YRH_SAMPLER_* */
    int32_t spp;     /* samples per pixel of the whole render */
    int32_t max_depth;
    uint64_t seed;
    float radiance_clamp;
    int32_t sample_index; /* this pass's first sample index (= completed spp) */
    int32_t sample_count; /* consecutive sample indices rendered by this pass (0 means 1); the
                             film receives them in index order, as that many passes would add them */
    int32_t pixel_list;   /* 1: render the pixels of the list the last yrh_film_update_active
                             wrote, sample_count samples each from the pixel's own next sample
                             index (its film sample count), added in index order until the pixel
                             converges under that call's adaptive settings or reaches spp: the film
                             equals sample_count one-sample passes with the list updated between
                             them (the render counters include the samples not added);
                             sample_index is ignored */
} yrh_render_desc;

/* yr::AdaptiveSamplingSettings (include/yaoray/scene/render_options.hpp:30-36). */
typedef struct yrh_adaptive_desc {
    int32_t enabled;
    int32_t min_spp;
    float relative_error;
    float absolute_error;
    float confidence;
} yrh_adaptive_desc;

/* ---- learned path guiding (no counterpart in the reference) -------------------
 * A hash-grid + MLP proposal of 8 vMF lobes over directions, conditioned on
 * position, shading normal and outgoing direction, trained online on the
 * integrator's own path vertices and combined with BSDF sampling by
 * one-sample MIS (so renders stay unbiased). */
typedef struct yrh_guide_desc {
    float bounds_min[3];      /* scene box the position encoding spans */
    float bounds_max[3];
    int32_t log2_table_size;  /* hash-grid entries per level = 2^log2_table_size (10..22) */
    int32_t min_resolution;   /* coarsest level resolution (>= 2) */
    int32_t max_resolution;   /* finest level resolution (>= min_resolution) */
    float learning_rate;      /* Adam step size */
    float guide_fraction;     /* probability of sampling the guide at a guidable vertex (0..1) */
    int32_t sample_capacity;  /* training samples kept between yrh_guide_train calls */
    uint64_t seed;            /* weight initialisation */
} yrh_guide_desc;

#define YRH_GUIDE_OFF 0     /* plain integrator (the reference's estimator) */
#define YRH_GUIDE_SAMPLE 1  /* guided sampling */
#define YRH_GUIDE_RECORD 2  /* guided sampling + training-sample recording */

/* yr::RenderStats counters (include/yaoray/runtime/backend.hpp:40-55) plus the
 * film's bad (non-finite) sample count. */
typedef struct yrh_render_stats {
    uint64_t rays_traced;
    uint64_t shadow_rays;
    uint64_t occluded_shadow_rays;
    uint64_t triangle_tests;
    uint64_t sphere_tests;
    uint64_t bvh_node_tests;
    uint64_t hits;
    uint64_t misses;
    uint64_t bad_samples;
} yrh_render_stats;

/* Per-pixel FP64 accumulator: yr::FilmPixel (include/yaoray/film/film.hpp:19-28)
 * with sums and Welford moments in double. */
typedef struct yrh_film_pixel {
    double sum[3];
    double luminance_mean;
    double luminance_m2;
    double albedo_sum[3];
    double normal_sum[3];
    double depth_sum;
    uint32_t samples;
    uint32_t aov_samples;
} yrh_film_pixel;

/* Allocates (or re-allocates) the film and zeroes it and the render stats. */
int yrh_film_reset(yrh_context* ctx, int32_t width, int32_t height);
/* Loads a film (e.g. a resumed render) from host pixels. */
int yrh_film_load(yrh_context* ctx, const yrh_film_pixel* pixels, int64_t count);
/* Copies the film to the host (synchronises). */
int yrh_film_read(yrh_context* ctx, yrh_film_pixel* pixels, int64_t count);
/* Runs one pass on the context's stream, accumulating into the film. */
int yrh_render_pass(yrh_context* ctx, const yrh_render_desc* desc);
/* The active pixels of RenderCpuPathTrace (src/backends/cpu/cpu_path_tracer.cpp:95-106,
 * 183-199) from the device film: those with fewer than spp samples that are not
 * converged by Film::IsConverged (src/film/film.cpp:113-121; only when
 * adaptive->enabled; the test runs on the pixel statistics rounded to FP32).
 * The list stays on the device for pixel-list passes. With active or converged
 * non-null it synchronises and returns the list length and the converged count. */
int yrh_film_update_active(yrh_context* ctx, const yrh_adaptive_desc* adaptive, int32_t spp, int64_t* active,
                           int64_t* converged);
/* Paths (pixels x sample indices) one pass should hold: as many as half of
 * the device memory not in use (plus what this context's path state already
 * holds) can keep, at ~460 bytes per path (+ the guiding buffers when a
 * proposal is configured), capped at 2^27. Every bounce of a pass waits for
 * its slowest paths, so fewer, larger passes spend less time in nearly empty
 * late bounces; the film is the same for any pass size. No counterpart in the
 * reference (its CPU passes are one sample index each). */
int yrh_suggest_paths_per_pass(yrh_context* ctx, int32_t max_depth, int64_t* paths);
/* Reads the accumulated counters (synchronises). */
int yrh_render_stats_read(yrh_context* ctx, yrh_render_stats* stats);

/* ---- Batched multi-camera rendering (no counterpart in the reference, which
 * renders one job per Render call) ------------------------------------------ */

typedef struct yrh_batch_job {
    float camera_origin[3];
    float camera_forward[3];
    float camera_right[3];
    float camera_up[3];
    float tan_half_fov; /* std::tan(fov_y_radians * 0.5f) */
    int32_t width;
    int32_t height;
    int32_t spp;
} yrh_batch_job;

typedef struct yrh_batch_desc {
    const yrh_batch_job* jobs;
    int32_t job_count;
    int32_t sampler; /* YRH_SAMPLER_*, shared by the jobs */
    int32_t max_depth;
    float radiance_clamp;
    uint64_t seed;
    int64_t paths_per_pass; /* wavefront size; 0 = min(2^23, total pixels) */
    int32_t use_graph;      /* 1: capture one pass as a HIP graph and replay it per pass */
    int32_t reserved;
} yrh_batch_desc;

/* Renders every job to its full spp in one pass schedule: each pass carries
 * consecutive pixel ranges / sample layers of the jobs, least remaining work
 * first (small jobs complete early), until it is full; a pixel receives its samples in index order, so each
 * job's FP64 film equals rendering that job alone with yrh_render_pass. Synchronises. job_ms (host, job_count entries,
 * or NULL) receives each job's completion time since the start in ms;
 * passes_out (or NULL) the number of passes. Counters: yrh_render_stats_read. */
int yrh_batch_render(yrh_context* ctx, const yrh_batch_desc* desc, float* job_ms, int32_t* passes_out);
/* Copies one job's film (width*height pixels) of the last batch to the host. */
int yrh_batch_film_read(yrh_context* ctx, int32_t job, yrh_film_pixel* pixels, int64_t count);

/* Batched sample generation: out[i] = Sampler(kind, seed, x[i], y[i], sample[i],
 * spp).Sample1D/2D at absolute dimension dim[i] (src/sampling/sampler.cpp:122-237),
 * computed on the device; dims = 1 or 2 floats per entry. Host arrays. */
int yrh_sample_values(yrh_context* ctx, int32_t kind, uint64_t seed, int32_t spp, int64_t n, const int32_t* x,
                      const int32_t* y, const int32_t* sample, const int32_t* dim, int32_t dims, float* out);

/* Batched BSDF queries on the device, the arithmetic the shading kernels run:
 * for query i, material materials[material_index[i]] and the 12 floats
 * queries[12i..] = wo (3), wi (3), shading normal (3), u (2), unused, out[12i..]
 * receives EvaluateBsdf(wo, wi, n) (3 floats), PdfBsdf(wo, wi, n), then
 * SampleBsdf(wo, n, u): wi (3), weight (3), pdf, and flags = valid + 2 specular
 * + 4 IsDeltaBsdf (src/shading/bsdf.cpp:17-329; include/yaoray/shading/bsdf.hpp:10-16).
 * Host arrays; synchronises. */
int yrh_bsdf_query(yrh_context* ctx, const yrh_material* materials, int32_t material_count, int64_t n,
                   const int32_t* material_index, const float* queries, float* out);

/* SampleTexture4 (src/shading/texture.cpp:312-322) of `texture` (its mip
 * chain built as at upload) at n points: uv[2i..] = (u, v), footprint[4i..] =
 * (dudx, dvdx, dudy, dvdy) or footprint NULL for none; out[4i..] receives
 * RGBA. Host arrays; synchronises. The texture lookups of the shading
 * kernels, exposed for parity tests. */
int yrh_texture_query(yrh_context* ctx, const yrh_texture* texture, int64_t n, const float* uv, const float* footprint,
                      float* out);

/* Device ray generation for one pass (GeneratePerspectiveCameraRay with the
 * sampler's Pixel dimension): writes width*height rays into device arrays. */
int yrh_generate_camera_rays(yrh_context* ctx, const yrh_render_desc* desc, float* org_tmin, float* dir_tmax);

/* ---- Learned path guiding (no counterpart in the reference) ---------------- */

/* Allocates and initialises the proposal (weights from desc->seed). */
int yrh_guide_configure(yrh_context* ctx, const yrh_guide_desc* desc);
/* YRH_GUIDE_OFF / _SAMPLE / _RECORD for the following render passes. */
int yrh_guide_set_mode(yrh_context* ctx, int32_t mode);
/* Device memory (bytes) that sampling passes may spend on per-cell copies of
 * the hash grid: levels are copied coarse first while they fit, each cell's 8
 * corners in one 64-byte record, so a lookup is one load instead of 8 gathers.
 * The features, and so the renders, are identical with any budget; 0 turns the
 * copies off. Default 512 MiB. */
int yrh_guide_set_cell_budget(yrh_context* ctx, int64_t bytes);
/* Guides only the vertices of the first `depths` bounces (0, the default:
 * every vertex that may continue); deeper vertices sample their BSDF as the
 * plain integrator does. Renders stay unbiased for any limit. */
int yrh_guide_set_depth_limit(yrh_context* ctx, int32_t depths);
/* `steps` Adam steps of `batch` recorded samples each (consecutive windows of
 * the recorded set), then forgets the recorded samples. */
int yrh_guide_train(yrh_context* ctx, int32_t steps, int32_t batch);
/* Appends n host samples of 13 floats (position, normal, outgoing, sampled
 * direction, weight > 0) to the recorded set, as if a render pass had
 * recorded them (samples beyond the capacity are dropped). */
int yrh_guide_add_samples(yrh_context* ctx, int64_t n, const float* samples);
/* Copies up to n recorded samples to the host as rows of 13 floats (position,
 * normal, outgoing, sampled direction, weight); returns the number copied in
 * *copied (synchronises). */
int yrh_guide_read_samples(yrh_context* ctx, int64_t n, float* samples, int64_t* copied);
/* Recorded training samples and their mean weight (synchronises). */
int yrh_guide_sample_stats(yrh_context* ctx, int64_t* count, float* mean_weight);
/* Sizes: MLP parameters (FP32 [in][out] matrices W0..W3), grid parameters
 * (levels x entries x 2), per-level resolutions and dense flags (8 each). */
int yrh_guide_info(yrh_context* ctx, int32_t* mlp_params, int64_t* grid_params, int32_t* resolutions,
                   int32_t* dense);
/* FP32 master parameters to / from host arrays. */
int yrh_guide_params_read(yrh_context* ctx, float* mlp, float* grid);
int yrh_guide_params_write(yrh_context* ctx, const float* mlp, const float* grid);
/* Network evaluation for n host rows of 9 floats (position, normal, outgoing
 * direction): writes 40 floats per row (pi[8], kappa[8], mu[8][3]). */
int yrh_guide_query(yrh_context* ctx, int64_t n, const float* inputs, float* mixture);
/* One forward + backward over n host samples of 13 floats (position, normal,
 * outgoing, sampled direction, weight): loss = mean(-(w / w_norm) log q(wi)).
 * Writes the FP32 gradients of every parameter; no update. */
int yrh_guide_gradients(yrh_context* ctx, int64_t n, const float* samples, float w_norm, float* grad_mlp,
                        float* grad_grid, float* loss);
/* Mixture pdf at dirs (n x 3) and mixture samples from u (n x 3) written as
 * (direction, pdf) rows of 4 floats. */
int yrh_guide_mixture_eval(yrh_context* ctx, int64_t n, const float* mixture, const float* dirs, const float* u,
                           float* pdf, float* samples);

/* ---- Timing on the context's stream --------------------------------------- */

/* HIP events recorded on the context's stream; elapsed milliseconds between
 * two recorded events (waits for the second one). */
int yrh_event_create(yrh_context* ctx, void** event);
int yrh_event_destroy(yrh_context* ctx, void* event);
int yrh_event_record(yrh_context* ctx, void* event);
int yrh_event_elapsed_ms(yrh_context* ctx, void* start, void* stop, float* ms);

/* Phase timing of the ray-cast calls (off by default). With timing on, each
 * yrh_trace_closest / yrh_trace_any records events on the context's stream
 * around its ray-ordering pass and around its traversal kernel;
 * yrh_trace_phase_ms waits for the last such call and returns both durations
 * in milliseconds (order_ms = 0 when the call ran no ordering pass). */
int yrh_set_trace_timing(yrh_context* ctx, int enable);
int yrh_trace_phase_ms(yrh_context* ctx, float* order_ms, float* trace_ms);
/* Render passes (yrh_render_pass) run each bounce's shadow rays on a second
 * stream, beside the next bounce's path-segment traversal (default on); the
 * film is the same either way. No counterpart in the reference. */
int yrh_set_render_overlap(yrh_context* ctx, int enable);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* YRH_H */
