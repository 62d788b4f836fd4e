/*
 * yaoray_host.h — C entry points of the C++20 host library libyaoray_host.so,
 * used by the Python bindings, tests and bench.py. The host library owns the
 * scene IR, the SAH builder (yr::BuildRenderAcceleration restated) and the
 * seeded procedural workloads, and reaches the GPU only through yrh.h.
 */
#ifndef YAORAY_HOST_H
#define YAORAY_HOST_H

#include <stdint.h>

#include "../../include/yrh.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

typedef struct yrhost_scene yrhost_scene;

#define YRHOST_SCENE_CORNELL 1   /* params: none (seed = material seed) */
#define YRHOST_SCENE_TRIANGLES 2 /* params: [triangle count] */
#define YRHOST_SCENE_INSTANCED 3 /* params: [instances, subdivisions, variants, env_w, env_h] */
#define YRHOST_SCENE_ROOM 4      /* params: [furniture triangles, env_w, env_h] */

typedef struct yrhost_scene_info {
    int64_t vertex_count;
    int64_t index_count;
    int32_t primitive_count;
    int32_t instance_count;
    int32_t material_count;
    int32_t emissive_count;
    int32_t accel_built;
    int32_t accel_kind;     /* YRH_ACCEL_* */
    int32_t node_count;     /* RenderAcceleration::NodeCount */
    int32_t max_depth;      /* RenderAcceleration::MaxDepth */
    int32_t blas_count;
    int32_t environment_active;
    float camera[13];       /* origin, forward, right, up, fov_y_radians */
    int32_t sphere_count;
    int32_t analytic_light_count;
    int32_t texture_count;
} yrhost_scene_info;

int yrhost_scene_create(int kind, const int64_t* params, int param_count, uint64_t seed, yrhost_scene** out);
/* Scene from caller arrays (copied). instance_primitive / instance_matrices
 * (16 floats each, column-major) may be NULL when instance_count == 0. */
int yrhost_scene_from_arrays(const yrh_vertex* vertices, int64_t vertex_count, const uint32_t* indices,
                             int64_t index_count, const yrh_primitive* primitives, int32_t primitive_count,
                             const int32_t* instance_primitive, const float* instance_matrices,
                             int32_t instance_count, yrhost_scene** out);
void yrhost_scene_destroy(yrhost_scene* scene);
const char* yrhost_scene_error(const yrhost_scene* scene);
int yrhost_scene_info_get(const yrhost_scene* scene, yrhost_scene_info* info);

/* Scene edits for in-code test jobs (the reference builds its test scenes the
 * same way, tests/backends/cpu/cpu_path_tracer_test_support.hpp:27-84).
 * camera: origin, forward, right, up, fov_y_radians. Emitters replace the
 * current list and rebuild the emissive alias table. */
int yrhost_scene_set_camera(yrhost_scene* scene, const float* camera13);
int yrhost_scene_set_materials(yrhost_scene* scene, const yrh_material* materials, int32_t count);
/* Appends a texture to the scene's texture list (RenderSceneIR::textures,
 * include/yaoray/scene/render_scene.hpp) and returns its index (>= 0), which
 * materials name in reflectance_texture / alpha_texture / normal_map; the
 * texels are copied. Returns a negative YRH_E_* code on bad arguments. */
int yrhost_scene_add_texture(yrhost_scene* scene, const yrh_texture* texture);
/* An active equirect environment on texture `texture_index` (RenderEnvironment,
 * include/yaoray/scene/light.hpp:10-17), with its importance-sampling
 * distribution (BuildEnvironmentDistribution, src/lighting/environment.cpp:73-123). */
int yrhost_scene_set_environment_map(yrhost_scene* scene, int32_t texture_index, const float* radiance3,
                                     float strength, float rotation_radians);
int yrhost_scene_set_emissives(yrhost_scene* scene, const yrh_emissive* emissives, int32_t count);
int yrhost_scene_set_constant_environment(yrhost_scene* scene, const float* radiance3, float strength);
/* Replaces the analytic spheres (copied; the acceleration structure must be
 * built again). */
int yrhost_scene_set_spheres(yrhost_scene* scene, const yrh_sphere* spheres, int32_t count);
/* Read-only view of the spheres (valid until the scene changes). */
int yrhost_scene_spheres(const yrhost_scene* scene, const yrh_sphere** spheres, int32_t* count);
/* Replaces the analytic (point / spot / distant) lights (copied) and rebuilds
 * their selection tables (power alias table, light tree). */
int yrhost_scene_set_analytic_lights(yrhost_scene* scene, const yrh_analytic_light* lights, int32_t count);

/* Read-only views of the scene arrays (valid until the scene is destroyed). */
int yrhost_scene_geometry(const yrhost_scene* scene, const yrh_vertex** vertices, const uint32_t** indices,
                          const yrh_primitive** primitives);

/* yr::BuildRenderAcceleration with the given build thread count (0 = all). */
int yrhost_build_accel(yrhost_scene* scene, int thread_count, double* seconds);
int yrhost_accel_flat(const yrhost_scene* scene, yrh_bvh_desc* out);
int yrhost_accel_blas(const yrhost_scene* scene, int32_t index, yrh_bvh_desc* out);
int yrhost_accel_tlas(const yrhost_scene* scene, yrh_bvh_desc* tlas, const yrh_tlas_primitive** primitives,
                      int32_t* primitive_count, const yrh_instance** instances, int32_t* instance_count);

/* The scene's geometry + acceleration as the C ABI describes it (pointers into
 * the scene, valid until it changes), and its upload to a context. */
int yrhost_geometry_desc(const yrhost_scene* scene, yrh_geometry_desc* out);
int yrhost_upload_geometry(const yrhost_scene* scene, yrh_context* ctx);
/* Materials, emitters (with the emissive alias table) and environment. */
int yrhost_shading_desc(yrhost_scene* scene, yrh_shading_desc* out);
int yrhost_upload_shading(yrhost_scene* scene, yrh_context* ctx);
/* Pass descriptor for the scene camera; tan(fov/2) is evaluated here. */
int yrhost_render_desc(const yrhost_scene* scene, int32_t width, int32_t height, int32_t sampler, int32_t spp,
                       int32_t max_depth, uint64_t seed, float radiance_clamp, int32_t sample_index,
                       yrh_render_desc* out);

/* ---- yr::HipBackend (the reference's RenderBackend Prepare/Render) --------- */

/* yr::FilmPixel, 56 bytes (include/yaoray/film/film.hpp:11-21). */
typedef struct yrhost_film_pixel {
    float sum[3];
    uint32_t samples;
    float luminance_mean;
    float luminance_m2;
    float albedo_sum[3];
    float normal_sum[3];
    float depth_sum;
    uint32_t aov_samples;
} yrhost_film_pixel;

/* yr::RenderSettings (include/yaoray/runtime/render_job.hpp). */
typedef struct yrhost_render_settings {
    int32_t width;
    int32_t height;
    int32_t sampler;    /* YRH_SAMPLER_* */
    int32_t spp;
    int32_t max_depth;
    float radiance_clamp;
    uint64_t seed;
    int32_t device;
    int32_t requested_backend; /* 0 = HIP; anything else is refused by Prepare */
    /* yr::AdaptiveSamplingSettings (include/yaoray/scene/render_options.hpp:30-36) */
    int32_t adaptive_enabled;
    int32_t adaptive_min_spp;
    float adaptive_relative_error;
    float adaptive_absolute_error;
    float adaptive_confidence;
} yrhost_render_settings;

/* yr::RenderRequest. The stop token and the progress callback are driven
 * from here: stop_before_start requests the stop before the first pass,
 * cancel_after_passes > 0 makes the progress callback cancel after that many
 * passes (with cancel_error as the message when non-empty); < 0 renders
 * without a progress callback. */
typedef struct yrhost_render_request {
    const yrhost_film_pixel* resume_film; /* width*height pixels or NULL */
    int32_t resume_width;
    int32_t resume_height;
    int32_t resume_completed_spp;
    int32_t stop_before_start;
    int32_t cancel_after_passes;
    const char* cancel_error;
} yrhost_render_request;

/* yr::RenderResult + yr::RenderStats. */
typedef struct yrhost_render_report {
    int32_t prepare_ok;
    int32_t ok;
    int32_t film_valid;
    int32_t film_width;
    int32_t film_height;
    int32_t progress_calls;
    int32_t threads;
    int32_t bvh_nodes;
    int32_t bvh_max_depth;
    int32_t last_progress_spp;
    uint64_t samples_rendered;
    uint64_t rays_traced;
    uint64_t shadow_rays;
    uint64_t occluded_shadow_rays;
    uint64_t triangle_tests;
    uint64_t sphere_tests;
    uint64_t bvh_node_tests;
    uint64_t hits;
    uint64_t misses;
    uint64_t bad_samples;
    uint64_t last_progress_samples;
    uint64_t converged_pixels;
    uint64_t last_progress_active;
    double elapsed_seconds;
    double prepare_seconds;
    double acceleration_build_seconds;
    char error[256];
} yrhost_render_report;

/* HipBackend::Prepare on a copy of the scene, then HipBackend::Render. The
 * film (when the result carries one) is copied to film_out (capacity in
 * pixels). Returns YRH_OK when the report is filled; the render's own
 * success is report->ok with report->error. */
int yrhost_backend_render(const yrhost_scene* scene, const yrhost_render_settings* settings,
                          const yrhost_render_request* request, yrhost_film_pixel* film_out,
                          int64_t film_capacity, yrhost_render_report* report);

/* Pinhole rays through pixel centres of the scene camera (row-major). */
int yrhost_camera_rays(const yrhost_scene* scene, int width, int height, float t_min, float t_max,
                       float* org_tmin, float* dir_tmax);
/* Rays with origins uniform in the unit cube and directions uniform on the sphere. */
int yrhost_random_rays(int64_t count, uint64_t seed, float t_min, float t_max, float* org_tmin, float* dir_tmax);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* YAORAY_HOST_H */
