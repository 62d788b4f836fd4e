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

/* Read-only views of the scene arrays (valid until the scene is destroyed). */
int yrhost_scene_geometry(const yrhost_scene* scene, const yrh_vertex** vertices, const uint32_t** indices,
                          const yrh_primitive** primitives);

/* yr::BuildRenderAcceleration with the given build thread count (0 = all). */
int yrhost_build_accel(yrhost_scene* scene, int thread_count, double* seconds);
int yrhost_accel_flat(const yrhost_scene* scene, yrh_bvh_desc* out);
int yrhost_accel_blas(const yrhost_scene* scene, int32_t index, yrh_bvh_desc* out);
int yrhost_accel_tlas(const yrhost_scene* scene, yrh_bvh_desc* tlas, const yrh_tlas_primitive** primitives,
                      int32_t* primitive_count, const yrh_instance** instances, int32_t* instance_count);

/* Uploads the scene's geometry and acceleration structure to a context. */
int yrhost_upload_geometry(const yrhost_scene* scene, yrh_context* ctx);

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
