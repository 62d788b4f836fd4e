// Seeded procedural workloads, built directly as RenderSceneIR (the reference
// builds its test scenes in code the same way: tests/accel/bvh_test_scene.hpp).
#pragma once

#include <cstdint>

#include "yr_ir.hpp"

namespace yrhip {

// PCG32 (O'Neill 2014, XSH-RR output), seeded with (seed, stream = seed).
class Pcg32 {
public:
    explicit Pcg32(std::uint64_t seed);
    std::uint32_t NextU32();
    float NextFloat();  // [0, 1) with 24 bits
private:
    std::uint64_t state_ = 0;
    std::uint64_t inc_ = 1;
};

yr::Vec3f UniformSphere(float u, float v);

// Pinhole camera looking from `eye` at `target` (forward, right = forward x up,
// up' = right x forward: the basis yr::LookAtMatrix builds, src/core/transform.cpp:171-196).
yr::RenderCamera LookAtCamera(yr::Point3f eye, yr::Point3f target, yr::Vec3f up, float fov_y_degrees);

// GeneratePerspectiveCameraRay (src/scene/camera.cpp:9-39) without differentials.
void CameraRay(const yr::RenderCamera& camera, int width, int height, int x, int y, float sx, float sy,
               yr::Point3f& origin, yr::Vec3f& direction);

// Cornell-style box: 5 wall quads, 2 yaw-rotated boxes, ceiling emitter (L = 17).
yr::RenderSceneIR MakeCornellScene(std::uint64_t material_seed);

// Ray-cast microbench geometry: `count` triangles, centres U[0,1]^3, vertices
// c + l*d with d uniform on the sphere and l log-uniform in [1e-3, 2e-2].
yr::RenderSceneIR MakeRandomTriangleScene(int count, std::uint64_t seed);

// Two-level scene: ground quad (identity instance) + `instance_count` random
// TRS instances of `variants` icosphere primitives subdivided `subdivisions`
// times; procedural sky+sun environment of env_width x env_height.
yr::RenderSceneIR MakeInstancedScene(int instance_count, int subdivisions, int variants, std::uint64_t seed,
                                     int env_width, int env_height);

// Interior for learned guiding: a 10 x 10 x 4 room lit through one 1 x 1
// window by the procedural sky + sun, and by a ceiling lamp hidden behind a
// hanging panel (indirect-dominated); tables and icosphere objects, diffuse
// and GGX conductor, up to `furniture_triangles` triangles; one flat BVH.
yr::RenderSceneIR MakeGuidingRoomScene(int furniture_triangles, std::uint64_t seed, int env_width, int env_height);

}  // namespace yrhip
