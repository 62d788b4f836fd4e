// Host-side check run under AddressSanitizer + UndefinedBehaviorSanitizer and
// under ThreadSanitizer (tests/sanitize/Makefile): the CPU parts of the
// library (scene construction, the SAH builder serial and with worker
// threads, light tables, environment distributions, texture marshalling
// into the C ABI descriptors) over small versions of every procedural scene.
// The serial and threaded builds must give byte-identical trees (the
// reference's race canary, tests/accel/bvh_parallel_builder_tests.cpp:35-49).
// Exit status 0 = pass; sanitizer reports abort the run.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../yaoray_amd/host/abi_marshal.hpp"
#include "../../yaoray_amd/host/scenes.hpp"

namespace {

int failures = 0;

void Check(bool ok, const std::string& what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
        ++failures;
    }
}

template <class T>
bool SameBytes(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

template <class Tree>
bool SameBvh(const Tree& a, const Tree& b) {
    return SameBytes(a.nodes, b.nodes) && SameBytes(a.primitive_indices, b.primitive_indices) &&
           SameBytes(a.primitives, b.primitives) && a.max_depth == b.max_depth;
}

void Exercise(const char* name, yr::RenderSceneIR scene) {
    yr::BvhBuildOptions serial;
    serial.thread_count = 1;
    yr::BvhBuildOptions threaded;
    threaded.thread_count = 4;
    threaded.parallel_min_subtree_size = 64;
    const yr::RenderAccelerationBuildResult a = yr::BuildRenderAcceleration(scene.Geometry(), serial);
    const yr::RenderAccelerationBuildResult b = yr::BuildRenderAcceleration(scene.Geometry(), threaded);
    Check(a.errors.empty() && b.errors.empty(), std::string(name) + ": build errors");
    Check(a.acceleration.kind == b.acceleration.kind, std::string(name) + ": acceleration kinds differ");
    if (a.acceleration.kind == yr::RenderAccelerationKind::FlatReference) {
        Check(SameBvh(a.acceleration.flat, b.acceleration.flat), std::string(name) + ": threaded flat BVH differs");
    } else {
        const yr::TwoLevelBvh& x = a.acceleration.two_level;
        const yr::TwoLevelBvh& y = b.acceleration.two_level;
        bool same = x.blases.size() == y.blases.size() && SameBvh(x.tlas, y.tlas);
        for (size_t k = 0; same && k < x.blases.size(); ++k) same = SameBvh(x.blases[k].bvh, y.blases[k].bvh);
        Check(same, std::string(name) + ": threaded two-level BVH differs");
    }
    yr::PrepareLightSampling(scene);
    // A small texture so the texture marshalling runs too.
    yr::RenderTexture tex;
    tex.width = 3;
    tex.height = 2;
    tex.texels.assign(6, yr::Color4f{0.25f, 0.5f, 0.75f, 1.0f});
    scene.textures.push_back(tex);
    if (!scene.materials.empty()) scene.materials[0].reflectance.texture = static_cast<int>(scene.textures.size()) - 1;
    yrhip::Marshal m;
    yrh_geometry_desc g{};
    yrh_shading_desc sd{};
    yrhip::DescribeGeometry(scene, a.acceleration, m, g);
    const std::string err = yrhip::DescribeShading(scene, m, sd);
    Check(err.empty(), std::string(name) + ": DescribeShading: " + err);
    Check(sd.texture_count == static_cast<int32_t>(scene.textures.size()), std::string(name) + ": texture count");
    std::printf("%s: %zu vertices, %d nodes, %d textures\n", name, scene.vertices.size(), a.acceleration.NodeCount(),
                sd.texture_count);
}

}  // namespace

int main() {
    Exercise("cornell", yrhip::MakeCornellScene(1));
    Exercise("triangles", yrhip::MakeRandomTriangleScene(20000, 2));
    Exercise("instanced", yrhip::MakeInstancedScene(200, 2, 4, 4, 32, 16));
    Exercise("room", yrhip::MakeGuidingRoomScene(5000, 6, 32, 16));
    if (failures != 0) return 1;
    std::printf("host check passed\n");
    return 0;
}
