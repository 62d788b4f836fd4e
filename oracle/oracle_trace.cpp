// ORACLE — test infrastructure only. Nothing in the product links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg use it, to check the product's results against.
//
// CPU restatement of the reference's ray casting over its own BVH arrays:
//   IntersectBounds      include/yaoray/core/bounds.hpp:64-90
//   IntersectTriangle    src/accel/bvh_traversal.cpp:52-121 (reads triangles
//                        through primitive_indices -> RenderBvhPrimitive ->
//                        RenderPrimitive -> indices -> vertices, as the reference)
//   IntersectBvh         src/accel/bvh_traversal.cpp:253-345
//   IntersectTwoLevelBvh src/accel/two_level_bvh.cpp:365-481
//   analytic spheres     src/accel/bvh_traversal.cpp:123-150,
//                        src/geometry/intersection.cpp:108-151
// plus a brute-force closest hit over every triangle (same Moller-Trumbore and
// tie rule) that pins the BVH result independently of any tree, and an
// early-exit occlusion traversal (same visit order; exits at the first hit).
// Pinned by tests/test_oracle_trace.py against the reference's known answers
// (tests/accel/bvh_traversal_tests.cpp), brute force, and oracle/_ref.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/yrh.h"

#include "oracle_trace.hpp"

namespace {

using namespace orc;

template <typename Fn>
void ParallelFor(int64_t n, int threads, Fn&& fn) {
    if (threads <= 1 || n < 1024) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = 256;
    std::atomic_int64_t next{0};
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&] {
            while (true) {
                const int64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                const int64_t e = std::min(n, b + chunk);
                for (int64_t i = b; i < e; ++i) fn(i);
            }
        });
    }
    for (auto& th : pool) th.join();
}

Ray LoadRay(const float* org, const float* dir, int64_t i, float& t_min, float& t_max) {
    t_min = org[4 * i + 3];
    t_max = dir[4 * i + 3];
    return Ray{V3{org[4 * i], org[4 * i + 1], org[4 * i + 2]}, V3{dir[4 * i], dir[4 * i + 1], dir[4 * i + 2]}};
}

void Store(const Hit& h, int64_t i, float* hit, int32_t* inst, uint32_t* occluded, const Counters& c,
           uint32_t* node_tests, uint32_t* tri_tests, uint32_t* sphere_tests = nullptr) {
    if (hit != nullptr) {
        const int id = h.hit ? h.id : -1;
        float fid;
        std::memcpy(&fid, &id, 4);
        hit[4 * i] = h.t;
        hit[4 * i + 1] = h.u;
        hit[4 * i + 2] = h.v;
        hit[4 * i + 3] = fid;
    }
    if (inst != nullptr) inst[i] = h.hit ? h.inst : -1;
    if (occluded != nullptr) occluded[i] = h.hit ? 1u : 0u;
    if (node_tests != nullptr) node_tests[i] = c.nodes;
    if (tri_tests != nullptr) tri_tests[i] = c.tris;
    if (sphere_tests != nullptr) sphere_tests[i] = c.spheres;
}

}  // namespace

extern "C" {
#pragma GCC visibility push(default)

// Flat closest hit (occlusion = 0) or early-exit occlusion (occlusion = 1).
int oracle_trace_flat(const yrh_vertex* v, int64_t nv, const uint32_t* idx, int64_t ni, const yrh_primitive* p,
                      int32_t np, const yrh_sphere* sph, int32_t nsph, const yrh_bvh_desc* bvh, const float* org_tmin,
                      const float* dir_tmax, int64_t n, int32_t occlusion, int32_t threads, float* hit,
                      uint32_t* occluded, uint32_t* node_tests, uint32_t* tri_tests, uint32_t* sphere_tests) {
    if (bvh == nullptr || n < 0) return -1;
    const Geometry g{v, nv, idx, ni, p, np, sph, nsph};
    ParallelFor(n, threads, [&](int64_t i) {
        float t_min, t_max;
        const Ray r = LoadRay(org_tmin, dir_tmax, i, t_min, t_max);
        Counters c;
        const Hit h = Bvh(g, *bvh, r, t_min, t_max, occlusion != 0, c);
        Store(h, i, hit, nullptr, occluded, c, node_tests, tri_tests, sphere_tests);
    });
    return 0;
}

int oracle_trace_two_level(const yrh_vertex* v, int64_t nv, const uint32_t* idx, int64_t ni, const yrh_primitive* p,
                           int32_t np, const yrh_sphere* sph, int32_t nsph, const yrh_bvh_desc* blases, int32_t nblas,
                           const yrh_instance* inst, int32_t ninst, const yrh_bvh_desc* tlas,
                           const yrh_tlas_primitive* tprim, int32_t ntprim, const float* org_tmin,
                           const float* dir_tmax, int64_t n, int32_t occlusion, int32_t threads, float* hit,
                           int32_t* inst_out, uint32_t* occluded, uint32_t* node_tests, uint32_t* tri_tests,
                           uint32_t* sphere_tests) {
    if (tlas == nullptr || n < 0) return -1;
    const Geometry g{v, nv, idx, ni, p, np, sph, nsph};
    const TwoLevel tl{blases, nblas, inst, ninst, tlas, tprim, ntprim};
    ParallelFor(n, threads, [&](int64_t i) {
        float t_min, t_max;
        const Ray r = LoadRay(org_tmin, dir_tmax, i, t_min, t_max);
        Counters c;
        const Hit h = Tlas(g, tl, r, t_min, t_max, occlusion != 0, c);
        Store(h, i, hit, inst_out, occluded, c, node_tests, tri_tests, sphere_tests);
    });
    return 0;
}

// Closest hit over every triangle in flat order (no tree): the tie rule makes
// the result independent of visit order, so it must equal the BVH result.
int oracle_brute_force(const yrh_vertex* v, int64_t nv, const uint32_t* idx, int64_t ni, const yrh_primitive* p,
                       int32_t np, const yrh_sphere* sph, int32_t nsph, const float* org_tmin, const float* dir_tmax,
                       int64_t n, int32_t threads, float* hit) {
    const Geometry g{v, nv, idx, ni, p, np, sph, nsph};
    ParallelFor(n, threads, [&](int64_t i) {
        float t_min, t_max;
        const Ray r = LoadRay(org_tmin, dir_tmax, i, t_min, t_max);
        Hit best;
        best.t = t_max;
        int flat = 0;
        for (int m = 0; m < np; ++m) {
            const int tris = static_cast<int>(p[m].index_count / 3);
            for (int k = 0; k < tris; ++k) Triangle(g, m, k, flat + k, r, t_min, true, best);
            flat += tris;
        }
        // A sphere takes only a strictly closer hit and a triangle takes an
        // equal-t hit from a sphere, so the closest hit does not depend on
        // the visit order except between two spheres at exactly equal t.
        for (int k = 0; k < nsph; ++k) Sphere(g, k, r, t_min, best);
        Counters c;
        Store(best, i, hit, nullptr, nullptr, c, nullptr, nullptr);
    });
    return 0;
}

#pragma GCC visibility pop
}  // extern "C"
