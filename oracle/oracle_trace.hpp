// ORACLE — test infrastructure only (see oracle_trace.cpp).
// Shared CPU restatement of the reference's traversal, used by the ray-cast
// oracle (oracle_trace.cpp) and the path-tracing oracle (oracle_render.cpp).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/yrh.h"

namespace orc {

struct V3 {
    float x, y, z;
};
inline V3 Sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 Cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct Geometry {
    const yrh_vertex* v;
    int64_t nv;
    const uint32_t* idx;
    int64_t ni;
    const yrh_primitive* prim;
    int32_t np;
    const yrh_sphere* sph = nullptr;
    int32_t nsph = 0;
};

struct Ray {
    V3 o, d;
};

struct Pre {
    float o[3], inv[3];
    unsigned char neg[3];
};

inline Pre Precompute(const Ray& r) {
    Pre p;
    const float o[3] = {r.o.x, r.o.y, r.o.z};
    const float d[3] = {r.d.x, r.d.y, r.d.z};
    for (int a = 0; a < 3; ++a) {
        p.o[a] = o[a];
        p.inv[a] = 1.0f / d[a];
        p.neg[a] = std::signbit(p.inv[a]) ? 1 : 0;
    }
    return p;
}

inline bool Bounds(const float* lo, const float* hi, const Pre& r, float t_min, float t_max, float& entry) {
    for (int a = 0; a < 3; ++a) {
        const float nb = r.neg[a] ? hi[a] : lo[a];
        const float fb = r.neg[a] ? lo[a] : hi[a];
        const float an = (nb - r.o[a]) * r.inv[a];
        const float af = (fb - r.o[a]) * r.inv[a];
        t_min = an > t_min ? an : t_min;
        t_max = af < t_max ? af : t_max;
    }
    entry = t_min;
    return t_max >= t_min;
}

struct Hit {
    bool hit = false;
    float t = 0;
    int id = -1;
    float u = 0, v = 0;
    int inst = -1;
};

// Moller-Trumbore exactly as the reference; `tie` enables the id tie-break.
inline void Triangle(const Geometry& g, int mesh, int local, int flat_id, const Ray& r, float t_min, bool tie, Hit& best) {
    if (mesh < 0 || mesh >= g.np || local < 0) return;
    const yrh_primitive& p = g.prim[mesh];
    const uint64_t base = static_cast<uint64_t>(p.first_index) + static_cast<uint64_t>(local) * 3u;
    if (base + 2 >= static_cast<uint64_t>(g.ni)) return;
    const uint32_t i0 = g.idx[base], i1 = g.idx[base + 1], i2 = g.idx[base + 2];
    if (i0 >= g.nv || i1 >= g.nv || i2 >= g.nv) return;
    const V3 p0{g.v[i0].position[0], g.v[i0].position[1], g.v[i0].position[2]};
    const V3 p1{g.v[i1].position[0], g.v[i1].position[1], g.v[i1].position[2]};
    const V3 p2{g.v[i2].position[0], g.v[i2].position[1], g.v[i2].position[2]};
    const V3 e1 = Sub(p1, p0), e2 = Sub(p2, p0);
    const V3 pvec = Cross(r.d, e2);
    const float det = Dot(e1, pvec);
    if (std::fabs(det) < 1.0e-8f) return;
    const float inv_det = 1.0f / det;
    const V3 tvec = Sub(r.o, p0);
    const float u = Dot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return;
    const V3 qvec = Cross(tvec, e1);
    const float v = Dot(r.d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return;
    const float t = Dot(e2, qvec) * inv_det;
    if (t <= t_min || t > best.t) return;
    if (tie && t == best.t && best.hit && best.id >= 0 && flat_id >= best.id) return;
    best.hit = true;
    best.t = t;
    best.id = flat_id;
    best.u = u;
    best.v = v;
}

// IntersectAnalyticSphere (src/accel/bvh_traversal.cpp:123-150) over
// IntersectSphere / SphereNormal / SphereUv (src/geometry/intersection.cpp:108-151):
// the nearer root in (t_min, best t), else the farther one; a hit records
// id = -2 - sphere (the device's encoding of a sphere hit) and the uv of the
// outward normal. A bad handle is skipped.
inline void Sphere(const Geometry& g, int sphere, const Ray& r, float t_min, Hit& best) {
    if (sphere < 0 || sphere >= g.nsph) return;
    const yrh_sphere& s = g.sph[sphere];
    const V3 c{s.center[0], s.center[1], s.center[2]};
    const V3 oc = Sub(r.o, c);
    const float a = Dot(r.d, r.d);
    const float b = 2.0f * Dot(oc, r.d);
    const float cc = Dot(oc, oc) - s.radius * s.radius;
    const float disc = b * b - 4.0f * a * cc;
    if (disc < 0.0f) return;
    const float sq = std::sqrt(disc);
    const float inv_2a = 0.5f / a;
    float t = (-b - sq) * inv_2a;
    if (!(t > t_min && t < best.t)) {
        t = (-b + sq) * inv_2a;
        if (!(t > t_min && t < best.t)) return;
    }
    if (t >= best.t) return;
    const V3 p{r.o.x + r.d.x * t, r.o.y + r.d.y * t, r.o.z + r.d.z * t};
    const V3 nv = Sub(p, c);
    const float len = std::sqrt(Dot(nv, nv));
    const V3 n = len > 0.0f ? V3{nv.x / len, nv.y / len, nv.z / len} : V3{0.0f, 0.0f, 0.0f};
    constexpr float kPi = 3.14159265358979323846f;
    const float y = std::clamp(n.y, -1.0f, 1.0f);
    best.hit = true;
    best.t = t;
    best.id = -2 - sphere;
    best.u = (std::atan2(n.x, n.z) + kPi) / (2.0f * kPi);
    best.v = (kPi - std::acos(y)) / kPi;
    best.inst = -1;
}

struct Counters {
    uint32_t nodes = 0, tris = 0, spheres = 0;
};

// IntersectBvh on one binary BVH. In occlusion mode the traversal stops at the
// first accepted triangle and the box upper bound stays t_max.
inline Hit Bvh(const Geometry& g, const yrh_bvh_desc& b, const Ray& r, float t_min, float t_max, bool occlusion, Counters& c) {
    Hit best;
    best.t = t_max;
    if (b.node_count <= 0) return best;
    const Pre pre = Precompute(r);
    float entry;
    ++c.nodes;
    if (!Bounds(b.nodes[0].bounds_min, b.nodes[0].bounds_max, pre, t_min, best.t, entry)) return best;
    std::vector<std::pair<int, float>> stack;
    stack.reserve(64);
    stack.emplace_back(0, entry);
    while (!stack.empty()) {
        const auto [node, e] = stack.back();
        stack.pop_back();
        if (e > best.t) continue;
        if (node < 0 || node >= b.node_count) continue;
        const yrh_bvh_node& n = b.nodes[node];
        if (n.primitive_count > 0) {
            for (int i = n.payload_offset; i < n.payload_offset + n.primitive_count; ++i) {
                if (i < 0 || i >= b.primitive_index_count) continue;
                const int pi = b.primitive_indices[i];
                if (pi < 0 || pi >= b.primitive_count) continue;
                const yrh_bvh_primitive& bp = b.primitives[pi];
                if (bp.kind == 0) {
                    ++c.tris;
                    Triangle(g, bp.mesh_primitive, bp.local_triangle, bp.flat_triangle_index, r, t_min, !occlusion,
                             best);
                } else {
                    ++c.spheres;
                    Sphere(g, bp.sphere, r, t_min, best);
                }
                if (occlusion && best.hit) return best;
            }
        } else {
            const int li = node + 1, ri = n.payload_offset;
            float le = t_min, re = t_min;
            bool hl = false, hr = false;
            if (li >= 0 && li < b.node_count) {
                ++c.nodes;
                hl = Bounds(b.nodes[li].bounds_min, b.nodes[li].bounds_max, pre, t_min, best.t, le);
            }
            if (ri >= 0 && ri < b.node_count) {
                ++c.nodes;
                hr = Bounds(b.nodes[ri].bounds_min, b.nodes[ri].bounds_max, pre, t_min, best.t, re);
            }
            if (hl && hr) {
                if (le <= re) {
                    stack.emplace_back(ri, re);
                    stack.emplace_back(li, le);
                } else {
                    stack.emplace_back(li, le);
                    stack.emplace_back(ri, re);
                }
            } else if (hl) {
                stack.emplace_back(li, le);
            } else if (hr) {
                stack.emplace_back(ri, re);
            }
        }
    }
    return best;
}

struct TwoLevel {
    const yrh_bvh_desc* blases;
    int32_t nblas;
    const yrh_instance* inst;
    int32_t ninst;
    const yrh_bvh_desc* tlas;
    const yrh_tlas_primitive* tprim;
    int32_t ntprim;
};

inline V3 XformPoint(const float* m, V3 p) {
    return V3{m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12], m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
              m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]};
}
inline V3 XformVector(const float* m, V3 v) {
    return V3{m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
              m[2] * v.x + m[6] * v.y + m[10] * v.z};
}

inline Hit Tlas(const Geometry& g, const TwoLevel& tl, const Ray& r, float t_min, float t_max, bool occlusion, Counters& c) {
    Hit best;
    best.t = t_max;
    const yrh_bvh_desc& t = *tl.tlas;
    if (t.node_count <= 0) return best;
    const Pre pre = Precompute(r);
    float entry;
    ++c.nodes;
    if (!Bounds(t.nodes[0].bounds_min, t.nodes[0].bounds_max, pre, t_min, best.t, entry)) return best;
    std::vector<std::pair<int, float>> stack;
    stack.emplace_back(0, entry);
    while (!stack.empty()) {
        const auto [node, e] = stack.back();
        stack.pop_back();
        if (e > best.t) continue;
        if (node < 0 || node >= t.node_count) continue;
        const yrh_bvh_node& n = t.nodes[node];
        if (n.primitive_count == 0) {
            const int li = node + 1, ri = n.payload_offset;
            float le = t_min, re = t_min;
            bool hl = false, hr = false;
            if (li >= 0 && li < t.node_count) {
                ++c.nodes;
                hl = Bounds(t.nodes[li].bounds_min, t.nodes[li].bounds_max, pre, t_min, best.t, le);
            }
            if (ri >= 0 && ri < t.node_count) {
                ++c.nodes;
                hr = Bounds(t.nodes[ri].bounds_min, t.nodes[ri].bounds_max, pre, t_min, best.t, re);
            }
            if (hl && hr) {
                if (le <= re) {
                    stack.emplace_back(ri, re);
                    stack.emplace_back(li, le);
                } else {
                    stack.emplace_back(li, le);
                    stack.emplace_back(ri, re);
                }
            } else if (hl) {
                stack.emplace_back(li, le);
            } else if (hr) {
                stack.emplace_back(ri, re);
            }
            continue;
        }
        for (int i = n.payload_offset; i < n.payload_offset + n.primitive_count; ++i) {
            if (i < 0 || i >= t.primitive_index_count) continue;
            const int pi = t.primitive_indices[i];
            if (pi < 0 || pi >= tl.ntprim) continue;
            const yrh_tlas_primitive& tp = tl.tprim[pi];
            if (tp.kind != 0) {  // IntersectTlasSphere (two_level_bvh.cpp:231-253): world ray, no instance
                ++c.spheres;
                Sphere(g, tp.sphere, r, t_min, best);
                if (occlusion && best.hit) return best;
                continue;
            }
            if (tp.instance < 0 || tp.instance >= tl.ninst) continue;
            const yrh_instance& in = tl.inst[tp.instance];
            if (in.blas_index < 0 || in.blas_index >= tl.nblas) continue;
            const Ray obj{XformPoint(in.world_to_object, r.o), XformVector(in.world_to_object, r.d)};
            Hit h = Bvh(g, tl.blases[in.blas_index], obj, t_min, best.t, occlusion, c);
            if (h.hit && (occlusion || h.t < best.t)) {
                h.inst = tp.instance;
                best = h;
                if (occlusion) return best;
            }
        }
    }
    return best;
}

}  // namespace orc
