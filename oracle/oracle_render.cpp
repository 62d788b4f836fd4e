// ORACLE — test infrastructure only (see oracle_trace.cpp).
//
// CPU restatement of the reference's path-tracing loop over the same inputs
// the GPU backend receives (yrh.h structs):
//   Sampler                 src/sampling/sampler.cpp:34-237
//   camera ray              src/scene/camera.cpp:9-39
//   TraceVisibleSurface     src/integrators/surface_query.cpp:101-200
//   material / normals      src/shading/material_evaluator.cpp:35-184,
//                           src/geometry/intersection.cpp:8-106
//   textures, alpha skip    oracle_texture.hpp; surface_query.cpp:35-97, 187-194
//   BSDFs                   src/shading/bsdf.cpp:17-329, bsdf_microfacet.cpp:20-298,
//                           bsdf_common.cpp:13-47
//   lights                  src/lighting/light_sampling.cpp:60-396,
//                           src/lighting/environment.cpp:54-219, src/shading/texture.cpp:116-130
//   MIS                     src/integrators/mis.cpp:14-31
//   path loop, NEE, shadows src/integrators/path_integrator.cpp:23-197,
//                           path_direct_lighting.cpp:23-283, path_integrator_common.cpp:15-201
//   film                    src/film/film.cpp:32-56 (FP32, as the reference)
//   pass driver             src/backends/cpu/cpu_path_tracer.cpp:109-180
//   adaptive sampling       cpu_path_tracer.cpp:95-106, 183-199 (active pixels),
//                           src/film/film.cpp:98-121 (IsConverged)
// Same operation order and the same libm calls as the reference, so on one
// machine its films equal the reference build's bit for bit (pinned by
// tests/test_oracle_render.py against oracle/_ref).
#include <algorithm>
#include <atomic>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "oracle_texture.hpp"
#include "oracle_trace.hpp"

namespace {

using orc::Cross;
using orc::Dot;
using orc::Sub;
using orc::V3;

constexpr float Pi = 3.14159265358979323846f;
constexpr float TwoPi = 2.0f * Pi;
const float kInf = std::numeric_limits<float>::infinity();

V3 Add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 Mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
V3 Div(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
V3 Had(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
V3 Neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }
float LengthSq(V3 v) { return Dot(v, v); }
float Length(V3 v) { return std::sqrt(LengthSq(v)); }
V3 Normalize(V3 v) {
    const float len = Length(v);
    return len > 0.0f ? Div(v, len) : V3{0, 0, 0};
}
bool NearBlack(V3 c) { return c.x <= 0.0f && c.y <= 0.0f && c.z <= 0.0f; }
float MaxComponent(V3 c) { return std::max(c.x, std::max(c.y, c.z)); }
bool Finite(V3 c) { return std::isfinite(c.x) && std::isfinite(c.y) && std::isfinite(c.z); }
V3 Load3(const float* p) { return V3{p[0], p[1], p[2]}; }
struct V2 {
    float x, y;
};

// ---- sampler ------------------------------------------------------------------

constexpr uint64_t DefaultSeed = 0x9E3779B97F4A7C15ull;

uint64_t Mix64(uint64_t v) {
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}
uint32_t Mix32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7FEB352Du;
    v ^= v >> 15;
    v *= 0x846CA68Bu;
    return v ^ (v >> 16);
}
uint32_t Hash32(uint64_t seed, uint32_t dim, int lane) {
    const uint64_t m = Mix64(seed ^ (static_cast<uint64_t>(dim) * 0xD1B54A32D192ED03ull) ^
                             (static_cast<uint64_t>(static_cast<uint32_t>(lane)) * 0x94D049BB133111EBull));
    return Mix32(static_cast<uint32_t>(m) ^ static_cast<uint32_t>(m >> 32));
}
float UnitFloat(uint32_t v) { return static_cast<float>(v >> 8) * (1.0f / 16777216.0f); }
uint32_t Reverse(uint32_t v) {
    v = ((v & 0x55555555u) << 1) | ((v >> 1) & 0x55555555u);
    v = ((v & 0x33333333u) << 2) | ((v >> 2) & 0x33333333u);
    v = ((v & 0x0F0F0F0Fu) << 4) | ((v >> 4) & 0x0F0F0F0Fu);
    v = ((v & 0x00FF00FFu) << 8) | ((v >> 8) & 0x00FF00FFu);
    return (v << 16) | (v >> 16);
}
uint32_t Sobol(uint32_t index, int axis) {
    if (axis == 0) return Reverse(index);
    uint32_t value = 0, direction = 0x80000000u;
    for (; index != 0; index >>= 1) {
        if (index & 1u) value ^= direction;
        direction ^= direction >> 1;
    }
    return value;
}
uint32_t Owen(uint32_t v, uint32_t seed) {
    v = Reverse(v);
    v += seed;
    v ^= v * 0x6C50B47Cu;
    v ^= v * 0xB82F1E52u;
    v ^= v * 0xC7AFE638u;
    v ^= v * 0x8D22F6E6u;
    return Reverse(v);
}
uint32_t Spread16(uint32_t v) {
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
float ClampUnit(float v) { return std::clamp(v, 0.0f, std::nextafter(1.0f, 0.0f)); }

struct Sampler {
    int kind = 0;
    uint64_t scene_seed = 0, pixel_seed = 0;
    uint32_t index = 0;
    int sample = 0, spp = 1, bounce = 0;

    Sampler(int k, uint64_t seed, int x, int y, int s, int n) : kind(k), scene_seed(seed), sample(s), spp(std::max(1, n)) {
        pixel_seed = Mix64(seed ^ (static_cast<uint64_t>(static_cast<uint32_t>(x)) << 32) ^ static_cast<uint32_t>(y));
        if (kind == YRH_SAMPLER_ZSOBOL) {
            const unsigned bits = std::bit_width(std::bit_ceil(static_cast<unsigned>(spp))) - 1u;
            index = ((Spread16(static_cast<uint32_t>(x)) | (Spread16(static_cast<uint32_t>(y)) << 1)) << bits) |
                    static_cast<uint32_t>(s);
        } else {
            index = static_cast<uint32_t>(s);
        }
    }
    uint32_t Abs(uint32_t d) const { return d == 0 ? d : d + static_cast<uint32_t>(bounce) * 16u; }
    float At1(uint32_t d, int lane) const {
        if (kind == YRH_SAMPLER_OWEN_SOBOL) return UnitFloat(Owen(Sobol(index, 0), Hash32(pixel_seed ^ DefaultSeed, d, lane)));
        if (kind == YRH_SAMPLER_ZSOBOL) return UnitFloat(Owen(Sobol(index, 0), Hash32(scene_seed, d, lane)));
        return UnitFloat(Hash32(pixel_seed ^ static_cast<uint64_t>(sample), d, lane));
    }
    void At2(uint32_t d, int lane, float& a, float& b) const {
        if (d == 0 && spp == 1) {
            a = b = 0.5f;
            return;
        }
        if (kind == YRH_SAMPLER_STRATIFIED) {
            const int count = d == 0 ? spp : 1;
            const int idx = d == 0 ? sample : lane;
            const int cols = static_cast<int>(std::ceil(std::sqrt(static_cast<float>(std::max(1, count)))));
            const int rows = (count + cols - 1) / cols;
            const int c = std::clamp(idx, 0, count - 1);
            a = ClampUnit((static_cast<float>(c % cols) + At1(d, lane)) / static_cast<float>(cols));
            b = ClampUnit((static_cast<float>(c / cols) + At1(d + 1, lane)) / static_cast<float>(rows));
            return;
        }
        if (kind == YRH_SAMPLER_OWEN_SOBOL || kind == YRH_SAMPLER_ZSOBOL) {
            const uint64_t s = kind == YRH_SAMPLER_OWEN_SOBOL ? pixel_seed : scene_seed;
            a = UnitFloat(Owen(Sobol(index, 0), Hash32(s, d, lane)));
            b = UnitFloat(Owen(Sobol(index, 1), Hash32(s, d + 1, lane)));
            return;
        }
        a = At1(d, lane);
        b = At1(d + 1, lane);
    }
    float Get1(uint32_t d) const { return At1(Abs(d), 0); }
    void Get2(uint32_t d, float& a, float& b) const { At2(Abs(d), 0, a, b); }
};

enum : uint32_t {
    DimPixel = 0,
    DimLightSelect = 2,
    DimLightSurface = 3,
    DimBsdf = 5,
    DimRoulette = 7,
    DimAnalyticSelect = 14
};

// ---- scene ------------------------------------------------------------------------

struct Scene {
    orc::Geometry g;
    const yrh_geometry_desc* gd;
    const yrh_shading_desc* sd;
    bool two_level;
    orc::TwoLevel tl;
    std::vector<ortex::Texture> tex;
    const ortex::Texture* Tex(int index) const {
        return index >= 0 && static_cast<size_t>(index) < tex.size() ? &tex[static_cast<size_t>(index)] : nullptr;
    }
};

struct Hit {
    bool hit = false;
    float t = 0, u = 0, v = 0;
    int tri = -1, inst = -1;
};

Hit Trace(const Scene& s, V3 o, V3 d, float t_min, float t_max, orc::Counters& c) {
    const orc::Ray r{o, d};
    const orc::Hit h = s.two_level ? orc::Tlas(s.g, s.tl, r, t_min, t_max, false, c)
                                   : orc::Bvh(s.g, s.gd->flat, r, t_min, t_max, false, c);
    Hit out;
    out.hit = h.hit;
    out.t = h.t;
    out.u = h.u;
    out.v = h.v;
    out.tri = h.id;
    out.inst = h.inst;
    return out;
}

struct TriRef {
    int prim = -1, local = -1;
};

TriRef Locate(const Scene& s, int flat) {
    int base = 0;
    for (int p = 0; p < s.g.np; ++p) {
        const int n = static_cast<int>(s.g.prim[p].index_count / 3);
        if (flat < base + n) return TriRef{p, flat - base};
        base += n;
    }
    return TriRef{};
}

V3 VertexPos(const Scene& s, uint32_t vi) { return Load3(s.g.v[vi].position); }

void TriVerts(const Scene& s, TriRef t, uint32_t& i0, uint32_t& i1, uint32_t& i2) {
    const uint32_t base = s.g.prim[t.prim].first_index + static_cast<uint32_t>(t.local) * 3;
    i0 = s.g.idx[base];
    i1 = s.g.idx[base + 1];
    i2 = s.g.idx[base + 2];
}

V3 GeometricNormal(const Scene& s, TriRef t) {
    uint32_t i0, i1, i2;
    TriVerts(s, t, i0, i1, i2);
    const V3 p0 = VertexPos(s, i0), p1 = VertexPos(s, i1), p2 = VertexPos(s, i2);
    return Normalize(Cross(Sub(p1, p0), Sub(p2, p0)));
}

V3 TransformNormal(const float* m, V3 n) {
    const float a = m[0], b = m[4], c = m[8], d = m[1], e = m[5], f = m[9], g = m[2], h = m[6], i = m[10];
    const float det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    if (std::fabs(det) <= 1.0e-12f) {
        const V3 fb = Normalize(orc::XformVector(m, n));
        return LengthSq(fb) > 0.0f ? fb : Normalize(n);
    }
    const float inv = 1.0f / det;
    return Normalize(V3{((e * i - f * h) * n.x + (f * g - d * i) * n.y + (d * h - e * g) * n.z) * inv,
                        ((c * h - b * i) * n.x + (a * i - c * g) * n.y + (b * g - a * h) * n.z) * inv,
                        ((b * f - c * e) * n.x + (c * d - a * f) * n.y + (a * e - b * d) * n.z) * inv});
}

V3 ResolveShadingNormal(const Scene& s, TriRef t, float u, float v, V3 geo) {
    V3 n = Normalize(geo);
    const yrh_primitive& p = s.g.prim[t.prim];
    if (p.has_normals) {
        uint32_t i0, i1, i2;
        TriVerts(s, t, i0, i1, i2);
        const float w = 1.0f - u - v;
        const float* n0 = s.g.v[i0].normal;
        const float* n1 = s.g.v[i1].normal;
        const float* n2 = s.g.v[i2].normal;
        const V3 interp = Normalize(V3{n0[0] * w + n1[0] * u + n2[0] * v, n0[1] * w + n1[1] * u + n2[1] * v,
                                       n0[2] * w + n1[2] * u + n2[2] * v});
        if (LengthSq(interp) > 0.0f) n = interp;
    }
    if (Dot(n, geo) < 0.0f) n = Neg(n);
    if (LengthSq(n) == 0.0f) return Normalize(geo);
    return n;
}

struct Surface {
    bool hit = false;
    bool exhausted = false;
    Hit h;
    yrh_material m{};  // the material with its textures applied
    V3 shading;
    int prim = -1;
    int sphere = -1;
    TriRef tri;
};

// A camera ray's differentials: the directions one pixel to the right and one
// pixel down, from the same origin (src/scene/camera.cpp:20-39).
struct Differentials {
    V3 dx, dy;
};

V2 InterpolateUv(const Scene& s, TriRef t, float u, float v) {
    uint32_t i0, i1, i2;
    TriVerts(s, t, i0, i1, i2);
    const float w = 1.0f - u - v;
    const float* a = s.g.v[i0].uv;
    const float* b = s.g.v[i1].uv;
    const float* c = s.g.v[i2].uv;
    return V2{a[0] * w + b[0] * u + c[0] * v, a[1] * w + b[1] * u + c[1] * v};
}

// ComputeTextureFootprint (src/integrators/surface_query.cpp:35-97): the uv
// change per pixel from where the differential rays meet the hit's plane.
ortex::Footprint TextureFootprint(const Scene& s, TriRef t, V3 o, const Differentials* rd, V3 hp, V3 geo,
                                  const float* o2w) {
    if (rd == nullptr || !s.g.prim[t.prim].has_uvs) return {};
    uint32_t i0, i1, i2;
    TriVerts(s, t, i0, i1, i2);
    V3 p0 = VertexPos(s, i0), p1 = VertexPos(s, i1), p2 = VertexPos(s, i2);
    if (o2w != nullptr) {
        p0 = orc::XformPoint(o2w, p0);
        p1 = orc::XformPoint(o2w, p1);
        p2 = orc::XformPoint(o2w, p2);
    }
    const float* uv0 = s.g.v[i0].uv;
    const float* uv1 = s.g.v[i1].uv;
    const float* uv2 = s.g.v[i2].uv;
    const float du1 = uv1[0] - uv0[0], dv1 = uv1[1] - uv0[1];
    const float du2 = uv2[0] - uv0[0], dv2 = uv2[1] - uv0[1];
    const float uv_det = du1 * dv2 - dv1 * du2;
    if (std::fabs(uv_det) < 1.0e-12f) return {};
    const float inv_uv = 1.0f / uv_det;
    const V3 e1 = Sub(p1, p0), e2 = Sub(p2, p0);
    const V3 dpdu = Mul(Sub(Mul(e1, dv2), Mul(e2, dv1)), inv_uv);
    const V3 dpdv = Mul(Sub(Mul(e2, du1), Mul(e1, du2)), inv_uv);
    auto on_plane = [&](V3 dir) {
        const float den = Dot(geo, dir);
        if (std::fabs(den) < 1.0e-12f) return hp;
        const float tt = Dot(geo, Sub(hp, o)) / den;
        return Add(o, Mul(dir, tt));
    };
    const V3 dpdx = Sub(on_plane(rd->dx), hp);
    const V3 dpdy = Sub(on_plane(rd->dy), hp);
    const float a = Dot(dpdu, dpdu), b = Dot(dpdu, dpdv), c = Dot(dpdv, dpdv);
    const float gram = a * c - b * b;
    if (std::fabs(gram) < 1.0e-20f) return {};
    const float inv_gram = 1.0f / gram;
    auto solve = [&](V3 dp, float& du, float& dv) {
        const float ur = Dot(dp, dpdu), vr = Dot(dp, dpdv);
        du = (ur * c - vr * b) * inv_gram;
        dv = (vr * a - ur * b) * inv_gram;
    };
    ortex::Footprint f;
    solve(dpdx, f.dudx, f.dvdx);
    solve(dpdy, f.dudy, f.dvdy);
    if (!std::isfinite(f.dudx) || !std::isfinite(f.dvdx) || !std::isfinite(f.dudy) || !std::isfinite(f.dvdy)) return {};
    return f;
}

// ResolveNormalMap (src/shading/material_evaluator.cpp:35-92): the map's
// tangent-space normal (x, y scaled) in the frame of the interpolated
// tangent, made orthogonal to the shading normal, and its handedness
// (negated for an instance whose transform mirrors).
V3 NormalMap(const Scene& s, TriRef t, float u, float v, const yrh_material& m, V2 uv, V3 sn, const float* o2w,
             ortex::Footprint fp) {
    const yrh_primitive& prim = s.g.prim[t.prim];
    const ortex::Texture* nm = s.Tex(m.normal_map);
    if (!prim.has_tangents || nm == nullptr) return sn;
    const ortex::C4 e = ortex::Sample(*nm, uv.x, uv.y, fp);
    V3 ts{e.x * 2.0f - 1.0f, e.y * 2.0f - 1.0f, e.z * 2.0f - 1.0f};
    ts.x *= m.normal_scale;
    ts.y *= m.normal_scale;
    ts = Normalize(ts);
    if (LengthSq(ts) == 0.0f) return sn;
    uint32_t i0, i1, i2;
    TriVerts(s, t, i0, i1, i2);
    const float w = 1.0f - u - v;
    const yrh_vertex &v0 = s.g.v[i0], &v1 = s.g.v[i1], &v2 = s.g.v[i2];
    V3 tan = Normalize(V3{v0.tangent[0] * w + v1.tangent[0] * u + v2.tangent[0] * v,
                          v0.tangent[1] * w + v1.tangent[1] * u + v2.tangent[1] * v,
                          v0.tangent[2] * w + v1.tangent[2] * u + v2.tangent[2] * v});
    if (o2w != nullptr) tan = orc::XformVector(o2w, tan);
    const V3 proj = Sub(tan, Mul(sn, Dot(tan, sn)));
    if (LengthSq(proj) <= 1.0e-12f) return sn;
    tan = Normalize(proj);
    if (LengthSq(tan) == 0.0f) return sn;
    float orientation = 1.0f;
    if (o2w != nullptr) {
        const float* q = o2w;
        const float det = q[0] * (q[5] * q[10] - q[9] * q[6]) - q[4] * (q[1] * q[10] - q[9] * q[2]) +
                          q[8] * (q[1] * q[6] - q[5] * q[2]);
        orientation = det < 0.0f ? -1.0f : 1.0f;
    }
    const float hand = (v0.tangent_handedness * w + v1.tangent_handedness * u + v2.tangent_handedness * v) * orientation;
    const V3 bit = Mul(Cross(sn, tan), hand);
    const V3 mapped = Normalize(Add(Add(Mul(tan, ts.x), Mul(bit, ts.y)), Mul(sn, ts.z)));
    return LengthSq(mapped) > 0.0f ? mapped : sn;
}

constexpr int MaxAlphaSkips = 64;         // surface_query.cpp:12
constexpr float AlphaSkipOffset = 1.0e-4f;  // surface_query.cpp:13

// TraceVisibleSurface (src/integrators/surface_query.cpp:101-200): the closest
// hit, its material with textures applied (EvaluateMaterialAtSurface,
// material_evaluator.cpp:112-180) and shading normal; a hit whose alpha is
// below 0.5 is skipped by tracing again from just past it, at most 64 times
// (then `exhausted`). Sphere hits (:129-147): material from the sphere,
// shading normal = SphereNormal (negated when flip_normals), not
// face-forwarded, always opaque. `rd`: the camera ray's differentials, for
// texture footprints (NULL for every other ray).
Surface TraceSurface(const Scene& s, V3 o, V3 d, float t_min, float t_max, orc::Counters& c,
                     const Differentials* rd = nullptr) {
    for (int skip = 0; skip < MaxAlphaSkips; ++skip) {
        Surface out;
        const Hit h = Trace(s, o, d, t_min, t_max, c);
        if (!h.hit || h.tri == -1) return out;
        out.hit = true;
        out.h = h;
        if (h.tri <= -2) {
            const yrh_sphere& sp = s.g.sph[-2 - h.tri];
            const V3 hp = Add(o, Mul(d, h.t));
            const V3 gn = Normalize(Sub(hp, Load3(sp.center)));
            out.sphere = -2 - h.tri;
            out.m = s.sd->materials[sp.material_index];
            out.shading = sp.flip_normals ? Neg(gn) : gn;
            return out;
        }
        out.tri = Locate(s, h.tri);
        out.prim = out.tri.prim;
        const yrh_primitive& prim = s.g.prim[out.prim];
        out.m = s.sd->materials[prim.material_index];
        const bool inst = h.inst >= 0;
        const float* o2w = inst ? s.tl.inst[h.inst].object_to_world : nullptr;
        V3 raw = GeometricNormal(s, out.tri);
        if (inst) raw = TransformNormal(o2w, raw);
        const V3 geo = Dot(raw, Neg(d)) < 0.0f ? Neg(raw) : raw;
        const ortex::Footprint fp = TextureFootprint(s, out.tri, o, rd, Add(o, Mul(d, h.t)), geo, o2w);
        const V2 uv = prim.has_uvs ? InterpolateUv(s, out.tri, h.u, h.v) : V2{0.0f, 0.0f};
        float alpha = out.m.alpha;
        if (prim.has_uvs) {
            if (const ortex::Texture* rt = s.Tex(out.m.reflectance_texture)) {
                const ortex::C4 tc = ortex::Sample(*rt, uv.x, uv.y, fp);
                out.m.reflectance[0] *= tc.x;
                out.m.reflectance[1] *= tc.y;
                out.m.reflectance[2] *= tc.z;
            }
            if (const ortex::Texture* at = s.Tex(out.m.alpha_texture)) alpha *= ortex::Sample(*at, uv.x, uv.y, fp).x;
        }
        V3 sn;
        if (inst) {
            sn = TransformNormal(o2w, ResolveShadingNormal(s, out.tri, h.u, h.v, GeometricNormal(s, out.tri)));
            if (Dot(sn, geo) < 0.0f) sn = Neg(sn);
        } else {
            sn = ResolveShadingNormal(s, out.tri, h.u, h.v, geo);
        }
        if (prim.has_uvs) sn = NormalMap(s, out.tri, h.u, h.v, out.m, uv, sn, o2w, fp);
        const V3 wo = Neg(d);
        out.shading = Dot(sn, wo) < 0.0f ? Neg(sn) : sn;
        if (alpha >= 0.5f) return out;
        t_min = h.t + AlphaSkipOffset;
        if (t_min >= t_max) return Surface{};
    }
    Surface out;
    out.exhausted = true;
    return out;
}

// ---- BSDFs ---------------------------------------------------------------------

constexpr float DeltaRoughness = 1.0e-3f;

bool Above(V3 d, V3 n) { return Dot(d, n) > 0.0f; }
V3 Reflect(V3 d, V3 n) { return Normalize(Sub(d, Mul(n, 2.0f * Dot(d, n)))); }
float AbsDot(V3 a, V3 b) { return std::fabs(Dot(a, b)); }
bool SameHemi(V3 a, V3 b, V3 n) { return Dot(a, n) * Dot(b, n) > 0.0f; }

float Fresnel(float ci, float ei, float et) {
    ci = std::clamp(ci, -1.0f, 1.0f);
    if (ci <= 0.0f) {
        std::swap(ei, et);
        ci = std::fabs(ci);
    }
    const float si = std::sqrt(std::max(0.0f, 1.0f - ci * ci));
    const float st = ei / et * si;
    if (st >= 1.0f) return 1.0f;
    const float ct = std::sqrt(std::max(0.0f, 1.0f - st * st));
    const float rpa = ((et * ci) - (ei * ct)) / ((et * ci) + (ei * ct));
    const float rpe = ((ei * ci) - (et * ct)) / ((ei * ci) + (et * ct));
    return 0.5f * (rpa * rpa + rpe * rpe);
}

bool Refract(V3 wo, V3 n, float eta, V3& wi) {
    const float ci = Dot(n, wo);
    const float s2i = std::max(0.0f, 1.0f - ci * ci);
    const float s2t = eta * eta * s2i;
    if (s2t >= 1.0f) return false;
    const float ct = std::sqrt(std::max(0.0f, 1.0f - s2t));
    wi = Normalize(Add(Mul(Neg(wo), eta), Mul(n, eta * ci - ct)));
    return true;
}

void Frame(V3 n, V3& t, V3& b) {
    const V3 helper = std::fabs(n.z) < 0.999f ? V3{0, 0, 1} : V3{1, 0, 0};
    t = Normalize(Cross(helper, n));
    b = Cross(n, t);
}

V3 CosineHemisphere(V3 n, float sx, float sy) {
    const float u1 = std::clamp(sx, 0.0f, 1.0f);
    const float u2 = std::clamp(sy, 0.0f, 1.0f);
    const float r = std::sqrt(u1);
    const float th = 2.0f * Pi * u2;
    const float lx = r * std::cos(th), ly = r * std::sin(th), lz = std::sqrt(std::max(0.0f, 1.0f - u1));
    V3 t, b;
    Frame(n, t, b);
    return Normalize(Add(Add(Mul(t, lx), Mul(b, ly)), Mul(n, lz)));
}

float Alpha(float r) {
    const float c = std::max(DeltaRoughness, r);
    return c * c;
}
float GgxD(float ch, float a) {
    const float c2 = ch * ch, a2 = a * a, d = c2 * (a2 - 1.0f) + 1.0f;
    return a2 / (Pi * d * d);
}
float G1(float c, float a) {
    const float c2 = c * c;
    if (c2 <= 0.0f) return 0.0f;
    const float t2 = (1.0f - c2) / c2;
    return 2.0f / (1.0f + std::sqrt(1.0f + a * a * t2));
}
V3 Schlick(V3 f0, float c) {
    const float t = std::pow(1.0f - std::clamp(c, 0.0f, 1.0f), 5.0f);
    return Add(Mul(f0, 1.0f - t), Mul(V3{1, 1, 1}, t));
}

V3 ConductorF(V3 f0, float r, V3 wo, V3 wi, V3 n) {
    if (!Above(wo, n) || !Above(wi, n)) return V3{0, 0, 0};
    const V3 h = Normalize(Add(wo, wi));
    if (LengthSq(h) <= 0.0f || !Above(h, n)) return V3{0, 0, 0};
    const float co = std::max(0.0f, Dot(n, wo)), ci = std::max(0.0f, Dot(n, wi));
    const float ch = std::max(0.0f, Dot(n, h)), coh = std::max(0.0f, Dot(wo, h));
    if (co <= 0.0f || ci <= 0.0f || coh <= 0.0f) return V3{0, 0, 0};
    const float a = Alpha(r);
    const float d = GgxD(ch, a), g = G1(co, a) * G1(ci, a);
    return Mul(Schlick(f0, coh), d * g / (4.0f * co * ci));
}
float ConductorPdf(float r, V3 wo, V3 wi, V3 n) {
    if (!Above(wo, n) || !Above(wi, n)) return 0.0f;
    const V3 h = Normalize(Add(wo, wi));
    if (LengthSq(h) <= 0.0f || !Above(h, n)) return 0.0f;
    const float ch = std::max(0.0f, Dot(n, h)), coh = std::max(0.0f, Dot(wo, h));
    if (ch <= 0.0f || coh <= 0.0f) return 0.0f;
    return GgxD(ch, Alpha(r)) * ch / (4.0f * coh);
}
V3 HalfVector(V3 n, float r, float sx, float sy) {
    const float u1 = std::clamp(sx, 0.0f, 0.999999f), u2 = std::clamp(sy, 0.0f, 1.0f);
    const float a = Alpha(r), a2 = a * a;
    const float t2 = a2 * u1 / std::max(1.0e-6f, 1.0f - u1);
    const float ct = 1.0f / std::sqrt(1.0f + t2);
    const float st = std::sqrt(std::max(0.0f, 1.0f - ct * ct));
    const float ph = 2.0f * Pi * u2;
    V3 t, b;
    Frame(n, t, b);
    return Normalize(Add(Add(Mul(t, st * std::cos(ph)), Mul(b, st * std::sin(ph))), Mul(n, ct)));
}
V3 DielReflF(V3 alb, float ior, float r, V3 wo, V3 wi, V3 n) {
    if (!SameHemi(wo, wi, n)) return V3{0, 0, 0};
    const V3 on = Dot(n, wo) >= 0.0f ? n : Neg(n);
    const V3 h = Normalize(Add(wo, wi));
    if (LengthSq(h) <= 0.0f || Dot(h, on) <= 0.0f) return V3{0, 0, 0};
    const float co = AbsDot(on, wo), ci = AbsDot(on, wi), ch = AbsDot(on, h), coh = AbsDot(wo, h);
    if (co <= 0.0f || ci <= 0.0f || coh <= 0.0f) return V3{0, 0, 0};
    const float a = Alpha(r);
    const float d = GgxD(ch, a), g = G1(co, a) * G1(ci, a), f = Fresnel(coh, 1.0f, std::max(1.0f, ior));
    return Mul(alb, f * d * g / (4.0f * co * ci));
}
float DielReflPdf(float ior, float r, V3 wo, V3 wi, V3 n) {
    if (!SameHemi(wo, wi, n)) return 0.0f;
    const V3 on = Dot(n, wo) >= 0.0f ? n : Neg(n);
    const V3 h = Normalize(Add(wo, wi));
    if (LengthSq(h) <= 0.0f || Dot(h, on) <= 0.0f) return 0.0f;
    const float ch = AbsDot(on, h), coh = AbsDot(wo, h);
    if (ch <= 0.0f || coh <= 0.0f) return 0.0f;
    const float f = Fresnel(coh, 1.0f, std::max(1.0f, ior));
    return f * GgxD(ch, Alpha(r)) * ch / (4.0f * coh);
}
V3 TransHalf(V3 wo, V3 wi, V3 n, float ior) {
    const bool ent = Dot(n, wo) >= 0.0f;
    const float ei = ent ? 1.0f : std::max(1.0f, ior), et = ent ? std::max(1.0f, ior) : 1.0f;
    V3 h = Normalize(Add(wo, Mul(wi, et / ei)));
    if (Dot(h, n) < 0.0f) h = Neg(h);
    return h;
}
V3 DielTransF(V3 alb, float ior, float r, V3 wo, V3 wi, V3 n) {
    if (SameHemi(wo, wi, n)) return V3{0, 0, 0};
    const bool ent = Dot(n, wo) >= 0.0f;
    const V3 on = ent ? n : Neg(n);
    const float ei = ent ? 1.0f : std::max(1.0f, ior), et = ent ? std::max(1.0f, ior) : 1.0f, eta = et / ei;
    const V3 h = TransHalf(wo, wi, on, ior);
    if (LengthSq(h) <= 0.0f || Dot(h, on) <= 0.0f) return V3{0, 0, 0};
    const float co = AbsDot(on, wo), ci = AbsDot(on, wi), ch = AbsDot(on, h), woh = Dot(wo, h), wih = Dot(wi, h);
    const float sd = woh + eta * wih;
    if (co <= 0.0f || ci <= 0.0f || ch <= 0.0f || woh <= 0.0f || wih >= 0.0f || sd == 0.0f) return V3{0, 0, 0};
    const float a = Alpha(r);
    const float d = GgxD(ch, a), g = G1(co, a) * G1(ci, a), f = Fresnel(std::fabs(woh), ei, et);
    const float num = std::fabs(wih) * std::fabs(woh) * eta * eta;
    const float den = co * ci * sd * sd;
    return Mul(alb, (1.0f - f) * d * g * num / den);
}
float DielTransPdf(float ior, float r, V3 wo, V3 wi, V3 n) {
    if (SameHemi(wo, wi, n)) return 0.0f;
    const bool ent = Dot(n, wo) >= 0.0f;
    const V3 on = ent ? n : Neg(n);
    const float ei = ent ? 1.0f : std::max(1.0f, ior), et = ent ? std::max(1.0f, ior) : 1.0f, eta = et / ei;
    const V3 h = TransHalf(wo, wi, on, ior);
    if (LengthSq(h) <= 0.0f || Dot(h, on) <= 0.0f) return 0.0f;
    const float ch = AbsDot(on, h), woh = Dot(wo, h), wih = Dot(wi, h), sd = woh + eta * wih;
    if (ch <= 0.0f || woh <= 0.0f || wih >= 0.0f || sd == 0.0f) return 0.0f;
    const float f = Fresnel(std::fabs(woh), ei, et);
    const float dwh = std::fabs((eta * eta * wih) / (sd * sd));
    return (1.0f - f) * GgxD(ch, Alpha(r)) * ch * dwh;
}

V3 Refl(const yrh_material& m) { return Load3(m.reflectance); }

V3 EvalBsdf(const yrh_material& m, V3 wo, V3 wi, V3 n) {
    switch (m.kind) {
        case YRH_MATERIAL_DIFFUSE:
            if (!Above(wo, n) || !Above(wi, n)) return V3{0, 0, 0};
            return Div(Refl(m), Pi);
        case YRH_MATERIAL_CONDUCTOR:
            if (m.uroughness <= DeltaRoughness && m.vroughness <= DeltaRoughness) return V3{0, 0, 0};
            return ConductorF(Refl(m), m.uroughness, wo, wi, n);
        case YRH_MATERIAL_THIN_DIELECTRIC: {
            if (m.uroughness <= DeltaRoughness) return V3{0, 0, 0};
            const V3 fw = Normalize(Neg(wo));
            if (std::max(0.0f, Dot(fw, wi)) <= 0.0f) return V3{0, 0, 0};
            const float f = Fresnel(std::fabs(Dot(n, wo)), 1.0f, std::max(1.0f, m.ior));
            return Mul(Div(Refl(m), Pi), 1.0f - f);
        }
        case YRH_MATERIAL_DIELECTRIC:
            if (m.uroughness <= DeltaRoughness) return V3{0, 0, 0};
            return SameHemi(wo, wi, n) ? DielReflF(Refl(m), m.ior, m.uroughness, wo, wi, n)
                                       : DielTransF(Refl(m), m.ior, m.uroughness, wo, wi, n);
    }
    return V3{0, 0, 0};
}

float PdfBsdf(const yrh_material& m, V3 wo, V3 wi, V3 n) {
    switch (m.kind) {
        case YRH_MATERIAL_DIFFUSE:
            if (!Above(wo, n) || !Above(wi, n)) return 0.0f;
            return std::max(0.0f, Dot(n, wi)) / Pi;
        case YRH_MATERIAL_CONDUCTOR:
            if (m.uroughness <= DeltaRoughness && m.vroughness <= DeltaRoughness) return 0.0f;
            return ConductorPdf(m.uroughness, wo, wi, n);
        case YRH_MATERIAL_THIN_DIELECTRIC:
            if (m.uroughness <= DeltaRoughness) return 0.0f;
            return std::max(0.0f, Dot(Normalize(Neg(wo)), wi)) / Pi;
        case YRH_MATERIAL_DIELECTRIC:
            if (m.uroughness <= DeltaRoughness) return 0.0f;
            return SameHemi(wo, wi, n) ? DielReflPdf(m.ior, m.uroughness, wo, wi, n)
                                       : DielTransPdf(m.ior, m.uroughness, wo, wi, n);
    }
    return 0.0f;
}

struct BS {
    V3 wi{0, 0, 0}, weight{0, 0, 0};
    float pdf = 0.0f;
    bool valid = false, specular = false;
};

BS MakeBS(V3 wi, V3 w, float pdf, bool spec) { return BS{wi, w, pdf, true, spec}; }

BS SampleConductor(V3 wo, V3 n, float sx, float sy, V3 f0, float r) {
    if (!Above(wo, n) || NearBlack(f0)) return BS{};
    const V3 h = HalfVector(n, r, sx, sy);
    const V3 wi = Reflect(Neg(wo), h);
    if (!Above(wi, n)) return BS{};
    const float pdf = ConductorPdf(r, wo, wi, n);
    if (pdf <= 0.0f) return BS{};
    return MakeBS(wi, Mul(ConductorF(f0, r, wo, wi, n), std::max(0.0f, Dot(n, wi)) / pdf), pdf, false);
}

BS SampleBsdf(const yrh_material& m, V3 wo, V3 n, float sx, float sy) {
    switch (m.kind) {
        case YRH_MATERIAL_DIFFUSE: {
            if (!Above(wo, n) || NearBlack(Refl(m))) return BS{};
            const V3 wi = CosineHemisphere(n, sx, sy);
            return MakeBS(wi, Refl(m), PdfBsdf(m, wo, wi, n), false);
        }
        case YRH_MATERIAL_CONDUCTOR:
            if (m.uroughness <= DeltaRoughness && m.vroughness <= DeltaRoughness) {
                if (!Above(wo, n) || NearBlack(Refl(m))) return BS{};
                return MakeBS(Reflect(Neg(wo), n), Refl(m), 1.0f, true);
            }
            return SampleConductor(wo, n, sx, sy, Refl(m), m.uroughness);
        case YRH_MATERIAL_THIN_DIELECTRIC: {
            if (NearBlack(Refl(m))) return BS{};
            const V3 on = Dot(n, wo) >= 0.0f ? n : Neg(n);
            const float f = Fresnel(std::fabs(Dot(on, wo)), 1.0f, std::max(1.0f, m.ior));
            if (m.uroughness <= DeltaRoughness) {
                if (sx < f) return MakeBS(Reflect(Neg(wo), on), Refl(m), 1.0f, true);
                return MakeBS(Normalize(Neg(wo)), Refl(m), 1.0f, true);
            }
            if (sx < f) {
                BS r = SampleConductor(wo, on, f > 0.0f ? sx / f : sx, sy, Mul(Refl(m), f), m.uroughness);
                if (r.valid) r.specular = false;
                return r;
            }
            const float rx = (1.0f - f) > 0.0f ? (sx - f) / (1.0f - f) : sx;
            const V3 fw = Normalize(Neg(wo));
            const V3 wi = CosineHemisphere(fw, rx, sy);
            const float pdf = std::max(0.0f, Dot(fw, wi)) / Pi;
            if (pdf <= 0.0f) return BS{};
            return MakeBS(wi, Mul(Refl(m), 1.0f - f), pdf, false);
        }
        case YRH_MATERIAL_DIELECTRIC: {
            if (NearBlack(Refl(m))) return BS{};
            const bool ent = Dot(n, wo) >= 0.0f;
            const V3 fn = ent ? n : Neg(n);
            const float ei = ent ? 1.0f : std::max(1.0f, m.ior), et = ent ? std::max(1.0f, m.ior) : 1.0f;
            const float eta = ei / et, co = std::max(0.0f, Dot(fn, wo));
            if (m.uroughness <= DeltaRoughness) {
                V3 refr;
                const bool can = Refract(wo, fn, eta, refr);
                const float f = can ? Fresnel(co, ei, et) : 1.0f;
                if (!can || sx < f) return MakeBS(Reflect(Neg(wo), fn), Refl(m), 1.0f, true);
                return MakeBS(refr, Refl(m), 1.0f, true);
            }
            float rx = sx;
            V3 h = HalfVector(fn, m.uroughness, rx, sy);
            if (Dot(h, wo) < 0.0f) h = Neg(h);
            const float f = Fresnel(std::fabs(Dot(wo, h)), ei, et);
            if (sx < f) {
                rx = f > 0.0f ? sx / f : sx;
                h = HalfVector(fn, m.uroughness, rx, sy);
                if (Dot(h, wo) < 0.0f) h = Neg(h);
                const V3 wi = Reflect(Neg(wo), h);
                const float pdf = PdfBsdf(m, wo, wi, n);
                if (pdf <= 0.0f) return BS{};
                return MakeBS(wi, Mul(EvalBsdf(m, wo, wi, n), AbsDot(n, wi) / pdf), pdf, false);
            }
            rx = (1.0f - f) > 0.0f ? (sx - f) / (1.0f - f) : sx;
            h = HalfVector(fn, m.uroughness, rx, sy);
            if (Dot(h, wo) < 0.0f) h = Neg(h);
            V3 wi;
            if (!Refract(wo, h, eta, wi)) return MakeBS(Reflect(Neg(wo), h), Refl(m), 1.0f, true);
            const float pdf = PdfBsdf(m, wo, wi, n);
            if (pdf <= 0.0f) return BS{};
            return MakeBS(wi, Mul(EvalBsdf(m, wo, wi, n), AbsDot(n, wi) / pdf), pdf, false);
        }
    }
    return BS{};
}

bool IsDelta(const yrh_material& m) {
    if (m.kind == YRH_MATERIAL_CONDUCTOR || m.kind == YRH_MATERIAL_DIELECTRIC || m.kind == YRH_MATERIAL_THIN_DIELECTRIC) {
        return m.uroughness == 0.0f && m.vroughness == 0.0f;
    }
    return false;
}

float PowerHeuristic(int na, float pa, int nb, float pb) {
    if (!(na > 0 && pa > 0.0f && std::isfinite(pa))) return 0.0f;
    if (!(nb > 0 && pb > 0.0f && std::isfinite(pb))) return 1.0f;
    const float a = static_cast<float>(na) * pa, b = static_cast<float>(nb) * pb;
    const float a2 = a * a, b2 = b * b, den = a2 + b2;
    if (!std::isfinite(den) || den <= 0.0f) return 0.0f;
    return a2 / den;
}

// ---- lights -----------------------------------------------------------------------

float ClampUnitSample(float s) { return std::isfinite(s) ? std::clamp(s, 0.0f, 1.0f) : 0.0f; }
float TriArea(V3 a, V3 b, V3 c) {
    const float area = Length(Cross(Sub(b, a), Sub(c, a))) * 0.5f;
    return std::isfinite(area) && area > 0.0f ? area : 0.0f;
}
float EmissiveArea(const yrh_emissive& e) { return std::isfinite(e.area) && e.area > 0.0f ? e.area : 0.0f; }

struct LS {
    V3 point, normal, radiance;
    float pdf = 0.0f;
    int index = -1;
    bool valid = false;
};

LS SampleEmissive(const Scene& s, float select, float tx, float ty) {
    LS out;
    const yrh_shading_desc& sd = *s.sd;
    const int n = sd.emissive_count;
    if (n == 0) return out;
    // SampleAliasTable (light_sampling.cpp:185-204)
    const float scaled = std::min(ClampUnitSample(select), std::nextafter(1.0f, 0.0f)) * static_cast<float>(n);
    const int column = std::min(static_cast<int>(scaled), n - 1);
    const float fraction = scaled - static_cast<float>(column);
    const float prob = sd.alias_probability[column];
    int index = column;
    float remapped;
    if (fraction < prob) {
        remapped = prob > 0.0f ? fraction / prob : 0.0f;
    } else {
        index = sd.alias_index[column];
        remapped = prob < 1.0f ? (fraction - prob) / (1.0f - prob) : 0.0f;
    }
    remapped = std::clamp(remapped, 0.0f, std::nextafter(1.0f, 0.0f));
    const float pmf = sd.alias_pmf[index];
    if (index < 0 || pmf <= 0.0f) return out;
    const yrh_emissive& ep = sd.emissives[index];
    const float parea = EmissiveArea(ep);
    if (parea <= 0.0f) return out;
    if (ep.primitive_index < 0 || ep.primitive_index >= s.g.np) return out;
    const yrh_primitive& prim = s.g.prim[ep.primitive_index];
    const int tris = static_cast<int>(prim.index_count / 3);
    if (tris == 0) return out;
    auto tri_pos = [&](int t, V3& p0, V3& p1, V3& p2) {
        const uint32_t base = prim.first_index + static_cast<uint32_t>(t) * 3;
        p0 = VertexPos(s, s.g.idx[base]);
        p1 = VertexPos(s, s.g.idx[base + 1]);
        p2 = VertexPos(s, s.g.idx[base + 2]);
    };
    float total = 0.0f;
    for (int t = 0; t < tris; ++t) {
        V3 a, b, c;
        tri_pos(t, a, b, c);
        total += TriArea(a, b, c);
    }
    if (!std::isfinite(total) || total <= 0.0f) return out;
    const float target = ClampUnitSample(remapped) * total;
    float cum = 0.0f;
    int chosen = -1, last = -1;
    for (int t = 0; t < tris; ++t) {
        V3 a, b, c;
        tri_pos(t, a, b, c);
        const float area = TriArea(a, b, c);
        if (area <= 0.0f) continue;
        cum += area;
        last = t;
        if (target < cum) {
            chosen = t;
            break;
        }
    }
    if (chosen < 0) chosen = last;
    if (chosen < 0) return out;
    V3 p0, p1, p2;
    tri_pos(chosen, p0, p1, p2);
    const float su = std::sqrt(tx);
    const float u = 1.0f - su;
    const float v = ty * su;
    const V3 point = Add(Add(Mul(p0, 1.0f - u - v), Mul(p1, u)), Mul(p2, v));
    const V3 normal = Normalize(Cross(Sub(p1, p0), Sub(p2, p0)));
    if (TriArea(p0, p1, p2) <= 0.0f || LengthSq(normal) == 0.0f) return out;
    out.point = point;
    out.normal = normal;
    out.radiance = Load3(ep.radiance);
    out.pdf = 1.0f / parea;
    out.pdf *= pmf;
    out.index = index;
    out.valid = true;
    return out;
}

float EmissivePdf(const Scene& s, V3 shading, V3 light, V3 ln, int index) {
    const yrh_shading_desc& sd = *s.sd;
    const int n = sd.emissive_count;
    if (n == 0) return 0.0f;
    float area_pdf;
    if (index >= 0 && index < n) {
        const float area = EmissiveArea(sd.emissives[index]);
        if (area <= 0.0f) return 0.0f;
        area_pdf = sd.alias_pmf[index] / area;
    } else {
        float total = 0.0f;
        for (int k = 0; k < n; ++k) total += EmissiveArea(sd.emissives[k]);
        if (!std::isfinite(total) || total <= 0.0f) return 0.0f;
        area_pdf = 1.0f / total;
    }
    const V3 dir = Sub(light, shading);
    const float d2 = LengthSq(dir);
    if (d2 <= 0.0f) return 0.0f;
    const float c = std::fabs(Dot(ln, Normalize(Sub(shading, light))));
    if (c <= 0.0f) return 0.0f;
    return area_pdf * d2 / c;
}

// environment
float Fract(float v) { return v - std::floor(v); }
void DirToUv(V3 d, float rot, float& u, float& v) {
    d = Normalize(d);
    const float phi = std::atan2(d.x, d.z);
    u = Fract(0.5f + (phi + rot) / TwoPi);
    const float th = std::acos(std::clamp(d.y, -1.0f, 1.0f));
    v = std::clamp(th / Pi, 0.0f, 1.0f);
}
V3 UvToDir(float u, float v, float rot) {
    const float phi = (u - 0.5f) * TwoPi - rot;
    const float th = std::clamp(v, 0.0f, 1.0f) * Pi;
    const float st = std::sin(th);
    return Normalize(V3{st * std::sin(phi), std::cos(th), st * std::cos(phi)});
}
V3 Texel(const yrh_environment_desc& e, int x, int y) {
    const float* t = e.texels + 4 * (static_cast<size_t>(ortex::WrapIndex(y, e.height, e.wrap_t)) *
                                         static_cast<size_t>(e.width) +
                                     static_cast<size_t>(ortex::WrapIndex(x, e.width, e.wrap_s)));
    return V3{t[0], t[1], t[2]};
}
// SampleTexture without a footprint on the map (texture.cpp:107-130, 312-322).
V3 Bilinear(const yrh_environment_desc& e, float u, float v) {
    if (e.width <= 0 || e.height <= 0 || e.texels == nullptr) return V3{0, 0, 0};
    if (e.nearest) {
        const int x = std::clamp(static_cast<int>(std::floor(ortex::WrapCoord(u, e.wrap_s) * static_cast<float>(e.width))),
                                 0, e.width - 1);
        const int y = std::clamp(static_cast<int>(std::floor(ortex::WrapCoord(v, e.wrap_t) * static_cast<float>(e.height))),
                                 0, e.height - 1);
        return Texel(e, x, y);
    }
    if (e.width == 1 && e.height == 1) return Load3(e.texels);
    const float s = ortex::WrapCoord(u, e.wrap_s) * static_cast<float>(e.width) - 0.5f;
    const float t = ortex::WrapCoord(v, e.wrap_t) * static_cast<float>(e.height) - 0.5f;
    const int s0 = static_cast<int>(std::floor(s)), t0 = static_cast<int>(std::floor(t));
    const float ds = s - static_cast<float>(s0), dt = t - static_cast<float>(t0);
    const V3 a = Add(Mul(Texel(e, s0, t0), 1.0f - ds), Mul(Texel(e, s0 + 1, t0), ds));
    const V3 b = Add(Mul(Texel(e, s0, t0 + 1), 1.0f - ds), Mul(Texel(e, s0 + 1, t0 + 1), ds));
    return Add(Mul(a, 1.0f - dt), Mul(b, dt));
}
V3 EnvColor(const Scene& s, V3 d) {
    const yrh_environment_desc& e = s.sd->environment;
    if (!e.active) return Mul(Load3(e.radiance), e.strength);
    if (e.strength <= 0.0f || e.texels == nullptr || e.width <= 0 || e.height <= 0) return V3{0, 0, 0};
    float u, v;
    DirToUv(d, e.rotation_radians, u, v);
    return Mul(Had(Bilinear(e, u, v), Load3(e.radiance)), e.strength);
}
bool EnvSampleable(const Scene& s) {
    const yrh_environment_desc& e = s.sd->environment;
    return e.active && e.strength > 0.0f && e.texels != nullptr && e.width > 0 && e.height > 0 && e.has_distribution;
}
float EnvPdf(const Scene& s, V3 d) {
    if (!EnvSampleable(s)) return 0.0f;
    const yrh_environment_desc& e = s.sd->environment;
    const int w = e.dist_width, h = e.dist_height;
    if (w <= 0 || h <= 0 || e.total_weight <= 0.0f) return 0.0f;
    float u, v;
    DirToUv(d, e.rotation_radians, u, v);
    const int x = std::clamp(static_cast<int>(std::floor(u * static_cast<float>(w))), 0, w - 1);
    const int y = std::clamp(static_cast<int>(std::floor(v * static_cast<float>(h))), 0, h - 1);
    const float prob = e.texel_weights[static_cast<size_t>(y) * static_cast<size_t>(w) + static_cast<size_t>(x)] /
                       e.total_weight;
    const float st = std::max(1.0e-6f, std::sin((static_cast<float>(y) + 0.5f) * Pi / static_cast<float>(h)));
    const float sa = (TwoPi / static_cast<float>(w)) * (Pi / static_cast<float>(h)) * st;
    return prob / sa;
}
struct ES {
    V3 dir, radiance;
    float pdf = 0.0f;
    bool valid = false;
};
ES SampleEnv(const Scene& s, float sx, float sy) {
    ES out;
    if (!EnvSampleable(s)) return out;
    const yrh_environment_desc& e = s.sd->environment;
    const int w = e.dist_width, h = e.dist_height;
    if (w <= 0 || h <= 0 || e.total_weight <= 0.0f) return out;
    const float rt = ClampUnit(sy) * e.total_weight;
    const float* rc = e.row_cdf;
    int y = static_cast<int>(std::lower_bound(rc, rc + h, rt) - rc);
    if (y == h) y = h - 1;
    const float rp = y == 0 ? 0.0f : rc[y - 1];
    const float rw = std::max(e.row_weights[y], 1.0e-6f);
    const float rf = std::clamp((rt - rp) / rw, 0.0f, 1.0f);
    const float* cc = e.conditional_cdfs + static_cast<size_t>(y) * static_cast<size_t>(w);
    const float ct = ClampUnit(sx) * rw;
    int x = static_cast<int>(std::lower_bound(cc, cc + w, ct) - cc);
    if (x == w) x = w - 1;
    const float cp = x == 0 ? 0.0f : cc[x - 1];
    const float tw = std::max(e.texel_weights[static_cast<size_t>(y) * static_cast<size_t>(w) + static_cast<size_t>(x)], 1.0e-6f);
    const float cf = std::clamp((ct - cp) / tw, 0.0f, 1.0f);
    const V3 d = UvToDir((static_cast<float>(x) + cf) / static_cast<float>(w), (static_cast<float>(y) + rf) / static_cast<float>(h),
                         e.rotation_radians);
    const float pdf = EnvPdf(s, d);
    if (pdf <= 0.0f) return out;
    out.dir = d;
    out.radiance = EnvColor(s, d);
    out.pdf = pdf;
    out.valid = true;
    return out;
}

// ---- integrator -----------------------------------------------------------------

float MaxAbs(V3 p) { return std::max(std::fabs(p.x), std::max(std::fabs(p.y), std::fabs(p.z))); }
float SurfaceBias(V3 p) { return std::max(1.0e-4f, MaxAbs(p) * 1.0e-5f); }
float ShadowBias(V3 o, V3 t, float d) {
    const float sc = std::max(MaxAbs(o), MaxAbs(t));
    return std::min(std::max(1.0e-4f, sc * 1.0e-5f), d * 1.0e-2f);
}
V3 BeerLambert(V3 col, float dist_scale, float d) {
    if (d <= 0.0f || dist_scale <= 0.0f) return V3{1, 1, 1};
    const float inv = 1.0f / dist_scale;
    auto ch = [&](float c) { return std::exp(std::log(std::clamp(c, 1.0e-6f, 1.0f)) * d * inv); };
    return V3{ch(col.x), ch(col.y), ch(col.z)};
}
bool TransBlack(V3 c) { return MaxComponent(c) <= 1.0e-6f; }

struct Medium {
    bool active = false;
    V3 color{1, 1, 1};
    float distance = 1.0f;
};

struct Stats {
    uint64_t rays = 0, shadow = 0, occluded = 0, tris = 0, spheres = 0, nodes = 0, hits = 0, misses = 0;
};

void AddTrace(Stats& st, const orc::Counters& c) {
    st.nodes += c.nodes;
    st.tris += c.tris;
    st.spheres += c.spheres;
}

bool ShadowVisible(const Scene& s, V3 o, V3 d, float max_dist, Stats& st, V3& T) {
    T = V3{1, 1, 1};
    Medium med;
    float remaining = max_dist;
    for (int k = 0; k < 16; ++k) {
        orc::Counters c;
        const Surface sf = TraceSurface(s, o, d, 1.0e-5f, remaining, c);
        AddTrace(st, c);
        const bool finite = std::isfinite(remaining);
        if (sf.exhausted) return false;
        if (!sf.hit) {
            if (med.active) {
                if (!finite) return false;
                T = Had(T, BeerLambert(med.color, med.distance, remaining));
            }
            return !TransBlack(T);
        }
        if (med.active) {
            T = Had(T, BeerLambert(med.color, med.distance, sf.h.t));
            if (TransBlack(T)) return false;
        }
        const V3 hp = Add(o, Mul(d, sf.h.t));
        const yrh_material& m = sf.m;
        if (m.kind != YRH_MATERIAL_DIELECTRIC && m.kind != YRH_MATERIAL_THIN_DIELECTRIC) return false;
        if (m.kind == YRH_MATERIAL_THIN_DIELECTRIC) {
            T = Had(T, V3{std::clamp(m.reflectance[0], 0.0f, 1.0f), std::clamp(m.reflectance[1], 0.0f, 1.0f),
                          std::clamp(m.reflectance[2], 0.0f, 1.0f)});
            if (TransBlack(T)) return false;
        } else if (med.active) {
            med = Medium{};
        } else {
            med = Medium{true, Load3(m.absorption_color), m.absorption_distance};
        }
        const float bias = SurfaceBias(hp);
        if (finite) {
            remaining -= sf.h.t + bias;
            if (remaining <= 0.0f) return !TransBlack(T);
        }
        o = Add(hp, Mul(d, bias));
    }
    return false;
}

// ---- analytic lights (src/lighting/analytic_light_sampling.cpp:9-140) -------------

struct AS {
    bool valid = false;
    V3 wi{0, 0, 0}, radiance{0, 0, 0};
    float distance = 0.0f, selection_pdf = 0.0f;
    int light_index = -1;
};

// DistanceSquaredToBounds / TreeImportance (:9-18).
float TreeImportance(const yrh_light_tree_node& nd, V3 p) {
    const float dx = std::max({nd.bounds_min[0] - p.x, 0.0f, p.x - nd.bounds_max[0]});
    const float dy = std::max({nd.bounds_min[1] - p.y, 0.0f, p.y - nd.bounds_max[1]});
    const float dz = std::max({nd.bounds_min[2] - p.z, 0.0f, p.z - nd.bounds_max[2]});
    return nd.power / std::max(1.0e-4f, dx * dx + dy * dy + dz * dz);
}

// SampleAnalyticLight (:80-140) with SampleAnalyticPoint / Distant / Spot (:22-78).
AS SampleAnalytic(const Scene& s, V3 p, float select) {
    const yrh_shading_desc& sd = *s.sd;
    AS out;
    const int n = sd.analytic_light_count;
    if (n <= 0) return out;
    int light = -1;
    float pdf = 1.0f;
    float u = std::clamp(select, 0.0f, std::nextafter(1.0f, 0.0f));
    if (sd.use_analytic_tree && sd.analytic_tree_root >= 0 && sd.analytic_tree_root < sd.analytic_tree_count) {
        int node = sd.analytic_tree_root;
        while (node >= 0) {
            const yrh_light_tree_node& nd = sd.analytic_tree[node];
            if (nd.light_index >= 0) {
                light = nd.light_index;
                break;
            }
            if (nd.left < 0 || nd.right < 0) return out;
            const float lw = TreeImportance(sd.analytic_tree[nd.left], p);
            const float rw = TreeImportance(sd.analytic_tree[nd.right], p);
            const float total = lw + rw;
            const float lp = total > 0.0f ? lw / total : 0.5f;
            if (u < lp) {
                pdf *= lp;
                u = lp > 0.0f ? u / lp : 0.0f;
                node = nd.left;
            } else {
                const float rp = 1.0f - lp;
                pdf *= rp;
                u = rp > 0.0f ? (u - lp) / rp : 0.0f;
                node = nd.right;
            }
        }
    } else if (sd.analytic_alias_pmf != nullptr) {
        // SampleAliasTable (src/lighting/light_sampling.cpp:184-204).
        const float scaled = std::min(ClampUnitSample(u), std::nextafter(1.0f, 0.0f)) * static_cast<float>(n);
        const int column = std::min(static_cast<int>(scaled), n - 1);
        const float fraction = scaled - static_cast<float>(column);
        light = fraction < sd.analytic_alias_probability[column] ? column : sd.analytic_alias_index[column];
        pdf = sd.analytic_alias_pmf[light];
    } else {
        light = std::min(static_cast<int>(u * static_cast<float>(n)), n - 1);
        pdf = 1.0f / static_cast<float>(n);
    }
    if (light < 0 || light >= n || pdf <= 0.0f) return out;
    const yrh_analytic_light& l = sd.analytic_lights[light];
    const V3 intensity = Load3(l.intensity);
    if (l.kind == YRH_LIGHT_POINT) {
        const V3 delta = Sub(Load3(l.position), p);
        const float d2 = Dot(delta, delta);
        if (d2 <= 0.0f) return out;
        out.distance = std::sqrt(d2);
        out.wi = Normalize(delta);
        out.radiance = Div(intensity, d2);
    } else if (l.kind == YRH_LIGHT_DISTANT) {
        const V3 to_light = Neg(Load3(l.direction));
        if (LengthSq(to_light) == 0.0f) return out;
        out.wi = Normalize(to_light);
        out.distance = 1.0e6f;
        out.radiance = intensity;
    } else {
        const V3 delta = Sub(Load3(l.position), p);
        const float d2 = Dot(delta, delta);
        if (d2 <= 0.0f) return out;
        const float dist = std::sqrt(d2);
        const V3 wi = Div(delta, dist);
        const V3 dir = Load3(l.direction);
        const V3 axis = LengthSq(dir) > 0.0f ? Normalize(dir) : V3{0.0f, 0.0f, 1.0f};
        const float cosine = -Dot(wi, axis);
        float falloff = 0.0f;
        if (cosine >= l.cone_cos_inner) {
            falloff = 1.0f;
        } else if (cosine > l.cone_angle) {
            const float t = (cosine - l.cone_angle) / (l.cone_cos_inner - l.cone_angle);
            falloff = t * t * (3.0f - 2.0f * t);
        }
        out.wi = wi;
        out.distance = dist;
        out.radiance = Mul(intensity, falloff / d2);
    }
    out.selection_pdf = pdf;
    out.light_index = light;
    out.valid = true;
    return out;
}

V3 DirectLight(const Scene& s, const yrh_material& m, V3 hp, V3 n, V3 wo, const Sampler& smp, Stats& st) {
    V3 rad{0, 0, 0};
    if (s.sd->emissive_count > 0) {
        float sx, sy;
        smp.Get2(DimLightSurface, sx, sy);
        const LS ls = SampleEmissive(s, smp.Get1(DimLightSelect), sx, sy);
        if (ls.valid && ls.pdf > 0.0f && !NearBlack(ls.radiance)) {
            const V3 to = Sub(ls.point, hp);
            const float d2 = LengthSq(to);
            if (d2 > 1.0e-4f * 1.0e-4f) {
                const float dist = std::sqrt(d2);
                const float bias = ShadowBias(hp, ls.point, dist);
                if (dist > bias) {
                    const V3 so = Add(hp, Mul(n, bias));
                    const V3 stl = Sub(ls.point, so);
                    const float sd2 = LengthSq(stl);
                    if (sd2 > 1.0e-4f * 1.0e-4f) {
                        const float sd = std::sqrt(sd2);
                        const V3 wi = Div(stl, sd);
                        const float c = std::max(0.0f, Dot(n, wi));
                        if (c > 0.0f) {
                            const float pl = EmissivePdf(s, hp, ls.point, ls.normal, ls.index);
                            if (pl > 0.0f) {
                                ++st.shadow;
                                V3 T;
                                if (!ShadowVisible(s, so, wi, sd - bias, st, T)) {
                                    ++st.occluded;
                                } else {
                                    const V3 f = EvalBsdf(m, wo, wi, n);
                                    if (!NearBlack(f)) {
                                        const float pb = PdfBsdf(m, wo, wi, n);
                                        const float w = PowerHeuristic(1, pl, 1, pb);
                                        rad = Add(rad, Mul(Had(f, Had(T, ls.radiance)), c * w / pl));
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
        rad = Mul(rad, 1.0f);
    }
    // Analytic lights: one delta sample, no MIS (path_direct_lighting.cpp:250-281).
    if (s.sd->analytic_light_count > 0) {
        const AS as = SampleAnalytic(s, hp, smp.Get1(DimAnalyticSelect));
        if (as.valid && as.selection_pdf > 0.0f && !NearBlack(as.radiance)) {
            const float c = std::max(0.0f, Dot(n, as.wi));
            if (c > 0.0f) {
                const V3 f = EvalBsdf(m, wo, as.wi, n);
                if (!NearBlack(f)) {
                    const V3 so = Add(hp, Mul(n, SurfaceBias(hp)));
                    ++st.shadow;
                    V3 T;
                    if (!ShadowVisible(s, so, as.wi, as.distance - SurfaceBias(hp), st, T)) {
                        ++st.occluded;
                    } else {
                        rad = Add(rad, Mul(Had(f, Had(T, as.radiance)), c / as.selection_pdf));
                    }
                }
            }
        }
    }
    if (EnvSampleable(s)) {
        V3 er{0, 0, 0};
        float sx, sy;
        smp.Get2(DimLightSurface, sx, sy);
        const ES es = SampleEnv(s, sx, sy);
        if (es.valid && es.pdf > 0.0f && !NearBlack(es.radiance)) {
            const V3 wi = es.dir;
            const float c = std::max(0.0f, Dot(n, wi));
            if (c > 0.0f) {
                const V3 f = EvalBsdf(m, wo, wi, n);
                if (!NearBlack(f)) {
                    const V3 so = Add(hp, Mul(n, SurfaceBias(hp)));
                    ++st.shadow;
                    V3 T;
                    if (!ShadowVisible(s, so, wi, kInf, st, T)) {
                        ++st.occluded;
                    } else {
                        const float pb = PdfBsdf(m, wo, wi, n);
                        const float w = PowerHeuristic(1, es.pdf, 1, pb);
                        er = Add(er, Mul(Had(f, Had(T, es.radiance)), c * w / es.pdf));
                    }
                }
            }
        }
        rad = Add(rad, Mul(er, 1.0f));
    }
    return rad;
}

struct PathOut {
    V3 beauty{0, 0, 0}, albedo{0, 0, 0}, normal{0, 0, 0};
    float depth = 0.0f;
    bool primary = false;
};

PathOut Integrate(const Scene& s, const yrh_render_desc& rd, V3 o, V3 d, const Differentials& camera_diff,
                  Sampler& smp, Stats& st) {
    PathOut out;
    V3 rad{0, 0, 0}, tp{1, 1, 1};
    bool prev_valid = false, prev_delta = false;
    V3 prev_origin{0, 0, 0};
    float prev_pdf = 0.0f;
    Medium med;
    const int max_depth = std::max(1, rd.max_depth);
    for (int depth = 0; depth < max_depth; ++depth) {
        smp.bounce = depth;
        ++st.rays;
        orc::Counters c;
        const Surface sf = TraceSurface(s, o, d, 1.0e-5f, kInf, c, depth == 0 ? &camera_diff : nullptr);
        AddTrace(st, c);
        if (!sf.hit) {  // a miss, or 64 alpha-skipped hits (exhausted)
            ++st.misses;
            float w = 1.0f;
            if (prev_valid && !prev_delta) w = PowerHeuristic(1, prev_pdf, 1, EnvPdf(s, d));
            rad = Add(rad, Mul(Had(tp, EnvColor(s, d)), w));
            break;
        }
        ++st.hits;
        if (med.active) tp = Had(tp, BeerLambert(med.color, med.distance, sf.h.t));
        if (NearBlack(tp)) break;
        const V3 hp = Add(o, Mul(d, sf.h.t));
        const V3 wo = Neg(d);
        const yrh_material& m = sf.m;
        const V3 n = sf.shading;
        if (depth == 0) {
            out.primary = true;
            out.albedo = Refl(m);
            out.normal = n;
            out.depth = sf.h.t;
        }
        // Sphere emission is not supported by the reference (zero contribution,
        // src/integrators/path_integrator.cpp:83-92).
        if (!NearBlack(Load3(m.emission)) && sf.sphere < 0) {
            const V3 ln = GeometricNormal(s, sf.tri);
            float w = 1.0f;
            if (prev_valid && !prev_delta) {
                int ei = -1;
                for (int k = 0; k < s.sd->emissive_count; ++k) {
                    if (s.sd->emissives[k].primitive_index == sf.prim) {
                        ei = k;
                        break;
                    }
                }
                w = PowerHeuristic(1, prev_pdf, 1, EmissivePdf(s, prev_origin, hp, ln, ei));
            }
            rad = Add(rad, Mul(Had(tp, Load3(m.emission)), w));
        }
        if (!IsDelta(m)) rad = Add(rad, Had(tp, DirectLight(s, m, hp, n, wo, smp, st)));
        if (depth + 1 >= max_depth) break;
        float bx, by;
        smp.Get2(DimBsdf, bx, by);
        const BS bs = SampleBsdf(m, wo, n, bx, by);
        if (!bs.valid || NearBlack(bs.weight)) break;
        tp = Had(tp, bs.weight);
        if (depth >= 3) {
            const float survival = std::clamp(MaxComponent(tp), 0.05f, 0.95f);
            if (smp.Get1(DimRoulette) >= survival) break;
            tp = Div(tp, survival);
        }
        if (m.kind == YRH_MATERIAL_DIELECTRIC && Dot(bs.wi, n) < 0.0f) {
            med = med.active ? Medium{} : Medium{true, Load3(m.absorption_color), m.absorption_distance};
        }
        prev_valid = true;
        prev_delta = bs.specular;
        prev_origin = hp;
        prev_pdf = bs.pdf;
        const V3 bn = Dot(bs.wi, n) >= 0.0f ? n : Neg(n);
        o = Add(hp, Mul(bn, SurfaceBias(hp)));
        d = bs.wi;
    }
    if (rd.radiance_clamp > 0.0f) {
        const float mx = MaxComponent(rad);
        if (!(mx <= rd.radiance_clamp || mx <= 0.0f)) rad = Mul(rad, rd.radiance_clamp / mx);
    }
    out.beauty = rad;
    return out;
}

struct FilmPixel32 {
    float sum[3];
    uint32_t samples;
    float mean, m2;
    float albedo[3], normal[3];
    float depth;
    uint32_t aov;
};
static_assert(sizeof(FilmPixel32) == 56);

void AddSample(FilmPixel32& p, const PathOut& s, uint64_t& bad) {
    if (!Finite(s.beauty)) {
        ++bad;
        return;
    }
    p.sum[0] = p.sum[0] + s.beauty.x;
    p.sum[1] = p.sum[1] + s.beauty.y;
    p.sum[2] = p.sum[2] + s.beauty.z;
    ++p.samples;
    const float lum = 0.2126f * s.beauty.x + 0.7152f * s.beauty.y + 0.0722f * s.beauty.z;
    const float delta = lum - p.mean;
    p.mean += delta / static_cast<float>(p.samples);
    p.m2 += delta * (lum - p.mean);
    if (s.primary && Finite(s.albedo) && Finite(s.normal) && std::isfinite(s.depth) && s.depth >= 0.0f) {
        p.albedo[0] = p.albedo[0] + s.albedo.x;
        p.albedo[1] = p.albedo[1] + s.albedo.y;
        p.albedo[2] = p.albedo[2] + s.albedo.z;
        p.normal[0] = p.normal[0] + s.normal.x;
        p.normal[1] = p.normal[1] + s.normal.y;
        p.normal[2] = p.normal[2] + s.normal.z;
        p.depth += s.depth;
        ++p.aov;
    }
}

}  // namespace

namespace {

Scene MakeScene(const yrh_geometry_desc* gd, const yrh_shading_desc* sd) {
    Scene s;
    s.g = orc::Geometry{gd->vertices, gd->vertex_count, gd->indices, gd->index_count, gd->primitives, gd->primitive_count,
                        gd->spheres, gd->sphere_count};
    s.gd = gd;
    s.sd = sd;
    s.two_level = gd->accel_kind == YRH_ACCEL_TWO_LEVEL;
    s.tl = orc::TwoLevel{gd->blases, gd->blas_count, gd->instances, gd->instance_count, &gd->tlas, gd->tlas_primitives,
                         gd->tlas_primitive_count};
    if (sd != nullptr && sd->textures != nullptr) {
        for (int32_t k = 0; k < sd->texture_count; ++k) s.tex.push_back(ortex::Build(sd->textures[k]));
    }
    return s;
}

// One pass: pixels[i] renders sample index sample_of(i) into out[pixels[i]]
// (camera ray from the Pixel sample, then the path loop), on `threads` workers.
template <class SampleOf>
void RenderPixels(const Scene& s, const yrh_render_desc& rd, const std::vector<int64_t>& pixels, SampleOf sample_of,
                  int threads, std::vector<Stats>& per_thread, std::vector<PathOut>& out) {
    const int W = rd.width, H = rd.height;
    const float w = static_cast<float>(std::max(1, W)), h = static_cast<float>(std::max(1, H));
    const V3 co = Load3(rd.camera_origin), cf = Load3(rd.camera_forward), cr = Load3(rd.camera_right),
             cu = Load3(rd.camera_up);
    const float hh = rd.tan_half_fov;
    const int64_t n = static_cast<int64_t>(pixels.size());
    std::atomic_int64_t next{0};
    auto work = [&](int tid) {
        Stats& st = per_thread[static_cast<size_t>(tid)];
        while (true) {
            const int64_t b = next.fetch_add(256);
            if (b >= n) break;
            const int64_t e = std::min(n, b + 256);
            for (int64_t i = b; i < e; ++i) {
                const int64_t p = pixels[static_cast<size_t>(i)];
                const int x = static_cast<int>(p % W), y = static_cast<int>(p / W);
                Sampler smp(rd.sampler, rd.seed, x, y, sample_of(i), rd.spp);
                float sx, sy;
                smp.At2(DimPixel, 0, sx, sy);
                const float fx = static_cast<float>(x) + sx, fy = static_cast<float>(y) + sy;
                const float aspect = w / h;
                auto direction = [&](float px, float py) {
                    const float scx = (2.0f * px / w - 1.0f) * aspect * hh;
                    const float scy = (1.0f - 2.0f * py / h) * hh;
                    return Normalize(Add(Add(cf, Mul(cr, scx)), Mul(cu, scy)));
                };
                const Differentials diff{direction(fx + 1.0f, fy), direction(fx, fy + 1.0f)};
                out[static_cast<size_t>(p)] = Integrate(s, rd, co, direction(fx, fy), diff, smp, st);
            }
        }
    };
    if (threads <= 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t) pool.emplace_back(work, t);
        for (auto& th : pool) th.join();
    }
}

void AddStats(const std::vector<Stats>& per_thread, uint64_t* stats) {
    for (const Stats& st : per_thread) {
        stats[0] += st.rays;
        stats[1] += st.shadow;
        stats[2] += st.occluded;
        stats[3] += st.tris;
        stats[4] += st.spheres;
        stats[5] += st.nodes;
        stats[6] += st.hits;
        stats[7] += st.misses;
    }
}

// Film::IsConverged (src/film/film.cpp:113-121) with LuminanceVariance and
// LuminanceStandardError (film.cpp:98-111), in FP32.
bool IsConverged(const FilmPixel32& p, int min_spp, float relative_error, float absolute_error, float confidence) {
    if (p.samples < static_cast<uint32_t>(std::max(2, min_spp))) return false;
    const float variance = p.samples > 1 ? p.m2 / static_cast<float>(p.samples - 1) : 0.0f;
    const float stderr_mean = p.samples > 1 ? std::sqrt(std::max(0.0f, variance) / static_cast<float>(p.samples))
                                            : std::numeric_limits<float>::infinity();
    const float bound = std::max(0.0f, confidence) * stderr_mean;
    const float tolerance = std::max(0.0f, absolute_error) + std::max(0.0f, relative_error) * std::fabs(p.mean);
    return std::isfinite(bound) && bound <= tolerance;
}

}  // namespace

extern "C" {
#pragma GCC visibility push(default)

// Renders pass_count passes of `rd` into `film` (width*height FilmPixel
// records of 56 bytes, accumulated in place): RenderCpuPathTrace's pass loop
// without adaptive sampling; a pixel's sample index is sample_begin plus its
// film sample count.
// stats: rays, shadow, occluded, triangle tests, sphere tests, node tests,
// hits, misses, bad samples (uint64 each, accumulated).
int oracle_render(const yrh_geometry_desc* gd, const yrh_shading_desc* sd, const yrh_render_desc* rd,
                  int32_t sample_begin, int32_t pass_count, int32_t threads, void* film, uint64_t* stats) {
    if (gd == nullptr || sd == nullptr || rd == nullptr || film == nullptr || stats == nullptr) return -1;
    const Scene s = MakeScene(gd, sd);
    const int64_t npix = static_cast<int64_t>(rd->width) * rd->height;
    auto* px = static_cast<FilmPixel32*>(film);
    std::vector<PathOut> pass(static_cast<size_t>(npix));
    std::vector<Stats> per_thread(static_cast<size_t>(std::max(1, threads)));
    std::vector<int64_t> all(static_cast<size_t>(npix));
    for (int64_t p = 0; p < npix; ++p) all[static_cast<size_t>(p)] = p;
    // Each pass renders every pixel's next sample: index = sample_begin + its
    // film sample count (cpu_path_tracer.cpp:126), so a pixel whose sample was
    // not finite (not added) renders the same index again.
    for (int pi = 0; pi < pass_count; ++pi) {
        RenderPixels(
            s, *rd, all, [&](int64_t p) { return sample_begin + static_cast<int>(px[p].samples); }, threads,
            per_thread, pass);
        uint64_t bad = 0;
        for (int64_t p = 0; p < npix; ++p) AddSample(px[p], pass[static_cast<size_t>(p)], bad);
        stats[8] += bad;
    }
    AddStats(per_thread, stats);
    return 0;
}

// RenderCpuPathTrace's pass loop with its active pixels
// (src/backends/cpu/cpu_path_tracer.cpp:95-106, 109-199): before each pass the
// active pixels are those below rd->spp samples that are not converged
// (IsConverged, when adaptive[0] != 0); each renders its next sample (index =
// its film sample count); the loop ends after pass rd->spp - 1 or when none is
// active. adaptive: enabled, min_spp, relative_error, absolute_error,
// confidence (as floats). first_pass = resume_completed_spp. passes: passes run;
// converged: converged pixels after the last pass.
int oracle_render_adaptive(const yrh_geometry_desc* gd, const yrh_shading_desc* sd, const yrh_render_desc* rd,
                           const float* adaptive, int32_t first_pass, int32_t threads, void* film, uint64_t* stats,
                           int32_t* passes, uint64_t* converged) {
    if (gd == nullptr || sd == nullptr || rd == nullptr || adaptive == nullptr || film == nullptr || stats == nullptr) {
        return -1;
    }
    const Scene s = MakeScene(gd, sd);
    const int64_t npix = static_cast<int64_t>(rd->width) * rd->height;
    auto* px = static_cast<FilmPixel32*>(film);
    const int spp = std::max(1, rd->spp);
    const bool on = adaptive[0] != 0.0f;
    const int min_spp = static_cast<int>(adaptive[1]);
    auto collect = [&](std::vector<int64_t>& list) {
        list.clear();
        uint64_t conv = 0;
        for (int64_t p = 0; p < npix; ++p) {
            const bool c = on && IsConverged(px[p], min_spp, adaptive[2], adaptive[3], adaptive[4]);
            if (c) ++conv;
            if (!c && px[p].samples < static_cast<uint32_t>(spp)) list.push_back(p);
        }
        return conv;
    };
    std::vector<PathOut> pass(static_cast<size_t>(npix));
    std::vector<Stats> per_thread(static_cast<size_t>(std::max(1, threads)));
    std::vector<int64_t> active;
    uint64_t conv = collect(active);
    int ran = 0;
    for (int pi = first_pass; pi < spp && !active.empty(); ++pi) {
        RenderPixels(s, *rd, active,
                     [&](int64_t i) { return static_cast<int>(px[active[static_cast<size_t>(i)]].samples); },
                     threads, per_thread, pass);
        uint64_t bad = 0;
        for (int64_t p : active) AddSample(px[p], pass[static_cast<size_t>(p)], bad);
        stats[8] += bad;
        conv = collect(active);
        ++ran;
    }
    AddStats(per_thread, stats);
    if (passes != nullptr) *passes = ran;
    if (converged != nullptr) *converged = conv;
    return 0;
}

// Camera rays of one pass (GeneratePerspectiveCameraRay with the Pixel sample).
int oracle_camera_rays(const yrh_render_desc* rd, float* org_tmin, float* dir_tmax) {
    const int W = rd->width, H = rd->height;
    const float w = static_cast<float>(std::max(1, W)), h = static_cast<float>(std::max(1, H));
    const V3 cf = Load3(rd->camera_forward), cr = Load3(rd->camera_right), cu = Load3(rd->camera_up);
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            Sampler smp(rd->sampler, rd->seed, x, y, rd->sample_index, rd->spp);
            float sx, sy;
            smp.At2(DimPixel, 0, sx, sy);
            const float fx = static_cast<float>(x) + sx, fy = static_cast<float>(y) + sy;
            const float aspect = w / h;
            const float scx = (2.0f * fx / w - 1.0f) * aspect * rd->tan_half_fov;
            const float scy = (1.0f - 2.0f * fy / h) * rd->tan_half_fov;
            const V3 d = Normalize(Add(Add(cf, Mul(cr, scx)), Mul(cu, scy)));
            const size_t i = 4 * (static_cast<size_t>(y) * static_cast<size_t>(W) + static_cast<size_t>(x));
            org_tmin[i] = rd->camera_origin[0];
            org_tmin[i + 1] = rd->camera_origin[1];
            org_tmin[i + 2] = rd->camera_origin[2];
            org_tmin[i + 3] = 1.0e-5f;
            dir_tmax[i] = d.x;
            dir_tmax[i + 1] = d.y;
            dir_tmax[i + 2] = d.z;
            dir_tmax[i + 3] = kInf;
        }
    }
    return 0;
}

// BSDF queries: for each i, material mats[mat[i]] and the 12 floats q[12i..]
// = wo(3), wi(3), normal(3), u(2), unused; writes 12 floats out[12i..] =
// EvaluateBsdf(wo, wi, n) (3), PdfBsdf(wo, wi, n), then SampleBsdf(wo, n, u):
// wi (3), weight (3), pdf, and flags = valid + 2 specular + 4 IsDeltaBsdf
// (src/shading/bsdf.cpp:17-329).
int oracle_bsdf_query(const yrh_material* mats, int32_t nmat, int64_t n, const int32_t* mat, const float* q,
                      float* out) {
    for (int64_t i = 0; i < n; ++i) {
        if (mat[i] < 0 || mat[i] >= nmat) return -1;
        const yrh_material& m = mats[mat[i]];
        const float* a = q + 12 * i;
        const V3 wo = Load3(a), wi = Load3(a + 3), nn = Load3(a + 6);
        const V3 f = EvalBsdf(m, wo, wi, nn);
        const float pdf = PdfBsdf(m, wo, wi, nn);
        const BS s = SampleBsdf(m, wo, nn, a[9], a[10]);
        float* o = out + 12 * i;
        const float vals[12] = {f.x, f.y, f.z, pdf, s.wi.x, s.wi.y, s.wi.z, s.weight.x, s.weight.y, s.weight.z, s.pdf,
                                static_cast<float>((s.valid ? 1 : 0) + (s.specular ? 2 : 0) + (IsDelta(m) ? 4 : 0))};
        std::memcpy(o, vals, sizeof(vals));
    }
    return 0;
}

// SampleAnalyticLight at points p (n x 3) with selection samples u (n): per
// row wi (3), distance, radiance (3), selection pdf, light index, valid.
int oracle_sample_analytic(const yrh_shading_desc* sd, int64_t n, const float* p, const float* u, float* out) {
    if (sd == nullptr) return -1;
    Scene s{};
    s.sd = sd;
    for (int64_t i = 0; i < n; ++i) {
        const AS a = SampleAnalytic(s, Load3(p + 3 * i), u[i]);
        const float vals[10] = {a.wi.x,         a.wi.y,        a.wi.z, a.distance, a.radiance.x, a.radiance.y,
                                a.radiance.z,   a.selection_pdf, static_cast<float>(a.light_index),
                                a.valid ? 1.0f : 0.0f};
        std::memcpy(out + 10 * i, vals, sizeof(vals));
    }
    return 0;
}

// SampleTexture4 of texture `t` at n points: uv (n x 2) and footprints
// (n x 4: dudx, dvdx, dudy, dvdy; NULL = none); out n x 4 (RGBA).
int oracle_sample_texture(const yrh_texture* t, int64_t n, const float* uv, const float* footprint, float* out) {
    if (t == nullptr) return -1;
    const ortex::Texture tex = ortex::Build(*t);
    for (int64_t i = 0; i < n; ++i) {
        ortex::Footprint f;
        if (footprint != nullptr) f = ortex::Footprint{footprint[4 * i], footprint[4 * i + 1], footprint[4 * i + 2],
                                                       footprint[4 * i + 3]};
        const ortex::C4 c = ortex::Sample(tex, uv[2 * i], uv[2 * i + 1], f);
        out[4 * i] = c.x;
        out[4 * i + 1] = c.y;
        out[4 * i + 2] = c.z;
        out[4 * i + 3] = c.w;
    }
    return 0;
}

// The mip chain of `t` (levels 1..N after level 0): writes the level count to
// *levels and, when out is not NULL, every level's texels (RGBA) one after
// another (out must hold them all: the caller asks for the count first).
int oracle_texture_mips(const yrh_texture* t, int32_t* levels, int64_t* texels, float* out) {
    if (t == nullptr || levels == nullptr || texels == nullptr) return -1;
    const ortex::Texture tex = ortex::Build(*t);
    *levels = tex.HasData() ? static_cast<int32_t>(tex.levels.size()) : 0;
    int64_t k = 0;
    for (const ortex::Level& l : tex.levels) {
        for (const ortex::C4& c : l.t) {
            if (out != nullptr) {
                out[4 * k] = c.x;
                out[4 * k + 1] = c.y;
                out[4 * k + 2] = c.z;
                out[4 * k + 3] = c.w;
            }
            ++k;
        }
    }
    *texels = k;
    return 0;
}

// Sampler values for the tests (host restatement of the device kernel).
int oracle_sample_values(int32_t kind, uint64_t seed, int32_t spp, int64_t n, const int32_t* x, const int32_t* y,
                         const int32_t* sample, const int32_t* dim, int32_t dims, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        Sampler s(kind, seed, x[i], y[i], sample[i], spp);
        if (dims == 1) {
            out[i] = s.At1(static_cast<uint32_t>(dim[i]), 0);
        } else {
            s.At2(static_cast<uint32_t>(dim[i]), 0, out[2 * i], out[2 * i + 1]);
        }
    }
    return 0;
}

#pragma GCC visibility pop
}  // extern "C"
