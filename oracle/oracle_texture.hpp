// ORACLE — test infrastructure only (see oracle_trace.cpp).
//
// CPU restatement of the reference's texture lookups over yrh_texture inputs:
//   wraps, texel fetch       src/shading/texture.cpp:20-41, 94-101
//   mip chain                texture.cpp:241-272 (BuildTextureMipChain)
//   nearest / bilinear       texture.cpp:107-130
//   trilinear / EWA          texture.cpp:132-237 (LOD from the footprint,
//                            128-entry Gaussian weight table, anisotropy <= 8)
//   filter choice            texture.cpp:312-322 (SampleTexture4)
// with the same operation order and libm calls, so the values equal the
// reference's bit for bit on one machine (pinned by tests/test_textures.py
// against oracle/_ref and closed forms). The sharing cache of
// BuildTextureSamplingCaches (texture.cpp:274-298) only deduplicates storage.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <vector>

#include "../include/yrh.h"

namespace ortex {

// yr::Color4f: default alpha 1 (include/yaoray/core/vec.hpp:33-48).
struct C4 {
    float x = 0.0f, y = 0.0f, z = 0.0f, w = 1.0f;
};
inline C4 operator+(C4 a, C4 b) { return C4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline C4 operator*(C4 a, float s) { return C4{a.x * s, a.y * s, a.z * s, a.w * s}; }
inline C4 operator/(C4 a, float s) { return C4{a.x / s, a.y / s, a.z / s, a.w / s}; }
inline C4 Blend(C4 a, C4 b, float t) { return a * (1.0f - t) + b * t; }

struct Footprint {
    float dudx = 0.0f, dvdx = 0.0f, dudy = 0.0f, dvdy = 0.0f;
    bool Zero() const { return dudx == 0.0f && dvdx == 0.0f && dudy == 0.0f && dvdy == 0.0f; }
};

struct Level {
    int w = 0, h = 0;
    std::vector<C4> t;
};

struct Texture {
    int width = 0, height = 0, filter = YRH_FILTER_BILINEAR, wrap_s = YRH_WRAP_REPEAT, wrap_t = YRH_WRAP_REPEAT;
    std::vector<Level> levels;  // [0] = the texels as given, then the mip chain
    bool HasData() const { return width > 0 && height > 0 && !levels.empty() && !levels[0].t.empty(); }
    int LevelCount() const { return std::max(1, static_cast<int>(levels.size())); }
    const Level& At(int level) const {
        if (level <= 0 || levels.size() <= 1) return levels[0];
        return levels[static_cast<size_t>(std::min(level, static_cast<int>(levels.size()) - 1))];
    }
};

inline float WrapCoord(float v, int wrap) {
    if (wrap == YRH_WRAP_CLAMP) return std::clamp(v, 0.0f, 1.0f);
    const float base = std::floor(v);
    const float f = v - base;
    if (wrap == YRH_WRAP_MIRRORED) return (static_cast<int>(base) & 1) == 0 ? f : 1.0f - f;
    return f < 0.0f ? f + 1.0f : f;
}

inline int WrapIndex(int i, int n, int wrap) {
    if (wrap == YRH_WRAP_CLAMP) return std::clamp(i, 0, n - 1);
    if (wrap == YRH_WRAP_MIRRORED) {
        const int period = 2 * n;
        int k = i % period;
        if (k < 0) k += period;
        return k < n ? k : period - 1 - k;
    }
    int k = i % n;
    if (k < 0) k += n;
    return k;
}

inline C4 Fetch(const Level& l, int x, int y, int ws, int wt) {
    if (l.w <= 0 || l.h <= 0 || l.t.empty()) return C4{};
    const size_t i = static_cast<size_t>(WrapIndex(y, l.h, wt)) * static_cast<size_t>(l.w) +
                     static_cast<size_t>(WrapIndex(x, l.w, ws));
    return i < l.t.size() ? l.t[i] : C4{};
}

inline Texture Build(const yrh_texture& d) {
    Texture t;
    t.width = d.width;
    t.height = d.height;
    t.filter = d.filter;
    t.wrap_s = d.wrap_s;
    t.wrap_t = d.wrap_t;
    Level l0;
    if (d.width > 0 && d.height > 0 && d.texels != nullptr) {
        l0.w = d.width;
        l0.h = d.height;
        const size_t n = static_cast<size_t>(d.width) * static_cast<size_t>(d.height);
        l0.t.resize(n);
        for (size_t i = 0; i < n; ++i) l0.t[i] = C4{d.texels[4 * i], d.texels[4 * i + 1], d.texels[4 * i + 2], d.texels[4 * i + 3]};
    }
    t.levels.push_back(std::move(l0));
    if (!t.HasData()) return t;
    // Box-filtered halvings down to 1x1; odd edges fetch through the wrap modes.
    while (t.levels.back().w > 1 || t.levels.back().h > 1) {
        const Level& src = t.levels.back();
        Level m;
        m.w = std::max(1, (src.w + 1) / 2);
        m.h = std::max(1, (src.h + 1) / 2);
        m.t.resize(static_cast<size_t>(m.w) * static_cast<size_t>(m.h));
        for (int y = 0; y < m.h; ++y) {
            for (int x = 0; x < m.w; ++x) {
                const C4 s = Fetch(src, 2 * x, 2 * y, t.wrap_s, t.wrap_t) + Fetch(src, 2 * x + 1, 2 * y, t.wrap_s, t.wrap_t) +
                             Fetch(src, 2 * x, 2 * y + 1, t.wrap_s, t.wrap_t) +
                             Fetch(src, 2 * x + 1, 2 * y + 1, t.wrap_s, t.wrap_t);
                m.t[static_cast<size_t>(y) * static_cast<size_t>(m.w) + static_cast<size_t>(x)] = s * 0.25f;
            }
        }
        t.levels.push_back(std::move(m));
    }
    return t;
}

inline C4 Nearest(const Texture& t, const Level& l, float u, float v) {
    if (l.w <= 0 || l.h <= 0 || l.t.empty()) return C4{};
    const int x = std::clamp(static_cast<int>(std::floor(WrapCoord(u, t.wrap_s) * static_cast<float>(l.w))), 0, l.w - 1);
    const int y = std::clamp(static_cast<int>(std::floor(WrapCoord(v, t.wrap_t) * static_cast<float>(l.h))), 0, l.h - 1);
    return Fetch(l, x, y, t.wrap_s, t.wrap_t);
}

inline C4 Bilinear(const Texture& t, const Level& l, float u, float v) {
    if (l.w <= 0 || l.h <= 0 || l.t.empty()) return C4{};
    if (l.w == 1 && l.h == 1) return l.t[0];
    const float s = WrapCoord(u, t.wrap_s) * static_cast<float>(l.w) - 0.5f;
    const float r = WrapCoord(v, t.wrap_t) * static_cast<float>(l.h) - 0.5f;
    const int s0 = static_cast<int>(std::floor(s)), r0 = static_cast<int>(std::floor(r));
    const float fs = s - static_cast<float>(s0), fr = r - static_cast<float>(r0);
    const C4 lo = Blend(Fetch(l, s0, r0, t.wrap_s, t.wrap_t), Fetch(l, s0 + 1, r0, t.wrap_s, t.wrap_t), fs);
    const C4 hi = Blend(Fetch(l, s0, r0 + 1, t.wrap_s, t.wrap_t), Fetch(l, s0 + 1, r0 + 1, t.wrap_s, t.wrap_t), fs);
    return Blend(lo, hi, fr);
}

// Footprint axes in level-0 texels.
inline void Axes(const Texture& t, Footprint f, float& ax, float& ay) {
    const float w = static_cast<float>(std::max(1, t.width)), h = static_cast<float>(std::max(1, t.height));
    ax = std::hypot(f.dudx * w, f.dvdx * h);
    ay = std::hypot(f.dudy * w, f.dvdy * h);
}

inline const std::array<float, 128>& GaussTable() {
    static const std::array<float, 128> table = [] {
        std::array<float, 128> v{};
        const float falloff = 2.0f;
        const float edge = std::exp(-falloff);
        for (size_t i = 0; i < v.size(); ++i) {
            const float r2 = static_cast<float>(i) / static_cast<float>(v.size() - 1);
            v[i] = std::exp(-falloff * r2) - edge;
        }
        return v;
    }();
    return table;
}

inline C4 EwaLevel(const Texture& t, float u, float v, Footprint f, int level) {
    const Level& l = t.At(level);
    if (l.w <= 0 || l.h <= 0 || l.t.empty()) return C4{};
    const float W = static_cast<float>(l.w), H = static_cast<float>(l.h);
    float ax = f.dudx * W, ay = f.dvdx * H, bx = f.dudy * W, by = f.dvdy * H;
    if (ax * ax + ay * ay < bx * bx + by * by) {
        std::swap(ax, bx);
        std::swap(ay, by);
    }
    const float major = std::hypot(ax, ay);
    float minor = std::hypot(bx, by);
    if (minor > 0.0f && major > minor * 8.0f) {
        const float k = major / (minor * 8.0f);
        bx *= k;
        by *= k;
        minor *= k;
    }
    if (minor == 0.0f) return Bilinear(t, l, u, v);
    const float s = WrapCoord(u, t.wrap_s) * W - 0.5f;
    const float r = WrapCoord(v, t.wrap_t) * H - 0.5f;
    // Implicit ellipse A ds^2 + B ds dr + C dr^2 < 1 over the two axes.
    float A = ay * ay + by * by + 1.0f;
    float B = -2.0f * (ax * ay + bx * by);
    float C = ax * ax + bx * bx + 1.0f;
    const float scale = 1.0f / (A * C - 0.25f * B * B);
    A *= scale;
    B *= scale;
    C *= scale;
    const float det = -B * B + 4.0f * A * C;
    if (det <= 0.0f) return Bilinear(t, l, u, v);
    const float inv_det = 1.0f / det;
    const float rs = 2.0f * std::sqrt(C * inv_det);
    const float rr = 2.0f * std::sqrt(A * inv_det);
    const int s0 = static_cast<int>(std::ceil(s - rs)), s1 = static_cast<int>(std::floor(s + rs));
    const int r0 = static_cast<int>(std::ceil(r - rr)), r1 = static_cast<int>(std::floor(r + rr));
    const std::array<float, 128>& g = GaussTable();
    C4 acc;
    float wsum = 0.0f;
    for (int ir = r0; ir <= r1; ++ir) {
        for (int is = s0; is <= s1; ++is) {
            const float ds = static_cast<float>(is) - s;
            const float dr = static_cast<float>(ir) - r;
            const float q = A * ds * ds + B * ds * dr + C * dr * dr;
            if (q >= 1.0f) continue;
            const size_t k = std::min(g.size() - 1, static_cast<size_t>(q * static_cast<float>(g.size())));
            acc = acc + Fetch(l, is, ir, t.wrap_s, t.wrap_t) * g[k];
            wsum += g[k];
        }
    }
    return wsum > 0.0f ? acc / wsum : Bilinear(t, l, u, v);
}

inline C4 Trilinear(const Texture& t, float u, float v, Footprint f) {
    float ax, ay;
    Axes(t, f, ax, ay);
    const float top = static_cast<float>(t.LevelCount() - 1);
    const float lod = std::min(std::max(0.0f, std::log2(std::max(1.0f, std::max(ax, ay)))), top);
    const int l0 = static_cast<int>(std::floor(lod));
    const int l1 = std::min(l0 + 1, t.LevelCount() - 1);
    return Blend(Bilinear(t, t.At(l0), u, v), Bilinear(t, t.At(l1), u, v), lod - static_cast<float>(l0));
}

inline C4 Ewa(const Texture& t, float u, float v, Footprint f) {
    float ax, ay;
    Axes(t, f, ax, ay);
    const float minor = std::max(1.0f, std::min(ax, ay));
    const float lod = std::min(std::max(0.0f, std::log2(minor)), static_cast<float>(t.LevelCount() - 1));
    const int l0 = static_cast<int>(std::floor(lod));
    const int l1 = std::min(l0 + 1, t.LevelCount() - 1);
    return Blend(EwaLevel(t, u, v, f, l0), EwaLevel(t, u, v, f, l1), lod - static_cast<float>(l0));
}

inline C4 Sample(const Texture& t, float u, float v, Footprint f) {
    if (!t.HasData()) return C4{};
    if (t.filter == YRH_FILTER_NEAREST) return Nearest(t, t.At(0), u, v);
    if (t.filter == YRH_FILTER_TRILINEAR && !f.Zero() && t.LevelCount() > 1) return Trilinear(t, u, v, f);
    if (t.filter == YRH_FILTER_EWA && !f.Zero() && t.LevelCount() > 1) return Ewa(t, u, v, f);
    return Bilinear(t, t.At(0), u, v);
}

}  // namespace ortex
