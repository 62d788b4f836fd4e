// Texture lookups for the shading kernels (gfx950): the reference's
// SampleTexture4 over a device texel pool.
//
//   wraps, texel fetch    src/shading/texture.cpp:20-41, 94-101
//   nearest / bilinear    texture.cpp:107-130
//   trilinear / EWA       texture.cpp:132-237
//   filter choice         texture.cpp:312-322
//
// Layout: every texture's levels (level 0 as given, then the box-filtered
// chain down to 1 x 1, built at upload) lie in one float4 pool; a texture
// record holds (width, height, filter | wrap_s << 4 | wrap_t << 8, level
// count) and the index of its first level record (width, height, texel
// offset). A texture without texels has level count 0 and samples as
// (0, 0, 0, 1).
//
// Arithmetic follows the reference operation by operation with contraction
// off. Its hypot and log2 are the host libm's; here they are evaluated in
// double and rounded once (hypot of floats is exact in double before the
// square root), which agrees with the host's results except, rarely, in the
// last bit of a level-of-detail weight.
#pragma once

#include "yrh_internal.h"

#pragma clang fp contract(off)

namespace yrh {
namespace tex {

constexpr int kNearest = 0;
constexpr int kBilinear = 1;
constexpr int kTrilinear = 2;
constexpr int kEwa = 3;
constexpr int kRepeat = 0;
constexpr int kClamp = 1;
constexpr int kMirrored = 2;
constexpr int kGaussEntries = 128;

struct TexView {
    const int4* records;  // 2 per texture: (w, h, modes, levels), (first level, 0, 0, 0)
    const int4* levels;   // (w, h, texel offset, 0)
    const float4* texels;
    const float* gauss;   // kGaussEntries EWA weights exp(-2 r^2) - exp(-2)
    int32_t count;
};

struct Level {
    const float4* t;
    int w, h;
};

struct Footprint {
    float dudx, dvdx, dudy, dvdy;
    __device__ __forceinline__ bool Zero() const { return dudx == 0.0f && dvdx == 0.0f && dudy == 0.0f && dvdy == 0.0f; }
};

__device__ __forceinline__ Footprint NoFootprint() { return Footprint{0.0f, 0.0f, 0.0f, 0.0f}; }

__device__ __forceinline__ float4 Blend(float4 a, float4 b, float t) {
    const float s = 1.0f - t;
    return make_float4(a.x * s + b.x * t, a.y * s + b.y * t, a.z * s + b.z * t, a.w * s + b.w * t);
}

__device__ __forceinline__ float WrapCoord(float v, int wrap) {
    if (wrap == kClamp) return v < 0.0f ? 0.0f : (1.0f < v ? 1.0f : v);  // std::clamp: NaN passes
    const float base = floorf(v);
    const float f = v - base;
    if (wrap == kMirrored) return (static_cast<int>(base) & 1) == 0 ? f : 1.0f - f;
    return f < 0.0f ? f + 1.0f : f;
}

__device__ __forceinline__ int WrapIndex(int i, int n, int wrap) {
    if (wrap == kClamp) return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    // In range, or one texel outside (every bilinear fetch, most EWA ones):
    // the same index the modulo forms below give, without integer division.
    if (i >= 0 && i < n) return i;
    if (i == -1) return wrap == kMirrored ? 0 : n - 1;
    if (i == n) return wrap == kMirrored ? n - 1 : 0;
    if (wrap == kMirrored) {
        const int period = 2 * n;
        int k = i % period;
        if (k < 0) k += period;
        return k < n ? k : period - 1 - k;
    }
    int k = i % n;
    if (k < 0) k += n;
    return k;
}

__device__ __forceinline__ float4 Fetch(const Level& l, int x, int y, int ws, int wt) {
    return l.t[static_cast<int64_t>(WrapIndex(y, l.h, wt)) * l.w + WrapIndex(x, l.w, ws)];
}

__device__ __forceinline__ float4 Nearest(const Level& l, float u, float v, int ws, int wt) {
    int x = static_cast<int>(floorf(WrapCoord(u, ws) * static_cast<float>(l.w)));
    int y = static_cast<int>(floorf(WrapCoord(v, wt) * static_cast<float>(l.h)));
    x = x < 0 ? 0 : (x > l.w - 1 ? l.w - 1 : x);
    y = y < 0 ? 0 : (y > l.h - 1 ? l.h - 1 : y);
    return Fetch(l, x, y, ws, wt);
}

__device__ __forceinline__ float4 Bilinear(const Level& l, float u, float v, int ws, int wt) {
    if (l.w == 1 && l.h == 1) return l.t[0];
    const float s = WrapCoord(u, ws) * static_cast<float>(l.w) - 0.5f;
    const float r = WrapCoord(v, wt) * static_cast<float>(l.h) - 0.5f;
    const int s0 = static_cast<int>(floorf(s));
    const int r0 = static_cast<int>(floorf(r));
    const float fs = s - static_cast<float>(s0);
    const float fr = r - static_cast<float>(r0);
    const float4 lo = Blend(Fetch(l, s0, r0, ws, wt), Fetch(l, s0 + 1, r0, ws, wt), fs);
    const float4 hi = Blend(Fetch(l, s0, r0 + 1, ws, wt), Fetch(l, s0 + 1, r0 + 1, ws, wt), fs);
    return Blend(lo, hi, fr);
}

__device__ __forceinline__ float HypotF(float a, float b) {
    const double x = static_cast<double>(a), y = static_cast<double>(b);
    return static_cast<float>(sqrt(x * x + y * y));
}

__device__ __forceinline__ float Log2F(float a) { return static_cast<float>(log2(static_cast<double>(a))); }

struct Tex {
    int w, h, filter, ws, wt, levels, first;
};

__device__ __forceinline__ Tex Load(const TexView& T, int index) {
    const int4 a = T.records[2 * index];
    const int4 b = T.records[2 * index + 1];
    return Tex{a.x, a.y, a.z & 15, (a.z >> 4) & 15, (a.z >> 8) & 15, a.w, b.x};
}

__device__ __forceinline__ Level LevelOf(const TexView& T, const Tex& t, int level) {
    level = level < 0 ? 0 : (level > t.levels - 1 ? t.levels - 1 : level);
    const int4 l = T.levels[t.first + level];
    return Level{T.texels + l.z, l.x, l.y};
}

// Footprint axes in level-0 texels (texture.cpp:132-138, 225-229).
__device__ __forceinline__ void Axes(const Tex& t, const Footprint& f, float& ax, float& ay) {
    const float w = static_cast<float>(t.w < 1 ? 1 : t.w), h = static_cast<float>(t.h < 1 ? 1 : t.h);
    ax = HypotF(f.dudx * w, f.dvdx * h);
    ay = HypotF(f.dudy * w, f.dvdy * h);
}

__device__ __forceinline__ float4 EwaLevel(const TexView& T, const Tex& t, float u, float v, const Footprint& f, int level) {
    const Level l = LevelOf(T, t, level);
    const float W = static_cast<float>(l.w), H = static_cast<float>(l.h);
    float ax = f.dudx * W, ay = f.dvdx * H, bx = f.dudy * W, by = f.dvdy * H;
    if (ax * ax + ay * ay < bx * bx + by * by) {
        float k = ax;
        ax = bx;
        bx = k;
        k = ay;
        ay = by;
        by = k;
    }
    const float major = HypotF(ax, ay);
    float minor = HypotF(bx, by);
    if (minor > 0.0f && major > minor * 8.0f) {
        const float k = major / (minor * 8.0f);
        bx *= k;
        by *= k;
        minor *= k;
    }
    if (minor == 0.0f) return Bilinear(l, u, v, t.ws, t.wt);
    const float s = WrapCoord(u, t.ws) * W - 0.5f;
    const float r = WrapCoord(v, t.wt) * H - 0.5f;
    float A = ay * ay + by * by + 1.0f;
    float B = -2.0f * (ax * ay + bx * by);
    float C = ax * ax + bx * bx + 1.0f;
    const float scale = 1.0f / (A * C - 0.25f * B * B);
    A *= scale;
    B *= scale;
    C *= scale;
    const float det = -B * B + 4.0f * A * C;
    if (det <= 0.0f) return Bilinear(l, u, v, t.ws, t.wt);
    const float inv_det = 1.0f / det;
    const float rs = 2.0f * sqrtf(C * inv_det);
    const float rr = 2.0f * sqrtf(A * inv_det);
    const int s0 = static_cast<int>(ceilf(s - rs)), s1 = static_cast<int>(floorf(s + rs));
    const int r0 = static_cast<int>(ceilf(r - rr)), r1 = static_cast<int>(floorf(r + rr));
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 1.0f);  // Color4f{} starts with alpha 1
    float wsum = 0.0f;
    for (int ir = r0; ir <= r1; ++ir) {
        const float dr = static_cast<float>(ir) - r;
        for (int is = s0; is <= s1; ++is) {
            const float ds = static_cast<float>(is) - s;
            const float q = A * ds * ds + B * ds * dr + C * dr * dr;
            if (q >= 1.0f) continue;
            // min(127, size_t(q * 128)): a (rounding-)negative q at or below
            // -1/128 converts to a huge size_t there, i.e. the last entry.
            const int k = static_cast<int>(q * static_cast<float>(kGaussEntries));
            const float g = T.gauss[k < 0 || k > kGaussEntries - 1 ? kGaussEntries - 1 : k];
            const float4 c = Fetch(l, is, ir, t.ws, t.wt);
            acc = make_float4(acc.x + c.x * g, acc.y + c.y * g, acc.z + c.z * g, acc.w + c.w * g);
            wsum += g;
        }
    }
    if (!(wsum > 0.0f)) return Bilinear(l, u, v, t.ws, t.wt);
    return make_float4(acc.x / wsum, acc.y / wsum, acc.z / wsum, acc.w / wsum);
}

// SampleTexture4 of texture `index` (texture.cpp:312-322).
__device__ __forceinline__ float4 Sample(const TexView& T, int index, float u, float v, const Footprint& f) {
    const Tex t = Load(T, index);
    if (t.levels <= 0) return make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (t.filter == kNearest) return Nearest(LevelOf(T, t, 0), u, v, t.ws, t.wt);
    if (t.filter >= kTrilinear && !f.Zero() && t.levels > 1) {
        float ax, ay;
        Axes(t, f, ax, ay);
        const float top = static_cast<float>(t.levels - 1);
        float lod;
        if (t.filter == kTrilinear) {
            lod = fminf(fmaxf(0.0f, Log2F(fmaxf(1.0f, fmaxf(ax, ay)))), top);
        } else {
            lod = fminf(fmaxf(0.0f, Log2F(fmaxf(1.0f, fminf(ax, ay)))), top);
        }
        const int l0 = static_cast<int>(floorf(lod));
        const int l1 = l0 + 1 < t.levels - 1 ? l0 + 1 : t.levels - 1;
        const float frac = lod - static_cast<float>(l0);
        if (t.filter == kTrilinear) {
            return Blend(Bilinear(LevelOf(T, t, l0), u, v, t.ws, t.wt), Bilinear(LevelOf(T, t, l1), u, v, t.ws, t.wt),
                         frac);
        }
        return Blend(EwaLevel(T, t, u, v, f, l0), EwaLevel(T, t, u, v, f, l1), frac);
    }
    return Bilinear(LevelOf(T, t, 0), u, v, t.ws, t.wt);
}

}  // namespace tex
}  // namespace yrh
