// Device-side shading math for the wavefront integrator (gfx950).
//
// Semantics follow the reference function by function (file:line on each);
// the integer sampler is bit-exact, camera rays are bit-exact given the
// host-computed tan(fov/2); functions that call transcendentals (cos, sin,
// pow, exp, log, atan2, acos) use the device's libm and agree statistically.
// Compiled with FP contraction off like the rest of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace yrh {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kOneMinusEps = 0x1.fffffep-1f;  // std::nextafter(1.0f, 0.0f)

// ---- vector helpers (include/yaoray/core/vec.hpp) ----------------------------

__device__ __forceinline__ float3 V(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 Add(float3 a, float3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 Sub(float3 a, float3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 Mul(float3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 Div(float3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float3 Hadamard(float3 a, float3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float3 Neg(float3 a) { return V(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float Dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 Cross(float3 a, float3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float LengthSq(float3 v) { return Dot(v, v); }
__device__ __forceinline__ float Length(float3 v) { return sqrtf(LengthSq(v)); }
__device__ __forceinline__ float3 Normalize(float3 v) {
    const float len = Length(v);
    return len > 0.0f ? Div(v, len) : V(0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ bool IsBlack(float3 c) { return c.x <= 0.0f && c.y <= 0.0f && c.z <= 0.0f; }
__device__ __forceinline__ float MaxComponent(float3 c) { return fmaxf(c.x, fmaxf(c.y, c.z)); }
__device__ __forceinline__ float Clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
__device__ __forceinline__ bool Finite3(float3 c) { return isfinite(c.x) && isfinite(c.y) && isfinite(c.z); }

// ---- sampler (src/sampling/sampler.cpp:34-237, 304-310) ----------------------

constexpr uint64_t kDefaultSeed = 0x9E3779B97F4A7C15ull;
constexpr uint32_t kBounceStride = 16;

enum SamplerKind : int { kIndependent = 0, kStratified = 1, kOwenSobol = 2, kZSobol = 3 };
enum Dim : uint32_t {
    kDimPixel = 0,
    kDimLightSelect = 2,
    kDimLightSurface = 3,
    kDimBsdf = 5,
    kDimRoulette = 7,
    kDimAnalyticSelect = 14,
    kDimGuide = 15,  // guide-or-BSDF choice (and lobe) of the guided proposal; the reference leaves it unused
};

__device__ __forceinline__ uint64_t Mix64(uint64_t v) {
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}

__device__ __forceinline__ uint32_t Mix32(uint32_t v) {
    v ^= v >> 16;
    v *= 0x7FEB352Du;
    v ^= v >> 15;
    v *= 0x846CA68Bu;
    return v ^ (v >> 16);
}

__device__ __forceinline__ uint32_t Hash32(uint64_t seed, uint32_t dim, int lane) {
    const uint64_t m = Mix64(seed ^ (static_cast<uint64_t>(dim) * 0xD1B54A32D192ED03ull) ^
                             (static_cast<uint64_t>(static_cast<uint32_t>(lane)) * 0x94D049BB133111EBull));
    return Mix32(static_cast<uint32_t>(m) ^ static_cast<uint32_t>(m >> 32));
}

__device__ __forceinline__ float UnitFloat(uint32_t v) { return static_cast<float>(v >> 8) * (1.0f / 16777216.0f); }

__device__ __forceinline__ uint32_t Reverse(uint32_t v) { return __builtin_bitreverse32(v); }

// Second Sobol axis: XOR of direction numbers d_0 = 2^31, d_{k+1} = d_k ^ (d_k >> 1)
// over the set bits of the index.
__device__ __forceinline__ uint32_t SobolAxis1(uint32_t index) {
    uint32_t value = 0;
    uint32_t direction = 0x80000000u;
    while (index != 0) {
        if (index & 1u) value ^= direction;
        direction ^= direction >> 1;
        index >>= 1;
    }
    return value;
}

__device__ __forceinline__ uint32_t Sobol(uint32_t index, int axis) {
    return axis == 0 ? Reverse(index) : SobolAxis1(index);
}

// Laine-Karras style nested scramble on the reversed value.
__device__ __forceinline__ uint32_t Owen(uint32_t value, uint32_t seed) {
    value = Reverse(value);
    value += seed;
    value ^= value * 0x6C50B47Cu;
    value ^= value * 0xB82F1E52u;
    value ^= value * 0xC7AFE638u;
    value ^= value * 0x8D22F6E6u;
    return Reverse(value);
}

__device__ __forceinline__ uint32_t Spread16(uint32_t v) {
    v &= 0x0000FFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

struct PixelSampler {
    int kind;
    uint64_t scene_seed;
    uint64_t pixel_seed;
    uint32_t sobol_index;
    int sample_index;
    int spp;
    int light_samples;
    int bounce;

    __device__ __forceinline__ void Init(int k, uint64_t seed, int x, int y, int sample, int samples_per_pixel) {
        kind = k;
        scene_seed = seed;
        pixel_seed = Mix64(seed ^ (static_cast<uint64_t>(static_cast<uint32_t>(x)) << 32) ^
                           static_cast<uint64_t>(static_cast<uint32_t>(y)));
        sample_index = sample;
        spp = samples_per_pixel < 1 ? 1 : samples_per_pixel;
        light_samples = 1;
        bounce = 0;
        if (kind == kZSobol) {
            uint32_t c = 1;
            while (c < static_cast<uint32_t>(spp)) c <<= 1;  // bit_ceil
            const uint32_t bits = 31u - static_cast<uint32_t>(__builtin_clz(c));
            const uint32_t morton = Spread16(static_cast<uint32_t>(x)) | (Spread16(static_cast<uint32_t>(y)) << 1);
            sobol_index = (morton << bits) | static_cast<uint32_t>(sample);
        } else {
            sobol_index = static_cast<uint32_t>(sample);
        }
    }

    __device__ __forceinline__ uint32_t Absolute(uint32_t dim) const {
        return dim == kDimPixel ? dim : dim + static_cast<uint32_t>(bounce) * kBounceStride;
    }

    __device__ __forceinline__ float At1D(uint32_t dim, int lane) const {
        if (kind == kOwenSobol) return UnitFloat(Owen(Sobol(sobol_index, 0), Hash32(pixel_seed ^ kDefaultSeed, dim, lane)));
        if (kind == kZSobol) return UnitFloat(Owen(Sobol(sobol_index, 0), Hash32(scene_seed, dim, lane)));
        return UnitFloat(Hash32(pixel_seed ^ static_cast<uint64_t>(static_cast<int64_t>(sample_index)), dim, lane));
    }

    __device__ __forceinline__ float2 At2D(uint32_t dim, int lane) const {
        if (dim == kDimPixel && spp == 1) return make_float2(0.5f, 0.5f);
        if (kind == kStratified) {
            const int count = dim == kDimPixel ? spp : light_samples;
            const int index = dim == kDimPixel ? sample_index : lane;
            const int columns = static_cast<int>(ceilf(sqrtf(static_cast<float>(count < 1 ? 1 : count))));
            const int rows = (count + columns - 1) / columns;
            const int c = index < 0 ? 0 : (index > count - 1 ? count - 1 : index);
            const int cx = c % columns;
            const int cy = c / columns;
            return make_float2(Clampf((static_cast<float>(cx) + At1D(dim, lane)) / static_cast<float>(columns), 0.0f, kOneMinusEps),
                               Clampf((static_cast<float>(cy) + At1D(dim + 1, lane)) / static_cast<float>(rows), 0.0f, kOneMinusEps));
        }
        if (kind == kOwenSobol || kind == kZSobol) {
            const uint64_t s = kind == kOwenSobol ? pixel_seed : scene_seed;
            return make_float2(UnitFloat(Owen(Sobol(sobol_index, 0), Hash32(s, dim, lane))),
                               UnitFloat(Owen(Sobol(sobol_index, 1), Hash32(s, dim + 1, lane))));
        }
        return make_float2(At1D(dim, lane), At1D(dim + 1, lane));
    }

    __device__ __forceinline__ float Get1D(uint32_t dim) const { return At1D(Absolute(dim), 0); }
    // A hashed uniform of (pixel, sample, dimension) outside the
    // low-discrepancy sequence. Every 1D dimension of the Sobol samplers is a
    // differently scrambled copy of Sobol dimension 0, so a choice drawn from
    // one is correlated with the first coordinate of any 2D sample of the same
    // bounce; a strategy choice that selects how that 2D sample is used must
    // come from here instead.
    __device__ __forceinline__ float Independent1D(uint32_t dim) const {
        return UnitFloat(Hash32(pixel_seed ^ static_cast<uint64_t>(static_cast<int64_t>(sample_index)), Absolute(dim), 0));
    }
    __device__ __forceinline__ float2 Get2D(uint32_t dim) const { return At2D(Absolute(dim), 0); }
};

// ---- camera (src/scene/camera.cpp:9-39) ----------------------------------------

struct DevCamera {
    float3 origin, forward, right, up;
    float tan_half_fov;  // std::tan(fov_y * 0.5f), computed on the host as the reference does
};

__device__ __forceinline__ float3 CameraDirection(const DevCamera& c, float width, float height, float fx, float fy) {
    const float aspect = width / height;
    const float h = c.tan_half_fov;
    const float sx = (2.0f * fx / width - 1.0f) * aspect * h;
    const float sy = (1.0f - 2.0f * fy / height) * h;
    return Normalize(Add(Add(c.forward, Mul(c.right, sx)), Mul(c.up, sy)));
}

// ---- materials / BSDFs (src/shading/bsdf.cpp, bsdf_microfacet.cpp, bsdf_common.cpp)

enum MaterialKind : int { kDiffuse = 0, kConductor = 1, kDielectric = 2, kThinDielectric = 3 };
constexpr float kDeltaRoughness = 1.0e-3f;

struct DevMaterial {
    int kind;
    float3 reflectance;
    float ior, urough, vrough, alpha;
    float3 emission;
    float absorption_distance;
    float3 absorption_color;
};

__device__ __forceinline__ bool Above(float3 d, float3 n) { return Dot(d, n) > 0.0f; }
__device__ __forceinline__ float3 Reflect(float3 d, float3 n) { return Normalize(Sub(d, Mul(n, 2.0f * Dot(d, n)))); }
__device__ __forceinline__ float AbsDot(float3 a, float3 b) { return fabsf(Dot(a, b)); }
__device__ __forceinline__ bool SameHemisphere(float3 a, float3 b, float3 n) { return Dot(a, n) * Dot(b, n) > 0.0f; }

__device__ __forceinline__ float FresnelDielectric(float cos_i, float eta_i, float eta_t) {
    cos_i = Clampf(cos_i, -1.0f, 1.0f);
    if (cos_i <= 0.0f) {
        const float tmp = eta_i;
        eta_i = eta_t;
        eta_t = tmp;
        cos_i = fabsf(cos_i);
    }
    const float sin_i = sqrtf(fmaxf(0.0f, 1.0f - cos_i * cos_i));
    const float sin_t = eta_i / eta_t * sin_i;
    if (sin_t >= 1.0f) return 1.0f;
    const float cos_t = sqrtf(fmaxf(0.0f, 1.0f - sin_t * sin_t));
    const float rpar = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
    const float rperp = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
    return 0.5f * (rpar * rpar + rperp * rperp);
}

__device__ __forceinline__ bool Refract(float3 wo, float3 n, float eta, float3& wi) {
    const float cos_i = Dot(n, wo);
    const float sin2_i = fmaxf(0.0f, 1.0f - cos_i * cos_i);
    const float sin2_t = eta * eta * sin2_i;
    if (sin2_t >= 1.0f) return false;
    const float cos_t = sqrtf(fmaxf(0.0f, 1.0f - sin2_t));
    wi = Normalize(Add(Mul(Neg(wo), eta), Mul(n, eta * cos_i - cos_t)));
    return true;
}

__device__ __forceinline__ void Frame(float3 n, float3& t, float3& b) {
    const float3 helper = fabsf(n.z) < 0.999f ? V(0.0f, 0.0f, 1.0f) : V(1.0f, 0.0f, 0.0f);
    t = Normalize(Cross(helper, n));
    b = Cross(n, t);
}

__device__ __forceinline__ float3 CosineHemisphere(float3 n, float2 s) {
    const float u1 = Clampf(s.x, 0.0f, 1.0f);
    const float u2 = Clampf(s.y, 0.0f, 1.0f);
    const float r = sqrtf(u1);
    const float theta = 2.0f * kPi * u2;
    const float lx = r * cosf(theta);
    const float ly = r * sinf(theta);
    const float lz = sqrtf(fmaxf(0.0f, 1.0f - u1));
    float3 t, b;
    Frame(n, t, b);
    return Normalize(Add(Add(Mul(t, lx), Mul(b, ly)), Mul(n, lz)));
}

__device__ __forceinline__ float RoughnessToAlpha(float r) {
    const float c = fmaxf(kDeltaRoughness, r);
    return c * c;
}

__device__ __forceinline__ float GgxD(float cos_h, float alpha) {
    const float c2 = cos_h * cos_h;
    const float a2 = alpha * alpha;
    const float d = c2 * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * d * d);
}

__device__ __forceinline__ float SmithG1(float cos_t, float alpha) {
    const float c2 = cos_t * cos_t;
    if (c2 <= 0.0f) return 0.0f;
    const float tan2 = (1.0f - c2) / c2;
    return 2.0f / (1.0f + sqrtf(1.0f + alpha * alpha * tan2));
}

__device__ __forceinline__ float3 Schlick(float3 f0, float cos_t) {
    const float t = powf(1.0f - Clampf(cos_t, 0.0f, 1.0f), 5.0f);
    return Add(Mul(f0, 1.0f - t), Mul(V(1.0f, 1.0f, 1.0f), t));
}

__device__ __forceinline__ float3 GgxConductorF(float3 f0, float rough, float3 wo, float3 wi, float3 n) {
    if (!Above(wo, n) || !Above(wi, n)) return V(0, 0, 0);
    const float3 h = Normalize(Add(wo, wi));
    if (LengthSq(h) <= 0.0f || !Above(h, n)) return V(0, 0, 0);
    const float co = fmaxf(0.0f, Dot(n, wo));
    const float ci = fmaxf(0.0f, Dot(n, wi));
    const float ch = fmaxf(0.0f, Dot(n, h));
    const float coh = fmaxf(0.0f, Dot(wo, h));
    if (co <= 0.0f || ci <= 0.0f || coh <= 0.0f) return V(0, 0, 0);
    const float a = RoughnessToAlpha(rough);
    const float d = GgxD(ch, a);
    const float g = SmithG1(co, a) * SmithG1(ci, a);
    return Mul(Schlick(f0, coh), d * g / (4.0f * co * ci));
}

__device__ __forceinline__ float GgxConductorPdf(float rough, float3 wo, float3 wi, float3 n) {
    if (!Above(wo, n) || !Above(wi, n)) return 0.0f;
    const float3 h = Normalize(Add(wo, wi));
    if (LengthSq(h) <= 0.0f || !Above(h, n)) return 0.0f;
    const float ch = fmaxf(0.0f, Dot(n, h));
    const float coh = fmaxf(0.0f, Dot(wo, h));
    if (ch <= 0.0f || coh <= 0.0f) return 0.0f;
    return GgxD(ch, RoughnessToAlpha(rough)) * ch / (4.0f * coh);
}

__device__ __forceinline__ float3 GgxHalfVector(float3 n, float rough, float2 s) {
    const float u1 = Clampf(s.x, 0.0f, 0.999999f);
    const float u2 = Clampf(s.y, 0.0f, 1.0f);
    const float a = RoughnessToAlpha(rough);
    const float a2 = a * a;
    const float tan2 = a2 * u1 / fmaxf(1.0e-6f, 1.0f - u1);
    const float ct = 1.0f / sqrtf(1.0f + tan2);
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    const float phi = 2.0f * kPi * u2;
    float3 t, b;
    Frame(n, t, b);
    return Normalize(Add(Add(Mul(t, st * cosf(phi)), Mul(b, st * sinf(phi))), Mul(n, ct)));
}

__device__ __forceinline__ float3 DielectricReflF(float3 albedo, float ior, float rough, float3 wo, float3 wi, float3 n) {
    if (!SameHemisphere(wo, wi, n)) return V(0, 0, 0);
    const float3 on = Dot(n, wo) >= 0.0f ? n : Neg(n);
    const float3 h = Normalize(Add(wo, wi));
    if (LengthSq(h) <= 0.0f || Dot(h, on) <= 0.0f) return V(0, 0, 0);
    const float co = AbsDot(on, wo), ci = AbsDot(on, wi), ch = AbsDot(on, h), coh = AbsDot(wo, h);
    if (co <= 0.0f || ci <= 0.0f || coh <= 0.0f) return V(0, 0, 0);
    const float a = RoughnessToAlpha(rough);
    const float d = GgxD(ch, a);
    const float g = SmithG1(co, a) * SmithG1(ci, a);
    const float f = FresnelDielectric(coh, 1.0f, fmaxf(1.0f, ior));
    return Mul(albedo, f * d * g / (4.0f * co * ci));
}

__device__ __forceinline__ float DielectricReflPdf(float ior, float rough, float3 wo, float3 wi, float3 n) {
    if (!SameHemisphere(wo, wi, n)) return 0.0f;
    const float3 on = Dot(n, wo) >= 0.0f ? n : Neg(n);
    const float3 h = Normalize(Add(wo, wi));
    if (LengthSq(h) <= 0.0f || Dot(h, on) <= 0.0f) return 0.0f;
    const float ch = AbsDot(on, h), coh = AbsDot(wo, h);
    if (ch <= 0.0f || coh <= 0.0f) return 0.0f;
    const float f = FresnelDielectric(coh, 1.0f, fmaxf(1.0f, ior));
    return f * GgxD(ch, RoughnessToAlpha(rough)) * ch / (4.0f * coh);
}

__device__ __forceinline__ float3 TransmissionHalf(float3 wo, float3 wi, float3 n, float ior) {
    const bool entering = Dot(n, wo) >= 0.0f;
    const float eta_i = entering ? 1.0f : fmaxf(1.0f, ior);
    const float eta_t = entering ? fmaxf(1.0f, ior) : 1.0f;
    float3 h = Normalize(Add(wo, Mul(wi, eta_t / eta_i)));
    if (Dot(h, n) < 0.0f) h = Neg(h);
    return h;
}

__device__ __forceinline__ float3 DielectricTransF(float3 albedo, float ior, float rough, float3 wo, float3 wi, float3 n) {
    if (SameHemisphere(wo, wi, n)) return V(0, 0, 0);
    const bool entering = Dot(n, wo) >= 0.0f;
    const float3 on = entering ? n : Neg(n);
    const float eta_i = entering ? 1.0f : fmaxf(1.0f, ior);
    const float eta_t = entering ? fmaxf(1.0f, ior) : 1.0f;
    const float eta = eta_t / eta_i;
    const float3 h = TransmissionHalf(wo, wi, on, ior);
    if (LengthSq(h) <= 0.0f || Dot(h, on) <= 0.0f) return V(0, 0, 0);
    const float co = AbsDot(on, wo), ci = AbsDot(on, wi), ch = AbsDot(on, h);
    const float woh = Dot(wo, h), wih = Dot(wi, h);
    const float sd = woh + eta * wih;
    if (co <= 0.0f || ci <= 0.0f || ch <= 0.0f || woh <= 0.0f || wih >= 0.0f || sd == 0.0f) return V(0, 0, 0);
    const float a = RoughnessToAlpha(rough);
    const float d = GgxD(ch, a);
    const float g = SmithG1(co, a) * SmithG1(ci, a);
    const float f = FresnelDielectric(fabsf(woh), eta_i, eta_t);
    const float num = fabsf(wih) * fabsf(woh) * eta * eta;
    const float den = co * ci * sd * sd;
    return Mul(albedo, (1.0f - f) * d * g * num / den);
}

__device__ __forceinline__ float DielectricTransPdf(float ior, float rough, float3 wo, float3 wi, float3 n) {
    if (SameHemisphere(wo, wi, n)) return 0.0f;
    const bool entering = Dot(n, wo) >= 0.0f;
    const float3 on = entering ? n : Neg(n);
    const float eta_i = entering ? 1.0f : fmaxf(1.0f, ior);
    const float eta_t = entering ? fmaxf(1.0f, ior) : 1.0f;
    const float eta = eta_t / eta_i;
    const float3 h = TransmissionHalf(wo, wi, on, ior);
    if (LengthSq(h) <= 0.0f || Dot(h, on) <= 0.0f) return 0.0f;
    const float ch = AbsDot(on, h);
    const float woh = Dot(wo, h), wih = Dot(wi, h);
    const float sd = woh + eta * wih;
    if (ch <= 0.0f || woh <= 0.0f || wih >= 0.0f || sd == 0.0f) return 0.0f;
    const float f = FresnelDielectric(fabsf(woh), eta_i, eta_t);
    const float dwh = fabsf((eta * eta * wih) / (sd * sd));
    return (1.0f - f) * GgxD(ch, RoughnessToAlpha(rough)) * ch * dwh;
}

struct DielectricFrame {
    float3 normal;
    float eta_i, eta_t, eta, cos_o;
};

__device__ __forceinline__ DielectricFrame MakeFrame(float3 wo, float3 n, float ior) {
    const bool entering = Dot(n, wo) >= 0.0f;
    DielectricFrame f;
    f.normal = entering ? n : Neg(n);
    f.eta_i = entering ? 1.0f : fmaxf(1.0f, ior);
    f.eta_t = entering ? fmaxf(1.0f, ior) : 1.0f;
    f.eta = f.eta_i / f.eta_t;
    f.cos_o = fmaxf(0.0f, Dot(f.normal, wo));
    return f;
}

// EvaluateBsdf (src/shading/bsdf.cpp:17-77) for the in-scope kinds.
__device__ __forceinline__ float3 EvalBsdf(const DevMaterial& m, float3 wo, float3 wi, float3 n) {
    switch (m.kind) {
        case kDiffuse:
            if (!Above(wo, n) || !Above(wi, n)) return V(0, 0, 0);
            return Div(m.reflectance, kPi);
        case kConductor:
            if (m.urough <= kDeltaRoughness && m.vrough <= kDeltaRoughness) return V(0, 0, 0);
            return GgxConductorF(m.reflectance, m.urough, wo, wi, n);
        case kThinDielectric: {
            if (m.urough <= kDeltaRoughness) return V(0, 0, 0);
            const float3 fwd = Normalize(Neg(wo));
            const float cf = fmaxf(0.0f, Dot(fwd, wi));
            if (cf <= 0.0f) return V(0, 0, 0);
            const float fr = FresnelDielectric(fabsf(Dot(n, wo)), 1.0f, fmaxf(1.0f, m.ior));
            return Mul(Div(m.reflectance, kPi), 1.0f - fr);
        }
        case kDielectric:
            if (m.urough <= kDeltaRoughness) return V(0, 0, 0);
            return SameHemisphere(wo, wi, n) ? DielectricReflF(m.reflectance, m.ior, m.urough, wo, wi, n)
                                             : DielectricTransF(m.reflectance, m.ior, m.urough, wo, wi, n);
    }
    return V(0, 0, 0);
}

// PdfBsdf (src/shading/bsdf.cpp:89-136).
__device__ __forceinline__ float PdfBsdf(const DevMaterial& m, float3 wo, float3 wi, float3 n) {
    switch (m.kind) {
        case kDiffuse:
            if (!Above(wo, n) || !Above(wi, n)) return 0.0f;
            return fmaxf(0.0f, Dot(n, wi)) / kPi;
        case kConductor:
            if (m.urough <= kDeltaRoughness && m.vrough <= kDeltaRoughness) return 0.0f;
            return GgxConductorPdf(m.urough, wo, wi, n);
        case kThinDielectric: {
            if (m.urough <= kDeltaRoughness) return 0.0f;
            return fmaxf(0.0f, Dot(Normalize(Neg(wo)), wi)) / kPi;
        }
        case kDielectric:
            if (m.urough <= kDeltaRoughness) return 0.0f;
            return SameHemisphere(wo, wi, n) ? DielectricReflPdf(m.ior, m.urough, wo, wi, n)
                                             : DielectricTransPdf(m.ior, m.urough, wo, wi, n);
    }
    return 0.0f;
}

struct BsdfSample {
    float3 wi;
    float3 weight;
    float pdf;
    bool valid;
    bool specular;
};

__device__ __forceinline__ BsdfSample Invalid() {
    BsdfSample s;
    s.wi = V(0, 0, 0);
    s.weight = V(0, 0, 0);
    s.pdf = 0.0f;
    s.valid = false;
    s.specular = false;
    return s;
}

__device__ __forceinline__ BsdfSample Make(float3 wi, float3 weight, float pdf, bool specular) {
    BsdfSample s;
    s.wi = wi;
    s.weight = weight;
    s.pdf = pdf;
    s.valid = true;
    s.specular = specular;
    return s;
}

__device__ __forceinline__ BsdfSample SampleGgxConductor(float3 wo, float3 n, float2 s, float3 f0, float rough) {
    if (!Above(wo, n) || IsBlack(f0)) return Invalid();
    const float3 h = GgxHalfVector(n, rough, s);
    const float3 wi = Reflect(Neg(wo), h);
    if (!Above(wi, n)) return Invalid();
    const float pdf = GgxConductorPdf(rough, wo, wi, n);
    if (pdf <= 0.0f) return Invalid();
    const float3 f = GgxConductorF(f0, rough, wo, wi, n);
    return Make(wi, Mul(f, fmaxf(0.0f, Dot(n, wi)) / pdf), pdf, false);
}

// SampleBsdf (src/shading/bsdf.cpp:148-297) for the in-scope kinds.
__device__ __forceinline__ BsdfSample SampleBsdf(const DevMaterial& m, float3 wo, float3 n, float2 s) {
    switch (m.kind) {
        case kDiffuse: {
            if (!Above(wo, n) || IsBlack(m.reflectance)) return Invalid();
            const float3 wi = CosineHemisphere(n, s);
            return Make(wi, m.reflectance, PdfBsdf(m, wo, wi, n), false);
        }
        case kConductor:
            if (m.urough <= kDeltaRoughness && m.vrough <= kDeltaRoughness) {
                if (!Above(wo, n) || IsBlack(m.reflectance)) return Invalid();
                return Make(Reflect(Neg(wo), n), m.reflectance, 1.0f, true);
            }
            return SampleGgxConductor(wo, n, s, m.reflectance, m.urough);
        case kThinDielectric: {
            if (IsBlack(m.reflectance)) return Invalid();
            const float3 on = Dot(n, wo) >= 0.0f ? n : Neg(n);
            const float fr = FresnelDielectric(fabsf(Dot(on, wo)), 1.0f, fmaxf(1.0f, m.ior));
            if (m.urough <= kDeltaRoughness) {
                if (s.x < fr) return Make(Reflect(Neg(wo), on), m.reflectance, 1.0f, true);
                return Make(Normalize(Neg(wo)), m.reflectance, 1.0f, true);
            }
            float2 r = s;
            if (s.x < fr) {
                r.x = fr > 0.0f ? s.x / fr : s.x;
                BsdfSample refl = SampleGgxConductor(wo, on, r, Mul(m.reflectance, fr), m.urough);
                if (refl.valid) refl.specular = false;
                return refl;
            }
            r.x = (1.0f - fr) > 0.0f ? (s.x - fr) / (1.0f - fr) : s.x;
            const float3 fwd = Normalize(Neg(wo));
            const float3 wi = CosineHemisphere(fwd, r);
            const float pdf = fmaxf(0.0f, Dot(fwd, wi)) / kPi;
            if (pdf <= 0.0f) return Invalid();
            return Make(wi, Mul(m.reflectance, 1.0f - fr), pdf, false);
        }
        case kDielectric: {
            if (IsBlack(m.reflectance)) return Invalid();
            const DielectricFrame f = MakeFrame(wo, n, m.ior);
            if (m.urough <= kDeltaRoughness) {
                float3 refr;
                const bool can = Refract(wo, f.normal, f.eta, refr);
                const float fr = can ? FresnelDielectric(f.cos_o, f.eta_i, f.eta_t) : 1.0f;
                if (!can || s.x < fr) return Make(Reflect(Neg(wo), f.normal), m.reflectance, 1.0f, true);
                return Make(refr, m.reflectance, 1.0f, true);
            }
            float2 r = s;
            float3 h = GgxHalfVector(f.normal, m.urough, r);
            if (Dot(h, wo) < 0.0f) h = Neg(h);
            const float fr = FresnelDielectric(fabsf(Dot(wo, h)), f.eta_i, f.eta_t);
            if (s.x < fr) {
                r.x = fr > 0.0f ? s.x / fr : s.x;
                h = GgxHalfVector(f.normal, m.urough, r);
                if (Dot(h, wo) < 0.0f) h = Neg(h);
                const float3 wi = Reflect(Neg(wo), h);
                const float pdf = PdfBsdf(m, wo, wi, n);
                if (pdf <= 0.0f) return Invalid();
                return Make(wi, Mul(EvalBsdf(m, wo, wi, n), AbsDot(n, wi) / pdf), pdf, false);
            }
            r.x = (1.0f - fr) > 0.0f ? (s.x - fr) / (1.0f - fr) : s.x;
            h = GgxHalfVector(f.normal, m.urough, r);
            if (Dot(h, wo) < 0.0f) h = Neg(h);
            float3 wi;
            if (!Refract(wo, h, f.eta, wi)) return Make(Reflect(Neg(wo), h), m.reflectance, 1.0f, true);
            const float pdf = PdfBsdf(m, wo, wi, n);
            if (pdf <= 0.0f) return Invalid();
            return Make(wi, Mul(EvalBsdf(m, wo, wi, n), AbsDot(n, wi) / pdf), pdf, false);
        }
    }
    return Invalid();
}

// IsDeltaBsdf (src/shading/bsdf.cpp:308-329): exact-zero roughness test.
__device__ __forceinline__ bool IsDelta(const DevMaterial& m) {
    if (m.kind == kConductor || m.kind == kDielectric || m.kind == kThinDielectric) {
        return m.urough == 0.0f && m.vrough == 0.0f;
    }
    return false;
}

// ---- MIS (src/integrators/mis.cpp:14-31) --------------------------------------------

__device__ __forceinline__ float PowerHeuristic(int na, float pa, int nb, float pb) {
    if (!(na > 0 && pa > 0.0f && isfinite(pa))) return 0.0f;
    if (!(nb > 0 && pb > 0.0f && isfinite(pb))) return 1.0f;
    const float a = static_cast<float>(na) * pa;
    const float b = static_cast<float>(nb) * pb;
    const float a2 = a * a;
    const float b2 = b * b;
    const float den = a2 + b2;
    if (!isfinite(den) || den <= 0.0f) return 0.0f;
    return a2 / den;
}

// ---- biases (src/integrators/path_integrator_common.cpp:15-20, 125-138) ------------

__device__ __forceinline__ float MaxAbs(float3 p) { return fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))); }
__device__ __forceinline__ float SurfaceBias(float3 p) { return fmaxf(1.0e-4f, MaxAbs(p) * 1.0e-5f); }
__device__ __forceinline__ float ShadowBias(float3 o, float3 t, float dist) {
    const float scale = fmaxf(MaxAbs(o), MaxAbs(t));
    return fminf(fmaxf(1.0e-4f, scale * 1.0e-5f), dist * 1.0e-2f);
}

__device__ __forceinline__ float SafeAbsorption(float v) { return Clampf(v, 1.0e-6f, 1.0f); }
__device__ __forceinline__ float3 BeerLambert(float3 color, float distance_scale, float dist) {
    if (dist <= 0.0f || distance_scale <= 0.0f) return V(1.0f, 1.0f, 1.0f);
    const float inv = 1.0f / distance_scale;
    return V(expf(logf(SafeAbsorption(color.x)) * dist * inv), expf(logf(SafeAbsorption(color.y)) * dist * inv),
             expf(logf(SafeAbsorption(color.z)) * dist * inv));
}

}  // namespace yrh
