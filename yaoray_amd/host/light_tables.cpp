// Light-sampling tables built on the host and consumed as-is by the device:
//   emissive power alias table   src/lighting/light_sampling.cpp:148-216
//   environment distribution     src/lighting/environment.cpp:73-123
// Both feed pdfs that the MIS weights compare against, so they are built with
// the reference's operation order (Vose's alias method with the same small /
// large work lists; row-major luminance x sin(theta) CDFs).
#include "yr_ir.hpp"

#include <algorithm>
#include <cmath>

namespace yr {
namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kMinSinTheta = 1.0e-6f;

float Luminance(Color3f c) { return std::max(0.0f, 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z); }

float EmissiveArea(const EmissivePrimitive& ep) { return std::isfinite(ep.area) && ep.area > 0.0f ? ep.area : 0.0f; }

}  // namespace

AliasTable BuildAliasTable(std::span<const float> weights) {
    AliasTable t;
    const size_t n = weights.size();
    if (n == 0) return t;
    t.probability.assign(n, 1.0f);
    t.alias.resize(n);
    t.pmf.resize(n);
    float total = 0.0f;
    for (float w : weights) {
        if (std::isfinite(w) && w > 0.0f) total += w;
    }
    std::vector<float> scaled(n);
    std::vector<int> small, large;
    small.reserve(n);
    large.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        const float w = std::isfinite(weights[i]) && weights[i] > 0.0f ? weights[i] : 0.0f;
        t.pmf[i] = total > 0.0f ? w / total : 1.0f / static_cast<float>(n);
        scaled[i] = t.pmf[i] * static_cast<float>(n);
        t.alias[i] = static_cast<int>(i);
        (scaled[i] < 1.0f ? small : large).push_back(static_cast<int>(i));
    }
    while (!small.empty() && !large.empty()) {
        const int lo = small.back();
        small.pop_back();
        const int hi = large.back();
        large.pop_back();
        t.probability[static_cast<size_t>(lo)] = scaled[static_cast<size_t>(lo)];
        t.alias[static_cast<size_t>(lo)] = hi;
        scaled[static_cast<size_t>(hi)] = scaled[static_cast<size_t>(hi)] + scaled[static_cast<size_t>(lo)] - 1.0f;
        (scaled[static_cast<size_t>(hi)] < 1.0f ? small : large).push_back(hi);
    }
    return t;
}

void PrepareLightSampling(RenderSceneIR& scene) {
    std::vector<float> w;
    w.reserve(scene.emissive_primitives.size());
    for (const EmissivePrimitive& ep : scene.emissive_primitives) w.push_back(EmissiveArea(ep) * Luminance(ep.radiance));
    scene.light_sampling.emissive_power = BuildAliasTable(w);
}

RenderEnvironmentDistribution BuildEnvironmentDistribution(const RenderTexture& tex) {
    RenderEnvironmentDistribution d;
    d.width = tex.width;
    d.height = tex.height;
    if (tex.width <= 0 || tex.height <= 0 || tex.texels.empty()) {
        d.uniform = true;
        d.width = std::max(1, tex.width);
        d.height = std::max(1, tex.height);
    }
    const int w = std::max(1, d.width);
    const int h = std::max(1, d.height);
    const size_t n = static_cast<size_t>(w) * static_cast<size_t>(h);
    d.texel_weights.assign(n, 0.0f);
    d.row_weights.assign(static_cast<size_t>(h), 0.0f);
    d.row_cdf.assign(static_cast<size_t>(h), 0.0f);
    d.conditional_cdfs.assign(n, 0.0f);
    for (int y = 0; y < h; ++y) {
        const float theta = (static_cast<float>(y) + 0.5f) * kPi / static_cast<float>(h);
        const float sin_theta = std::max(kMinSinTheta, std::sin(theta));
        float row = 0.0f;
        for (int x = 0; x < w; ++x) {
            const size_t i = static_cast<size_t>(y) * static_cast<size_t>(w) + static_cast<size_t>(x);
            const Color4f c = i < tex.texels.size() ? tex.texels[i] : Color4f{0, 0, 0, 1};
            const float weight = Luminance(Color3f{c.x, c.y, c.z}) * sin_theta;
            d.texel_weights[i] = weight;
            row += weight;
            d.conditional_cdfs[i] = row;
        }
        d.row_weights[static_cast<size_t>(y)] = row;
        d.total_weight += row;
        d.row_cdf[static_cast<size_t>(y)] = d.total_weight;
    }
    if (d.total_weight <= 0.0f || !std::isfinite(d.total_weight)) {
        d.uniform = true;
        d.total_weight = static_cast<float>(w * h);
        for (int y = 0; y < h; ++y) {
            float row = 0.0f;
            for (int x = 0; x < w; ++x) {
                const size_t i = static_cast<size_t>(y) * static_cast<size_t>(w) + static_cast<size_t>(x);
                d.texel_weights[i] = 1.0f;
                row += 1.0f;
                d.conditional_cdfs[i] = row;
            }
            d.row_weights[static_cast<size_t>(y)] = row;
            d.row_cdf[static_cast<size_t>(y)] = static_cast<float>((y + 1) * w);
        }
    }
    return d;
}

}  // namespace yr
