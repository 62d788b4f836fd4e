// Light-sampling tables built on the host and consumed as-is by the device:
//   emissive power alias table   src/lighting/light_sampling.cpp:148-216
//   analytic power alias table + analytic light tree   :206-269
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

namespace {

// Analytic light tree over the local (point / spot) lights: nodes in
// pre-order, each splitting its lights at the median position along the
// widest axis of their positions (src/lighting/light_sampling.cpp:235-268).
// The node bounds and powers are what light selection's importance reads,
// and the split order fixes which light each descent reaches, so the steps
// are the reference's.
struct AnalyticTreeBuilder {
    const std::vector<AnalyticLight>& lights;
    const std::vector<float>& power;
    std::vector<int>& order;
    std::vector<LightTreeNode>& nodes;

    int Build(size_t begin, size_t end) {
        Bounds3f bounds, centroids;
        float sum = 0.0f;
        for (size_t i = begin; i < end; ++i) {
            const int l = order[i];
            const Point3f p = lights[static_cast<size_t>(l)].position;
            bounds = Union(bounds, p);
            centroids = Union(centroids, p);
            sum += power[static_cast<size_t>(l)];
        }
        const int index = static_cast<int>(nodes.size());
        nodes.push_back(LightTreeNode{bounds, sum});
        if (end - begin == 1) {
            nodes[static_cast<size_t>(index)].light_index = order[begin];
            return index;
        }
        const Vec3f e = centroids.max - centroids.min;
        const int axis = e.x >= e.y && e.x >= e.z ? 0 : (e.y >= e.z ? 1 : 2);
        const size_t mid = begin + (end - begin) / 2;
        const auto key = [&](int l) {
            const Point3f p = lights[static_cast<size_t>(l)].position;
            return axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
        };
        std::nth_element(order.begin() + static_cast<std::ptrdiff_t>(begin),
                         order.begin() + static_cast<std::ptrdiff_t>(mid),
                         order.begin() + static_cast<std::ptrdiff_t>(end), [&](int a, int b) { return key(a) < key(b); });
        const int left = Build(begin, mid);
        const int right = Build(mid, end);
        nodes[static_cast<size_t>(index)].left = left;
        nodes[static_cast<size_t>(index)].right = right;
        return index;
    }

    static Bounds3f Union(const Bounds3f& b, Point3f p) {
        return Bounds3f{Point3f{std::min(b.min.x, p.x), std::min(b.min.y, p.y), std::min(b.min.z, p.z)},
                        Point3f{std::max(b.max.x, p.x), std::max(b.max.y, p.y), std::max(b.max.z, p.z)}};
    }
};

}  // namespace

void PrepareLightSampling(RenderSceneIR& scene) {
    std::vector<float> w;
    w.reserve(scene.emissive_primitives.size());
    for (const EmissivePrimitive& ep : scene.emissive_primitives) w.push_back(EmissiveArea(ep) * Luminance(ep.radiance));
    LightSamplingCache& cache = scene.light_sampling;
    cache.emissive_power = BuildAliasTable(w);

    // Analytic power: luminance x solid-angle extent (4 pi for a point light,
    // the cone's 2 pi (1 - cos) for a spot light, 1 for a distant light).
    std::vector<float> power;
    std::vector<int> local;
    power.reserve(scene.analytic_lights.size());
    for (size_t i = 0; i < scene.analytic_lights.size(); ++i) {
        const AnalyticLight& l = scene.analytic_lights[i];
        float p = Luminance(l.intensity);
        if (l.kind == AnalyticLightKind::Spot) {
            p *= std::max(0.0f, 2.0f * kPi * (1.0f - l.cone_angle));
        } else if (l.kind == AnalyticLightKind::Point) {
            p *= 4.0f * kPi;
        }
        power.push_back(p);
        if (l.kind != AnalyticLightKind::Distant && p > 0.0f) local.push_back(static_cast<int>(i));
    }
    cache.analytic_power = BuildAliasTable(power);
    cache.analytic_tree.clear();
    cache.analytic_tree_root = -1;
    cache.use_analytic_tree = local.size() == scene.analytic_lights.size() && local.size() > 8;
    if (!cache.use_analytic_tree) return;
    cache.analytic_tree.reserve(local.size() * 2);
    AnalyticTreeBuilder b{scene.analytic_lights, power, local, cache.analytic_tree};
    cache.analytic_tree_root = b.Build(0, local.size());
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
