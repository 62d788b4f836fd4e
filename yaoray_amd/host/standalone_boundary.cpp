// Definitions of the boundary types the standalone library declares itself
// (hip_backend.hpp without YAORAY_HIP_DROPIN): yr::Film as
// src/film/film.cpp defines it, and the backend factory's Hip case. The
// drop-in build links the reference's own film.cpp and its patched factory
// (src/runtime/backend.cpp) instead of this file.
#include <cmath>
#include <stdexcept>

#include "hip_backend.hpp"

namespace yr {

namespace {

bool Finite(Color3f c) { return std::isfinite(c.x) && std::isfinite(c.y) && std::isfinite(c.z); }

}  // namespace

Film::Film(int width, int height) : width_(width), height_(height) {
    if (width <= 0 || height <= 0) throw std::invalid_argument("Film dimensions must be positive");
    pixels_.resize(static_cast<size_t>(width) * static_cast<size_t>(height));
}

Color3f Film::LinearPixel(int x, int y) const {
    if (x < 0 || y < 0 || x >= width_ || y >= height_) throw std::out_of_range("Film pixel coordinate out of range");
    const FilmPixel& p = pixels_[static_cast<size_t>(y) * static_cast<size_t>(width_) + static_cast<size_t>(x)];
    return p.samples == 0 ? Color3f{} : p.sum / static_cast<float>(p.samples);
}

std::uint32_t Film::SampleCount(int x, int y) const {
    if (x < 0 || y < 0 || x >= width_ || y >= height_) throw std::out_of_range("Film pixel coordinate out of range");
    return pixels_[static_cast<size_t>(y) * static_cast<size_t>(width_) + static_cast<size_t>(x)].samples;
}

float Film::LuminanceVariance(int x, int y) const {
    const FilmPixel& p = pixels_[static_cast<size_t>(y) * static_cast<size_t>(width_) + static_cast<size_t>(x)];
    return p.samples > 1 ? p.luminance_m2 / static_cast<float>(p.samples - 1) : 0.0f;
}

std::uint64_t Film::TotalSampleCount() const {
    std::uint64_t total = 0;
    for (const FilmPixel& p : pixels_) total += p.samples;
    return total;
}

void Film::SetPixelForCheckpoint(int x, int y, FilmPixel pixel) {
    if (x < 0 || y < 0 || x >= width_ || y >= height_) return;
    pixels_[static_cast<size_t>(y) * static_cast<size_t>(width_) + static_cast<size_t>(x)] = pixel;
}

// Film::AddSample (src/film/film.cpp:32-56): non-finite beauty only counts;
// FP32 sums, Welford on luminance (0.2126, 0.7152, 0.0722), AOVs on primary hits.
void Film::AddSample(int x, int y, const FilmSample& sample) {
    if (x < 0 || y < 0 || x >= width_ || y >= height_) throw std::out_of_range("Film sample coordinate out of range");
    if (!Finite(sample.beauty)) {
        ++bad_sample_count_;
        return;
    }
    FilmPixel& p = pixels_[static_cast<size_t>(y) * static_cast<size_t>(width_) + static_cast<size_t>(x)];
    p.sum = p.sum + sample.beauty;
    ++p.samples;
    const float lum = 0.2126f * sample.beauty.x + 0.7152f * sample.beauty.y + 0.0722f * sample.beauty.z;
    const float delta = lum - p.luminance_mean;
    p.luminance_mean += delta / static_cast<float>(p.samples);
    p.luminance_m2 += delta * (lum - p.luminance_mean);
    if (sample.primary_hit && Finite(sample.albedo) && Finite(sample.normal) && std::isfinite(sample.depth) &&
        sample.depth >= 0.0f) {
        p.albedo_sum = p.albedo_sum + sample.albedo;
        p.normal_sum = p.normal_sum + sample.normal;
        p.depth_sum += sample.depth;
        ++p.aov_samples;
    }
}

std::unique_ptr<RenderBackend> CreateRenderBackend(RenderBackendKind kind) {
    if (kind == RenderBackendKind::Hip) return std::make_unique<HipBackend>(0);
    return nullptr;
}

}  // namespace yr
