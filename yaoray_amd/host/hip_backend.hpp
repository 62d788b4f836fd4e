// The reference's runtime boundary (include/yaoray/runtime/backend.hpp:15-100,
// include/yaoray/film/film.hpp:11-69, include/yaoray/runtime/render_job.hpp)
// with the MI355X backend behind it. Same names, argument meaning and error
// behaviour: `ok = false` + message, Film's constructor throws on a
// non-positive size, the progress callback runs on the caller's thread after
// each complete pass and may cancel, a stop token is polled between passes,
// and a resumed film must carry exactly `resume_completed_spp` samples.
//
// Two builds share this file: the standalone library (yaoray_amd/build.py)
// declares the reference's boundary types itself (yr_ir.hpp and below); the
// drop-in build inside the reference tree (-DYAORAY_HIP_DROPIN, oracle/Makefile
// `dropin`) takes them from the reference's own headers instead, and the
// reference's factory (src/runtime/backend.cpp) gains the Hip case.
#pragma once

#if defined(YAORAY_HIP_DROPIN)

#include <yaoray/accel/acceleration.hpp>
#include <yaoray/lighting/light_sampling.hpp>
#include <yaoray/runtime/backend.hpp>

#else

#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <stop_token>
#include <string>
#include <vector>

#include "yr_ir.hpp"

namespace yr {

struct FilmSample {
    Color3f beauty;
    Color3f albedo;
    Vec3f normal;
    float depth = 0.0f;
    bool primary_hit = false;
};

struct FilmPixel {
    Color3f sum;
    std::uint32_t samples = 0;
    float luminance_mean = 0.0f;
    float luminance_m2 = 0.0f;
    Color3f albedo_sum;
    Vec3f normal_sum;
    float depth_sum = 0.0f;
    std::uint32_t aov_samples = 0;
};
static_assert(sizeof(FilmPixel) == 56);

class Film {
public:
    Film(int width, int height);
    int Width() const { return width_; }
    int Height() const { return height_; }
    Color3f LinearPixel(int x, int y) const;
    std::uint32_t SampleCount(int x, int y) const;
    float LuminanceVariance(int x, int y) const;
    std::uint64_t TotalSampleCount() const;
    std::uint64_t BadSampleCount() const { return bad_sample_count_; }
    const std::vector<FilmPixel>& Pixels() const { return pixels_; }
    void SetPixelForCheckpoint(int x, int y, FilmPixel pixel);
    void AddSample(int x, int y, const FilmSample& sample);

private:
    int width_ = 0;
    int height_ = 0;
    std::vector<FilmPixel> pixels_;
    std::uint64_t bad_sample_count_ = 0;
};

struct RenderJob {
    RenderSceneIR scene;
    RenderSettings settings;
};

struct RenderProgress {
    int completed_spp = 0;
    int target_spp = 0;
    std::uint64_t completed_samples = 0;
    std::uint64_t target_samples = 0;
    std::uint64_t active_pixels = 0;
    std::uint64_t converged_pixels = 0;
    std::uint64_t rays_traced = 0;
    double elapsed_seconds = 0.0;
};

struct RenderProgressDecision {
    bool cancel = false;
    std::string error;
};

using RenderProgressCallback = std::function<RenderProgressDecision(const RenderProgress&, const Film&)>;

struct RenderRequest {
    const Film* resume_film = nullptr;
    int resume_completed_spp = 0;
    std::stop_token stop_token;
    RenderProgressCallback progress_callback;
};

struct RenderStats {
    std::uint64_t samples_rendered = 0;
    std::uint64_t converged_pixels = 0;
    std::uint64_t rays_traced = 0;
    std::uint64_t shadow_rays = 0;
    std::uint64_t occluded_shadow_rays = 0;
    std::uint64_t triangle_tests = 0;
    std::uint64_t sphere_tests = 0;
    std::uint64_t bvh_node_tests = 0;
    int bvh_nodes = 0;
    int bvh_max_depth = 0;
    std::uint64_t hits = 0;
    std::uint64_t misses = 0;
    int threads = 1;
    double elapsed_seconds = 0.0;
};

struct RenderResult {
    bool ok = false;
    std::string error;
    std::optional<Film> film;
    RenderStats stats;
};

class PreparedScene {
public:
    virtual ~PreparedScene() = default;
    virtual RenderBackendKind Kind() const = 0;
    virtual const RenderSceneIR& SourceScene() const = 0;
    virtual const RenderSettings& Settings() const = 0;
};

struct BackendPrepareResult {
    bool ok = false;
    std::string error;
    std::unique_ptr<PreparedScene> scene;
    double elapsed_seconds = 0.0;
    double acceleration_build_seconds = 0.0;
};

struct RenderBackendCapabilities {
    RenderBackendKind kind = RenderBackendKind::Hip;
    bool runnable = false;
    bool supports_debug_direct = false;
    bool supports_path = false;
    bool supports_offline_progress = false;
    bool supports_texture_materials = false;
};

class RenderBackend {
public:
    virtual ~RenderBackend() = default;
    virtual RenderBackendKind Kind() const = 0;
    virtual RenderBackendCapabilities Capabilities() const = 0;
    virtual BackendPrepareResult Prepare(RenderJob job) = 0;
    virtual RenderResult Render(const PreparedScene& scene, const RenderRequest& request) = 0;
};

// The factory's new case: RenderBackendKind::Hip -> HipBackend on device 0
// (src/runtime/backend.cpp:10-18 dispatches the other kinds).
std::unique_ptr<RenderBackend> CreateRenderBackend(RenderBackendKind kind);

}  // namespace yr

#endif  // YAORAY_HIP_DROPIN

struct yrh_context;

namespace yr {

class HipPreparedScene;

class HipBackend final : public RenderBackend {
public:
    explicit HipBackend(int device = 0) : device_(device) {}
    RenderBackendKind Kind() const override { return RenderBackendKind::Hip; }
    RenderBackendCapabilities Capabilities() const override;
    BackendPrepareResult Prepare(RenderJob job) override;
    RenderResult Render(const PreparedScene& scene, const RenderRequest& request) override;

private:
    int device_ = 0;
};

}  // namespace yr
