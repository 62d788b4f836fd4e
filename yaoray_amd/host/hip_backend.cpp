// HipBackend: Prepare / Render over the C ABI, with the pass, progress,
// cancel and resume semantics of the reference's CPU backend
// (src/backends/cpu/cpu_backend.cpp:54-102, cpu_path_tracer.cpp:39-229).
#include "hip_backend.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>

#include "../../include/yrh.h"
#include "abi_marshal.hpp"

namespace yr {

class HipPreparedScene final : public PreparedScene {
public:
    HipPreparedScene(RenderJob j, RenderAcceleration a, yrh_context* c, int cus)
        : job(std::move(j)), accel(std::move(a)), ctx(c), compute_units(cus) {}
    ~HipPreparedScene() override { yrh_destroy(ctx); }
    HipPreparedScene(const HipPreparedScene&) = delete;
    HipPreparedScene& operator=(const HipPreparedScene&) = delete;
    RenderBackendKind Kind() const override { return RenderBackendKind::Hip; }
    const RenderSceneIR& SourceScene() const override { return job.scene; }
    const RenderSettings& Settings() const override { return job.settings; }

    RenderJob job;
    RenderAcceleration accel;
    yrh_context* ctx = nullptr;
    int compute_units = 0;
    mutable bool rendering = false;
};

namespace {

std::string LastError(yrh_context* ctx, const char* what) {
    return std::string(what) + ": " + (ctx != nullptr ? yrh_last_error(ctx) : "no context");
}

void ToFilm(const std::vector<yrh_film_pixel>& src, Film& film, std::uint64_t bad) {
    for (int y = 0; y < film.Height(); ++y) {
        for (int x = 0; x < film.Width(); ++x) {
            const yrh_film_pixel& s = src[static_cast<size_t>(y) * static_cast<size_t>(film.Width()) + static_cast<size_t>(x)];
            FilmPixel p;
            p.sum = Color3f{static_cast<float>(s.sum[0]), static_cast<float>(s.sum[1]), static_cast<float>(s.sum[2])};
            p.samples = s.samples;
            p.luminance_mean = static_cast<float>(s.luminance_mean);
            p.luminance_m2 = static_cast<float>(s.luminance_m2);
            p.albedo_sum = Color3f{static_cast<float>(s.albedo_sum[0]), static_cast<float>(s.albedo_sum[1]),
                                   static_cast<float>(s.albedo_sum[2])};
            p.normal_sum = Vec3f{static_cast<float>(s.normal_sum[0]), static_cast<float>(s.normal_sum[1]),
                                 static_cast<float>(s.normal_sum[2])};
            p.depth_sum = static_cast<float>(s.depth_sum);
            p.aov_samples = s.aov_samples;
            film.SetPixelForCheckpoint(x, y, p);
        }
    }
    // Film keeps its bad-sample count private; a non-finite sample only
    // counts, so adding the new ones brings it to the device's total.
    FilmSample nan_sample;
    nan_sample.beauty = Color3f{std::numeric_limits<float>::quiet_NaN(), 0.0f, 0.0f};
    for (std::uint64_t k = film.BadSampleCount(); k < bad; ++k) film.AddSample(0, 0, nan_sample);
}

std::vector<yrh_film_pixel> FromFilm(const Film& film) {
    std::vector<yrh_film_pixel> out(film.Pixels().size());
    for (size_t i = 0; i < out.size(); ++i) {
        const FilmPixel& p = film.Pixels()[i];
        yrh_film_pixel& d = out[i];
        d.sum[0] = p.sum.x;
        d.sum[1] = p.sum.y;
        d.sum[2] = p.sum.z;
        d.samples = p.samples;
        d.luminance_mean = p.luminance_mean;
        d.luminance_m2 = p.luminance_m2;
        d.albedo_sum[0] = p.albedo_sum.x;
        d.albedo_sum[1] = p.albedo_sum.y;
        d.albedo_sum[2] = p.albedo_sum.z;
        d.normal_sum[0] = p.normal_sum.x;
        d.normal_sum[1] = p.normal_sum.y;
        d.normal_sum[2] = p.normal_sum.z;
        d.depth_sum = p.depth_sum;
        d.aov_samples = p.aov_samples;
    }
    return out;
}

int SamplerCode(RenderSamplerKind k) {
    switch (k) {
        case RenderSamplerKind::Independent: return YRH_SAMPLER_INDEPENDENT;
        case RenderSamplerKind::Stratified: return YRH_SAMPLER_STRATIFIED;
        case RenderSamplerKind::OwenSobol: return YRH_SAMPLER_OWEN_SOBOL;
        case RenderSamplerKind::ZSobol: return YRH_SAMPLER_ZSOBOL;
    }
    return YRH_SAMPLER_INDEPENDENT;
}

}  // namespace

RenderBackendCapabilities HipBackend::Capabilities() const {
    return RenderBackendCapabilities{RenderBackendKind::Hip, true, false, true, true, false};
}

BackendPrepareResult HipBackend::Prepare(RenderJob job) {
    BackendPrepareResult result;
    if (job.settings.requested_backend != RenderBackendKind::Hip) {
        result.error = "HIP backend cannot prepare a scene requested for another backend";
        return result;
    }
    const auto start = std::chrono::steady_clock::now();
    PrepareLightSampling(job.scene);
    const auto bvh_start = std::chrono::steady_clock::now();
    RenderAccelerationBuildResult build = BuildRenderAcceleration(job.scene.Geometry());
    const auto bvh_end = std::chrono::steady_clock::now();
    result.acceleration_build_seconds = std::chrono::duration<double>(bvh_end - bvh_start).count();
    if (!build.errors.empty()) {
        result.error = build.errors.front();
        return result;
    }
    yrh_context* ctx = nullptr;
    if (yrh_create(device_, &ctx) != YRH_OK) {
        result.error = "no HIP device " + std::to_string(device_);
        return result;
    }
    yrhip::Marshal m;
    yrh_geometry_desc g{};
    yrhip::DescribeGeometry(job.scene, build.acceleration, m, g);
    if (yrh_upload_geometry(ctx, &g) != YRH_OK) {
        result.error = LastError(ctx, "geometry upload");
        yrh_destroy(ctx);
        return result;
    }
    yrh_shading_desc sd{};
    const std::string err = yrhip::DescribeShading(job.scene, m, sd);
    if (!err.empty() || yrh_upload_shading(ctx, &sd) != YRH_OK) {
        result.error = err.empty() ? LastError(ctx, "shading upload") : err;
        yrh_destroy(ctx);
        return result;
    }
    int cus = 0;
    yrh_device_info(ctx, nullptr, 0, &cus);
    result.scene = std::make_unique<HipPreparedScene>(std::move(job), std::move(build.acceleration), ctx, cus);
    result.ok = true;
    result.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    return result;
}

RenderResult HipBackend::Render(const PreparedScene& scene, const RenderRequest& request) {
    RenderResult result;
    const auto* hip = dynamic_cast<const HipPreparedScene*>(&scene);
    if (scene.Kind() != RenderBackendKind::Hip || hip == nullptr) {
        result.error = "HIP backend received a non-HIP prepared scene.";
        return result;
    }
    if (hip->rendering) throw std::logic_error("HIP prepared scene is already rendering");
    struct Guard {
        const HipPreparedScene* s;
        ~Guard() { s->rendering = false; }
    } guard{hip};
    hip->rendering = true;

    const RenderSettings& settings = hip->Settings();
    yrh_context* ctx = hip->ctx;
    Film film = request.resume_film != nullptr ? *request.resume_film : Film(settings.width, settings.height);
    result.stats.bvh_nodes = hip->accel.NodeCount();
    result.stats.bvh_max_depth = hip->accel.MaxDepth();
    result.stats.threads = hip->compute_units;
    auto fail = [&](std::string message) {
        result.ok = false;
        result.error = std::move(message);
        result.film.emplace(film);
        return result;
    };
    const int spp = std::max(1, settings.spp);
    if (request.resume_completed_spp < 0 || request.resume_completed_spp > spp) return fail("invalid resume completed spp");
    if (request.resume_film != nullptr) {
        if (request.resume_film->Width() != settings.width || request.resume_film->Height() != settings.height) {
            return fail("resume film dimensions do not match render scene");
        }
        // With adaptive sampling converged pixels may hold fewer samples.
        const bool adaptive_resume = settings.adaptive_sampling.enabled;
        for (const FilmPixel& p : request.resume_film->Pixels()) {
            const auto done = static_cast<std::uint32_t>(request.resume_completed_spp);
            if (adaptive_resume ? p.samples > done : p.samples != done) {
                return fail("resume film sample counts do not match resume completed spp");
            }
        }
    }
    const auto start = std::chrono::steady_clock::now();
    if (request.stop_token.stop_requested()) return fail("render cancelled");
    if (yrh_film_reset(ctx, settings.width, settings.height) != YRH_OK) return fail(LastError(ctx, "film reset"));
    const int64_t npix = static_cast<int64_t>(settings.width) * settings.height;
    std::vector<yrh_film_pixel> device_film(static_cast<size_t>(npix));
    if (request.resume_film != nullptr) {
        device_film = FromFilm(film);
        if (yrh_film_load(ctx, device_film.data(), npix) != YRH_OK) return fail(LastError(ctx, "film load"));
    }
    yrh_render_desc d{};
    const RenderCamera& c = hip->job.scene.camera;
    const float cam[12] = {c.origin.x, c.origin.y, c.origin.z, c.forward.x, c.forward.y, c.forward.z,
                           c.right.x,  c.right.y,  c.right.z,  c.up.x,      c.up.y,      c.up.z};
    std::copy(cam, cam + 3, d.camera_origin);
    std::copy(cam + 3, cam + 6, d.camera_forward);
    std::copy(cam + 6, cam + 9, d.camera_right);
    std::copy(cam + 9, cam + 12, d.camera_up);
    d.tan_half_fov = std::tan(c.fov_y_radians * 0.5f);
    d.width = settings.width;
    d.height = settings.height;
    d.sampler = SamplerCode(settings.sampler);
    d.spp = spp;
    d.max_depth = settings.max_depth;
    d.seed = settings.seed;
    d.radiance_clamp = settings.radiance_clamp;

    yrh_render_stats st{};
    const std::uint64_t resumed_bad = request.resume_film != nullptr ? request.resume_film->BadSampleCount() : 0;
    auto sync_film = [&]() -> bool {
        if (yrh_film_read(ctx, device_film.data(), npix) != YRH_OK) return false;
        if (yrh_render_stats_read(ctx, &st) != YRH_OK) return false;
        ToFilm(device_film, film, resumed_bad + st.bad_samples);
        return true;
    };
    auto fill_stats = [&]() {
        result.stats.rays_traced = st.rays_traced;
        result.stats.shadow_rays = st.shadow_rays;
        result.stats.occluded_shadow_rays = st.occluded_shadow_rays;
        result.stats.triangle_tests = st.triangle_tests;
        result.stats.sphere_tests = st.sphere_tests;
        result.stats.bvh_node_tests = st.bvh_node_tests;
        result.stats.hits = st.hits;
        result.stats.misses = st.misses;
        result.stats.samples_rendered = film.TotalSampleCount();
        result.stats.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    };
    // RenderCpuPathTrace's active-pixel loop (cpu_path_tracer.cpp:95-106,
    // 120-157, 183-199): the device keeps the list of pixels below spp samples
    // (that Film::IsConverged has not stopped, with adaptive sampling), and
    // each pass renders the next sample of every listed pixel, its index the
    // pixel's film sample count. A non-finite sample is not added, so that
    // pixel renders the same index again, as in the reference.
    const AdaptiveSamplingSettings& as = settings.adaptive_sampling;
    const bool adaptive = as.enabled;
    const yrh_adaptive_desc ad{adaptive ? 1 : 0, as.min_spp, as.relative_error, as.absolute_error, as.confidence};
    int64_t active = npix, converged = 0;
    if (yrh_film_update_active(ctx, &ad, spp, &active, &converged) != YRH_OK) return fail(LastError(ctx, "active pixels"));
    result.stats.converged_pixels = static_cast<std::uint64_t>(converged);
    d.pixel_list = 1;
    // Without adaptive sampling or a progress callback nothing observes the
    // film between passes, so consecutive passes are batched into one
    // wavefront as large as the device memory suggests (each listed pixel's
    // samples are added in index order, so the film is the same).
    int64_t batch_paths = int64_t{1} << 23;
    if (!request.progress_callback && yrh_suggest_paths_per_pass(ctx, settings.max_depth, &batch_paths) != YRH_OK) {
        return fail(LastError(ctx, "suggest paths per pass"));
    }
    // YAORAY_HIP_MAX_PATHS_PER_PASS caps the wavefront (device memory shared
    // with other work); any cap gives the same film.
    if (const char* cap = std::getenv("YAORAY_HIP_MAX_PATHS_PER_PASS"); cap != nullptr && std::atoll(cap) > 0) {
        batch_paths = std::min<int64_t>(batch_paths, std::atoll(cap));
    }
    const int batch = adaptive || request.progress_callback
        ? 1
        : static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(spp, batch_paths / std::max<int64_t>(1, npix))));
    // Adaptive passes without a progress callback render kAdaptiveBatch samples
    // of each active pixel, added in order until the pixel converges (the same
    // film as one-sample passes; the first ones cover min_spp, before which no
    // pixel can converge). A pass holds at most batch_paths paths (active
    // pixels x samples), so large frames take more, smaller passes; splitting
    // a pixel-list pass leaves the film unchanged.
    constexpr int kAdaptiveBatch = 16;
    for (int pass = request.resume_completed_spp; pass < spp && active > 0; pass += d.sample_count) {
        if (request.stop_token.stop_requested()) {
            if (!sync_film()) return fail(LastError(ctx, "film read"));
            fill_stats();
            return fail("render cancelled");
        }
        d.sample_index = pass;
        int take = batch;
        if (adaptive && !request.progress_callback) {
            take = std::max(kAdaptiveBatch, as.min_spp - pass);
            take = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(take, batch_paths / std::max<int64_t>(1, active))));
        }
        d.sample_count = std::min(take, spp - pass);
        if (yrh_render_pass(ctx, &d) != YRH_OK) return fail(LastError(ctx, "render pass"));
        if (yrh_film_update_active(ctx, &ad, spp, &active, &converged) != YRH_OK) {
            return fail(LastError(ctx, "active pixels"));
        }
        if (adaptive) result.stats.converged_pixels = static_cast<std::uint64_t>(converged);
        if (request.progress_callback) {
            if (!sync_film()) return fail(LastError(ctx, "film read"));
            fill_stats();
            const RenderProgress progress{pass + 1,
                                          spp,
                                          film.TotalSampleCount(),
                                          static_cast<std::uint64_t>(spp) * static_cast<std::uint64_t>(npix),
                                          static_cast<std::uint64_t>(active),
                                          static_cast<std::uint64_t>(converged),
                                          st.rays_traced,
                                          result.stats.elapsed_seconds};
            const RenderProgressDecision decision = request.progress_callback(progress, film);
            if (decision.cancel) {
                return fail(decision.error.empty() ? "render cancelled by progress callback" : decision.error);
            }
        }
    }
    if (!sync_film()) return fail(LastError(ctx, "film read"));
    fill_stats();
    result.ok = true;
    result.film.emplace(std::move(film));
    return result;
}

}  // namespace yr
