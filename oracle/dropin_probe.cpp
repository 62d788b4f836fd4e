// ORACLE — test infrastructure only (see oracle/oracle_trace.cpp).
//
// Drives the HIP backend as the reference itself would: compiled against
// the reference's headers with yaoray_amd/host/dropin/reference_hip_backend.patch
// applied, linked with the reference's own frontend, factory, film and
// checkpoint code (oracle/Makefile `dropin`). Contains no reference code.
//
//   dropin_probe render <scene.pbrt> <cpu|hip> <out.bin> [spp] [max_depth]
//       LoadPbrtScene -> CompilePbrtScene -> CreateRenderBackend(kind) ->
//       Prepare -> Render (src/app/main.cpp:100-175 flow), then writes the
//       film and stats.
//   dropin_probe checkpoint <scene.pbrt> <cpu|hip> <ckpt.yrc> <out.bin> <passes>
//       Render cancelled by its progress callback after <passes> passes,
//       WriteFilmCheckpoint (src/film/film_checkpoint.cpp:71), LoadFilmCheckpoint
//       (:120) with the settings hash, then Render resumed from the loaded film.
//
// out.bin: "YRPF", int32 ok, width, height, spp, uint64 stats[10]
// (rays, shadow, occluded, triangle tests, sphere tests, node tests, hits,
// misses, bad samples, converged pixels), int32 error length + bytes, then
// width * height FilmPixel records (56 bytes).
#include <yaoray/film/film_checkpoint.hpp>
#include <yaoray/frontend/pbrt/pbrt_scene.hpp>
#include <yaoray/frontend/pbrt/scene_compiler.hpp>
#include <yaoray/runtime/backend.hpp>
#include <yaoray/scene/diagnostic.hpp>
#include <yaoray/scene/render_scene_hash.hpp>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <optional>
#include <string>

namespace {

bool WriteOut(const std::string& path, const yr::RenderResult& r, int spp, int width, int height) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    const std::int32_t head[4] = {r.ok ? 1 : 0, r.film ? r.film->Width() : width, r.film ? r.film->Height() : height,
                                  spp};
    f.write("YRPF", 4);
    f.write(reinterpret_cast<const char*>(head), sizeof(head));
    const std::uint64_t stats[10] = {r.stats.rays_traced,    r.stats.shadow_rays,   r.stats.occluded_shadow_rays,
                                     r.stats.triangle_tests, r.stats.sphere_tests,  r.stats.bvh_node_tests,
                                     r.stats.hits,           r.stats.misses,
                                     r.film ? r.film->BadSampleCount() : 0, r.stats.converged_pixels};
    f.write(reinterpret_cast<const char*>(stats), sizeof(stats));
    const std::int32_t n = static_cast<std::int32_t>(r.error.size());
    f.write(reinterpret_cast<const char*>(&n), sizeof(n));
    f.write(r.error.data(), n);
    if (r.film) {
        const auto& px = r.film->Pixels();
        f.write(reinterpret_cast<const char*>(px.data()), static_cast<std::streamsize>(px.size() * sizeof(px[0])));
    }
    return static_cast<bool>(f);
}

std::optional<yr::RenderBackendKind> Kind(const std::string& name) { return yr::ParseRenderBackendName(name); }

struct Loaded {
    yr::RenderSceneIR scene;
    yr::RenderSettings settings;
};

std::optional<Loaded> Load(const char* path) {
    yr::PbrtSceneLoadResult parsed = yr::LoadPbrtScene(path);
    if (yr::HasSceneErrors(parsed.diagnostics) || !parsed.scene) {
        std::cerr << yr::FormatSceneDiagnostics(parsed.diagnostics) << '\n';
        return std::nullopt;
    }
    yr::SceneCompileResult compiled = yr::CompilePbrtScene(*parsed.scene);
    if (yr::HasSceneErrors(compiled.diagnostics) || !compiled.scene) {
        std::cerr << yr::FormatSceneDiagnostics(compiled.diagnostics) << '\n';
        return std::nullopt;
    }
    return Loaded{std::move(*compiled.scene), compiled.settings};
}

int Render(int argc, char** argv) {
    if (argc < 5) return 2;
    auto loaded = Load(argv[2]);
    const auto kind = Kind(argv[3]);
    if (!loaded || !kind) return 2;
    if (argc > 5) loaded->settings.spp = std::atoi(argv[5]);
    if (argc > 6) loaded->settings.max_depth = std::atoi(argv[6]);
    loaded->settings.requested_backend = *kind;
    const int spp = loaded->settings.spp, w = loaded->settings.width, h = loaded->settings.height;
    const auto backend = yr::CreateRenderBackend(*kind);
    if (!backend) {
        std::cerr << "Render backend not available: " << argv[3] << '\n';
        return 1;
    }
    yr::BackendPrepareResult prep = backend->Prepare(yr::RenderJob{std::move(loaded->scene), loaded->settings});
    yr::RenderResult r;
    if (!prep.ok || !prep.scene) {
        r.error = "Render backend preparation failed: " + prep.error;
    } else {
        r = backend->Render(*prep.scene, yr::RenderRequest{});
    }
    std::cout << "backend " << argv[3] << " ok " << r.ok << " error '" << r.error << "' rays " << r.stats.rays_traced
              << '\n';
    return WriteOut(argv[4], r, spp, w, h) ? 0 : 1;
}

int Checkpoint(int argc, char** argv) {
    if (argc < 7) return 2;
    auto loaded = Load(argv[2]);
    const auto kind = Kind(argv[3]);
    if (!loaded || !kind) return 2;
    loaded->settings.requested_backend = *kind;
    const int passes = std::atoi(argv[6]);
    const std::uint64_t hash = yr::ComputeRenderSettingsHash(loaded->scene, loaded->settings);
    const auto backend = yr::CreateRenderBackend(*kind);
    if (!backend) return 1;
    yr::BackendPrepareResult prep = backend->Prepare(yr::RenderJob{std::move(loaded->scene), loaded->settings});
    if (!prep.ok || !prep.scene) {
        std::cerr << "prepare: " << prep.error << '\n';
        return 1;
    }
    const yr::RenderSettings& st = prep.scene->Settings();
    yr::RenderRequest first;
    first.progress_callback = [&](const yr::RenderProgress& p, const yr::Film&) {
        return yr::RenderProgressDecision{p.completed_spp >= passes, "checkpoint"};
    };
    const yr::RenderResult part = backend->Render(*prep.scene, first);
    if (part.ok || !part.film || part.error != "checkpoint") {
        std::cerr << "interrupted render: " << part.error << '\n';
        return 1;
    }
    const yr::FilmCheckpointWriteResult wr =
        yr::WriteFilmCheckpoint(argv[4], *part.film, yr::FilmCheckpointMetadata{st.width, st.height, st.spp, passes, hash});
    if (!wr.ok) {
        std::cerr << "checkpoint write: " << wr.error << '\n';
        return 1;
    }
    yr::FilmCheckpointLoadResult lr = yr::LoadFilmCheckpoint(argv[4], st.width, st.height, st.spp, hash);
    if (!lr.ok || !lr.film || lr.metadata.completed_spp != passes) {
        std::cerr << "checkpoint load: " << lr.error << '\n';
        return 1;
    }
    yr::RenderRequest resume;
    resume.resume_film = &*lr.film;
    resume.resume_completed_spp = lr.metadata.completed_spp;
    const yr::RenderResult r = backend->Render(*prep.scene, resume);
    std::cout << "resumed " << argv[3] << " from " << passes << " passes: ok " << r.ok << " error '" << r.error
              << "'\n";
    return WriteOut(argv[5], r, st.spp, st.width, st.height) ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "render") == 0) return Render(argc, argv);
    if (argc >= 2 && std::strcmp(argv[1], "checkpoint") == 0) return Checkpoint(argc, argv);
    std::cerr << "usage: dropin_probe render <scene.pbrt> <cpu|hip> <out.bin> [spp] [max_depth]\n"
                 "       dropin_probe checkpoint <scene.pbrt> <cpu|hip> <ckpt> <out.bin> <passes>\n";
    return 2;
}
