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
//   dropin_probe dump <scene.pbrt> <out.bin>
//       The compiled RenderSceneIR and RenderSettings as the C ABI carries them
//       (raw vertex / index / primitive / sphere / instance arrays, materials,
//       emitters and analytic lights through the backend's own marshalling,
//       textures with their level-0 texels, the environment map's index):
//       "YRIR", then sections (16-byte name, int64 count, int32 element size,
//       bytes) up to a section named "end".
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
#include <yaoray/lighting/light_sampling.hpp>

#include "../yaoray_amd/host/abi_marshal.hpp"

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

void Section(std::ofstream& f, const char* name, const void* data, int64_t count, int32_t elem) {
    char tag[16] = {};
    std::strncpy(tag, name, 15);
    f.write(tag, 16);
    f.write(reinterpret_cast<const char*>(&count), 8);
    f.write(reinterpret_cast<const char*>(&elem), 4);
    if (count > 0) f.write(static_cast<const char*>(data), static_cast<std::streamsize>(count * elem));
}

int Dump(int argc, char** argv) {
    if (argc < 4) return 2;
    auto loaded = Load(argv[2]);
    if (!loaded) return 2;
    yr::RenderSceneIR& ir = loaded->scene;
    const yr::RenderSettings& st = loaded->settings;
    yr::PrepareLightSampling(ir);
    yrhip::Marshal m;
    yrh_shading_desc sd{};
    const std::string err = yrhip::DescribeShading(ir, m, sd);
    if (!err.empty()) {
        std::cerr << "dump: " << err << '\n';
        return 1;
    }
    std::ofstream f(argv[3], std::ios::binary);
    f.write("YRIR", 4);
    const yr::RenderCamera& c = ir.camera;
    const float cam[13] = {c.origin.x, c.origin.y, c.origin.z, c.forward.x, c.forward.y, c.forward.z, c.right.x,
                           c.right.y,  c.right.z,  c.up.x,     c.up.y,      c.up.z,      c.fov_y_radians};
    Section(f, "camera", cam, 13, 4);
    const int64_t settings[6] = {st.width, st.height, st.spp, st.max_depth, static_cast<int64_t>(st.sampler),
                                 static_cast<int64_t>(st.seed)};
    Section(f, "settings", settings, 6, 8);
    Section(f, "vertices", ir.vertices.data(), static_cast<int64_t>(ir.vertices.size()), 48);
    Section(f, "indices", ir.indices.data(), static_cast<int64_t>(ir.indices.size()), 4);
    Section(f, "primitives", ir.primitives.data(), static_cast<int64_t>(ir.primitives.size()), 16);
    std::vector<yrh_sphere> sph(ir.spheres.size());
    for (size_t k = 0; k < sph.size(); ++k) {
        const yr::RenderSphere& r = ir.spheres[k];
        sph[k] = yrh_sphere{{r.center.x, r.center.y, r.center.z}, r.radius, r.material_index, r.area_light_index,
                            static_cast<uint8_t>(r.flip_normals ? 1 : 0), {0, 0, 0}};
    }
    Section(f, "spheres", sph.data(), static_cast<int64_t>(sph.size()), sizeof(yrh_sphere));
    std::vector<int32_t> inst_prim;
    std::vector<float> inst_mat;
    for (const yr::RenderInstance& in : ir.instances) {
        inst_prim.push_back(in.primitive.Value());
        inst_mat.insert(inst_mat.end(), in.object_to_world.m.begin(), in.object_to_world.m.end());
    }
    Section(f, "inst_prim", inst_prim.data(), static_cast<int64_t>(inst_prim.size()), 4);
    Section(f, "inst_mat", inst_mat.data(), static_cast<int64_t>(inst_mat.size()), 4);
    Section(f, "materials", sd.materials, sd.material_count, sizeof(yrh_material));
    Section(f, "emissives", sd.emissives, sd.emissive_count, sizeof(yrh_emissive));
    Section(f, "analytic", sd.analytic_lights, sd.analytic_light_count, sizeof(yrh_analytic_light));
    const float env[4] = {ir.environment.radiance.x, ir.environment.radiance.y, ir.environment.radiance.z,
                          ir.environment.strength};
    Section(f, "environment", env, 4, 4);
    // Textures (level 0; RenderSceneIR::textures order, the environment map
    // among them) and the environment's texture index and rotation.
    std::vector<yrh_texture> tex(ir.textures.size());
    std::vector<float> texels;
    for (size_t k = 0; k < tex.size(); ++k) {
        const yr::RenderTexture& t = ir.textures[k];
        tex[k] = yrh_texture{t.width, t.height, static_cast<int32_t>(t.filter), static_cast<int32_t>(t.wrap_s),
                             static_cast<int32_t>(t.wrap_t), 0, nullptr};
        const std::vector<yr::Color4f>& px = t.sampling_cache ? t.sampling_cache->texels : t.texels;
        if (static_cast<int64_t>(px.size()) != static_cast<int64_t>(t.width) * t.height) {
            std::cerr << "dump: texture " << k << " holds " << px.size() << " texels for " << t.width << " x "
                      << t.height << '\n';
            return 1;
        }
        for (const yr::Color4f& c : px) texels.insert(texels.end(), {c.x, c.y, c.z, c.w});
    }
    // 24-byte records: width, height, filter, wrap_s, wrap_t, reserved.
    std::vector<int32_t> tex_rec;
    for (const yrh_texture& t : tex) tex_rec.insert(tex_rec.end(), {t.width, t.height, t.filter, t.wrap_s, t.wrap_t, 0});
    Section(f, "textures", tex_rec.data(), static_cast<int64_t>(tex.size()), 24);
    Section(f, "texels", texels.data(), static_cast<int64_t>(texels.size()), 4);
    int32_t env_map[4] = {ir.environment.active ? 1 : 0, ir.environment.texture_index, 0, 0};
    std::memcpy(&env_map[2], &ir.environment.rotation_radians, 4);
    Section(f, "env_map", env_map, 4, 4);
    Section(f, "end", nullptr, 0, 0);
    std::cout << "dumped " << ir.vertices.size() << " vertices, " << ir.spheres.size() << " spheres, "
              << ir.analytic_lights.size() << " analytic lights, " << ir.textures.size() << " textures\n";
    return f ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "render") == 0) return Render(argc, argv);
    if (argc >= 2 && std::strcmp(argv[1], "checkpoint") == 0) return Checkpoint(argc, argv);
    if (argc >= 2 && std::strcmp(argv[1], "dump") == 0) return Dump(argc, argv);
    std::cerr << "usage: dropin_probe render <scene.pbrt> <cpu|hip> <out.bin> [spp] [max_depth]\n"
                 "       dropin_probe checkpoint <scene.pbrt> <cpu|hip> <ckpt> <out.bin> <passes>\n"
                 "       dropin_probe dump <scene.pbrt> <out.bin>\n";
    return 2;
}
