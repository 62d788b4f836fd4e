#include "scenes.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <utility>

namespace yrhip {

using namespace yr;

namespace {
constexpr float kPi = 3.14159265358979323846f;
}

Pcg32::Pcg32(std::uint64_t seed) {
    state_ = 0u;
    inc_ = (seed << 1u) | 1u;
    (void)NextU32();
    state_ += seed;
    (void)NextU32();
}

std::uint32_t Pcg32::NextU32() {
    const std::uint64_t old = state_;
    state_ = old * 6364136223846793005ULL + inc_;
    const std::uint32_t xorshifted = static_cast<std::uint32_t>(((old >> 18u) ^ old) >> 27u);
    const std::uint32_t rot = static_cast<std::uint32_t>(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}

float Pcg32::NextFloat() { return static_cast<float>(NextU32() >> 8) * (1.0f / 16777216.0f); }

Vec3f UniformSphere(float u, float v) {
    const float z = 1.0f - 2.0f * u;
    const float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
    const float phi = 2.0f * kPi * v;
    return Vec3f{r * std::cos(phi), r * std::sin(phi), z};
}

RenderCamera LookAtCamera(Point3f eye, Point3f target, Vec3f up, float fov_y_degrees) {
    RenderCamera c;
    c.origin = eye;
    c.forward = Normalize(target - eye);
    c.right = Normalize(Cross(c.forward, Normalize(up)));
    c.up = Cross(c.right, c.forward);
    c.fov_y_radians = fov_y_degrees * kPi / 180.0f;
    return c;
}

void CameraRay(const RenderCamera& cam, int width, int height, int x, int y, float sx, float sy, Point3f& origin,
               Vec3f& direction) {
    const float w = static_cast<float>(std::max(1, width));
    const float h = static_cast<float>(std::max(1, height));
    const float fx = static_cast<float>(x) + sx;
    const float fy = static_cast<float>(y) + sy;
    const float aspect = w / h;
    const float half_height = std::tan(cam.fov_y_radians * 0.5f);
    const float screen_x = (2.0f * fx / w - 1.0f) * aspect * half_height;
    const float screen_y = (1.0f - 2.0f * fy / h) * half_height;
    origin = cam.origin;
    direction = Normalize(cam.forward + cam.right * screen_x + cam.up * screen_y);
}

namespace {

struct MeshBuilder {
    RenderSceneIR& s;
    int BeginPrimitive(int material) {
        RenderPrimitive p;
        p.first_index = static_cast<std::uint32_t>(s.indices.size());
        p.material_index = material;
        s.primitives.push_back(p);
        return static_cast<int>(s.primitives.size()) - 1;
    }
    std::uint32_t Vertex(Point3f p, Vec3f n = {}) {
        RenderVertex v;
        v.position = p;
        v.normal = n;
        s.vertices.push_back(v);
        return static_cast<std::uint32_t>(s.vertices.size() - 1);
    }
    void Tri(std::uint32_t a, std::uint32_t b, std::uint32_t c) {
        s.indices.push_back(a);
        s.indices.push_back(b);
        s.indices.push_back(c);
        s.primitives.back().index_count += 3;
    }
    // Quad a-b-c-d (counter-clockwise seen from the front side).
    void Quad(Point3f a, Point3f b, Point3f c, Point3f d) {
        const std::uint32_t ia = Vertex(a), ib = Vertex(b), ic = Vertex(c), id = Vertex(d);
        Tri(ia, ib, ic);
        Tri(ia, ic, id);
    }
    // Axis-aligned box of half size h, rotated about +y by yaw, centred at c.
    void Box(Point3f c, Vec3f h, float yaw_degrees) {
        const Mat4f rot = RotationAxisMatrix(yaw_degrees, Vec3f{0, 1, 0});
        auto P = [&](float sx, float sy, float sz) {
            return TransformPoint(rot, Point3f{sx * h.x, sy * h.y, sz * h.z}) + c;
        };
        Quad(P(-1, -1, 1), P(1, -1, 1), P(1, 1, 1), P(-1, 1, 1));      // +z
        Quad(P(1, -1, -1), P(-1, -1, -1), P(-1, 1, -1), P(1, 1, -1));  // -z
        Quad(P(1, -1, 1), P(1, -1, -1), P(1, 1, -1), P(1, 1, 1));      // +x
        Quad(P(-1, -1, -1), P(-1, -1, 1), P(-1, 1, 1), P(-1, 1, -1));  // -x
        Quad(P(-1, 1, 1), P(1, 1, 1), P(1, 1, -1), P(-1, 1, -1));      // +y
        Quad(P(-1, -1, -1), P(1, -1, -1), P(1, -1, 1), P(-1, -1, 1));  // -y
    }
};

float TriangleArea(Point3f a, Point3f b, Point3f c) { return Length(Cross(b - a, c - a)) * 0.5f; }

// Unit icosphere subdivided `subdivisions` times, scaled by `radius` about
// `centre`, with its smooth normals.
void AddIcosphere(MeshBuilder& mb, int subdivisions, Point3f centre = {}, float radius = 1.0f) {
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    std::vector<Vec3f> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                            {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (Vec3f& p : v) p = Normalize(p);
    std::vector<std::array<int, 3>> f = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                         {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                         {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                         {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    for (int s = 0; s < subdivisions; ++s) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            const auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            v.push_back(Normalize((v[static_cast<size_t>(a)] + v[static_cast<size_t>(b)]) * 0.5f));
            const int idx = static_cast<int>(v.size()) - 1;
            mid.emplace(key, idx);
            return idx;
        };
        std::vector<std::array<int, 3>> nf;
        nf.reserve(f.size() * 4);
        for (const auto& tri : f) {
            const int a = midpoint(tri[0], tri[1]), b = midpoint(tri[1], tri[2]), c = midpoint(tri[2], tri[0]);
            nf.push_back({tri[0], a, c});
            nf.push_back({tri[1], b, a});
            nf.push_back({tri[2], c, b});
            nf.push_back({a, b, c});
        }
        f.swap(nf);
    }
    const std::uint32_t base = static_cast<std::uint32_t>(mb.s.vertices.size());
    for (const Vec3f& p : v) mb.Vertex(centre + p * radius, p);
    for (const auto& tri : f) {
        mb.Tri(base + static_cast<std::uint32_t>(tri[0]), base + static_cast<std::uint32_t>(tri[1]),
               base + static_cast<std::uint32_t>(tri[2]));
    }
    mb.s.primitives.back().has_normals = true;
}

Color3f SkyRadiance(Vec3f d) {
    // Procedural sky: horizon-to-zenith gradient above, dark ground below, and
    // a sun disc of 2 degrees radius.
    const Vec3f sun = Normalize(Vec3f{0.4f, 0.6f, -0.5f});
    Color3f c;
    if (d.y >= 0.0f) {
        const float k = std::pow(d.y, 0.5f);
        c = Color3f{0.85f, 0.9f, 1.0f} * (1.0f - k) + Color3f{0.25f, 0.45f, 0.95f} * k;
    } else {
        c = Color3f{0.12f, 0.11f, 0.1f};
    }
    const float cos_sun = Dot(d, sun);
    if (cos_sun > std::cos(2.0f * kPi / 180.0f)) c = c + Color3f{200.0f, 190.0f, 170.0f};
    return c;
}

}  // namespace

RenderSceneIR MakeCornellScene(std::uint64_t material_seed) {
    RenderSceneIR s;
    Pcg32 rng(material_seed);
    for (int m = 0; m < 8; ++m) {
        RenderMaterial mat;
        mat.kind = RenderMaterialKind::Diffuse;
        const float r = 0.1f + 0.7f * rng.NextFloat();
        const float g = 0.1f + 0.7f * rng.NextFloat();
        const float b = 0.1f + 0.7f * rng.NextFloat();
        mat.reflectance.value = Color3f{r, g, b};
        s.materials.push_back(mat);
    }
    const float yaw_tall = -30.0f + 60.0f * rng.NextFloat();
    const float yaw_short = -30.0f + 60.0f * rng.NextFloat();
    MeshBuilder mb{s};
    mb.BeginPrimitive(0);  // floor y = 0
    mb.Quad({-1, 0, 1}, {1, 0, 1}, {1, 0, -1}, {-1, 0, -1});
    mb.BeginPrimitive(1);  // ceiling y = 2
    mb.Quad({-1, 2, -1}, {1, 2, -1}, {1, 2, 1}, {-1, 2, 1});
    mb.BeginPrimitive(2);  // back wall z = -1
    mb.Quad({-1, 0, -1}, {1, 0, -1}, {1, 2, -1}, {-1, 2, -1});
    mb.BeginPrimitive(3);  // left wall x = -1
    mb.Quad({-1, 0, 1}, {-1, 0, -1}, {-1, 2, -1}, {-1, 2, 1});
    mb.BeginPrimitive(4);  // right wall x = 1
    mb.Quad({1, 0, -1}, {1, 0, 1}, {1, 2, 1}, {1, 2, -1});
    mb.BeginPrimitive(5);
    mb.Box({-0.35f, 0.6f, -0.3f}, {0.3f, 0.6f, 0.3f}, yaw_tall);
    mb.BeginPrimitive(6);
    mb.Box({0.35f, 0.3f, 0.3f}, {0.3f, 0.3f, 0.3f}, yaw_short);
    // Ceiling emitter 1 cm below the ceiling, facing down.
    s.materials[7].emission = Color3f{17.0f, 17.0f, 17.0f};
    const int light = mb.BeginPrimitive(7);
    const Point3f a{-0.25f, 1.99f, 0.25f}, b{0.25f, 1.99f, 0.25f}, c{0.25f, 1.99f, -0.25f}, d{-0.25f, 1.99f, -0.25f};
    mb.Quad(a, b, c, d);
    s.emissive_primitives.push_back(
        EmissivePrimitive{light, Color3f{17.0f, 17.0f, 17.0f}, TriangleArea(a, b, c) + TriangleArea(a, c, d)});
    s.camera = LookAtCamera({0.0f, 1.0f, 3.5f}, {0.0f, 1.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, 40.0f);
    PrepareLightSampling(s);
    return s;
}

RenderSceneIR MakeRandomTriangleScene(int count, std::uint64_t seed) {
    RenderSceneIR s;
    s.materials.push_back(RenderMaterial{});
    Pcg32 rng(seed);
    s.vertices.reserve(static_cast<size_t>(count) * 3);
    s.indices.reserve(static_cast<size_t>(count) * 3);
    const float ln_lo = std::log(1.0e-3f);
    const float ln_hi = std::log(2.0e-2f);
    RenderPrimitive p;
    p.first_index = 0;
    p.index_count = static_cast<std::uint32_t>(count) * 3u;
    p.material_index = 0;
    for (int t = 0; t < count; ++t) {
        const float cx = rng.NextFloat(), cy = rng.NextFloat(), cz = rng.NextFloat();
        for (int k = 0; k < 3; ++k) {
            const Vec3f d = UniformSphere(rng.NextFloat(), rng.NextFloat());
            const float len = std::exp(ln_lo + (ln_hi - ln_lo) * rng.NextFloat());
            RenderVertex v;
            v.position = Point3f{cx, cy, cz} + d * len;
            s.vertices.push_back(v);
            s.indices.push_back(static_cast<std::uint32_t>(s.vertices.size() - 1));
        }
    }
    s.primitives.push_back(p);
    s.camera = LookAtCamera({0.5f, 0.5f, -1.5f}, {0.5f, 0.5f, 0.5f}, {0.0f, 1.0f, 0.0f}, 40.0f);
    return s;
}

void AddSkyEnvironment(RenderSceneIR& s, int env_width, int env_height);

RenderSceneIR MakeInstancedScene(int instance_count, int subdivisions, int variants, std::uint64_t seed, int env_width,
                                 int env_height) {
    RenderSceneIR s;
    Pcg32 rng(seed);
    MeshBuilder mb{s};
    // Primitive 0: ground, grey diffuse.
    RenderMaterial ground;
    ground.reflectance.value = Color3f{0.5f, 0.5f, 0.5f};
    s.materials.push_back(ground);
    mb.BeginPrimitive(0);
    mb.Quad({-100, 0, 100}, {100, 0, 100}, {100, 0, -100}, {-100, 0, -100});
    // Icosphere variants, one material each: diffuse / GGX conductor / smooth dielectric.
    for (int v = 0; v < variants; ++v) {
        RenderMaterial m;
        const int kind = v % 3;
        if (kind == 0) {
            m.kind = RenderMaterialKind::Diffuse;
            m.reflectance.value = Color3f{0.1f + 0.8f * rng.NextFloat(), 0.1f + 0.8f * rng.NextFloat(),
                                    0.1f + 0.8f * rng.NextFloat()};
        } else if (kind == 1) {
            m.kind = RenderMaterialKind::Conductor;
            m.reflectance.value = Color3f{0.5f + 0.5f * rng.NextFloat(), 0.5f + 0.5f * rng.NextFloat(),
                                    0.5f + 0.5f * rng.NextFloat()};
            const float alpha = 0.05f + 0.45f * rng.NextFloat();
            m.uroughness.value = std::sqrt(alpha);
            m.vroughness.value = m.uroughness.value;
        } else {
            m.kind = RenderMaterialKind::Dielectric;
            m.reflectance.value = Color3f{1.0f, 1.0f, 1.0f};
            m.ior = 1.5f;
        }
        s.materials.push_back(m);
        mb.BeginPrimitive(static_cast<int>(s.materials.size()) - 1);
        AddIcosphere(mb, subdivisions);
    }
    // Ground as an explicit identity instance (instances are non-empty).
    s.instances.push_back(RenderInstance{0, Mat4f{}});
    for (int i = 0; i < instance_count; ++i) {
        const int prim = 1 + static_cast<int>(rng.NextU32() % static_cast<std::uint32_t>(std::max(1, variants)));
        const float scale = 0.1f + 0.9f * rng.NextFloat();
        const Vec3f axis = UniformSphere(rng.NextFloat(), rng.NextFloat());
        const float angle = 360.0f * rng.NextFloat();
        const float x = -30.0f + 60.0f * rng.NextFloat();
        const float z = -30.0f + 60.0f * rng.NextFloat();
        const float y = scale + 3.0f * rng.NextFloat();
        const Mat4f o2w = Multiply(TranslationMatrix({x, y, z}),
                                   Multiply(RotationAxisMatrix(angle, axis), ScaleMatrix({scale, scale, scale})));
        s.instances.push_back(RenderInstance{prim, o2w});
    }
    AddSkyEnvironment(s, env_width, env_height);
    s.camera = LookAtCamera({0.0f, 6.0f, 40.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, 40.0f);
    PrepareLightSampling(s);
    return s;
}

// Procedural sky + sun (equirect, u from atan2(x, z), v = theta / pi) as the
// scene's environment, with its sampling distribution.
void AddSkyEnvironment(RenderSceneIR& s, int env_width, int env_height) {
    if (env_width > 0 && env_height > 0) {
        RenderTexture tex;
        tex.width = env_width;
        tex.height = env_height;
        tex.texels.resize(static_cast<size_t>(env_width) * static_cast<size_t>(env_height));
        for (int y = 0; y < env_height; ++y) {
            const float theta = (static_cast<float>(y) + 0.5f) / static_cast<float>(env_height) * kPi;
            for (int x = 0; x < env_width; ++x) {
                const float phi = ((static_cast<float>(x) + 0.5f) / static_cast<float>(env_width) - 0.5f) * 2.0f * kPi;
                const Vec3f d{std::sin(theta) * std::sin(phi), std::cos(theta), std::sin(theta) * std::cos(phi)};
                const Color3f c = SkyRadiance(d);
                tex.texels[static_cast<size_t>(y) * static_cast<size_t>(env_width) + static_cast<size_t>(x)] =
                    Color4f{c.x, c.y, c.z, 1.0f};
            }
        }
        s.textures.push_back(std::move(tex));
        s.environment_distributions.push_back(BuildEnvironmentDistribution(s.textures.back()));
        s.environment.active = true;
        s.environment.radiance = Color3f{1.0f, 1.0f, 1.0f};
        s.environment.strength = 1.0f;
        s.environment.texture_index = static_cast<int>(s.textures.size()) - 1;
        s.environment.distribution_index = static_cast<int>(s.environment_distributions.size()) - 1;
    }
}

RenderSceneIR MakeGuidingRoomScene(int furniture_triangles, std::uint64_t seed, int env_width, int env_height) {
    RenderSceneIR s;
    Pcg32 rng(seed);
    MeshBuilder mb{s};
    auto diffuse = [&](Color3f c) {
        RenderMaterial m;
        m.kind = RenderMaterialKind::Diffuse;
        m.reflectance.value = c;
        s.materials.push_back(m);
        return static_cast<int>(s.materials.size()) - 1;
    };
    const int walls = diffuse({0.75f, 0.75f, 0.72f});
    const int floor = diffuse({0.55f, 0.5f, 0.45f});
    const int ceiling = diffuse({0.8f, 0.8f, 0.8f});
    const int blocker = diffuse({0.6f, 0.6f, 0.6f});
    RenderMaterial lamp;
    lamp.reflectance.value = Color3f{0.0f, 0.0f, 0.0f};
    lamp.emission = Color3f{30.0f, 28.0f, 25.0f};
    s.materials.push_back(lamp);
    const int lamp_mat = static_cast<int>(s.materials.size()) - 1;
    // Furniture materials: diffuse and GGX conductors (alpha in [0.05, 0.5]).
    std::vector<int> furniture;
    for (int k = 0; k < 8; ++k) {
        if (k % 2 == 0) {
            furniture.push_back(diffuse({0.15f + 0.7f * rng.NextFloat(), 0.15f + 0.7f * rng.NextFloat(),
                                         0.15f + 0.7f * rng.NextFloat()}));
        } else {
            RenderMaterial m;
            m.kind = RenderMaterialKind::Conductor;
            m.reflectance.value = Color3f{0.5f + 0.5f * rng.NextFloat(), 0.5f + 0.5f * rng.NextFloat(),
                                    0.5f + 0.5f * rng.NextFloat()};
            const float alpha = 0.05f + 0.45f * rng.NextFloat();
            m.uroughness.value = std::sqrt(alpha);
            m.vroughness.value = m.uroughness.value;
            s.materials.push_back(m);
            furniture.push_back(static_cast<int>(s.materials.size()) - 1);
        }
    }
    // Room 10 x 10 x 4: floor, ceiling, three solid walls and a wall at x = 10
    // with a 1 x 1 window (y in [2, 3], z in [4.5, 5.5]).
    mb.BeginPrimitive(floor);
    mb.Quad({0, 0, 10}, {10, 0, 10}, {10, 0, 0}, {0, 0, 0});
    mb.BeginPrimitive(ceiling);
    mb.Quad({0, 4, 0}, {10, 4, 0}, {10, 4, 10}, {0, 4, 10});
    mb.BeginPrimitive(walls);
    mb.Quad({0, 0, 0}, {10, 0, 0}, {10, 4, 0}, {0, 4, 0});
    mb.Quad({10, 0, 10}, {0, 0, 10}, {0, 4, 10}, {10, 4, 10});
    mb.Quad({0, 0, 10}, {0, 0, 0}, {0, 4, 0}, {0, 4, 10});
    mb.Quad({10, 0, 0}, {10, 0, 10}, {10, 2, 10}, {10, 2, 0});
    mb.Quad({10, 3, 0}, {10, 3, 10}, {10, 4, 10}, {10, 4, 0});
    mb.Quad({10, 2, 0}, {10, 2, 4.5f}, {10, 3, 4.5f}, {10, 3, 0});
    mb.Quad({10, 2, 5.5f}, {10, 2, 10}, {10, 3, 10}, {10, 3, 5.5f});
    // Ceiling lamp hidden behind a panel hanging below it: it lights the room
    // only through the ceiling and the gap around the panel.
    const int light = mb.BeginPrimitive(lamp_mat);
    const Point3f la{2, 3.98f, 8}, lb{3, 3.98f, 8}, lc{3, 3.98f, 7}, ld{2, 3.98f, 7};
    mb.Quad(la, lb, lc, ld);
    s.emissive_primitives.push_back(
        EmissivePrimitive{light, lamp.emission, TriangleArea(la, lb, lc) + TriangleArea(la, lc, ld)});
    mb.BeginPrimitive(blocker);
    mb.Box({2.5f, 3.6f, 7.5f}, {0.9f, 0.02f, 0.9f}, 0.0f);
    // Tables (top + four legs) and spheres on the floor and on the tables, up
    // to `furniture_triangles` triangles.
    int tris = 0;
    std::vector<Point3f> tops;
    for (int t = 0; t < 10 && tris < furniture_triangles; ++t) {
        const Point3f c{1.5f + 7.0f * rng.NextFloat(), 0.0f, 1.5f + 7.0f * rng.NextFloat()};
        const float yaw = 90.0f * rng.NextFloat();
        const Vec3f half{0.5f + 0.4f * rng.NextFloat(), 0.03f, 0.35f + 0.3f * rng.NextFloat()};
        mb.BeginPrimitive(furniture[static_cast<size_t>(t) % furniture.size()]);
        mb.Box({c.x, 0.75f, c.z}, half, yaw);
        const Mat4f rot = RotationAxisMatrix(yaw, Vec3f{0, 1, 0});
        for (int l = 0; l < 4; ++l) {
            const float sx = (l & 1) ? 1.0f : -1.0f, sz = (l & 2) ? 1.0f : -1.0f;
            const Point3f off = TransformPoint(rot, Point3f{sx * (half.x - 0.05f), 0.0f, sz * (half.z - 0.05f)});
            mb.Box({c.x + off.x, 0.36f, c.z + off.z}, {0.03f, 0.36f, 0.03f}, yaw);
        }
        tops.push_back(Point3f{c.x, 0.78f, c.z});
        tris += 60;
    }
    while (tris < furniture_triangles) {
        const float radius = 0.15f + 0.35f * rng.NextFloat();
        Point3f c;
        if (!tops.empty() && rng.NextFloat() < 0.4f) {
            const Point3f& top = tops[rng.NextU32() % static_cast<std::uint32_t>(tops.size())];
            c = Point3f{top.x + 0.3f * (rng.NextFloat() - 0.5f), top.y + radius, top.z + 0.3f * (rng.NextFloat() - 0.5f)};
        } else {
            c = Point3f{0.8f + 8.4f * rng.NextFloat(), radius, 0.8f + 8.4f * rng.NextFloat()};
        }
        mb.BeginPrimitive(furniture[rng.NextU32() % static_cast<std::uint32_t>(furniture.size())]);
        AddIcosphere(mb, 4, c, radius);
        tris += 5120;
    }
    AddSkyEnvironment(s, env_width, env_height);
    s.camera = LookAtCamera({1.0f, 1.6f, 1.0f}, {6.0f, 1.2f, 6.5f}, {0.0f, 1.0f, 0.0f}, 65.0f);
    PrepareLightSampling(s);
    return s;
}

}  // namespace yrhip
