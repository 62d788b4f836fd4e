// Wavefront path tracer for MI355X (gfx950): the GPU counterpart of
// yr::RenderCpuPathTrace (src/backends/cpu/cpu_path_tracer.cpp:39-229) driving
// yr::IntegratePathSample (src/integrators/path_integrator.cpp:23-197).
//
// One pass = one sample for every pixel. Per bounce, with every count kept in
// device memory (no host synchronisation inside a pass):
//   trace    persistent closest-hit kernel over the compacted ray queue
//   shade    per path: miss/environment, emission x MIS, next-event estimation
//            (emissive + environment samples -> shadow queue), BSDF sampling,
//            Russian roulette, next ray appended to the other queue
//   shadow   any-hit kernel over the shadow queue (opaque scenes), or the
//            reference's transparent-shadow loop (dielectric scenes)
//   resolve  per path: adds the visible NEE contributions in reference order
// and finally `accumulate` adds each path's sample to the FP64 film.
// Path state is SoA in HBM; queues are appended with one atomic per wave.
#include "yrh_internal.h"
#include "yrh_guide.h"
#include "yrh_shading.h"
#include "yrh_texture.h"
#include "yrh_traverse.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#pragma clang fp contract(off)

namespace yrh {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxTransparentShadowHits = 16;  // src/integrators/path_direct_lighting.cpp:19
constexpr int kMaxAlphaSkips = 64;              // src/integrators/surface_query.cpp:12
constexpr float kAlphaSkipOffset = 1.0e-4f;     // surface_query.cpp:13
// prim_flags bits
constexpr int kHasNormals = 1;
constexpr int kHasUvs = 2;
constexpr int kHasTangents = 4;

// Counter words of one pass (uint32).
enum Counter : int {
    kCountQueue0 = 0,
    kCountQueue1 = 1,
    kCountShadow = 2,
    kCountPassPaths = 3,  // paths of the current pass (batched passes)
    kPassIndex = 4,       // current pass of a batch schedule (advanced by EndPassKernel)
    // Per-launch ray counters of the trace kernels: one per claim range
    // (trav::kParts ranges, trav::kPartStride words apart).
    kWorkExtend = 32,
    kWorkShadow = kWorkExtend + trav::kParts * trav::kPartStride,
    kCounterWords = kWorkShadow + trav::kParts * trav::kPartStride,
};

// 64-bit statistics words.
enum Stat : int {
    kStatRays = 0,
    kStatShadow,
    kStatOccluded,
    kStatTriTests,
    kStatSphereTests,
    kStatNodeTests,
    kStatHits,
    kStatMisses,
    kStatBad,
    kStatTraceNodes,  // scratch totals written by the trace kernels: [node, tri, sphere]
    kStatTraceTris,
    kStatTraceSpheres,
    kStatWords,
};
// The statistics live in kStatSlots copies kStatStride words (128 B) apart;
// workgroup b adds to copy b % kStatSlots and the readback sums them: one
// counter word for every wave of a 100-million-path launch made its atomics
// a quarter of the render time.
constexpr int kStatSlots = 64;
constexpr int kStatStride = 16;
static_assert(kStatWords <= kStatStride, "statistics words fit a slot");
constexpr size_t kStatBytes = sizeof(unsigned long long) * kStatStride * kStatSlots;

}  // namespace

struct RenderState {
    // ---- surface geometry (shading copy) ----
    DevBuf positions;    // float4 per vertex
    DevBuf normals;      // float4 per vertex
    DevBuf indices;      // uint32
    DevBuf tri_prim;     // int32 per flat triangle: mesh primitive
    DevBuf spheres;      // float4 per sphere: centre, radius
    DevBuf sphere_info;  // int2 per sphere: material index, flip_normals
    DevBuf prim_info;    // int4 per primitive: material, emissive index, first index, flat offset
    DevBuf prim_flags;   // int32 per primitive: kHasNormals | kHasUvs | kHasTangents
    DevBuf uvs;          // float2 per vertex (when a primitive has uvs)
    DevBuf tangents;     // float4 per vertex: tangent, handedness (when a primitive has tangents)
    // float4[6] per instance: normal-matrix rows (.w: 1 / det, singular flag,
    // orientation sign), object-to-world rows (.w: translation)
    DevBuf inst_normal;
    int32_t primitive_count = 0;
    int64_t triangle_count = 0;
    int64_t vertex_count = 0;
    std::vector<int32_t> host_prim_first;  // first index per primitive (emitter tables)
    std::vector<int32_t> host_prim_count;
    std::vector<float> host_positions;     // xyz per vertex (emitter tables)
    std::vector<uint32_t> host_indices;
    std::vector<int32_t> host_prim_material;
    std::vector<int32_t> host_sphere_material;
    // ---- shading ----
    bool shading_loaded = false;
    bool transparent = false;  // any dielectric / thin dielectric material
    DevBuf materials;          // DevMaterial
    int32_t material_count = 0;
    // textures (yrh_texture.h layout) and per material (reflectance texture,
    // alpha texture, normal map, normal scale bits); texture indices are -1
    // where the material has none
    DevBuf tex_records, tex_levels, tex_texels, tex_gauss, mat_tex;
    int32_t texture_count = 0;
    bool textured = false;    // a material reads a texture
    bool alpha = false;       // a material's hits may be alpha-skipped
    bool footprints = false;  // a material's texture filters with camera-ray footprints
    DevBuf emit;               // float4 (radiance, stored area)
    DevBuf emit_info;          // int4 (primitive, cdf offset, cdf count, 0)
    DevBuf emit_area;          // float: summed triangle area (SelectTriangleByArea)
    DevBuf emit_cdf;           // float: running triangle areas per emitter
    DevBuf alias_prob, alias_index, alias_pmf;
    int32_t emissive_count = 0;
    float total_emissive_area = 0.0f;
    // analytic lights: three float4 per light (position + kind bits,
    // direction + cone_angle, intensity + cone_cos_inner), the power alias
    // table, and the light tree (three float4 per node: bounds min + power,
    // bounds max, (left, right, light index) bits)
    DevBuf alight, alight_prob, alight_alias, alight_pmf, alight_tree;
    int32_t analytic_count = 0;
    int32_t analytic_pmf_count = 0;
    int32_t analytic_tree_count = 0;
    int32_t analytic_tree_root = -1;
    int32_t use_analytic_tree = 0;
    // Shadow-queue entries per path slot: emissive, analytic (scenes with
    // analytic lights), environment.
    int shadow_slots = 2;
    int cap_shadow_slots = 0;
    // environment
    int32_t env_active = 0;
    float env_radiance[3] = {0, 0, 0};
    float env_strength = 1.0f;
    float env_rotation = 0.0f;
    DevBuf env_texels;  // float4
    int32_t env_w = 0, env_h = 0;
    int32_t env_nearest = 0, env_ws = 0, env_wt = 0;
    int32_t env_sampleable = 0;
    DevBuf env_weights, env_row_weights;
    // CDFs padded with +inf to a multiple of kCdfBlock entries (rows of the
    // conditional CDFs env_cstride apart), and every kCdfBlock-th entry of
    // each (padded with +inf to a multiple of 4) for the two-level search.
    DevBuf env_row_cdf, env_row_top, env_cond_cdf, env_cond_top;
    int32_t env_rtop = 0, env_cstride = 0, env_ctop = 0;
    int32_t env_dw = 0, env_dh = 0;
    float env_total = 0.0f;
    // ---- film ----
    DevBuf film;  // yrh_film_pixel
    int32_t film_w = 0, film_h = 0;
    // Active pixels (yrh_film_update_active): the list and the words
    // [0] list length, [1] converged pixels; act_ready once a list matches the
    // film; act_desc: the adaptive settings pixel-list passes accept samples by.
    DevBuf act_list, act_count;
    bool act_ready = false;
    yrh_adaptive_desc act_desc{};
    // ---- pass buffers (capacity = paths per pass) ----
    int64_t capacity = 0;
    DevBuf p_pixel, p_tp, p_rad, p_prev, p_flags, p_med, p_alb, p_nrm, p_nee, p_slot;  // p_slot: int4 shadow entries
    DevBuf p_diff;  // float4[2] per path: the camera ray's differential directions (footprints only)
    DevBuf q_org[2], q_dir[2], q_path[2];
    DevBuf hits, hit_inst;
    DevBuf s_org, s_dir, s_li, s_bsdf, s_occ, s_trans;
    DevBuf l_state, l_med;
    DevBuf counters;  // uint32[kCounterWords]
    DevBuf stats;     // uint64[kStatStride * kStatSlots]: kStatSlots copies of the statistics words
    // batched passes
    DevBuf p_job, p_sample;
    DevBuf b_jobs, b_items, b_pass_items, b_film;
    std::vector<int64_t> b_job_pixels;  // per job: pixel count
    std::vector<int64_t> b_job_offset;  // per job: first pixel in b_film
};

namespace {

template <typename T>
int UploadVec(yrh_context* ctx, DevBuf& buf, const T* data, size_t count) {
    buf.reset();
    const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    int rc = CheckHip(ctx, hipMalloc(&buf.ptr, bytes), "hipMalloc");
    if (rc != YRH_OK) return rc;
    buf.bytes = bytes;
    if (count > 0) rc = yrh_copy_to_device(ctx, buf.ptr, data, count * sizeof(T));
    return rc;
}

int Alloc(yrh_context* ctx, DevBuf& buf, size_t bytes) {
    buf.reset();
    bytes = std::max<size_t>(bytes, 16);
    int rc = CheckHip(ctx, hipMalloc(&buf.ptr, bytes), "hipMalloc");
    if (rc == YRH_OK) buf.bytes = bytes;
    return rc;
}

RenderState& State(yrh_context* ctx) {
    if (ctx->render == nullptr) ctx->render = new RenderState();
    return *ctx->render;
}

// ---- kernel parameter blocks ------------------------------------------------------

struct SurfaceView {
    const float4* positions;
    const float4* normals;
    const uint32_t* indices;
    const int32_t* tri_prim;
    const int4* prim_info;
    const int32_t* prim_flags;
    const float4* inst_normal;
    const float4* spheres;     // centre, radius
    const int2* sphere_info;   // material index, flip_normals
    const float2* uvs;
    const float4* tangents;
};

struct LightView {
    const DevMaterial* materials;
    const int4* mat_tex;  // per material: reflectance texture, alpha texture, normal map, normal scale bits
    tex::TexView tex;
    const float4* emit;
    const int4* emit_info;
    const float* emit_area;
    const float* emit_cdf;
    const float* alias_prob;
    const int32_t* alias_index;
    const float* alias_pmf;
    int32_t emissive_count;
    float total_emissive_area;
    const float4* alight;
    const float* alight_prob;
    const int32_t* alight_alias;
    const float* alight_pmf;
    const float4* alight_tree;
    int32_t analytic_count;
    int32_t analytic_pmf_count;
    int32_t analytic_tree_count;
    int32_t analytic_tree_root;
    int32_t use_analytic_tree;
    int32_t env_active;
    float3 env_radiance;
    float env_strength;
    float env_rotation;
    const float4* env_texels;
    int32_t env_w, env_h;
    int32_t env_nearest, env_ws, env_wt;
    int32_t env_sampleable;
    const float* env_weights;
    const float* env_row_weights;
    const float* env_row_cdf;
    const float* env_row_top;
    const float* env_cond_cdf;
    const float* env_cond_top;
    int32_t env_rtop, env_cstride, env_ctop;
    int32_t env_dw, env_dh;
    float env_total;
    const float* positions_xyz;  // unused on device (kept for symmetry)
};

// Batched multi-camera passes: the camera jobs, and per pass a list of work
// items (job, sample index, pixel range) laid end to end over the pass's paths.
struct DevJob {
    DevCamera camera;
    int32_t width, height, spp;
    int64_t film_offset;  // first pixel of the job in the batch film
};

struct DevItem {
    int32_t job, sample, pixel_begin, pixel_count;
    int32_t occurrence;  // earlier items of the same job in this pass (their pixels may repeat)
    int32_t pad;
    int64_t path_begin;  // first path of the item within its pass
};

struct PassView {
    int32_t width, height;
    int32_t sampler, spp, max_depth, sample_index;
    int32_t sample_count;  // paths = npix * sample_count; path p renders sample sample_index + p / npix
    int64_t npix;
    uint64_t seed;
    float radiance_clamp;
    DevCamera camera;
    // path state
    int32_t* p_pixel;
    float4* p_tp;
    float4* p_rad;
    float4* p_prev;
    int32_t* p_flags;
    float4* p_med;
    float4* p_alb;
    float4* p_nrm;
    float4* p_nee;
    int4* p_slot;  // shadow-queue entries of the path's vertex: emissive, analytic, environment
    float4* p_diff;  // camera-ray differential directions (2 per path), or nullptr
    // queues
    float4* q_org[2];
    float4* q_dir[2];
    int32_t* q_path[2];
    float4* hits;
    int32_t* hit_inst;
    float4* s_org;
    float4* s_dir;
    float4* s_li;
    float4* s_bsdf;
    uint32_t* s_occ;
    float4* s_trans;
    int4* l_state;
    float4* l_med;
    uint32_t* counters;
    unsigned long long* stats;
    yrh_film_pixel* film;
    // batched passes (jobs == nullptr: one camera, width x height x sample_count paths)
    const DevJob* jobs;
    const DevItem* items;
    const int32_t* pass_items;  // items of pass k: [pass_items[k], pass_items[k + 1])
    int32_t* p_job;
    int32_t* p_sample;
    // pixel-list passes (adaptive sampling): sample_count paths per listed
    // pixel at its next sample indices, accepted in order until the pixel
    // converges (act_*: Film::IsConverged settings); nullptr otherwise
    const int32_t* active_list;
    const uint32_t* active_count;
    int32_t act_on, act_min_spp;
    float act_rel, act_abs, act_conf;
};

enum PathFlag : int { kPrevValid = 1, kPrevDelta = 2, kMediumActive = 4, kPrimaryHit = 8 };

__device__ __forceinline__ int Lane() { return static_cast<int>(threadIdx.x & 63u); }

// Appends `count` (0..2) entries per thread to a queue with one atomic per
// workgroup (all threads must call; uses `lds`, kBlock / 64 + 1 words);
// returns the first slot of this thread's entries, consecutive, or -1. The
// queue counters see one atomic per workgroup instead of one per wave and
// entry kind, which on 100-million-path launches contended on one word.
__device__ __forceinline__ int32_t BlockAppend(uint32_t* counter, uint32_t count, uint32_t* lds) {
    constexpr int kWaves = kBlock / 64;
    const int lane = Lane();
    const int wave = static_cast<int>(threadIdx.x >> 6);
    uint32_t inc = count;  // inclusive prefix over the wave
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t up = __shfl_up(inc, off);
        if (lane >= off) inc += up;
    }
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int w = 0; w < kWaves; ++w) total += lds[w];
        lds[kWaves] = total != 0u ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    uint32_t base = lds[kWaves];
    for (int w = 0; w < wave; ++w) base += lds[w];
    __syncthreads();  // lds is reused by the next call
    return count != 0u ? static_cast<int32_t>(base + inc - count) : -1;
}

// Wave-summed statistics word `word`, one atomic per wave into this
// workgroup's copy (all lanes must call).
__device__ __forceinline__ void WaveStat(unsigned long long* stats, int word, unsigned v) {
    unsigned long long s = v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (Lane() == 0 && s != 0) atomicAdd(stats + word + kStatStride * (blockIdx.x % kStatSlots), s);
}

__device__ __forceinline__ float3 XYZ(float4 v) { return V(v.x, v.y, v.z); }
__device__ __forceinline__ float4 F4(float3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

__device__ __forceinline__ float MaxC(float3 c) {
    // std::max(c.x, std::max(c.y, c.z))
    const float yz = c.y < c.z ? c.z : c.y;
    return c.x < yz ? yz : c.x;
}

__device__ __forceinline__ DevMaterial LoadMaterial(const DevMaterial* m, int idx) { return m[idx]; }

// TransformNormal (src/core/transform.cpp:46-73) with host-precomputed cofactors.
__device__ __forceinline__ float3 TransformNormalDev(const float4* rows, float3 n) {
    const float4 r0 = rows[0], r1 = rows[1], r2 = rows[2];
    if (r1.w != 0.0f) {
        const float4 m0 = rows[3], m1 = rows[4], m2 = rows[5];
        const float3 fb = Normalize(V(m0.x * n.x + m0.y * n.y + m0.z * n.z, m1.x * n.x + m1.y * n.y + m1.z * n.z,
                                      m2.x * n.x + m2.y * n.y + m2.z * n.z));
        return LengthSq(fb) > 0.0f ? fb : Normalize(n);
    }
    const float inv = r0.w;
    return Normalize(V((r0.x * n.x + r0.y * n.y + r0.z * n.z) * inv, (r1.x * n.x + r1.y * n.y + r1.z * n.z) * inv,
                       (r2.x * n.x + r2.y * n.y + r2.z * n.z) * inv));
}

// ---- environment (src/lighting/environment.cpp:54-219) ------------------------------

__device__ __forceinline__ float Fract(float v) { return v - floorf(v); }

__device__ __forceinline__ float2 DirToUv(float3 d, float rot) {
    d = Normalize(d);
    const float phi = atan2f(d.x, d.z);
    const float u = Fract(0.5f + (phi + rot) / kTwoPi);
    const float theta = acosf(Clampf(d.y, -1.0f, 1.0f));
    return make_float2(u, Clampf(theta / kPi, 0.0f, 1.0f));
}

__device__ __forceinline__ float3 UvToDir(float2 uv, float rot) {
    const float phi = (uv.x - 0.5f) * kTwoPi - rot;
    const float theta = Clampf(uv.y, 0.0f, 1.0f) * kPi;
    const float st = sinf(theta);
    return Normalize(V(st * sinf(phi), cosf(theta), st * cosf(phi)));
}

// SampleTexture without a footprint on the map (src/shading/texture.cpp:107-130,
// 312-322): bilinear, or nearest, with the map's wrap modes.
__device__ __forceinline__ float3 EnvLookup(const LightView& L, float2 uv) {
    if (L.env_w <= 0 || L.env_h <= 0) return V(0, 0, 0);
    const tex::Level l{L.env_texels, L.env_w, L.env_h};
    return XYZ(L.env_nearest ? tex::Nearest(l, uv.x, uv.y, L.env_ws, L.env_wt)
                             : tex::Bilinear(l, uv.x, uv.y, L.env_ws, L.env_wt));
}

__device__ __forceinline__ float3 EnvColor(const LightView& L, float3 d) {
    if (!L.env_active) return Mul(L.env_radiance, L.env_strength);
    if (L.env_strength <= 0.0f || L.env_texels == nullptr) return V(0, 0, 0);
    const float3 c = EnvLookup(L, DirToUv(d, L.env_rotation));
    return Mul(Hadamard(c, L.env_radiance), L.env_strength);
}

__device__ __forceinline__ float EnvPdf(const LightView& L, float3 d) {
    if (!L.env_sampleable) return 0.0f;
    const int w = L.env_dw, h = L.env_dh;
    if (w <= 0 || h <= 0 || L.env_total <= 0.0f) return 0.0f;
    const float2 uv = DirToUv(d, L.env_rotation);
    int x = static_cast<int>(floorf(uv.x * static_cast<float>(w)));
    int y = static_cast<int>(floorf(uv.y * static_cast<float>(h)));
    x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
    y = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
    const float prob = L.env_weights[static_cast<int64_t>(y) * w + x] / L.env_total;
    const float st = fmaxf(1.0e-6f, sinf((static_cast<float>(y) + 0.5f) * kPi / static_cast<float>(h)));
    const float texel_sa = (kTwoPi / static_cast<float>(w)) * (kPi / static_cast<float>(h)) * st;
    return prob / texel_sa;
}

// std::lower_bound over a non-decreasing float array, in two rounds of
// independent loads instead of log2(n) dependent ones. `a` holds the n values
// and +inf up to a multiple of kCdfBlock; top[j] = a[kCdfBlock * j] for the
// ntop (a multiple of 4) block starts, +inf-padded. The index is the number
// of entries below the target: k block starts lie below it, so it falls in
// block k - 1, whose entries below the target are counted. (A finite target
// never counts the +inf padding; the caller clamps n to n - 1 as the
// reference does for a target above every entry.)
constexpr int kCdfBlock = 32;

__device__ __forceinline__ int Below4(const float4& v, float t) {
    return static_cast<int>(v.x < t) + static_cast<int>(v.y < t) + static_cast<int>(v.z < t) + static_cast<int>(v.w < t);
}

__device__ __forceinline__ int BlockLowerBound(const float* __restrict__ a, const float* __restrict__ top, int ntop,
                                               float target) {
    const float4* t4 = reinterpret_cast<const float4*>(top);
    int k = 0;
#pragma unroll 4
    for (int j = 0; j < ntop / 4; ++j) k += Below4(t4[j], target);
    if (k == 0) return 0;
    const int base = kCdfBlock * (k - 1);
    const float4* a4 = reinterpret_cast<const float4*>(a + base);
    int c = 0;
#pragma unroll
    for (int j = 0; j < kCdfBlock / 4; ++j) c += Below4(a4[j], target);
    return base + c;
}

struct EnvSample {
    float3 dir;
    float3 radiance;
    float pdf;
    bool valid;
};

__device__ __forceinline__ EnvSample SampleEnv(const LightView& L, float2 s) {
    EnvSample out;
    out.valid = false;
    out.pdf = 0.0f;
    if (!L.env_sampleable) return out;
    const int w = L.env_dw, h = L.env_dh;
    if (w <= 0 || h <= 0 || L.env_total <= 0.0f) return out;
    const float row_target = Clampf(s.y, 0.0f, kOneMinusEps) * L.env_total;
    int y = BlockLowerBound(L.env_row_cdf, L.env_row_top, L.env_rtop, row_target);
    if (y == h) y = h - 1;
    const float row_prev = y == 0 ? 0.0f : L.env_row_cdf[y - 1];
    const float row_w = fmaxf(L.env_row_weights[y], 1.0e-6f);
    const float row_frac = Clampf((row_target - row_prev) / row_w, 0.0f, 1.0f);
    const float* cond = L.env_cond_cdf + static_cast<int64_t>(y) * L.env_cstride;
    const float col_target = Clampf(s.x, 0.0f, kOneMinusEps) * row_w;
    int x = BlockLowerBound(cond, L.env_cond_top + static_cast<int64_t>(y) * L.env_ctop, L.env_ctop, col_target);
    if (x == w) x = w - 1;
    const float col_prev = x == 0 ? 0.0f : cond[x - 1];
    const float tex_w = fmaxf(L.env_weights[static_cast<int64_t>(y) * w + x], 1.0e-6f);
    const float col_frac = Clampf((col_target - col_prev) / tex_w, 0.0f, 1.0f);
    const float2 uv = make_float2((static_cast<float>(x) + col_frac) / static_cast<float>(w),
                                  (static_cast<float>(y) + row_frac) / static_cast<float>(h));
    const float3 d = UvToDir(uv, L.env_rotation);
    const float pdf = EnvPdf(L, d);
    if (pdf <= 0.0f) return out;
    out.dir = d;
    out.radiance = EnvColor(L, d);
    out.pdf = pdf;
    out.valid = true;
    return out;
}

// ---- emissive lights (src/lighting/light_sampling.cpp:148-396) ------------------------

__device__ __forceinline__ float ClampUnitSample(float s) { return isfinite(s) ? Clampf(s, 0.0f, 1.0f) : 0.0f; }

__device__ __forceinline__ float TriArea(float3 a, float3 b, float3 c) {
    const float area = Length(Cross(Sub(b, a), Sub(c, a))) * 0.5f;
    return isfinite(area) && area > 0.0f ? area : 0.0f;
}

__device__ __forceinline__ float3 Pos(const SurfaceView& S, uint32_t vi) { return XYZ(S.positions[vi]); }

struct LightSample {
    float3 point, normal, radiance;
    float pdf;
    int index;
    bool valid;
};

__device__ __forceinline__ LightSample SampleEmissive(const SurfaceView& S, const LightView& L, float select, float2 ts) {
    LightSample out;
    out.valid = false;
    out.pdf = 0.0f;
    out.index = -1;
    const int n = L.emissive_count;
    if (n == 0) return out;
    // SampleAliasTable
    const float scaled = fminf(ClampUnitSample(select), kOneMinusEps) * static_cast<float>(n);
    const int column = min(static_cast<int>(scaled), n - 1);
    const float fraction = scaled - static_cast<float>(column);
    const float prob = L.alias_prob[column];
    int index = column;
    float remapped;
    if (fraction < prob) {
        remapped = prob > 0.0f ? fraction / prob : 0.0f;
    } else {
        index = L.alias_index[column];
        remapped = prob < 1.0f ? (fraction - prob) / (1.0f - prob) : 0.0f;
    }
    remapped = Clampf(remapped, 0.0f, kOneMinusEps);
    const float pmf = L.alias_pmf[index];
    if (index < 0 || pmf <= 0.0f) return out;
    // SampleEmissivePrimitive
    const float4 e = L.emit[index];
    const float area = isfinite(e.w) && e.w > 0.0f ? e.w : 0.0f;
    if (area <= 0.0f) return out;
    const int4 info = L.emit_info[index];
    if (info.x < 0 || info.z <= 0) return out;
    const float prim_area = L.emit_area[index];
    if (!isfinite(prim_area) || prim_area <= 0.0f) return out;
    // SelectTriangleByArea: first running area above the target (zero-area
    // triangles never start a new running value, so they are never picked).
    const float target = ClampUnitSample(remapped) * prim_area;
    const float* cdf = L.emit_cdf + info.y;
    int tri = 0;
    {
        int lo = 0, hi = info.z;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (!(target < cdf[mid])) {
                lo = mid + 1;
            } else {
                hi = mid;
            }
        }
        tri = lo;
        if (tri >= info.z) {
            // no running area exceeds the target: the last positive-area triangle
            tri = info.z - 1;
            while (tri > 0 && !(cdf[tri] > (tri > 0 ? cdf[tri - 1] : 0.0f))) --tri;
        }
    }
    const int4 pi = S.prim_info[info.x];
    const int64_t base = static_cast<int64_t>(pi.z) + static_cast<int64_t>(tri) * 3;
    const float3 p0 = Pos(S, S.indices[base]), p1 = Pos(S, S.indices[base + 1]), p2 = Pos(S, S.indices[base + 2]);
    const float su = sqrtf(ts.x);
    const float u = 1.0f - su;
    const float v = ts.y * su;
    const float3 point = Add(Add(Mul(p0, 1.0f - u - v), Mul(p1, u)), Mul(p2, v));
    const float3 nrm = Normalize(Cross(Sub(p1, p0), Sub(p2, p0)));
    const float ta = TriArea(p0, p1, p2);
    if (ta <= 0.0f || LengthSq(nrm) == 0.0f) return out;
    out.point = point;
    out.normal = nrm;
    out.radiance = XYZ(e);
    out.pdf = (1.0f / area) * pmf;
    out.index = index;
    out.valid = true;
    return out;
}

__device__ __forceinline__ float EmissivePdfSolidAngle(const LightView& L, float3 shading, float3 light, float3 ln,
                                                       int index) {
    const int n = L.emissive_count;
    if (n == 0) return 0.0f;
    float area_pdf;
    if (index >= 0 && index < n) {
        const float w = L.emit[index].w;
        const float area = isfinite(w) && w > 0.0f ? w : 0.0f;
        if (area <= 0.0f) return 0.0f;
        area_pdf = L.alias_pmf[index] / area;
    } else {
        const float total = L.total_emissive_area;
        if (!isfinite(total) || total <= 0.0f) return 0.0f;
        area_pdf = 1.0f / total;
    }
    const float3 dir = Sub(light, shading);
    const float d2 = LengthSq(dir);
    if (d2 <= 0.0f) return 0.0f;
    const float c = fabsf(Dot(ln, Normalize(Sub(shading, light))));
    if (c <= 0.0f) return 0.0f;
    return area_pdf * d2 / c;
}

// ---- analytic lights (src/lighting/analytic_light_sampling.cpp:22-140) --------------

struct AnalyticSample {
    float3 wi, radiance;
    float distance;
    float selection_pdf;
    bool valid;
};

__device__ __forceinline__ int4 AsInt4(float4 v) {
    return make_int4(__float_as_int(v.x), __float_as_int(v.y), __float_as_int(v.z), __float_as_int(v.w));
}

// TreeImportance: power over the squared distance to the node's bounds.
__device__ __forceinline__ float TreeImportance(const float4* node, float3 p) {
    const float4 lo = node[0], hi = node[1];
    const float dx = fmaxf(fmaxf(lo.x - p.x, 0.0f), p.x - hi.x);
    const float dy = fmaxf(fmaxf(lo.y - p.y, 0.0f), p.y - hi.y);
    const float dz = fmaxf(fmaxf(lo.z - p.z, 0.0f), p.z - hi.z);
    return lo.w / fmaxf(1.0e-4f, dx * dx + dy * dy + dz * dz);
}

// SampleAnalyticLight: pick one light (light tree descent by importance, else
// the power alias table, else uniformly), then its delta sample.
__device__ __forceinline__ AnalyticSample SampleAnalytic(const LightView& L, float3 p, float select) {
    AnalyticSample out;
    out.valid = false;
    out.selection_pdf = 0.0f;
    const int n = L.analytic_count;
    if (n <= 0) return out;
    int light = -1;
    float pdf = 1.0f;
    float u = Clampf(select, 0.0f, kOneMinusEps);
    if (L.use_analytic_tree && L.analytic_tree_root >= 0 && L.analytic_tree_root < L.analytic_tree_count) {
        // Children sit after their parent (checked at upload), so the descent ends.
        int node = L.analytic_tree_root;
        while (node >= 0) {
            const int4 link = AsInt4(L.alight_tree[3 * node + 2]);
            if (link.z >= 0) {
                light = link.z;
                break;
            }
            if (link.x < 0 || link.y < 0) return out;
            const float lw = TreeImportance(L.alight_tree + 3 * link.x, p);
            const float rw = TreeImportance(L.alight_tree + 3 * link.y, p);
            const float total = lw + rw;
            const float lp = total > 0.0f ? lw / total : 0.5f;
            if (u < lp) {
                pdf *= lp;
                u = lp > 0.0f ? u / lp : 0.0f;
                node = link.x;
            } else {
                const float rp = 1.0f - lp;
                pdf *= rp;
                u = rp > 0.0f ? (u - lp) / rp : 0.0f;
                node = link.y;
            }
        }
    } else if (L.analytic_pmf_count == n) {
        // SampleAliasTable over the analytic power table.
        const float scaled = fminf(ClampUnitSample(u), kOneMinusEps) * static_cast<float>(n);
        const int column = min(static_cast<int>(scaled), n - 1);
        const float fraction = scaled - static_cast<float>(column);
        const float prob = L.alight_prob[column];
        light = fraction < prob ? column : L.alight_alias[column];
        pdf = L.alight_pmf[light];
    } else {
        light = min(static_cast<int>(u * static_cast<float>(n)), n - 1);
        pdf = 1.0f / static_cast<float>(n);
    }
    if (light < 0 || light >= n || pdf <= 0.0f) return out;
    const float4 a = L.alight[3 * light], b = L.alight[3 * light + 1], c = L.alight[3 * light + 2];
    const int kind = __float_as_int(a.w);
    const float3 intensity = XYZ(c);
    if (kind == YRH_LIGHT_DISTANT) {
        const float3 to_light = Neg(XYZ(b));
        if (LengthSq(to_light) == 0.0f) return out;
        out.wi = Normalize(to_light);
        out.distance = 1.0e6f;
        out.radiance = intensity;
    } else {
        const float3 delta = Sub(XYZ(a), p);
        const float d2 = Dot(delta, delta);
        if (d2 <= 0.0f) return out;
        if (kind == YRH_LIGHT_POINT) {
            out.distance = sqrtf(d2);
            out.wi = Normalize(delta);
            out.radiance = Div(intensity, d2);
        } else {
            const float dist = sqrtf(d2);
            const float3 wi = Div(delta, dist);
            const float3 axis = LengthSq(XYZ(b)) > 0.0f ? Normalize(XYZ(b)) : V(0.0f, 0.0f, 1.0f);
            const float cosine = -Dot(wi, axis);
            const float outer = b.w, inner = c.w;
            float falloff = 0.0f;
            if (cosine >= inner) {
                falloff = 1.0f;
            } else if (cosine > outer) {
                const float t = (cosine - outer) / (inner - outer);
                falloff = t * t * (3.0f - 2.0f * t);
            }
            out.wi = wi;
            out.distance = dist;
            out.radiance = Mul(intensity, falloff / d2);
        }
    }
    out.selection_pdf = pdf;
    out.valid = true;
    return out;
}

// ---- kernels -----------------------------------------------------------------------

__global__ void RaygenKernel(PassView P) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t pixel;
    int x, y, sample, spp, width, height;
    DevCamera camera;
    if (P.jobs != nullptr) {
        // Batched pass: find the work item holding path p (items are ordered by path_begin).
        const int32_t pass = static_cast<int32_t>(P.counters[kPassIndex]);
        const int32_t ib = P.pass_items[pass], ie = P.pass_items[pass + 1];
        const DevItem last = P.items[ie - 1];
        const int64_t total = last.path_begin + last.pixel_count;
        if (p == 0) {
            P.counters[kCountQueue0] = static_cast<uint32_t>(total);
            P.counters[kCountPassPaths] = static_cast<uint32_t>(total);
        }
        if (p >= total) return;
        int32_t lo = ib, hi = ie - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (P.items[mid].path_begin <= p) {
                lo = mid;
            } else {
                hi = mid - 1;
            }
        }
        const DevItem it = P.items[lo];
        const DevJob& J = P.jobs[it.job];
        const int64_t local = it.pixel_begin + (p - it.path_begin);
        x = static_cast<int>(local % J.width);
        y = static_cast<int>(local / J.width);
        sample = it.sample;
        spp = J.spp;
        width = J.width;
        height = J.height;
        camera = J.camera;
        pixel = J.film_offset + local;
        P.p_job[p] = it.job | (it.occurrence << 16);
        P.p_sample[p] = it.sample;
    } else if (P.active_list != nullptr) {
        // RenderCpuPathTrace's active-pixel passes: listed pixel e renders its
        // next sample_count samples (indices from its film sample count) as
        // paths e, e + n, e + 2n, ... (n listed pixels).
        const int64_t n = *P.active_count;
        const int64_t total = n * P.sample_count;
        if (p == 0) P.counters[kCountQueue0] = static_cast<uint32_t>(total);
        if (p >= total) return;
        pixel = P.active_list[p % n];
        x = static_cast<int>(pixel % P.width);
        y = static_cast<int>(pixel / P.width);
        sample = static_cast<int>(P.film[pixel].samples) + static_cast<int>(p / n);
        P.p_sample[p] = sample;  // the bounces re-create the path's sampler from it
        spp = P.spp;
        width = P.width;
        height = P.height;
        camera = P.camera;
    } else {
        const int64_t total = P.npix * P.sample_count;
        if (p == 0) P.counters[kCountQueue0] = static_cast<uint32_t>(total);
        if (p >= total) return;
        pixel = p % P.npix;
        x = static_cast<int>(pixel % P.width);
        y = static_cast<int>(pixel / P.width);
        sample = P.sample_index + static_cast<int>(p / P.npix);
        spp = P.spp;
        width = P.width;
        height = P.height;
        camera = P.camera;
    }
    PixelSampler smp;
    smp.Init(P.sampler, P.seed, x, y, sample, spp);
    const float2 s = smp.At2D(kDimPixel, 0);
    const float w = static_cast<float>(width < 1 ? 1 : width);
    const float h = static_cast<float>(height < 1 ? 1 : height);
    const float3 d = CameraDirection(camera, w, h, static_cast<float>(x) + s.x, static_cast<float>(y) + s.y);
    P.q_org[0][p] = F4(camera.origin, 1.0e-5f);
    P.q_dir[0][p] = F4(d, __int_as_float(0x7f800000));
    if (P.p_diff != nullptr) {
        // The ray differentials one pixel right and down (GeneratePerspectiveCameraRay,
        // src/scene/camera.cpp:20-39), for texture footprints at the first hit.
        const float fx = static_cast<float>(x) + s.x, fy = static_cast<float>(y) + s.y;
        P.p_diff[2 * p] = F4(CameraDirection(camera, w, h, fx + 1.0f, fy), 0.0f);
        P.p_diff[2 * p + 1] = F4(CameraDirection(camera, w, h, fx, fy + 1.0f), 0.0f);
    }
    P.q_path[0][p] = static_cast<int32_t>(p);
    P.p_pixel[p] = static_cast<int32_t>(pixel);
    P.p_tp[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    P.p_rad[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.p_prev[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.p_flags[p] = 0;
    P.p_med[p] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    P.p_alb[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.p_nrm[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.p_slot[p] = make_int4(-1, -1, -1, -1);
}

// Zeroes the counters a bounce appends to before anything uses them.
// Zeroes the counters a bounce uses before anything uses them: kBeginTrace
// the path-segment claim counters (before the bounce's traversal), kBeginShade
// the next queue's count, the shadow queue's count and its claim counters
// (before Shade appends to them).
constexpr int kBeginTrace = 1;
constexpr int kBeginShade = 2;

__global__ void BeginBounceKernel(uint32_t* counters, int next_queue, int what) {
    const int i = static_cast<int>(threadIdx.x);
    if (what & kBeginShade) {
        if (i == 0) counters[next_queue] = 0u;
        if (i == 1) counters[kCountShadow] = 0u;
        if (i >= 2 && i < 2 + trav::kParts) counters[kWorkShadow + (i - 2) * trav::kPartStride] = 0u;
    }
    if ((what & kBeginTrace) && i < trav::kParts) counters[kWorkExtend + i * trav::kPartStride] = 0u;
}

__global__ void CopyTraceTotalsKernel(unsigned long long* stats) {
    if (threadIdx.x == 0) {
        stats[kStatNodeTests] += stats[kStatTraceNodes];
        stats[kStatTriTests] += stats[kStatTraceTris];
        stats[kStatSphereTests] += stats[kStatTraceSpheres];
        stats[kStatTraceNodes] = 0;
        stats[kStatTraceTris] = 0;
        stats[kStatTraceSpheres] = 0;
    }
}

struct SurfacePoint {
    float3 hp, wo, sn, obj_geo;
    int4 pinfo;
    DevMaterial mat;
};

// ---- textures at a triangle hit (material_evaluator.cpp:35-149, surface_query.cpp:35-97)

__device__ __forceinline__ float2 HitUv(const SurfaceView& S, uint32_t i0, uint32_t i1, uint32_t i2, float u, float v) {
    const float w = 1.0f - u - v;
    const float2 a = S.uvs[i0], b = S.uvs[i1], c = S.uvs[i2];
    return make_float2(a.x * w + b.x * u + c.x * v, a.y * w + b.y * u + c.y * v);
}

// TransformPoint / TransformVector (src/core/transform.cpp:30-44) with the
// instance's object-to-world rows.
__device__ __forceinline__ float3 XformPointRows(const float4* rows, float3 p) {
    const float4 m0 = rows[3], m1 = rows[4], m2 = rows[5];
    return V(m0.x * p.x + m0.y * p.y + m0.z * p.z + m0.w, m1.x * p.x + m1.y * p.y + m1.z * p.z + m1.w,
             m2.x * p.x + m2.y * p.y + m2.z * p.z + m2.w);
}
__device__ __forceinline__ float3 XformVectorRows(const float4* rows, float3 v) {
    const float4 m0 = rows[3], m1 = rows[4], m2 = rows[5];
    return V(m0.x * v.x + m0.y * v.y + m0.z * v.z, m1.x * v.x + m1.y * v.y + m1.z * v.z,
             m2.x * v.x + m2.y * v.y + m2.z * v.z);
}

// ComputeTextureFootprint (surface_query.cpp:35-97): the uv change per pixel
// where the camera ray's differential rays (directions diff[0], diff[1] from
// the same origin o) meet the hit's plane. p0..p2: object-space corners.
__device__ __forceinline__ tex::Footprint HitFootprint(const SurfaceView& S, const float4* nrows, uint32_t i0, uint32_t i1, uint32_t i2,
                                       float3 p0, float3 p1, float3 p2, float3 o, const float4* diff, float3 hp,
                                       float3 geo) {
    const tex::Footprint none = tex::NoFootprint();
    if (nrows != nullptr) {
        p0 = XformPointRows(nrows, p0);
        p1 = XformPointRows(nrows, p1);
        p2 = XformPointRows(nrows, p2);
    }
    const float2 uv0 = S.uvs[i0], uv1 = S.uvs[i1], uv2 = S.uvs[i2];
    const float du1 = uv1.x - uv0.x, dv1 = uv1.y - uv0.y;
    const float du2 = uv2.x - uv0.x, dv2 = uv2.y - uv0.y;
    const float uv_det = du1 * dv2 - dv1 * du2;
    if (fabsf(uv_det) < 1.0e-12f) return none;
    const float inv_uv = 1.0f / uv_det;
    const float3 e1 = Sub(p1, p0), e2 = Sub(p2, p0);
    const float3 dpdu = Mul(Sub(Mul(e1, dv2), Mul(e2, dv1)), inv_uv);
    const float3 dpdv = Mul(Sub(Mul(e2, du1), Mul(e1, du2)), inv_uv);
    float3 on[2];
    for (int k = 0; k < 2; ++k) {
        const float3 dir = XYZ(diff[k]);
        const float den = Dot(geo, dir);
        if (fabsf(den) < 1.0e-12f) {
            on[k] = hp;
        } else {
            const float tt = Dot(geo, Sub(hp, o)) / den;
            on[k] = Add(o, Mul(dir, tt));
        }
    }
    const float3 dpdx = Sub(on[0], hp), dpdy = Sub(on[1], hp);
    const float a = Dot(dpdu, dpdu), b = Dot(dpdu, dpdv), c = Dot(dpdv, dpdv);
    const float gram = a * c - b * b;
    if (fabsf(gram) < 1.0e-20f) return none;
    const float inv_gram = 1.0f / gram;
    const float ux = Dot(dpdx, dpdu), vx = Dot(dpdx, dpdv), uy = Dot(dpdy, dpdu), vy = Dot(dpdy, dpdv);
    const tex::Footprint f{(ux * c - vx * b) * inv_gram, (vx * a - ux * b) * inv_gram, (uy * c - vy * b) * inv_gram,
                           (vy * a - uy * b) * inv_gram};
    if (!isfinite(f.dudx) || !isfinite(f.dvdx) || !isfinite(f.dudy) || !isfinite(f.dvdy)) return none;
    return f;
}

// ResolveNormalMap (material_evaluator.cpp:35-92): the map's tangent-space
// normal (x, y scaled) in the frame of the interpolated tangent made
// orthogonal to the shading normal, with the interpolated handedness
// (negated under a mirroring instance transform).
__device__ __forceinline__ float3 NormalMapped(const SurfaceView& S, const LightView& L, int map, float scale, uint32_t i0, uint32_t i1,
                               uint32_t i2, float u, float v, float2 uv, float3 sn, const float4* nrows,
                               const tex::Footprint& fp) {
    const float4 e = tex::Sample(L.tex, map, uv.x, uv.y, fp);
    float3 ts = V(e.x * 2.0f - 1.0f, e.y * 2.0f - 1.0f, e.z * 2.0f - 1.0f);
    ts.x *= scale;
    ts.y *= scale;
    ts = Normalize(ts);
    if (LengthSq(ts) == 0.0f) return sn;
    const float w = 1.0f - u - v;
    const float4 t0 = S.tangents[i0], t1 = S.tangents[i1], t2 = S.tangents[i2];
    float3 tan = Normalize(V(t0.x * w + t1.x * u + t2.x * v, t0.y * w + t1.y * u + t2.y * v,
                             t0.z * w + t1.z * u + t2.z * v));
    if (nrows != nullptr) tan = XformVectorRows(nrows, tan);
    const float3 proj = Sub(tan, Mul(sn, Dot(tan, sn)));
    if (LengthSq(proj) <= 1.0e-12f) return sn;
    tan = Normalize(proj);
    if (LengthSq(tan) == 0.0f) return sn;
    const float orientation = nrows != nullptr ? nrows[2].w : 1.0f;
    const float hand = (t0.w * w + t1.w * u + t2.w * v) * orientation;
    const float3 bit = Mul(Cross(sn, tan), hand);
    const float3 mapped = Normalize(Add(Add(Mul(tan, ts.x), Mul(bit, ts.y)), Mul(sn, ts.z)));
    return LengthSq(mapped) > 0.0f ? mapped : sn;
}

// The alpha EvaluateMaterialAtSurface gives a triangle hit (material_evaluator.cpp:135-149):
// the material's alpha, times the alpha texture's first channel where the
// primitive has uvs. `diff`: the camera ray's differentials (footprint), or nullptr.
__device__ __forceinline__ float HitAlpha(const SurfaceView& S, const LightView& L, int tri, int inst, float u, float v, float3 o,
                          float3 d, float t, const float4* diff) {
    const int prim = S.tri_prim[tri];
    const int4 pinfo = S.prim_info[prim];
    const float alpha = L.materials[pinfo.x].alpha;
    const int at = L.mat_tex[pinfo.x].y;
    if (at < 0 || !(S.prim_flags[prim] & kHasUvs)) return alpha;
    const int64_t base = static_cast<int64_t>(pinfo.z) + static_cast<int64_t>(tri - pinfo.w) * 3;
    const uint32_t i0 = S.indices[base], i1 = S.indices[base + 1], i2 = S.indices[base + 2];
    const float2 uv = HitUv(S, i0, i1, i2, u, v);
    tex::Footprint fp = tex::NoFootprint();
    if (diff != nullptr) {
        const float3 p0 = Pos(S, i0), p1 = Pos(S, i1), p2 = Pos(S, i2);
        const float3 obj_geo = Normalize(Cross(Sub(p1, p0), Sub(p2, p0)));
        const float4* nrows = inst >= 0 ? S.inst_normal + 6 * static_cast<int64_t>(inst) : nullptr;
        const float3 raw = nrows != nullptr ? TransformNormalDev(nrows, obj_geo) : obj_geo;
        const float3 geo = Dot(raw, Neg(d)) < 0.0f ? Neg(raw) : raw;
        fp = HitFootprint(S, nrows, i0, i1, i2, p0, p1, p2, o, diff, Add(o, Mul(d, t)), geo);
    }
    return alpha * tex::Sample(L.tex, at, uv.x, uv.y, fp).x;
}

// TraceVisibleSurface's hit resolution (src/integrators/surface_query.cpp:101-200):
// material (kTex: with its reflectance texture applied), hit point,
// object-space geometric normal and the face-forwarded shading normal (kTex:
// normal-mapped). `diff`: the camera ray's differentials, or nullptr.
template <bool kTex>
__device__ __forceinline__ SurfacePoint ResolveSurface(const SurfaceView& S, const LightView& L, float3 org,
                                                       float3 dir, float4 h, int tri, int inst, const float4* diff) {
    const float t = h.x;
    if (trav::IsSphereId(tri)) {
        // Sphere hit (surface_query.cpp:129-146): outward normal at ray.At(t)
        // (flipped for flip_normals) is both the geometric and the shading
        // normal, not face-forwarded; spheres carry no emission index.
        const int k = -2 - tri;
        const float4 sph = S.spheres[k];
        const int2 info = S.sphere_info[k];
        SurfacePoint out;
        out.hp = Add(org, Mul(dir, t));
        out.wo = Neg(dir);
        const float3 n = Normalize(Sub(out.hp, XYZ(sph)));
        out.sn = info.y ? Neg(n) : n;
        out.obj_geo = out.sn;
        out.pinfo = make_int4(info.x, -1, 0, 0);
        out.mat = LoadMaterial(L.materials, info.x);
        return out;
    }
    const int prim = S.tri_prim[tri];
    const int4 pinfo = S.prim_info[prim];
    DevMaterial mat = LoadMaterial(L.materials, pinfo.x);
    const float3 hp = Add(org, Mul(dir, t));
    const float3 wo = Neg(dir);
    const int64_t base = static_cast<int64_t>(pinfo.z) + static_cast<int64_t>(tri - pinfo.w) * 3;
    const uint32_t i0 = S.indices[base], i1 = S.indices[base + 1], i2 = S.indices[base + 2];
    const float3 p0 = Pos(S, i0), p1 = Pos(S, i1), p2 = Pos(S, i2);
    const float3 obj_geo = Normalize(Cross(Sub(p1, p0), Sub(p2, p0)));
    const float4* nrows = inst >= 0 ? S.inst_normal + 6 * static_cast<int64_t>(inst) : nullptr;
    const float3 raw = inst >= 0 ? TransformNormalDev(nrows, obj_geo) : obj_geo;
    const float3 geo = Dot(raw, wo) < 0.0f ? Neg(raw) : raw;
    const int pflags = S.prim_flags[prim];
    int normal_map = -1;
    float normal_scale = 1.0f;
    float2 uv = make_float2(0.0f, 0.0f);
    tex::Footprint fp = tex::NoFootprint();
    if (kTex && (pflags & kHasUvs)) {
        const int4 mt = L.mat_tex[pinfo.x];
        if (mt.x >= 0 || mt.z >= 0) {
            uv = HitUv(S, i0, i1, i2, h.y, h.z);
            if (diff != nullptr) fp = HitFootprint(S, nrows, i0, i1, i2, p0, p1, p2, org, diff, hp, geo);
            if (mt.x >= 0) mat.reflectance = Hadamard(mat.reflectance, XYZ(tex::Sample(L.tex, mt.x, uv.x, uv.y, fp)));
            if (pflags & kHasTangents) {
                normal_map = mt.z;
                normal_scale = __int_as_float(mt.w);
            }
        }
    }
    // ResolveShadingNormal (src/geometry/intersection.cpp:84-106) on the object-space
    // normal for instances, on the face-forwarded normal otherwise.
    const float3 ref = inst >= 0 ? obj_geo : geo;
    float3 sn = Normalize(ref);
    if (pflags & kHasNormals) {
        const float u = h.y, v = h.z, w = 1.0f - u - v;
        const float3 n0 = XYZ(S.normals[i0]), n1 = XYZ(S.normals[i1]), n2 = XYZ(S.normals[i2]);
        const float3 ni = Normalize(V(n0.x * w + n1.x * u + n2.x * v, n0.y * w + n1.y * u + n2.y * v,
                                      n0.z * w + n1.z * u + n2.z * v));
        if (LengthSq(ni) > 0.0f) sn = ni;
    }
    if (Dot(sn, ref) < 0.0f) sn = Neg(sn);
    if (LengthSq(sn) == 0.0f) sn = Normalize(ref);
    if (inst >= 0) {
        sn = TransformNormalDev(nrows, sn);
        if (Dot(sn, geo) < 0.0f) sn = Neg(sn);
    }
    if (kTex && normal_map >= 0) sn = NormalMapped(S, L, normal_map, normal_scale, i0, i1, i2, h.y, h.z, uv, sn, nrows, fp);
    if (Dot(sn, wo) < 0.0f) sn = Neg(sn);
    SurfacePoint out;
    out.hp = hp;
    out.wo = wo;
    out.sn = sn;
    out.obj_geo = obj_geo;
    out.pinfo = pinfo;
    out.mat = mat;
    return out;
}

// Shadow entries of a vertex, in EstimateDirectLight's summation order.
constexpr int kShadowEmissive = 0;
constexpr int kShadowAnalytic = 1;
constexpr int kShadowEnv = 2;
constexpr int kShadowKinds = 3;

struct ShadeOut {
    bool next;
    float3 next_org, next_dir;
    bool shadow[kShadowKinds];
    float3 s_org[kShadowKinds], s_dir[kShadowKinds];
    float s_tmax[kShadowKinds];
    float3 s_li[kShadowKinds], s_bsdf[kShadowKinds];
    float s_scale[kShadowKinds];
};

// ---- guided sampling (learned proposal, yrh_guide.h) ----------------------------

__device__ __forceinline__ float Lum(float3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// Vertices the proposal guides: non-delta reflective BSDFs.
__device__ __forceinline__ bool Guidable(const DevMaterial& m) {
    return (m.kind == kDiffuse || m.kind == kConductor) && !IsDelta(m);
}

__device__ __forceinline__ float GuidePdf(const yrg::MixtureView& gm, float3 w) { return yrg::MixturePdfWithNorm(gm, w); }

// One-sample MIS of the proposal (chosen with probability alpha) and the BSDF:
// the direction comes from the chosen strategy, the pdf is the mixture
// alpha q + (1 - alpha) p_bsdf, so the estimate stays unbiased.
__device__ __forceinline__ BsdfSample SampleGuided(const DevMaterial& m, float3 wo, float3 n, const yrg::MixtureView& gm,
                                                   float alpha, float u, float2 s) {
    BsdfSample b = Invalid();
    float3 wi;
    if (u < alpha) {
        wi = Normalize(yrg::MixtureSample(gm, u / alpha, s.x, s.y));
    } else {
        const BsdfSample bs = SampleBsdf(m, wo, n, s);
        if (!bs.valid) return b;
        wi = bs.wi;
    }
    const float c = Dot(wi, n);
    if (!(c > 0.0f)) return b;
    const float pdf = alpha * GuidePdf(gm, wi) + (1.0f - alpha) * PdfBsdf(m, wo, wi, n);
    if (!(pdf > 0.0f)) return b;
    b.wi = wi;
    b.weight = Mul(EvalBsdf(m, wo, wi, n), c / pdf);
    b.pdf = pdf;
    b.valid = true;
    b.specular = false;
    return b;
}

// kAnalytic: the scene has analytic lights (a third shadow entry per vertex).
template <bool kGuided, bool kAnalytic, bool kTex>
__device__ __forceinline__ void ShadePath(const PassView& P, const SurfaceView& S, const LightView& L,
                                          const yrg::GuideView& G, int depth, int i, int cur, ShadeOut& o,
                                          unsigned& hit_count, unsigned& miss_count) {
    o.next = false;
    o.shadow[kShadowEmissive] = o.shadow[kShadowAnalytic] = o.shadow[kShadowEnv] = false;
    const int p = P.q_path[cur][i];
    const float4 o4 = P.q_org[cur][i];
    const float4 d4 = P.q_dir[cur][i];
    const float3 org = XYZ(o4);
    const float3 dir = XYZ(d4);
    const float4 h = P.hits[i];
    const int tri = __float_as_int(h.w);
    const int inst = P.hit_inst != nullptr ? P.hit_inst[i] : -1;
    int flags = P.p_flags[p];
    float3 tp = XYZ(P.p_tp[p]);
    float3 rad = XYZ(P.p_rad[p]);
    const float4 prev = P.p_prev[p];
    P.p_slot[p] = make_int4(-1, -1, -1, -1);
    const int pix = P.p_pixel[p];
    PixelSampler smp;
    if (P.jobs != nullptr) {
        const DevJob& J = P.jobs[P.p_job[p] & 0xFFFF];
        const int64_t local = pix - J.film_offset;
        smp.Init(P.sampler, P.seed, static_cast<int>(local % J.width), static_cast<int>(local / J.width), P.p_sample[p],
                 J.spp);
    } else if (P.active_list != nullptr) {
        smp.Init(P.sampler, P.seed, pix % P.width, pix / P.width, P.p_sample[p], P.spp);
    } else {
        smp.Init(P.sampler, P.seed, pix % P.width, pix / P.width, P.sample_index + static_cast<int>(p / P.npix), P.spp);
    }
    smp.bounce = depth;

    if (tri == -1) {
        ++miss_count;
        float w = 1.0f;
        if ((flags & kPrevValid) && !(flags & kPrevDelta)) w = PowerHeuristic(1, prev.w, 1, EnvPdf(L, dir));
        rad = Add(rad, Mul(Hadamard(tp, EnvColor(L, dir)), w));
        P.p_rad[p] = F4(rad, 0.0f);
        return;
    }
    ++hit_count;
    const float t = h.x;
    if (flags & kMediumActive) {
        const float4 med = P.p_med[p];
        tp = Hadamard(tp, BeerLambert(XYZ(med), med.w, t));
    }
    if (IsBlack(tp)) {
        P.p_tp[p] = F4(tp, 0.0f);
        return;
    }
    // P.p_diff is set for the camera rays' bounce only (texture footprints).
    const SurfacePoint sp = ResolveSurface<kTex>(S, L, org, dir, h, tri, inst, P.p_diff != nullptr ? P.p_diff + 2 * p : nullptr);
    const int4 pinfo = sp.pinfo;
    const DevMaterial& mat = sp.mat;
    const float3 hp = sp.hp;
    const float3 wo = sp.wo;
    const float3 obj_geo = sp.obj_geo;
    const float3 sn = sp.sn;
    // Guided vertex: the proposal's mixture of this queue row (GuideQueryKernel).
    // Only vertices that may continue are guided: at the last bounce no
    // continuation ray is traced, so next-event MIS keeps the BSDF pdf there,
    // and the guided estimator has the plain one's expectation at every depth.
    const bool guided = kGuided && G.valid[i] != 0;
    const yrg::MixtureView gm{guided ? G.mix + i : nullptr, G.mix_stride};
    const float alpha = guided ? G.alpha : 0.0f;

    if (depth == 0) {
        flags |= kPrimaryHit;
        P.p_alb[p] = F4(mat.reflectance, t);
        P.p_nrm[p] = F4(sn, 0.0f);
    }
    // Emission x MIS (path_integrator.cpp:86-93; light normal in object space
    // as GeometricNormal returns it). Sphere hits emit nothing (:83-86).
    if (!IsBlack(mat.emission) && tri >= 0) {
        float w = 1.0f;
        if ((flags & kPrevValid) && !(flags & kPrevDelta)) {
            const float pdf_light = EmissivePdfSolidAngle(L, XYZ(prev), hp, obj_geo, pinfo.y);
            w = PowerHeuristic(1, prev.w, 1, pdf_light);
        }
        rad = Add(rad, Mul(Hadamard(tp, mat.emission), w));
    }
    // Next-event estimation (path_direct_lighting.cpp:115-283): the emissive and
    // environment samples share the DirectLightSurface 2D sample.
    if (!IsDelta(mat)) {
        P.p_nee[p] = F4(tp, 0.0f);
        const float2 surf = smp.Get2D(kDimLightSurface);
        if (L.emissive_count > 0) {
            const LightSample ls = SampleEmissive(S, L, smp.Get1D(kDimLightSelect), surf);
            if (ls.valid && ls.pdf > 0.0f && !IsBlack(ls.radiance)) {
                const float3 to_light = Sub(ls.point, hp);
                const float d2 = LengthSq(to_light);
                if (d2 > 1.0e-4f * 1.0e-4f) {
                    const float dist = sqrtf(d2);
                    const float bias = ShadowBias(hp, ls.point, dist);
                    if (dist > bias) {
                        const float3 so = Add(hp, Mul(sn, bias));
                        const float3 stl = Sub(ls.point, so);
                        const float sd2 = LengthSq(stl);
                        if (sd2 > 1.0e-4f * 1.0e-4f) {
                            const float sd = sqrtf(sd2);
                            const float3 wi = Div(stl, sd);
                            const float c = fmaxf(0.0f, Dot(sn, wi));
                            if (c > 0.0f) {
                                const float pdf_light = EmissivePdfSolidAngle(L, hp, ls.point, ls.normal, ls.index);
                                if (pdf_light > 0.0f) {
                                    const float3 f = EvalBsdf(mat, wo, wi, sn);
                                    float pdf_bsdf = PdfBsdf(mat, wo, wi, sn);
                                    if (guided) pdf_bsdf = alpha * GuidePdf(gm, wi) + (1.0f - alpha) * pdf_bsdf;
                                    const float mis = PowerHeuristic(1, pdf_light, 1, pdf_bsdf);
                                    o.shadow[kShadowEmissive] = true;
                                    o.s_org[kShadowEmissive] = so;
                                    o.s_dir[kShadowEmissive] = wi;
                                    o.s_tmax[kShadowEmissive] = sd - bias;
                                    o.s_li[kShadowEmissive] = ls.radiance;
                                    o.s_bsdf[kShadowEmissive] = f;
                                    o.s_scale[kShadowEmissive] = c * mis / pdf_light;
                                }
                            }
                        }
                    }
                }
            }
        }
        // Analytic lights: one delta sample, no MIS (path_direct_lighting.cpp:250-281).
        if (kAnalytic && L.analytic_count > 0) {
            const AnalyticSample as = SampleAnalytic(L, hp, smp.Get1D(kDimAnalyticSelect));
            if (as.valid && as.selection_pdf > 0.0f && !IsBlack(as.radiance)) {
                const float c = fmaxf(0.0f, Dot(sn, as.wi));
                if (c > 0.0f) {
                    const float3 f = EvalBsdf(mat, wo, as.wi, sn);
                    if (!IsBlack(f)) {
                        const float bias = SurfaceBias(hp);
                        o.shadow[kShadowAnalytic] = true;
                        o.s_org[kShadowAnalytic] = Add(hp, Mul(sn, bias));
                        o.s_dir[kShadowAnalytic] = as.wi;
                        o.s_tmax[kShadowAnalytic] = as.distance - bias;
                        o.s_li[kShadowAnalytic] = as.radiance;
                        o.s_bsdf[kShadowAnalytic] = f;
                        o.s_scale[kShadowAnalytic] = c / as.selection_pdf;
                    }
                }
            }
        }
        if (L.env_sampleable) {
            const EnvSample es = SampleEnv(L, surf);
            if (es.valid && es.pdf > 0.0f && !IsBlack(es.radiance)) {
                const float c = fmaxf(0.0f, Dot(sn, es.dir));
                if (c > 0.0f) {
                    const float3 f = EvalBsdf(mat, wo, es.dir, sn);
                    if (!IsBlack(f)) {
                        float pdf_bsdf = PdfBsdf(mat, wo, es.dir, sn);
                        if (guided) pdf_bsdf = alpha * GuidePdf(gm, es.dir) + (1.0f - alpha) * pdf_bsdf;
                        const float mis = PowerHeuristic(1, es.pdf, 1, pdf_bsdf);
                        o.shadow[kShadowEnv] = true;
                        o.s_org[kShadowEnv] = Add(hp, Mul(sn, SurfaceBias(hp)));
                        o.s_dir[kShadowEnv] = es.dir;
                        o.s_tmax[kShadowEnv] = __int_as_float(0x7f800000);
                        o.s_li[kShadowEnv] = es.radiance;
                        o.s_bsdf[kShadowEnv] = f;
                        o.s_scale[kShadowEnv] = c * mis / es.pdf;
                    }
                }
            }
        }
    }
    const int max_depth = P.max_depth < 1 ? 1 : P.max_depth;
    bool cont = depth + 1 < max_depth;
    BsdfSample bs = Invalid();
    if (cont) {
        if (guided) {
            bs = SampleGuided(mat, wo, sn, gm, alpha, smp.Independent1D(kDimGuide), smp.Get2D(kDimBsdf));
        } else {
            bs = SampleBsdf(mat, wo, sn, smp.Get2D(kDimBsdf));
        }
        cont = bs.valid && !IsBlack(bs.weight);
    }
    const float beta_lum = Lum(tp);  // throughput arriving at this vertex
    if (cont) {
        tp = Hadamard(tp, bs.weight);
        if (depth >= 3) {
            const float survival = Clampf(MaxC(tp), 0.05f, 0.95f);
            // A guided throughput depends on the sampled direction, so the
            // roulette decision must not come from a Sobol dimension that
            // correlates with that direction's 2D sample (see Independent1D).
            const float xi = guided ? smp.Independent1D(kDimRoulette) : smp.Get1D(kDimRoulette);
            if (xi >= survival) {
                cont = false;
            } else {
                tp = Div(tp, survival);
            }
        }
    }
    if (cont) {
        if (mat.kind == kDielectric && Dot(bs.wi, sn) < 0.0f) {
            if (flags & kMediumActive) {
                flags &= ~kMediumActive;
                P.p_med[p] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            } else {
                flags |= kMediumActive;
                P.p_med[p] = F4(mat.absorption_color, mat.absorption_distance);
            }
        }
        flags |= kPrevValid;
        flags = bs.specular ? (flags | kPrevDelta) : (flags & ~kPrevDelta);
        P.p_prev[p] = F4(hp, bs.pdf);
        const float3 bn = Dot(bs.wi, sn) >= 0.0f ? sn : Neg(sn);
        if (guided && G.rec != nullptr) {
            // Training record: the radiance this direction brings back is
            // known once the path ends (GuideCollectKernel).
            float4* r = G.rec + 4 * (static_cast<int64_t>(depth) * G.rec_paths + p);
            r[0] = F4(hp, beta_lum);
            r[1] = F4(sn, Lum(rad));
            r[2] = F4(wo, __int_as_float(G.stamp));
            r[3] = F4(bs.wi, 0.0f);
        }
        o.next = true;
        o.next_org = Add(hp, Mul(bn, SurfaceBias(hp)));
        o.next_dir = bs.wi;
    }
    P.p_tp[p] = F4(tp, 0.0f);
    P.p_rad[p] = F4(rad, 0.0f);
    P.p_flags[p] = flags;
}

template <bool kGuided, bool kAnalytic, bool kTex>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void ShadeKernel(PassView P, SurfaceView S, LightView L, yrg::GuideView G,
                                                     int depth, int cur) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int n = static_cast<int>(P.counters[cur]);
    ShadeOut o;
    o.next = false;
    o.shadow[kShadowEmissive] = o.shadow[kShadowAnalytic] = o.shadow[kShadowEnv] = false;
    unsigned hit_count = 0, miss_count = 0;
    const bool valid = i < n;
    if (valid) ShadePath<kGuided, kAnalytic, kTex>(P, S, L, G, depth, i, cur, o, hit_count, miss_count);
    const int nxt = 1 - cur;
    __shared__ uint32_t append_lds[kBlock / 64 + 1];
    const int32_t slot = BlockAppend(P.counters + nxt, o.next ? 1u : 0u, append_lds);
    if (slot >= 0) {
        const int p = P.q_path[cur][i];
        P.q_org[nxt][slot] = F4(o.next_org, 1.0e-5f);
        P.q_dir[nxt][slot] = F4(o.next_dir, __int_as_float(0x7f800000));
        P.q_path[nxt][slot] = p;
    }
    int slot_of[kShadowKinds] = {-1, -1, -1};
    uint32_t wanted = 0;
    for (int k = 0; k < kShadowKinds; ++k) wanted += (k != kShadowAnalytic || kAnalytic) && o.shadow[k] ? 1u : 0u;
    int32_t s = BlockAppend(P.counters + kCountShadow, wanted, append_lds);
    for (int k = 0; k < kShadowKinds; ++k) {
        if ((k != kShadowAnalytic || kAnalytic) && o.shadow[k]) {
            P.s_org[s] = F4(o.s_org[k], 1.0e-5f);
            P.s_dir[s] = F4(o.s_dir[k], o.s_tmax[k]);
            P.s_li[s] = F4(o.s_li[k], o.s_scale[k]);
            P.s_bsdf[s] = F4(o.s_bsdf[k], 0.0f);
            P.s_trans[s] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            slot_of[k] = s;
            ++s;
        }
    }
    if (valid && wanted > 0) {
        P.p_slot[P.q_path[cur][i]] = make_int4(slot_of[kShadowEmissive], slot_of[kShadowAnalytic], slot_of[kShadowEnv], -1);
    }
    WaveStat(P.stats, kStatHits, hit_count);
    WaveStat(P.stats, kStatMisses, miss_count);
    if (i == 0) atomicAdd(P.stats + kStatRays, static_cast<unsigned long long>(n));
}

// Proposal for every queue row of this bounce: each lane resolves its row's
// surface (position, shading normal, outgoing direction), the wave evaluates
// the network on its 64 rows with MFMA, and each guidable row gets its mixture.
// Per-bounce proposal inference: kQueryBlock lanes share one copy of the
// weight fragments in LDS; each wave runs its 64 rows through the MLP 16 at a
// time, and all 64 lanes then form the 16 rows' mixtures (lane: row l % 16,
// lobes 2(l / 16) and 2(l / 16) + 1, with the same operations as
// HeadFromOutputs), written column-wise (value k of row i at
// mix[k * mix_stride + i]).
constexpr int kQueryBlock = 512;

// Orders a wave's LDS writes before its own later LDS reads (the staging
// areas below are private to one wave, so no workgroup barrier is needed).
__device__ __forceinline__ void WaveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
constexpr int kQueryWaves = kQueryBlock / 64;

template <bool kTex>
__global__ __launch_bounds__(kQueryBlock) void GuideQueryKernel(PassView P, SurfaceView S, LightView L,
                                                               yrg::GuideView G, int cur) {
    // Per wave, first the inputs of its 64 rows (stride 11: odd, no bank
    // conflicts), then the network outputs of 16 rows at a time.
    __shared__ float lds_stage[kQueryWaves][16 * yrg::kOutStride];
    static_assert(16 * yrg::kOutStride >= 64 * 11, "staging area holds the inputs");
    __shared__ yrg::bf16x8_t lds_w[yrg::kFwdFragCount * 64];  // forward weight fragments, shared by the block
    const int n = static_cast<int>(P.counters[cur]);
    if (static_cast<int64_t>(blockIdx.x) * blockDim.x >= n) return;  // whole block idle
    for (int k = static_cast<int>(threadIdx.x); k < yrg::kFwdFragCount * 64; k += kQueryBlock) lds_w[k] = G.frag[k];
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int wave = static_cast<int>(threadIdx.x >> 6);
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    bool guidable = false;
    float3 hp = V(0.0f, 0.0f, 0.0f), sn = hp, wo = hp;
    if (i < n) {
        const float4 h = P.hits[i];
        const int tri = __float_as_int(h.w);
        if (tri != -1) {
            const int inst = P.hit_inst != nullptr ? P.hit_inst[i] : -1;
            const float4* diff = P.p_diff != nullptr ? P.p_diff + 2 * P.q_path[cur][i] : nullptr;
            const SurfacePoint sp =
                ResolveSurface<kTex>(S, L, XYZ(P.q_org[cur][i]), XYZ(P.q_dir[cur][i]), h, tri, inst, diff);
            guidable = Guidable(sp.mat);
            hp = sp.hp;
            sn = sp.sn;
            wo = sp.wo;
        }
    }
    float* in = lds_stage[wave] + lane * 11;
    in[0] = hp.x; in[1] = hp.y; in[2] = hp.z;
    in[3] = sn.x; in[4] = sn.y; in[5] = sn.z;
    in[6] = wo.x; in[7] = wo.y; in[8] = wo.z;
    in[9] = guidable ? 1.0f : 0.0f;
    if (i < n) G.valid[i] = guidable ? 1 : 0;
    __syncthreads();
    const int q = lane / 16;
    const int row = lane % 16;
    float* lo = lds_stage[wave];
    const int64_t wave_row0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + wave * 64;
    // Encodings of all 4 row blocks first, so their grid loads are in flight
    // together.
    yrg::bf16x4_t xin[4][2];
    bool use[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        const float* r = lds_stage[wave] + (rb * 16 + row) * 11;
        float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        use[rb] = r[9] != 0.0f;
        if (use[rb]) {
            const float pp[3] = {r[0], r[1], r[2]};
            const float nn[3] = {r[3], r[4], r[5]};
            const float ww[3] = {r[6], r[7], r[8]};
            yrg::EncodeShareView(G, pp, q, f);
            xin[rb][1] = yrg::DirectionShare(q, nn, ww);
        } else {
            xin[rb][1] = yrg::PackBf16(0.0f, 0.0f, 0.0f, 0.0f);
        }
        xin[rb][0] = yrg::PackBf16(f[0], f[1], f[2], f[3]);
    }
    WaveSync();  // the inputs are read; the area now takes the outputs
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        const yrg::bf16x4_t x[1][2] = {{xin[rb][0], xin[rb][1]}};
        yrg::bf16x4_t h0[1][4], h1[1][4], h2[1][4];
        yrg::f32x4_t o[1][3];
        yrg::MlpForward<1>(lds_w, lane, x, h0, h1, h2, o);
        for (int mb = 0; mb < 3; ++mb) {
            for (int j = 0; j < 4; ++j) lo[row * yrg::kOutStride + mb * 16 + 4 * q + j] = o[0][mb][j];
        }
        WaveSync();
        const int64_t ri = wave_row0 + rb * 16 + row;
        if (ri < n && use[rb]) {
            const float* out = lo + row * yrg::kOutStride;
            float mx = out[0];
            for (int k = 1; k < yrg::kGuideLobes; ++k) mx = fmaxf(mx, out[5 * k]);
            float sum = 0.0f;
            for (int k = 0; k < yrg::kGuideLobes; ++k) sum += expf(out[5 * k] - mx);
            float* col = G.mix + ri;
            for (int k = 2 * q; k < 2 * q + 2; ++k) {
                col[k * G.mix_stride] = expf(out[5 * k] - mx) / sum;
                const float s = fminf(fmaxf(out[5 * k + 1], yrg::kLogKappaMin), yrg::kLogKappaMax);
                const float kappa = expf(s);
                col[(8 + k) * G.mix_stride] = kappa;
                col[(yrg::kGuideMixture + k) * G.mix_stride] = yrg::VmfNorm(kappa);
                const float ux = out[5 * k + 2], uy = out[5 * k + 3], uz = out[5 * k + 4];
                const float len = sqrtf(ux * ux + uy * uy + uz * uz);
                const bool ok = len > 1.0e-6f;
                col[(16 + 3 * k) * G.mix_stride] = ok ? ux / len : 0.0f;
                col[(16 + 3 * k + 1) * G.mix_stride] = ok ? uy / len : 0.0f;
                col[(16 + 3 * k + 2) * G.mix_stride] = ok ? uz / len : 1.0f;
            }
        }
        WaveSync();
    }
}

// After the last bounce: each record of this pass becomes a training sample
// (position, normal, outgoing, sampled direction, weight) with weight =
// luminance of the radiance the path gathered after the vertex over the
// luminance of the throughput arriving at it, i.e. f cos L / p of the sampled
// direction. Zero-weight samples carry no gradient and are skipped.
// Each workgroup takes kCollectPer runs of kBlock consecutive slots and
// appends its kept samples with one atomic (a single counter word taken by
// every wave had cost more than the reads).
constexpr int kCollectPer = 8;

__global__ __launch_bounds__(kBlock) void GuideCollectKernel(PassView P, yrg::GuideView G) {
    __shared__ uint32_t append_lds[kBlock / 64 + 1];
    __shared__ float wsum_lds[kBlock / 64];
    const int64_t total = G.rec_paths * G.rec_depths;
    const int64_t first = static_cast<int64_t>(blockIdx.x) * (kBlock * kCollectPer) + threadIdx.x;
    float w[kCollectPer];
    uint32_t keep = 0;
    for (int j = 0; j < kCollectPer; ++j) {
        const int64_t slot = first + j * kBlock;
        w[j] = 0.0f;
        if (slot < total) {
            const float4* r = G.rec + 4 * slot;
            if (__float_as_int(r[2].w) == G.stamp) {
                const float final_lum = Lum(XYZ(P.p_rad[slot % G.rec_paths]));
                w[j] = (final_lum - r[1].w) / r[0].w;
                if (isfinite(w[j]) && w[j] > 0.0f) keep |= 1u << j;
            }
        }
    }
    const int32_t base = BlockAppend(G.sample_count, static_cast<uint32_t>(__popc(keep)), append_lds);
    float ws = 0.0f;
    uint32_t idx = static_cast<uint32_t>(base);
    for (int j = 0; j < kCollectPer; ++j) {
        if (!((keep >> j) & 1u)) continue;
        if (idx < static_cast<uint32_t>(G.sample_capacity)) {
            const float4* r = G.rec + 4 * (first + j * kBlock);
            float4* d = G.samples + 4 * static_cast<int64_t>(idx);
            d[0] = F4(XYZ(r[0]), w[j]);
            d[1] = F4(XYZ(r[1]), 0.0f);
            d[2] = F4(XYZ(r[2]), 0.0f);
            d[3] = F4(XYZ(r[3]), 0.0f);
            ws += w[j];
        }
        ++idx;
    }
    // The weight sum of the stored samples: one atomic per workgroup.
    for (int off = 32; off > 0; off >>= 1) ws += __shfl_xor(ws, off);
    if ((threadIdx.x & 63u) == 0) wsum_lds[threadIdx.x >> 6] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int k = 0; k < kBlock / 64; ++k) t += wsum_lds[k];
        if (t != 0.0f) atomicAdd(G.sample_wsum, t);
    }
}

// Adds the visible next-event contributions of this bounce, emissive, then
// analytic, then environment, as EstimateDirectLight sums them.
__global__ __launch_bounds__(kBlock) void ResolveKernel(PassView P, yrg::GuideView G, int depth, int cur) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int n = static_cast<int>(P.counters[cur]);
    unsigned shadow = 0, occluded = 0;
    if (i < n) {
        const int p = P.q_path[cur][i];
        const int4 slots = P.p_slot[p];
        if (slots.x >= 0 || slots.y >= 0 || slots.z >= 0) {
            float3 direct = V(0.0f, 0.0f, 0.0f);
            for (int k = 0; k < kShadowKinds; ++k) {
                const int s = k == kShadowEmissive ? slots.x : (k == kShadowAnalytic ? slots.y : slots.z);
                if (s < 0) continue;
                ++shadow;
                if (P.s_occ[s] != 0u) {
                    ++occluded;
                    continue;
                }
                const float4 li = P.s_li[s];
                const float3 c = Mul(Hadamard(XYZ(P.s_bsdf[s]), Hadamard(XYZ(P.s_trans[s]), XYZ(li))), li.w);
                direct = Add(direct, c);
            }
            const float3 tp = XYZ(P.p_nee[p]);
            const float3 rad = Add(XYZ(P.p_rad[p]), Hadamard(tp, direct));
            P.p_rad[p] = F4(rad, 0.0f);
            if (G.rec != nullptr) {
                // This vertex's own next-event estimate is not radiance along
                // its sampled direction: move the record's snapshot past it.
                float4* r = G.rec + 4 * (static_cast<int64_t>(depth) * G.rec_paths + p);
                if (__float_as_int(r[2].w) == G.stamp) r[1].w = Lum(rad);
            }
        }
    }
    WaveStat(P.stats, kStatShadow, shadow);
    WaveStat(P.stats, kStatOccluded, occluded);
}

// ---- transparent shadows (TraceShadowVisibility, path_direct_lighting.cpp:23-113) ----
// One persistent launch walks every shadow ray through its dielectric surfaces:
// the traversal hands each closest hit to ShadowWalk::Finish, which applies the
// reference's per-hit rule and returns the next segment (same lane) until the
// ray is decided. l_state[e] = (0, transparent hits so far, finite, 0);
// l_med[e] = medium (color, distance; distance < 0 = inactive); s_trans holds
// the transmittance and the remaining distance in .w.

__device__ __forceinline__ bool TransBlack(float3 c) { return MaxC(c) <= 1.0e-6f; }

template <bool kAlpha, bool kTex>
struct ShadowWalk {
    const PassView* P;
    const SurfaceView* S;
    const LightView* L;

    __device__ __forceinline__ void Begin(int32_t e, float4& o, float4& d) const {
        o = P->s_org[e];
        d = P->s_dir[e];
        P->l_state[e] = make_int4(0, 0, isfinite(d.w) ? 1 : 0, 0);
        P->l_med[e] = make_float4(1.0f, 1.0f, 1.0f, -1.0f);
        P->s_trans[e] = make_float4(1.0f, 1.0f, 1.0f, d.w);
    }

    __device__ __forceinline__ bool Finish(int32_t e, const trav::Best& best, int32_t inst, uint32_t, uint32_t,
                                           uint32_t, const float4& lo, const float4& ld, float4& next_org,
                                           float4& next_dir) const {
        int4 st = P->l_state[e];
        const float4 tr = P->s_trans[e];
        float3 T = XYZ(tr);
        float remaining = tr.w;
        float4 med = P->l_med[e];
        const bool finite = st.z != 0;
        uint32_t occ = 1u;
        bool again = false;
        bool hit = best.hit;
        if (kAlpha && hit && !trav::IsSphereId(best.id) &&
            !(HitAlpha(*S, *L, best.id, inst, best.u, best.v, XYZ(lo), XYZ(ld), best.t, nullptr) >= 0.5f)) {
            // TraceVisibleSurface's alpha skip within this segment
            // (surface_query.cpp:187-199): trace again from just past the
            // hit; past the segment's end it is a miss, and the 64th skipped
            // hit leaves the segment exhausted, which blocks the shadow ray.
            const float t_min = best.t + kAlphaSkipOffset;
            st.w += 1;
            if (t_min >= remaining) {
                hit = false;
            } else if (st.w >= kMaxAlphaSkips) {
                P->s_occ[e] = 1u;
                P->l_state[e] = st;
                return false;
            } else {
                P->l_state[e] = st;
                next_org = make_float4(lo.x, lo.y, lo.z, t_min);
                next_dir = make_float4(ld.x, ld.y, ld.z, remaining);
                return true;
            }
        }
        if (!hit) {
            bool ok = true;
            if (med.w >= 0.0f) {
                if (!finite) {
                    ok = false;
                } else {
                    T = Hadamard(T, BeerLambert(XYZ(med), med.w, remaining));
                }
            }
            occ = ok && !TransBlack(T) ? 0u : 1u;
        } else {
            bool ok = true;
            if (med.w >= 0.0f) {
                T = Hadamard(T, BeerLambert(XYZ(med), med.w, best.t));
                if (TransBlack(T)) ok = false;
            }
            const int mat_index = trav::IsSphereId(best.id) ? S->sphere_info[-2 - best.id].x
                                                             : S->prim_info[S->tri_prim[best.id]].x;
            DevMaterial mat = L->materials[mat_index];
            if (ok && mat.kind != kDielectric && mat.kind != kThinDielectric) ok = false;
            if (kTex && ok && mat.kind == kThinDielectric && !trav::IsSphereId(best.id)) {
                // The hit's material has its reflectance texture applied
                // (EvaluateMaterialAtSurface); shadow rays carry no footprint.
                const int rt = L->mat_tex[mat_index].x;
                const int prim = S->tri_prim[best.id];
                if (rt >= 0 && (S->prim_flags[prim] & kHasUvs)) {
                    const int4 pinfo = S->prim_info[prim];
                    const int64_t base = static_cast<int64_t>(pinfo.z) + static_cast<int64_t>(best.id - pinfo.w) * 3;
                    const float2 uv = HitUv(*S, S->indices[base], S->indices[base + 1], S->indices[base + 2], best.u,
                                            best.v);
                    mat.reflectance =
                        Hadamard(mat.reflectance, XYZ(tex::Sample(L->tex, rt, uv.x, uv.y, tex::NoFootprint())));
                }
            }
            if (ok) {
                if (mat.kind == kThinDielectric) {
                    T = Hadamard(T, V(Clampf(mat.reflectance.x, 0.0f, 1.0f), Clampf(mat.reflectance.y, 0.0f, 1.0f),
                                      Clampf(mat.reflectance.z, 0.0f, 1.0f)));
                    if (TransBlack(T)) ok = false;
                } else {
                    med = med.w >= 0.0f ? make_float4(1.0f, 1.0f, 1.0f, -1.0f)
                                        : F4(mat.absorption_color, mat.absorption_distance);
                }
            }
            if (ok) {
                const float3 hp = Add(XYZ(lo), Mul(XYZ(ld), best.t));
                const float bias = SurfaceBias(hp);
                bool done = false;
                if (finite) {
                    remaining -= best.t + bias;
                    if (remaining <= 0.0f) {
                        occ = !TransBlack(T) ? 0u : 1u;
                        done = true;
                    }
                }
                if (!done) {
                    st.y += 1;
                    if (st.y >= kMaxTransparentShadowHits) {
                        occ = 1u;
                    } else {
                        again = true;
                        st.w = 0;  // a new segment: its own alpha skips
                        next_org = F4(Add(hp, Mul(XYZ(ld), bias)), 1.0e-5f);
                        next_dir = make_float4(ld.x, ld.y, ld.z, remaining);
                    }
                }
            }
        }
        P->s_occ[e] = occ;
        P->s_trans[e] = F4(T, remaining);
        P->l_med[e] = med;
        P->l_state[e] = st;
        return again;
    }
};

// kAlpha: the scene has alpha cut-outs, which the walk skips as
// TraceVisibleSurface does within each segment; kTex: materials have
// textures (thin glass's reflectance texture scales its shadow
// transmittance). Scenes without either run the variant without that code
// (fewer registers).
template <bool kTwoLevel, bool kSpheres, bool kAlpha, bool kTex>
__global__ __launch_bounds__(trav::kMaxBlock) __attribute__((amdgpu_waves_per_eu(kAlpha || kTex ? 1 : 5))) void ShadowWalkKernel(trav::TraceArgs a, PassView P, SurfaceView S,
                                                                    LightView L) {
    extern __shared__ int2 lds_stack[];
    const ShadowWalk<kAlpha, kTex> pol{&P, &S, &L};
    if (kTwoLevel) {
        trav::TraceTwoLevelLoop<false, kSpheres>(a, pol, lds_stack);
    } else {
        trav::TraceFlatLoop<false, kSpheres>(a, pol, lds_stack);
    }
}

template <bool kTwoLevel, bool kSpheres>
void LaunchShadowWalkAs(bool alpha, bool tex, int grid, int block, size_t lds, hipStream_t st,
                        const trav::TraceArgs& a, const PassView& P, const SurfaceView& S, const LightView& L) {
    if (alpha && tex) {
        hipLaunchKernelGGL((ShadowWalkKernel<kTwoLevel, kSpheres, true, true>), dim3(grid), dim3(block), lds, st, a, P, S, L);
    } else if (alpha) {
        hipLaunchKernelGGL((ShadowWalkKernel<kTwoLevel, kSpheres, true, false>), dim3(grid), dim3(block), lds, st, a, P, S, L);
    } else if (tex) {
        hipLaunchKernelGGL((ShadowWalkKernel<kTwoLevel, kSpheres, false, true>), dim3(grid), dim3(block), lds, st, a, P, S, L);
    } else {
        hipLaunchKernelGGL((ShadowWalkKernel<kTwoLevel, kSpheres, false, false>), dim3(grid), dim3(block), lds, st, a, P, S, L);
    }
}

template <bool kTwoLevel>
void LaunchShadowWalk(bool spheres, bool alpha, bool tex, int grid, int block, size_t lds, hipStream_t st,
                      const trav::TraceArgs& a, const PassView& P, const SurfaceView& S, const LightView& L) {
    if (spheres) {
        LaunchShadowWalkAs<kTwoLevel, true>(alpha, tex, grid, block, lds, st, a, P, S, L);
    } else {
        LaunchShadowWalkAs<kTwoLevel, false>(alpha, tex, grid, block, lds, st, a, P, S, L);
    }
}

// ---- path segments through alpha cut-outs (TraceVisibleSurface, surface_query.cpp:101-200)
// The closest-hit traversal of a bounce's path queue hands each closest hit
// to AlphaPath::Finish: a triangle hit whose alpha is below 0.5 is skipped by
// tracing again from just past it (same lane), at most 64 times; running out
// of skips, or past t_max, the segment is a miss, which is what the path loop
// makes of both. hit[ray].w holds the skip count until the ray is done.
struct AlphaPath {
    const trav::TraceArgs* a;
    const SurfaceView* S;
    const LightView* L;
    const float4* diff;     // camera-ray differentials per path (first bounce only), or nullptr
    const int32_t* q_path;  // queue entry -> path

    __device__ __forceinline__ void Begin(int32_t ray, float4& o, float4& d) const {
        o = a->org[ray];
        d = a->dir[ray];
        a->hit[ray] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
    }

    __device__ __forceinline__ bool Finish(int32_t ray, const trav::Best& best, int32_t inst, uint32_t, uint32_t,
                                           uint32_t, const float4& lo, const float4& ld, float4& next_org,
                                           float4& next_dir) const {
        bool hit = best.hit;
        if (hit && !trav::IsSphereId(best.id)) {
            const float4* df = diff != nullptr ? diff + 2 * q_path[ray] : nullptr;
            if (!(HitAlpha(*S, *L, best.id, inst, best.u, best.v, XYZ(lo), XYZ(ld), best.t, df) >= 0.5f)) {
                const int skips = __float_as_int(a->hit[ray].w) + 1;
                const float t_min = best.t + kAlphaSkipOffset;
                const float t_max = a->dir[ray].w;
                if (t_min < t_max && skips < kMaxAlphaSkips) {
                    a->hit[ray] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(skips));
                    next_org = make_float4(lo.x, lo.y, lo.z, t_min);
                    next_dir = make_float4(ld.x, ld.y, ld.z, t_max);
                    return true;
                }
                hit = false;
            }
        }
        a->hit[ray] = make_float4(best.t, best.u, best.v, __int_as_float(hit ? best.id : -1));
        if (a->inst_out != nullptr) a->inst_out[ray] = hit ? inst : -1;
        return false;
    }
};

template <bool kTwoLevel, bool kSpheres>
__global__ __launch_bounds__(trav::kMaxBlock) void AlphaTraceKernel(trav::TraceArgs a, SurfaceView S, LightView L,
                                                                    const float4* diff, const int32_t* q_path) {
    extern __shared__ int2 lds_stack[];
    const AlphaPath pol{&a, &S, &L, diff, q_path};
    if (kTwoLevel) {
        trav::TraceTwoLevelLoop<false, kSpheres>(a, pol, lds_stack);
    } else {
        trav::TraceFlatLoop<false, kSpheres>(a, pol, lds_stack);
    }
}

void LaunchAlphaTrace(bool two_level, bool spheres, int grid, int block, size_t lds, hipStream_t st,
                      const trav::TraceArgs& a, const SurfaceView& S, const LightView& L, const float4* diff,
                      const int32_t* q_path) {
    if (two_level) {
        if (spheres) {
            hipLaunchKernelGGL((AlphaTraceKernel<true, true>), dim3(grid), dim3(block), lds, st, a, S, L, diff, q_path);
        } else {
            hipLaunchKernelGGL((AlphaTraceKernel<true, false>), dim3(grid), dim3(block), lds, st, a, S, L, diff, q_path);
        }
    } else if (spheres) {
        hipLaunchKernelGGL((AlphaTraceKernel<false, true>), dim3(grid), dim3(block), lds, st, a, S, L, diff, q_path);
    } else {
        hipLaunchKernelGGL((AlphaTraceKernel<false, false>), dim3(grid), dim3(block), lds, st, a, S, L, diff, q_path);
    }
}

template <bool kGuided, bool kAnalytic>
void LaunchShade(bool textured, int grid, hipStream_t st, const PassView& P, const SurfaceView& S, const LightView& L,
                 const yrg::GuideView& G, int depth, int cur) {
    if (textured) {
        hipLaunchKernelGGL((ShadeKernel<kGuided, kAnalytic, true>), dim3(grid), dim3(kBlock), 0, st, P, S, L, G, depth,
                           cur);
    } else {
        hipLaunchKernelGGL((ShadeKernel<kGuided, kAnalytic, false>), dim3(grid), dim3(kBlock), 0, st, P, S, L, G, depth,
                           cur);
    }
}

// Film::AddSample (src/film/film.cpp:32-56) of path p's sample into pixel px,
// in FP64; returns 1 for a dropped non-finite sample.
__device__ __forceinline__ unsigned AddSample(const PassView& P, int64_t p, yrh_film_pixel& px) {
    unsigned bad = 0;
    {
        float3 b = XYZ(P.p_rad[p]);
        if (P.radiance_clamp > 0.0f) {
            const float m = MaxC(b);
            if (!(m <= P.radiance_clamp || m <= 0.0f)) b = Mul(b, P.radiance_clamp / m);
        }
        if (!Finite3(b)) {
            ++bad;
        } else {
            px.sum[0] += static_cast<double>(b.x);
            px.sum[1] += static_cast<double>(b.y);
            px.sum[2] += static_cast<double>(b.z);
            px.samples += 1u;
            const double lum = 0.2126 * static_cast<double>(b.x) + 0.7152 * static_cast<double>(b.y) +
                               0.0722 * static_cast<double>(b.z);
            const double delta = lum - px.luminance_mean;
            px.luminance_mean += delta / static_cast<double>(px.samples);
            px.luminance_m2 += delta * (lum - px.luminance_mean);
            const int flags = P.p_flags[p];
            if (flags & kPrimaryHit) {
                const float4 a = P.p_alb[p];
                const float4 nm = P.p_nrm[p];
                if (Finite3(XYZ(a)) && Finite3(XYZ(nm)) && isfinite(a.w) && a.w >= 0.0f) {
                    px.albedo_sum[0] += a.x;
                    px.albedo_sum[1] += a.y;
                    px.albedo_sum[2] += a.z;
                    px.normal_sum[0] += nm.x;
                    px.normal_sum[1] += nm.y;
                    px.normal_sum[2] += nm.z;
                    px.depth_sum += a.w;
                    px.aov_samples += 1u;
                }
            }
        }
    }
    return bad;
}

// One thread per pixel, the pass's samples in index order.
// Film::IsConverged (src/film/film.cpp:113-121) on the pixel's statistics
// rounded to FP32 (the reference keeps its film in FP32): the standard error
// of the luminance mean, scaled by the confidence, within the absolute plus
// relative tolerance, after at least max(2, min_spp) samples.
__device__ __forceinline__ bool PixelConverged(const yrh_film_pixel& px, int32_t min_spp, float rel, float abs_err,
                                               float confidence) {
    const uint32_t n = px.samples;
    if (n < static_cast<uint32_t>(min_spp > 2 ? min_spp : 2)) return false;
    const float mean = static_cast<float>(px.luminance_mean);
    const float m2 = static_cast<float>(px.luminance_m2);
    const float variance = m2 / static_cast<float>(n - 1);  // LuminanceVariance, n > 1
    const float stderr_mean = sqrtf(fmaxf(0.0f, variance) / static_cast<float>(n));
    const float bound = fmaxf(0.0f, confidence) * stderr_mean;
    const float tolerance = fmaxf(0.0f, abs_err) + fmaxf(0.0f, rel) * fabsf(mean);
    return isfinite(bound) && bound <= tolerance;
}

__global__ __launch_bounds__(kBlock) void AccumulateKernel(PassView P) {
    const int64_t pixel = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned bad = 0;
    if (P.active_list != nullptr) {
        // Pixel-list pass: thread e takes listed pixel e and adds its samples in
        // index order, as that many one-sample passes would: it stops once the
        // pixel converges (the reference would stop rendering it) or reaches
        // spp. A non-finite sample is not added, so the reference would render
        // the same index in each remaining pass: those count as bad too.
        const int64_t e = pixel;
        const int64_t n = *P.active_count;
        if (e < n) {
            // The pixel's record stays in registers over its samples (one
            // load, one store: a read-modify-write in memory per sample made
            // each sample wait on the previous one's store).
            yrh_film_pixel* dst = P.film + P.active_list[e];
            yrh_film_pixel px = *dst;
            for (int k = 0; k < P.sample_count; ++k) {
                if (px.samples >= static_cast<uint32_t>(P.spp)) break;
                if (AddSample(P, static_cast<int64_t>(k) * n + e, px) != 0u) {
                    bad += static_cast<unsigned>(P.sample_count - k);
                    break;
                }
                if (P.act_on && PixelConverged(px, P.act_min_spp, P.act_rel, P.act_abs, P.act_conf)) break;
            }
            *dst = px;
        }
    } else if (pixel < P.npix) {
        yrh_film_pixel px = P.film[pixel];
        for (int k = 0; k < P.sample_count; ++k) bad += AddSample(P, static_cast<int64_t>(k) * P.npix + pixel, px);
        P.film[pixel] = px;
    }
    WaveStat(P.stats, kStatBad, bad);
}

__global__ void ZeroWordsKernel(uint32_t* w, int n) {
    if (static_cast<int>(threadIdx.x) < n) w[threadIdx.x] = 0u;
}

// The active-pixel list of RenderCpuPathTrace (src/backends/cpu/cpu_path_tracer.cpp:95-106,
// 183-199): pixels that are not converged (adaptive sampling on) and have
// fewer than spp samples; counts[1] = converged pixels. One atomic per
// workgroup for each count; the list holds each workgroup's pixels in order.
__global__ __launch_bounds__(kBlock) void UpdateActiveKernel(const yrh_film_pixel* film, int64_t npix, int32_t adaptive,
                                                             int32_t min_spp, float rel, float abs_err,
                                                             float confidence, uint32_t spp, int32_t* list,
                                                             uint32_t* counts) {
    __shared__ uint32_t append_lds[kBlock / 64 + 1];
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool converged = false, active = false;
    if (k < npix) {
        const yrh_film_pixel& px = film[k];
        converged = adaptive != 0 && PixelConverged(px, min_spp, rel, abs_err, confidence);
        active = !converged && px.samples < spp;
    }
    const int32_t slot = BlockAppend(counts, active ? 1u : 0u, append_lds);
    if (active) list[slot] = static_cast<int32_t>(k);
    (void)BlockAppend(counts + 1, converged ? 1u : 0u, append_lds);
}

// Batched pass: one thread per path of the given occurrence. The items of a
// job within a pass are consecutive sample layers, numbered by occurrence;
// within one occurrence a (job, pixel) appears once, and the occurrences are
// accumulated in order, so the film updates never collide and every pixel
// receives its samples in index order.
__global__ __launch_bounds__(kBlock) void AccumulateBatchKernel(PassView P, int occurrence) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned bad = 0;
    if (p < static_cast<int64_t>(P.counters[kCountPassPaths]) && (P.p_job[p] >> 16) == occurrence) {
        bad = AddSample(P, p, P.film[P.p_pixel[p]]);
    }
    WaveStat(P.stats, kStatBad, bad);
}

__global__ void EndPassKernel(uint32_t* counters) {
    if (threadIdx.x == 0) counters[kPassIndex] += 1u;
}

__global__ void SampleValuesKernel(int kind, uint64_t seed, int spp, int64_t n, const int32_t* x, const int32_t* y,
                                   const int32_t* sample, const int32_t* dim, int dims, float* out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PixelSampler s;
    s.Init(kind, seed, x[i], y[i], sample[i], spp);
    if (dims == 1) {
        out[i] = s.At1D(static_cast<uint32_t>(dim[i]), 0);
    } else {
        const float2 v = s.At2D(static_cast<uint32_t>(dim[i]), 0);
        out[2 * i] = v.x;
        out[2 * i + 1] = v.y;
    }
}

int Grid(int64_t n, int block = kBlock) { return static_cast<int>(std::max<int64_t>(1, (n + block - 1) / block)); }

PassView MakePassView(RenderState& R, const yrh_render_desc& d) {
    PassView P{};
    P.width = d.width;
    P.height = d.height;
    P.sampler = d.sampler;
    P.spp = d.spp;
    P.max_depth = d.max_depth;
    P.sample_index = d.sample_index;
    P.sample_count = d.sample_count < 1 ? 1 : d.sample_count;
    P.npix = static_cast<int64_t>(d.width) * d.height;
    P.seed = d.seed;
    P.radiance_clamp = d.radiance_clamp;
    P.camera.origin = make_float3(d.camera_origin[0], d.camera_origin[1], d.camera_origin[2]);
    P.camera.forward = make_float3(d.camera_forward[0], d.camera_forward[1], d.camera_forward[2]);
    P.camera.right = make_float3(d.camera_right[0], d.camera_right[1], d.camera_right[2]);
    P.camera.up = make_float3(d.camera_up[0], d.camera_up[1], d.camera_up[2]);
    P.camera.tan_half_fov = d.tan_half_fov;
    P.p_pixel = R.p_pixel.as<int32_t>();
    P.p_tp = R.p_tp.as<float4>();
    P.p_rad = R.p_rad.as<float4>();
    P.p_prev = R.p_prev.as<float4>();
    P.p_flags = R.p_flags.as<int32_t>();
    P.p_med = R.p_med.as<float4>();
    P.p_alb = R.p_alb.as<float4>();
    P.p_nrm = R.p_nrm.as<float4>();
    P.p_nee = R.p_nee.as<float4>();
    P.p_slot = R.p_slot.as<int4>();
    P.p_diff = R.footprints ? R.p_diff.as<float4>() : nullptr;
    for (int k = 0; k < 2; ++k) {
        P.q_org[k] = R.q_org[k].as<float4>();
        P.q_dir[k] = R.q_dir[k].as<float4>();
        P.q_path[k] = R.q_path[k].as<int32_t>();
    }
    P.hits = R.hits.as<float4>();
    P.hit_inst = R.hit_inst.as<int32_t>();
    P.s_org = R.s_org.as<float4>();
    P.s_dir = R.s_dir.as<float4>();
    P.s_li = R.s_li.as<float4>();
    P.s_bsdf = R.s_bsdf.as<float4>();
    P.s_occ = R.s_occ.as<uint32_t>();
    P.s_trans = R.s_trans.as<float4>();
    P.l_state = R.l_state.as<int4>();
    P.l_med = R.l_med.as<float4>();
    P.counters = R.counters.as<uint32_t>();
    P.stats = R.stats.as<unsigned long long>();
    P.film = R.film.as<yrh_film_pixel>();
    P.p_job = R.p_job.as<int32_t>();
    P.p_sample = R.p_sample.as<int32_t>();
    if (d.pixel_list != 0) {
        P.active_list = R.act_list.as<int32_t>();
        P.active_count = R.act_count.as<uint32_t>();
        P.act_on = R.act_desc.enabled;
        P.act_min_spp = R.act_desc.min_spp;
        P.act_rel = R.act_desc.relative_error;
        P.act_abs = R.act_desc.absolute_error;
        P.act_conf = R.act_desc.confidence;
    }
    return P;
}

SurfaceView MakeSurfaceView(RenderState& R) {
    SurfaceView S{};
    S.positions = R.positions.as<float4>();
    S.normals = R.normals.as<float4>();
    S.indices = R.indices.as<uint32_t>();
    S.tri_prim = R.tri_prim.as<int32_t>();
    S.spheres = R.spheres.as<float4>();
    S.sphere_info = R.sphere_info.as<int2>();
    S.prim_info = R.prim_info.as<int4>();
    S.prim_flags = R.prim_flags.as<int32_t>();
    S.inst_normal = R.inst_normal.as<float4>();
    S.uvs = R.uvs.as<float2>();
    S.tangents = R.tangents.as<float4>();
    return S;
}

LightView MakeLightView(RenderState& R) {
    LightView L{};
    L.materials = R.materials.as<DevMaterial>();
    L.mat_tex = R.textured || R.alpha ? R.mat_tex.as<int4>() : nullptr;
    L.tex = tex::TexView{R.tex_records.as<int4>(), R.tex_levels.as<int4>(), R.tex_texels.as<float4>(),
                         R.tex_gauss.as<float>(), R.texture_count};
    L.emit = R.emit.as<float4>();
    L.emit_info = R.emit_info.as<int4>();
    L.emit_area = R.emit_area.as<float>();
    L.emit_cdf = R.emit_cdf.as<float>();
    L.alias_prob = R.alias_prob.as<float>();
    L.alias_index = R.alias_index.as<int32_t>();
    L.alias_pmf = R.alias_pmf.as<float>();
    L.emissive_count = R.emissive_count;
    L.total_emissive_area = R.total_emissive_area;
    L.alight = R.alight.as<float4>();
    L.alight_prob = R.alight_prob.as<float>();
    L.alight_alias = R.alight_alias.as<int32_t>();
    L.alight_pmf = R.alight_pmf.as<float>();
    L.alight_tree = R.alight_tree.as<float4>();
    L.analytic_count = R.analytic_count;
    L.analytic_pmf_count = R.analytic_pmf_count;
    L.analytic_tree_count = R.analytic_tree_count;
    L.analytic_tree_root = R.analytic_tree_root;
    L.use_analytic_tree = R.use_analytic_tree;
    L.env_active = R.env_active;
    L.env_radiance = make_float3(R.env_radiance[0], R.env_radiance[1], R.env_radiance[2]);
    L.env_strength = R.env_strength;
    L.env_rotation = R.env_rotation;
    L.env_texels = R.env_w > 0 && R.env_h > 0 ? R.env_texels.as<float4>() : nullptr;
    L.env_w = R.env_w;
    L.env_h = R.env_h;
    L.env_nearest = R.env_nearest;
    L.env_ws = R.env_ws;
    L.env_wt = R.env_wt;
    L.env_sampleable = R.env_sampleable;
    L.env_weights = R.env_weights.as<float>();
    L.env_row_weights = R.env_row_weights.as<float>();
    L.env_row_cdf = R.env_row_cdf.as<float>();
    L.env_row_top = R.env_row_top.as<float>();
    L.env_cond_cdf = R.env_cond_cdf.as<float>();
    L.env_cond_top = R.env_cond_top.as<float>();
    L.env_rtop = R.env_rtop;
    L.env_cstride = R.env_cstride;
    L.env_ctop = R.env_ctop;
    L.env_dw = R.env_dw;
    L.env_dh = R.env_dh;
    L.env_total = R.env_total;
    return L;
}

DevMaterial ToDevMaterial(const yrh_material& m) {
    DevMaterial d;
    d.kind = m.kind;
    d.reflectance = make_float3(m.reflectance[0], m.reflectance[1], m.reflectance[2]);
    d.ior = m.ior;
    d.urough = m.uroughness;
    d.vrough = m.vroughness;
    d.alpha = m.alpha;
    d.emission = make_float3(m.emission[0], m.emission[1], m.emission[2]);
    d.absorption_distance = m.absorption_distance;
    d.absorption_color = make_float3(m.absorption_color[0], m.absorption_color[1], m.absorption_color[2]);
    return d;
}

// One BSDF query per thread (yrh_bsdf_query): the device EvalBsdf, PdfBsdf,
// SampleBsdf and IsDelta the shading kernels run, on caller tuples.
// SampleTexture4 of texture 0 at n points (uv 2 floats, footprint 4 floats or none).
__global__ void TextureQueryKernel(tex::TexView T, int64_t n, const float* uv, const float* fp, float4* out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const tex::Footprint f = fp != nullptr ? tex::Footprint{fp[4 * i], fp[4 * i + 1], fp[4 * i + 2], fp[4 * i + 3]}
                                           : tex::NoFootprint();
    out[i] = tex::Sample(T, 0, uv[2 * i], uv[2 * i + 1], f);
}

__global__ void BsdfQueryKernel(const DevMaterial* mats, int64_t n, const int32_t* mat, const float* q, float* out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevMaterial m = mats[mat[i]];
    const float* a = q + 12 * i;
    const float3 wo = V(a[0], a[1], a[2]), wi = V(a[3], a[4], a[5]), nn = V(a[6], a[7], a[8]);
    const float3 f = EvalBsdf(m, wo, wi, nn);
    const float pdf = PdfBsdf(m, wo, wi, nn);
    const BsdfSample s = SampleBsdf(m, wo, nn, make_float2(a[9], a[10]));
    float* o = out + 12 * i;
    o[0] = f.x;
    o[1] = f.y;
    o[2] = f.z;
    o[3] = pdf;
    o[4] = s.wi.x;
    o[5] = s.wi.y;
    o[6] = s.wi.z;
    o[7] = s.weight.x;
    o[8] = s.weight.y;
    o[9] = s.weight.z;
    o[10] = s.pdf;
    o[11] = static_cast<float>((s.valid ? 1 : 0) + (s.specular ? 2 : 0) + (IsDelta(m) ? 4 : 0));
}

// Bytes per path slot and per shadow slot of the arrays EnsureCapacity
// allocates, and per guided row (mixture + valid flag).
constexpr size_t kPathSlotBytes = 4 + 4 + 4 + 16 + 16 + 16 + 4 + 16 + 16 + 16 + 16 + 16 + 2 * 16 + 2 * 16 + 2 * 4 + 16 + 4;
constexpr size_t kShadowSlotBytes = 16 + 16 + 16 + 16 + 4 + 16 + 16 + 16;
constexpr size_t kGuideRowBytes = yrg::kGuideMixtureRow * 4 + 4;

// The camera-ray differentials (two float4 per path) of scenes whose textures
// filter with footprints.
int EnsureDifferentials(yrh_context* ctx, RenderState& R) {
    const size_t bytes = 32 * static_cast<size_t>(R.capacity);
    if (!R.footprints || R.p_diff.bytes >= bytes) return YRH_OK;
    int rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return rc != YRH_OK ? rc : Alloc(ctx, R.p_diff, bytes);
}

int EnsureCapacity(yrh_context* ctx, RenderState& R, int64_t paths) {
    if (R.capacity >= paths && R.cap_shadow_slots >= R.shadow_slots && R.counters.ptr != nullptr) {
        return EnsureDifferentials(ctx, R);
    }
    int rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (rc != YRH_OK) return rc;
    paths = std::max(paths, R.capacity);
    const size_t P = static_cast<size_t>(paths);
    const size_t S = static_cast<size_t>(R.shadow_slots) * P;
    struct Item {
        DevBuf* buf;
        size_t bytes;
    } items[] = {
        {&R.p_pixel, 4 * P}, {&R.p_job, 4 * P}, {&R.p_sample, 4 * P}, {&R.p_tp, 16 * P}, {&R.p_rad, 16 * P}, {&R.p_prev, 16 * P}, {&R.p_flags, 4 * P},
        {&R.p_med, 16 * P}, {&R.p_alb, 16 * P}, {&R.p_nrm, 16 * P}, {&R.p_nee, 16 * P}, {&R.p_slot, 16 * P},
        {&R.q_org[0], 16 * P}, {&R.q_org[1], 16 * P}, {&R.q_dir[0], 16 * P}, {&R.q_dir[1], 16 * P},
        {&R.q_path[0], 4 * P}, {&R.q_path[1], 4 * P}, {&R.hits, 16 * P}, {&R.hit_inst, 4 * P},
        {&R.s_org, 16 * S}, {&R.s_dir, 16 * S}, {&R.s_li, 16 * S}, {&R.s_bsdf, 16 * S}, {&R.s_occ, 4 * S},
        {&R.s_trans, 16 * S}, {&R.l_state, 16 * S}, {&R.l_med, 16 * S},
        {&R.counters, 4 * kCounterWords},
    };
    for (const Item& it : items) {
        if ((rc = Alloc(ctx, *it.buf, it.bytes)) != YRH_OK) return rc;
    }
    if (R.stats.ptr == nullptr) {
        if ((rc = Alloc(ctx, R.stats, kStatBytes)) != YRH_OK) return rc;
        if ((rc = CheckHip(ctx, hipMemsetAsync(R.stats.ptr, 0, kStatBytes, ctx->stream), "hipMemsetAsync")) != YRH_OK) {
            return rc;
        }
    }
    rc = CheckHip(ctx, hipMemsetAsync(R.counters.ptr, 0, 4 * kCounterWords, ctx->stream), "hipMemsetAsync");
    R.capacity = paths;
    R.cap_shadow_slots = R.shadow_slots;
    return rc != YRH_OK ? rc : EnsureDifferentials(ctx, R);
}

// Raygen and the bounces of one pass over `paths` path slots (the queue
// capacity; the kernels read the live counts from the counter words).
int LaunchPassSequence(yrh_context* ctx, RenderState& R, const PassView& P, const SurfaceView& S, const LightView& L,
                       const yrg::GuideView& GV, int64_t paths, int max_depth, bool overlap) {
    hipStream_t st = ctx->stream;
    const bool guided = GV.mode != YRH_GUIDE_OFF;
    int rc;
    // Overlap: a bounce's shadow rays and Resolve run on the second stream
    // while the main stream traces the next bounce's path segments (and
    // queries the guide); Shade of the next bounce waits for that Resolve.
    // Every per-path operation keeps its order, so the film is the same.
    hipStream_t ss = st;
    if (overlap) {
        if (ctx->aux_stream == nullptr &&
            (rc = CheckHip(ctx, hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking), "hipStreamCreate")) !=
                YRH_OK) {
            return rc;
        }
        for (hipEvent_t& e : ctx->bounce_events) {
            if (e == nullptr &&
                (rc = CheckHip(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate")) != YRH_OK) {
                return rc;
            }
        }
        ss = ctx->aux_stream;
    }
    hipEvent_t shaded = overlap ? ctx->bounce_events[0] : nullptr;
    hipEvent_t resolved = overlap ? ctx->bounce_events[1] : nullptr;
    {
        const StageRange r("raygen");
        hipLaunchKernelGGL(RaygenKernel, dim3(Grid(paths)), dim3(kBlock), 0, st, P);
    }
    for (int depth = 0; depth < max_depth; ++depth) {
        const StageRange bounce_range("bounce");
        const int cur = depth & 1;
        hipLaunchKernelGGL(BeginBounceKernel, dim3(1), dim3(64), 0, st, P.counters, 1 - cur, kBeginTrace);
        TraceRequest tr;
        tr.any = false;
        tr.org = P.q_org[cur];
        tr.dir = P.q_dir[cur];
        tr.n_device = P.counters + cur;
        tr.capacity = paths;
        tr.expected = paths;
        tr.work_counter = P.counters + kWorkExtend;
        tr.parts = trav::kParts;
        tr.hit = P.hits;
        tr.instance = ctx->geo.accel_kind == YRH_ACCEL_TWO_LEVEL ? P.hit_inst : nullptr;
        tr.totals = P.stats + kStatTraceNodes;
        PassView Pk = P;
        if (ctx->geo.accel_kind != YRH_ACCEL_TWO_LEVEL) Pk.hit_inst = nullptr;
        if (depth > 0) Pk.p_diff = nullptr;  // only camera rays carry differentials
        if (R.alpha) {
            // Alpha cut-outs: the traversal skips invisible hits itself.
            trav::TraceArgs ta;
            int tgrid = 0, tblock = 0;
            if ((rc = PrepareTraceArgs(ctx, tr, ta, tgrid, tblock)) != YRH_OK) return rc;
            LaunchAlphaTrace(ctx->geo.accel_kind == YRH_ACCEL_TWO_LEVEL, ctx->geo.has_spheres, tgrid, tblock,
                             LdsBytes(ta, tblock), st, ta, S, L, Pk.p_diff, P.q_path[cur]);
        } else if ((rc = LaunchTrace(ctx, tr)) != YRH_OK) {
            return rc;
        }
        // Guided bounces: all that may continue, or the first depth_limit. A
        // plain vertex after guided ones is the plain estimator's vertex (its
        // NEE and emission weights use the pdf its direction was drawn from).
        const bool guide_here = guided && depth + 1 < max_depth && (GV.depth_limit <= 0 || depth < GV.depth_limit);
        const bool textured = L.mat_tex != nullptr;
        if (guide_here) {
            if (textured) {
                hipLaunchKernelGGL(GuideQueryKernel<true>, dim3(Grid(paths, kQueryBlock)), dim3(kQueryBlock), 0, st, Pk,
                                   S, L, GV, cur);
            } else {
                hipLaunchKernelGGL(GuideQueryKernel<false>, dim3(Grid(paths, kQueryBlock)), dim3(kQueryBlock), 0, st, Pk,
                                   S, L, GV, cur);
            }
        }
        // The previous bounce's Resolve reads this bounce's path queue count,
        // the path slots and the shadow queue, which Shade rewrites.
        if (overlap && depth > 0 && (rc = CheckHip(ctx, hipStreamWaitEvent(st, resolved, 0), "hipStreamWaitEvent")) != YRH_OK) {
            return rc;
        }
        hipLaunchKernelGGL(BeginBounceKernel, dim3(1), dim3(64), 0, st, P.counters, 1 - cur, kBeginShade);
        const bool analytic = L.analytic_count > 0;
        const int sgrid = Grid(paths);
        if (guide_here) {
            if (analytic) {
                LaunchShade<true, true>(textured, sgrid, st, Pk, S, L, GV, depth, cur);
            } else {
                LaunchShade<true, false>(textured, sgrid, st, Pk, S, L, GV, depth, cur);
            }
        } else if (analytic) {
            LaunchShade<false, true>(textured, sgrid, st, Pk, S, L, GV, depth, cur);
        } else {
            LaunchShade<false, false>(textured, sgrid, st, Pk, S, L, GV, depth, cur);
        }
        if (overlap) {
            if ((rc = CheckHip(ctx, hipEventRecord(shaded, st), "hipEventRecord")) != YRH_OK) return rc;
            if ((rc = CheckHip(ctx, hipStreamWaitEvent(ss, shaded, 0), "hipStreamWaitEvent")) != YRH_OK) return rc;
        }
        if (!R.transparent) {
            TraceRequest sr;
            sr.any = true;
            sr.org = P.s_org;
            sr.dir = P.s_dir;
            sr.n_device = P.counters + kCountShadow;
            sr.capacity = R.shadow_slots * paths;
            sr.expected = paths;  // about one shadow ray per path in flight
            sr.work_counter = P.counters + kWorkShadow;
            sr.parts = trav::kParts;
            sr.occluded = P.s_occ;
            sr.totals = P.stats + kStatTraceNodes;
            sr.aux = overlap;
            if ((rc = LaunchTrace(ctx, sr)) != YRH_OK) return rc;
        } else {
            TraceRequest wr;
            wr.any = false;
            wr.org = P.s_org;
            wr.dir = P.s_dir;
            wr.n_device = P.counters + kCountShadow;
            wr.capacity = R.shadow_slots * paths;
            wr.expected = paths;
            wr.work_counter = P.counters + kWorkShadow;
            wr.parts = trav::kParts;
            wr.hit = P.s_trans;  // not written by the walk; LaunchTrace's checks want an output
            wr.totals = P.stats + kStatTraceNodes;
            wr.aux = overlap;
            trav::TraceArgs wa;
            int wgrid = 0, wblock = 0;
            if ((rc = PrepareTraceArgs(ctx, wr, wa, wgrid, wblock)) != YRH_OK) return rc;
            if (ctx->geo.accel_kind == YRH_ACCEL_TWO_LEVEL) {
                LaunchShadowWalk<true>(ctx->geo.has_spheres, R.alpha, R.textured, wgrid, wblock, LdsBytes(wa, wblock), ss,
                                       wa, P, S, L);
            } else {
                LaunchShadowWalk<false>(ctx->geo.has_spheres, R.alpha, R.textured, wgrid, wblock, LdsBytes(wa, wblock), ss,
                                        wa, P, S, L);
            }
        }
        hipLaunchKernelGGL(ResolveKernel, dim3(Grid(paths)), dim3(kBlock), 0, ss, P, GV, depth, cur);
        if (overlap && (rc = CheckHip(ctx, hipEventRecord(resolved, ss), "hipEventRecord")) != YRH_OK) return rc;
    }
    if (overlap && (rc = CheckHip(ctx, hipStreamWaitEvent(st, resolved, 0), "hipStreamWaitEvent")) != YRH_OK) return rc;
    if ((rc = CheckHip(ctx, hipGetLastError(), "pass sequence launch")) != YRH_OK) return rc;
    if (GV.rec != nullptr) {
        hipLaunchKernelGGL(GuideCollectKernel, dim3(Grid(paths * max_depth, kBlock * kCollectPer)), dim3(kBlock), 0, st,
                           P, GV);
    }
    return YRH_OK;
}

// The device texture layout (yrh_texture.h) of `n` textures: level 0 as
// given, then the box-filtered chain down to 1 x 1 (BuildTextureMipChain,
// src/shading/texture.cpp:241-272), all in one pool; the EWA weight table as
// the reference tabulates it (texture.cpp:140-152).
struct TexturePool {
    std::vector<int4> records, levels;
    std::vector<float4> texels;
    std::vector<char> footprint;  // per texture: filters with footprints (trilinear / EWA with mips)
    float gauss[tex::kGaussEntries];
    std::string error;
};

bool BuildTexturePool(const yrh_texture* textures, int32_t ntex, TexturePool& pool) {
    std::vector<int4>& trec = pool.records;
    std::vector<int4>& tlev = pool.levels;
    std::vector<float4>& tpool = pool.texels;
    std::vector<char>& tex_footprint = pool.footprint;
    trec.assign(2 * static_cast<size_t>(ntex), make_int4(0, 0, 0, 0));
    tex_footprint.assign(static_cast<size_t>(ntex), 0);
    for (int32_t k = 0; k < ntex; ++k) {
        const yrh_texture& t = textures[k];
        if (t.filter < YRH_FILTER_NEAREST || t.filter > YRH_FILTER_EWA || t.wrap_s < YRH_WRAP_REPEAT ||
            t.wrap_s > YRH_WRAP_MIRRORED || t.wrap_t < YRH_WRAP_REPEAT || t.wrap_t > YRH_WRAP_MIRRORED) {
            pool.error = "texture filter or wrap mode out of range";
            return false;
        }
        const int first = static_cast<int>(tlev.size());
        int levels = 0;
        if (t.width > 0 && t.height > 0 && t.texels != nullptr) {
            int w = t.width, h = t.height;
            size_t off = tpool.size();
            tlev.push_back(make_int4(w, h, static_cast<int>(off), 0));
            const float4* src = reinterpret_cast<const float4*>(t.texels);
            tpool.insert(tpool.end(), src, src + static_cast<size_t>(w) * static_cast<size_t>(h));
            levels = 1;
            auto wrap = [](int i, int n, int mode) {
                if (mode == YRH_WRAP_CLAMP) return std::clamp(i, 0, n - 1);
                if (mode == YRH_WRAP_MIRRORED) {
                    int q = i % (2 * n);
                    if (q < 0) q += 2 * n;
                    return q < n ? q : 2 * n - 1 - q;
                }
                int q = i % n;
                return q < 0 ? q + n : q;
            };
            while (w > 1 || h > 1) {
                const int nw = std::max(1, (w + 1) / 2), nh = std::max(1, (h + 1) / 2);
                std::vector<float4> next(static_cast<size_t>(nw) * static_cast<size_t>(nh));
                auto at = [&](int x, int y) {
                    return tpool[off + static_cast<size_t>(wrap(y, h, t.wrap_t)) * static_cast<size_t>(w) +
                                 static_cast<size_t>(wrap(x, w, t.wrap_s))];
                };
                for (int y = 0; y < nh; ++y) {
                    for (int x = 0; x < nw; ++x) {
                        const float4 a = at(2 * x, 2 * y), b = at(2 * x + 1, 2 * y), c = at(2 * x, 2 * y + 1),
                                     d = at(2 * x + 1, 2 * y + 1);
                        next[static_cast<size_t>(y) * static_cast<size_t>(nw) + static_cast<size_t>(x)] =
                            make_float4((a.x + b.x + c.x + d.x) * 0.25f, (a.y + b.y + c.y + d.y) * 0.25f,
                                        (a.z + b.z + c.z + d.z) * 0.25f, (a.w + b.w + c.w + d.w) * 0.25f);
                    }
                }
                off = tpool.size();
                tpool.insert(tpool.end(), next.begin(), next.end());
                tlev.push_back(make_int4(nw, nh, static_cast<int>(off), 0));
                w = nw;
                h = nh;
                ++levels;
            }
            if (tpool.size() > static_cast<size_t>(std::numeric_limits<int32_t>::max())) {
                pool.error = "textures exceed 2^31 texels with their mip chains";
                return false;
            }
        }
        trec[2 * static_cast<size_t>(k)] = make_int4(t.width, t.height, t.filter | (t.wrap_s << 4) | (t.wrap_t << 8), levels);
        trec[2 * static_cast<size_t>(k) + 1] = make_int4(first, 0, 0, 0);
        tex_footprint[static_cast<size_t>(k)] = (t.filter == YRH_FILTER_TRILINEAR || t.filter == YRH_FILTER_EWA) && levels > 1;
    }
    {
        const float falloff = 2.0f;
        const float edge = std::exp(-falloff);
        for (int k = 0; k < tex::kGaussEntries; ++k) {
            const float r2 = static_cast<float>(k) / static_cast<float>(tex::kGaussEntries - 1);
            pool.gauss[k] = std::exp(-falloff * r2) - edge;
        }
    }
    return true;
}

}  // namespace

int UploadSurfaceGeometry(yrh_context* ctx, const yrh_geometry_desc* g) {
    RenderState& R = State(ctx);
    R.shading_loaded = false;
    const int64_t nv = g->vertex_count;
    // The shading kernels index vertices through every primitive's range.
    for (int64_t k = 0; k < g->index_count; ++k) {
        if (static_cast<int64_t>(g->indices[k]) >= nv) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "vertex index out of range");
    }
    for (int32_t p = 0; p < g->primitive_count; ++p) {
        const yrh_primitive& mp = g->primitives[p];
        if (mp.index_count % 3 != 0 ||
            static_cast<int64_t>(mp.first_index) + static_cast<int64_t>(mp.index_count) > g->index_count) {
            return Fail(ctx, YRH_E_INVALID_ARGUMENT, "primitive index range out of bounds");
        }
    }
    if (g->accel_kind == YRH_ACCEL_TWO_LEVEL && g->instance_count > 0 && g->instances == nullptr) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "instances missing");
    }
    std::vector<float4> pos(static_cast<size_t>(nv)), nrm(static_cast<size_t>(nv));
    R.host_positions.resize(static_cast<size_t>(3 * nv));
    for (int64_t v = 0; v < nv; ++v) {
        const yrh_vertex& vx = g->vertices[v];
        pos[static_cast<size_t>(v)] = make_float4(vx.position[0], vx.position[1], vx.position[2], 0.0f);
        nrm[static_cast<size_t>(v)] = make_float4(vx.normal[0], vx.normal[1], vx.normal[2], 0.0f);
        for (int c = 0; c < 3; ++c) R.host_positions[static_cast<size_t>(3 * v + c)] = vx.position[c];
    }
    R.host_indices.assign(g->indices, g->indices + g->index_count);
    std::vector<int4> info(static_cast<size_t>(g->primitive_count));
    std::vector<int32_t> flags(static_cast<size_t>(g->primitive_count));
    std::vector<int32_t> tri_prim;
    R.host_prim_first.assign(static_cast<size_t>(g->primitive_count), 0);
    R.host_prim_count.assign(static_cast<size_t>(g->primitive_count), 0);
    R.host_prim_material.assign(static_cast<size_t>(g->primitive_count), 0);
    int32_t flat = 0;
    bool any_uvs = false, any_tangents = false;
    for (int32_t p = 0; p < g->primitive_count; ++p) {
        const yrh_primitive& mp = g->primitives[p];
        const int32_t tris = static_cast<int32_t>(mp.index_count / 3);
        info[static_cast<size_t>(p)] = make_int4(mp.material_index, -1, static_cast<int32_t>(mp.first_index), flat);
        flags[static_cast<size_t>(p)] =
            (mp.has_normals ? kHasNormals : 0) | (mp.has_uvs ? kHasUvs : 0) | (mp.has_tangents ? kHasTangents : 0);
        any_uvs = any_uvs || mp.has_uvs;
        any_tangents = any_tangents || mp.has_tangents;
        R.host_prim_first[static_cast<size_t>(p)] = static_cast<int32_t>(mp.first_index);
        R.host_prim_count[static_cast<size_t>(p)] = tris;
        R.host_prim_material[static_cast<size_t>(p)] = mp.material_index;
        for (int32_t t = 0; t < tris; ++t) tri_prim.push_back(p);
        flat += tris;
    }
    R.primitive_count = g->primitive_count;
    R.triangle_count = flat;
    R.vertex_count = nv;
    // Instance normal transforms: cofactor rows of the upper 3x3 of object_to_world
    // and 1/det, as TransformNormal computes them per call.
    std::vector<float4> inst;
    if (g->accel_kind == YRH_ACCEL_TWO_LEVEL) {
        inst.resize(6 * static_cast<size_t>(g->instance_count));
        for (int32_t k = 0; k < g->instance_count; ++k) {
            const float* m = g->instances[k].object_to_world;
            const float a = m[0], b = m[4], c = m[8], d = m[1], e = m[5], f = m[9], gg = m[2], h = m[6], i = m[10];
            const float det = a * (e * i - f * h) - b * (d * i - f * gg) + c * (d * h - e * gg);
            const bool singular = std::fabs(det) <= 1.0e-12f;
            const float inv = singular ? 0.0f : 1.0f / det;
            float4* o = &inst[6 * static_cast<size_t>(k)];
            // .w of row 2: the orientation ResolveNormalMap gives the
            // tangent frame (material_evaluator.cpp:75-82), same determinant
            o[0] = make_float4(e * i - f * h, f * gg - d * i, d * h - e * gg, inv);
            o[1] = make_float4(c * h - b * i, a * i - c * gg, b * gg - a * h, singular ? 1.0f : 0.0f);
            o[2] = make_float4(b * f - c * e, c * d - a * f, a * e - b * d, det < 0.0f ? -1.0f : 1.0f);
            o[3] = make_float4(m[0], m[4], m[8], m[12]);
            o[4] = make_float4(m[1], m[5], m[9], m[13]);
            o[5] = make_float4(m[2], m[6], m[10], m[14]);
        }
    }
    int rc;
    if ((rc = UploadVec(ctx, R.positions, pos.data(), pos.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.normals, nrm.data(), nrm.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.indices, g->indices, static_cast<size_t>(g->index_count))) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.tri_prim, tri_prim.data(), tri_prim.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.prim_info, info.data(), info.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.prim_flags, flags.data(), flags.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.inst_normal, inst.data(), inst.size())) != YRH_OK) return rc;
    // Texture coordinates and tangent frames, when some primitive has them.
    std::vector<float2> uvs(any_uvs ? static_cast<size_t>(nv) : 0);
    std::vector<float4> tangents(any_tangents ? static_cast<size_t>(nv) : 0);
    for (int64_t v = 0; v < nv && (any_uvs || any_tangents); ++v) {
        const yrh_vertex& vx = g->vertices[v];
        if (any_uvs) uvs[static_cast<size_t>(v)] = make_float2(vx.uv[0], vx.uv[1]);
        if (any_tangents) {
            tangents[static_cast<size_t>(v)] =
                make_float4(vx.tangent[0], vx.tangent[1], vx.tangent[2], vx.tangent_handedness);
        }
    }
    if ((rc = UploadVec(ctx, R.uvs, uvs.data(), uvs.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.tangents, tangents.data(), tangents.size())) != YRH_OK) return rc;
    std::vector<float4> sph(static_cast<size_t>(g->sphere_count));
    std::vector<int2> sph_info(static_cast<size_t>(g->sphere_count));
    R.host_sphere_material.assign(static_cast<size_t>(g->sphere_count), 0);
    for (int32_t k = 0; k < g->sphere_count; ++k) {
        const yrh_sphere& sp = g->spheres[k];
        sph[static_cast<size_t>(k)] = make_float4(sp.center[0], sp.center[1], sp.center[2], sp.radius);
        sph_info[static_cast<size_t>(k)] = make_int2(sp.material_index, sp.flip_normals ? 1 : 0);
        R.host_sphere_material[static_cast<size_t>(k)] = sp.material_index;
    }
    if ((rc = UploadVec(ctx, R.spheres, sph.data(), sph.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.sphere_info, sph_info.data(), sph_info.size())) != YRH_OK) return rc;
    return YRH_OK;
}

void DestroyRenderState(yrh_context* ctx) {
    delete ctx->render;
    ctx->render = nullptr;
}

}  // namespace yrh

using namespace yrh;

extern "C" {

int yrh_upload_shading(yrh_context* ctx, const yrh_shading_desc* sd) {
    const StageRange stage_range("yrh_upload_shading");
    if (ctx == nullptr || sd == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null argument");
    if (!ctx->geo.loaded || ctx->render == nullptr) return Fail(ctx, YRH_E_NO_SCENE, "upload geometry first");
    RenderState& R = *ctx->render;
    if (sd->material_count <= 0 || sd->materials == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "no materials");
    for (int32_t m : R.host_prim_material) {
        if (m < 0 || m >= sd->material_count) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "primitive material index out of range");
    }
    for (int32_t m : R.host_sphere_material) {
        if (m < 0 || m >= sd->material_count) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "sphere material index out of range");
    }
    (void)hipSetDevice(ctx->device);
    int rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (rc != YRH_OK) return rc;
    const int32_t ntex = sd->textures != nullptr ? std::max(0, sd->texture_count) : 0;
    TexturePool pool;
    if (!BuildTexturePool(sd->textures, ntex, pool)) return Fail(ctx, YRH_E_INVALID_ARGUMENT, pool.error);
    std::vector<int4>& trec = pool.records;
    std::vector<int4>& tlev = pool.levels;
    std::vector<float4>& tpool = pool.texels;
    std::vector<char>& tex_footprint = pool.footprint;
    float* gauss = pool.gauss;
    std::vector<DevMaterial> mats(static_cast<size_t>(sd->material_count));
    std::vector<int4> mtex(static_cast<size_t>(sd->material_count));
    R.transparent = false;
    R.textured = false;
    R.alpha = false;
    R.footprints = false;
    for (int32_t k = 0; k < sd->material_count; ++k) {
        const yrh_material& m = sd->materials[k];
        if (m.kind < YRH_MATERIAL_DIFFUSE || m.kind > YRH_MATERIAL_THIN_DIELECTRIC) {
            return Fail(ctx, YRH_E_UNSUPPORTED, "material kind not supported by the HIP backend");
        }
        // TextureIndexValid (material_evaluator.cpp:19-21): other indices mean none.
        auto valid = [&](int32_t t) { return t >= 0 && t < ntex ? t : -1; };
        const int32_t rt = valid(m.reflectance_texture), at = valid(m.alpha_texture), nm = valid(m.normal_map);
        float scale = m.normal_scale;
        int32_t scale_bits;
        std::memcpy(&scale_bits, &scale, 4);
        mtex[static_cast<size_t>(k)] = make_int4(rt, at, nm, scale_bits);
        mats[static_cast<size_t>(k)] = ToDevMaterial(m);
        if (m.kind == YRH_MATERIAL_DIELECTRIC || m.kind == YRH_MATERIAL_THIN_DIELECTRIC) R.transparent = true;
        if (rt >= 0 || nm >= 0) R.textured = true;
        if (at >= 0 || !(m.alpha >= 0.5f)) R.alpha = true;
        for (int32_t t : {rt, at, nm}) {
            if (t >= 0 && tex_footprint[static_cast<size_t>(t)]) R.footprints = true;
        }
    }
    // Alpha cut-outs make shadow rays take the transparent walk (closest hits).
    if (R.alpha) R.transparent = true;
    if ((rc = UploadVec(ctx, R.materials, mats.data(), mats.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.mat_tex, mtex.data(), mtex.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.tex_records, trec.data(), trec.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.tex_levels, tlev.data(), tlev.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.tex_texels, tpool.data(), tpool.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.tex_gauss, gauss, tex::kGaussEntries)) != YRH_OK) return rc;
    R.texture_count = ntex;
    R.material_count = sd->material_count;

    // Emitters: per-emitter running triangle areas in index order (the
    // reference's two linear passes in SelectTriangleByArea).
    const int32_t ne = std::max(0, sd->emissive_count);
    if (ne > 0 && (sd->emissives == nullptr || sd->alias_probability == nullptr || sd->alias_index == nullptr ||
                   sd->alias_pmf == nullptr)) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "emitter tables missing");
    }
    for (int32_t k = 0; k < ne; ++k) {
        if (sd->alias_index[k] < 0 || sd->alias_index[k] >= ne) {
            return Fail(ctx, YRH_E_INVALID_ARGUMENT, "alias table index out of range");
        }
    }
    if (sd->environment.texels != nullptr && sd->environment.has_distribution &&
        (sd->environment.dist_width <= 0 || sd->environment.dist_height <= 0)) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "environment distribution has no texels");
    }
    std::vector<float4> emit(static_cast<size_t>(ne));
    std::vector<int4> einfo(static_cast<size_t>(ne));
    std::vector<float> earea(static_cast<size_t>(ne)), cdf;
    std::vector<int4> prim_info_update;
    std::vector<int32_t> prim_emissive(static_cast<size_t>(R.primitive_count), -1);
    float total_area = 0.0f;
    auto tri_area = [&](int32_t prim, int32_t t) {
        const size_t base = static_cast<size_t>(R.host_prim_first[static_cast<size_t>(prim)]) + static_cast<size_t>(t) * 3;
        float p[3][3];
        for (int k = 0; k < 3; ++k) {
            const uint32_t vi = R.host_indices[base + static_cast<size_t>(k)];
            for (int c = 0; c < 3; ++c) p[k][c] = R.host_positions[3 * static_cast<size_t>(vi) + static_cast<size_t>(c)];
        }
        const float e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        const float e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        const float cx = e1[1] * e2[2] - e1[2] * e2[1];
        const float cy = e1[2] * e2[0] - e1[0] * e2[2];
        const float cz = e1[0] * e2[1] - e1[1] * e2[0];
        const float area = std::sqrt(cx * cx + cy * cy + cz * cz) * 0.5f;
        return std::isfinite(area) && area > 0.0f ? area : 0.0f;
    };
    for (int32_t k = 0; k < ne; ++k) {
        const yrh_emissive& ep = sd->emissives[k];
        emit[static_cast<size_t>(k)] = make_float4(ep.radiance[0], ep.radiance[1], ep.radiance[2], ep.area);
        const float ea = std::isfinite(ep.area) && ep.area > 0.0f ? ep.area : 0.0f;
        total_area += ea;
        const int32_t prim = ep.primitive_index;
        if (prim < 0 || prim >= R.primitive_count) {
            einfo[static_cast<size_t>(k)] = make_int4(-1, static_cast<int32_t>(cdf.size()), 0, 0);
            continue;
        }
        if (prim_emissive[static_cast<size_t>(prim)] < 0) prim_emissive[static_cast<size_t>(prim)] = k;
        const int32_t tris = R.host_prim_count[static_cast<size_t>(prim)];
        float sum = 0.0f;
        for (int32_t t = 0; t < tris; ++t) sum += tri_area(prim, t);
        earea[static_cast<size_t>(k)] = sum;
        einfo[static_cast<size_t>(k)] = make_int4(prim, static_cast<int32_t>(cdf.size()), tris, 0);
        float running = 0.0f;
        for (int32_t t = 0; t < tris; ++t) {
            const float a = tri_area(prim, t);
            if (a > 0.0f) running += a;
            cdf.push_back(running);
        }
    }
    R.emissive_count = ne;
    R.total_emissive_area = total_area;
    if ((rc = UploadVec(ctx, R.emit, emit.data(), emit.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.emit_info, einfo.data(), einfo.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.emit_area, earea.data(), earea.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.emit_cdf, cdf.data(), cdf.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.alias_prob, sd->alias_probability, static_cast<size_t>(ne))) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.alias_index, sd->alias_index, static_cast<size_t>(ne))) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.alias_pmf, sd->alias_pmf, static_cast<size_t>(ne))) != YRH_OK) return rc;
    // Analytic lights and their selection tables (alias table, light tree).
    {
        const auto bits = [](int32_t v) {
            float f;
            std::memcpy(&f, &v, 4);
            return f;
        };
        const int32_t na = std::max(0, sd->analytic_light_count);
        if (na > 0 && sd->analytic_lights == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "analytic lights missing");
        const bool has_alias = na > 0 && sd->analytic_alias_probability != nullptr &&
                               sd->analytic_alias_index != nullptr && sd->analytic_alias_pmf != nullptr;
        if (has_alias) {
            for (int32_t k = 0; k < na; ++k) {
                if (sd->analytic_alias_index[k] < 0 || sd->analytic_alias_index[k] >= na) {
                    return Fail(ctx, YRH_E_INVALID_ARGUMENT, "analytic alias table index out of range");
                }
            }
        }
        const int32_t nt = na > 0 && sd->use_analytic_tree && sd->analytic_tree != nullptr
                               ? std::max(0, sd->analytic_tree_count) : 0;
        std::vector<float4> lights(3 * static_cast<size_t>(na)), tree(3 * static_cast<size_t>(nt));
        for (int32_t k = 0; k < na; ++k) {
            const yrh_analytic_light& l = sd->analytic_lights[k];
            if (l.kind < YRH_LIGHT_POINT || l.kind > YRH_LIGHT_DISTANT) {
                return Fail(ctx, YRH_E_INVALID_ARGUMENT, "unknown analytic light kind");
            }
            float4* o = lights.data() + 3 * static_cast<size_t>(k);
            o[0] = make_float4(l.position[0], l.position[1], l.position[2], bits(l.kind));
            o[1] = make_float4(l.direction[0], l.direction[1], l.direction[2], l.cone_angle);
            o[2] = make_float4(l.intensity[0], l.intensity[1], l.intensity[2], l.cone_cos_inner);
        }
        for (int32_t k = 0; k < nt; ++k) {
            const yrh_light_tree_node& t = sd->analytic_tree[k];
            // A leaf names a light; an interior node's children come after it
            // (the reference's recursive build order), so a descent ends.
            if (t.light_index >= na || (t.light_index < 0 && ((t.left >= 0 && (t.left <= k || t.left >= nt)) ||
                                                               (t.right >= 0 && (t.right <= k || t.right >= nt))))) {
                return Fail(ctx, YRH_E_INVALID_ARGUMENT, "analytic light tree node out of range");
            }
            float4* o = tree.data() + 3 * static_cast<size_t>(k);
            o[0] = make_float4(t.bounds_min[0], t.bounds_min[1], t.bounds_min[2], t.power);
            o[1] = make_float4(t.bounds_max[0], t.bounds_max[1], t.bounds_max[2], 0.0f);
            o[2] = make_float4(bits(t.left), bits(t.right), bits(t.light_index), 0.0f);
        }
        if ((rc = UploadVec(ctx, R.alight, lights.data(), lights.size())) != YRH_OK) return rc;
        if ((rc = UploadVec(ctx, R.alight_tree, tree.data(), tree.size())) != YRH_OK) return rc;
        const size_t nalias = has_alias ? static_cast<size_t>(na) : 0;
        if ((rc = UploadVec(ctx, R.alight_prob, sd->analytic_alias_probability, nalias)) != YRH_OK) return rc;
        if ((rc = UploadVec(ctx, R.alight_alias, sd->analytic_alias_index, nalias)) != YRH_OK) return rc;
        if ((rc = UploadVec(ctx, R.alight_pmf, sd->analytic_alias_pmf, nalias)) != YRH_OK) return rc;
        R.analytic_count = na;
        R.analytic_pmf_count = static_cast<int32_t>(nalias);
        R.analytic_tree_count = nt;
        R.analytic_tree_root = nt > 0 ? sd->analytic_tree_root : -1;
        R.use_analytic_tree = nt > 0 ? 1 : 0;
        R.shadow_slots = na > 0 ? 3 : 2;
    }
    // Primitive -> first emitter record (the reference's linear scan).
    {
        std::vector<int4> info(static_cast<size_t>(R.primitive_count));
        rc = CheckHip(ctx, hipMemcpy(info.data(), R.prim_info.ptr, info.size() * sizeof(int4), hipMemcpyDeviceToHost),
                      "hipMemcpy");
        if (rc != YRH_OK) return rc;
        for (size_t p = 0; p < info.size(); ++p) info[p].y = prim_emissive[p];
        if ((rc = UploadVec(ctx, R.prim_info, info.data(), info.size())) != YRH_OK) return rc;
    }
    // Environment.
    const yrh_environment_desc& env = sd->environment;
    R.env_active = env.active ? 1 : 0;
    for (int c = 0; c < 3; ++c) R.env_radiance[c] = env.radiance[c];
    R.env_strength = env.strength;
    R.env_rotation = env.rotation_radians;
    R.env_w = env.texels != nullptr ? std::max(0, env.width) : 0;
    R.env_h = env.texels != nullptr ? std::max(0, env.height) : 0;
    if (env.wrap_s < YRH_WRAP_REPEAT || env.wrap_s > YRH_WRAP_MIRRORED || env.wrap_t < YRH_WRAP_REPEAT ||
        env.wrap_t > YRH_WRAP_MIRRORED) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "environment wrap mode out of range");
    }
    R.env_nearest = env.nearest ? 1 : 0;
    R.env_ws = env.wrap_s;
    R.env_wt = env.wrap_t;
    const size_t env_texels = static_cast<size_t>(R.env_w) * static_cast<size_t>(R.env_h);
    if ((rc = UploadVec(ctx, R.env_texels, reinterpret_cast<const float4*>(env.texels), env_texels)) != YRH_OK) return rc;
    const bool has_tex = env_texels > 0;
    R.env_sampleable = (env.active && env.strength > 0.0f && has_tex && env.has_distribution) ? 1 : 0;
    R.env_dw = env.has_distribution ? env.dist_width : 0;
    R.env_dh = env.has_distribution ? env.dist_height : 0;
    R.env_total = env.total_weight;
    const size_t nd = static_cast<size_t>(std::max(0, R.env_dw)) * static_cast<size_t>(std::max(0, R.env_dh));
    if (env.has_distribution && (env.texel_weights == nullptr || env.row_weights == nullptr || env.row_cdf == nullptr ||
                                 env.conditional_cdfs == nullptr)) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "environment distribution arrays missing");
    }
    if ((rc = UploadVec(ctx, R.env_weights, env.texel_weights, env.has_distribution ? nd : 0)) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.env_row_weights, env.row_weights, env.has_distribution ? static_cast<size_t>(R.env_dh) : 0)) != YRH_OK) return rc;
    {
        // Padded CDFs and their block starts for BlockLowerBound.
        const float inf = std::numeric_limits<float>::infinity();
        const int h = env.has_distribution ? std::max(0, R.env_dh) : 0;
        const int w = env.has_distribution ? std::max(0, R.env_dw) : 0;
        const auto padded = [](int n) { return (n + kCdfBlock - 1) / kCdfBlock * kCdfBlock; };
        const auto tops = [](int n) { return ((n + kCdfBlock - 1) / kCdfBlock + 3) / 4 * 4; };
        std::vector<float> row(static_cast<size_t>(padded(h)), inf), row_top(static_cast<size_t>(tops(h)), inf);
        for (int i = 0; i < h; ++i) row[static_cast<size_t>(i)] = env.row_cdf[i];
        for (int j = 0; j * kCdfBlock < h; ++j) row_top[static_cast<size_t>(j)] = row[static_cast<size_t>(j) * kCdfBlock];
        R.env_cstride = padded(w);
        R.env_ctop = tops(w);
        R.env_rtop = tops(h);
        std::vector<float> cond(static_cast<size_t>(R.env_cstride) * static_cast<size_t>(h), inf);
        std::vector<float> cond_top(static_cast<size_t>(R.env_ctop) * static_cast<size_t>(h), inf);
        for (int y = 0; y < h; ++y) {
            float* c = cond.data() + static_cast<size_t>(y) * static_cast<size_t>(R.env_cstride);
            for (int x = 0; x < w; ++x) c[x] = env.conditional_cdfs[static_cast<size_t>(y) * static_cast<size_t>(w) + static_cast<size_t>(x)];
            float* t = cond_top.data() + static_cast<size_t>(y) * static_cast<size_t>(R.env_ctop);
            for (int j = 0; j * kCdfBlock < w; ++j) t[j] = c[j * kCdfBlock];
        }
        if ((rc = UploadVec(ctx, R.env_row_cdf, row.data(), row.size())) != YRH_OK) return rc;
        if ((rc = UploadVec(ctx, R.env_row_top, row_top.data(), row_top.size())) != YRH_OK) return rc;
        if ((rc = UploadVec(ctx, R.env_cond_cdf, cond.data(), cond.size())) != YRH_OK) return rc;
        if ((rc = UploadVec(ctx, R.env_cond_top, cond_top.data(), cond_top.size())) != YRH_OK) return rc;
    }
    R.shading_loaded = true;
    return YRH_OK;
}

int yrh_film_reset(yrh_context* ctx, int32_t width, int32_t height) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (width <= 0 || height <= 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "Film dimensions must be positive");
    (void)hipSetDevice(ctx->device);
    RenderState& R = State(ctx);
    int rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (rc != YRH_OK) return rc;
    const size_t bytes = static_cast<size_t>(width) * static_cast<size_t>(height) * sizeof(yrh_film_pixel);
    if (R.film.bytes < bytes) {
        if ((rc = Alloc(ctx, R.film, bytes)) != YRH_OK) return rc;
    }
    // The context stream is non-blocking: clears go on it, so they are ordered
    // before a following film load or pass (a null-stream memset is not).
    if ((rc = CheckHip(ctx, hipMemsetAsync(R.film.ptr, 0, bytes, ctx->stream), "hipMemsetAsync")) != YRH_OK) return rc;
    R.film_w = width;
    R.film_h = height;
    R.act_ready = false;
    if (R.stats.ptr == nullptr) {
        if ((rc = Alloc(ctx, R.stats, kStatBytes)) != YRH_OK) return rc;
    }
    return CheckHip(ctx, hipMemsetAsync(R.stats.ptr, 0, kStatBytes, ctx->stream), "hipMemsetAsync");
}

int yrh_film_load(yrh_context* ctx, const yrh_film_pixel* pixels, int64_t count) {
    if (ctx == nullptr || ctx->render == nullptr || pixels == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "no film");
    RenderState& R = *ctx->render;
    if (count != static_cast<int64_t>(R.film_w) * R.film_h) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "film size mismatch");
    R.act_ready = false;  // an active list describes the film it was built from
    return yrh_copy_to_device(ctx, R.film.ptr, pixels, static_cast<size_t>(count) * sizeof(yrh_film_pixel));
}

int yrh_film_read(yrh_context* ctx, yrh_film_pixel* pixels, int64_t count) {
    if (ctx == nullptr || ctx->render == nullptr || pixels == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "no film");
    RenderState& R = *ctx->render;
    if (count != static_cast<int64_t>(R.film_w) * R.film_h) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "film size mismatch");
    const int rc = yrh_copy_to_host(ctx, pixels, R.film.ptr, static_cast<size_t>(count) * sizeof(yrh_film_pixel));
    return rc != YRH_OK ? rc : CheckStackOverflow(ctx);
}

int yrh_film_update_active(yrh_context* ctx, const yrh_adaptive_desc* adaptive, int32_t spp, int64_t* active,
                           int64_t* converged) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (ctx->render == nullptr || ctx->render->film.ptr == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "no film");
    if (spp < 1) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "spp must be positive");
    (void)hipSetDevice(ctx->device);
    RenderState& R = *ctx->render;
    const int64_t npix = static_cast<int64_t>(R.film_w) * R.film_h;
    int rc;
    if (R.act_list.bytes < static_cast<size_t>(npix) * 4) {
        if ((rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")) != YRH_OK) return rc;
        if ((rc = Alloc(ctx, R.act_list, static_cast<size_t>(npix) * 4)) != YRH_OK) return rc;
    }
    if (R.act_count.ptr == nullptr && (rc = Alloc(ctx, R.act_count, 16)) != YRH_OK) return rc;
    const bool on = adaptive != nullptr && adaptive->enabled != 0;
    hipLaunchKernelGGL(ZeroWordsKernel, dim3(1), dim3(64), 0, ctx->stream, R.act_count.as<uint32_t>(), 2);
    hipLaunchKernelGGL(UpdateActiveKernel, dim3(Grid(npix)), dim3(kBlock), 0, ctx->stream, R.film.as<yrh_film_pixel>(),
                       npix, on ? 1 : 0, on ? adaptive->min_spp : 0, on ? adaptive->relative_error : 0.0f,
                       on ? adaptive->absolute_error : 0.0f, on ? adaptive->confidence : 0.0f,
                       static_cast<uint32_t>(spp), R.act_list.as<int32_t>(), R.act_count.as<uint32_t>());
    if ((rc = CheckHip(ctx, hipGetLastError(), "update active launch")) != YRH_OK) return rc;
    R.act_ready = true;
    R.act_desc = on ? *adaptive : yrh_adaptive_desc{};
    if (active != nullptr || converged != nullptr) {
        uint32_t c[2] = {0, 0};
        if ((rc = yrh_copy_to_host(ctx, c, R.act_count.ptr, 8)) != YRH_OK) return rc;
        if (active != nullptr) *active = c[0];
        if (converged != nullptr) *converged = c[1];
    }
    return YRH_OK;
}

int yrh_render_stats_read(yrh_context* ctx, yrh_render_stats* stats) {
    if (ctx == nullptr || stats == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null argument");
    if (ctx->render == nullptr || ctx->render->stats.ptr == nullptr) {
        *stats = yrh_render_stats{};
        return YRH_OK;
    }
    std::vector<unsigned long long> slots(static_cast<size_t>(kStatStride) * kStatSlots);
    int rc = yrh_copy_to_host(ctx, slots.data(), ctx->render->stats.ptr, kStatBytes);
    if (rc == YRH_OK) rc = CheckStackOverflow(ctx);
    if (rc != YRH_OK) return rc;
    unsigned long long w[kStatWords] = {};
    for (int k = 0; k < kStatSlots; ++k) {
        for (int j = 0; j < kStatWords; ++j) w[j] += slots[static_cast<size_t>(k) * kStatStride + j];
    }
    stats->rays_traced = w[kStatRays];
    stats->shadow_rays = w[kStatShadow];
    stats->occluded_shadow_rays = w[kStatOccluded];
    stats->triangle_tests = w[kStatTriTests] + w[kStatTraceTris];
    stats->sphere_tests = w[kStatSphereTests] + w[kStatTraceSpheres];
    stats->bvh_node_tests = w[kStatNodeTests] + w[kStatTraceNodes];
    stats->hits = w[kStatHits];
    stats->misses = w[kStatMisses];
    stats->bad_samples = w[kStatBad];
    return YRH_OK;
}

int yrh_suggest_paths_per_pass(yrh_context* ctx, int32_t max_depth, int64_t* paths) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (paths == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null output");
    (void)hipSetDevice(ctx->device);
    size_t free_bytes = 0, total_bytes = 0;
    int rc = CheckHip(ctx, hipMemGetInfo(&free_bytes, &total_bytes), "hipMemGetInfo");
    if (rc != YRH_OK) return rc;
    // Per path: the path slot state (EnsureCapacity) and its shadow slots.
    const int shadow_slots = ctx->render != nullptr ? ctx->render->shadow_slots : 2;
    size_t per_path = kPathSlotBytes + static_cast<size_t>(shadow_slots) * kShadowSlotBytes;
    if (ctx->guide != nullptr) per_path += kGuideRowBytes + 64 * static_cast<size_t>(std::max(1, max_depth));
    const size_t held = ctx->render != nullptr ? static_cast<size_t>(ctx->render->capacity) * per_path : 0;
    const int64_t fit = static_cast<int64_t>((free_bytes + held) / 2 / per_path);
    *paths = std::max<int64_t>(int64_t{1} << 16, std::min<int64_t>(fit, int64_t{1} << 27));
    return YRH_OK;
}

int yrh_render_pass(yrh_context* ctx, const yrh_render_desc* d) {
    const StageRange stage_range("yrh_render_pass");
    if (ctx == nullptr || d == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null argument");
    if (ctx->render == nullptr || !ctx->render->shading_loaded) return Fail(ctx, YRH_E_NO_SCENE, "upload geometry and shading first");
    RenderState& R = *ctx->render;
    if (d->width != R.film_w || d->height != R.film_h) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "film size does not match the pass");
    if (d->sampler < YRH_SAMPLER_INDEPENDENT || d->sampler > YRH_SAMPLER_ZSOBOL) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "unknown sampler");
    if (d->sample_index < 0 || d->sample_count < 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "negative sample index or count");
    if (d->pixel_list != 0 && !R.act_ready) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "pixel-list pass without an active list (yrh_film_update_active)");
    }
    const int64_t count = d->sample_count < 1 ? 1 : d->sample_count;
    if (static_cast<int64_t>(d->sample_index) + count > (int64_t{1} << 31) - 1) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "sample index overflow");
    }
    const int64_t npix = static_cast<int64_t>(d->width) * d->height;
    const int64_t paths = npix * count;
    if (paths > (int64_t{1} << 30)) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "pass too large (pixels x samples > 2^30)");
    (void)hipSetDevice(ctx->device);
    int rc = EnsureCapacity(ctx, R, paths);
    if (rc != YRH_OK) return rc;
    PassView P = MakePassView(R, *d);
    const SurfaceView S = MakeSurfaceView(R);
    const LightView L = MakeLightView(R);
    hipStream_t st = ctx->stream;
    const int max_depth = std::max(1, d->max_depth);
    yrg::GuideView GV;
    if ((rc = GuidePrepareView(ctx, paths, max_depth, &GV)) != YRH_OK) return rc;
    if ((rc = LaunchPassSequence(ctx, R, P, S, L, GV, paths, max_depth, ctx->render_overlap)) != YRH_OK) return rc;
    hipLaunchKernelGGL(AccumulateKernel, dim3(Grid(npix)), dim3(kBlock), 0, st, P);
    hipLaunchKernelGGL(CopyTraceTotalsKernel, dim3(1), dim3(64), 0, st, P.stats);
    return CheckHip(ctx, hipGetLastError(), "render pass launch");
}

// ---- batched multi-camera rendering ----------------------------------------------

int yrh_batch_render(yrh_context* ctx, const yrh_batch_desc* d, float* job_ms, int32_t* passes_out) {
    const StageRange stage_range("yrh_batch_render");
    if (ctx == nullptr || d == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null argument");
    if (ctx->render == nullptr || !ctx->render->shading_loaded) return Fail(ctx, YRH_E_NO_SCENE, "upload geometry and shading first");
    if (d->job_count <= 0 || d->jobs == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "no jobs");
    if (d->job_count > 0xFFFF) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "too many jobs (at most 65535)");
    if (d->sampler < YRH_SAMPLER_INDEPENDENT || d->sampler > YRH_SAMPLER_ZSOBOL) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "unknown sampler");
    RenderState& R = *ctx->render;
    // Jobs and their film layout.
    std::vector<DevJob> jobs(static_cast<size_t>(d->job_count));
    R.b_job_pixels.assign(jobs.size(), 0);
    R.b_job_offset.assign(jobs.size(), 0);
    int64_t film_pixels = 0, largest = 0;
    for (int32_t j = 0; j < d->job_count; ++j) {
        const yrh_batch_job& b = d->jobs[j];
        if (b.width <= 0 || b.height <= 0 || b.spp <= 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad job size or spp");
        DevJob& J = jobs[static_cast<size_t>(j)];
        J.camera.origin = make_float3(b.camera_origin[0], b.camera_origin[1], b.camera_origin[2]);
        J.camera.forward = make_float3(b.camera_forward[0], b.camera_forward[1], b.camera_forward[2]);
        J.camera.right = make_float3(b.camera_right[0], b.camera_right[1], b.camera_right[2]);
        J.camera.up = make_float3(b.camera_up[0], b.camera_up[1], b.camera_up[2]);
        J.camera.tan_half_fov = b.tan_half_fov;
        J.width = b.width;
        J.height = b.height;
        J.spp = b.spp;
        J.film_offset = film_pixels;
        const int64_t npix = static_cast<int64_t>(b.width) * b.height;
        R.b_job_pixels[static_cast<size_t>(j)] = npix;
        R.b_job_offset[static_cast<size_t>(j)] = film_pixels;
        film_pixels += npix;
        largest = std::max(largest, npix);
    }
    if (film_pixels > (int64_t{1} << 31) - 1) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "batch film too large");
    const int64_t capacity = d->paths_per_pass > 0 ? d->paths_per_pass : std::min<int64_t>(int64_t{1} << 23, film_pixels);
    if (capacity > (int64_t{1} << 30)) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "paths_per_pass too large");
    // Schedule: each pass takes consecutive pixel ranges / sample layers of the
    // jobs, the job with the least remaining work first, until it is full:
    // small cameras complete in the first passes instead of waiting behind the
    // large ones (every pass but the last is full either way, so the pass count
    // is the same). Items of one job in a pass are numbered by occurrence; the
    // pass accumulates occurrence by occurrence, so each pixel sees its samples
    // in index order.
    constexpr int32_t kMaxOccurrences = 64;
    int32_t max_occurrence = 0;
    std::vector<DevItem> items;
    std::vector<int32_t> pass_items{0};
    std::vector<int32_t> job_last_pass(jobs.size(), 0);
    std::vector<int32_t> sample(jobs.size(), 0);
    std::vector<int64_t> next_pixel(jobs.size(), 0);
    std::vector<int32_t> order(jobs.size());
    for (size_t j = 0; j < order.size(); ++j) order[j] = static_cast<int32_t>(j);
    auto remaining = [&](int32_t j) {
        const size_t k = static_cast<size_t>(j);
        return R.b_job_pixels[k] * (jobs[k].spp - sample[k]) - next_pixel[k];
    };
    while (true) {
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return remaining(a) < remaining(b); });
        if (remaining(order.back()) <= 0) break;
        int64_t used = 0;
        const int32_t pass = static_cast<int32_t>(pass_items.size()) - 1;
        for (const int32_t j : order) {
            const size_t k = static_cast<size_t>(j);
            for (int32_t occ = 0; used < capacity && remaining(j) > 0 && occ < kMaxOccurrences; ++occ) {
                const int64_t take = std::min(capacity - used, R.b_job_pixels[k] - next_pixel[k]);
                items.push_back(DevItem{j, sample[k], static_cast<int32_t>(next_pixel[k]), static_cast<int32_t>(take), occ,
                                        0, used});
                used += take;
                next_pixel[k] += take;
                if (next_pixel[k] == R.b_job_pixels[k]) {
                    next_pixel[k] = 0;
                    ++sample[k];
                }
                job_last_pass[k] = pass;
                max_occurrence = std::max(max_occurrence, occ);
            }
        }
        pass_items.push_back(static_cast<int32_t>(items.size()));
    }
    const int32_t passes = static_cast<int32_t>(pass_items.size()) - 1;
    if (passes_out != nullptr) *passes_out = passes;
    (void)hipSetDevice(ctx->device);
    int rc;
    if ((rc = EnsureCapacity(ctx, R, capacity)) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.b_jobs, jobs.data(), jobs.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.b_items, items.data(), items.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, R.b_pass_items, pass_items.data(), pass_items.size())) != YRH_OK) return rc;
    if ((rc = Alloc(ctx, R.b_film, static_cast<size_t>(film_pixels) * sizeof(yrh_film_pixel))) != YRH_OK) return rc;
    // Stream-ordered clears (the context stream does not wait for the null stream).
    if ((rc = CheckHip(ctx, hipMemsetAsync(R.b_film.ptr, 0, R.b_film.bytes, ctx->stream), "hipMemsetAsync")) != YRH_OK) {
        return rc;
    }
    if ((rc = CheckHip(ctx, hipMemsetAsync(R.stats.ptr, 0, kStatBytes, ctx->stream), "hipMemsetAsync")) != YRH_OK) return rc;
    if ((rc = CheckHip(ctx, hipMemsetAsync(R.counters.ptr, 0, 4 * kCounterWords, ctx->stream), "hipMemsetAsync")) !=
        YRH_OK) {
        return rc;
    }
    yrh_render_desc rd{};
    rd.sampler = d->sampler;
    rd.max_depth = d->max_depth;
    rd.seed = d->seed;
    rd.radiance_clamp = d->radiance_clamp;
    rd.width = 1;
    rd.height = 1;
    rd.spp = 1;
    PassView P = MakePassView(R, rd);
    P.jobs = R.b_jobs.as<DevJob>();
    P.items = R.b_items.as<DevItem>();
    P.pass_items = R.b_pass_items.as<int32_t>();
    P.film = R.b_film.as<yrh_film_pixel>();
    const SurfaceView S = MakeSurfaceView(R);
    const LightView L = MakeLightView(R);
    const int max_depth = std::max(1, d->max_depth);
    yrg::GuideView GV;
    if ((rc = GuidePrepareView(ctx, capacity, max_depth, &GV)) != YRH_OK) return rc;
    // Size the traversal scratch now: nothing may allocate inside a capture.
    {
        TraceRequest warm;
        warm.org = P.s_org;
        warm.dir = P.s_dir;
        warm.n_device = P.counters + kCountShadow;
        warm.capacity = R.shadow_slots * capacity;
        trav::TraceArgs wa;
        int g = 0, b = 0;
        if ((rc = PrepareTraceArgs(ctx, warm, wa, g, b)) != YRH_OK) return rc;
    }
    hipStream_t st = ctx->stream;
    auto launch_pass = [&]() -> int {
        int r = LaunchPassSequence(ctx, R, P, S, L, GV, capacity, max_depth, false);
        if (r != YRH_OK) return r;
        for (int32_t occ = 0; occ <= max_occurrence; ++occ) {
            hipLaunchKernelGGL(AccumulateBatchKernel, dim3(Grid(capacity)), dim3(kBlock), 0, st, P, occ);
        }
        hipLaunchKernelGGL(CopyTraceTotalsKernel, dim3(1), dim3(64), 0, st, P.stats);
        hipLaunchKernelGGL(EndPassKernel, dim3(1), dim3(64), 0, st, P.counters);
        return CheckHip(ctx, hipGetLastError(), "batch pass launch");
    };
    std::vector<hipEvent_t> ev;
    if (job_ms != nullptr) {
        ev.resize(static_cast<size_t>(passes) + 1);
        for (hipEvent_t& e : ev) {
            if ((rc = CheckHip(ctx, hipEventCreate(&e), "hipEventCreate")) != YRH_OK) return rc;
        }
        if ((rc = CheckHip(ctx, hipEventRecord(ev[0], st), "hipEventRecord")) != YRH_OK) return rc;
    }
    if (d->use_graph) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if ((rc = CheckHip(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) != YRH_OK) {
            return rc;
        }
        const int lrc = launch_pass();
        const int crc = CheckHip(ctx, hipStreamEndCapture(st, &graph), "hipStreamEndCapture");
        if (lrc != YRH_OK || crc != YRH_OK) {
            if (graph != nullptr) (void)hipGraphDestroy(graph);
            return lrc != YRH_OK ? lrc : crc;
        }
        rc = CheckHip(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        for (int32_t k = 0; rc == YRH_OK && k < passes; ++k) {
            rc = CheckHip(ctx, hipGraphLaunch(exec, st), "hipGraphLaunch");
            if (rc == YRH_OK && job_ms != nullptr) rc = CheckHip(ctx, hipEventRecord(ev[static_cast<size_t>(k) + 1], st), "hipEventRecord");
        }
        const int src = CheckHip(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
        if (exec != nullptr) (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        if (rc == YRH_OK) rc = src;
    } else {
        rc = YRH_OK;
        for (int32_t k = 0; rc == YRH_OK && k < passes; ++k) {
            rc = launch_pass();
            if (rc == YRH_OK && job_ms != nullptr) rc = CheckHip(ctx, hipEventRecord(ev[static_cast<size_t>(k) + 1], st), "hipEventRecord");
        }
        const int src = CheckHip(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
        if (rc == YRH_OK) rc = src;
    }
    if (job_ms != nullptr) {
        for (int32_t j = 0; rc == YRH_OK && j < d->job_count; ++j) {
            float ms = 0.0f;
            rc = CheckHip(ctx, hipEventElapsedTime(&ms, ev[0], ev[static_cast<size_t>(job_last_pass[static_cast<size_t>(j)]) + 1]),
                          "hipEventElapsedTime");
            job_ms[j] = ms;
        }
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    return rc != YRH_OK ? rc : CheckStackOverflow(ctx);
}

int yrh_batch_film_read(yrh_context* ctx, int32_t job, yrh_film_pixel* pixels, int64_t count) {
    if (ctx == nullptr || pixels == nullptr) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null argument");
    if (ctx->render == nullptr || ctx->render->b_film.ptr == nullptr) return Fail(ctx, YRH_E_NO_SCENE, "no batch rendered");
    RenderState& R = *ctx->render;
    if (job < 0 || job >= static_cast<int32_t>(R.b_job_pixels.size())) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad job");
    if (count != R.b_job_pixels[static_cast<size_t>(job)]) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "pixel count does not match the job");
    return yrh_copy_to_host(ctx, pixels, R.b_film.as<yrh_film_pixel>() + R.b_job_offset[static_cast<size_t>(job)],
                            static_cast<size_t>(count) * sizeof(yrh_film_pixel));
}

int yrh_sample_values(yrh_context* ctx, int32_t kind, uint64_t seed, int32_t spp, int64_t n, const int32_t* x,
                      const int32_t* y, const int32_t* sample, const int32_t* dim, int32_t dims, float* out) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (n < 0 || (n > 0 && (x == nullptr || y == nullptr || sample == nullptr || dim == nullptr || out == nullptr)) ||
        (dims != 1 && dims != 2) || kind < 0 || kind > 3) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad sample_values arguments");
    }
    if (n == 0) return YRH_OK;
    (void)hipSetDevice(ctx->device);
    DevBuf bx, by, bs, bd, bo;
    int rc;
    if ((rc = UploadVec(ctx, bx, x, static_cast<size_t>(n))) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, by, y, static_cast<size_t>(n))) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, bs, sample, static_cast<size_t>(n))) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, bd, dim, static_cast<size_t>(n))) != YRH_OK) return rc;
    if ((rc = Alloc(ctx, bo, static_cast<size_t>(n) * static_cast<size_t>(dims) * sizeof(float))) != YRH_OK) return rc;
    hipLaunchKernelGGL(SampleValuesKernel, dim3(Grid(n)), dim3(kBlock), 0, ctx->stream, kind, seed, spp, n,
                       bx.as<int32_t>(), by.as<int32_t>(), bs.as<int32_t>(), bd.as<int32_t>(), dims, bo.as<float>());
    if ((rc = CheckHip(ctx, hipGetLastError(), "sample kernel launch")) != YRH_OK) return rc;
    return yrh_copy_to_host(ctx, out, bo.ptr, static_cast<size_t>(n) * static_cast<size_t>(dims) * sizeof(float));
}

int yrh_bsdf_query(yrh_context* ctx, const yrh_material* materials, int32_t material_count, int64_t n,
                   const int32_t* material_index, const float* queries, float* out) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (n < 0 || material_count <= 0 || materials == nullptr ||
        (n > 0 && (material_index == nullptr || queries == nullptr || out == nullptr))) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad bsdf_query arguments");
    }
    std::vector<DevMaterial> mats(static_cast<size_t>(material_count));
    for (int32_t k = 0; k < material_count; ++k) {
        if (materials[k].kind < YRH_MATERIAL_DIFFUSE || materials[k].kind > YRH_MATERIAL_THIN_DIELECTRIC) {
            return Fail(ctx, YRH_E_UNSUPPORTED, "material kind not supported by the HIP backend");
        }
        mats[static_cast<size_t>(k)] = ToDevMaterial(materials[k]);
    }
    for (int64_t i = 0; i < n; ++i) {
        if (material_index[i] < 0 || material_index[i] >= material_count) {
            return Fail(ctx, YRH_E_INVALID_ARGUMENT, "material index out of range");
        }
    }
    if (n == 0) return YRH_OK;
    (void)hipSetDevice(ctx->device);
    DevBuf bm, bi, bq, bo;
    int rc;
    if ((rc = UploadVec(ctx, bm, mats.data(), mats.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, bi, material_index, static_cast<size_t>(n))) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, bq, queries, static_cast<size_t>(12 * n))) != YRH_OK) return rc;
    if ((rc = Alloc(ctx, bo, static_cast<size_t>(12 * n) * sizeof(float))) != YRH_OK) return rc;
    hipLaunchKernelGGL(BsdfQueryKernel, dim3(Grid(n)), dim3(kBlock), 0, ctx->stream, bm.as<DevMaterial>(), n,
                       bi.as<int32_t>(), bq.as<float>(), bo.as<float>());
    if ((rc = CheckHip(ctx, hipGetLastError(), "bsdf query launch")) != YRH_OK) return rc;
    return yrh_copy_to_host(ctx, out, bo.ptr, static_cast<size_t>(12 * n) * sizeof(float));
}

int yrh_texture_query(yrh_context* ctx, const yrh_texture* texture, int64_t n, const float* uv, const float* footprint,
                      float* out) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (texture == nullptr || n < 0 || (n > 0 && (uv == nullptr || out == nullptr))) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad texture_query arguments");
    }
    TexturePool pool;
    if (!BuildTexturePool(texture, 1, pool)) return Fail(ctx, YRH_E_INVALID_ARGUMENT, pool.error);
    if (n == 0) return YRH_OK;
    (void)hipSetDevice(ctx->device);
    DevBuf br, bl, bt, bg, bu, bf, bo;
    int rc;
    if ((rc = UploadVec(ctx, br, pool.records.data(), pool.records.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, bl, pool.levels.data(), pool.levels.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, bt, pool.texels.data(), pool.texels.size())) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, bg, pool.gauss, tex::kGaussEntries)) != YRH_OK) return rc;
    if ((rc = UploadVec(ctx, bu, uv, static_cast<size_t>(2 * n))) != YRH_OK) return rc;
    if (footprint != nullptr && (rc = UploadVec(ctx, bf, footprint, static_cast<size_t>(4 * n))) != YRH_OK) return rc;
    if ((rc = Alloc(ctx, bo, static_cast<size_t>(n) * sizeof(float4))) != YRH_OK) return rc;
    const tex::TexView T{br.as<int4>(), bl.as<int4>(), bt.as<float4>(), bg.as<float>(), 1};
    hipLaunchKernelGGL(TextureQueryKernel, dim3(Grid(n)), dim3(kBlock), 0, ctx->stream, T, n, bu.as<float>(),
                       footprint != nullptr ? bf.as<float>() : nullptr, bo.as<float4>());
    if ((rc = CheckHip(ctx, hipGetLastError(), "texture query launch")) != YRH_OK) return rc;
    return yrh_copy_to_host(ctx, out, bo.ptr, static_cast<size_t>(n) * sizeof(float4));
}

int yrh_generate_camera_rays(yrh_context* ctx, const yrh_render_desc* d, float* org_tmin, float* dir_tmax) {
    if (ctx == nullptr || d == nullptr || org_tmin == nullptr || dir_tmax == nullptr) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "null argument");
    }
    if (d->width <= 0 || d->height <= 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad frame size");
    (void)hipSetDevice(ctx->device);
    RenderState& R = State(ctx);
    const int64_t paths = static_cast<int64_t>(d->width) * d->height;
    int rc = EnsureCapacity(ctx, R, paths);
    if (rc != YRH_OK) return rc;
    PassView P = MakePassView(R, *d);
    P.sample_count = 1;  // the caller's arrays hold one ray per pixel
    P.q_org[0] = reinterpret_cast<float4*>(org_tmin);
    P.q_dir[0] = reinterpret_cast<float4*>(dir_tmax);
    hipLaunchKernelGGL(RaygenKernel, dim3(Grid(paths)), dim3(kBlock), 0, ctx->stream, P);
    return CheckHip(ctx, hipGetLastError(), "raygen launch");
}

}  // extern "C"
