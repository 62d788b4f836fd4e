// Learned path-guiding proposal: shared definitions for the render kernels
// and the network kernels.
//
// Model (the reference has no path guiding; this is the project's own):
//   * position encoding: multiresolution hash grid, kGuideLevels levels x 2
//     features, trilinear interpolation over the scene bounds; level l has
//     resolution floor(n_min * (n_max / n_min)^(l / 7) + 0.5);
//     a level whose (res + 1)^3 corners fit the table is indexed densely,
//     otherwise by the spatial hash (x * 1) ^ (y * 2654435761) ^ (z * 805459861)
//     modulo the table size;
//   * network input (32): 16 grid features, shading normal (3), outgoing
//     direction (3), a constant 1 (the only bias), zero padding;
//   * 4 layers: 32 -> 64 -> 64 -> 64 -> 48, ReLU between, no other biases;
//     BF16 weights and activations, FP32 accumulation (MFMA);
//   * head: 8 von Mises-Fisher lobes from outputs [5k .. 5k+4] =
//     (logit, log kappa, mu xyz): weights softmax(logits), kappa =
//     exp(clamp(s, kLogKappaMin, kLogKappaMax)), mu = m / |m|;
//   * q(w) = sum_k pi_k * kappa_k / (2 pi (1 - exp(-2 kappa_k))) * exp(kappa_k (mu_k . w - 1)).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace yrg {

constexpr int kGuideLevels = 8;
constexpr int kGuideFeatures = 2;
constexpr int kGuideIn = 32;
constexpr int kGuideHidden = 64;
constexpr int kGuideOut = 48;
constexpr int kGuideLobes = 8;
constexpr int kGuideMixture = 5 * kGuideLobes;  // per row: pi[8], kappa[8], mu[8][3]
constexpr float kLogKappaMin = -3.0f;
constexpr float kLogKappaMax = 6.2f;

// Flat FP32 parameter layout, each matrix row-major [in][out].
constexpr int kW0 = 0;
constexpr int kW1 = kW0 + kGuideIn * kGuideHidden;
constexpr int kW2 = kW1 + kGuideHidden * kGuideHidden;
constexpr int kW3 = kW2 + kGuideHidden * kGuideHidden;
constexpr int kMlpParams = kW3 + kGuideHidden * kGuideOut;  // 13312

struct GuideGrid {
    float bmin[3];
    float inv_extent[3];
    int32_t res[kGuideLevels];
    int32_t dense[kGuideLevels];
    uint32_t table_size;  // entries per level (power of two)
};

__host__ __device__ inline uint32_t GridIndex(const GuideGrid& g, int level, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t r = static_cast<uint32_t>(g.res[level]) + 1u;
    if (g.dense[level]) return x + r * (y + r * z);
    return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & (g.table_size - 1u);
}

// Corner c (bit0 x, bit1 y, bit2 z) of the cell holding p at `level`:
// returns the table entry and the trilinear weight.
struct GridCell {
    uint32_t base[3];
    float frac[3];
};

__host__ __device__ inline GridCell LocateCell(const GuideGrid& g, int level, const float p[3]) {
    GridCell c;
    const float res = static_cast<float>(g.res[level]);
    for (int a = 0; a < 3; ++a) {
        float u = (p[a] - g.bmin[a]) * g.inv_extent[a];
        u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
        const float s = u * res;
        float f = floorf(s);
        if (f > res - 1.0f) f = res - 1.0f;
        c.base[a] = static_cast<uint32_t>(f);
        c.frac[a] = s - f;
    }
    return c;
}

__host__ __device__ inline float CornerWeight(const GridCell& c, int corner) {
    const float wx = (corner & 1) ? c.frac[0] : 1.0f - c.frac[0];
    const float wy = (corner & 2) ? c.frac[1] : 1.0f - c.frac[1];
    const float wz = (corner & 4) ? c.frac[2] : 1.0f - c.frac[2];
    return wx * wy * wz;
}

__host__ __device__ inline uint32_t CornerIndex(const GuideGrid& g, int level, const GridCell& c, int corner) {
    return GridIndex(g, level, c.base[0] + (corner & 1), c.base[1] + ((corner >> 1) & 1),
                     c.base[2] + ((corner >> 2) & 1));
}

__host__ __device__ inline uint16_t ToBf16(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return static_cast<uint16_t>(u >> 16);
}

__host__ __device__ inline float FromBf16(uint16_t h) {
    const uint32_t u = static_cast<uint32_t>(h) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// ---- mixture head ---------------------------------------------------------------

struct Mixture {
    float pi[kGuideLobes];
    float kappa[kGuideLobes];
    float mu[kGuideLobes][3];
};

__device__ inline void HeadFromOutputs(const float* o, Mixture& m) {
    float mx = o[0];
    for (int k = 1; k < kGuideLobes; ++k) mx = fmaxf(mx, o[5 * k]);
    float sum = 0.0f;
    for (int k = 0; k < kGuideLobes; ++k) {
        m.pi[k] = expf(o[5 * k] - mx);
        sum += m.pi[k];
    }
    for (int k = 0; k < kGuideLobes; ++k) {
        m.pi[k] /= sum;
        const float s = fminf(fmaxf(o[5 * k + 1], kLogKappaMin), kLogKappaMax);
        m.kappa[k] = expf(s);
        const float x = o[5 * k + 2], y = o[5 * k + 3], z = o[5 * k + 4];
        const float len = sqrtf(x * x + y * y + z * z);
        if (len > 1.0e-6f) {
            m.mu[k][0] = x / len;
            m.mu[k][1] = y / len;
            m.mu[k][2] = z / len;
        } else {
            m.mu[k][0] = 0.0f;
            m.mu[k][1] = 0.0f;
            m.mu[k][2] = 1.0f;
        }
    }
}

// vMF density with the normaliser written as kappa / (2 pi (1 - e^{-2 kappa})).
__device__ inline float VmfNorm(float kappa) { return kappa / (2.0f * 3.14159265358979f * (-expm1f(-2.0f * kappa))); }
__device__ inline float VmfPdf(float kappa, float cos_mu) { return VmfNorm(kappa) * expf(kappa * (cos_mu - 1.0f)); }

// A mixture's 40 values: pi[8], kappa[8], mu[8][3], as a row (stride 1) or
// as column i of per-row arrays stride apart (the render's per-queue-row
// mixtures, so a wave's reads of one value are contiguous).
struct MixtureView {
    const float* p;
    int64_t stride;
    __device__ float operator[](int k) const { return p[k * stride]; }
};

__device__ inline float MixturePdf(const MixtureView& m, float3 w) {
    float q = 0.0f;
    for (int k = 0; k < kGuideLobes; ++k) {
        const float c = m[16 + 3 * k] * w.x + m[16 + 3 * k + 1] * w.y + m[16 + 3 * k + 2] * w.z;
        q += m[k] * VmfPdf(m[8 + k], c);
    }
    return q;
}

// The same density from a row that also carries each lobe's normaliser
// VmfNorm(kappa) at 40 + k (the render's per-queue-row mixtures): the same
// operations, so the same value, with one exponential per lobe.
constexpr int kGuideMixtureRow = kGuideMixture + kGuideLobes;
// Row stride of network outputs staged in LDS: odd, so the 16 rows of a
// block read one output column from 16 different banks (a stride of 48
// words put them in 2 banks, an 8-way conflict).
constexpr int kOutStride = kGuideOut + 1;
__device__ inline float MixturePdfWithNorm(const MixtureView& m, float3 w) {
    float q = 0.0f;
    for (int k = 0; k < kGuideLobes; ++k) {
        const float c = m[16 + 3 * k] * w.x + m[16 + 3 * k + 1] * w.y + m[16 + 3 * k + 2] * w.z;
        q += m[k] * (m[kGuideMixture + k] * expf(m[8 + k] * (c - 1.0f)));
    }
    return q;
}

// Samples the mixture: u0 picks the lobe, (u1, u2) the direction within it
// (W. Jakob, "Numerically stable sampling of the von Mises-Fisher
// distribution on S^2", 2012; frame of Duff et al., JCGT 2017).
__device__ inline float3 MixtureSample(const MixtureView& m, float u0, float u1, float u2) {
    int k = 0;
    float acc = m[0];
    while (k < kGuideLobes - 1 && u0 >= acc) {
        ++k;
        acc += m[k];
    }
    const float kap = m[8 + k];
    const float e = expf(-2.0f * kap);
    float w = 1.0f + logf(fmaxf(u1 + (1.0f - u1) * e, 1.0e-30f)) / kap;
    w = fminf(fmaxf(w, -1.0f), 1.0f);
    const float st = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    const float phi = 2.0f * 3.14159265358979f * u2;
    const float lx = st * cosf(phi), ly = st * sinf(phi);
    const float nx = m[16 + 3 * k], ny = m[16 + 3 * k + 1], nz = m[16 + 3 * k + 2];
    const float sign = nz >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + nz);
    const float b = nx * ny * a;
    const float3 t = make_float3(1.0f + sign * nx * nx * a, sign * b, -sign * nx);
    const float3 bt = make_float3(b, sign + ny * ny * a, -ny);
    return make_float3(t.x * lx + bt.x * ly + nx * w, t.y * lx + bt.y * ly + ny * w, t.z * lx + bt.z * ly + nz * w);
}

}  // namespace yrg

// ---- network on one wave (MFMA BF16 -> FP32) ---------------------------------------
//
// Formulation: Z^T = W^T H^T, so a layer's output block in the MFMA result
// layout (lane l holds features 4(l/16)+r, r = 0..3, of row l%16) is exactly
// the B operand of the next layer; activations never leave registers.
// The forward pass uses gfx950's 16x16x32 BF16 MFMA: its B operand (8 BF16
// per lane) is a lane's values of two consecutive 16-feature output blocks,
// 2kb2 and 2kb2+1, side by side. The K index a lane's j-th value stands for
// only has to agree between A and B, so the weight fragments take the same
// order and no value moves between lanes. Weights come as pre-packed
// A-operand fragments:
//   forward fragment (layer, mb over outputs, kb2 over 32-input blocks), 8 BF16
//     per lane: lane l holds W[32 kb2 + 4(l/16) + j][mb*16 + l%16] for j = 0..3
//     and W[32 kb2 + 16 + 4(l/16) + j - 4][mb*16 + l%16] for j = 4..7;
//   backward fragment (layer, mb over inputs, kb over outputs; 16x16x16 BF16
//     MFMA), 4 BF16 per lane: lane l holds W[mb*16 + l%16][kb*16 + 4(l/16) + j].

namespace yrg {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef short bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_mfma_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kLayerIn[4] = {kGuideIn, kGuideHidden, kGuideHidden, kGuideHidden};
constexpr int kLayerOut[4] = {kGuideHidden, kGuideHidden, kGuideHidden, kGuideOut};
constexpr int kLayerParam[4] = {kW0, kW1, kW2, kW3};
constexpr int kFragBase[4] = {0, 8, 24, 40};
constexpr int kFragCount = 52;  // backward fragments (16 x 16 blocks)
constexpr int kFwdFragBase[4] = {0, 4, 12, 20};
constexpr int kFwdFragCount = 26;  // forward fragments (16 outputs x 32 inputs): the same bytes

__device__ __forceinline__ f32x4_t Mfma(bf16x4_t a, bf16x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4_t Mfma32(bf16x8_t a, bf16x8_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_mfma_t, a),
                                                   __builtin_bit_cast(bf16x8_mfma_t, b), c, 0, 0, 0);
}

// B operand of a 32-input block: a lane's values of two 16-feature blocks.
__device__ __forceinline__ bf16x8_t Cat(bf16x4_t lo, bf16x4_t hi) {
    bf16x8_t r;
    r[0] = lo[0];
    r[1] = lo[1];
    r[2] = lo[2];
    r[3] = lo[3];
    r[4] = hi[0];
    r[5] = hi[1];
    r[6] = hi[2];
    r[7] = hi[3];
    return r;
}

__device__ __forceinline__ bf16x4_t PackBf16(f32x4_t d, bool relu) {
    bf16x4_t r;
    for (int j = 0; j < 4; ++j) {
        const float v = relu ? fmaxf(d[j], 0.0f) : d[j];
        r[j] = static_cast<short>(ToBf16(v));
    }
    return r;
}

__device__ __forceinline__ bf16x4_t PackBf16(float a, float b, float c, float d) {
    bf16x4_t r;
    r[0] = static_cast<short>(ToBf16(a));
    r[1] = static_cast<short>(ToBf16(b));
    r[2] = static_cast<short>(ToBf16(c));
    r[3] = static_cast<short>(ToBf16(d));
    return r;
}

__device__ __forceinline__ f32x4_t Zero4() {
    f32x4_t z = {0.0f, 0.0f, 0.0f, 0.0f};
    return z;
}

// Forward pass of RB row blocks. x: input B operands (2 blocks of 16 inputs);
// h0..h2: hidden activations (B operands); o: FP32 outputs (3 blocks of 16).
template <int RB>
__device__ __forceinline__ void MlpForward(const bf16x8_t* __restrict__ frag, int lane, const bf16x4_t (&x)[RB][2],
                                           bf16x4_t (&h0)[RB][4], bf16x4_t (&h1)[RB][4], bf16x4_t (&h2)[RB][4],
                                           f32x4_t (&o)[RB][3]) {
    for (int mb = 0; mb < 4; ++mb) {
        const bf16x8_t a = frag[(kFwdFragBase[0] + mb) * 64 + lane];
        for (int r = 0; r < RB; ++r) h0[r][mb] = PackBf16(Mfma32(a, Cat(x[r][0], x[r][1]), Zero4()), true);
    }
    for (int mb = 0; mb < 4; ++mb) {
        f32x4_t acc[RB];
        for (int r = 0; r < RB; ++r) acc[r] = Zero4();
        for (int kb2 = 0; kb2 < 2; ++kb2) {
            const bf16x8_t a = frag[(kFwdFragBase[1] + mb * 2 + kb2) * 64 + lane];
            for (int r = 0; r < RB; ++r) acc[r] = Mfma32(a, Cat(h0[r][2 * kb2], h0[r][2 * kb2 + 1]), acc[r]);
        }
        for (int r = 0; r < RB; ++r) h1[r][mb] = PackBf16(acc[r], true);
    }
    for (int mb = 0; mb < 4; ++mb) {
        f32x4_t acc[RB];
        for (int r = 0; r < RB; ++r) acc[r] = Zero4();
        for (int kb2 = 0; kb2 < 2; ++kb2) {
            const bf16x8_t a = frag[(kFwdFragBase[2] + mb * 2 + kb2) * 64 + lane];
            for (int r = 0; r < RB; ++r) acc[r] = Mfma32(a, Cat(h1[r][2 * kb2], h1[r][2 * kb2 + 1]), acc[r]);
        }
        for (int r = 0; r < RB; ++r) h2[r][mb] = PackBf16(acc[r], true);
    }
    for (int mb = 0; mb < 3; ++mb) {
        f32x4_t acc[RB];
        for (int r = 0; r < RB; ++r) acc[r] = Zero4();
        for (int kb2 = 0; kb2 < 2; ++kb2) {
            const bf16x8_t a = frag[(kFwdFragBase[3] + mb * 2 + kb2) * 64 + lane];
            for (int r = 0; r < RB; ++r) acc[r] = Mfma32(a, Cat(h2[r][2 * kb2], h2[r][2 * kb2 + 1]), acc[r]);
        }
        for (int r = 0; r < RB; ++r) o[r][mb] = acc[r];
    }
}

// Grid features 4q..4q+3 (levels 2q and 2q+1) of position p: the encoding
// share of the lanes l with l/16 == q.
__device__ __forceinline__ void EncodeShare(const GuideGrid& g, const float2* __restrict__ grid, const float p[3],
                                            int q, float f[4]) {
    for (int lv = 0; lv < 2; ++lv) {
        const int level = 2 * q + lv;
        const GridCell c = LocateCell(g, level, p);
        const float2* table = grid + static_cast<size_t>(level) * g.table_size;
        float a = 0.0f, b = 0.0f;
        for (int k = 0; k < 8; ++k) {
            const float w = CornerWeight(c, k);
            const float2 v = table[CornerIndex(g, level, c, k)];
            a += w * v.x;
            b += w * v.y;
        }
        f[2 * lv] = a;
        f[2 * lv + 1] = b;
    }
}

// The second input block (inputs 16..31) for lane group q: normal, outgoing
// direction, the constant 1, zero padding.
__device__ __forceinline__ bf16x4_t DirectionShare(int q, const float n[3], const float wo[3]) {
    if (q == 0) return PackBf16(n[0], n[1], n[2], wo[0]);
    if (q == 1) return PackBf16(wo[1], wo[2], 1.0f, 0.0f);
    return PackBf16(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace yrg

namespace yrg {

// What a render pass's kernels see of the proposal.
struct GuideView {
    int32_t mode;  // YRH_GUIDE_*
    float alpha;   // probability of sampling the proposal at a guidable vertex
    GuideGrid grid;
    const bf16x8_t* frag;  // forward fragments
    const float2* table;   // hash-grid features
    float* mix;            // per queue row i: value k (pi[8], kappa[8], mu[8][3], VmfNorm(kappa)[8]) at mix[k * mix_stride + i]
    int64_t mix_stride;
    int32_t* valid;        // per queue row: 1 when the row carries a mixture
    float4* rec;           // vertex records [depth][path][4]
    int64_t rec_paths;
    int32_t rec_depths;
    int32_t stamp;         // pass stamp written into the records of this pass
    float4* samples;       // training set [capacity][4]
    uint32_t* sample_count;
    float* sample_wsum;
    int32_t sample_capacity;
    // Per-cell copies of the grid for sampling passes (nullptr: none). Level l
    // is copied when bit l of cell_mask is set: its cell (x, y, z), with res
    // cells per axis, holds the features of its 8 corners in corner order as
    // 16 floats (one 64-byte line) at cells + cell_off[l] + 4 (x + res (y + res z)).
    const float4* cells;
    int64_t cell_off[kGuideLevels];
    uint32_t cell_mask;
    int32_t depth_limit;  // guided bounces: depth < depth_limit (0: every bounce that may continue)
};

// EncodeShare over the view: a copied level is one 64-byte load of the cell's
// 8 corners instead of 8 gathers from the table. The corner values, weights
// and summation order are EncodeShare's, so the features are identical.
__device__ __forceinline__ void EncodeShareView(const GuideView& G, const float p[3], int q, float f[4]) {
    for (int lv = 0; lv < 2; ++lv) {
        const int level = 2 * q + lv;
        const GridCell c = LocateCell(G.grid, level, p);
        float a = 0.0f, b = 0.0f;
        if ((G.cell_mask >> level) & 1u) {
            const uint32_t r = static_cast<uint32_t>(G.grid.res[level]);
            const float4* rec = G.cells + G.cell_off[level] +
                                4 * static_cast<int64_t>(c.base[0] + r * (c.base[1] + r * c.base[2]));
            const float4 v0 = rec[0], v1 = rec[1], v2 = rec[2], v3 = rec[3];
            const float vx[8] = {v0.x, v0.z, v1.x, v1.z, v2.x, v2.z, v3.x, v3.z};
            const float vy[8] = {v0.y, v0.w, v1.y, v1.w, v2.y, v2.w, v3.y, v3.w};
            for (int k = 0; k < 8; ++k) {
                const float w = CornerWeight(c, k);
                a += w * vx[k];
                b += w * vy[k];
            }
        } else {
            const float2* table = G.table + static_cast<size_t>(level) * G.grid.table_size;
            for (int k = 0; k < 8; ++k) {
                const float w = CornerWeight(c, k);
                const float2 v = table[CornerIndex(G.grid, level, c, k)];
                a += w * v.x;
                b += w * v.y;
            }
        }
        f[2 * lv] = a;
        f[2 * lv + 1] = b;
    }
}

}  // namespace yrg
