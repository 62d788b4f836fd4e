// Learned path-guiding proposal on MI355X (gfx950): parameters, online
// training and the yrh_guide_* entry points. The model is specified in
// yrh_guide.h; the reference has no path guiding.
//
// Training step over a batch of recorded samples (rows):
//   GuideBackwardKernel   one wave per 16 rows: hash-grid encoding, MFMA
//                         forward, mixture negative log-likelihood gradient,
//                         MFMA backward to the inputs (hash-grid gradients by
//                         atomic add); keeps every layer's input and output
//                         gradient (BF16) for the weight-gradient GEMM
//   GuideWeightGradKernel dW = H^T dZ per 16x16 tile with MFMA, K = rows,
//                         split over row chunks (FP32 partial sums)
//   GuideAdamMlpKernel    sum of the partials + Adam on the FP32 master weights
//   GuideAdamGridKernel   Adam on the touched hash-grid entries
//   GuidePackKernel       master weights -> BF16 A-operand fragments
//
// Before the first sampling pass after a change of the parameters,
// GuideCellPackKernel copies each level that fits the cell budget into
// per-cell records (the 8 corners of a cell in one 64-byte line).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "yrh_guide.h"
#include "yrh_internal.h"

namespace yrh {

using yrg::bf16x4_t;
using yrg::bf16x8_t;
using yrg::f32x4_t;

// Device memory the per-cell grid copies may take (levels coarse first).
constexpr int64_t kGuideCellBudgetDefault = static_cast<int64_t>(512) << 20;

struct GuideState {
    yrh_guide_desc desc{};
    yrg::GuideGrid grid{};
    int32_t mode = YRH_GUIDE_OFF;
    int64_t grid_params = 0;  // levels * entries * 2
    DevBuf mlp, mlp_m, mlp_v;
    DevBuf frag_fwd, frag_bwd;
    DevBuf table, table_m, table_v, table_grad;
    DevBuf samples;     // float4 x 4 per recorded sample: (p, w), (n, 0), (wo, 0), (wi, 0)
    DevBuf sample_hdr;  // uint32 [0] count, [1] dropped; float [2] weight sum, [3] loss
    DevBuf partial;     // FP32 [chunks][kMlpParams]
    int64_t partial_chunks = 0;
    DevBuf act, grad;  // BF16 per 16-row block: 14 activation blocks, 15 gradient blocks of 16x16
    int64_t scratch_blocks = 0;
    DevBuf mix, valid;
    int64_t rows = 0;
    DevBuf rec;
    int64_t rec_slots = 0;
    int32_t stamp = 0;
    int64_t adam_step = 0;
    // Per-cell copies of the grid levels in cell_mask; valid while
    // cells_fresh (cleared by every change of the grid parameters).
    DevBuf cells;
    int64_t cell_off[yrg::kGuideLevels] = {};
    uint32_t cell_mask = 0;
    bool cells_fresh = false;
    int64_t cell_budget = kGuideCellBudgetDefault;
    int32_t depth_limit = 0;
};

namespace {

constexpr int kBlock = 256;
constexpr int kActBlocks = 14;   // X (2), H0 (4), H1 (4), H2 (4)
constexpr int kGradBlocks = 15;  // dZ0 (4), dZ1 (4), dZ2 (4), dZ3 (3)
constexpr int kTiles = 52;       // 16x16 tiles of dW0..dW3
constexpr int kChunkBlocks = 64;  // row blocks per weight-gradient partial

int Grid(int64_t n, int block = kBlock) { return static_cast<int>(std::max<int64_t>(1, (n + block - 1) / block)); }

int Alloc(yrh_context* ctx, DevBuf& b, size_t bytes) {
    b.reset();
    bytes = std::max<size_t>(bytes, 16);
    int rc = CheckHip(ctx, hipMalloc(&b.ptr, bytes), "hipMalloc");
    if (rc == YRH_OK) b.bytes = bytes;
    return rc;
}

int Ensure(yrh_context* ctx, DevBuf& b, size_t bytes) {
    if (b.ptr != nullptr && b.bytes >= bytes) return YRH_OK;
    int rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return rc != YRH_OK ? rc : Alloc(ctx, b, bytes);
}

// Layer of a fragment / tile index and its local index.
__device__ __forceinline__ int LayerOfFrag(int f) { return f < 8 ? 0 : (f < 24 ? 1 : (f < 40 ? 2 : 3)); }

// Cell i of `level` (x fastest): the table entries of its 8 corners in corner
// order, 4 float4 per cell.
__global__ void GuideCellPackKernel(const float2* __restrict__ table, yrg::GuideGrid g, int level, uint32_t cells,
                                    float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const uint32_t r = static_cast<uint32_t>(g.res[level]);
    const uint32_t x = i % r, y = (i / r) % r, z = i / (r * r);
    const float2* t = table + static_cast<size_t>(level) * g.table_size;
    float2 v[8];
    for (int k = 0; k < 8; ++k) v[k] = t[yrg::GridIndex(g, level, x + (k & 1), y + ((k >> 1) & 1), z + ((k >> 2) & 1))];
    float4* o = out + 4 * static_cast<size_t>(i);
    for (int j = 0; j < 4; ++j) o[j] = make_float4(v[2 * j].x, v[2 * j].y, v[2 * j + 1].x, v[2 * j + 1].y);
}

// Layer of a forward (16 x 32) fragment index.
__device__ __forceinline__ int LayerOfFwdFrag(int f) { return f < 4 ? 0 : (f < 12 ? 1 : (f < 20 ? 2 : 3)); }

__global__ void GuidePackKernel(const float* __restrict__ mlp, bf16x8_t* __restrict__ fwd, bf16x4_t* __restrict__ bwd) {
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= yrg::kFragCount * 64) return;
    const int f = t / 64, lane = t % 64;
    const int r = lane % 16, g = 4 * (lane / 16);
    if (f < yrg::kFwdFragCount) {
        // Forward fragment (mb, kb2): the values of input blocks 2 kb2 and 2 kb2 + 1.
        const int l = LayerOfFwdFrag(f);
        const int local = f - yrg::kFwdFragBase[l];
        const int in = yrg::kLayerIn[l], out = yrg::kLayerOut[l];
        const float* w = mlp + yrg::kLayerParam[l];
        const int nkb2 = in / 32;
        const int mb = local / nkb2, kb2 = local % nkb2;
        bf16x8_t v;
        for (int j = 0; j < 8; ++j) {
            const int feature = 32 * kb2 + (j < 4 ? g + j : 16 + g + j - 4);
            v[j] = static_cast<short>(yrg::ToBf16(w[feature * out + mb * 16 + r]));
        }
        fwd[t] = v;
    }
    const int l = LayerOfFrag(f);
    const int local = f - yrg::kFragBase[l];
    const int out = yrg::kLayerOut[l];
    const float* w = mlp + yrg::kLayerParam[l];
    {
        const int nkb = out / 16;
        const int mb = local / nkb, kb = local % nkb;
        bf16x4_t v;
        for (int j = 0; j < 4; ++j) v[j] = static_cast<short>(yrg::ToBf16(w[(mb * 16 + r) * out + kb * 16 + g + j]));
        bwd[t] = v;
    }
}

struct TrainArgs {
    const float4* samples;
    const uint32_t* count;  // recorded samples (device; may exceed capacity when samples were dropped)
    uint32_t capacity;
    const float* wsum;      // their weight sum (device), used when w_norm <= 0
    float w_norm;
    int64_t offset;  // first sample of this batch (mod count)
    int64_t rows;    // batch rows
    int32_t spread;  // 1: row k of the run takes sample floor(frac(k phi) count), k = offset + row
    const bf16x8_t* fwd;
    const bf16x4_t* bwd;
    const float2* table;
    float* table_grad;
    yrg::GuideGrid grid;
    uint16_t* act;
    uint16_t* grad;
    float* loss;
};

__device__ __forceinline__ void StoreBlock(uint16_t* base, bf16x4_t v, int lane) {
    const int row = lane % 16, g = 4 * (lane / 16);
    for (int j = 0; j < 4; ++j) base[(g + j) * 16 + row] = static_cast<uint16_t>(v[j]);
}

__device__ __forceinline__ bf16x4_t MaskRelu(f32x4_t d, bf16x4_t h) {
    bf16x4_t r;
    for (int j = 0; j < 4; ++j) {
        const float keep = yrg::FromBf16(static_cast<uint16_t>(h[j])) > 0.0f ? d[j] : 0.0f;
        r[j] = static_cast<short>(yrg::ToBf16(keep));
    }
    return r;
}

// Gradient of -c log q with respect to the 48 outputs of one row (40 used).
__device__ void HeadGradient(const float* o, float wx, float wy, float wz, float c, float* d, float& loss) {
    yrg::Mixture m;
    yrg::HeadFromOutputs(o, m);
    float v[yrg::kGuideLobes];
    float q = 0.0f;
    for (int k = 0; k < yrg::kGuideLobes; ++k) {
        const float cm = m.mu[k][0] * wx + m.mu[k][1] * wy + m.mu[k][2] * wz;
        v[k] = yrg::VmfPdf(m.kappa[k], cm);
        q += m.pi[k] * v[k];
    }
    q = fmaxf(q, 1.0e-30f);
    loss = -c * logf(q);
    for (int j = 0; j < yrg::kGuideOut; ++j) d[j] = 0.0f;
    for (int k = 0; k < yrg::kGuideLobes; ++k) {
        const float r = m.pi[k] * v[k] / q;
        d[5 * k] = -c * (r - m.pi[k]);
        const float kap = m.kappa[k];
        const float cm = m.mu[k][0] * wx + m.mu[k][1] * wy + m.mu[k][2] * wz;
        const float s = o[5 * k + 1];
        if (s >= yrg::kLogKappaMin && s <= yrg::kLogKappaMax) {
            const float dlogv = 1.0f / kap - 2.0f / expm1f(2.0f * kap) + (cm - 1.0f);
            d[5 * k + 1] = -c * r * dlogv * kap;
        }
        const float x = o[5 * k + 2], y = o[5 * k + 3], z = o[5 * k + 4];
        const float len = sqrtf(x * x + y * y + z * z);
        if (len > 1.0e-6f) {
            const float f = -c * r * kap / len;
            d[5 * k + 2] = f * (wx - cm * m.mu[k][0]);
            d[5 * k + 3] = f * (wy - cm * m.mu[k][1]);
            d[5 * k + 4] = f * (wz - cm * m.mu[k][2]);
        }
    }
}

__global__ __launch_bounds__(kBlock) void GuideBackwardKernel(TrainArgs A) {
    __shared__ float lds_o[4][16 * yrg::kOutStride];
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int wave = static_cast<int>(threadIdx.x >> 6);
    const int64_t rb = static_cast<int64_t>(blockIdx.x) * 4 + wave;
    const int64_t nrb = (A.rows + 15) / 16;
    const int row = lane % 16, q = lane / 16;
    const int64_t idx = rb * 16 + row;
    const uint32_t count = min(*A.count, A.capacity);
    const bool valid = rb < nrb && idx < A.rows && count > 0;
    float p[3] = {0, 0, 0}, n[3] = {0, 0, 0}, wo[3] = {0, 0, 0}, wi[3] = {0, 0, 1};
    float w = 0.0f;
    if (valid) {
        // Spread batches: the recorded set is in path-slot order (depth-major,
        // pixels in order), so a contiguous batch would see one depth of one
        // image band. Row k takes frac(k phi) of the set instead (a Weyl
        // sequence: every batch and every run of batches is spread evenly).
        const uint64_t k = static_cast<uint64_t>(A.offset + idx);
        const int64_t s = A.spread ? static_cast<int64_t>((((k * 0x9E3779B97F4A7C15ull) >> 32) * count) >> 32)
                                   : static_cast<int64_t>(k % count);
        const float4 a = A.samples[4 * s], b = A.samples[4 * s + 1], c = A.samples[4 * s + 2], d = A.samples[4 * s + 3];
        p[0] = a.x; p[1] = a.y; p[2] = a.z; w = a.w;
        n[0] = b.x; n[1] = b.y; n[2] = b.z;
        wo[0] = c.x; wo[1] = c.y; wo[2] = c.z;
        wi[0] = d.x; wi[1] = d.y; wi[2] = d.z;
    }
    bf16x4_t x[1][2];
    {
        float f[4] = {0, 0, 0, 0};
        if (valid) yrg::EncodeShare(A.grid, A.table, p, q, f);
        x[0][0] = yrg::PackBf16(f[0], f[1], f[2], f[3]);
        x[0][1] = valid ? yrg::DirectionShare(q, n, wo) : yrg::PackBf16(0.0f, 0.0f, 0.0f, 0.0f);
    }
    bf16x4_t h0[1][4], h1[1][4], h2[1][4];
    f32x4_t o[1][3];
    yrg::MlpForward<1>(A.fwd, lane, x, h0, h1, h2, o);
    float* lo = lds_o[wave];
    for (int mb = 0; mb < 3; ++mb) {
        for (int j = 0; j < 4; ++j) lo[row * yrg::kOutStride + mb * 16 + 4 * q + j] = o[0][mb][j];
    }
    __syncthreads();
    if (q == 0) {
        float out[yrg::kGuideOut], d[yrg::kGuideOut];
        for (int j = 0; j < yrg::kGuideOut; ++j) out[j] = lo[row * yrg::kOutStride + j];
        float norm = A.w_norm;
        if (!(norm > 0.0f)) norm = count > 0 ? *A.wsum / static_cast<float>(count) : 1.0f;
        const float c = valid && norm > 0.0f ? w / norm / static_cast<float>(A.rows) : 0.0f;
        float loss = 0.0f;
        HeadGradient(out, wi[0], wi[1], wi[2], c, d, loss);
        for (int j = 0; j < yrg::kGuideOut; ++j) lo[row * yrg::kOutStride + j] = d[j];
        // One atomic per wave for the loss (lanes 0..15 hold the 16 rows):
        // every row adding to the one word serialised the whole step.
        if (!valid) loss = 0.0f;
        for (int off = 8; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
        if (lane == 0 && loss != 0.0f) atomicAdd(A.loss, loss);
    }
    __syncthreads();
    bf16x4_t dz3[3], dz2[4], dz1[4], dz0[4];
    for (int kb = 0; kb < 3; ++kb) {
        const float* src = lo + row * yrg::kOutStride + kb * 16 + 4 * q;
        dz3[kb] = yrg::PackBf16(src[0], src[1], src[2], src[3]);
    }
    for (int mb = 0; mb < 4; ++mb) {
        f32x4_t acc = yrg::Zero4();
        for (int kb = 0; kb < 3; ++kb) acc = yrg::Mfma(A.bwd[(yrg::kFragBase[3] + mb * 3 + kb) * 64 + lane], dz3[kb], acc);
        dz2[mb] = MaskRelu(acc, h2[0][mb]);
    }
    for (int mb = 0; mb < 4; ++mb) {
        f32x4_t acc = yrg::Zero4();
        for (int kb = 0; kb < 4; ++kb) acc = yrg::Mfma(A.bwd[(yrg::kFragBase[2] + mb * 4 + kb) * 64 + lane], dz2[kb], acc);
        dz1[mb] = MaskRelu(acc, h1[0][mb]);
    }
    for (int mb = 0; mb < 4; ++mb) {
        f32x4_t acc = yrg::Zero4();
        for (int kb = 0; kb < 4; ++kb) acc = yrg::Mfma(A.bwd[(yrg::kFragBase[1] + mb * 4 + kb) * 64 + lane], dz1[kb], acc);
        dz0[mb] = MaskRelu(acc, h0[0][mb]);
    }
    // Input gradient of the grid features (inputs 0..15): lane l holds features
    // 4q..4q+3 of its row, i.e. its own encoding share (levels 2q, 2q+1).
    f32x4_t dx = yrg::Zero4();
    for (int kb = 0; kb < 4; ++kb) dx = yrg::Mfma(A.bwd[(yrg::kFragBase[0] + 0 * 4 + kb) * 64 + lane], dz0[kb], dx);
    // Hash-grid gradients. The 16 lanes of a quarter (lanes 16q..16q+15) hold
    // 16 rows' shares of the same two levels. Where all 16 touch the same
    // corner entry (or add nothing) — rows from one cell of a coarse level,
    // e.g. the rows of yrh_guide_gradients in set order — the quarter sums its
    // contributions and adds them with one atomic per feature instead of 16.
    // (Training batches are spread over the recorded set, where this is rare;
    // runs of 16 neighbouring samples instead measured no faster.)
    {
        const uint64_t quarter = 0xFFFFull << (16 * q);
        const int quarter_lead = lane & ~15;
        for (int lv = 0; lv < 2; ++lv) {
            const int level = 2 * q + lv;
            const float ga = valid ? dx[2 * lv] : 0.0f;
            const float gb = valid ? dx[2 * lv + 1] : 0.0f;
            const yrg::GridCell cell = yrg::LocateCell(A.grid, level, p);
            float* g = A.table_grad + static_cast<size_t>(level) * A.grid.table_size * 2;
            for (int k = 0; k < 8; ++k) {
                const float cw = yrg::CornerWeight(cell, k);
                float va = cw * ga, vb = cw * gb;
                const bool adds = va != 0.0f || vb != 0.0f;
                const uint32_t e = adds ? yrg::CornerIndex(A.grid, level, cell, k) : 0xFFFFFFFFu;
                uint32_t e0 = e;
                for (int off = 1; off < 16; off <<= 1) e0 = min(e0, static_cast<uint32_t>(__shfl_xor(static_cast<int>(e0), off)));
                const uint64_t same = __ballot(!adds || e == e0);
                if ((same & quarter) == quarter) {
                    for (int off = 8; off > 0; off >>= 1) {
                        va += __shfl_xor(va, off);
                        vb += __shfl_xor(vb, off);
                    }
                    if (lane == quarter_lead && e0 != 0xFFFFFFFFu) {
                        unsafeAtomicAdd(g + 2 * e0, va);
                        unsafeAtomicAdd(g + 2 * e0 + 1, vb);
                    }
                } else if (adds) {
                    unsafeAtomicAdd(g + 2 * e, va);
                    unsafeAtomicAdd(g + 2 * e + 1, vb);
                }
            }
        }
    }
    if (rb < nrb) {
        uint16_t* act = A.act + rb * kActBlocks * 256;
        uint16_t* grd = A.grad + rb * kGradBlocks * 256;
        StoreBlock(act + 0 * 256, x[0][0], lane);
        StoreBlock(act + 1 * 256, x[0][1], lane);
        for (int b = 0; b < 4; ++b) {
            StoreBlock(act + (2 + b) * 256, h0[0][b], lane);
            StoreBlock(act + (6 + b) * 256, h1[0][b], lane);
            StoreBlock(act + (10 + b) * 256, h2[0][b], lane);
            StoreBlock(grd + b * 256, dz0[b], lane);
            StoreBlock(grd + (4 + b) * 256, dz1[b], lane);
            StoreBlock(grd + (8 + b) * 256, dz2[b], lane);
        }
        for (int b = 0; b < 3; ++b) StoreBlock(grd + (12 + b) * 256, dz3[b], lane);
    }
}

// Tile t of dW: layer, block over inputs (mb) and over outputs (nb), and the
// activation / gradient blocks it multiplies.
__device__ __forceinline__ void TileOf(int t, int& l, int& mb, int& nb, int& ablk, int& gblk) {
    if (t < 8) {
        l = 0; mb = t / 4; nb = t % 4; ablk = mb; gblk = nb;
    } else if (t < 24) {
        l = 1; mb = (t - 8) / 4; nb = (t - 8) % 4; ablk = 2 + mb; gblk = 4 + nb;
    } else if (t < 40) {
        l = 2; mb = (t - 24) / 4; nb = (t - 24) % 4; ablk = 6 + mb; gblk = 8 + nb;
    } else {
        l = 3; mb = (t - 40) / 3; nb = (t - 40) % 3; ablk = 10 + mb; gblk = 12 + nb;
    }
}

__global__ __launch_bounds__(kBlock) void GuideWeightGradKernel(const uint16_t* __restrict__ act,
                                                                const uint16_t* __restrict__ grad, int64_t nrb,
                                                                int64_t chunks, float* __restrict__ partial) {
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int64_t item = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int t = static_cast<int>(item % kTiles);
    const int64_t chunk = item / kTiles;
    if (chunk >= chunks) return;
    int l, mb, nb, ablk, gblk;
    TileOf(t, l, mb, nb, ablk, gblk);
    const int r = lane % 16, g = 4 * (lane / 16);
    f32x4_t acc = yrg::Zero4();
    const int64_t rb_end = std::min<int64_t>(nrb, (chunk + 1) * kChunkBlocks);
    for (int64_t rb = chunk * kChunkBlocks; rb < rb_end; ++rb) {
        const bf16x4_t a = *reinterpret_cast<const bf16x4_t*>(act + (rb * kActBlocks + ablk) * 256 + r * 16 + g);
        const bf16x4_t b = *reinterpret_cast<const bf16x4_t*>(grad + (rb * kGradBlocks + gblk) * 256 + r * 16 + g);
        acc = yrg::Mfma(a, b, acc);
    }
    const int out = yrg::kLayerOut[l];
    float* dst = partial + chunk * yrg::kMlpParams + yrg::kLayerParam[l];
    for (int j = 0; j < 4; ++j) dst[(mb * 16 + g + j) * out + nb * 16 + r] = acc[j];
}

struct AdamArgs {
    float lr, beta1, beta2, eps, c1, c2;  // c1 = 1 - beta1^t, c2 = 1 - beta2^t
};

// Gradient sum over the partials; `grad_out` (optional) receives it; Adam
// updates the master weights when `update` is set and samples exist.
__global__ void GuideAdamMlpKernel(const float* __restrict__ partial, int64_t chunks, float* __restrict__ mlp,
                                   float* __restrict__ m, float* __restrict__ v, float* __restrict__ grad_out,
                                   const uint32_t* count, int update, AdamArgs a) {
    const int p = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= yrg::kMlpParams) return;
    float g = 0.0f;
    for (int64_t c = 0; c < chunks; ++c) g += partial[c * yrg::kMlpParams + p];
    // W3 columns 40..47 feed no head output.
    if (p >= yrg::kW3 && (p - yrg::kW3) % yrg::kGuideOut >= yrg::kGuideMixture) g = 0.0f;
    if (grad_out != nullptr) grad_out[p] = g;
    if (!update || *count == 0) return;
    const float mm = a.beta1 * m[p] + (1.0f - a.beta1) * g;
    const float vv = a.beta2 * v[p] + (1.0f - a.beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    mlp[p] -= a.lr * (mm / a.c1) / (sqrtf(vv / a.c2) + a.eps);
}

// Adam on the hash-grid entries with a nonzero gradient (sparse update);
// clears the gradient.
__global__ void GuideAdamGridKernel(float* __restrict__ table, float* __restrict__ m, float* __restrict__ v,
                                    float* __restrict__ grad, float* __restrict__ grad_out, int64_t n, int update,
                                    AdamArgs a) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float g = grad[p];
    if (grad_out != nullptr) grad_out[p] = g;
    if (g == 0.0f) return;
    grad[p] = 0.0f;
    if (!update) return;
    const float mm = a.beta1 * m[p] + (1.0f - a.beta1) * g;
    const float vv = a.beta2 * v[p] + (1.0f - a.beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    table[p] -= a.lr * (mm / a.c1) / (sqrtf(vv / a.c2) + a.eps);
}

__global__ void FillKernel(float* p, int64_t n, float value) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// Appends rows of (p, n, wo, wi, w) to the recorded set.
__global__ void GuideAddSamplesKernel(const float* __restrict__ rows, int64_t n, float4* __restrict__ samples,
                                      uint32_t* hdr, int32_t capacity) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float* s = rows + 13 * r;
    if (!(s[12] > 0.0f) || !isfinite(s[12])) return;
    const uint32_t idx = atomicAdd(hdr, 1u);
    if (idx >= static_cast<uint32_t>(capacity)) return;
    float4* d = samples + 4 * static_cast<int64_t>(idx);
    d[0] = make_float4(s[0], s[1], s[2], s[12]);
    d[1] = make_float4(s[3], s[4], s[5], 0.0f);
    d[2] = make_float4(s[6], s[7], s[8], 0.0f);
    d[3] = make_float4(s[9], s[10], s[11], 0.0f);
    atomicAdd(reinterpret_cast<float*>(hdr) + 2, s[12]);
}

__global__ void ResetHeaderKernel(uint32_t* hdr) {
    if (threadIdx.x < 4) hdr[threadIdx.x] = 0u;
}

// Network evaluation for host-supplied rows (position, normal, outgoing):
// one wave per 64 rows.
__global__ __launch_bounds__(kBlock) void GuideQueryRowsKernel(const float* __restrict__ rows, int64_t n,
                                                               const bf16x8_t* __restrict__ fwd,
                                                               const float2* __restrict__ table, yrg::GuideGrid grid,
                                                               float* __restrict__ mixture) {
    __shared__ float lds_o[4][64 * yrg::kOutStride];
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int wave = static_cast<int>(threadIdx.x >> 6);
    const int64_t base = (static_cast<int64_t>(blockIdx.x) * 4 + wave) * 64;
    const int q = lane / 16;
    bf16x4_t x[4][2];
    for (int rb = 0; rb < 4; ++rb) {
        const int64_t r = base + rb * 16 + lane % 16;
        float f[4] = {0, 0, 0, 0};
        if (r < n) {
            const float* src = rows + 9 * r;
            const float p[3] = {src[0], src[1], src[2]};
            const float nn[3] = {src[3], src[4], src[5]};
            const float wo[3] = {src[6], src[7], src[8]};
            yrg::EncodeShare(grid, table, p, q, f);
            x[rb][1] = yrg::DirectionShare(q, nn, wo);
        } else {
            x[rb][1] = yrg::PackBf16(0.0f, 0.0f, 0.0f, 0.0f);
        }
        x[rb][0] = yrg::PackBf16(f[0], f[1], f[2], f[3]);
    }
    bf16x4_t h0[4][4], h1[4][4], h2[4][4];
    f32x4_t o[4][3];
    yrg::MlpForward<4>(fwd, lane, x, h0, h1, h2, o);
    float* lo = lds_o[wave];
    for (int rb = 0; rb < 4; ++rb) {
        for (int mb = 0; mb < 3; ++mb) {
            for (int j = 0; j < 4; ++j) lo[(rb * 16 + lane % 16) * yrg::kOutStride + mb * 16 + 4 * q + j] = o[rb][mb][j];
        }
    }
    __syncthreads();
    const int64_t r = base + lane;
    if (r < n) {
        yrg::Mixture m;
        yrg::HeadFromOutputs(lo + lane * yrg::kOutStride, m);
        float* dst = mixture + 40 * r;
        for (int k = 0; k < yrg::kGuideLobes; ++k) {
            dst[k] = m.pi[k];
            dst[8 + k] = m.kappa[k];
            dst[16 + 3 * k] = m.mu[k][0];
            dst[16 + 3 * k + 1] = m.mu[k][1];
            dst[16 + 3 * k + 2] = m.mu[k][2];
        }
    }
}

__global__ void GuideMixtureEvalKernel(const float* __restrict__ mixture, const float* __restrict__ dirs,
                                       const float* __restrict__ u, int64_t n, float* __restrict__ pdf,
                                       float* __restrict__ samples) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float* m = mixture + 40 * r;
    const yrg::MixtureView mv{m, 1};
    pdf[r] = yrg::MixturePdf(mv, make_float3(dirs[3 * r], dirs[3 * r + 1], dirs[3 * r + 2]));
    float3 w = yrg::MixtureSample(mv, u[3 * r], u[3 * r + 1], u[3 * r + 2]);
    const float len = sqrtf(w.x * w.x + w.y * w.y + w.z * w.z);
    w = make_float3(w.x / len, w.y / len, w.z / len);
    samples[4 * r] = w.x;
    samples[4 * r + 1] = w.y;
    samples[4 * r + 2] = w.z;
    samples[4 * r + 3] = yrg::MixturePdf(mv, w);
}

// ---- host helpers ------------------------------------------------------------------

uint64_t SplitMix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

float Uniform(uint64_t& s, float lo, float hi) {
    const float u = static_cast<float>(SplitMix(s) >> 40) * (1.0f / 16777216.0f);
    return lo + (hi - lo) * u;
}

GuideState* Guide(yrh_context* ctx) { return ctx != nullptr ? ctx->guide : nullptr; }

int NeedGuide(yrh_context* ctx) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (ctx->guide == nullptr) return Fail(ctx, YRH_E_NO_SCENE, "guiding is not configured");
    (void)hipSetDevice(ctx->device);
    return YRH_OK;
}

int Pack(yrh_context* ctx, GuideState& G) {
    hipLaunchKernelGGL(GuidePackKernel, dim3(Grid(yrg::kFragCount * 64)), dim3(kBlock), 0, ctx->stream,
                       G.mlp.as<float>(), G.frag_fwd.as<bf16x8_t>(), G.frag_bwd.as<bf16x4_t>());
    return CheckHip(ctx, hipGetLastError(), "guide pack");
}

// Per-cell copies for the sampling passes: levels are taken coarse first while
// their records fit the budget. If the device cannot hold them the levels stay
// on the table (a slower lookup, the same features).
int PackCells(yrh_context* ctx, GuideState& G) {
    G.cells_fresh = true;
    G.cell_mask = 0;
    int64_t total = 0;
    uint32_t mask = 0;
    for (int l = 0; l < yrg::kGuideLevels; ++l) {
        const int64_t r = G.grid.res[l];
        const int64_t bytes = r * r * r * 64;
        if (r * r * r > 0x7FFFFFFF || total + bytes > G.cell_budget) break;
        G.cell_off[l] = total / 16;
        total += bytes;
        mask |= 1u << l;
    }
    if (mask == 0) return YRH_OK;
    if (G.cells.bytes < static_cast<size_t>(total)) {
        G.cells.reset();
        void* p = nullptr;
        if (hipMalloc(&p, static_cast<size_t>(total)) != hipSuccess) {
            (void)hipGetLastError();
            return YRH_OK;
        }
        G.cells.ptr = p;
        G.cells.bytes = static_cast<size_t>(total);
    }
    for (int l = 0; l < yrg::kGuideLevels; ++l) {
        if (!((mask >> l) & 1u)) continue;
        const int64_t r = G.grid.res[l];
        const uint32_t n = static_cast<uint32_t>(r * r * r);
        hipLaunchKernelGGL(GuideCellPackKernel, dim3(Grid(n)), dim3(kBlock), 0, ctx->stream, G.table.as<float2>(),
                           G.grid, l, n, G.cells.as<float4>() + G.cell_off[l]);
    }
    G.cell_mask = mask;
    return CheckHip(ctx, hipGetLastError(), "guide cell pack");
}

// Phase A + B of one step over `rows` rows from `samples` (count on device).
int BackwardPass(yrh_context* ctx, GuideState& G, const float4* samples, const uint32_t* count, uint32_t capacity,
                 const float* wsum, float w_norm, int64_t offset, int64_t rows, float* loss, int spread) {
    const int64_t nrb = (rows + 15) / 16;
    int rc;
    if (G.scratch_blocks < nrb) {
        if ((rc = Ensure(ctx, G.act, static_cast<size_t>(nrb) * kActBlocks * 512)) != YRH_OK) return rc;
        if ((rc = Ensure(ctx, G.grad, static_cast<size_t>(nrb) * kGradBlocks * 512)) != YRH_OK) return rc;
        G.scratch_blocks = nrb;
    }
    const int64_t chunks = (nrb + kChunkBlocks - 1) / kChunkBlocks;
    if (G.partial_chunks < chunks) {
        if ((rc = Ensure(ctx, G.partial, static_cast<size_t>(chunks) * yrg::kMlpParams * 4)) != YRH_OK) return rc;
        G.partial_chunks = chunks;
    }
    TrainArgs A{};
    A.samples = samples;
    A.count = count;
    A.capacity = capacity;
    A.wsum = wsum;
    A.w_norm = w_norm;
    A.offset = offset;
    A.spread = spread;
    A.rows = rows;
    A.fwd = G.frag_fwd.as<bf16x8_t>();
    A.bwd = G.frag_bwd.as<bf16x4_t>();
    A.table = G.table.as<float2>();
    A.table_grad = G.table_grad.as<float>();
    A.grid = G.grid;
    A.act = G.act.as<uint16_t>();
    A.grad = G.grad.as<uint16_t>();
    A.loss = loss;
    hipLaunchKernelGGL(GuideBackwardKernel, dim3(Grid(nrb, 4)), dim3(kBlock), 0, ctx->stream, A);
    hipLaunchKernelGGL(GuideWeightGradKernel, dim3(Grid(kTiles * chunks, 4)), dim3(kBlock), 0, ctx->stream,
                       G.act.as<uint16_t>(), G.grad.as<uint16_t>(), nrb, chunks, G.partial.as<float>());
    return CheckHip(ctx, hipGetLastError(), "guide backward");
}

AdamArgs MakeAdam(const GuideState& G, int64_t step) {
    AdamArgs a;
    a.lr = G.desc.learning_rate;
    a.beta1 = 0.9f;
    a.beta2 = 0.99f;
    a.eps = 1.0e-10f;
    a.c1 = static_cast<float>(1.0 - std::pow(0.9, static_cast<double>(step)));
    a.c2 = static_cast<float>(1.0 - std::pow(0.99, static_cast<double>(step)));
    return a;
}

}  // namespace

int GuidePrepareView(yrh_context* ctx, int64_t paths, int max_depth, yrg::GuideView* view) {
    *view = yrg::GuideView{};
    GuideState* G = ctx->guide;
    if (G == nullptr || G->mode == YRH_GUIDE_OFF) return YRH_OK;
    int rc;
    if (G->rows < paths) {
        if ((rc = Ensure(ctx, G->mix, static_cast<size_t>(paths) * yrg::kGuideMixtureRow * 4)) != YRH_OK) return rc;
        if ((rc = Ensure(ctx, G->valid, static_cast<size_t>(paths) * 4)) != YRH_OK) return rc;
        G->rows = paths;
    }
    if (G->mode == YRH_GUIDE_RECORD) {
        const int64_t slots = paths * std::max(1, max_depth);
        if (G->rec_slots < slots) {
            if ((rc = Ensure(ctx, G->rec, static_cast<size_t>(slots) * 64)) != YRH_OK) return rc;
            // A fresh record region holds no stamp of a past pass.
            hipLaunchKernelGGL(FillKernel, dim3(Grid(slots * 16)), dim3(kBlock), 0, ctx->stream, G->rec.as<float>(),
                               slots * 16, 0.0f);
            G->rec_slots = slots;
        }
    }
    if (G->mode == YRH_GUIDE_SAMPLE && !G->cells_fresh && (rc = PackCells(ctx, *G)) != YRH_OK) return rc;
    G->stamp = G->stamp == 0x3FFFFFFF ? 1 : G->stamp + 1;
    view->mode = G->mode;
    view->alpha = G->desc.guide_fraction;
    view->grid = G->grid;
    view->frag = G->frag_fwd.as<bf16x8_t>();
    view->table = G->table.as<float2>();
    view->mix = G->mix.as<float>();
    view->mix_stride = G->rows;
    view->valid = G->valid.as<int32_t>();
    view->rec = G->mode == YRH_GUIDE_RECORD ? G->rec.as<float4>() : nullptr;
    view->rec_paths = paths;
    view->rec_depths = std::max(1, max_depth);
    view->stamp = G->stamp;
    view->samples = G->samples.as<float4>();
    view->sample_count = G->sample_hdr.as<uint32_t>();
    view->sample_wsum = G->sample_hdr.as<float>() + 2;
    view->sample_capacity = G->desc.sample_capacity;
    view->depth_limit = G->depth_limit;
    // Recording passes interleave with training steps: they read the table.
    if (G->mode == YRH_GUIDE_SAMPLE && G->cell_mask != 0) {
        view->cells = G->cells.as<float4>();
        for (int l = 0; l < yrg::kGuideLevels; ++l) view->cell_off[l] = G->cell_off[l];
        view->cell_mask = G->cell_mask;
    }
    return CheckHip(ctx, hipGetLastError(), "guide view");
}

void DestroyGuideState(yrh_context* ctx) {
    delete ctx->guide;
    ctx->guide = nullptr;
}

}  // namespace yrh

using namespace yrh;

extern "C" {

int yrh_guide_configure(yrh_context* ctx, const yrh_guide_desc* d) {
    if (ctx == nullptr || d == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (d->log2_table_size < 10 || d->log2_table_size > 22 || d->min_resolution < 2 ||
        d->max_resolution < d->min_resolution || d->max_resolution > 4096 || !(d->guide_fraction >= 0.0f) ||
        d->guide_fraction > 1.0f || !(d->learning_rate > 0.0f) || d->sample_capacity < 0) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "invalid guide descriptor");
    }
    for (int a = 0; a < 3; ++a) {
        if (!(d->bounds_max[a] > d->bounds_min[a])) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "empty guide bounds");
    }
    (void)hipSetDevice(ctx->device);
    int rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (rc != YRH_OK) return rc;
    DestroyGuideState(ctx);
    auto* G = new GuideState();
    ctx->guide = G;
    G->desc = *d;
    yrg::GuideGrid& g = G->grid;
    g.table_size = 1u << d->log2_table_size;
    const double ratio = static_cast<double>(d->max_resolution) / static_cast<double>(d->min_resolution);
    for (int l = 0; l < yrg::kGuideLevels; ++l) {
        g.res[l] = static_cast<int32_t>(std::floor(d->min_resolution * std::pow(ratio, l / 7.0) + 0.5));
        g.dense[l] = std::pow(static_cast<double>(g.res[l]) + 1.0, 3.0) <= static_cast<double>(g.table_size) ? 1 : 0;
    }
    for (int a = 0; a < 3; ++a) {
        g.bmin[a] = d->bounds_min[a];
        g.inv_extent[a] = 1.0f / (d->bounds_max[a] - d->bounds_min[a]);
    }
    G->grid_params = static_cast<int64_t>(yrg::kGuideLevels) * g.table_size * 2;
    // He-uniform MLP weights; the output layer starts small so the first
    // proposal is a broad mixture; hash-grid features start near zero.
    uint64_t s = d->seed;
    std::vector<float> w(yrg::kMlpParams, 0.0f);
    for (int l = 0; l < 4; ++l) {
        const int in = yrg::kLayerIn[l], out = yrg::kLayerOut[l];
        const float a = std::sqrt(6.0f / static_cast<float>(in)) * (l == 3 ? 0.1f : 1.0f);
        for (int i = 0; i < in; ++i) {
            for (int o = 0; o < out; ++o) {
                const bool pad = l == 3 && o >= yrg::kGuideMixture;
                w[yrg::kLayerParam[l] + i * out + o] = pad ? 0.0f : Uniform(s, -a, a);
            }
        }
    }
    std::vector<float> t(static_cast<size_t>(G->grid_params));
    for (float& v : t) v = Uniform(s, -1.0e-4f, 1.0e-4f);
    const size_t mb = yrg::kMlpParams * 4, tb = static_cast<size_t>(G->grid_params) * 4;
    if ((rc = Alloc(ctx, G->mlp, mb)) != YRH_OK || (rc = Alloc(ctx, G->mlp_m, mb)) != YRH_OK ||
        (rc = Alloc(ctx, G->mlp_v, mb)) != YRH_OK ||
        (rc = Alloc(ctx, G->frag_fwd, yrg::kFragCount * 64 * 8)) != YRH_OK ||
        (rc = Alloc(ctx, G->frag_bwd, yrg::kFragCount * 64 * 8)) != YRH_OK || (rc = Alloc(ctx, G->table, tb)) != YRH_OK ||
        (rc = Alloc(ctx, G->table_m, tb)) != YRH_OK || (rc = Alloc(ctx, G->table_v, tb)) != YRH_OK ||
        (rc = Alloc(ctx, G->table_grad, tb)) != YRH_OK ||
        (rc = Alloc(ctx, G->samples, static_cast<size_t>(std::max(1, d->sample_capacity)) * 64)) != YRH_OK ||
        (rc = Alloc(ctx, G->sample_hdr, 16)) != YRH_OK) {
        return rc;
    }
    if ((rc = yrh_copy_to_device(ctx, G->mlp.ptr, w.data(), mb)) != YRH_OK) return rc;
    if ((rc = yrh_copy_to_device(ctx, G->table.ptr, t.data(), tb)) != YRH_OK) return rc;
    for (DevBuf* b : {&G->mlp_m, &G->mlp_v, &G->table_m, &G->table_v, &G->table_grad, &G->sample_hdr}) {
        if ((rc = CheckHip(ctx, hipMemsetAsync(b->ptr, 0, b->bytes, ctx->stream), "hipMemsetAsync")) != YRH_OK) return rc;
    }
    if ((rc = Pack(ctx, *G)) != YRH_OK) return rc;
    return CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int yrh_guide_set_mode(yrh_context* ctx, int32_t mode) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    if (mode < YRH_GUIDE_OFF || mode > YRH_GUIDE_RECORD) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "unknown guide mode");
    ctx->guide->mode = mode;
    return YRH_OK;
}

int yrh_guide_set_depth_limit(yrh_context* ctx, int32_t depths) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    if (depths < 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "negative guide depth limit");
    ctx->guide->depth_limit = depths;
    return YRH_OK;
}

int yrh_guide_set_cell_budget(yrh_context* ctx, int64_t bytes) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    if (bytes < 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "negative cell budget");
    GuideState& G = *ctx->guide;
    if (bytes != G.cell_budget) {
        // Earlier passes may still read the copies.
        if ((rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")) != YRH_OK) return rc;
        G.cell_budget = bytes;
        G.cells.reset();
        G.cell_mask = 0;
        G.cells_fresh = false;
    }
    return YRH_OK;
}

int yrh_guide_train(yrh_context* ctx, int32_t steps, int32_t batch) {
    const StageRange stage_range("yrh_guide_train");
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    if (steps < 0 || batch <= 0) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad steps or batch");
    GuideState& G = *ctx->guide;
    uint32_t* hdr = G.sample_hdr.as<uint32_t>();
    float* fh = G.sample_hdr.as<float>();
    for (int s = 0; s < steps; ++s) {
        if ((rc = BackwardPass(ctx, G, G.samples.as<float4>(), hdr, static_cast<uint32_t>(G.desc.sample_capacity),
                               fh + 2, 0.0f, static_cast<int64_t>(s) * batch, batch, fh + 3, 1)) != YRH_OK) {
            return rc;
        }
        ++G.adam_step;
        const AdamArgs a = MakeAdam(G, G.adam_step);
        hipLaunchKernelGGL(GuideAdamMlpKernel, dim3(Grid(yrg::kMlpParams)), dim3(kBlock), 0, ctx->stream,
                           G.partial.as<float>(), (static_cast<int64_t>(batch) + 16 * kChunkBlocks - 1) / (16 * kChunkBlocks),
                           G.mlp.as<float>(), G.mlp_m.as<float>(), G.mlp_v.as<float>(), static_cast<float*>(nullptr),
                           hdr, 1, a);
        hipLaunchKernelGGL(GuideAdamGridKernel, dim3(Grid(G.grid_params)), dim3(kBlock), 0, ctx->stream,
                           G.table.as<float>(), G.table_m.as<float>(), G.table_v.as<float>(), G.table_grad.as<float>(),
                           static_cast<float*>(nullptr), G.grid_params, 1, a);
        if ((rc = Pack(ctx, G)) != YRH_OK) return rc;
    }
    if (steps > 0) G.cells_fresh = false;
    hipLaunchKernelGGL(ResetHeaderKernel, dim3(1), dim3(64), 0, ctx->stream, hdr);
    return CheckHip(ctx, hipGetLastError(), "guide train");
}

int yrh_guide_add_samples(yrh_context* ctx, int64_t n, const float* samples) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    if (n < 0 || (n > 0 && samples == nullptr)) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad samples");
    if (n == 0) return YRH_OK;
    GuideState& G = *ctx->guide;
    DevBuf buf;
    if ((rc = Alloc(ctx, buf, static_cast<size_t>(n) * 52)) != YRH_OK) return rc;
    if ((rc = yrh_copy_to_device(ctx, buf.ptr, samples, static_cast<size_t>(n) * 52)) != YRH_OK) return rc;
    hipLaunchKernelGGL(GuideAddSamplesKernel, dim3(Grid(n)), dim3(kBlock), 0, ctx->stream, buf.as<float>(), n,
                       G.samples.as<float4>(), G.sample_hdr.as<uint32_t>(), G.desc.sample_capacity);
    if ((rc = CheckHip(ctx, hipGetLastError(), "guide add samples")) != YRH_OK) return rc;
    return CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int yrh_guide_read_samples(yrh_context* ctx, int64_t n, float* samples, int64_t* copied) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    if (n < 0 || (n > 0 && samples == nullptr)) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad samples");
    int64_t count = 0;
    if ((rc = yrh_guide_sample_stats(ctx, &count, nullptr)) != YRH_OK) return rc;
    const int64_t m = std::min(n, count);
    std::vector<float> raw(static_cast<size_t>(m) * 16);
    if (m > 0 && (rc = yrh_copy_to_host(ctx, raw.data(), ctx->guide->samples.ptr, raw.size() * 4)) != YRH_OK) return rc;
    for (int64_t r = 0; r < m; ++r) {
        const float* s = raw.data() + 16 * r;
        float* d = samples + 13 * r;
        d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
        d[3] = s[4]; d[4] = s[5]; d[5] = s[6];
        d[6] = s[8]; d[7] = s[9]; d[8] = s[10];
        d[9] = s[12]; d[10] = s[13]; d[11] = s[14];
        d[12] = s[3];
    }
    if (copied != nullptr) *copied = m;
    return YRH_OK;
}

int yrh_guide_sample_stats(yrh_context* ctx, int64_t* count, float* mean_weight) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    uint32_t h[4];
    if ((rc = CheckHip(ctx, hipMemcpyAsync(h, ctx->guide->sample_hdr.ptr, 16, hipMemcpyDeviceToHost, ctx->stream),
                       "hipMemcpyAsync")) != YRH_OK) {
        return rc;
    }
    if ((rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")) != YRH_OK) return rc;
    float wsum;
    std::memcpy(&wsum, &h[2], 4);
    const uint32_t kept = std::min<uint32_t>(h[0], static_cast<uint32_t>(ctx->guide->desc.sample_capacity));
    if (count != nullptr) *count = kept;
    if (mean_weight != nullptr) *mean_weight = kept > 0 ? wsum / static_cast<float>(kept) : 0.0f;
    return YRH_OK;
}

int yrh_guide_info(yrh_context* ctx, int32_t* mlp_params, int64_t* grid_params, int32_t* resolutions, int32_t* dense) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    const GuideState& G = *ctx->guide;
    if (mlp_params != nullptr) *mlp_params = yrg::kMlpParams;
    if (grid_params != nullptr) *grid_params = G.grid_params;
    for (int l = 0; l < yrg::kGuideLevels; ++l) {
        if (resolutions != nullptr) resolutions[l] = G.grid.res[l];
        if (dense != nullptr) dense[l] = G.grid.dense[l];
    }
    return YRH_OK;
}

int yrh_guide_params_read(yrh_context* ctx, float* mlp, float* grid) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    GuideState& G = *ctx->guide;
    if (mlp != nullptr && (rc = yrh_copy_to_host(ctx, mlp, G.mlp.ptr, yrg::kMlpParams * 4)) != YRH_OK) return rc;
    if (grid != nullptr && (rc = yrh_copy_to_host(ctx, grid, G.table.ptr, static_cast<size_t>(G.grid_params) * 4)) != YRH_OK) {
        return rc;
    }
    return YRH_OK;
}

int yrh_guide_params_write(yrh_context* ctx, const float* mlp, const float* grid) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    GuideState& G = *ctx->guide;
    if (mlp != nullptr && (rc = yrh_copy_to_device(ctx, G.mlp.ptr, mlp, yrg::kMlpParams * 4)) != YRH_OK) return rc;
    if (grid != nullptr && (rc = yrh_copy_to_device(ctx, G.table.ptr, grid, static_cast<size_t>(G.grid_params) * 4)) != YRH_OK) {
        return rc;
    }
    G.cells_fresh = false;
    if ((rc = Pack(ctx, G)) != YRH_OK) return rc;
    return CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int yrh_guide_query(yrh_context* ctx, int64_t n, const float* inputs, float* mixture) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    if (n < 0 || (n > 0 && (inputs == nullptr || mixture == nullptr))) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad rows");
    if (n == 0) return YRH_OK;
    GuideState& G = *ctx->guide;
    DevBuf in, out;
    if ((rc = Alloc(ctx, in, static_cast<size_t>(n) * 36)) != YRH_OK) return rc;
    if ((rc = Alloc(ctx, out, static_cast<size_t>(n) * 160)) != YRH_OK) return rc;
    if ((rc = yrh_copy_to_device(ctx, in.ptr, inputs, static_cast<size_t>(n) * 36)) != YRH_OK) return rc;
    hipLaunchKernelGGL(GuideQueryRowsKernel, dim3(Grid(n, 256)), dim3(kBlock), 0, ctx->stream, in.as<float>(), n,
                       G.frag_fwd.as<bf16x8_t>(), G.table.as<float2>(), G.grid, out.as<float>());
    if ((rc = CheckHip(ctx, hipGetLastError(), "guide query")) != YRH_OK) return rc;
    return yrh_copy_to_host(ctx, mixture, out.ptr, static_cast<size_t>(n) * 160);
}

int yrh_guide_gradients(yrh_context* ctx, int64_t n, const float* samples, float w_norm, float* grad_mlp,
                        float* grad_grid, float* loss) {
    int rc = NeedGuide(ctx);
    if (rc != YRH_OK) return rc;
    if (n <= 0 || samples == nullptr || !(w_norm > 0.0f)) return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad samples");
    GuideState& G = *ctx->guide;
    std::vector<float> rows(static_cast<size_t>(n) * 16, 0.0f);
    for (int64_t r = 0; r < n; ++r) {
        const float* s = samples + 13 * r;
        float* d = rows.data() + 16 * r;
        d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; d[3] = s[12];
        d[4] = s[3]; d[5] = s[4]; d[6] = s[5];
        d[8] = s[6]; d[9] = s[7]; d[10] = s[8];
        d[12] = s[9]; d[13] = s[10]; d[14] = s[11];
    }
    DevBuf buf, hdr, gout;
    if ((rc = Alloc(ctx, buf, rows.size() * 4)) != YRH_OK || (rc = Alloc(ctx, hdr, 16)) != YRH_OK ||
        (rc = Alloc(ctx, gout, yrg::kMlpParams * 4)) != YRH_OK) {
        return rc;
    }
    uint32_t h[4] = {static_cast<uint32_t>(n), 0, 0, 0};
    if ((rc = yrh_copy_to_device(ctx, buf.ptr, rows.data(), rows.size() * 4)) != YRH_OK) return rc;
    if ((rc = yrh_copy_to_device(ctx, hdr.ptr, h, 16)) != YRH_OK) return rc;
    if ((rc = BackwardPass(ctx, G, buf.as<float4>(), hdr.as<uint32_t>(), static_cast<uint32_t>(n), hdr.as<float>() + 2,
                           w_norm, 0, n, hdr.as<float>() + 3, 0)) != YRH_OK) {
        return rc;
    }
    const int64_t chunks = ((n + 15) / 16 + kChunkBlocks - 1) / kChunkBlocks;
    const AdamArgs a = MakeAdam(G, 1);
    hipLaunchKernelGGL(GuideAdamMlpKernel, dim3(Grid(yrg::kMlpParams)), dim3(kBlock), 0, ctx->stream,
                       G.partial.as<float>(), chunks, G.mlp.as<float>(), G.mlp_m.as<float>(), G.mlp_v.as<float>(),
                       gout.as<float>(), hdr.as<uint32_t>(), 0, a);
    DevBuf ggrid;
    if ((rc = Alloc(ctx, ggrid, static_cast<size_t>(G.grid_params) * 4)) != YRH_OK) return rc;
    hipLaunchKernelGGL(GuideAdamGridKernel, dim3(Grid(G.grid_params)), dim3(kBlock), 0, ctx->stream,
                       G.table.as<float>(), G.table_m.as<float>(), G.table_v.as<float>(), G.table_grad.as<float>(),
                       ggrid.as<float>(), G.grid_params, 0, a);
    if ((rc = CheckHip(ctx, hipGetLastError(), "guide gradients")) != YRH_OK) return rc;
    if (grad_mlp != nullptr && (rc = yrh_copy_to_host(ctx, grad_mlp, gout.ptr, yrg::kMlpParams * 4)) != YRH_OK) return rc;
    if (grad_grid != nullptr &&
        (rc = yrh_copy_to_host(ctx, grad_grid, ggrid.ptr, static_cast<size_t>(G.grid_params) * 4)) != YRH_OK) {
        return rc;
    }
    if ((rc = yrh_copy_to_host(ctx, h, hdr.ptr, 16)) != YRH_OK) return rc;
    if (loss != nullptr) std::memcpy(loss, &h[3], 4);
    return YRH_OK;
}

int yrh_guide_mixture_eval(yrh_context* ctx, int64_t n, const float* mixture, const float* dirs, const float* u,
                           float* pdf, float* samples) {
    if (ctx == nullptr) return YRH_E_INVALID_ARGUMENT;
    if (n < 0 || (n > 0 && (mixture == nullptr || dirs == nullptr || u == nullptr || pdf == nullptr || samples == nullptr))) {
        return Fail(ctx, YRH_E_INVALID_ARGUMENT, "bad rows");
    }
    if (n == 0) return YRH_OK;
    (void)hipSetDevice(ctx->device);
    DevBuf m, d, uu, p, s;
    int rc;
    if ((rc = Alloc(ctx, m, static_cast<size_t>(n) * 160)) != YRH_OK || (rc = Alloc(ctx, d, static_cast<size_t>(n) * 12)) != YRH_OK ||
        (rc = Alloc(ctx, uu, static_cast<size_t>(n) * 12)) != YRH_OK || (rc = Alloc(ctx, p, static_cast<size_t>(n) * 4)) != YRH_OK ||
        (rc = Alloc(ctx, s, static_cast<size_t>(n) * 16)) != YRH_OK) {
        return rc;
    }
    if ((rc = yrh_copy_to_device(ctx, m.ptr, mixture, static_cast<size_t>(n) * 160)) != YRH_OK ||
        (rc = yrh_copy_to_device(ctx, d.ptr, dirs, static_cast<size_t>(n) * 12)) != YRH_OK ||
        (rc = yrh_copy_to_device(ctx, uu.ptr, u, static_cast<size_t>(n) * 12)) != YRH_OK) {
        return rc;
    }
    hipLaunchKernelGGL(GuideMixtureEvalKernel, dim3(Grid(n)), dim3(kBlock), 0, ctx->stream, m.as<float>(), d.as<float>(),
                       uu.as<float>(), n, p.as<float>(), s.as<float>());
    if ((rc = CheckHip(ctx, hipGetLastError(), "mixture eval")) != YRH_OK) return rc;
    if ((rc = yrh_copy_to_host(ctx, pdf, p.ptr, static_cast<size_t>(n) * 4)) != YRH_OK) return rc;
    return yrh_copy_to_host(ctx, samples, s.ptr, static_cast<size_t>(n) * 16);
}

}  // extern "C"
