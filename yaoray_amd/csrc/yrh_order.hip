// Ray ordering for the batched ray-cast entry points (MI355X, gfx950).
//
// Rays whose origins lie close together visit the same upper BVH nodes, so a
// wave whose lanes hold such rays shares cache lines in L1/L2. Before a large
// ray-cast launch the ray indices are put in a counting-sort order by
// a 12-bit Morton code of the origin's cell in the scene box (16 cells per
// axis); the traversal kernel then claims rays in that order and writes every
// result to the ray's own index, so outputs are unchanged. A pinhole camera's
// rays (one cell) stay in runs of 64 consecutive pixels, and with them their
// coherence.
//
// Three passes over one launch's n rays, split into chunks of consecutive
// rays, one workgroup per chunk, each of its waves on a contiguous run of
// 64-ray steps:
//   histogram  per chunk, counts per cell in LDS -> H[cell * chunks + chunk]
//   scan       exclusive prefix sum of H (cell-major: cells in order, chunks
//              in order within a cell) in three kernels (tile sums, scan of
//              the tile sums, tiles)
//   scatter    per chunk, 64 rays at a time: free positions of the ray's
//              cell from LDS offsets seeded from H -> order[position] = ray
// The histogram pass stores each ray's cell (16 bits) for the scatter pass.
// 64 rays of one cell (a camera's rays) take 64 consecutive positions in lane
// order with one LDS atomic, so pixel runs stay whole; rays of mixed cells
// take positions with one LDS atomic each. Runs from different waves of a
// chunk may interleave within a cell (results never depend on the order,
// only cache behaviour does). Loads run kBatch 64-ray steps ahead.
#include "yrh_internal.h"

namespace yrh {
namespace {

constexpr int kCellBits = 12;
constexpr int kCells = 1 << kCellBits;
constexpr int kOrderBlock = 256;
constexpr int kScanPerThread = 16;
constexpr int kScanTile = kOrderBlock * kScanPerThread;  // H elements per scan tile
constexpr int kMaxChunks = 1024;
constexpr int64_t kMinChunk = 4096;
constexpr int kBatch = 8;  // 64-ray steps whose origins are loaded together
constexpr int kWaves = kOrderBlock / 64;  // waves per chunk, each on a contiguous run of steps

struct CellMap {
    float lo_x, lo_y, lo_z;
    float sx, sy, sz;  // cells per unit length (0 on a flat axis)
};

__device__ __forceinline__ uint32_t CellCoord(float o, float lo, float s) {
    // NaN -> 0 through fmaxf; outside the box -> the border cell.
    return static_cast<uint32_t>(fminf(fmaxf((o - lo) * s, 0.0f), 15.0f));
}

__device__ __forceinline__ uint32_t CellKey(const CellMap& m, const float4& o) {
    const uint32_t x = CellCoord(o.x, m.lo_x, m.sx);
    const uint32_t y = CellCoord(o.y, m.lo_y, m.sy);
    const uint32_t z = CellCoord(o.z, m.lo_z, m.sz);
    uint32_t key = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        key |= ((x >> b) & 1u) << (3 * b);
        key |= ((y >> b) & 1u) << (3 * b + 1);
        key |= ((z >> b) & 1u) << (3 * b + 2);
    }
    return key;
}

// Keys of `count` (<= kBatch) consecutive 64-ray steps from `step`, loaded together.
__device__ __forceinline__ void LoadKeys(const float4* __restrict__ org, const CellMap& m, int64_t begin,
                                         int64_t end, int64_t step, uint32_t (&key)[kBatch],
                                         bool (&valid)[kBatch]) {
    const int lane = threadIdx.x & 63;
    float4 o[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; ++j) {
        const int64_t i = begin + (step + j) * 64 + lane;
        valid[j] = i < end;
        o[j] = valid[j] ? org[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int j = 0; j < kBatch; ++j) key[j] = CellKey(m, o[j]);
}

// True when every valid lane of the wave holds the same cell (a camera's rays).
__device__ __forceinline__ bool OneCell(uint32_t key, bool valid, uint64_t& live) {
    live = __ballot(valid);
    const uint32_t first = __shfl(key, __ffsll(static_cast<long long>(live)) - 1);
    return __ballot(valid && key == first) == live;
}

// The 64-ray steps of a chunk taken by wave `wave` of kWaves: a contiguous run.
__device__ __forceinline__ void WaveSteps(int64_t begin, int64_t end, int wave, int64_t& s0, int64_t& s1) {
    const int64_t steps = end > begin ? (end - begin + 63) / 64 : 0;
    const int64_t per = (steps + kWaves - 1) / kWaves;
    s0 = per * wave < steps ? per * wave : steps;
    s1 = s0 + per < steps ? s0 + per : steps;
}

__global__ __launch_bounds__(kOrderBlock) void CellHistogramKernel(const float4* __restrict__ org, int64_t n,
                                                                   int64_t chunk, CellMap m,
                                                                   uint32_t* __restrict__ hist, int chunks,
                                                                   uint16_t* __restrict__ keys) {
    __shared__ uint32_t count[kCells];
    for (int c = threadIdx.x; c < kCells; c += kOrderBlock) count[c] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t begin = static_cast<int64_t>(blockIdx.x) * chunk;
    const int64_t end = begin + chunk < n ? begin + chunk : n;
    int64_t s0, s1;
    WaveSteps(begin, end, static_cast<int>(threadIdx.x >> 6), s0, s1);
    for (int64_t step = s0; step < s1; step += kBatch) {
        uint32_t key[kBatch];
        bool valid[kBatch];
        LoadKeys(org, m, begin, begin + s1 * 64 < end ? begin + s1 * 64 : end, step, key, valid);
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            if (valid[j]) keys[begin + (step + j) * 64 + lane] = static_cast<uint16_t>(key[j]);
            uint64_t live;
            if (OneCell(key[j], valid[j], live)) {
                if (live != 0ull && lane == __ffsll(static_cast<long long>(live)) - 1) {
                    atomicAdd(&count[key[j]], static_cast<uint32_t>(__popcll(live)));
                }
            } else if (valid[j]) {
                atomicAdd(&count[key[j]], 1u);
            }
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < kCells; c += kOrderBlock) {
        hist[static_cast<int64_t>(c) * chunks + blockIdx.x] = count[c];
    }
}

// Exclusive scan of a block's values (one per thread); returns the block total.
__device__ __forceinline__ uint32_t BlockExclusiveScan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t wave_sums[kOrderBlock / 64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t inc = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t up = __shfl_up(inc, off);
        if (lane >= off) inc += up;
    }
    if (lane == 63) wave_sums[wave] = inc;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int w = 0; w < kOrderBlock / 64; ++w) {
        if (w < wave) before += wave_sums[w];
        total += wave_sums[w];
    }
    __syncthreads();  // wave_sums is reused by the next call
    return before + inc - v;
}

__global__ __launch_bounds__(kOrderBlock) void ScanTileSumsKernel(const uint32_t* __restrict__ h, int64_t len,
                                                                  uint32_t* __restrict__ tile_sums) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanPerThread;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        if (base + k < len) s += h[base + k];
    }
    uint32_t total;
    (void)BlockExclusiveScan(s, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the tile sums in place.
__global__ __launch_bounds__(kOrderBlock) void ScanSumsKernel(uint32_t* __restrict__ tile_sums, int tiles) {
    uint32_t carry = 0;
    for (int start = 0; start < tiles; start += kOrderBlock) {
        const int i = start + static_cast<int>(threadIdx.x);
        const uint32_t v = i < tiles ? tile_sums[i] : 0u;
        uint32_t total;
        const uint32_t ex = BlockExclusiveScan(v, total);
        if (i < tiles) tile_sums[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(kOrderBlock) void ScanTilesKernel(uint32_t* __restrict__ h, int64_t len,
                                                               const uint32_t* __restrict__ tile_sums) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        v[k] = base + k < len ? h[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = BlockExclusiveScan(s, total) + tile_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        if (base + k < len) h[base + k] = run;
        run += v[k];
    }
}

__global__ __launch_bounds__(kOrderBlock) void CellScatterKernel(const uint16_t* __restrict__ keys, int64_t n,
                                                                 int64_t chunk, const uint32_t* __restrict__ offsets,
                                                                 int chunks, int32_t* __restrict__ order) {
    __shared__ uint32_t next[kCells];
    for (int c = threadIdx.x; c < kCells; c += kOrderBlock) {
        next[c] = offsets[static_cast<int64_t>(c) * chunks + blockIdx.x];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    const int64_t begin = static_cast<int64_t>(blockIdx.x) * chunk;
    const int64_t end = begin + chunk < n ? begin + chunk : n;
    int64_t s0, s1;
    WaveSteps(begin, end, static_cast<int>(threadIdx.x >> 6), s0, s1);
    const int64_t wave_end = begin + s1 * 64 < end ? begin + s1 * 64 : end;
    for (int64_t step = s0; step < s1; step += kBatch) {
        uint32_t key[kBatch];
        bool valid[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int64_t i = begin + (step + j) * 64 + lane;
            valid[j] = i < wave_end;
            key[j] = valid[j] ? keys[i] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int64_t i = begin + (step + j) * 64 + lane;
            uint32_t pos = 0;
            uint64_t live;
            if (OneCell(key[j], valid[j], live)) {
                // The step's rays take consecutive positions in lane order.
                const int leader = __ffsll(static_cast<long long>(live)) - 1;
                uint32_t base = 0;
                if (live != 0ull && lane == leader) base = atomicAdd(&next[key[j]], static_cast<uint32_t>(__popcll(live)));
                base = __shfl(base, leader < 0 ? 0 : leader);
                pos = base + static_cast<uint32_t>(__popcll(live & below));
            } else if (valid[j]) {
                pos = atomicAdd(&next[key[j]], 1u);
            }
            if (valid[j] && pos < static_cast<uint64_t>(n)) order[pos] = static_cast<int32_t>(i);
        }
    }
}

}  // namespace

int OrderRaysByOrigin(yrh_context* ctx, const float4* org, int64_t n, const int32_t** order_out) {
    *order_out = nullptr;
    const int chunks = static_cast<int>(std::min<int64_t>(kMaxChunks, std::max<int64_t>(1, (n + kMinChunk - 1) / kMinChunk)));
    const int64_t chunk = (n + chunks - 1) / chunks;
    const int64_t len = static_cast<int64_t>(kCells) * chunks;
    const int tiles = static_cast<int>((len + kScanTile - 1) / kScanTile);
    const size_t order_bytes = static_cast<size_t>(n) * sizeof(int32_t);
    const size_t key_bytes = (static_cast<size_t>(n) * sizeof(uint16_t) + 255) & ~static_cast<size_t>(255);
    const size_t scratch_bytes = static_cast<size_t>(len + tiles) * sizeof(uint32_t) + key_bytes;
    if (ctx->ray_order.bytes < order_bytes || ctx->ray_order_scratch.bytes < scratch_bytes) {
        // Earlier launches on the stream may still read the old buffers.
        int rc = CheckHip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        if (rc != YRH_OK) return rc;
        if (ctx->ray_order.bytes < order_bytes) {
            ctx->ray_order.reset();
            rc = CheckHip(ctx, hipMalloc(&ctx->ray_order.ptr, order_bytes), "hipMalloc ray order");
            if (rc != YRH_OK) return rc;
            ctx->ray_order.bytes = order_bytes;
        }
        if (ctx->ray_order_scratch.bytes < scratch_bytes) {
            ctx->ray_order_scratch.reset();
            rc = CheckHip(ctx, hipMalloc(&ctx->ray_order_scratch.ptr, scratch_bytes), "hipMalloc ray order scratch");
            if (rc != YRH_OK) return rc;
            ctx->ray_order_scratch.bytes = scratch_bytes;
        }
    }
    const DevGeometry& g = ctx->geo;
    CellMap m;
    m.lo_x = g.root_lo[0];
    m.lo_y = g.root_lo[1];
    m.lo_z = g.root_lo[2];
    const float ex = g.root_hi[0] - g.root_lo[0], ey = g.root_hi[1] - g.root_lo[1], ez = g.root_hi[2] - g.root_lo[2];
    m.sx = ex > 0.0f ? 16.0f / ex : 0.0f;
    m.sy = ey > 0.0f ? 16.0f / ey : 0.0f;
    m.sz = ez > 0.0f ? 16.0f / ez : 0.0f;
    uint32_t* hist = ctx->ray_order_scratch.as<uint32_t>();
    uint32_t* tile_sums = hist + len;
    uint16_t* keys = reinterpret_cast<uint16_t*>(tile_sums + tiles);
    int32_t* order = ctx->ray_order.as<int32_t>();
    hipLaunchKernelGGL(CellHistogramKernel, dim3(chunks), dim3(kOrderBlock), 0, ctx->stream, org, n, chunk, m, hist,
                       chunks, keys);
    hipLaunchKernelGGL(ScanTileSumsKernel, dim3(tiles), dim3(kOrderBlock), 0, ctx->stream, hist, len, tile_sums);
    hipLaunchKernelGGL(ScanSumsKernel, dim3(1), dim3(kOrderBlock), 0, ctx->stream, tile_sums, tiles);
    hipLaunchKernelGGL(ScanTilesKernel, dim3(tiles), dim3(kOrderBlock), 0, ctx->stream, hist, len, tile_sums);
    hipLaunchKernelGGL(CellScatterKernel, dim3(chunks), dim3(kOrderBlock), 0, ctx->stream, keys, n, chunk, hist,
                       chunks, order);
    const int rc = CheckHip(ctx, hipGetLastError(), "ray order launch");
    if (rc == YRH_OK) *order_out = order;
    return rc;
}

}  // namespace yrh
