// BVH traversal for MI355X (gfx950): the closest-hit / any-hit machinery
// shared by the ray-cast kernels (yrh_trace.hip) and the render kernels that
// trace inside a longer per-ray loop (yrh_render.hip).
//
// Arithmetic follows the reference operation by operation, with FP
// contraction disabled (the reference x86-64 build has no FMA):
//   slab test      include/yaoray/core/bounds.hpp:50-90
//   triangle test  src/accel/bvh_traversal.cpp:52-121 (Moller-Trumbore, two-sided,
//                  |det| < 1e-8 reject, equal-t tie -> lower flat triangle index)
//   traversal      src/accel/bvh_traversal.cpp:253-345 (root box, pop-skip when
//                  entry_t > best t, both child boxes, push far then near, ties
//                  near = left)
//   two-level      src/accel/two_level_bvh.cpp:365-481 (object ray through the
//                  host inverse, unnormalised direction so t is shared, BLAS run
//                  with t_max = best t, accepted only if strictly closer)
// so hit ids, t, barycentrics and the node/triangle test counters equal the
// reference's on the same tree.
//
// Execution model: a persistent grid sized to the device. Each wave64 refills
// idle lanes from a global ray counter with one atomic per wave (ballot +
// mbcnt), so divergent ray lengths do not idle the wave. The traversal stack
// lives in LDS (the first lds_levels entries per lane, a launch option up to
// kMaxLdsStack; laid out [level][lane] so a wave's pushes hit consecutive
// banks) and spills deeper levels to a per-lane global region; its capacity is the proven bound max_depth + 1 (one pending
// sibling per level plus the top pair).
//
// A policy object supplies each ray's first segment (Begin) and receives each
// finished segment with its origin and direction (Finish); Finish may hand
// back a further segment of the same ray, which the lane traverses next
// without giving up its slot. The plain ray-cast policy writes the hit and
// ends the ray; the transparent-shadow policy walks a shadow ray through
// dielectric surfaces this way.
#pragma once

#include "yrh_internal.h"

#pragma clang fp contract(off)

namespace yrh {
namespace trav {

constexpr int kMaxLdsStack = 16;  // LDS stack levels per lane, at most (launch option)
constexpr int kMaxBlock = 256;
constexpr int kRefillIdleLanes = 32;  // refill once half the wave is idle (swept: 8..56)

struct TraceArgs {
    const float4* org;
    const float4* dir;
    int32_t n;
    const uint32_t* n_device;    // when set, the ray count is read from device memory
    unsigned long long* totals;  // optional [0] node, [1] triangle, [2] sphere tests of the launch
    // outputs
    float4* hit;
    int32_t* inst_out;
    uint32_t* node_tests;
    uint32_t* triangle_tests;
    uint32_t* sphere_tests;
    uint32_t* occluded;
    // flat BVH or concatenated BLASes
    const DevNode* nodes;
    const float4* slots;
    const int32_t* leaf_count;
    int32_t root_code;
    float root_lo_x, root_lo_y, root_lo_z, root_hi_x, root_hi_y, root_hi_z;
    int32_t empty;
    // two-level
    const DevNode* tnodes;
    const int32_t* tleaf_count;
    const DevInstanceSlot* instance_slots;  // per TLAS leaf slot
    // scheduling
    uint32_t* counter;       // this launch's ray counter (zero at launch)
    uint32_t* next_counter;  // the next launch's counter(s), zeroed here (may be null)
    int32_t parts;           // 1: one counter; kParts: per-range counters kPartStride words apart
    const int32_t* order;    // claim position -> ray index (null = identity)
    uint32_t* overflow;      // stack overflow flag
    int2* spill;
    int64_t spill_stride;
    int32_t capacity;
    int32_t lds_levels;  // stack levels held in LDS (dynamic shared memory: lds_levels x block int2)
};

struct RayPre {
    float ox, oy, oz;
    float ix, iy, iz;
    bool nx, ny, nz;
};

__device__ __forceinline__ RayPre Precompute(float ox, float oy, float oz, float dx, float dy, float dz) {
    RayPre r;
    r.ox = ox;
    r.oy = oy;
    r.oz = oz;
    r.ix = 1.0f / dx;
    r.iy = 1.0f / dy;
    r.iz = 1.0f / dz;
    r.nx = signbit(r.ix);
    r.ny = signbit(r.iy);
    r.nz = signbit(r.iz);
    return r;
}

// IntersectBounds. The reference narrows the interval with comparisons
// (t_min = an > t_min ? an : t_min, t_max = af < t_max ? af : t_max), which
// keep the bound when a 0*inf product is NaN. fmaxf / fminf (IEEE maxNum /
// minNum, v_max3_f32 / v_min3_f32 here) return the same value for every
// input: a NaN operand yields the other operand, and the bounds themselves are
// never NaN; only the sign of a zero result may differ, which no comparison
// below can observe. Near / far planes are still chosen by direction sign, as
// the reference does, not by min / max of the two plane distances (that
// would differ on a NaN distance).
__device__ __forceinline__ bool Slab(const RayPre& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                     float t_min, float t_max, float& entry) {
    const float anx = ((r.nx ? hx : lx) - r.ox) * r.ix;
    const float afx = ((r.nx ? lx : hx) - r.ox) * r.ix;
    const float any = ((r.ny ? hy : ly) - r.oy) * r.iy;
    const float afy = ((r.ny ? ly : hy) - r.oy) * r.iy;
    const float anz = ((r.nz ? hz : lz) - r.oz) * r.iz;
    const float afz = ((r.nz ? lz : hz) - r.oz) * r.iz;
    t_min = fmaxf(fmaxf(anx, any), fmaxf(anz, t_min));
    t_max = fminf(fminf(afx, afy), fminf(afz, t_max));
    entry = t_min;
    return t_max >= t_min;
}

struct Best {
    float t;
    float u;
    float v;
    int32_t id;
    bool hit;
};

// IntersectTriangle on a (p0, e1, e2) slot. Returns true when `best` changed.
__device__ __forceinline__ bool TriangleTest(const float4& s0, const float4& s1, const float4& s2,
                                             float ox, float oy, float oz, float dx, float dy, float dz,
                                             float t_min, bool any_hit, Best& best) {
    const float e1x = s1.x, e1y = s1.y, e1z = s1.z;
    const float e2x = s2.x, e2y = s2.y, e2z = s2.z;
    // pvec = Cross(d, e2)
    const float px = dy * e2z - dz * e2y;
    const float py = dz * e2x - dx * e2z;
    const float pz = dx * e2y - dy * e2x;
    const float det = e1x * px + e1y * py + e1z * pz;
    if (fabsf(det) < 1.0e-8f) return false;
    const float inv_det = 1.0f / det;
    const float tx = ox - s0.x;
    const float ty = oy - s0.y;
    const float tz = oz - s0.z;
    const float u = (tx * px + ty * py + tz * pz) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    // qvec = Cross(tvec, e1)
    const float qx = ty * e1z - tz * e1y;
    const float qy = tz * e1x - tx * e1z;
    const float qz = tx * e1y - ty * e1x;
    const float v = (dx * qx + dy * qy + dz * qz) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    const float t = (e2x * qx + e2y * qy + e2z * qz) * inv_det;
    if (t <= t_min || t > best.t) return false;
    const int32_t id = __float_as_int(s0.w);
    if (!any_hit && t == best.t && best.hit && best.id >= 0 && id >= best.id) return false;
    best.hit = true;
    best.t = t;
    best.id = id;
    best.u = u;
    best.v = v;
    return true;
}

// Sphere slot of a leaf (device layout): (centre, id bits) (radius, 0, 0, -1)
// (0, 0, 0, -1) with id = -2 - sphere index; triangle slots carry their flat
// triangle index (>= 0) in the same word.
__device__ __forceinline__ bool IsSphereId(int32_t id) { return id <= -2; }

// SphereUv (src/geometry/intersection.cpp:136-151) of the outward normal at
// ray.At(t): the (u, v) the reference stores in a sphere hit.
__device__ __forceinline__ void SphereUv(float cx, float cy, float cz, float ox, float oy, float oz, float dx, float dy,
                                         float dz, float t, float& u, float& v) {
    const float px = ox + dx * t, py = oy + dy * t, pz = oz + dz * t;
    const float nx = px - cx, ny = py - cy, nz = pz - cz;
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
    if (len > 0.0f) {
        n0 = nx / len;
        n1 = ny / len;
        n2 = nz / len;
    }
    constexpr float kPi = 3.14159265358979323846f;
    const float y = fminf(fmaxf(n1, -1.0f), 1.0f);
    u = (atan2f(n0, n2) + kPi) / (2.0f * kPi);
    v = (kPi - acosf(y)) / kPi;
}

// IntersectAnalyticSphere (src/accel/bvh_traversal.cpp:123-150) over
// IntersectSphere (src/geometry/intersection.cpp:108-128), same operations in
// the same order: the nearer root if it lies in (t_min, best t), else the
// farther one; a hit is strictly closer than `best` and clears the triangle
// id. (u, v) = SphereUv when `with_uv` (ray casts report it; renders
// recompute the normal). Returns true when `best` changed.
__device__ __forceinline__ bool SphereTest(float cx, float cy, float cz, float radius, int32_t id, float ox, float oy,
                                           float oz, float dx, float dy, float dz, float t_min, bool with_uv,
                                           Best& best) {
    const float ocx = ox - cx, ocy = oy - cy, ocz = oz - cz;
    const float a = dx * dx + dy * dy + dz * dz;
    const float b = 2.0f * (ocx * dx + ocy * dy + ocz * dz);
    const float c = (ocx * ocx + ocy * ocy + ocz * ocz) - radius * radius;
    const float disc = b * b - 4.0f * a * c;
    if (disc < 0.0f) return false;
    const float sq = sqrtf(disc);
    const float inv_2a = 0.5f / a;
    float t = (-b - sq) * inv_2a;
    if (!(t > t_min && t < best.t)) {
        t = (-b + sq) * inv_2a;
        if (!(t > t_min && t < best.t)) return false;
    }
    best.hit = true;
    best.t = t;
    best.id = id;
    best.u = 0.0f;
    best.v = 0.0f;
    if (with_uv) SphereUv(cx, cy, cz, ox, oy, oz, dx, dy, dz, t, best.u, best.v);
    return true;
}

// LDS word of the traversal stack, in the LDS address space so that its
// stores stay ds_write (a generic pointer lets the compiler merge them with
// the spill stores into flat stores).
typedef __attribute__((address_space(3))) uint64_t LdsWord;

// Per-lane traversal stack. The top entry is held in registers: a push writes
// the previous top to memory and a pop of the register top reads nothing, so
// the common interior step (push far, push near, pop near) costs one LDS
// write instead of two writes and a read. Entries 0 .. sp-1; when top_valid
// the entry sp-1 is `top` and only 0 .. sp-2 are in memory.
struct Stack {
    LdsWord* lds;  // this lane's column: lds[level * stride], (code, entry bits) packed in 64 bits
    int lds_stride;
    int2* spill;  // this lane's column in the global spill region
    int64_t spill_stride;
    int capacity;
    int lds_levels;
    int sp;
    bool overflow;
    bool top_valid;
    int2 top;

    __device__ __forceinline__ void Store(int level, int2 e) {
        if (level < lds_levels) {
            lds[level * lds_stride] = static_cast<uint64_t>(static_cast<uint32_t>(e.x)) |
                                      (static_cast<uint64_t>(static_cast<uint32_t>(e.y)) << 32);
        } else {
            spill[static_cast<int64_t>(level - lds_levels) * spill_stride] = e;
        }
    }
    __device__ __forceinline__ void Push(int32_t code, float entry) {
        if (sp >= capacity) {
            overflow = true;
            return;
        }
        if (top_valid) Store(sp - 1, top);
        top = make_int2(code, __float_as_int(entry));
        top_valid = true;
        ++sp;
    }
    __device__ __forceinline__ int2 Pop() {
        --sp;
        if (top_valid) {
            top_valid = false;
            return top;
        }
        if (sp < lds_levels) {
            const uint64_t w = lds[sp * lds_stride];
            return make_int2(static_cast<int32_t>(static_cast<uint32_t>(w)), static_cast<int32_t>(w >> 32));
        }
        return spill[static_cast<int64_t>(sp - lds_levels) * spill_stride];
    }
    // The children of an interior node, pushed as the reference does: with
    // both hit, the far one first and the near one (popped next) on top;
    // ties go left. Branch-light form of the two Push calls, with the same
    // stack contents; on overflow both are dropped (the launch then fails).
    __device__ __forceinline__ void PushChildren(bool hl, bool hr, int32_t cl, float el, int32_t cr, float er) {
        const int k = static_cast<int>(hl) + static_cast<int>(hr);
        if (k == 0) return;
        if (sp + k > capacity) {
            overflow = true;
            return;
        }
        const bool left_near = hl && (!hr || el <= er);
        const int2 left = make_int2(cl, __float_as_int(el));
        const int2 right = make_int2(cr, __float_as_int(er));
        if (top_valid) Store(sp - 1, top);
        if (k == 2) Store(sp, left_near ? right : left);
        top = left_near ? left : right;
        top_valid = true;
        sp += k;
    }
    // Drops every entry at or above `level`.
    __device__ __forceinline__ void Truncate(int level) {
        if (level < sp) {
            sp = level;
            top_valid = false;
        }
    }
};

__device__ __forceinline__ void DecodeLeaf(int32_t code, const int32_t* leaf_count, int32_t& first, int32_t& count) {
    first = code & kMaxSlot;
    const int32_t nib = (code >> 27) & 0xF;
    count = nib < kLeafCountEscape ? nib + 1 : leaf_count[first];
}

// One interior step on a fetched node record (a, b, c = both child boxes,
// d = child codes): test both boxes and push the children as the reference
// does (far first, so the near child is popped next; ties go left).
__device__ __forceinline__ void InteriorStep(const float4& na, const float4& nb, const float4& nc, const float4& nd,
                                             const RayPre& r, float t_min, float t_max, uint32_t& nt, Stack& st) {
    nt += 2;
    float el, er;
    const bool hl = Slab(r, na.x, na.y, na.z, na.w, nb.x, nb.y, t_min, t_max, el);
    const bool hr = Slab(r, nb.z, nb.w, nc.x, nc.y, nc.z, nc.w, t_min, t_max, er);
    st.PushChildren(hl, hr, __float_as_int(nd.x), el, __float_as_int(nd.y), er);
}

__device__ __forceinline__ void InteriorStep(const DevNode* __restrict__ nodes, int32_t code, const RayPre& r,
                                             float t_min, float t_max, uint32_t& nt, Stack& st) {
    const float4* q = reinterpret_cast<const float4*>(nodes + code);
    InteriorStep(q[0], q[1], q[2], q[3], r, t_min, t_max, nt, st);
}

// The record a lane fetches in one traversal iteration: a 64-byte node or a
// 48-byte triangle slot, loaded before either is processed, so a wave whose
// lanes are split between interior nodes and leaves waits for one memory
// latency per iteration instead of one per branch and per triangle.
struct Fetch {
    float4 q0, q1, q2, q3;
    __device__ __forceinline__ void Load(const float4* p, bool node) {
        q0 = p[0];
        q1 = p[1];
        q2 = p[2];
        if (node) q3 = p[3];
    }
};

__device__ __forceinline__ uint64_t LanesBelow(int lane) { return (1ull << lane) - 1ull; }

// Wave-level sum of the per-lane test counts (node, triangle, sphere), one
// atomic per wave and count.
__device__ __forceinline__ void AddTotals(unsigned long long* totals, uint64_t nt, uint64_t tt, uint64_t ts,
                                          int lane) {
    for (int off = 32; off > 0; off >>= 1) {
        nt += __shfl_xor(static_cast<unsigned long long>(nt), off);
        tt += __shfl_xor(static_cast<unsigned long long>(tt), off);
        ts += __shfl_xor(static_cast<unsigned long long>(ts), off);
    }
    if (lane == 0) {
        atomicAdd(totals, static_cast<unsigned long long>(nt));
        atomicAdd(totals + 1, static_cast<unsigned long long>(tt));
        if (ts != 0) atomicAdd(totals + 2, static_cast<unsigned long long>(ts));
    }
}

// Ray claiming split into kParts ranges of the claim order, one counter each
// (kPartStride words apart, so no two share a cache line). Workgroups are
// dispatched to the 8 XCDs round-robin, so workgroup b starts on range
// b % kParts: each XCD works through its own contiguous stretch of the
// (origin-sorted) rays and its L2 keeps the nodes those rays share. A wave
// whose range runs out moves on to the next range, so every ray is taken.
constexpr int kParts = 8;
constexpr int kPartStride = 32;

__device__ __forceinline__ int64_t PartBegin(int part, int32_t n) {
    return static_cast<int64_t>(part) * n / kParts;
}

__device__ __forceinline__ int32_t ClaimRayParts(uint32_t* counters, int& part_state, bool want, int lane, int32_t n,
                                                 bool& exhausted) {
    int part = part_state & 0xFF;
    int parts_left = part_state >> 8;
    uint64_t mask = __ballot(want);
    int32_t result = -1;
    while (mask != 0 && parts_left > 0) {
        const int64_t lo = PartBegin(part, n);
        const int64_t size = PartBegin(part + 1, n) - lo;
        const int leader = __ffsll(static_cast<long long>(mask)) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(counters + part * kPartStride, static_cast<uint32_t>(__popcll(mask)));
        base = __shfl(base, leader);
        if ((mask >> lane) & 1ull) {
            const uint64_t mine = static_cast<uint64_t>(base) + static_cast<uint64_t>(__popcll(mask & LanesBelow(lane)));
            if (mine < static_cast<uint64_t>(size)) result = static_cast<int32_t>(lo + static_cast<int64_t>(mine));
        }
        mask = __ballot(want && result < 0);
        if (mask != 0) {  // this range is used up
            part = part + 1 == kParts ? 0 : part + 1;
            --parts_left;
        }
    }
    part_state = part | (parts_left << 8);
    if (want && result < 0) exhausted = true;
    return result;
}

// Claims one ray index per wanting lane with a single atomic per wave.
__device__ __forceinline__ int32_t ClaimRay(uint32_t* counter, bool want, int lane, int32_t n, bool& exhausted) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return -1;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
    base = __shfl(base, leader);
    if (!want) return -1;
    const uint32_t mine = base + static_cast<uint32_t>(__popcll(mask & LanesBelow(lane)));
    if (mine >= static_cast<uint32_t>(n)) {
        exhausted = true;
        return -1;
    }
    return static_cast<int32_t>(mine);
}

__device__ __forceinline__ void TransformRay(const float4& r0, const float4& r1, const float4& r2, float ox, float oy,
                                             float oz, float dx, float dy, float dz, float& rox, float& roy,
                                             float& roz, float& rdx, float& rdy, float& rdz) {
    // TransformPoint / TransformVector (src/core/transform.cpp:30-44), left to right.
    rox = r0.x * ox + r0.y * oy + r0.z * oz + r0.w;
    roy = r1.x * ox + r1.y * oy + r1.z * oz + r1.w;
    roz = r2.x * ox + r2.y * oy + r2.z * oz + r2.w;
    rdx = r0.x * dx + r0.y * dy + r0.z * dz;
    rdy = r1.x * dx + r1.y * dy + r1.z * dz;
    rdz = r2.x * dx + r2.y * dy + r2.z * dz;
}

// The ray-cast policy: rays from (org, dir) arrays, results to the output arrays.
template <bool kAny>
struct RayArrayPolicy {
    const TraceArgs* a;
    __device__ __forceinline__ void Begin(int32_t ray, float4& o, float4& d) const {
        o = a->org[ray];
        d = a->dir[ray];
    }
    __device__ __forceinline__ bool Finish(int32_t ray, const Best& best, int32_t inst, uint32_t nt, uint32_t tt,
                                           uint32_t ts, const float4&, const float4&, float4&, float4&) const {
        if (kAny) {
            a->occluded[ray] = best.hit ? 1u : 0u;
        } else {
            a->hit[ray] = make_float4(best.t, best.u, best.v, __int_as_float(best.hit ? best.id : -1));
            if (a->inst_out != nullptr) a->inst_out[ray] = best.hit ? inst : -1;
        }
        if (a->node_tests != nullptr) a->node_tests[ray] = nt;
        if (a->triangle_tests != nullptr) a->triangle_tests[ray] = tt;
        if (a->sphere_tests != nullptr) a->sphere_tests[ray] = ts;
        return false;
    }
};

// kSpheres: the tree may hold sphere slots (a scene without spheres runs the
// variant without the sphere branch, which costs its triangle tests ~2 %).
template <bool kAny, bool kSpheres, class Policy>
__device__ __forceinline__ void TraceFlatLoop(const TraceArgs& a, const Policy& pol, int2* lds_stack) {
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int64_t gtid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    Stack st{(LdsWord*)(lds_stack) + threadIdx.x, static_cast<int>(blockDim.x), a.spill + gtid, a.spill_stride,
             a.capacity, a.lds_levels, 0, false, false, make_int2(0, 0)};

    if (a.next_counter != nullptr && blockIdx.x == 0 && threadIdx.x < static_cast<unsigned>(a.parts)) {
        a.next_counter[threadIdx.x * (a.parts > 1 ? kPartStride : 1)] = 0u;
    }
    const int32_t n = a.n_device != nullptr ? static_cast<int32_t>(*a.n_device) : a.n;
    // Claim range (low byte) and ranges left to try (next byte), wave-uniform.
    int part_state = static_cast<int>(blockIdx.x % kParts) | (kParts << 8);
    uint32_t sum_nt = 0, sum_tt = 0, sum_ts = 0;  // this lane's test counts since its last flush
    int32_t ray = -1;
    bool exhausted = false;
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, t_min = 0;
    RayPre pre{};
    Best best{0, 0, 0, -1, false};
    uint32_t nt = 0, tt = 0, ts = 0;
    int32_t cursor = 0, leaf_end = 0;  // primitive slots of the leaf being tested

    auto start = [&](const float4& o4, const float4& d4) {
        ox = o4.x; oy = o4.y; oz = o4.z; t_min = o4.w;
        dx = d4.x; dy = d4.y; dz = d4.z;
        best = Best{d4.w, 0.0f, 0.0f, -1, false};
        nt = 0;
        tt = 0;
        ts = 0;
        cursor = leaf_end = 0;
        pre = Precompute(ox, oy, oz, dx, dy, dz);
        if (!a.empty) {
            ++nt;
            float entry;
            if (Slab(pre, a.root_lo_x, a.root_lo_y, a.root_lo_z, a.root_hi_x, a.root_hi_y, a.root_hi_z, t_min,
                     best.t, entry)) {
                st.Push(a.root_code, entry);
            }
        }
    };

    while (true) {
        if (st.sp == 0 && cursor >= leaf_end && ray >= 0) {
            sum_nt += nt;
            sum_tt += tt;
            sum_ts += ts;
            if (a.totals != nullptr && (sum_nt | sum_tt | sum_ts) >= 0x80000000u) {  // keep 32-bit sums from wrapping
                atomicAdd(a.totals, static_cast<unsigned long long>(sum_nt));
                atomicAdd(a.totals + 1, static_cast<unsigned long long>(sum_tt));
                atomicAdd(a.totals + 2, static_cast<unsigned long long>(sum_ts));
                sum_nt = sum_tt = sum_ts = 0;
            }
            float4 no, nd;
            if (pol.Finish(ray, best, -1, nt, tt, ts, make_float4(ox, oy, oz, t_min), make_float4(dx, dy, dz, 0.0f),
                           no, nd)) {
                start(no, nd);
            } else {
                ray = -1;
            }
        }
        const bool want = ray < 0 && !exhausted;
        const int32_t claimed = a.parts > 1 ? ClaimRayParts(a.counter, part_state, want, lane, n, exhausted)
                                            : ClaimRay(a.counter, want, lane, n, exhausted);
        if (claimed >= 0) {
            ray = claimed;
            if (a.order != nullptr) {
                const int32_t mapped = a.order[claimed];
                if (static_cast<uint32_t>(mapped) < static_cast<uint32_t>(n)) ray = mapped;
            }
            float4 o4, d4;
            pol.Begin(ray, o4, d4);
            start(o4, d4);
        }
        if (__ballot(st.sp > 0 || cursor < leaf_end || ray >= 0 || !exhausted) == 0) break;

        while (true) {
            // Pick this iteration's record: the next triangle of the current
            // leaf, else the next stack entry (a popped leaf starts its first
            // triangle in the same iteration; an entry beyond the best hit
            // is dropped as the reference's pop-skip does).
            bool node = false, tri = false;
            const float4* src = nullptr;
            if (cursor < leaf_end) {
                tri = true;
            } else if (st.sp > 0) {
                const int2 e = st.Pop();
                if (!(__int_as_float(e.y) > best.t)) {
                    if (e.x < 0) {
                        int32_t first, count;
                        DecodeLeaf(e.x, a.leaf_count, first, count);
                        cursor = first;
                        leaf_end = first + count;
                        tri = count > 0;
                    } else {
                        node = true;
                        src = reinterpret_cast<const float4*>(a.nodes + e.x);
                    }
                }
            }
            if (tri) src = a.slots + 3 * static_cast<int64_t>(cursor);
            Fetch f;
            if (node || tri) f.Load(src, node);
            if (node) {
                InteriorStep(f.q0, f.q1, f.q2, f.q3, pre, t_min, best.t, nt, st);
            } else if (tri) {
                ++cursor;
                bool changed;
                if (kSpheres && IsSphereId(__float_as_int(f.q0.w))) {
                    ++ts;
                    changed = SphereTest(f.q0.x, f.q0.y, f.q0.z, f.q1.x, __float_as_int(f.q0.w), ox, oy, oz, dx, dy,
                                         dz, t_min, !kAny, best);
                } else {
                    ++tt;
                    changed = TriangleTest(f.q0, f.q1, f.q2, ox, oy, oz, dx, dy, dz, t_min, kAny, best);
                }
                if (changed && kAny) {
                    st.Truncate(0);
                    cursor = leaf_end;
                }
            }
            const bool lane_busy = cursor < leaf_end || st.sp > 0;
            const uint64_t active = __ballot(lane_busy);
            if (active == 0) break;
            const uint64_t idle = __ballot(!lane_busy && (ray >= 0 || !exhausted));
            if (__popcll(idle) >= kRefillIdleLanes) break;
        }
    }
    if (st.overflow) atomicOr(a.overflow, 1u);
    if (a.totals != nullptr) AddTotals(a.totals, sum_nt, sum_tt, sum_ts, lane);
}

template <bool kAny, bool kSpheres, class Policy>
__device__ __forceinline__ void TraceTwoLevelLoop(const TraceArgs& a, const Policy& pol, int2* lds_stack) {
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const int64_t gtid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    Stack st{(LdsWord*)(lds_stack) + threadIdx.x, static_cast<int>(blockDim.x), a.spill + gtid, a.spill_stride,
             a.capacity, a.lds_levels, 0, false, false, make_int2(0, 0)};

    if (a.next_counter != nullptr && blockIdx.x == 0 && threadIdx.x < static_cast<unsigned>(a.parts)) {
        a.next_counter[threadIdx.x * (a.parts > 1 ? kPartStride : 1)] = 0u;
    }
    const int32_t n = a.n_device != nullptr ? static_cast<int32_t>(*a.n_device) : a.n;
    // Claim range (low byte) and ranges left to try (next byte), wave-uniform.
    int part_state = static_cast<int>(blockIdx.x % kParts) | (kParts << 8);
    uint32_t sum_nt = 0, sum_tt = 0, sum_ts = 0;  // this lane's test counts since its last flush
    int32_t ray = -1;
    bool exhausted = false;
    float ox = 0, oy = 0, oz = 0, dx = 0, dy = 0, dz = 0, t_min = 0;
    // Precomputed reciprocals of the ray being traversed: the world ray in the
    // TLAS, the object ray inside a BLAS (recomputed, bit-identically, when
    // the BLAS is left), so one set of registers serves both levels.
    RayPre pre{};
    Best best{0, 0, 0, -1, false};
    int32_t best_inst = -1;
    uint32_t nt = 0, tt = 0, ts = 0;
    // object-space state of the BLAS being traversed
    bool in_blas = false;
    int blas_base = 0;
    int32_t cur_inst = -1;
    float bx = 0, by = 0, bz = 0, bdx = 0, bdy = 0, bdz = 0;
    Best bbest{0, 0, 0, -1, false};
    // TLAS leaf being expanded
    int32_t cursor = 0, leaf_end = 0;
    // triangle slots of the BLAS leaf being tested
    int32_t tcur = 0, tend = 0;

    auto start = [&](const float4& o4, const float4& d4) {
        ox = o4.x; oy = o4.y; oz = o4.z; t_min = o4.w;
        dx = d4.x; dy = d4.y; dz = d4.z;
        best = Best{d4.w, 0.0f, 0.0f, -1, false};
        best_inst = -1;
        nt = 0;
        tt = 0;
        ts = 0;
        in_blas = false;
        cursor = leaf_end = 0;
        tcur = tend = 0;
        pre = Precompute(ox, oy, oz, dx, dy, dz);
        if (!a.empty) {
            ++nt;
            float entry;
            if (Slab(pre, a.root_lo_x, a.root_lo_y, a.root_lo_z, a.root_hi_x, a.root_hi_y, a.root_hi_z, t_min,
                     best.t, entry)) {
                st.Push(a.root_code, entry);
            }
        }
    };

    while (true) {
        const bool busy = in_blas || cursor < leaf_end || st.sp > 0;
        if (!busy && ray >= 0) {
            sum_nt += nt;
            sum_tt += tt;
            sum_ts += ts;
            if (a.totals != nullptr && (sum_nt | sum_tt | sum_ts) >= 0x80000000u) {  // keep 32-bit sums from wrapping
                atomicAdd(a.totals, static_cast<unsigned long long>(sum_nt));
                atomicAdd(a.totals + 1, static_cast<unsigned long long>(sum_tt));
                atomicAdd(a.totals + 2, static_cast<unsigned long long>(sum_ts));
                sum_nt = sum_tt = sum_ts = 0;
            }
            float4 no, nd;
            if (pol.Finish(ray, best, best_inst, nt, tt, ts, make_float4(ox, oy, oz, t_min),
                           make_float4(dx, dy, dz, 0.0f), no, nd)) {
                start(no, nd);
            } else {
                ray = -1;
            }
        }
        const bool want = ray < 0 && !exhausted;
        const int32_t claimed = a.parts > 1 ? ClaimRayParts(a.counter, part_state, want, lane, n, exhausted)
                                            : ClaimRay(a.counter, want, lane, n, exhausted);
        if (claimed >= 0) {
            ray = claimed;
            if (a.order != nullptr) {
                const int32_t mapped = a.order[claimed];
                if (static_cast<uint32_t>(mapped) < static_cast<uint32_t>(n)) ray = mapped;
            }
            float4 o4, d4;
            pol.Begin(ray, o4, d4);
            start(o4, d4);
        }
        if (__ballot(in_blas || cursor < leaf_end || st.sp > 0 || ray >= 0 || !exhausted) == 0) break;

        while (true) {
            // One record per iteration, as in the flat loop: a BLAS leaf's
            // triangles one by one, a node from the stack, or a BLAS's root.
            // A lane that finishes a BLAS or expands a TLAS leaf tries its
            // next instance in the same iteration, and an instance whose root
            // box is hit goes straight to its root record (no stack round
            // trip): fewer iterations per instance visit, the same
            // operations in the same order.
            bool node = false, tri = false;
            const float4* src = nullptr;
            bool try_instance = false;
            if (in_blas && tcur < tend) {
                tri = true;
            } else if (in_blas && st.sp == blas_base) {
                // BLAS finished: accept only a strictly closer hit (two_level_bvh.cpp:474);
                // an occlusion query accepts any hit within [t_min, t_max].
                if (bbest.hit && (kAny || bbest.t < best.t)) {
                    best = bbest;
                    best_inst = cur_inst;
                    if (kAny) {
                        st.Truncate(0);
                        cursor = leaf_end;
                    }
                }
                in_blas = false;
                pre = Precompute(ox, oy, oz, dx, dy, dz);
                ++cursor;
                try_instance = true;
            } else if (!in_blas && cursor < leaf_end) {
                try_instance = true;
            } else if (st.sp > 0) {
                const int2 e = st.Pop();
                const int32_t code = e.x;
                const float entry = __int_as_float(e.y);
                if (in_blas) {
                    if (!(entry > bbest.t)) {
                        if (code < 0) {
                            int32_t first, count;
                            DecodeLeaf(code, a.leaf_count, first, count);
                            tcur = first;
                            tend = first + count;
                            tri = count > 0;
                        } else {
                            node = true;
                            src = reinterpret_cast<const float4*>(a.nodes + code);
                        }
                    }
                } else if (!(entry > best.t)) {
                    if (code < 0) {
                        DecodeLeaf(code, a.tleaf_count, cursor, leaf_end);
                        leaf_end += cursor;
                        try_instance = true;
                    } else {
                        node = true;
                        src = reinterpret_cast<const float4*>(a.tnodes + code);
                    }
                }
            }
            if (try_instance && cursor < leaf_end) {
                const DevInstanceSlot m = a.instance_slots[cursor];
                const int32_t slot_id = __float_as_int(m.hi.w);
                if (kSpheres && IsSphereId(slot_id)) {
                    // A TLAS sphere (IntersectTlasSphere, two_level_bvh.cpp:231-253):
                    // tested against the world ray, no instance.
                    ++ts;
                    ++cursor;
                    if (SphereTest(m.r0.x, m.r0.y, m.r0.z, m.r0.w, slot_id, ox, oy, oz, dx, dy, dz, t_min, !kAny,
                                   best)) {
                        best_inst = -1;
                        if (kAny) {
                            st.Truncate(0);
                            cursor = leaf_end;
                        }
                    }
                } else {
                    // Next instance of the TLAS leaf: object ray, BLAS root box.
                    cur_inst = slot_id;
                    const DevInstanceSlot& root = m;
                    TransformRay(m.r0, m.r1, m.r2, ox, oy, oz, dx, dy, dz, bx, by, bz, bdx, bdy, bdz);
                    const RayPre bpre = Precompute(bx, by, bz, bdx, bdy, bdz);
                    ++nt;
                    float entry;
                    if (Slab(bpre, root.lo.x, root.lo.y, root.lo.z, root.hi.x, root.hi.y, root.hi.z, t_min, best.t,
                             entry)) {
                        // The root's entry distance is within [t_min, best t], so the
                        // stack pop it replaces would always have taken it.
                        pre = bpre;
                        bbest = Best{best.t, 0.0f, 0.0f, -1, false};
                        blas_base = st.sp;
                        in_blas = true;
                        const int32_t code = __float_as_int(root.lo.w);
                        if (code < 0) {
                            int32_t first, count;
                            DecodeLeaf(code, a.leaf_count, first, count);
                            tcur = first;
                            tend = first + count;
                            tri = count > 0;
                        } else {
                            node = true;
                            src = reinterpret_cast<const float4*>(a.nodes + code);
                        }
                    } else {
                        ++cursor;
                    }
                }
            }
            if (tri) src = a.slots + 3 * static_cast<int64_t>(tcur);
            Fetch f;
            if (node || tri) f.Load(src, node);
            if (node) {
                InteriorStep(f.q0, f.q1, f.q2, f.q3, pre, t_min, in_blas ? bbest.t : best.t, nt, st);
            } else if (tri) {
                ++tt;
                ++tcur;
                if (TriangleTest(f.q0, f.q1, f.q2, bx, by, bz, bdx, bdy, bdz, t_min, kAny, bbest) && kAny) {
                    tcur = tend;
                    st.Truncate(blas_base);  // finish this BLAS now
                }
            }
            const bool lane_busy = in_blas || cursor < leaf_end || st.sp > 0;
            const uint64_t active = __ballot(lane_busy);
            if (active == 0) break;
            const uint64_t idle = __ballot(!lane_busy && (ray >= 0 || !exhausted));
            if (__popcll(idle) >= kRefillIdleLanes) break;
        }
    }
    if (st.overflow) atomicOr(a.overflow, 1u);
    if (a.totals != nullptr) AddTotals(a.totals, sum_nt, sum_tt, sum_ts, lane);
}

}  // namespace trav

// Fills the launch arguments (grid, stack spill region, counters) of one
// traversal over `r`'s queue (yrh_trace.hip).
int PrepareTraceArgs(yrh_context* ctx, const TraceRequest& r, trav::TraceArgs& a, int& grid, int& block);
// Dynamic LDS bytes of a traversal launch with `block` threads.
size_t LdsBytes(const trav::TraceArgs& a, int block);

}  // namespace yrh
