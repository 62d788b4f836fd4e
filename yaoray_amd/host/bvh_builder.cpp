// SAH binned BVH builder, flat and two-level.
//
// The device traverses the reference's own tree, so this builder has to emit
// exactly the arrays yr::BuildBvh / yr::BuildTwoLevelBvh emit: same bucket
// count, cost model, partition and median-fallback calls on the same element
// sequence, same depth-first preorder node numbering. Reference:
//   primitive refs        src/accel/bvh_builder.cpp:34-140 (triangles, then spheres)
//   split choice          src/accel/bvh_split_policy.cpp:13-182
//   subtree recursion     src/accel/bvh_serial_builder.cpp:8-57,
//                         src/accel/bvh_parallel_builder.cpp:9-106 (same result)
//   leaf / merge          src/accel/bvh_builder_common.cpp:55-112
//   two-level             src/accel/two_level_bvh.cpp:25-109, 257-331
// The element order after std::partition / std::nth_element depends only on
// the comparison sequence, so sorting this file's own reference records
// reproduces the reference permutation. Parallel subtrees are built into
// private arrays and spliced in preorder, which yields the serial result.
#include "yr_ir.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <thread>

namespace yr {
namespace {

constexpr int kSahBuckets = 12;
constexpr float kSahTraversalCost = 0.5f;
constexpr float kDegenerateExtent = 1.0e-8f;

struct PrimRef {
    int flat_index = -1;
    Bounds3f bounds;
    Point3f centroid;
};

bool Finite(Point3f p) { return std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z); }
bool Finite(const Bounds3f& b) { return Finite(b.min) && Finite(b.max); }

// Argument order of every min/max matches the reference so signed zeros land
// in the same bounds.
Bounds3f Grow(const Bounds3f& acc, const Bounds3f& b) {
    return Bounds3f{Point3f{std::min(acc.min.x, b.min.x), std::min(acc.min.y, b.min.y), std::min(acc.min.z, b.min.z)},
                    Point3f{std::max(acc.max.x, b.max.x), std::max(acc.max.y, b.max.y), std::max(acc.max.z, b.max.z)}};
}
Bounds3f Grow(const Bounds3f& acc, Point3f p) {
    return Bounds3f{Point3f{std::min(acc.min.x, p.x), std::min(acc.min.y, p.y), std::min(acc.min.z, p.z)},
                    Point3f{std::max(acc.max.x, p.x), std::max(acc.max.y, p.y), std::max(acc.max.z, p.z)}};
}

float Axis(Point3f p, int axis) { return axis == 0 ? p.x : (axis == 1 ? p.y : p.z); }
Vec3f Extent(const Bounds3f& b) { return Vec3f{b.max.x - b.min.x, b.max.y - b.min.y, b.max.z - b.min.z}; }

int WidestAxis(const Bounds3f& b) {
    const Vec3f e = Extent(b);
    if (e.x >= e.y && e.x >= e.z) return 0;
    return e.y >= e.z ? 1 : 2;
}

float HalfArea2(const Bounds3f& b) {
    if (b.min.x > b.max.x || b.min.y > b.max.y || b.min.z > b.max.z) return 0.0f;
    const Vec3f e = Extent(b);
    return 2.0f * (e.x * e.y + e.y * e.z + e.z * e.x);
}

int BucketOf(float centroid, float lo, float extent) {
    const float offset = (centroid - lo) / extent;
    return std::clamp(static_cast<int>(offset * static_cast<float>(kSahBuckets)), 0, kSahBuckets - 1);
}

void MedianSplit(std::vector<PrimRef>& refs, int begin, int end, int axis, int mid) {
    std::nth_element(refs.begin() + begin, refs.begin() + mid, refs.begin() + end,
                     [axis](const PrimRef& a, const PrimRef& b) { return Axis(a.centroid, axis) < Axis(b.centroid, axis); });
}

// Returns the split position, or -1 for "make a leaf".
int ChooseSplit(std::vector<PrimRef>& refs, int begin, int end, const Bounds3f& bounds, const Bounds3f& cbounds,
                const BvhBuildOptions& opt) {
    const int count = end - begin;
    if (count <= opt.max_leaf_triangles) return -1;
    const int half = begin + count / 2;
    if (opt.split_method == BvhSplitMethod::LongestAxisMedian) {
        MedianSplit(refs, begin, end, WidestAxis(cbounds), half);
        return half;
    }
    int best_axis = -1;
    int best_split = -1;
    float best_cost = std::numeric_limits<float>::infinity();
    const float parent_area = HalfArea2(bounds);
    if (parent_area > 0.0f) {
        const Vec3f cext = Extent(cbounds);
        for (int axis = 0; axis < 3; ++axis) {
            const float extent = Axis(cext, axis);
            if (extent < kDegenerateExtent) continue;
            const float lo = Axis(cbounds.min, axis);
            int cnt[kSahBuckets] = {};
            Bounds3f box[kSahBuckets];
            for (int i = begin; i < end; ++i) {
                const PrimRef& r = refs[static_cast<size_t>(i)];
                const int b = BucketOf(Axis(r.centroid, axis), lo, extent);
                ++cnt[b];
                box[b] = Grow(box[b], r.bounds);
            }
            int lcnt[kSahBuckets], rcnt[kSahBuckets];
            Bounds3f lbox[kSahBuckets], rbox[kSahBuckets];
            lcnt[0] = cnt[0];
            lbox[0] = box[0];
            for (int i = 1; i < kSahBuckets; ++i) {
                lcnt[i] = lcnt[i - 1] + cnt[i];
                lbox[i] = Grow(lbox[i - 1], box[i]);
            }
            rcnt[kSahBuckets - 1] = cnt[kSahBuckets - 1];
            rbox[kSahBuckets - 1] = box[kSahBuckets - 1];
            for (int i = kSahBuckets - 2; i >= 0; --i) {
                rcnt[i] = rcnt[i + 1] + cnt[i];
                rbox[i] = Grow(rbox[i + 1], box[i]);
            }
            for (int s = 0; s < kSahBuckets - 1; ++s) {
                const int lc = lcnt[s];
                const int rc = rcnt[s + 1];
                if (lc == 0 || rc == 0) continue;
                const float cost = kSahTraversalCost +
                    (static_cast<float>(lc) * HalfArea2(lbox[s]) + static_cast<float>(rc) * HalfArea2(rbox[s + 1])) /
                        parent_area;
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = s;
                }
            }
        }
    }
    const bool found = best_axis >= 0;
    if (found && best_cost >= static_cast<float>(count)) return -1;
    if (!found) {
        MedianSplit(refs, begin, end, WidestAxis(cbounds), half);
        return half;
    }
    const float extent = Axis(Extent(cbounds), best_axis);
    const float lo = Axis(cbounds.min, best_axis);
    const auto pivot = std::partition(refs.begin() + begin, refs.begin() + end, [&](const PrimRef& r) {
        return BucketOf(Axis(r.centroid, best_axis), lo, extent) <= best_split;
    });
    int mid = static_cast<int>(pivot - refs.begin());
    if (mid == begin || mid == end) {
        mid = half;
        MedianSplit(refs, begin, end, best_axis, mid);
    }
    return mid;
}

struct Builder {
    std::vector<PrimRef>& refs;
    const BvhBuildOptions& opt;
    std::vector<std::string>& errors;

    bool RangeBounds(int begin, int end, Bounds3f& bounds, Bounds3f& cbounds) {
        for (int i = begin; i < end; ++i) {
            bounds = Grow(bounds, refs[static_cast<size_t>(i)].bounds);
            cbounds = Grow(cbounds, refs[static_cast<size_t>(i)].centroid);
        }
        if (!Finite(bounds) || !Finite(cbounds)) {
            errors.push_back("BVH build encountered non-finite primitive bounds");
            return false;
        }
        return true;
    }

    // Builds [begin, end) into `out` (appending) and returns the subtree root index.
    int Build(int begin, int end, int depth, int fork_budget, RenderBvh& out) {
        if (begin >= end) {
            errors.push_back("BVH build produced an empty primitive range");
            return -1;
        }
        const int node = static_cast<int>(out.nodes.size());
        out.nodes.push_back(RenderBvhNode{});
        Bounds3f bounds, cbounds;
        if (!RangeBounds(begin, end, bounds, cbounds)) return -1;
        const int mid = ChooseSplit(refs, begin, end, bounds, cbounds, opt);
        if (mid < 0) {
            const int first = static_cast<int>(out.primitive_indices.size());
            for (int i = begin; i < end; ++i) out.primitive_indices.push_back(refs[static_cast<size_t>(i)].flat_index);
            out.nodes[static_cast<size_t>(node)] = RenderBvhNode{bounds, first, end - begin};
            out.max_depth = std::max(out.max_depth, depth);
            return node;
        }
        if (mid <= begin || mid >= end) {
            errors.push_back("BVH split produced an empty child range");
            return -1;
        }
        int left = -1, right = -1;
        const bool fork = fork_budget > 1 && (end - begin) > opt.parallel_min_subtree_size;
        if (!fork) {
            left = Build(begin, mid, depth + 1, 1, out);
            right = Build(mid, end, depth + 1, 1, out);
        } else {
            // Disjoint ranges of `refs`; each side builds privately, then both
            // are spliced in preorder (left first) exactly as a serial build.
            RenderBvh lsub, rsub;
            std::vector<std::string> lerr, rerr;
            const int lbudget = (fork_budget + 1) / 2;
            const int rbudget = fork_budget - lbudget;
            Builder lb{refs, opt, lerr}, rb{refs, opt, rerr};
            std::thread lt([&] { lb.Build(begin, mid, depth + 1, lbudget, lsub); });
            std::thread rt([&] { rb.Build(mid, end, depth + 1, rbudget, rsub); });
            lt.join();
            rt.join();
            errors.insert(errors.end(), lerr.begin(), lerr.end());
            errors.insert(errors.end(), rerr.begin(), rerr.end());
            if (!lerr.empty() || !rerr.empty() || lsub.nodes.empty() || rsub.nodes.empty()) return -1;
            left = Splice(lsub, out);
            right = Splice(rsub, out);
        }
        if (left < 0 || right < 0) return -1;
        out.nodes[static_cast<size_t>(node)] = RenderBvhNode{bounds, right, 0};
        out.max_depth = std::max(out.max_depth, depth);
        return node;
    }

    static int Splice(const RenderBvh& sub, RenderBvh& out) {
        const int node_base = static_cast<int>(out.nodes.size());
        const int prim_base = static_cast<int>(out.primitive_indices.size());
        for (RenderBvhNode n : sub.nodes) {
            if (n.IsLeaf()) {
                n.payload_offset += prim_base;
            } else if (n.payload_offset >= 0) {
                n.payload_offset += node_base;
            }
            out.nodes.push_back(n);
        }
        out.primitive_indices.insert(out.primitive_indices.end(), sub.primitive_indices.begin(),
                                     sub.primitive_indices.end());
        out.max_depth = std::max(out.max_depth, sub.max_depth);
        return node_base;
    }
};

bool ValidOptions(const BvhBuildOptions& o, std::vector<std::string>& errors) {
    if (o.max_leaf_triangles < 1) errors.push_back("BVH max_leaf_triangles must be at least 1");
    if (o.thread_count < 0) errors.push_back("BVH thread_count must be >= 0");
    if (o.parallel_min_subtree_size < 1) errors.push_back("BVH parallel_min_subtree_size must be >= 1");
    return errors.empty();
}

int Threads(int requested) {
    if (requested > 0) return requested;
    const int hw = static_cast<int>(std::thread::hardware_concurrency());
    return hw > 0 ? hw : 1;
}

bool BuildFromRefs(std::vector<PrimRef>& refs, const BvhBuildOptions& opt, RenderBvh& bvh,
                   std::vector<std::string>& errors) {
    if (!ValidOptions(opt, errors)) return false;
    if (refs.empty()) return true;
    Builder b{refs, opt, errors};
    const int root = b.Build(0, static_cast<int>(refs.size()), 1, Threads(opt.thread_count), bvh);
    return root == 0 && errors.empty();
}

Bounds3f TransformBounds(const Mat4f& t, const Bounds3f& b) {
    Bounds3f r;
    for (int c = 0; c < 8; ++c) {
        const Point3f p{(c & 1) ? b.max.x : b.min.x, (c & 2) ? b.max.y : b.min.y, (c & 4) ? b.max.z : b.min.z};
        r = Grow(r, TransformPoint(t, p));
    }
    return r;
}

// SphereBounds (src/geometry/intersection.cpp:130-134).
Bounds3f SphereBounds(Point3f c, float r) {
    return Bounds3f{Point3f{c.x - r, c.y - r, c.z - r}, Point3f{c.x + r, c.y + r, c.z + r}};
}

bool InvertibleAffine(const Mat4f& t) {
    for (float v : t.m) {
        if (!std::isfinite(v)) return false;
    }
    const float det = t.m[0] * (t.m[5] * t.m[10] - t.m[9] * t.m[6]) - t.m[4] * (t.m[1] * t.m[10] - t.m[9] * t.m[2]) +
                      t.m[8] * (t.m[1] * t.m[6] - t.m[5] * t.m[2]);
    return std::isfinite(det) && std::fabs(det) > 1.0e-12f;
}

}  // namespace

BvhBuildResult BuildBvh(GeometryView g, const BvhBuildOptions& opt) {
    BvhBuildResult result;
    std::vector<PrimRef> refs;
    refs.reserve(g.indices.size() / 3);
    std::vector<RenderBvhPrimitive> prims;
    prims.reserve(g.indices.size() / 3);
    int flat = 0;
    for (int p = 0; p < static_cast<int>(g.primitives.size()); ++p) {
        const RenderPrimitive& mp = g.primitives[static_cast<size_t>(p)];
        const size_t first = mp.first_index;
        const size_t count = mp.index_count;
        if (first > g.indices.size() || count > g.indices.size() - first) {
            result.errors.push_back("BVH build encountered a primitive with an invalid index range");
            return result;
        }
        if (count % 3 != 0) {
            result.errors.push_back("BVH build encountered a primitive with a non-triangular index count");
            return result;
        }
        for (int tri = 0; tri < static_cast<int>(count / 3); ++tri) {
            const size_t o = first + static_cast<size_t>(tri) * 3;
            const uint32_t i0 = g.indices[o], i1 = g.indices[o + 1], i2 = g.indices[o + 2];
            if (i0 >= g.vertices.size() || i1 >= g.vertices.size() || i2 >= g.vertices.size()) {
                result.errors.push_back("BVH build encountered a triangle with an invalid vertex index");
                return result;
            }
            const Point3f p0 = g.vertices[i0].position, p1 = g.vertices[i1].position, p2 = g.vertices[i2].position;
            const Bounds3f b = Grow(Grow(Bounds3f{p0, p0}, p1), p2);
            const Point3f c = (p0 + p1 + p2) * (1.0f / 3.0f);
            if (!Finite(b) || !Finite(c)) {
                result.errors.push_back("BVH build encountered non-finite triangle data");
                return result;
            }
            refs.push_back(PrimRef{static_cast<int>(refs.size()), b, c});
            prims.push_back(RenderBvhPrimitive{BvhPrimitiveKind::Triangle, p, -1, tri, flat++});
        }
    }
    // Spheres after every triangle, centroid = centre (BuildPrimitiveRefs).
    for (int k = 0; k < static_cast<int>(g.spheres.size()); ++k) {
        const RenderSphere& sp = g.spheres[static_cast<size_t>(k)];
        const Bounds3f b = SphereBounds(sp.center, sp.radius);
        if (!(sp.radius > 0.0f) || !Finite(b) || !Finite(sp.center)) {
            result.errors.push_back("BVH build encountered invalid sphere data");
            return result;
        }
        refs.push_back(PrimRef{static_cast<int>(refs.size()), b, sp.center});
        prims.push_back(RenderBvhPrimitive{BvhPrimitiveKind::Sphere, -1, k, -1, -1});
    }
    RenderBvh& bvh = result.bvh;
    if (refs.empty()) {
        ValidOptions(opt, result.errors);
        return result;
    }
    bvh.primitives = std::move(prims);
    bvh.total_primitives = static_cast<int>(refs.size());
    bvh.total_spheres = static_cast<int>(g.spheres.size());
    bvh.total_triangles = bvh.total_primitives - bvh.total_spheres;
    if (!BuildFromRefs(refs, opt, bvh, result.errors)) bvh = RenderBvh{};
    return result;
}

int RenderAcceleration::NodeCount() const {
    if (kind == RenderAccelerationKind::FlatReference) return static_cast<int>(flat.nodes.size());
    int n = static_cast<int>(two_level.tlas.nodes.size());
    for (const RenderBlas& b : two_level.blases) n += static_cast<int>(b.bvh.nodes.size());
    return n;
}

int RenderAcceleration::MaxDepth() const {
    if (kind == RenderAccelerationKind::FlatReference) return flat.max_depth;
    int d = 0;
    for (const RenderBlas& b : two_level.blases) d = std::max(d, b.bvh.max_depth);
    return two_level.tlas.max_depth + d;
}

RenderAccelerationBuildResult BuildRenderAcceleration(GeometryView g, const BvhBuildOptions& opt) {
    RenderAccelerationBuildResult result;
    if (g.instances.empty()) {
        BvhBuildResult flat = BuildBvh(g, opt);
        result.errors = std::move(flat.errors);
        result.acceleration.kind = RenderAccelerationKind::FlatReference;
        result.acceleration.flat = std::move(flat.bvh);
        return result;
    }
    result.acceleration.kind = RenderAccelerationKind::TwoLevel;
    TwoLevelBvh& tl = result.acceleration.two_level;
    // One BLAS per mesh primitive with flat ids offset to global numbering.
    int flat_offset = 0;
    for (int p = 0; p < static_cast<int>(g.primitives.size()); ++p) {
        const RenderPrimitive& mp = g.primitives[static_cast<size_t>(p)];
        GeometryView one{g.vertices, g.indices, std::span<const RenderPrimitive>(&mp, 1), {}, {}};
        BvhBuildResult b = BuildBvh(one, opt);
        if (!b.errors.empty()) {
            result.errors = std::move(b.errors);
            result.acceleration.two_level = TwoLevelBvh{};
            return result;
        }
        if (b.bvh.nodes.empty()) {
            result.errors.push_back("BLAS build encountered an empty mesh primitive");
            result.acceleration.two_level = TwoLevelBvh{};
            return result;
        }
        for (RenderBvhPrimitive& t : b.bvh.primitives) {
            t.mesh_primitive = p;
            t.flat_triangle_index += flat_offset;
        }
        const Bounds3f ob = b.bvh.nodes.front().bounds;
        tl.blases.push_back(RenderBlas{p, ob, std::move(b.bvh)});
        flat_offset += static_cast<int>(mp.index_count / 3);
    }
    tl.instances.reserve(g.instances.size());
    for (const RenderInstance& src : g.instances) {
        const int p = src.primitive;
        if (p < 0 || p >= static_cast<int>(tl.blases.size())) {
            result.errors.push_back("instance references an invalid mesh primitive");
            result.acceleration.two_level = TwoLevelBvh{};
            return result;
        }
        if (!InvertibleAffine(src.object_to_world)) {
            result.errors.push_back("instance transform is singular or non-finite");
            result.acceleration.two_level = TwoLevelBvh{};
            return result;
        }
        tl.instances.push_back(PreparedMeshInstance{p, p, src.object_to_world, Inverse(src.object_to_world),
                                                    TransformBounds(src.object_to_world,
                                                                    tl.blases[static_cast<size_t>(p)].object_bounds)});
    }
    // TLAS over instance world bounds (centroid = box centre), then the
    // spheres (centroid = centre), two_level_bvh.cpp:51-109.
    std::vector<PrimRef> refs;
    refs.reserve(tl.instances.size() + g.spheres.size());
    for (int i = 0; i < static_cast<int>(tl.instances.size()); ++i) {
        const Bounds3f& wb = tl.instances[static_cast<size_t>(i)].world_bounds;
        tl.tlas.primitives.push_back(RenderTlasPrimitive{TlasPrimitiveKind::Instance, i, -1, wb});
        refs.push_back(PrimRef{i, wb, (wb.min + wb.max) * 0.5f});
    }
    for (int k = 0; k < static_cast<int>(g.spheres.size()); ++k) {
        const RenderSphere& sp = g.spheres[static_cast<size_t>(k)];
        const Bounds3f b = SphereBounds(sp.center, sp.radius);
        if (!(sp.radius > 0.0f) || !Finite(b)) {
            result.errors.push_back("TLAS build encountered invalid sphere data");
            result.acceleration.two_level = TwoLevelBvh{};
            return result;
        }
        const int flat_index = static_cast<int>(tl.tlas.primitives.size());
        tl.tlas.primitives.push_back(RenderTlasPrimitive{TlasPrimitiveKind::Sphere, -1, k, b});
        refs.push_back(PrimRef{flat_index, b, sp.center});
    }
    RenderBvh hierarchy;
    if (!BuildFromRefs(refs, opt, hierarchy, result.errors)) {
        result.acceleration.two_level = TwoLevelBvh{};
        return result;
    }
    tl.tlas.nodes = std::move(hierarchy.nodes);
    tl.tlas.primitive_indices = std::move(hierarchy.primitive_indices);
    tl.tlas.max_depth = hierarchy.max_depth;
    return result;
}

}  // namespace yr
