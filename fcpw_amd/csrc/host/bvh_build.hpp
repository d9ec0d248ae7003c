// bvh_build.hpp -- binned SAH BVH construction on the host (the CPU half of
// the reference's GPU path: GPUScene::transferToGPU consumes a BVH built by
// Scene::build, README.md:84-86).
//
// The tree is the one the reference builds (include/fcpw/aggregates/bvh.inl:3-234
// with fcpw.inl:796-867: binned split over DIM x 8 buckets, flat axes skipped,
// centroid partition with a midpoint fallback, leaf at <= 4 references or at
// depth 62, flat DFS layout with left = i + 1 and right = i + offset).  Floating
// point follows the reference's Eigen expressions (a fixed 3-vector reduces as
// c0 op (c1 op c2)), so node boxes and split choices come out bit-identical.
//
// Construction is parallel over subtrees: child offsets are relative, so a
// subtree built into its own array can be appended after its left sibling
// unchanged; leaves record absolute primitive ranges, which the in-place
// partition fixes before the subtrees split.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <future>
#include <limits>
#include <stdexcept>
#include <vector>

namespace fcpw_amd {

enum class CostHeuristic { LongestAxisCenter, SurfaceArea, OverlapSurfaceArea, Volume, OverlapVolume };

constexpr int kBvhMaxDepth = 64; // FCPW_BVH_MAX_DEPTH (bvh.h:5)

template <int DIM>
struct Box {
    float lo[DIM], hi[DIM];
};

template <int DIM>
struct BvhNode {
    Box<DIM> box;
    int32_t offset; // first leaf slot (leaf) or right-child offset (internal)
    int32_t nrefs;  // 0 for internal nodes
};

template <int DIM>
struct FlatBvh {
    std::vector<BvhNode<DIM>> nodes;
    std::vector<int32_t> prim_order; // leaf slot -> original primitive index
    int n_leaves = 0;
    int max_depth = 0;
};

namespace detail {

constexpr float kMax = std::numeric_limits<float>::max();
constexpr float kLow = std::numeric_limits<float>::lowest();
constexpr float kEps = std::numeric_limits<float>::epsilon();

inline float mn(float a, float b) { return (b < a) ? b : a; } // std::min
inline float mx(float a, float b) { return (a < b) ? b : a; } // std::max

template <int DIM>
inline Box<DIM> empty_box()
{
    Box<DIM> b;
    for (int i = 0; i < DIM; i++) { b.lo[i] = kMax; b.hi[i] = kLow; }
    return b;
}
template <int DIM>
inline void grow(Box<DIM> &b, const Box<DIM> &o)
{
    for (int i = 0; i < DIM; i++) { b.lo[i] = mn(b.lo[i], o.lo[i]); b.hi[i] = mx(b.hi[i], o.hi[i]); }
}
// expandToInclude(point) inflates by FLT_EPSILON (bounding_volumes.h:45-49)
template <int DIM>
inline void grow_point(Box<DIM> &b, const float *p)
{
    for (int i = 0; i < DIM; i++) { b.lo[i] = mn(b.lo[i], p[i] - kEps); b.hi[i] = mx(b.hi[i], p[i] + kEps); }
}
// Eigen reduction order for fixed sizes 2 and 3
inline float red_sum(const float *c, int dim) { return dim == 3 ? c[0] + (c[1] + c[2]) : c[0] + c[1]; }
inline float red_prod(const float *c, int dim) { return dim == 3 ? c[0] * (c[1] * c[2]) : c[0] * c[1]; }

template <int DIM>
inline float surface_area(const Box<DIM> &b)
{
    float e[DIM], q[DIM];
    for (int i = 0; i < DIM; i++) e[i] = mx(b.hi[i] - b.lo[i], 1e-5f);
    float P = red_prod(e, DIM);
    for (int i = 0; i < DIM; i++) q[i] = P / e[i];
    return 2.0f * red_sum(q, DIM);
}
template <int DIM>
inline float volume(const Box<DIM> &b)
{
    float e[DIM];
    for (int i = 0; i < DIM; i++) e[i] = b.hi[i] - b.lo[i];
    return red_prod(e, DIM);
}
template <int DIM>
inline bool valid(const Box<DIM> &b)
{
    for (int i = 0; i < DIM; i++)
        if (!(b.hi[i] >= b.lo[i])) return false;
    return true;
}
template <int DIM>
inline Box<DIM> overlap_box(const Box<DIM> &a, const Box<DIM> &b)
{
    Box<DIM> r;
    for (int i = 0; i < DIM; i++) { r.lo[i] = mx(a.lo[i], b.lo[i]); r.hi[i] = mn(a.hi[i], b.hi[i]); }
    return r;
}

template <int DIM>
struct Ref {
    Box<DIM> box;
    float c[DIM];
    int32_t prim;
};

template <int DIM>
class Builder {
public:
    Builder(std::vector<Ref<DIM>> &refs, CostHeuristic h, int leaf_size, int n_buckets, int parallel_cutoff)
        : refs_(refs), h_(h), leaf_size_(leaf_size), nb_(n_buckets), cutoff_(parallel_cutoff) {}

    // builds the subtree over refs_[start, end) into out (DFS order, root first);
    // returns {max depth, leaf count}
    std::pair<int, int> build(std::vector<BvhNode<DIM>> &out, int start, int end, int depth)
    {
        Box<DIM> bb = empty_box<DIM>(), bc = empty_box<DIM>();
        for (int p = start; p < end; p++) {
            grow(bb, refs_[p].box);
            grow_point(bc, refs_[p].c);
        }
        int n = end - start;
        int idx = (int)out.size();
        BvhNode<DIM> node;
        node.box = bb;
        if (n <= leaf_size_ || depth == kBvhMaxDepth - 2) {
            node.offset = start;
            node.nrefs = n;
            out.push_back(node);
            return {depth, 1};
        }
        node.offset = 0;
        node.nrefs = 0;
        out.push_back(node);

        int split_dim;
        float split_coord;
        split(bb, bc, start, end, split_dim, split_coord);
        int mid = partition(start, end, split_dim, split_coord);

        std::pair<int, int> l, r;
        if (n >= cutoff_ && depth < 6) {
            std::vector<BvhNode<DIM>> right;
            right.reserve((size_t)(end - mid) * 2);
            Builder<DIM> rb(refs_, h_, leaf_size_, nb_, cutoff_);
            auto fut = std::async(std::launch::async, [&]() { return rb.build(right, mid, end, depth + 1); });
            l = build(out, start, mid, depth + 1);
            r = fut.get();
            out[idx].offset = (int32_t)(out.size() - idx);
            out.insert(out.end(), right.begin(), right.end());
        } else {
            l = build(out, start, mid, depth + 1);
            out[idx].offset = (int32_t)(out.size() - idx);
            r = build(out, mid, end, depth + 1);
        }
        return {std::max(l.first, r.first), l.second + r.second};
    }

private:
    float cost(const Box<DIM> &bl, const Box<DIM> &br, int nl, int nr) const
    {
        switch (h_) {
        case CostHeuristic::SurfaceArea:
            return (float)nl * surface_area(bl) + (float)nr * surface_area(br);
        case CostHeuristic::OverlapSurfaceArea: {
            Box<DIM> bi = overlap_box(bl, br);
            float c = ((float)nl / surface_area(br) + (float)nr / surface_area(bl)) * std::fabs(surface_area(bi));
            return valid(bi) ? c : -c;
        }
        case CostHeuristic::Volume:
            return (float)nl * volume(bl) + (float)nr * volume(br);
        case CostHeuristic::OverlapVolume: {
            Box<DIM> bi = overlap_box(bl, br);
            float c = ((float)nl / volume(br) + (float)nr / volume(bl)) * std::fabs(volume(bi));
            return valid(bi) ? c : -c;
        }
        default:
            return kMax;
        }
    }

    // computeObjectSplit (bvh.inl:39-111)
    void split(const Box<DIM> &bb, const Box<DIM> &bc, int start, int end, int &split_dim, float &split_coord)
    {
        float best = kMax;
        split_dim = -1;
        split_coord = 0.0f;
        if (h_ != CostHeuristic::LongestAxisCenter) {
            std::vector<Box<DIM>> bbox(nb_), rbox(nb_);
            std::vector<int> bcnt(nb_), rcnt(nb_);
            for (int d = 0; d < DIM; d++) {
                float extent = bb.hi[d] - bb.lo[d];
                if (extent < 1e-6f) continue;
                float width = extent / nb_;
                for (int k = 0; k < nb_; k++) { bbox[k] = empty_box<DIM>(); bcnt[k] = 0; }
                for (int p = start; p < end; p++) {
                    int k = (int)((refs_[p].c[d] - bb.lo[d]) / width);
                    k = std::clamp(k, 0, nb_ - 1);
                    grow(bbox[k], refs_[p].box);
                    bcnt[k] += 1;
                }
                Box<DIM> acc = empty_box<DIM>();
                for (int k = nb_ - 1; k > 0; k--) {
                    grow(acc, bbox[k]);
                    rbox[k] = acc;
                    rcnt[k] = bcnt[k] + (k != nb_ - 1 ? rcnt[k + 1] : 0);
                }
                Box<DIM> left = empty_box<DIM>();
                int nl = 0;
                for (int k = 1; k < nb_; k++) {
                    grow(left, bbox[k - 1]);
                    nl += bcnt[k - 1];
                    if (nl > 0 && rcnt[k] > 0) {
                        float c = cost(left, rbox[k], nl, rcnt[k]);
                        if (c < best) {
                            best = c;
                            split_dim = d;
                            split_coord = bb.lo[d] + k * width;
                        }
                    }
                }
            }
        }
        if (split_dim == -1) {
            int md = 0;
            float me = bc.hi[0] - bc.lo[0];
            for (int d = 1; d < DIM; d++) {
                float e = bc.hi[d] - bc.lo[d];
                if (e > me) { me = e; md = d; }
            }
            split_dim = md;
            split_coord = (bc.lo[md] + bc.hi[md]) * 0.5f;
        }
    }

    // performObjectSplit (bvh.inl:113-140): the same in-place swap sequence, so
    // leaves list their primitives in the reference order
    int partition(int start, int end, int d, float coord)
    {
        int mid = start;
        for (int i = start; i < end; i++) {
            if (refs_[i].c[d] < coord) {
                std::swap(refs_[i], refs_[mid]);
                mid++;
            }
        }
        if (mid == start || mid == end) mid = start + (end - start) / 2;
        return mid;
    }

    std::vector<Ref<DIM>> &refs_;
    CostHeuristic h_;
    int leaf_size_, nb_, cutoff_;
};

} // namespace detail

// the tree over references with precomputed boxes and centroids (primitives,
// or the instances of a scene of several objects); refs is permuted into leaf
// order, and prim_order lists each leaf slot's ref.prim
template <int DIM>
FlatBvh<DIM> build_bvh_refs(std::vector<detail::Ref<DIM>> &refs, CostHeuristic heuristic, int leaf_size = 4,
                            int n_buckets = 8, int parallel_cutoff = 65536)
{
    using namespace detail;
    const int n = (int)refs.size();
    if (n <= 0) throw std::invalid_argument("Cannot construct a BVH without primitives");
    FlatBvh<DIM> out;
    out.nodes.reserve((size_t)n * 2);
    Builder<DIM> b(refs, heuristic, leaf_size, n_buckets, parallel_cutoff);
    auto res = b.build(out.nodes, 0, n, 0);
    out.max_depth = res.first;
    out.n_leaves = res.second;
    out.prim_order.resize((size_t)n);
    for (int i = 0; i < n; i++) out.prim_order[i] = refs[i].prim;
    return out;
}

// positions: n_vertices x DIM; indices: n_prims x DIM (triangles in 3D,
// segments in 2D)
template <int DIM>
FlatBvh<DIM> build_bvh(const float *positions, int n_vertices, const int32_t *indices, int n_prims,
                       CostHeuristic heuristic, int leaf_size = 4, int n_buckets = 8, int parallel_cutoff = 65536)
{
    using namespace detail;
    if (n_prims <= 0) throw std::invalid_argument("Cannot construct a BVH without primitives");
    std::vector<Ref<DIM>> refs((size_t)n_prims);
    for (int i = 0; i < n_prims; i++) {
        const float *v[DIM];
        for (int k = 0; k < DIM; k++) {
            int32_t vi = indices[(size_t)i * DIM + k];
            if (vi < 0 || vi >= n_vertices) throw std::out_of_range("primitive vertex index out of range");
            v[k] = positions + (size_t)vi * DIM;
        }
        Ref<DIM> &r = refs[i];
        r.box = empty_box<DIM>();
        for (int k = 0; k < DIM; k++) grow_point(r.box, v[k]);
        for (int d = 0; d < DIM; d++)
            r.c[d] = DIM == 3 ? ((v[0][d] + v[1][d]) + v[DIM - 1][d]) / 3.0f : (v[0][d] + v[1][d]) * 0.5f;
        r.prim = i;
    }
    return build_bvh_refs<DIM>(refs, heuristic, leaf_size, n_buckets, parallel_cutoff);
}

// Bvh::refit (bvh.inl:548-599): leaf boxes from the primitives' inflated
// boxes, internal boxes as the union of the children, bottom-up
template <int DIM>
void refit_bvh(FlatBvh<DIM> &bvh, const float *positions, const int32_t *indices)
{
    using namespace detail;
    auto &nodes = bvh.nodes;
    for (int i = (int)nodes.size() - 1; i >= 0; i--) {
        BvhNode<DIM> &n = nodes[i];
        if (n.nrefs > 0) {
            n.box = empty_box<DIM>();
            for (int p = 0; p < n.nrefs; p++) {
                int32_t prim = bvh.prim_order[n.offset + p];
                Box<DIM> pb = empty_box<DIM>();
                for (int k = 0; k < DIM; k++) grow_point(pb, positions + (size_t)indices[(size_t)prim * DIM + k] * DIM);
                grow(n.box, pb);
            }
        } else {
            n.box = nodes[i + 1].box;
            grow(n.box, nodes[i + n.offset].box);
        }
    }
}

} // namespace fcpw_amd
