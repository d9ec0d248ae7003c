// snch_build.hpp -- host half of the closest-silhouette query: the silhouette
// objects, per-leaf silhouette references and bounding cones that FCPW attaches
// to its BVH when Scene::computeSilhouettes was called (the SNCH node type),
// and the GPU records GPUScene uploads.
//
// Follows, step for step:
//   Scene<DIM>::computeSilhouettes        include/fcpw/fcpw.inl:405-521
//   sortSoupPositions<DIM, SnchNode>      include/fcpw/fcpw.inl:709-756
//   Bvh::assignGeometricDataToNodes       include/fcpw/aggregates/bvh.inl:306-420
//   computeBoundingConesRecursive         include/fcpw/aggregates/bvh.inl:236-304
//   SilhouetteVertex / SilhouetteEdge     include/fcpw/geometry/silhouette_vertices.inl:1-60,
//                                         include/fcpw/geometry/silhouette_edges.inl:1-82
//   extractSilhouetteVertices / Edges     include/fcpw/gpu/bvh_interop_structures.h:237-300
// Arithmetic follows the Eigen expressions of those files (fixed-size reductions
// c0 + (c1 + c2); normalized() leaves a zero vector unchanged), so cones and
// records are bit-identical to the reference host code compiled without FMA.
//
// Vertex numbering.  The reference renumbers vertices in leaf order before it
// builds 2D silhouette vertices, and reports that sorted id as the silhouette
// index.  This backend keeps the caller's vertex order everywhere else, so the
// 2D objects are built in the sorted numbering and read positions through
// sorted_to_user.  The reference fills a sorted vertex's fields only from the
// segments that reference it after sorting; a field no segment writes keeps
// the value computeSilhouettes stored for the unsorted vertex of the same id
// (possible only at the ends of open curves).  That is reproduced as is.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <utility>
#include <vector>

#include "../../../include/fcpw_hip.h"
#include "bvh_build.hpp"

namespace fcpw_amd {

// ignoreSilhouette(angle, primitive index): fcpw.h:78-83
using IgnoreSilhouetteFn = std::function<bool(float, int)>;

constexpr float kPiF = 3.14159265358979323846f;

template <int DIM>
struct SnchData {
    // silhouette objects: 2D one per sorted vertex id {prev, self, next, -},
    // 3D one per edge id {opposite vertex of face 1, a, b, opposite vertex of face 0}
    std::vector<std::array<int32_t, 4>> ind;
    std::vector<int32_t> pindex;          // index reported by the query
    std::vector<int32_t> sorted_to_user;  // 2D: sorted vertex id -> caller's vertex id (-1: none)
    std::vector<int32_t> ref_obj;         // per silhouette reference (leaf order): object id
    std::vector<float> cone;              // per node: axis[3], halfAngle, radius
    std::vector<int32_t> sil_offset, sil_count; // per node
};

namespace snch_detail {

inline float emin(float a, float b) { return (b < a) ? b : a; } // std::min
inline float emax(float a, float b) { return (a < b) ? b : a; } // std::max
inline float edot(const float *a, const float *b, int dim)
{
    float c[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < dim; i++) c[i] = a[i] * b[i];
    return detail::red_sum(c, dim);
}
inline float enorm(const float *a, int dim) { return std::sqrt(edot(a, a, dim)); }
inline void enormalize(float *v, int dim)
{
    float z = edot(v, v, dim);
    if (z > 0.0f) {
        float r = std::sqrt(z);
        for (int i = 0; i < dim; i++) v[i] = v[i] / r;
    }
}
inline void ecross(const float *a, const float *b, float *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// positions of the objects' vertex ids
template <int DIM>
struct PosView {
    const float *user;
    const std::vector<int32_t> *sorted_to_user; // 2D only
    void get(int32_t v, float *out) const
    {
        if (DIM == 3) {
            for (int i = 0; i < 3; i++) out[i] = user[(size_t)v * 3 + i];
            return;
        }
        int32_t u = (v >= 0 && (size_t)v < sorted_to_user->size()) ? (*sorted_to_user)[v] : -1;
        for (int i = 0; i < DIM; i++) out[i] = u >= 0 ? user[(size_t)u * DIM + i] : 0.0f;
    }
};

template <int DIM>
inline bool has_face(const std::array<int32_t, 4> &ind, int f)
{
    if (DIM == 2) return f == 0 ? ind[2] != -1 : ind[0] != -1;
    return f == 0 ? ind[3] != -1 : ind[0] != -1;
}

// SilhouetteVertex::normal(fIndex, normalize) (silhouette_vertices.inl:37-47),
// SilhouetteEdge::normal(fIndex, normalize) (silhouette_edges.inl:47-68)
template <int DIM>
inline void face_normal(const PosView<DIM> &P, const std::array<int32_t, 4> &ind, int f, bool normalize, float *n)
{
    if (DIM == 2) {
        int i = f == 0 ? 1 : 0;
        float pa[2], pb[2];
        P.get(ind[i], pa);
        P.get(ind[i + 1], pb);
        float s0 = pb[0] - pa[0], s1 = pb[1] - pa[1];
        n[0] = s1;
        n[1] = -s0;
        if (normalize) enormalize(n, 2);
        return;
    }
    int i = f == 0 ? 3 : 0, j = f == 0 ? 1 : 2, k = f == 0 ? 2 : 1;
    float pa[3], pb[3], pc[3], v1[3], v2[3];
    P.get(ind[j], pa);
    P.get(ind[k], pb);
    P.get(ind[i], pc);
    for (int c = 0; c < 3; c++) {
        v1[c] = pb[c] - pa[c];
        v2[c] = pc[c] - pa[c];
    }
    ecross(v1, v2, n);
    if (normalize) enormalize(n, 3);
}

// normal() with no precomputed soup normals: the normalized sum of the
// unnormalized adjacent face normals (silhouette_vertices.inl:49-58,
// silhouette_edges.inl:70-82)
template <int DIM>
inline void object_normal(const PosView<DIM> &P, const std::array<int32_t, 4> &ind, float *n)
{
    float t[3];
    n[0] = n[1] = n[2] = 0.0f;
    for (int f = 0; f < 2; f++) {
        if (!has_face<DIM>(ind, f)) continue;
        face_normal<DIM>(P, ind, f, false, t);
        for (int c = 0; c < DIM; c++) n[c] = n[c] + t[c];
    }
    enormalize(n, DIM);
}

template <int DIM>
inline void object_centroid(const PosView<DIM> &P, const std::array<int32_t, 4> &ind, float *c)
{
    if (DIM == 2) {
        P.get(ind[1], c);
        return;
    }
    float pa[3], pb[3];
    P.get(ind[1], pa);
    P.get(ind[2], pb);
    for (int i = 0; i < 3; i++) c[i] = (pa[i] + pb[i]) * 0.5f;
}

// assignEdgeIndices (fcpw.inl:405-426): edges numbered in order of first appearance
inline int assign_edge_indices(const int32_t *idx, int n_tris, std::vector<int32_t> &eidx)
{
    struct Key {
        int32_t a, b, pos;
    };
    int n = 3 * n_tris;
    std::vector<Key> k((size_t)n);
    for (int i = 0; i < n_tris; i++)
        for (int j = 0; j < 3; j++) {
            int32_t I = idx[3 * i + j], J = idx[3 * i + (j + 1) % 3];
            if (I > J) std::swap(I, J);
            k[(size_t)(3 * i + j)] = {I, J, 3 * i + j};
        }
    std::sort(k.begin(), k.end(), [](const Key &p, const Key &q) {
        if (p.a != q.a) return p.a < q.a;
        if (p.b != q.b) return p.b < q.b;
        return p.pos < q.pos;
    });
    std::vector<std::pair<int32_t, int32_t>> groups; // (first position, start in k)
    for (int i = 0; i < n; i++)
        if (i == 0 || k[i].a != k[i - 1].a || k[i].b != k[i - 1].b) groups.push_back({k[i].pos, i});
    std::sort(groups.begin(), groups.end());
    eidx.assign((size_t)n, -1);
    for (size_t g = 0; g < groups.size(); g++) {
        int i = groups[g].second;
        int32_t a = k[i].a, b = k[i].b;
        for (; i < n && k[i].a == a && k[i].b == b; i++) eidx[(size_t)k[i].pos] = (int32_t)g;
    }
    return (int)groups.size();
}

// subtree end of node i in the flat DFS array
template <int DIM>
inline void subtree_ends(const std::vector<BvhNode<DIM>> &nodes, std::vector<int32_t> &end)
{
    end.assign(nodes.size(), 0);
    for (int i = (int)nodes.size() - 1; i >= 0; i--)
        end[i] = nodes[i].nrefs > 0 ? i + 1 : end[i + nodes[i].offset];
}

} // namespace snch_detail

// Builds the silhouette objects, references and cones for a built BVH.
// positions / indices are the caller's arrays (n_vertices x DIM, n_prims x DIM).
template <int DIM>
SnchData<DIM> build_snch(const FlatBvh<DIM> &bvh, const float *positions, int n_vertices, const int32_t *indices,
                         int n_prims, const IgnoreSilhouetteFn &ignore)
{
    using namespace snch_detail;
    SnchData<DIM> S;
    const auto &nodes = bvh.nodes;
    const auto &order = bvh.prim_order;
    std::vector<int32_t> eidx;
    PosView<DIM> P{positions, &S.sorted_to_user};
    const std::array<int32_t, 4> none = {-1, -1, -1, -1};

    if (DIM == 2) {
        // leaf-order vertex numbering (sortLineSegmentSoupPositions, fcpw.inl:613-650)
        std::vector<int32_t> user_to_sorted((size_t)n_vertices, -1);
        S.sorted_to_user.assign((size_t)n_vertices, -1);
        int next = 0;
        for (int s = 0; s < n_prims; s++)
            for (int k = 0; k < 2; k++) {
                int32_t v = indices[(size_t)order[s] * 2 + k];
                if (user_to_sorted[v] == -1) {
                    user_to_sorted[v] = next;
                    S.sorted_to_user[next] = v;
                    next++;
                }
            }
        // Scene<2>::computeSilhouettes on the caller's ids, then the re-assignment of
        // sortSoupPositions<2, SnchNode> on sorted ids, in leaf-slot order
        S.ind.assign((size_t)n_vertices, none);
        S.pindex.assign((size_t)n_vertices, -1);
        auto assign = [&](int32_t a, int32_t b) {
            S.ind[a][1] = a; S.ind[a][2] = b; S.pindex[a] = a;
            S.ind[b][0] = a; S.ind[b][1] = b; S.pindex[b] = b;
        };
        for (int i = 0; i < n_prims; i++) assign(indices[2 * i], indices[2 * i + 1]);
        for (int s = 0; s < n_prims; s++)
            assign(user_to_sorted[indices[(size_t)order[s] * 2]], user_to_sorted[indices[(size_t)order[s] * 2 + 1]]);
    } else {
        // Scene<3>::computeSilhouettes (fcpw.inl:476-521); the reference then renames the
        // vertex ids into its sorted numbering, which leaves every position unchanged
        int E = assign_edge_indices(indices, n_prims, eidx);
        S.ind.assign((size_t)E, none);
        S.pindex.assign((size_t)E, -1);
        for (int i = 0; i < n_prims; i++) {
            const int32_t *t = indices + (size_t)i * 3;
            for (int j = 0; j < 3; j++) {
                int I = j == 0 ? 2 : j - 1, J = j, K = j == 2 ? 0 : j + 1;
                int e = eidx[(size_t)(3 * i + j)];
                bool forward = true;
                if (t[J] > t[K]) {
                    std::swap(J, K);
                    forward = false;
                }
                S.ind[e][forward ? 0 : 3] = t[I];
                S.ind[e][1] = t[J];
                S.ind[e][2] = t[K];
                S.pindex[e] = e;
            }
        }
    }

    // per-leaf references, deduplicated within a leaf (bvh.inl:313-420)
    std::vector<float> ref_n, ref_fn; // normal() and the two unit face normals per reference
    S.sil_offset.assign(nodes.size(), 0);
    S.sil_count.assign(nodes.size(), 0);
    std::vector<int32_t> seen;
    std::vector<int32_t> user_to_sorted_2d;
    if (DIM == 2) {
        user_to_sorted_2d.assign((size_t)n_vertices, -1);
        for (size_t s = 0; s < S.sorted_to_user.size(); s++)
            if (S.sorted_to_user[s] >= 0) user_to_sorted_2d[S.sorted_to_user[s]] = (int32_t)s;
    }
    for (size_t i = 0; i < nodes.size(); i++) {
        int start = (int)S.ref_obj.size();
        seen.clear();
        for (int j = 0; j < nodes[i].nrefs; j++) {
            int slot = nodes[i].offset + j;
            int prim = order[slot];
            for (int k = 0; k < DIM; k++) {
                int32_t o = DIM == 2 ? user_to_sorted_2d[indices[(size_t)prim * 2 + k]] : eidx[(size_t)prim * 3 + k];
                if (std::find(seen.begin(), seen.end(), o) != seen.end()) continue;
                seen.push_back(o);
                const auto &ind = S.ind[o];
                bool f0 = has_face<DIM>(ind, 0), f1 = has_face<DIM>(ind, 1);
                float n0[3] = {0.0f, 0.0f, 0.0f}, n1[3] = {0.0f, 0.0f, 0.0f}, n[3];
                if (f0) face_normal<DIM>(P, ind, 0, true, n0);
                if (f1) face_normal<DIM>(P, ind, 1, true, n1);
                object_normal<DIM>(P, ind, n);
                if (f0 && f1 && ignore) {
                    float angle;
                    if (DIM == 2) {
                        angle = n0[0] * n1[1] - n0[1] * n1[0];
                    } else {
                        float pa[3], pb[3], ed[3], cr[3];
                        P.get(ind[1], pa);
                        P.get(ind[2], pb);
                        for (int c = 0; c < 3; c++) ed[c] = pb[c] - pa[c];
                        enormalize(ed, 3);
                        ecross(n0, n1, cr);
                        angle = std::atan2(edot(ed, cr, 3), edot(n0, n1, 3));
                    }
                    if (ignore(angle, prim)) continue;
                }
                S.ref_obj.push_back(o);
                ref_n.insert(ref_n.end(), n, n + DIM);
                ref_fn.insert(ref_fn.end(), n0, n0 + DIM);
                ref_fn.insert(ref_fn.end(), n1, n1 + DIM);
            }
        }
        S.sil_offset[i] = start;
        S.sil_count[i] = (int)S.ref_obj.size() - start;
    }

    // cones: the SnchNode default (axis 0, halfAngle pi, radius 0), then
    // computeBoundingConesRecursive for every node over its whole subtree
    S.cone.assign(nodes.size() * 5, 0.0f);
    for (size_t i = 0; i < nodes.size(); i++) S.cone[i * 5 + 3] = kPiF;
    std::vector<int32_t> end;
    subtree_ends<DIM>(nodes, end);
    parallel_rows((int64_t)nodes.size(), [&](int64_t s0, int64_t s1) {
        for (int64_t start = s0; start < s1; start++) {
            const auto &node = nodes[start];
            float centroid[3], axis[3] = {0.0f, 0.0f, 0.0f}, radius = 0.0f;
            for (int c = 0; c < DIM; c++) centroid[c] = (node.box.lo[c] + node.box.hi[c]) * 0.5f;
            bool any = false, two = true;
            for (int32_t i = (int32_t)start; i < end[start]; i++)
                for (int j = 0; j < S.sil_count[i]; j++) {
                    int r = S.sil_offset[i] + j;
                    const auto &ind = S.ind[S.ref_obj[r]];
                    for (int c = 0; c < DIM; c++) axis[c] = axis[c] + ref_n[(size_t)r * DIM + c];
                    float sc[3], dv[3];
                    object_centroid<DIM>(P, ind, sc);
                    for (int c = 0; c < DIM; c++) dv[c] = sc[c] - centroid[c];
                    radius = emax(radius, enorm(dv, DIM));
                    two = two && has_face<DIM>(ind, 0) && has_face<DIM>(ind, 1);
                    any = true;
                }
            float *cone = &S.cone[(size_t)start * 5];
            if (!any) {
                cone[3] = -kPiF;
            } else if (!two) {
                cone[3] = kPiF;
            } else {
                float axis_norm = enorm(axis, DIM);
                if (axis_norm > detail::kEps) {
                    for (int c = 0; c < DIM; c++) axis[c] = axis[c] / axis_norm;
                    float half = 0.0f;
                    for (int32_t i = (int32_t)start; i < end[start]; i++)
                        for (int j = 0; j < S.sil_count[i]; j++) {
                            int r = S.sil_offset[i] + j;
                            for (int f = 0; f < 2; f++) {
                                const float *nf = &ref_fn[((size_t)r * 2 + f) * DIM];
                                half = emax(half, std::acos(emax(-1.0f, emin(1.0f, edot(axis, nf, DIM)))));
                            }
                        }
                    cone[0] = axis[0];
                    cone[1] = axis[1];
                    cone[2] = DIM == 3 ? axis[2] : 0.0f;
                    cone[3] = half;
                    cone[4] = radius;
                }
            }
        }
    });
    return S;
}

// Scene::refit of the cones (refitRecursive, bvh.inl:468-581): after the boxes
// of bvh were refit, each leaf recomputes its cone from its own references
// (computeBoundingCone, :483-531) and each internal node merges its children's
// (mergeBoundingCones, bounding_volumes.h:334-387, with rotate, core.h:107-149).
template <int DIM>
void refit_snch_cones(SnchData<DIM> &S, const FlatBvh<DIM> &bvh, const float *positions)
{
    using namespace snch_detail;
    PosView<DIM> P{positions, &S.sorted_to_user};
    const auto &nodes = bvh.nodes;
    auto centroid = [&](int i, float *c) {
        for (int k = 0; k < DIM; k++) c[k] = (nodes[i].box.lo[k] + nodes[i].box.hi[k]) * 0.5f;
    };
    // rotate<DIM>(u, v, theta)
    auto rotate = [](const float *u, const float *v, float theta, float *out) {
        if (DIM == 2) {
            float det = u[0] * v[1] - u[1] * v[0];
            theta *= std::copysign(1.0f, det);
            float c = std::cos(theta), s = std::sin(theta);
            out[0] = c * u[0] - s * u[1];
            out[1] = s * u[0] + c * u[1];
            return;
        }
        float c = std::cos(theta), s = std::sin(theta), w[3];
        ecross(u, v, w);
        enormalize(w, 3);
        float o[3] = {(1.0f - c) * w[0], (1.0f - c) * w[1], (1.0f - c) * w[2]};
        float R[3][3] = {{c + o[0] * w[0], o[1] * w[0] - s * w[2], o[2] * w[0] + s * w[1]},
                         {o[0] * w[1] + s * w[2], c + o[1] * w[1], o[2] * w[1] - s * w[0]},
                         {o[0] * w[2] - s * w[1], o[1] * w[2] + s * w[0], c + o[2] * w[2]}};
        for (int i = 0; i < 3; i++) {
            float t[3] = {R[i][0] * u[0], R[i][1] * u[1], R[i][2] * u[2]};
            out[i] = detail::red_sum(t, 3);
        }
    };
    for (int i = (int)nodes.size() - 1; i >= 0; i--) {
        float *C = &S.cone[(size_t)i * 5];
        float axis[3] = {0.0f, 0.0f, 0.0f}, half = kPiF, radius = 0.0f;
        if (nodes[i].nrefs == 0) {
            int l = i + 1, r = i + nodes[i].offset;
            const float *A = &S.cone[(size_t)l * 5], *B = &S.cone[(size_t)r * 5];
            if (A[3] >= 0.0f && B[3] >= 0.0f) {
                float axA[3] = {A[0], A[1], A[2]}, axB[3] = {B[0], B[1], B[2]}, hA = A[3], hB = B[3];
                float oA[3], oB[3], oN[3], dA[3], dB[3];
                centroid(l, oA);
                centroid(r, oB);
                centroid(i, oN);
                for (int k = 0; k < DIM; k++) { dA[k] = oN[k] - oA[k]; dB[k] = oN[k] - oB[k]; }
                radius = std::sqrt(emax(A[4] * A[4] + edot(dA, dA, DIM), B[4] * B[4] + edot(dB, dB, DIM)));
                if (hB > hA) {
                    std::swap(axA, axB);
                    std::swap(hA, hB);
                }
                float theta = std::acos(emax(-1.0f, emin(1.0f, edot(axA, axB, DIM))));
                if (emin(theta + hB, kPiF) <= hA) {
                    std::memcpy(axis, axA, sizeof(axis));
                    half = hA;
                } else {
                    float o_theta = (hA + theta + hB) / 2.0f;
                    if (o_theta >= kPiF) {
                        std::memcpy(axis, axA, sizeof(axis));
                    } else {
                        rotate(axA, axB, o_theta - hA, axis);
                        half = o_theta;
                    }
                }
            } else if (A[3] >= 0.0f) {
                std::memcpy(axis, A, 3 * sizeof(float)); half = A[3]; radius = A[4];
            } else if (B[3] >= 0.0f) {
                std::memcpy(axis, B, 3 * sizeof(float)); half = B[3]; radius = B[4];
            } else {
                half = -kPiF;
            }
        } else {
            float c[3];
            centroid(i, c);
            bool any = false, two = true;
            for (int j = 0; j < S.sil_count[i]; j++) {
                const auto &ind = S.ind[S.ref_obj[S.sil_offset[i] + j]];
                float n[3], sc[3], dv[3];
                object_normal<DIM>(P, ind, n);
                for (int k = 0; k < DIM; k++) axis[k] = axis[k] + n[k];
                object_centroid<DIM>(P, ind, sc);
                for (int k = 0; k < DIM; k++) dv[k] = sc[k] - c[k];
                radius = emax(radius, enorm(dv, DIM));
                two = two && has_face<DIM>(ind, 0) && has_face<DIM>(ind, 1);
                any = true;
            }
            if (!any) {
                half = -kPiF;
            } else if (!two) {
                half = kPiF;
            } else {
                float an = enorm(axis, DIM);
                if (an > detail::kEps) {
                    for (int k = 0; k < DIM; k++) axis[k] = axis[k] / an;
                    half = 0.0f;
                    for (int j = 0; j < S.sil_count[i]; j++) {
                        const auto &ind = S.ind[S.ref_obj[S.sil_offset[i] + j]];
                        for (int f = 0; f < 2; f++) {
                            float nf[3] = {0.0f, 0.0f, 0.0f};
                            face_normal<DIM>(P, ind, f, true, nf);
                            half = emax(half, std::acos(emax(-1.0f, emin(1.0f, edot(axis, nf, DIM)))));
                        }
                    }
                }
            }
        }
        C[0] = axis[0];
        C[1] = axis[1];
        C[2] = DIM == 3 ? axis[2] : 0.0f;
        C[3] = half;
        C[4] = radius;
    }
}

// GPU records (extractSilhouetteVertices / extractSilhouetteEdges) from the
// given positions (the caller's order): fcpw_hip_vertex in 2D, fcpw_hip_edge in 3D
template <int DIM>
std::vector<uint8_t> snch_records(const SnchData<DIM> &S, const float *positions)
{
    using namespace snch_detail;
    PosView<DIM> P{positions, &S.sorted_to_user};
    size_t rs = DIM == 2 ? sizeof(fcpw_hip_vertex) : sizeof(fcpw_hip_edge);
    std::vector<uint8_t> out(S.ref_obj.size() * rs);
    for (size_t r = 0; r < S.ref_obj.size(); r++) {
        int o = S.ref_obj[r];
        const auto &ind = S.ind[o];
        bool f0 = has_face<DIM>(ind, 0), f1 = has_face<DIM>(ind, 1);
        float n0[3] = {0.0f, 0.0f, 0.0f}, n1[3] = {0.0f, 0.0f, 0.0f}, pa[3] = {0.0f, 0.0f, 0.0f},
              pb[3] = {0.0f, 0.0f, 0.0f};
        if (f0) face_normal<DIM>(P, ind, 0, true, n0);
        if (f1) face_normal<DIM>(P, ind, 1, true, n1);
        P.get(ind[1], pa);
        if (DIM == 3) P.get(ind[2], pb);
        uint32_t index = (uint32_t)S.pindex[o], one = (f0 && f1) ? 0u : 1u;
        if (DIM == 2) {
            fcpw_hip_vertex v = {{pa[0], pa[1], 0.0f}, {n0[0], n0[1], 0.0f}, {n1[0], n1[1], 0.0f}, index, one};
            std::memcpy(&out[r * rs], &v, rs);
        } else {
            fcpw_hip_edge e = {{pa[0], pa[1], pa[2]}, {pb[0], pb[1], pb[2]}, {n0[0], n0[1], n0[2]},
                               {n1[0], n1[1], n1[2]}, index, one};
            std::memcpy(&out[r * rs], &e, rs);
        }
    }
    return out;
}

} // namespace fcpw_amd
