// instances.hpp -- the instance layer of a scene with several objects or
// instance transforms, built on the host: Scene::build's instance list
// (fcpw.inl:1066-1085), the inverse of each transform, and the BVH over the
// instances (makeAggregate over the objects' aggregates, fcpw.inl:1103-1112)
// with the instance bounds of Bvh / TransformedAggregate.
//
// An instance is bounded by its object's root box (Bvh::boundingBox,
// bvh.inl:614-618), or for a transformed one by the 2^DIM root-box corners
// mapped by t, each expanded into the box with +-FLT_EPSILON
// (BoundingBox::transform, bounding_volumes.h:183-202).  Its split centroid is
// the mean of its primitives' centroids (Bvh::centroid, bvh.inl:620-631), mapped
// by t when transformed (primitive.h:300-303).  A point is mapped as the GPU
// twin maps it, mul(t, float4(p, 1)) with each row summed left to right
// (transform.slang:18-21), so the device refit recomputes these boxes bit for bit.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "../../../include/fcpw_hip.h"
#include "bvh_build.hpp"

namespace fcpw_amd {

// transformPoint on a 3x4 row-major transform (2D: rows 0-1, columns 0, 1, 3)
template <int DIM>
inline void xf_point(const float *t, const float *p, float *o)
{
    for (int i = 0; i < DIM; i++) {
        const float *row = t + 4 * i;
        float s = row[0] * p[0] + row[1] * p[1];
        if (DIM == 3) s = s + row[2] * p[DIM - 1];
        o[i] = s + row[3];
    }
}

// (DIM+1) x (DIM+1) row-major affine matrix -> 3x4 record (2D: row 2 = e_z)
template <int DIM>
inline void transform_from_matrix(const float *m, float *t)
{
    for (int i = 0; i < 12; i++) t[i] = 0.0f;
    for (int i = 0; i < DIM; i++) {
        for (int j = 0; j < DIM; j++) t[i * 4 + j] = m[i * (DIM + 1) + j];
        t[i * 4 + 3] = m[i * (DIM + 1) + DIM];
    }
    if (DIM == 2) t[10] = 1.0f;
}

// Transform::inverse of an affine transform: the linear part inverted by its
// cofactors (adjugate times 1/det), the translation -(Linv * v)
template <int DIM>
inline void invert_transform(const float *t, float *ti)
{
    for (int i = 0; i < 12; i++) ti[i] = 0.0f;
    if (DIM == 3) {
        float cof[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                cof[i][j] = t[i1 * 4 + j1] * t[i2 * 4 + j2] - t[i1 * 4 + j2] * t[i2 * 4 + j1];
            }
        float det = (t[0] * cof[0][0] + t[1] * cof[0][1]) + t[2] * cof[0][2];
        if (!(det != 0.0f) || !std::isfinite(det)) throw std::invalid_argument("instance transform is singular");
        float inv = 1.0f / det;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) ti[i * 4 + j] = cof[j][i] * inv;
        for (int i = 0; i < 3; i++)
            ti[i * 4 + 3] = -((ti[i * 4 + 0] * t[3] + ti[i * 4 + 1] * t[7]) + ti[i * 4 + 2] * t[11]);
    } else {
        float det = t[0] * t[5] - t[1] * t[4];
        if (!(det != 0.0f) || !std::isfinite(det)) throw std::invalid_argument("instance transform is singular");
        float inv = 1.0f / det;
        ti[0] = t[5] * inv;
        ti[1] = -t[1] * inv;
        ti[4] = -t[4] * inv;
        ti[5] = t[0] * inv;
        for (int i = 0; i < 2; i++) ti[i * 4 + 3] = -(ti[i * 4 + 0] * t[3] + ti[i * 4 + 1] * t[7]);
        ti[10] = 1.0f;
    }
}

template <int DIM>
inline Box<DIM> instance_box(const Box<DIM> &root, const fcpw_hip_instance &I)
{
    using namespace detail;
    if (!I.transformed) return root;
    Box<DIM> b = empty_box<DIM>();
    for (int i = 0; i < (1 << DIM); i++) {
        float p[DIM], q[DIM];
        int temp = i;
        for (int j = 0; j < DIM; j++) {
            p[j] = (temp % 2) == 0 ? root.lo[j] : root.hi[j];
            temp /= 2;
        }
        xf_point<DIM>(I.t, p, q);
        grow_point(b, q);
    }
    return b;
}

// Bvh::centroid: the running sum of the primitive centroids in leaf order,
// divided by the count
template <int DIM>
inline void object_centroid(const FlatBvh<DIM> &b, const float *positions, const int32_t *indices, float *c)
{
    float s[DIM];
    for (int d = 0; d < DIM; d++) s[d] = 0.0f;
    const int n = (int)b.prim_order.size();
    for (int slot = 0; slot < n; slot++) {
        const int32_t *f = indices + (size_t)b.prim_order[slot] * DIM;
        for (int d = 0; d < DIM; d++) {
            float pc = DIM == 3 ? ((positions[(size_t)f[0] * DIM + d] + positions[(size_t)f[1] * DIM + d]) +
                                   positions[(size_t)f[DIM - 1] * DIM + d]) / 3.0f
                                : (positions[(size_t)f[0] * DIM + d] + positions[(size_t)f[1] * DIM + d]) * 0.5f;
            s[d] = s[d] + pc;
        }
    }
    for (int d = 0; d < DIM; d++) c[d] = s[d] / (float)n;
}

struct ObjectGeometry {
    const float *positions;
    const int32_t *indices;
};

// the BVH over the instances; prim_order[j] = instance of leaf slot j
template <int DIM>
inline FlatBvh<DIM> build_instance_tree(const std::vector<const FlatBvh<DIM> *> &objects,
                                        const std::vector<ObjectGeometry> &geometry,
                                        const std::vector<fcpw_hip_instance> &instances, CostHeuristic heuristic)
{
    using namespace detail;
    std::vector<std::vector<float>> cent(objects.size());
    std::vector<Ref<DIM>> refs(instances.size());
    for (size_t i = 0; i < instances.size(); i++) {
        const fcpw_hip_instance &I = instances[i];
        if (I.object < 0 || (size_t)I.object >= objects.size())
            throw std::out_of_range("instance object index out of range");
        const FlatBvh<DIM> &b = *objects[I.object];
        if (cent[I.object].empty()) {
            cent[I.object].resize(DIM);
            object_centroid<DIM>(b, geometry[I.object].positions, geometry[I.object].indices, cent[I.object].data());
        }
        Ref<DIM> &r = refs[i];
        r.box = instance_box<DIM>(b.nodes[0].box, I);
        if (I.transformed) xf_point<DIM>(I.t, cent[I.object].data(), r.c);
        else
            for (int d = 0; d < DIM; d++) r.c[d] = cent[I.object][d];
        r.prim = (int32_t)i;
    }
    return build_bvh_refs<DIM>(refs, heuristic);
}

// Bvh::refit over the instances once the objects are refit: leaf boxes from
// the instance boxes, internal boxes as the union of the children
template <int DIM>
inline void refit_instance_tree(FlatBvh<DIM> &tree, const std::vector<const FlatBvh<DIM> *> &objects,
                                const std::vector<fcpw_hip_instance> &instances)
{
    using namespace detail;
    auto &nodes = tree.nodes;
    for (int i = (int)nodes.size() - 1; i >= 0; i--) {
        BvhNode<DIM> &n = nodes[i];
        if (n.nrefs > 0) {
            n.box = empty_box<DIM>();
            for (int p = 0; p < n.nrefs; p++) {
                const fcpw_hip_instance &I = instances[tree.prim_order[n.offset + p]];
                grow(n.box, instance_box<DIM>(objects[I.object]->nodes[0].box, I));
            }
        } else {
            n.box = nodes[i + 1].box;
            grow(n.box, nodes[i + n.offset].box);
        }
    }
}

} // namespace fcpw_amd
