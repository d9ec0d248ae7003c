// geometry.hpp -- FP32 primitive routines of the MI355X query backend.
//
// Each routine restates the reference GPU arithmetic
// (include/fcpw/gpu/geometry.slang, include/fcpw/gpu/bounding-volumes.slang):
// float3 dot products are summed left to right, (x*x' + y*y') + z*z', and the
// library is compiled with -ffp-contract=off so no product is fused into an
// add.  Division and sqrt are IEEE correctly rounded (hipcc's default
// -fhip-fp32-correctly-rounded-divide-sqrt).  A 2D primitive is the 3D one
// with z = 0; adding the zero product changes no sum, so the 2D routines drop z.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace fcpw_hip {

struct v3 {
    float x, y, z;
};
struct v2 {
    float x, y;
};

__host__ __device__ __forceinline__ v3 sub(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ v2 sub(v2 a, v2 b) { return {a.x - b.x, a.y - b.y}; }
__host__ __device__ __forceinline__ float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__host__ __device__ __forceinline__ float dot(v2 a, v2 b) { return a.x * b.x + a.y * b.y; }
__host__ __device__ __forceinline__ v3 cross(v3 a, v3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// A value whose square is min(u, v)^2, for non-NaN u, v: on the device the
// median of (u, v, -FLT_MAX), one v_med3_f32 that (unlike fminf) needs no NaN
// canonicalisation of operands computed in another basic block.  It differs
// from min(u, v) only when that is -inf, where it gives -FLT_MAX, whose square
// rounds to +inf as well.
__host__ __device__ __forceinline__ float min_med3(float u, float v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(u, v, -FLT_MAX);
#else
    return fminf(u, v);
#endif
}

// BoundingBox.overlap (bounding-volumes.slang:43-53): squared distance to the
// box (0 inside) and to its farthest corner
__host__ __device__ __forceinline__ void box_sqdist(v3 lo, v3 hi, v3 c, float &d2min, float &d2max)
{
    v3 u = sub(lo, c), v = sub(c, hi);
    v3 a = {fmaxf(fmaxf(u.x, v.x), 0.0f), fmaxf(fmaxf(u.y, v.y), 0.0f), fmaxf(fmaxf(u.z, v.z), 0.0f)};
    v3 b = {min_med3(u.x, v.x), min_med3(u.y, v.y), min_med3(u.z, v.z)};
    d2min = dot(a, a);
    d2max = dot(b, b);
}
__host__ __device__ __forceinline__ void box_sqdist(v2 lo, v2 hi, v2 c, float &d2min, float &d2max)
{
    v2 u = sub(lo, c), v = sub(c, hi);
    v2 a = {fmaxf(fmaxf(u.x, v.x), 0.0f), fmaxf(fmaxf(u.y, v.y), 0.0f)};
    v2 b = {min_med3(u.x, v.x), min_med3(u.y, v.y)};
    d2min = dot(a, a);
    d2max = dot(b, b);
}

// BoundingBox.intersect (bounding-volumes.slang:67-106), slab test with the
// parallel-slab fix; a 2D ray has an unconstrained z slab, so 2D skips z
__host__ __device__ __forceinline__ bool slab1(float lo, float hi, float o, float d, float dinv, float &tn, float &tf)
{
    float t0 = (lo - o) * dinv;
    float t1 = (hi - o) * dinv;
    bool parallel = (d == 0.0f);
    bool within = (o >= lo) && (o <= hi);
    tn = parallel ? -FLT_MAX : fminf(t0, t1);
    tf = parallel ? FLT_MAX : fmaxf(t0, t1);
    return !(parallel && !within);
}
__host__ __device__ __forceinline__ bool box_ray(v3 lo, v3 hi, v3 o, v3 d, v3 dinv, float tmax, float &tmin_out)
{
    float n0, f0, n1, f1, n2, f2;
    bool ok = slab1(lo.x, hi.x, o.x, d.x, dinv.x, n0, f0);
    ok &= slab1(lo.y, hi.y, o.y, d.y, dinv.y, n1, f1);
    ok &= slab1(lo.z, hi.z, o.z, d.z, dinv.z, n2, f2);
    float tNearMax = fmaxf(0.0f, fmaxf(n0, fmaxf(n1, n2)));
    float tFarMin = fminf(tmax, fminf(f0, fminf(f1, f2)));
    tmin_out = tNearMax;
    return ok && !(tNearMax > tFarMin);
}
__host__ __device__ __forceinline__ bool box_ray(v2 lo, v2 hi, v2 o, v2 d, v2 dinv, float tmax, float &tmin_out)
{
    float n0, f0, n1, f1;
    bool ok = slab1(lo.x, hi.x, o.x, d.x, dinv.x, n0, f0);
    ok &= slab1(lo.y, hi.y, o.y, d.y, dinv.y, n1, f1);
    // the z slab of a 2D box contributes tNear = -FLT_MAX, tFar = FLT_MAX
    float tNearMax = fmaxf(0.0f, fmaxf(n0, fmaxf(n1, -FLT_MAX)));
    float tFarMin = fminf(tmax, fminf(f0, fminf(f1, FLT_MAX)));
    tmin_out = tNearMax;
    return ok && !(tNearMax > tFarMin);
}

// findClosestPointTriangle (geometry.slang:251-332): Ericson's region test;
// t = barycentric weights of (pa, pb).  The regions are tested in the
// reference's order with its expressions; the edge and face regions share one
// division (num / den with the region's operands: v = d1 / (d1 - d3),
// w = d2 / (d2 - d6), w = (d4 - d3) / ((d4 - d3) + (d5 - d6)),
// denom = 1 / (va + vb + vc)) and every region one distance, so a wave whose
// lanes fall in several regions runs each of these once.
__host__ __device__ __forceinline__ float closest_point_triangle(v3 pa, v3 pb, v3 pc, v3 x, v3 &p, float &t0, float &t1)
{
    const v3 ab = sub(pb, pa), ac = sub(pc, pa), ax = sub(x, pa);
    const float d1 = dot(ab, ax), d2 = dot(ac, ax);
    // region: 0 vertex (p set), 1 edge ab, 2 edge ac, 3 edge bc, 4 face
    int region = 0;
    float num = 1.0f, den = 1.0f, vb = 0.0f, vc = 0.0f;
    if (d1 <= 0.0f && d2 <= 0.0f) {
        t0 = 1.0f; t1 = 0.0f; p = pa;
    } else {
        const v3 bx = sub(x, pb);
        const float d3 = dot(ab, bx), d4 = dot(ac, bx);
        if (d3 >= 0.0f && d4 <= d3) {
            t0 = 0.0f; t1 = 1.0f; p = pb;
        } else {
            const v3 cx = sub(x, pc);
            const float d5 = dot(ab, cx), d6 = dot(ac, cx);
            if (d6 >= 0.0f && d5 <= d6) {
                t0 = 0.0f; t1 = 0.0f; p = pc;
            } else {
                vc = d1 * d4 - d3 * d2;
                if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
                    region = 1; num = d1; den = d1 - d3;
                } else {
                    vb = d5 * d2 - d1 * d6;
                    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
                        region = 2; num = d2; den = d2 - d6;
                    } else {
                        const float va = d3 * d6 - d5 * d4;
                        if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
                            region = 3; num = d4 - d3; den = (d4 - d3) + (d5 - d6);
                        } else {
                            region = 4; den = va + vb + vc;
                        }
                    }
                }
            }
        }
    }
    if (region) {
        const float q = num / den;
        if (region == 1) {
            t0 = 1.0f - q; t1 = q;
            p = {pa.x + ab.x * q, pa.y + ab.y * q, pa.z + ab.z * q};
        } else if (region == 2) {
            t0 = 1.0f - q; t1 = 0.0f;
            p = {pa.x + ac.x * q, pa.y + ac.y * q, pa.z + ac.z * q};
        } else if (region == 3) {
            t0 = 0.0f; t1 = 1.0f - q;
            const v3 bc = sub(pc, pb);
            p = {pb.x + bc.x * q, pb.y + bc.y * q, pb.z + bc.z * q};
        } else {
            const float v = vb * q, w = vc * q;
            t0 = 1.0f - v - w; t1 = v;
            p = {(pa.x + ab.x * v) + ac.x * w, (pa.y + ab.y * v) + ac.y * w, (pa.z + ab.z * v) + ac.z * w};
        }
    }
    const v3 r = sub(x, p);
    return sqrtf(dot(r, r));
}


// findClosestPointLineSegment (geometry.slang:79-103)
__host__ __device__ __forceinline__ float closest_point_segment(v2 pa, v2 pb, v2 x, v2 &p, float &t)
{
    v2 u = sub(pb, pa), v = sub(x, pa);
    float c1 = dot(u, v);
    if (c1 <= 0.0f) {
        t = 0.0f; p = pa;
    } else {
        float c2 = dot(u, u);
        if (c2 <= c1) {
            t = 1.0f; p = pb;
        } else {
            t = c1 / c2;
            p = {pa.x + u.x * t, pa.y + u.y * t};
        }
    }
    v2 r = sub(x, p);
    return sqrtf(dot(r, r));
}

// intersectTriangle (geometry.slang:202-249), Moller-Trumbore
__host__ __device__ __forceinline__ bool ray_triangle(v3 pa, v3 pb, v3 pc, v3 o, v3 d, float tmax, float &t_out, float &v_out,
                                             float &w_out)
{
    v3 v1 = sub(pb, pa), v2 = sub(pc, pa);
    v3 q = cross(d, v2);
    float det = dot(v1, q);
    if (fabsf(det) <= FLT_EPSILON) return false;
    float invDet = 1.0f / det;
    v3 r = sub(o, pa);
    float v = dot(r, q) * invDet;
    if (v < 0.0f || v > 1.0f) return false;
    v3 s = cross(r, v1);
    float w = dot(d, s) * invDet;
    if (w < 0.0f || v + w > 1.0f) return false;
    float t = dot(v2, s) * invDet;
    if (t >= 0.0f && t <= tmax) {
        t_out = t; v_out = v; w_out = w;
        return true;
    }
    return false;
}

// intersectLineSegment (geometry.slang:37-77)
__host__ __device__ __forceinline__ bool ray_segment(v2 pa, v2 pb, v2 o, v2 d, float tmax, float &t_out, float &s_out)
{
    v2 u = sub(pa, o), v = sub(pb, pa);
    float dv = d.x * v.y - d.y * v.x;
    if (fabsf(dv) <= FLT_EPSILON) return false;
    float ud = u.x * d.y - u.y * d.x;
    float s = ud / dv;
    if (s >= 0.0f && s <= 1.0f) {
        float t = (u.x * v.y - u.y * v.x) / dv;
        if (t >= 0.0f && t <= tmax) {
            t_out = t; s_out = s;
            return true;
        }
    }
    return false;
}

} // namespace fcpw_hip
