#pragma once
// device.hpp -- device code of the MI355X (gfx950) query backend: record and
// primitive access, traversal stacks, the closest-point / ray / refit / adjoint
// kernels and the query-ordering kernels.  Included once, by fcpw_hip.hip.
//
// Device layout (HBM), built on the host from the reference's flat DFS BVH:
//   records  one per internal node, in DFS order.  A record holds the boxes of
//            BOTH children plus a 32-bit link per child, so expanding a node is
//            one 64-byte (3D) / 48-byte (2D) contiguous fetch:
//              3D: float4 {loA.xyz, linkA} {hiA.xyz, 0} {loB.xyz, linkB} {hiB.xyz, 0}
//              2D: float4 {loA.xy, hiA.xy} {loB.xy, hiB.xy} {linkA, linkB, 0, 0}
//            link = internal record index, or LEAF_BIT | (count-1) << 27 | first leaf slot
//   root     the root box and link in the same slot format (2 float4)
//   prims    leaf-slot order; 3D: float4 {pa,0} {pb,0} {pc,0}; 2D: float4 {pa.xy, pb.xy}
//   prim_index / prim_vids   original primitive index and vertex ids per leaf slot
//   rec_parent, leaf_refs, counters   the bottom-up refit schedule
//
// Query kernels run one query per lane.  The traversal order is the reference
// GPU order (include/fcpw/gpu/bvh.slang:144-254,395-502): test both children,
// shrink r^2 by each child's farthest-corner distance, continue with the nearer
// child and push the farther one.  Each stack entry carries the child's link and
// its squared (or parametric) distance, so a pop is pruned without refetching
// the node.  The newest K entries of the stack live in LDS (K x 8 bytes per
// lane, lane-interleaved so a wave's accesses hit distinct banks) as a sliding
// window; older entries spill to a private array.  Depth <= 63 (leaf forced at
// depth 62, bvh.inl:170), so 64 entries always suffice.
#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../../include/fcpw_hip.h"
#include "geometry.hpp"

namespace fcpw_hip {

constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t CNT_SHIFT = 27;
constexpr uint32_t CNT_MASK = 0xFu;
constexpr uint32_t OFF_MASK = (1u << 27) - 1u;
constexpr int MAX_LEAF_PRIMS = 16;
constexpr int STACK_TOTAL = 64;

struct DevScene {
    const float4 *records;
    const float4 *root;
    const float4 *prims;
    const int32_t *prim_index;
    int *flags;                // bit 0: traversal stack overflow
    unsigned long long *stats; // traversal counters: internal nodes, primitive tests, queries, and per wave:
                               // loop iterations, leaf passes, 64-item evaluation groups
};

// one instance of a scene of several objects (fcpw_hip_instance on the device):
// rows of t and of its inverse, the object drawn and whether it is transformed
struct DevInstance {
    float4 t[3];
    float4 ti[3];
    int32_t object, transformed, pad0, pad1;
};

__device__ __forceinline__ uint32_t link_of(float f) { return __float_as_uint(f); }

// Asks the compiler to fit a kernel in 64 VGPRs, so that a SIMD can hold eight
// of its waves.  Applied where that measured faster: the cooperative ray kernel
// (from 71 VGPRs) and the cooperative silhouette kernel (from 65) fit without
// spilling registers to scratch, the batch kernel (from 79) with 32 bytes of
// spills per lane.  The instanced ray kernel (101 VGPRs) spilled 144 bytes per
// lane and ran 2.2x slower, so it goes without.
#define FCPW_EIGHT_WAVES __attribute__((amdgpu_waves_per_eu(8, 8)))

// min(a, b) of two squared distances: for floats in [+0, +inf] the bit
// patterns order as unsigned integers, so an integer min gives fminf's value
// (and a NaN b, whose pattern lies above +inf, leaves a, as fminf does)
// without the NaN canonicalisation fminf compiles to
__device__ __forceinline__ float fmin_sq(float a, float b)
{
    return __uint_as_float(min(__float_as_uint(a), __float_as_uint(b)));
}

// ---------------------------------------------------------------------------
// traversal stack: the NEWEST (up to) K entries live in LDS, lane-interleaved
// (the entry of stack index i of lane t at lds[(i mod K) * BLOCK + t], so a
// wave's accesses fall in distinct banks); older entries sit in a private
// spill array at their stack index.  Entries [stack_lo, sp) are in LDS and
// [0, stack_lo) in the spill array, with sp - stack_lo <= K.  A push onto a
// full window first moves the window's oldest entry -- the one in the LDS slot
// the push is about to write -- to the spill array; a pop below the window
// reads the spill array and leaves the window empty.  Pushes and pops at the
// top of a deep stack therefore stay in LDS, and each entry is spilled at most
// once per push.  The stack pointers stay in registers and the LDS array is
// declared in the traversal function itself, so the accesses compile to
// ds_read/ds_write rather than generic (flat) loads.  K = 0 keeps every entry
// in the spill array.
// ---------------------------------------------------------------------------
#define FCPW_STACK_DECL(BLOCK, K)                                                                              \
    __shared__ uint64_t lds_stack[(K) > 0 ? (K) * (BLOCK) : 1];                                              \
    uint32_t stack_spill[2 * (STACK_TOTAL - (K))];                                                           \
    const int stack_lane = threadIdx.x;                                                                       \
    int sp = 0;                                                                                               \
    int stack_lo = 0

// An entry packs (link, distance bits) into 64 bits.  The cold spill path moves
// the entry as two 32-bit halves: an access of a different type from the LDS
// one cannot be merged with it into a single generic (flat) access, which would
// send every LDS push and pop down the slower vector-memory path.  Spilled
// entries are ordinary cached scratch accesses.
#define FCPW_PUSH(BLOCK, K, LINK, D, ON_FULL)                                                                  \
    if (sp >= STACK_TOTAL) {                                                                                  \
        ON_FULL;                                                                                              \
    } else {                                                                                                  \
        const uint32_t l_ = (uint32_t)(LINK), d_ = __float_as_uint(D);                                        \
        if ((K) == 0) {                                                                                       \
            stack_spill[2 * sp] = l_;                                                                         \
            stack_spill[2 * sp + 1] = d_;                                                                     \
        } else {                                                                                              \
            const int slot_ = (int)((uint32_t)sp % (uint32_t)((K) > 0 ? (K) : 1)) * (BLOCK) + stack_lane;    \
            if (sp - stack_lo == (K)) {                                                                       \
                const uint64_t old_ = lds_stack[slot_];                                                       \
                stack_spill[2 * stack_lo] = (uint32_t)old_;                                                   \
                stack_spill[2 * stack_lo + 1] = (uint32_t)(old_ >> 32);                                       \
                stack_lo++;                                                                                   \
            }                                                                                                 \
            lds_stack[slot_] = (uint64_t)l_ | ((uint64_t)d_ << 32);                                           \
        }                                                                                                     \
        sp++;                                                                                                 \
    }

#define FCPW_POP(BLOCK, K, LINK, D)                                                                            \
    do {                                                                                                      \
        sp--;                                                                                                 \
        if ((K) > 0 && sp >= stack_lo) {                                                                     \
            const uint64_t e_ = lds_stack[(int)((uint32_t)sp % (uint32_t)((K) > 0 ? (K) : 1)) * (BLOCK) + stack_lane]; \
            (LINK) = (uint32_t)e_;                                                                            \
            (D) = __uint_as_float((uint32_t)(e_ >> 32));                                                      \
        } else {                                                                                              \
            (LINK) = stack_spill[2 * sp];                                                                     \
            (D) = __uint_as_float(stack_spill[2 * sp + 1]);                                                   \
            stack_lo = sp;                                                                                    \
        }                                                                                                     \
    } while (0)

// ---------------------------------------------------------------------------
// per-dimension record/primitive access
// ---------------------------------------------------------------------------
template <int DIM>
struct Geo;

template <>
struct Geo<3> {
    using vec = v3;
    static constexpr int REC = 4;
    static __device__ __forceinline__ void root(const DevScene &S, vec &lo, vec &hi, uint32_t &link)
    {
        float4 a = S.root[0], b = S.root[1];
        lo = {a.x, a.y, a.z};
        hi = {b.x, b.y, b.z};
        link = link_of(a.w);
    }
    static __device__ __forceinline__ void record(const DevScene &S, uint32_t r, vec &loA, vec &hiA, uint32_t &lA,
                                                  vec &loB, vec &hiB, uint32_t &lB)
    {
        const float4 *p = S.records + (size_t)r * REC;
        float4 a0 = p[0], a1 = p[1], b0 = p[2], b1 = p[3];
        loA = {a0.x, a0.y, a0.z};
        hiA = {a1.x, a1.y, a1.z};
        lA = link_of(a0.w);
        loB = {b0.x, b0.y, b0.z};
        hiB = {b1.x, b1.y, b1.z};
        lB = link_of(b0.w);
    }
    static __device__ __forceinline__ float closest(const DevScene &S, uint32_t slot, vec x, vec &p, float &u0,
                                                    float &u1)
    {
        const float4 *t = S.prims + (size_t)slot * 3;
        float4 a = t[0], b = t[1], c = t[2];
        return closest_point_triangle({a.x, a.y, a.z}, {b.x, b.y, b.z}, {c.x, c.y, c.z}, x, p, u0, u1);
    }
    // Triangle.getSurfaceArea (geometry.slang:356-359)
    static __device__ __forceinline__ float area(const DevScene &S, uint32_t slot)
    {
        const float4 *t = S.prims + (size_t)slot * 3;
        float4 a = t[0], b = t[1], c = t[2];
        v3 m = cross({b.x - a.x, b.y - a.y, b.z - a.z}, {c.x - a.x, c.y - a.y, c.z - a.z});
        return 0.5f * sqrtf(dot(m, m));
    }
    // Triangle.samplePoint (geometry.slang:409-423): uniform point, returns its pdf
    static __device__ __forceinline__ float sample(const DevScene &S, uint32_t slot, float r1, float r2, float uv[2],
                                                   float p[3], float n[3])
    {
        const float4 *t = S.prims + (size_t)slot * 3;
        float4 a = t[0], b = t[1], c = t[2];
        v3 m = cross({b.x - a.x, b.y - a.y, b.z - a.z}, {c.x - a.x, c.y - a.y, c.z - a.z});
        float ar = sqrtf(dot(m, m));
        float u1 = sqrtf(r1), u = 1.0f - u1, v = r2 * u1, w = (1.0f - u) - v;
        uv[0] = u;
        uv[1] = v;
        p[0] = (a.x * u + b.x * v) + c.x * w;
        p[1] = (a.y * u + b.y * v) + c.y * w;
        p[2] = (a.z * u + b.z * v) + c.z * w;
        n[0] = m.x / ar; n[1] = m.y / ar; n[2] = m.z / ar;
        return 2.0f / ar;
    }
    static __device__ __forceinline__ void normal(const DevScene &S, uint32_t slot, float n[3])
    {
        const float4 *t = S.prims + (size_t)slot * 3;
        float4 a = t[0], b = t[1], c = t[2];
        v3 m = cross({b.x - a.x, b.y - a.y, b.z - a.z}, {c.x - a.x, c.y - a.y, c.z - a.z});
        float l = sqrtf(dot(m, m));
        n[0] = m.x / l; n[1] = m.y / l; n[2] = m.z / l;
    }
    static __device__ __forceinline__ vec load_point(const float *q) { return {q[0], q[1], q[2]}; }
};

template <>
struct Geo<2> {
    using vec = v2;
    static constexpr int REC = 3;
    static __device__ __forceinline__ void root(const DevScene &S, vec &lo, vec &hi, uint32_t &link)
    {
        float4 a = S.root[0], b = S.root[1];
        lo = {a.x, a.y};
        hi = {a.z, a.w};
        link = link_of(b.x);
    }
    static __device__ __forceinline__ void record(const DevScene &S, uint32_t r, vec &loA, vec &hiA, uint32_t &lA,
                                                  vec &loB, vec &hiB, uint32_t &lB)
    {
        const float4 *p = S.records + (size_t)r * REC;
        float4 a = p[0], b = p[1], l = p[2];
        loA = {a.x, a.y};
        hiA = {a.z, a.w};
        loB = {b.x, b.y};
        hiB = {b.z, b.w};
        lA = link_of(l.x);
        lB = link_of(l.y);
    }
    static __device__ __forceinline__ float closest(const DevScene &S, uint32_t slot, vec x, vec &p, float &u0,
                                                    float &u1)
    {
        float4 s = S.prims[slot];
        u1 = 0.0f;
        return closest_point_segment({s.x, s.y}, {s.z, s.w}, x, p, u0);
    }
    // LineSegment.getSurfaceArea (geometry.slang:133-136)
    static __device__ __forceinline__ float area(const DevScene &S, uint32_t slot)
    {
        float4 s = S.prims[slot];
        v2 m = {s.z - s.x, s.w - s.y};
        return sqrtf(dot(m, m));
    }
    // LineSegment.samplePoint (geometry.slang:180-191)
    static __device__ __forceinline__ float sample(const DevScene &S, uint32_t slot, float r1, float r2, float uv[2],
                                                   float p[3], float n[3])
    {
        (void)r2;
        float4 s = S.prims[slot];
        v2 m = {s.z - s.x, s.w - s.y};
        float ar = sqrtf(dot(m, m));
        uv[0] = r1;
        uv[1] = 0.0f;
        p[0] = s.x + r1 * m.x;
        p[1] = s.y + r1 * m.y;
        p[2] = 0.0f;
        n[0] = m.y / ar; n[1] = -m.x / ar; n[2] = 0.0f;
        return 1.0f / ar;
    }

    static __device__ __forceinline__ void normal(const DevScene &S, uint32_t slot, float n[3])
    {
        float4 s = S.prims[slot];
        v2 m = {s.w - s.y, -(s.z - s.x)};
        float l = sqrtf(dot(m, m));
        n[0] = m.x / l; n[1] = m.y / l; n[2] = 0.0f;
    }
    static __device__ __forceinline__ vec load_point(const float *q) { return {q[0], q[1]}; }
};

__device__ __forceinline__ void store_p(v3 p, float o[3]) { o[0] = p.x; o[1] = p.y; o[2] = p.z; }
__device__ __forceinline__ void store_p(v2 p, float o[3]) { o[0] = p.x; o[1] = p.y; o[2] = 0.0f; }

// ---------------------------------------------------------------------------
// closest point traversal (bvh.slang:395-502 order)
// ---------------------------------------------------------------------------
struct CpResult {
    int32_t slot; // -1: not found
    float d;
    float p[3];
    float uv[2];
    uint32_t n_internal, n_prims; // records expanded and primitives tested (coop kernel)
    uint32_t n_iter, n_pass, n_group; // wave loop iterations, leaf passes, 64-item evaluation groups (wave-uniform)
};

// r2 is updated in place, as the reference's inout BoundingSphere
template <int DIM, int BLOCK, int K>
__device__ __forceinline__ void closest_point_query(const DevScene &S, typename Geo<DIM>::vec x, float &r2,
                                                    CpResult &res)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    res.slot = -1;
    res.d = FLT_MAX;
    res.p[0] = res.p[1] = res.p[2] = 0.0f;
    res.uv[0] = res.uv[1] = 0.0f;

    vec lo, hi;
    uint32_t cur;
    G::root(S, lo, hi, cur);
    float d2min, d2max;
    box_sqdist(lo, hi, x, d2min, d2max);
    if (!(d2min <= r2)) return;
    r2 = fminf(r2, d2max);
    float cur_d2 = d2min;

    FCPW_STACK_DECL(BLOCK, K);
    vec bp = {};
    while (true) {
        if (cur_d2 <= r2) {
            if (cur & LEAF_BIT) {
                uint32_t off = cur & OFF_MASK;
                uint32_t cnt = ((cur >> CNT_SHIFT) & CNT_MASK) + 1u;
                for (uint32_t k = 0; k < cnt; k++) {
                    vec p;
                    float u0, u1;
                    float d = G::closest(S, off + k, x, p, u0, u1);
                    if (d * d <= r2) {
                        r2 = fminf(r2, d * d);
                        res.slot = (int32_t)(off + k);
                        res.d = d;
                        bp = p;
                        res.uv[0] = u0;
                        res.uv[1] = u1;
                    }
                }
            } else {
                vec loA, hiA, loB, hiB;
                uint32_t lA, lB;
                G::record(S, cur, loA, hiA, lA, loB, hiB, lB);
                float a0, a1, b0, b1;
                box_sqdist(loA, hiA, x, a0, a1);
                bool oA = a0 <= r2;
                r2 = fminf(r2, a1);
                box_sqdist(loB, hiB, x, b0, b1);
                bool oB = b0 <= r2;
                r2 = fminf(r2, b1);
                if (oA && oB) {
                    uint32_t nearL = lA, farL = lB;
                    float nearD = a0, farD = b0;
                    if (b0 < a0) { nearL = lB; farL = lA; nearD = b0; farD = a0; }
                    FCPW_PUSH(BLOCK, K, farL, farD, { atomicOr(S.flags, 1); break; });
                    cur = nearL;
                    cur_d2 = nearD;
                    continue;
                } else if (oA) {
                    cur = lA; cur_d2 = a0;
                    continue;
                } else if (oB) {
                    cur = lB; cur_d2 = b0;
                    continue;
                }
            }
        }
        if (sp == 0) break;
        FCPW_POP(BLOCK, K, cur, cur_d2);
    }
    if (res.slot >= 0) store_p(bp, res.p);
}


// ---------------------------------------------------------------------------
// ray traversal (bvh.slang:144-254 order)
// ---------------------------------------------------------------------------
struct RayResult {
    int32_t slot;
    float t;
    float p[3];
    float n[3];
    float uv[2];
    uint32_t n_internal, n_prims; // records expanded and primitives tested
};

template <int DIM>
struct RayGeo;

template <>
struct RayGeo<3> {
    static __device__ __forceinline__ bool hit(const DevScene &S, uint32_t slot, v3 o, v3 d, float tmax,
                                               RayResult &r, bool fill)
    {
        const float4 *t = S.prims + (size_t)slot * 3;
        float4 a = t[0], b = t[1], c = t[2];
        float tt, v, w;
        v3 pa = {a.x, a.y, a.z}, pb = {b.x, b.y, b.z}, pc = {c.x, c.y, c.z};
        if (!ray_triangle(pa, pb, pc, o, d, tmax, tt, v, w)) return false;
        if (fill) {
            v3 v1 = sub(pb, pa), v2 = sub(pc, pa);
            r.p[0] = (pa.x + v1.x * v) + v2.x * w;
            r.p[1] = (pa.y + v1.y * v) + v2.y * w;
            r.p[2] = (pa.z + v1.z * v) + v2.z * w;
            v3 n = cross(v1, v2);
            float l = sqrtf(dot(n, n));
            r.n[0] = n.x / l; r.n[1] = n.y / l; r.n[2] = n.z / l;
            r.uv[0] = 1.0f - v - w;
            r.uv[1] = v;
            r.t = tt;
        }
        return true;
    }
};

template <>
struct RayGeo<2> {
    static __device__ __forceinline__ bool hit(const DevScene &S, uint32_t slot, v2 o, v2 d, float tmax,
                                               RayResult &r, bool fill)
    {
        float4 s4 = S.prims[slot];
        v2 pa = {s4.x, s4.y}, pb = {s4.z, s4.w};
        float tt, s;
        if (!ray_segment(pa, pb, o, d, tmax, tt, s)) return false;
        if (fill) {
            v2 v = sub(pb, pa);
            r.p[0] = pa.x + s * v.x;
            r.p[1] = pa.y + s * v.y;
            r.p[2] = 0.0f;
            v2 n = {v.y, -v.x};
            float l = sqrtf(dot(n, n));
            r.n[0] = n.x / l; r.n[1] = n.y / l; r.n[2] = 0.0f;
            r.uv[0] = s;
            r.uv[1] = 0.0f;
            r.t = tt;
        }
        return true;
    }
};

// tmax is updated in place, as the reference's inout Ray
template <int DIM, int BLOCK, int K>
__device__ __forceinline__ void ray_query(const DevScene &S, typename Geo<DIM>::vec o, typename Geo<DIM>::vec d,
                                          typename Geo<DIM>::vec dinv, float &tmax, bool occlusion,
                                          RayResult &res)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    res.slot = -1;
    res.t = FLT_MAX;
    for (int i = 0; i < 3; i++) { res.p[i] = 0.0f; res.n[i] = 0.0f; }
    res.uv[0] = res.uv[1] = 0.0f;
    res.n_internal = res.n_prims = 0;

    vec lo, hi;
    uint32_t cur;
    G::root(S, lo, hi, cur);
    float cur_t;
    if (!box_ray(lo, hi, o, d, dinv, tmax, cur_t)) return;

    FCPW_STACK_DECL(BLOCK, K);
    while (true) {
        if (!(cur_t > tmax)) {
            if (cur & LEAF_BIT) {
                uint32_t off = cur & OFF_MASK;
                uint32_t cnt = ((cur >> CNT_SHIFT) & CNT_MASK) + 1u;
                bool stop = false;
                for (uint32_t k = 0; k < cnt; k++) {
                    res.n_prims++;
                    if (RayGeo<DIM>::hit(S, off + k, o, d, tmax, res, !occlusion)) {
                        res.slot = (int32_t)(off + k);
                        if (occlusion) { stop = true; break; }
                        tmax = fminf(tmax, res.t);
                    }
                }
                if (stop) break;
            } else {
                vec loA, hiA, loB, hiB;
                uint32_t lA, lB;
                G::record(S, cur, loA, hiA, lA, loB, hiB, lB);
                res.n_internal++;
                float ta, tb;
                bool hA = box_ray(loA, hiA, o, d, dinv, tmax, ta);
                bool hB = box_ray(loB, hiB, o, d, dinv, tmax, tb);
                if (hA && hB) {
                    uint32_t nearL = lA, farL = lB;
                    float nearT = ta, farT = tb;
                    if (tb < ta) { nearL = lB; farL = lA; nearT = tb; farT = ta; }
                    FCPW_PUSH(BLOCK, K, farL, farT, { atomicOr(S.flags, 1); break; });
                    cur = nearL;
                    cur_t = nearT;
                    continue;
                } else if (hA) {
                    cur = lA; cur_t = ta;
                    continue;
                } else if (hB) {
                    cur = lB; cur_t = tb;
                    continue;
                }
            }
        }
        if (sp == 0) break;
        FCPW_POP(BLOCK, K, cur, cur_t);
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// interaction record written as five 8-byte stores (the record is 8-byte aligned)
__device__ __forceinline__ void write_interaction(float2 *o, const float p[3], const float n[3], const float uv[2],
                                                  float d, uint32_t index)
{
    o[0] = make_float2(p[0], p[1]);
    o[1] = make_float2(p[2], n[0]);
    o[2] = make_float2(n[1], n[2]);
    o[3] = make_float2(uv[0], uv[1]);
    o[4] = make_float2(d, __uint_as_float(index));
}

// Workgroups are dispatched round-robin over the XCDs (block b runs on XCD
// b mod xcds).  Renumbering blocks so that XCD k owns one contiguous range of
// tiles keeps each XCD's L2 on one region of the (Morton-ordered) queries.
// The tile is formed in 32 bits and widened once at the end: a version that
// returned a 64-bit value from two branches was compiled (ROCm 7.2 clang) with
// the high word of the pass-through branch left holding a stale quotient, so
// the blocks past the last full group of xcds computed out-of-range tiles.
// The argument packs the XCD count (bits 0-7) and a chunk shift c (bits 8-12).
// With c = 0 each XCD owns one contiguous range of tiles; with c > 0 XCD j takes
// chunks j, j + xcds, j + 2 xcds, ... of 2^c consecutive tiles, which keeps
// some L2 locality while spreading every region of the curve over all XCDs.
__device__ __forceinline__ int64_t xcd_tile(int xcd_arg)
{
    const uint32_t b = blockIdx.x, G = gridDim.x;
    const uint32_t xa = (uint32_t)xcd_arg & 0xFFu, cs = ((uint32_t)xcd_arg >> 8) & 31u;
    const uint32_t x = xa > 1 ? xa : 1u;
    if (cs == 0) {
        const uint32_t full = (G / x) * x;
        const uint32_t t = b < full ? (b % x) * (full / x) + b / x : b;
        return (int64_t)(uint64_t)t;
    }
    const uint32_t C = 1u << cs, span = x * C;
    const uint32_t full = (G / span) * span;
    const uint32_t l = b / x;
    const uint32_t t = b < full ? ((l / C) * x + b % x) * C + l % C : b;
    return (int64_t)(uint64_t)t;
}

// query input / output policies shared by the closest-point kernels
struct CpIO {
    const float *in;     // spheres (4 floats: c, r2) or points (DIM floats)
    const float *r2;     // compact input: per-query r^2 or nullptr (unbounded)
    int64_t n;
    const uint32_t *order;
    // interaction output
    float2 *out;
    int record_normals;
    // compact output (any may be nullptr)
    float *d, *p;
    int32_t *prim, *slot;
};
enum { CP_OUT_INTERACTION = 0, CP_OUT_COMPACT = 1 };

template <int DIM, int OUT>
__device__ __forceinline__ void cp_load(const CpIO &io, int64_t i, float q[3], float &r2)
{
    if (OUT == CP_OUT_INTERACTION) {
        float4 s = reinterpret_cast<const float4 *>(io.in)[i];
        q[0] = s.x; q[1] = s.y; q[2] = s.z;
        r2 = s.w;
    } else {
        const float *x = io.in + i * DIM;
        q[0] = x[0]; q[1] = x[1]; q[2] = DIM == 3 ? x[DIM - 1] : 0.0f;
        r2 = io.r2 ? io.r2[i] : INFINITY;
    }
}

template <int DIM, int OUT>
__device__ __forceinline__ void cp_store(const DevScene &S, const CpIO &io, int64_t i, const CpResult &r)
{
    if (OUT == CP_OUT_INTERACTION) {
        float nrm[3] = {0.0f, 0.0f, 0.0f};
        uint32_t index = FCPW_HIP_UINT_MAX;
        if (r.slot >= 0) {
            index = (uint32_t)S.prim_index[r.slot];
            if (io.record_normals) Geo<DIM>::normal(S, (uint32_t)r.slot, nrm);
        }
        write_interaction(io.out + i * 5, r.p, nrm, r.uv, r.d, index);
    } else {
        if (io.d) io.d[i] = r.d;
        if (io.p)
            for (int k = 0; k < DIM; k++) io.p[i * DIM + k] = r.p[k];
        if (io.prim) io.prim[i] = r.slot >= 0 ? S.prim_index[r.slot] : -1;
        if (io.slot) io.slot[i] = r.slot;
    }
}

// one query per lane, the reference loop
template <int DIM, int BLOCK, int K, int OUT>
__global__ __launch_bounds__(BLOCK) void k_closest_points(DevScene S, CpIO io, int xcds)
{
    int64_t k = xcd_tile(xcds) * BLOCK + threadIdx.x;
    if (k >= io.n) return;
    int64_t i = io.order ? (int64_t)io.order[k] : k;
    float q[3], r2;
    cp_load<DIM, OUT>(io, i, q, r2);
    CpResult r;
    closest_point_query<DIM, BLOCK, K>(S, Geo<DIM>::load_point(q), r2, r);
    cp_store<DIM, OUT>(S, io, i, r);
}

// ---------------------------------------------------------------------------
// Wave-cooperative leaf processing (one wave per workgroup).  Lanes descend
// the tree independently; a lane that reaches a leaf holds it.  Once at least
// `hold_min` lanes hold a leaf (or every lane still working does), the wave
// lays all (lane, primitive) pairs of the held leaves out in LDS and all 64
// lanes -- including lanes whose own query has finished -- evaluate up to 64
// primitive distances per pass.  Each owner then replays the reference's
// acceptance test (d*d <= r2, r2 = d*d, later equal distances win) over its
// own primitives in leaf order.  A distance does not depend on r2, so the
// replay gives the reference's result bit for bit; the closest point and uv
// of the winning primitive are recomputed once at the end.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Per-lane state of a cooperative traversal.
template <int DIM>
struct CoopLane {
    typename Geo<DIM>::vec x;
    float r2, cur_d2;
    uint32_t cur, loff, lcnt;
    int32_t best;
    bool active, live, holding;
    uint32_t n_internal, n_prims;
};

// LDS shared by the wave for its leaf passes (a lane's query point and held
// leaf are read from its registers with cross-lane permutes)
struct CoopShared {
    uint16_t item[256];   // (lane, primitive) pairs of one pass
    float d[256];         // their distances
};

// Starts lane state L on query x with squared radius r2 (valid = the lane has a query).
template <int DIM>
__device__ __forceinline__ void coop_begin(const DevScene &S, CoopShared &sh, CoopLane<DIM> &L,
                                           typename Geo<DIM>::vec x, float r2, bool valid)
{
    using G = Geo<DIM>;
    (void)sh;
    typename G::vec lo, hi;
    G::root(S, lo, hi, L.cur);
    float d2min, d2max;
    box_sqdist(lo, hi, x, d2min, d2max);
    L.x = x;
    L.active = valid && (d2min <= r2);
    // an active lane has r2 >= d2min >= +0; -0 becomes +0, which every comparison treats alike
    L.r2 = L.active ? fmin_sq(r2 == 0.0f ? 0.0f : r2, d2max) : r2;
    L.cur_d2 = L.active ? d2min : 0.0f;
    L.live = L.active;
    L.holding = false;
    L.loff = L.lcnt = 0;
    L.best = -1;
    L.n_internal = L.n_prims = 0;
}

// One traversal step of an active lane that holds no leaf: pop until a node
// survives the r^2 test, then either hold it (leaf) or expand it (internal:
// test both children, shrink r^2 by their d2Max, continue with the nearer,
// push the farther).  The stack variables are those FCPW_STACK_DECL declares.
#define FCPW_COOP_STEP(DIM, BLOCK, K, S, L, OVERFLOW)                                                        \
    do {                                                                                                      \
        using G_ = Geo<DIM>;                                                                                  \
        using vec_ = typename G_::vec;                                                                        \
        if ((L).active && !(L).holding) {                                                                     \
            if (!((L).live && (L).cur_d2 <= (L).r2)) {                                                       \
                (L).live = false;                                                                             \
                while (sp > 0) {                                                                              \
                    FCPW_POP(BLOCK, K, (L).cur, (L).cur_d2);                                                  \
                    if ((L).cur_d2 <= (L).r2) { (L).live = true; break; }                                     \
                }                                                                                             \
            }                                                                                                 \
            if (!(L).live) {                                                                                  \
                (L).active = false;                                                                           \
            } else if ((L).cur & LEAF_BIT) {                                                                  \
                (L).holding = true;                                                                           \
                (L).live = false;                                                                             \
                (L).loff = (L).cur & OFF_MASK;                                                                \
                (L).lcnt = (((L).cur >> CNT_SHIFT) & CNT_MASK) + 1u;                                          \
            } else {                                                                                          \
                vec_ loA, hiA, loB, hiB;                                                                      \
                uint32_t lA, lB;                                                                              \
                G_::record(S, (L).cur, loA, hiA, lA, loB, hiB, lB);                                           \
                (L).n_internal++;                                                                             \
                float a0, a1, b0, b1;                                                                         \
                box_sqdist(loA, hiA, (L).x, a0, a1);                                                          \
                const bool oA = a0 <= (L).r2;                                                                 \
                (L).r2 = fmin_sq((L).r2, a1);                                                                 \
                box_sqdist(loB, hiB, (L).x, b0, b1);                                                          \
                const bool oB = b0 <= (L).r2;                                                                 \
                (L).r2 = fmin_sq((L).r2, b1);                                                                 \
                /* continue with B when only B overlaps, or both do and B is strictly nearer; */            \
                /* with neither, live turns false and the next step pops */                                 \
                const bool bFirst = oB && (!oA || b0 < a0);                                                   \
                (L).live = oA || oB;                                                                          \
                (L).cur = bFirst ? lB : lA;                                                                   \
                (L).cur_d2 = bFirst ? b0 : a0;                                                                \
                if (oA && oB) {                                                                               \
                    const uint32_t farL = bFirst ? lA : lB;                                                   \
                    const float farD = bFirst ? a0 : b0;                                                      \
                    FCPW_PUSH(BLOCK, K, farL, farD, { OVERFLOW = true; (L).active = false; (L).live = false; }); \
                }                                                                                             \
            }                                                                                                 \
        }                                                                                                     \
    } while (0)

// The leaf pass, run by the whole wave: all (lane, primitive) pairs of the held
// leaves are evaluated by all 64 lanes, then each owner replays the reference's
// acceptance test (d*d <= r2, later equal distances win) in leaf order.
template <int DIM>
__device__ __forceinline__ uint32_t coop_leaf_pass(const DevScene &S, CoopShared &sh, CoopLane<DIM> &L)
{
    uint32_t groups = 0;
    using G = Geo<DIM>;
    const int lane = threadIdx.x & 63;
    float xs[3];
    store_p(L.x, xs);
    for (uint32_t j0 = 0; j0 < (uint32_t)MAX_LEAF_PRIMS; j0 += 4) {
        bool h0 = L.holding && L.lcnt > j0, h1 = L.holding && L.lcnt > j0 + 1;
        bool h2 = L.holding && L.lcnt > j0 + 2, h3 = L.holding && L.lcnt > j0 + 3;
        uint64_t m0 = __ballot(h0), m1 = __ballot(h1), m2 = __ballot(h2), m3 = __ballot(h3);
        if (m0 == 0) break;
        uint32_t c0 = __popcll(m0), c1 = __popcll(m1), c2 = __popcll(m2);
        uint32_t total = c0 + c1 + c2 + __popcll(m3);
        uint32_t i0 = lanes_below(m0), i1 = c0 + lanes_below(m1), i2 = c0 + c1 + lanes_below(m2),
                 i3 = c0 + c1 + c2 + lanes_below(m3);
        if (h0) sh.item[i0] = (uint16_t)((lane << 4) | j0);
        if (h1) sh.item[i1] = (uint16_t)((lane << 4) | (j0 + 1));
        if (h2) sh.item[i2] = (uint16_t)((lane << 4) | (j0 + 2));
        if (h3) sh.item[i3] = (uint16_t)((lane << 4) | (j0 + 3));
        __syncthreads();
        for (uint32_t pass = 0; pass < total; pass += 64) {
            groups++;
            const uint32_t k = pass + (uint32_t)lane;
            // every lane takes part in the permutes; lanes past the end read item 0
            const uint32_t it = sh.item[k < total ? k : 0u];
            const int o = (int)(it >> 4);
            float xo[3];
            xo[0] = __shfl(xs[0], o, 64);
            xo[1] = __shfl(xs[1], o, 64);
            xo[2] = DIM == 3 ? __shfl(xs[2], o, 64) : 0.0f;
            const uint32_t off = __shfl(L.loff, o, 64);
            if (k < total) {
                typename G::vec pp;
                float u0, u1;
                sh.d[k] = G::closest(S, off + (it & 15u), G::load_point(xo), pp, u0, u1);
            }
        }
        __syncthreads();
        if (L.holding) {
            const uint32_t idx[4] = {i0, i1, i2, i3};
            for (uint32_t jj = 0; jj < 4; jj++) {
                if (L.lcnt > j0 + jj) {
                    L.n_prims++;
                    float d = sh.d[idx[jj]];
                    if (d * d <= L.r2) {
                        L.r2 = d * d; // min(r2, d * d) when d * d <= r2
                        L.best = (int32_t)(L.loff + j0 + jj);
                    }
                }
            }
        }
        __syncthreads();
        if (m3 == 0) break;
    }
    L.holding = false;
    return groups;
}

// The lane's result: the winning primitive's closest point and uv recomputed.
template <int DIM>
__device__ __forceinline__ void coop_result(const DevScene &S, const CoopLane<DIM> &L, CpResult &res)
{
    res.slot = -1;
    res.d = FLT_MAX;
    res.p[0] = res.p[1] = res.p[2] = 0.0f;
    res.uv[0] = res.uv[1] = 0.0f;
    res.n_internal = L.n_internal;
    res.n_prims = L.n_prims;
    if (L.best >= 0) {
        typename Geo<DIM>::vec p;
        float u0, u1;
        res.d = Geo<DIM>::closest(S, (uint32_t)L.best, L.x, p, u0, u1);
        res.slot = L.best;
        res.uv[0] = u0;
        res.uv[1] = u1;
        store_p(p, res.p);
    }
}

template <int DIM, int K>
__device__ __forceinline__ void closest_point_query_coop(const DevScene &S, typename Geo<DIM>::vec x, float r2,
                                                         bool valid, int hold_min, CpResult &res)
{
    constexpr int BLOCK = 64;
    __shared__ CoopShared sh;
    CoopLane<DIM> L;
    coop_begin<DIM>(S, sh, L, x, r2, valid);
    bool overflow = false;
    uint32_t n_iter = 0, n_pass = 0, n_group = 0;
    FCPW_STACK_DECL(BLOCK, K);
    __syncthreads();
    while (true) {
        // two traversal steps per wave iteration: fewer ballots and leaf-pass checks per
        // node (a lane that reaches a leaf in the first holds it through the second);
        // measured on the 1M-triangle torus, 16M queries: 57.7 ms per forward step against 60.0
        // with one step per iteration and 58.8 with three (wall time per step, an earlier build)
        FCPW_COOP_STEP(DIM, BLOCK, K, S, L, overflow);
        FCPW_COOP_STEP(DIM, BLOCK, K, S, L, overflow);
        uint64_t act = __ballot(L.active);
        if (act == 0) break;
        n_iter++;
        uint64_t hold = __ballot(L.holding);
        if (hold == 0 || ((int)__popcll(hold) < hold_min && hold != act)) continue;
        n_pass++;
        n_group += coop_leaf_pass<DIM>(S, sh, L);
    }
    if (overflow) atomicOr(S.flags, 1);
    coop_result<DIM>(S, L, res);
    res.n_iter = n_iter;
    res.n_pass = n_pass;
    res.n_group = n_group;
}

// Adds a wave's traversal counts to S.stats (called by every lane of the wave).
__device__ __forceinline__ void wave_stats_add(const DevScene &S, uint32_t n_internal, uint32_t n_prims,
                                               uint32_t n_queries)
{
    for (int off = 32; off > 0; off >>= 1) {
        n_internal += __shfl_xor(n_internal, off, 64);
        n_prims += __shfl_xor(n_prims, off, 64);
        n_queries += __shfl_xor(n_queries, off, 64);
    }
    if ((threadIdx.x & 63u) == 0) {
        atomicAdd(&S.stats[0], (unsigned long long)n_internal);
        atomicAdd(&S.stats[1], (unsigned long long)n_prims);
        atomicAdd(&S.stats[2], (unsigned long long)n_queries);
    }
}

// Adds a wave's loop counts (wave-uniform values) to S.stats.
__device__ __forceinline__ void wave_loop_stats_add(const DevScene &S, uint32_t n_iter, uint32_t n_pass,
                                                    uint32_t n_group)
{
    if ((threadIdx.x & 63u) == 0) {
        atomicAdd(&S.stats[3], (unsigned long long)n_iter);
        atomicAdd(&S.stats[4], (unsigned long long)n_pass);
        atomicAdd(&S.stats[5], (unsigned long long)n_group);
        atomicAdd(&S.stats[6], 1ull);
    }
}

template <int DIM, int K, int OUT, bool STATS>
__global__ __launch_bounds__(64) void k_closest_points_coop(DevScene S, CpIO io, int xcds, int hold_min)
{
    int64_t k = xcd_tile(xcds) * 64 + threadIdx.x;
    bool valid = k < io.n;
    int64_t i = valid ? (io.order ? (int64_t)io.order[k] : k) : 0;
    float q[3] = {0.0f, 0.0f, 0.0f}, r2 = -1.0f;
    if (valid) cp_load<DIM, OUT>(io, i, q, r2);
    CpResult r;
    closest_point_query_coop<DIM, K>(S, Geo<DIM>::load_point(q), r2, valid, hold_min, r);
    if (valid) cp_store<DIM, OUT>(S, io, i, r);
    if (STATS) {
        wave_stats_add(S, r.n_internal, r.n_prims, valid ? 1u : 0u);
        wave_loop_stats_add(S, r.n_iter, r.n_pass, r.n_group);
    }
}

// ---------------------------------------------------------------------------
// Wave-cooperative ray traversal: the leaf passes of the closest-point kernel
// applied to rays.  A ray-primitive test computes t without reference to
// tmax and accepts the hit when 0 <= t <= tmax (geometry.slang:37-77,
// 123-160), so the wave evaluates t for every (lane, primitive) pair of the
// held leaves with no upper limit, and each owner replays the acceptance
// test, tmax = t, in leaf order; an occlusion query stops at its first
// accepted hit.  The winner's point, normal and uv are recomputed at the end.
// ---------------------------------------------------------------------------
template <int DIM>
struct RayCoopLane {
    typename Geo<DIM>::vec o, d, dinv;
    float tmax, cur_t;
    uint32_t cur, loff, lcnt;
    int32_t best;
    bool active, live, holding;
    uint32_t n_internal, n_prims;
};

// t of a ray-primitive hit with no upper limit on t; -1 when the primitive is missed
template <int DIM>
__device__ __forceinline__ float ray_t(const DevScene &S, uint32_t slot, typename Geo<DIM>::vec o,
                                       typename Geo<DIM>::vec d)
{
    if constexpr (DIM == 3) {
        const float4 *t = S.prims + (size_t)slot * 3;
        const float4 a = t[0], b = t[1], c = t[2];
        float tt, v, w;
        const v3 pa = {a.x, a.y, a.z}, pb = {b.x, b.y, b.z}, pc = {c.x, c.y, c.z};
        return ray_triangle(pa, pb, pc, o, d, __builtin_inff(), tt, v, w) ? tt : -1.0f;
    } else {
        const float4 s4 = S.prims[slot];
        const v2 pa = {s4.x, s4.y}, pb = {s4.z, s4.w};
        float tt, s;
        return ray_segment(pa, pb, o, d, __builtin_inff(), tt, s) ? tt : -1.0f;
    }
}

// One traversal step of an active ray lane that holds no leaf (the order of
// ray_query): pop until a node's entry distance is within tmax, then hold it
// (leaf) or expand it (internal: continue with the nearer child hit, push the
// farther).
#define FCPW_RAY_COOP_STEP(DIM, BLOCK, K, S, L, OVERFLOW)                                                    \
    do {                                                                                                      \
        using G_ = Geo<DIM>;                                                                                  \
        using vec_ = typename G_::vec;                                                                        \
        if ((L).active && !(L).holding) {                                                                     \
            if (!((L).live && !((L).cur_t > (L).tmax))) {                                                    \
                (L).live = false;                                                                             \
                while (sp > 0) {                                                                              \
                    FCPW_POP(BLOCK, K, (L).cur, (L).cur_t);                                                   \
                    if (!((L).cur_t > (L).tmax)) { (L).live = true; break; }                                  \
                }                                                                                             \
            }                                                                                                 \
            if (!(L).live) {                                                                                  \
                (L).active = false;                                                                           \
            } else if ((L).cur & LEAF_BIT) {                                                                  \
                (L).holding = true;                                                                           \
                (L).live = false;                                                                             \
                (L).loff = (L).cur & OFF_MASK;                                                                \
                (L).lcnt = (((L).cur >> CNT_SHIFT) & CNT_MASK) + 1u;                                          \
            } else {                                                                                          \
                vec_ loA, hiA, loB, hiB;                                                                      \
                uint32_t lA, lB;                                                                              \
                G_::record(S, (L).cur, loA, hiA, lA, loB, hiB, lB);                                           \
                (L).n_internal++;                                                                             \
                float ta, tb;                                                                                 \
                const bool hA = box_ray(loA, hiA, (L).o, (L).d, (L).dinv, (L).tmax, ta);                      \
                const bool hB = box_ray(loB, hiB, (L).o, (L).d, (L).dinv, (L).tmax, tb);                      \
                (L).live = hA || hB;                                                                          \
                if (hA && hB) {                                                                               \
                    uint32_t nearL = lA, farL = lB;                                                           \
                    float nearT = ta, farT = tb;                                                              \
                    if (tb < ta) { nearL = lB; farL = lA; nearT = tb; farT = ta; }                           \
                    FCPW_PUSH(BLOCK, K, farL, farT, { OVERFLOW = true; (L).active = false; (L).live = false; }); \
                    (L).cur = nearL;                                                                          \
                    (L).cur_t = nearT;                                                                        \
                } else if (hA) {                                                                              \
                    (L).cur = lA;                                                                             \
                    (L).cur_t = ta;                                                                           \
                } else if (hB) {                                                                              \
                    (L).cur = lB;                                                                             \
                    (L).cur_t = tb;                                                                           \
                }                                                                                             \
            }                                                                                                 \
        }                                                                                                     \
    } while (0)

// The ray leaf pass, run by the whole wave: t of every (lane, primitive) pair
// of the held leaves, then each owner's in-order acceptance (0 <= t <= tmax,
// tmax = t; an occlusion query ends at its first hit).
template <int DIM>
__device__ __forceinline__ uint32_t ray_coop_leaf_pass(const DevScene &S, CoopShared &sh, RayCoopLane<DIM> &L,
                                                       bool occlusion)
{
    uint32_t groups = 0;
    using G = Geo<DIM>;
    const int lane = threadIdx.x & 63;
    float os[3], ds[3];
    store_p(L.o, os);
    store_p(L.d, ds);
    for (uint32_t j0 = 0; j0 < (uint32_t)MAX_LEAF_PRIMS; j0 += 4) {
        bool h0 = L.holding && L.lcnt > j0, h1 = L.holding && L.lcnt > j0 + 1;
        bool h2 = L.holding && L.lcnt > j0 + 2, h3 = L.holding && L.lcnt > j0 + 3;
        uint64_t m0 = __ballot(h0), m1 = __ballot(h1), m2 = __ballot(h2), m3 = __ballot(h3);
        if (m0 == 0) break;
        uint32_t c0 = __popcll(m0), c1 = __popcll(m1), c2 = __popcll(m2);
        uint32_t total = c0 + c1 + c2 + __popcll(m3);
        uint32_t i0 = lanes_below(m0), i1 = c0 + lanes_below(m1), i2 = c0 + c1 + lanes_below(m2),
                 i3 = c0 + c1 + c2 + lanes_below(m3);
        if (h0) sh.item[i0] = (uint16_t)((lane << 4) | j0);
        if (h1) sh.item[i1] = (uint16_t)((lane << 4) | (j0 + 1));
        if (h2) sh.item[i2] = (uint16_t)((lane << 4) | (j0 + 2));
        if (h3) sh.item[i3] = (uint16_t)((lane << 4) | (j0 + 3));
        __syncthreads();
        for (uint32_t pass = 0; pass < total; pass += 64) {
            groups++;
            const uint32_t k = pass + (uint32_t)lane;
            // every lane takes part in the permutes; lanes past the end read item 0
            const uint32_t it = sh.item[k < total ? k : 0u];
            const int src = (int)(it >> 4);
            float oo[3], dd[3];
            oo[0] = __shfl(os[0], src, 64);
            oo[1] = __shfl(os[1], src, 64);
            oo[2] = DIM == 3 ? __shfl(os[2], src, 64) : 0.0f;
            dd[0] = __shfl(ds[0], src, 64);
            dd[1] = __shfl(ds[1], src, 64);
            dd[2] = DIM == 3 ? __shfl(ds[2], src, 64) : 0.0f;
            const uint32_t off = __shfl(L.loff, src, 64);
            if (k < total) sh.d[k] = ray_t<DIM>(S, off + (it & 15u), G::load_point(oo), G::load_point(dd));
        }
        __syncthreads();
        if (L.holding) {
            const uint32_t idx[4] = {i0, i1, i2, i3};
            for (uint32_t jj = 0; jj < 4; jj++) {
                if (L.lcnt > j0 + jj) {
                    L.n_prims++;
                    const float t = sh.d[idx[jj]];
                    if (t >= 0.0f && t <= L.tmax) {
                        L.best = (int32_t)(L.loff + j0 + jj);
                        if (occlusion) {
                            // the query ends here: no later group of this leaf, no further traversal
                            L.lcnt = 0;
                            L.active = false;
                            break;
                        }
                        L.tmax = t; // min(tmax, t) when t <= tmax
                    }
                }
            }
        }
        __syncthreads();
        if (m3 == 0) break;
    }
    L.holding = false;
    return groups;
}

template <int DIM, int K, bool STATS>
__global__ __launch_bounds__(64) FCPW_EIGHT_WAVES void k_intersect_coop(DevScene S, const float2 *rays, int64_t n, int occlusion,
                                                       float2 *out, const uint32_t *order, int xcds, int hold_min)
{
    using G = Geo<DIM>;
    constexpr int BLOCK = 64;
    __shared__ CoopShared sh;
    const int64_t k = xcd_tile(xcds) * 64 + threadIdx.x;
    const bool valid = k < n;
    const int64_t i = valid ? (order ? (int64_t)order[k] : k) : 0;
    float o3[3] = {0.0f, 0.0f, 0.0f}, d3[3] = {0.0f, 0.0f, 0.0f}, di3[3] = {0.0f, 0.0f, 0.0f};
    float tmax = -1.0f;
    if (valid) {
        const float2 *rr = rays + i * 5; // GPURay: o(3) d(3) dinv(3) tmax
        const float2 r0 = rr[0], r1 = rr[1], r2 = rr[2], r3 = rr[3], r4 = rr[4];
        o3[0] = r0.x; o3[1] = r0.y; o3[2] = r1.x;
        d3[0] = r1.y; d3[1] = r2.x; d3[2] = r2.y;
        di3[0] = r3.x; di3[1] = r3.y; di3[2] = r4.x;
        tmax = r4.y;
    }
    RayCoopLane<DIM> L;
    L.o = G::load_point(o3);
    L.d = G::load_point(d3);
    L.dinv = G::load_point(di3);
    L.tmax = tmax;
    L.holding = false;
    L.loff = L.lcnt = 0;
    L.best = -1;
    L.n_internal = L.n_prims = 0;
    {
        typename G::vec lo, hi;
        G::root(S, lo, hi, L.cur);
        L.active = valid && box_ray(lo, hi, L.o, L.d, L.dinv, L.tmax, L.cur_t);
        L.live = L.active;
    }
    bool overflow = false;
    const bool occ = occlusion != 0;
    uint32_t n_iter = 0, n_pass = 0, n_group = 0;
    FCPW_STACK_DECL(BLOCK, K);
    __syncthreads();
    while (true) {
        // two traversal steps per wave iteration, as in closest_point_query_coop
        FCPW_RAY_COOP_STEP(DIM, BLOCK, K, S, L, overflow);
        FCPW_RAY_COOP_STEP(DIM, BLOCK, K, S, L, overflow);
        const uint64_t act = __ballot(L.active);
        if (act == 0) break;
        n_iter++;
        const uint64_t hold = __ballot(L.holding);
        if (hold == 0 || ((int)__popcll(hold) < hold_min && hold != act)) continue;
        n_pass++;
        n_group += ray_coop_leaf_pass<DIM>(S, sh, L, occ);
    }
    if (overflow) atomicOr(S.flags, 1);
    if (valid) {
        const uint32_t index = L.best >= 0 ? (uint32_t)S.prim_index[L.best] : FCPW_HIP_UINT_MAX;
        float z[3] = {0.0f, 0.0f, 0.0f}, uv[2] = {0.0f, 0.0f};
        if (occ) {
            // only the occluder's index is written (bvh.slang:183-187)
            write_interaction(out + i * 5, z, z, uv, FLT_MAX, index);
        } else if (L.best >= 0) {
            RayResult r;
            RayGeo<DIM>::hit(S, (uint32_t)L.best, L.o, L.d, __builtin_inff(), r, true);
            write_interaction(out + i * 5, r.p, r.n, r.uv, r.t, index);
        } else {
            write_interaction(out + i * 5, z, z, uv, FLT_MAX, index);
        }
    }
    if (STATS) {
        wave_stats_add(S, L.n_internal, L.n_prims, valid ? 1u : 0u);
        wave_loop_stats_add(S, n_iter, n_pass, n_group);
    }
}

// Sphere intersection (bvh.slang:257-392) with the constant branch weight of
// the sphereIntersection entry point (bvh.cs.slang:59-82): one random walk
// from the root that picks, with equal probability, one of the children
// overlapping the sphere (no stack), then an area-weighted reservoir choice
// among the leaf's primitives inside the sphere (all of them when the leaf box
// lies inside), then a uniform point on the chosen primitive.  d = the pdf of
// that point.  Random numbers: rnd[3i] drives the walk and the choice,
// rnd[3i+1], rnd[3i+2] the point.
// One sphere (x, r2) with random numbers u (walk and choice) and r1, r2s (the
// point); returns the chosen leaf slot (-1: none) and fills p, nrm, uv and d
// (the pdf; FLT_MAX when nothing is chosen).
template <int DIM>
__device__ __forceinline__ int32_t sphere_sample(const DevScene &S, typename Geo<DIM>::vec x, float r2, float u,
                                                 float r1, float r2s, float p[3], float nrm[3], float uv[2], float &d)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    p[0] = p[1] = p[2] = 0.0f;
    nrm[0] = nrm[1] = nrm[2] = 0.0f;
    uv[0] = uv[1] = 0.0f;
    float dsel = FLT_MAX;
    int32_t selected = -1;
    vec lo, hi;
    uint32_t cur;
    G::root(S, lo, hi, cur);
    float d0, d1;
    box_sqdist(lo, hi, x, d0, d1);
    if (d0 <= r2) {
        float maxd = d1, pdf = 1.0f;
        while (true) {
            if (cur & LEAF_BIT) {
                const uint32_t off = cur & OFF_MASK, cnt = ((cur >> CNT_SHIFT) & CNT_MASK) + 1u;
                float total = 0.0f;
                for (uint32_t k = 0; k < cnt; k++) {
                    bool hitp = maxd <= r2;
                    if (!hitp) {
                        vec pp;
                        float u0, u1;
                        float dd = G::closest(S, off + k, x, pp, u0, u1);
                        hitp = dd * dd <= r2;
                    }
                    if (hitp) {
                        const float a = G::area(S, off + k);
                        total += a;
                        const float prob = a / total;
                        if (u < prob) {
                            u = u / prob;
                            dsel = a * pdf;
                            selected = (int32_t)(off + k);
                        } else {
                            u = (u - prob) / (1.0f - prob);
                        }
                    }
                }
                if (total > 0.0f) dsel = dsel / total;
                break;
            }
            vec loA, hiA, loB, hiB;
            uint32_t lA, lB;
            G::record(S, cur, loA, hiA, lA, loB, hiB, lB);
            float a0, a1, b0, b1;
            box_sqdist(loA, hiA, x, a0, a1);
            box_sqdist(loB, hiB, x, b0, b1);
            const float wl = a0 <= r2 ? 1.0f : 0.0f, wr = b0 <= r2 ? 1.0f : 0.0f;
            const float tot = wl + wr;
            if (!(tot > 0.0f)) break;
            const float pl = wl / tot, pr = 1.0f - pl;
            if (u < pl) {
                u = u / pl;
                cur = lA;
                pdf *= pl;
                maxd = a1;
            } else {
                u = (u - pl) / pr;
                cur = lB;
                pdf *= pr;
                maxd = b1;
            }
        }
    }
    if (selected >= 0) dsel *= G::sample(S, (uint32_t)selected, r1, r2s, uv, p, nrm);
    else dsel = FLT_MAX;
    d = dsel;
    return selected;
}

template <int DIM>
__global__ __launch_bounds__(128) void k_intersect_spheres(DevScene S, const float4 *spheres, const float *rnd,
                                                           int64_t n, float2 *out)
{
    using G = Geo<DIM>;
    const int64_t i = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (i >= n) return;
    const float4 s4 = spheres[i];
    const float q[3] = {s4.x, s4.y, s4.z};
    float p[3], nrm[3], uv[2], d;
    const int32_t selected =
        sphere_sample<DIM>(S, G::load_point(q), s4.w, rnd[i * 3], rnd[i * 3 + 1], rnd[i * 3 + 2], p, nrm, uv, d);
    const uint32_t index = selected >= 0 ? (uint32_t)S.prim_index[selected] : FCPW_HIP_UINT_MAX;
    write_interaction(out + i * 5, p, nrm, uv, d, index);
}

template <int DIM, int BLOCK, int K, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_intersect(DevScene S, const float2 *rays, int64_t n, int occlusion,
                                                     float2 *out, const uint32_t *order, int xcds)
{
    using vec = typename Geo<DIM>::vec;
    int64_t k = xcd_tile(xcds) * BLOCK + threadIdx.x;
    uint32_t n_internal = 0, n_prims = 0;
    if (k < n) {
        int64_t i = order ? (int64_t)order[k] : k;
        const float2 *rr = rays + i * 5; // GPURay: o(3) d(3) dinv(3) tmax
        float2 r0 = rr[0], r1 = rr[1], r2 = rr[2], r3 = rr[3], r4 = rr[4];
        float o3[3] = {r0.x, r0.y, r1.x}, d3[3] = {r1.y, r2.x, r2.y}, di3[3] = {r3.x, r3.y, r4.x};
        float tmax = r4.y;
        vec o = Geo<DIM>::load_point(o3), d = Geo<DIM>::load_point(d3), di = Geo<DIM>::load_point(di3);
        RayResult r;
        ray_query<DIM, BLOCK, K>(S, o, d, di, tmax, occlusion != 0, r);
        n_internal = r.n_internal;
        n_prims = r.n_prims;
        uint32_t index = r.slot >= 0 ? (uint32_t)S.prim_index[r.slot] : FCPW_HIP_UINT_MAX;
        if (occlusion) {
            // only the occluder's index is written (bvh.slang:183-187)
            float z[3] = {0.0f, 0.0f, 0.0f}, uv[2] = {0.0f, 0.0f};
            write_interaction(out + i * 5, z, z, uv, FLT_MAX, index);
        } else {
            write_interaction(out + i * 5, r.p, r.n, r.uv, r.t, index);
        }
    }
    if (STATS) wave_stats_add(S, n_internal, n_prims, k < n ? 1u : 0u);
}

// ---------------------------------------------------------------------------
// Morton ordering of a query batch: a counting sort of the queries over a grid
// of 2^18, 2^21 or 2^24 cells (6, 7 or 8 bits per axis in 3D; 9, 10 or 12 in
// 2D) laid along a Z-order curve, with a three-phase multi-block scan.
// Neighbouring lanes then traverse neighbouring parts of the tree, which keeps
// a wave's branches together and its node fetches in L1/L2.  The order inside
// a cell is whatever the atomics produce; results do not depend on it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread3(uint32_t v) // 10 bits -> every third bit
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t spread2(uint32_t v) // 16 bits -> every second bit
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
__device__ __forceinline__ uint32_t quant(float x, float lo, float inv, uint32_t cells)
{
    float t = (x - lo) * inv; // NaN and out-of-range values clamp to the border cells
    int c = (t >= 0.0f) ? (int)fminf(t * (float)cells, (float)(cells - 1)) : 0;
    return (uint32_t)c;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// cell of a point on a grid of 2^axis_bits cells per axis, Z-order
template <int DIM>
__device__ __forceinline__ uint32_t morton_cell(const float *q, float4 lo, float4 inv, int axis_bits)
{
    const uint32_t cells = 1u << axis_bits;
    if (DIM == 3)
        return spread3(quant(q[0], lo.x, inv.x, cells)) | (spread3(quant(q[1], lo.y, inv.y, cells)) << 1) |
               (spread3(quant(q[2], lo.z, inv.z, cells)) << 2);
    return spread2(quant(q[0], lo.x, inv.x, cells)) | (spread2(quant(q[1], lo.y, inv.y, cells)) << 1);
}

// dir_off >= 0: the records also hold a direction at that float offset, and the
// signs of its components (the octant, or quadrant in 2D) form the code's top bits
template <int DIM>
__global__ void k_morton_hist(const float *pts, int stride, int64_t n, float4 lo, float4 inv, int axis_bits,
                              uint32_t *codes, uint32_t *ranks, uint32_t *hist, int dir_off)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t code = morton_cell<DIM>(pts + i * stride, lo, inv, axis_bits);
    if (dir_off >= 0) {
        const float *d = pts + i * stride + dir_off;
        uint32_t oct = (d[0] < 0.0f ? 1u : 0u) | (d[1] < 0.0f ? 2u : 0u);
        if (DIM == 3) oct |= d[2] < 0.0f ? 4u : 0u;
        code |= oct << (DIM * axis_bits);
    }
    codes[i] = code;
    ranks[i] = atomicAdd(&hist[code], 1u); // the query's place within its cell
}

// exclusive scan of n counters in place, three phases: per-segment sums, a
// one-block scan of the sums, then each segment rescanned from its base.
// Global memory is read and written lane-contiguously; the per-thread runs of
// SCAN_PER consecutive counters are formed in LDS (padded one word per 32 so
// the runs of a wave's lanes fall in distinct banks).
constexpr int SCAN_THREADS = 256, SCAN_PER = 32, SCAN_SEG = SCAN_THREADS * SCAN_PER;

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_segment_sums(const uint32_t *v, int64_t n, uint32_t *sums)
{
    __shared__ uint32_t part[SCAN_THREADS];
    const int64_t seg = (int64_t)blockIdx.x * SCAN_SEG;
    uint32_t s = 0;
    for (int k = 0; k < SCAN_PER; k++) {
        const int64_t i = seg + (int64_t)k * SCAN_THREADS + threadIdx.x;
        if (i < n) s += v[i];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = SCAN_THREADS / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = part[0];
}

// exclusive scan of m <= 1024 * 16 segment sums by one block
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t *sums, int m)
{
    constexpr int PER = 16;
    __shared__ uint32_t part[1024];
    const int base = threadIdx.x * PER;
    uint32_t s = 0;
    for (int k = 0; k < PER; k++)
        if (base + k < m) s += sums[base + k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (int k = 0; k < PER; k++)
        if (base + k < m) {
            uint32_t c = sums[base + k];
            sums[base + k] = run;
            run += c;
        }
}

__device__ __forceinline__ int scan_pad(int j) { return j + (j >> 5); }

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_segment_apply(uint32_t *v, int64_t n, const uint32_t *sums)
{
    __shared__ uint32_t part[SCAN_THREADS];
    __shared__ uint32_t seg_v[SCAN_SEG + SCAN_SEG / 32];
    const int64_t seg = (int64_t)blockIdx.x * SCAN_SEG;
    for (int k = 0; k < SCAN_PER; k++) {
        const int j = k * SCAN_THREADS + (int)threadIdx.x;
        seg_v[scan_pad(j)] = seg + j < n ? v[seg + j] : 0u;
    }
    __syncthreads();
    const int base = (int)threadIdx.x * SCAN_PER;
    uint32_t s = 0;
    for (int k = 0; k < SCAN_PER; k++) s += seg_v[scan_pad(base + k)];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        uint32_t x = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = sums[blockIdx.x] + part[threadIdx.x] - s;
    for (int k = 0; k < SCAN_PER; k++) {
        const int p = scan_pad(base + k);
        const uint32_t c = seg_v[p];
        seg_v[p] = run;
        run += c;
    }
    __syncthreads();
    for (int k = 0; k < SCAN_PER; k++) {
        const int j = k * SCAN_THREADS + (int)threadIdx.x;
        if (seg + j < n) v[seg + j] = seg_v[scan_pad(j)];
    }
}

// each query to its cell's first position plus its place within the cell
__global__ void k_morton_place(const uint32_t *codes, const uint32_t *ranks, int64_t n, const uint32_t *offsets,
                               uint32_t *perm)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[offsets[codes[i]] + ranks[i]] = (uint32_t)i;
}

// rebuild the primitive records from vertex positions (caller's vertex order)
template <int DIM>
__global__ void k_rebuild_prims(float4 *prims, const int32_t *vids, const float *V, int n_prims)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_prims) return;
    if (DIM == 3) {
        for (int k = 0; k < 3; k++) {
            const float *v = V + (size_t)vids[(size_t)s * 3 + k] * 3;
            prims[(size_t)s * 3 + k] = make_float4(v[0], v[1], v[2], 0.0f);
        }
    } else {
        const float *a = V + (size_t)vids[(size_t)s * 2] * 2;
        const float *b = V + (size_t)vids[(size_t)s * 2 + 1] * 2;
        prims[s] = make_float4(a[0], a[1], b[0], b[1]);
    }
}

// Bottom-up refit in one launch: one thread per leaf computes the leaf box,
// writes it into its parent record's slot and counts the arrival; the second
// arrival at a record unions the two slot boxes and climbs.  Box values only
// involve min, max and the +-FLT_EPSILON inflation, so the result is bit-equal
// to the reference refit (bvh.slang:27-141, bvh.inl:548-581) in any order.
// The second arrival resets the counter, leaving all counters 0 for the next call.
// refit_upward is the climb from one leaf box (ref = its slot, -1 for a root leaf).
template <int DIM>
__device__ __forceinline__ void refit_upward(float4 *records, float4 *root, const int32_t *rec_parent,
                                             unsigned int *counters, int32_t ref, float lo[3], float hi[3])
{
    while (true) {
        if (ref < 0) { // the box belongs to the root
            if (DIM == 3) {
                float4 a = root[0];
                root[0] = make_float4(lo[0], lo[1], lo[2], a.w);
                root[1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
            } else {
                root[0] = make_float4(lo[0], lo[1], hi[0], hi[1]);
            }
            return;
        }
        int32_t rec = ref >> 1, slot = ref & 1;
        if (DIM == 3) {
            float4 *p = records + (size_t)rec * 4 + slot * 2;
            float w = p[0].w;
            p[0] = make_float4(lo[0], lo[1], lo[2], w);
            p[1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        } else {
            records[(size_t)rec * 3 + slot] = make_float4(lo[0], lo[1], hi[0], hi[1]);
        }
        __threadfence();
        unsigned int prev = atomicAdd(&counters[rec], 1u);
        if (prev == 0) return;
        counters[rec] = 0u;
        __threadfence();
        if (DIM == 3) {
            const float4 *p = records + (size_t)rec * 4;
            float4 a0 = p[0], a1 = p[1], b0 = p[2], b1 = p[3];
            lo[0] = fminf(a0.x, b0.x); lo[1] = fminf(a0.y, b0.y); lo[2] = fminf(a0.z, b0.z);
            hi[0] = fmaxf(a1.x, b1.x); hi[1] = fmaxf(a1.y, b1.y); hi[2] = fmaxf(a1.z, b1.z);
        } else {
            const float4 *p = records + (size_t)rec * 3;
            float4 a = p[0], b = p[1];
            lo[0] = fminf(a.x, b.x); lo[1] = fminf(a.y, b.y);
            hi[0] = fmaxf(a.z, b.z); hi[1] = fmaxf(a.w, b.w);
        }
        ref = rec_parent[rec];
    }
}

template <int DIM>
__global__ void k_refit(float4 *records, float4 *root, const float4 *prims, const int32_t *leaf_refs,
                        const uint32_t *leaf_links, const int32_t *rec_parent, unsigned int *counters, int n_leaves)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_leaves) return;
    uint32_t link = leaf_links[j];
    uint32_t off = link & OFF_MASK, cnt = ((link >> CNT_SHIFT) & CNT_MASK) + 1u;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t k = 0; k < cnt; k++) {
        if (DIM == 3) {
            const float4 *t = prims + (size_t)(off + k) * 3;
            float4 a = t[0], b = t[1], c = t[2];
            // Triangle.getBoundingBox (geometry.slang:342-346)
            lo[0] = fminf(lo[0], fminf(fminf(a.x, b.x), c.x) - FLT_EPSILON);
            lo[1] = fminf(lo[1], fminf(fminf(a.y, b.y), c.y) - FLT_EPSILON);
            lo[2] = fminf(lo[2], fminf(fminf(a.z, b.z), c.z) - FLT_EPSILON);
            hi[0] = fmaxf(hi[0], fmaxf(fmaxf(a.x, b.x), c.x) + FLT_EPSILON);
            hi[1] = fmaxf(hi[1], fmaxf(fmaxf(a.y, b.y), c.y) + FLT_EPSILON);
            hi[2] = fmaxf(hi[2], fmaxf(fmaxf(a.z, b.z), c.z) + FLT_EPSILON);
        } else {
            float4 s = prims[off + k];
            // LineSegment.getBoundingBox (geometry.slang:112-116)
            lo[0] = fminf(lo[0], fminf(s.x, s.z) - FLT_EPSILON);
            lo[1] = fminf(lo[1], fminf(s.y, s.w) - FLT_EPSILON);
            hi[0] = fmaxf(hi[0], fmaxf(s.x, s.z) + FLT_EPSILON);
            hi[1] = fmaxf(hi[1], fmaxf(s.y, s.w) + FLT_EPSILON);
        }
    }
    refit_upward<DIM>(records, root, rec_parent, counters, leaf_refs[j], lo, hi);
}

// Adjoint of d(x, mesh): recompute p and its barycentric weights from the saved
// leaf slot, then d(d)/dx = (x - p)/d and d(d)/dV_k = -b_k (x - p)/d.
// The terms of one query: the slot (-1 when it contributes nothing), dL/dx and,
// per vertex k of the primitive, its id and -b_k dL/dx.
template <int DIM>
struct AdjointTerms {
    int32_t slot;
    float g[DIM];
    float c[DIM][DIM];
    int32_t vid[DIM];
};

template <int DIM>
__device__ __forceinline__ void adjoint_terms(const float4 *prims, const int32_t *prim_vids, int n_prims,
                                              const float *points, const int32_t *slots, const float *grad_d,
                                              int64_t i, AdjointTerms<DIM> &T)
{
    int32_t s = slots[i];
    if (s >= n_prims) s = -1; // not a slot of this scene
    for (int k = 0; k < DIM; k++) T.g[k] = 0.0f;
    T.slot = -1;
    float gd = grad_d[i];
    if (s < 0) return;
    float dist, w[3];
    float diff[3] = {0.0f, 0.0f, 0.0f};
    if (DIM == 3) {
        const float4 *t = prims + (size_t)s * 3;
        float4 a = t[0], b = t[1], c = t[2];
        v3 x = {points[i * 3], points[i * 3 + 1], points[i * 3 + 2]}, p;
        float t0, t1;
        dist = closest_point_triangle({a.x, a.y, a.z}, {b.x, b.y, b.z}, {c.x, c.y, c.z}, x, p, t0, t1);
        w[0] = t0; w[1] = t1; w[2] = 1.0f - t0 - t1;
        diff[0] = x.x - p.x; diff[1] = x.y - p.y; diff[2] = x.z - p.z;
    } else {
        float4 q = prims[s];
        v2 x = {points[i * 2], points[i * 2 + 1]}, p;
        float t;
        dist = closest_point_segment({q.x, q.y}, {q.z, q.w}, x, p, t);
        w[0] = 1.0f - t; w[1] = t;
        diff[0] = x.x - p.x; diff[1] = x.y - p.y;
    }
    if (!(dist > 0.0f)) return;
    float sc = gd / dist;
    for (int k = 0; k < DIM; k++) T.g[k] = sc * diff[k];
    for (int v = 0; v < DIM; v++) {
        T.vid[v] = prim_vids[(size_t)s * DIM + v];
        for (int k = 0; k < DIM; k++) T.c[v][k] = -w[v] * T.g[k];
    }
    T.slot = s;
}

// hardware FP32 atomic add (global_atomic_add_f32, no return value) rather
// than the compare-and-swap loop atomicAdd compiles to without
// -munsafe-fp-atomics; grad_vertices is device (coarse-grained) memory
template <int DIM>
__device__ __forceinline__ void adjoint_scatter(const AdjointTerms<DIM> &T, float *grad_vertices)
{
    for (int v = 0; v < DIM; v++)
        for (int k = 0; k < DIM; k++)
            (void)unsafeAtomicAdd(&grad_vertices[(size_t)T.vid[v] * DIM + k], T.c[v][k]);
}

// one query of the adjoint (shared by the single-scene and batched kernels)
template <int DIM>
__device__ __forceinline__ void distance_backward_one(const float4 *prims, const int32_t *prim_vids, int n_prims,
                                                      const float *points, const int32_t *slots,
                                                      const float *grad_d, int64_t i, float *grad_points,
                                                      float *grad_vertices)
{
    AdjointTerms<DIM> T;
    adjoint_terms<DIM>(prims, prim_vids, n_prims, points, slots, grad_d, i, T);
    if (grad_vertices && T.slot >= 0) adjoint_scatter<DIM>(T, grad_vertices);
    if (grad_points)
        for (int k = 0; k < DIM; k++) grad_points[i * DIM + k] = T.g[k];
}

// The adjoint over the queries in a given order -- the Morton order of the
// preceding forward pass, so that a wave's queries read neighbouring primitives
// and scatter into neighbouring vertices (entry k handles query order[k]).
// REDUCE 1: a run of consecutive lanes on the same primitive first sums its
// vertex terms (a segmented scan across the wave) and only the run's last lane
// issues the atomics.  REDUCE 2: every vertex term of the wave goes into a
// per-wave hash table in LDS keyed by vertex id (LDS atomics), and each vertex
// the wave touched gets one global atomic per coordinate.
constexpr int BWD_HASH = 256; // entries per wave: more than the 64 x 3 keys a wave inserts

// Sums the vertex terms of a wave's lanes per vertex id in the wave's LDS hash
// table (hkey / hval, BWD_HASH entries), then adds each touched vertex to
// grad_vertices with one global atomic per coordinate.  Every thread of the
// block calls it (it holds two block barriers); a lane with T.slot < 0 adds
// nothing.
template <int DIM>
__device__ __forceinline__ void adjoint_hash_scatter(const AdjointTerms<DIM> &T, uint32_t *hkey, float (*hval)[DIM],
                                                     float *grad_vertices)
{
    const int lane = (int)(threadIdx.x & 63u);
    for (int e = lane; e < BWD_HASH; e += 64) {
        hkey[e] = 0xFFFFFFFFu;
        for (int c = 0; c < DIM; c++) hval[e][c] = 0.0f;
    }
    __syncthreads();
    if (T.slot >= 0) {
        for (int v = 0; v < DIM; v++) {
            const uint32_t key = (uint32_t)T.vid[v];
            uint32_t h = (key * 2654435761u) >> 24; // 8 bits: BWD_HASH = 256
            // at most 64 * DIM < BWD_HASH keys, so an empty entry is always found
            while (true) {
                const uint32_t old = atomicCAS(&hkey[h], 0xFFFFFFFFu, key);
                if (old == 0xFFFFFFFFu || old == key) break;
                h = (h + 1u) & (BWD_HASH - 1);
            }
            for (int c = 0; c < DIM; c++) atomicAdd(&hval[h][c], T.c[v][c]);
        }
    }
    __syncthreads();
    for (int e = lane; e < BWD_HASH; e += 64) {
        const uint32_t key = hkey[e];
        if (key != 0xFFFFFFFFu)
            for (int c = 0; c < DIM; c++) (void)unsafeAtomicAdd(&grad_vertices[(size_t)key * DIM + c], hval[e][c]);
    }
}
template <int DIM, int REDUCE>
__global__ __launch_bounds__(256) void k_distance_backward_ordered(const float4 *prims, const int32_t *prim_vids,
                                                                   const float *points, const int32_t *slots,
                                                                   const float *grad_d, int64_t n, int n_prims,
                                                                   const uint32_t *order, float *grad_points,
                                                                   float *grad_vertices)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = k < n ? (int64_t)order[k] : n;
    const bool valid = i < n;
    AdjointTerms<DIM> T;
    T.slot = -1;
    if (valid) {
        adjoint_terms<DIM>(prims, prim_vids, n_prims, points, slots, grad_d, i, T);
        if (grad_points)
            for (int c = 0; c < DIM; c++) grad_points[i * DIM + c] = T.g[c];
    }
    if (!grad_vertices) return;
    if constexpr (REDUCE == 0) {
        if (T.slot >= 0) adjoint_scatter<DIM>(T, grad_vertices);
    } else if constexpr (REDUCE == 2) {
        // one table per wave of the 256-thread block
        __shared__ uint32_t hkey[4][BWD_HASH];
        __shared__ float hval[4][BWD_HASH][DIM];
        const int w = (int)(threadIdx.x >> 6);
        adjoint_hash_scatter<DIM>(T, hkey[w], hval[w], grad_vertices);
    } else {
        // every lane of the wave takes part in the shuffles; a lane with nothing
        // to add carries slot -1 and zero terms
        const int lane = (int)(threadIdx.x & 63u);
        if (T.slot < 0)
            for (int v = 0; v < DIM; v++)
                for (int c = 0; c < DIM; c++) T.c[v][c] = 0.0f;
        const int prev = __shfl_up(T.slot, 1, 64);
        const bool head = lane == 0 || prev != T.slot;
        const uint64_t heads = __ballot(head);
        const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const int start = 63 - __clzll(heads & upto); // the first lane of this lane's run
        for (int off = 1; off < 64; off <<= 1) {
            for (int v = 0; v < DIM; v++)
                for (int c = 0; c < DIM; c++) {
                    const float t = __shfl_up(T.c[v][c], off, 64);
                    if (lane - off >= start) T.c[v][c] += t;
                }
        }
        const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
        if (tail && T.slot >= 0) adjoint_scatter<DIM>(T, grad_vertices);
    }
}

template <int DIM>
__global__ void k_distance_backward(const float4 *prims, const int32_t *prim_vids, const float *points,
                                    const int32_t *slots, const float *grad_d, int64_t n, int n_prims,
                                    float *grad_points, float *grad_vertices)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    distance_backward_one<DIM>(prims, prim_vids, n_prims, points, slots, grad_d, i, grad_points, grad_vertices);
}

} // namespace fcpw_hip
