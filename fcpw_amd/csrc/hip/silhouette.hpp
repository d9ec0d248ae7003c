#pragma once
// silhouette.hpp -- device code of the closest-silhouette query and of the cone
// refit.  Included once, by fcpw_hip.hip, after device.hpp.
//
// Device layout (next to the closest-point records of device.hpp):
//   cones     per record two child entries, then one entry for the root; an
//             entry is two float4 {axis.xyz, halfAngle} {radius, sin halfAngle,
//             cos halfAngle, max(1, |axis|)} (GPUBoundingCone,
//             bvh_interop_structures.h:41-52, plus three values of its own
//             for cone_overlap_bounded)
//   sil_leaf  per leaf, indexed by the leaf's first primitive slot: uint2
//             {first silhouette record, count}
//   sils      silhouette records in reference order (a silhouette shared by two
//             leaves appears in both): 2D vertex float4 {p.xy, n0.xy}
//             {n1.xy, index, hasOneAdjacentFace}; 3D edge float4 {pa, index}
//             {pb, hasOneAdjacentFace} {n0, 0} {n1, 0}
//
// All silhouette arithmetic is the reference's float3 arithmetic
// (include/fcpw/gpu/bvh.slang, geometry.slang, bounding-volumes.slang): dots
// summed left to right, 2D carried with z = 0.  acosf/asinf/atan2f/cosf/sinf
// are the device library's.

namespace fcpw_hip {

struct SilScene {
    const float4 *cones;
    const uint2 *sil_leaf;
    const float4 *sils;
};

constexpr float PI_F = 3.14159265358979323846f;   // M_PI (math-constants.slang)
constexpr float PI_2_F = 1.57079632679489661923f; // M_PI_2

__device__ __forceinline__ v3 to3(v3 a) { return a; }
__device__ __forceinline__ v3 to3(v2 a) { return {a.x, a.y, 0.0f}; }
__device__ __forceinline__ bool in_range(float v, float lo, float hi) { return v >= lo && v <= hi; }

// computeOrthonormalBasis + projectToPlane (bounding-volumes.slang:122-142)
__device__ __forceinline__ float project_to_plane(v3 n, v3 e)
{
    const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    const v3 b1 = {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    const v3 b2 = {b, sign + n.y * n.y * a, -n.y};
    const float r1 = dot(e, v3{fabsf(b1.x), fabsf(b1.y), fabsf(b1.z)});
    const float r2 = dot(e, v3{fabsf(b2.x), fabsf(b2.y), fabsf(b2.z)});
    return sqrtf(r1 * r1 + r2 * r2);
}

// BoundingCone.overlap (bounding-volumes.slang:170-236): whether some direction
// inside the cone is orthogonal to some direction from o into the box
__device__ __forceinline__ bool cone_overlap(float4 cone, float radius, v3 o, v3 lo, v3 hi, float dist_to_box)
{
    const float half = cone.w;
    if (half >= PI_2_F || dist_to_box < FLT_EPSILON) return true;
    const v3 c = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    v3 v = sub(c, o);
    const float l = sqrtf(dot(v, v));
    v = {v.x / l, v.y / l, v.z / l};
    const float d_axis = acosf(fmaxf(-1.0f, fminf(1.0f, dot(v3{cone.x, cone.y, cone.z}, v))));
    if (in_range(PI_2_F, d_axis - half, d_axis + half)) return true;
    if (l > radius) {
        const float hs = half + asinf(radius / l);
        return hs >= PI_2_F ? true : in_range(PI_2_F, d_axis - hs, d_axis + hs);
    }
    const v3 e = sub(hi, c);
    float d = dot(e, v3{fabsf(v.x), fabsf(v.y), fabsf(v.z)});
    const float s = l - d;
    if (s <= 0.0f) return true;
    d = project_to_plane(v, e);
    const float hs = half + atan2f(d, s);
    return hs >= PI_2_F ? true : in_range(PI_2_F, d_axis - hs, d_axis + hs);
}

// The second float4 of a cone entry (written by the upload and by the device
// cone refit): the reference's radius, then sin and cos of the half angle and
// max(1, |axis|) (NaN for a NaN axis, which sends every test to the reference form).
__host__ __device__ inline float4 cone_aux(float ax, float ay, float az, float half, float radius)
{
    const float n = sqrtf(ax * ax + ay * ay + az * az);
    return make_float4(radius, sinf(half), cosf(half), n <= 1.0f ? 1.0f : n);
}

// BoundingCone.overlap decided without inverse trigonometry or division where
// the decision is clear by a margin.  With v = c - o, l = |v| and
// cc = dot(axis, v / l), the reference accepts (past its two early accepts)
// iff hs = half + beta >= pi/2 or |acos(cc) - pi/2| <= hs, where
//   beta = asin(radius / l)                         when l > radius,
//   beta = atan2(d, s), s = l - dot(e, |v / l|),     when l <= radius
//          d = projectToPlane(v / l, e)             (accept at once if s <= 0);
// its first test, |acos(cc) - pi/2| <= half, is implied, also in floats, as
// hs >= half and rounding is monotone.  In sines, with sh = sin(half),
// ch = cos(half):
//   l > radius, w = sqrt(l^2 - radius^2):
//     hs >= pi/2            <=>  radius - l ch >= 0
//     asin|cc| <= hs < pi/2 <=>  (sh w + ch radius) - |dot(axis, v)| >= 0
//   l <= radius, rho = sqrt(d^2 + s^2):
//     s <= 0                <=>  l^2 - dot(e, |v|) <= 0
//     hs >= pi/2            <=>  d sh - s ch >= 0
//     asin|cc| <= hs < pi/2 <=>  l (sh s + ch d) - |dot(axis, v)| rho >= 0
// (sin(half + beta) times l, or times l rho).  d, s and 1/l are computed here
// with the hardware reciprocal; a side decides only when it clears zero by
// CONE_MARGIN times its scale (l, l^2 or l (rho + l), times max(1, |axis|)).
// The float error of these sides, and the reference's own (rounding of cc and
// of the angle sums, the last-place errors of acos, asin and atan2, mapped to
// sines, where no error is larger than in angles and near |cc| = 1 the error of
// acos shrinks to 2 |dcc|), stay below 1e-6 of that scale; inside the margin,
// or when a side is NaN, the reference's form decides.  So the decision is the
// reference's everywhere.  c, v and l are computed as the reference computes
// them, so both take the same branch on l > radius, and the analysis starts
// from that v: its own rounding is shared.  Past the early accept on
// dist_to_box the query lies outside the box, so l >= sqrt(FLT_EPSILON) and
// no margin underflows.
constexpr float CONE_MARGIN = 2e-5f;

// the reference's form out of line: rarely reached, and kept out of the
// bounded test's register allocation
__device__ __noinline__ bool cone_overlap_ref(float4 cone, float radius, v3 o, v3 lo, v3 hi, float dist_to_box)
{
    return cone_overlap(cone, radius, o, lo, hi, dist_to_box);
}

__device__ __forceinline__ bool cone_overlap_bounded(float4 cone, float4 aux, v3 o, v3 lo, v3 hi, float dist_to_box)
{
    if (cone.w >= PI_2_F || dist_to_box < FLT_EPSILON) return true;
    const v3 c = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    const v3 v = sub(c, o);
    const float l = sqrtf(dot(v, v));
    const float radius = aux.x, sh = aux.y, ch = aux.z;
    const float a = fabsf(dot(v3{cone.x, cone.y, cone.z}, v));
    if (l > radius) {
        const float m = CONE_MARGIN * l * aux.w;
        const float t = radius - l * ch;
        const float w = __builtin_amdgcn_sqrtf((l - radius) * (l + radius));
        const float g = (sh * w + ch * radius) - a;
        if (t >= m || g >= m) return true;
        if (t <= -m && g <= -m) return false;
    } else {
        const v3 e = sub(hi, c);
        const float sl = l * l - dot(e, v3{fabsf(v.x), fabsf(v.y), fabsf(v.z)});
        const float ms = CONE_MARGIN * l * l;
        if (sl <= -ms) return true;
        if (sl >= ms) {
            // projectToPlane(v / l, e) with the reciprocal
            const float il = __builtin_amdgcn_rcpf(l);
            const v3 n = {v.x * il, v.y * il, v.z * il};
            const float sg = n.z >= 0.0f ? 1.0f : -1.0f;
            const float pa = -__builtin_amdgcn_rcpf(sg + n.z);
            const float pb = n.x * n.y * pa;
            const v3 b1 = {1.0f + sg * n.x * n.x * pa, sg * pb, -sg * n.x};
            const v3 b2 = {pb, sg + n.y * n.y * pa, -n.y};
            const float r1 = dot(e, v3{fabsf(b1.x), fabsf(b1.y), fabsf(b1.z)});
            const float r2 = dot(e, v3{fabsf(b2.x), fabsf(b2.y), fabsf(b2.z)});
            const float d = __builtin_amdgcn_sqrtf(r1 * r1 + r2 * r2);
            const float s = sl * il;
            const float rho = __builtin_amdgcn_sqrtf(d * d + s * s);
            const float sc = CONE_MARGIN * (rho + l) * aux.w;
            const float t = d * sh - s * ch;
            const float g = l * (sh * s + ch * d) - a * rho;
            if (t >= sc || g >= sc * l) return true;
            if (t <= -sc && g <= -sc * l) return false;
        }
    }
    return cone_overlap_ref(cone, radius, o, lo, hi, dist_to_box);
}

struct SilRecord {
    v3 pa, pb, n0, n1;
    uint32_t index, one;
};

template <int DIM>
__device__ __forceinline__ SilRecord load_silhouette(const float4 *sils, uint32_t r)
{
    SilRecord s;
    if (DIM == 2) {
        const float4 a = sils[(size_t)r * 2], b = sils[(size_t)r * 2 + 1];
        s.pa = {a.x, a.y, 0.0f};
        s.pb = {0.0f, 0.0f, 0.0f};
        s.n0 = {a.z, a.w, 0.0f};
        s.n1 = {b.x, b.y, 0.0f};
        s.index = __float_as_uint(b.z);
        s.one = __float_as_uint(b.w);
    } else {
        const float4 *p = sils + (size_t)r * 4;
        const float4 a = p[0], b = p[1], c = p[2], d = p[3];
        s.pa = {a.x, a.y, a.z};
        s.pb = {b.x, b.y, b.z};
        s.n0 = {c.x, c.y, c.z};
        s.n1 = {d.x, d.y, d.z};
        s.index = __float_as_uint(a.w);
        s.one = __float_as_uint(b.w);
    }
    return s;
}

// isSilhouetteVertex / isSilhouetteEdge (geometry.slang:490-522,599-633)
template <int DIM>
__device__ __forceinline__ bool is_silhouette(const SilRecord &s, v3 view, float d, bool flip, float prec)
{
    const float sign = flip ? 1.0f : -1.0f;
    if (d <= prec) {
        if (DIM == 2) {
            const float det = s.n0.x * s.n1.y - s.n1.x * s.n0.y;
            return sign * det > prec;
        }
        v3 e = sub(s.pb, s.pa);
        const float l = sqrtf(dot(e, e));
        e = {e.x / l, e.y / l, e.z / l};
        const float angle = atan2f(dot(e, cross(s.n0, s.n1)), dot(s.n0, s.n1));
        return sign * angle > prec;
    }
    const v3 u = {view.x / d, view.y / d, view.z / d};
    const float dot0 = dot(u, s.n0), dot1 = dot(u, s.n1);
    if (fabsf(dot0) <= prec) return sign * dot1 > prec;
    if (fabsf(dot1) <= prec) return sign * dot0 > prec;
    return dot0 * dot1 < 0.0f;
}

// Vertex / Edge .findClosestSilhouettePoint (geometry.slang:555-590,667-701)
// The closest point p (segment parameter t) of a silhouette record to x; returns |x - p|.
template <int DIM>
__device__ __forceinline__ float silhouette_point(const SilRecord &s, v3 x, v3 &p, float &t)
{
    t = 0.0f;
    if (DIM == 2) {
        p = s.pa;
    } else {
        // findClosestPointLineSegment (geometry.slang:79-103) on float3
        const v3 u = sub(s.pb, s.pa), v = sub(x, s.pa);
        const float c1 = dot(u, v);
        if (c1 <= 0.0f) {
            p = s.pa;
        } else {
            const float c2 = dot(u, u);
            if (c2 <= c1) {
                p = s.pb;
                t = 1.0f;
            } else {
                t = c1 / c2;
                p = {s.pa.x + u.x * t, s.pa.y + u.y * t, s.pa.z + u.z * t};
            }
        }
    }
    const v3 view = sub(x, p);
    return sqrtf(dot(view, view));
}

template <int DIM>
__device__ __forceinline__ bool silhouette_test(const SilRecord &s, v3 x, float r2, bool flip, float r2min,
                                                float prec, v3 &p, float &t, float &d)
{
    if (r2min >= r2) return false;
    d = silhouette_point<DIM>(s, x, p, t);
    if (d * d > r2) return false;
    bool process = s.one == 1u;
    if (!process) process = is_silhouette<DIM>(s, sub(x, p), d, flip, prec);
    return process && d * d <= r2;
}

// The distance of a silhouette record to x when the record is on the silhouette
// seen from x, NaN otherwise: silhouette_test without its radius checks, which
// a NaN fails as the test's own rejection does.
template <int DIM>
__device__ __forceinline__ float silhouette_distance(const SilRecord &s, v3 x, bool flip, float prec)
{
    v3 p;
    float t;
    const float d = silhouette_point<DIM>(s, x, p, t);
    bool process = s.one == 1u;
    if (!process) process = is_silhouette<DIM>(s, sub(x, p), d, flip, prec);
    return process ? d : __builtin_nanf("");
}

// The children of record cur that the silhouette traversal visits (bvh.slang:580-627):
// a child with a valid cone whose box overlaps the sphere and whose cone passes
// BoundingCone.overlap (the reference's form, or cone_overlap_bounded, which
// decides the same); a0, b0 are the children's squared box distances.
template <int DIM, bool BOUNDED>
__device__ __forceinline__ void silhouette_children(const DevScene &S, const SilScene &T, uint32_t cur, v3 x,
                                                    typename Geo<DIM>::vec xb, float r2, bool &oA, bool &oB,
                                                    float &a0, float &b0, uint32_t &lA, uint32_t &lB)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    vec loA, hiA, loB, hiB;
    G::record(S, cur, loA, hiA, lA, loB, hiB, lB);
    const float4 *cp = T.cones + (size_t)cur * 4;
    const float4 cA = cp[0], rA = cp[1], cB = cp[2], rB = cp[3];
    float unused;
    a0 = 0.0f;
    b0 = 0.0f;
    oA = cA.w >= 0.0f;
    if (oA) {
        box_sqdist(loA, hiA, xb, a0, unused);
        oA = a0 <= r2;
        if (oA) oA = BOUNDED ? cone_overlap_bounded(cA, rA, x, to3(loA), to3(hiA), a0)
                             : cone_overlap(cA, rA.x, x, to3(loA), to3(hiA), a0);
    }
    oB = cB.w >= 0.0f;
    if (oB) {
        box_sqdist(loB, hiB, xb, b0, unused);
        oB = b0 <= r2;
        if (oB) oB = BOUNDED ? cone_overlap_bounded(cB, rB, x, to3(loB), to3(hiB), b0)
                             : cone_overlap(cB, rB.x, x, to3(loB), to3(hiB), b0);
    }
}

// Bvh.findClosestSilhouettePoint (bvh.slang:503-646) for one query per lane.
// A stack entry carries the child's squared box distance, which is what the
// reference recomputes when it pops the node.
template <int DIM, int BLOCK, int K>
__global__ __launch_bounds__(BLOCK) void k_closest_silhouettes(DevScene S, SilScene T, const float4 *spheres,
                                                               const uint32_t *flip, int64_t n, float r2min,
                                                               float prec, float2 *out, const uint32_t *order,
                                                               int xcds)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    const int64_t k = xcd_tile(xcds) * BLOCK + threadIdx.x;
    if (k >= n) return;
    const int64_t i = order ? (int64_t)order[k] : k;
    const float4 s4 = spheres[i];
    const v3 x = {s4.x, s4.y, DIM == 3 ? s4.z : 0.0f};
    const float q[3] = {s4.x, s4.y, s4.z};
    const vec xb = G::load_point(q);
    float r2 = s4.w;
    const bool fl = flip[i] != 0u;

    uint32_t best = FCPW_HIP_UINT_MAX;
    float best_d = FLT_MAX, best_t = 0.0f;
    v3 best_p = {0.0f, 0.0f, 0.0f};

    FCPW_STACK_DECL(BLOCK, K);
    vec lo, hi;
    uint32_t cur;
    G::root(S, lo, hi, cur);
    float cur_d2, unused;
    box_sqdist(lo, hi, xb, cur_d2, unused);
    if (!(r2min >= r2) && cur_d2 <= r2) {
        while (true) {
            if (cur_d2 <= r2) {
                if (cur & LEAF_BIT) {
                    const uint2 rng = T.sil_leaf[cur & OFF_MASK];
                    for (uint32_t j = 0; j < rng.y; j++) {
                        const SilRecord s = load_silhouette<DIM>(T.sils, rng.x + j);
                        if (s.index == best) continue;
                        v3 p;
                        float t, d;
                        if (silhouette_test<DIM>(s, x, r2, fl, r2min, prec, p, t, d)) {
                            r2 = fminf(r2, d * d);
                            best = s.index;
                            best_d = d;
                            best_p = p;
                            best_t = t;
                            if (r2min >= r2) break;
                        }
                    }
                } else {
                    uint32_t lA, lB;
                    float a0, b0;
                    bool oA, oB;
                    silhouette_children<DIM, false>(S, T, cur, x, xb, r2, oA, oB, a0, b0, lA, lB);
                    if (oA && oB) {
                        uint32_t nearL = lA, farL = lB;
                        float nearD = a0, farD = b0;
                        if (b0 < a0) { nearL = lB; farL = lA; nearD = b0; farD = a0; }
                        FCPW_PUSH(BLOCK, K, farL, farD, { atomicOr(S.flags, 1); break; });
                        cur = nearL;
                        cur_d2 = nearD;
                        continue;
                    } else if (oA) {
                        cur = lA;
                        cur_d2 = a0;
                        continue;
                    } else if (oB) {
                        cur = lB;
                        cur_d2 = b0;
                        continue;
                    }
                }
            }
            if (sp == 0) break;
            FCPW_POP(BLOCK, K, cur, cur_d2);
        }
    }
    const float p3[3] = {best_p.x, best_p.y, DIM == 3 ? best_p.z : 0.0f};
    const float nz[3] = {0.0f, 0.0f, 0.0f}, uv[2] = {best_t, 0.0f};
    write_interaction(out + i * 5, p3, nz, uv, best_d, best);
}

// ---------------------------------------------------------------------------
// Wave-cooperative closest silhouette points (one wave per workgroup).  Lanes
// descend as in k_closest_silhouettes; a lane that reaches a leaf with
// silhouette records holds it, and the wave evaluates the distance of every
// (lane, record) pair of the held leaves together (NaN for a record that is
// not on the silhouette seen from its lane's query).  Each holder then replays
// the reference's leaf loop in order: skip the record of the current best,
// nothing passes once r2min >= r2, accept when d*d <= r2.  Neither distance nor
// classification depends on r2, so the replay is exact; the winner's point is
// recomputed at the end.  Items are (lane << 10 | record within the leaf), so
// leaves hold at most 1023 records (the launch falls back to the per-lane
// kernel otherwise).  The cone tests take cone_overlap_bounded; the per-lane
// kernel keeps the reference's form, so equal results of the two kernels check
// one form against the other.
// ---------------------------------------------------------------------------
struct SilShared {
    uint16_t item[256];
    float d[256];
    uint32_t index[256];
};

template <int DIM, int K>
__global__ __launch_bounds__(64) FCPW_EIGHT_WAVES void k_closest_silhouettes_coop(DevScene S, SilScene T, const float4 *spheres,
                                                                 const uint32_t *flip, int64_t n, float r2min,
                                                                 float prec, float2 *out, const uint32_t *order,
                                                                 int xcds, int hold_min)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    constexpr int BLOCK = 64;
    __shared__ SilShared sh;
    const int lane = threadIdx.x & 63;
    const int64_t k = xcd_tile(xcds) * 64 + threadIdx.x;
    const bool valid = k < n;
    const int64_t i = valid ? (order ? (int64_t)order[k] : k) : 0;
    float4 s4 = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    bool fl = false;
    if (valid) {
        s4 = spheres[i];
        fl = flip[i] != 0u;
    }
    const v3 x = {s4.x, s4.y, DIM == 3 ? s4.z : 0.0f};
    const float q[3] = {s4.x, s4.y, s4.z};
    const vec xb = G::load_point(q);
    float r2 = s4.w;
    uint32_t best = FCPW_HIP_UINT_MAX;
    int32_t best_rec = -1;
    uint32_t cur, loff = 0, lcnt = 0;
    float cur_d2, unused;
    {
        vec lo, hi;
        G::root(S, lo, hi, cur);
        box_sqdist(lo, hi, xb, cur_d2, unused);
    }
    bool active = valid && !(r2min >= r2) && cur_d2 <= r2, live = active, holding = false, overflow = false;
    FCPW_STACK_DECL(BLOCK, K);
    __syncthreads();
    while (true) {
        // two traversal steps per wave iteration, as in closest_point_query_coop
        for (int rep = 0; rep < 2; rep++) if (active && !holding) {
            if (!(live && cur_d2 <= r2)) {
                live = false;
                while (sp > 0) {
                    FCPW_POP(BLOCK, K, cur, cur_d2);
                    if (cur_d2 <= r2) { live = true; break; }
                }
            }
            if (!live) {
                active = false;
            } else if (cur & LEAF_BIT) {
                const uint2 rng = T.sil_leaf[cur & OFF_MASK];
                loff = rng.x;
                lcnt = rng.y;
                holding = lcnt > 0;
                live = false;
            } else {
                uint32_t lA, lB;
                float a0, b0;
                bool oA, oB;
                silhouette_children<DIM, true>(S, T, cur, x, xb, r2, oA, oB, a0, b0, lA, lB);
                live = oA || oB;
                if (oA && oB) {
                    uint32_t nearL = lA, farL = lB;
                    float nearD = a0, farD = b0;
                    if (b0 < a0) { nearL = lB; farL = lA; nearD = b0; farD = a0; }
                    FCPW_PUSH(BLOCK, K, farL, farD, { overflow = true; active = false; live = false; });
                    cur = nearL;
                    cur_d2 = nearD;
                } else if (oA) {
                    cur = lA;
                    cur_d2 = a0;
                } else if (oB) {
                    cur = lB;
                    cur_d2 = b0;
                }
            }
        }
        const uint64_t act = __ballot(active);
        if (act == 0) break;
        const uint64_t hold = __ballot(holding);
        if (hold == 0 || ((int)__popcll(hold) < hold_min && hold != act)) continue;
        const uint32_t flu = fl ? 1u : 0u;
        for (uint32_t j0 = 0;; j0 += 4) {
            const bool h0 = holding && lcnt > j0, h1 = holding && lcnt > j0 + 1;
            const bool h2 = holding && lcnt > j0 + 2, h3 = holding && lcnt > j0 + 3;
            const uint64_t m0 = __ballot(h0), m1 = __ballot(h1), m2 = __ballot(h2), m3 = __ballot(h3);
            if (m0 == 0) break;
            const uint32_t c0 = __popcll(m0), c1 = __popcll(m1), c2 = __popcll(m2);
            const uint32_t total = c0 + c1 + c2 + __popcll(m3);
            const uint32_t i0 = lanes_below(m0), i1 = c0 + lanes_below(m1), i2 = c0 + c1 + lanes_below(m2),
                           i3 = c0 + c1 + c2 + lanes_below(m3);
            if (h0) sh.item[i0] = (uint16_t)((lane << 10) | j0);
            if (h1) sh.item[i1] = (uint16_t)((lane << 10) | (j0 + 1));
            if (h2) sh.item[i2] = (uint16_t)((lane << 10) | (j0 + 2));
            if (h3) sh.item[i3] = (uint16_t)((lane << 10) | (j0 + 3));
            __syncthreads();
            for (uint32_t pass = 0; pass < total; pass += 64) {
                const uint32_t kk = pass + (uint32_t)lane;
                // every lane takes part in the permutes; lanes past the end read item 0
                const uint32_t it = sh.item[kk < total ? kk : 0u];
                const int src = (int)(it >> 10);
                const v3 xo = {__shfl(x.x, src, 64), __shfl(x.y, src, 64), DIM == 3 ? __shfl(x.z, src, 64) : 0.0f};
                const bool fo = __shfl(flu, src, 64) != 0u;
                const uint32_t off = __shfl(loff, src, 64);
                if (kk < total) {
                    const SilRecord rec = load_silhouette<DIM>(T.sils, off + (it & 1023u));
                    sh.d[kk] = silhouette_distance<DIM>(rec, xo, fo, prec);
                    sh.index[kk] = rec.index;
                }
            }
            __syncthreads();
            if (holding) {
                const uint32_t idx[4] = {i0, i1, i2, i3};
                for (uint32_t jj = 0; jj < 4; jj++) {
                    if (lcnt > j0 + jj) {
                        const uint32_t index = sh.index[idx[jj]];
                        if (index == best || r2min >= r2) continue;
                        const float d = sh.d[idx[jj]];
                        if (d * d <= r2) {
                            r2 = d * d; // min(r2, d * d) when d * d <= r2
                            best = index;
                            best_rec = (int32_t)(loff + j0 + jj);
                        }
                    }
                }
            }
            __syncthreads();
            if (m3 == 0) break;
        }
        holding = false;
    }
    if (overflow) atomicOr(S.flags, 1);
    if (!valid) return;
    float p3[3] = {0.0f, 0.0f, 0.0f}, uv[2] = {0.0f, 0.0f}, d = FLT_MAX;
    const float nz[3] = {0.0f, 0.0f, 0.0f};
    if (best_rec >= 0) {
        v3 p;
        float t;
        d = silhouette_point<DIM>(load_silhouette<DIM>(T.sils, (uint32_t)best_rec), x, p, t);
        p3[0] = p.x;
        p3[1] = p.y;
        p3[2] = DIM == 3 ? p.z : 0.0f;
        uv[0] = t;
    }
    write_interaction(out + i * 5, p3, nz, uv, d, best);
}

// ---------------------------------------------------------------------------
// cone refit (Bvh.refit, bvh.slang:27-141): runs after k_refit has rewritten
// every box.  Leaves rebuild their cone from their silhouette records
// (refitLeafNode); the last child to arrive at a record merges the two child
// cones into the parent's (refitInternalNode + mergeBoundingCones,
// bounding-volumes.slang:262-318, with rotate at :238-260).
// ---------------------------------------------------------------------------
struct Cone {
    v3 axis;
    float half, radius;
};

// rotate (bounding-volumes.slang:238-260); w is length(cross(u, v)) in all
// three components, as the reference declares it
__device__ __forceinline__ v3 rotate_axis(v3 u, v3 v, float theta)
{
    const float c = cosf(theta), s = sinf(theta);
    const v3 cr = cross(u, v);
    const float L = sqrtf(dot(cr, cr));
    const v3 w = {L, L, L};
    const v3 o = {(1.0f - c) * w.x, (1.0f - c) * w.y, (1.0f - c) * w.z};
    const float R00 = c + o.x * w.x, R01 = o.y * w.x - s * w.z, R02 = o.z * w.x + s * w.y;
    const float R10 = o.x * w.y + s * w.z, R11 = c + o.y * w.y, R12 = o.z * w.y - s * w.x;
    const float R20 = o.x * w.z - s * w.y, R21 = o.y * w.z + s * w.x, R22 = c + o.z * w.z;
    return {(R00 * u.x + R01 * u.y) + R02 * u.z, (R10 * u.x + R11 * u.y) + R12 * u.z,
            (R20 * u.x + R21 * u.y) + R22 * u.z};
}

__device__ __forceinline__ Cone merge_cones(Cone A, Cone B, v3 oA, v3 oB, v3 oN)
{
    Cone c = {{0.0f, 0.0f, 0.0f}, PI_F, 0.0f};
    if (A.half >= 0.0f && B.half >= 0.0f) {
        v3 axA = A.axis, axB = B.axis;
        float hA = A.half, hB = B.half;
        const v3 dA = sub(oN, oA), dB = sub(oN, oB);
        c.radius = sqrtf(fmaxf(A.radius * A.radius + dot(dA, dA), B.radius * B.radius + dot(dB, dB)));
        if (hB > hA) {
            const v3 t = axA; axA = axB; axB = t;
            const float th = hA; hA = hB; hB = th;
        }
        const float theta = acosf(fmaxf(-1.0f, fminf(1.0f, dot(axA, axB))));
        if (fminf(theta + hB, PI_F) <= hA) {
            c.axis = axA;
            c.half = hA;
            return c;
        }
        const float o_theta = (hA + theta + hB) / 2.0f;
        if (o_theta >= PI_F) {
            c.axis = axA;
            return c;
        }
        c.axis = rotate_axis(axA, axB, o_theta - hA);
        c.half = o_theta;
    } else if (A.half >= 0.0f) {
        c = A;
    } else if (B.half >= 0.0f) {
        c = B;
    } else {
        c.half = -PI_F;
    }
    return c;
}

__device__ __forceinline__ Cone load_cone(const float4 *cones, int64_t k)
{
    const float4 a = cones[k * 2], b = cones[k * 2 + 1];
    return {{a.x, a.y, a.z}, a.w, b.x};
}
__device__ __forceinline__ void store_cone(float4 *cones, int64_t k, const Cone &c)
{
    cones[k * 2] = make_float4(c.axis.x, c.axis.y, c.axis.z, c.half);
    cones[k * 2 + 1] = cone_aux(c.axis.x, c.axis.y, c.axis.z, c.half, c.radius);
}

// box of child slot ref (rec * 2 + slot) or of the root (ref < 0), z = 0 in 2D
template <int DIM>
__device__ __forceinline__ void slot_box(const float4 *records, const float4 *root, int32_t ref, v3 &lo, v3 &hi)
{
    if (DIM == 3) {
        const float4 *p = ref < 0 ? root : records + (size_t)(ref >> 1) * 4 + (ref & 1) * 2;
        const float4 a = p[0], b = p[1];
        lo = {a.x, a.y, a.z};
        hi = {b.x, b.y, b.z};
    } else {
        const float4 a = ref < 0 ? root[0] : records[(size_t)(ref >> 1) * 3 + (ref & 1)];
        lo = {a.x, a.y, 0.0f};
        hi = {a.z, a.w, 0.0f};
    }
}

template <int DIM>
__global__ void k_refit_cones(const float4 *records, const float4 *root, float4 *cones, const uint2 *sil_leaf,
                              const float4 *sils, const int32_t *leaf_refs, const uint32_t *leaf_links,
                              const int32_t *rec_parent, unsigned int *counters, int n_leaves, int n_records)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_leaves) return;
    const uint2 rng = sil_leaf[leaf_links[j] & OFF_MASK];
    int32_t ref = leaf_refs[j];
    v3 lo, hi;
    slot_box<DIM>(records, root, ref, lo, hi);
    // refitLeafNode (bvh.slang:27-95)
    Cone c = {{0.0f, 0.0f, 0.0f}, 0.0f, 0.0f};
    const v3 centroid = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    bool any = false, two = true;
    for (uint32_t k = 0; k < rng.y; k++) {
        const SilRecord s = load_silhouette<DIM>(sils, rng.x + k);
        c.axis = {c.axis.x + s.n0.x, c.axis.y + s.n0.y, c.axis.z + s.n0.z};
        c.axis = {c.axis.x + s.n1.x, c.axis.y + s.n1.y, c.axis.z + s.n1.z};
        const v3 sc = DIM == 2 ? s.pa : v3{0.5f * (s.pa.x + s.pb.x), 0.5f * (s.pa.y + s.pb.y), 0.5f * (s.pa.z + s.pb.z)};
        const v3 dv = sub(sc, centroid);
        c.radius = fmaxf(c.radius, sqrtf(dot(dv, dv)));
        two = two && s.one == 0u;
        any = true;
    }
    if (!any) {
        c.half = -PI_F;
    } else if (!two) {
        c.half = PI_F;
    } else {
        const float axis_norm = sqrtf(dot(c.axis, c.axis));
        if (axis_norm > FLT_EPSILON) {
            c.axis = {c.axis.x / axis_norm, c.axis.y / axis_norm, c.axis.z / axis_norm};
            for (uint32_t k = 0; k < rng.y; k++) {
                const SilRecord s = load_silhouette<DIM>(sils, rng.x + k);
                c.half = fmaxf(c.half, acosf(fmaxf(-1.0f, fminf(1.0f, dot(c.axis, s.n0)))));
                c.half = fmaxf(c.half, acosf(fmaxf(-1.0f, fminf(1.0f, dot(c.axis, s.n1)))));
            }
        }
    }
    while (true) {
        store_cone(cones, ref < 0 ? 2 * (int64_t)n_records : ref, c);
        if (ref < 0) return;
        const int32_t rec = ref >> 1;
        __threadfence();
        if (atomicAdd(&counters[rec], 1u) == 0u) return;
        counters[rec] = 0u;
        __threadfence();
        // refitInternalNode (bvh.slang:97-126)
        v3 loA, hiA, loB, hiB;
        slot_box<DIM>(records, root, rec * 2, loA, hiA);
        slot_box<DIM>(records, root, rec * 2 + 1, loB, hiB);
        const v3 loM = {fminf(loA.x, loB.x), fminf(loA.y, loB.y), fminf(loA.z, loB.z)};
        const v3 hiM = {fmaxf(hiA.x, hiB.x), fmaxf(hiA.y, hiB.y), fmaxf(hiA.z, hiB.z)};
        const v3 cA = {0.5f * (loA.x + hiA.x), 0.5f * (loA.y + hiA.y), 0.5f * (loA.z + hiA.z)};
        const v3 cB = {0.5f * (loB.x + hiB.x), 0.5f * (loB.y + hiB.y), 0.5f * (loB.z + hiB.z)};
        const v3 cM = {0.5f * (loM.x + hiM.x), 0.5f * (loM.y + hiM.y), 0.5f * (loM.z + hiM.z)};
        c = merge_cones(load_cone(cones, rec * 2), load_cone(cones, rec * 2 + 1), cA, cB, cM);
        ref = rec_parent[rec];
    }
}

} // namespace fcpw_hip
