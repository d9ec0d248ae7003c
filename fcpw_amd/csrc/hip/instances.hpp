// instances.hpp -- scenes of several objects and transformed instances.
// Included once, by fcpw_hip.hip, after batch.hpp.
//
// The reference GPU path serves one object (fcpw_gpu.h:17-20); its CPU Scene
// puts a BVH over the objects' aggregates and wraps instanced objects in
// TransformedAggregate (fcpw.inl:1052-1126, primitive.h:289-440), whose GPU
// twin is aggregate.slang:45-145 with transform.slang:1-56.  Here:
//   objects  contexts of their own, each with its uploaded tree
//   inst     one DevInstance per instance: rows of t and tinv, object, flag
//   table    one DevScene per object (refreshed when an object is re-uploaded)
//   tree     the instance tree in the scene context's records / root, leaf
//            slot j holding instance prim_index[j]
// A query walks the instance tree in the reference GPU order and, at each
// instance of a reached leaf, runs the object's traversal on the query mapped
// by tinv, maps r^2 (or tMax) back by t and, on a hit, maps the point and
// normal and recomputes d = |p - x| (TransformedAggregate.findClosestPoint /
// intersect).  An object without instance transforms is traversed unmapped.
// Point mapping is mul(t, float4(p, 1)) with each row summed left to right.
#pragma once

namespace fcpw_hip {

__device__ __forceinline__ v3 xf(const float4 *r, v3 p)
{
    return {((r[0].x * p.x + r[0].y * p.y) + r[0].z * p.z) + r[0].w,
            ((r[1].x * p.x + r[1].y * p.y) + r[1].z * p.z) + r[1].w,
            ((r[2].x * p.x + r[2].y * p.y) + r[2].z * p.z) + r[2].w};
}
__device__ __forceinline__ v2 xf(const float4 *r, v2 p)
{
    return {(r[0].x * p.x + r[0].y * p.y) + r[0].w, (r[1].x * p.x + r[1].y * p.y) + r[1].w};
}
__device__ __forceinline__ v3 shift_x(v3 c, float s) { return {c.x + s, c.y + 0.0f, c.z + 0.0f}; }
__device__ __forceinline__ v2 shift_x(v2 c, float s) { return {c.x + s, c.y + 0.0f}; }
__device__ __forceinline__ v3 axpy(v3 o, v3 d, float s) { return {o.x + d.x * s, o.y + d.y * s, o.z + d.z * s}; }
__device__ __forceinline__ v2 axpy(v2 o, v2 d, float s) { return {o.x + d.x * s, o.y + d.y * s}; }
__device__ __forceinline__ v3 divs(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ v2 divs(v2 a, float s) { return {a.x / s, a.y / s}; }
__device__ __forceinline__ v3 recip(v3 a) { return {1.0f / a.x, 1.0f / a.y, 1.0f / a.z}; }
__device__ __forceinline__ v2 recip(v2 a) { return {1.0f / a.x, 1.0f / a.y}; }

// transformSphere's radius (transform.slang:31-42): sphere (c, r2), tc = xf(c)
template <typename V>
__device__ __forceinline__ float xf_r2(const float4 *r, V c, V tc, float r2)
{
    if (!(r2 < FLT_MAX)) return FLT_MAX;
    V dd = sub(xf(r, shift_x(c, sqrtf(r2))), tc);
    return dot(dd, dd);
}

// normalize(transpose(linear(tinv)) * n), as v / length(v) (transform.slang:48-49)
template <int DIM>
__device__ __forceinline__ void xf_normal(const float4 *ti, float n[3])
{
    if (DIM == 3) {
        v3 m = {(ti[0].x * n[0] + ti[1].x * n[1]) + ti[2].x * n[2], (ti[0].y * n[0] + ti[1].y * n[1]) + ti[2].y * n[2],
                (ti[0].z * n[0] + ti[1].z * n[1]) + ti[2].z * n[2]};
        float l = sqrtf(dot(m, m));
        n[0] = m.x / l; n[1] = m.y / l; n[2] = m.z / l;
    } else {
        v2 m = {ti[0].x * n[0] + ti[1].x * n[1], ti[0].y * n[0] + ti[1].y * n[1]};
        float l = sqrtf(dot(m, m));
        n[0] = m.x / l; n[1] = m.y / l; n[2] = 0.0f;
    }
}

// the inverse rows and flags of an instance; the forward rows are read when
// needed (after the object's traversal), so they are not held in registers
// across it
__device__ __forceinline__ void load_instance(const DevInstance *inst, int i, float4 ti[3], int &object,
                                              int &transformed)
{
    const DevInstance *I = inst + i;
    for (int r = 0; r < 3; r++) ti[r] = I->ti[r];
    object = I->object;
    transformed = I->transformed;
}
__device__ __forceinline__ void load_forward(const DevInstance *inst, int i, float4 t[3])
{
    const DevInstance *I = inst + i;
    for (int r = 0; r < 3; r++) t[r] = I->t[r];
}

// TransformedAggregate.findClosestPoint on one instance; r2 is updated in place
template <int DIM, int BLOCK, int K>
__device__ __forceinline__ bool instance_closest(const DevInstance *inst, const DevScene *objs, int ii,
                                                 typename Geo<DIM>::vec x, float &r2, int &object, int32_t &slot,
                                                 float &d, float p[3], float uv[2])
{
    using vec = typename Geo<DIM>::vec;
    float4 t[3], ti[3];
    int transformed;
    load_instance(inst, ii, ti, object, transformed);
    const DevScene S = global_scene(objs[object]);
    vec xl = x;
    float r2l = r2;
    if (transformed) {
        xl = xf(ti, x);
        r2l = xf_r2(ti, x, xl, r2);
    }
    CpResult c;
    closest_point_query<DIM, BLOCK, K>(S, xl, r2l, c);
    if (transformed) load_forward(inst, ii, t);
    r2 = transformed ? xf_r2(t, xl, xf(t, xl), r2l) : r2l;
    if (c.slot < 0) return false;
    slot = c.slot;
    uv[0] = c.uv[0];
    uv[1] = c.uv[1];
    if (transformed) {
        vec pw = xf(t, Geo<DIM>::load_point(c.p));
        vec e = sub(pw, x);
        d = sqrtf(dot(e, e));
        store_p(pw, p);
    } else {
        d = c.d;
        p[0] = c.p[0]; p[1] = c.p[1]; p[2] = c.p[2];
    }
    return true;
}

// Closest points on a scene of instances run as two launches over the queries'
// Morton order when "inst_seed" = B > 0.  The order is cut into blocks of B
// queries; the first launch answers, per block, the 64 consecutive queries in
// its middle (the whole block when it is shorter: its representatives), with
// their own radii and one coherent wave per block.  The second launch answers
// the other queries, each with its squared radius lowered to
//   U = ((d_rep + |x - x_rep|) (1 + 1e-3) + 1e-4 (1 + max_k |x_k|))^2,
// x_rep being the representative nearest in the order and d_rep its reported
// distance (no lowering when that query found nothing).  The records are those
// of the reference's walk:
//  - a lower starting radius never adds an acceptance.  Along the reference's
//    visit order the lowered walk's radius stays at or below the reference's
//    (it skips only nodes whose box distance exceeds its own radius, and those
//    nodes' bounds and primitives could not lower it further), and every
//    acceptance test is monotone in the radius: the object walk's radius is the
//    world radius mapped by xf_r2, whose rounded steps are each monotone;
//  - the reference reports its last acceptance W, so the lowered walk reports W
//    when it still accepts W, i.e. when U maps to at least W's squared distance
//    in its object's space.  By the triangle inequality a point of the scene lies
//    within d_rep + |x - x_rep| of x (up to the rounding of the mapped points,
//    a few units in the last place of the coordinates), so W's distance, which
//    is the smallest to that rounding, is below the unmargined bound; the 1e-3
//    relative and 1e-4 (1 + |x|) absolute margins exceed that rounding and the
//    rounding of xf_r2 by orders of magnitude.
// A lowered radius prunes most of the object walks that the reference's
// box-ordered instance walk starts with loose radii.
//
// Block b's representatives: positions [r0, r0 + nr) of the order.
__device__ __forceinline__ void seed_reps(int64_t b, int block, int64_t n, int64_t &r0, int64_t &nr)
{
    const int64_t len = std::min<int64_t>(block, n - b * block);
    nr = std::min<int64_t>(64, len);
    r0 = b * block + (len - nr) / 2;
}

// The ordered position lane k answers in launch `pass` (false: nothing to answer).
__device__ __forceinline__ bool seed_position(int64_t k, int64_t n_launch, int64_t n, int block, int pass,
                                              int64_t &pos)
{
    pos = k;
    if (k >= n_launch) return false;
    if (!block) return true;
    int64_t r0, nr;
    if (pass == 0) {
        seed_reps(k >> 6, block, n, r0, nr);
        pos = r0 + (k & 63);
        return (k & 63) < nr;
    }
    seed_reps(k / block, block, n, r0, nr);
    return k < r0 || k >= r0 + nr;
}

template <int DIM>
__device__ __forceinline__ float seeded_r2(const float4 *spheres, const float2 *out, const uint32_t *order,
                                           int64_t pos, int64_t n, int block, float r2, typename Geo<DIM>::vec x)
{
    int64_t r0, nr;
    seed_reps(pos / block, block, n, r0, nr);
    const uint32_t rep = order[std::min(std::max(pos, r0), r0 + nr - 1)];
    const float dr = out[(size_t)rep * 5 + 4].x;
    if (!(dr < FLT_MAX)) return r2;
    const float4 sr = spheres[rep];
    float q[3];
    store_p(x, q);
    const float e0 = q[0] - sr.x, e1 = q[1] - sr.y, e2 = DIM == 3 ? q[2] - sr.z : 0.0f;
    const float m = fmaxf(fmaxf(fabsf(q[0]), fabsf(q[1])), DIM == 3 ? fabsf(q[2]) : 0.0f);
    const float b = (dr + sqrtf(e0 * e0 + e1 * e1 + e2 * e2)) * 1.001f + 1e-4f * (1.0f + m);
    return fminf(r2, b * b);
}

// __launch_bounds__(BLOCK, 6): six waves per SIMD (80 VGPRs); measured 36 ms
// against 42 ms unbounded (113 VGPRs, four waves) on 256 instances
template <int DIM, int BLOCK, int K, int KT>
__global__ __launch_bounds__(BLOCK, 6) void k_instances_closest(DevScene T, const DevInstance *inst, const DevScene *objs,
                                                             int n_inst, const float4 *spheres, int64_t n,
                                                             int record_normals, float2 *out, int32_t *object_out,
                                                             const uint32_t *order, int xcds, int64_t n_all,
                                                             int block, int pass)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    const int64_t k = xcd_tile(xcds) * BLOCK + threadIdx.x;
    int64_t pos;
    if (!seed_position(k, n, n_all, block, pass, pos)) return;
    const int64_t i = order ? (int64_t)order[pos] : pos;
    const float4 s4 = spheres[i];
    const float q[3] = {s4.x, s4.y, s4.z};
    const vec x = G::load_point(q);
    float r2 = s4.w;
    if (block && pass == 1) r2 = seeded_r2<DIM>(spheres, out, order, pos, n_all, block, r2, x);
    int best_inst = -1, best_obj = -1;
    int32_t best_slot = -1;
    float best_d = FLT_MAX, bp[3] = {0.0f, 0.0f, 0.0f}, buv[2] = {0.0f, 0.0f};
    int obj;
    int32_t slot;
    float d, p[3], uv[2];
    if (n_inst == 1) {
        // a single instance is the scene's aggregate itself (fcpw.inl:1088-1092)
        const int ii = T.prim_index[0];
        if (instance_closest<DIM, BLOCK, K>(inst, objs, ii, x, r2, obj, slot, d, p, uv)) {
            best_inst = ii; best_obj = obj; best_slot = slot; best_d = d;
            bp[0] = p[0]; bp[1] = p[1]; bp[2] = p[2];
            buv[0] = uv[0]; buv[1] = uv[1];
        }
    } else {
        vec lo, hi;
        uint32_t cur;
        G::root(T, lo, hi, cur);
        float d2min, d2max;
        box_sqdist(lo, hi, x, d2min, d2max);
        if (d2min <= r2) {
            r2 = fminf(r2, d2max);
            float cur_d2 = d2min;
            FCPW_STACK_DECL(BLOCK, KT);
            while (true) {
                if (cur_d2 <= r2) {
                    if (cur & LEAF_BIT) {
                        const uint32_t off = cur & OFF_MASK, cnt = ((cur >> CNT_SHIFT) & CNT_MASK) + 1u;
                        for (uint32_t j = 0; j < cnt; j++) {
                            const int ii = T.prim_index[off + j];
                            if (instance_closest<DIM, BLOCK, K>(inst, objs, ii, x, r2, obj, slot, d, p, uv)) {
                                r2 = fminf(r2, d * d);
                                best_inst = ii; best_obj = obj; best_slot = slot; best_d = d;
                                bp[0] = p[0]; bp[1] = p[1]; bp[2] = p[2];
                                buv[0] = uv[0]; buv[1] = uv[1];
                            }
                        }
                    } else {
                        vec loA, hiA, loB, hiB;
                        uint32_t lA, lB;
                        G::record(T, cur, loA, hiA, lA, loB, hiB, lB);
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
                            FCPW_PUSH(BLOCK, KT, farL, farD, { atomicOr(T.flags, 1); break; });
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
                FCPW_POP(BLOCK, KT, cur, cur_d2);
            }
        }
    }
    float nrm[3] = {0.0f, 0.0f, 0.0f};
    uint32_t index = FCPW_HIP_UINT_MAX;
    if (best_inst >= 0) {
        const DevScene S = global_scene(objs[best_obj]);
        index = (uint32_t)S.prim_index[best_slot];
        if (record_normals) {
            G::normal(S, (uint32_t)best_slot, nrm);
            if (inst[best_inst].transformed) xf_normal<DIM>(inst[best_inst].ti, nrm);
        }
    }
    write_interaction(out + i * 5, bp, nrm, buv, best_d, index);
    if (object_out) object_out[i] = best_obj;
}

// ---------------------------------------------------------------------------
// Wave-cooperative closest points on a scene of instances.  Each lane walks the
// instance tree and, per instance of a reached leaf, the object's tree on the
// mapped query, with one stack for both levels (an object's entries sit above
// the index `base` where its walk began).  Object leaves are evaluated by the
// whole wave as in k_closest_points_coop: every (lane, primitive) pair of the
// held leaves, from the holder's object and mapped query, then each holder
// replays d*d <= r2 in leaf order.  Instances are visited in the order of
// k_instances_closest and the winner is recomputed from its instance and slot,
// so the records are that kernel's bit for bit.
// ---------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ float closest_prim(const float4 *prims, uint32_t slot, typename Geo<DIM>::vec x,
                                              typename Geo<DIM>::vec &p, float &u0, float &u1)
{
    DevScene S = {};
    S.prims = prims;
    return Geo<DIM>::closest(S, slot, x, p, u0, u1);
}

template <int DIM, int K>
__global__ __launch_bounds__(64) void k_instances_closest_coop(DevScene T, const DevInstance *inst,
                                                               const DevScene *objs, int n_inst,
                                                               const float4 *spheres, int64_t n, int record_normals,
                                                               float2 *out, int32_t *object_out,
                                                               const uint32_t *order, int xcds, int hold_min,
                                                               int64_t n_all, int block, int pass)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    constexpr int BLOCK = 64;
    __shared__ CoopShared sh;
    const int lane = threadIdx.x & 63;
    const int64_t k = xcd_tile(xcds) * 64 + threadIdx.x;
    int64_t pos;
    const bool valid = seed_position(k, n, n_all, block, pass, pos);
    const int64_t i = valid ? (order ? (int64_t)order[pos] : pos) : 0;
    float q[3] = {0.0f, 0.0f, 0.0f};
    float r2 = -1.0f;
    if (valid) {
        const float4 s4 = spheres[i];
        q[0] = s4.x; q[1] = s4.y; q[2] = s4.z;
        r2 = s4.w;
    }
    const vec x = G::load_point(q);
    if (valid && block && pass == 1) r2 = seeded_r2<DIM>(spheres, out, order, pos, n_all, block, r2, x);
    int best_inst = -1;
    int32_t best_slot = -1;
    // traversal state of either level
    uint32_t cur = 0;
    float cur_d2 = 0.0f;
    bool live = false, active = false, holding = false, start = false;
    bool inner = false;                              // walking an object's tree
    uint32_t leaf_off = 0, leaf_cnt = 0, leaf_j = 0; // the instance leaf being visited
    int base = 0;
    // the instance being visited
    int ii = 0, transformed = 0;
    vec xl = x;
    float r2l = 0.0f;
    DevScene O = {};
    uint32_t loff = 0, lcnt = 0;
    int32_t lbest = -1;
    bool overflow = false;
    FCPW_STACK_DECL(BLOCK, K);
    if (valid) {
        if (n_inst == 1) {
            // a single instance is the scene's aggregate itself (fcpw.inl:1088-1092)
            active = start = true;
            leaf_off = 0;
            leaf_cnt = 1;
        } else {
            vec lo, hi;
            G::root(T, lo, hi, cur);
            float d2min, d2max;
            box_sqdist(lo, hi, x, d2min, d2max);
            if (d2min <= r2) {
                r2 = fminf(r2, d2max);
                cur_d2 = d2min;
                active = live = true;
            }
        }
    }
    __syncthreads();
    while (true) {
        // two traversal steps per wave iteration, as in closest_point_query_coop
        for (int rep = 0; rep < 2; rep++) if (active && !holding) {
            if (start) {
                // TransformedAggregate.findClosestPoint: map the query into the object's space
                start = false;
                ii = T.prim_index[leaf_off + leaf_j];
                float4 ti[3];
                int object;
                load_instance(inst, ii, ti, object, transformed);
                O = global_scene(objs[object]);
                xl = x;
                r2l = r2;
                if (transformed) {
                    xl = xf(ti, x);
                    r2l = xf_r2(ti, x, xl, r2);
                }
                inner = true;
                base = sp;
                lbest = -1;
                vec lo, hi;
                G::root(O, lo, hi, cur);
                float d2min, d2max;
                box_sqdist(lo, hi, xl, d2min, d2max);
                live = d2min <= r2l;
                if (live) {
                    r2l = fminf(r2l, d2max);
                    cur_d2 = d2min;
                }
            }
            // the first step into the object is taken in the iteration that mapped the query
            {
                const float lim = inner ? r2l : r2;
                if (!(live && cur_d2 <= lim)) {
                    live = false;
                    while (sp > base) {
                        FCPW_POP(BLOCK, K, cur, cur_d2);
                        if (cur_d2 <= lim) { live = true; break; }
                    }
                }
                if (!live) {
                    if (inner) {
                        // the object's walk is over: map its radius and result back
                        float4 t[3];
                        if (transformed) load_forward(inst, ii, t);
                        r2 = transformed ? xf_r2(t, xl, xf(t, xl), r2l) : r2l;
                        if (lbest >= 0) {
                            vec pl;
                            float u0, u1;
                            float d = closest_prim<DIM>(O.prims, (uint32_t)lbest, xl, pl, u0, u1);
                            if (transformed) {
                                const vec e = sub(xf(t, pl), x);
                                d = sqrtf(dot(e, e));
                            }
                            r2 = fminf(r2, d * d);
                            best_inst = ii;
                            best_slot = lbest;
                        }
                        inner = false;
                        base = 0;
                        if (++leaf_j < leaf_cnt) start = true;
                        else if (n_inst == 1) active = false;
                    } else {
                        active = false;
                    }
                } else if (cur & LEAF_BIT) {
                    if (inner) {
                        holding = true;
                        loff = cur & OFF_MASK;
                        lcnt = ((cur >> CNT_SHIFT) & CNT_MASK) + 1u;
                    } else {
                        leaf_off = cur & OFF_MASK;
                        leaf_cnt = ((cur >> CNT_SHIFT) & CNT_MASK) + 1u;
                        leaf_j = 0;
                        start = true;
                    }
                    live = false;
                } else {
                    DevScene R = {};
                    R.records = inner ? O.records : T.records;
                    const vec xq = inner ? xl : x;
                    float rr = inner ? r2l : r2;
                    vec loA, hiA, loB, hiB;
                    uint32_t lA, lB;
                    G::record(R, cur, loA, hiA, lA, loB, hiB, lB);
                    float a0, a1, b0, b1;
                    box_sqdist(loA, hiA, xq, a0, a1);
                    const bool oA = a0 <= rr;
                    rr = fminf(rr, a1);
                    box_sqdist(loB, hiB, xq, b0, b1);
                    const bool oB = b0 <= rr;
                    rr = fminf(rr, b1);
                    if (inner) r2l = rr;
                    else r2 = rr;
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
        }
        const uint64_t act = __ballot(active);
        if (act == 0) break;
        const uint64_t hold = __ballot(holding);
        if (hold == 0 || ((int)__popcll(hold) < hold_min && hold != act)) continue;
        // the leaf pass over the held object leaves
        float xs[3];
        store_p(xl, xs);
        const uint64_t prims_ptr = (uint64_t)(uintptr_t)O.prims;
        const uint32_t plo = (uint32_t)prims_ptr, phi = (uint32_t)(prims_ptr >> 32);
        for (uint32_t j0 = 0; j0 < (uint32_t)MAX_LEAF_PRIMS; j0 += 4) {
            const bool h0 = holding && lcnt > j0, h1 = holding && lcnt > j0 + 1;
            const bool h2 = holding && lcnt > j0 + 2, h3 = holding && lcnt > j0 + 3;
            const uint64_t m0 = __ballot(h0), m1 = __ballot(h1), m2 = __ballot(h2), m3 = __ballot(h3);
            if (m0 == 0) break;
            const uint32_t c0 = __popcll(m0), c1 = __popcll(m1), c2 = __popcll(m2);
            const uint32_t total = c0 + c1 + c2 + __popcll(m3);
            const uint32_t i0 = lanes_below(m0), i1 = c0 + lanes_below(m1), i2 = c0 + c1 + lanes_below(m2),
                           i3 = c0 + c1 + c2 + lanes_below(m3);
            if (h0) sh.item[i0] = (uint16_t)((lane << 4) | j0);
            if (h1) sh.item[i1] = (uint16_t)((lane << 4) | (j0 + 1));
            if (h2) sh.item[i2] = (uint16_t)((lane << 4) | (j0 + 2));
            if (h3) sh.item[i3] = (uint16_t)((lane << 4) | (j0 + 3));
            __syncthreads();
            for (uint32_t pass = 0; pass < total; pass += 64) {
                const uint32_t kk = pass + (uint32_t)lane;
                // every lane takes part in the permutes; lanes past the end read item 0
                const uint32_t it = sh.item[kk < total ? kk : 0u];
                const int src = (int)(it >> 4);
                float xo[3];
                xo[0] = __shfl(xs[0], src, 64);
                xo[1] = __shfl(xs[1], src, 64);
                xo[2] = DIM == 3 ? __shfl(xs[2], src, 64) : 0.0f;
                const uint32_t off = __shfl(loff, src, 64);
                const uint32_t lo32 = __shfl(plo, src, 64), hi32 = __shfl(phi, src, 64);
                if (kk < total) {
                    const float4 *pr = (const float4 *)(uintptr_t)(((uint64_t)hi32 << 32) | lo32);
                    vec pp;
                    float u0, u1;
                    sh.d[kk] = closest_prim<DIM>(pr, off + (it & 15u), G::load_point(xo), pp, u0, u1);
                }
            }
            __syncthreads();
            if (holding) {
                const uint32_t idx[4] = {i0, i1, i2, i3};
                for (uint32_t jj = 0; jj < 4; jj++) {
                    if (lcnt > j0 + jj) {
                        const float d = sh.d[idx[jj]];
                        if (d * d <= r2l) {
                            r2l = fminf(r2l, d * d);
                            lbest = (int32_t)(loff + j0 + jj);
                        }
                    }
                }
            }
            __syncthreads();
            if (m3 == 0) break;
        }
        holding = false;
    }
    if (overflow) atomicOr(T.flags, 1);
    if (!valid) return;
    float bp[3] = {0.0f, 0.0f, 0.0f}, buv[2] = {0.0f, 0.0f}, nrm[3] = {0.0f, 0.0f, 0.0f};
    float best_d = FLT_MAX;
    uint32_t index = FCPW_HIP_UINT_MAX;
    int best_obj = -1;
    if (best_inst >= 0) {
        float4 ti[3];
        int tr;
        load_instance(inst, best_inst, ti, best_obj, tr);
        const DevScene S = global_scene(objs[best_obj]);
        const vec xb = tr ? xf(ti, x) : x;
        vec pl;
        float u0, u1;
        best_d = G::closest(S, (uint32_t)best_slot, xb, pl, u0, u1);
        buv[0] = u0;
        buv[1] = u1;
        if (tr) {
            float4 t[3];
            load_forward(inst, best_inst, t);
            const vec pw = xf(t, pl);
            const vec e = sub(pw, x);
            best_d = sqrtf(dot(e, e));
            store_p(pw, bp);
        } else {
            store_p(pl, bp);
        }
        index = (uint32_t)S.prim_index[best_slot];
        if (record_normals) {
            G::normal(S, (uint32_t)best_slot, nrm);
            if (tr) xf_normal<DIM>(ti, nrm);
        }
    }
    write_interaction(out + i * 5, bp, nrm, buv, best_d, index);
    if (object_out) object_out[i] = best_obj;
}

// TransformedAggregate.intersect on one instance; tmax is updated in place
template <int DIM, int BLOCK, int K>
__device__ __forceinline__ bool instance_ray(const DevInstance *inst, const DevScene *objs, int ii,
                                             typename Geo<DIM>::vec o, typename Geo<DIM>::vec dir,
                                             typename Geo<DIM>::vec dinv, float &tmax, bool occlusion, int &object,
                                             RayResult &r)
{
    using vec = typename Geo<DIM>::vec;
    float4 t[3], ti[3];
    int transformed;
    load_instance(inst, ii, ti, object, transformed);
    const DevScene S = global_scene(objs[object]);
    vec ol = o, dl = dir, il = dinv;
    float tl = tmax;
    if (transformed) {
        // transformRay(tinv, r) (transform.slang:22-29) and the Ray constructor's 1/d
        ol = xf(ti, o);
        vec m = sub(xf(ti, axpy(o, dir, tmax < FLT_MAX ? tmax : 1.0f)), ol);
        float nrm = sqrtf(dot(m, m));
        dl = divs(m, nrm);
        il = recip(dl);
        tl = tmax < FLT_MAX ? nrm : FLT_MAX;
    }
    ray_query<DIM, BLOCK, K>(S, ol, dl, il, tl, occlusion, r);
    if (transformed) {
        load_forward(inst, ii, t);
        load_instance(inst, ii, ti, object, transformed);
        if (tl < FLT_MAX) {
            vec dd = sub(xf(t, axpy(ol, dl, tl)), xf(t, ol));
            tmax = sqrtf(dot(dd, dd));
        } else {
            tmax = FLT_MAX;
        }
    } else {
        tmax = tl;
    }
    if (r.slot < 0) return false;
    if (transformed && !occlusion) {
        vec pw = xf(t, Geo<DIM>::load_point(r.p));
        vec e = sub(pw, o);
        r.t = sqrtf(dot(e, e));
        store_p(pw, r.p);
        xf_normal<DIM>(ti, r.n);
    }
    return true;
}

template <int DIM, int BLOCK, int K, int KT>
__global__ __launch_bounds__(BLOCK) void k_instances_intersect(DevScene T, const DevInstance *inst,
                                                               const DevScene *objs, int n_inst, const float2 *rays,
                                                               int64_t n, int occlusion, float2 *out,
                                                               int32_t *object_out, const uint32_t *order, int xcds)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    const int64_t k = xcd_tile(xcds) * BLOCK + threadIdx.x;
    if (k >= n) return;
    const int64_t i = order ? (int64_t)order[k] : k;
    const float2 *rr = rays + i * 5; // GPURay: o(3) d(3) dinv(3) tmax
    const float2 r0 = rr[0], r1 = rr[1], r2 = rr[2], r3 = rr[3], r4 = rr[4];
    const float o3[3] = {r0.x, r0.y, r1.x}, d3[3] = {r1.y, r2.x, r2.y}, di3[3] = {r3.x, r3.y, r4.x};
    float tmax = r4.y;
    const vec o = G::load_point(o3), dir = G::load_point(d3), di = G::load_point(di3);
    RayResult best, r;
    best.slot = -1;
    best.t = FLT_MAX;
    for (int c = 0; c < 3; c++) { best.p[c] = 0.0f; best.n[c] = 0.0f; }
    best.uv[0] = best.uv[1] = 0.0f;
    int best_obj = -1, obj;
    const bool occl = occlusion != 0;
    if (n_inst == 1) {
        if (instance_ray<DIM, BLOCK, K>(inst, objs, T.prim_index[0], o, dir, di, tmax, occl, obj, r)) {
            best = r;
            best_obj = obj;
        }
    } else {
        vec lo, hi;
        uint32_t cur;
        G::root(T, lo, hi, cur);
        float cur_t;
        if (box_ray(lo, hi, o, dir, di, tmax, cur_t)) {
            FCPW_STACK_DECL(BLOCK, KT);
            bool done = false;
            while (!done) {
                if (!(cur_t > tmax)) {
                    if (cur & LEAF_BIT) {
                        const uint32_t off = cur & OFF_MASK, cnt = ((cur >> CNT_SHIFT) & CNT_MASK) + 1u;
                        for (uint32_t j = 0; j < cnt; j++) {
                            if (instance_ray<DIM, BLOCK, K>(inst, objs, T.prim_index[off + j], o, dir, di, tmax,
                                                            occl, obj, r)) {
                                best = r;
                                best_obj = obj;
                                if (occl) { done = true; break; }
                                tmax = fminf(tmax, r.t);
                            }
                        }
                        if (done) break;
                    } else {
                        vec loA, hiA, loB, hiB;
                        uint32_t lA, lB;
                        G::record(T, cur, loA, hiA, lA, loB, hiB, lB);
                        float ta, tb;
                        bool hA = box_ray(loA, hiA, o, dir, di, tmax, ta);
                        bool hB = box_ray(loB, hiB, o, dir, di, tmax, tb);
                        if (hA && hB) {
                            uint32_t nearL = lA, farL = lB;
                            float nearT = ta, farT = tb;
                            if (tb < ta) { nearL = lB; farL = lA; nearT = tb; farT = ta; }
                            FCPW_PUSH(BLOCK, KT, farL, farT, { atomicOr(T.flags, 1); break; });
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
                FCPW_POP(BLOCK, KT, cur, cur_t);
            }
        }
    }
    uint32_t index = FCPW_HIP_UINT_MAX;
    if (best.slot >= 0) index = (uint32_t)global_scene(objs[best_obj]).prim_index[best.slot];
    write_interaction(out + i * 5, best.p, best.n, best.uv, best.t, index);
    if (object_out) object_out[i] = best_obj;
}

// ---------------------------------------------------------------------------
// Wave-cooperative rays on a scene of instances: the structure of
// k_instances_closest_coop (one stack for both levels) with the ray leaf pass
// of k_intersect_coop on the mapped rays.  The winner is recomputed from its
// instance, its slot and the world tMax at the start of its instance's visit
// (the mapped direction depends on it, transform.slang:22-29).
// ---------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ void map_ray(const float4 ti[3], typename Geo<DIM>::vec o, typename Geo<DIM>::vec dir,
                                        float tmax, typename Geo<DIM>::vec &ol, typename Geo<DIM>::vec &dl,
                                        typename Geo<DIM>::vec &il, float &tl, float &nrm)
{
    // transformRay(tinv, r) (transform.slang:22-29) and the Ray constructor's 1/d
    ol = xf(ti, o);
    const typename Geo<DIM>::vec m = sub(xf(ti, axpy(o, dir, tmax < FLT_MAX ? tmax : 1.0f)), ol);
    nrm = sqrtf(dot(m, m));
    dl = divs(m, nrm);
    il = recip(dl);
    tl = tmax < FLT_MAX ? nrm : FLT_MAX;
}

template <int DIM, int K>
__global__ __launch_bounds__(64) void k_instances_intersect_coop(DevScene T, const DevInstance *inst,
                                                                 const DevScene *objs, int n_inst, const float2 *rays,
                                                                 int64_t n, int occlusion, float2 *out,
                                                                 int32_t *object_out, const uint32_t *order, int xcds,
                                                                 int hold_min)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    constexpr int BLOCK = 64;
    __shared__ CoopShared sh;
    const int lane = threadIdx.x & 63;
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
    const vec o = G::load_point(o3), dir = G::load_point(d3), di = G::load_point(di3);
    const bool occl = occlusion != 0;
    int best_inst = -1;
    int32_t best_slot = -1;
    float best_tv = 0.0f; // world tMax when the winner's instance was entered
    uint32_t cur = 0;
    float cur_t = 0.0f;
    bool live = false, active = false, holding = false, start = false, inner = false;
    uint32_t leaf_off = 0, leaf_cnt = 0, leaf_j = 0;
    int base = 0, ii = 0, transformed = 0;
    vec ol = o, dl = dir, il = di;
    float tl = 0.0f, tv = 0.0f;
    DevScene O = {};
    uint32_t loff = 0, lcnt = 0;
    int32_t lbest = -1;
    bool overflow = false;
    FCPW_STACK_DECL(BLOCK, K);
    if (valid) {
        if (n_inst == 1) {
            // a single instance is the scene's aggregate itself (fcpw.inl:1088-1092)
            active = start = true;
            leaf_cnt = 1;
        } else {
            vec lo, hi;
            G::root(T, lo, hi, cur);
            active = live = box_ray(lo, hi, o, dir, di, tmax, cur_t);
        }
    }
    __syncthreads();
    while (true) {
        // two traversal steps per wave iteration, as in closest_point_query_coop
        for (int rep = 0; rep < 2; rep++) if (active && !holding) {
            if (start) {
                // TransformedAggregate.intersect: map the ray into the object's space
                start = false;
                ii = T.prim_index[leaf_off + leaf_j];
                float4 ti[3];
                int object;
                load_instance(inst, ii, ti, object, transformed);
                O = global_scene(objs[object]);
                tv = tmax;
                ol = o;
                dl = dir;
                il = di;
                tl = tmax;
                if (transformed) {
                    float nrm;
                    map_ray<DIM>(ti, o, dir, tmax, ol, dl, il, tl, nrm);
                }
                inner = true;
                base = sp;
                lbest = -1;
                vec lo, hi;
                G::root(O, lo, hi, cur);
                live = box_ray(lo, hi, ol, dl, il, tl, cur_t);
            }
            {
                const float lim = inner ? tl : tmax;
                if (!(live && !(cur_t > lim))) {
                    live = false;
                    while (sp > base) {
                        FCPW_POP(BLOCK, K, cur, cur_t);
                        if (!(cur_t > lim)) { live = true; break; }
                    }
                }
                if (!live) {
                    if (inner) {
                        // the object's walk is over: map its tMax and hit back
                        float4 t[3];
                        if (transformed) {
                            load_forward(inst, ii, t);
                            if (tl < FLT_MAX) {
                                const vec dd = sub(xf(t, axpy(ol, dl, tl)), xf(t, ol));
                                tmax = sqrtf(dot(dd, dd));
                            } else {
                                tmax = FLT_MAX;
                            }
                        } else {
                            tmax = tl;
                        }
                        if (lbest >= 0) {
                            RayResult r;
                            RayGeo<DIM>::hit(O, (uint32_t)lbest, ol, dl, __builtin_inff(), r, true);
                            float tw = r.t;
                            if (transformed) {
                                const vec e = sub(xf(t, G::load_point(r.p)), o);
                                tw = sqrtf(dot(e, e));
                            }
                            tmax = fminf(tmax, tw);
                            best_inst = ii;
                            best_slot = lbest;
                            best_tv = tv;
                        }
                        inner = false;
                        base = 0;
                        if (++leaf_j < leaf_cnt) start = true;
                        else if (n_inst == 1) active = false;
                    } else {
                        active = false;
                    }
                } else if (cur & LEAF_BIT) {
                    if (inner) {
                        holding = true;
                        loff = cur & OFF_MASK;
                        lcnt = ((cur >> CNT_SHIFT) & CNT_MASK) + 1u;
                    } else {
                        leaf_off = cur & OFF_MASK;
                        leaf_cnt = ((cur >> CNT_SHIFT) & CNT_MASK) + 1u;
                        leaf_j = 0;
                        start = true;
                    }
                    live = false;
                } else {
                    DevScene R = {};
                    R.records = inner ? O.records : T.records;
                    const vec oq = inner ? ol : o, dq = inner ? dl : dir, iq = inner ? il : di;
                    vec loA, hiA, loB, hiB;
                    uint32_t lA, lB;
                    G::record(R, cur, loA, hiA, lA, loB, hiB, lB);
                    float ta, tb;
                    const bool hA = box_ray(loA, hiA, oq, dq, iq, lim, ta);
                    const bool hB = box_ray(loB, hiB, oq, dq, iq, lim, tb);
                    live = hA || hB;
                    if (hA && hB) {
                        uint32_t nearL = lA, farL = lB;
                        float nearT = ta, farT = tb;
                        if (tb < ta) { nearL = lB; farL = lA; nearT = tb; farT = ta; }
                        FCPW_PUSH(BLOCK, K, farL, farT, { overflow = true; active = false; live = false; });
                        cur = nearL;
                        cur_t = nearT;
                    } else if (hA) {
                        cur = lA;
                        cur_t = ta;
                    } else if (hB) {
                        cur = lB;
                        cur_t = tb;
                    }
                }
            }
        }
        const uint64_t act = __ballot(active);
        if (act == 0) break;
        const uint64_t hold = __ballot(holding);
        if (hold == 0 || ((int)__popcll(hold) < hold_min && hold != act)) continue;
        // the leaf pass over the held object leaves
        float os[3], ds[3];
        store_p(ol, os);
        store_p(dl, ds);
        const uint64_t prims_ptr = (uint64_t)(uintptr_t)O.prims;
        const uint32_t plo = (uint32_t)prims_ptr, phi = (uint32_t)(prims_ptr >> 32);
        for (uint32_t j0 = 0; j0 < (uint32_t)MAX_LEAF_PRIMS; j0 += 4) {
            const bool h0 = holding && lcnt > j0, h1 = holding && lcnt > j0 + 1;
            const bool h2 = holding && lcnt > j0 + 2, h3 = holding && lcnt > j0 + 3;
            const uint64_t m0 = __ballot(h0), m1 = __ballot(h1), m2 = __ballot(h2), m3 = __ballot(h3);
            if (m0 == 0) break;
            const uint32_t c0 = __popcll(m0), c1 = __popcll(m1), c2 = __popcll(m2);
            const uint32_t total = c0 + c1 + c2 + __popcll(m3);
            const uint32_t i0 = lanes_below(m0), i1 = c0 + lanes_below(m1), i2 = c0 + c1 + lanes_below(m2),
                           i3 = c0 + c1 + c2 + lanes_below(m3);
            if (h0) sh.item[i0] = (uint16_t)((lane << 4) | j0);
            if (h1) sh.item[i1] = (uint16_t)((lane << 4) | (j0 + 1));
            if (h2) sh.item[i2] = (uint16_t)((lane << 4) | (j0 + 2));
            if (h3) sh.item[i3] = (uint16_t)((lane << 4) | (j0 + 3));
            __syncthreads();
            for (uint32_t pass = 0; pass < total; pass += 64) {
                const uint32_t kk = pass + (uint32_t)lane;
                // every lane takes part in the permutes; lanes past the end read item 0
                const uint32_t it = sh.item[kk < total ? kk : 0u];
                const int src = (int)(it >> 4);
                float oo[3], dd[3];
                oo[0] = __shfl(os[0], src, 64);
                oo[1] = __shfl(os[1], src, 64);
                oo[2] = DIM == 3 ? __shfl(os[2], src, 64) : 0.0f;
                dd[0] = __shfl(ds[0], src, 64);
                dd[1] = __shfl(ds[1], src, 64);
                dd[2] = DIM == 3 ? __shfl(ds[2], src, 64) : 0.0f;
                const uint32_t off = __shfl(loff, src, 64);
                const uint32_t lo32 = __shfl(plo, src, 64), hi32 = __shfl(phi, src, 64);
                if (kk < total) {
                    DevScene P = {};
                    P.prims = (const float4 *)(uintptr_t)(((uint64_t)hi32 << 32) | lo32);
                    sh.d[kk] = ray_t<DIM>(P, off + (it & 15u), G::load_point(oo), G::load_point(dd));
                }
            }
            __syncthreads();
            if (holding) {
                const uint32_t idx[4] = {i0, i1, i2, i3};
                for (uint32_t jj = 0; jj < 4; jj++) {
                    if (lcnt > j0 + jj) {
                        const float t = sh.d[idx[jj]];
                        if (t >= 0.0f && t <= tl) {
                            lbest = (int32_t)(loff + j0 + jj);
                            if (occl) {
                                // the first occluder ends the query
                                best_inst = ii;
                                best_slot = lbest;
                                lcnt = 0;
                                active = false;
                                break;
                            }
                            tl = fminf(tl, t);
                        }
                    }
                }
            }
            __syncthreads();
            if (m3 == 0) break;
        }
        holding = false;
    }
    if (overflow) atomicOr(T.flags, 1);
    if (!valid) return;
    float p[3] = {0.0f, 0.0f, 0.0f}, nv[3] = {0.0f, 0.0f, 0.0f}, uv[2] = {0.0f, 0.0f}, tw = FLT_MAX;
    uint32_t index = FCPW_HIP_UINT_MAX;
    int best_obj = -1;
    if (best_inst >= 0) {
        float4 ti[3];
        int tr;
        load_instance(inst, best_inst, ti, best_obj, tr);
        const DevScene S = global_scene(objs[best_obj]);
        index = (uint32_t)S.prim_index[best_slot];
        if (!occl) {
            vec ob = o, db = dir, ib = di;
            float tb = best_tv, nrm;
            if (tr) map_ray<DIM>(ti, o, dir, best_tv, ob, db, ib, tb, nrm);
            RayResult r;
            RayGeo<DIM>::hit(S, (uint32_t)best_slot, ob, db, __builtin_inff(), r, true);
            tw = r.t;
            p[0] = r.p[0]; p[1] = r.p[1]; p[2] = r.p[2];
            nv[0] = r.n[0]; nv[1] = r.n[1]; nv[2] = r.n[2];
            uv[0] = r.uv[0]; uv[1] = r.uv[1];
            if (tr) {
                float4 t[3];
                load_forward(inst, best_inst, t);
                const vec pw = xf(t, G::load_point(r.p));
                const vec e = sub(pw, o);
                tw = sqrtf(dot(e, e));
                store_p(pw, p);
                xf_normal<DIM>(ti, nv);
            }
        }
    }
    write_interaction(out + i * 5, p, nv, uv, tw, index);
    if (object_out) object_out[i] = best_obj;
}

// TransformedAggregate.intersect(sphere) (aggregate.slang:81-97) on one
// instance: the sphere mapped by tinv, the object's walk and point sample with
// (u, r1, r2s), the sampled point and normal mapped back; d (the pdf) is kept
template <int DIM>
__device__ __forceinline__ bool instance_sphere(const DevInstance *inst, const DevScene *objs, int ii,
                                                typename Geo<DIM>::vec x, float r2, float u, float r1, float r2s,
                                                int &object, int32_t &slot, float p[3], float n[3], float uv[2],
                                                float &d)
{
    using vec = typename Geo<DIM>::vec;
    float4 t[3], ti[3];
    int transformed;
    load_instance(inst, ii, ti, object, transformed);
    const DevScene S = global_scene(objs[object]);
    vec xl = x;
    float r2l = r2;
    if (transformed) {
        xl = xf(ti, x);
        r2l = xf_r2(ti, x, xl, r2);
    }
    slot = sphere_sample<DIM>(S, xl, r2l, u, r1, r2s, p, n, uv, d);
    if (slot < 0) return false;
    if (transformed) {
        load_forward(inst, ii, t);
        store_p(xf(t, Geo<DIM>::load_point(p)), p);
        xf_normal<DIM>(ti, n);
    }
    return true;
}

// Sphere intersection over the instances: the random walk of
// Bvh::processSubtreeForIntersection over aggregates (bvh.inl:1140-1250) with
// the constant branch weight -- children overlapping the sphere chosen with
// equal probability, then, in the leaf, each instance queried with the walk's
// current u and the instances that find a primitive chosen uniformly
// (reservoir weight 1); d = the instance's pdf x the walk's pdf / the number
// of instances that found one.
template <int DIM>
__global__ __launch_bounds__(128) void k_instances_spheres(DevScene T, const DevInstance *inst, const DevScene *objs,
                                                           int n_inst, const float4 *spheres, const float *rnd,
                                                           int64_t n, float2 *out, int32_t *object_out)
{
    using G = Geo<DIM>;
    using vec = typename G::vec;
    const int64_t i = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (i >= n) return;
    const float4 s4 = spheres[i];
    const float q[3] = {s4.x, s4.y, s4.z};
    const vec x = G::load_point(q);
    const float r2 = s4.w, r1 = rnd[i * 3 + 1], r2s = rnd[i * 3 + 2];
    float u = rnd[i * 3];
    float bp[3] = {0.0f, 0.0f, 0.0f}, bn[3] = {0.0f, 0.0f, 0.0f}, buv[2] = {0.0f, 0.0f}, bd = FLT_MAX;
    int best_obj = -1, obj;
    int32_t best_slot = -1, slot;
    float p[3], nn[3], uv[2], d;
    if (n_inst == 1) {
        if (instance_sphere<DIM>(inst, objs, T.prim_index[0], x, r2, u, r1, r2s, obj, slot, p, nn, uv, d)) {
            best_obj = obj; best_slot = slot; bd = d;
            for (int c = 0; c < 3; c++) { bp[c] = p[c]; bn[c] = nn[c]; }
            buv[0] = uv[0]; buv[1] = uv[1];
        }
    } else {
        vec lo, hi;
        uint32_t cur;
        G::root(T, lo, hi, cur);
        float d0, d1;
        box_sqdist(lo, hi, x, d0, d1);
        if (d0 <= r2) {
            float pdf = 1.0f;
            while (true) {
                if (cur & LEAF_BIT) {
                    const uint32_t off = cur & OFF_MASK, cnt = ((cur >> CNT_SHIFT) & CNT_MASK) + 1u;
                    float total = 0.0f;
                    for (uint32_t k = 0; k < cnt; k++) {
                        if (instance_sphere<DIM>(inst, objs, T.prim_index[off + k], x, r2, u, r1, r2s, obj, slot, p,
                                                 nn, uv, d)) {
                            total += 1.0f;
                            const float prob = 1.0f / total;
                            if (u < prob) {
                                u = u / prob;
                                best_obj = obj; best_slot = slot; bd = d * pdf;
                                for (int c = 0; c < 3; c++) { bp[c] = p[c]; bn[c] = nn[c]; }
                                buv[0] = uv[0]; buv[1] = uv[1];
                            } else {
                                u = (u - prob) / (1.0f - prob);
                            }
                        }
                    }
                    if (total > 0.0f) bd = bd / total;
                    break;
                }
                vec loA, hiA, loB, hiB;
                uint32_t lA, lB;
                G::record(T, cur, loA, hiA, lA, loB, hiB, lB);
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
                } else {
                    u = (u - pl) / pr;
                    cur = lB;
                    pdf *= pr;
                }
            }
        }
    }
    const uint32_t index = best_slot >= 0 ? (uint32_t)objs[best_obj].prim_index[best_slot] : FCPW_HIP_UINT_MAX;
    write_interaction(out + i * 5, bp, bn, buv, bd, index);
    if (object_out) object_out[i] = best_obj;
}

// Bvh::refit over the instances (bvh.inl:548-599) once the objects are refit:
// each leaf bounds its instances (root box of the object, or its 2^DIM
// corners mapped by t and expanded with +-FLT_EPSILON), then the boxes are
// merged bottom-up exactly as k_refit does
template <int DIM>
__global__ void k_instances_refit(float4 *records, float4 *root, const DevInstance *inst, const DevScene *objs,
                                  const int32_t *leaf_inst, const int32_t *leaf_refs, const uint32_t *leaf_links,
                                  const int32_t *rec_parent, unsigned int *counters, int n_leaves)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_leaves) return;
    uint32_t link = leaf_links[j];
    uint32_t off = link & OFF_MASK, cnt = ((link >> CNT_SHIFT) & CNT_MASK) + 1u;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t k = 0; k < cnt; k++) {
        const DevInstance *I = inst + leaf_inst[off + k];
        const float4 *rb = objs[I->object].root;
        float blo[3], bhi[3];
        if (DIM == 3) {
            float4 a = rb[0], b = rb[1];
            blo[0] = a.x; blo[1] = a.y; blo[2] = a.z;
            bhi[0] = b.x; bhi[1] = b.y; bhi[2] = b.z;
        } else {
            float4 a = rb[0];
            blo[0] = a.x; blo[1] = a.y; blo[2] = 0.0f;
            bhi[0] = a.z; bhi[1] = a.w; bhi[2] = 0.0f;
        }
        if (I->transformed) {
            float4 t[3] = {I->t[0], I->t[1], I->t[2]};
            float tlo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, thi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            for (int c = 0; c < (1 << DIM); c++) {
                float q[3];
                if (DIM == 3) {
                    v3 pc = {(c & 1) ? bhi[0] : blo[0], (c & 2) ? bhi[1] : blo[1], (c & 4) ? bhi[2] : blo[2]};
                    v3 m = xf(t, pc);
                    q[0] = m.x; q[1] = m.y; q[2] = m.z;
                } else {
                    v2 pc = {(c & 1) ? bhi[0] : blo[0], (c & 2) ? bhi[1] : blo[1]};
                    v2 m = xf(t, pc);
                    q[0] = m.x; q[1] = m.y; q[2] = 0.0f;
                }
                for (int a = 0; a < DIM; a++) {
                    tlo[a] = fminf(tlo[a], q[a] - FLT_EPSILON);
                    thi[a] = fmaxf(thi[a], q[a] + FLT_EPSILON);
                }
            }
            for (int a = 0; a < DIM; a++) { blo[a] = tlo[a]; bhi[a] = thi[a]; }
        }
        for (int a = 0; a < DIM; a++) {
            lo[a] = fminf(lo[a], blo[a]);
            hi[a] = fmaxf(hi[a], bhi[a]);
        }
    }
    refit_upward<DIM>(records, root, rec_parent, counters, leaf_refs[j], lo, hi);
}

// object index of the hits of a one-object scene: 0, or -1 on a miss
__global__ void k_single_object_index(const float2 *out, int64_t n, int32_t *object_out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    object_out[i] = __float_as_uint(out[i * 5 + 4].y) == FCPW_HIP_UINT_MAX ? -1 : 0;
}

} // namespace fcpw_hip

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
constexpr int INST_BLOCK = 128, INST_K = 8, INST_KT = 4;
// LDS stack entries of the 3D cooperative closest-point kernel, whose single stack holds both levels
// (measured: 29.3 ms against 31.7 with 8 for 4M queries over 256 instances)
constexpr int INST_COOP_K3 = 12;

// the object table, rebuilt when a member object was re-uploaded
static int instances_ready(fcpw_hip_ctx *c, const char *fn)
{
    bool dirty = false;
    for (size_t o = 0; o < c->inst_objects.size(); o++) {
        fcpw_hip_ctx *m = c->inst_objects[o];
        if (!m->uploaded || m->instanced || m->dim != c->dim)
            FAIL(c, FCPW_HIP_ERR_NOT_UPLOADED, "%s: object %zu has no uploaded single-object BVH", fn, o);
        if (m->generation != c->inst_generations[o]) dirty = true;
    }
    if (!dirty) return FCPW_HIP_OK;
    std::vector<DevScene> table;
    int depth = 0;
    for (size_t o = 0; o < c->inst_objects.size(); o++) {
        DevScene S = dev_scene(c->inst_objects[o]);
        S.flags = c->flags; // overflow of any level is reported by the scene
        depth = std::max(depth, c->inst_objects[o]->max_depth);
        table.push_back(S);
        c->inst_generations[o] = c->inst_objects[o]->generation;
    }
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, hipMemcpy(c->obj_table, table.data(), table.size() * sizeof(DevScene), hipMemcpyHostToDevice));
    c->inst_obj_depth = depth;
    return FCPW_HIP_OK;
}

static int instances_refit(fcpw_hip_ctx *c, hipStream_t s)
{
    int st = instances_ready(c, "refit");
    if (st) return st;
    std::vector<fcpw_hip_ctx *> done;
    for (fcpw_hip_ctx *m : c->inst_objects) {
        if (std::find(done.begin(), done.end(), m) != done.end()) continue;
        done.push_back(m);
        st = fcpw_hip_refit(m, s);
        if (st) FAIL(c, st, "refit (object): %s", m->err.c_str());
    }
    int threads = 256, blocks = (c->n_leaves + threads - 1) / threads;
    if (c->dim == 3)
        hipLaunchKernelGGL(k_instances_refit<3>, dim3(blocks), dim3(threads), 0, s, c->records, c->root, c->inst,
                           c->obj_table, c->prim_index, c->leaf_refs, c->leaf_links, c->rec_parent, c->counters,
                           c->n_leaves);
    else
        hipLaunchKernelGGL(k_instances_refit<2>, dim3(blocks), dim3(threads), 0, s, c->records, c->root, c->inst,
                           c->obj_table, c->prim_index, c->leaf_refs, c->leaf_links, c->rec_parent, c->counters,
                           c->n_leaves);
    HIPCHECK(c, hipGetLastError());
    return FCPW_HIP_OK;
}

static int launch_instances_closest(fcpw_hip_ctx *c, const void *spheres, int64_t n, int record_normals, void *out,
                                    int32_t *object_out, hipStream_t s)
{
    int st = instances_ready(c, "closest_points");
    if (st) return st;
    const uint32_t *order;
    st = morton_order(c, (const float *)spheres, 4, n, s, &order);
    if (st) return st;
    if ((st = timing_begin(c, s))) return st;
    DevScene T = dev_scene(c);
    // with ordered queries and inst_seed = B: each block's representatives, then
    // the rest with radii lowered from them (seeded_r2); otherwise one launch
    const int block = order && c->inst_seed > 0 ? c->inst_seed : 0;
    const bool coop = c->coop && c->max_depth + c->inst_obj_depth < STACK_TOTAL;
    for (int pass = block ? 0 : 1; pass < 2; pass++) {
        const int64_t m = pass == 0 ? (n + block - 1) / block * 64 : n;
        if (coop) {
            // one stack serves both levels, so the two trees' depths must fit it together
            const dim3 g = grid_for(c, m, 64);
#define FCPW_ICOOP(D, K)                                                                                       \
    hipLaunchKernelGGL((k_instances_closest_coop<D, K>), g, dim3(64), 0, s, T, c->inst, c->obj_table,          \
                       c->n_instances, (const float4 *)spheres, m, record_normals, (float2 *)out, object_out, \
                       order, xcd_arg(c), c->hold_min, n, block, pass)
            if (c->dim == 3 && c->inst_stack == 8) FCPW_ICOOP(3, 8);
            else if (c->dim == 3) FCPW_ICOOP(3, INST_COOP_K3);
            else FCPW_ICOOP(2, INST_K);
#undef FCPW_ICOOP
        } else {
            const dim3 g = grid_for(c, m, INST_BLOCK);
            if (c->dim == 3)
                hipLaunchKernelGGL((k_instances_closest<3, INST_BLOCK, INST_K, INST_KT>), g, dim3(INST_BLOCK), 0, s, T,
                                   c->inst, c->obj_table, c->n_instances, (const float4 *)spheres, m, record_normals,
                                   (float2 *)out, object_out, order, xcd_arg(c), n, block, pass);
            else
                hipLaunchKernelGGL((k_instances_closest<2, INST_BLOCK, INST_K, INST_KT>), g, dim3(INST_BLOCK), 0, s, T,
                                   c->inst, c->obj_table, c->n_instances, (const float4 *)spheres, m, record_normals,
                                   (float2 *)out, object_out, order, xcd_arg(c), n, block, pass);
        }
        HIPCHECK(c, hipGetLastError());
    }
    return timing_end(c, s);
}

static int launch_instances_intersect(fcpw_hip_ctx *c, const void *rays, int64_t n, int occlusion, void *out,
                                      int32_t *object_out, hipStream_t s)
{
    int st = instances_ready(c, "intersect");
    if (st) return st;
    const uint32_t *order;
    // ordered by origin, and by direction octant when ray_sort_dir is set
    st = morton_order(c, (const float *)rays, 10, n, s, &order, c->ray_sort_dir ? 3 : -1);
    if (st) return st;
    if ((st = timing_begin(c, s))) return st;
    DevScene T = dev_scene(c);
    if (c->ray_coop && c->max_depth + c->inst_obj_depth < STACK_TOTAL) {
        // one stack serves both levels, so the two trees' depths must fit it together
        const dim3 g = grid_for(c, n, 64);
        if (c->dim == 3)
            hipLaunchKernelGGL((k_instances_intersect_coop<3, INST_K>), g, dim3(64), 0, s, T, c->inst, c->obj_table,
                               c->n_instances, (const float2 *)rays, n, occlusion, (float2 *)out, object_out, order,
                               xcd_arg(c), c->ray_hold_min);
        else
            hipLaunchKernelGGL((k_instances_intersect_coop<2, INST_K>), g, dim3(64), 0, s, T, c->inst, c->obj_table,
                               c->n_instances, (const float2 *)rays, n, occlusion, (float2 *)out, object_out, order,
                               xcd_arg(c), c->ray_hold_min);
        HIPCHECK(c, hipGetLastError());
        return timing_end(c, s);
    }
    dim3 g = grid_for(c, n, INST_BLOCK);
    if (c->dim == 3)
        hipLaunchKernelGGL((k_instances_intersect<3, INST_BLOCK, INST_K, INST_KT>), g, dim3(INST_BLOCK), 0, s, T,
                           c->inst, c->obj_table, c->n_instances, (const float2 *)rays, n, occlusion, (float2 *)out,
                           object_out, order, xcd_arg(c));
    else
        hipLaunchKernelGGL((k_instances_intersect<2, INST_BLOCK, INST_K, INST_KT>), g, dim3(INST_BLOCK), 0, s, T,
                           c->inst, c->obj_table, c->n_instances, (const float2 *)rays, n, occlusion, (float2 *)out,
                           object_out, order, xcd_arg(c));
    HIPCHECK(c, hipGetLastError());
    return timing_end(c, s);
}

static int launch_instances_spheres(fcpw_hip_ctx *c, const void *spheres, const float *rnd, int64_t n, void *out,
                                    int32_t *object_out, hipStream_t s)
{
    int st = instances_ready(c, "intersect_spheres");
    if (st) return st;
    DevScene T = dev_scene(c);
    const unsigned blocks = (unsigned)((n + 127) / 128);
    if (c->dim == 3)
        hipLaunchKernelGGL(k_instances_spheres<3>, dim3(blocks), dim3(128), 0, s, T, c->inst, c->obj_table,
                           c->n_instances, (const float4 *)spheres, rnd, n, (float2 *)out, object_out);
    else
        hipLaunchKernelGGL(k_instances_spheres<2>, dim3(blocks), dim3(128), 0, s, T, c->inst, c->obj_table,
                           c->n_instances, (const float4 *)spheres, rnd, n, (float2 *)out, object_out);
    HIPCHECK(c, hipGetLastError());
    return FCPW_HIP_OK;
}

static int fill_single_object_index(fcpw_hip_ctx *c, const void *out, int64_t n, int32_t *object_out, hipStream_t s)
{
    if (!object_out) return FCPW_HIP_OK;
    hipLaunchKernelGGL(k_single_object_index, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const float2 *)out,
                       n, object_out);
    HIPCHECK(c, hipGetLastError());
    return FCPW_HIP_OK;
}

extern "C" {

int fcpw_hip_upload_instances(fcpw_hip_ctx *c, fcpw_hip_ctx *const *objects, int n_objects,
                              const fcpw_hip_instance *instances, int n_instances, int n_nodes, const float *pmin,
                              const float *pmax, const int32_t *nrefs, const int32_t *offset,
                              const int32_t *leaf_instances)
{
    if (!c) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    if (n_objects <= 0 || n_instances <= 0 || n_nodes <= 0 || !objects || !instances || !pmin || !pmax || !nrefs ||
        !offset || !leaf_instances)
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_instances: empty scene or NULL array");
    const int dim = objects[0] ? objects[0]->dim : 0;
    for (int o = 0; o < n_objects; o++) {
        fcpw_hip_ctx *m = objects[o];
        if (!m || m == c || !m->uploaded || m->instanced)
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_instances: object %d has no uploaded single-object BVH", o);
        if (m->dim != dim || m->device != c->device)
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_instances: objects must share dimension and device");
    }
    std::vector<int> seen(n_instances, 0);
    for (int i = 0; i < n_instances; i++) {
        if (instances[i].object < 0 || instances[i].object >= n_objects)
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_instances: instance %d draws object %d of %d", i,
                 instances[i].object, n_objects);
        int li = leaf_instances[i];
        if (li < 0 || li >= n_instances || seen[li]++)
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_instances: leaf_instances is not a permutation");
    }
    TreeRecords T;
    int st = tree_records(c, "upload_instances", dim, n_nodes, pmin, pmax, nrefs, offset, n_instances, T);
    if (st) return st;
    std::vector<DevInstance> inst((size_t)n_instances);
    for (int i = 0; i < n_instances; i++) {
        const fcpw_hip_instance &I = instances[i];
        DevInstance &D = inst[i];
        for (int r = 0; r < 3; r++) {
            D.t[r] = make_float4(I.t[r * 4], I.t[r * 4 + 1], I.t[r * 4 + 2], I.t[r * 4 + 3]);
            D.ti[r] = make_float4(I.tinv[r * 4], I.tinv[r * 4 + 1], I.tinv[r * 4 + 2], I.tinv[r * 4 + 3]);
        }
        D.object = I.object;
        D.transformed = I.transformed ? 1 : 0;
        D.pad0 = D.pad1 = 0;
    }
    HIPCHECK(c, hipSetDevice(c->device));
    free_scene(c);
    st = upload_tree(c, T, leaf_instances, n_instances);
    if (st) return st;
    HIPCHECK(c, hipMalloc((void **)&c->inst, inst.size() * sizeof(DevInstance)));
    HIPCHECK(c, hipMemcpy(c->inst, inst.data(), inst.size() * sizeof(DevInstance), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMalloc((void **)&c->obj_table, (size_t)n_objects * sizeof(DevScene)));
    c->inst_objects.assign(objects, objects + n_objects);
    c->inst_generations.assign((size_t)n_objects, ~(uint64_t)0);
    set_query_grid(c, pmin, pmax, dim);
    c->dim = dim;
    c->n_records = T.n_rec;
    c->n_leaves = T.n_leaf;
    c->n_prims = n_instances;
    c->n_vertices = 0;
    c->max_depth = T.max_depth;
    c->node_slot = std::move(T.node_slot);
    c->node_leaf_off = std::move(T.node_leaf_off);
    c->n_instances = n_instances;
    c->instanced = true;
    c->uploaded = true;
    c->generation++;
    st = instances_ready(c, "upload_instances");
    if (st) return st;
    HIPCHECK(c, hipDeviceSynchronize());
    return FCPW_HIP_OK;
}

int fcpw_hip_closest_points_objects_device(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres, int64_t n,
                                           int record_normals, fcpw_hip_interaction *out, int32_t *object_index,
                                           void *stream)
{
    int st = check_ready(c, "closest_points_device");
    if (st) return st;
    if (n < 0 || (n > 0 && (!spheres || !out)))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "closest_points_device: bad arguments");
    if (((uintptr_t)spheres & 15) || ((uintptr_t)out & 7) || ((uintptr_t)object_index & 3))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT,
             "closest_points_device: spheres must be 16-byte, out 8-byte and object_index 4-byte aligned");
    if (n == 0) return FCPW_HIP_OK;
    hipStream_t s = (hipStream_t)stream;
    if (c->instanced) return launch_instances_closest(c, spheres, n, record_normals, out, object_index, s);
    st = launch_closest_records(c, spheres, n, record_normals, out, s);
    if (st) return st;
    return fill_single_object_index(c, out, n, object_index, s);
}

int fcpw_hip_closest_points_objects(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres, int64_t n,
                                    int record_normals, fcpw_hip_interaction *out, int32_t *object_index)
{
    int st = check_ready(c, "closest_points");
    if (st) return st;
    if (n < 0 || (n > 0 && (!spheres || !out))) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "closest_points: bad arguments");
    if (n == 0) return FCPW_HIP_OK;
    HIPCHECK(c, hipSetDevice(c->device));
    const size_t rec = (size_t)n * sizeof(fcpw_hip_interaction);
    st = grow(c, &c->stage_in, &c->stage_in_bytes, (size_t)n * sizeof(fcpw_hip_bounding_sphere));
    if (st) return st;
    st = grow(c, &c->stage_out, &c->stage_out_bytes, rec + (size_t)n * sizeof(int32_t));
    if (st) return st;
    int32_t *obj_dev = object_index ? (int32_t *)((char *)c->stage_out + rec) : nullptr;
    if ((st = stage_copy_in(c, c->stage_in, spheres, (size_t)n * sizeof(fcpw_hip_bounding_sphere), 0))) return st;
    st = fcpw_hip_closest_points_objects_device(c, (const fcpw_hip_bounding_sphere *)c->stage_in, n, record_normals,
                                                (fcpw_hip_interaction *)c->stage_out, obj_dev, nullptr);
    if (st) return st;
    if ((st = stage_copy_out(c, out, c->stage_out, rec, 0))) return st;
    if (object_index)
        if ((st = stage_copy_out(c, object_index, obj_dev, (size_t)n * sizeof(int32_t), 0))) return st;
    return check_overflow(c);
}

int fcpw_hip_intersect_objects_device(fcpw_hip_ctx *c, const fcpw_hip_ray *rays, int64_t n, int occlusion,
                                      fcpw_hip_interaction *out, int32_t *object_index, void *stream)
{
    int st = check_ready(c, "intersect_device");
    if (st) return st;
    if (n < 0 || (n > 0 && (!rays || !out))) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "intersect_device: bad arguments");
    if (((uintptr_t)rays & 7) || ((uintptr_t)out & 7) || ((uintptr_t)object_index & 3))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT,
             "intersect_device: rays and out must be 8-byte and object_index 4-byte aligned");
    if (n == 0) return FCPW_HIP_OK;
    hipStream_t s = (hipStream_t)stream;
    if (c->instanced) return launch_instances_intersect(c, rays, n, occlusion, out, object_index, s);
    st = launch_intersect(c, rays, n, occlusion, out, s);
    if (st) return st;
    return fill_single_object_index(c, out, n, object_index, s);
}

int fcpw_hip_intersect_objects(fcpw_hip_ctx *c, const fcpw_hip_ray *rays, int64_t n, int occlusion,
                               fcpw_hip_interaction *out, int32_t *object_index)
{
    int st = check_ready(c, "intersect");
    if (st) return st;
    if (n < 0 || (n > 0 && (!rays || !out))) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "intersect: bad arguments");
    if (n == 0) return FCPW_HIP_OK;
    HIPCHECK(c, hipSetDevice(c->device));
    const size_t rec = (size_t)n * sizeof(fcpw_hip_interaction);
    st = grow(c, &c->stage_in, &c->stage_in_bytes, (size_t)n * sizeof(fcpw_hip_ray));
    if (st) return st;
    st = grow(c, &c->stage_out, &c->stage_out_bytes, rec + (size_t)n * sizeof(int32_t));
    if (st) return st;
    int32_t *obj_dev = object_index ? (int32_t *)((char *)c->stage_out + rec) : nullptr;
    if ((st = stage_copy_in(c, c->stage_in, rays, (size_t)n * sizeof(fcpw_hip_ray), 0))) return st;
    st = fcpw_hip_intersect_objects_device(c, (const fcpw_hip_ray *)c->stage_in, n, occlusion,
                                           (fcpw_hip_interaction *)c->stage_out, obj_dev, nullptr);
    if (st) return st;
    if ((st = stage_copy_out(c, out, c->stage_out, rec, 0))) return st;
    if (object_index)
        if ((st = stage_copy_out(c, object_index, obj_dev, (size_t)n * sizeof(int32_t), 0))) return st;
    return check_overflow(c);
}

int fcpw_hip_intersect_spheres_objects_device(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres,
                                             const fcpw_hip_float3 *rand_nums, int64_t n, fcpw_hip_interaction *out,
                                             int32_t *object_index, void *stream)
{
    int st = check_ready(c, "intersect_spheres_device");
    if (st) return st;
    if (n < 0 || (n > 0 && (!spheres || !rand_nums || !out)))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "intersect_spheres_device: bad arguments");
    if (((uintptr_t)spheres & 15) || ((uintptr_t)rand_nums & 3) || ((uintptr_t)out & 7) ||
        ((uintptr_t)object_index & 3))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT,
             "intersect_spheres_device: spheres must be 16-byte, out 8-byte, rand_nums and object_index 4-byte "
             "aligned");
    if (n == 0) return FCPW_HIP_OK;
    hipStream_t s = (hipStream_t)stream;
    if (c->instanced) return launch_instances_spheres(c, spheres, (const float *)rand_nums, n, out, object_index, s);
    st = fcpw_hip_intersect_spheres_device(c, spheres, rand_nums, n, out, stream);
    if (st) return st;
    return fill_single_object_index(c, out, n, object_index, s);
}

int fcpw_hip_intersect_spheres_objects(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres,
                                      const fcpw_hip_float3 *rand_nums, int64_t n, fcpw_hip_interaction *out,
                                      int32_t *object_index)
{
    int st = check_ready(c, "intersect_spheres");
    if (st) return st;
    if (n < 0 || (n > 0 && (!spheres || !rand_nums || !out)))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "intersect_spheres: bad arguments");
    if (n == 0) return FCPW_HIP_OK;
    HIPCHECK(c, hipSetDevice(c->device));
    const size_t sb = (size_t)n * sizeof(fcpw_hip_bounding_sphere), rb = (size_t)n * sizeof(fcpw_hip_float3);
    const size_t rec = (size_t)n * sizeof(fcpw_hip_interaction);
    st = grow(c, &c->stage_in, &c->stage_in_bytes, sb + rb);
    if (st) return st;
    st = grow(c, &c->stage_out, &c->stage_out_bytes, rec + (size_t)n * sizeof(int32_t));
    if (st) return st;
    int32_t *obj_dev = object_index ? (int32_t *)((char *)c->stage_out + rec) : nullptr;
    if ((st = stage_copy_in(c, c->stage_in, spheres, sb, 0))) return st;
    if ((st = stage_copy_in(c, (char *)c->stage_in + sb, rand_nums, rb, 0))) return st;
    st = fcpw_hip_intersect_spheres_objects_device(c, (const fcpw_hip_bounding_sphere *)c->stage_in,
                                                   (const fcpw_hip_float3 *)((char *)c->stage_in + sb), n,
                                                   (fcpw_hip_interaction *)c->stage_out, obj_dev, nullptr);
    if (st) return st;
    if ((st = stage_copy_out(c, out, c->stage_out, rec, 0))) return st;
    if (object_index)
        if ((st = stage_copy_out(c, object_index, obj_dev, (size_t)n * sizeof(int32_t), 0))) return st;
    return FCPW_HIP_OK;
}

} // extern "C"
