/*
 * fcpw_oracle_mbvh.cpp -- TEST INFRASTRUCTURE ONLY: the vectorized leg of the
 * CPU baseline, and the checks that pin it.  The product never loads it.
 *
 * A restatement of the reference's vectorized CPU aggregate, the path
 * Scene::build(aggregateType, vectorize = true) selects when FCPW is built
 * with Enoki (reference files relative to the fcpw source tree):
 *   - leaves packed to FCPW_SIMD_WIDTH by the SAH build
 *         include/fcpw/fcpw.inl:812-827, include/fcpw/aggregates/bvh.inl:3-37,113-140
 *         (oracle_build_packed in fcpw_oracle.c)
 *   - collapse of the binary tree into 4-wide nodes
 *         include/fcpw/aggregates/mbvh.inl:56-147, mbvh.h:4-24
 *   - leaf groups of WIDTH primitives in structure-of-arrays form
 *         include/fcpw/aggregates/mbvh.inl:171-245 (populateLeafNode), mbvh.h:44-49
 *   - closest-point traversal
 *         include/fcpw/aggregates/mbvh.inl:2034-2197 (findClosestPointPrimitives,
 *         findClosestPointFromNode), :745-803 (enqueueNodes<4>, sortOrder4)
 *   - wide box overlap and wide closest point on triangles / segments
 *         include/fcpw/core/wide_query_operations.h:127-268
 * Enoki itself (deps/enoki) is not compiled here.  The wide arithmetic is
 * written with GCC vector extensions; rcp() is computed as Enoki computes it
 * on x86: the hardware reciprocal estimate refined by one Newton-Raphson step
 * (deps/enoki/include/enoki/array_avx.h:324-352 for 8 lanes without AVX-512VL,
 * array_avx512.h:312-330 for 16 lanes).  Enoki's exact summation order (and
 * use of FMA) in dot() is a library detail this file does not claim to
 * reproduce; results are pinned against the scalar restatement and the brute
 * force within the reference's own tolerance (|dd| <= 1e-6,
 * tests/aggregate_tests.cpp:562,790-807), not bit for bit.
 *
 * The file is compiled twice: MBVH_WIDTH = 8 with AVX2 + FMA (FCPW_SIMD_WIDTH 8,
 * CMakeLists.txt:91-94) and MBVH_WIDTH = 16 with AVX-512 (FCPW_SIMD_WIDTH 16,
 * CMakeLists.txt:95-98); every exported name carries the width.
 */
#include <immintrin.h>
#include <pthread.h>
#include <unistd.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fcpw_oracle.h"

#ifndef MBVH_WIDTH
#error "MBVH_WIDTH must be 8 or 16"
#endif

namespace {

constexpr int W = MBVH_WIDTH;       // FCPW_SIMD_WIDTH: primitives per leaf group
constexpr int B = 4;                // FCPW_MBVH_BRANCHING_FACTOR (mbvh.h:4-10)
constexpr int MAX_DEPTH = 96;       // FCPW_MBVH_MAX_DEPTH for 4-wide nodes
constexpr int MAX_INT = 0x7fffffff; // maxInt (core.h:80): an unused child slot
constexpr float MIN_FLOAT = -FLT_MAX;

typedef float vf __attribute__((vector_size(W * 4)));
typedef int32_t vi __attribute__((vector_size(W * 4)));
typedef float vf4 __attribute__((vector_size(16)));
typedef int32_t vi4 __attribute__((vector_size(16)));

// Enoki rcp on x86: estimate, one Newton-Raphson step r' = 2r - (r m) r
#if MBVH_WIDTH == 8
inline vf rcp(vf m)
{
    __m256 x = (__m256)m;
    __m256 r = _mm256_rcp_ps(x);
    __m256 t0 = _mm256_add_ps(r, r), t1 = _mm256_mul_ps(r, x), ro = r;
    r = _mm256_fnmadd_ps(t1, r, t0);
    return (vf)_mm256_blendv_ps(r, ro, t1); // keep the estimate where r*m is NaN (m = 0 or inf)
}
inline bool all(vi m) { return _mm256_movemask_ps((__m256)m) == 0xff; }
inline bool any(vi m) { return _mm256_movemask_ps((__m256)m) != 0; }
#else
inline vf rcp(vf m)
{
    __m512 x = (__m512)m;
    __m512 r = _mm512_rcp14_ps(x);
    __m512 t0 = _mm512_add_ps(r, r), t1 = _mm512_mul_ps(r, x);
    r = _mm512_fnmadd_ps(t1, r, t0);
    return (vf)_mm512_fixupimm_ps(r, x, _mm512_set1_epi32(0x0087A622), 0);
}
inline bool all(vi m) { return _mm512_cmpneq_epi32_mask((__m512i)m, _mm512_setzero_si512()) == 0xffff; }
inline bool any(vi m) { return _mm512_cmpneq_epi32_mask((__m512i)m, _mm512_setzero_si512()) != 0; }
#endif

inline vf splat(float a)
{
    vf r;
    for (int w = 0; w < W; w++) r[w] = a;
    return r;
}
inline vf sel(vi m, vf a, vf b) { return m ? a : b; }

struct Vec {
    vf c[3];
};

// 4-wide inner node (MbvhNode, mbvh.h:14-24): child >= 0 an inner node, < 0 a
// leaf, MAX_INT an unused slot; a leaf node's child = {-(first group + 1),
// groups, reference offset, references}
struct alignas(16) Node {
    float bmin[3][B], bmax[3][B];
    int32_t child[B];
};

// MbvhLeafNode (mbvh.h:44-49): vertex k, coordinate i, lane w
struct alignas(64) Leaf {
    float pos[3][3][W];
    int32_t prim[W];
};

struct Mbvh {
    int dim, k;
    std::vector<Node> nodes;
    std::vector<Leaf> leaves;
    int n_leaf_groups = 0, max_depth = 0;
    // the binary tree being collapsed
    std::vector<float> pmin, pmax;
    std::vector<int32_t> nrefs, offset;
};

float surface_area(const Mbvh &m, int i)
{
    return oracle_box_surface_area(&m.pmin[3 * i], &m.pmax[3 * i], m.dim);
}

// collapseBvh (mbvh.inl:56-147)
int collapse(Mbvh &m, int bi, int depth)
{
    m.max_depth = std::max(m.max_depth, depth);
    const int mi = (int)m.nodes.size();
    Node nd;
    for (int j = 0; j < 3; j++)
        for (int w = 0; w < B; w++) {
            nd.bmin[j][w] = FLT_MAX;
            nd.bmax[j][w] = MIN_FLOAT;
        }
    for (int w = 0; w < B; w++) nd.child[w] = MAX_INT;
    m.nodes.push_back(nd);
    if (m.nrefs[bi] > 0) {
        Node &n = m.nodes[mi];
        n.child[0] = -(m.n_leaf_groups + 1);
        n.child[1] = m.nrefs[bi] / W + (m.nrefs[bi] % W != 0 ? 1 : 0);
        n.child[2] = m.offset[bi];
        n.child[3] = m.nrefs[bi];
        m.n_leaf_groups += n.child[1];
        return mi;
    }
    int cnt = 2, list[B];
    list[0] = bi + m.offset[bi];
    list[1] = bi + 1;
    while (cnt < B) {
        float best = MIN_FLOAT;
        int bestIdx = -1;
        for (int i = 0; i < cnt; i++) {
            if (m.nrefs[list[i]] == 0) {
                float sa = surface_area(m, list[i]);
                if (best < sa) {
                    best = sa;
                    bestIdx = i;
                }
            }
        }
        if (bestIdx == -1) break;
        const int e = list[bestIdx];
        list[bestIdx] = e + m.offset[e];
        list[cnt++] = e + 1;
    }
    std::sort(list, list + cnt);
    for (int i = 0; i < cnt; i++) {
        const int c = list[i];
        for (int j = 0; j < m.dim; j++) {
            m.nodes[mi].bmin[j][i] = m.pmin[3 * c + j];
            m.nodes[mi].bmax[j][i] = m.pmax[3 * c + j];
        }
        const int child = collapse(m, c, depth + 1);
        m.nodes[mi].child[i] = child;
    }
    return mi;
}

// wide closest point on a triangle (wide_query_operations.h:176-268): the
// region tests of Ericson's method on W triangles at once, with the
// reference's early returns once every lane has found its region
vf closest_wide_triangle(const Leaf &L, const float x[3], Vec &pt, vf t[2])
{
    Vec pa, pb, pc, ab, ac, ax, bx, cx;
    for (int i = 0; i < 3; i++) {
        pa.c[i] = *(const vf *)L.pos[0][i];
        pb.c[i] = *(const vf *)L.pos[1][i];
        pc.c[i] = *(const vf *)L.pos[2][i];
        ab.c[i] = pb.c[i] - pa.c[i];
        ac.c[i] = pc.c[i] - pa.c[i];
        ax.c[i] = x[i] - pa.c[i];
    }
    auto dot = [](const Vec &a, const Vec &b) { return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2]; };
    auto norm_to_x = [&](const Vec &p) {
        vf d0 = x[0] - p.c[0], d1 = x[1] - p.c[1], d2 = x[2] - p.c[2];
        vf s = d0 * d0 + d1 * d1 + d2 * d2;
        vf r;
        for (int w = 0; w < W; w++) r[w] = sqrtf(s[w]);
        return r;
    };
    auto set_pt = [&](vi m, const Vec &v) {
        for (int i = 0; i < 3; i++) pt.c[i] = sel(m, v.c[i], pt.c[i]);
    };
    for (int i = 0; i < 3; i++) pt.c[i] = splat(0.0f);
    t[0] = t[1] = splat(0.0f);
    const vf zero = splat(0.0f), one = splat(1.0f);
    vf d1 = dot(ab, ax), d2 = dot(ac, ax);
    vi a1 = (d1 <= 0.0f) & (d2 <= 0.0f);
    vi a7 = a1;
    set_pt(a1, pa);
    t[0] = sel(a1, one, t[0]);
    t[1] = sel(a1, zero, t[1]);
    if (all(a7)) return norm_to_x(pt);

    for (int i = 0; i < 3; i++) bx.c[i] = x[i] - pb.c[i];
    vf d3 = dot(ab, bx), d4 = dot(ac, bx);
    vi a2 = (d3 >= 0.0f) & (d4 <= d3);
    a7 |= a2;
    set_pt(a2, pb);
    t[0] = sel(a2, zero, t[0]);
    t[1] = sel(a2, one, t[1]);
    if (all(a7)) return norm_to_x(pt);

    for (int i = 0; i < 3; i++) cx.c[i] = x[i] - pc.c[i];
    vf d5 = dot(ab, cx), d6 = dot(ac, cx);
    vi a3 = (d6 >= 0.0f) & (d5 <= d6);
    a7 |= a3;
    set_pt(a3, pc);
    t[0] = sel(a3, zero, t[0]);
    t[1] = sel(a3, zero, t[1]);
    if (all(a7)) return norm_to_x(pt);

    vf vc = d1 * d4 - d3 * d2;
    vi a4 = (vc <= 0.0f) & (d1 >= 0.0f) & (d3 <= 0.0f);
    a7 |= a4;
    vf v = d1 * rcp(d1 - d3);
    Vec e;
    for (int i = 0; i < 3; i++) e.c[i] = pa.c[i] + ab.c[i] * v;
    set_pt(a4, e);
    t[0] = sel(a4, 1.0f - v, t[0]);
    t[1] = sel(a4, v, t[1]);
    if (all(a7)) return norm_to_x(pt);

    vf vb = d5 * d2 - d1 * d6;
    vi a5 = (vb <= 0.0f) & (d2 >= 0.0f) & (d6 <= 0.0f);
    a7 |= a5;
    vf w = d2 * rcp(d2 - d6);
    for (int i = 0; i < 3; i++) e.c[i] = pa.c[i] + ac.c[i] * w;
    set_pt(a5, e);
    t[0] = sel(a5, 1.0f - w, t[0]);
    t[1] = sel(a5, zero, t[1]);
    if (all(a7)) return norm_to_x(pt);

    vf va = d3 * d6 - d5 * d4;
    vi a6 = (va <= 0.0f) & ((d4 - d3) >= 0.0f) & ((d5 - d6) >= 0.0f);
    a7 |= a6;
    w = (d4 - d3) * rcp((d4 - d3) + (d5 - d6));
    for (int i = 0; i < 3; i++) e.c[i] = pb.c[i] + (pc.c[i] - pb.c[i]) * w;
    set_pt(a6, e);
    t[0] = sel(a6, zero, t[0]);
    t[1] = sel(a6, 1.0f - w, t[1]);
    if (all(a7)) return norm_to_x(pt);

    vf denom = rcp(va + vb + vc);
    v = vb * denom;
    w = vc * denom;
    vi inside = ~a7;
    for (int i = 0; i < 3; i++) e.c[i] = pa.c[i] + ab.c[i] * v + ac.c[i] * w;
    set_pt(inside, e);
    t[0] = sel(inside, 1.0f - v - w, t[0]);
    t[1] = sel(inside, v, t[1]);
    return norm_to_x(pt);
}

// wide closest point on a line segment (wide_query_operations.h:150-173)
vf closest_wide_segment(const Leaf &L, const float x[3], Vec &pt, vf &t)
{
    Vec pa, u, v;
    for (int i = 0; i < 2; i++) {
        pa.c[i] = *(const vf *)L.pos[0][i];
        u.c[i] = *(const vf *)L.pos[1][i] - pa.c[i];
        v.c[i] = x[i] - pa.c[i];
    }
    vf c1 = u.c[0] * v.c[0] + u.c[1] * v.c[1];
    vf c2 = u.c[0] * u.c[0] + u.c[1] * u.c[1];
    vi a1 = c1 <= 0.0f, a2 = c2 <= c1;
    t = c1 * rcp(c2);
    t = sel(a1, splat(0.0f), t);
    t = sel(a2, splat(1.0f), t);
    vf s = splat(0.0f);
    for (int i = 0; i < 2; i++) {
        pt.c[i] = pa.c[i] + u.c[i] * t;
        vf d = x[i] - pt.c[i];
        s = s + d * d;
    }
    pt.c[2] = splat(0.0f);
    vf r;
    for (int w = 0; w < W; w++) r[w] = sqrtf(s[w]);
    return r;
}

// sortOrder4 (mbvh.inl:779-788): a, b, c, d ordered by decreasing t
inline void sort_order4(const float t[4], int &a, int &b, int &c, int &d)
{
    int tmp;
    if (t[a] < t[b]) { tmp = a; a = b; b = tmp; }
    if (t[c] < t[d]) { tmp = c; c = d; d = tmp; }
    if (t[a] < t[c]) { tmp = a; a = c; c = tmp; }
    if (t[b] < t[d]) { tmp = b; b = d; d = tmp; }
    if (t[b] < t[c]) { tmp = b; b = c; c = tmp; }
}

struct Entry {
    int node;
    float dist;
};

// findClosestPointFromNode (mbvh.inl:2097-2197) for one query, with
// findClosestPointPrimitives (:2034-2095) at the leaves
void closest_one(const Mbvh &m, const float *xq, float r2, oracle_hit &out, int64_t *nodes_visited)
{
    float x[3] = {xq[0], xq[1], m.dim == 3 ? xq[2] : 0.0f};
    out = oracle_hit{};
    out.d = FLT_MAX;
    out.index = -1;
    Entry st[MAX_DEPTH];
    st[0].node = 0;
    st[0].dist = MIN_FLOAT;
    int sp = 0;
    int64_t visited = 0;
    while (sp >= 0) {
        const int ni = st[sp].node;
        const float cur = st[sp].dist;
        sp--;
        if (cur > r2) continue;
        const Node &n = m.nodes[ni];
        if (n.child[0] < 0) {
            const int first = -n.child[0] - 1, groups = n.child[1], refs = n.child[3];
            int start = 0;
            for (int l = 0; l < groups; l++) {
                const Leaf &L = m.leaves[first + l];
                Vec pt;
                vf t[2];
                vf d;
                if (m.dim == 3) {
                    d = closest_wide_triangle(L, x, pt, t);
                } else {
                    d = closest_wide_segment(L, x, pt, t[0]);
                    t[1] = splat(0.0f);
                }
                vf d2 = d * d;
                int ci = -1;
                const int lanes = std::min(W, refs - start);
                for (int w = 0; w < lanes; w++) {
                    if (d2[w] <= r2) {
                        ci = w;
                        r2 = d2[w];
                    }
                }
                if (ci != -1) {
                    out.d = d[ci];
                    for (int i = 0; i < 3; i++) out.p[i] = pt.c[i][ci];
                    out.uv[0] = t[0][ci];
                    out.uv[1] = m.dim == 3 ? t[1][ci] : 0.0f;
                    out.index = L.prim[ci];
                }
                start += W;
            }
            visited++;
        } else {
            // overlapWideBox (wide_query_operations.h:127-139) on the 4 children
            vf4 d2min = {0, 0, 0, 0}, d2max = {0, 0, 0, 0};
            for (int j = 0; j < m.dim; j++) {
                vf4 lo = *(const vf4 *)n.bmin[j], hi = *(const vf4 *)n.bmax[j];
                vf4 u = lo - x[j], v = x[j] - hi;
                vf4 mx = u > v ? u : v;
                mx = mx > 0.0f ? mx : (vf4){0, 0, 0, 0};
                vf4 mn = u < v ? u : v;
                d2min += mx * mx;
                d2max += mn * mn;
            }
            vi4 mask = (d2min <= r2) & (*(const vi4 *)n.child != MAX_INT);
            visited++;
            if (mask[0] | mask[1] | mask[2] | mask[3]) {
                // enqueueNodes<4> (mbvh.inl:790-803): nearest pushed last
                float dm[4] = {d2min[0], d2min[1], d2min[2], d2min[3]};
                int o[4] = {0, 1, 2, 3};
                sort_order4(dm, o[0], o[1], o[2], o[3]);
                for (int k = 0; k < 4; k++) {
                    const int c = o[k];
                    if (mask[c]) {
                        sp++;
                        st[sp].node = n.child[c];
                        st[sp].dist = dm[c];
                        r2 = std::min(r2, d2max[c]);
                    }
                }
            }
        }
    }
    if (nodes_visited) *nodes_visited = visited;
}

struct Job {
    const Mbvh *m;
    const float *q, *r2;
    oracle_hit *out;
    int64_t *visited;
    int begin, end;
};

void *run_job(void *p)
{
    Job *j = (Job *)p;
    int64_t v = 0;
    for (int i = j->begin; i < j->end; i++) {
        int64_t nv = 0;
        closest_one(*j->m, j->q + (size_t)i * j->m->dim, j->r2[i], j->out[i], &nv);
        v += nv;
    }
    if (j->visited) *j->visited = v;
    return nullptr;
}

} // namespace

#define MBVH_CAT2(a, b) a##b
#define MBVH_CAT(a, b) MBVH_CAT2(a, b)
#define MBVH_NAME(x) MBVH_CAT(x, MBVH_CAT(_w, MBVH_WIDTH))

extern "C" {

/* Collapses h (built by oracle_build_packed with leaf_size = width) into the
   4-wide tree with leaf groups of `width` primitives (mbvh.inl:56-147,171-245). */
void *MBVH_NAME(oracle_mbvh_build)(const oracle_bvh *h)
{
    Mbvh *m = new Mbvh();
    m->dim = oracle_dim(h);
    m->k = m->dim;
    const int nn = oracle_num_nodes(h), np = oracle_num_prims(h);
    m->pmin.resize(3 * (size_t)nn);
    m->pmax.resize(3 * (size_t)nn);
    m->nrefs.resize(nn);
    m->offset.resize(nn);
    oracle_export_nodes(h, m->pmin.data(), m->pmax.data(), m->nrefs.data(), m->offset.data());
    collapse(*m, 0, 0);
    // populateLeafNode: reference slot order, the unused lanes of a group
    // repeat its last primitive (they are never read back)
    std::vector<float> verts((size_t)np * m->k * m->dim);
    std::vector<int32_t> prim(np);
    oracle_export_prim_vertices(h, verts.data());
    oracle_export_prim_order(h, prim.data());
    m->leaves.resize(m->n_leaf_groups);
    for (const Node &n : m->nodes) {
        if (n.child[0] >= 0) continue;
        const int first = -n.child[0] - 1, off = n.child[2], refs = n.child[3];
        for (int g = 0; g < n.child[1]; g++) {
            Leaf &L = m->leaves[first + g];
            std::memset(&L, 0, sizeof(L));
            for (int w = 0; w < W; w++) {
                const int p = std::min(g * W + w, refs - 1);
                const int slot = off + p;
                L.prim[w] = prim[slot];
                for (int v = 0; v < m->k; v++)
                    for (int i = 0; i < m->dim; i++)
                        L.pos[v][i][w] = verts[((size_t)slot * m->k + v) * m->dim + i];
            }
        }
    }
    m->pmin.clear(); m->pmax.clear(); m->nrefs.clear(); m->offset.clear();
    m->pmin.shrink_to_fit(); m->pmax.shrink_to_fit(); m->nrefs.shrink_to_fit(); m->offset.shrink_to_fit();
    return m;
}

void MBVH_NAME(oracle_mbvh_free)(void *m) { delete (Mbvh *)m; }

/* nodes, leaf groups and depth of the 4-wide tree */
void MBVH_NAME(oracle_mbvh_info)(const void *vm, int64_t *n_nodes, int64_t *n_leaf_groups, int64_t *max_depth)
{
    const Mbvh *m = (const Mbvh *)vm;
    *n_nodes = (int64_t)m->nodes.size();
    *n_leaf_groups = m->n_leaf_groups;
    *max_depth = m->max_depth;
}

/* the 4-wide tree: per node the child boxes (bmin / bmax: n_nodes x 3 x 4, z = 0
   in 2D) and child slots (n_nodes x 4), per leaf group the primitive of each
   lane (n_leaf_groups x width; lanes past a leaf's references repeat its last
   primitive) */
void MBVH_NAME(oracle_mbvh_export)(const void *vm, float *bmin, float *bmax, int32_t *child, int32_t *leaf_prim)
{
    const Mbvh *m = (const Mbvh *)vm;
    for (size_t i = 0; i < m->nodes.size(); i++) {
        const Node &n = m->nodes[i];
        for (int j = 0; j < 3; j++)
            for (int w = 0; w < B; w++) {
                bmin[(i * 3 + j) * B + w] = j < m->dim ? n.bmin[j][w] : 0.0f;
                bmax[(i * 3 + j) * B + w] = j < m->dim ? n.bmax[j][w] : 0.0f;
            }
        for (int w = 0; w < B; w++) child[i * B + w] = n.child[w];
    }
    for (size_t g = 0; g < m->leaves.size(); g++)
        for (int w = 0; w < W; w++) leaf_prim[g * W + w] = m->leaves[g].prim[w];
}

/* Scene<DIM>::findClosestPoints on the vectorized aggregate: nthreads <= 0 uses
   every online core, with the static partition of fcpw.inl:1547-1559.  Records:
   d, p, uv and the original primitive index (-1 and d = FLT_MAX on a miss);
   normals are not computed.  visited (optional) receives the total number of
   nodes visited (the reference's nodesVisited count). */
void MBVH_NAME(oracle_mbvh_closest_points)(const void *vm, const float *queries, const float *r2, int n,
                                           int nthreads, oracle_hit *out, int64_t *visited)
{
    const Mbvh *m = (const Mbvh *)vm;
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads < 1) nthreads = 1;
    if (n < 2 * nthreads) nthreads = 1;
    std::vector<Job> jobs(nthreads);
    std::vector<int64_t> counts(nthreads, 0);
    std::vector<pthread_t> th(nthreads);
    const int per = n / nthreads;
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = Job{m, queries, r2, out, &counts[t], t * per, t == nthreads - 1 ? n : (t + 1) * per};
        if (nthreads > 1) pthread_create(&th[t], nullptr, run_job, &jobs[t]);
        else run_job(&jobs[t]);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], nullptr);
    if (visited) {
        int64_t s = 0;
        for (int64_t c : counts) s += c;
        *visited = s;
    }
}

} // extern "C"
