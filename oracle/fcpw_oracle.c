/*
 * fcpw_oracle.c -- TEST INFRASTRUCTURE ONLY (see fcpw_oracle.h).
 *
 * Arithmetic conventions, chosen to follow each reference file:
 *   - CPU-order functions follow the Eigen expressions in the reference C++.
 *     Eigen reduces a fixed-size 3-vector as c0 + (c1 + c2) (its unroller
 *     splits the range in halves), so dot, squaredNorm, sum and prod use that
 *     order here; 2-vectors reduce as c0 + c1.  min/max follow std::min/std::max.
 *   - GPU-order functions follow the Slang source, whose float3 dot is written
 *     left to right: (x*x' + y*y') + z*z'.
 *   - Built with -ffp-contract=off, so no multiply-add is fused.
 * Which Eigen commit the reference pins is unknown (the submodule is empty), so
 * bit equality with the reference binary is not claimed; the file:line citations
 * name the code each function restates.
 */
#include "fcpw_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define MAXF FLT_MAX
#define LOWF (-FLT_MAX)
#define EPSF FLT_EPSILON
#define BVH_MAX_DEPTH 64
/* M_PI and M_PI_2 rounded to float, as the reference stores them */
#define PI_F ((float)3.14159265358979323846)
#define PI_2_F ((float)1.57079632679489661923)

typedef struct {
    float pmin[3], pmax[3];
} box_t;

typedef struct {
    box_t box;
    int32_t offset; /* referenceOffset (leaf) or secondChildOffset (internal) */
    int32_t nrefs;  /* 0 for internal nodes */
} node_t;

struct oracle_bvh {
    int dim, k; /* k = vertices per primitive */
    int n_vertices, n_prims;
    float *positions;     /* sorted (leaf-order) vertex positions, dim floats each */
    int32_t *prim_vidx;   /* per leaf slot: k sorted vertex ids */
    int32_t *prim_orig;   /* per leaf slot: original primitive index (pIndex) */
    int32_t *index_map;   /* user vertex -> sorted vertex */
    node_t *nodes;
    int n_nodes, n_leaves, max_depth;
    /* silhouettes, set by oracle_build_silhouettes (NULL until then) */
    struct sil_data *sil;
};

/* ----------------------------------------------------------------------------
 * scalar helpers
 * ------------------------------------------------------------------------- */
static inline float smin(float a, float b) { return (b < a) ? b : a; } /* std::min */
static inline float smax(float a, float b) { return (a < b) ? b : a; } /* std::max */

/* Eigen fixed-size reduction order: 3 -> c0 + (c1 + c2), 2 -> c0 + c1 */
static inline float esum(const float *c, int dim)
{
    return dim == 3 ? c[0] + (c[1] + c[2]) : c[0] + c[1];
}
static inline float eprod(const float *c, int dim)
{
    return dim == 3 ? c[0] * (c[1] * c[2]) : c[0] * c[1];
}
static inline float edot(const float *a, const float *b, int dim)
{
    float c[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < dim; i++) c[i] = a[i] * b[i];
    return esum(c, dim);
}
static inline float enorm(const float *a, int dim) { return sqrtf(edot(a, a, dim)); }

/* Slang float3 dot, left to right; a 2D float3 has z = 0, and adding the zero
   product leaves the sum unchanged, so the 2-term form is used for dim == 2 */
static inline float gdot(const float *a, const float *b, int dim)
{
    float s = a[0] * b[0] + a[1] * b[1];
    if (dim == 3) s = s + a[2] * b[2];
    return s;
}

/* ----------------------------------------------------------------------------
 * bounding boxes (bounding_volumes.h:31-168)
 * ------------------------------------------------------------------------- */
static inline void box_init(box_t *b)
{
    for (int i = 0; i < 3; i++) { b->pmin[i] = MAXF; b->pmax[i] = LOWF; }
}
/* expandToInclude(point): inflated by +-FLT_EPSILON (bounding_volumes.h:45-49) */
static inline void box_expand_point(box_t *b, const float *p, int dim)
{
    for (int i = 0; i < dim; i++) {
        b->pmin[i] = smin(b->pmin[i], p[i] - EPSF);
        b->pmax[i] = smax(b->pmax[i], p[i] + EPSF);
    }
}
static inline void box_expand_box(box_t *b, const box_t *o, int dim)
{
    for (int i = 0; i < dim; i++) {
        b->pmin[i] = smin(b->pmin[i], o->pmin[i]);
        b->pmax[i] = smax(b->pmax[i], o->pmax[i]);
    }
}
/* surfaceArea: extent clamped to 1e-5 (bounding_volumes.h:165-168) */
static float box_surface_area(const box_t *b, int dim)
{
    float e[3] = {0.0f, 0.0f, 0.0f}, q[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < dim; i++) e[i] = smax(b->pmax[i] - b->pmin[i], 1e-5f);
    float P = eprod(e, dim);
    for (int i = 0; i < dim; i++) q[i] = P / e[i];
    return 2.0f * esum(q, dim);
}
static float box_volume(const box_t *b, int dim)
{
    float e[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < dim; i++) e[i] = b->pmax[i] - b->pmin[i];
    return eprod(e, dim);
}
static int box_valid(const box_t *b, int dim)
{
    for (int i = 0; i < dim; i++)
        if (!(b->pmax[i] >= b->pmin[i])) return 0;
    return 1;
}
static box_t box_intersection(const box_t *a, const box_t *b, int dim)
{
    box_t r;
    box_init(&r);
    for (int i = 0; i < dim; i++) {
        r.pmin[i] = smax(a->pmin[i], b->pmin[i]);
        r.pmax[i] = smin(a->pmax[i], b->pmax[i]);
    }
    return r;
}
/* maxDimension: first index of the largest extent (Eigen maxCoeff(&index)) */
static int box_max_dimension(const box_t *b, int dim)
{
    int idx = 0;
    float best = b->pmax[0] - b->pmin[0];
    for (int i = 1; i < dim; i++) {
        float e = b->pmax[i] - b->pmin[i];
        if (e > best) { best = e; idx = i; }
    }
    return idx;
}
/* computeSquaredDistance (bounding_volumes.h:64-69), Eigen order */
static inline void box_sqdist_cpu(const box_t *b, const float *p, int dim, float *d2min, float *d2max)
{
    float a[3], c[3];
    for (int i = 0; i < dim; i++) {
        float u = b->pmin[i] - p[i];
        float v = p[i] - b->pmax[i];
        a[i] = smax(smax(u, v), 0.0f);
        c[i] = smin(u, v);
    }
    *d2min = edot(a, a, dim);
    *d2max = edot(c, c, dim);
}
/* BoundingBox.overlap (bounding-volumes.slang:43-53), left-to-right dot */
static inline void box_sqdist_gpu(const box_t *b, const float *p, int dim, float *d2min, float *d2max)
{
    float a[3], c[3];
    for (int i = 0; i < dim; i++) {
        float u = b->pmin[i] - p[i];
        float v = p[i] - b->pmax[i];
        a[i] = fmaxf(fmaxf(u, v), 0.0f);
        c[i] = fminf(u, v);
    }
    *d2min = gdot(a, a, dim);
    *d2max = gdot(c, c, dim);
}
/* CPU slab test without the parallel-slab fix (bounding_volumes.h:99-114) */
static inline int box_ray_cpu(const box_t *b, const float *o, const float *invd, float tmax, int dim,
                              float *tmin_out, float *tmax_out)
{
    float tn[3], tf[3];
    for (int i = 0; i < dim; i++) {
        float t0 = (b->pmin[i] - o[i]) * invd[i];
        float t1 = (b->pmax[i] - o[i]) * invd[i];
        tn[i] = smin(t0, t1);
        tf[i] = smax(t0, t1);
    }
    /* Eigen maxCoeff/minCoeff reduce as c0 op (c1 op c2) */
    float tnm = dim == 3 ? smax(tn[0], smax(tn[1], tn[2])) : smax(tn[0], tn[1]);
    float tfm = dim == 3 ? smin(tf[0], smin(tf[1], tf[2])) : smin(tf[0], tf[1]);
    float tNearMax = smax(0.0f, tnm);
    float tFarMin = smin(tmax, tfm);
    if (tNearMax > tFarMin) return 0;
    *tmin_out = tNearMax;
    *tmax_out = tFarMin;
    return 1;
}
/* GPU slab test with the parallel-slab fix (bounding-volumes.slang:67-106);
   a 2D float3 box has z = [0,0] and a 2D ray has o.z = d.z = 0, which the fix
   turns into an unconstrained z slab, so dim == 2 simply skips z */
static inline int box_ray_gpu(const box_t *b, const float *o, const float *d, const float *dinv, float tmax,
                              int dim, float *tmin_out, float *tmax_out)
{
    float tn[3] = {-MAXF, -MAXF, -MAXF}, tf[3] = {MAXF, MAXF, MAXF};
    for (int i = 0; i < dim; i++) {
        float t0 = (b->pmin[i] - o[i]) * dinv[i];
        float t1 = (b->pmax[i] - o[i]) * dinv[i];
        int parallel = (d[i] == 0.0f);
        int within = (o[i] >= b->pmin[i]) && (o[i] <= b->pmax[i]);
        if (parallel && !within) { *tmin_out = MAXF; *tmax_out = MAXF; return 0; }
        tn[i] = parallel ? -MAXF : fminf(t0, t1);
        tf[i] = parallel ? MAXF : fmaxf(t0, t1);
    }
    float tNearMax = fmaxf(0.0f, fmaxf(tn[0], fmaxf(tn[1], tn[2])));
    float tFarMin = fminf(tmax, fminf(tf[0], fminf(tf[1], tf[2])));
    if (tNearMax > tFarMin) { *tmin_out = MAXF; *tmax_out = MAXF; return 0; }
    *tmin_out = tNearMax;
    *tmax_out = tFarMin;
    return 1;
}

/* ----------------------------------------------------------------------------
 * primitives
 * ------------------------------------------------------------------------- */
static inline const float *vtx(const oracle_bvh *h, int slot, int c)
{
    return h->positions + (size_t)h->prim_vidx[(size_t)slot * h->k + c] * h->dim;
}

/* Triangle::boundingBox / LineSegment::boundingBox (triangles.inl:12-23) */
static box_t prim_box(const float *const *v, int k, int dim)
{
    box_t b;
    box_init(&b);
    for (int c = 0; c < k; c++) box_expand_point(&b, v[c], dim);
    return b;
}
/* Triangle::centroid = (pa + pb + pc)/3 (triangles.inl:25-32); segment (pa + pb)*0.5 */
static void prim_centroid(const float *const *v, int k, int dim, float *out)
{
    for (int i = 0; i < dim; i++)
        out[i] = (k == 3) ? ((v[0][i] + v[1][i]) + v[2][i]) / 3.0f : (v[0][i] + v[1][i]) * 0.5f;
}

/* findClosestPointTriangle (triangles.inl:374-458), `dot` selects Eigen or Slang order */
typedef float (*dot_fn)(const float *, const float *, int);
static float closest_point_triangle(const float *pa, const float *pb, const float *pc, const float *x,
                                    float *pt, float *t, dot_fn dot)
{
    float ab[3], ac[3], ax[3], bx[3], cx[3], r[3];
    for (int i = 0; i < 3; i++) { ab[i] = pb[i] - pa[i]; ac[i] = pc[i] - pa[i]; ax[i] = x[i] - pa[i]; }
    float d1 = dot(ab, ax, 3), d2 = dot(ac, ax, 3);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        t[0] = 1.0f; t[1] = 0.0f;
        for (int i = 0; i < 3; i++) pt[i] = pa[i];
        goto done;
    }
    for (int i = 0; i < 3; i++) bx[i] = x[i] - pb[i];
    float d3 = dot(ab, bx, 3), d4 = dot(ac, bx, 3);
    if (d3 >= 0.0f && d4 <= d3) {
        t[0] = 0.0f; t[1] = 1.0f;
        for (int i = 0; i < 3; i++) pt[i] = pb[i];
        goto done;
    }
    for (int i = 0; i < 3; i++) cx[i] = x[i] - pc[i];
    float d5 = dot(ab, cx, 3), d6 = dot(ac, cx, 3);
    if (d6 >= 0.0f && d5 <= d6) {
        t[0] = 0.0f; t[1] = 0.0f;
        for (int i = 0; i < 3; i++) pt[i] = pc[i];
        goto done;
    }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        t[0] = 1.0f - v; t[1] = v;
        for (int i = 0; i < 3; i++) pt[i] = pa[i] + ab[i] * v;
        goto done;
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float w = d2 / (d2 - d6);
        t[0] = 1.0f - w; t[1] = 0.0f;
        for (int i = 0; i < 3; i++) pt[i] = pa[i] + ac[i] * w;
        goto done;
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        t[0] = 0.0f; t[1] = 1.0f - w;
        for (int i = 0; i < 3; i++) pt[i] = pb[i] + (pc[i] - pb[i]) * w;
        goto done;
    }
    {
        float denom = 1.0f / (va + vb + vc);
        float v = vb * denom, w = vc * denom;
        t[0] = 1.0f - v - w; t[1] = v;
        for (int i = 0; i < 3; i++) pt[i] = (pa[i] + ab[i] * v) + ac[i] * w;
    }
done:
    for (int i = 0; i < 3; i++) r[i] = x[i] - pt[i];
    return sqrtf(dot(r, r, 3));
}

/* findClosestPointLineSegment (line_segments.inl:200-227; geometry.slang:79-103) */
static float closest_point_segment(const float *pa, const float *pb, const float *x, float *pt, float *t,
                                   dot_fn dot)
{
    float u[2], v[2], r[2];
    for (int i = 0; i < 2; i++) { u[i] = pb[i] - pa[i]; v[i] = x[i] - pa[i]; }
    float c1 = dot(u, v, 2);
    if (c1 <= 0.0f) {
        pt[0] = pa[0]; pt[1] = pa[1]; *t = 0.0f;
    } else {
        float c2 = dot(u, u, 2);
        if (c2 <= c1) {
            pt[0] = pb[0]; pt[1] = pb[1]; *t = 1.0f;
        } else {
            *t = c1 / c2;
            for (int i = 0; i < 2; i++) pt[i] = pa[i] + u[i] * (*t);
        }
    }
    for (int i = 0; i < 2; i++) r[i] = x[i] - pt[i];
    return sqrtf(dot(r, r, 2));
}

/* closest point to one primitive slot; fills p (3, z = 0 in 2D) and uv (2) */
static float prim_closest(const oracle_bvh *h, int slot, const float *x, float *p, float *uv, dot_fn dot)
{
    if (h->dim == 3)
        return closest_point_triangle(vtx(h, slot, 0), vtx(h, slot, 1), vtx(h, slot, 2), x, p, uv, dot);
    p[2] = 0.0f;
    uv[1] = 0.0f;
    return closest_point_segment(vtx(h, slot, 0), vtx(h, slot, 1), x, p, &uv[0], dot);
}

/* normals: CPU Triangle::normal(true) (triangles.inl:48-59) = cross().normalized();
   LineSegment::normal(true) (line_segments.inl:46-55) = (s.y, -s.x).normalized() */
static void prim_normal_cpu(const oracle_bvh *h, int slot, float *n)
{
    if (h->dim == 3) {
        const float *pa = vtx(h, slot, 0), *pb = vtx(h, slot, 1), *pc = vtx(h, slot, 2);
        float a[3], b[3];
        for (int i = 0; i < 3; i++) { a[i] = pb[i] - pa[i]; b[i] = pc[i] - pa[i]; }
        n[0] = a[1] * b[2] - a[2] * b[1];
        n[1] = a[2] * b[0] - a[0] * b[2];
        n[2] = a[0] * b[1] - a[1] * b[0];
        float z = edot(n, n, 3);
        if (z > 0.0f) { float s = sqrtf(z); for (int i = 0; i < 3; i++) n[i] = n[i] / s; }
    } else {
        const float *pa = vtx(h, slot, 0), *pb = vtx(h, slot, 1);
        n[0] = pb[1] - pa[1];
        n[1] = -(pb[0] - pa[0]);
        n[2] = 0.0f;
        float z = edot(n, n, 2);
        if (z > 0.0f) { float s = sqrtf(z); n[0] = n[0] / s; n[1] = n[1] / s; }
    }
}
/* GPU getNormal (geometry.slang:125-131,361-366): normalize(v) written as v / length(v) */
static void prim_normal_gpu(const oracle_bvh *h, int slot, float *n)
{
    if (h->dim == 3) {
        const float *pa = vtx(h, slot, 0), *pb = vtx(h, slot, 1), *pc = vtx(h, slot, 2);
        float a[3], b[3];
        for (int i = 0; i < 3; i++) { a[i] = pb[i] - pa[i]; b[i] = pc[i] - pa[i]; }
        n[0] = a[1] * b[2] - a[2] * b[1];
        n[1] = a[2] * b[0] - a[0] * b[2];
        n[2] = a[0] * b[1] - a[1] * b[0];
        float l = sqrtf(gdot(n, n, 3));
        for (int i = 0; i < 3; i++) n[i] = n[i] / l;
    } else {
        const float *pa = vtx(h, slot, 0), *pb = vtx(h, slot, 1);
        n[0] = pb[1] - pa[1];
        n[1] = -(pb[0] - pa[0]);
        n[2] = 0.0f;
        float l = sqrtf(gdot(n, n, 2));
        n[0] = n[0] / l;
        n[1] = n[1] / l;
    }
}

/* Triangle::intersect (triangles.inl:222-260), Eigen order; returns hit and fills out */
static int ray_triangle_cpu(const float *pa, const float *pb, const float *pc, const float *o, const float *d,
                            float tmax, int occlusion, oracle_hit *out, int *occluded)
{
    float v1[3], v2[3], p[3], s[3], q[3];
    for (int i = 0; i < 3; i++) { v1[i] = pb[i] - pa[i]; v2[i] = pc[i] - pa[i]; }
    p[0] = d[1] * v2[2] - d[2] * v2[1];
    p[1] = d[2] * v2[0] - d[0] * v2[2];
    p[2] = d[0] * v2[1] - d[1] * v2[0];
    float det = edot(v1, p, 3);
    if (fabsf(det) <= EPSF) return 0;
    float invDet = 1.0f / det;
    for (int i = 0; i < 3; i++) s[i] = o[i] - pa[i];
    float v = edot(s, p, 3) * invDet;
    if (v < 0 || v > 1) return 0;
    q[0] = s[1] * v1[2] - s[2] * v1[1];
    q[1] = s[2] * v1[0] - s[0] * v1[2];
    q[2] = s[0] * v1[1] - s[1] * v1[0];
    float w = edot(d, q, 3) * invDet;
    if (w < 0 || v + w > 1) return 0;
    float t = edot(v2, q, 3) * invDet;
    if (t >= 0.0f && t <= tmax) {
        if (occlusion) { *occluded = 1; return 1; }
        out->d = t;
        for (int i = 0; i < 3; i++) out->p[i] = (pa[i] + v1[i] * v) + v2[i] * w;
        float n[3];
        n[0] = v1[1] * v2[2] - v1[2] * v2[1];
        n[1] = v1[2] * v2[0] - v1[0] * v2[2];
        n[2] = v1[0] * v2[1] - v1[1] * v2[0];
        float z = edot(n, n, 3);
        if (z > 0.0f) { float sq = sqrtf(z); for (int i = 0; i < 3; i++) n[i] = n[i] / sq; }
        for (int i = 0; i < 3; i++) out->n[i] = n[i];
        out->uv[0] = 1.0f - v - w;
        out->uv[1] = v;
        return 1;
    }
    return 0;
}
/* intersectTriangle (geometry.slang:202-249), left-to-right dot */
static int ray_triangle_gpu(const float *pa, const float *pb, const float *pc, const float *o, const float *d,
                            float tmax, int occlusion, oracle_hit *out)
{
    float v1[3], v2[3], q[3], r[3], s[3];
    for (int i = 0; i < 3; i++) { v1[i] = pb[i] - pa[i]; v2[i] = pc[i] - pa[i]; }
    q[0] = d[1] * v2[2] - d[2] * v2[1];
    q[1] = d[2] * v2[0] - d[0] * v2[2];
    q[2] = d[0] * v2[1] - d[1] * v2[0];
    float det = gdot(v1, q, 3);
    if (fabsf(det) <= EPSF) return 0;
    float invDet = 1.0f / det;
    for (int i = 0; i < 3; i++) r[i] = o[i] - pa[i];
    float v = gdot(r, q, 3) * invDet;
    if (v < 0.0f || v > 1.0f) return 0;
    s[0] = r[1] * v1[2] - r[2] * v1[1];
    s[1] = r[2] * v1[0] - r[0] * v1[2];
    s[2] = r[0] * v1[1] - r[1] * v1[0];
    float w = gdot(d, s, 3) * invDet;
    if (w < 0.0f || v + w > 1.0f) return 0;
    float t = gdot(v2, s, 3) * invDet;
    if (t >= 0.0f && t <= tmax) {
        if (occlusion) return 1;
        for (int i = 0; i < 3; i++) out->p[i] = (pa[i] + v1[i] * v) + v2[i] * w;
        float n[3];
        n[0] = v1[1] * v2[2] - v1[2] * v2[1];
        n[1] = v1[2] * v2[0] - v1[0] * v2[2];
        n[2] = v1[0] * v2[1] - v1[1] * v2[0];
        float l = sqrtf(gdot(n, n, 3));
        for (int i = 0; i < 3; i++) out->n[i] = n[i] / l;
        out->uv[0] = 1.0f - v - w;
        out->uv[1] = v;
        out->d = t;
        return 1;
    }
    return 0;
}
/* LineSegment::intersect (line_segments.inl:140-175) and intersectLineSegment
   (geometry.slang:37-77) use the same 2D cross products */
static int ray_segment(const float *pa, const float *pb, const float *o, const float *d, float tmax,
                       int occlusion, int cpu_normal, oracle_hit *out, int *occluded)
{
    float u[2] = {pa[0] - o[0], pa[1] - o[1]};
    float v[2] = {pb[0] - pa[0], pb[1] - pa[1]};
    float dv = d[0] * v[1] - d[1] * v[0];
    if (fabsf(dv) <= EPSF) return 0;
    float ud = u[0] * d[1] - u[1] * d[0];
    float s = ud / dv;
    if (s >= 0.0f && s <= 1.0f) {
        float uvx = u[0] * v[1] - u[1] * v[0];
        float t = uvx / dv;
        if (t >= 0.0f && t <= tmax) {
            if (occlusion) { if (occluded) *occluded = 1; return 1; }
            out->d = t;
            out->p[0] = pa[0] + s * v[0];
            out->p[1] = pa[1] + s * v[1];
            out->p[2] = 0.0f;
            float n[2] = {v[1], -v[0]};
            float z = cpu_normal ? edot(n, n, 2) : gdot(n, n, 2);
            float l = sqrtf(z);
            if (!cpu_normal || z > 0.0f) { n[0] = n[0] / l; n[1] = n[1] / l; }
            out->n[0] = n[0]; out->n[1] = n[1]; out->n[2] = 0.0f;
            out->uv[0] = s;
            out->uv[1] = 0.0f;
            return 1;
        }
    }
    return 0;
}

/* ----------------------------------------------------------------------------
 * BVH build (bvh.inl:3-234)
 * ------------------------------------------------------------------------- */
typedef struct {
    oracle_bvh *h;
    int heuristic, leaf_size, n_buckets;
    int pack_leaves, depth_guess; /* packLeaves / depthGuess (bvh.inl:440-455) */
    box_t *ref_boxes;
    float *ref_cent;  /* dim floats per reference */
    box_t *bkt_box, *rbkt_box;
    int *bkt_cnt, *rbkt_cnt;
    int cap;
} build_ctx;

static void swap_refs(build_ctx *b, int i, int j)
{
    int dim = b->h->dim;
    int32_t t = b->h->prim_orig[i]; b->h->prim_orig[i] = b->h->prim_orig[j]; b->h->prim_orig[j] = t;
    box_t tb = b->ref_boxes[i]; b->ref_boxes[i] = b->ref_boxes[j]; b->ref_boxes[j] = tb;
    for (int d = 0; d < dim; d++) {
        float tc = b->ref_cent[(size_t)i * dim + d];
        b->ref_cent[(size_t)i * dim + d] = b->ref_cent[(size_t)j * dim + d];
        b->ref_cent[(size_t)j * dim + d] = tc;
    }
}

/* computeSplitCost (bvh.inl:3-37); packLeaves is set for the vectorized
   aggregate's SAH build (fcpw.inl:812-827) */
static float split_cost(build_ctx *b, const box_t *bl, const box_t *br, int nl, int nr, int depth)
{
    int dim = b->h->dim;
    float cost = MAXF;
    if (b->pack_leaves && depth > 0 && ((float)b->depth_guess / depth) < 1.5f && nl % b->leaf_size != 0 &&
        nr % b->leaf_size != 0)
        return cost;
    if (b->heuristic == ORACLE_SURFACE_AREA) {
        cost = (float)nl * box_surface_area(bl, dim) + (float)nr * box_surface_area(br, dim);
    } else if (b->heuristic == ORACLE_OVERLAP_SURFACE_AREA) {
        box_t bi = box_intersection(bl, br, dim);
        cost = ((float)nl / box_surface_area(br, dim) + (float)nr / box_surface_area(bl, dim)) *
               fabsf(box_surface_area(&bi, dim));
        if (!box_valid(&bi, dim)) cost *= -1;
    } else if (b->heuristic == ORACLE_VOLUME) {
        cost = (float)nl * box_volume(bl, dim) + (float)nr * box_volume(br, dim);
    } else if (b->heuristic == ORACLE_OVERLAP_VOLUME) {
        box_t bi = box_intersection(bl, br, dim);
        cost = ((float)nl / box_volume(br, dim) + (float)nr / box_volume(bl, dim)) * fabsf(box_volume(&bi, dim));
        if (!box_valid(&bi, dim)) cost *= -1;
    }
    return cost;
}

/* computeObjectSplit (bvh.inl:39-111) */
static void object_split(build_ctx *b, const box_t *bb, const box_t *bc, int start, int end, int depth,
                         int *split_dim, float *split_coord)
{
    int dim = b->h->dim, nb = b->n_buckets;
    float split_cost_best = MAXF;
    *split_dim = -1;
    *split_coord = 0.0f;
    if (b->heuristic != ORACLE_LONGEST_AXIS_CENTER) {
        float extent[3];
        for (int d = 0; d < dim; d++) extent[d] = bb->pmax[d] - bb->pmin[d];
        for (int d = 0; d < dim; d++) {
            if (extent[d] < 1e-6f) continue;
            float bucket_width = extent[d] / nb;
            for (int k = 0; k < nb; k++) { box_init(&b->bkt_box[k]); b->bkt_cnt[k] = 0; }
            for (int p = start; p < end; p++) {
                int bi = (int)((b->ref_cent[(size_t)p * dim + d] - bb->pmin[d]) / bucket_width);
                if (bi < 0) bi = 0;
                if (bi > nb - 1) bi = nb - 1;
                box_expand_box(&b->bkt_box[bi], &b->ref_boxes[p], dim);
                b->bkt_cnt[bi] += 1;
            }
            box_t right;
            box_init(&right);
            for (int k = nb - 1; k > 0; k--) {
                box_expand_box(&right, &b->bkt_box[k], dim);
                b->rbkt_box[k] = right;
                b->rbkt_cnt[k] = b->bkt_cnt[k];
                if (k != nb - 1) b->rbkt_cnt[k] += b->rbkt_cnt[k + 1];
            }
            box_t left;
            box_init(&left);
            int nl = 0;
            for (int k = 1; k < nb; k++) {
                box_expand_box(&left, &b->bkt_box[k - 1], dim);
                nl += b->bkt_cnt[k - 1];
                if (nl > 0 && b->rbkt_cnt[k] > 0) {
                    float cost = split_cost(b, &left, &b->rbkt_box[k], nl, b->rbkt_cnt[k], depth);
                    if (cost < split_cost_best) {
                        split_cost_best = cost;
                        *split_dim = d;
                        *split_coord = bb->pmin[d] + k * bucket_width;
                    }
                }
            }
        }
    }
    if (*split_dim == -1) {
        *split_dim = box_max_dimension(bc, dim);
        *split_coord = (bc->pmin[*split_dim] + bc->pmax[*split_dim]) * 0.5f;
    }
}

/* performObjectSplit (bvh.inl:113-140) */
static int object_partition(build_ctx *b, int start, int end, int split_dim, float split_coord)
{
    int dim = b->h->dim, mid = start;
    for (int i = start; i < end; i++) {
        if (b->ref_cent[(size_t)i * dim + split_dim] < split_coord) {
            swap_refs(b, i, mid);
            mid++;
        }
    }
    if (mid == start || mid == end) {
        mid = start + (end - start) / 2;
        if (b->pack_leaves) { /* one branch a multiple of the leaf size (bvh.inl:132-136) */
            while ((mid - start) % b->leaf_size != 0 && mid < end) mid++;
            if (mid == end) mid = start + (end - start) / 2;
        }
    }
    return mid;
}

static int push_node(build_ctx *b, const node_t *n)
{
    oracle_bvh *h = b->h;
    if (h->n_nodes == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        h->nodes = (node_t *)realloc(h->nodes, (size_t)b->cap * sizeof(node_t));
    }
    h->nodes[h->n_nodes] = *n;
    return h->n_nodes++;
}

/* buildRecursive (bvh.inl:142-209); the parent's secondChildOffset is set when
   the right child is created, which is what the "touched twice" bookkeeping does */
static void build_recursive(build_ctx *b, int parent, int is_right, int start, int end, int depth)
{
    oracle_bvh *h = b->h;
    int dim = h->dim;
    if (depth > h->max_depth) h->max_depth = depth;
    box_t bb, bc;
    box_init(&bb);
    box_init(&bc);
    for (int p = start; p < end; p++) {
        box_expand_box(&bb, &b->ref_boxes[p], dim);
        box_expand_point(&bc, &b->ref_cent[(size_t)p * dim], dim);
    }
    node_t node;
    node.box = bb;
    int nrefs = end - start;
    int leaf = (nrefs <= b->leaf_size || depth == BVH_MAX_DEPTH - 2);
    if (leaf) {
        node.offset = start;
        node.nrefs = nrefs;
        h->n_leaves++;
    } else {
        node.offset = 0;
        node.nrefs = 0;
    }
    int cur = push_node(b, &node);
    if (parent >= 0 && is_right) h->nodes[parent].offset = cur - parent;
    if (leaf) return;
    int split_dim;
    float split_coord;
    object_split(b, &bb, &bc, start, end, depth, &split_dim, &split_coord);
    int mid = object_partition(b, start, end, split_dim, split_coord);
    build_recursive(b, cur, 0, start, mid, depth + 1);
    build_recursive(b, cur, 1, mid, end, depth + 1);
}

/* sortTriangleSoupPositions / sortLineSegmentSoupPositions (fcpw.inl:613-689) */
static void sort_soup_positions(oracle_bvh *h, const float *user_pos, const int32_t *user_idx)
{
    int V = h->n_vertices, dim = h->dim, k = h->k;
    for (int v = 0; v < V; v++) h->index_map[v] = -1;
    int next = 0;
    for (int i = 0; i < h->n_nodes; i++) {
        const node_t *n = &h->nodes[i];
        for (int j = 0; j < n->nrefs; j++) {
            int slot = n->offset + j;
            int orig = h->prim_orig[slot];
            for (int c = 0; c < k; c++) {
                int vi = user_idx[(size_t)orig * k + c];
                if (h->index_map[vi] == -1) {
                    memcpy(h->positions + (size_t)next * dim, user_pos + (size_t)vi * dim, dim * sizeof(float));
                    h->index_map[vi] = next++;
                }
                h->prim_vidx[(size_t)slot * k + c] = h->index_map[vi];
            }
        }
    }
}

oracle_bvh *oracle_build(int dim, const float *positions, int n_vertices, const int32_t *indices, int n_prims,
                         int heuristic, int leaf_size, int n_buckets)
{
    return oracle_build_packed(dim, positions, n_vertices, indices, n_prims, heuristic, leaf_size, n_buckets, 0);
}

oracle_bvh *oracle_build_packed(int dim, const float *positions, int n_vertices, const int32_t *indices,
                                int n_prims, int heuristic, int leaf_size, int n_buckets, int pack_leaves)
{
    if ((dim != 2 && dim != 3) || n_prims <= 0 || n_vertices <= 0) return NULL;
    oracle_bvh *h = (oracle_bvh *)calloc(1, sizeof(oracle_bvh));
    h->dim = dim;
    h->k = dim;
    h->n_vertices = n_vertices;
    h->n_prims = n_prims;
    h->positions = (float *)calloc((size_t)n_vertices * dim, sizeof(float));
    h->prim_vidx = (int32_t *)malloc((size_t)n_prims * h->k * sizeof(int32_t));
    h->prim_orig = (int32_t *)malloc((size_t)n_prims * sizeof(int32_t));
    h->index_map = (int32_t *)malloc((size_t)n_vertices * sizeof(int32_t));

    build_ctx b;
    memset(&b, 0, sizeof(b));
    b.h = h;
    b.heuristic = heuristic;
    b.leaf_size = leaf_size;
    b.n_buckets = n_buckets;
    b.pack_leaves = pack_leaves;
    b.depth_guess = (int)log2((double)n_prims); /* depthGuess = std::log2(primitives.size()) */
    b.ref_boxes = (box_t *)malloc((size_t)n_prims * sizeof(box_t));
    b.ref_cent = (float *)malloc((size_t)n_prims * dim * sizeof(float));
    b.bkt_box = (box_t *)malloc((size_t)n_buckets * sizeof(box_t));
    b.rbkt_box = (box_t *)malloc((size_t)n_buckets * sizeof(box_t));
    b.bkt_cnt = (int *)malloc((size_t)n_buckets * sizeof(int));
    b.rbkt_cnt = (int *)malloc((size_t)n_buckets * sizeof(int));
    for (int i = 0; i < n_prims; i++) {
        const float *v[3];
        for (int c = 0; c < h->k; c++) v[c] = positions + (size_t)indices[(size_t)i * h->k + c] * dim;
        h->prim_orig[i] = i;
        b.ref_boxes[i] = prim_box(v, h->k, dim);
        prim_centroid(v, h->k, dim, b.ref_cent + (size_t)i * dim);
    }
    build_recursive(&b, -1, 0, 0, n_prims, 0);
    sort_soup_positions(h, positions, indices);
    free(b.ref_boxes); free(b.ref_cent); free(b.bkt_box); free(b.rbkt_box); free(b.bkt_cnt); free(b.rbkt_cnt);
    return h;
}

static void sil_free(struct sil_data *s);

void oracle_free(oracle_bvh *h)
{
    if (!h) return;
    sil_free(h->sil);
    free(h->positions); free(h->prim_vidx); free(h->prim_orig); free(h->index_map); free(h->nodes);
    free(h);
}

int oracle_num_nodes(const oracle_bvh *h) { return h->n_nodes; }
int oracle_dim(const oracle_bvh *h) { return h->dim; }
int oracle_num_prims(const oracle_bvh *h) { return h->n_prims; }
float oracle_box_surface_area(const float *pmin, const float *pmax, int dim)
{
    box_t b;
    for (int d = 0; d < 3; d++) {
        b.pmin[d] = d < dim ? pmin[d] : 0.0f;
        b.pmax[d] = d < dim ? pmax[d] : 0.0f;
    }
    return box_surface_area(&b, dim);
}
void oracle_export_prim_vertices(const oracle_bvh *h, float *out)
{
    for (int s = 0; s < h->n_prims; s++)
        for (int c = 0; c < h->k; c++)
            memcpy(out + ((size_t)s * h->k + c) * h->dim,
                   h->positions + (size_t)h->prim_vidx[(size_t)s * h->k + c] * h->dim, h->dim * sizeof(float));
}
int oracle_num_leaves(const oracle_bvh *h) { return h->n_leaves; }
int oracle_max_depth(const oracle_bvh *h) { return h->max_depth; }

void oracle_export_nodes(const oracle_bvh *h, float *pmin, float *pmax, int32_t *nrefs, int32_t *offset)
{
    for (int i = 0; i < h->n_nodes; i++) {
        for (int d = 0; d < 3; d++) {
            pmin[3 * i + d] = d < h->dim ? h->nodes[i].box.pmin[d] : 0.0f;
            pmax[3 * i + d] = d < h->dim ? h->nodes[i].box.pmax[d] : 0.0f;
        }
        nrefs[i] = h->nodes[i].nrefs;
        offset[i] = h->nodes[i].offset;
    }
}
void oracle_export_prim_order(const oracle_bvh *h, int32_t *prim_order)
{
    memcpy(prim_order, h->prim_orig, (size_t)h->n_prims * sizeof(int32_t));
}
void oracle_export_index_map(const oracle_bvh *h, int32_t *index_map)
{
    memcpy(index_map, h->index_map, (size_t)h->n_vertices * sizeof(int32_t));
}

/* ----------------------------------------------------------------------------
 * refit (bvh.inl:548-599) after Scene::updateObjectVertices (fcpw.inl:1128-1143)
 * ------------------------------------------------------------------------- */
void oracle_update_vertices(oracle_bvh *h, const float *pos, int n_vertices)
{
    for (int v = 0; v < n_vertices && v < h->n_vertices; v++) {
        int s = h->index_map[v];
        if (s >= 0) memcpy(h->positions + (size_t)s * h->dim, pos + (size_t)v * h->dim, h->dim * sizeof(float));
    }
}
static void refit_recursive(oracle_bvh *h, int ni)
{
    node_t *n = &h->nodes[ni];
    if (n->nrefs == 0) {
        refit_recursive(h, ni + 1);
        refit_recursive(h, ni + n->offset);
        n->box = h->nodes[ni + 1].box;
        box_expand_box(&n->box, &h->nodes[ni + n->offset].box, h->dim);
    } else {
        box_init(&n->box);
        for (int p = 0; p < n->nrefs; p++) {
            const float *v[3];
            for (int c = 0; c < h->k; c++) v[c] = vtx(h, n->offset + p, c);
            box_t pb = prim_box(v, h->k, h->dim);
            box_expand_box(&n->box, &pb, h->dim);
        }
    }
}
void oracle_refit(oracle_bvh *h)
{
    if (h->n_nodes > 0) refit_recursive(h, 0);
}

/* ----------------------------------------------------------------------------
 * thread pool: static partition over nthreads (fcpw.inl:1547-1559)
 * ------------------------------------------------------------------------- */
typedef void (*range_fn)(void *ctx, int start, int end);
typedef struct { range_fn fn; void *ctx; int start, end; } range_job;
static void *range_thread(void *arg)
{
    range_job *j = (range_job *)arg;
    j->fn(j->ctx, j->start, j->end);
    return NULL;
}
static void parallel_for(int n, int nthreads, range_fn fn, void *ctx)
{
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads < 1) nthreads = 1;
    if (nthreads == 1 || n < 2 * nthreads) { fn(ctx, 0, n); return; }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    range_job *jobs = (range_job *)malloc(sizeof(range_job) * nthreads);
    int per = n / nthreads;
    for (int i = 0; i < nthreads; i++) {
        jobs[i].fn = fn; jobs[i].ctx = ctx;
        jobs[i].start = i * per;
        jobs[i].end = (i == nthreads - 1) ? n : (i + 1) * per;
        pthread_create(&th[i], NULL, range_thread, &jobs[i]);
    }
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
    free(jobs);
}

static void hit_default(oracle_hit *o)
{
    memset(o, 0, sizeof(*o));
    o->d = MAXF;
    o->index = -1;
}

typedef struct { int node; float dist; } stack_entry;

/* ----------------------------------------------------------------------------
 * CPU closest point (bvh.inl:1290-1418 + Triangle/LineSegment::findClosestPoint)
 * ------------------------------------------------------------------------- */
typedef struct {
    const oracle_bvh *h;
    const float *q, *r2;
    int record_normal;
    oracle_hit *out;
    int32_t *n_internal, *n_leaf_prims, *n_pops;
} cp_ctx;

static void cp_cpu_range(void *vctx, int start, int end)
{
    cp_ctx *c = (cp_ctx *)vctx;
    const oracle_bvh *h = c->h;
    int dim = h->dim;
    stack_entry st[BVH_MAX_DEPTH];
    for (int qi = start; qi < end; qi++) {
        float x[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) x[d] = c->q[(size_t)qi * dim + d];
        float r2 = c->r2[qi];
        oracle_hit best;
        hit_default(&best);
        int best_slot = -1;
        float h0, h1, h2, h3;
        box_sqdist_cpu(&h->nodes[0].box, x, dim, &h0, &h1);
        if (h0 <= r2) {
            r2 = smin(r2, h1);
            int sp = 0;
            st[0].node = 0;
            st[0].dist = h0;
            while (sp >= 0) {
                int ni = st[sp].node;
                float dist = st[sp].dist;
                sp--;
                if (dist > r2) continue;
                const node_t *n = &h->nodes[ni];
                if (n->nrefs > 0) {
                    for (int p = 0; p < n->nrefs; p++) {
                        int slot = n->offset + p;
                        float pt[3], uv[2];
                        float dd = prim_closest(h, slot, x, pt, uv, edot);
                        if (dd * dd <= r2) {
                            r2 = smin(r2, dd * dd);
                            hit_default(&best);
                            best.d = dd;
                            memcpy(best.p, pt, sizeof(pt));
                            memcpy(best.uv, uv, sizeof(uv));
                            best.index = h->prim_orig[slot];
                            best_slot = slot;
                        }
                    }
                } else {
                    int l = ni + 1, r = ni + n->offset;
                    box_sqdist_cpu(&h->nodes[l].box, x, dim, &h0, &h1);
                    int hit0 = h0 <= r2;
                    r2 = smin(r2, h1);
                    box_sqdist_cpu(&h->nodes[r].box, x, dim, &h2, &h3);
                    int hit1 = h2 <= r2;
                    r2 = smin(r2, h3);
                    if (hit0 && hit1) {
                        int closer = l, other = r;
                        if (h0 == 0.0f && h2 == 0.0f) {
                            if (h3 < h1) { closer = r; other = l; }
                        } else if (h2 < h0) {
                            float t = h0; h0 = h2; h2 = t;
                            closer = r; other = l;
                        }
                        st[++sp].node = other; st[sp].dist = h2;
                        st[++sp].node = closer; st[sp].dist = h0;
                    } else if (hit0) {
                        st[++sp].node = l; st[sp].dist = h0;
                    } else if (hit1) {
                        st[++sp].node = r; st[sp].dist = h2;
                    }
                }
            }
        }
        if (best_slot >= 0 && c->record_normal) prim_normal_cpu(h, best_slot, best.n);
        c->out[qi] = best;
    }
}

void oracle_closest_points_cpu(const oracle_bvh *h, const float *queries, const float *r2, int n,
                               int record_normal, int nthreads, oracle_hit *out)
{
    cp_ctx c = {h, queries, r2, record_normal, out, NULL, NULL, NULL};
    parallel_for(n, nthreads, cp_cpu_range, &c);
}

/* ----------------------------------------------------------------------------
 * GPU-order closest point (bvh.slang:395-502), with traversal counters
 * ------------------------------------------------------------------------- */
/* One query: x (3 floats, z = 0 in 2D) against h in the GPU order.  r2 is
   updated in place, as the reference's inout BoundingSphere; returns the leaf
   slot of the result (best filled) or -1 (best untouched). */
static int cp_gpu_one(const oracle_bvh *h, const float *x, float *r2io, oracle_hit *best, int *n_int_out,
                      int *n_prim_out, int *n_pop_out)
{
    int dim = h->dim;
    int st[BVH_MAX_DEPTH];
    float r2 = *r2io;
    int best_slot = -1, n_int = 0, n_prim = 0, n_pop = 0;
    float d0, d1, d2, d3;
    box_sqdist_gpu(&h->nodes[0].box, x, dim, &d0, &d1);
    if (d0 <= r2) {
        r2 = fminf(r2, d1);
        int sp = 0;
        st[0] = 0;
        while (sp >= 0) {
            int ni = st[sp--];
            const node_t *n = &h->nodes[ni];
            n_pop++;
            box_sqdist_gpu(&n->box, x, dim, &d0, &d1);
            if (!(d0 <= r2)) continue;
            if (n->nrefs > 0) {
                for (int p = 0; p < n->nrefs; p++) {
                    int slot = n->offset + p;
                    float pt[3], uv[2];
                    n_prim++;
                    float dd = prim_closest(h, slot, x, pt, uv, gdot);
                    if (dd * dd <= r2) {
                        r2 = fminf(r2, dd * dd);
                        hit_default(best);
                        best->d = dd;
                        memcpy(best->p, pt, sizeof(pt));
                        memcpy(best->uv, uv, sizeof(uv));
                        best->index = h->prim_orig[slot];
                        best_slot = slot;
                    }
                }
            } else {
                n_int++;
                int l = ni + 1, r = ni + n->offset;
                box_sqdist_gpu(&h->nodes[l].box, x, dim, &d0, &d1);
                int ol = d0 <= r2;
                r2 = fminf(r2, d1);
                box_sqdist_gpu(&h->nodes[r].box, x, dim, &d2, &d3);
                int orr = d2 <= r2;
                r2 = fminf(r2, d3);
                if (ol && orr) {
                    int closer = l, other = r;
                    if (d2 < d0) { closer = r; other = l; }
                    st[++sp] = other;
                    st[++sp] = closer;
                } else if (ol) {
                    st[++sp] = l;
                } else if (orr) {
                    st[++sp] = r;
                }
            }
        }
    }
    *r2io = r2;
    if (n_int_out) *n_int_out = n_int;
    if (n_prim_out) *n_prim_out = n_prim;
    if (n_pop_out) *n_pop_out = n_pop;
    return best_slot;
}

static void cp_gpu_range(void *vctx, int start, int end)
{
    cp_ctx *c = (cp_ctx *)vctx;
    const oracle_bvh *h = c->h;
    int dim = h->dim;
    for (int qi = start; qi < end; qi++) {
        float x[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) x[d] = c->q[(size_t)qi * dim + d];
        float r2 = c->r2[qi];
        oracle_hit best;
        hit_default(&best);
        int n_int, n_prim, n_pop;
        int best_slot = cp_gpu_one(h, x, &r2, &best, &n_int, &n_prim, &n_pop);
        if (best_slot >= 0 && c->record_normal) prim_normal_gpu(h, best_slot, best.n);
        c->out[qi] = best;
        if (c->n_internal) c->n_internal[qi] = n_int;
        if (c->n_leaf_prims) c->n_leaf_prims[qi] = n_prim;
        if (c->n_pops) c->n_pops[qi] = n_pop;
    }
}

void oracle_closest_points_gpuorder(const oracle_bvh *h, const float *queries, const float *r2, int n,
                                    int record_normal, int nthreads, oracle_hit *out, int32_t *n_internal,
                                    int32_t *n_leaf_prims, int32_t *n_pops)
{
    cp_ctx c = {h, queries, r2, record_normal, out, n_internal, n_leaf_prims, n_pops};
    parallel_for(n, nthreads, cp_gpu_range, &c);
}

/* ----------------------------------------------------------------------------
 * Baseline aggregate (baseline.inl:342-384): scene order, FP32, Eigen order
 * ------------------------------------------------------------------------- */
typedef struct {
    const oracle_bvh *h;
    const float *q, *r2;
    oracle_hit *out;
    int32_t *slot_of; /* original prim -> leaf slot */
} bf_ctx;

static void bf_range(void *vctx, int start, int end)
{
    bf_ctx *c = (bf_ctx *)vctx;
    const oracle_bvh *h = c->h;
    int dim = h->dim;
    for (int qi = start; qi < end; qi++) {
        float x[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) x[d] = c->q[(size_t)qi * dim + d];
        float r2 = c->r2[qi];
        oracle_hit best;
        hit_default(&best);
        for (int p = 0; p < h->n_prims; p++) {
            int slot = c->slot_of[p];
            float pt[3], uv[2];
            float dd = prim_closest(h, slot, x, pt, uv, edot);
            if (dd * dd <= r2) {
                r2 = smin(r2, dd * dd);
                best.d = dd;
                memcpy(best.p, pt, sizeof(pt));
                memcpy(best.uv, uv, sizeof(uv));
                best.index = p;
            }
        }
        c->out[qi] = best;
    }
}

void oracle_brute_closest_f32(const oracle_bvh *h, const float *queries, const float *r2, int n, int nthreads,
                              oracle_hit *out)
{
    int32_t *slot_of = (int32_t *)malloc((size_t)h->n_prims * sizeof(int32_t));
    for (int s = 0; s < h->n_prims; s++) slot_of[h->prim_orig[s]] = s;
    bf_ctx c = {h, queries, r2, out, slot_of};
    parallel_for(n, nthreads, bf_range, &c);
    free(slot_of);
}

/* ----------------------------------------------------------------------------
 * rays: CPU order (bvh.inl:655-771) and GPU order (bvh.slang:144-254)
 * ------------------------------------------------------------------------- */
typedef struct {
    const oracle_bvh *h;
    const float *o, *d, *tmax;
    int occlusion;
    oracle_hit *out;
    int32_t *n_internal, *n_leaf_prims;
} ray_ctx;

static void ray_cpu_range(void *vctx, int start, int end)
{
    ray_ctx *c = (ray_ctx *)vctx;
    const oracle_bvh *h = c->h;
    int dim = h->dim;
    stack_entry st[BVH_MAX_DEPTH];
    for (int qi = start; qi < end; qi++) {
        float o[3] = {0, 0, 0}, d[3] = {0, 0, 0}, invd[3] = {0, 0, 0};
        for (int k = 0; k < dim; k++) {
            o[k] = c->o[(size_t)qi * dim + k];
            d[k] = c->d[(size_t)qi * dim + k];
            invd[k] = 1.0f / d[k]; /* Ray ctor: d.cwiseInverse() (ray.h:10-11) */
        }
        float tmax = c->tmax[qi];
        oracle_hit best;
        hit_default(&best);
        float b0, b1, b2, b3;
        if (box_ray_cpu(&h->nodes[0].box, o, invd, tmax, dim, &b0, &b1)) {
            int sp = 0, done = 0;
            st[0].node = 0;
            st[0].dist = b0;
            while (sp >= 0 && !done) {
                int ni = st[sp].node;
                float dist = st[sp].dist;
                sp--;
                if (dist > tmax) continue;
                const node_t *n = &h->nodes[ni];
                if (n->nrefs > 0) {
                    for (int p = 0; p < n->nrefs; p++) {
                        int slot = n->offset + p, occl = 0, hit;
                        oracle_hit cand = best;
                        if (dim == 3)
                            hit = ray_triangle_cpu(vtx(h, slot, 0), vtx(h, slot, 1), vtx(h, slot, 2), o, d, tmax,
                                                   c->occlusion, &cand, &occl);
                        else
                            hit = ray_segment(vtx(h, slot, 0), vtx(h, slot, 1), o, d, tmax, c->occlusion, 1, &cand,
                                              &occl);
                        if (hit) {
                            if (c->occlusion) { done = 1; break; }
                            cand.index = h->prim_orig[slot];
                            best = cand;
                            tmax = smin(tmax, best.d);
                        }
                    }
                } else {
                    int l = ni + 1, r = ni + n->offset;
                    int hit0 = box_ray_cpu(&h->nodes[l].box, o, invd, tmax, dim, &b0, &b1);
                    int hit1 = box_ray_cpu(&h->nodes[r].box, o, invd, tmax, dim, &b2, &b3);
                    if (hit0 && hit1) {
                        int closer = l, other = r;
                        if (b2 < b0) {
                            float t = b0; b0 = b2; b2 = t;
                            t = b1; b1 = b3; b3 = t;
                            closer = r; other = l;
                        }
                        st[++sp].node = other; st[sp].dist = b2;
                        st[++sp].node = closer; st[sp].dist = b0;
                    } else if (hit0) {
                        st[++sp].node = l; st[sp].dist = b0;
                    } else if (hit1) {
                        st[++sp].node = r; st[sp].dist = b2;
                    }
                }
            }
        }
        c->out[qi] = best;
    }
}

void oracle_intersect_cpu(const oracle_bvh *h, const float *origins, const float *dirs, const float *tmax, int n,
                          int occlusion, int nthreads, oracle_hit *out)
{
    ray_ctx c = {h, origins, dirs, tmax, occlusion, out, NULL, NULL};
    parallel_for(n, nthreads, ray_cpu_range, &c);
}

/* One ray (o, d, dinv: 3 floats, z = 0 in 2D) against h in the GPU order; tmax
   is updated in place (the reference's inout Ray).  Returns the leaf slot of
   the hit (best filled; with occlusion only its index) or -1. */
static int ray_gpu_one(const oracle_bvh *h, const float *o, const float *d, const float *dinv, float *tmax_io,
                       int occlusion, oracle_hit *best, int *n_int_out, int *n_prim_out)
{
    int dim = h->dim;
    stack_entry st[BVH_MAX_DEPTH];
    float tmax = *tmax_io;
    int best_slot = -1, n_int = 0, n_prim = 0;
    float b0, b1, b2, b3;
    if (box_ray_gpu(&h->nodes[0].box, o, d, dinv, tmax, dim, &b0, &b1)) {
        int sp = 0, done = 0;
        st[0].node = 0;
        st[0].dist = b0;
        while (sp >= 0 && !done) {
            int ni = st[sp].node;
            float dist = st[sp].dist;
            sp--;
            if (dist > tmax) continue;
            const node_t *n = &h->nodes[ni];
            if (n->nrefs > 0) {
                for (int p = 0; p < n->nrefs; p++) {
                    int slot = n->offset + p, hit;
                    oracle_hit cand;
                    hit_default(&cand);
                    n_prim++;
                    if (dim == 3)
                        hit = ray_triangle_gpu(vtx(h, slot, 0), vtx(h, slot, 1), vtx(h, slot, 2), o, d, tmax,
                                               occlusion, &cand);
                    else
                        hit = ray_segment(vtx(h, slot, 0), vtx(h, slot, 1), o, d, tmax, occlusion, 0, &cand, NULL);
                    if (hit) {
                        best_slot = slot;
                        if (occlusion) {
                            /* bvh.slang:183-187: only the index of the occluder is written */
                            best->index = h->prim_orig[slot];
                            done = 1;
                            break;
                        }
                        cand.index = h->prim_orig[slot];
                        tmax = fminf(tmax, cand.d);
                        *best = cand;
                    }
                }
            } else {
                n_int++;
                int l = ni + 1, r = ni + n->offset;
                int hl = box_ray_gpu(&h->nodes[l].box, o, d, dinv, tmax, dim, &b0, &b1);
                int hr = box_ray_gpu(&h->nodes[r].box, o, d, dinv, tmax, dim, &b2, &b3);
                if (hl && hr) {
                    int closer = l, other = r;
                    if (b2 < b0) {
                        float t = b0; b0 = b2; b2 = t;
                        closer = r; other = l;
                    }
                    st[++sp].node = other; st[sp].dist = b2;
                    st[++sp].node = closer; st[sp].dist = b0;
                } else if (hl) {
                    st[++sp].node = l; st[sp].dist = b0;
                } else if (hr) {
                    st[++sp].node = r; st[sp].dist = b2;
                }
            }
        }
    }
    *tmax_io = tmax;
    if (n_int_out) *n_int_out = n_int;
    if (n_prim_out) *n_prim_out = n_prim;
    return best_slot;
}

static void ray_gpu_range(void *vctx, int start, int end)
{
    ray_ctx *c = (ray_ctx *)vctx;
    const oracle_bvh *h = c->h;
    int dim = h->dim;
    for (int qi = start; qi < end; qi++) {
        float o[3] = {0, 0, 0}, d[3] = {0, 0, 0}, dinv[3] = {0, 0, 0};
        for (int k = 0; k < dim; k++) {
            o[k] = c->o[(size_t)qi * dim + k];
            d[k] = c->d[(size_t)qi * dim + k];
            dinv[k] = 1.0f / d[k]; /* GPURay dInv (bvh_interop_structures.h:737-741) */
        }
        float tmax = c->tmax[qi];
        oracle_hit best;
        hit_default(&best);
        int n_int, n_prim;
        ray_gpu_one(h, o, d, dinv, &tmax, c->occlusion, &best, &n_int, &n_prim);
        c->out[qi] = best;
        if (c->n_internal) c->n_internal[qi] = n_int;
        if (c->n_leaf_prims) c->n_leaf_prims[qi] = n_prim;
    }
}

void oracle_intersect_gpuorder(const oracle_bvh *h, const float *origins, const float *dirs, const float *tmax,
                               int n, int occlusion, int nthreads, oracle_hit *out, int32_t *n_internal,
                               int32_t *n_leaf_prims)
{
    ray_ctx c = {h, origins, dirs, tmax, occlusion, out, n_internal, n_leaf_prims};
    parallel_for(n, nthreads, ray_gpu_range, &c);
}

/* ----------------------------------------------------------------------------
 * Sphere intersection, GPU semantics (bvh.slang:257-392) with the constant
 * branch weight of the sphereIntersection entry point (bvh.cs.slang:59-82,
 * aggregate.slang:13-20): one random walk from the root, choosing among the
 * children that overlap the sphere with equal probability, then an
 * area-weighted reservoir choice among the leaf's primitives inside the
 * sphere, then a uniform point on the chosen primitive.  Primitive area,
 * sphere test and point sampling follow geometry.slang (LineSegment
 * 133-136,152-162,180-191; Triangle 356-359,380-391,409-423).
 * ------------------------------------------------------------------------- */
static float prim_area_gpu(const oracle_bvh *h, int slot)
{
    if (h->dim == 3) {
        const float *pa = vtx(h, slot, 0), *pb = vtx(h, slot, 1), *pc = vtx(h, slot, 2);
        float a[3], b[3], n[3];
        for (int i = 0; i < 3; i++) { a[i] = pb[i] - pa[i]; b[i] = pc[i] - pa[i]; }
        n[0] = a[1] * b[2] - a[2] * b[1];
        n[1] = a[2] * b[0] - a[0] * b[2];
        n[2] = a[0] * b[1] - a[1] * b[0];
        return 0.5f * sqrtf(gdot(n, n, 3));
    }
    const float *pa = vtx(h, slot, 0), *pb = vtx(h, slot, 1);
    float sv[2] = {pb[0] - pa[0], pb[1] - pa[1]};
    return sqrtf(gdot(sv, sv, 2));
}

/* samplePoint: writes uv, p, n and returns the sampling pdf */
static float prim_sample_gpu(const oracle_bvh *h, int slot, float r1, float r2, float *uv, float *p, float *n)
{
    if (h->dim == 3) {
        const float *pa = vtx(h, slot, 0), *pb = vtx(h, slot, 1), *pc = vtx(h, slot, 2);
        float a[3], b[3];
        for (int i = 0; i < 3; i++) { a[i] = pb[i] - pa[i]; b[i] = pc[i] - pa[i]; }
        n[0] = a[1] * b[2] - a[2] * b[1];
        n[1] = a[2] * b[0] - a[0] * b[2];
        n[2] = a[0] * b[1] - a[1] * b[0];
        float area = sqrtf(gdot(n, n, 3));
        float u1 = sqrtf(r1), u = 1.0f - u1, v = r2 * u1, w = (1.0f - u) - v;
        uv[0] = u;
        uv[1] = v;
        for (int i = 0; i < 3; i++) {
            p[i] = (pa[i] * u + pb[i] * v) + pc[i] * w;
            n[i] = n[i] / area;
        }
        return 2.0f / area;
    }
    const float *pa = vtx(h, slot, 0), *pb = vtx(h, slot, 1);
    float sv[2] = {pb[0] - pa[0], pb[1] - pa[1]};
    float area = sqrtf(gdot(sv, sv, 2));
    uv[0] = r1;
    uv[1] = 0.0f;
    p[0] = pa[0] + r1 * sv[0];
    p[1] = pa[1] + r1 * sv[1];
    p[2] = 0.0f;
    n[0] = sv[1] / area;
    n[1] = -sv[0] / area;
    n[2] = 0.0f;
    return 1.0f / area;
}

typedef struct {
    const oracle_bvh *h;
    const float *spheres, *rnd; /* n x 4 (c, r2) and n x 3 */
    oracle_hit *out;
} sph_ctx;

/* One sphere (x, r2) with random numbers rn[0] (walk and choice), rn[1], rn[2]
   (the point); returns 1 and fills out, or 0 with out the default record. */
static int sph_gpu_one(const oracle_bvh *h, const float *x, float r2, const float *rn, oracle_hit *out)
{
    int dim = h->dim;
    oracle_hit it;
    hit_default(&it);
    int selected = -1, did = 0;
    float d0, d1, d2, d3;
    box_sqdist_gpu(&h->nodes[0].box, x, dim, &d0, &d1);
    if (d0 <= r2) {
        float maxd = d1, pdf = 1.0f, u = rn[0];
        int cur = 0;
        for (;;) {
            const node_t *nd = &h->nodes[cur];
            if (nd->nrefs > 0) {
                float total = 0.0f;
                for (int k = 0; k < nd->nrefs; k++) {
                    int slot = nd->offset + k, hitp = 0;
                    oracle_hit cand;
                    hit_default(&cand);
                    if (maxd <= r2) {
                        hitp = 1;
                    } else {
                        float pt[3], uv[2];
                        float dd = prim_closest(h, slot, x, pt, uv, gdot);
                        if (dd * dd <= r2) {
                            hitp = 1;
                            memcpy(cand.p, pt, sizeof(pt));
                            memcpy(cand.uv, uv, sizeof(uv));
                        }
                    }
                    if (hitp) {
                        cand.d = prim_area_gpu(h, slot);
                        cand.index = h->prim_orig[slot];
                        did = 1;
                        total += cand.d;
                        float prob = cand.d / total;
                        if (u < prob) {
                            u = u / prob;
                            it = cand;
                            it.d *= pdf;
                            selected = slot;
                        } else {
                            u = (u - prob) / (1.0f - prob);
                        }
                    }
                }
                if (total > 0.0f) it.d /= total;
                break;
            }
            int l = cur + 1, r = cur + nd->offset;
            box_sqdist_gpu(&h->nodes[l].box, x, dim, &d0, &d1);
            float wl = (d0 <= r2) ? 1.0f : 0.0f;
            box_sqdist_gpu(&h->nodes[r].box, x, dim, &d2, &d3);
            float wr = (d2 <= r2) ? 1.0f : 0.0f;
            float tot = wl + wr;
            if (!(tot > 0.0f)) break;
            float pl = wl / tot, pr = 1.0f - pl;
            if (u < pl) {
                u = u / pl;
                cur = l;
                pdf *= pl;
                maxd = d1;
            } else {
                u = (u - pl) / pr;
                cur = r;
                pdf *= pr;
                maxd = d3;
            }
        }
    }
    if (did && (it.index == -1 || selected < 0)) did = 0;
    if (did) {
        float spdf = prim_sample_gpu(h, selected, rn[1], rn[2], it.uv, it.p, it.n);
        it.d *= spdf;
    } else {
        hit_default(&it);
    }
    *out = it;
    return did;
}

static void sph_range(void *vctx, int start, int end)
{
    sph_ctx *c = (sph_ctx *)vctx;
    const oracle_bvh *h = c->h;
    int dim = h->dim;
    for (int qi = start; qi < end; qi++) {
        float x[3] = {c->spheres[(size_t)qi * 4], c->spheres[(size_t)qi * 4 + 1], 0.0f};
        if (dim == 3) x[2] = c->spheres[(size_t)qi * 4 + 2];
        sph_gpu_one(h, x, c->spheres[(size_t)qi * 4 + 3], c->rnd + (size_t)qi * 3, &c->out[qi]);
    }
}

void oracle_intersect_spheres(const oracle_bvh *h, const float *spheres, const float *rand_nums, int n,
                              int nthreads, oracle_hit *out)
{
    sph_ctx c = {h, spheres, rand_nums, out};
    parallel_for(n, nthreads, sph_range, &c);
}

/* ----------------------------------------------------------------------------
 * FP64 brute force, independent method: project onto the supporting line or
 * plane, keep the projection when it falls inside the primitive, otherwise take
 * the nearest boundary point (edge clamping).  No code is shared with the
 * region test above, so an error in either shows up as a disagreement.
 * ------------------------------------------------------------------------- */
static double seg_dist_f64(const double *a, const double *b, const double *x, int dim, double *p, double *t_out)
{
    double ab[3] = {0, 0, 0}, ax[3] = {0, 0, 0};
    double L = 0.0, s = 0.0;
    for (int i = 0; i < dim; i++) { ab[i] = b[i] - a[i]; ax[i] = x[i] - a[i]; L += ab[i] * ab[i]; s += ab[i] * ax[i]; }
    double t = L > 0.0 ? s / L : 0.0;
    if (t < 0.0) t = 0.0;
    if (t > 1.0) t = 1.0;
    double dd = 0.0;
    for (int i = 0; i < dim; i++) { p[i] = a[i] + t * ab[i]; double r = x[i] - p[i]; dd += r * r; }
    if (t_out) *t_out = t;
    return sqrt(dd);
}

static double tri_dist_f64(const double *a, const double *b, const double *c, const double *x, double *p)
{
    double e0[3], e1[3], n[3], ax[3];
    for (int i = 0; i < 3; i++) { e0[i] = b[i] - a[i]; e1[i] = c[i] - a[i]; ax[i] = x[i] - a[i]; }
    n[0] = e0[1] * e1[2] - e0[2] * e1[1];
    n[1] = e0[2] * e1[0] - e0[0] * e1[2];
    n[2] = e0[0] * e1[1] - e0[1] * e1[0];
    double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (nn > 0.0) {
        double s = (ax[0] * n[0] + ax[1] * n[1] + ax[2] * n[2]) / nn;
        double q[3];
        for (int i = 0; i < 3; i++) q[i] = x[i] - s * n[i];
        /* signed sub-triangle areas relative to n decide containment */
        const double *V[3] = {a, b, c};
        int inside = 1;
        for (int k = 0; k < 3 && inside; k++) {
            const double *u = V[k], *v = V[(k + 1) % 3];
            double uv[3], uq[3], cr[3];
            for (int i = 0; i < 3; i++) { uv[i] = v[i] - u[i]; uq[i] = q[i] - u[i]; }
            cr[0] = uv[1] * uq[2] - uv[2] * uq[1];
            cr[1] = uv[2] * uq[0] - uv[0] * uq[2];
            cr[2] = uv[0] * uq[1] - uv[1] * uq[0];
            if (cr[0] * n[0] + cr[1] * n[1] + cr[2] * n[2] < 0.0) inside = 0;
        }
        if (inside) {
            double dd = 0.0;
            for (int i = 0; i < 3; i++) { p[i] = q[i]; double r = x[i] - q[i]; dd += r * r; }
            return sqrt(dd);
        }
    }
    double best = INFINITY, pp[3];
    const double *V[3] = {a, b, c};
    for (int k = 0; k < 3; k++) {
        double dk = seg_dist_f64(V[k], V[(k + 1) % 3], x, 3, pp, NULL);
        if (dk < best) { best = dk; p[0] = pp[0]; p[1] = pp[1]; p[2] = pp[2]; }
    }
    return best;
}

typedef struct {
    int dim;
    const float *pos;
    const int32_t *idx;
    int np;
    const float *q, *dirs;
    double *d, *p;
    int32_t *index;
} f64_ctx;

static void load_prim_f64(const f64_ctx *c, int prim, double v[3][3])
{
    for (int k = 0; k < c->dim; k++)
        for (int i = 0; i < 3; i++)
            v[k][i] = i < c->dim ? (double)c->pos[(size_t)c->idx[(size_t)prim * c->dim + k] * c->dim + i] : 0.0;
}

static void f64_cp_range(void *vctx, int start, int end)
{
    f64_ctx *c = (f64_ctx *)vctx;
    for (int qi = start; qi < end; qi++) {
        double x[3] = {0, 0, 0}, best = INFINITY, bp[3] = {0, 0, 0}, pp[3];
        int bi = -1;
        for (int i = 0; i < c->dim; i++) x[i] = c->q[(size_t)qi * c->dim + i];
        for (int pr = 0; pr < c->np; pr++) {
            double v[3][3];
            load_prim_f64(c, pr, v);
            double dd = c->dim == 3 ? tri_dist_f64(v[0], v[1], v[2], x, pp) : seg_dist_f64(v[0], v[1], x, 2, pp, NULL);
            if (dd < best) { best = dd; bi = pr; bp[0] = pp[0]; bp[1] = pp[1]; bp[2] = c->dim == 3 ? pp[2] : 0.0; }
        }
        c->d[qi] = best;
        for (int i = 0; i < 3; i++) c->p[(size_t)qi * 3 + i] = bp[i];
        c->index[qi] = bi;
    }
}

void oracle_brute_closest_f64(int dim, const float *positions, int n_vertices, const int32_t *indices, int n_prims,
                              const float *queries, int n, int nthreads, double *d, double *p, int32_t *index)
{
    (void)n_vertices;
    f64_ctx c = {dim, positions, indices, n_prims, queries, NULL, d, p, index};
    parallel_for(n, nthreads, f64_cp_range, &c);
}

/* FP64 ray: exact plane/line intersection with an inside test by signed areas */
static void f64_ray_range(void *vctx, int start, int end)
{
    f64_ctx *c = (f64_ctx *)vctx;
    for (int qi = start; qi < end; qi++) {
        double o[3] = {0, 0, 0}, dr[3] = {0, 0, 0}, best = INFINITY;
        int bi = -1;
        for (int i = 0; i < c->dim; i++) {
            o[i] = c->q[(size_t)qi * c->dim + i];
            dr[i] = c->dirs[(size_t)qi * c->dim + i];
        }
        for (int pr = 0; pr < c->np; pr++) {
            double v[3][3] = {{0}};
            load_prim_f64(c, pr, v);
            double t = -1.0;
            if (c->dim == 3) {
                double e0[3], e1[3], n[3];
                for (int i = 0; i < 3; i++) { e0[i] = v[1][i] - v[0][i]; e1[i] = v[2][i] - v[0][i]; }
                n[0] = e0[1] * e1[2] - e0[2] * e1[1];
                n[1] = e0[2] * e1[0] - e0[0] * e1[2];
                n[2] = e0[0] * e1[1] - e0[1] * e1[0];
                double dn = dr[0] * n[0] + dr[1] * n[1] + dr[2] * n[2];
                if (dn == 0.0) continue;
                t = ((v[0][0] - o[0]) * n[0] + (v[0][1] - o[1]) * n[1] + (v[0][2] - o[2]) * n[2]) / dn;
                if (t < 0.0) continue;
                double q[3];
                for (int i = 0; i < 3; i++) q[i] = o[i] + t * dr[i];
                int inside = 1;
                for (int k = 0; k < 3 && inside; k++) {
                    const double *u = v[k], *w = v[(k + 1) % 3];
                    double uw[3], uq[3], cr[3];
                    for (int i = 0; i < 3; i++) { uw[i] = w[i] - u[i]; uq[i] = q[i] - u[i]; }
                    cr[0] = uw[1] * uq[2] - uw[2] * uq[1];
                    cr[1] = uw[2] * uq[0] - uw[0] * uq[2];
                    cr[2] = uw[0] * uq[1] - uw[1] * uq[0];
                    if (cr[0] * n[0] + cr[1] * n[1] + cr[2] * n[2] < 0.0) inside = 0;
                }
                if (!inside) continue;
            } else {
                double e[2] = {v[1][0] - v[0][0], v[1][1] - v[0][1]};
                double den = dr[0] * e[1] - dr[1] * e[0];
                if (den == 0.0) continue;
                double ao[2] = {v[0][0] - o[0], v[0][1] - o[1]};
                double s = (ao[0] * dr[1] - ao[1] * dr[0]) / den;
                t = (ao[0] * e[1] - ao[1] * e[0]) / den;
                if (s < 0.0 || s > 1.0 || t < 0.0) continue;
            }
            if (t < best) { best = t; bi = pr; }
        }
        c->d[qi] = best;
        c->index[qi] = bi;
    }
}

void oracle_brute_ray_f64(int dim, const float *positions, int n_vertices, const int32_t *indices, int n_prims,
                          const float *origins, const float *dirs, int n, int nthreads, double *t, int32_t *index)
{
    (void)n_vertices;
    f64_ctx c = {dim, positions, indices, n_prims, origins, dirs, t, NULL, index};
    parallel_for(n, nthreads, f64_ray_range, &c);
}

void oracle_prim_dist_f64(int dim, const float *positions, const int32_t *indices, const float *queries,
                          const int32_t *prims, int n, double *d)
{
    f64_ctx c = {dim, positions, indices, 0, NULL, NULL, NULL, NULL, NULL};
    c.pos = positions;
    c.idx = indices;
    for (int qi = 0; qi < n; qi++) {
        double x[3] = {0, 0, 0}, v[3][3], pp[3];
        for (int i = 0; i < dim; i++) x[i] = queries[(size_t)qi * dim + i];
        if (prims[qi] < 0) { d[qi] = INFINITY; continue; }
        load_prim_f64(&c, prims[qi], v);
        d[qi] = dim == 3 ? tri_dist_f64(v[0], v[1], v[2], x, pp) : seg_dist_f64(v[0], v[1], x, 2, pp, NULL);
    }
}

/* ============================================================================
 * Silhouettes (the SNCH variant of the BVH):
 *   - silhouette objects of Scene::computeSilhouettes (fcpw.inl:405-521) and
 *     their renumbering by sortSoupPositions (fcpw.inl:709-756);
 *   - per-leaf silhouette references and bounding cones of
 *     Bvh::assignGeometricDataToNodes (bvh.inl:236-420);
 *   - GPU records of extractSilhouetteVertices / extractSilhouetteEdges
 *     (bvh_interop_structures.h:237-300);
 *   - the GPU traversal Bvh.findClosestSilhouettePoint (bvh.slang:503-646) with
 *     Vertex/Edge.findClosestSilhouettePoint (geometry.slang:490-708) and
 *     BoundingCone.overlap (bounding-volumes.slang:122-236).
 * Host-side quantities follow the Eigen expressions (c0 + (c1 + c2) reductions,
 * normalized() leaves a zero vector unchanged); GPU-side ones use float3
 * arithmetic summed left to right with z = 0 in 2D.
 * ========================================================================= */
struct sil_data {
    int n_obj, n_refs;
    int32_t (*ind)[4];  /* per object: 2D {prev, self, next, -}; 3D {opp. of face 1, a, b, opp. of face 0} */
    int32_t *pindex;    /* per object: reported index */
    int32_t *ref_obj;   /* per reference (leaf order): object id */
    float *ref_n;       /* per reference: normal() used for the cone axis (dim floats) */
    float *ref_fn;      /* per reference: the two unit face normals (2 x dim floats) */
    float *cone;        /* per node: axis[3], halfAngle, radius */
    int32_t *off, *cnt; /* per node: silhouetteReferenceOffset, nSilhouetteReferences */
    oracle_silhouette *rec; /* per reference: GPU record */
    int32_t *eidx;      /* 3D: edge id of (original triangle, local edge) */
};

static void sil_free(struct sil_data *s)
{
    if (!s) return;
    free(s->ind); free(s->pindex); free(s->ref_obj); free(s->ref_n); free(s->ref_fn); free(s->cone);
    free(s->off); free(s->cnt); free(s->rec); free(s->eidx);
    free(s);
}

static const float *spos(const oracle_bvh *h, int v) { return h->positions + (size_t)v * h->dim; }

/* Eigen normalized(): v / sqrt(squaredNorm) when squaredNorm > 0 */
static void enormalize(float *v, int dim)
{
    float z = edot(v, v, dim);
    if (z > 0.0f) {
        float r = sqrtf(z);
        for (int i = 0; i < dim; i++) v[i] = v[i] / r;
    }
}
static void ecross(const float *a, const float *b, float *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static int sil_has_face(const oracle_bvh *h, const int32_t *ind, int f)
{
    if (h->dim == 2) return f == 0 ? ind[2] != -1 : ind[0] != -1;
    return f == 0 ? ind[3] != -1 : ind[0] != -1;
}
/* SilhouetteVertex::normal(fIndex, normalize) (silhouette_vertices.inl:37-47),
   SilhouetteEdge::normal(fIndex, normalize) (silhouette_edges.inl:47-68) */
static void sil_face_normal(const oracle_bvh *h, const int32_t *ind, int f, int normalize, float *n)
{
    if (h->dim == 2) {
        int i = f == 0 ? 1 : 0;
        const float *pa = spos(h, ind[i]), *pb = spos(h, ind[i + 1]);
        float s0 = pb[0] - pa[0], s1 = pb[1] - pa[1];
        n[0] = s1; n[1] = -s0;
        if (normalize) enormalize(n, 2);
        return;
    }
    int i = f == 0 ? 3 : 0, j = f == 0 ? 1 : 2, k = f == 0 ? 2 : 1;
    const float *pa = spos(h, ind[j]), *pb = spos(h, ind[k]), *pc = spos(h, ind[i]);
    float v1[3], v2[3];
    for (int c = 0; c < 3; c++) { v1[c] = pb[c] - pa[c]; v2[c] = pc[c] - pa[c]; }
    ecross(v1, v2, n);
    if (normalize) enormalize(n, 3);
}
/* normal() without precomputed soup normals: the normalized sum of the
   unnormalized adjacent face normals */
static void sil_normal(const oracle_bvh *h, const int32_t *ind, float *n)
{
    int dim = h->dim;
    float t[3];
    n[0] = n[1] = n[2] = 0.0f;
    for (int f = 0; f < 2; f++) {
        if (!sil_has_face(h, ind, f)) continue;
        sil_face_normal(h, ind, f, 0, t);
        for (int c = 0; c < dim; c++) n[c] = n[c] + t[c];
    }
    enormalize(n, dim);
}
static void sil_centroid(const oracle_bvh *h, const int32_t *ind, float *c)
{
    if (h->dim == 2) {
        memcpy(c, spos(h, ind[1]), 2 * sizeof(float));
        return;
    }
    const float *pa = spos(h, ind[1]), *pb = spos(h, ind[2]);
    for (int i = 0; i < 3; i++) c[i] = (pa[i] + pb[i]) * 0.5f;
}

/* assignEdgeIndices (fcpw.inl:405-426): edges numbered by first appearance */
typedef struct { int32_t a, b, pos; } edge_key;
static int edge_key_cmp(const void *x, const void *y)
{
    const edge_key *p = (const edge_key *)x, *q = (const edge_key *)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    return p->pos < q->pos ? -1 : (p->pos > q->pos);
}
static int int2_cmp(const void *x, const void *y)
{
    const int32_t *p = (const int32_t *)x, *q = (const int32_t *)y;
    return p[0] < q[0] ? -1 : (p[0] > q[0]);
}
static int assign_edge_indices(const int32_t *idx, int n_tris, int32_t *eidx)
{
    int n = 3 * n_tris;
    edge_key *k = (edge_key *)malloc((size_t)n * sizeof(edge_key));
    for (int i = 0; i < n_tris; i++)
        for (int j = 0; j < 3; j++) {
            int32_t I = idx[3 * i + j], J = idx[3 * i + (j + 1) % 3];
            if (I > J) { int32_t t = I; I = J; J = t; }
            k[3 * i + j].a = I; k[3 * i + j].b = J; k[3 * i + j].pos = 3 * i + j;
        }
    qsort(k, n, sizeof(edge_key), edge_key_cmp);
    /* (first position, group start) per distinct edge, then number groups by first position */
    int32_t *grp = (int32_t *)malloc((size_t)n * 2 * sizeof(int32_t));
    int ng = 0;
    for (int i = 0; i < n; i++)
        if (i == 0 || k[i].a != k[i - 1].a || k[i].b != k[i - 1].b) { grp[2 * ng] = k[i].pos; grp[2 * ng + 1] = i; ng++; }
    qsort(grp, ng, 2 * sizeof(int32_t), int2_cmp);
    for (int g = 0; g < ng; g++) {
        int i = grp[2 * g + 1];
        int32_t a = k[i].a, b = k[i].b;
        for (; i < n && k[i].a == a && k[i].b == b; i++) eidx[k[i].pos] = g;
    }
    free(grp);
    free(k);
    return ng;
}

/* computeBoundingConesRecursive (bvh.inl:236-304) */
static void cones_recursive(const oracle_bvh *h, struct sil_data *S, int start, int end)
{
    int dim = h->dim;
    const node_t *node = &h->nodes[start];
    float centroid[3], axis[3] = {0.0f, 0.0f, 0.0f}, radius = 0.0f;
    for (int i = 0; i < dim; i++) centroid[i] = (node->box.pmin[i] + node->box.pmax[i]) * 0.5f;
    int any = 0, two = 1;
    for (int i = start; i < end; i++) {
        for (int j = 0; j < S->cnt[i]; j++) {
            int r = S->off[i] + j;
            const int32_t *ind = S->ind[S->ref_obj[r]];
            for (int c = 0; c < dim; c++) axis[c] = axis[c] + S->ref_n[(size_t)r * dim + c];
            float sc[3], dv[3];
            sil_centroid(h, ind, sc);
            for (int c = 0; c < dim; c++) dv[c] = sc[c] - centroid[c];
            radius = smax(radius, enorm(dv, dim));
            two = two && sil_has_face(h, ind, 0) && sil_has_face(h, ind, 1);
            any = 1;
        }
    }
    float *cone = S->cone + (size_t)start * 5;
    if (!any) {
        cone[3] = -PI_F;
    } else if (!two) {
        cone[3] = PI_F;
    } else {
        float axis_norm = enorm(axis, dim);
        if (axis_norm > EPSF) {
            for (int c = 0; c < dim; c++) axis[c] = axis[c] / axis_norm;
            float half = 0.0f;
            for (int i = start; i < end; i++)
                for (int j = 0; j < S->cnt[i]; j++) {
                    int r = S->off[i] + j;
                    for (int f = 0; f < 2; f++) {
                        const float *n = S->ref_fn + ((size_t)r * 2 + f) * dim;
                        float a = acosf(smax(-1.0f, smin(1.0f, edot(axis, n, dim))));
                        half = smax(half, a);
                    }
                }
            cone[0] = axis[0]; cone[1] = axis[1]; cone[2] = dim == 3 ? axis[2] : 0.0f;
            cone[3] = half;
            cone[4] = radius;
        }
    }
    if (node->nrefs == 0) {
        cones_recursive(h, S, start + 1, start + node->offset);
        cones_recursive(h, S, start + node->offset, end);
    }
}

/* GPU records (bvh_interop_structures.h:237-300) from the current positions */
static void sil_make_records(const oracle_bvh *h, struct sil_data *S)
{
    int dim = h->dim;
    for (int r = 0; r < S->n_refs; r++) {
        int o = S->ref_obj[r];
        const int32_t *ind = S->ind[o];
        oracle_silhouette *g = &S->rec[r];
        memset(g, 0, sizeof(*g));
        int f0 = sil_has_face(h, ind, 0), f1 = sil_has_face(h, ind, 1);
        memcpy(g->pa, spos(h, ind[1]), dim * sizeof(float));
        if (dim == 3) memcpy(g->pb, spos(h, ind[2]), 3 * sizeof(float));
        if (f0) sil_face_normal(h, ind, 0, 1, g->n0);
        if (f1) sil_face_normal(h, ind, 1, 1, g->n1);
        g->index = (uint32_t)S->pindex[o];
        g->has_one_adjacent_face = (f0 && f1) ? 0u : 1u;
    }
}

int oracle_build_silhouettes(oracle_bvh *h, const int32_t *user_idx, oracle_ignore_fn ignore, void *user)
{
    int dim = h->dim, np = h->n_prims;
    sil_free(h->sil);
    struct sil_data *S = (struct sil_data *)calloc(1, sizeof(struct sil_data));
    h->sil = S;
    if (dim == 2) {
        /* Scene<2>::computeSilhouettes (fcpw.inl:436-473) on the user's vertex ids, then the
           re-assignment of sortSoupPositions<2, SnchNode> (fcpw.inl:709-731) on sorted ids;
           a field no sorted segment writes keeps its value from the first pass */
        S->n_obj = h->n_vertices;
        S->ind = (int32_t(*)[4])malloc((size_t)S->n_obj * sizeof(*S->ind));
        S->pindex = (int32_t *)malloc((size_t)S->n_obj * sizeof(int32_t));
        for (int v = 0; v < S->n_obj; v++) { S->ind[v][0] = S->ind[v][1] = S->ind[v][2] = S->ind[v][3] = -1; S->pindex[v] = -1; }
        for (int i = 0; i < np; i++) {
            int a = user_idx[2 * i], b = user_idx[2 * i + 1];
            S->ind[a][1] = a; S->ind[a][2] = b; S->pindex[a] = a;
            S->ind[b][0] = a; S->ind[b][1] = b; S->pindex[b] = b;
        }
        for (int s = 0; s < np; s++) {
            int a = h->prim_vidx[2 * s], b = h->prim_vidx[2 * s + 1];
            S->ind[a][1] = a; S->ind[a][2] = b; S->pindex[a] = a;
            S->ind[b][0] = a; S->ind[b][1] = b; S->pindex[b] = b;
        }
    } else {
        /* Scene<3>::computeSilhouettes (fcpw.inl:476-521) on the user's vertex ids, then the
           remap of sortSoupPositions<3, SnchNode> (fcpw.inl:743-756) */
        S->eidx = (int32_t *)malloc((size_t)np * 3 * sizeof(int32_t));
        S->n_obj = assign_edge_indices(user_idx, np, S->eidx);
        S->ind = (int32_t(*)[4])malloc((size_t)S->n_obj * sizeof(*S->ind));
        S->pindex = (int32_t *)malloc((size_t)S->n_obj * sizeof(int32_t));
        for (int e = 0; e < S->n_obj; e++) { S->ind[e][0] = S->ind[e][1] = S->ind[e][2] = S->ind[e][3] = -1; S->pindex[e] = -1; }
        for (int i = 0; i < np; i++) {
            const int32_t *t = user_idx + 3 * i;
            for (int j = 0; j < 3; j++) {
                int I = j == 0 ? 2 : j - 1, J = j, K = j == 2 ? 0 : j + 1;
                int e = S->eidx[3 * i + j];
                int orient = 1;
                if (t[J] > t[K]) { int tmp = J; J = K; K = tmp; orient = -1; }
                S->ind[e][orient == 1 ? 0 : 3] = t[I];
                S->ind[e][1] = t[J];
                S->ind[e][2] = t[K];
                S->pindex[e] = e;
            }
        }
        for (int e = 0; e < S->n_obj; e++)
            for (int j = 0; j < 4; j++)
                if (S->ind[e][j] != -1) S->ind[e][j] = h->index_map[S->ind[e][j]];
    }

    /* references per leaf, deduplicated within the leaf (bvh.inl:313-420) */
    int cap = np * (dim == 3 ? 3 : 2);
    S->ref_obj = (int32_t *)malloc((size_t)cap * sizeof(int32_t));
    S->ref_n = (float *)malloc((size_t)cap * dim * sizeof(float));
    S->ref_fn = (float *)malloc((size_t)cap * 2 * dim * sizeof(float));
    S->off = (int32_t *)malloc((size_t)h->n_nodes * sizeof(int32_t));
    S->cnt = (int32_t *)malloc((size_t)h->n_nodes * sizeof(int32_t));
    int32_t seen[3 * 64];
    for (int i = 0; i < h->n_nodes; i++) {
        const node_t *node = &h->nodes[i];
        int start = S->n_refs, n_seen = 0;
        for (int j = 0; j < node->nrefs; j++) {
            int slot = node->offset + j;
            int prim = h->prim_orig[slot];
            for (int k = 0; k < dim; k++) {
                int o = dim == 2 ? h->prim_vidx[2 * slot + k] : S->eidx[3 * prim + k];
                int dup = 0;
                for (int q = 0; q < n_seen; q++) dup |= seen[q] == o;
                if (dup) continue;
                if (n_seen < 3 * 64) seen[n_seen++] = o;
                const int32_t *ind = S->ind[o];
                int f0 = sil_has_face(h, ind, 0), f1 = sil_has_face(h, ind, 1);
                float n0[3] = {0.0f, 0.0f, 0.0f}, n1[3] = {0.0f, 0.0f, 0.0f}, n[3];
                if (f0) sil_face_normal(h, ind, 0, 1, n0);
                if (f1) sil_face_normal(h, ind, 1, 1, n1);
                sil_normal(h, ind, n);
                int ign = 0;
                if (f0 && f1 && ignore) {
                    float angle;
                    if (dim == 2) {
                        angle = n0[0] * n1[1] - n0[1] * n1[0];
                    } else {
                        const float *pa = spos(h, ind[1]), *pb = spos(h, ind[2]);
                        float ed[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]}, cr[3];
                        enormalize(ed, 3);
                        ecross(n0, n1, cr);
                        angle = atan2f(edot(ed, cr, 3), edot(n0, n1, 3));
                    }
                    ign = ignore(angle, prim, user) != 0;
                }
                if (ign) continue;
                int r = S->n_refs++;
                S->ref_obj[r] = o;
                memcpy(S->ref_n + (size_t)r * dim, n, dim * sizeof(float));
                memcpy(S->ref_fn + (size_t)r * 2 * dim, n0, dim * sizeof(float));
                memcpy(S->ref_fn + (size_t)r * 2 * dim + dim, n1, dim * sizeof(float));
            }
        }
        S->off[i] = start;
        S->cnt[i] = S->n_refs - start;
    }

    /* cones: SnchNode default (axis 0, halfAngle pi, radius 0), then bvh.inl:236-304 */
    S->cone = (float *)calloc((size_t)h->n_nodes * 5, sizeof(float));
    for (int i = 0; i < h->n_nodes; i++) S->cone[(size_t)i * 5 + 3] = PI_F;
    cones_recursive(h, S, 0, h->n_nodes);

    S->rec = (oracle_silhouette *)calloc((size_t)(S->n_refs > 0 ? S->n_refs : 1), sizeof(oracle_silhouette));
    sil_make_records(h, S);
    return S->n_refs;
}

int oracle_num_silhouette_refs(const oracle_bvh *h) { return h->sil ? h->sil->n_refs : 0; }

void oracle_export_silhouettes(const oracle_bvh *h, float *cones, int32_t *n_silhouettes, int32_t *silhouette_offset,
                               oracle_silhouette *records)
{
    const struct sil_data *S = h->sil;
    if (!S) return;
    if (cones) memcpy(cones, S->cone, (size_t)h->n_nodes * 5 * sizeof(float));
    if (n_silhouettes) memcpy(n_silhouettes, S->cnt, (size_t)h->n_nodes * sizeof(int32_t));
    if (silhouette_offset) memcpy(silhouette_offset, S->off, (size_t)h->n_nodes * sizeof(int32_t));
    if (records) memcpy(records, S->rec, (size_t)S->n_refs * sizeof(oracle_silhouette));
}

void oracle_update_silhouette_records(oracle_bvh *h)
{
    if (h->sil) sil_make_records(h, h->sil);
}

/* ---- GPU-side geometry, float3 arithmetic summed left to right ---- */
static inline float g3dot(const float *a, const float *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static inline float g3len(const float *a) { return sqrtf(g3dot(a, a)); }

/* isSilhouetteVertex / isSilhouetteEdge (geometry.slang:490-522,599-633) */
static int is_silhouette(const oracle_silhouette *s, int dim, const float *view, float d, int flip, float prec)
{
    float sign = flip ? 1.0f : -1.0f;
    if (d <= prec) {
        if (dim == 2) {
            float det = s->n0[0] * s->n1[1] - s->n1[0] * s->n0[1];
            return sign * det > prec;
        }
        float e[3] = {s->pb[0] - s->pa[0], s->pb[1] - s->pa[1], s->pb[2] - s->pa[2]}, cr[3];
        float l = g3len(e);
        for (int i = 0; i < 3; i++) e[i] = e[i] / l;
        ecross(s->n0, s->n1, cr);
        float angle = atan2f(g3dot(e, cr), g3dot(s->n0, s->n1));
        return sign * angle > prec;
    }
    float u[3] = {view[0] / d, view[1] / d, view[2] / d};
    float dot0 = g3dot(u, s->n0), dot1 = g3dot(u, s->n1);
    if (fabsf(dot0) <= prec) return sign * dot1 > prec;
    if (fabsf(dot1) <= prec) return sign * dot0 > prec;
    return dot0 * dot1 < 0.0f;
}

/* Vertex/Edge.findClosestSilhouettePoint (geometry.slang:555-590,667-701) */
static int silhouette_test(const oracle_silhouette *s, int dim, const float *x, float r2, int flip, float r2min,
                           float prec, oracle_hit *c)
{
    if (r2min >= r2) return 0;
    float p[3], t = 0.0f, d;
    if (dim == 2) {
        memcpy(p, s->pa, sizeof(p));
    } else {
        /* findClosestPointLineSegment (geometry.slang:79-103) on float3 */
        float u[3], v[3];
        for (int i = 0; i < 3; i++) { u[i] = s->pb[i] - s->pa[i]; v[i] = x[i] - s->pa[i]; }
        float c1 = g3dot(u, v);
        if (c1 <= 0.0f) {
            memcpy(p, s->pa, sizeof(p));
        } else {
            float c2 = g3dot(u, u);
            if (c2 <= c1) {
                memcpy(p, s->pb, sizeof(p));
                t = 1.0f;
            } else {
                t = c1 / c2;
                for (int i = 0; i < 3; i++) p[i] = s->pa[i] + u[i] * t;
            }
        }
    }
    float view[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
    d = g3len(view);
    if (d * d > r2) return 0;
    int process = s->has_one_adjacent_face == 1;
    if (!process) process = is_silhouette(s, dim, view, d, flip, prec);
    if (process && d * d <= r2) {
        hit_default(c);
        memcpy(c->p, p, sizeof(p));
        c->uv[0] = t;
        c->d = d;
        c->index = (int32_t)s->index;
        return 1;
    }
    return 0;
}

/* computeOrthonormalBasis + projectToPlane (bounding-volumes.slang:122-142) */
static float project_to_plane(const float *n, const float *e)
{
    float sign = n[2] >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n[2]);
    float b = n[0] * n[1] * a;
    float b1[3] = {1.0f + sign * n[0] * n[0] * a, sign * b, -sign * n[0]};
    float b2[3] = {b, sign + n[1] * n[1] * a, -n[1]};
    float ab1[3] = {fabsf(b1[0]), fabsf(b1[1]), fabsf(b1[2])}, ab2[3] = {fabsf(b2[0]), fabsf(b2[1]), fabsf(b2[2])};
    float r1 = g3dot(e, ab1), r2 = g3dot(e, ab2);
    return sqrtf(r1 * r1 + r2 * r2);
}
static inline int in_range(float v, float lo, float hi) { return v >= lo && v <= hi; }

/* BoundingCone.overlap (bounding-volumes.slang:170-236); box in 3 floats (z = 0 in 2D) */
static int cone_overlap(const float *cone, const float *o, const float *lo, const float *hi, float dist_to_box)
{
    const float half_pi = PI_2_F;
    float half = cone[3], radius = cone[4];
    if (half >= half_pi || dist_to_box < EPSF) return 1;
    float c[3], v[3];
    for (int i = 0; i < 3; i++) { c[i] = 0.5f * (lo[i] + hi[i]); v[i] = c[i] - o[i]; }
    float l = g3len(v);
    for (int i = 0; i < 3; i++) v[i] = v[i] / l;
    float d_axis = acosf(fmaxf(-1.0f, fminf(1.0f, g3dot(cone, v))));
    if (in_range(half_pi, d_axis - half, d_axis + half)) return 1;
    if (l > radius) {
        float vh = asinf(radius / l);
        float hs = half + vh;
        return hs >= half_pi ? 1 : in_range(half_pi, d_axis - hs, d_axis + hs);
    }
    float e[3] = {hi[0] - c[0], hi[1] - c[1], hi[2] - c[2]}, av[3] = {fabsf(v[0]), fabsf(v[1]), fabsf(v[2])};
    float d = g3dot(e, av);
    float s = l - d;
    if (s <= 0.0f) return 1;
    d = project_to_plane(v, e);
    float vh = atan2f(d, s);
    float hs = half + vh;
    return hs >= half_pi ? 1 : in_range(half_pi, d_axis - hs, d_axis + hs);
}

static float node_d2min(const oracle_bvh *h, int ni, const float *x)
{
    float d2min, d2max;
    box_sqdist_gpu(&h->nodes[ni].box, x, h->dim, &d2min, &d2max);
    return d2min;
}

typedef struct {
    const oracle_bvh *h;
    const float *spheres;
    const uint32_t *flip;
    float r2min, prec;
    int brute;
    oracle_hit *out;
} csp_ctx;

static void csp_range(void *vctx, int start, int end)
{
    const csp_ctx *C = (const csp_ctx *)vctx;
    const oracle_bvh *h = C->h;
    const struct sil_data *S = h->sil;
    int dim = h->dim;
    for (int qi = start; qi < end; qi++) {
        const float *sp = C->spheres + (size_t)qi * 4;
        float x[3] = {sp[0], sp[1], dim == 3 ? sp[2] : 0.0f};
        float r2 = sp[3];
        int flip = C->flip[qi] != 0;
        oracle_hit best;
        hit_default(&best);
        if (!(C->r2min >= r2)) {
            if (C->brute) {
                /* every reference in order, the leaf loop of bvh.slang:540-566 */
                for (int r = 0; r < S->n_refs; r++) {
                    if ((int32_t)S->rec[r].index == best.index) continue;
                    oracle_hit c;
                    if (silhouette_test(&S->rec[r], dim, x, r2, flip, C->r2min, C->prec, &c)) {
                        r2 = fminf(r2, c.d * c.d);
                        best = c;
                        if (C->r2min >= r2) break;
                    }
                }
            } else if (node_d2min(h, 0, x) <= r2) {
                int stack[BVH_MAX_DEPTH * 2 + 2];
                int sp_ = 0;
                stack[0] = 0;
                while (sp_ >= 0) {
                    int ni = stack[sp_--];
                    const node_t *node = &h->nodes[ni];
                    if (!(node_d2min(h, ni, x) <= r2)) continue;
                    if (node->nrefs > 0) {
                        for (int p = 0; p < S->cnt[ni]; p++) {
                            const oracle_silhouette *s = &S->rec[S->off[ni] + p];
                            if ((int32_t)s->index == best.index) continue;
                            oracle_hit c;
                            if (silhouette_test(s, dim, x, r2, flip, C->r2min, C->prec, &c)) {
                                r2 = fminf(r2, c.d * c.d);
                                best = c;
                                if (C->r2min >= r2) break;
                            }
                        }
                    } else {
                        int kid[2] = {ni + 1, ni + node->offset};
                        int ov[2];
                        float dk[2] = {0.0f, 0.0f};
                        for (int k = 0; k < 2; k++) {
                            const float *cone = S->cone + (size_t)kid[k] * 5;
                            ov[k] = cone[3] >= 0.0f;
                            if (ov[k]) {
                                dk[k] = node_d2min(h, kid[k], x);
                                ov[k] = dk[k] <= r2;
                                if (ov[k]) {
                                    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
                                    memcpy(lo, h->nodes[kid[k]].box.pmin, dim * sizeof(float));
                                    memcpy(hi, h->nodes[kid[k]].box.pmax, dim * sizeof(float));
                                    ov[k] = cone_overlap(cone, x, lo, hi, dk[k]);
                                }
                            }
                        }
                        if (ov[0] && ov[1]) {
                            int closer = kid[0], other = kid[1];
                            if (dk[1] < dk[0]) { closer = kid[1]; other = kid[0]; }
                            stack[++sp_] = other;
                            stack[++sp_] = closer;
                        } else if (ov[0]) {
                            stack[++sp_] = kid[0];
                        } else if (ov[1]) {
                            stack[++sp_] = kid[1];
                        }
                    }
                }
            }
        }
        C->out[qi] = best;
    }
}

void oracle_closest_silhouettes_gpuorder(const oracle_bvh *h, const float *spheres, const uint32_t *flip, int n,
                                         float squared_min_radius, float precision, int brute, int nthreads,
                                         oracle_hit *out)
{
    if (!h->sil) {
        for (int i = 0; i < n; i++) hit_default(&out[i]);
        return;
    }
    csp_ctx c = {h, spheres, flip, squared_min_radius, precision, brute, out};
    parallel_for(n, nthreads, csp_range, &c);
}

/* ----------------------------------------------------------------------------
 * FP64 closest silhouette point on raw geometry, classified by an independent
 * orientation test: a vertex (2D) is on the silhouette seen from x when its two
 * neighbours lie on the same side of the line through x and the vertex; an edge
 * (3D) when the opposite vertices of its two triangles lie on the same side of
 * the plane through x and the edge.  A vertex or edge with one neighbour/face is
 * always on the silhouette.  Candidates whose normalised orientation is within
 * tol of zero, or that lie within tol_d of x, are "ambiguous" (the FP32 path
 * classifies those with its precision rules).  d_strict is the nearest certain
 * silhouette point and d_loose the nearest certain-or-ambiguous one.
 * ------------------------------------------------------------------------- */
typedef struct {
    int dim, n_cand;
    const float *pos;
    int32_t *a, *b, *c, *d; /* per candidate; 2D: vertex a with neighbours c (prev), d (next) */
    double tol, tol_d;
    const float *q;
    double *d_strict, *d_loose, *p_strict;
} sf64_ctx;

static void sf64_range(void *vctx, int start, int end)
{
    const sf64_ctx *C = (const sf64_ctx *)vctx;
    int dim = C->dim;
    for (int qi = start; qi < end; qi++) {
        double x[3] = {C->q[(size_t)qi * dim], C->q[(size_t)qi * dim + 1], dim == 3 ? C->q[(size_t)qi * dim + 2] : 0.0};
        double ds = INFINITY, dl = INFINITY, ps[3] = {0, 0, 0};
        for (int k = 0; k < C->n_cand; k++) {
            double A[3] = {0, 0, 0}, P[3], dist;
            for (int i = 0; i < dim; i++) A[i] = C->pos[(size_t)C->a[k] * dim + i];
            if (dim == 2) {
                memcpy(P, A, sizeof(P));
            } else {
                double B[3], u[3], w[3];
                for (int i = 0; i < 3; i++) { B[i] = C->pos[(size_t)C->b[k] * 3 + i]; u[i] = B[i] - A[i]; w[i] = x[i] - A[i]; }
                double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
                double t = uu > 0 ? (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / uu : 0.0;
                t = t < 0 ? 0 : (t > 1 ? 1 : t);
                for (int i = 0; i < 3; i++) P[i] = A[i] + t * u[i];
            }
            double r[3] = {x[0] - P[0], x[1] - P[1], x[2] - P[2]};
            dist = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            int strict, ambiguous = 0;
            if (C->c[k] < 0 || C->d[k] < 0) {
                strict = 1;
            } else if (dist <= C->tol_d) {
                strict = 0;
                ambiguous = 1;
            } else {
                double s[2];
                int nb[2] = {C->c[k], C->d[k]};
                for (int j = 0; j < 2; j++) {
                    double N[3] = {0, 0, 0};
                    for (int i = 0; i < dim; i++) N[i] = C->pos[(size_t)nb[j] * dim + i];
                    if (dim == 2) {
                        /* cross(A - x, N - A) / (|A - x| |N - A|) */
                        double e0 = A[0] - x[0], e1 = A[1] - x[1], f0 = N[0] - A[0], f1 = N[1] - A[1];
                        double den = sqrt(e0 * e0 + e1 * e1) * sqrt(f0 * f0 + f1 * f1);
                        s[j] = den > 0 ? (e0 * f1 - e1 * f0) / den : 0.0;
                    } else {
                        /* det(x - A, B - A, N - A) / (|x - P| |(B - A) x (N - A)|) */
                        double B[3], e[3], f[3], g[3], m[3];
                        for (int i = 0; i < 3; i++) {
                            B[i] = C->pos[(size_t)C->b[k] * 3 + i];
                            e[i] = x[i] - A[i]; f[i] = B[i] - A[i]; g[i] = N[i] - A[i];
                        }
                        m[0] = f[1] * g[2] - f[2] * g[1];
                        m[1] = f[2] * g[0] - f[0] * g[2];
                        m[2] = f[0] * g[1] - f[1] * g[0];
                        double den = dist * sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
                        s[j] = den > 0 ? (e[0] * m[0] + e[1] * m[1] + e[2] * m[2]) / den : 0.0;
                    }
                }
                strict = s[0] * s[1] > 0;
                ambiguous = fabs(s[0]) <= C->tol || fabs(s[1]) <= C->tol;
                if (ambiguous) strict = 0;
            }
            if (strict && dist < ds) { ds = dist; memcpy(ps, P, sizeof(ps)); }
            if ((strict || ambiguous) && dist < dl) dl = dist;
        }
        C->d_strict[qi] = ds;
        C->d_loose[qi] = dl;
        memcpy(C->p_strict + (size_t)qi * 3, ps, sizeof(ps));
    }
}

int oracle_closest_silhouette_f64(int dim, const float *positions, int n_vertices, const int32_t *indices,
                                  int n_prims, const float *queries, int n, double tol, double tol_d, int nthreads,
                                  double *d_strict, double *d_loose, double *p_strict)
{
    sf64_ctx C;
    memset(&C, 0, sizeof(C));
    C.dim = dim; C.pos = positions; C.tol = tol; C.tol_d = tol_d; C.q = queries;
    C.d_strict = d_strict; C.d_loose = d_loose; C.p_strict = p_strict;
    int nonmanifold = 0;
    if (dim == 2) {
        C.a = (int32_t *)malloc((size_t)n_vertices * sizeof(int32_t));
        C.c = (int32_t *)malloc((size_t)n_vertices * sizeof(int32_t));
        C.d = (int32_t *)malloc((size_t)n_vertices * sizeof(int32_t));
        int32_t *used = (int32_t *)calloc((size_t)n_vertices, sizeof(int32_t));
        for (int v = 0; v < n_vertices; v++) { C.a[v] = v; C.c[v] = -1; C.d[v] = -1; }
        for (int i = 0; i < n_prims; i++) {
            int a = indices[2 * i], b = indices[2 * i + 1];
            if (C.d[a] >= 0 || C.c[b] >= 0) nonmanifold = 1;
            C.d[a] = b; C.c[b] = a;
            used[a] = used[b] = 1;
        }
        /* only vertices some segment uses are candidates */
        int m = 0;
        for (int v = 0; v < n_vertices; v++)
            if (used[v]) { C.a[m] = C.a[v]; C.c[m] = C.c[v]; C.d[m] = C.d[v]; m++; }
        C.n_cand = m;
        free(used);
    } else {
        int32_t *eidx = (int32_t *)malloc((size_t)n_prims * 3 * sizeof(int32_t));
        int E = assign_edge_indices(indices, n_prims, eidx);
        C.n_cand = E;
        C.a = (int32_t *)malloc((size_t)E * sizeof(int32_t));
        C.b = (int32_t *)malloc((size_t)E * sizeof(int32_t));
        C.c = (int32_t *)malloc((size_t)E * sizeof(int32_t));
        C.d = (int32_t *)malloc((size_t)E * sizeof(int32_t));
        for (int e = 0; e < E; e++) C.c[e] = C.d[e] = -1;
        for (int i = 0; i < n_prims; i++)
            for (int j = 0; j < 3; j++) {
                int e = eidx[3 * i + j];
                C.a[e] = indices[3 * i + j];
                C.b[e] = indices[3 * i + (j + 1) % 3];
                int opp = indices[3 * i + (j + 2) % 3];
                if (C.c[e] < 0) C.c[e] = opp;
                else if (C.d[e] < 0) C.d[e] = opp;
                else nonmanifold = 1;
            }
        free(eidx);
    }
    if (!nonmanifold) parallel_for(n, nthreads, sf64_range, &C);
    free(C.a); free(C.b); free(C.c); free(C.d);
    return nonmanifold ? -1 : 0;
}

/* ----------------------------------------------------------------------------
 * GPU refit of the cones (Bvh.refit, bvh.slang:27-141), after oracle_refit has
 * recomputed the boxes and oracle_update_silhouette_records the records.
 * Leaves rebuild their cone from their records (refitLeafNode); internal nodes
 * merge their children's (refitInternalNode, mergeBoundingCones and rotate,
 * bounding-volumes.slang:238-318).  rotate's w is length(cross(u, v)) in every
 * component, as the reference declares it.
 * ------------------------------------------------------------------------- */
static void g3rotate(const float *u, const float *v, float theta, float *out)
{
    float c = cosf(theta), s = sinf(theta), cr[3];
    ecross(u, v, cr);
    float L = g3len(cr);
    float w[3] = {L, L, L}, o[3];
    for (int i = 0; i < 3; i++) o[i] = (1.0f - c) * w[i];
    float R[3][3] = {{c + o[0] * w[0], o[1] * w[0] - s * w[2], o[2] * w[0] + s * w[1]},
                     {o[0] * w[1] + s * w[2], c + o[1] * w[1], o[2] * w[1] - s * w[0]},
                     {o[0] * w[2] - s * w[1], o[1] * w[2] + s * w[0], c + o[2] * w[2]}};
    for (int i = 0; i < 3; i++) out[i] = (R[i][0] * u[0] + R[i][1] * u[1]) + R[i][2] * u[2];
}

static void g3merge_cones(const float *A, const float *B, const float *oA, const float *oB, const float *oN,
                          float *C)
{
    C[0] = C[1] = C[2] = 0.0f;
    C[3] = PI_F;
    C[4] = 0.0f;
    if (A[3] >= 0.0f && B[3] >= 0.0f) {
        float axA[3] = {A[0], A[1], A[2]}, axB[3] = {B[0], B[1], B[2]};
        float hA = A[3], hB = B[3];
        float dA[3], dB[3];
        for (int i = 0; i < 3; i++) { dA[i] = oN[i] - oA[i]; dB[i] = oN[i] - oB[i]; }
        C[4] = sqrtf(fmaxf(A[4] * A[4] + g3dot(dA, dA), B[4] * B[4] + g3dot(dB, dB)));
        if (hB > hA) {
            for (int i = 0; i < 3; i++) { float t = axA[i]; axA[i] = axB[i]; axB[i] = t; }
            float t = hA; hA = hB; hB = t;
        }
        float theta = acosf(fmaxf(-1.0f, fminf(1.0f, g3dot(axA, axB))));
        if (fminf(theta + hB, PI_F) <= hA) {
            memcpy(C, axA, sizeof(axA));
            C[3] = hA;
            return;
        }
        float o_theta = (hA + theta + hB) / 2.0f;
        if (o_theta >= PI_F) {
            memcpy(C, axA, sizeof(axA));
            return;
        }
        g3rotate(axA, axB, o_theta - hA, C);
        C[3] = o_theta;
    } else if (A[3] >= 0.0f) {
        memcpy(C, A, 5 * sizeof(float));
    } else if (B[3] >= 0.0f) {
        memcpy(C, B, 5 * sizeof(float));
    } else {
        C[3] = -PI_F;
    }
}

static void node_centroid3(const oracle_bvh *h, int ni, float *c)
{
    c[2] = 0.0f;
    for (int i = 0; i < h->dim; i++) c[i] = 0.5f * (h->nodes[ni].box.pmin[i] + h->nodes[ni].box.pmax[i]);
}

static void refit_cones_recursive(oracle_bvh *h, int ni)
{
    struct sil_data *S = h->sil;
    const node_t *node = &h->nodes[ni];
    float *C = S->cone + (size_t)ni * 5;
    if (node->nrefs == 0) {
        int l = ni + 1, r = ni + node->offset;
        refit_cones_recursive(h, l);
        refit_cones_recursive(h, r);
        float cl[3], cr[3], cn[3];
        node_centroid3(h, l, cl);
        node_centroid3(h, r, cr);
        node_centroid3(h, ni, cn);
        g3merge_cones(S->cone + (size_t)l * 5, S->cone + (size_t)r * 5, cl, cr, cn, C);
        return;
    }
    float axis[3] = {0.0f, 0.0f, 0.0f}, centroid[3], half = 0.0f, radius = 0.0f;
    node_centroid3(h, ni, centroid);
    int any = 0, two = 1;
    for (int p = 0; p < S->cnt[ni]; p++) {
        const oracle_silhouette *s = &S->rec[S->off[ni] + p];
        for (int i = 0; i < 3; i++) axis[i] = axis[i] + s->n0[i];
        for (int i = 0; i < 3; i++) axis[i] = axis[i] + s->n1[i];
        float sc[3], dv[3];
        for (int i = 0; i < 3; i++) sc[i] = h->dim == 2 ? s->pa[i] : 0.5f * (s->pa[i] + s->pb[i]);
        for (int i = 0; i < 3; i++) dv[i] = sc[i] - centroid[i];
        radius = fmaxf(radius, g3len(dv));
        two = two && s->has_one_adjacent_face == 0;
        any = 1;
    }
    if (!any) {
        half = -PI_F;
    } else if (!two) {
        half = PI_F;
    } else {
        float an = g3len(axis);
        if (an > EPSF) {
            for (int i = 0; i < 3; i++) axis[i] = axis[i] / an;
            for (int p = 0; p < S->cnt[ni]; p++) {
                const oracle_silhouette *s = &S->rec[S->off[ni] + p];
                half = fmaxf(half, acosf(fmaxf(-1.0f, fminf(1.0f, g3dot(axis, s->n0)))));
                half = fmaxf(half, acosf(fmaxf(-1.0f, fminf(1.0f, g3dot(axis, s->n1)))));
            }
        }
    }
    C[0] = axis[0]; C[1] = axis[1]; C[2] = axis[2]; C[3] = half; C[4] = radius;
}

void oracle_refit_silhouette_cones_gpu(oracle_bvh *h)
{
    if (h->sil) refit_cones_recursive(h, 0);
}

/* ----------------------------------------------------------------------------
 * Scenes of several objects and transformed instances.
 *
 * Scene::build (fcpw.inl:1052-1126) wraps each object in an aggregate; an
 * object with instance transforms yields one TransformedAggregate per
 * transform (primitive.h:289-440), in object order then transform order.  With
 * more than one instance, a BVH over the instances is built by the same
 * builder (makeAggregate over the aggregate pointers, fcpw.inl:1103-1112),
 * each instance bounded by its object's root box (bvh.inl:614-618) and
 * centred on the mean of its primitive centroids (bvh.inl:620-631); a
 * transformed instance maps the 2^DIM root-box corners with epsilon-inflated
 * expansion (bounding_volumes.h:183-202, primitive.h:295-298) and its
 * centroid by t (primitive.h:300-303).
 *
 * Queries follow the GPU order at both levels (bvh.slang:144-254,395-502).
 * A transformed instance is queried as TransformedAggregate is on the GPU
 * (aggregate.slang:64-117, transform.slang:1-56): the sphere / ray is mapped
 * by tInv, the object's tree traversed, r^2 / tMax mapped back by t, and on a
 * hit the point is mapped by t, the normal by transpose(tInv) and normalized,
 * and d = |p - x|.  An object without instance transforms is queried directly
 * (bvh.inl:1317-1320,680-683).  Found results shrink r^2 / tMax as primitives
 * do (bvh.inl:1322-1327,697-704).
 *
 * Transforms are 3x4 row-major affine matrices; a 2D transform uses rows 0-1
 * and columns 0, 1 (linear) and 3 (translation).  transformPoint is
 * mul(t, float4(p, 1)) with the row sums formed left to right.
 * ------------------------------------------------------------------------- */
static void xf_point(const float *t, const float *p, int dim, float *o)
{
    float r[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < dim; i++) {
        const float *row = t + 4 * i;
        float s = row[0] * p[0] + row[1] * p[1];
        if (dim == 3) s = s + row[2] * p[2];
        r[i] = s + row[3];
    }
    o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
}

/* transformSphere's radius (transform.slang:31-42): c is the sphere centre,
   tc = transformPoint(t, c) */
static float xf_r2(const float *t, const float *c, const float *tc, float r2, int dim)
{
    if (!(r2 < MAXF)) return MAXF;
    float e[3] = {c[0] + sqrtf(r2), c[1] + 0.0f, c[2] + 0.0f}, m[3], dd[3];
    xf_point(t, e, dim, m);
    for (int i = 0; i < 3; i++) dd[i] = m[i] - tc[i];
    return gdot(dd, dd, dim);
}

/* normal mapping of transformInteraction (transform.slang:45-56):
   normalize(transpose(linear(tInv)) * n), written as v / length(v) */
static void xf_normal(const float *tinv, const float *n, int dim, float *o)
{
    float r[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < dim; i++) {
        float s = tinv[0 * 4 + i] * n[0] + tinv[1 * 4 + i] * n[1];
        if (dim == 3) s = s + tinv[2 * 4 + i] * n[2];
        r[i] = s;
    }
    float l = sqrtf(gdot(r, r, dim));
    for (int i = 0; i < dim; i++) r[i] = r[i] / l;
    o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
}

/* Transform::inverse for an affine transform: the linear part inverted by
   cofactors (adjugate times 1/det), the translation -(Linv * v) */
void oracle_invert_transform(int dim, const float *t, float *tinv)
{
    for (int i = 0; i < 12; i++) tinv[i] = 0.0f;
    if (dim == 3) {
        float cof[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                cof[i][j] = t[i1 * 4 + j1] * t[i2 * 4 + j2] - t[i1 * 4 + j2] * t[i2 * 4 + j1];
            }
        float det = (t[0] * cof[0][0] + t[1] * cof[0][1]) + t[2] * cof[0][2];
        float inv = 1.0f / det;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) tinv[i * 4 + j] = cof[j][i] * inv;
        for (int i = 0; i < 3; i++) {
            float s = (tinv[i * 4 + 0] * t[3] + tinv[i * 4 + 1] * t[7]) + tinv[i * 4 + 2] * t[11];
            tinv[i * 4 + 3] = -s;
        }
    } else {
        float det = t[0] * t[5] - t[1] * t[4];
        float inv = 1.0f / det;
        tinv[0] = t[5] * inv;
        tinv[1] = -t[1] * inv;
        tinv[4] = -t[4] * inv;
        tinv[5] = t[0] * inv;
        for (int i = 0; i < 2; i++) tinv[i * 4 + 3] = -(tinv[i * 4 + 0] * t[3] + tinv[i * 4 + 1] * t[7]);
        tinv[10] = 1.0f;
    }
}

static box_t instance_box(oracle_bvh *const *objects, const oracle_instance *I, int dim)
{
    box_t r = objects[I->object]->nodes[0].box;
    if (!I->transformed) return r;
    box_t b;
    box_init(&b);
    for (int i = 0; i < (1 << dim); i++) {
        float p[3] = {0.0f, 0.0f, 0.0f}, q[3];
        int temp = i;
        for (int j = 0; j < dim; j++) {
            p[j] = (temp % 2) == 0 ? r.pmin[j] : r.pmax[j];
            temp /= 2;
        }
        xf_point(I->t, p, dim, q);
        box_expand_point(&b, q, dim);
    }
    return b;
}

static void instance_centroid(oracle_bvh *const *objects, const oracle_instance *I, int dim, float *c)
{
    const oracle_bvh *h = objects[I->object];
    float s[3] = {0.0f, 0.0f, 0.0f};
    for (int slot = 0; slot < h->n_prims; slot++) {
        const float *v[3];
        for (int k = 0; k < h->k; k++) v[k] = vtx(h, slot, k);
        float pc[3];
        prim_centroid(v, h->k, dim, pc);
        for (int d = 0; d < dim; d++) s[d] = s[d] + pc[d];
    }
    float m[3] = {0.0f, 0.0f, 0.0f};
    for (int d = 0; d < dim; d++) m[d] = s[d] / (float)h->n_prims;
    if (I->transformed) xf_point(I->t, m, dim, m);
    for (int d = 0; d < dim; d++) c[d] = m[d];
}

oracle_bvh *oracle_build_instances(int dim, oracle_bvh *const *objects, int n_objects, const oracle_instance *inst,
                                   int n_inst, int heuristic, int leaf_size, int n_buckets)
{
    if ((dim != 2 && dim != 3) || n_inst <= 0) return NULL;
    for (int i = 0; i < n_inst; i++)
        if (inst[i].object < 0 || inst[i].object >= n_objects || objects[inst[i].object]->dim != dim) return NULL;
    oracle_bvh *h = (oracle_bvh *)calloc(1, sizeof(oracle_bvh));
    h->dim = dim;
    h->k = 0;
    h->n_prims = n_inst;
    h->prim_orig = (int32_t *)malloc((size_t)n_inst * sizeof(int32_t));
    build_ctx b;
    memset(&b, 0, sizeof(b));
    b.h = h;
    b.heuristic = heuristic;
    b.leaf_size = leaf_size;
    b.n_buckets = n_buckets;
    b.ref_boxes = (box_t *)malloc((size_t)n_inst * sizeof(box_t));
    b.ref_cent = (float *)malloc((size_t)n_inst * dim * sizeof(float));
    b.bkt_box = (box_t *)malloc((size_t)n_buckets * sizeof(box_t));
    b.rbkt_box = (box_t *)malloc((size_t)n_buckets * sizeof(box_t));
    b.bkt_cnt = (int *)malloc((size_t)n_buckets * sizeof(int));
    b.rbkt_cnt = (int *)malloc((size_t)n_buckets * sizeof(int));
    for (int i = 0; i < n_inst; i++) {
        h->prim_orig[i] = i;
        b.ref_boxes[i] = instance_box(objects, &inst[i], dim);
        instance_centroid(objects, &inst[i], dim, b.ref_cent + (size_t)i * dim);
    }
    build_recursive(&b, -1, 0, 0, n_inst, 0);
    free(b.ref_boxes); free(b.ref_cent); free(b.bkt_box); free(b.rbkt_box); free(b.bkt_cnt); free(b.rbkt_cnt);
    return h;
}

/* Bvh::refit over aggregates (bvh.inl:548-599): the objects are refit first
   (by the caller), then the instance boxes and the tree above them */
static void refit_instances_recursive(oracle_bvh *h, oracle_bvh *const *objects, const oracle_instance *inst, int ni)
{
    node_t *n = &h->nodes[ni];
    if (n->nrefs == 0) {
        refit_instances_recursive(h, objects, inst, ni + 1);
        refit_instances_recursive(h, objects, inst, ni + n->offset);
        n->box = h->nodes[ni + 1].box;
        box_expand_box(&n->box, &h->nodes[ni + n->offset].box, h->dim);
    } else {
        box_init(&n->box);
        for (int p = 0; p < n->nrefs; p++) {
            box_t ib = instance_box(objects, &inst[h->prim_orig[n->offset + p]], h->dim);
            box_expand_box(&n->box, &ib, h->dim);
        }
    }
}

void oracle_refit_instances(oracle_bvh *tlas, oracle_bvh *const *objects, const oracle_instance *inst)
{
    if (tlas->n_nodes > 0) refit_instances_recursive(tlas, objects, inst, 0);
}

typedef struct {
    const oracle_bvh *tlas;
    oracle_bvh *const *objects;
    const oracle_instance *inst;
    const float *q, *r2, *o, *d, *tmax;
    int record_normal, occlusion;
    oracle_hit *out;
    int32_t *object_out;
} inst_ctx;

/* TransformedAggregate::findClosestPoint (aggregate.slang:95-117); r2 inout */
static int inst_cp(const inst_ctx *c, int ii, const float *x, float *r2, oracle_hit *hit, int *slot_out)
{
    const oracle_instance *I = &c->inst[ii];
    const oracle_bvh *h = c->objects[I->object];
    int dim = h->dim;
    float xl[3] = {x[0], x[1], x[2]}, r2l = *r2;
    if (I->transformed) {
        xf_point(I->tinv, x, dim, xl);
        r2l = xf_r2(I->tinv, x, xl, *r2, dim);
    }
    oracle_hit cand;
    hit_default(&cand);
    int slot = cp_gpu_one(h, xl, &r2l, &cand, NULL, NULL, NULL);
    if (I->transformed) {
        float tc[3];
        xf_point(I->t, xl, dim, tc);
        *r2 = xf_r2(I->t, xl, tc, r2l, dim);
    } else {
        *r2 = r2l;
    }
    if (slot < 0) return 0;
    if (I->transformed) {
        float p[3], e[3];
        xf_point(I->t, cand.p, dim, p);
        memcpy(cand.p, p, sizeof(p));
        for (int i = 0; i < 3; i++) e[i] = p[i] - x[i];
        cand.d = sqrtf(gdot(e, e, dim));
    }
    *hit = cand;
    *slot_out = slot;
    return 1;
}

static void inst_cp_range(void *vctx, int start, int end)
{
    inst_ctx *c = (inst_ctx *)vctx;
    const oracle_bvh *T = c->tlas;
    int dim = T->dim;
    int st[BVH_MAX_DEPTH];
    for (int qi = start; qi < end; qi++) {
        float x[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) x[d] = c->q[(size_t)qi * dim + d];
        float r2 = c->r2[qi];
        oracle_hit best, cand;
        hit_default(&best);
        int best_inst = -1, best_slot = -1, slot;
        if (T->n_prims == 1) {
            /* a single instance is the scene's aggregate itself (fcpw.inl:1088-1092) */
            if (inst_cp(c, T->prim_orig[0], x, &r2, &cand, &slot)) {
                best = cand;
                best_inst = T->prim_orig[0];
                best_slot = slot;
            }
        } else {
            float d0, d1, d2, d3;
            box_sqdist_gpu(&T->nodes[0].box, x, dim, &d0, &d1);
            if (d0 <= r2) {
                r2 = fminf(r2, d1);
                int sp = 0;
                st[0] = 0;
                while (sp >= 0) {
                    int ni = st[sp--];
                    const node_t *n = &T->nodes[ni];
                    box_sqdist_gpu(&n->box, x, dim, &d0, &d1);
                    if (!(d0 <= r2)) continue;
                    if (n->nrefs > 0) {
                        for (int p = 0; p < n->nrefs; p++) {
                            int ii = T->prim_orig[n->offset + p];
                            if (inst_cp(c, ii, x, &r2, &cand, &slot)) {
                                r2 = fminf(r2, cand.d * cand.d);
                                best = cand;
                                best_inst = ii;
                                best_slot = slot;
                            }
                        }
                    } else {
                        int l = ni + 1, r = ni + n->offset;
                        box_sqdist_gpu(&T->nodes[l].box, x, dim, &d0, &d1);
                        int ol = d0 <= r2;
                        r2 = fminf(r2, d1);
                        box_sqdist_gpu(&T->nodes[r].box, x, dim, &d2, &d3);
                        int orr = d2 <= r2;
                        r2 = fminf(r2, d3);
                        if (ol && orr) {
                            int closer = l, other = r;
                            if (d2 < d0) { closer = r; other = l; }
                            st[++sp] = other;
                            st[++sp] = closer;
                        } else if (ol) {
                            st[++sp] = l;
                        } else if (orr) {
                            st[++sp] = r;
                        }
                    }
                }
            }
        }
        if (best_inst >= 0 && c->record_normal) {
            const oracle_instance *I = &c->inst[best_inst];
            float n[3];
            prim_normal_gpu(c->objects[I->object], best_slot, n);
            if (I->transformed) xf_normal(I->tinv, n, dim, n);
            memcpy(best.n, n, sizeof(n));
        }
        c->out[qi] = best;
        if (c->object_out) c->object_out[qi] = best_inst >= 0 ? c->inst[best_inst].object : -1;
    }
}

void oracle_instances_closest_points(const oracle_bvh *tlas, oracle_bvh *const *objects, const oracle_instance *inst,
                                     const float *queries, const float *r2, int n, int record_normal, int nthreads,
                                     oracle_hit *out, int32_t *object_out)
{
    inst_ctx c;
    memset(&c, 0, sizeof(c));
    c.tlas = tlas;
    c.objects = objects;
    c.inst = inst;
    c.q = queries;
    c.r2 = r2;
    c.record_normal = record_normal;
    c.out = out;
    c.object_out = object_out;
    parallel_for(n, nthreads, inst_cp_range, &c);
}

/* TransformedAggregate::intersect (aggregate.slang:59-79); tmax inout */
static int inst_ray(const inst_ctx *c, int ii, const float *o, const float *d, const float *dinv, float *tmax,
                    oracle_hit *hit)
{
    const oracle_instance *I = &c->inst[ii];
    const oracle_bvh *h = c->objects[I->object];
    int dim = h->dim;
    float ol[3] = {o[0], o[1], o[2]}, dl[3] = {d[0], d[1], d[2]}, il[3] = {dinv[0], dinv[1], dinv[2]};
    float tl = *tmax;
    if (I->transformed) {
        /* transformRay(tInv, r) (transform.slang:22-29) and the Ray constructor's 1/d */
        float s = *tmax < MAXF ? *tmax : 1.0f, e[3], m[3];
        xf_point(I->tinv, o, dim, ol);
        for (int i = 0; i < 3; i++) e[i] = o[i] + d[i] * s;
        xf_point(I->tinv, e, dim, m);
        for (int i = 0; i < 3; i++) dl[i] = m[i] - ol[i];
        float nrm = sqrtf(gdot(dl, dl, dim));
        for (int i = 0; i < 3; i++) dl[i] = dl[i] / nrm;
        tl = *tmax < MAXF ? nrm : MAXF;
        for (int i = 0; i < dim; i++) il[i] = 1.0f / dl[i];
    }
    oracle_hit cand;
    hit_default(&cand);
    int slot = ray_gpu_one(h, ol, dl, il, &tl, c->occlusion, &cand, NULL, NULL);
    if (I->transformed) {
        if (tl < MAXF) {
            float a[3], e[3], b[3], dd[3];
            xf_point(I->t, ol, dim, a);
            for (int i = 0; i < 3; i++) e[i] = ol[i] + dl[i] * tl;
            xf_point(I->t, e, dim, b);
            for (int i = 0; i < 3; i++) dd[i] = b[i] - a[i];
            *tmax = sqrtf(gdot(dd, dd, dim));
        } else {
            *tmax = MAXF;
        }
    } else {
        *tmax = tl;
    }
    if (slot < 0) return 0;
    if (I->transformed && !c->occlusion) {
        float p[3], e[3], n[3];
        xf_point(I->t, cand.p, dim, p);
        memcpy(cand.p, p, sizeof(p));
        xf_normal(I->tinv, cand.n, dim, n);
        memcpy(cand.n, n, sizeof(n));
        for (int i = 0; i < 3; i++) e[i] = p[i] - o[i];
        cand.d = sqrtf(gdot(e, e, dim));
    }
    *hit = cand;
    return 1;
}

static void inst_ray_range(void *vctx, int start, int end)
{
    inst_ctx *c = (inst_ctx *)vctx;
    const oracle_bvh *T = c->tlas;
    int dim = T->dim;
    stack_entry st[BVH_MAX_DEPTH];
    for (int qi = start; qi < end; qi++) {
        float o[3] = {0, 0, 0}, d[3] = {0, 0, 0}, dinv[3] = {0, 0, 0};
        for (int k = 0; k < dim; k++) {
            o[k] = c->o[(size_t)qi * dim + k];
            d[k] = c->d[(size_t)qi * dim + k];
            dinv[k] = 1.0f / d[k];
        }
        float tmax = c->tmax[qi];
        oracle_hit best, cand;
        hit_default(&best);
        int best_inst = -1;
        if (T->n_prims == 1) {
            if (inst_ray(c, T->prim_orig[0], o, d, dinv, &tmax, &cand)) {
                best = cand;
                best_inst = T->prim_orig[0];
            }
        } else {
            float b0, b1, b2, b3;
            if (box_ray_gpu(&T->nodes[0].box, o, d, dinv, tmax, dim, &b0, &b1)) {
                int sp = 0, done = 0;
                st[0].node = 0;
                st[0].dist = b0;
                while (sp >= 0 && !done) {
                    int ni = st[sp].node;
                    float dist = st[sp].dist;
                    sp--;
                    if (dist > tmax) continue;
                    const node_t *n = &T->nodes[ni];
                    if (n->nrefs > 0) {
                        for (int p = 0; p < n->nrefs; p++) {
                            int ii = T->prim_orig[n->offset + p];
                            if (inst_ray(c, ii, o, d, dinv, &tmax, &cand)) {
                                best = cand;
                                best_inst = ii;
                                if (c->occlusion) { done = 1; break; }
                                tmax = fminf(tmax, cand.d);
                            }
                        }
                    } else {
                        int l = ni + 1, r = ni + n->offset;
                        int hl = box_ray_gpu(&T->nodes[l].box, o, d, dinv, tmax, dim, &b0, &b1);
                        int hr = box_ray_gpu(&T->nodes[r].box, o, d, dinv, tmax, dim, &b2, &b3);
                        if (hl && hr) {
                            int closer = l, other = r;
                            if (b2 < b0) {
                                float t = b0; b0 = b2; b2 = t;
                                closer = r; other = l;
                            }
                            st[++sp].node = other; st[sp].dist = b2;
                            st[++sp].node = closer; st[sp].dist = b0;
                        } else if (hl) {
                            st[++sp].node = l; st[sp].dist = b0;
                        } else if (hr) {
                            st[++sp].node = r; st[sp].dist = b2;
                        }
                    }
                }
            }
        }
        c->out[qi] = best;
        if (c->object_out) c->object_out[qi] = best_inst >= 0 ? c->inst[best_inst].object : -1;
    }
}

void oracle_instances_intersect(const oracle_bvh *tlas, oracle_bvh *const *objects, const oracle_instance *inst,
                                const float *origins, const float *dirs, const float *tmax, int n, int occlusion,
                                int nthreads, oracle_hit *out, int32_t *object_out)
{
    inst_ctx c;
    memset(&c, 0, sizeof(c));
    c.tlas = tlas;
    c.objects = objects;
    c.inst = inst;
    c.o = origins;
    c.d = dirs;
    c.tmax = tmax;
    c.occlusion = occlusion;
    c.out = out;
    c.object_out = object_out;
    parallel_for(n, nthreads, inst_ray_range, &c);
}

/* TransformedAggregate.intersect(sphere) (aggregate.slang:81-97): the sphere
   mapped by tInv, the object's walk and point sample, point and normal mapped
   back; d (the pdf) is kept */
static int inst_sph(const inst_ctx *c, int ii, const float *x, float r2, const float *rn, oracle_hit *hit)
{
    const oracle_instance *I = &c->inst[ii];
    const oracle_bvh *h = c->objects[I->object];
    int dim = h->dim;
    float xl[3] = {x[0], x[1], x[2]}, r2l = r2;
    if (I->transformed) {
        xf_point(I->tinv, x, dim, xl);
        r2l = xf_r2(I->tinv, x, xl, r2, dim);
    }
    if (!sph_gpu_one(h, xl, r2l, rn, hit)) return 0;
    if (I->transformed) {
        float p[3], n[3];
        xf_point(I->t, hit->p, dim, p);
        memcpy(hit->p, p, sizeof(p));
        xf_normal(I->tinv, hit->n, dim, n);
        memcpy(hit->n, n, sizeof(n));
    }
    return 1;
}

/* Bvh::processSubtreeForIntersection over aggregates (bvh.inl:1140-1250) with
   the constant branch weight, in GPU arithmetic: the walk picks, with equal
   probability, a child overlapping the sphere; in the reached leaf every
   instance is sampled with the walk's current u (and rn[1], rn[2]) and those
   that find a primitive are chosen among uniformly (reservoir weight 1); the
   chosen record's d is multiplied by the walk's pdf and divided by that count */
static void inst_sph_range(void *vctx, int start, int end)
{
    inst_ctx *c = (inst_ctx *)vctx;
    const oracle_bvh *T = c->tlas;
    int dim = T->dim;
    for (int qi = start; qi < end; qi++) {
        float x[3] = {c->q[(size_t)qi * 4], c->q[(size_t)qi * 4 + 1], 0.0f};
        if (dim == 3) x[2] = c->q[(size_t)qi * 4 + 2];
        float r2 = c->q[(size_t)qi * 4 + 3];
        const float *rn = c->r2 + (size_t)qi * 3;
        oracle_hit best, cand;
        hit_default(&best);
        int best_inst = -1;
        if (T->n_prims == 1) {
            if (inst_sph(c, T->prim_orig[0], x, r2, rn, &cand)) {
                best = cand;
                best_inst = T->prim_orig[0];
            }
        } else {
            float d0, d1, d2, d3;
            box_sqdist_gpu(&T->nodes[0].box, x, dim, &d0, &d1);
            if (d0 <= r2) {
                float pdf = 1.0f, u = rn[0];
                int cur = 0;
                for (;;) {
                    const node_t *nd = &T->nodes[cur];
                    if (nd->nrefs > 0) {
                        float total = 0.0f;
                        for (int k = 0; k < nd->nrefs; k++) {
                            int ii = T->prim_orig[nd->offset + k];
                            float mr[3] = {u, rn[1], rn[2]};
                            if (inst_sph(c, ii, x, r2, mr, &cand)) {
                                total += 1.0f;
                                float prob = 1.0f / total;
                                if (u < prob) {
                                    u = u / prob;
                                    best = cand;
                                    best.d *= pdf;
                                    best_inst = ii;
                                } else {
                                    u = (u - prob) / (1.0f - prob);
                                }
                            }
                        }
                        if (total > 0.0f) best.d /= total;
                        break;
                    }
                    int l = cur + 1, r = cur + nd->offset;
                    box_sqdist_gpu(&T->nodes[l].box, x, dim, &d0, &d1);
                    float wl = (d0 <= r2) ? 1.0f : 0.0f;
                    box_sqdist_gpu(&T->nodes[r].box, x, dim, &d2, &d3);
                    float wr = (d2 <= r2) ? 1.0f : 0.0f;
                    float tot = wl + wr;
                    if (!(tot > 0.0f)) break;
                    float pl = wl / tot, pr = 1.0f - pl;
                    if (u < pl) {
                        u = u / pl;
                        cur = l;
                        pdf *= pl;
                    } else {
                        u = (u - pl) / pr;
                        cur = r;
                        pdf *= pr;
                    }
                }
            }
        }
        c->out[qi] = best;
        if (c->object_out) c->object_out[qi] = best_inst >= 0 ? c->inst[best_inst].object : -1;
    }
}

void oracle_instances_intersect_spheres(const oracle_bvh *tlas, oracle_bvh *const *objects,
                                        const oracle_instance *inst, const float *spheres, const float *rand_nums,
                                        int n, int nthreads, oracle_hit *out, int32_t *object_out)
{
    inst_ctx c;
    memset(&c, 0, sizeof(c));
    c.tlas = tlas;
    c.objects = objects;
    c.inst = inst;
    c.q = spheres;
    c.r2 = rand_nums;
    c.out = out;
    c.object_out = object_out;
    parallel_for(n, nthreads, inst_sph_range, &c);
}
