/*
 * fcpw_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference FCPW query path, used as the parity
 * checker for the HIP kernels.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product (fcpw_amd) never does.
 *
 * What is restated (reference files relative to the fcpw source tree):
 *   - binned SAH BVH build          include/fcpw/aggregates/bvh.inl:3-234,427-466
 *   - leaf-order vertex renumbering include/fcpw/fcpw.inl:613-689 (indexMap)
 *   - CPU closest-point traversal   include/fcpw/aggregates/bvh.inl:1290-1418
 *   - CPU ray traversal             include/fcpw/aggregates/bvh.inl:655-771
 *   - CPU refit                     include/fcpw/aggregates/bvh.inl:548-599
 *   - brute-force aggregate         include/fcpw/aggregates/baseline.inl:342-384
 *   - GPU traversal order           include/fcpw/gpu/bvh.slang:144-254,395-502
 *   - GPU sphere intersection       include/fcpw/gpu/bvh.slang:257-392
 *   - silhouettes and SNCH cones    include/fcpw/fcpw.inl:405-521,709-756,
 *                                   include/fcpw/aggregates/bvh.inl:236-420
 *   - GPU closest silhouette point  include/fcpw/gpu/bvh.slang:503-646,
 *                                   include/fcpw/gpu/geometry.slang:490-708,
 *                                   include/fcpw/gpu/bounding-volumes.slang:122-236
 *   - several objects, instances    include/fcpw/fcpw.inl:1052-1126,
 *                                   include/fcpw/core/primitive.h:289-440,
 *                                   include/fcpw/gpu/aggregate.slang:45-145,
 *                                   include/fcpw/gpu/transform.slang:1-56
 *   - primitive routines            include/fcpw/geometry/triangles.inl:12-32,222-260,374-487,
 *                                   include/fcpw/geometry/line_segments.inl:11-28,140-175,200-255,
 *                                   include/fcpw/gpu/geometry.slang:37-103,202-332
 * plus an FP64 brute force written with an independent method (plane projection
 * and edge clamping), which is the error reference for every FP32 path.
 */
#ifndef FCPW_ORACLE_H
#define FCPW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 40-byte record, the same layout as the reference GPUInteraction
   (include/fcpw/gpu/bvh_interop_structures.h:760-774).  For the CPU-order
   paths `index` is Interaction::primitiveIndex (-1 on a miss, whose bit
   pattern equals FCPW_GPU_UINT_MAX). 2D results keep p.z = n.z = 0, uv.y = 0. */
typedef struct {
    float p[3];
    float n[3];
    float uv[2];
    float d;
    int32_t index;
} oracle_hit;

/* cost heuristics, numbered as fcpw::CostHeuristic (bvh.h:11-17) */
enum { ORACLE_LONGEST_AXIS_CENTER = 0, ORACLE_SURFACE_AREA = 1, ORACLE_OVERLAP_SURFACE_AREA = 2,
       ORACLE_VOLUME = 3, ORACLE_OVERLAP_VOLUME = 4 };

typedef struct oracle_bvh oracle_bvh;

/* dim = 2 (line segments, 2 indices per primitive) or 3 (triangles, 3 indices) */
oracle_bvh *oracle_build(int dim, const float *positions, int n_vertices,
                         const int32_t *indices, int n_prims,
                         int heuristic, int leaf_size, int n_buckets);
/* the same with packLeaves (bvh.inl:3-37,113-140): the build of the vectorized
   aggregate passes pack_leaves = 1 for the SAH heuristics and leaf_size =
   FCPW_SIMD_WIDTH (fcpw.inl:812-827) */
oracle_bvh *oracle_build_packed(int dim, const float *positions, int n_vertices, const int32_t *indices,
                                int n_prims, int heuristic, int leaf_size, int n_buckets, int pack_leaves);
void oracle_free(oracle_bvh *h);

int oracle_num_nodes(const oracle_bvh *h);
int oracle_dim(const oracle_bvh *h);
int oracle_num_prims(const oracle_bvh *h);
/* BoundingBox::surfaceArea (bounding_volumes.h:164-167) of a box given as 3 floats */
float oracle_box_surface_area(const float *pmin, const float *pmax, int dim);
/* per leaf slot, the positions of its dim vertices (n_prims x dim x dim floats) */
void oracle_export_prim_vertices(const oracle_bvh *h, float *out);
int oracle_num_leaves(const oracle_bvh *h);
int oracle_max_depth(const oracle_bvh *h);

/* flat tree export: boxes padded to 3 floats (z = 0 in 2D), offset = reference
   offset for leaves or secondChildOffset for internal nodes (nrefs == 0) */
void oracle_export_nodes(const oracle_bvh *h, float *pmin, float *pmax, int32_t *nrefs, int32_t *offset);
/* prim_order[j] = original primitive index stored at leaf slot j */
void oracle_export_prim_order(const oracle_bvh *h, int32_t *prim_order);
/* index_map[user vertex] = sorted vertex (fcpw.inl:652-689); -1 for unreferenced */
void oracle_export_index_map(const oracle_bvh *h, int32_t *index_map);

/* Scene::updateObjectVertices (fcpw.inl:1128-1143) followed by Scene::refit */
void oracle_update_vertices(oracle_bvh *h, const float *positions_user_order, int n_vertices);
void oracle_refit(oracle_bvh *h);

/* nthreads <= 0: one thread per online core, static partition as fcpw.inl:1547-1559 */
void oracle_closest_points_cpu(const oracle_bvh *h, const float *queries, const float *r2, int n,
                               int record_normal, int nthreads, oracle_hit *out);
/* reference GPU order; n_internal / n_leaf_prims (optional) receive per-query
   counts of expanded internal nodes and primitive tests; n_pops the node pops */
void oracle_closest_points_gpuorder(const oracle_bvh *h, const float *queries, const float *r2, int n,
                                    int record_normal, int nthreads, oracle_hit *out,
                                    int32_t *n_internal, int32_t *n_leaf_prims, int32_t *n_pops);
void oracle_intersect_cpu(const oracle_bvh *h, const float *origins, const float *dirs,
                          const float *tmax, int n, int occlusion, int nthreads, oracle_hit *out);
void oracle_intersect_gpuorder(const oracle_bvh *h, const float *origins, const float *dirs,
                               const float *tmax, int n, int occlusion, int nthreads, oracle_hit *out,
                               int32_t *n_internal, int32_t *n_leaf_prims);
/* Baseline aggregate: every primitive in scene order, FP32 (baseline.inl:342-384) */
/* GPU sphere intersection (bvh.slang:257-392, constant branch weight):
   spheres n x 4 (center, r^2), rand_nums n x 3 */
void oracle_intersect_spheres(const oracle_bvh *h, const float *spheres, const float *rand_nums, int n,
                              int nthreads, oracle_hit *out);
void oracle_brute_closest_f32(const oracle_bvh *h, const float *queries, const float *r2, int n,
                              int nthreads, oracle_hit *out);

/* ---- silhouettes (SNCH) ----
   GPU silhouette record in the reference GPUEdge layout (bvh_interop_structures.h:141-162);
   in 2D a GPUVertex (:121-139) is stored with p in pa, pb = 0, and z = 0 throughout. */
typedef struct {
    float pa[3], pb[3], n0[3], n1[3];
    uint32_t index, has_one_adjacent_face;
} oracle_silhouette;
/* Scene::computeSilhouettes' ignoreSilhouette(angle, primitive index) callback */
typedef int (*oracle_ignore_fn)(float angle, int32_t index, void *user);
/* Scene::computeSilhouettes + the SNCH data of Bvh::assignGeometricDataToNodes
   (fcpw.inl:405-521,709-756; bvh.inl:236-420).  user_idx: the primitives'
   vertex ids as passed to oracle_build.  Returns the number of references. */
int oracle_build_silhouettes(oracle_bvh *h, const int32_t *user_idx, oracle_ignore_fn ignore, void *user);
int oracle_num_silhouette_refs(const oracle_bvh *h);
/* per node: cones (axis[3], halfAngle, radius), nSilhouettes, silhouetteOffset;
   per reference: the GPU record (any pointer may be NULL) */
void oracle_export_silhouettes(const oracle_bvh *h, float *cones, int32_t *n_silhouettes, int32_t *silhouette_offset,
                               oracle_silhouette *records);
/* recompute the GPU records from the current positions (refit with updateGeometry) */
void oracle_update_silhouette_records(oracle_bvh *h);
/* GPU refit of the cones (bvh.slang:27-141) from the current boxes and records */
void oracle_refit_silhouette_cones_gpu(oracle_bvh *h);
/* GPU closest silhouette point (bvh.slang:503-646): spheres n x 4, flip n;
   brute != 0 tests every reference in order instead of traversing */
void oracle_closest_silhouettes_gpuorder(const oracle_bvh *h, const float *spheres, const uint32_t *flip, int n,
                                         float squared_min_radius, float precision, int brute, int nthreads,
                                         oracle_hit *out);

/* FP64 closest silhouette point on raw geometry (user order) with an independent
   orientation classification; returns -1 (and computes nothing) for
   non-manifold input.  d_strict / d_loose bracket the exact answer (see the
   definition beside the implementation); p_strict is n x 3. */
int oracle_closest_silhouette_f64(int dim, const float *positions, int n_vertices, const int32_t *indices,
                                  int n_prims, const float *queries, int n, double tol, double tol_d, int nthreads,
                                  double *d_strict, double *d_loose, double *p_strict);

/* ---- scenes of several objects and transformed instances ----
   Scene::build's instance list (fcpw.inl:1066-1085): per object, in order, one
   untransformed instance when it has no instance transforms, else one
   transformed instance per transform.  t / tinv: 3x4 row-major affine matrices
   (2D: rows 0-1, columns 0, 1 and 3). */
typedef struct {
    int32_t object;
    int32_t transformed;
    float t[12];
    float tinv[12];
} oracle_instance;
/* Transform::inverse of an affine transform (cofactor inverse of the linear part) */
void oracle_invert_transform(int dim, const float *t, float *tinv);
/* the BVH over the instances (makeAggregate over aggregates, fcpw.inl:1103-1112);
   its "primitives" are instance ids: prim_order gives the instance of each leaf slot */
oracle_bvh *oracle_build_instances(int dim, oracle_bvh *const *objects, int n_objects, const oracle_instance *inst,
                                   int n_inst, int heuristic, int leaf_size, int n_buckets);
/* Bvh::refit over aggregates, after the objects themselves were refit */
void oracle_refit_instances(oracle_bvh *tlas, oracle_bvh *const *objects, const oracle_instance *inst);
/* GPU order at both levels; object_out (optional) = Interaction::objectIndex, -1 on a miss */
void oracle_instances_closest_points(const oracle_bvh *tlas, oracle_bvh *const *objects, const oracle_instance *inst,
                                     const float *queries, const float *r2, int n, int record_normal, int nthreads,
                                     oracle_hit *out, int32_t *object_out);
void oracle_instances_intersect(const oracle_bvh *tlas, oracle_bvh *const *objects, const oracle_instance *inst,
                                const float *origins, const float *dirs, const float *tmax, int n, int occlusion,
                                int nthreads, oracle_hit *out, int32_t *object_out);
/* spheres n x 4 (center, r^2), rand_nums n x 3: the walk over the instance tree
   of bvh.inl:1140-1250, each instance sampled as aggregate.slang:81-97 */
void oracle_instances_intersect_spheres(const oracle_bvh *tlas, oracle_bvh *const *objects,
                                        const oracle_instance *inst, const float *spheres, const float *rand_nums,
                                        int n, int nthreads, oracle_hit *out, int32_t *object_out);

/* FP64 brute force on raw geometry (positions in user order) */
void oracle_brute_closest_f64(int dim, const float *positions, int n_vertices, const int32_t *indices,
                              int n_prims, const float *queries, int n, int nthreads,
                              double *d, double *p, int32_t *index);
void oracle_brute_ray_f64(int dim, const float *positions, int n_vertices, const int32_t *indices,
                          int n_prims, const float *origins, const float *dirs, int n, int nthreads,
                          double *t, int32_t *index);
/* FP64 distance from each query to one given primitive (tie classification) */
void oracle_prim_dist_f64(int dim, const float *positions, const int32_t *indices,
                          const float *queries, const int32_t *prims, int n, double *d);

/* ---- the vectorized CPU aggregate (fcpw_oracle_mbvh.cpp), compiled for 8
   lanes (AVX2 + FMA, suffix _w8) and 16 lanes (AVX-512, suffix _w16) ---- */
#define ORACLE_MBVH_DECL(W)                                                                                 \
    void *oracle_mbvh_build_w##W(const oracle_bvh *h);                                                     \
    void oracle_mbvh_free_w##W(void *m);                                                                    \
    void oracle_mbvh_info_w##W(const void *m, int64_t *n_nodes, int64_t *n_leaf_groups, int64_t *max_depth); \
    void oracle_mbvh_export_w##W(const void *m, float *bmin, float *bmax, int32_t *child, int32_t *leaf_prim); \
    void oracle_mbvh_closest_points_w##W(const void *m, const float *queries, const float *r2, int n,       \
                                         int nthreads, oracle_hit *out, int64_t *visited);
ORACLE_MBVH_DECL(8)
ORACLE_MBVH_DECL(16)

#ifdef __cplusplus
}
#endif
#endif
