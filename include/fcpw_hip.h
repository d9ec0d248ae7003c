/*
 * fcpw_hip.h -- C ABI of the MI355X query backend for FCPW (libfcpw_hip.so).
 *
 * The backend replaces the Slang/slang-rhi GPU path of FCPW's GPUScene<DIM>
 * (include/fcpw/fcpw_gpu.h:9-93, include/fcpw/fcpw_gpu.inl,
 *  include/fcpw/gpu/bvh.cs.slang, include/fcpw/gpu/bvh.slang).
 * Plain pointers and sizes only; every function returns an fcpw_hip_status
 * and records a message readable through fcpw_hip_last_error().
 *
 * Conventions
 *  - The caller owns every pointer it passes; the context owns what it uploads.
 *  - "_device" entry points take device pointers and a hipStream_t (passed as
 *    void*, NULL = default stream); they are stream-ordered and never
 *    synchronize.  The other query entry points take host pointers and return
 *    after the results are back on the host (as the reference's submit +
 *    waitOnHost, include/fcpw/gpu/slang_rhi_utils.h:582-583).
 *  - One context per thread: like the reference (fcpw_gpu.h:30), a context is
 *    not thread-safe.
 *  - Record layouts are those of include/fcpw/gpu/bvh_interop_structures.h:730-774.
 */
#ifndef FCPW_HIP_H
#define FCPW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FCPW_HIP_UINT_MAX 4294967295u /* FCPW_GPU_UINT_MAX, bvh_interop_structures.h:10 */

typedef enum {
    FCPW_HIP_OK = 0,
    FCPW_HIP_ERR_INVALID_ARGUMENT = 1,
    FCPW_HIP_ERR_NOT_UPLOADED = 2,   /* query or refit before fcpw_hip_upload_bvh */
    FCPW_HIP_ERR_HIP = 3,            /* a HIP runtime call failed */
    FCPW_HIP_ERR_OUT_OF_MEMORY = 4,
    FCPW_HIP_ERR_UNSUPPORTED = 5,    /* a reference feature this backend does not build */
    FCPW_HIP_ERR_TRAVERSAL_OVERFLOW = 6
} fcpw_hip_status;

/* float3 / float2 as bvh_interop_structures.h:14-20 (float2 is 8-byte aligned) */
typedef struct { float x, y, z; } fcpw_hip_float3;
typedef struct { float x, y; } __attribute__((aligned(8))) fcpw_hip_float2;

/* GPUBoundingSphere (bvh_interop_structures.h:749-758), 16 bytes */
typedef struct { fcpw_hip_float3 c; float r2; } fcpw_hip_bounding_sphere;

/* GPURay (bvh_interop_structures.h:730-747), 40 bytes; dinv = 1/d per component */
typedef struct { fcpw_hip_float3 o, d, dinv; float tmax; } fcpw_hip_ray;

/* GPUInteraction (bvh_interop_structures.h:760-774), 40 bytes; a miss keeps
   p = n = uv = 0, d = FLT_MAX, index = FCPW_HIP_UINT_MAX */
typedef struct {
    fcpw_hip_float3 p, n;
    fcpw_hip_float2 uv;
    float d;
    uint32_t index;
} fcpw_hip_interaction;

/* GPUBoundingCone (bvh_interop_structures.h:41-52), 20 bytes; half_angle < 0
   marks a node with no silhouette */
typedef struct { fcpw_hip_float3 axis; float half_angle; float radius; } fcpw_hip_bounding_cone;

/* GPUVertex (bvh_interop_structures.h:121-139), 44 bytes: a 2D silhouette
   vertex with the unit normals of its two adjacent segments (z = 0) */
typedef struct {
    fcpw_hip_float3 p, n0, n1;
    uint32_t index, has_one_adjacent_face;
} fcpw_hip_vertex;

/* GPUEdge (bvh_interop_structures.h:141-162), 56 bytes: a 3D silhouette edge
   with the unit normals of its two adjacent triangles */
typedef struct {
    fcpw_hip_float3 pa, pb, n0, n1;
    uint32_t index, has_one_adjacent_face;
} fcpw_hip_edge;

typedef struct fcpw_hip_ctx fcpw_hip_ctx;

/* Stands in for GPUContext::initDevice (slang_rhi_utils.h:23-62) at
   GPUScene construction: binds the context to a HIP device. */
int fcpw_hip_create(int device, fcpw_hip_ctx **out);
int fcpw_hip_destroy(fcpw_hip_ctx *ctx);
/* Last error message of ctx, or of the failed create when ctx is NULL. */
const char *fcpw_hip_last_error(const fcpw_hip_ctx *ctx);

/*
 * Replaces GPUBvhBuffers::allocate<DIM> (bvh_interop_structures.h:656-728),
 * called from GPUScene::transferToGPU (fcpw_gpu.inl:18-62).
 * The BVH arrives as the reference's flat DFS array (bvh.h:19-31, the data
 * extractBvhNodes reads, bvh_interop_structures.h:172-190): node i has box
 * [pmin, pmax] (3 floats each, z = 0 in 2D), nrefs > 0 for a leaf whose
 * primitives occupy leaf slots [offset, offset + nrefs), nrefs == 0 for an
 * internal node whose children are i + 1 and i + offset.
 * Primitives arrive in leaf-slot order (extractTriangles / extractLineSegments,
 * bvh_interop_structures.h:220-275): prim_vertices holds dim vertex ids per
 * slot into positions (n_vertices x dim floats, the caller's vertex order) and
 * prim_index the original primitive index reported in results.
 * dim = 3: triangles; dim = 2: line segments.  Replaces any previous upload.
 */
int fcpw_hip_upload_bvh(fcpw_hip_ctx *ctx, int dim, int n_nodes, const float *node_pmin,
                        const float *node_pmax, const int32_t *node_nrefs, const int32_t *node_offset,
                        int n_prims, const int32_t *prim_vertices, const int32_t *prim_index,
                        int n_vertices, const float *positions);

/*
 * Builds the BVH on the device instead of the host: the reference's binned SAH
 * build (Bvh::build, bvh.inl:3-234, with the leaf size 4 and 8 buckets of a
 * non-vectorized Scene::build, fcpw.inl:796-867) followed by the upload of
 * fcpw_hip_upload_bvh, in one call.  The tree is the reference's: the same
 * nodes in the same DFS order, the same boxes, and each leaf's primitives in
 * the same order.  positions: n_vertices x dim floats; indices: n_prims x dim
 * vertex ids (caller's order); heuristic numbered as fcpw::CostHeuristic
 * (0 LongestAxisCenter, 1 SurfaceArea, 2 OverlapSurfaceArea, 3 Volume,
 * 4 OverlapVolume).  Optional outputs (NULL to skip) receive the flat tree as
 * fcpw_hip_upload_bvh takes it: node_pmin / node_pmax (n x 3), node_nrefs,
 * node_offset (n entries each, room for 2 * n_prims - 1 nodes), prim_order
 * (n_prims: original primitive per leaf slot) and n_nodes.  Silhouette data
 * is built on the host from that exported tree: fcpw_hip_upload_silhouettes
 * needs node_nrefs and node_offset to have been requested.  Replaces any
 * previous upload.  (No reference counterpart: the
 * reference builds on the CPU and transfers, README.md:84-86.)
 */
int fcpw_hip_build_bvh(fcpw_hip_ctx *ctx, int dim, const float *positions, int n_vertices, const int32_t *indices,
                       int n_prims, int heuristic, float *node_pmin, float *node_pmax, int32_t *node_nrefs,
                       int32_t *node_offset, int32_t *prim_order, int *n_nodes);

/*
 * Replaces the geometry re-upload of GPUScene::refit(scene, updateGeometry=true)
 * (fcpw_gpu.inl:64-80 -> bvh_interop_structures.h:582-603): new vertex
 * positions in the caller's vertex order; the primitive records are rebuilt on
 * the device.  Call fcpw_hip_refit afterwards to update the boxes.
 */
int fcpw_hip_update_positions(fcpw_hip_ctx *ctx, const float *positions, int n_vertices);
int fcpw_hip_update_positions_device(fcpw_hip_ctx *ctx, const float *positions, int n_vertices,
                                     void *stream);

/* Replaces runBvhUpdate + the refit entry point (bvh_interop_structures.h:891-934,
   bvh.cs.slang:135-149): one bottom-up launch recomputes every box; with
   silhouettes uploaded, a second one recomputes every cone from the current
   silhouette records (Bvh.refit, bvh.slang:27-141). */
int fcpw_hip_refit(fcpw_hip_ctx *ctx, void *stream);

/*
 * Replaces GPUScene::findClosestPoints (fcpw_gpu.inl:255-274) and the
 * closestPoint entry point (bvh.cs.slang:84-105).  record_normals != 0 writes
 * the face normal of the closest primitive (bvh.slang:496-499).
 */
int fcpw_hip_closest_points(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres, int64_t n,
                            int record_normals, fcpw_hip_interaction *out);
int fcpw_hip_closest_points_device(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres, int64_t n,
                                   int record_normals, fcpw_hip_interaction *out, void *stream);

/*
 * fcpw_hip_closest_points_device without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory during the call
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as fcpw_hip_closest_points_device
 * on a single-object scene; the query-ordering workspace lives in a
 * file-scope device array of the library (about 25 MiB per GPU, held until
 * the process ends) instead of the context's grow-on-demand buffers, with a
 * 2^18-cell ordering grid.  At most 2,097,152 queries per call (more returns
 * FCPW_HIP_ERR_INVALID_ARGUMENT before anything is launched).  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams, or from different contexts) race.  The "timing" and
 * "stats" switches do not apply.
 */
int fcpw_hip_closest_points_device_noalloc(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres, int64_t n,
                                           int record_normals, fcpw_hip_interaction *out, void *stream);

/*
 * Replaces GPUScene::intersect(rays) (fcpw_gpu.inl:137-156) and the
 * rayIntersection entry point (bvh.cs.slang:36-57).  With check_for_occlusion
 * the first hit ends the query and only its index is written (bvh.slang:183-187).
 */
int fcpw_hip_intersect(fcpw_hip_ctx *ctx, const fcpw_hip_ray *rays, int64_t n, int check_for_occlusion,
                       fcpw_hip_interaction *out);
int fcpw_hip_intersect_device(fcpw_hip_ctx *ctx, const fcpw_hip_ray *rays, int64_t n, int check_for_occlusion,
                              fcpw_hip_interaction *out, void *stream);

/*
 * Replaces GPUScene::intersect(spheres, randNums) (fcpw_gpu.inl:158-217) and the
 * sphereIntersection entry point (bvh.cs.slang:59-82, bvh.slang:257-392): per
 * sphere, one primitive inside it is chosen at random (one random walk down the
 * overlapping children with equal probability, then area-weighted among the
 * leaf's primitives) and a uniform point on it is sampled.  out: p, n, uv of the
 * point, d = its pdf, index = the primitive; a sphere with no primitive keeps
 * the default record.  rand_nums: three numbers in [0, 1) per sphere (the first
 * drives the choice, the other two the point; the third is unused in 2D).
 */
int fcpw_hip_intersect_spheres(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres,
                               const fcpw_hip_float3 *rand_nums, int64_t n, fcpw_hip_interaction *out);
int fcpw_hip_intersect_spheres_device(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres,
                                      const fcpw_hip_float3 *rand_nums, int64_t n, fcpw_hip_interaction *out,
                                      void *stream);

/*
 * Replaces the silhouette half of GPUBvhBuffers::allocate<DIM> for a scene with
 * silhouettes (the SNCH node type): the cone and silhouette range of every
 * node (extractSnchNodes, bvh_interop_structures.h:193-218) and the silhouette
 * records in reference order (extractSilhouetteVertices / extractSilhouetteEdges,
 * :237-300).  Call after fcpw_hip_upload_bvh with the same n_nodes; node i's
 * silhouettes are records [silhouette_offset[i], + n_silhouettes[i]) and only
 * leaves have any.  silhouettes: n_refs fcpw_hip_vertex (dim 2) or
 * fcpw_hip_edge (dim 3).  Calling it again replaces the records and cones (the
 * refit with updateGeometry re-extracts the records from new positions).
 */
int fcpw_hip_upload_silhouettes(fcpw_hip_ctx *ctx, int n_nodes, const fcpw_hip_bounding_cone *cones,
                                const int32_t *n_silhouettes, const int32_t *silhouette_offset, int n_refs,
                                const void *silhouettes);

/*
 * Replaces GPUScene::findClosestSilhouettePoints (fcpw_gpu.inl:276-338) and the
 * closestSilhouettePoint entry point (bvh.cs.slang:107-133, bvh.slang:503-646):
 * per sphere, the closest point on the visibility silhouette (2D: vertices where
 * one adjacent segment faces the query and the other does not; 3D: such edges;
 * boundary vertices/edges always count) within sqrt(r2), pruned by the node
 * cones.  flip_normal_orientation[i] != 0 flips the orientation used for
 * queries on the geometry (distance <= precision).  The search stops once a
 * silhouette within sqrt(squared_min_radius) is found.  out: p, uv (edge
 * parameter in uv.x), d, index (2D: the vertex id in leaf order, the
 * reference's sorted numbering; 3D: the edge id); n = 0.  A sphere with no
 * silhouette keeps the default record.  Needs fcpw_hip_upload_silhouettes
 * (FCPW_HIP_ERR_UNSUPPORTED otherwise, as the reference kernel built without
 * silhouette data finds nothing).  2D queries use z = 0.
 */
int fcpw_hip_closest_silhouette_points(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres,
                                       const uint32_t *flip_normal_orientation, int64_t n,
                                       float squared_min_radius, float precision, fcpw_hip_interaction *out);
int fcpw_hip_closest_silhouette_points_device(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres,
                                              const uint32_t *flip_normal_orientation, int64_t n,
                                              float squared_min_radius, float precision,
                                              fcpw_hip_interaction *out, void *stream);

/*
 * Compact closest-point query for differentiable use (no reference
 * counterpart; the reference has no gradients).  points: n x dim floats;
 * r2: n floats or NULL for an unbounded search.  Writes d (n), p (n x dim),
 * prim (n, original primitive index or -1) and slot (n, leaf slot or -1, the
 * handle fcpw_hip_distance_backward_device recomputes the primitive from).
 * Any output pointer may be NULL.
 */
int fcpw_hip_closest_points_compact_device(fcpw_hip_ctx *ctx, const float *points, const float *r2, int64_t n,
                                           float *d, float *p, int32_t *prim, int32_t *slot, void *stream);

/*
 * Adjoint of the unsigned distance d(x, mesh) (no reference counterpart).
 * From x, the saved leaf slot and grad_d = dL/dd it recomputes the closest
 * point p and its barycentric weights b, and writes
 *   grad_points[i]  = grad_d[i] * (x - p) / d                 (n x dim, overwritten)
 *   grad_vertices[v] += -grad_d[i] * b_k * (x - p) / d        (n_vertices x dim, FP32 atomics)
 * in the caller's vertex order.  Queries with d == 0 or slot < 0 contribute 0.
 * Either output may be NULL.  grad_vertices must be device memory from
 * hipMalloc (coarse-grained): the accumulation uses the hardware FP32 atomic
 * add, whose summation order is not fixed.
 */
int fcpw_hip_distance_backward_device(fcpw_hip_ctx *ctx, const float *points, const int32_t *slot,
                                      const float *grad_d, int64_t n, float *grad_points, float *grad_vertices,
                                      void *stream);

/*
 * Scenes of several objects and transformed instances.  Replaces, for the GPU,
 * the instance layer Scene::build puts above the objects' trees
 * (fcpw.inl:1052-1126; TransformedAggregate, primitive.h:289-440, whose GPU
 * twin is aggregate.slang:45-145 with transform.slang:1-56); the reference's
 * own GPU path serves one object (fcpw_gpu.h:17-20).
 *
 * Each object is a context with its BVH uploaded (fcpw_hip_upload_bvh or
 * fcpw_hip_build_bvh).  fcpw_hip_upload_instances turns ctx into the scene:
 * instances[i] draws object instances[i].object, transformed by t when
 * .transformed != 0 (an object without instance transforms is one
 * untransformed instance), and the flat tree over the instances (same layout
 * as fcpw_hip_upload_bvh; leaf slot j holds instance leaf_instances[j]) bounds
 * them.  The object contexts must stay alive while ctx is used; re-uploading
 * one is picked up by the next call.  A transform must be affine with uniform
 * scaling (fcpw.h:68): 3x4 row-major, 2D in rows 0-1 with columns 0, 1
 * (linear part) and 3 (translation); tinv is its inverse.
 *
 * On such a ctx, fcpw_hip_closest_points(_device), fcpw_hip_intersect(_device)
 * and fcpw_hip_intersect_spheres(_device) answer over all instances; index is
 * the primitive index within its object, and the _objects variants also write
 * the object index (Interaction::objectIndex, interaction.h:67) or -1 on a
 * miss (0 for every hit of a one-object scene).  Per instance the query follows
 * TransformedAggregate: sphere / ray mapped by tinv, the object's tree
 * traversed, r^2 / tMax mapped back by t; a hit's point is mapped by t, its
 * normal by transpose(tinv) (normalized) and d = |p - x| (rays: |p - o|, which
 * equals t for unit directions).  With occlusion only index and object are
 * written.  Sphere intersection walks the instance tree as
 * Bvh::processSubtreeForIntersection does over aggregates (bvh.inl:1140-1250):
 * in the reached leaf, each instance is sampled with the walk's current random
 * number and one of those that find a primitive is chosen uniformly; d = the
 * instance's pdf (in its own space, as TransformedAggregate keeps it) times the
 * walk's pdf over that count.  fcpw_hip_refit refits every object, then the
 * instance boxes and the tree above them (bvh.inl:585-599).  Silhouette queries
 * return FCPW_HIP_ERR_UNSUPPORTED.
 */
typedef struct {
    int32_t object;
    int32_t transformed;
    float t[12];
    float tinv[12];
} fcpw_hip_instance;

int fcpw_hip_upload_instances(fcpw_hip_ctx *ctx, fcpw_hip_ctx *const *objects, int n_objects,
                              const fcpw_hip_instance *instances, int n_instances, int n_nodes,
                              const float *node_pmin, const float *node_pmax, const int32_t *node_nrefs,
                              const int32_t *node_offset, const int32_t *leaf_instances);
int fcpw_hip_closest_points_objects(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres, int64_t n,
                                    int record_normals, fcpw_hip_interaction *out, int32_t *object_index);
int fcpw_hip_closest_points_objects_device(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres, int64_t n,
                                           int record_normals, fcpw_hip_interaction *out, int32_t *object_index,
                                           void *stream);
int fcpw_hip_intersect_objects(fcpw_hip_ctx *ctx, const fcpw_hip_ray *rays, int64_t n, int check_for_occlusion,
                               fcpw_hip_interaction *out, int32_t *object_index);
int fcpw_hip_intersect_objects_device(fcpw_hip_ctx *ctx, const fcpw_hip_ray *rays, int64_t n,
                                      int check_for_occlusion, fcpw_hip_interaction *out, int32_t *object_index,
                                      void *stream);
int fcpw_hip_intersect_spheres_objects(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres,
                                       const fcpw_hip_float3 *rand_nums, int64_t n, fcpw_hip_interaction *out,
                                       int32_t *object_index);
int fcpw_hip_intersect_spheres_objects_device(fcpw_hip_ctx *ctx, const fcpw_hip_bounding_sphere *spheres,
                                              const fcpw_hip_float3 *rand_nums, int64_t n, fcpw_hip_interaction *out,
                                              int32_t *object_index, void *stream);

/*
 * Batched queries over several uploaded contexts in one persistent launch (no
 * reference counterpart: the reference GPU path serves one object per
 * GPUScene, fcpw_gpu.h:17-20).  The scenes must share dim and device and stay
 * alive while the batch is used; a re-uploaded member is picked up by the next
 * call.  Queries are concatenated in scene order: counts[s] queries for scene
 * s (fcpw_hip_batch_set_queries, which also sizes every workspace).  After
 * that, the query entry points allocate nothing and only enqueue kernels on
 * the stream, so a caller may capture them in a HIP graph.  Results index each
 * scene's own primitives; grad_vertices is the concatenation of the scenes'
 * vertex gradients (each scene's vertices in its caller order).
 */
typedef struct fcpw_hip_batch fcpw_hip_batch;
int fcpw_hip_batch_create(fcpw_hip_ctx *const *scenes, int n_scenes, fcpw_hip_batch **out);
int fcpw_hip_batch_destroy(fcpw_hip_batch *batch);
/* Last error of batch, or of the failed create when batch is NULL. */
const char *fcpw_hip_batch_last_error(const fcpw_hip_batch *batch);
int fcpw_hip_batch_set_queries(fcpw_hip_batch *batch, const int64_t *counts);
int fcpw_hip_batch_closest_points_device(fcpw_hip_batch *batch, const fcpw_hip_bounding_sphere *spheres,
                                         int record_normals, fcpw_hip_interaction *out, void *stream);
int fcpw_hip_batch_closest_points_compact_device(fcpw_hip_batch *batch, const float *points, const float *r2,
                                                 float *d, float *p, int32_t *prim, int32_t *slot, void *stream);
int fcpw_hip_batch_distance_backward_device(fcpw_hip_batch *batch, const float *points, const int32_t *slot,
                                            const float *grad_d, float *grad_points, float *grad_vertices,
                                            void *stream);
/* Batch knobs (none changes a result bit): "stack" (0, 4, 6, 8, 12, 16; default 6), "hold_min"
   (1..64), "chunk" (queries per work item: 64, 128, 256, 512), "sort" (0, 1),
   "waves_per_cu" (0 = as many as fit, else 1..32). */
int fcpw_hip_batch_set_option(fcpw_hip_batch *batch, const char *key, int value);
int fcpw_hip_batch_get_option(const fcpw_hip_batch *batch, const char *key, int *value);

/* Tuning knobs read by the kernels (values chosen by fcpw_amd.autotune).  None
   of them changes a result bit: every setting visits nodes in the same order.
     "coop"      1 = wave-cooperative leaf processing (default), 0 = one query per
                 lane with the leaf loop inline (closest points, on single-object and
                 instanced scenes, and closest silhouette points)
     "ray_coop"  the same choice for ray queries (default 1)
     "ray_stack", "ray_hold_min"  the cooperative ray kernels' LDS stack entries (4, 8
                 or 16, default 4) and leaf-pass threshold (1..64, default 8)
     "sil_hold_min"  the cooperative silhouette kernel's leaf-pass threshold (default 3)
     "layout"    the order of the tree's records in HBM: 0 = DFS order of the internal
                 nodes (default), 1 = treelets (a node, its internal children and
                 grandchildren contiguous, treelets in DFS order); setting it on an
                 uploaded single-object scene reorders the records in place, and later
                 uploads and device builds of the context use it
     "sil_stack", "inst_stack"  LDS stack entries of the 3D cooperative silhouette kernel
                 (4, 8 or 12, default 4) and of the 3D cooperative instanced closest-point
                 kernel (8 or 12, default 12)
     "ray_sort_dir"  1 = rays ordered by direction octant (quadrant), then by origin
                 along the Morton curve (default); 0 = by origin only
     "hold_min"  coop: lanes that must hold a leaf before the wave processes
                 leaves together (1..64, default 1)
     "stack"     traversal-stack entries per lane kept in LDS (default 6): the newest
                 entries, as a sliding window; older ones are spilled once to cached
                 scratch: 0, 4, 6, 8, 12, 16 or 24 (the 2D coop kernel uses 16 for 24,
                 the one-query-per-lane kernel 8 for 0, 4, 6 and 12)
     "block"     threads per block when coop = 0 (64, 128, 256)
     "sort"      1 = order queries along a Morton curve before traversal (default)
     "morton_bits"  cells of that ordering grid, 2^18, 2^21 or 2^24 (default 24)
     "xcds"      number of groups the block -> tile mapping splits the Morton-ordered
                 tiles into (1..64, default 4: on the 8 XCDs each range is shared by two
                 XCDs, measured faster than one range per XCD on the 1M-triangle torus)
     "sil_xcds"  the same for the silhouette kernels (default 8)
     "xcd_chunk" 0 = one contiguous range per group (default); c > 0 = groups take
                 chunks of 2^c tiles in turn (0..16)
     "inst_seed" closest points on a scene of instances with ordered queries run
                 as two launches: per block of B queries of the Morton order, the
                 64 in its middle first, then the others with their radius lowered
                 to an upper bound from the nearest of those (its distance plus the
                 distance between the two queries, with margins), which never
                 changes a record; B = 1024 by default, 64..2^20, 0 = one launch
     "bwd_order" 1 = the adjoint walks the queries in the Morton order the latest
                 query call on this context computed, when that batch had the same
                 size (default); 0 = in the caller's order
     "bwd_reduce"  how a wave of that order combines its vertex terms before the
                 global FP32 atomics: 2 = sums per vertex in an LDS hash table, one
                 atomic per vertex and coordinate (default); 1 = consecutive queries
                 on the same primitive summed by a segmented wave scan; 0 = none
                 (both options change only the FP32 summation order of grad_vertices)
   Measurement switches (off by default; results are unchanged):
     "stats"     1 = run the counting variant of the traversal kernels, which adds
                 the internal nodes expanded, primitives tested and queries to
                 counters read back through fcpw_hip_get_info; setting 1 clears them
     "timing"    1 = record a HIP event pair on the launch stream around every
                 traversal kernel (up to 1024 launches); setting 1 starts a new series
   Unknown keys or values return FCPW_HIP_ERR_INVALID_ARGUMENT. */
int fcpw_hip_set_option(fcpw_hip_ctx *ctx, const char *key, int value);
int fcpw_hip_get_option(const fcpw_hip_ctx *ctx, const char *key, int *value);

/* Uploaded BVH facts: n_records (internal nodes), n_leaves, n_prims, n_vertices,
   dim, max_depth, n_cus, device (the HIP device the context was created on),
   generation (bumped by every upload, which moves the device buffers),
   positions_version (bumped by every upload and vertex-position update), n_objects, n_instances (0 unless instanced; on a scene of
   several objects the tree facts are the instance tree's: n_prims = instances,
   n_vertices = 0).  The instanced kernels run with fixed shapes (128 threads, 8
   LDS stack entries per object traversal, 4 for the instance tree); of the
   tuning knobs below only "sort", "morton_bits" and "xcds" apply to them.
   Measurement: stat_internal_nodes, stat_prim_tests,
   stat_queries and, for closest-point launches, stat_waves, stat_wave_iterations,
   stat_leaf_passes, stat_leaf_groups (cooperative leaf passes and their 64-item
   evaluation groups); totals since "stats" was set to 1; synchronizes the device;
   kernel_launches and kernel_ns (count and summed duration of the traversal
   launches timed since "timing" was set to 1; waits for the last one). */
int fcpw_hip_get_info(const fcpw_hip_ctx *ctx, const char *key, int64_t *value);

#ifdef __cplusplus
}
#endif
#endif /* FCPW_HIP_H */
