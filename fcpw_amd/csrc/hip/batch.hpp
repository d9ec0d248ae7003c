// batch.hpp -- many uploaded scenes queried in one persistent launch.  Included
// once, by fcpw_hip.hip, after the single-scene entry points.
//
// The reference GPU path serves one object per GPUScene (fcpw_gpu.h:17-20);
// batching has no reference counterpart.  Layout:
//   table    one BatchScene per member scene, in longest-first order (estimated
//            cost = queries x log2(primitives)), holding the scene's device
//            buffers, its Morton grid and its ranges in the query arrays
//   queries  concatenated in the caller's scene order (caller index i)
//   perm     per-scene Morton order: positions [p_begin, p_begin + n_q) of the
//            permutation hold the caller indices of table scene t
//   queue    one device counter; waves pop work items (chunk_q queries of one
//            scene) from it until the items run out
// Every launch is stream-ordered and allocates nothing once set_queries has
// sized the workspaces, and the counter and histogram are cleared by fill
// kernels, so a caller may capture the calls in a HIP graph.
#pragma once

namespace fcpw_hip {

constexpr int BATCH_CELL_BITS = 15; // Morton cells per scene: 32^3 (3D), 128^2 (2D)
constexpr int BATCH_MAX_SCENES = 1024;

// A pointer loaded from memory is generic to the compiler; passing it through
// the global address space lets the loads behind it compile to global rather
// than flat instructions.
template <typename T>
__device__ __forceinline__ T *as_global(T *p)
{
    return (T *)(__attribute__((address_space(1))) T *)(uintptr_t)p;
}

__device__ __forceinline__ DevScene global_scene(const DevScene &S)
{
    return DevScene{as_global(S.records), as_global(S.root), as_global(S.prims), as_global(S.prim_index),
                    as_global(S.flags), as_global(S.stats)};
}

struct BatchScene {
    DevScene S;
    const int32_t *prim_vids;
    float4 qlo, qinv;
    int64_t q_begin;    // first query of the scene in the caller's arrays
    int64_t p_begin;    // first position of the scene in the Morton permutation
    int64_t n_q;        // queries of the scene
    int64_t item_begin; // first work item of the scene (items of chunk_q queries)
    int64_t c64_begin;  // first 64-query block of the scene (ordering and adjoint kernels)
    int32_t n_prims;
    int32_t vert_offset; // first row of the scene in the concatenated vertex gradient
};

// last table entry whose begin (field selected by F) is <= key; empty scenes
// share their begin with the next scene, so the last match is the owner
template <typename F>
__device__ __forceinline__ int batch_find(const BatchScene *T, int n, int64_t key, F begin_of)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (begin_of(T[mid]) <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int DIM>
__global__ __launch_bounds__(64) void k_batch_morton_hist(const BatchScene *T, int n_scenes, const float *pts,
                                                          int stride, uint32_t *codes, uint32_t *ranks, uint32_t *hist)
{
    const int t = batch_find(T, n_scenes, (int64_t)blockIdx.x, [](const BatchScene &b) { return b.c64_begin; });
    const BatchScene &B = T[t];
    int64_t local = ((int64_t)blockIdx.x - B.c64_begin) * 64 + threadIdx.x;
    if (local >= B.n_q) return;
    int64_t i = B.q_begin + local;
    const float *q = pts + i * stride;
    uint32_t code;
    code = morton_cell<DIM>(q, B.qlo, B.qinv, DIM == 3 ? BATCH_CELL_BITS / 3 : BATCH_CELL_BITS / 2);
    code += (uint32_t)t << BATCH_CELL_BITS;
    codes[i] = code;
    ranks[i] = atomicAdd(&hist[code], 1u); // the query's place within its cell
}

// Persistent closest-point launch: each wave pops work items until none are
// left, so waves that drew short queries take more items and the batch ends
// when the total work does, not when the slowest scene's launch would.
template <int DIM, int K, int OUT>
__global__ __launch_bounds__(64) FCPW_EIGHT_WAVES void k_batch_closest_coop(const BatchScene *T, int n_scenes, CpIO io,
                                                           const uint32_t *perm, uint32_t *queue, int64_t n_items,
                                                           int chunk_q, int hold_min)
{
    while (true) {
        uint32_t item = 0;
        if (threadIdx.x == 0) item = atomicAdd(queue, 1u);
        item = __shfl(item, 0, 64);
        if ((int64_t)item >= n_items) break;
        const int t = batch_find(T, n_scenes, (int64_t)item, [](const BatchScene &b) { return b.item_begin; });
        const BatchScene &B = T[t];
        const DevScene S = global_scene(B.S);
        const int64_t first = ((int64_t)item - B.item_begin) * chunk_q;
        for (int off = 0; off < chunk_q; off += 64) {
            const int64_t local = first + off + threadIdx.x;
            const bool valid = local < B.n_q;
            if (__ballot(valid) == 0) break;
            const int64_t i = valid ? (perm ? (int64_t)perm[B.p_begin + local] : B.q_begin + local) : 0;
            float q[3] = {0.0f, 0.0f, 0.0f}, r2 = -1.0f;
            if (valid) cp_load<DIM, OUT>(io, i, q, r2);
            CpResult r;
            closest_point_query_coop<DIM, K>(S, Geo<DIM>::load_point(q), r2, valid, hold_min, r);
            if (valid) cp_store<DIM, OUT>(S, io, i, r);
        }
    }
}

// The adjoint of the batch: one wave per 64 queries of one scene, taken in the
// scene's Morton order when perm holds it (the permutation of the latest
// forward; queries of a wave then share primitives and vertices), else in the
// caller's order.  The wave's vertex terms are summed per vertex in LDS before
// one global atomic per touched vertex and coordinate.
template <int DIM>
__global__ __launch_bounds__(64) void k_batch_distance_backward(const BatchScene *T, int n_scenes,
                                                                const float *points, const int32_t *slots,
                                                                const float *grad_d, const uint32_t *perm,
                                                                float *grad_points, float *grad_vertices)
{
    __shared__ uint32_t hkey[BWD_HASH];
    __shared__ float hval[BWD_HASH][DIM];
    const int t = batch_find(T, n_scenes, (int64_t)blockIdx.x, [](const BatchScene &b) { return b.c64_begin; });
    const BatchScene &B = T[t];
    const int64_t local = ((int64_t)blockIdx.x - B.c64_begin) * 64 + threadIdx.x;
    AdjointTerms<DIM> A;
    A.slot = -1;
    if (local < B.n_q) {
        const int64_t i = perm ? (int64_t)perm[B.p_begin + local] : B.q_begin + local;
        adjoint_terms<DIM>(as_global(B.S.prims), as_global(B.prim_vids), B.n_prims, points, slots, grad_d, i, A);
        if (grad_points)
            for (int c = 0; c < DIM; c++) grad_points[i * DIM + c] = A.g[c];
    }
    if (grad_vertices) adjoint_hash_scatter<DIM>(A, hkey, hval, grad_vertices + (size_t)B.vert_offset * DIM);
}

} // namespace fcpw_hip

// ===========================================================================
// host side of the batch
// ===========================================================================
struct fcpw_hip_batch {
    int device = 0, dim = 0;
    std::string err;
    std::vector<fcpw_hip_ctx *> scenes; // caller order
    std::vector<uint64_t> generations;  // of the scenes when the table was built
    std::vector<int64_t> counts;        // queries per scene, caller order
    bool have_counts = false, dirty = true;
    bool perm_valid = false; // perm holds a Morton order of the current table's query ranges
    std::vector<BatchScene> host_table; // table order
    BatchScene *table = nullptr;
    int64_t n_total = 0, n_items = 0, n_c64 = 0, n_vertices_total = 0;
    uint32_t *queue = nullptr;
    void *codes = nullptr, *perm = nullptr, *hist = nullptr, *sums = nullptr;
    size_t codes_bytes = 0, perm_bytes = 0, hist_bytes = 0, sums_bytes = 0;
    int n_cus = 256;
    int occupancy[2][6][2] = {}; // resident waves per CU by [dim 2/3][stack 0/4/8/16/12/6][output], 0 = not queried
    // options
    int stack = 6, hold_min = 1, chunk = 64, sort = 1, waves_per_cu = 0;
};

static int batch_table(fcpw_hip_batch *b)
{
    const int n = (int)b->scenes.size();
    b->perm_valid = false; // the scenes' ranges in the permutation may move
    std::vector<int> order(n);
    std::vector<double> cost(n);
    for (int s = 0; s < n; s++) {
        order[s] = s;
        cost[s] = (double)b->counts[s] * std::log2((double)b->scenes[s]->n_prims + 2.0);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
    std::vector<int64_t> q_begin(n + 1, 0);
    std::vector<int32_t> v_begin(n + 1, 0);
    for (int s = 0; s < n; s++) {
        q_begin[s + 1] = q_begin[s] + b->counts[s];
        v_begin[s + 1] = v_begin[s] + b->scenes[s]->n_vertices;
    }
    b->host_table.assign(n, BatchScene{});
    int64_t p = 0, item = 0, c64 = 0;
    for (int t = 0; t < n; t++) {
        const int s = order[t];
        fcpw_hip_ctx *c = b->scenes[s];
        BatchScene &B = b->host_table[t];
        B.S = dev_scene(c);
        B.prim_vids = c->prim_vids;
        B.qlo = make_float4(c->qlo[0], c->qlo[1], c->qlo[2], 0.0f);
        B.qinv = make_float4(c->qinv[0], c->qinv[1], c->qinv[2], 0.0f);
        B.q_begin = q_begin[s];
        B.p_begin = p;
        B.n_q = b->counts[s];
        B.item_begin = item;
        B.c64_begin = c64;
        B.n_prims = c->n_prims;
        B.vert_offset = v_begin[s];
        p += B.n_q;
        item += (B.n_q + b->chunk - 1) / b->chunk;
        c64 += (B.n_q + 63) / 64;
    }
    b->n_total = q_begin[n];
    b->n_items = item;
    b->n_c64 = c64;
    b->n_vertices_total = v_begin[n];
    HIPCHECK(b, hipSetDevice(b->device));
    if (!b->table) HIPCHECK(b, hipMalloc((void **)&b->table, sizeof(BatchScene) * BATCH_MAX_SCENES));
    HIPCHECK(b, hipMemcpy(b->table, b->host_table.data(), sizeof(BatchScene) * n, hipMemcpyHostToDevice));
    b->generations.resize(n);
    for (int s = 0; s < n; s++) b->generations[s] = b->scenes[s]->generation;
    b->dirty = false;
    return FCPW_HIP_OK;
}

static int batch_grow(fcpw_hip_batch *b, void **buf, size_t *cap, size_t bytes)
{
    if (*cap >= bytes) return FCPW_HIP_OK;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    HIPCHECK(b, hipMalloc(buf, bytes));
    *cap = bytes;
    return FCPW_HIP_OK;
}

// table up to date with the member scenes (a re-upload moves their buffers).  A
// member may have been re-transferred since fcpw_hip_batch_create, so the
// checks made there are repeated: an instanced upload (its records index
// instances, not primitives) or a change of dimension or device is rejected
// here, on the host, before anything is launched.
static int batch_ready(fcpw_hip_batch *b, const char *fn)
{
    if (!b) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    if (!b->have_counts) FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "%s: call fcpw_hip_batch_set_queries first", fn);
    for (size_t s = 0; s < b->scenes.size(); s++) {
        const fcpw_hip_ctx *m = b->scenes[s];
        if (!m->uploaded || m->instanced)
            FAIL(b, FCPW_HIP_ERR_NOT_UPLOADED, "%s: member scene %zu has no uploaded single-object BVH", fn, s);
        if (m->dim != b->dim || m->device != b->device)
            FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "%s: member scene %zu changed dimension or device", fn, s);
        if (m->generation != b->generations[s]) b->dirty = true;
    }
    if (b->dirty) return batch_table(b);
    return FCPW_HIP_OK;
}

static int batch_order(fcpw_hip_batch *b, const float *pts, int stride, hipStream_t s, const uint32_t **perm_out)
{
    *perm_out = nullptr;
    if (!b->sort || b->n_total < 4096) return FCPW_HIP_OK;
    const int n = (int)b->scenes.size();
    const int64_t cells = (int64_t)n << BATCH_CELL_BITS;
    const int64_t segs = (cells + SCAN_SEG - 1) / SCAN_SEG;
    uint32_t *codes = (uint32_t *)b->codes, *perm = (uint32_t *)b->perm, *hist = (uint32_t *)b->hist;
    uint32_t *ranks = codes + b->n_total;
    uint32_t *sums = (uint32_t *)b->sums;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, hist, 0u, cells);
    if (b->dim == 3)
        hipLaunchKernelGGL(k_batch_morton_hist<3>, dim3((unsigned)b->n_c64), dim3(64), 0, s, b->table, n, pts, stride,
                           codes, ranks, hist);
    else
        hipLaunchKernelGGL(k_batch_morton_hist<2>, dim3((unsigned)b->n_c64), dim3(64), 0, s, b->table, n, pts, stride,
                           codes, ranks, hist);
    hipLaunchKernelGGL(k_scan_segment_sums, dim3((unsigned)segs), dim3(SCAN_THREADS), 0, s, hist, cells, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, (int)segs);
    hipLaunchKernelGGL(k_scan_segment_apply, dim3((unsigned)segs), dim3(SCAN_THREADS), 0, s, hist, cells, sums);
    hipLaunchKernelGGL(k_morton_place, dim3((unsigned)((b->n_total + 255) / 256)), dim3(256), 0, s, codes, ranks,
                       b->n_total, hist, perm);
    HIPCHECK(b, hipGetLastError());
    *perm_out = perm;
    b->perm_valid = true;
    return FCPW_HIP_OK;
}

template <int DIM, int K, int OUT>
static int batch_launch_k(fcpw_hip_batch *b, const CpIO &io, const uint32_t *perm, hipStream_t s)
{
    auto kern = k_batch_closest_coop<DIM, K, OUT>;
    int per_cu = b->waves_per_cu;
    if (per_cu <= 0) {
        int &occ = b->occupancy[DIM - 2][K == 0 ? 0 : K == 4 ? 1 : K == 8 ? 2 : K == 16 ? 3 : K == 12 ? 4 : 5][OUT];
        if (occ == 0) {
            int blocks = 0;
            HIPCHECK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kern, 64, 0));
            occ = std::max(1, blocks);
        }
        per_cu = occ;
    }
    int64_t waves = std::min<int64_t>((int64_t)b->n_cus * per_cu, std::max<int64_t>(1, b->n_items));
    hipLaunchKernelGGL(kern, dim3((unsigned)waves), dim3(64), 0, s, b->table, (int)b->scenes.size(), io, perm,
                       b->queue, b->n_items, b->chunk, b->hold_min);
    HIPCHECK(b, hipGetLastError());
    return FCPW_HIP_OK;
}

template <int OUT>
static int batch_launch(fcpw_hip_batch *b, CpIO io, hipStream_t s)
{
    int st = batch_order(b, io.in, OUT == CP_OUT_INTERACTION ? 4 : b->dim, s, &io.order);
    if (st) return st;
    const uint32_t *perm = io.order;
    io.order = nullptr; // the batch kernel maps positions itself
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, b->queue, 0u, (int64_t)1);
    const int k = b->stack;
    if (b->dim == 3) {
        if (k == 0) return batch_launch_k<3, 0, OUT>(b, io, perm, s);
        if (k == 4) return batch_launch_k<3, 4, OUT>(b, io, perm, s);
        if (k == 6) return batch_launch_k<3, 6, OUT>(b, io, perm, s);
        if (k == 8) return batch_launch_k<3, 8, OUT>(b, io, perm, s);
        if (k == 12) return batch_launch_k<3, 12, OUT>(b, io, perm, s);
        return batch_launch_k<3, 16, OUT>(b, io, perm, s);
    }
    if (k == 0) return batch_launch_k<2, 0, OUT>(b, io, perm, s);
    if (k == 4) return batch_launch_k<2, 4, OUT>(b, io, perm, s);
    if (k == 6) return batch_launch_k<2, 6, OUT>(b, io, perm, s);
    if (k == 8) return batch_launch_k<2, 8, OUT>(b, io, perm, s);
    if (k == 12) return batch_launch_k<2, 12, OUT>(b, io, perm, s);
    return batch_launch_k<2, 16, OUT>(b, io, perm, s);
}

extern "C" {

int fcpw_hip_batch_create(fcpw_hip_ctx *const *scenes, int n_scenes, fcpw_hip_batch **out)
{
    if (!out) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!scenes || n_scenes <= 0 || n_scenes > BATCH_MAX_SCENES) {
        g_create_error = "batch_create: need 1.." + std::to_string(BATCH_MAX_SCENES) + " scenes";
        return FCPW_HIP_ERR_INVALID_ARGUMENT;
    }
    for (int s = 0; s < n_scenes; s++) {
        if (!scenes[s] || !scenes[s]->uploaded || scenes[s]->instanced) {
            g_create_error = "batch_create: scene " + std::to_string(s) + " has no uploaded single-object BVH";
            return FCPW_HIP_ERR_NOT_UPLOADED;
        }
        if (scenes[s]->dim != scenes[0]->dim || scenes[s]->device != scenes[0]->device) {
            g_create_error = "batch_create: scenes must share dimension and device";
            return FCPW_HIP_ERR_INVALID_ARGUMENT;
        }
    }
    fcpw_hip_batch *b = new fcpw_hip_batch();
    b->scenes.assign(scenes, scenes + n_scenes);
    b->counts.assign(n_scenes, 0);
    b->device = scenes[0]->device;
    b->dim = scenes[0]->dim;
    b->n_cus = scenes[0]->n_cus;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipMalloc((void **)&b->queue, 256);
    if (e != hipSuccess) {
        g_create_error = hipGetErrorString(e);
        delete b;
        return FCPW_HIP_ERR_HIP;
    }
    *out = b;
    return FCPW_HIP_OK;
}

int fcpw_hip_batch_destroy(fcpw_hip_batch *b)
{
    if (!b) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(b->device);
    void *bufs[] = {b->table, b->queue, b->codes, b->perm, b->hist, b->sums};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    delete b;
    return FCPW_HIP_OK;
}

const char *fcpw_hip_batch_last_error(const fcpw_hip_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int fcpw_hip_batch_set_queries(fcpw_hip_batch *b, const int64_t *counts)
{
    if (!b || !counts) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    int64_t total = 0;
    for (size_t s = 0; s < b->scenes.size(); s++) {
        if (counts[s] < 0) FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "batch_set_queries: negative count for scene %zu", s);
        total += counts[s];
    }
    if (total > (int64_t)UINT32_MAX) FAIL(b, FCPW_HIP_ERR_UNSUPPORTED, "batch_set_queries: more than 2^32 queries");
    b->counts.assign(counts, counts + b->scenes.size());
    b->have_counts = true;
    int st = batch_table(b);
    if (st) return st;
    const size_t cells = (size_t)b->scenes.size() << BATCH_CELL_BITS;
    const size_t segs = (cells + SCAN_SEG - 1) / SCAN_SEG;
    st = batch_grow(b, &b->codes, &b->codes_bytes, std::max<size_t>(8, (size_t)total * 8)); // codes, then ranks
    if (!st) st = batch_grow(b, &b->perm, &b->perm_bytes, std::max<size_t>(4, (size_t)total * 4));
    if (!st) st = batch_grow(b, &b->hist, &b->hist_bytes, cells * 4);
    if (!st) st = batch_grow(b, &b->sums, &b->sums_bytes, segs * 4);
    return st;
}

int fcpw_hip_batch_closest_points_device(fcpw_hip_batch *b, const fcpw_hip_bounding_sphere *spheres,
                                         int record_normals, fcpw_hip_interaction *out, void *stream)
{
    int st = batch_ready(b, "batch_closest_points_device");
    if (st) return st;
    if (b->n_total == 0) return FCPW_HIP_OK;
    if (!spheres || !out) FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "batch_closest_points_device: null pointer");
    if (((uintptr_t)spheres & 15) || ((uintptr_t)out & 7))
        FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "batch_closest_points_device: spheres must be 16-byte and out 8-byte aligned");
    CpIO io = {};
    io.in = (const float *)spheres;
    io.n = b->n_total;
    io.out = (float2 *)out;
    io.record_normals = record_normals;
    return batch_launch<CP_OUT_INTERACTION>(b, io, (hipStream_t)stream);
}

int fcpw_hip_batch_closest_points_compact_device(fcpw_hip_batch *b, const float *points, const float *r2, float *d,
                                                 float *p, int32_t *prim, int32_t *slot, void *stream)
{
    int st = batch_ready(b, "batch_closest_points_compact_device");
    if (st) return st;
    if (b->n_total == 0) return FCPW_HIP_OK;
    if (!points) FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "batch_closest_points_compact_device: null points");
    CpIO io = {};
    io.in = points;
    io.r2 = r2;
    io.n = b->n_total;
    io.d = d;
    io.p = p;
    io.prim = prim;
    io.slot = slot;
    return batch_launch<CP_OUT_COMPACT>(b, io, (hipStream_t)stream);
}

int fcpw_hip_batch_distance_backward_device(fcpw_hip_batch *b, const float *points, const int32_t *slot,
                                            const float *grad_d, float *grad_points, float *grad_vertices,
                                            void *stream)
{
    int st = batch_ready(b, "batch_distance_backward_device");
    if (st) return st;
    if (b->n_total == 0) return FCPW_HIP_OK;
    if (!points || !slot || !grad_d) FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "batch_distance_backward_device: null input");
    hipStream_t s = (hipStream_t)stream;
    const int n = (int)b->scenes.size();
    // any permutation of a scene's range serves; the latest forward's Morton order is the fast one
    const uint32_t *perm = b->perm_valid ? (const uint32_t *)b->perm : nullptr;
    if (b->dim == 3)
        hipLaunchKernelGGL(k_batch_distance_backward<3>, dim3((unsigned)b->n_c64), dim3(64), 0, s, b->table, n,
                           points, slot, grad_d, perm, grad_points, grad_vertices);
    else
        hipLaunchKernelGGL(k_batch_distance_backward<2>, dim3((unsigned)b->n_c64), dim3(64), 0, s, b->table, n,
                           points, slot, grad_d, perm, grad_points, grad_vertices);
    HIPCHECK(b, hipGetLastError());
    return FCPW_HIP_OK;
}

int fcpw_hip_batch_set_option(fcpw_hip_batch *b, const char *key, int value)
{
    if (!b || !key) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    std::string k(key);
    if (k == "stack") {
        if (value != 0 && value != 4 && value != 6 && value != 8 && value != 12 && value != 16)
            FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "stack must be 0, 4, 6, 8, 12 or 16");
        b->stack = value;
    } else if (k == "hold_min") {
        if (value < 1 || value > 64) FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "hold_min must be in [1, 64]");
        b->hold_min = value;
    } else if (k == "chunk") {
        if (value != 64 && value != 128 && value != 256 && value != 512)
            FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "chunk must be 64, 128, 256 or 512");
        if (value != b->chunk) b->dirty = true;
        b->chunk = value;
    } else if (k == "sort") {
        if (value != 0 && value != 1) FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "sort must be 0 or 1");
        b->sort = value;
    } else if (k == "waves_per_cu") {
        if (value < 0 || value > 32) FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "waves_per_cu must be in [0, 32]");
        b->waves_per_cu = value;
    } else {
        FAIL(b, FCPW_HIP_ERR_INVALID_ARGUMENT, "unknown batch option '%s'", key);
    }
    return FCPW_HIP_OK;
}

int fcpw_hip_batch_get_option(const fcpw_hip_batch *b, const char *key, int *value)
{
    if (!b || !key || !value) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    std::string k(key);
    if (k == "stack") *value = b->stack;
    else if (k == "hold_min") *value = b->hold_min;
    else if (k == "chunk") *value = b->chunk;
    else if (k == "sort") *value = b->sort;
    else if (k == "waves_per_cu") *value = b->waves_per_cu;
    else return FCPW_HIP_ERR_INVALID_ARGUMENT;
    return FCPW_HIP_OK;
}

} // extern "C"
