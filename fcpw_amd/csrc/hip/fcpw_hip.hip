// fcpw_hip.hip -- MI355X (gfx950) query backend for FCPW behind the C ABI in
// include/fcpw_hip.h: the single-scene entry points (upload, refit, queries,
// adjoint, options), from batch.hpp the batched multi-scene launch and from
// instances.hpp scenes of several objects and transformed instances.
// Device code lives in device.hpp, the per-scene context in context.hpp.
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>

#include "device.hpp"
#include "silhouette.hpp"
#include "context.hpp"
#include "gpu_build.hpp"

// the allocation-free entry point's scratch lookup (defined with the array below)
static void *noalloc_scratch_ptr(fcpw_hip_ctx *c);
// the record layout of an uploaded tree (defined before upload_tree)
static int relayout_records(fcpw_hip_ctx *c, int layout);

// scenes of several objects (instances.hpp)
static int instances_refit(fcpw_hip_ctx *c, hipStream_t s);

// in-place exclusive scan of n uint32 values (segments of SCAN_SEG, at most 16384 of them)
static int exclusive_scan(fcpw_hip_ctx *c, uint32_t *v, int64_t n, uint32_t *sums, hipStream_t s)
{
    const int64_t segs = (n + SCAN_SEG - 1) / SCAN_SEG;
    if (segs > 1024 * 16) FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "scan of %lld values exceeds the scan capacity", (long long)n);
    hipLaunchKernelGGL(k_scan_segment_sums, dim3((unsigned)segs), dim3(SCAN_THREADS), 0, s, v, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, (int)segs);
    hipLaunchKernelGGL(k_scan_segment_apply, dim3((unsigned)segs), dim3(SCAN_THREADS), 0, s, v, n, sums);
    HIPCHECK(c, hipGetLastError());
    return FCPW_HIP_OK;
}

// device scratch of the build: numbered slots kept in the context and grown on
// demand, so a rebuild of a scene of the same size allocates nothing
struct BuildScratch {
    fcpw_hip_ctx *c;
    int next = 0;
    template <typename T>
    hipError_t alloc(T **p, size_t count)
    {
        *p = nullptr;
        if (next >= fcpw_hip_ctx::BUILD_SLOTS) return hipErrorOutOfMemory;
        const int k = next++;
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if (c->build_bytes[k] < bytes) {
            if (c->build_buf[k]) (void)hipFree(c->build_buf[k]);
            c->build_buf[k] = nullptr;
            c->build_bytes[k] = 0;
            hipError_t e = hipMalloc(&c->build_buf[k], bytes);
            if (e != hipSuccess) return e;
            c->build_bytes[k] = bytes;
        }
        *p = (T *)c->build_buf[k];
        return hipSuccess;
    }
};

static unsigned blocks_for(int64_t n, int t) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

template <int DIM>
static int gpu_build(fcpw_hip_ctx *c, const float *positions, int n_vertices, const int32_t *indices, int n_prims,
                     int heuristic, float *out_pmin, float *out_pmax, int32_t *out_nrefs, int32_t *out_offset,
                     int32_t *out_prim_order, int *out_n_nodes)
{
    using namespace gbuild;
    const int n = n_prims;
    const int cap_nodes = 2 * n - 1;
    hipStream_t st = 0;
    BuildScratch T{c};
    float *V = nullptr;
    int32_t *idx = nullptr;
    HIPCHECK(c, hipMalloc((void **)&V, (size_t)n_vertices * DIM * sizeof(float)));
    HIPCHECK(c, T.alloc(&idx, (size_t)n * DIM));
    HIPCHECK(c, hipMemcpy(V, positions, (size_t)n_vertices * DIM * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(idx, indices, (size_t)n * DIM * sizeof(int32_t), hipMemcpyHostToDevice));
    c->V = V; // owned by the context from here on

    Build B = {};
    B.n = n;
    B.heuristic = heuristic;
    float4 *plo, *phi, *pcen;
    HIPCHECK(c, T.alloc(&plo, n));
    HIPCHECK(c, T.alloc(&phi, n));
    HIPCHECK(c, T.alloc(&pcen, n));
    B.pbox_lo = plo; B.pbox_hi = phi; B.pcen = pcen;
    HIPCHECK(c, T.alloc(&B.ref, n));
    HIPCHECK(c, T.alloc(&B.ref2, n));
    HIPCHECK(c, T.alloc(&B.elem_node, n));
    HIPCHECK(c, T.alloc(&B.scan_r, (size_t)n + 1));
    HIPCHECK(c, T.alloc(&B.scan_k, (size_t)n + 1));
    for (int k = 0; k < 2; k++) {
        HIPCHECK(c, T.alloc(&B.st_val[k], n));
        HIPCHECK(c, T.alloc(&B.st_nxt[k], n));
    }
    int32_t **nd_arrays[] = {&B.nd_start, &B.nd_n, &B.nd_depth, &B.nd_left, &B.nd_dim, &B.nd_mid, &B.nd_rot,
                             &B.nd_right, &B.nd_flags, &B.nd_size, &B.nd_leaves, &B.nd_idx, &B.nd_lb};
    for (int32_t **a : nd_arrays) HIPCHECK(c, T.alloc(a, cap_nodes));
    HIPCHECK(c, T.alloc(&B.nd_box, (size_t)cap_nodes * 6));
    HIPCHECK(c, T.alloc(&B.nd_cbox, (size_t)cap_nodes * 6));
    HIPCHECK(c, T.alloc(&B.nd_coord, cap_nodes));
    HIPCHECK(c, T.alloc(&B.level_info, 4));
    uint32_t *sums = nullptr;
    HIPCHECK(c, T.alloc(&sums, 16384));
    // a level holds at most two children per internal node of the previous level,
    // and an internal node has more than LEAF references
    const int max_level = std::max(1, 2 * (n / (LEAF + 1)) + 2);
    HIPCHECK(c, T.alloc(&B.bins, (size_t)max_level * BIN_NODE));
    HIPCHECK(c, T.alloc(&B.internal_scan, (size_t)max_level + 1));

    const int t = 256;
    hipLaunchKernelGGL(k_prim_data<DIM>, dim3(blocks_for(n, t)), dim3(t), 0, st, V, idx, n, plo, phi, pcen);
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, t)), dim3(t), 0, st, B.ref, n);
    const int32_t root_init[3] = {0, n, 0};
    HIPCHECK(c, hipMemcpy(B.nd_start, &root_init[0], sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(B.nd_n, &root_init[1], sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(B.nd_depth, &root_init[2], sizeof(int32_t), hipMemcpyHostToDevice));

    std::vector<std::pair<int, int>> levels; // (first node, count)
    int g0 = 0, m = 1, total = 1;
    while (m > 0) {
        if (m > max_level) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "build_bvh: level size %d exceeds %d", m, max_level);
        hipLaunchKernelGGL(k_level_init, dim3(blocks_for(m, t)), dim3(t), 0, st, B, g0, m);
        hipLaunchKernelGGL(k_elem_node, dim3(blocks_for(n, t)), dim3(t), 0, st, B, g0, m);
        hipLaunchKernelGGL(k_small<DIM>, dim3(blocks_for(m, 64)), dim3(64), 0, st, B, g0, m);
        hipLaunchKernelGGL(k_bounds<DIM>, dim3(blocks_for(n, ACC_THREADS)), dim3(ACC_THREADS), 0, st, B);
        hipLaunchKernelGGL(k_bins<DIM>, dim3(blocks_for(n, ACC_THREADS)), dim3(ACC_THREADS), 0, st, B, g0);
        hipLaunchKernelGGL(k_split<DIM>, dim3(blocks_for(m, 64)), dim3(64), 0, st, B, g0, m);
        hipLaunchKernelGGL(k_flag_right, dim3(blocks_for((int64_t)n + 1, t)), dim3(t), 0, st, B);
        int e = exclusive_scan(c, B.scan_r, (int64_t)n + 1, sums, st);
        if (e) return e;
        hipLaunchKernelGGL(k_flag_rotate, dim3(blocks_for((int64_t)n + 1, t)), dim3(t), 0, st, B);
        e = exclusive_scan(c, B.scan_k, (int64_t)n + 1, sums, st);
        if (e) return e;
        HIPCHECK(c, hipMemsetAsync(B.level_info, 0, 4 * sizeof(int), st));
        hipLaunchKernelGGL(k_node_part, dim3(blocks_for((int64_t)m + 1, t)), dim3(t), 0, st, B, g0, m);
        e = exclusive_scan(c, B.internal_scan, (int64_t)m + 1, sums, st);
        if (e) return e;
        uint32_t n_internal = 0;
        int max_stream = 0;
        HIPCHECK(c, hipMemcpy(&n_internal, B.internal_scan + m, sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIPCHECK(c, hipMemcpy(&max_stream, B.level_info, sizeof(int), hipMemcpyDeviceToHost));
        hipLaunchKernelGGL(k_stream, dim3(blocks_for(n, t)), dim3(t), 0, st, B);
        int in = 0;
        if (max_stream > 0) {
            int rounds = 1;
            while ((1 << rounds) <= max_stream) rounds++;
            for (int r = 0; r < rounds; r++) {
                hipLaunchKernelGGL(k_jump, dim3(blocks_for(n, t)), dim3(t), 0, st, B, in);
                in ^= 1;
            }
            hipLaunchKernelGGL(k_place_right, dim3(blocks_for(n, t)), dim3(t), 0, st, B, in);
        }
        const int next = g0 + m;
        if ((int64_t)next + 2 * (int64_t)n_internal > cap_nodes)
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "build_bvh: node capacity exceeded");
        hipLaunchKernelGGL(k_children, dim3(blocks_for(m, t)), dim3(t), 0, st, B, g0, m, next);
        HIPCHECK(c, hipGetLastError());
        std::swap(B.ref, B.ref2);
        levels.push_back({g0, m});
        g0 = next;
        m = 2 * (int)n_internal;
        total += m;
    }
    for (int l = (int)levels.size() - 1; l >= 0; l--)
        hipLaunchKernelGGL(k_sizes, dim3(blocks_for(levels[l].second, t)), dim3(t), 0, st, B, levels[l].first,
                           levels[l].second);
    HIPCHECK(c, hipMemsetAsync(B.nd_idx, 0, sizeof(int32_t), st));
    HIPCHECK(c, hipMemsetAsync(B.nd_lb, 0, sizeof(int32_t), st));
    for (size_t l = 0; l < levels.size(); l++)
        hipLaunchKernelGGL(k_dfs, dim3(blocks_for(levels[l].second, t)), dim3(t), 0, st, B, levels[l].first,
                           levels[l].second);
    HIPCHECK(c, hipGetLastError());

    const int n_leaf = (total + 1) / 2, n_rec = n_leaf - 1;
    const int REC = DIM == 3 ? 4 : 3, PR = DIM == 3 ? 3 : 1;
    HIPCHECK(c, hipMalloc((void **)&c->records, (size_t)std::max(n_rec, 1) * REC * sizeof(float4)));
    HIPCHECK(c, hipMemset(c->records, 0, (size_t)std::max(n_rec, 1) * REC * sizeof(float4)));
    HIPCHECK(c, hipMalloc((void **)&c->root, 2 * sizeof(float4)));
    HIPCHECK(c, hipMalloc((void **)&c->prims, (size_t)n * PR * sizeof(float4)));
    HIPCHECK(c, hipMalloc((void **)&c->prim_index, (size_t)n * sizeof(int32_t)));
    HIPCHECK(c, hipMalloc((void **)&c->prim_vids, (size_t)n * DIM * sizeof(int32_t)));
    HIPCHECK(c, hipMalloc((void **)&c->rec_parent, (size_t)std::max(n_rec, 1) * sizeof(int32_t)));
    HIPCHECK(c, hipMalloc((void **)&c->leaf_refs, (size_t)n_leaf * sizeof(int32_t)));
    HIPCHECK(c, hipMalloc((void **)&c->leaf_links, (size_t)n_leaf * sizeof(uint32_t)));
    HIPCHECK(c, hipMalloc((void **)&c->counters, (size_t)std::max(n_rec, 1) * sizeof(unsigned int)));
    HIPCHECK(c, hipMemset(c->counters, 0, (size_t)std::max(n_rec, 1) * sizeof(unsigned int)));
    Emit E = {};
    E.records = c->records;
    E.root = c->root;
    E.rec_parent = c->rec_parent;
    E.leaf_refs = c->leaf_refs;
    E.leaf_links = c->leaf_links;
    HIPCHECK(c, T.alloc(&E.pmin, (size_t)total * 3));
    HIPCHECK(c, T.alloc(&E.pmax, (size_t)total * 3));
    HIPCHECK(c, T.alloc(&E.nrefs, total));
    HIPCHECK(c, T.alloc(&E.offset, total));
    HIPCHECK(c, T.alloc(&E.status, 2));
    HIPCHECK(c, hipMemset(E.status, 0, 2 * sizeof(int)));
    hipLaunchKernelGGL(k_emit<DIM>, dim3(blocks_for(total, t)), dim3(t), 0, st, B, E, total);
    hipLaunchKernelGGL(k_emit_prims<DIM>, dim3(blocks_for(n, t)), dim3(t), 0, st, B.ref, V, idx, n, c->prims,
                       c->prim_index, c->prim_vids);
    HIPCHECK(c, hipGetLastError());
    int status[2] = {0, 0};
    HIPCHECK(c, hipMemcpy(status, E.status, sizeof(status), hipMemcpyDeviceToHost));
    if (status[0] & 1)
        FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "build_bvh: a leaf at the depth cap holds more than %d primitives",
             MAX_LEAF_PRIMS);
    if (status[1] > STACK_TOTAL - 1)
        FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "build_bvh: tree depth %d exceeds %d", status[1], STACK_TOTAL - 1);
    float root_box[6];
    {
        float4 r[2];
        HIPCHECK(c, hipMemcpy(r, c->root, sizeof(r), hipMemcpyDeviceToHost));
        if (DIM == 3) {
            const float v[6] = {r[0].x, r[0].y, r[0].z, r[1].x, r[1].y, r[1].z};
            memcpy(root_box, v, sizeof(v));
        } else {
            const float v[6] = {r[0].x, r[0].y, 0.0f, r[0].z, r[0].w, 0.0f};
            memcpy(root_box, v, sizeof(v));
        }
    }
    if (out_pmin) HIPCHECK(c, hipMemcpy(out_pmin, E.pmin, (size_t)total * 3 * sizeof(float), hipMemcpyDeviceToHost));
    if (out_pmax) HIPCHECK(c, hipMemcpy(out_pmax, E.pmax, (size_t)total * 3 * sizeof(float), hipMemcpyDeviceToHost));
    if (out_nrefs) HIPCHECK(c, hipMemcpy(out_nrefs, E.nrefs, (size_t)total * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (out_offset)
        HIPCHECK(c, hipMemcpy(out_offset, E.offset, (size_t)total * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (out_prim_order)
        HIPCHECK(c, hipMemcpy(out_prim_order, c->prim_index, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (out_n_nodes) *out_n_nodes = total;
    // with the flat tree on the host, map its nodes to record slots as fcpw_hip_upload_bvh
    // does (internal nodes numbered in DFS order), so fcpw_hip_upload_silhouettes can follow
    std::vector<int32_t> node_slot, node_leaf_off;
    if (out_nrefs && out_offset) {
        node_slot.assign((size_t)total, 2 * n_rec);
        node_leaf_off.assign((size_t)total, -1);
        int rank = 0;
        for (int i = 0; i < total; i++) {
            if (out_nrefs[i] > 0) {
                node_leaf_off[i] = out_offset[i];
                continue;
            }
            node_slot[i + 1] = rank * 2;
            node_slot[i + out_offset[i]] = rank * 2 + 1;
            rank++;
        }
    }
    HIPCHECK(c, hipDeviceSynchronize());
    for (int d = 0; d < 3; d++) {
        float lo = d < DIM ? root_box[d] : 0.0f, hi = d < DIM ? root_box[3 + d] : 0.0f;
        float ext = std::max(hi - lo, 1e-6f);
        c->qlo[d] = 0.5f * (lo + hi) - 0.75f * ext;
        c->qinv[d] = 1.0f / (1.5f * ext);
    }
    c->dim = DIM;
    c->n_records = n_rec;
    c->n_leaves = n_leaf;
    c->n_prims = n;
    c->n_vertices = n_vertices;
    c->max_depth = status[1];
    c->node_slot = std::move(node_slot);
    c->node_leaf_off = std::move(node_leaf_off);
    c->uploaded = true;
    c->generation++;
    c->positions_version++;
    return c->layout ? relayout_records(c, c->layout) : FCPW_HIP_OK;
}

extern "C" {

int fcpw_hip_create(int device, fcpw_hip_ctx **out)
{
    if (!out) {
        g_create_error = "out is NULL";
        return FCPW_HIP_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        g_create_error = std::string("no HIP device (") + hipGetErrorString(e) + ")";
        return FCPW_HIP_ERR_HIP;
    }
    if (device < 0 || device >= count) {
        g_create_error = "device index out of range";
        return FCPW_HIP_ERR_INVALID_ARGUMENT;
    }
    fcpw_hip_ctx *c = new fcpw_hip_ctx();
    c->device = device;
    e = hipSetDevice(device);
    if (e == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cus = prop.multiProcessorCount;
        e = hipMalloc((void **)&c->flags, sizeof(int));
    }
    if (e == hipSuccess) e = hipMemset(c->flags, 0, sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&c->stats, 8 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->stats, 0, 8 * sizeof(unsigned long long));
    // the allocation-free entry point's scratch array, looked up here once per
    // device; a failure only disables that entry point, whose calls report it
    if (e == hipSuccess) {
        c->scratch = noalloc_scratch_ptr(c);
        if (!c->scratch) {
            c->scratch_err = c->err;
            c->err.clear();
        }
    }
    if (e != hipSuccess) {
        g_create_error = hipGetErrorString(e);
        if (c->flags) (void)hipFree(c->flags);
        if (c->stats) (void)hipFree(c->stats);
        delete c;
        return FCPW_HIP_ERR_HIP;
    }
    *out = c;
    return FCPW_HIP_OK;
}

int fcpw_hip_destroy(fcpw_hip_ctx *c)
{
    if (!c) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(c->device);
    free_scene(c);
    if (c->flags) (void)hipFree(c->flags);
    if (c->stats) (void)hipFree(c->stats);
    void *ws[] = {c->stage_in, c->stage_out, c->ws_codes, c->ws_perm, c->ws_hist, c->ws_sums};
    for (void *b : ws)
        if (b) (void)hipFree(b);
    if (c->pin_in) (void)hipHostFree(c->pin_in);
    if (c->pin_out) (void)hipHostFree(c->pin_out);
    for (hipEvent_t e : c->pin_ev) (void)hipEventDestroy(e);
    for (void *b : c->build_buf)
        if (b) (void)hipFree(b);
    for (hipEvent_t e : c->ev_begin) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->ev_end) (void)hipEventDestroy(e);
    delete c;
    return FCPW_HIP_OK;
}

const char *fcpw_hip_last_error(const fcpw_hip_ctx *c)
{
    return c ? c->err.c_str() : g_create_error.c_str();
}

static uint32_t make_link(bool leaf, int32_t offset, int32_t nrefs, const std::vector<int32_t> &rec_of, int node)
{
    if (!leaf) return (uint32_t)rec_of[node];
    return LEAF_BIT | ((uint32_t)(nrefs - 1) << CNT_SHIFT) | (uint32_t)offset;
}

static int check_ready(fcpw_hip_ctx *c, const char *fn)
{
    if (!c) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    if (!c->uploaded) FAIL(c, FCPW_HIP_ERR_NOT_UPLOADED, "%s: no BVH has been uploaded", fn);
    return FCPW_HIP_OK;
}

// entry points served for one object's tree only
static int check_single(fcpw_hip_ctx *c, const char *fn)
{
    int st = check_ready(c, fn);
    if (st) return st;
    if (c->instanced)
        FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "%s: not supported on a scene of several objects or instances", fn);
    return FCPW_HIP_OK;
}

} // extern "C"

// The child-pair records of a flat DFS tree (the layout fcpw_hip_upload_bvh
// documents) with its refit schedule; validation errors name fn.
struct TreeRecords {
    std::vector<float4> rec;
    float4 root[2];
    std::vector<int32_t> parent, leaf_refs;
    std::vector<uint32_t> leaf_links;
    std::vector<int32_t> node_slot, node_leaf_off;
    int n_rec = 0, n_leaf = 0, max_depth = 0;
};

static int tree_records(fcpw_hip_ctx *c, const char *fn, int dim, int n_nodes, const float *pmin, const float *pmax,
                        const int32_t *nrefs, const int32_t *offset, int n_prims, TreeRecords &T)
{
    // validate the reference layout and number internal nodes in DFS order
    std::vector<int32_t> rec_of(n_nodes, -1);
    int n_rec = 0, n_leaf = 0;
    for (int i = 0; i < n_nodes; i++) {
        if (nrefs[i] > 0) {
            if (nrefs[i] > MAX_LEAF_PRIMS)
                FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "%s: leaf %d holds %d primitives (at most %d)", fn, i, nrefs[i],
                     MAX_LEAF_PRIMS);
            if (offset[i] < 0 || (int64_t)offset[i] + nrefs[i] > n_prims)
                FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "%s: leaf %d primitive range out of bounds", fn, i);
            n_leaf++;
        } else if (nrefs[i] == 0) {
            if (offset[i] <= 1 || (int64_t)i + offset[i] >= n_nodes)
                FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "%s: internal node %d has a bad child offset %d", fn, i,
                     offset[i]);
            rec_of[i] = n_rec++;
        } else {
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "%s: node %d has a negative primitive count", fn, i);
        }
    }
    if (n_leaf != n_rec + 1) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "%s: not a binary tree", fn);

    const int REC = dim == 3 ? 4 : 3;
    T.rec.assign((size_t)std::max(n_rec, 1) * REC, make_float4(0, 0, 0, 0));
    T.parent.assign((size_t)std::max(n_rec, 1), -1);
    T.leaf_refs.clear();
    T.leaf_links.clear();
    T.leaf_refs.reserve(n_leaf);
    T.leaf_links.reserve(n_leaf);
    std::vector<int32_t> depth(n_nodes, 0);
    int max_depth = 0;
    auto &rec = T.rec;

    auto put_slot = [&](int r, int slot, int node, uint32_t link) {
        const float *lo = pmin + (size_t)node * 3, *hi = pmax + (size_t)node * 3;
        if (dim == 3) {
            rec[(size_t)r * 4 + slot * 2] = make_float4(lo[0], lo[1], lo[2], bits_as_float(link));
            rec[(size_t)r * 4 + slot * 2 + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        } else {
            rec[(size_t)r * 3 + slot] = make_float4(lo[0], lo[1], hi[0], hi[1]);
            if (slot == 0) rec[(size_t)r * 3 + 2].x = bits_as_float(link);
            else rec[(size_t)r * 3 + 2].y = bits_as_float(link);
        }
    };
    T.node_slot.assign((size_t)n_nodes, 2 * n_rec);
    T.node_leaf_off.assign((size_t)n_nodes, -1);
    for (int i = 0; i < n_nodes; i++)
        if (nrefs[i] > 0) T.node_leaf_off[i] = offset[i];
    for (int i = 0; i < n_nodes; i++) {
        if (nrefs[i] > 0) continue;
        int r = rec_of[i];
        int kids[2] = {i + 1, i + offset[i]};
        for (int sd = 0; sd < 2; sd++) {
            int k = kids[sd];
            T.node_slot[k] = r * 2 + sd;
            depth[k] = depth[i] + 1;
            max_depth = std::max(max_depth, depth[k]);
            bool leaf = nrefs[k] > 0;
            uint32_t link = make_link(leaf, offset[k], nrefs[k], rec_of, k);
            put_slot(r, sd, k, link);
            if (leaf) {
                T.leaf_refs.push_back(r * 2 + sd);
                T.leaf_links.push_back(link);
            } else {
                T.parent[rec_of[k]] = r * 2 + sd;
            }
        }
    }
    if (max_depth > STACK_TOTAL - 1)
        FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "%s: tree depth %d exceeds %d", fn, max_depth, STACK_TOTAL - 1);

    {
        bool leaf = nrefs[0] > 0;
        uint32_t link = make_link(leaf, offset[0], nrefs[0], rec_of, 0);
        if (leaf) {
            T.leaf_refs.push_back(-1);
            T.leaf_links.push_back(link);
        }
        if (dim == 3) {
            T.root[0] = make_float4(pmin[0], pmin[1], pmin[2], bits_as_float(link));
            T.root[1] = make_float4(pmax[0], pmax[1], pmax[2], 0.0f);
        } else {
            T.root[0] = make_float4(pmin[0], pmin[1], pmax[0], pmax[1]);
            T.root[1] = make_float4(bits_as_float(link), 0.0f, 0.0f, 0.0f);
        }
    }
    if ((int)T.leaf_refs.size() != n_leaf) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "%s: inconsistent leaves", fn);
    T.n_rec = n_rec;
    T.n_leaf = n_leaf;
    T.max_depth = max_depth;
    return FCPW_HIP_OK;
}

// ---------------------------------------------------------------------------
// Record layouts.  Records are numbered by their internal node; a layout is an
// order of those numbers in memory:
//   0  DFS: internal nodes in the preorder of the flat tree (first child first),
//      as fcpw_hip_upload_bvh and fcpw_hip_build_bvh produce them;
//   1  treelets: a node, its internal children and grandchildren stored
//      contiguously (three levels, up to 7 records = 448 bytes in 3D) in
//      breadth-first order, the treelets in DFS order of their roots, so that
//      both children of a record lie next to it rather than only the first.
// Relayout renumbers the records in place on the device: every internal link,
// the root link, the refit schedule (parents, leaf slots), the silhouette cones
// and the host table of node slots.  Only record indices change, so the
// traversal visits the same nodes in the same order and every result and count
// is unchanged.
// ---------------------------------------------------------------------------
static int relayout_records(fcpw_hip_ctx *c, int layout)
{
    const int n_rec = c->n_records;
    if (n_rec <= 1) return FCPW_HIP_OK; // nothing to reorder
    const int REC = c->dim == 3 ? 4 : 3;
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, hipDeviceSynchronize());
    std::vector<float4> rec((size_t)n_rec * REC), root(2);
    std::vector<int32_t> parent(n_rec), leaf_refs(c->n_leaves);
    HIPCHECK(c, hipMemcpy(rec.data(), c->records, rec.size() * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(root.data(), c->root, 2 * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(parent.data(), c->rec_parent, parent.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(leaf_refs.data(), c->leaf_refs, leaf_refs.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    auto link_at = [&](int r, int side) -> uint32_t & {
        float &f = c->dim == 3 ? (side == 0 ? rec[(size_t)r * 4].w : rec[(size_t)r * 4 + 2].w)
                               : (side == 0 ? rec[(size_t)r * 3 + 2].x : rec[(size_t)r * 3 + 2].y);
        return *reinterpret_cast<uint32_t *>(&f);
    };
    uint32_t &root_link = *reinterpret_cast<uint32_t *>(c->dim == 3 ? &root[0].w : &root[1].x);
    if (root_link & LEAF_BIT) return FCPW_HIP_OK;
    // the target order, from the topology
    std::vector<int32_t> order;
    order.reserve(n_rec);
    std::vector<int32_t> todo{(int32_t)root_link};
    while (!todo.empty()) {
        const int t = todo.back();
        todo.pop_back();
        if (layout == 0) {
            order.push_back(t);
            for (int side = 1; side >= 0; side--) { // first child on top of the stack
                const uint32_t l = link_at(t, side);
                if (!(l & LEAF_BIT)) todo.push_back((int32_t)l);
            }
        } else {
            std::vector<int32_t> level{t}, frontier;
            for (int d = 0; d < 3 && !level.empty(); d++) {
                std::vector<int32_t> next;
                for (int r : level) {
                    order.push_back(r);
                    for (int side = 0; side < 2; side++) {
                        const uint32_t l = link_at(r, side);
                        if (!(l & LEAF_BIT)) next.push_back((int32_t)l);
                    }
                }
                if (d == 2) frontier = next;
                level.swap(next);
            }
            for (auto it = frontier.rbegin(); it != frontier.rend(); ++it) todo.push_back(*it);
        }
    }
    if ((int)order.size() != n_rec) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "relayout: the records do not form one tree");
    std::vector<int32_t> new_of(n_rec, -1);
    for (int k = 0; k < n_rec; k++) new_of[order[k]] = k;
    auto slot_map = [&](int32_t slot) { return slot < 0 || slot >= 2 * n_rec ? slot : new_of[slot >> 1] * 2 + (slot & 1); };
    std::vector<float4> rec2(rec.size());
    std::vector<int32_t> parent2(n_rec);
    for (int r = 0; r < n_rec; r++) {
        for (int side = 0; side < 2; side++) {
            uint32_t &l = link_at(r, side);
            if (!(l & LEAF_BIT)) l = (uint32_t)new_of[l];
        }
        for (int j = 0; j < REC; j++) rec2[(size_t)new_of[r] * REC + j] = rec[(size_t)r * REC + j];
        parent2[new_of[r]] = slot_map(parent[r]);
    }
    root_link = (uint32_t)new_of[root_link];
    for (int32_t &l : leaf_refs) l = slot_map(l);
    for (int32_t &sl : c->node_slot) sl = slot_map(sl);
    HIPCHECK(c, hipMemcpy(c->records, rec2.data(), rec2.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->root, root.data(), 2 * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->rec_parent, parent2.data(), parent2.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->leaf_refs, leaf_refs.data(), leaf_refs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (c->has_silhouettes) {
        // two float4 per child slot, the root's slot (2 * n_rec) last and unmoved
        std::vector<float4> cones((size_t)(2 * n_rec + 1) * 2), cones2(cones.size());
        HIPCHECK(c, hipMemcpy(cones.data(), c->cones, cones.size() * sizeof(float4), hipMemcpyDeviceToHost));
        for (int32_t sl = 0; sl <= 2 * n_rec; sl++) {
            const int32_t to = slot_map(sl);
            cones2[(size_t)to * 2] = cones[(size_t)sl * 2];
            cones2[(size_t)to * 2 + 1] = cones[(size_t)sl * 2 + 1];
        }
        HIPCHECK(c, hipMemcpy(c->cones, cones2.data(), cones2.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
    HIPCHECK(c, hipDeviceSynchronize());
    c->generation++; // batches re-read their tables
    return FCPW_HIP_OK;
}

// allocates and fills the tree buffers of c (records, root, leaf-slot index,
// refit schedule); the scene's previous buffers must have been freed
static int upload_tree(fcpw_hip_ctx *c, const TreeRecords &T, const int32_t *slot_index, int n_slots)
{
    size_t rec_bytes = T.rec.size() * sizeof(float4);
    HIPCHECK(c, hipMalloc((void **)&c->records, rec_bytes));
    HIPCHECK(c, hipMalloc((void **)&c->root, 2 * sizeof(float4)));
    HIPCHECK(c, hipMalloc((void **)&c->prim_index, (size_t)n_slots * sizeof(int32_t)));
    HIPCHECK(c, hipMalloc((void **)&c->rec_parent, T.parent.size() * sizeof(int32_t)));
    HIPCHECK(c, hipMalloc((void **)&c->leaf_refs, T.leaf_refs.size() * sizeof(int32_t)));
    HIPCHECK(c, hipMalloc((void **)&c->leaf_links, T.leaf_links.size() * sizeof(uint32_t)));
    HIPCHECK(c, hipMalloc((void **)&c->counters, T.parent.size() * sizeof(unsigned int)));
    HIPCHECK(c, hipMemcpy(c->records, T.rec.data(), rec_bytes, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->root, T.root, sizeof(T.root), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->prim_index, slot_index, (size_t)n_slots * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->rec_parent, T.parent.data(), T.parent.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->leaf_refs, T.leaf_refs.data(), T.leaf_refs.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->leaf_links, T.leaf_links.data(), T.leaf_links.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemset(c->counters, 0, T.parent.size() * sizeof(unsigned int)));
    return FCPW_HIP_OK;
}

// Morton grid of the query ordering: 1.5x the scene box
static void set_query_grid(fcpw_hip_ctx *c, const float *pmin, const float *pmax, int dim)
{
    for (int d = 0; d < 3; d++) {
        float lo = d < dim ? pmin[d] : 0.0f, hi = d < dim ? pmax[d] : 0.0f;
        float ext = std::max(hi - lo, 1e-6f);
        c->qlo[d] = 0.5f * (lo + hi) - 0.75f * ext;
        c->qinv[d] = 1.0f / (1.5f * ext);
    }
}

extern "C" {

int fcpw_hip_upload_bvh(fcpw_hip_ctx *c, int dim, int n_nodes, const float *pmin, const float *pmax,
                        const int32_t *nrefs, const int32_t *offset, int n_prims, const int32_t *prim_vertices,
                        const int32_t *prim_index, int n_vertices, const float *positions)
{
    if (!c) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    if (dim != 2 && dim != 3) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_bvh: dim must be 2 or 3, got %d", dim);
    if (n_nodes <= 0 || n_prims <= 0 || n_vertices <= 0 || !pmin || !pmax || !nrefs || !offset || !prim_vertices ||
        !prim_index || !positions)
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_bvh: empty BVH or NULL array");
    if ((size_t)n_prims > (size_t)OFF_MASK)
        FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "upload_bvh: %d primitives exceed the 2^27 leaf-slot range", n_prims);
    TreeRecords T;
    int st = tree_records(c, "upload_bvh", dim, n_nodes, pmin, pmax, nrefs, offset, n_prims, T);
    if (st) return st;
    for (int64_t k = 0; k < (int64_t)n_prims * dim; k++)
        if (prim_vertices[k] < 0 || prim_vertices[k] >= n_vertices)
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_bvh: vertex id out of range");

    // primitive records in leaf-slot order
    const int PR = dim == 3 ? 3 : 1;
    std::vector<float4> prims((size_t)n_prims * PR);
    for (int s = 0; s < n_prims; s++) {
        if (dim == 3) {
            for (int k = 0; k < 3; k++) {
                const float *v = positions + (size_t)prim_vertices[(size_t)s * 3 + k] * 3;
                prims[(size_t)s * 3 + k] = make_float4(v[0], v[1], v[2], 0.0f);
            }
        } else {
            const float *a = positions + (size_t)prim_vertices[(size_t)s * 2] * 2;
            const float *b = positions + (size_t)prim_vertices[(size_t)s * 2 + 1] * 2;
            prims[s] = make_float4(a[0], a[1], b[0], b[1]);
        }
    }

    HIPCHECK(c, hipSetDevice(c->device));
    free_scene(c);
    st = upload_tree(c, T, prim_index, n_prims);
    if (st) return st;
    HIPCHECK(c, hipMalloc((void **)&c->prims, prims.size() * sizeof(float4)));
    HIPCHECK(c, hipMalloc((void **)&c->prim_vids, (size_t)n_prims * dim * sizeof(int32_t)));
    HIPCHECK(c, hipMalloc((void **)&c->V, (size_t)n_vertices * dim * sizeof(float)));
    HIPCHECK(c, hipMemcpy(c->prims, prims.data(), prims.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->prim_vids, prim_vertices, (size_t)n_prims * dim * sizeof(int32_t),
                          hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->V, positions, (size_t)n_vertices * dim * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(c, hipDeviceSynchronize());
    set_query_grid(c, pmin, pmax, dim);
    c->dim = dim;
    c->n_records = T.n_rec;
    c->n_leaves = T.n_leaf;
    c->n_prims = n_prims;
    c->n_vertices = n_vertices;
    c->max_depth = T.max_depth;
    c->node_slot = std::move(T.node_slot);
    c->node_leaf_off = std::move(T.node_leaf_off);
    c->uploaded = true;
    c->generation++;
    c->positions_version++;
    return c->layout ? relayout_records(c, c->layout) : FCPW_HIP_OK;
}

int fcpw_hip_build_bvh(fcpw_hip_ctx *c, int dim, const float *positions, int n_vertices, const int32_t *indices,
                       int n_prims, int heuristic, float *node_pmin, float *node_pmax, int32_t *node_nrefs,
                       int32_t *node_offset, int32_t *prim_order, int *n_nodes)
{
    if (!c) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    if (dim != 2 && dim != 3) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "build_bvh: dim must be 2 or 3, got %d", dim);
    if (n_prims <= 0 || n_vertices <= 0 || !positions || !indices)
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "build_bvh: Cannot construct a BVH without primitives");
    if (heuristic < 0 || heuristic > 4)
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "build_bvh: unknown heuristic %d", heuristic);
    if ((size_t)n_prims > (size_t)OFF_MASK || (int64_t)n_prims * 2 > (int64_t)INT32_MAX / 4)
        FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "build_bvh: %d primitives exceed the supported range", n_prims);
    for (int64_t k = 0; k < (int64_t)n_prims * dim; k++)
        if (indices[k] < 0 || indices[k] >= n_vertices)
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "build_bvh: vertex id out of range");
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, hipDeviceSynchronize());
    free_scene(c);
    int st = dim == 3 ? gpu_build<3>(c, positions, n_vertices, indices, n_prims, heuristic, node_pmin, node_pmax,
                                     node_nrefs, node_offset, prim_order, n_nodes)
                      : gpu_build<2>(c, positions, n_vertices, indices, n_prims, heuristic, node_pmin, node_pmax,
                                     node_nrefs, node_offset, prim_order, n_nodes);
    if (st) free_scene(c);
    return st;
}

int fcpw_hip_upload_silhouettes(fcpw_hip_ctx *c, int n_nodes, const fcpw_hip_bounding_cone *cones,
                                const int32_t *n_silhouettes, const int32_t *silhouette_offset, int n_refs,
                                const void *silhouettes)
{
    int st = check_single(c, "upload_silhouettes");
    if (st) return st;
    if (c->node_slot.empty())
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT,
             "upload_silhouettes: the node layout is unknown (a device build without the exported tree)");
    if (n_nodes != (int)c->node_slot.size() || !cones || !n_silhouettes || !silhouette_offset || n_refs < 0 ||
        (n_refs > 0 && !silhouettes))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_silhouettes: expected %d nodes and non-NULL arrays",
             (int)c->node_slot.size());
    const int n_slots = 2 * c->n_records + 1;
    std::vector<float4> cone_buf((size_t)n_slots * 2, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    std::vector<uint2> leaf_buf((size_t)c->n_prims, make_uint2(0u, 0u));
    int leaf_max = 0;
    for (int i = 0; i < n_nodes; i++) {
        const fcpw_hip_bounding_cone &k = cones[i];
        const int slot = c->node_slot[i];
        cone_buf[(size_t)slot * 2] = make_float4(k.axis.x, k.axis.y, k.axis.z, k.half_angle);
        cone_buf[(size_t)slot * 2 + 1] = cone_aux(k.axis.x, k.axis.y, k.axis.z, k.half_angle, k.radius);
        const int32_t cnt = n_silhouettes[i], off = silhouette_offset[i];
        if (cnt < 0 || off < 0 || (int64_t)off + cnt > n_refs)
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_silhouettes: node %d silhouette range out of bounds", i);
        if (c->node_leaf_off[i] < 0) {
            if (cnt != 0) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "upload_silhouettes: internal node %d has silhouettes", i);
        } else {
            leaf_buf[c->node_leaf_off[i]] = make_uint2((uint32_t)off, (uint32_t)cnt);
            leaf_max = std::max(leaf_max, (int)cnt);
        }
    }
    const int RW = c->dim == 3 ? 4 : 2;
    std::vector<float4> rec_buf((size_t)std::max(n_refs, 1) * RW, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (int r = 0; r < n_refs; r++) {
        if (c->dim == 2) {
            const fcpw_hip_vertex &v = ((const fcpw_hip_vertex *)silhouettes)[r];
            rec_buf[(size_t)r * 2] = make_float4(v.p.x, v.p.y, v.n0.x, v.n0.y);
            rec_buf[(size_t)r * 2 + 1] =
                make_float4(v.n1.x, v.n1.y, bits_as_float(v.index), bits_as_float(v.has_one_adjacent_face));
        } else {
            const fcpw_hip_edge &e = ((const fcpw_hip_edge *)silhouettes)[r];
            rec_buf[(size_t)r * 4] = make_float4(e.pa.x, e.pa.y, e.pa.z, bits_as_float(e.index));
            rec_buf[(size_t)r * 4 + 1] = make_float4(e.pb.x, e.pb.y, e.pb.z, bits_as_float(e.has_one_adjacent_face));
            rec_buf[(size_t)r * 4 + 2] = make_float4(e.n0.x, e.n0.y, e.n0.z, 0.0f);
            rec_buf[(size_t)r * 4 + 3] = make_float4(e.n1.x, e.n1.y, e.n1.z, 0.0f);
        }
    }
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, hipDeviceSynchronize());
    free_silhouettes(c);
    HIPCHECK(c, hipMalloc((void **)&c->cones, cone_buf.size() * sizeof(float4)));
    HIPCHECK(c, hipMalloc((void **)&c->sil_leaf, leaf_buf.size() * sizeof(uint2)));
    HIPCHECK(c, hipMalloc((void **)&c->sils, rec_buf.size() * sizeof(float4)));
    HIPCHECK(c, hipMemcpy(c->cones, cone_buf.data(), cone_buf.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->sil_leaf, leaf_buf.data(), leaf_buf.size() * sizeof(uint2), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->sils, rec_buf.data(), rec_buf.size() * sizeof(float4), hipMemcpyHostToDevice));
    c->n_sil_refs = n_refs;
    c->sil_leaf_max = leaf_max;
    c->has_silhouettes = true;
    return FCPW_HIP_OK;
}

static int launch_rebuild(fcpw_hip_ctx *c, hipStream_t s)
{
    int threads = 256, blocks = (c->n_prims + threads - 1) / threads;
    if (c->dim == 3)
        hipLaunchKernelGGL(k_rebuild_prims<3>, dim3(blocks), dim3(threads), 0, s, c->prims, c->prim_vids, c->V,
                           c->n_prims);
    else
        hipLaunchKernelGGL(k_rebuild_prims<2>, dim3(blocks), dim3(threads), 0, s, c->prims, c->prim_vids, c->V,
                           c->n_prims);
    HIPCHECK(c, hipGetLastError());
    return FCPW_HIP_OK;
}

int fcpw_hip_update_positions(fcpw_hip_ctx *c, const float *positions, int n_vertices)
{
    int st = check_single(c, "update_positions");
    if (st) return st;
    if (!positions || n_vertices != c->n_vertices)
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "update_positions: expected %d vertices, got %d", c->n_vertices,
             n_vertices);
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, hipMemcpy(c->V, positions, (size_t)n_vertices * c->dim * sizeof(float), hipMemcpyHostToDevice));
    c->positions_version++;
    st = launch_rebuild(c, 0);
    if (st) return st;
    HIPCHECK(c, hipDeviceSynchronize());
    return FCPW_HIP_OK;
}

int fcpw_hip_update_positions_device(fcpw_hip_ctx *c, const float *positions, int n_vertices, void *stream)
{
    int st = check_single(c, "update_positions_device");
    if (st) return st;
    if (!positions || n_vertices != c->n_vertices)
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "update_positions_device: expected %d vertices, got %d",
             c->n_vertices, n_vertices);
    hipStream_t s = (hipStream_t)stream;
    HIPCHECK(c, hipMemcpyAsync(c->V, positions, (size_t)n_vertices * c->dim * sizeof(float),
                               hipMemcpyDeviceToDevice, s));
    c->positions_version++;
    return launch_rebuild(c, s);
}

int fcpw_hip_refit(fcpw_hip_ctx *c, void *stream)
{
    int st = check_ready(c, "refit");
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    if (c->instanced) return instances_refit(c, s);
    int threads = 256, blocks = (c->n_leaves + threads - 1) / threads;
    if (c->dim == 3)
        hipLaunchKernelGGL(k_refit<3>, dim3(blocks), dim3(threads), 0, s, c->records, c->root, c->prims,
                           c->leaf_refs, c->leaf_links, c->rec_parent, c->counters, c->n_leaves);
    else
        hipLaunchKernelGGL(k_refit<2>, dim3(blocks), dim3(threads), 0, s, c->records, c->root, c->prims,
                           c->leaf_refs, c->leaf_links, c->rec_parent, c->counters, c->n_leaves);
    HIPCHECK(c, hipGetLastError());
    if (c->has_silhouettes) {
        if (c->dim == 3)
            hipLaunchKernelGGL(k_refit_cones<3>, dim3(blocks), dim3(threads), 0, s, c->records, c->root, c->cones,
                               c->sil_leaf, c->sils, c->leaf_refs, c->leaf_links, c->rec_parent, c->counters,
                               c->n_leaves, c->n_records);
        else
            hipLaunchKernelGGL(k_refit_cones<2>, dim3(blocks), dim3(threads), 0, s, c->records, c->root, c->cones,
                               c->sil_leaf, c->sils, c->leaf_refs, c->leaf_links, c->rec_parent, c->counters,
                               c->n_leaves, c->n_records);
        HIPCHECK(c, hipGetLastError());
    }
    return FCPW_HIP_OK;
}

// ---------------------------------------------------------------------------
// launch helpers: (block, stack) pairs are compiled in; the grid is capped and
// the kernels stride over the remaining queries
// ---------------------------------------------------------------------------
static dim3 grid_for(fcpw_hip_ctx *, int64_t n, int block)
{
    return dim3((unsigned)std::max<int64_t>(1, (n + block - 1) / block));
}

// "timing": a HIP event pair on the launch stream around each traversal launch
static int timing_begin(fcpw_hip_ctx *c, hipStream_t s)
{
    if (!c->timing || c->ev_used >= fcpw_hip_ctx::MAX_TIMED) return FCPW_HIP_OK;
    if (c->ev_used == (int)c->ev_begin.size()) {
        hipEvent_t a = nullptr, b = nullptr;
        HIPCHECK(c, hipEventCreate(&a));
        hipError_t e = hipEventCreate(&b);
        if (e != hipSuccess) (void)hipEventDestroy(a);
        HIPCHECK(c, e);
        c->ev_begin.push_back(a);
        c->ev_end.push_back(b);
    }
    HIPCHECK(c, hipEventRecord(c->ev_begin[c->ev_used], s));
    return FCPW_HIP_OK;
}

static int timing_end(fcpw_hip_ctx *c, hipStream_t s)
{
    if (!c->timing || c->ev_used >= fcpw_hip_ctx::MAX_TIMED) return FCPW_HIP_OK;
    HIPCHECK(c, hipEventRecord(c->ev_end[c->ev_used], s));
    c->ev_used++;
    return FCPW_HIP_OK;
}

// Morton order of n query points (pts + i * stride floats); returns the
// permutation, or nullptr when ordering is off or the batch is small
// Cells of the ordering grid: morton_bits in total, of which the direction's
// octant (quadrant) takes dim bits for rays (dir_off >= 0)
static int64_t morton_cells(int dim, int morton_bits, int dir_off, int *axis_bits)
{
    const int dir_bits = dir_off >= 0 ? dim : 0;
    *axis_bits = (morton_bits - dir_bits) / dim;
    return (int64_t)1 << (*axis_bits * dim + dir_bits);
}

// Enqueues the counting sort of n queries along the Morton curve into the given
// buffers: codes and ranks (2n words), perm (n), hist (cells), sums (segments).
static int morton_order_into(fcpw_hip_ctx *c, const float *pts, int stride, int64_t n, hipStream_t s, int morton_bits,
                             int dir_off, uint32_t *codes, uint32_t *perm, uint32_t *hist, uint32_t *sums);

static int morton_order(fcpw_hip_ctx *c, const float *pts, int stride, int64_t n, hipStream_t s,
                        const uint32_t **perm_out, int dir_off = -1)
{
    *perm_out = nullptr;
    if (!c->sort || n < 4096) return FCPW_HIP_OK;
    if (n > (int64_t)UINT32_MAX) FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "query batches above 2^32 are not supported");
    int axis_bits;
    const int64_t cells = morton_cells(c->dim, c->morton_bits, dir_off, &axis_bits);
    const int64_t segs = (cells + SCAN_SEG - 1) / SCAN_SEG;
    c->ws_perm_n = -1;
    int st = grow(c, &c->ws_codes, &c->ws_codes_bytes, (size_t)n * 8); // codes, then ranks
    if (!st) st = grow(c, &c->ws_perm, &c->ws_perm_bytes, (size_t)n * 4);
    if (!st) st = grow(c, &c->ws_hist, &c->ws_hist_bytes, (size_t)cells * 4);
    if (!st) st = grow(c, &c->ws_sums, &c->ws_sums_bytes, (size_t)segs * 4);
    if (st) return st;
    st = morton_order_into(c, pts, stride, n, s, c->morton_bits, dir_off, (uint32_t *)c->ws_codes,
                           (uint32_t *)c->ws_perm, (uint32_t *)c->ws_hist, (uint32_t *)c->ws_sums);
    if (st) return st;
    c->ws_perm_n = n;
    *perm_out = (const uint32_t *)c->ws_perm;
    return FCPW_HIP_OK;
}

static int morton_order_into(fcpw_hip_ctx *c, const float *pts, int stride, int64_t n, hipStream_t s, int morton_bits,
                             int dir_off, uint32_t *codes, uint32_t *perm, uint32_t *hist, uint32_t *sums)
{
    int axis_bits;
    const int64_t cells = morton_cells(c->dim, morton_bits, dir_off, &axis_bits);
    const int64_t segs = (cells + SCAN_SEG - 1) / SCAN_SEG;
    uint32_t *ranks = codes + n;
    unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, hist, 0u, cells);
    float4 lo = make_float4(c->qlo[0], c->qlo[1], c->qlo[2], 0.0f);
    float4 inv = make_float4(c->qinv[0], c->qinv[1], c->qinv[2], 0.0f);
    if (c->dim == 3)
        hipLaunchKernelGGL(k_morton_hist<3>, dim3(nb), dim3(256), 0, s, pts, stride, n, lo, inv, axis_bits, codes,
                           ranks, hist, dir_off);
    else
        hipLaunchKernelGGL(k_morton_hist<2>, dim3(nb), dim3(256), 0, s, pts, stride, n, lo, inv, axis_bits, codes,
                           ranks, hist, dir_off);
    hipLaunchKernelGGL(k_scan_segment_sums, dim3((unsigned)segs), dim3(SCAN_THREADS), 0, s, hist, cells, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, (int)segs);
    hipLaunchKernelGGL(k_scan_segment_apply, dim3((unsigned)segs), dim3(SCAN_THREADS), 0, s, hist, cells, sums);
    hipLaunchKernelGGL(k_morton_place, dim3(nb), dim3(256), 0, s, codes, ranks, n, hist, perm);
    HIPCHECK(c, hipGetLastError());
    return FCPW_HIP_OK;
}

#define FCPW_DISPATCH_BK(c, V, ...)                                                                            \
    do {                                                                                                       \
        const int _b = (c)->block, _k = (c)->stack < 8 || (c)->stack == 12 ? 8 : (c)->stack;                  \
        if (_b == 64 && _k == 8) { constexpr int B = 64, K = 8; __VA_ARGS__; }                               \
        else if (_b == 64 && _k == 16) { constexpr int B = 64, K = 16; __VA_ARGS__; }                        \
        else if (_b == 64 && _k == 24) { constexpr int B = 64, K = 24; __VA_ARGS__; }                        \
        else if (_b == 128 && _k == 8) { constexpr int B = 128, K = 8; __VA_ARGS__; }                        \
        else if (_b == 128 && _k == 16) { constexpr int B = 128, K = 16; __VA_ARGS__; }                      \
        else if (_b == 128 && _k == 24) { constexpr int B = 128, K = 24; __VA_ARGS__; }                      \
        else if (_b == 256 && _k == 8) { constexpr int B = 256, K = 8; __VA_ARGS__; }                        \
        else if (_b == 256 && _k == 16) { constexpr int B = 256, K = 16; __VA_ARGS__; }                      \
        else { constexpr int B = 256, K = 24; __VA_ARGS__; }                                                 \
    } while (0)



} // extern "C"

// closest-point launch for either output policy; io.in holds spheres (4
// floats per query) for interaction records, points (dim floats) otherwise
template <int OUT>
static int launch_closest_kernel(fcpw_hip_ctx *c, const CpIO &io, hipStream_t s, bool stats);

template <int OUT>
static int launch_closest(fcpw_hip_ctx *c, CpIO io, hipStream_t s)
{
    int st = morton_order(c, io.in, OUT == CP_OUT_INTERACTION ? 4 : c->dim, io.n, s, &io.order);
    if (st) return st;
    if ((st = timing_begin(c, s))) return st;
    if ((st = launch_closest_kernel<OUT>(c, io, s, c->stats_on))) return st;
    return timing_end(c, s);
}

// the traversal kernel alone (the queries ordered by io.order when set)
template <int OUT>
static int launch_closest_kernel(fcpw_hip_ctx *c, const CpIO &io, hipStream_t s, bool stats)
{
    DevScene S = dev_scene(c);
    const int64_t n = io.n;
    if (stats) {
        // counting pass: the coop kernel with per-wave counters (same traversal order)
        dim3 g = grid_for(c, n, 64);
        if (c->dim == 3)
            hipLaunchKernelGGL((k_closest_points_coop<3, 16, OUT, true>), g, dim3(64), 0, s, S, io, xcd_arg(c),
                               c->hold_min);
        else
            hipLaunchKernelGGL((k_closest_points_coop<2, 16, OUT, true>), g, dim3(64), 0, s, S, io, xcd_arg(c),
                               c->hold_min);
    } else if (c->coop) {
        dim3 g = grid_for(c, n, 64);
#define FCPW_COOP(D, K) \
    hipLaunchKernelGGL((k_closest_points_coop<D, K, OUT, false>), g, dim3(64), 0, s, S, io, xcd_arg(c), c->hold_min)
        const int k = c->stack; // 0, 4, 6, 8, 12, 16 or 24 LDS entries (2D: 24 runs as 16)
        if (c->dim == 3) {
            if (k == 0) FCPW_COOP(3, 0);
            else if (k == 4) FCPW_COOP(3, 4);
            else if (k == 6) FCPW_COOP(3, 6);
            else if (k == 8) FCPW_COOP(3, 8);
            else if (k == 12) FCPW_COOP(3, 12);
            else if (k == 16) FCPW_COOP(3, 16);
            else FCPW_COOP(3, 24);
        } else {
            if (k == 0) FCPW_COOP(2, 0);
            else if (k == 4) FCPW_COOP(2, 4);
            else if (k == 6) FCPW_COOP(2, 6);
            else if (k == 8) FCPW_COOP(2, 8);
            else if (k == 12) FCPW_COOP(2, 12);
            else FCPW_COOP(2, 16);
        }
#undef FCPW_COOP
    } else if (c->dim == 3) {
        FCPW_DISPATCH_BK(c, 0, hipLaunchKernelGGL((k_closest_points<3, B, K, OUT>), grid_for(c, n, B), dim3(B), 0, s,
                                                  S, io, xcd_arg(c)));
    } else {
        FCPW_DISPATCH_BK(c, 0, hipLaunchKernelGGL((k_closest_points<2, B, K, OUT>), grid_for(c, n, B), dim3(B), 0, s,
                                                  S, io, xcd_arg(c)));
    }
    HIPCHECK(c, hipGetLastError());
    return FCPW_HIP_OK;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free closest-point entry point: the query-ordering
// workspace in a file-scope device array instead of the context's hipMalloc'ed,
// grow-on-demand buffers.  Every GPU the library runs on holds its own copy
// until the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for that device (noalloc_scratch_ptr), never
// through the array's name.
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_MAX_QUERIES = (int64_t)1 << 21; // 2,097,152 queries per call
constexpr int NOALLOC_MORTON_BITS = 18;                  // a 2^18-cell ordering grid
constexpr int64_t NOALLOC_CELLS = (int64_t)1 << NOALLOC_MORTON_BITS;
constexpr int64_t NOALLOC_SEGS = (NOALLOC_CELLS + SCAN_SEG - 1) / SCAN_SEG;
// 32-bit words: codes and ranks (2 per query), perm (1 per query), histogram, segment sums
constexpr int64_t NOALLOC_WORDS = 3 * NOALLOC_MAX_QUERIES + NOALLOC_CELLS + NOALLOC_SEGS;
__device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

static std::mutex g_scratch_mutex;
static std::vector<void *> g_scratch_ptr; // per device; null until looked up

// The device address of `scratch` on c's device, looked up once per device with
// that device current; null (and c's error set) when the lookup fails.
static void *noalloc_scratch_ptr(fcpw_hip_ctx *c)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= c->device) g_scratch_ptr.resize(c->device + 1, nullptr);
    if (!g_scratch_ptr[c->device]) {
        void *ptr = nullptr;
        hipError_t e = hipSetDevice(c->device);
        if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            c->err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_ptr[c->device] = ptr;
    }
    return g_scratch_ptr[c->device];
}

static int launch_closest_records(fcpw_hip_ctx *c, const void *spheres, int64_t n, int record_normals, void *out,
                                  hipStream_t s)
{
    CpIO io = {};
    io.in = (const float *)spheres;
    io.n = n;
    io.out = (float2 *)out;
    io.record_normals = record_normals;
    return launch_closest<CP_OUT_INTERACTION>(c, io, s);
}

extern "C" {

static int launch_intersect(fcpw_hip_ctx *c, const void *rays, int64_t n, int occlusion, void *out, hipStream_t s)
{
    DevScene S = dev_scene(c);
    const uint32_t *order;
    // ordered by origin, and by direction octant when ray_sort_dir is set
    int st = morton_order(c, (const float *)rays, 10, n, s, &order, c->ray_sort_dir ? 3 : -1);
    if (st) return st;
    if ((st = timing_begin(c, s))) return st;
    if (c->ray_coop) {
        const dim3 g = grid_for(c, n, 64);
        const int k = c->ray_stack;
#define FCPW_RCOOP(D, K, ST)                                                                                  \
    hipLaunchKernelGGL((k_intersect_coop<D, K, ST>), g, dim3(64), 0, s, S, (const float2 *)rays, n, occlusion, \
                       (float2 *)out, order, xcd_arg(c), c->ray_hold_min)
        if (c->stats_on) {
            if (c->dim == 3) FCPW_RCOOP(3, 16, true);
            else FCPW_RCOOP(2, 16, true);
        } else if (c->dim == 3) {
            if (k == 4) FCPW_RCOOP(3, 4, false);
            else if (k == 8) FCPW_RCOOP(3, 8, false);
            else FCPW_RCOOP(3, 16, false);
        } else {
            if (k == 4) FCPW_RCOOP(2, 4, false);
            else if (k == 8) FCPW_RCOOP(2, 8, false);
            else FCPW_RCOOP(2, 16, false);
        }
#undef FCPW_RCOOP
    } else if (c->stats_on && c->dim == 3)
        hipLaunchKernelGGL((k_intersect<3, 128, 16, true>), grid_for(c, n, 128), dim3(128), 0, s, S,
                           (const float2 *)rays, n, occlusion, (float2 *)out, order, xcd_arg(c));
    else if (c->stats_on)
        hipLaunchKernelGGL((k_intersect<2, 128, 16, true>), grid_for(c, n, 128), dim3(128), 0, s, S,
                           (const float2 *)rays, n, occlusion, (float2 *)out, order, xcd_arg(c));
    else if (c->dim == 3)
        FCPW_DISPATCH_BK(c, 0, hipLaunchKernelGGL((k_intersect<3, B, K, false>), grid_for(c, n, B), dim3(B), 0, s, S,
                                            (const float2 *)rays, n, occlusion, (float2 *)out, order, xcd_arg(c)));
    else
        FCPW_DISPATCH_BK(c, 0, hipLaunchKernelGGL((k_intersect<2, B, K, false>), grid_for(c, n, B), dim3(B), 0, s, S,
                                            (const float2 *)rays, n, occlusion, (float2 *)out, order, xcd_arg(c)));
    HIPCHECK(c, hipGetLastError());
    return timing_end(c, s);
}

static int check_overflow(fcpw_hip_ctx *c)
{
    int flag = 0;
    HIPCHECK(c, hipMemcpy(&flag, c->flags, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
        HIPCHECK(c, hipMemset(c->flags, 0, sizeof(int)));
        FAIL(c, FCPW_HIP_ERR_TRAVERSAL_OVERFLOW, "traversal stack exceeded %d entries", STACK_TOTAL);
    }
    return FCPW_HIP_OK;
}

int fcpw_hip_closest_points_device(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres, int64_t n,
                                   int record_normals, fcpw_hip_interaction *out, void *stream)
{
    int st = check_ready(c, "closest_points_device");
    if (st) return st;
    if (c->instanced) return fcpw_hip_closest_points_objects_device(c, spheres, n, record_normals, out, nullptr, stream);
    if (n < 0 || (n > 0 && (!spheres || !out))) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "closest_points_device: bad arguments");
    if (((uintptr_t)spheres & 15) || ((uintptr_t)out & 7))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "closest_points_device: spheres must be 16-byte and out 8-byte aligned");
    if (n == 0) return FCPW_HIP_OK;
    return launch_closest_records(c, spheres, n, record_normals, out, (hipStream_t)stream);
}

int fcpw_hip_closest_points_device_noalloc(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres, int64_t n,
                                           int record_normals, fcpw_hip_interaction *out, void *stream)
{
    int st = check_single(c, "closest_points_device_noalloc");
    if (st) return st;
    if (n < 0 || (n > 0 && (!spheres || !out)))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "closest_points_device_noalloc: bad arguments");
    if (n > NOALLOC_MAX_QUERIES)
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT,
             "closest_points_device_noalloc: %lld queries exceed the scratch maximum of %lld per call",
             (long long)n, (long long)NOALLOC_MAX_QUERIES);
    if (((uintptr_t)spheres & 15) || ((uintptr_t)out & 7))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT,
             "closest_points_device_noalloc: spheres must be 16-byte and out 8-byte aligned");
    if (!c->scratch)
        FAIL(c, FCPW_HIP_ERR_HIP, "closest_points_device_noalloc: the scratch array's address is unknown (%s)",
             c->scratch_err.c_str());
    if (n == 0) return FCPW_HIP_OK;
    hipStream_t s = (hipStream_t)stream;
    CpIO io = {};
    io.in = (const float *)spheres;
    io.n = n;
    io.out = (float2 *)out;
    io.record_normals = record_normals;
    if (c->sort && n >= 4096) {
        // codes | ranks | perm | histogram | segment sums; the histogram is cleared
        // by the fill kernel morton_order_into starts with
        uint32_t *codes = (uint32_t *)c->scratch;
        uint32_t *perm = codes + 2 * NOALLOC_MAX_QUERIES, *hist = perm + NOALLOC_MAX_QUERIES;
        uint32_t *sums = hist + NOALLOC_CELLS;
        st = morton_order_into(c, io.in, 4, n, s, NOALLOC_MORTON_BITS, -1, codes, perm, hist, sums);
        if (st) return st;
        io.order = perm;
    }
    return launch_closest_kernel<CP_OUT_INTERACTION>(c, io, s, false);
}

int fcpw_hip_closest_points(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres, int64_t n, int record_normals,
                            fcpw_hip_interaction *out)
{
    int st = check_ready(c, "closest_points");
    if (st) return st;
    if (c->instanced) return fcpw_hip_closest_points_objects(c, spheres, n, record_normals, out, nullptr);
    if (n < 0 || (n > 0 && (!spheres || !out))) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "closest_points: bad arguments");
    if (n == 0) return FCPW_HIP_OK;
    HIPCHECK(c, hipSetDevice(c->device));
    st = grow(c, &c->stage_in, &c->stage_in_bytes, (size_t)n * sizeof(fcpw_hip_bounding_sphere));
    if (st) return st;
    st = grow(c, &c->stage_out, &c->stage_out_bytes, (size_t)n * sizeof(fcpw_hip_interaction));
    if (st) return st;
    st = stage_copy_in(c, c->stage_in, spheres, (size_t)n * sizeof(fcpw_hip_bounding_sphere), 0);
    if (st) return st;
    st = launch_closest_records(c, c->stage_in, n, record_normals, c->stage_out, 0);
    if (st) return st;
    if ((st = stage_copy_out(c, out, c->stage_out, (size_t)n * sizeof(fcpw_hip_interaction), 0))) return st;
    return check_overflow(c);
}

int fcpw_hip_intersect_device(fcpw_hip_ctx *c, const fcpw_hip_ray *rays, int64_t n, int occlusion,
                              fcpw_hip_interaction *out, void *stream)
{
    int st = check_ready(c, "intersect_device");
    if (st) return st;
    if (c->instanced) return fcpw_hip_intersect_objects_device(c, rays, n, occlusion, out, nullptr, stream);
    if (n < 0 || (n > 0 && (!rays || !out))) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "intersect_device: bad arguments");
    if (((uintptr_t)rays & 7) || ((uintptr_t)out & 7))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "intersect_device: rays and out must be 8-byte aligned");
    if (n == 0) return FCPW_HIP_OK;
    return launch_intersect(c, rays, n, occlusion, out, (hipStream_t)stream);
}

int fcpw_hip_intersect(fcpw_hip_ctx *c, const fcpw_hip_ray *rays, int64_t n, int occlusion, fcpw_hip_interaction *out)
{
    int st = check_ready(c, "intersect");
    if (st) return st;
    if (c->instanced) return fcpw_hip_intersect_objects(c, rays, n, occlusion, out, nullptr);
    if (n < 0 || (n > 0 && (!rays || !out))) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "intersect: bad arguments");
    if (n == 0) return FCPW_HIP_OK;
    HIPCHECK(c, hipSetDevice(c->device));
    st = grow(c, &c->stage_in, &c->stage_in_bytes, (size_t)n * sizeof(fcpw_hip_ray));
    if (st) return st;
    st = grow(c, &c->stage_out, &c->stage_out_bytes, (size_t)n * sizeof(fcpw_hip_interaction));
    if (st) return st;
    st = stage_copy_in(c, c->stage_in, rays, (size_t)n * sizeof(fcpw_hip_ray), 0);
    if (st) return st;
    st = launch_intersect(c, c->stage_in, n, occlusion, c->stage_out, 0);
    if (st) return st;
    if ((st = stage_copy_out(c, out, c->stage_out, (size_t)n * sizeof(fcpw_hip_interaction), 0))) return st;
    return check_overflow(c);
}

static int launch_spheres(fcpw_hip_ctx *c, const void *spheres, const float *rnd, int64_t n, void *out, hipStream_t s)
{
    DevScene S = dev_scene(c);
    const unsigned blocks = (unsigned)((n + 127) / 128);
    if (c->dim == 3)
        hipLaunchKernelGGL(k_intersect_spheres<3>, dim3(blocks), dim3(128), 0, s, S, (const float4 *)spheres, rnd, n,
                           (float2 *)out);
    else
        hipLaunchKernelGGL(k_intersect_spheres<2>, dim3(blocks), dim3(128), 0, s, S, (const float4 *)spheres, rnd, n,
                           (float2 *)out);
    HIPCHECK(c, hipGetLastError());
    return FCPW_HIP_OK;
}

int fcpw_hip_intersect_spheres_device(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres,
                                      const fcpw_hip_float3 *rand_nums, int64_t n, fcpw_hip_interaction *out,
                                      void *stream)
{
    int st = check_ready(c, "intersect_spheres_device");
    if (st) return st;
    if (c->instanced)
        return fcpw_hip_intersect_spheres_objects_device(c, spheres, rand_nums, n, out, nullptr, stream);
    if (n < 0 || (n > 0 && (!spheres || !rand_nums || !out)))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "intersect_spheres_device: bad arguments");
    if (((uintptr_t)spheres & 15) || ((uintptr_t)rand_nums & 3) || ((uintptr_t)out & 7))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT,
             "intersect_spheres_device: spheres must be 16-byte, rand_nums 4-byte and out 8-byte aligned");
    if (n == 0) return FCPW_HIP_OK;
    return launch_spheres(c, spheres, (const float *)rand_nums, n, out, (hipStream_t)stream);
}

int fcpw_hip_intersect_spheres(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres,
                               const fcpw_hip_float3 *rand_nums, int64_t n, fcpw_hip_interaction *out)
{
    int st = check_ready(c, "intersect_spheres");
    if (st) return st;
    if (c->instanced) return fcpw_hip_intersect_spheres_objects(c, spheres, rand_nums, n, out, nullptr);
    if (n < 0 || (n > 0 && (!spheres || !rand_nums || !out)))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "intersect_spheres: bad arguments");
    if (n == 0) return FCPW_HIP_OK;
    HIPCHECK(c, hipSetDevice(c->device));
    const size_t sb = (size_t)n * sizeof(fcpw_hip_bounding_sphere), rb = (size_t)n * sizeof(fcpw_hip_float3);
    st = grow(c, &c->stage_in, &c->stage_in_bytes, sb + rb);
    if (st) return st;
    st = grow(c, &c->stage_out, &c->stage_out_bytes, (size_t)n * sizeof(fcpw_hip_interaction));
    if (st) return st;
    if ((st = stage_copy_in(c, c->stage_in, spheres, sb, 0))) return st;
    if ((st = stage_copy_in(c, (char *)c->stage_in + sb, rand_nums, rb, 0))) return st;
    st = launch_spheres(c, c->stage_in, (const float *)((char *)c->stage_in + sb), n, c->stage_out, 0);
    if (st) return st;
    if ((st = stage_copy_out(c, out, c->stage_out, (size_t)n * sizeof(fcpw_hip_interaction), 0))) return st;
    return FCPW_HIP_OK;
}

// LDS stack entries of the 3D cooperative silhouette kernel (4M queries on the
// torus: 118.7 ms with 4, 120.1 with 0, 129.4 with 8, 132.0 with 12)
constexpr int SIL_COOP_K3 = 4;

static int launch_silhouettes(fcpw_hip_ctx *c, const void *spheres, const uint32_t *flip, int64_t n, float r2min,
                              float prec, void *out, hipStream_t s)
{
    DevScene S = dev_scene(c);
    SilScene T = {c->cones, c->sil_leaf, c->sils};
    const uint32_t *order = nullptr;
    int st = morton_order(c, (const float *)spheres, 4, n, s, &order);
    if (st) return st;
    const unsigned blocks = (unsigned)((n + 127) / 128);
    st = timing_begin(c, s);
    if (st) return st;
    if (c->coop && c->sil_leaf_max <= 1023) {
        // record indices within a leaf travel in 10 bits of a leaf-pass item
        const dim3 g = grid_for(c, n, 64);
#define FCPW_SIL3(K)                                                                                          \
    hipLaunchKernelGGL((k_closest_silhouettes_coop<3, K>), g, dim3(64), 0, s, S, T, (const float4 *)spheres, flip, n, \
                       r2min, prec, (float2 *)out, order, sil_xcd_arg(c), c->sil_hold_min)
        if (c->dim == 3) {
            if (c->sil_stack == 8) FCPW_SIL3(8);
            else if (c->sil_stack == 12) FCPW_SIL3(12);
            else FCPW_SIL3(SIL_COOP_K3);
        } else
#undef FCPW_SIL3
            hipLaunchKernelGGL((k_closest_silhouettes_coop<2, 8>), g, dim3(64), 0, s, S, T, (const float4 *)spheres,
                               flip, n, r2min, prec, (float2 *)out, order, sil_xcd_arg(c), c->sil_hold_min);
    } else if (c->dim == 3)
        hipLaunchKernelGGL((k_closest_silhouettes<3, 128, 8>), dim3(blocks), dim3(128), 0, s, S, T,
                           (const float4 *)spheres, flip, n, r2min, prec, (float2 *)out, order, sil_xcd_arg(c));
    else
        hipLaunchKernelGGL((k_closest_silhouettes<2, 128, 8>), dim3(blocks), dim3(128), 0, s, S, T,
                           (const float4 *)spheres, flip, n, r2min, prec, (float2 *)out, order, sil_xcd_arg(c));
    HIPCHECK(c, hipGetLastError());
    return timing_end(c, s);
}

static int check_silhouettes(fcpw_hip_ctx *c, const char *fn)
{
    int st = check_single(c, fn);
    if (st) return st;
    if (!c->has_silhouettes)
        FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "%s: the scene was transferred without silhouettes (computeSilhouettes)", fn);
    return FCPW_HIP_OK;
}

int fcpw_hip_closest_silhouette_points_device(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres,
                                              const uint32_t *flip, int64_t n, float squared_min_radius,
                                              float precision, fcpw_hip_interaction *out, void *stream)
{
    int st = check_silhouettes(c, "closest_silhouette_points_device");
    if (st) return st;
    if (n < 0 || (n > 0 && (!spheres || !flip || !out)))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "closest_silhouette_points_device: bad arguments");
    if (((uintptr_t)spheres & 15) || ((uintptr_t)flip & 3) || ((uintptr_t)out & 7))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT,
             "closest_silhouette_points_device: spheres must be 16-byte, flags 4-byte and out 8-byte aligned");
    if (n == 0) return FCPW_HIP_OK;
    return launch_silhouettes(c, spheres, flip, n, squared_min_radius, precision, out, (hipStream_t)stream);
}

int fcpw_hip_closest_silhouette_points(fcpw_hip_ctx *c, const fcpw_hip_bounding_sphere *spheres, const uint32_t *flip,
                                       int64_t n, float squared_min_radius, float precision,
                                       fcpw_hip_interaction *out)
{
    int st = check_silhouettes(c, "closest_silhouette_points");
    if (st) return st;
    if (n < 0 || (n > 0 && (!spheres || !flip || !out)))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "closest_silhouette_points: bad arguments");
    if (n == 0) return FCPW_HIP_OK;
    HIPCHECK(c, hipSetDevice(c->device));
    const size_t sb = (size_t)n * sizeof(fcpw_hip_bounding_sphere), fb = (size_t)n * sizeof(uint32_t);
    st = grow(c, &c->stage_in, &c->stage_in_bytes, sb + fb);
    if (st) return st;
    st = grow(c, &c->stage_out, &c->stage_out_bytes, (size_t)n * sizeof(fcpw_hip_interaction));
    if (st) return st;
    if ((st = stage_copy_in(c, c->stage_in, spheres, sb, 0))) return st;
    if ((st = stage_copy_in(c, (char *)c->stage_in + sb, flip, fb, 0))) return st;
    st = launch_silhouettes(c, c->stage_in, (const uint32_t *)((char *)c->stage_in + sb), n, squared_min_radius,
                            precision, c->stage_out, 0);
    if (st) return st;
    if ((st = stage_copy_out(c, out, c->stage_out, (size_t)n * sizeof(fcpw_hip_interaction), 0))) return st;
    return FCPW_HIP_OK;
}

int fcpw_hip_closest_points_compact_device(fcpw_hip_ctx *c, const float *points, const float *r2, int64_t n,
                                           float *d, float *p, int32_t *prim, int32_t *slot, void *stream)
{
    int st = check_single(c, "closest_points_compact_device");
    if (st) return st;
    if (n < 0 || (n > 0 && !points)) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "closest_points_compact_device: bad arguments");
    if (n == 0) return FCPW_HIP_OK;
    CpIO io = {};
    io.in = points;
    io.r2 = r2;
    io.n = n;
    io.d = d;
    io.p = p;
    io.prim = prim;
    io.slot = slot;
    return launch_closest<CP_OUT_COMPACT>(c, io, (hipStream_t)stream);
}

int fcpw_hip_distance_backward_device(fcpw_hip_ctx *c, const float *points, const int32_t *slot, const float *grad_d,
                                      int64_t n, float *grad_points, float *grad_vertices, void *stream)
{
    int st = check_single(c, "distance_backward_device");
    if (st) return st;
    if (n < 0 || (n > 0 && (!points || !slot || !grad_d)))
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "distance_backward_device: bad arguments");
    if (n == 0) return FCPW_HIP_OK;
    hipStream_t s = (hipStream_t)stream;
    int threads = 256;
    unsigned blocks = (unsigned)((n + threads - 1) / threads);
    // The Morton permutation the forward pass of a batch of this size left in the
    // workspace (any permutation of [0, n) gives the same sums up to FP32
    // rounding order; this one keeps a wave's primitives and vertices together).
    // Stream order makes it the one of the latest query call on this context.
    const uint32_t *order = (c->bwd_order && c->ws_perm_n == n) ? (const uint32_t *)c->ws_perm : nullptr;
#define FCPW_BWD(D)                                                                                           \
    do {                                                                                                      \
        if (!order)                                                                                           \
            hipLaunchKernelGGL(k_distance_backward<D>, dim3(blocks), dim3(threads), 0, s, c->prims, c->prim_vids, \
                               points, slot, grad_d, n, c->n_prims, grad_points, grad_vertices);                  \
        else if (c->bwd_reduce == 2)                                                                          \
            hipLaunchKernelGGL((k_distance_backward_ordered<D, 2>), dim3(blocks), dim3(threads), 0, s, c->prims,  \
                               c->prim_vids, points, slot, grad_d, n, c->n_prims, order, grad_points, grad_vertices); \
        else if (c->bwd_reduce == 1)                                                                          \
            hipLaunchKernelGGL((k_distance_backward_ordered<D, 1>), dim3(blocks), dim3(threads), 0, s, c->prims,  \
                               c->prim_vids, points, slot, grad_d, n, c->n_prims, order, grad_points, grad_vertices); \
        else                                                                                                  \
            hipLaunchKernelGGL((k_distance_backward_ordered<D, 0>), dim3(blocks), dim3(threads), 0, s, c->prims,  \
                               c->prim_vids, points, slot, grad_d, n, c->n_prims, order, grad_points, grad_vertices); \
    } while (0)
    if (c->dim == 3) FCPW_BWD(3);
    else FCPW_BWD(2);
#undef FCPW_BWD
    HIPCHECK(c, hipGetLastError());
    return FCPW_HIP_OK;
}

int fcpw_hip_set_option(fcpw_hip_ctx *c, const char *key, int value)
{
    if (!c || !key) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    std::string k(key);
    if (k == "block") {
        if (value != 64 && value != 128 && value != 256) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "block must be 64, 128 or 256");
        c->block = value;
    } else if (k == "layout") {
        if (value != 0 && value != 1) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "layout must be 0 or 1");
        if (c->uploaded && c->instanced)
            FAIL(c, FCPW_HIP_ERR_UNSUPPORTED, "layout: not supported on a scene of several objects or instances");
        if (c->uploaded && value != c->layout) {
            int st = relayout_records(c, value);
            if (st) return st;
        }
        c->layout = value;
    } else if (k == "sil_stack" || k == "inst_stack") {
        if (k == "sil_stack" ? (value != 4 && value != 8 && value != 12) : (value != 8 && value != 12))
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "%s must be %s", key, k == "sil_stack" ? "4, 8 or 12" : "8 or 12");
        (k == "sil_stack" ? c->sil_stack : c->inst_stack) = value;
    } else if (k == "bwd_order" || k == "bwd_reduce") {
        if (value != 0 && value != 1 && !(k == "bwd_reduce" && value == 2))
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "%s must be 0 or 1%s", key, k == "bwd_reduce" ? " or 2" : "");
        (k == "bwd_order" ? c->bwd_order : c->bwd_reduce) = value;
    } else if (k == "stack") {
        if (value != 0 && value != 4 && value != 6 && value != 8 && value != 12 && value != 16 && value != 24)
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "stack must be 0, 4, 6, 8, 12, 16 or 24");
        c->stack = value;
    } else if (k == "sort") {
        if (value != 0 && value != 1) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "sort must be 0 or 1");
        c->sort = value;
    } else if (k == "coop") {
        if (value != 0 && value != 1) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "coop must be 0 or 1");
        c->coop = value;
    } else if (k == "hold_min") {
        if (value < 1 || value > 64) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "hold_min must be in [1, 64]");
        c->hold_min = value;
    } else if (k == "ray_coop") {
        if (value != 0 && value != 1) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "ray_coop must be 0 or 1");
        c->ray_coop = value;
    } else if (k == "ray_stack") {
        if (value != 4 && value != 8 && value != 16) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "ray_stack must be 4, 8 or 16");
        c->ray_stack = value;
    } else if (k == "ray_hold_min" || k == "sil_hold_min") {
        if (value < 1 || value > 64) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "%s must be in [1, 64]", key);
        (k == "ray_hold_min" ? c->ray_hold_min : c->sil_hold_min) = value;
    } else if (k == "xcds") {
        if (value < 1 || value > 64) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "xcds must be in [1, 64]");
        c->xcds = value;
    } else if (k == "sil_xcds") {
        if (value < 1 || value > 64) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "sil_xcds must be in [1, 64]");
        c->sil_xcds = value;
    } else if (k == "ray_sort_dir") {
        if (value != 0 && value != 1) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "ray_sort_dir must be 0 or 1");
        c->ray_sort_dir = value;
    } else if (k == "inst_seed") {
        if (value != 0 && (value < 64 || value > (1 << 20)))
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "inst_seed must be 0 or in [64, 2^20]");
        c->inst_seed = value;
    } else if (k == "xcd_chunk") {
        if (value < 0 || value > 16) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "xcd_chunk must be in [0, 16]");
        c->xcd_chunk = value;
    } else if (k == "morton_bits") {
        if (value != 18 && value != 21 && value != 24)
            FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "morton_bits must be 18, 21 or 24");
        c->morton_bits = value;
    } else if (k == "stats") {
        if (value != 0 && value != 1) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "stats must be 0 or 1");
        if (value) {
            HIPCHECK(c, hipSetDevice(c->device));
            HIPCHECK(c, hipDeviceSynchronize());
            HIPCHECK(c, hipMemset(c->stats, 0, 8 * sizeof(unsigned long long)));
        }
        c->stats_on = value;
    } else if (k == "timing") {
        if (value != 0 && value != 1) FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "timing must be 0 or 1");
        if (value) c->ev_used = 0;
        c->timing = value;
    } else {
        FAIL(c, FCPW_HIP_ERR_INVALID_ARGUMENT, "unknown option '%s'", key);
    }
    return FCPW_HIP_OK;
}

int fcpw_hip_get_option(const fcpw_hip_ctx *c, const char *key, int *value)
{
    if (!c || !key || !value) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    std::string k(key);
    if (k == "block") *value = c->block;
    else if (k == "stack") *value = c->stack;
    else if (k == "bwd_order") *value = c->bwd_order;
    else if (k == "sil_stack") *value = c->sil_stack;
    else if (k == "layout") *value = c->layout;
    else if (k == "inst_stack") *value = c->inst_stack;
    else if (k == "bwd_reduce") *value = c->bwd_reduce;
    else if (k == "sort") *value = c->sort;
    else if (k == "coop") *value = c->coop;
    else if (k == "xcds") *value = c->xcds;
    else if (k == "xcd_chunk") *value = c->xcd_chunk;
    else if (k == "ray_sort_dir") *value = c->ray_sort_dir;
    else if (k == "sil_xcds") *value = c->sil_xcds;
    else if (k == "inst_seed") *value = c->inst_seed;
    else if (k == "morton_bits") *value = c->morton_bits;
    else if (k == "hold_min") *value = c->hold_min;
    else if (k == "ray_coop") *value = c->ray_coop;
    else if (k == "ray_stack") *value = c->ray_stack;
    else if (k == "ray_hold_min") *value = c->ray_hold_min;
    else if (k == "sil_hold_min") *value = c->sil_hold_min;
    else if (k == "stats") *value = c->stats_on;
    else if (k == "timing") *value = c->timing;
    else return FCPW_HIP_ERR_INVALID_ARGUMENT;
    return FCPW_HIP_OK;
}

int fcpw_hip_get_info(const fcpw_hip_ctx *c, const char *key, int64_t *value)
{
    if (!c || !key || !value) return FCPW_HIP_ERR_INVALID_ARGUMENT;
    std::string k(key);
    if (k == "n_records") *value = c->n_records;
    else if (k == "n_leaves") *value = c->n_leaves;
    else if (k == "n_prims") *value = c->n_prims;
    else if (k == "n_vertices") *value = c->n_vertices;
    else if (k == "dim") *value = c->dim;
    else if (k == "max_depth") *value = c->max_depth;
    else if (k == "n_cus") *value = c->n_cus;
    else if (k == "uploaded") *value = c->uploaded ? 1 : 0;
    else if (k == "generation") *value = (int64_t)c->generation;
    else if (k == "positions_version") *value = (int64_t)c->positions_version;
    else if (k == "device") *value = c->device;
    else if (k == "n_instances") *value = c->instanced ? c->n_instances : 0;
    else if (k == "n_objects") *value = c->instanced ? (int64_t)c->inst_objects.size() : (c->uploaded ? 1 : 0);
    else if (k.rfind("stat_", 0) == 0) {
        static const char *names[7] = {"stat_internal_nodes", "stat_prim_tests", "stat_queries", "stat_wave_iterations",
                                       "stat_leaf_passes", "stat_leaf_groups", "stat_waves"};
        int which = -1;
        for (int j = 0; j < 7; j++)
            if (k == names[j]) which = j;
        if (which < 0) return FCPW_HIP_ERR_INVALID_ARGUMENT;
        fcpw_hip_ctx *m = const_cast<fcpw_hip_ctx *>(c);
        unsigned long long v[8];
        HIPCHECK(m, hipSetDevice(c->device));
        HIPCHECK(m, hipDeviceSynchronize());
        HIPCHECK(m, hipMemcpy(v, c->stats, sizeof(v), hipMemcpyDeviceToHost));
        *value = (int64_t)v[which];
    } else if (k == "kernel_launches") {
        *value = c->ev_used;
    } else if (k == "kernel_ns") {
        fcpw_hip_ctx *m = const_cast<fcpw_hip_ctx *>(c);
        double total = 0.0;
        if (c->ev_used > 0) HIPCHECK(m, hipEventSynchronize(c->ev_end[c->ev_used - 1]));
        for (int i = 0; i < c->ev_used; i++) {
            float ms = 0.0f;
            HIPCHECK(m, hipEventElapsedTime(&ms, c->ev_begin[i], c->ev_end[i]));
            total += ms;
        }
        *value = (int64_t)(total * 1e6);
    } else return FCPW_HIP_ERR_INVALID_ARGUMENT;
    return FCPW_HIP_OK;
}

} // extern "C"
#include "batch.hpp"
#include "instances.hpp"
