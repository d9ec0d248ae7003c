// context.hpp -- host-side state of one uploaded scene (fcpw_hip_ctx) and the
// error-reporting helpers shared by the entry points.  Included once, by
// fcpw_hip.hip, after device.hpp.
#pragma once
using namespace fcpw_hip;

static float bits_as_float(uint32_t u)
{
    float f;
    memcpy(&f, &u, sizeof(f));
    return f;
}

struct fcpw_hip_ctx {
    int device = 0;
    int dim = 0;
    bool uploaded = false;
    uint64_t generation = 0; // bumped by every upload (batches re-read the buffers)
    uint64_t positions_version = 0; // bumped by every upload and vertex-position update
    std::string err;
    // device buffers
    float4 *records = nullptr, *root = nullptr, *prims = nullptr;
    int32_t *prim_index = nullptr, *prim_vids = nullptr;
    float *V = nullptr;
    int32_t *rec_parent = nullptr, *leaf_refs = nullptr;
    uint32_t *leaf_links = nullptr;
    unsigned int *counters = nullptr;
    int *flags = nullptr;
    unsigned long long *stats = nullptr; // internal nodes, primitive tests, queries (option "stats")
    // silhouettes (fcpw_hip_upload_silhouettes): cones per child slot and root, leaf ranges, records
    float4 *cones = nullptr, *sils = nullptr;
    uint2 *sil_leaf = nullptr;
    bool has_silhouettes = false;
    int n_sil_refs = 0;
    int sil_leaf_max = 0; // most silhouette records in one leaf
    // scene of several objects (fcpw_hip_upload_instances): records / root /
    // prim_index then hold the tree over the instances (prim_index: leaf slot ->
    // instance); the objects are contexts of their own
    bool instanced = false;
    std::vector<fcpw_hip_ctx *> inst_objects;
    std::vector<uint64_t> inst_generations;
    DevInstance *inst = nullptr;
    DevScene *obj_table = nullptr;
    int n_instances = 0;
    int inst_obj_depth = 0; // deepest object tree
    std::vector<int32_t> node_slot;     // flat node -> child slot (record * 2 + side), root = 2 * n_records
    std::vector<int32_t> node_leaf_off; // flat node -> first leaf slot, -1 for internal nodes
    // staging for the host-pointer entry points: device buffers, and pinned host
    // bounce buffers with one event per chunk of a device-to-host copy
    void *stage_in = nullptr, *stage_out = nullptr;
    size_t stage_in_bytes = 0, stage_out_bytes = 0;
    void *pin_in = nullptr, *pin_out = nullptr;
    size_t pin_in_bytes = 0, pin_out_bytes = 0;
    std::vector<hipEvent_t> pin_ev;
    // query ordering workspace (grown on demand, kept for the context's lifetime)
    void *ws_codes = nullptr, *ws_perm = nullptr, *ws_hist = nullptr, *ws_sums = nullptr;
    size_t ws_codes_bytes = 0, ws_perm_bytes = 0, ws_hist_bytes = 0, ws_sums_bytes = 0;
    // number of queries of the batch whose Morton permutation ws_perm holds (-1: none);
    // the adjoint walks the queries in that order when its batch has the same size
    int64_t ws_perm_n = -1;
    // device address of the allocation-free entry point's scratch array on this
    // context's device (shared by every context on the device), or null with the
    // reason in scratch_err
    void *scratch = nullptr;
    std::string scratch_err;
    int bwd_order = 1, bwd_reduce = 2; // adjoint: the forward's query order, per-wave vertex sums in LDS
    float qlo[3] = {0, 0, 0}, qinv[3] = {1, 1, 1}; // Morton grid: 1.5x the scene box
    // facts
    int n_records = 0, n_leaves = 0, n_prims = 0, n_vertices = 0, max_depth = 0;
    // options
    int block = 128, stack = 6, sort = 1, xcds = 4, coop = 1, hold_min = 1, morton_bits = 24;
    int xcd_chunk = 0; // log2 tiles per XCD chunk of the tile mapping (0: one range per XCD)
    int ray_coop = 1; // wave-cooperative leaves for ray queries
    int ray_sort_dir = 1; // rays ordered by direction octant, then origin
    int ray_stack = 4, ray_hold_min = 8; // the cooperative ray kernels' LDS stack entries and leaf-pass threshold
    int sil_stack = 4, inst_stack = 12;  // LDS stack entries of the 3D cooperative silhouette and instance kernels
    int layout = 0;                      // record order in memory: 0 DFS, 1 treelets of three levels
    int sil_hold_min = 3;                // the cooperative silhouette kernel's leaf-pass threshold
    int sil_xcds = 8;                    // the silhouette kernels' XCD grouping
    int inst_seed = 1024;                // instanced closest points: ordered queries per representative wave (0: off)
    int stats_on = 0, timing = 0;
    int n_cus = 256;
    // device scratch of fcpw_hip_build_bvh, kept (grow-only) for the next build
    static constexpr int BUILD_SLOTS = 48;
    void *build_buf[BUILD_SLOTS] = {};
    size_t build_bytes[BUILD_SLOTS] = {};
    // HIP event pairs around each traversal launch while "timing" is on
    static constexpr int MAX_TIMED = 1024;
    std::vector<hipEvent_t> ev_begin, ev_end;
    int ev_used = 0;
};

static std::string g_create_error;

#define FAIL(ctx, code, ...)                                                                                   \
    do {                                                                                                       \
        char _b[512];                                                                                          \
        snprintf(_b, sizeof(_b), __VA_ARGS__);                                                                 \
        (ctx)->err = _b;                                                                                       \
        return (code);                                                                                         \
    } while (0)

#define HIPCHECK(ctx, expr)                                                                                    \
    do {                                                                                                       \
        hipError_t _e = (expr);                                                                                \
        if (_e != hipSuccess) {                                                                                \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(_e);                                   \
            return _e == hipErrorOutOfMemory ? FCPW_HIP_ERR_OUT_OF_MEMORY : FCPW_HIP_ERR_HIP;                 \
        }                                                                                                      \
    } while (0)

static void free_silhouettes(fcpw_hip_ctx *c)
{
    void *bufs[] = {c->cones, c->sils, c->sil_leaf};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    c->cones = c->sils = nullptr;
    c->sil_leaf = nullptr;
    c->has_silhouettes = false;
    c->n_sil_refs = 0;
    c->sil_leaf_max = 0;
}

static void free_scene(fcpw_hip_ctx *c)
{
    void *bufs[] = {c->records, c->root, c->prims, c->prim_index, c->prim_vids, c->V,
                    c->rec_parent, c->leaf_refs, c->leaf_links, c->counters, c->inst, c->obj_table};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    free_silhouettes(c);
    c->inst = nullptr;
    c->obj_table = nullptr;
    c->instanced = false;
    c->inst_objects.clear();
    c->inst_generations.clear();
    c->n_instances = 0;
    c->records = c->root = c->prims = nullptr;
    c->prim_index = c->prim_vids = nullptr;
    c->V = nullptr;
    c->rec_parent = c->leaf_refs = nullptr;
    c->leaf_links = nullptr;
    c->counters = nullptr;
    c->uploaded = false;
}

// the tile-mapping argument of the traversal kernels (see xcd_tile)
static inline int xcd_arg(const fcpw_hip_ctx *c) { return c->xcds | (c->xcd_chunk << 8); }
static inline int sil_xcd_arg(const fcpw_hip_ctx *c) { return c->sil_xcds | (c->xcd_chunk << 8); }

static DevScene dev_scene(fcpw_hip_ctx *c)
{
    return DevScene{c->records, c->root, c->prims, c->prim_index, c->flags, c->stats};
}

static int grow(fcpw_hip_ctx *c, void **buf, size_t *cap, size_t bytes)
{
    if (*cap >= bytes) return FCPW_HIP_OK;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    HIPCHECK(c, hipMalloc(buf, bytes));
    *cap = bytes;
    return FCPW_HIP_OK;
}

// ---------------------------------------------------------------------------
// Host <-> device copies of query and result arrays for the host-pointer entry
// points.  Caller memory is pageable, so a copy goes through pinned bounce
// buffers of the context (hipHostMalloc, grown on demand): the array is cut
// into chunks, host threads copy the chunks between the caller's memory and
// the bounce buffer, and each chunk's DMA overlaps the host copies of the
// others.  Small arrays take a plain hipMemcpy.
// ---------------------------------------------------------------------------
// 8 MiB: 4 MiB pieces left the copy engine idle between pieces (73.7-74.2 ms
// against 72.5-72.6 for 16M closest points through the host API; 16 MiB: 72.7-73.2)
constexpr size_t STAGE_CHUNK = (size_t)8 << 20;

// host threads for the bounce copies: the CPUs this process may run on, at most 16
static int stage_threads()
{
    static const int n = [] {
        int k = 0;
        cpu_set_t set;
        if (sched_getaffinity(0, sizeof(set), &set) == 0) k = CPU_COUNT(&set);
        if (k < 1) k = (int)std::thread::hardware_concurrency();
        return std::max(1, std::min(k, 16));
    }();
    return n;
}

static int pin_grow(fcpw_hip_ctx *c, void **buf, size_t *cap, size_t bytes)
{
    if (*cap >= bytes) return FCPW_HIP_OK;
    if (*buf) (void)hipHostFree(*buf);
    *buf = nullptr;
    *cap = 0;
    HIPCHECK(c, hipHostMalloc(buf, bytes, hipHostMallocDefault));
    *cap = bytes;
    return FCPW_HIP_OK;
}

// runs body(k) for chunks k = 0 .. n_chunks - 1 on the stage threads (chunk k on
// thread k mod threads, in increasing k); returns the first nonzero status
template <typename F>
static int stage_parallel(int64_t n_chunks, F &&body)
{
    const int nt = (int)std::min<int64_t>(stage_threads(), n_chunks);
    std::atomic<int> status{0};
    auto work = [&](int t) {
        for (int64_t k = t; k < n_chunks; k += nt) {
            const int st = body(k);
            if (st) {
                int expected = 0;
                status.compare_exchange_strong(expected, st);
                return;
            }
        }
    };
    std::vector<std::thread> th;
    int started = 1;
    try {
        for (; started < nt; started++) th.emplace_back(work, started);
    } catch (...) {
        // no more threads: this thread also takes the shares that found none
    }
    for (int t = started; t < nt; t++) work(t);
    work(0);
    for (auto &t : th) t.join();
    return status.load();
}

// host src -> device dst, stream-ordered on s before whatever is launched after it
static int stage_copy_in(fcpw_hip_ctx *c, void *dst, const void *src, size_t bytes, hipStream_t s)
{
    if (bytes <= STAGE_CHUNK) {
        HIPCHECK(c, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
        return FCPW_HIP_OK;
    }
    // the bounce buffer may still feed the DMAs of an earlier copy queued on s
    HIPCHECK(c, hipStreamSynchronize(s));
    int st = pin_grow(c, &c->pin_in, &c->pin_in_bytes, bytes);
    if (st) return st;
    const int64_t n_chunks = (int64_t)((bytes + STAGE_CHUNK - 1) / STAGE_CHUNK);
    char *pin = (char *)c->pin_in;
    st = stage_parallel(n_chunks, [&](int64_t k) {
        const size_t off = (size_t)k * STAGE_CHUNK, len = std::min(STAGE_CHUNK, bytes - off);
        memcpy(pin + off, (const char *)src + off, len);
        return hipMemcpyAsync((char *)dst + off, pin + off, len, hipMemcpyHostToDevice, s) == hipSuccess
                   ? FCPW_HIP_OK
                   : FCPW_HIP_ERR_HIP;
    });
    if (st) FAIL(c, st, "host-to-device copy of %zu bytes failed", bytes);
    return FCPW_HIP_OK;
}

// device src -> host dst after the work already queued on s; returns when dst holds it
static int stage_copy_out(fcpw_hip_ctx *c, void *dst, const void *src, size_t bytes, hipStream_t s)
{
    if (bytes <= STAGE_CHUNK) {
        HIPCHECK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        HIPCHECK(c, hipStreamSynchronize(s));
        return FCPW_HIP_OK;
    }
    int st = pin_grow(c, &c->pin_out, &c->pin_out_bytes, bytes);
    if (st) return st;
    const int64_t n_chunks = (int64_t)((bytes + STAGE_CHUNK - 1) / STAGE_CHUNK);
    while ((int64_t)c->pin_ev.size() < n_chunks) {
        hipEvent_t e = nullptr;
        HIPCHECK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->pin_ev.push_back(e);
    }
    char *pin = (char *)c->pin_out;
    for (int64_t k = 0; k < n_chunks; k++) {
        const size_t off = (size_t)k * STAGE_CHUNK, len = std::min(STAGE_CHUNK, bytes - off);
        HIPCHECK(c, hipMemcpyAsync(pin + off, (const char *)src + off, len, hipMemcpyDeviceToHost, s));
        HIPCHECK(c, hipEventRecord(c->pin_ev[k], s));
    }
    st = stage_parallel(n_chunks, [&](int64_t k) {
        const size_t off = (size_t)k * STAGE_CHUNK, len = std::min(STAGE_CHUNK, bytes - off);
        if (hipEventSynchronize(c->pin_ev[k]) != hipSuccess) return (int)FCPW_HIP_ERR_HIP;
        memcpy((char *)dst + off, pin + off, len);
        return (int)FCPW_HIP_OK;
    });
    if (st) FAIL(c, st, "device-to-host copy of %zu bytes failed", bytes);
    return FCPW_HIP_OK;
}

