#pragma once
// gpu_build.hpp -- the reference's binned SAH BVH build (bvh.inl:3-234), run
// on the device level by level, producing the reference's tree: the same
// nodes in the same DFS order with the same boxes, and each leaf's primitives
// in the same order.  Included once, by fcpw_hip.hip, after context.hpp.
//
// Per level, every node of the level is processed at once:
//   bounds     node box and centroid box: min/max atomics on order-preserving
//              integer images of the floats (exact, order independent);
//   bins       8 buckets per axis (box + count), the same bucket arithmetic;
//   split      one thread per node: the SAH sweep of computeObjectSplit, in the
//              reference's evaluation order, or the longest-axis fallback;
//   partition  the reference's in-place partition (performObjectSplit) keeps
//              the "left" references in order and cycles the "right" block
//              like a queue: a left reference met after k right ones moves the
//              right block's first entry to its end.  That sequence is written
//              as a stream (appends and copies of earlier entries) whose
//              positions come from two exclusive scans; pointer jumping
//              resolves the copies, so the final order equals the sequential
//              swap sequence's;
//   children   two per internal node, numbered in level order by a scan.
// Then subtree sizes (bottom-up) give every node its DFS index, and the
// traversal records, refit schedule and primitive records are written in
// place, as fcpw_hip_upload_bvh would have written them from the host tree.

namespace fcpw_hip {
namespace gbuild {

constexpr int NB = 8;             // buckets (nBuckets, fcpw.inl:805-840)
constexpr int LEAF = 4;           // leafSize of a non-vectorized BVH (fcpw.inl:812-819)
constexpr int DEPTH_CAP = 62;     // FCPW_BVH_MAX_DEPTH - 2 (bvh.inl:170)
constexpr int BIN_WORDS = 7;      // lo.xyz, hi.xyz (ordered ints), count
constexpr int BIN_NODE = 3 * NB * BIN_WORDS;

enum Heuristic { LONGEST_AXIS_CENTER = 0, SURFACE_AREA = 1, OVERLAP_SURFACE_AREA = 2, VOLUME = 3, OVERLAP_VOLUME = 4 };

// order-preserving image of a float in an int (for atomic min / max)
__device__ __forceinline__ int f2o(float f)
{
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

struct Build {
    // primitives (by original index)
    const float4 *pbox_lo, *pbox_hi, *pcen;
    // references in the current order and the next order
    int32_t *ref, *ref2;
    int32_t *elem_node;        // per position: node of the current level, or -1
    uint32_t *scan_r, *scan_k; // n + 1 entries each
    int32_t *st_val[2], *st_nxt[2];
    // nodes, in level order
    int32_t *nd_start, *nd_n, *nd_depth, *nd_left;
    int *nd_box;               // 6 ordered ints per node: box lo, hi
    int *nd_cbox;              // 6 ordered ints per node: centroid box lo, hi
    int32_t *nd_dim;           // split axis (-1: leaf)
    float *nd_coord;
    int32_t *nd_mid, *nd_rot, *nd_right, *nd_flags; // flags bit 0: order kept (fallback split)
    int *bins;                 // BIN_NODE ints per node of the level
    uint32_t *internal_scan;   // per node of the level (+1)
    int *level_info;           // [0] max stream length of the level
    int32_t *nd_size, *nd_leaves, *nd_idx, *nd_lb;
    int n;
    int heuristic;
};

template <int DIM>
__global__ void k_prim_data(const float *V, const int32_t *idx, int n, float4 *lo, float4 *hi, float4 *cen)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[DIM][3];
    for (int k = 0; k < DIM; k++)
        for (int d = 0; d < 3; d++) v[k][d] = d < DIM ? V[(size_t)idx[(size_t)i * DIM + k] * DIM + d] : 0.0f;
    float l[3] = {0.0f, 0.0f, 0.0f}, h[3] = {0.0f, 0.0f, 0.0f}, c[3] = {0.0f, 0.0f, 0.0f};
    for (int d = 0; d < DIM; d++) {
        // expandToInclude(vertex) with the FLT_EPSILON inflation (bounding_volumes.h:45-49)
        float a = FLT_MAX, b = -FLT_MAX;
        for (int k = 0; k < DIM; k++) {
            const float lo_k = v[k][d] - FLT_EPSILON, hi_k = v[k][d] + FLT_EPSILON;
            a = (lo_k < a) ? lo_k : a;
            b = (b < hi_k) ? hi_k : b;
        }
        l[d] = a;
        h[d] = b;
        // Triangle::centroid (triangles.inl:39-46), LineSegment::centroid (line_segments.inl:38-44)
        c[d] = DIM == 3 ? ((v[0][d] + v[1][d]) + v[DIM - 1][d]) / 3.0f : (v[0][d] + v[1][d]) * 0.5f;
    }
    lo[i] = make_float4(l[0], l[1], l[2], 0.0f);
    hi[i] = make_float4(h[0], h[1], h[2], 0.0f);
    cen[i] = make_float4(c[0], c[1], c[2], 0.0f);
}

__global__ void k_iota(int32_t *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// empty boxes and bins for the level's nodes [g0, g0 + m)
__global__ void k_level_init(Build B, int g0, int m)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int g = g0 + j;
    for (int d = 0; d < 3; d++) {
        B.nd_box[g * 6 + d] = f2o(FLT_MAX);
        B.nd_box[g * 6 + 3 + d] = f2o(-FLT_MAX);
        B.nd_cbox[g * 6 + d] = f2o(FLT_MAX);
        B.nd_cbox[g * 6 + 3 + d] = f2o(-FLT_MAX);
    }
    int *b = B.bins + (size_t)j * BIN_NODE;
    for (int k = 0; k < 3 * NB; k++) {
        for (int d = 0; d < 3; d++) {
            b[k * BIN_WORDS + d] = f2o(FLT_MAX);
            b[k * BIN_WORDS + 3 + d] = f2o(-FLT_MAX);
        }
        b[k * BIN_WORDS + 6] = 0;
    }
    B.nd_dim[g] = -1;
    B.nd_flags[g] = 0;
}

// node of the level owning each position (level nodes are sorted by start and disjoint)
__global__ void k_elem_node(Build B, int g0, int m)
{
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= B.n) return;
    int lo = 0, hi = m - 1, found = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (B.nd_start[g0 + mid] <= pos) {
            found = mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    int g = -1;
    if (found >= 0) {
        const int c = g0 + found;
        if (pos < B.nd_start[c] + B.nd_n[c]) g = c;
    }
    B.elem_node[pos] = g;
}

__device__ __forceinline__ bool needs_split(const Build &B, int g)
{
    return B.nd_n[g] > LEAF && B.nd_depth[g] != DEPTH_CAP;
}

// nodes of at most SMALL references are bounded, binned and split by one thread
// (k_small); larger ones by the element-parallel kernels below
constexpr int SMALL = 64;
__device__ __forceinline__ bool is_small(const Build &B, int g) { return B.nd_n[g] <= SMALL; }

// Element-parallel accumulation for the large nodes.  The nodes a block touches
// are consecutive in level order (node ranges are sorted and contiguous); when
// they span fewer than SLOTS node ids the block accumulates in LDS, one slot per
// node, and flushes each used slot with one set of global atomics.  Otherwise
// every element goes to global memory directly.
constexpr int ACC_THREADS = 256;
constexpr int SLOTS = 8;

struct BlockNodes {
    int gmin, gmax;
};

__device__ __forceinline__ void block_node_range(const Build &B, int g, bool valid, BlockNodes &r, int *smin,
                                                 int *smax)
{
    if (threadIdx.x == 0) {
        *smin = INT_MAX;
        *smax = -1;
    }
    __syncthreads();
    if (valid) {
        atomicMin(smin, g);
        atomicMax(smax, g);
    }
    __syncthreads();
    r.gmin = *smin;
    r.gmax = *smax;
}

template <int DIM>
__global__ __launch_bounds__(ACC_THREADS) void k_bounds(Build B)
{
    __shared__ int acc[SLOTS][12];
    __shared__ int used[SLOTS];
    __shared__ int smin, smax;
    const int pos = blockIdx.x * ACC_THREADS + threadIdx.x;
    for (int k = threadIdx.x; k < SLOTS * 12; k += ACC_THREADS) {
        const int w = k % 12;
        acc[k / 12][w] = (w / 3) % 2 == 0 ? f2o(FLT_MAX) : f2o(-FLT_MAX);
    }
    if (threadIdx.x < SLOTS) used[threadIdx.x] = 0;
    const int g = pos < B.n ? B.elem_node[pos] : -1;
    const bool valid = g >= 0 && !is_small(B, g);
    BlockNodes r;
    block_node_range(B, g, valid, r, &smin, &smax);
    const bool lds = r.gmax >= 0 && r.gmax - r.gmin < SLOTS;
    if (valid) {
        const int p = B.ref[pos];
        const float4 l = B.pbox_lo[p], h = B.pbox_hi[p], c = B.pcen[p];
        const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z}, cv[3] = {c.x, c.y, c.z};
        if (lds) used[g - r.gmin] = 1;
        for (int d = 0; d < DIM; d++) {
            // the centroid box grows by the inflated point (expandToInclude(point))
            const int a = f2o(lv[d]), b = f2o(hv[d]), ca = f2o(cv[d] - FLT_EPSILON), cb = f2o(cv[d] + FLT_EPSILON);
            if (lds) {
                int *q = acc[g - r.gmin];
                atomicMin(&q[d], a); atomicMax(&q[3 + d], b);
                atomicMin(&q[6 + d], ca); atomicMax(&q[9 + d], cb);
            } else {
                atomicMin(&B.nd_box[g * 6 + d], a); atomicMax(&B.nd_box[g * 6 + 3 + d], b);
                atomicMin(&B.nd_cbox[g * 6 + d], ca); atomicMax(&B.nd_cbox[g * 6 + 3 + d], cb);
            }
        }
    }
    __syncthreads();
    if (lds) {
        for (int k = threadIdx.x; k < SLOTS * 12; k += ACC_THREADS) {
            const int slot = k / 12, w = k % 12, d = w % 3;
            if (!used[slot] || d >= DIM) continue;
            const int node = r.gmin + slot;
            int *dst = w < 6 ? &B.nd_box[node * 6 + w] : &B.nd_cbox[node * 6 + (w - 6)];
            if ((w / 3) % 2 == 0) atomicMin(dst, acc[slot][w]);
            else atomicMax(dst, acc[slot][w]);
        }
    }
}

// computeObjectSplit's bucket index (bvh.inl:59-64)
__device__ __forceinline__ int bucket_of(float c, float lo, float width)
{
    int k = (int)((c - lo) / width);
    return k < 0 ? 0 : (k > NB - 1 ? NB - 1 : k);
}

template <int DIM>
__global__ __launch_bounds__(ACC_THREADS) void k_bins(Build B, int g0)
{
    __shared__ int acc[SLOTS * BIN_NODE];
    __shared__ int used[SLOTS];
    __shared__ int smin, smax;
    const int pos = blockIdx.x * ACC_THREADS + threadIdx.x;
    for (int k = threadIdx.x; k < SLOTS * 3 * NB; k += ACC_THREADS) {
        for (int d = 0; d < 3; d++) {
            acc[k * BIN_WORDS + d] = f2o(FLT_MAX);
            acc[k * BIN_WORDS + 3 + d] = f2o(-FLT_MAX);
        }
        acc[k * BIN_WORDS + 6] = 0;
    }
    if (threadIdx.x < SLOTS) used[threadIdx.x] = 0;
    const int g = pos < B.n ? B.elem_node[pos] : -1;
    const bool valid = B.heuristic != LONGEST_AXIS_CENTER && g >= 0 && needs_split(B, g) && !is_small(B, g);
    BlockNodes r;
    block_node_range(B, g, valid, r, &smin, &smax);
    const bool lds = r.gmax >= 0 && r.gmax - r.gmin < SLOTS;
    if (valid) {
        const int p = B.ref[pos];
        const float4 l = B.pbox_lo[p], h = B.pbox_hi[p], c = B.pcen[p];
        const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z}, cv[3] = {c.x, c.y, c.z};
        int *dst = lds ? acc + (g - r.gmin) * BIN_NODE : B.bins + (size_t)(g - g0) * BIN_NODE;
        if (lds) used[g - r.gmin] = 1;
        for (int d = 0; d < DIM; d++) {
            const float nlo = o2f(B.nd_box[g * 6 + d]), nhi = o2f(B.nd_box[g * 6 + 3 + d]);
            const float extent = nhi - nlo;
            if (extent < 1e-6f) continue;
            const float width = extent / (float)NB;
            int *bin = dst + (d * NB + bucket_of(cv[d], nlo, width)) * BIN_WORDS;
            for (int e = 0; e < DIM; e++) {
                atomicMin(&bin[e], f2o(lv[e]));
                atomicMax(&bin[3 + e], f2o(hv[e]));
            }
            atomicAdd(&bin[6], 1);
        }
    }
    __syncthreads();
    if (lds) {
        for (int k = threadIdx.x; k < SLOTS * 3 * NB; k += ACC_THREADS) {
            const int slot = k / (3 * NB);
            const int *a = acc + k * BIN_WORDS;
            if (!used[slot] || a[6] == 0) continue;
            int *dst = B.bins + (size_t)(r.gmin + slot - g0) * BIN_NODE + (k % (3 * NB)) * BIN_WORDS;
            for (int e = 0; e < DIM; e++) {
                atomicMin(&dst[e], a[e]);
                atomicMax(&dst[3 + e], a[3 + e]);
            }
            atomicAdd(&dst[6], a[6]);
        }
    }
}

// ---- the reference's split cost (bvh.inl:3-37) with the Eigen reductions ----
template <int DIM>
__device__ __forceinline__ float red_sum(const float *c) { return DIM == 3 ? c[0] + (c[1] + c[2]) : c[0] + c[1]; }
template <int DIM>
__device__ __forceinline__ float red_prod(const float *c) { return DIM == 3 ? c[0] * (c[1] * c[2]) : c[0] * c[1]; }
__device__ __forceinline__ float smax_(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float smin_(float a, float b) { return (b < a) ? b : a; }

template <int DIM>
__device__ __forceinline__ float surface_area(const float *lo, const float *hi)
{
    float e[3], q[3];
    for (int i = 0; i < DIM; i++) e[i] = smax_(hi[i] - lo[i], 1e-5f);
    const float P = red_prod<DIM>(e);
    for (int i = 0; i < DIM; i++) q[i] = P / e[i];
    return 2.0f * red_sum<DIM>(q);
}
template <int DIM>
__device__ __forceinline__ float volume(const float *lo, const float *hi)
{
    float e[3];
    for (int i = 0; i < DIM; i++) e[i] = hi[i] - lo[i];
    return red_prod<DIM>(e);
}
template <int DIM>
__device__ __forceinline__ float split_cost(int h, const float *llo, const float *lhi, const float *rlo,
                                            const float *rhi, int nl, int nr)
{
    if (h == SURFACE_AREA) return (float)nl * surface_area<DIM>(llo, lhi) + (float)nr * surface_area<DIM>(rlo, rhi);
    if (h == VOLUME) return (float)nl * volume<DIM>(llo, lhi) + (float)nr * volume<DIM>(rlo, rhi);
    float ilo[3], ihi[3];
    bool valid = true;
    for (int i = 0; i < DIM; i++) {
        ilo[i] = smax_(llo[i], rlo[i]);
        ihi[i] = smin_(lhi[i], rhi[i]);
        if (!(ihi[i] >= ilo[i])) valid = false;
    }
    float c;
    if (h == OVERLAP_SURFACE_AREA)
        c = ((float)nl / surface_area<DIM>(rlo, rhi) + (float)nr / surface_area<DIM>(llo, lhi)) *
            fabsf(surface_area<DIM>(ilo, ihi));
    else
        c = ((float)nl / volume<DIM>(rlo, rhi) + (float)nr / volume<DIM>(llo, lhi)) * fabsf(volume<DIM>(ilo, ihi));
    return valid ? c : -c;
}

// the bucket sweep of computeObjectSplit for one axis (bvh.inl:66-100): bucket
// boxes blo/bhi (empty buckets hold the empty box) and counts bcnt
template <int DIM>
__device__ __forceinline__ void sweep_axis(int h, int d, float lo_d, float width, const float (*blo)[3],
                                           const float (*bhi)[3], const int *bcnt, float &best, int &dim,
                                           float &coord)
{
    float rlo[NB][3], rhi[NB][3];
    int rcnt[NB];
    float alo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, ahi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int k = NB - 1; k > 0; k--) {
        for (int e = 0; e < DIM; e++) {
            alo[e] = smin_(alo[e], blo[k][e]);
            ahi[e] = smax_(ahi[e], bhi[k][e]);
            rlo[k][e] = alo[e];
            rhi[k][e] = ahi[e];
        }
        rcnt[k] = bcnt[k] + (k != NB - 1 ? rcnt[k + 1] : 0);
    }
    float llo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    int nl = 0;
    for (int k = 1; k < NB; k++) {
        for (int e = 0; e < DIM; e++) {
            llo[e] = smin_(llo[e], blo[k - 1][e]);
            lhi[e] = smax_(lhi[e], bhi[k - 1][e]);
        }
        nl += bcnt[k - 1];
        if (nl > 0 && rcnt[k] > 0) {
            const float c = split_cost<DIM>(h, llo, lhi, rlo[k], rhi[k], nl, rcnt[k]);
            if (c < best) {
                best = c;
                dim = d;
                coord = lo_d + (float)k * width;
            }
        }
    }
}

// LongestAxisCenter fallback on the centroid box (bvh.inl:104-108)
template <int DIM>
__device__ __forceinline__ void fallback_split(const float *clo, const float *chi, int &dim, float &coord)
{
    int md = 0;
    float me = chi[0] - clo[0];
    for (int d = 1; d < DIM; d++) {
        const float e = chi[d] - clo[d];
        if (e > me) {
            me = e;
            md = d;
        }
    }
    dim = md;
    coord = (clo[md] + chi[md]) * 0.5f;
}

// computeObjectSplit (bvh.inl:39-111) for every large node of the level that splits
template <int DIM>
__global__ void k_split(Build B, int g0, int m)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int g = g0 + j;
    if (!needs_split(B, g) || is_small(B, g)) return;
    float lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = o2f(B.nd_box[g * 6 + d]);
        hi[d] = o2f(B.nd_box[g * 6 + 3 + d]);
    }
    float best = FLT_MAX, coord = 0.0f;
    int dim = -1;
    if (B.heuristic != LONGEST_AXIS_CENTER) {
        const int *bins = B.bins + (size_t)j * BIN_NODE;
        for (int d = 0; d < DIM; d++) {
            const float extent = hi[d] - lo[d];
            if (extent < 1e-6f) continue;
            const float width = extent / (float)NB;
            float blo[NB][3], bhi[NB][3];
            int bcnt[NB];
            for (int k = 0; k < NB; k++) {
                const int *b = bins + (d * NB + k) * BIN_WORDS;
                for (int e = 0; e < 3; e++) {
                    blo[k][e] = o2f(b[e]);
                    bhi[k][e] = o2f(b[3 + e]);
                }
                bcnt[k] = b[6];
            }
            sweep_axis<DIM>(B.heuristic, d, lo[d], width, blo, bhi, bcnt, best, dim, coord);
        }
    }
    if (dim == -1) {
        float clo[3], chi[3];
        for (int d = 0; d < 3; d++) {
            clo[d] = o2f(B.nd_cbox[g * 6 + d]);
            chi[d] = o2f(B.nd_cbox[g * 6 + 3 + d]);
        }
        fallback_split<DIM>(clo, chi, dim, coord);
    }
    B.nd_dim[g] = dim;
    B.nd_coord[g] = coord;
}

// small nodes: node box, centroid box, bins and split by one thread, growing
// the boxes element by element in range order as the reference does
template <int DIM>
__global__ void k_small(Build B, int g0, int m)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int g = g0 + j;
    if (!is_small(B, g)) return;
    const int s = B.nd_start[g], n = B.nd_n[g];
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = s; i < s + n; i++) {
        const int p = B.ref[i];
        const float4 l = B.pbox_lo[p], h = B.pbox_hi[p], c = B.pcen[p];
        const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z}, cv[3] = {c.x, c.y, c.z};
        for (int d = 0; d < DIM; d++) {
            lo[d] = smin_(lo[d], lv[d]);
            hi[d] = smax_(hi[d], hv[d]);
            clo[d] = smin_(clo[d], cv[d] - FLT_EPSILON);
            chi[d] = smax_(chi[d], cv[d] + FLT_EPSILON);
        }
    }
    for (int d = 0; d < 3; d++) {
        B.nd_box[g * 6 + d] = f2o(lo[d]);
        B.nd_box[g * 6 + 3 + d] = f2o(hi[d]);
        B.nd_cbox[g * 6 + d] = f2o(clo[d]);
        B.nd_cbox[g * 6 + 3 + d] = f2o(chi[d]);
    }
    if (!needs_split(B, g)) return;
    float best = FLT_MAX, coord = 0.0f;
    int dim = -1;
    if (B.heuristic != LONGEST_AXIS_CENTER) {
        for (int d = 0; d < DIM; d++) {
            const float extent = hi[d] - lo[d];
            if (extent < 1e-6f) continue;
            const float width = extent / (float)NB;
            float blo[NB][3], bhi[NB][3];
            int bcnt[NB];
            for (int k = 0; k < NB; k++) {
                for (int e = 0; e < 3; e++) {
                    blo[k][e] = FLT_MAX;
                    bhi[k][e] = -FLT_MAX;
                }
                bcnt[k] = 0;
            }
            for (int i = s; i < s + n; i++) {
                const int p = B.ref[i];
                const float4 l = B.pbox_lo[p], h = B.pbox_hi[p], c = B.pcen[p];
                const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z}, cv[3] = {c.x, c.y, c.z};
                const int k = bucket_of(cv[d], lo[d], width);
                for (int e = 0; e < DIM; e++) {
                    blo[k][e] = smin_(blo[k][e], lv[e]);
                    bhi[k][e] = smax_(bhi[k][e], hv[e]);
                }
                bcnt[k] += 1;
            }
            sweep_axis<DIM>(B.heuristic, d, lo[d], width, blo, bhi, bcnt, best, dim, coord);
        }
    }
    if (dim == -1) fallback_split<DIM>(clo, chi, dim, coord);
    B.nd_dim[g] = dim;
    B.nd_coord[g] = coord;
}

__device__ __forceinline__ bool goes_left(const Build &B, int g, int pos)
{
    const float4 c = B.pcen[B.ref[pos]];
    const int d = B.nd_dim[g];
    const float v = d == 0 ? c.x : (d == 1 ? c.y : c.z);
    return v < B.nd_coord[g];
}

// right flags (scanned into the number of right references before each position)
__global__ void k_flag_right(Build B)
{
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos > B.n) return;
    uint32_t r = 0u;
    if (pos < B.n) {
        const int g = B.elem_node[pos];
        if (g >= 0 && B.nd_dim[g] >= 0) r = goes_left(B, g, pos) ? 0u : 1u;
    }
    B.scan_r[pos] = r;
}

// a left reference met after at least one right one rotates the right block
__global__ void k_flag_rotate(Build B)
{
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos > B.n) return;
    uint32_t k = 0u;
    if (pos < B.n) {
        const int g = B.elem_node[pos];
        if (g >= 0 && B.nd_dim[g] >= 0 && goes_left(B, g, pos))
            k = (B.scan_r[pos] - B.scan_r[B.nd_start[g]]) > 0u ? 1u : 0u;
    }
    B.scan_k[pos] = k;
}

// per node: split position, stream length; fallback split keeps the order
__global__ void k_node_part(Build B, int g0, int m)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > m) return;
    if (j == m) { // one past the last node: the scan's total slot
        B.internal_scan[m] = 0u;
        return;
    }
    const int g = g0 + j;
    uint32_t internal = 0u;
    if (B.nd_dim[g] >= 0) {
        const int s = B.nd_start[g], n = B.nd_n[g], e = s + n;
        const int right = (int)(B.scan_r[e] - B.scan_r[s]);
        const int rot = (int)(B.scan_k[e] - B.scan_k[s]);
        const int left = n - right;
        int mid = s + left, flags = 0;
        if (left == 0 || right == 0) { // performObjectSplit's "bad split" (bvh.inl:127-136)
            mid = s + n / 2;
            flags = 1;
        } else {
            atomicMax(&B.level_info[0], right + rot);
        }
        B.nd_mid[g] = mid;
        B.nd_rot[g] = rot;
        B.nd_right[g] = right;
        B.nd_flags[g] = flags;
        internal = 1u;
    }
    B.internal_scan[j] = internal;
}

// left references to their final places; right references and rotations into the stream
__global__ void k_stream(Build B)
{
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= B.n) return;
    const int g = B.elem_node[pos];
    const int p = B.ref[pos];
    if (g < 0 || B.nd_dim[g] < 0 || (B.nd_flags[g] & 1)) {
        B.ref2[pos] = p;
        return;
    }
    const int s = B.nd_start[g];
    const int R = (int)(B.scan_r[pos] - B.scan_r[s]);
    const int K = (int)(B.scan_k[pos] - B.scan_k[s]);
    if (goes_left(B, g, pos)) {
        B.ref2[s + (pos - s - R)] = p;
        if (R > 0) {
            B.st_val[0][s + R + K] = -1;
            B.st_nxt[0][s + R + K] = s + K;
        }
    } else {
        B.st_val[0][s + R + K] = p;
        B.st_nxt[0][s + R + K] = -1;
    }
}

// one pointer-jumping round over the stream entries of the level's splitting nodes
__global__ void k_jump(Build B, int in)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= B.n) return;
    const int g = B.elem_node[q];
    if (g < 0 || B.nd_dim[g] < 0 || (B.nd_flags[g] & 1)) return;
    const int s = B.nd_start[g];
    if (q - s >= B.nd_right[g] + B.nd_rot[g]) return;
    const int32_t *vi = B.st_val[in], *ni = B.st_nxt[in];
    int32_t *vo = B.st_val[in ^ 1], *no = B.st_nxt[in ^ 1];
    const int x = ni[q];
    if (x < 0) {
        vo[q] = vi[q];
        no[q] = -1;
    } else if (ni[x] < 0) {
        vo[q] = vi[x];
        no[q] = -1;
    } else {
        vo[q] = vi[q];
        no[q] = ni[x];
    }
}

// the right block: entries [rot, rot + right) of the resolved stream
__global__ void k_place_right(Build B, int in)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= B.n) return;
    const int g = B.elem_node[q];
    if (g < 0 || B.nd_dim[g] < 0 || (B.nd_flags[g] & 1)) return;
    const int s = B.nd_start[g], rot = B.nd_rot[g], right = B.nd_right[g];
    const int k = q - s - rot;
    if (k < 0 || k >= right) return;
    B.ref2[B.nd_mid[g] + k] = B.st_val[in][q];
}

// children of the level's internal nodes, numbered in level order
__global__ void k_children(Build B, int g0, int m, int next)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int g = g0 + j;
    if (B.nd_dim[g] < 0) {
        B.nd_left[g] = -1;
        return;
    }
    const int c = next + 2 * (int)B.internal_scan[j];
    const int s = B.nd_start[g], n = B.nd_n[g], mid = B.nd_mid[g], dep = B.nd_depth[g] + 1;
    B.nd_left[g] = c;
    B.nd_start[c] = s;
    B.nd_n[c] = mid - s;
    B.nd_depth[c] = dep;
    B.nd_start[c + 1] = mid;
    B.nd_n[c + 1] = s + n - mid;
    B.nd_depth[c + 1] = dep;
}

// subtree sizes and leaf counts, one level at a time from the deepest
__global__ void k_sizes(Build B, int g0, int m)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int g = g0 + j, l = B.nd_left[g];
    if (l < 0) {
        B.nd_size[g] = 1;
        B.nd_leaves[g] = 1;
    } else {
        B.nd_size[g] = 1 + B.nd_size[l] + B.nd_size[l + 1];
        B.nd_leaves[g] = B.nd_leaves[l] + B.nd_leaves[l + 1];
    }
}

// DFS index and the number of leaves before each node, one level at a time from the root
__global__ void k_dfs(Build B, int g0, int m)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int g = g0 + j, l = B.nd_left[g];
    if (l < 0) return;
    B.nd_idx[l] = B.nd_idx[g] + 1;
    B.nd_lb[l] = B.nd_lb[g];
    B.nd_idx[l + 1] = B.nd_idx[g] + 1 + B.nd_size[l];
    B.nd_lb[l + 1] = B.nd_lb[g] + B.nd_leaves[l];
}

struct Emit {
    float4 *records, *root;
    int32_t *rec_parent, *leaf_refs;
    uint32_t *leaf_links;
    float *pmin, *pmax;     // flat DFS export (3 floats per node)
    int32_t *nrefs, *offset;
    int *status;            // bit 0: a leaf holds more than MAX_LEAF_PRIMS primitives; [1]: max depth
};

__device__ __forceinline__ uint32_t link_of_node(const Build &B, int c)
{
    if (B.nd_left[c] < 0)
        return LEAF_BIT | ((uint32_t)(B.nd_n[c] - 1) << CNT_SHIFT) | (uint32_t)B.nd_start[c];
    return (uint32_t)(B.nd_idx[c] - B.nd_lb[c]);
}

template <int DIM>
__device__ __forceinline__ void put_slot(const Build &B, const Emit &E, int r, int side, int c, uint32_t link)
{
    float lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = d < DIM ? o2f(B.nd_box[c * 6 + d]) : 0.0f;
        hi[d] = d < DIM ? o2f(B.nd_box[c * 6 + 3 + d]) : 0.0f;
    }
    if (DIM == 3) {
        E.records[(size_t)r * 4 + side * 2] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(link));
        E.records[(size_t)r * 4 + side * 2 + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    } else {
        E.records[(size_t)r * 3 + side] = make_float4(lo[0], lo[1], hi[0], hi[1]);
        if (side == 0) E.records[(size_t)r * 3 + 2].x = __uint_as_float(link);
        else E.records[(size_t)r * 3 + 2].y = __uint_as_float(link);
    }
}

// the flat DFS tree (export) and the traversal records / refit schedule
template <int DIM>
__global__ void k_emit(Build B, Emit E, int total)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int i = B.nd_idx[g];
    const int l = B.nd_left[g];
    for (int d = 0; d < 3; d++) {
        E.pmin[(size_t)i * 3 + d] = d < DIM ? o2f(B.nd_box[g * 6 + d]) : 0.0f;
        E.pmax[(size_t)i * 3 + d] = d < DIM ? o2f(B.nd_box[g * 6 + 3 + d]) : 0.0f;
    }
    atomicMax(&E.status[1], B.nd_depth[g]);
    if (l < 0) {
        E.nrefs[i] = B.nd_n[g];
        E.offset[i] = B.nd_start[g];
        if (B.nd_n[g] > MAX_LEAF_PRIMS) atomicOr(&E.status[0], 1);
    } else {
        E.nrefs[i] = 0;
        E.offset[i] = 1 + B.nd_size[l];
        const int r = i - B.nd_lb[g];
        for (int side = 0; side < 2; side++) {
            const int c = l + side;
            const uint32_t link = link_of_node(B, c);
            put_slot<DIM>(B, E, r, side, c, link);
            if (B.nd_left[c] < 0) {
                E.leaf_refs[B.nd_lb[c]] = r * 2 + side;
                E.leaf_links[B.nd_lb[c]] = link;
            } else {
                E.rec_parent[B.nd_idx[c] - B.nd_lb[c]] = r * 2 + side;
            }
        }
    }
    if (g == 0) {
        const uint32_t link = link_of_node(B, 0);
        float lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = d < DIM ? o2f(B.nd_box[d]) : 0.0f;
            hi[d] = d < DIM ? o2f(B.nd_box[3 + d]) : 0.0f;
        }
        if (DIM == 3) {
            E.root[0] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(link));
            E.root[1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        } else {
            E.root[0] = make_float4(lo[0], lo[1], hi[0], hi[1]);
            E.root[1] = make_float4(__uint_as_float(link), 0.0f, 0.0f, 0.0f);
        }
        if (l < 0) {
            E.leaf_refs[0] = -1;
            E.leaf_links[0] = link;
        } else {
            E.rec_parent[0] = -1;
        }
    }
}

// primitive records in leaf-slot order (as fcpw_hip_upload_bvh lays them out)
template <int DIM>
__global__ void k_emit_prims(const int32_t *order, const float *V, const int32_t *idx, int n, float4 *prims,
                             int32_t *prim_index, int32_t *prim_vids)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int p = order[s];
    prim_index[s] = p;
    for (int k = 0; k < DIM; k++) prim_vids[(size_t)s * DIM + k] = idx[(size_t)p * DIM + k];
    if (DIM == 3) {
        for (int k = 0; k < 3; k++) {
            const float *v = V + (size_t)idx[(size_t)p * 3 + k] * 3;
            prims[(size_t)s * 3 + k] = make_float4(v[0], v[1], v[2], 0.0f);
        }
    } else {
        const float *a = V + (size_t)idx[(size_t)p * 2] * 2, *b = V + (size_t)idx[(size_t)p * 2 + 1] * 2;
        prims[s] = make_float4(a[0], a[1], b[0], b[1]);
    }
}

} // namespace gbuild
} // namespace fcpw_hip
