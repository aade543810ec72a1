// BVH construction for gfx950: leaf (swept) boxes + Morton codes, a stable LDS radix
// sort of (code, triangle id), the Karras radix tree and a bottom-up refit.
// Morton modes (per context): the reference's (infinite scene box: every code is 0, so the
// sort is the identity and is skipped), or a scene box fitted to the leaf boxes with 30-bit
// (10 bits per axis) or 63-bit (21 bits per axis) codes.
// Output is the reference wire format (src/build_bvh.cu:309-378):
//   bboxes [B, 2n-1, 6] (min xyz, max xyz), tree [B, 2n-1, 4] int64 (parent, left, right,
//   id), internal rows 0..n-2 (row 0 = root), leaf rows n-1..2n-2 in sorted order.
#include <float.h>

#include "ccc_geom.h"
#include "ccc_internal.h"

namespace ccc {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;

template <typename T> struct Lim;
template <> struct Lim<double> { static constexpr double max() { return DBL_MAX; } };
template <> struct Lim<float> { static constexpr float max() { return FLT_MAX; } };

// Spread the low 10 bits of v to every third bit (bit k -> bit 3k).
__device__ __forceinline__ uint32_t spread_bits10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// bvh.cu:54-62: each normalised coordinate is scaled by 1024, clamped to [0, 1023] with
// fmin/fmax (NaN -> 0) and truncated; code = x-bits * 4 + y-bits * 2 + z-bits.
__device__ __forceinline__ uint32_t morton30(double x, double y, double z) {
    x = fmin(fmax(x * 1024.0, 0.0), 1023.0);
    y = fmin(fmax(y * 1024.0, 0.0), 1023.0);
    z = fmin(fmax(z * 1024.0, 0.0), 1023.0);
    return spread_bits10((uint32_t)x) * 4u + spread_bits10((uint32_t)y) * 2u + spread_bits10((uint32_t)z);
}
__device__ __forceinline__ uint32_t morton30(float x, float y, float z) {
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
    z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
    return spread_bits10((uint32_t)x) * 4u + spread_bits10((uint32_t)y) * 2u + spread_bits10((uint32_t)z);
}

// 63-bit codes (not in the reference): 21 bits per axis, the same quantisation (scale by
// 2^21, clamp to [0, 2^21 - 1] with fmin/fmax, truncate) and interleave order.
__device__ __forceinline__ uint64_t spread_bits21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x001f00000000ffffull;
    v = (v | (v << 16)) & 0x001f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}
__device__ __forceinline__ uint64_t morton63(double x, double y, double z) {
    x = fmin(fmax(x * 2097152.0, 0.0), 2097151.0);
    y = fmin(fmax(y * 2097152.0, 0.0), 2097151.0);
    z = fmin(fmax(z * 2097152.0, 0.0), 2097151.0);
    return spread_bits21((uint64_t)x) * 4u + spread_bits21((uint64_t)y) * 2u + spread_bits21((uint64_t)z);
}
__device__ __forceinline__ uint64_t morton63(float x, float y, float z) {
    x = fminf(fmaxf(x * 2097152.0f, 0.0f), 2097151.0f);
    y = fminf(fmaxf(y * 2097152.0f, 0.0f), 2097151.0f);
    z = fminf(fmaxf(z * 2097152.0f, 0.0f), 2097151.0f);
    return spread_bits21((uint64_t)x) * 4u + spread_bits21((uint64_t)y) * 2u + spread_bits21((uint64_t)z);
}

// Leaf boxes (build_bvh.cu:212-219, :452-460), the triangle ids in input order and, for the
// fitted modes, the centroids the Morton codes are computed from (build_bvh.cu:26-71).
// The reference's scene box is the union of all leaf boxes with the reduction's initial
// value AABB<T>() = [-MAX, +MAX] (build_bvh.cu:243-244, :505-506; aabb.cu:25-33); min/max
// against +-MAX leave that initial value unchanged whatever the boxes are, so the scene box
// is [-MAX, +MAX], every normalised centroid is ~0.5 * 2^-1023 or smaller and every code is
// 0: the stable sort then keeps the input order, and the compat mode writes no codes.
template <typename T, bool MOTION, bool FIT>
__global__ void k_leaf_boxes(const T *__restrict__ t0, const T *__restrict__ t1, int64_t total,
                             int64_t n, T *__restrict__ boxes, T *__restrict__ cent,
                             int32_t *__restrict__ ids) {
    int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total) return;
    const T *a = t0 + 9 * gi;
    V3<T> a0 = ld3(a), a1 = ld3(a + 3), a2 = ld3(a + 6);
    Box<T> bx = tri_box(a0, a1, a2);
    V3<T> c = mk((a0.x + a1.x + a2.x) / (T)3.0, (a0.y + a1.y + a2.y) / (T)3.0, (a0.z + a1.z + a2.z) / (T)3.0);
    if (MOTION) {
        const T *b = t1 + 9 * gi;
        V3<T> b0 = ld3(b), b1 = ld3(b + 3), b2 = ld3(b + 6);
        bx = unite(bx, tri_box(b0, b1, b2));
        V3<T> c2 = mk((b0.x + b1.x + b2.x) / (T)3.0, (b0.y + b1.y + b2.y) / (T)3.0, (b0.z + b1.z + b2.z) / (T)3.0);
        c = mk((c.x + c2.x) / (T)2.0, (c.y + c2.y) / (T)2.0, (c.z + c2.z) / (T)2.0);
    }
    T *o = boxes + 6 * gi;
    o[0] = bx.mn.x; o[1] = bx.mn.y; o[2] = bx.mn.z; o[3] = bx.mx.x; o[4] = bx.mx.y; o[5] = bx.mx.z;
    if (FIT) {
        T *q = cent + 3 * gi;
        q[0] = c.x; q[1] = c.y; q[2] = c.z;
    }
    ids[gi] = (int32_t)(gi % n);
}

// Fitted scene box of each batch element: the union of its leaf boxes (min / max are exact
// and order-independent, NaN coordinates are ignored as by the leaf boxes' fmin / fmax).
// Stage 1: each block of element b reduces a strided share of its leaves to one box.
constexpr int kSceneBlocks = 64;
template <typename T>
__global__ __launch_bounds__(256) void k_scene_partial(const T *__restrict__ boxes_all, int64_t n, T *__restrict__ part_all) {
    __shared__ T red[6][256];
    const int b = blockIdx.y, tid = threadIdx.x;
    const T *boxes = boxes_all + (int64_t)b * n * 6;
    // the empty box (+inf, -inf) is the identity of the union
    const T inf = (T)INFINITY;
    T acc[6] = {inf, inf, inf, -inf, -inf, -inf};
    for (int64_t i = (int64_t)blockIdx.x * 256 + tid; i < n; i += (int64_t)gridDim.x * 256) {
        const T *p = boxes + 6 * i;
#pragma unroll
        for (int k = 0; k < 3; ++k) { acc[k] = tmin(acc[k], p[k]); acc[k + 3] = tmax(acc[k + 3], p[k + 3]); }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][tid] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                red[k][tid] = tmin(red[k][tid], red[k][tid + w]);
                red[k + 3][tid] = tmax(red[k + 3][tid], red[k + 3][tid + w]);
            }
        }
        __syncthreads();
    }
    if (tid < 6) part_all[((int64_t)b * gridDim.x + blockIdx.x) * 6 + tid] = red[tid][0];
}

// Stage 2 and the codes: one block per batch element unions the partial boxes, then every
// leaf's centroid is normalised by the scene box and quantised (bvh.cu:54-62; NaN -> 0).
template <typename T, typename K>
__global__ __launch_bounds__(256) void k_codes(const T *__restrict__ part_all, int nparts, const T *__restrict__ cent_all,
                                               int64_t n, K *__restrict__ codes_all) {
    __shared__ T scene[6];
    const int b = blockIdx.y;
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        const T *p = part_all + (int64_t)b * nparts * 6;
        T v = p[k];
        for (int j = 1; j < nparts; ++j) v = k < 3 ? tmin(v, p[6 * j + k]) : tmax(v, p[6 * j + k]);
        scene[k] = v;
    }
    __syncthreads();
    const V3<T> smin = mk(scene[0], scene[1], scene[2]), smax = mk(scene[3], scene[4], scene[5]);
    const T *cent = cent_all + (int64_t)b * n * 3;
    K *codes = codes_all + (int64_t)b * n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T *c = cent + 3 * i;
        const T x = (c[0] - smin.x) / (smax.x - smin.x);
        const T y = (c[1] - smin.y) / (smax.y - smin.y);
        const T z = (c[2] - smin.z) / (smax.z - smin.z);
        if (sizeof(K) == 8) codes[i] = (K)morton63(x, y, z);
        else codes[i] = (K)morton30(x, y, z);
    }
}

// ---- stable LSD radix sort of (code, id), 8 bits per pass --------------------------------
// Pass = per-tile digit histogram -> per-element exclusive scan (digit-major) -> stable
// scatter.  The scatter ranks keys inside a wave with a ballot multisplit (8 ballots give
// the lanes holding the same digit) and across the block's four waves with LDS counters.

template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const K *__restrict__ keys, int64_t n, int shift,
                                                            uint32_t *__restrict__ hist, int nb) {
    __shared__ uint32_t h[256];
    const int b = blockIdx.y, blk = blockIdx.x, tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const K *k = keys + (int64_t)b * n;
    for (int r = 0; r < kSortRounds; ++r) {
        int64_t i = (int64_t)blk * kSortTile + r * kSortThreads + tid;
        if (i < n) atomicAdd(&h[(uint32_t)(k[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[((int64_t)b * 256 + tid) * nb + blk] = h[tid];
}

// Exclusive scan of one element's [256][nb] histogram, one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_sort_scan(uint32_t *__restrict__ hist, int nb) {
    __shared__ uint32_t part[1024];
    const int tid = threadIdx.x;
    uint32_t *h = hist + (int64_t)blockIdx.x * 256 * nb;
    const int64_t total = (int64_t)256 * nb;
    const int64_t chunk = (total + 1023) / 1024;
    const int64_t lo = tid * chunk, hi = lo + chunk < total ? lo + chunk : total;
    uint32_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += h[i];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = part[tid] - s;
    for (int64_t i = lo; i < hi; ++i) {
        uint32_t v = h[i];
        h[i] = run;
        run += v;
    }
}

template <typename K>
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const K *__restrict__ kin, const int32_t *__restrict__ vin,
                                                               K *__restrict__ kout, int32_t *__restrict__ vout,
                                                               int64_t n, int shift, const uint32_t *__restrict__ hist, int nb) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t wcnt[kSortThreads / 64][256];
    const int b = blockIdx.y, blk = blockIdx.x, tid = threadIdx.x, w = tid >> 6;
    const int64_t off = (int64_t)b * n;
    gbase[tid] = hist[((int64_t)b * 256 + tid) * nb + blk];
    run[tid] = 0;
    for (int r = 0; r < kSortRounds; ++r) {
#pragma unroll
        for (int q = 0; q < kSortThreads / 64; ++q) wcnt[q][tid] = 0;
        __syncthreads();
        const int64_t i = (int64_t)blk * kSortTile + r * kSortThreads + tid;
        const bool valid = i < n;
        const K key = valid ? kin[off + i] : (K)0;
        const int32_t val = valid ? vin[off + i] : 0;
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            unsigned long long bb = __ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        const unsigned rank = lanes_below(peers);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t base = run[tid], acc = 0;
#pragma unroll
            for (int q = 0; q < kSortThreads / 64; ++q) {
                uint32_t c = wcnt[q][tid];
                wcnt[q][tid] = base + acc;
                acc += c;
            }
            run[tid] = base + acc;
        }
        __syncthreads();
        if (valid) {
            const int64_t dst = (int64_t)gbase[d] + wcnt[w][d] + rank;
            kout[off + dst] = key;
            vout[off + dst] = val;
        }
        __syncthreads();
    }
}

// ---- Karras radix tree ----------------------------------------------------------------------
// bvh.cu:70-89 LongestCommonPrefix: -1 outside [0, n-1]; equal codes extend the prefix with
// the ids' common prefix (32 + clz(id_i ^ id_j) for 32-bit keys, 64 + ... for 64-bit ones).
// codes == NULL: every code is 0 (the reference's scene box), only the ids decide.
template <typename K> __device__ __forceinline__ int clz_key(K x);
template <> __device__ __forceinline__ int clz_key<uint32_t>(uint32_t x) { return __clz((int)x); }
template <> __device__ __forceinline__ int clz_key<uint64_t>(uint64_t x) { return __clzll((long long)x); }

template <typename K>
__device__ __forceinline__ int lcp_delta(const K *codes, const int32_t *ids, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j > n - 1) return -1;
    if (codes) {
        const K a = codes[i], c = codes[j];
        if (a != c) return clz_key<K>(a ^ c);
    }
    return 8 * (int)sizeof(K) + __clz((int)((uint32_t)ids[i] ^ (uint32_t)ids[j]));
}

// Karras 2012, Algorithm 1 (what build_bvh.cu:74-143 computes): node i finds its range
// direction, the far end j by exponential + binary search and the split by binary search.
// Writes left/right/self of row i and the parent field of both children's rows.
template <typename K>
__global__ void k_karras(const K *__restrict__ codes_all, const int32_t *__restrict__ ids_all, int64_t n,
                         int64_t *__restrict__ tree_all, int2 *__restrict__ range_all) {
    const int b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const K *codes = codes_all ? codes_all + (int64_t)b * n : nullptr;
    const int32_t *ids = ids_all + (int64_t)b * n;
    int64_t *tree = tree_all + (int64_t)b * (2 * n - 1) * 4;
    int2 *range = range_all + (int64_t)b * (n - 1);
    const int d = (lcp_delta(codes, ids, n, i, i + 1) - lcp_delta(codes, ids, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = lcp_delta(codes, ids, n, i, i - d);
    int64_t lmax = 2;
    while (lcp_delta(codes, ids, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (lcp_delta(codes, ids, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = lcp_delta(codes, ids, n, i, j);
    int64_t s = 0;
    for (int64_t div = 2;; div *= 2) {
        const int64_t t = (l + div - 1) / div;
        if (lcp_delta(codes, ids, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int64_t g = i + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    const int64_t left = (lo == g) ? (n - 1 + g) : g;
    const int64_t right = (hi == g + 1) ? (n - 1 + g + 1) : (g + 1);
    tree[4 * i + 1] = left;
    tree[4 * i + 2] = right;
    tree[4 * i + 3] = -1;
    if (i == 0) tree[0] = -1;
    tree[4 * left + 0] = i;
    tree[4 * right + 0] = i;
    range[i] = make_int2((int)lo, (int)hi);
}

// Node boxes (build_bvh.cu:145-209 computes them by an atomic bottom-up climb).  Every
// internal node of the Karras tree covers a contiguous range [lo, hi] of sorted leaf
// positions, so its box is the union of the leaf boxes of that range.  Min / max are exact
// and order-independent, so any grouping gives the climb's result; here the union is read
// off a hierarchy of chunk boxes (64, 64^2, ... leaves), with no atomics or fences.
constexpr int kChunkBits = 6;
constexpr int kChunk = 1 << kChunkBits;
constexpr int kChunkLevels = 5;      // 64^5 = 2^30 > 2^28 leaves

// Leaf rows (build_bvh.cu:309-378 layout): the box of triangle ids[p] and (-1, -1, id).
template <typename T>
__global__ void k_leaf_rows(const T *__restrict__ boxes_all, const int32_t *__restrict__ ids_all, int64_t n,
                            T *__restrict__ bboxes_all, int64_t *__restrict__ tree_all) {
    const int b = blockIdx.y;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t rows = 2 * n - 1;
    const T *boxes = boxes_all + (int64_t)b * n * 6;
    T *bb = bboxes_all + (int64_t)b * rows * 6;
    int64_t *tree = tree_all + (int64_t)b * rows * 4;
    const int64_t row = n - 1 + p;
    const int32_t id = ids_all[(int64_t)b * n + p];
#pragma unroll
    for (int k = 0; k < 6; ++k) bb[6 * row + k] = boxes[6 * (int64_t)id + k];
    if (n == 1) tree[4 * row + 0] = -1;
    tree[4 * row + 1] = -1;
    tree[4 * row + 2] = -1;
    tree[4 * row + 3] = id;
}

template <typename T> __device__ __forceinline__ void box_union_into(T *acc, const T *src) {
    acc[0] = tmin(acc[0], src[0]);
    acc[1] = tmin(acc[1], src[1]);
    acc[2] = tmin(acc[2], src[2]);
    acc[3] = tmax(acc[3], src[3]);
    acc[4] = tmax(acc[4], src[4]);
    acc[5] = tmax(acc[5], src[5]);
}

// chunk level L box c = union of the (at most 64) level L-1 boxes 64c .. 64c+63; level 0 is
// the leaf rows of bboxes.
template <typename T>
__global__ void k_chunk_boxes(const T *__restrict__ src_all, int64_t src_stride, int64_t src_count, T *__restrict__ dst_all,
                              int64_t dst_stride, int64_t dst_count) {
    const int b = blockIdx.y;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= dst_count) return;
    const T *src = src_all + (int64_t)b * src_stride;
    T acc[6];
    const int64_t first = c << kChunkBits;
    const int64_t last = first + kChunk < src_count ? first + kChunk : src_count;
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] = src[6 * first + k];
    for (int64_t j = first + 1; j < last; ++j) box_union_into(acc, src + 6 * j);
    T *dst = dst_all + (int64_t)b * dst_stride + 6 * c;
#pragma unroll
    for (int k = 0; k < 6; ++k) dst[k] = acc[k];
}

struct ChunkLevels {
    const void *box[kChunkLevels + 1];   // [0]: leaf boxes (bboxes leaf rows), [L]: level L chunks
    int64_t stride[kChunkLevels + 1];    // elements between batch elements
    int levels;                          // chunk levels in use
};

template <typename T>
__global__ void k_node_boxes(const int2 *__restrict__ range_all, int64_t n, ChunkLevels lv, T *__restrict__ bboxes_all) {
    const int b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int2 r = range_all[(int64_t)b * (n - 1) + i];
    int64_t p = r.x;
    const int64_t end = (int64_t)r.y + 1;
    T acc[6];
    bool first = true;
    while (p < end) {
        // the largest aligned chunk that starts at p and ends inside the range
        int L = 0;
        while (L < lv.levels && (p & ((int64_t)kChunk << (kChunkBits * L)) - 1) == 0 &&
               p + ((int64_t)kChunk << (kChunkBits * L)) <= end)
            ++L;
        const T *src = (const T *)lv.box[L] + (int64_t)b * lv.stride[L] + 6 * (p >> (kChunkBits * L));
        if (first) {
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[k] = src[k];
            first = false;
        } else {
            box_union_into(acc, src);
        }
        p += (int64_t)1 << (kChunkBits * L);
    }
    T *o = bboxes_all + (int64_t)b * (2 * n - 1) * 6 + 6 * i;
#pragma unroll
    for (int k = 0; k < 6; ++k) o[k] = acc[k];
}

// Stable LSD radix sort of B rows of n (key, id) pairs, 8 bits per pass over the key's width;
// returns which buffers hold the result.
template <typename K>
static ccc_status radix_sort(K *&ka, K *&kb, int32_t *&va, int32_t *&vb, int64_t B, int64_t n, int bits, uint32_t *hist,
                             hipStream_t s) {
    const int nb = (int)((n + kSortTile - 1) / kSortTile);
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL((k_sort_hist<K>), dim3(nb, B), dim3(kSortThreads), 0, s, ka, n, shift, hist, nb);
        hipLaunchKernelGGL(k_sort_scan, dim3(B), dim3(1024), 0, s, hist, nb);
        hipLaunchKernelGGL((k_sort_scatter<K>), dim3(nb, B), dim3(kSortThreads), 0, s, ka, va, kb, vb, n, shift, hist, nb);
        CCC_LAUNCH_CHECK("radix sort");
        K *tk = ka; ka = kb; kb = tk;
        int32_t *tv = va; va = vb; vb = tv;
    }
    return CCC_OK;
}

template <typename T>
static ccc_status build(ccc_context *ctx, const T *t0, const T *t1, int64_t B, int64_t n, int morton, T *bboxes, int64_t *tree,
                        hipStream_t s, void *arena, size_t arena_bytes) {
    Workspace ws(ctx, s, arena, arena_bytes);
    KernelTimer tm(ctx, s);
    const int64_t total = B * n;
    const int nb = (int)((n + kSortTile - 1) / kSortTile);
    const bool fit = morton != CCC_MORTON_COMPAT;
    const size_t ksz = morton == CCC_MORTON_63 ? 8 : 4;
    T *boxes = ws.get_as<T>(total * 6);
    int32_t *va = ws.get_as<int32_t>(total);
    T *cent = fit ? ws.get_as<T>(total * 3) : nullptr;
    T *part = fit ? ws.get_as<T>((size_t)B * kSceneBlocks * 6) : nullptr;
    void *ka = fit ? ws.get(total * ksz) : nullptr;
    void *kb = fit ? ws.get(total * ksz) : nullptr;
    int32_t *vb = fit ? ws.get_as<int32_t>(total) : nullptr;
    uint32_t *hist = fit ? ws.get_as<uint32_t>((size_t)B * 256 * nb) : nullptr;
    int2 *range = ws.get_as<int2>((size_t)B * (n > 1 ? n - 1 : 1));
    ChunkLevels lv{};
    lv.box[0] = bboxes + (size_t)6 * (n - 1);
    lv.stride[0] = (2 * n - 1) * 6;
    int64_t count = n;
    while (lv.levels < kChunkLevels && count > kChunk) {
        count = (count + kChunk - 1) / kChunk;
        ++lv.levels;
        lv.box[lv.levels] = ws.get_as<T>((size_t)B * count * 6);
        lv.stride[lv.levels] = count * 6;
    }
    if (ws.failed()) return CCC_ERR_NO_MEMORY;

    tm.begin("bvh_leaf_boxes");
#define CCC_LEAF(M, F) hipLaunchKernelGGL((k_leaf_boxes<T, M, F>), dim3(grid_for(total, 256)), dim3(256), 0, s, t0, t1, total, n, boxes, cent, va)
    if (t1) { if (fit) CCC_LEAF(true, true); else CCC_LEAF(true, false); }
    else { if (fit) CCC_LEAF(false, true); else CCC_LEAF(false, false); }
#undef CCC_LEAF
    tm.end();
    CCC_LAUNCH_CHECK("k_leaf_boxes");

    // compat: every code is 0 and the stable sort keeps the input order (no codes, no sort)
    if (fit) {
        tm.begin("bvh_sort");
        const unsigned gparts = (unsigned)std::min<int64_t>(kSceneBlocks, (n + 255) / 256);
        hipLaunchKernelGGL((k_scene_partial<T>), dim3(gparts, B), dim3(256), 0, s, boxes, n, part);
        const unsigned gc = (unsigned)std::min<int64_t>(1024, (n + 255) / 256);
        if (ksz == 8) {
            hipLaunchKernelGGL((k_codes<T, uint64_t>), dim3(gc, B), dim3(256), 0, s, part, (int)gparts, cent, n, (uint64_t *)ka);
            CCC_LAUNCH_CHECK("morton codes");
            uint64_t *a = (uint64_t *)ka, *b = (uint64_t *)kb;
            ccc_status st = radix_sort<uint64_t>(a, b, va, vb, B, n, 64, hist, s);
            if (st != CCC_OK) return st;
            ka = a;
        } else {
            hipLaunchKernelGGL((k_codes<T, uint32_t>), dim3(gc, B), dim3(256), 0, s, part, (int)gparts, cent, n, (uint32_t *)ka);
            CCC_LAUNCH_CHECK("morton codes");
            uint32_t *a = (uint32_t *)ka, *b = (uint32_t *)kb;
            ccc_status st = radix_sort<uint32_t>(a, b, va, vb, B, n, 32, hist, s);
            if (st != CCC_OK) return st;
            ka = a;
        }
        tm.end();
    }
    if (n > 1) {
        tm.begin("bvh_karras");
        if (ksz == 8) hipLaunchKernelGGL((k_karras<uint64_t>), dim3(grid_for(n - 1, 256), B), dim3(256), 0, s, (const uint64_t *)ka, va, n, tree, range);
        else hipLaunchKernelGGL((k_karras<uint32_t>), dim3(grid_for(n - 1, 256), B), dim3(256), 0, s, (const uint32_t *)ka, va, n, tree, range);
        tm.end();
        CCC_LAUNCH_CHECK("k_karras");
    }
    tm.begin("bvh_refit");
    hipLaunchKernelGGL((k_leaf_rows<T>), dim3(grid_for(n, 256), B), dim3(256), 0, s, boxes, va, n, bboxes, tree);
    {
        int64_t src_count = n;
        for (int L = 1; L <= lv.levels; ++L) {
            const int64_t dst_count = (src_count + kChunk - 1) / kChunk;
            hipLaunchKernelGGL((k_chunk_boxes<T>), dim3(grid_for(dst_count, 256), B), dim3(256), 0, s, (const T *)lv.box[L - 1],
                               lv.stride[L - 1], src_count, (T *)lv.box[L], lv.stride[L], dst_count);
            src_count = dst_count;
        }
    }
    if (n > 1) hipLaunchKernelGGL((k_node_boxes<T>), dim3(grid_for(n - 1, 256), B), dim3(256), 0, s, range, n, lv, bboxes);
    tm.end();
    CCC_LAUNCH_CHECK("node boxes");
    return CCC_OK;
}

ccc_status build_bvh_device(ccc_context *ctx, const void *t0, const void *t1, int64_t B, int64_t n, int32_t dtype, int morton,
                            void *bboxes, int64_t *tree, hipStream_t s, void *arena, size_t arena_bytes) {
    if (morton != CCC_MORTON_COMPAT && morton != CCC_MORTON_30 && morton != CCC_MORTON_63) return bad_arg("unknown Morton mode");
    if (dtype == CCC_FLOAT64)
        return build<double>(ctx, (const double *)t0, (const double *)t1, B, n, morton, (double *)bboxes, tree, s, arena, arena_bytes);
    if (dtype == CCC_FLOAT32)
        return build<float>(ctx, (const float *)t0, (const float *)t1, B, n, morton, (float *)bboxes, tree, s, arena, arena_bytes);
    return bad_arg("dtype must be CCC_FLOAT32 or CCC_FLOAT64");
}

size_t build_bvh_arena_bytes(int64_t B, int64_t n, int32_t dtype, int morton) {
    auto r = [](size_t b) { return ((b ? b : 256) + 255) & ~(size_t)255; };
    const size_t esz = dtype == CCC_FLOAT64 ? 8 : 4, total = (size_t)(B * n);
    const bool fit = morton != CCC_MORTON_COMPAT;
    const size_t ksz = morton == CCC_MORTON_63 ? 8 : 4;
    const int nb = (int)((n + kSortTile - 1) / kSortTile);
    size_t tot = r(esz * total * 6) + r(4 * total);
    if (fit) tot += r(esz * total * 3) + r(esz * B * kSceneBlocks * 6) + 2 * r(ksz * total) + r(4 * total) + r(4 * (size_t)B * 256 * nb);
    tot += r(sizeof(int2) * B * (n > 1 ? n - 1 : 1));
    int64_t count = n;
    for (int levels = 0; levels < kChunkLevels && count > kChunk; ++levels) {
        count = (count + kChunk - 1) / kChunk;
        tot += r(esz * B * count * 6);
    }
    return tot;
}

static ccc_status build_entry(ccc_context *ctx, const void *t0, const void *t1, int64_t B, int64_t n, int32_t dtype,
                              void *bboxes, int64_t *tree, void *stream) {
    clear_error();
    if (!ctx) return bad_arg("context is NULL");
    if (B < 0 || n < 1) return bad_arg("need batch >= 0 and n >= 1 triangles");
    if (n > (int64_t)INT32_MAX / 2) return bad_arg("too many triangles");
    if (B == 0) return CCC_OK;
    if (!t0 || !bboxes || !tree) return bad_arg("NULL tensor pointer");
    CCC_HIP_TRY(hipSetDevice(ctx->device));
    return build_bvh_device(ctx, t0, t1, B, n, dtype, ctx->morton, bboxes, tree, (hipStream_t)stream);
}

}  // namespace ccc

extern "C" {

ccc_status ccc_bvh(ccc_context *ctx, const void *triangles, int64_t batch, int64_t n, int32_t dtype, void *bboxes,
                   int64_t *tree, void *stream) {
    return ccc::build_entry(ctx, triangles, nullptr, batch, n, dtype, bboxes, tree, stream);
}

ccc_status ccc_bvh_motion(ccc_context *ctx, const void *triangles, const void *triangles_next, int64_t batch, int64_t n,
                          int32_t dtype, void *bboxes, int64_t *tree, void *stream) {
    if (!triangles_next) {
        ccc::set_error("triangles_next is NULL");
        return CCC_ERR_BAD_ARG;
    }
    return ccc::build_entry(ctx, triangles, triangles_next, batch, n, dtype, bboxes, tree, stream);
}

}  // extern "C"
