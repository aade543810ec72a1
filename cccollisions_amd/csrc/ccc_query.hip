// Queries over the BVH wire format for gfx950: tree preparation, traversal, the 15-way
// sub-primitive box filter, the fused CCD (coefficients -> cubic -> filter -> collision or
// impulse) kernel, static proximity, static triangle-triangle SAT and impulses from
// collision lists.
//
// Pipeline of one query call (all batch elements at once, batch index = blockIdx.y):
//   check tree -> rightmost leaf per internal node -> 64-byte traversal nodes (children's
//   fp32 boxes rounded outward and grown by the threshold) -> traversal (one lane per
//   query leaf, LDS stack, exact fp64 leaf test, wave-staged appends) -> candidate pairs
//   -> per pair: 15-bit sub-primitive mask -> packed (x, y, type) items -> per item:
//   narrow phase.
#include <float.h>
#include <limits.h>

#include <algorithm>

#include "ccc_geom.h"
#include "ccc_internal.h"

namespace ccc {

constexpr int kThreads = 256;
constexpr int kTravThreads = 64;   // one wave per traversal block (128: 20.2 ms per step, 64: 18.7)
// Pending subtrees per lane (LDS).  A four-wide visit (two binary levels) pushes at most
// three entries, so with start points kFrontierDepth = 7 levels down the short stack covers
// the balanced trees built here (index radix trees, ceil(log2 N) + 1 deep) up to 2^22
// leaves; a traversal that overflows it runs again with the deep one.
constexpr int kStackShort = 24;
constexpr int kStackDeep = 32;
constexpr int kStage = 512;       // staged pairs per wave before a flush
// queries start from the subtrees this deep below the root (k_traverse per pass on the 130k
// cloth: depth 6 2.85 ms, 7 2.02 ms, 8 2.60 ms)
constexpr int kFrontierDepth = 7;
constexpr int kFrontier = 1 << kFrontierDepth;

// Workgroups are dispatched to the eight XCDs round-robin by workgroup ID, and each XCD has
// its own L2.  Persistent kernels walk their tiles with CCC_FOR_TILES, which gives the
// workgroups of one XCD one contiguous range of tiles (neighbouring tiles read the same
// triangles); their grids are a multiple of kXcds.
constexpr int kXcds = 8;
#define CCC_FOR_TILES(tile, T)                                                                          \
    for (int64_t tile##_k = blockIdx.x / kXcds, tile##_per = ((T) + kXcds - 1) / kXcds,                 \
                 tile = (int64_t)(blockIdx.x % kXcds) * tile##_per + tile##_k;                          \
         tile##_k < tile##_per && tile < (T);                                                           \
         tile##_k += gridDim.x / kXcds, tile = (int64_t)(blockIdx.x % kXcds) * tile##_per + tile##_k)

struct alignas(16) TNode {
    float4 a;  // L.min.x, L.min.y, L.min.z, L.max.x
    float4 b;  // L.max.y, L.max.z, R.min.x, R.min.y
    float4 c;  // R.min.z, R.max.x, R.max.y, R.max.z
    int4 d;    // left ref, right ref, rightmost leaf position of left / right child
};               // ref >= 0: internal row; ref < 0: leaf at position ~ref

__device__ __forceinline__ float f32_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    return nextafterf(f, -INFINITY);
}
__device__ __forceinline__ float f32_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return nextafterf(f, INFINITY);
}

__device__ __forceinline__ int64_t pack_item(int64_t x, int64_t y, int type) {
    return (int64_t)(((uint64_t)x << 32) | ((uint64_t)y << 4) | (uint64_t)type);
}
__device__ __forceinline__ void unpack_item(int64_t it, int64_t &x, int64_t &y, int &type) {
    uint64_t u = (uint64_t)it;
    x = (int64_t)(u >> 32);
    y = (int64_t)((u & 0xffffffffull) >> 4);
    type = (int)(u & 15ull);
}

// Block-wide reservation of `cnt` (< 16) consecutive slots per thread with one atomic.
// Every thread of the block (of NT threads) must call it.
template <int NT = kThreads>
__device__ __forceinline__ unsigned long long block_reserve(unsigned long long *counter, unsigned cnt) {
    __shared__ unsigned long long s_wave[NT / 64];
    __shared__ unsigned long long s_base;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned long long b0 = __ballot(cnt & 1u), b1 = __ballot(cnt & 2u), b2 = __ballot(cnt & 4u), b3 = __ballot(cnt & 8u);
    unsigned below = lanes_below(b0) + 2u * lanes_below(b1) + 4u * lanes_below(b2) + 8u * lanes_below(b3);
    unsigned long long wt = (unsigned long long)__popcll(b0) + 2ull * __popcll(b1) + 4ull * __popcll(b2) + 8ull * __popcll(b3);
    if (lane == 0) s_wave[w] = wt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int q = 0; q < NT / 64; ++q) {
            unsigned long long v = s_wave[q];
            s_wave[q] = tot;
            tot += v;
        }
        s_base = tot ? atomicAdd(counter, tot) : 0ull;
    }
    __syncthreads();
    unsigned long long r = s_base + s_wave[w] + below;
    __syncthreads();
    return r;
}

__device__ __forceinline__ void block_add(unsigned long long *counter, unsigned long long v) {
    __shared__ unsigned long long s_sum;
    if (threadIdx.x == 0) s_sum = 0;
    __syncthreads();
    if (v) atomicAdd(&s_sum, v);
    __syncthreads();
    if (threadIdx.x == 0 && s_sum) atomicAdd(counter, s_sum);
}

// ---- tree preparation -------------------------------------------------------------------

// Checks that `tree` is a binary tree over n leaves in the wire format.  Violations set
// DEV_BAD_TREE; nothing else reads an index before it has been range-checked here.
__global__ void k_check_tree(const int64_t *__restrict__ tree_all, int64_t n, int *__restrict__ claims_all,
                             DevCounters *__restrict__ ctr) {
    const int b = blockIdx.y;
    const int64_t rows = 2 * n - 1;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int64_t *tree = tree_all + (int64_t)b * rows * 4;
    int *claims = claims_all + (int64_t)b * n;
    const int64_t p = tree[4 * r], l = tree[4 * r + 1], rr = tree[4 * r + 2], id = tree[4 * r + 3];
    bool ok = true;
    if (n == 1) {
        ok = p == -1 && l == -1 && rr == -1 && id == 0;
    } else if (r < n - 1) {
        ok = l >= 1 && l < rows && rr >= 1 && rr < rows && l != rr;
        if (ok) ok = tree[4 * l] == r && tree[4 * rr] == r;
        ok = ok && (r == 0 ? p == -1 : (p >= 0 && p < n - 1));
    } else {
        ok = l == -1 && rr == -1 && id >= 0 && id < n && p >= 0 && p < n - 1;
        if (ok) ok = atomicAdd(&claims[id], 1) == 0;
    }
    if (ok && n > 1 && r != 0) ok = tree[4 * p + 1] == r || tree[4 * p + 2] == r;
    if (!ok) atomicOr(&ctr[b].flags, DEV_BAD_TREE);
}

// Rightmost leaf position under every internal node (build_bvh.cu:707-743): bottom-up
// climb, the second thread to arrive at a node combines its children.  Bounded by n steps
// (a longer climb can only come from a cycle, reported as DEV_BAD_TREE).
// Rightmost leaf position of every internal node (build_bvh.cu:707-741 set_rightmost_nodes:
// rightmost = the larger of the children's rightmost leaves, compared by leaf address).  Trees whose subtrees cover contiguous
// ordered leaf ranges -- every tree bvh / bvh_motion builds -- take the walk below; any
// other tree falls back to an atomic bottom-up climb.
//
// Walk: lo / hi of a node = position of the leaf reached by following left / right
// children (at most 64 steps).
__global__ void k_ranges(const int64_t *__restrict__ tree_all, int64_t n, int *__restrict__ lo_all, int *__restrict__ rm_all,
                         DevCounters *__restrict__ ctr) {
    const int b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    if (ctr[b].flags & DEV_BAD_TREE) return;
    const int64_t *tree = tree_all + (int64_t)b * (2 * n - 1) * 4;
    int64_t c = i, d = i;
    for (int step = 0; c < n - 1 || d < n - 1; ++step) {
        if (step >= 64) {
            atomicOr(&ctr[b].flags, DEV_UNORDERED);
            return;
        }
        if (c < n - 1) c = tree[4 * c + 1];
        if (d < n - 1) d = tree[4 * d + 2];
    }
    lo_all[(int64_t)b * (n - 1) + i] = (int)(c - (n - 1));
    rm_all[(int64_t)b * (n - 1) + i] = (int)(d - (n - 1));
}

// The walk's hi is the subtree maximum only if, at every node, the left child's range ends
// right before the right child's begins.
__global__ void k_order_check(const int64_t *__restrict__ tree_all, int64_t n, const int *__restrict__ lo_all,
                              const int *__restrict__ rm_all, DevCounters *__restrict__ ctr) {
    const int b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    if (ctr[b].flags & (DEV_BAD_TREE | DEV_UNORDERED)) return;
    const int64_t *tree = tree_all + (int64_t)b * (2 * n - 1) * 4;
    const int *lo = lo_all + (int64_t)b * (n - 1);
    const int *rm = rm_all + (int64_t)b * (n - 1);
    const int64_t l = tree[4 * i + 1], r = tree[4 * i + 2];
    const int hl = l >= n - 1 ? (int)(l - (n - 1)) : rm[l];
    const int lr = r >= n - 1 ? (int)(r - (n - 1)) : lo[r];
    if (hl + 1 != lr) atomicOr(&ctr[b].flags, DEV_UNORDERED);
}

// Climb for the remaining trees: the second thread to reach a node combines its children.
__global__ void k_rightmost(const int64_t *__restrict__ tree_all, int64_t n, int *__restrict__ counters_all,
                            int *__restrict__ rm_all, DevCounters *__restrict__ ctr) {
    const int b = blockIdx.y;
    const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= n || n < 2) return;
    if ((ctr[b].flags & DEV_BAD_TREE) || !(ctr[b].flags & DEV_UNORDERED)) return;
    const int64_t rows = 2 * n - 1;
    const int64_t *tree = tree_all + (int64_t)b * rows * 4;
    int *counters = counters_all + (int64_t)b * (n - 1);
    int *rm = rm_all + (int64_t)b * (n - 1);
    int64_t cur = tree[4 * (n - 1 + pos)];
    for (int64_t step = 0; cur >= 0; ++step) {
        if (step > n) { atomicOr(&ctr[b].flags, DEV_BAD_TREE); return; }
        __threadfence();
        const int old = __hip_atomic_fetch_add(&counters[cur], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) return;
        __threadfence();
        const int64_t l = tree[4 * cur + 1], r = tree[4 * cur + 2];
        const int a = l >= n - 1 ? (int)(l - (n - 1)) : __hip_atomic_load(&rm[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int c = r >= n - 1 ? (int)(r - (n - 1)) : __hip_atomic_load(&rm[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&rm[cur], a > c ? a : c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        cur = tree[4 * cur];
    }
}

// One 64-byte traversal node per internal row: both children's boxes in fp32, rounded
// outward, grown by the threshold and by one more float ulp on each side, so that the
// fp32 test passes whenever the reference's fp64 test passes for either orientation.
template <typename T>
__global__ void k_nodes(const T *__restrict__ bb_all, const int64_t *__restrict__ tree_all, const int *__restrict__ rm_all,
                        int64_t n, double thr, TNode *__restrict__ nodes_all, const DevCounters *__restrict__ ctr) {
    const int b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    if (ctr[b].flags & DEV_BAD_TREE) return;
    const int64_t rows = 2 * n - 1;
    const T *bb = bb_all + (int64_t)b * rows * 6;
    const int64_t *tree = tree_all + (int64_t)b * rows * 4;
    const int *rm = rm_all + (int64_t)b * (n - 1);
    const int64_t l = tree[4 * i + 1], r = tree[4 * i + 2];
    float lo[6], ro[6];
    const T *lb = bb + 6 * l, *rb = bb + 6 * r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = f32_down((double)lb[k] - thr);
        lo[k + 3] = f32_up((double)lb[k + 3] + thr);
        ro[k] = f32_down((double)rb[k] - thr);
        ro[k + 3] = f32_up((double)rb[k + 3] + thr);
    }
    TNode nd;
    nd.a = make_float4(lo[0], lo[1], lo[2], lo[3]);
    nd.b = make_float4(lo[4], lo[5], ro[0], ro[1]);
    nd.c = make_float4(ro[2], ro[3], ro[4], ro[5]);
    const int lref = l >= n - 1 ? ~(int)(l - (n - 1)) : (int)l;
    const int rref = r >= n - 1 ? ~(int)(r - (n - 1)) : (int)r;
    const int lrm = l >= n - 1 ? (int)(l - (n - 1)) : rm[l];
    const int rrm = r >= n - 1 ? (int)(r - (n - 1)) : rm[r];
    nd.d = make_int4(lref, rref, lrm, rrm);
    nodes_all[(int64_t)b * (n - 1) + i] = nd;
}

// Four-wide view of the binary tree for the traversal.  For internal row i: the up to four
// grandchildren of i, left to right (a child that is a leaf stands for itself), with their
// TNode fp32 boxes as structure of arrays, refs and rightmost leaf positions; unused entries
// (at the end) have a NaN box and rightmost INT_MIN.  Entry k > 0 starts at the leaf after
// entry k-1's rightmost one (trees built here cover contiguous position ranges).  Testing a
// grandchild's box without its parent's prunes nothing that matters when boxes contain their
// children's (the traversal's assumption throughout: it starts below the root, and its node
// tests are conservative fp32): a pair needs the leaf's exact box to overlap the query.  One
// node is 128 bytes, one cache line.
struct alignas(16) QNode {
    float4 mnx, mny, mnz, mxx, mxy, mxz;
    int4 ref;   // >= 0 internal row, < 0 leaf at position ~ref
    int4 rm;    // rightmost leaf position; INT_MIN: unused entry
};

// side 0 / 1 of a TNode: box, ref and rightmost leaf of that child
__device__ __forceinline__ void tnode_child(const TNode &nd, int side, float *mn, float *mx, int &ref, int &rm) {
    if (side == 0) {
        mn[0] = nd.a.x; mn[1] = nd.a.y; mn[2] = nd.a.z; mx[0] = nd.a.w; mx[1] = nd.b.x; mx[2] = nd.b.y;
        ref = nd.d.x; rm = nd.d.z;
    } else {
        mn[0] = nd.b.z; mn[1] = nd.b.w; mn[2] = nd.c.x; mx[0] = nd.c.y; mx[1] = nd.c.z; mx[2] = nd.c.w;
        ref = nd.d.y; rm = nd.d.w;
    }
}

struct QEntry {
    float mn[3], mx[3];
    int ref, rm;
};

__global__ void k_qnodes(const TNode *__restrict__ nodes_all, int64_t n, QNode *__restrict__ q_all,
                         const DevCounters *__restrict__ ctr) {
    const int b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    if (ctr[b].flags & DEV_BAD_TREE) return;
    const TNode *nodes = nodes_all + (int64_t)b * (n - 1);
    const TNode nd = nodes[i];
    QEntry none;    // unused entry: a NaN box fails every comparison
#pragma unroll
    for (int a = 0; a < 3; ++a) { none.mn[a] = NAN; none.mx[a] = NAN; }
    none.ref = 0;
    none.rm = INT_MIN;
    // side s of row i -> its entries g[s][0], g[s][1] (g[s][1] unused when the child is a leaf)
    QEntry g[2][2];
    bool internal[2];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        QEntry c;
        tnode_child(nd, side, c.mn, c.mx, c.ref, c.rm);
        internal[side] = c.ref >= 0;
        if (internal[side]) {
            const TNode cn = nodes[c.ref];
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2) tnode_child(cn, s2, g[side][s2].mn, g[side][s2].mx, g[side][s2].ref, g[side][s2].rm);
        } else {
            g[side][0] = c;
            g[side][1] = none;
        }
    }
    // left to right, unused entries last
    const QEntry e0 = g[0][0];
    const QEntry e1 = internal[0] ? g[0][1] : g[1][0];
    const QEntry e2 = internal[0] ? g[1][0] : (internal[1] ? g[1][1] : none);
    const QEntry e3 = internal[0] && internal[1] ? g[1][1] : none;
    QNode q;
    q.mnx = make_float4(e0.mn[0], e1.mn[0], e2.mn[0], e3.mn[0]);
    q.mny = make_float4(e0.mn[1], e1.mn[1], e2.mn[1], e3.mn[1]);
    q.mnz = make_float4(e0.mn[2], e1.mn[2], e2.mn[2], e3.mn[2]);
    q.mxx = make_float4(e0.mx[0], e1.mx[0], e2.mx[0], e3.mx[0]);
    q.mxy = make_float4(e0.mx[1], e1.mx[1], e2.mx[1], e3.mx[1]);
    q.mxz = make_float4(e0.mx[2], e1.mx[2], e2.mx[2], e3.mx[2]);
    q.ref = make_int4(e0.ref, e1.ref, e2.ref, e3.ref);
    q.rm = make_int4(e0.rm, e1.rm, e2.rm, e3.rm);
    q_all[(int64_t)b * (n - 1) + i] = q;
}

// Leaf record by position: the leaf box rounded inward to fp32 and the leaf's triangle id.
// A query box that overlaps the inward box overlaps the exact box (fp rounding is monotone,
// so fl(inward.max + thr) <= fl(max + thr) and fl(inward.min - thr) >= fl(min - thr)); only
// when the inward test fails does the traversal read the exact box from `bboxes`.
struct alignas(32) LeafRec {
    float mn[3], mx[3];
    int id, pad;
};

template <typename T>
__global__ void k_leaf_recs(const T *__restrict__ bb_all, const int64_t *__restrict__ tree_all, int64_t n,
                            LeafRec *__restrict__ recs_all, const DevCounters *__restrict__ ctr) {
    const int b = blockIdx.y;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (ctr[b].flags & DEV_BAD_TREE) return;
    const int64_t rows = 2 * n - 1, row = n - 1 + p;
    const T *bx = bb_all + (int64_t)b * rows * 6 + 6 * row;
    LeafRec r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = f32_up((double)bx[k]);
        r.mx[k] = f32_down((double)bx[k + 3]);
    }
    r.id = (int)tree_all[(int64_t)b * rows * 4 + 4 * row + 3];
    r.pad = 0;
    recs_all[(int64_t)b * n + p] = r;
}

template <typename T> __device__ __forceinline__ bool overlap_inward(const Box<T> &q, const LeafRec &r, T thr) {
    return (q.mn.x <= ((T)r.mx[0] + thr)) && (q.mx.x >= ((T)r.mn[0] - thr)) && (q.mn.y <= ((T)r.mx[1] + thr)) &&
           (q.mx.y >= ((T)r.mn[1] - thr)) && (q.mn.z <= ((T)r.mx[2] + thr)) && (q.mx.z >= ((T)r.mn[2] - thr));
}

// Start points of the traversal: the subtrees kFrontierDepth levels below the root, each
// with its fp32 box grown outward by the threshold (as in TNode), its rightmost leaf and
// its first leaf position.  Slot i follows the path of the bits of i from the root (1 =
// right); a leaf reached above that depth stands for itself in the slot of its leftmost
// path, and the other slots below it are empty (first = INT_MAX: no query enters them).
// A query runs one traversal per start point, so the launch has kFrontier times as many
// lanes.
struct Frontier {
    int count;               // slots in use: kFrontier, or 0 (no traversal)
    int ref[kFrontier];      // internal node index, or ~leaf position
    int rightmost[kFrontier];
    int first[kFrontier];    // first leaf position (in-order trees; 0 otherwise); INT_MAX: empty
    float box[kFrontier][6];
};

template <typename T>
__global__ __launch_bounds__(kFrontier) void k_frontier(const T *__restrict__ bb_all, const int64_t *__restrict__ tree_all,
                                                      const int *__restrict__ rm_all, const int *__restrict__ lo_all, int64_t n,
                                                      double thr, Frontier *__restrict__ fr_all, const DevCounters *__restrict__ ctr) {
    const int b = blockIdx.x, i = threadIdx.x;     // one thread per slot
    Frontier &f = fr_all[b];
    if (n < 2 || (ctr[b].flags & DEV_BAD_TREE)) {
        if (i == 0) f.count = 0;
        return;
    }
    if (i == 0) f.count = kFrontier;
    const int64_t rows = 2 * n - 1;
    const T *bb = bb_all + (int64_t)b * rows * 6;
    const int64_t *tree = tree_all + (int64_t)b * rows * 4;
    const int *rm = rm_all + (int64_t)b * (n - 1);
    const int *lo = lo_all + (int64_t)b * (n - 1);
    const bool ordered = !(ctr[b].flags & DEV_UNORDERED);
    int ref = 0;            // the root row
    bool used = true;
    for (int level = 0; level < kFrontierDepth; ++level) {
        if (ref < 0) {      // a leaf above the frontier depth: only its leftmost slot keeps it
            used = (i & ((1 << (kFrontierDepth - level)) - 1)) == 0;
            break;
        }
        const int64_t c = tree[4 * ref + 1 + ((i >> (kFrontierDepth - 1 - level)) & 1)];
        ref = c >= n - 1 ? ~(int)(c - (n - 1)) : (int)c;
    }
    f.ref[i] = ref;
    if (!used) {
        f.rightmost[i] = 0;
        f.first[i] = INT_MAX;
        for (int k = 0; k < 3; ++k) { f.box[i][k] = INFINITY; f.box[i][k + 3] = -INFINITY; }
        return;
    }
    const int64_t row = ref >= 0 ? ref : n - 1 + (int64_t)(~ref);
    f.rightmost[i] = ref >= 0 ? rm[ref] : ~ref;
    f.first[i] = ref >= 0 ? (ordered ? lo[ref] : 0) : ~ref;
    for (int k = 0; k < 3; ++k) {
        f.box[i][k] = f32_down((double)bb[6 * row + k] - thr);
        f.box[i][k + 3] = f32_up((double)bb[6 * row + k + 3] + thr);
    }
}

// fp32 box of each block of kTravThreads query leaves (the union of the lanes' outward-
// rounded query boxes; empty past the last leaf), so that a traversal block whose queries
// cannot enter its start point ends before it starts.
template <typename T>
__global__ __launch_bounds__(kTravThreads) void k_qblock_boxes(const T *__restrict__ bb_all, int64_t n, int nqb,
                                                               float *__restrict__ qbox_all, const DevCounters *__restrict__ ctr) {
    const int b = blockIdx.y;
    const int64_t p = (int64_t)blockIdx.x * kTravThreads + threadIdx.x;
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (p < n && !(ctr[b].flags & DEV_BAD_TREE)) {
        const T *q = bb_all + (int64_t)b * (2 * n - 1) * 6 + 6 * (n - 1 + p);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = f32_down((double)q[k]);
            v[k + 3] = f32_up((double)q[k + 3]);
        }
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = fminf(v[k], __shfl_xor(v[k], d, 64));
            v[k + 3] = fmaxf(v[k + 3], __shfl_xor(v[k + 3], d, 64));
        }
    }
    static_assert(kTravThreads == 64, "one wave per query block");
    if (threadIdx.x == 0) {
        float *o = qbox_all + ((int64_t)b * nqb + blockIdx.x) * 6;
#pragma unroll
        for (int k = 0; k < 6; ++k) o[k] = v[k];
    }
}

// ---- traversal ----------------------------------------------------------------------------
// collisions_static.cu:10-120 (ordering on) and collisions_continuous.cu:127-230,
// :1079-1100 (PARTIAL: ordering off, queries with to_check[position]).  A leaf child is
// reported after the reference's exact test checkOverlap(query box, leaf box, threshold) on
// the boxes stored in `bboxes`; the pair is (min id, max id).  The walk runs over the
// four-wide nodes (QNode): each visit tests up to four grandchildren, which halves the
// dependent node reads of a query against the binary nodes.
template <typename T, bool PARTIAL, int DEPTH>
__global__ __launch_bounds__(kTravThreads) void k_traverse(const T *__restrict__ bb_all, const int64_t *__restrict__ tree_all,
                                                           const QNode *__restrict__ qnodes_all,
                                                           const LeafRec *__restrict__ recs_all, int64_t n, T thr,
                                                           const uint8_t *__restrict__ to_check_all,
                                                           const int32_t *__restrict__ mesh_end,
                                                           int2 *__restrict__ pairs_all, int64_t cap,
                                                           DevCounters *__restrict__ ctr_all, const Frontier *__restrict__ fr_all,
                                                           const float *__restrict__ qbox_all) {
    // Pending right children: with the short stack as 16-bit offsets from the start point's
    // node (a subtree's nodes lie close together in the trees built here; a farther node
    // ends the traversal like an overflow, and the call runs again with the deep stack).
    using StackT = typename std::conditional<(DEPTH < kStackDeep), short, int>::type;
    __shared__ StackT stack[DEPTH][kTravThreads];
    constexpr int kStageCap = kStage + 128;   // an iteration that would overflow it flushes first
    __shared__ int stage[kTravThreads / 64][kStageCap];                             // the other triangle of each pair
    __shared__ int lqid[kTravThreads / 64][64];                                     // each lane's query triangle
    __shared__ unsigned short stag[kTravThreads / 64][kStageCap];                  // lane << 10 | rank within the lane
    __shared__ int loff[kTravThreads / 64][64];                         // a flush's offset of each lane
    __shared__ int fill[kTravThreads / 64];
    const int b = blockIdx.y, tid = threadIdx.x, w = tid >> 6;
    const int64_t rows = 2 * n - 1;
    const T *bb = bb_all + (int64_t)b * rows * 6;
    const int64_t *tree = tree_all + (int64_t)b * rows * 4;
    const QNode *qnodes = qnodes_all + (int64_t)b * (n - 1);
    const LeafRec *recs = recs_all + (int64_t)b * n;
    DevCounters *ctr = ctr_all + b;
    int2 *pairs = pairs_all + (int64_t)b * cap;

    // block -> (start point s, query block); query blocks in XCD-contiguous order (the
    // query-block count nqb is a multiple of kXcds)
    const int nqb = gridDim.x / kFrontier;
    const int s = blockIdx.x / nqb, local = blockIdx.x % nqb;
    const int64_t qblock = (int64_t)(local % kXcds) * (nqb / kXcds) + local / kXcds;
    const int64_t p = qblock * blockDim.x + tid;
    const Frontier &fr = fr_all[b];
    {
        // the whole block ends here when none of its queries can enter the start point (the
        // same box and rightmost tests as each lane's below, on the block's union box)
        const float *qb = qbox_all + ((int64_t)b * nqb + qblock) * 6;
        bool dead = s >= fr.count;
        if (!dead) {
            const float *fb = fr.box[s];
            dead = !(qb[0] <= fb[3] && qb[3] >= fb[0] && qb[1] <= fb[4] && qb[4] >= fb[1] && qb[2] <= fb[5] && qb[5] >= fb[2]);
            if (!PARTIAL) dead = dead || fr.rightmost[s] <= (int)(qblock * kTravThreads);
        }
        if (dead) return;
    }
    if ((tid & 63) == 0) fill[w] = 0;
    __syncthreads();
    bool active = p < n && n > 1 && !(ctr->flags & DEV_BAD_TREE) && s < fr.count;
    if (PARTIAL && active) active = to_check_all[(int64_t)b * n + p] != 0;
    const uint8_t *to_check = PARTIAL ? to_check_all + (int64_t)b * n : nullptr;

    // Query setup (inactive lanes: no query; they stay in the loop below as idle lanes).
    Box<T> q{};
    int qid = 0, lim = INT_MAX;
    float qx0 = 0, qy0 = 0, qz0 = 0, qx1 = 0, qy1 = 0, qz1 = 0;
    bool alive = false;
    if (active) {
        const int64_t qrow = n - 1 + p;
        q = ld_box(bb + 6 * qrow);
        qid = recs[p].id;
        qx0 = f32_down((double)q.mn.x); qy0 = f32_down((double)q.mn.y); qz0 = f32_down((double)q.mn.z);
        qx1 = f32_up((double)q.mx.x); qy1 = f32_up((double)q.mx.y); qz1 = f32_up((double)q.mx.z);
        // packed meshes: leaves at positions >= lim belong to later meshes
        lim = mesh_end ? mesh_end[p] : INT_MAX;
        // the start point itself: box, rightmost and first-leaf tests as for a child
        const float *fb = fr.box[s];
        alive = qx0 <= fb[3] && qx1 >= fb[0] && qy0 <= fb[4] && qy1 >= fb[1] && qz0 <= fb[5] && qz1 >= fb[2];
        if (!PARTIAL) alive = alive && fr.rightmost[s] > (int)p;
        alive = alive && fr.first[s] < lim;
    }
    const int lane = tid & 63;
    const int sbase = fr.ref[s];
    int sp = 0, node = sbase;
    int mine = 0;   // pairs this lane has staged since the last flush
    // Writes the wave's nf staged pairs grouped by lane (lane order, each lane's pairs in the
    // order found), so that the pair list keeps one query's pairs together.  Full wave.
    auto flush_stage = [&](int nf) {
        int incl = mine;    // inclusive scan of `mine` over the wave
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        loff[w][lane] = incl - mine;
        lqid[w][lane] = qid;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&ctr->cand, (unsigned long long)nf);
        base = readfirstlane_u64(base);
        for (int j = lane; j < nf; j += 64) {
            const unsigned t = stag[w][j];
            const unsigned long long d = base + (unsigned long long)(loff[w][t >> 10] + (int)(t & 1023u));
            if (d < (unsigned long long)cap) {
                const int qq = lqid[w][t >> 10], oo = stage[w][j];
                pairs[d] = make_int2(min(qq, oo), max(qq, oo));
            }
        }
        mine = 0;
    };
    // a pending subtree onto the stack; an overflow (or, with 16-bit entries, a node too far
    // from the start node) ends this lane and flags the call for the deep stack
    auto push = [&](int ref) {
        const int v = sizeof(StackT) < sizeof(int) ? ref - sbase : ref;
        const bool fits = sizeof(StackT) == sizeof(int) || (v >= -32768 && v <= 32767);
        if (sp >= DEPTH || !fits) { atomicOr(&ctr->flags, DEV_STACK); alive = false; }
        else stack[sp++][tid] = (StackT)v;
    };
    while (__any(alive)) {
        int o0 = -1, o1 = -1, o2 = -1, o3 = -1;     // pairs found in this iteration
        int r0 = 0, r1 = 0, r2 = 0, r3 = 0;         // children (entries) that may be descended into
        unsigned im = 0;                            // bit k: descend into r_k
        if (alive) {
            {
                QNode qn;
                if (node >= 0) {
                    qn = qnodes[node];
                } else {
                    // a start point that is a leaf: entry 0 is the leaf, with an unbounded box
                    qn.mnx = make_float4(-INFINITY, NAN, NAN, NAN);
                    qn.mxx = make_float4(INFINITY, NAN, NAN, NAN);
                    qn.mny = qn.mnx; qn.mnz = qn.mnx; qn.mxy = qn.mxx; qn.mxz = qn.mxx;
                    qn.ref = make_int4(node, 0, 0, 0);
                    qn.rm = make_int4(~node, INT_MIN, INT_MIN, INT_MIN);
                }
                const float mnx[4] = {qn.mnx.x, qn.mnx.y, qn.mnx.z, qn.mnx.w}, mxx[4] = {qn.mxx.x, qn.mxx.y, qn.mxx.z, qn.mxx.w};
                const float mny[4] = {qn.mny.x, qn.mny.y, qn.mny.z, qn.mny.w}, mxy[4] = {qn.mxy.x, qn.mxy.y, qn.mxy.z, qn.mxy.w};
                const float mnz[4] = {qn.mnz.x, qn.mnz.y, qn.mnz.z, qn.mnz.w}, mxz[4] = {qn.mxz.x, qn.mxz.y, qn.mxz.z, qn.mxz.w};
                const int rf[4] = {qn.ref.x, qn.ref.y, qn.ref.z, qn.ref.w}, rmv[4] = {qn.rm.x, qn.rm.y, qn.rm.z, qn.rm.w};
                int oid[4];
                bool leaf[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    bool hit = qx0 <= mxx[k] && qx1 >= mnx[k] && qy0 <= mxy[k] && qy1 >= mny[k] && qz0 <= mxz[k] && qz1 >= mnz[k];
                    if (!PARTIAL) hit = hit && rmv[k] > (int)p;
                    // entry k starts after entry k-1's last leaf; the node itself (so entry 0)
                    // is only entered when its first leaf is below lim
                    if (k > 0) hit = hit && rmv[k - 1] + 1 < lim;
                    leaf[k] = hit && rf[k] < 0;
                    if (hit && rf[k] >= 0) im |= 1u << k;
                }
                // The four leaf records are requested before the first is tested, so that
                // their loads are in flight together (record 0 stands in for entries that are
                // not hit leaves: no branch around the loads).  A leaf whose inward-rounded
                // record does not decide is tested on its exact box from `bboxes`, under a
                // wave-uniform branch with every lane loading a valid row: with that fallback
                // in a lane-divergent branch per leaf, the compiler lost the record's id in
                // the lanes that took it (their pairs were dropped).
                LeafRec lr[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) lr[k] = recs[leaf[k] ? ~rf[k] : 0];
                bool ov[4];
                bool need = false;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    ov[k] = leaf[k] && overlap_inward(q, lr[k], thr);
                    need = need || (leaf[k] && !ov[k]);
                }
                if (__any(need)) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const bool nk = leaf[k] && !ov[k];
                        const Box<T> eb = ld_box(bb + 6 * (n - 1 + (int64_t)(nk ? ~rf[k] : 0)));
                        ov[k] = ov[k] || (nk && overlap(q, eb, thr));
                    }
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int id = lr[k].id;
                    oid[k] = ov[k] && (!PARTIAL || !to_check[id] || qid < id) ? id : -1;
                }
                o0 = oid[0]; o1 = oid[1]; o2 = oid[2]; o3 = oid[3];
                r0 = rf[0]; r1 = rf[1]; r2 = rf[2]; r3 = rf[3];
            }
        }
        // stage this iteration's pairs in the wave's LDS buffer (tagged with lane and rank)
        {
            const unsigned c = (o0 >= 0 ? 1u : 0u) + (o1 >= 0 ? 1u : 0u) + (o2 >= 0 ? 1u : 0u) + (o3 >= 0 ? 1u : 0u);
            const unsigned long long b0 = __ballot(c & 1u), b1 = __ballot(c & 2u), b2 = __ballot(c & 4u);
            const unsigned below = lanes_below(b0) + 2u * lanes_below(b1) + 4u * lanes_below(b2);
            const unsigned tot = (unsigned)__popcll(b0) + 2u * (unsigned)__popcll(b1) + 4u * (unsigned)__popcll(b2);
            if (tot) {
                int f = fill[w];
                if (f + (int)tot > kStageCap) {     // would not fit: write out what is staged first
                    flush_stage(f);
                    f = 0;
                }
                int pos = f + (int)below;
                const unsigned lane_tag = (unsigned)lane << 10;
                if (o0 >= 0) { stage[w][pos] = o0; stag[w][pos++] = (unsigned short)(lane_tag | mine++); }
                if (o1 >= 0) { stage[w][pos] = o1; stag[w][pos++] = (unsigned short)(lane_tag | mine++); }
                if (o2 >= 0) { stage[w][pos] = o2; stag[w][pos++] = (unsigned short)(lane_tag | mine++); }
                if (o3 >= 0) { stage[w][pos] = o3; stag[w][pos] = (unsigned short)(lane_tag | mine++); }
                const int nf = f + (int)tot;
                if (nf >= kStage) {
                    flush_stage(nf);
                    if (lane == 0) fill[w] = 0;
                } else {
                    if (lane == 0) fill[w] = nf;
                }
            }
        }
        if (alive) {
            if (im == 0) {
                if (sp == 0) alive = false;
                else node = sizeof(StackT) < sizeof(int) ? sbase + (int)stack[--sp][tid] : (int)stack[--sp][tid];
            } else {
                // descend into the leftmost; the others wait on the stack, right to left, so
                // that the next one to the right comes off first
                const int lo = __ffs(im) - 1;
                node = lo == 0 ? r0 : lo == 1 ? r1 : lo == 2 ? r2 : r3;
                if ((im & 8u) && lo != 3) push(r3);
                if (alive && (im & 4u) && lo != 2) push(r2);
                if (alive && (im & 2u) && lo != 1) push(r1);
            }
        }
    }
    // final flush of the wave's staged pairs
    {
        const int f = fill[w];
        if (f > 0) flush_stage(f);
    }
}

// ---- sub-primitive filter -----------------------------------------------------------------
// Candidate pairs -> the pairs of sub-primitives (9 edge-edge, 6 vertex-face) whose boxes
// overlap, with PP pairs per thread and one reservation per PP * kThreads pairs (the item
// counter is one contended device address: every reservation is a returning atomic on it).
//   SUB_MOVING: swept boxes (collisions_continuous.cu:234-318, :340-396) -> packed items;
//   SUB_STATIC: boxes grown by the threshold (proximity_static.cu:11-118) -> packed items;
//   SUB_MINCOL: swept boxes, written as (x, y, type) rows (find_minimal_candidates, :321-338).
// Items stay in pair order: pair u * kThreads + t of a tile comes after every pair
// u' * kThreads + t' < u * kThreads + t.  SUB_MOVING writes a tile's edge-edge items first and
// its vertex-face items after them, each in pair order, so that the narrow phase's 64-item
// tiles hold one kind except where the kind changes (about 3% of them on the 130k cloth),
// while a pair's items of both kinds stay close together in the list (their vertex records
// are read again while still in the caches).  The moving masks come from the vertex records
// (moving_mask_axes).  A tile's items are assembled in LDS (a 2-byte entry per item: pair in
// tile and type) and written as one contiguous run.
enum { SUB_MOVING = 0, SUB_STATIC = 1, SUB_MINCOL = 2 };
// Exclusive block-wide scan of cnt (< 16) over the NT threads in thread order, without a
// reservation: the thread's offset; the block's sum in `total`.  Every thread calls it.
template <int NT = kThreads>
__device__ __forceinline__ unsigned block_scan_excl(unsigned cnt, unsigned &total) {
    __shared__ unsigned s_part[NT / 64 + 1];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned long long b0 = __ballot(cnt & 1u), b1 = __ballot(cnt & 2u), b2 = __ballot(cnt & 4u), b3 = __ballot(cnt & 8u);
    unsigned below = lanes_below(b0) + 2u * lanes_below(b1) + 4u * lanes_below(b2) + 8u * lanes_below(b3);
    if (lane == 0) s_part[w] = (unsigned)(__popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2) + 8 * __popcll(b3));
    __syncthreads();
    unsigned before = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < NT / 64; ++q) {
        const unsigned v = s_part[q];
        if (q < w) before += v;
        tot += v;
    }
    __syncthreads();
    total = tot;
    return before + below;
}

// Two exclusive block-wide scans in one pass: counts a < 16 and b < 8 per thread; returns the
// thread's offsets packed as (a offset) | (b offset) << 16 and the block's sums packed the
// same way in `total` (a block sum stays below 2^16: at most 15 * NT).  Every thread calls it.
template <int NT = kThreads>
__device__ __forceinline__ unsigned block_scan_excl2(unsigned a, unsigned b, unsigned &total) {
    static_assert(15 * NT < 65536, "packed halves do not carry");
    __shared__ unsigned s_part[NT / 64 + 1];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned below = 0, mine = 0;
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) {
        const unsigned long long m = __ballot((a >> bit) & 1u);
        below += lanes_below(m) << bit;
        mine += (unsigned)__popcll(m) << bit;
    }
#pragma unroll
    for (int bit = 0; bit < 3; ++bit) {
        const unsigned long long m = __ballot((b >> bit) & 1u);
        below += lanes_below(m) << (16 + bit);
        mine += (unsigned)__popcll(m) << (16 + bit);
    }
    if (lane == 0) s_part[w] = mine;
    __syncthreads();
    unsigned before = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < NT / 64; ++q) {
        const unsigned v = s_part[q];
        if (q < w) before += v;
        tot += v;
    }
    __syncthreads();
    total = tot;
    return before + below;
}

template <typename T, int KIND, int PP>
__global__ __launch_bounds__(kThreads) void k_subprim(const int2 *__restrict__ pairs_all, int64_t capK, int64_t c0, int64_t c1,
                                                      const T *__restrict__ src_all, int64_t n, T thr,
                                                      void *__restrict__ out_all, int64_t capM,
                                                      DevCounters *__restrict__ ctr_all) {
    static_assert(PP >= 1 && PP * kThreads <= 4096, "pair index and type fit 16 bits");
    // (pair in tile) << 4 | type; SUB_MOVING: edge-edge entries from the front, vertex-face
    // entries from the back (at most 9 and 6 per pair: the two runs never meet)
    constexpr int kEnt = 15 * PP * kThreads;
    constexpr bool SPLIT = KIND == SUB_MOVING;
    __shared__ unsigned short s_ent[kEnt];
    __shared__ int2 s_pair[PP * kThreads];
    __shared__ unsigned long long s_base;
    const int b = blockIdx.y, tid = threadIdx.x;
    DevCounters *ctr = ctr_all + b;
    const unsigned long long kc = ctr->cand;
    int64_t K = (int64_t)(kc < (unsigned long long)capK ? kc : (unsigned long long)capK);
    if (K > c1) K = c1;     // this launch covers the pairs [c0, c1)
    const int2 *pairs = pairs_all + (int64_t)b * capK;
    // SUB_STATIC: the t0 soup; otherwise the vertex records [x0 y0 z0 dx dy dz] per corner
    const T *src = src_all + (int64_t)b * n * (KIND == SUB_STATIC ? 9 : 18);
    // SUB_MOVING: 32-bit items (pair index within the chunk) << 4 | type; the others 64-bit
    int64_t *out = SPLIT ? nullptr : static_cast<int64_t *>(out_all) + (int64_t)b * capM * (KIND == SUB_MINCOL ? 3 : 1);
    uint32_t *out32 = SPLIT ? static_cast<uint32_t *>(out_all) + (int64_t)b * capM : nullptr;
    constexpr int kTile = PP * kThreads;
    const int64_t ntiles = K > c0 ? (K - c0 + kTile - 1) / kTile : 0;
    CCC_FOR_TILES(tile, ntiles) {
        unsigned tot = 0, tot_vf = 0;
#pragma unroll
        for (int u = 0; u < PP; ++u) {
            const int64_t i = c0 + tile * kTile + u * kThreads + tid;
            uint32_t mask = 0;
            int x = 0, y = 0;
            if (i < K) {
                const int2 pr = pairs[i];
                x = pr.x;
                y = pr.y;
                if (KIND == SUB_STATIC) {
                    V3<T> a[3], c[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) { a[k] = ld3(src + 9 * (int64_t)x + 3 * k); c[k] = ld3(src + 9 * (int64_t)y + 3 * k); }
                    if (!share_vertex(a, c)) mask = static_mask(a, c, thr);
                } else {
                    mask = moving_mask_axes(src, x, y);
                }
            }
            if (!SPLIT) s_pair[u * kThreads + tid] = make_int2(x, y);
            if (SPLIT) {
                uint32_t mee = mask & 0x1ffu, mvf = mask >> 9;
                unsigned t;
                const unsigned k2 = block_scan_excl2((unsigned)__popc(mee), (unsigned)__popc(mvf), t);
                unsigned ke = tot + (k2 & 0xffffu), kv = tot_vf + (k2 >> 16);
                tot += t & 0xffffu;
                tot_vf += t >> 16;
                while (mee) {
                    const int type = __ffs(mee) - 1;
                    mee &= mee - 1;
                    s_ent[ke++] = (unsigned short)(((u * kThreads + tid) << 4) | type);
                }
                while (mvf) {
                    const int type = 9 + __ffs(mvf) - 1;
                    mvf &= mvf - 1;
                    s_ent[kEnt - 1 - kv++] = (unsigned short)(((u * kThreads + tid) << 4) | type);
                }
            } else {
                unsigned t;
                unsigned k = tot + block_scan_excl((unsigned)__popc(mask), t);
                tot += t;
                while (mask) {
                    const int type = __ffs(mask) - 1;
                    mask &= mask - 1;
                    s_ent[k++] = (unsigned short)(((u * kThreads + tid) << 4) | type);
                }
            }
        }
        if (tid == 0) {
            const unsigned all = tot + tot_vf;
            s_base = all ? atomicAdd(&ctr->items, (unsigned long long)all) : 0ull;
        }
        __syncthreads();
        const unsigned long long base = s_base;
        const unsigned long long room = base < (unsigned long long)capM ? (unsigned long long)capM - base : 0ull;
        const unsigned lim = room < tot ? (unsigned)room : tot;
        const uint32_t tile_pair = (uint32_t)(tile * kTile);     // < 2^28: the chunk's pairs
        for (unsigned k = tid; k < lim; k += kThreads) {
            const unsigned e = s_ent[k];
            const unsigned long long slot = base + k;
            if (SPLIT) {
                out32[slot] = ((tile_pair + (e >> 4)) << 4) | (e & 15u);
            } else {
                const int2 pr = s_pair[e >> 4];
                if (KIND == SUB_MINCOL) {
                    out[3 * slot] = pr.x; out[3 * slot + 1] = pr.y; out[3 * slot + 2] = (int64_t)(e & 15u);
                } else {
                    out[slot] = pack_item(pr.x, pr.y, (int)(e & 15u));
                }
            }
        }
        if (SPLIT) {
            // vertex-face item j of the tile follows its tot edge-edge items
            const unsigned long long bv = base + tot;
            const unsigned long long rv = bv < (unsigned long long)capM ? (unsigned long long)capM - bv : 0ull;
            const unsigned limv = rv < tot_vf ? (unsigned)rv : tot_vf;
            for (unsigned k = tid; k < limv; k += kThreads) {
                const unsigned e = s_ent[kEnt - 1 - k];
                out32[bv + k] = ((tile_pair + (e >> 4)) << 4) | (e & 15u);
            }
        }
        __syncthreads();
    }
}

// ---- fused narrow phase --------------------------------------------------------------------
// Per (x, y, type) item: tetrahedron (collisions_continuous.cu:36-124), coefficients
// (:409-445), solveCubic (:511-613), flatten of roots in [0, 1] (:732-765, including
// duplicates), filterCollisions (:640-684); then
//   COLLECT: rows (x, y, type) + root, truncated at the output capacity (:1028-1063);
//   IMPULSE: the accepted collisions of the block are staged in LDS and, once at least a
//            block's worth has gathered, every thread takes one and scatters its impulse
//            with atomics into dv / dx / cnt (collisions_impulses.cu:1312-1366).
// The tetrahedron is gathered again for the filter so that it is not live across the
// cubic solver (fewer VGPRs, more waves per SIMD).
// Vertex records [x0 y0 z0 dx dy dz] (d = x1 - x0) per triangle corner from the t0 / t1 soups.
template <typename T>
__global__ void k_vertex_records(const T *__restrict__ t0_all, const T *__restrict__ t1_all, int64_t corners,
                                 T *__restrict__ tv_all) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // corner of batch element blockIdx.y
    if (i >= corners) return;
    const int64_t b = blockIdx.y;
    const T *a = t0_all + (b * corners + i) * 3, *c = t1_all + (b * corners + i) * 3;
    VRec<T> r;
    r.a[0] = a[0]; r.a[1] = a[1]; r.a[2] = a[2];
    r.a[3] = c[0] - a[0]; r.a[4] = c[1] - a[1]; r.a[5] = c[2] - a[2];
    reinterpret_cast<VRec<T> *>(tv_all)[b * corners + i] = r;
}

enum { MODE_COLLECT = 0, MODE_IMPULSE = 1 };
// Threads (items) per k_ccd tile, in both modes: one wave per block, so that its phases need
// no workgroup barrier (k_ccd per step on the 130k cloth: 4 waves per block 206 ms, 2 waves
// 193 then 175, 1 wave 172.5 with barriers; the collect call 20.8 -> 17.1 ms once its tiles
// were one wave too).
constexpr int kCcdThreads = 64;
template <int MODE> constexpr int ccd_threads() { return kCcdThreads; }   // one wavefront per workgroup in both modes
// Staged collisions are scattered kFlush at a time, one lane each for their impulse
// parameters and in two halves of kFlush / 2 for the corner sums; fewer than kFlush stay
// staged between tiles and a tile adds at most 3 * kCcdThreads.
constexpr int kFlush = kCcdThreads;
constexpr int kHalfFlush = kFlush / 2;
constexpr int kImpStage = kFlush + 3 * kCcdThreads;

template <typename T>
__device__ __forceinline__ void scatter_impulse(const MTet<T> &m, int type, T root, int64_t x, int64_t y, const T *masses,
                                                T *dv, T *dx, int *cnt) {
    T q[4];
    int side[4], idx[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        type_corner(type, k, side[k], idx[k]);
        q[k] = masses[3 * (side[k] ? y : x) + idx[k]];
    }
    const ImpulseParams<T> ip = collision_impulse(m, type, root, q);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        V3<T> imp, imd;
        corner_impulse(ip, type, k, m.dv[k], imp, imd);
        const int64_t c = 3 * (side[k] ? y : x) + idx[k];
        atomicAdd(dv + 3 * c + 0, imp.x);
        atomicAdd(dv + 3 * c + 1, imp.y);
        atomicAdd(dv + 3 * c + 2, imp.z);
        atomicAdd(dx + 3 * c + 0, imd.x);
        atomicAdd(dx + 3 * c + 1, imd.y);
        atomicAdd(dx + 3 * c + 2, imd.z);
        atomicAdd(cnt + c, 1);
    }
}

// ---- order-independent impulse sums (deterministic mode) ----------------------------------
// A finite value v with |v| < 2^45 is held as the integer trunc(|v| * 2^80) with v's sign,
// split into three 42-bit limbs; every limb of every contribution is added with a 64-bit
// integer atomic, which is exact and associative, so the sums do not depend on the order of
// the additions (2^21 additions per limb fit in 64 bits).  Values that cannot be held set
// flag bits in a fourth word instead: 1 NaN, 2 +inf, 4 -inf, 8 out of range; the result
// is then what an fp64 sum gives (NaN, or the infinity) and NaN when out of range.
// Per triangle corner: 6 components (dv xyz, dx xyz) x 4 words.
constexpr int kDetWords = kDetWordsPerCorner;
constexpr double kTwo42 = 4398046511104.0, kTwo45 = 35184372088832.0, kTwo80 = 1208925819614629174706176.0;

// The three limbs of v (returns 0), or its flag bit when v cannot be held.
__device__ __forceinline__ unsigned det_split(double v, long long &l0, long long &l1, long long &l2) {
    if (!(fabs(v) < kTwo45))       // NaN fails this test too
        return v != v ? 1u : v == (double)INFINITY ? 2u : v == -(double)INFINITY ? 4u : 8u;
    const double sc = fabs(v) * kTwo80;                  // exact: a power-of-two scaling, < 2^125
    const double hi = floor(sc / (kTwo42 * kTwo42));    // < 2^41; exact
    const double r1 = sc - hi * (kTwo42 * kTwo42);      // the bits of sc below 2^84: exact
    const double mid = floor(r1 / kTwo42);              // < 2^42
    const double lo = floor(r1 - mid * kTwo42);         // < 2^42, fraction truncated
    l0 = (long long)lo;
    l1 = (long long)mid;
    l2 = (long long)hi;
    if (v < 0) { l0 = -l0; l1 = -l1; l2 = -l2; }
    return 0u;
}

__device__ __forceinline__ void det_add1(long long *a, double v) {
    long long l0, l1, l2;
    unsigned long long *u = reinterpret_cast<unsigned long long *>(a);
    const unsigned f = det_split(v, l0, l1, l2);
    if (f) {
        atomicOr(u + 3, (unsigned long long)f);
        return;
    }
    if (l0) atomicAdd(u + 0, (unsigned long long)l0);
    if (l1) atomicAdd(u + 1, (unsigned long long)l1);
    if (l2) atomicAdd(u + 2, (unsigned long long)l2);
}

// Limbs -> dv / dx of every corner: carries normalise the two low limbs into [0, 2^42),
// then hi * 2^84 + mid * 2^42 + lo is formed in double in that order and scaled by 2^-80.
template <typename T>
__global__ void k_det_finalize(const long long *__restrict__ acc, int64_t corners, T *__restrict__ dv, T *__restrict__ dx) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= corners) return;
    const long long *a = acc + kDetWords * c;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        long long l0 = a[4 * k], l1 = a[4 * k + 1], l2 = a[4 * k + 2];
        const unsigned long long f = (unsigned long long)a[4 * k + 3];
        double v;
        if (f & 9ull) v = (double)NAN;
        else if ((f & 6ull) == 6ull) v = (double)NAN;
        else if (f & 2ull) v = (double)INFINITY;
        else if (f & 4ull) v = -(double)INFINITY;
        else {
            const long long c0 = l0 >> 42;
            l0 -= c0 * (1ll << 42);
            l1 += c0;
            const long long c1 = l1 >> 42;
            l1 -= c1 * (1ll << 42);
            l2 += c1;
            v = ((double)l2 * (kTwo42 * kTwo42) + (double)l1 * kTwo42 + (double)l0) / kTwo80;
        }
        T *o = k < 3 ? dv + 3 * c + k : dx + 3 * c + (k - 3);
        *o = (T)v;
    }
}

// Per-block sums of corner impulses: open addressing on the corner key 3 * triangle +
// corner, kImpSlots slots (a power of two), linear probing up to kImpProbe slots; a
// contribution that finds no slot goes to the global arrays directly.  A half flush adds
// 4 kHalfFlush = 128 corner shares, on fewer distinct corners (neighbouring collisions share
// corners); a whole flush at once overfilled the table (every share that found no slot
// spent the whole probe window first): k_ccd 16.2 -> 23.4 ms per pass on the 130k cloth.
constexpr int kImpSlots = 64;
constexpr int kImpProbe = 16;
constexpr int log2_exact(int v) { return v <= 1 ? 0 : 1 + log2_exact(v / 2); }
constexpr int kImpBits = log2_exact(kImpSlots);

template <typename T> struct ImpTable {
    int *key;
    int *count;
    T *acc;     // [slot][dv.xyz, dx.xyz]

    static constexpr size_t bytes() { return (size_t)kImpSlots * (2 * sizeof(int) + 6 * sizeof(T)); }
    __device__ static ImpTable over(unsigned char *raw) {
        ImpTable t;
        t.acc = reinterpret_cast<T *>(raw);
        t.key = reinterpret_cast<int *>(raw + (size_t)kImpSlots * 6 * sizeof(T));
        t.count = t.key + kImpSlots;
        return t;
    }
    __device__ void clear(int tid) {
        for (int s = tid; s < kImpSlots; s += kCcdThreads) {
            key[s] = -1;
            count[s] = 0;
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[6 * s + c] = (T)0;
        }
    }
    // slot of `k`, inserting it if absent; -1 when the probe window is full
    __device__ int slot(int k) {
        int h = (int)(((unsigned)k * 2654435761u) >> (32 - kImpBits)) & (kImpSlots - 1);
        for (int p = 0; p < kImpProbe; ++p) {
            const int old = atomicCAS(&key[h], -1, k);
            if (old == -1 || old == k) return h;
            h = (h + 1) & (kImpSlots - 1);
        }
        return -1;
    }
    // adds one corner's impulse and position correction; to the global arrays directly
    // when the probe window is full
    __device__ void add(int64_t c, const V3<T> &imp, const V3<T> &imd, T *dv, T *dx, int *cnt) {
        const int s = slot((int)c);
        if (s >= 0) {
            atomicAdd(&acc[6 * s + 0], imp.x);
            atomicAdd(&acc[6 * s + 1], imp.y);
            atomicAdd(&acc[6 * s + 2], imp.z);
            atomicAdd(&acc[6 * s + 3], imd.x);
            atomicAdd(&acc[6 * s + 4], imd.y);
            atomicAdd(&acc[6 * s + 5], imd.z);
            atomicAdd(&count[s], 1);
        } else {
            atomicAdd(dv + 3 * c + 0, imp.x);
            atomicAdd(dv + 3 * c + 1, imp.y);
            atomicAdd(dv + 3 * c + 2, imp.z);
            atomicAdd(dx + 3 * c + 0, imd.x);
            atomicAdd(dx + 3 * c + 1, imd.y);
            atomicAdd(dx + 3 * c + 2, imd.z);
            atomicAdd(cnt + c, 1);
        }
    }
    __device__ void drain(int tid, T *dv, T *dx, int *cnt) {
        for (int s = tid; s < kImpSlots; s += kCcdThreads) {
            const int k = key[s];
            if (k < 0) continue;
            atomicAdd(dv + 3 * (int64_t)k + 0, acc[6 * s + 0]);
            atomicAdd(dv + 3 * (int64_t)k + 1, acc[6 * s + 1]);
            atomicAdd(dv + 3 * (int64_t)k + 2, acc[6 * s + 2]);
            atomicAdd(dx + 3 * (int64_t)k + 0, acc[6 * s + 3]);
            atomicAdd(dx + 3 * (int64_t)k + 1, acc[6 * s + 4]);
            atomicAdd(dx + 3 * (int64_t)k + 2, acc[6 * s + 5]);
            atomicAdd(cnt + k, count[s]);
        }
    }
};
static_assert(kImpSlots == 1 << kImpBits, "the slot count is a power of two");

// The deterministic mode's per-block table: the same slots and probing as ImpTable, holding
// the two low limbs of each component as 64-bit integers (LDS integer atomics: exact, so the
// table's merge is order-independent too).  High limbs and flags, which cloth-scale
// corrections rarely have, go to the global limbs directly; each occupied slot then adds its
// two limbs per component and its count with global atomics.
struct DetTable {
    int *key;
    int *count;
    unsigned long long *acc;    // [slot][component][lo, mid]

    static constexpr size_t bytes() { return (size_t)kImpSlots * (2 * sizeof(int) + 12 * sizeof(unsigned long long)); }
    __device__ static DetTable over(unsigned char *raw) {
        DetTable t;
        t.acc = reinterpret_cast<unsigned long long *>(raw);
        t.key = reinterpret_cast<int *>(raw + (size_t)kImpSlots * 12 * sizeof(unsigned long long));
        t.count = t.key + kImpSlots;
        return t;
    }
    __device__ void clear(int tid) {
        for (int s = tid; s < kImpSlots; s += kCcdThreads) {
            key[s] = -1;
            count[s] = 0;
#pragma unroll
            for (int c = 0; c < 12; ++c) acc[12 * s + c] = 0ull;
        }
    }
    __device__ int slot(int k) {
        int h = (int)(((unsigned)k * 2654435761u) >> (32 - kImpBits)) & (kImpSlots - 1);
        for (int p = 0; p < kImpProbe; ++p) {
            const int old = atomicCAS(&key[h], -1, k);
            if (old == -1 || old == k) return h;
            h = (h + 1) & (kImpSlots - 1);
        }
        return -1;
    }
    __device__ void add(int64_t c, const double v[6], long long *det, int *cnt) {
        const int s = slot((int)c);
        unsigned long long *g = reinterpret_cast<unsigned long long *>(det + kDetWords * c);
        if (s < 0) {
#pragma unroll
            for (int k = 0; k < 6; ++k) det_add1(det + kDetWords * c + 4 * k, v[k]);
            atomicAdd(cnt + c, 1);
            return;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            long long l0, l1, l2;
            const unsigned f = det_split(v[k], l0, l1, l2);
            if (f) {
                atomicOr(g + 4 * k + 3, (unsigned long long)f);
                continue;
            }
            if (l0) atomicAdd(&acc[12 * s + 2 * k], (unsigned long long)l0);
            if (l1) atomicAdd(&acc[12 * s + 2 * k + 1], (unsigned long long)l1);
            if (l2) atomicAdd(g + 4 * k + 2, (unsigned long long)l2);
        }
        atomicAdd(&count[s], 1);
    }
    __device__ void drain(int tid, long long *det, int *cnt) {
        for (int s = tid; s < kImpSlots; s += kCcdThreads) {
            const int k = key[s];
            if (k < 0) continue;
            unsigned long long *g = reinterpret_cast<unsigned long long *>(det + kDetWords * (int64_t)k);
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const unsigned long long lo = acc[12 * s + 2 * c], mid = acc[12 * s + 2 * c + 1];
                if (lo) atomicAdd(g + 4 * c, lo);
                if (mid) atomicAdd(g + 4 * c + 1, mid);
            }
            atomicAdd(cnt + k, count[s]);
        }
    }
};

// The same pointer, offset by a zero the optimiser cannot see through, so that loads
// through it are not merged with earlier loads of the same addresses.  The pointer itself
// keeps its provenance (a kernel argument), so its loads stay global-memory loads rather
// than flat ones.
template <typename P> __device__ __forceinline__ P *opaque_ptr(P *p) {
    int64_t z;
    asm volatile("s_mov_b64 %0, 0" : "=s"(z));
    return p + z;
}

// One tile of kCcdThreads items per step, in four phases.  Every workgroup of this kernel is
// one wavefront, so its list counters are wave-uniform registers (a lane's place in a list
// comes from a ballot) and the LDS exchanges between phases need only wave-level ordering
// (wave_sync), not a workgroup barrier, whose fence would also wait for the wave's
// outstanding global loads and atomics.
//   A  item ((pair index within the chunk) << 4 | type, its candidate pair read one tile
//      ahead) -> tetrahedron -> coefficients -> cubic_prepare, in registers;
//   B  the lane's roots inside [0, 1] (at most three) go to an LDS root list (item, root),
//      placed by a wave-wide scan of the counts;
//   C  filterCollisions per root in whole rounds of kCcdThreads roots (the tetrahedron is
//      gathered again); fewer than a round are carried to the next tile;
//   D  (impulse mode) accepted collisions wait in an LDS stage; kFlush at a time they are
//      scattered through an LDS hash table that reuses the root-list memory: one lane per
//      collision computes its impulse parameters (gathering the tetrahedron once for the
//      weights and once for the rest, so that this step needs no more registers than the
//      CCD phases); lanes l and l ^ 32 exchange them, and each half of the chunk adds its
//      corners' shares to the table and drains it.  The scatter stays in this
//      kernel: run as a kernel of its own over per-workgroup collision lists it took 14 ms
//      per pass on the 130k cloth against ~6 here, where other workgroups' fp64 work hides
//      its atomic latency.
//      (collect mode) accepted rows are written kCollectFlush at a time with one
//      reservation (the row counter is one contended device address).
// Each lane's item of the next tile is loaded while this tile is solved.  The cubic's
// branches diverge within the wave whatever the order of the items, so the items are not
// grouped by branch.  Occupancy: the impulse mode four waves per SIMD (128 VGPRs).  At five
// waves the impulse mode spilled (once about a hundred registers and 40% slower; with the
// current code 32 per lane and 40% slower again); at three 13% slower.  Staging each wave's
// query triangle in LDS (so that half the corner reads leave the vector memory path) made
// the impulse mode spill and took 196-209 ms per step instead of 177.
constexpr int kCollectFlush = 64;

// Orders the wave's LDS accesses before this point against those after it.  The lanes of a
// wavefront run in lockstep and the LDS serves one wave's accesses in issue order, so no
// instruction is needed: the fences and the barrier only keep the compiler from moving
// memory accesses across the point.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Waves per SIMD asked of the compiler: the deterministic table's LDS (12.3 KB) allows only
// 3.25 waves, and at four its registers spilled (37.28 against 37.06 ms per pass at three).
template <typename T, int MODE, bool DET>
__global__ __launch_bounds__(kCcdThreads, MODE == MODE_IMPULSE ? (DET ? 3 : 4) : 5) void k_ccd(const uint32_t *__restrict__ items_all, int64_t capM,
                                                  const int2 *__restrict__ pairs_all, int64_t capK, int64_t c0,
                                                  const T *__restrict__ tv_all,
                                                  const T *__restrict__ masses_all, int64_t n,
                                                  int64_t *__restrict__ coll_all, T *__restrict__ roots_all, int64_t out_cap,
                                                  T *__restrict__ dv_all, T *__restrict__ dx_all, int *__restrict__ cnt_all,
                                                  long long *__restrict__ det_all, DevCounters *__restrict__ ctr_all) {
    constexpr int NT = kCcdThreads;     // items per tile: the lanes of the one wavefront
    static_assert(NT == 64, "one wavefront per workgroup");
    constexpr int kRootCapT = 3 * NT;
    // the root list (phases B-C) shares its LDS with the impulse table of phase D
    constexpr size_t kListBytes = (sizeof(int64_t) + sizeof(T)) * kRootCapT;
    constexpr size_t kTableBytes = DET ? DetTable::bytes() : ImpTable<T>::bytes();
    constexpr size_t kScatterBytes = MODE == MODE_IMPULSE ? kTableBytes : 0;
    __shared__ __align__(16) unsigned char s_raw[kListBytes > kScatterBytes ? kListBytes : kScatterBytes];
    int64_t *s_ritem = reinterpret_cast<int64_t *>(s_raw);
    T *s_rroot = reinterpret_cast<T *>(s_raw + sizeof(int64_t) * kRootCapT);
    // accepted collisions wait here: impulse mode until kFlush of them are scattered,
    // collect mode until kCollectFlush of them are written as output rows
    constexpr int kStageCap = MODE == MODE_IMPULSE ? kImpStage : kCollectFlush + 3 * NT;
    __shared__ int64_t s_item[kStageCap];
    __shared__ T s_root[kStageCap];
    // roots waiting for a whole round of the filter phase: fewer than NT, carried from tile to
    // tile as (item, root) so that the filter runs full waves
    __shared__ int64_t s_citem[NT];
    __shared__ T s_croot[NT];
    const int b = blockIdx.y, tid = threadIdx.x;
    DevCounters *ctr = ctr_all + b;
    const unsigned long long mc = ctr->items;
    const int64_t M = (int64_t)(mc < (unsigned long long)capM ? mc : (unsigned long long)capM);
    const uint32_t *items = items_all + (int64_t)b * capM;
    const int2 *pairs = pairs_all + (int64_t)b * capK + c0;     // this chunk's candidate pairs
    const T *tv = tv_all + (int64_t)b * n * 18;
    const T *masses = MODE == MODE_IMPULSE ? masses_all + (int64_t)b * n * 3 : nullptr;
    T *dv = MODE == MODE_IMPULSE ? dv_all + (int64_t)b * n * 9 : nullptr;
    T *dx = MODE == MODE_IMPULSE ? dx_all + (int64_t)b * n * 9 : nullptr;
    int *cnt = MODE == MODE_IMPULSE ? cnt_all + (int64_t)b * n * 3 : nullptr;
    long long *det = DET ? det_all + (int64_t)b * n * 3 * kDetWords : nullptr;
    // wave-uniform counts: roots found, collisions scattered, staged collisions, carried roots
    unsigned long long nroots = 0, ncoll = 0;
    int fill = 0, ncarry = 0;

    // Staged collisions are scattered kFlush at a time: their corner impulses are first
    // summed per corner in an LDS hash table, then each occupied slot adds its sums to dv /
    // dx / cnt with one set of global atomics.
    static_assert(kScatterBytes <= sizeof(s_raw), "the impulse table fits the root-list buffer");
    static_assert(kFlush == NT && kHalfFlush == 32, "one lane per collision of a flush; halves of 32 lanes");
    ImpTable<T> tab = ImpTable<T>::over(s_raw);
    DetTable dtab = DetTable::over(s_raw);
    // Scatters staged collisions [0, f) in chunks of kFlush.  One lane per collision computes
    // its impulse parameters; the parameters of lanes l and l ^ 32 are then exchanged, and
    // each half of the chunk is summed in the LDS table and drained: collision j of the half
    // has corners 0 and 2 added by lane j and corners 1 and 3 by lane j + 32 (the corner
    // steps hold few registers, so the parameter step alone bounds this phase's register
    // use).
    auto flush = [&](int f) {
        for (int base = 0; base < f; base += kFlush) {
            const int nf = f - base < kFlush ? f - base : kFlush;
            ImpulseParams<T> ip{};
            if (tid < nf) {
                int64_t x, y;
                int type;
                unpack_item(s_item[base + tid], x, y, type);
                const T root = s_root[base + tid];
                // the tetrahedron is gathered once per step (from L2) rather than held
                // across both: that keeps this phase within the CCD phases' registers
                const ImpulseWeights<T> w = impulse_weights(gather_mtet_rec(opaque_ptr(tv), x, y, type), type, root);
                T q[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    int side, idx;
                    type_corner(type, k, side, idx);
                    q[k] = masses[3 * (side ? y : x) + idx];
                }
                ip = impulse_params(gather_mtet_rec(opaque_ptr(tv), x, y, type), type, root, w, q);
            }
            ImpulseParams<T> pp;      // lane l ^ 32's parameters
            pp.nrm.x = __shfl_xor(ip.nrm.x, 32);
            pp.nrm.y = __shfl_xor(ip.nrm.y, 32);
            pp.nrm.z = __shfl_xor(ip.nrm.z, 32);
            pp.im0 = __shfl_xor(ip.im0, 32);
            pp.im1 = __shfl_xor(ip.im1, 32);
            pp.keep = __shfl_xor(ip.keep, 32);
            const int j = tid & (kHalfFlush - 1), hi = tid >> 5;
            for (int h = 0; h * kHalfFlush < nf; ++h) {
                const int nh = nf - h * kHalfFlush < kHalfFlush ? nf - h * kHalfFlush : kHalfFlush;
                if (DET) dtab.clear(tid);
                else tab.clear(tid);
                wave_sync();
                if (j < nh) {
                    // collision h * 32 + j: this lane's own when hi == h, else lane l ^ 32's
                    const bool own = hi == h;
                    ImpulseParams<T> cp;
                    cp.nrm.x = own ? ip.nrm.x : pp.nrm.x;
                    cp.nrm.y = own ? ip.nrm.y : pp.nrm.y;
                    cp.nrm.z = own ? ip.nrm.z : pp.nrm.z;
                    cp.im0 = own ? ip.im0 : pp.im0;
                    cp.im1 = own ? ip.im1 : pp.im1;
                    cp.keep = own ? ip.keep : pp.keep;
                    int64_t x, y;
                    int type;
                    unpack_item(s_item[base + h * kHalfFlush + j], x, y, type);
#pragma unroll
                    for (int r = 0; r < 2; ++r) {
                        const int k = 2 * r + hi;
                        int side, idx;
                        type_corner(type, k, side, idx);
                        const int64_t c = 3 * (side ? y : x) + idx;
                        V3<T> v, dvk;
                        ld_vrec(opaque_ptr(tv), c, v, dvk);
                        V3<T> imp, imd;
                        corner_impulse(cp, type, k, dvk, imp, imd);
                        if (DET) {
                            const double v[6] = {(double)imp.x, (double)imp.y, (double)imp.z, (double)imd.x, (double)imd.y, (double)imd.z};
                            dtab.add(c, v, det, cnt);
                        } else {
                            tab.add(c, imp, imd, dv, dx, cnt);
                        }
                    }
                }
                wave_sync();
                if (DET) dtab.drain(tid, det, cnt);
                else tab.drain(tid, dv, dx, cnt);
                wave_sync();
            }
        }
    };

    // Collect mode: the staged collisions [0, f) to output rows (rows past out_cap are
    // counted, not written).
    auto write_rows = [&](int f) {
        unsigned long long ob = 0;
        if (tid == 0 && f) ob = atomicAdd(&ctr->coll, (unsigned long long)f);
        ob = readfirstlane_u64(ob);
        int64_t *coll = coll_all + (int64_t)b * out_cap * 3;
        for (int j = tid; j < f; j += NT) {
            const unsigned long long slot = ob + (unsigned long long)j;
            if (slot < (unsigned long long)out_cap) {
                int64_t x, y;
                int type;
                unpack_item(s_item[j], x, y, type);
                coll[3 * slot] = x; coll[3 * slot + 1] = y; coll[3 * slot + 2] = type;
                roots_all[(int64_t)b * out_cap + slot] = s_root[j];
            }
        }
        wave_sync();
    };

    // Filter of one root per lane where `on` (filterCollisions); accepted collisions join the
    // stage in lane order.  Every lane calls it.
    auto filter_round = [&](bool on, int64_t rit, T root) {
        bool acc = false;
        if (on) {
            int64_t x, y;
            int type;
            unpack_item(rit, x, y, type);
            acc = filter_collision(gather_mtet_rec(opaque_ptr(tv), x, y, type), type, root);
        }
        const unsigned long long m = __ballot(acc);
        if (acc) {
            const int p = fill + (int)lanes_below(m);
            s_item[p] = rit;
            s_root[p] = root;
        }
        fill += (int)__popcll(m);
    };

    // The tiles of CCC_FOR_TILES, walked by hand so that each lane's item of the next tile is
    // loaded one tile ahead (after this tile's gathers are issued) and its candidate pair
    // after phase B; kNoItem: no item.  An item is (pair index within the chunk) << 4 | type;
    // the packed (x, y, type) form is made when the tile starts.  The moving items come in
    // runs of one kind (edge-edge, then vertex-face, per k_subprim tile), so nearly every
    // tile's roots and collisions take one branch of the filter and of the impulse in the
    // whole wave.
    constexpr uint32_t kNoItem = 0xffffffffu;   // type 15 does not exist
    // tile counts in 32 bits (at most 15 * 2^28 / 64 tiles): the kernel's uniform values fill
    // the scalar registers
    const int ntiles = (int)((M + NT - 1) / NT);
    const int tper = (ntiles + kXcds - 1) / kXcds, tstep = (int)gridDim.x / kXcds;
    const int tbase = (int)(blockIdx.x % kXcds) * tper;
    auto item_of = [&](int k) -> uint32_t {
        const int64_t i = (int64_t)(tbase + k) * NT + tid;
        return (k < tper && tbase + k < ntiles && i < M) ? items[i] : kNoItem;
    };
    // the next tile's item waits as its pair (registers) and its type (LDS, 15: no item), so
    // that it holds no more registers than the packed item through phases C and D
    __shared__ unsigned char s_ntype[NT];
    __shared__ int64_t s_myitem[NT];
    uint32_t next_item = item_of(blockIdx.x / kXcds);
    int2 next_pair = make_int2(0, 0);
    if (next_item != kNoItem) next_pair = pairs[next_item >> 4];
    s_ntype[tid] = (unsigned char)(next_item & 15u);
    for (int tk = (int)blockIdx.x / kXcds; tk < tper && tbase + tk < ntiles; tk += tstep) {
        // ---- A: coefficients and branch of the lane's item
        const int ntype = s_ntype[tid];
        const int64_t it = ntype != 15 ? pack_item(next_pair.x, next_pair.y, ntype) : (int64_t)-1;
        int code = CUBIC_NONE;
        CubicState st;
        if (it >= 0) {
            int64_t x, y;
            int type;
            unpack_item(it, x, y, type);
            T cf[4];
            {
                const MTet<T> m = gather_mtet_rec(tv, x, y, type);
                cubic_coeffs(m, cf);
            }
            code = cubic_prepare(cf[0], cf[1], cf[2], cf[3], st);
        }
        s_myitem[tid] = it;     // read back for the root list, not held through the solver
        next_item = item_of(tk + tstep);
        // ---- B: the lane's roots in [0, 1], in the order found, to the root list
        T r0, r1, r2;
        const int nk = cubic_roots<T>(code, st, r0, r1, r2);
        {
            const unsigned long long b0 = __ballot(nk & 1), b1 = __ballot(nk & 2);
            const int off = (int)(lanes_below(b0) + 2u * lanes_below(b1));
            // the item as parked in LDS before the solver (the empty asm keeps the compiler
            // from holding it in registers instead: two more VGPRs spilled the solver)
            asm volatile("" ::: "memory");
            const int64_t itb = s_myitem[tid];
            if (nk > 0) { s_ritem[off] = itb; s_rroot[off] = r0; }
            if (nk > 1) { s_ritem[off + 1] = itb; s_rroot[off + 1] = r1; }
            if (nk > 2) { s_ritem[off + 2] = itb; s_rroot[off + 2] = r2; }
            const int nr = (int)(__popcll(b0) + 2 * __popcll(b1));
            nroots += (unsigned long long)nr;
            wave_sync();
            // the next tile's candidate pair (its item arrived during phase B)
            if (next_item != kNoItem) next_pair = pairs[next_item >> 4];
            s_ntype[tid] = (unsigned char)(next_item & 15u);
            // ---- C: filter one root per lane in whole rounds over the carried roots
            // followed by this tile's; the rest (fewer than NT) is carried to the next tile
            const int total = ncarry + nr, full = total / NT * NT;
            for (int base = 0; base < full; base += NT) {
                const int j = base + tid;
                int64_t rit;
                T root;
                if (j < ncarry) {
                    rit = s_citem[j];
                    root = s_croot[j];
                } else {
                    rit = s_ritem[j - ncarry];
                    root = s_rroot[j - ncarry];
                }
                filter_round(true, rit, root);
            }
            wave_sync();      // the carried roots have been read
            // with a whole round the carry was consumed and the rest is this tile's tail;
            // without one, this tile's roots join the carry
            const int from = full > 0 ? full - ncarry : 0, to = full > 0 ? 0 : ncarry;
            for (int k = tid; k < total - full - to; k += NT) {
                s_citem[to + k] = s_ritem[from + k];
                s_croot[to + k] = s_rroot[from + k];
            }
            ncarry = total - full;
            wave_sync();      // the root list has been read (its memory holds the table)
        }
        // ---- D: scatter the staged impulses in whole chunks; the rest moves to the front
        if (MODE == MODE_IMPULSE) {
            if (fill >= kFlush) {
                const int done = fill - fill % kFlush;
                flush(done);
                ncoll += (unsigned long long)done;
                // [done, fill) and [0, fill - done) do not overlap: fill - done < kFlush <= done
                if (tid < fill - done) {
                    s_item[tid] = s_item[done + tid];
                    s_root[tid] = s_root[done + tid];
                }
                wave_sync();
                fill -= done;
            }
        } else if (fill >= kCollectFlush) {
            // ---- D (collect): write the staged rows once kCollectFlush have gathered
            write_rows(fill);
            fill = 0;
        }
    }
    // the carried roots' last, partial round
    filter_round(tid < ncarry, s_citem[tid], s_croot[tid]);
    wave_sync();
    if (MODE == MODE_IMPULSE) {
        flush(fill);
        ncoll += (unsigned long long)fill;
    } else {
        write_rows(fill);
    }
    if (tid == 0) {
        if (nroots) atomicAdd(&ctr->roots, nroots);
        if (MODE == MODE_IMPULSE && ncoll) atomicAdd(&ctr->coll, ncoll);
    }
}

// The narrow phase's device arithmetic on given inputs, for tests: what 0-4 evaluates the
// portable cbrt / acos / sin / cos / cbrtf of ccc_libm.h at x[i]; what 5 runs cubic_prepare +
// cubic_roots<double> on the coefficients c3 c2 c1 c0 = x[4i..4i+3] and writes the number of
// roots kept in [0, 1] and those roots to y[4i..4i+3].
__global__ void k_device_math(int what, const double *__restrict__ x, double *__restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (what == 5) {
        CubicState st;
        const int code = cubic_prepare(x[4 * i], x[4 * i + 1], x[4 * i + 2], x[4 * i + 3], st);
        double r0, r1, r2;
        const int k = cubic_roots<double>(code, st, r0, r1, r2);
        y[4 * i] = (double)k;
        y[4 * i + 1] = k > 0 ? r0 : 0.0;
        y[4 * i + 2] = k > 1 ? r1 : 0.0;
        y[4 * i + 3] = k > 2 ? r2 : 0.0;
        return;
    }
    const double v = x[i];
    y[i] = what == 0 ? pm_cbrt(v) : what == 1 ? pm_acos(v) : what == 2 ? pm_sin(v) : what == 3 ? pm_cos(v)
         : what == 4 ? (double)pm_cbrtf((float)v) : what == 6 ? div_exact<3>(v) : what == 7 ? div_exact<27>(v)
         : what == 8 ? div_exact_finite<3>(v) : what == 9 ? sqrt_no_rescale(v)
         : what == 10 ? (unit_ratio(v, x[n + i]) ? 1.0 : 0.0)
                      : (unit_ratio((float)v, (float)x[n + i]) ? 1.0 : 0.0);
}

// compute_impulses_from_collisions (collisions_impulses.cu:1736-1880) over given rows;
// rows with a negative or out-of-range index are skipped.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_from_collisions(const int64_t *__restrict__ coll_all, const T *__restrict__ roots_all,
                                                              int64_t m, const T *__restrict__ t0_all, const T *__restrict__ t1_all,
                                                              const T *__restrict__ masses_all, int64_t n, T *__restrict__ dv_all,
                                                              T *__restrict__ dx_all, int *__restrict__ cnt_all) {
    const int b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t *row = coll_all + ((int64_t)b * m + i) * 3;
    const int64_t x = row[0], y = row[1], type = row[2];
    if (x < 0 || y < 0 || type < 0 || type > 14 || x >= n || y >= n) return;
    const T root = roots_all[(int64_t)b * m + i];
    const T *t0 = t0_all + (int64_t)b * n * 9;
    const T *t1 = t1_all + (int64_t)b * n * 9;
    const MTet<T> mt = gather_mtet(t0, t1, x, y, (int)type);
    scatter_impulse(mt, (int)type, root, x, y, masses_all + (int64_t)b * n * 3, dv_all + (int64_t)b * n * 9,
                    dx_all + (int64_t)b * n * 9, cnt_all + (int64_t)b * n * 3);
}

// Static proximity per item (proximity_static.cu:435-508): distance by type, kept when
// 0 <= d <= threshold; rows (x, y, type) truncated at the output capacity.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_proximity(const int64_t *__restrict__ items_all, int64_t capM,
                                                        const T *__restrict__ t0_all, int64_t n, T thr,
                                                        int64_t *__restrict__ out_all, int64_t out_cap,
                                                        DevCounters *__restrict__ ctr_all) {
    const int b = blockIdx.y;
    DevCounters *ctr = ctr_all + b;
    const unsigned long long mc = ctr->items;
    const int64_t M = (int64_t)(mc < (unsigned long long)capM ? mc : (unsigned long long)capM);
    const int64_t *items = items_all + (int64_t)b * capM;
    const T *t0 = t0_all + (int64_t)b * n * 9;
    int64_t *out = out_all + (int64_t)b * out_cap * 3;
    const int64_t ntiles = (M + kThreads - 1) / kThreads;
    CCC_FOR_TILES(tile, ntiles) {
        const int64_t i = tile * kThreads + threadIdx.x;
        int64_t x = 0, y = 0;
        int type = 0;
        unsigned keep = 0;
        if (i < M) {
            unpack_item(items[i], x, y, type);
            V3<T> p[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int side, idx;
                type_corner(type, k, side, idx);
                p[k] = ld3(t0 + 9 * (side ? y : x) + 3 * idx);
            }
            T d;
            if (type < 9) d = dist_ee(p[0], p[1], p[2], p[3]);
            else if (type < 12) d = dist_vf(p[0], p[1], p[2], p[3]);
            else d = dist_vf(p[1], p[2], p[3], p[0]);
            keep = (d <= thr && d >= 0) ? 1u : 0u;
        }
        const unsigned long long slot = block_reserve(&ctr->coll, keep);
        if (keep && slot < (unsigned long long)out_cap) {
            out[3 * slot] = x; out[3 * slot + 1] = y; out[3 * slot + 2] = type;
        }
    }
}

// Static triangle-triangle test per candidate pair (collisions_static.cu:123-159):
// 17-axis SAT with threshold and no shared vertex; rows (x, y).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_sat(const int2 *__restrict__ pairs_all, int64_t capK, const T *__restrict__ t0_all,
                                                  int64_t n, T thr, int64_t *__restrict__ out_all, int64_t out_cap,
                                                  DevCounters *__restrict__ ctr_all) {
    const int b = blockIdx.y;
    DevCounters *ctr = ctr_all + b;
    const unsigned long long kc = ctr->cand;
    const int64_t K = (int64_t)(kc < (unsigned long long)capK ? kc : (unsigned long long)capK);
    const int2 *pairs = pairs_all + (int64_t)b * capK;
    const T *t0 = t0_all + (int64_t)b * n * 9;
    int64_t *out = out_all + (int64_t)b * out_cap * 2;
    const int64_t ntiles = (K + kThreads - 1) / kThreads;
    CCC_FOR_TILES(tile, ntiles) {
        const int64_t i = tile * kThreads + threadIdx.x;
        unsigned keep = 0;
        int x = 0, y = 0;
        if (i < K) {
            const int2 pr = pairs[i];
            x = pr.x;
            y = pr.y;
            V3<T> a[3], c[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) { a[k] = ld3(t0 + 9 * (int64_t)x + 3 * k); c[k] = ld3(t0 + 9 * (int64_t)y + 3 * k); }
            keep = (sat_overlap(a, c, thr) && !(count_shared(a, c) > 0)) ? 1u : 0u;
        }
        const unsigned long long slot = block_reserve(&ctr->coll, keep);
        if (keep && slot < (unsigned long long)out_cap) { out[2 * slot] = x; out[2 * slot + 1] = y; }
    }
}

// ---- host orchestration ---------------------------------------------------------------------

// Triangle-edge queries over the candidate pairs of a threshold-0 traversal, one pair per
// thread: SAT with at most `allow` shared vertices (collisions_static.cu:123-159), then the
// six edges cid 0..2 = (v0v1, v1v2, v2v0) of x against y and 3..5 of y against x.
//   TE_COLLISIONS  (allow 0) rows (x, y, cid, loop) + edge coefficient for every edge that
//                  pierces the other triangle (collisions_static.cu:445-500, :612-779);
//   TE_CANDIDATES  (allow 1) rows (x, y, cid, loop) for every edge that does not end on a
//                  shared corner (collisions_static.cu:520-604, :782-943);
//   TE_CANDIDATES2 (allow 1) one row (x, y) per pair with bit masks (mx, my) of the same
//                  edges (icontour.cu:11-108, :110-282).
// A thread reserves all of its rows with one block-wide reservation.
enum { TE_COLLISIONS = 0, TE_CANDIDATES = 1, TE_CANDIDATES2 = 2 };

template <typename T, int MODE>
__global__ __launch_bounds__(kThreads) void k_tri_edge(const int2 *__restrict__ pairs_all, int64_t capK,
                                                       const T *__restrict__ t0_all, int64_t n,
                                                       int64_t *__restrict__ rows_all, T *__restrict__ coef_all,
                                                       int32_t *__restrict__ masks_all, int64_t out_cap,
                                                       DevCounters *__restrict__ ctr_all) {
    const int b = blockIdx.y;
    DevCounters *ctr = ctr_all + b;
    const unsigned long long kc = ctr->cand;
    const int64_t K = (int64_t)(kc < (unsigned long long)capK ? kc : (unsigned long long)capK);
    const int2 *pairs = pairs_all + (int64_t)b * capK;
    const T *t0 = t0_all + (int64_t)b * n * 9;
    constexpr int W = MODE == TE_CANDIDATES2 ? 2 : 4;
    int64_t *rows = rows_all + (int64_t)b * out_cap * W;
    const int64_t ntiles = (K + kThreads - 1) / kThreads;
    CCC_FOR_TILES(tile, ntiles) {
        const int64_t i = tile * kThreads + threadIdx.x;
        int x = 0, y = 0;
        unsigned emit = 0;          // bit cid: row for edge cid
        unsigned loops = 0;         // bit cid: loop node (loop id 0)
        int mx = 0, my = 0;
        T coefs[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) coefs[c] = (T)-1;
        bool valid = false;
        if (i < K) {
            const int2 pr = pairs[i];
            x = pr.x;
            y = pr.y;
            V3<T> a[3], c3[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) { a[k] = ld3(t0 + 9 * (int64_t)x + 3 * k); c3[k] = ld3(t0 + 9 * (int64_t)y + 3 * k); }
            const int allow = MODE == TE_COLLISIONS ? 0 : 1;
            valid = sat_overlap(a, c3, (T)0) && !(count_shared(a, c3) > allow);
            if (valid) {
#pragma unroll
                for (int cid = 0; cid < 6; ++cid) {
                    const V3<T> *t1 = cid < 3 ? a : c3;
                    const V3<T> *t2 = cid < 3 ? c3 : a;
                    const int e = cid % 3;
                    const V3<T> e0 = t1[e], e1 = t1[(e + 1) % 3];
                    if (MODE == TE_COLLISIONS) {
                        int64_t loop;
                        T cf;
                        if (te_intersect(t2, e0, e1, cid, loop, cf)) {
                            emit |= 1u << cid;
                            if (loop == 0) loops |= 1u << cid;
                            coefs[cid] = cf;
                        }
                    } else {
                        const bool s0 = has_vertex(t2, e0), s1 = has_vertex(t2, e1);
                        if (s0) {
                            if (cid < 3) { emit |= 1u << cid; loops |= 1u << cid; mx |= 1 << cid; my |= 1 << cid; }
                        } else if (!s1) {
                            // the reference compares t2 with the box of its own edge (collisions_static.cu:594-596)
                            const Box<T> b0 = tri_box(t2[0], t2[1], t2[2]);
                            const Box<T> b1 = unite(point_box(t2[e]), point_box(t2[(e + 1) % 3]));
                            if (overlap(b0, b1, (T)0)) { emit |= 1u << cid; mx |= 1 << cid; }
                        }
                    }
                }
            }
        }
        const unsigned cnt = MODE == TE_CANDIDATES2 ? (valid ? 1u : 0u) : (unsigned)__popc(emit);
        unsigned long long slot = block_reserve(&ctr->coll, cnt);
        if (MODE == TE_CANDIDATES2) {
            if (valid && slot < (unsigned long long)out_cap) {
                rows[2 * slot] = x;
                rows[2 * slot + 1] = y;
                int32_t *masks = masks_all + (int64_t)b * out_cap * 2;
                masks[2 * slot] = mx;
                masks[2 * slot + 1] = my;
            }
        } else {
#pragma unroll
            for (int cid = 0; cid < 6; ++cid) {
                if (!(emit & (1u << cid))) continue;
                if (slot < (unsigned long long)out_cap) {
                    rows[4 * slot] = x;
                    rows[4 * slot + 1] = y;
                    rows[4 * slot + 2] = cid;
                    rows[4 * slot + 3] = (loops & (1u << cid)) ? 0 : -1;
                    if (MODE == TE_COLLISIONS) coef_all[(int64_t)b * out_cap + slot] = coefs[cid];
                }
                ++slot;
            }
        }
    }
}

enum class Narrow { SAT, PROXIMITY, CCD_COLLECT, CCD_IMPULSE, MINCOL, TRI_EDGE };

struct Query {
    ccc_context *ctx;
    const void *bboxes;
    const int64_t *tree;
    const void *t0, *t1, *masses;
    const uint8_t *to_check;
    int64_t B, n;
    int32_t dtype;
    double thr;             // threshold already converted from the binding's float
    int64_t cap_candidates; // per batch element (max_candidates * n)
    int64_t out_cap;        // rows per batch element of the output
    Narrow narrow;
    int64_t *out_rows;      // SAT [B,cap,2] / PROXIMITY, CCD, MINCOL [B,cap,3]
    void *out_roots;        // CCD_COLLECT [B,cap]
    void *dv, *dx;          // CCD_IMPULSE
    int32_t *cnt;
    int te_mode;            // TRI_EDGE: TE_COLLISIONS / TE_CANDIDATES / TE_CANDIDATES2
    const int32_t *mesh_end;  // packed meshes (B = 1): per leaf position, end of its mesh
    int32_t *masks;         // TRI_EDGE candidates2 [B,cap,2]
    ccc_counts *counts;
    uint32_t flags;
    hipStream_t s;
    void *arena;            // non-NULL: every workspace buffer is carved from this block
    size_t arena_bytes;
    // Resident passes (the resident simulation): counters, the candidate list and the
    // deterministic impulse limbs live in the caller's memory, and with `nosync` nothing
    // waits on the host -- the list keeps at most cap_candidates pairs (more are counted),
    // the narrow phase walks [0, cap_candidates) in chunks whose kernels stop at the
    // device's counts, and the counters stay on the device.  Such a pass is capturable.
    bool nosync;
    DevCounters *ctr_ext;   // B counters (zeroed by the pass) or NULL
    int2 *pairs_ext;        // B * cap_candidates pairs or NULL
    int64_t chunk_ext;      // pairs per narrow-phase chunk (0: the context's)
    int ccd_blocks_per_cu;  // grid cap of k_ccd in blocks per CU (0: 128)
    bool stack_deep;        // start with the deep traversal stack
    long long *det_acc;     // non-NULL: impulse sums in fixed-point limbs (n * 3 * 6 * 4 per element)
};

// End of a pair chunk: its sub-primitive items are done; the next chunk refills the list.
__global__ void k_fold_items(DevCounters *__restrict__ ctr, int64_t B) {
    for (int64_t b = threadIdx.x; b < B; b += blockDim.x) {
        ctr[b].items_done += ctr[b].items;
        ctr[b].items = 0;
    }
}

// Pairs per narrow-phase chunk over all batch elements of a call (the context may set
// another size) for 4-byte items, half as many for 8-byte ones: the item list holds at most
// 15 items per pair of a chunk, about 4 GiB either way.  On the 130k cloth one chunk of 2^26
// instead of two of 2^25 took k_ccd from 16.32 to 16.17 ms per pass.
constexpr int64_t kPairChunk = (int64_t)1 << 26;
// Largest chunk: at most 15 * 2^28 items per batch element; the continuous narrow phase's
// 32-bit items hold the pair index within the chunk in 28 bits.
constexpr int64_t kMaxPairChunk = (int64_t)1 << 28;
// Bytes per sub-primitive item: the continuous narrow phase keeps (pair index within the
// chunk) << 4 | type in 32 bits (chunks hold at most 2^28 pairs); the static one (x, y, type)
// in 64.
static size_t item_bytes(Narrow nw) { return nw == Narrow::PROXIMITY ? sizeof(int64_t) : sizeof(uint32_t); }
static int64_t pair_chunk_of(const ccc_context *ctx, int64_t chunk_ext, int64_t pairs, int64_t B, size_t item_size) {
    const int64_t dflt = kPairChunk * (int64_t)sizeof(uint32_t) / (int64_t)item_size;
    const int64_t budget = chunk_ext > 0 ? chunk_ext * B : ctx->pair_chunk > 0 ? ctx->pair_chunk : dflt;
    return std::max<int64_t>(std::min<int64_t>(std::min<int64_t>(pairs, budget / B), kMaxPairChunk), 1);
}
constexpr int kSubprimPairs = 2;   // pairs per thread of k_subprim

// Grid of the tile-looping kernels: one block per tile up to 128 blocks per CU, each block
// then walking several tiles.  Measured on the 130k cloth (k_ccd per step): 4 blocks per CU
// (all resident at once) 232 ms, 16 215 ms, 32 210, 128 207, 512 214 -- many short blocks
// let the dispatcher balance the uneven tiles; past that the per-block start-up dominates.
static unsigned persistent_grid(const ccc_context *ctx, int64_t items, int64_t B, int block = kThreads, int per_cu = 128) {
    int64_t want = (items + block - 1) / block;
    if (want < 1) want = 1;                        // no items: a grid of idle blocks, never 0
    int64_t limit = (int64_t)ctx->num_cus * per_cu / (B > 0 ? B : 1);
    if (limit < kXcds) limit = kXcds;
    if (want > limit) want = limit;
    want = (want + kXcds - 1) / kXcds * kXcds;    // CCC_FOR_TILES needs a multiple of kXcds
    return (unsigned)want;
}

template <typename T>
static ccc_status run_query(const Query &q) {
    hipStream_t s = q.s;
    const int64_t B = q.B, n = q.n, rows = 2 * n - 1;
    const int64_t nint = n > 1 ? n - 1 : 1;
    Workspace ws(q.ctx, s, q.arena, q.arena_bytes);
    // no event objects in the allocation-free entry point (arena without resident mode)
    KernelTimer tm(q.arena && !q.nosync ? nullptr : q.ctx, s);
    DevCounters *ctr = q.ctr_ext ? q.ctr_ext : ws.get_as<DevCounters>(B);
    int *claims = ws.get_as<int>(B * n);
    int *climb = ws.get_as<int>(B * nint);
    int *rm = ws.get_as<int>(B * nint);
    int *lo = ws.get_as<int>(B * nint);
    TNode *nodes = ws.get_as<TNode>(B * nint);
    QNode *qnodes = ws.get_as<QNode>(B * nint);
    LeafRec *recs = ws.get_as<LeafRec>(B * n);
    Frontier *frontier = ws.get_as<Frontier>(B);
    if (ws.failed()) return CCC_ERR_NO_MEMORY;
    ccc_status st;
    if ((st = fill_u64(ctr, B * (int64_t)(sizeof(DevCounters) / 8), 0ull, s)) != CCC_OK) return st;
    if ((st = fill_u32(claims, B * n, 0u, s)) != CCC_OK) return st;
    if ((st = fill_u32(climb, B * nint, 0u, s)) != CCC_OK) return st;

    // outputs start at -1 (index rows, roots) or 0 (impulses)
    switch (q.narrow) {
        case Narrow::SAT: st = fill_u64(q.out_rows, B * q.out_cap * 2, ~0ull, s); break;
        case Narrow::TRI_EDGE:
            st = fill_u64(q.out_rows, B * q.out_cap * (q.te_mode == TE_CANDIDATES2 ? 2 : 4), ~0ull, s);
            if (st == CCC_OK && q.te_mode == TE_COLLISIONS) {
                if (sizeof(T) == 8) st = fill_u64(q.out_roots, B * q.out_cap, 0xBFF0000000000000ull, s);
                else st = fill_u32(q.out_roots, B * q.out_cap, 0xBF800000u, s);
            }
            if (st == CCC_OK && q.te_mode == TE_CANDIDATES2) st = fill_u32(q.masks, B * q.out_cap * 2, 0u, s);
            break;
        case Narrow::PROXIMITY:
        case Narrow::MINCOL: st = fill_u64(q.out_rows, B * q.out_cap * 3, ~0ull, s); break;
        case Narrow::CCD_COLLECT:
            st = fill_u64(q.out_rows, B * q.out_cap * 3, ~0ull, s);
            if (st == CCC_OK) {
                if (sizeof(T) == 8) st = fill_u64(q.out_roots, B * q.out_cap, 0xBFF0000000000000ull, s);
                else st = fill_u32(q.out_roots, B * q.out_cap, 0xBF800000u, s);
            }
            break;
        case Narrow::CCD_IMPULSE:
            if (sizeof(T) == 8) { st = fill_u64(q.dv, B * n * 9, 0ull, s); if (st == CCC_OK) st = fill_u64(q.dx, B * n * 9, 0ull, s); }
            else { st = fill_u32(q.dv, B * n * 9, 0u, s); if (st == CCC_OK) st = fill_u32(q.dx, B * n * 9, 0u, s); }
            if (st == CCC_OK) st = fill_u32(q.cnt, B * n * 3, 0u, s);
            if (st == CCC_OK && q.det_acc) st = fill_u64(q.det_acc, B * n * 3 * kDetWords, 0ull, s);
            break;
    }
    if (st != CCC_OK) return st;

    const T *bb = (const T *)q.bboxes;
    tm.begin("tree_prep");
    hipLaunchKernelGGL(k_check_tree, dim3(grid_for(rows, 256), B), dim3(256), 0, s, q.tree, n, claims, ctr);
    if (n > 1) {
        hipLaunchKernelGGL(k_ranges, dim3(grid_for(n - 1, 256), B), dim3(256), 0, s, q.tree, n, lo, rm, ctr);
        hipLaunchKernelGGL(k_order_check, dim3(grid_for(n - 1, 256), B), dim3(256), 0, s, q.tree, n, lo, rm, ctr);
    }
    hipLaunchKernelGGL(k_rightmost, dim3(grid_for(n, 256), B), dim3(256), 0, s, q.tree, n, climb, rm, ctr);
    if (n > 1) hipLaunchKernelGGL((k_nodes<T>), dim3(grid_for(n - 1, 256), B), dim3(256), 0, s, bb, q.tree, rm, n, q.thr, nodes, ctr);
    if (n > 1) hipLaunchKernelGGL(k_qnodes, dim3(grid_for(n - 1, 256), B), dim3(256), 0, s, nodes, n, qnodes, ctr);
    hipLaunchKernelGGL((k_leaf_recs<T>), dim3(grid_for(n, 256), B), dim3(256), 0, s, bb, q.tree, n, recs, ctr);
    hipLaunchKernelGGL((k_frontier<T>), dim3(B), dim3(kFrontier), 0, s, bb, q.tree, rm, lo, n, q.thr, frontier, ctr);
    tm.end();
    CCC_LAUNCH_CHECK("tree preparation");

    // traversal, re-run once with the exact size if the capacity was exceeded
    std::vector<DevCounters> h(B);
    int64_t capK = q.cap_candidates > 0 ? q.cap_candidates : 1;
    bool warn_cand = false;
    int2 *pairs = nullptr;
    // one block per (start point, query block); query blocks rounded to a multiple of kXcds
    const unsigned gT = kFrontier * (unsigned)((grid_for(n, kTravThreads) + kXcds - 1) / kXcds * kXcds);
    const int nqb = (int)(gT / kFrontier);
    float *qbox = ws.get_as<float>(B * (int64_t)nqb * 6);
    if (ws.failed()) return CCC_ERR_NO_MEMORY;
    hipLaunchKernelGGL((k_qblock_boxes<T>), dim3(nqb, B), dim3(kTravThreads), 0, s, bb, n, nqb, qbox, ctr);
    CCC_LAUNCH_CHECK("k_qblock_boxes");
    bool deep = q.stack_deep;   // the short stack overflowed once: run with the deep one
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (q.pairs_ext) {
            pairs = q.pairs_ext;
        } else {
            if (pairs) ws.release(pairs);     // a re-run: the first attempt's pairs are not kept
            pairs = ws.get_as<int2>(B * capK);
            if (ws.failed()) return CCC_ERR_NO_MEMORY;
        }
        tm.begin("traverse");
#define CCC_TRAVERSE(P, D)                                                                                   \
    hipLaunchKernelGGL((k_traverse<T, P, D>), dim3(gT, B), dim3(kTravThreads), 0, s, bb, q.tree, qnodes, recs, n, \
                       (T)q.thr, q.to_check, q.mesh_end, pairs, capK, ctr, frontier, qbox)
        if (q.to_check) {
            if (deep) CCC_TRAVERSE(true, kStackDeep);
            else CCC_TRAVERSE(true, kStackShort);
        } else {
            if (deep) CCC_TRAVERSE(false, kStackDeep);
            else CCC_TRAVERSE(false, kStackShort);
        }
#undef CCC_TRAVERSE
        tm.end();
        CCC_LAUNCH_CHECK("k_traverse");
        if (q.nosync) break;          // the caller checks the counters after its step
        CCC_HIP_TRY(hipMemcpyAsync(h.data(), ctr, sizeof(DevCounters) * B, hipMemcpyDeviceToHost, s));
        CCC_HIP_TRY(hipStreamSynchronize(s));
        unsigned flags = 0;
        unsigned long long kmax = 0;
        for (int64_t b = 0; b < B; ++b) {
            flags |= h[b].flags;
            kmax = std::max(kmax, h[b].cand);
        }
        if (flags & DEV_BAD_TREE) { set_error("tree tensor is not a valid BVH over n leaves"); return CCC_ERR_BAD_TREE; }
        if (flags & DEV_STACK) {
            if (deep) { set_error("BVH deeper than the traversal stack"); return CCC_ERR_STACK; }
            // again with the deep stack (not a capacity attempt)
            deep = true;
            --attempt;
            for (int64_t b = 0; b < B; ++b) {
                h[b].cand = 0;
                h[b].flags &= ~DEV_STACK;
            }
            CCC_HIP_TRY(hipMemcpyAsync(ctr, h.data(), sizeof(DevCounters) * B, hipMemcpyHostToDevice, s));
            continue;
        }
        if (kmax <= (unsigned long long)capK) break;
        warn_cand = true;
        if (q.flags & CCC_FLAG_FIXED_CAPACITY) break;
        if (attempt == 1 || q.pairs_ext) break;   // a caller-owned list does not grow here
        capK = (int64_t)kmax;
        for (int64_t b = 0; b < B; ++b) h[b].cand = 0;
        CCC_HIP_TRY(hipMemcpyAsync(ctr, h.data(), sizeof(DevCounters) * B, hipMemcpyHostToDevice, s));
    }
    int64_t kmax = 0;
    if (q.nosync) kmax = capK;        // sizes must not depend on the counts
    else
        for (int64_t b = 0; b < B; ++b) kmax = std::max<int64_t>(kmax, std::min<int64_t>((int64_t)h[b].cand, capK));

    const T thr = (T)q.thr;
    const unsigned gK = persistent_grid(q.ctx, kmax, B);
    if (q.narrow == Narrow::SAT) {
        CCC_TIMED(tm, "sat", hipLaunchKernelGGL((k_sat<T>), dim3(gK, B), dim3(kThreads), 0, s, pairs, capK, (const T *)q.t0, n, thr,
                                                q.out_rows, q.out_cap, ctr));
        CCC_LAUNCH_CHECK("k_sat");
    } else if (q.narrow == Narrow::TRI_EDGE) {
        tm.begin("tri_edge");
        if (q.te_mode == TE_COLLISIONS)
            hipLaunchKernelGGL((k_tri_edge<T, TE_COLLISIONS>), dim3(gK, B), dim3(kThreads), 0, s, pairs, capK, (const T *)q.t0, n,
                               q.out_rows, (T *)q.out_roots, (int32_t *)nullptr, q.out_cap, ctr);
        else if (q.te_mode == TE_CANDIDATES)
            hipLaunchKernelGGL((k_tri_edge<T, TE_CANDIDATES>), dim3(gK, B), dim3(kThreads), 0, s, pairs, capK, (const T *)q.t0, n,
                               q.out_rows, (T *)nullptr, (int32_t *)nullptr, q.out_cap, ctr);
        else
            hipLaunchKernelGGL((k_tri_edge<T, TE_CANDIDATES2>), dim3(gK, B), dim3(kThreads), 0, s, pairs, capK, (const T *)q.t0, n,
                               q.out_rows, (T *)nullptr, q.masks, q.out_cap, ctr);
        tm.end();
        CCC_LAUNCH_CHECK("k_tri_edge");
    } else if (q.narrow == Narrow::MINCOL) {
        T *tv = ws.get_as<T>(B * n * 18);
        if (ws.failed()) return CCC_ERR_NO_MEMORY;
        const unsigned gS = persistent_grid(q.ctx, (kmax + kSubprimPairs - 1) / kSubprimPairs, B);
        tm.begin("mincol");
        hipLaunchKernelGGL((k_vertex_records<T>), dim3(grid_for(3 * n, 256), B), dim3(256), 0, s, (const T *)q.t0,
                           (const T *)q.t1, 3 * n, tv);
        hipLaunchKernelGGL((k_subprim<T, SUB_MINCOL, kSubprimPairs>), dim3(gS, B), dim3(kThreads), 0, s, pairs, capK,
                           (int64_t)0, capK, (const T *)tv, n, thr, q.out_rows, q.out_cap, ctr);
        tm.end();
        CCC_LAUNCH_CHECK("k_subprim");
    } else {
        // sub-primitive items and the narrow phase run over chunks of the pair list; the
        // chunk bounds the pairs of all batch elements together (the item buffer holds B
        // chunks), so the workspace does not grow with the batch
        int64_t chunk = pair_chunk_of(q.ctx, q.chunk_ext, kmax, B, item_bytes(q.narrow));
        const bool ccd = q.narrow == Narrow::CCD_COLLECT || q.narrow == Narrow::CCD_IMPULSE;
        T *tv = ccd ? ws.get_as<T>(B * n * 18) : nullptr;
        if (ws.failed()) return CCC_ERR_NO_MEMORY;
        if (ws.arena()) {
            // the item list takes what the arena has left (results do not depend on the chunk)
            const int64_t fit = (int64_t)(ws.arena_left() / ((size_t)B * 15 * item_bytes(q.narrow)));
            if (fit < 1) {
                set_error("the scratch buffer has no room left for the sub-primitive items of one candidate pair");
                return CCC_ERR_NO_MEMORY;
            }
            chunk = std::min(chunk, fit);
        }
        const int64_t capM = 15 * chunk;
        void *items = ws.get_as<unsigned char>(B * capM * (int64_t)item_bytes(q.narrow));
        const unsigned gS = persistent_grid(q.ctx, (chunk + kSubprimPairs - 1) / kSubprimPairs, B);
        const unsigned gM = persistent_grid(q.ctx, capM, B);
        const unsigned gCcdC = persistent_grid(q.ctx, capM, B, ccd_threads<MODE_COLLECT>());
        // one-wave impulse blocks, up to 128 per CU: on the 130k cloth 256 per CU took 171.6 ms
        // per step against 173.7 and 64 took 178.3, but 256 cost the 122k garment scene (fewer
        // items per pass) 23% of its step rate
        const unsigned gCcdI = persistent_grid(q.ctx, capM, B, ccd_threads<MODE_IMPULSE>(),
                                               q.ccd_blocks_per_cu > 0 ? q.ccd_blocks_per_cu : 128);
        if (ws.failed()) return CCC_ERR_NO_MEMORY;
        if (ccd) {
            hipLaunchKernelGGL((k_vertex_records<T>), dim3(grid_for(3 * n, 256), B), dim3(256), 0, s, (const T *)q.t0,
                               (const T *)q.t1, 3 * n, tv);
            CCC_LAUNCH_CHECK("k_vertex_records");
        }
        auto narrow_phase = [&]() -> ccc_status {
            for (int64_t c0 = 0; c0 < kmax; c0 += chunk) {
                const int64_t c1 = c0 + chunk;
                tm.begin("subprim");
                if (q.narrow == Narrow::PROXIMITY)
                    hipLaunchKernelGGL((k_subprim<T, SUB_STATIC, kSubprimPairs>), dim3(gS, B), dim3(kThreads), 0, s, pairs, capK,
                                       c0, c1, (const T *)q.t0, n, thr, items, capM, ctr);
                else
                    hipLaunchKernelGGL((k_subprim<T, SUB_MOVING, kSubprimPairs>), dim3(gS, B), dim3(kThreads), 0, s, pairs, capK,
                                       c0, c1, (const T *)tv, n, thr, items, capM, ctr);
                tm.end();
                CCC_LAUNCH_CHECK("k_subprim");
                tm.begin(q.narrow == Narrow::PROXIMITY ? "proximity" : q.narrow == Narrow::CCD_COLLECT ? "ccd_collect" : "ccd_impulse");
                if (q.narrow == Narrow::PROXIMITY)
                    hipLaunchKernelGGL((k_proximity<T>), dim3(gM, B), dim3(kThreads), 0, s, (const int64_t *)items, capM, (const T *)q.t0, n, thr,
                                       q.out_rows, q.out_cap, ctr);
                else if (q.narrow == Narrow::CCD_COLLECT)
                    hipLaunchKernelGGL((k_ccd<T, MODE_COLLECT, false>), dim3(gCcdC, B), dim3(ccd_threads<MODE_COLLECT>()), 0, s, (const uint32_t *)items, capM, (const int2 *)pairs, capK, c0,
                                       (const T *)tv, (const T *)nullptr, n, q.out_rows, (T *)q.out_roots, q.out_cap,
                                       (T *)nullptr, (T *)nullptr, (int *)nullptr, (long long *)nullptr, ctr);
                else if (q.det_acc)
                    hipLaunchKernelGGL((k_ccd<T, MODE_IMPULSE, true>), dim3(gCcdI, B), dim3(ccd_threads<MODE_IMPULSE>()), 0, s, (const uint32_t *)items, capM, (const int2 *)pairs, capK, c0,
                                       (const T *)tv, (const T *)q.masses, n, (int64_t *)nullptr, (T *)nullptr, (int64_t)0,
                                       (T *)q.dv, (T *)q.dx, q.cnt, q.det_acc, ctr);
                else
                    hipLaunchKernelGGL((k_ccd<T, MODE_IMPULSE, false>), dim3(gCcdI, B), dim3(ccd_threads<MODE_IMPULSE>()), 0, s, (const uint32_t *)items, capM, (const int2 *)pairs, capK, c0,
                                       (const T *)tv, (const T *)q.masses, n, (int64_t *)nullptr, (T *)nullptr, (int64_t)0,
                                       (T *)q.dv, (T *)q.dx, q.cnt, (long long *)nullptr, ctr);
                tm.end();
                CCC_LAUNCH_CHECK("narrow phase");
                hipLaunchKernelGGL(k_fold_items, dim3(1), dim3(64), 0, s, ctr, B);
                CCC_LAUNCH_CHECK("k_fold_items");
            }
            return CCC_OK;
        };
        if ((st = narrow_phase()) != CCC_OK) return st;
        if (q.det_acc) {
            hipLaunchKernelGGL((k_det_finalize<T>), dim3(grid_for(B * n * 3, 256)), dim3(256), 0, s, q.det_acc, B * n * 3,
                               (T *)q.dv, (T *)q.dx);
            CCC_LAUNCH_CHECK("k_det_finalize");
        }
    }
    if (q.nosync) return CCC_OK;
    CCC_HIP_TRY(hipMemcpyAsync(h.data(), ctr, sizeof(DevCounters) * B, hipMemcpyDeviceToHost, s));
    CCC_HIP_TRY(hipStreamSynchronize(s));
    bool truncated = false;
    for (int64_t b = 0; b < B; ++b) {
        const int64_t K = std::min<int64_t>((int64_t)h[b].cand, capK);
        if (q.counts) {
            q.counts[b].candidates = (int64_t)h[b].cand;
            q.counts[b].subprims = (int64_t)(q.narrow == Narrow::MINCOL ? h[b].items : h[b].items_done);
            q.counts[b].roots = (int64_t)h[b].roots;
            q.counts[b].collisions = q.narrow == Narrow::MINCOL ? (int64_t)h[b].items : (int64_t)h[b].coll;
        }
        (void)K;
        const int64_t produced = q.narrow == Narrow::MINCOL ? (int64_t)h[b].items : (int64_t)h[b].coll;
        if (q.narrow != Narrow::CCD_IMPULSE && produced > q.out_cap) truncated = true;
    }
    ccc_status out = CCC_OK;
    if (warn_cand) out |= CCC_WARN_CANDIDATES;
    if (truncated) out |= CCC_WARN_TRUNCATED;
    return out;
}

// Workspace bytes of a resident CCD_IMPULSE pass (nosync, caller-owned counters and pair
// list): the same buffers, sizes and 256-byte rounding as run_query's requests.
static size_t resident_bytes(const Query &q) {
    const int64_t B = q.B, n = q.n, nint = n > 1 ? n - 1 : 1;
    const size_t esz = q.dtype == CCC_FLOAT64 ? 8 : 4;
    auto r = [](size_t b) { return ((b ? b : 256) + 255) & ~(size_t)255; };
    size_t tot = 0;
    if (!q.ctr_ext) tot += r(sizeof(DevCounters) * B);
    tot += r(4 * B * n) + 3 * r(4 * B * nint);
    tot += r(sizeof(TNode) * B * nint) + r(sizeof(QNode) * B * nint) + r(sizeof(LeafRec) * B * n) + r(sizeof(Frontier) * B);
    const unsigned gT = kFrontier * (unsigned)((grid_for(n, kTravThreads) + kXcds - 1) / kXcds * kXcds);
    tot += r(sizeof(float) * B * (int64_t)(gT / kFrontier) * 6);
    const int64_t capK = q.cap_candidates > 0 ? q.cap_candidates : 1;
    if (!q.pairs_ext) tot += r(sizeof(int2) * B * capK);
    const int64_t chunk = pair_chunk_of(q.ctx, q.chunk_ext, capK, B, item_bytes(q.narrow));
    tot += r(esz * B * n * 18) + r(item_bytes(q.narrow) * B * 15 * chunk);
    return tot;
}

static Query resident_query(ccc_context *ctx, const ResidentPass &p, hipStream_t s) {
    Query q{};
    q.ctx = ctx;
    q.bboxes = p.bboxes;
    q.tree = p.tree;
    q.t0 = p.t0;
    q.t1 = p.t1;
    q.masses = p.masses;
    q.B = 1;
    q.n = p.n;
    q.dtype = p.dtype;
    q.thr = 0.0;
    q.cap_candidates = p.cap_pairs;
    q.narrow = Narrow::CCD_IMPULSE;
    q.dv = p.dv;
    q.dx = p.dx;
    q.cnt = p.cnt;
    q.mesh_end = p.mesh_end;
    q.flags = CCC_FLAG_FIXED_CAPACITY;
    q.s = s;
    q.nosync = true;
    q.ctr_ext = p.ctr;
    q.pairs_ext = p.pairs;
    q.chunk_ext = p.chunk;
    q.ccd_blocks_per_cu = p.ccd_blocks_per_cu;
    q.stack_deep = p.stack_deep;
    q.det_acc = p.det_acc;
    return q;
}

size_t resident_pass_arena_bytes(ccc_context *ctx, const ResidentPass &p) { return resident_bytes(resident_query(ctx, p, nullptr)); }

ccc_status resident_pass(ccc_context *ctx, const ResidentPass &p, void *arena, size_t arena_bytes, hipStream_t s) {
    Query q = resident_query(ctx, p, s);
    q.arena = arena;
    q.arena_bytes = arena_bytes;
    if (p.dtype == CCC_FLOAT64) return run_query<double>(q);
    if (p.dtype == CCC_FLOAT32) return run_query<float>(q);
    return bad_arg("dtype must be CCC_FLOAT32 or CCC_FLOAT64");
}

static ccc_status run(const Query &q) {
    clear_error();
    if (!q.ctx) return bad_arg("context is NULL");
    if (q.B < 0 || q.n < 1) return bad_arg("need batch >= 0 and n >= 1 triangles");
    if (q.n >= ((int64_t)1 << 28)) return bad_arg("at most 2^28 - 1 triangles per mesh");
    if (q.B == 0) return CCC_OK;
    if (!q.bboxes || !q.tree || !q.t0) return bad_arg("NULL tensor pointer");
    if ((q.narrow == Narrow::CCD_COLLECT || q.narrow == Narrow::CCD_IMPULSE || q.narrow == Narrow::MINCOL) && !q.t1)
        return bad_arg("triangles_next is NULL");
    if (q.narrow == Narrow::CCD_IMPULSE && (!q.masses || !q.dv || !q.dx || !q.cnt)) return bad_arg("NULL impulse tensor");
    if (q.narrow != Narrow::CCD_IMPULSE && (!q.out_rows || q.out_cap < 0)) return bad_arg("NULL output tensor");
    if (q.narrow == Narrow::CCD_COLLECT && !q.out_roots) return bad_arg("NULL roots tensor");
    if (q.narrow == Narrow::TRI_EDGE && q.te_mode == TE_COLLISIONS && !q.out_roots) return bad_arg("NULL coefficient tensor");
    if (q.narrow == Narrow::TRI_EDGE && q.te_mode == TE_CANDIDATES2 && !q.masks) return bad_arg("NULL mask tensor");
    CCC_HIP_TRY(hipSetDevice(q.ctx->device));
    if (q.dtype == CCC_FLOAT64) return run_query<double>(q);
    if (q.dtype == CCC_FLOAT32) return run_query<float>(q);
    return bad_arg("dtype must be CCC_FLOAT32 or CCC_FLOAT64");
}

template <typename T>
static ccc_status from_collisions(ccc_context *ctx, const int64_t *coll, const void *roots, int64_t m, const void *t0,
                                  const void *t1, const void *masses, int64_t B, int64_t n, void *dv, void *dx, int32_t *cnt,
                                  hipStream_t s) {
    ccc_status st;
    if (sizeof(T) == 8) { st = fill_u64(dv, B * n * 9, 0ull, s); if (st == CCC_OK) st = fill_u64(dx, B * n * 9, 0ull, s); }
    else { st = fill_u32(dv, B * n * 9, 0u, s); if (st == CCC_OK) st = fill_u32(dx, B * n * 9, 0u, s); }
    if (st == CCC_OK) st = fill_u32(cnt, B * n * 3, 0u, s);
    if (st != CCC_OK) return st;
    KernelTimer tm(ctx, s);
    if (m > 0) {
        tm.begin("from_collisions");
        hipLaunchKernelGGL((k_from_collisions<T>), dim3(grid_for(m, kThreads), B), dim3(kThreads), 0, s, coll, (const T *)roots, m,
                           (const T *)t0, (const T *)t1, (const T *)masses, n, (T *)dv, (T *)dx, cnt);
        tm.end();
        CCC_LAUNCH_CHECK("k_from_collisions");
    }
    CCC_HIP_TRY(hipStreamSynchronize(s));
    return CCC_OK;
}

// mesh_end[i] = offsets[m + 1] for the mesh m that owns triangle i (offsets ascending).
__global__ void k_mesh_end(const int64_t *__restrict__ offsets, int64_t meshes, int64_t total, int32_t *__restrict__ mesh_end) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int64_t lo = 0, hi = meshes - 1;      // first m with offsets[m + 1] > i
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (offsets[mid + 1] > i) hi = mid;
        else lo = mid + 1;
    }
    mesh_end[i] = (int32_t)offsets[lo + 1];
}

// Independent meshes of different sizes stored back to back, handled as one soup: one BVH
// over the concatenation (compat Morton codes are all zero, so leaf position = triangle
// index and every subtree covers a contiguous index range) and one traversal whose queries
// never enter index ranges of later meshes.  Every stage after the traversal works on the
// resulting pairs unchanged.  The pair set of each mesh is exactly its own single-mesh set.
static ccc_status packed_impulses(ccc_context *ctx, const void *t0, const void *t1, const void *masses, const int64_t *offsets,
                                  int64_t meshes, int32_t dtype, int32_t max_candidates, void *dv, void *dx, int32_t *cnt,
                                  ccc_counts *counts, uint32_t flags, hipStream_t s) {
    clear_error();
    if (!ctx) return bad_arg("context is NULL");
    if (meshes < 0 || (meshes > 0 && !offsets)) return bad_arg("offsets is NULL");
    if (meshes == 0) return CCC_OK;
    if (offsets[0] != 0) return bad_arg("offsets[0] must be 0");
    for (int64_t m = 0; m < meshes; ++m)
        if (offsets[m + 1] <= offsets[m]) return bad_arg("every mesh needs at least one triangle (offsets strictly increasing)");
    const int64_t n = offsets[meshes];
    if (n >= ((int64_t)1 << 28)) return bad_arg("at most 2^28 - 1 triangles in a packed batch");
    if (!t0 || !t1 || !masses || !dv || !dx || !cnt) return bad_arg("NULL tensor pointer");
    if (max_candidates < 1) return bad_arg("max_candidates must be >= 1");
    if (dtype != CCC_FLOAT32 && dtype != CCC_FLOAT64) return bad_arg("dtype must be CCC_FLOAT32 or CCC_FLOAT64");
    CCC_HIP_TRY(hipSetDevice(ctx->device));
    const size_t esz = dtype == CCC_FLOAT64 ? 8 : 4;
    Workspace ws(ctx, s);
    void *bboxes = ws.get((size_t)(2 * n - 1) * 6 * esz);
    int64_t *tree = ws.get_as<int64_t>((2 * n - 1) * 4);
    int32_t *mesh_end = ws.get_as<int32_t>(n);
    int64_t *d_off = ws.get_as<int64_t>(meshes + 1);
    if (ws.failed()) return CCC_ERR_NO_MEMORY;
    CCC_HIP_TRY(hipMemcpyAsync(d_off, offsets, sizeof(int64_t) * (meshes + 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mesh_end, dim3(grid_for(n, 256)), dim3(256), 0, s, d_off, meshes, n, mesh_end);
    CCC_LAUNCH_CHECK("k_mesh_end");
    // the packed traversal needs leaf position = triangle index (the reference's codes)
    ccc_status st = build_bvh_device(ctx, t0, t1, 1, n, dtype, CCC_MORTON_COMPAT, bboxes, tree, s);
    if ((st & CCC_ERR_MASK) != CCC_OK) return st;
    Query q{};
    q.ctx = ctx;
    q.bboxes = bboxes;
    q.tree = tree;
    q.t0 = t0;
    q.t1 = t1;
    q.masses = masses;
    q.B = 1;
    q.n = n;
    q.dtype = dtype;
    q.flags = flags;
    q.s = s;
    q.cap_candidates = (int64_t)max_candidates * n;
    q.narrow = Narrow::CCD_IMPULSE;
    q.dv = dv;
    q.dx = dx;
    q.cnt = cnt;
    q.counts = counts;
    q.mesh_end = mesh_end;
    if (dtype == CCC_FLOAT64) return run_query<double>(q);
    return run_query<float>(q);
}

// Scratch of the allocation-free entry point (ccc_compute_impulses_noalloc): every buffer
// of a call is carved from this array.  Each GPU that loads this code object holds its own
// copy until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (the array's name on the host is a
// stand-in, not a device address).
static __device__ __align__(256) unsigned char scratch[CCC_NOALLOC_SCRATCH_BYTES];

ccc_status scratch_address(int device, void **out) {
    static std::mutex mu;
    static std::map<int, void *> cached;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cached.find(device);
    if (it != cached.end()) {
        *out = it->second;
        return CCC_OK;
    }
    int prev = 0;
    CCC_HIP_TRY(hipGetDevice(&prev));
    CCC_HIP_TRY(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    cached[device] = p;
    *out = p;
    return CCC_OK;
}

}  // namespace ccc

using ccc::Narrow;
using ccc::Query;

static Query base_query(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *t0, int64_t B, int64_t n,
                        int32_t dtype, uint32_t flags, void *stream) {
    Query q{};
    q.ctx = ctx;
    q.bboxes = bboxes;
    q.tree = tree;
    q.t0 = t0;
    q.B = B;
    q.n = n;
    q.dtype = dtype;
    q.flags = flags;
    q.s = (hipStream_t)stream;
    return q;
}

static ccc_status tri_edge_query(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles,
                                 int64_t batch, int64_t n, int32_t dtype, int32_t max_collisions, int mode, int64_t *out,
                                 void *coefs, int32_t *masks, ccc_counts *counts, uint32_t flags, void *stream) {
    Query q = base_query(ctx, bboxes, tree, triangles, batch, n, dtype, flags, stream);
    q.thr = 0.0;
    q.cap_candidates = (int64_t)max_collisions * n;
    q.out_cap = (int64_t)max_collisions * n;
    q.narrow = Narrow::TRI_EDGE;
    q.te_mode = mode;
    q.out_rows = out;
    q.out_roots = coefs;
    q.masks = masks;
    q.counts = counts;
    if (max_collisions < 1) return ccc::bad_arg("max_collisions must be >= 1");
    return ccc::run(q);
}

extern "C" {

ccc_status ccc_find_collisions(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles, int64_t batch,
                               int64_t n, int32_t dtype, float threshold, int32_t max_collisions, int64_t *out,
                               ccc_counts *counts, uint32_t flags, void *stream) {
    Query q = base_query(ctx, bboxes, tree, triangles, batch, n, dtype, flags, stream);
    q.thr = (double)threshold;
    q.cap_candidates = (int64_t)max_collisions * n;
    q.out_cap = (int64_t)max_collisions * n;
    q.narrow = Narrow::SAT;
    q.out_rows = out;
    q.counts = counts;
    if (max_collisions < 1) return ccc::bad_arg("max_collisions must be >= 1");
    return ccc::run(q);
}

ccc_status ccc_find_triangle_edge_collisions(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles,
                                             int64_t batch, int64_t n, int32_t dtype, int32_t max_collisions, int64_t *out,
                                             void *coefs, ccc_counts *counts, uint32_t flags, void *stream) {
    return tri_edge_query(ctx, bboxes, tree, triangles, batch, n, dtype, max_collisions, ccc::TE_COLLISIONS, out, coefs, nullptr,
                          counts, flags, stream);
}

ccc_status ccc_find_triangle_edge_candidates(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles,
                                             int64_t batch, int64_t n, int32_t dtype, int32_t max_collisions, int64_t *out,
                                             ccc_counts *counts, uint32_t flags, void *stream) {
    return tri_edge_query(ctx, bboxes, tree, triangles, batch, n, dtype, max_collisions, ccc::TE_CANDIDATES, out, nullptr, nullptr,
                          counts, flags, stream);
}

ccc_status ccc_find_triangle_edge_candidates2(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles,
                                              int64_t batch, int64_t n, int32_t dtype, int32_t max_collisions, int64_t *pairs,
                                              int32_t *masks, ccc_counts *counts, uint32_t flags, void *stream) {
    return tri_edge_query(ctx, bboxes, tree, triangles, batch, n, dtype, max_collisions, ccc::TE_CANDIDATES2, pairs, nullptr, masks,
                          counts, flags, stream);
}

ccc_status ccc_compute_impulses_packed(ccc_context *ctx, const void *triangles, const void *triangles_next, const void *masses,
                                       const int64_t *offsets, int64_t num_meshes, int32_t dtype, int32_t max_candidates,
                                       void *dv, void *dx, int32_t *cnt, ccc_counts *counts, uint32_t flags, void *stream) {
    return ccc::packed_impulses(ctx, triangles, triangles_next, masses, offsets, num_meshes, dtype, max_candidates, dv, dx, cnt,
                                counts, flags, (hipStream_t)stream);
}

ccc_status ccc_find_proximity(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles, int64_t batch,
                              int64_t n, int32_t dtype, float threshold, int32_t max_candidates, int64_t *out,
                              ccc_counts *counts, uint32_t flags, void *stream) {
    Query q = base_query(ctx, bboxes, tree, triangles, batch, n, dtype, flags, stream);
    q.thr = (double)threshold;
    q.cap_candidates = (int64_t)max_candidates * n;
    q.out_cap = (int64_t)max_candidates * n;
    q.narrow = Narrow::PROXIMITY;
    q.out_rows = out;
    q.counts = counts;
    if (max_candidates < 1) return ccc::bad_arg("max_candidates_per_triangle must be >= 1");
    return ccc::run(q);
}

static ccc_status continuous(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *t0, const void *t1,
                             const uint8_t *to_check, int64_t batch, int64_t n, int32_t dtype, int32_t max_candidates,
                             int32_t max_collisions, int64_t *collisions, void *roots, ccc_counts *counts, uint32_t flags,
                             void *stream) {
    Query q = base_query(ctx, bboxes, tree, t0, batch, n, dtype, flags, stream);
    q.t1 = t1;
    q.to_check = to_check;
    q.thr = 0.0;
    q.cap_candidates = (int64_t)max_candidates * n;
    q.out_cap = (int64_t)max_collisions * n;
    q.narrow = Narrow::CCD_COLLECT;
    q.out_rows = collisions;
    q.out_roots = roots;
    q.counts = counts;
    if (max_candidates < 1 || max_collisions < 1) return ccc::bad_arg("capacities must be >= 1");
    return ccc::run(q);
}

ccc_status ccc_find_collisions_continuous(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles,
                                          const void *triangles_next, int64_t batch, int64_t n, int32_t dtype,
                                          int32_t max_candidates, int32_t max_collisions, int64_t *collisions, void *roots,
                                          ccc_counts *counts, uint32_t flags, void *stream) {
    return continuous(ctx, bboxes, tree, triangles, triangles_next, nullptr, batch, n, dtype, max_candidates, max_collisions,
                      collisions, roots, counts, flags, stream);
}

ccc_status ccc_find_collisions_continuous_partial(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                                                  const void *triangles, const void *triangles_next, const uint8_t *to_check,
                                                  int64_t batch, int64_t n, int32_t dtype, int32_t max_candidates,
                                                  int32_t max_collisions, int64_t *collisions, void *roots,
                                                  ccc_counts *counts, uint32_t flags, void *stream) {
    if (!to_check) { ccc::set_error("triangles_to_check is NULL"); return CCC_ERR_BAD_ARG; }
    return continuous(ctx, bboxes, tree, triangles, triangles_next, to_check, batch, n, dtype, max_candidates, max_collisions,
                      collisions, roots, counts, flags, stream);
}

ccc_status ccc_find_mincol_cand(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles,
                                const void *triangles_next, int64_t batch, int64_t n, int32_t dtype, int32_t max_collisions,
                                int64_t *out, ccc_counts *counts, uint32_t flags, void *stream) {
    Query q = base_query(ctx, bboxes, tree, triangles, batch, n, dtype, flags, stream);
    q.t1 = triangles_next;
    q.thr = 0.0;
    q.cap_candidates = (int64_t)max_collisions * n;
    q.out_cap = (int64_t)max_collisions * n * 15;
    q.narrow = Narrow::MINCOL;
    q.out_rows = out;
    q.counts = counts;
    if (max_collisions < 1) return ccc::bad_arg("max_collisions must be >= 1");
    return ccc::run(q);
}

static ccc_status impulses(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *t0, const void *t1,
                           const void *masses, const uint8_t *to_check, int64_t batch, int64_t n, int32_t dtype,
                           int32_t max_candidates, void *dv, void *dx, int32_t *cnt, ccc_counts *counts, uint32_t flags,
                           void *stream) {
    Query q = base_query(ctx, bboxes, tree, t0, batch, n, dtype, flags, stream);
    q.t1 = t1;
    q.masses = masses;
    q.to_check = to_check;
    q.thr = 0.0;
    q.cap_candidates = (int64_t)max_candidates * n;
    q.out_cap = 0;
    q.narrow = Narrow::CCD_IMPULSE;
    q.dv = dv;
    q.dx = dx;
    q.cnt = cnt;
    q.counts = counts;
    if (max_candidates < 1) return ccc::bad_arg("max_candidates_per_triangle must be >= 1");
    return ccc::run(q);
}

ccc_status ccc_compute_impulses(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles,
                                const void *triangles_next, const void *masses, int64_t batch, int64_t n, int32_t dtype,
                                int32_t max_candidates, int32_t max_collisions, void *dv, void *dx, int32_t *cnt,
                                ccc_counts *counts, uint32_t flags, void *stream) {
    (void)max_collisions;
    return impulses(ctx, bboxes, tree, triangles, triangles_next, masses, nullptr, batch, n, dtype, max_candidates, dv, dx, cnt,
                    counts, flags, stream);
}

ccc_status ccc_compute_impulses_noalloc(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles,
                                        const void *triangles_next, const void *masses, int64_t batch, int64_t n, int32_t dtype,
                                        int32_t max_candidates, int32_t max_collisions, void *dv, void *dx, int32_t *cnt,
                                        ccc_counts *counts, uint32_t flags, void *stream) {
    (void)max_collisions;
    ccc::clear_error();
    if (!ctx) return ccc::bad_arg("context is NULL");
    if (!ctx->scratch) { ccc::set_error("the context has no scratch address"); return CCC_ERR_HIP; }
    Query q = base_query(ctx, bboxes, tree, triangles, batch, n, dtype, flags, stream);
    q.t1 = triangles_next;
    q.masses = masses;
    q.thr = 0.0;
    q.cap_candidates = (int64_t)max_candidates * n;
    q.out_cap = 0;
    q.narrow = Narrow::CCD_IMPULSE;
    q.dv = dv;
    q.dx = dx;
    q.cnt = cnt;
    q.counts = counts;
    q.arena = ctx->scratch;
    q.arena_bytes = CCC_NOALLOC_SCRATCH_BYTES;
    if (max_candidates < 1) return ccc::bad_arg("max_candidates_per_triangle must be >= 1");
    return ccc::run(q);
}

ccc_status ccc_compute_impulses_partial(ccc_context *ctx, const void *bboxes, const int64_t *tree, const void *triangles,
                                        const void *triangles_next, const void *masses, const uint8_t *to_check, int64_t batch,
                                        int64_t n, int32_t dtype, int32_t max_candidates, int32_t max_collisions, void *dv,
                                        void *dx, int32_t *cnt, ccc_counts *counts, uint32_t flags, void *stream) {
    (void)max_collisions;
    if (!to_check) { ccc::set_error("triangles_to_check is NULL"); return CCC_ERR_BAD_ARG; }
    return impulses(ctx, bboxes, tree, triangles, triangles_next, masses, to_check, batch, n, dtype, max_candidates, dv, dx, cnt,
                    counts, flags, stream);
}

ccc_status ccc_compute_impulses_from_collisions(ccc_context *ctx, const int64_t *collisions, const void *roots, int64_t m,
                                                const void *triangles, const void *triangles_next, const void *masses,
                                                int64_t batch, int64_t n, int32_t dtype, void *dv, void *dx, int32_t *cnt,
                                                uint32_t flags, void *stream) {
    (void)flags;
    ccc::clear_error();
    if (!ctx) return ccc::bad_arg("context is NULL");
    if (batch < 0 || n < 1 || m < 0) return ccc::bad_arg("invalid sizes");
    if (batch == 0) return CCC_OK;
    if (!triangles || !triangles_next || !masses || !dv || !dx || !cnt || (m > 0 && (!collisions || !roots)))
        return ccc::bad_arg("NULL tensor pointer");
    CCC_HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    if (dtype == CCC_FLOAT64)
        return ccc::from_collisions<double>(ctx, collisions, roots, m, triangles, triangles_next, masses, batch, n, dv, dx, cnt, s);
    if (dtype == CCC_FLOAT32)
        return ccc::from_collisions<float>(ctx, collisions, roots, m, triangles, triangles_next, masses, batch, n, dv, dx, cnt, s);
    return ccc::bad_arg("dtype must be CCC_FLOAT32 or CCC_FLOAT64");
}

}  // extern "C"

extern "C" ccc_status ccc_debug_device_math(ccc_context *ctx, int what, const double *x, double *y, int64_t n,
                                            void *stream) {
    ccc::clear_error();
    if (!ctx) return ccc::bad_arg("context is NULL");
    if (what < 0 || what > 11) return ccc::bad_arg("what must be 0-11");
    if (n < 0 || (n > 0 && (!x || !y))) return ccc::bad_arg("x and y must be device arrays of n doubles (what 5: 4 n each; what 10, 11: x 2 n)");
    if (n == 0) return CCC_OK;
    CCC_HIP_TRY(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(ccc::k_device_math, dim3(ccc::grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, what, x, y, n);
    CCC_LAUNCH_CHECK("k_device_math");
    return CCC_OK;
}
