// Resident simulations: the caller's side of the reference's impulse path, kept on the
// device and replayed as one HIP graph per step.
//
// A simulation (one mesh, or several meshes of different sizes packed back to back) holds
// its state resident in HBM: start-of-step positions x0, predicted end-of-step positions x1,
// faces, per-corner masses, the triangle soups, its BVH, its candidate-pair list, the
// impulse outputs and a vertex -> corner index (CSR).  One step runs `iters` impulse passes,
// each: gather the end-of-step soup from x1, bvh_motion (src/build_bvh.cu:570-648),
// compute_impulses (src/collisions_impulses.cu:1069-1397), then move every vertex of x1 by
// the mean of the finite position corrections of its triangle corners.  The reference leaves
// that update to its caller (the simulator is not part of it); it is defined here as
// cccollisions_amd.scenes.apply_corrections defines it, summing each vertex's corners in
// corner order, so the kernel gives the same bits as that numpy function.
//
// Nothing in a step waits on the host or allocates: workspace comes from an arena (the
// simulation's own, or one the caller passes, e.g. a serving slot's), the candidate list has
// a fixed capacity and the narrow phase a fixed number of chunks.  The step can therefore be
// captured once in a HIP graph and replayed.  After each step the host reads the per-pass
// counters once: if a pass found more candidate pairs than the list holds, or its traversal
// needed the deep stack, the capacity (or stack) grows and the step runs again from the
// saved x1.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <tuple>

#include "ccc_geom.h"
#include "ccc_internal.h"

namespace ccc {

// soup[f, k] = x[faces[f, k]] for every face corner.
template <typename T>
__global__ void k_gather_soup(const T *__restrict__ x, const int64_t *__restrict__ faces, int64_t corners, T *__restrict__ soup) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= corners) return;
    const int64_t v = faces[c];
    soup[3 * c + 0] = x[3 * v + 0];
    soup[3 * c + 1] = x[3 * v + 1];
    soup[3 * c + 2] = x[3 * v + 2];
}

// x1[v] += (sum of the finite dx of v's corners) / (sum of their cnt), over v's corners in
// ascending corner order starting from zero (the order numpy.add.at uses); a corner whose
// correction has a non-finite component contributes neither.
template <typename T>
__global__ void k_apply_corrections(T *__restrict__ x1, const T *__restrict__ dx, const int32_t *__restrict__ cnt,
                                    const int32_t *__restrict__ off, const int32_t *__restrict__ corner, int64_t nv) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    T sx = 0, sy = 0, sz = 0;
    double k = 0.0;
    for (int32_t j = off[v]; j < off[v + 1]; ++j) {
        const int64_t c = corner[j];
        const T a = dx[3 * c + 0], b = dx[3 * c + 1], d = dx[3 * c + 2];
        if (!(isfinite(a) && isfinite(b) && isfinite(d))) continue;
        sx = sx + a;
        sy = sy + b;
        sz = sz + d;
        k = k + (double)cnt[c];
    }
    if (k > 0.0) {
        x1[3 * v + 0] = (T)((double)x1[3 * v + 0] + (double)sx / k);
        x1[3 * v + 1] = (T)((double)x1[3 * v + 1] + (double)sy / k);
        x1[3 * v + 2] = (T)((double)x1[3 * v + 2] + (double)sz / k);
    }
}

template <typename T>
__global__ void k_copy(T *__restrict__ dst, const T *__restrict__ src, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

// Next step's state: x0 <- x1 (the corrected positions), x1 <- x1 + (x1 - x0) (the same
// displacement again: a constant-velocity prediction).
template <typename T>
__global__ void k_advance(T *__restrict__ x0, T *__restrict__ x1, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const T a = x0[i], b = x1[i];
        x0[i] = b;
        x1[i] = b + (b - a);
    }
}

static unsigned stride_grid(int64_t count) {
    int64_t g = (count + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 65536));
}

// Autotuned launch configurations, shared by the simulations of one context whose meshes
// fall in the same size bucket.
struct TunedKey {
    int bucket, dtype, det, packed;
    bool operator<(const TunedKey &o) const {
        return std::tie(bucket, dtype, det, packed) < std::tie(o.bucket, o.dtype, o.det, o.packed);
    }
};
struct TunedCfg {
    int ccd_blocks_per_cu, morton;
};
static std::mutex g_tune_mu;
static std::map<std::pair<ccc_context *, TunedKey>, TunedCfg> g_tuned;

}  // namespace ccc

using namespace ccc;

struct ccc_sim {
    ccc_context *ctx = nullptr;
    int32_t dtype = CCC_FLOAT64;
    size_t esz = 8;
    int64_t nv = 0, nf = 0, meshes = 1;
    ccc_sim_opts o{};
    // resident device state
    void *x0 = nullptr, *x1 = nullptr, *x1_save = nullptr, *masses = nullptr, *t0 = nullptr, *t1 = nullptr;
    void *bboxes = nullptr, *dv = nullptr, *dx = nullptr;
    int64_t *faces = nullptr, *tree = nullptr;
    int32_t *cnt = nullptr, *csr_off = nullptr, *csr_corner = nullptr, *mesh_end = nullptr;
    int2 *pairs = nullptr;
    int64_t cap_pairs = 0;
    long long *det_acc = nullptr;
    DevCounters *ctr = nullptr;
    DevCounters *h_ctr = nullptr;         // pinned host copy of the per-pass counters
    void *arena = nullptr;                // own workspace (allocated at the first step that needs it)
    size_t arena_bytes = 0;
    bool deep = false;
    // graphs, one per workspace arena the steps ran with
    hipStream_t cs = nullptr;             // capture stream
    std::map<void *, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
    std::vector<ccc_counts> counts;       // per pass of the last step
    double tuned_ms[8] = {0};
    int tuned_n = 0;
};

namespace ccc {

static ccc_status dmalloc(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        set_error("device allocation of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
        return CCC_ERR_NO_MEMORY;
    }
    return CCC_OK;
}
static void dfree(void *p) {
    if (p) (void)hipFree(p);
}

static void drop_graphs(ccc_sim *s) {
    for (auto &kv : s->graphs) {
        (void)hipGraphExecDestroy(kv.second.second);
        (void)hipGraphDestroy(kv.second.first);
    }
    s->graphs.clear();
}

static ResidentPass pass_of(const ccc_sim *s, int it) {
    ResidentPass p{};
    p.bboxes = s->bboxes;
    p.tree = s->tree;
    p.t0 = s->t0;
    p.t1 = s->t1;
    p.masses = s->masses;
    p.n = s->nf;
    p.dtype = s->dtype;
    p.mesh_end = s->mesh_end;
    p.pairs = s->pairs;
    p.cap_pairs = s->cap_pairs;
    p.chunk = s->o.pair_chunk > 0 ? s->o.pair_chunk : 0;
    p.ccd_blocks_per_cu = s->o.ccd_blocks_per_cu;
    p.stack_deep = s->deep;
    p.det_acc = s->det_acc;
    p.dv = s->dv;
    p.dx = s->dx;
    p.cnt = s->cnt;
    p.ctr = s->ctr + it;
    return p;
}

static size_t arena_need(const ccc_sim *s) {
    const int morton = s->meshes > 1 ? CCC_MORTON_COMPAT : s->o.morton;
    return std::max(build_bvh_arena_bytes(1, s->nf, s->dtype, morton), resident_pass_arena_bytes(s->ctx, pass_of(s, 0)));
}

// The kernels of one step, enqueued on `st`.
template <typename T>
static ccc_status enqueue_step(ccc_sim *s, void *arena, size_t bytes, hipStream_t st) {
    const int64_t corners = 3 * s->nf, coords = 3 * s->nv;
    KernelTimer tm(s->ctx, st);     // records only while the context profiles (eager steps)
    hipLaunchKernelGGL((k_copy<T>), dim3(stride_grid(coords)), dim3(256), 0, st, (T *)s->x1_save, (const T *)s->x1, coords);
    hipLaunchKernelGGL((k_gather_soup<T>), dim3(grid_for(corners, 256)), dim3(256), 0, st, (const T *)s->x0, s->faces, corners,
                       (T *)s->t0);
    CCC_LAUNCH_CHECK("step preparation");
    const int morton = s->meshes > 1 ? CCC_MORTON_COMPAT : s->o.morton;
    for (int it = 0; it < s->o.iters; ++it) {
        CCC_TIMED(tm, "sim_gather", hipLaunchKernelGGL((k_gather_soup<T>), dim3(grid_for(corners, 256)), dim3(256), 0, st,
                                                       (const T *)s->x1, s->faces, corners, (T *)s->t1));
        CCC_LAUNCH_CHECK("k_gather_soup");
        ccc_status r = build_bvh_device(s->ctx, s->t0, s->t1, 1, s->nf, s->dtype, morton, s->bboxes, s->tree, st, arena, bytes);
        if (r != CCC_OK) return r;
        r = resident_pass(s->ctx, pass_of(s, it), arena, bytes, st);
        if (r != CCC_OK) return r;
        CCC_TIMED(tm, "sim_update", hipLaunchKernelGGL((k_apply_corrections<T>), dim3(grid_for(s->nv, 256)), dim3(256), 0, st,
                                                       (T *)s->x1, (const T *)s->dx, s->cnt, s->csr_off, s->csr_corner, s->nv));
        CCC_LAUNCH_CHECK("k_apply_corrections");
    }
    return CCC_OK;
}


static ccc_status enqueue(ccc_sim *s, void *arena, size_t bytes, hipStream_t st) {
    return s->dtype == CCC_FLOAT64 ? enqueue_step<double>(s, arena, bytes, st) : enqueue_step<float>(s, arena, bytes, st);
}

// Launches one step on `st` (through the graph cached for `arena`, capturing it first).
static ccc_status launch_step(ccc_sim *s, void *arena, size_t bytes, hipStream_t st) {
    // a profiled step times its kernels with events and runs eagerly
    if (!s->o.graph || s->ctx->profiling) return enqueue(s, arena, bytes, st);
    auto it = s->graphs.find(arena);
    if (it == s->graphs.end()) {
        CCC_HIP_TRY(hipStreamBeginCapture(s->cs, hipStreamCaptureModeThreadLocal));
        const ccc_status r = enqueue(s, arena, bytes, s->cs);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s->cs, &g);
        if (r != CCC_OK) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            return hip_fail(ei, "hipGraphInstantiate");
        }
        it = s->graphs.emplace(arena, std::make_pair(g, x)).first;
    }
    CCC_HIP_TRY(hipGraphLaunch(it->second.second, st));
    return CCC_OK;
}

template <typename T>
static ccc_status restore_x1(ccc_sim *s, hipStream_t st) {
    hipLaunchKernelGGL((k_copy<T>), dim3(stride_grid(3 * s->nv)), dim3(256), 0, st, (T *)s->x1, (const T *)s->x1_save, 3 * s->nv);
    CCC_LAUNCH_CHECK("restore x1");
    return CCC_OK;
}

static ccc_status grow_pairs(ccc_sim *s, int64_t cap) {
    dfree(s->pairs);
    s->pairs = nullptr;
    s->cap_pairs = 0;
    ccc_status r = dmalloc((void **)&s->pairs, sizeof(int2) * (size_t)cap);
    if (r != CCC_OK) return r;
    s->cap_pairs = cap;
    return CCC_OK;
}

// One step with the given workspace; re-runs it (from the saved x1) after a capacity or
// stack overflow.  Synchronises `st`.
static ccc_status run_step(ccc_sim *s, void *arena, size_t bytes, hipStream_t st) {
    for (int attempt = 0; attempt < 4; ++attempt) {
        const bool own = arena == nullptr;
        const size_t need = arena_need(s);
        void *a = arena;
        size_t ab = bytes;
        if (own) {
            if (s->arena_bytes < need) {
                CCC_HIP_TRY(hipStreamSynchronize(st));
                drop_graphs(s);
                dfree(s->arena);
                s->arena = nullptr;
                s->arena_bytes = 0;
                ccc_status r = dmalloc(&s->arena, need);
                if (r != CCC_OK) return r;
                s->arena_bytes = need;
            }
            a = s->arena;
            ab = s->arena_bytes;
        } else if (bytes < need) {
            set_error("the workspace arena holds " + std::to_string(bytes) + " bytes; this simulation's step needs " +
                      std::to_string(need));
            return CCC_ERR_NO_MEMORY;
        }
        ccc_status r = launch_step(s, a, ab, st);
        if (r != CCC_OK) return r;
        const int iters = s->o.iters;
        CCC_HIP_TRY(hipMemcpyAsync(s->h_ctr, s->ctr, sizeof(DevCounters) * iters, hipMemcpyDeviceToHost, st));
        CCC_HIP_TRY(hipStreamSynchronize(st));
        unsigned flags = 0;
        unsigned long long kmax = 0;
        for (int i = 0; i < iters; ++i) {
            flags |= s->h_ctr[i].flags;
            kmax = std::max(kmax, s->h_ctr[i].cand);
        }
        if (flags & DEV_BAD_TREE) {
            set_error("internal error: the step's tree is not a valid BVH");
            return CCC_ERR_BAD_TREE;
        }
        const bool stack = (flags & DEV_STACK) != 0;
        const bool over = kmax > (unsigned long long)s->cap_pairs;
        if (!stack && !over) {
            s->counts.assign(iters, ccc_counts{});
            for (int i = 0; i < iters; ++i) {
                s->counts[i].candidates = (int64_t)s->h_ctr[i].cand;
                s->counts[i].subprims = (int64_t)s->h_ctr[i].items_done;
                s->counts[i].roots = (int64_t)s->h_ctr[i].roots;
                s->counts[i].collisions = (int64_t)s->h_ctr[i].coll;
            }
            return CCC_OK;
        }
        if (stack && s->deep) {
            set_error("BVH deeper than the traversal stack");
            return CCC_ERR_STACK;
        }
        // the step ran on a truncated list or a failed traversal: grow, restore x1, again
        drop_graphs(s);
        if (stack) s->deep = true;
        if (over) {
            r = grow_pairs(s, (int64_t)(kmax + kmax / 4 + 1024));
            if (r != CCC_OK) return r;
        }
        r = s->dtype == CCC_FLOAT64 ? restore_x1<double>(s, st) : restore_x1<float>(s, st);
        if (r != CCC_OK) return r;
    }
    set_error("the step kept overflowing its candidate capacity");
    return CCC_ERR_NO_MEMORY;
}

}  // namespace ccc

extern "C" {

void ccc_sim_default_opts(ccc_sim_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->iters = 10;
    o->max_candidates = 64;
    o->deterministic = 0;
    o->graph = 1;
    o->morton = CCC_MORTON_COMPAT;
    o->ccd_blocks_per_cu = 128;
    o->pair_chunk = 0;
}

ccc_status ccc_sim_create(ccc_context *ctx, const void *x0, const void *x1, int64_t num_vertices, const int64_t *faces,
                          const void *masses, int64_t num_faces, const int64_t *mesh_face_offsets, int64_t num_meshes,
                          int32_t dtype, const ccc_sim_opts *opts, void *stream, ccc_sim **out) {
    clear_error();
    if (!out) return bad_arg("out is NULL");
    *out = nullptr;
    if (!ctx) return bad_arg("context is NULL");
    if (!x0 || !x1 || !faces || !masses) return bad_arg("NULL tensor pointer");
    if (num_vertices < 1 || num_faces < 1) return bad_arg("need at least one vertex and one face");
    if (num_faces >= ((int64_t)1 << 28) || 3 * num_faces >= INT32_MAX) return bad_arg("at most 2^28 - 1 faces");
    if (num_vertices >= INT32_MAX) return bad_arg("too many vertices");
    if (dtype != CCC_FLOAT32 && dtype != CCC_FLOAT64) return bad_arg("dtype must be CCC_FLOAT32 or CCC_FLOAT64");
    ccc_sim_opts o;
    ccc_sim_default_opts(&o);
    if (opts) o = *opts;
    if (o.iters < 1) return bad_arg("iters must be >= 1");
    if (o.max_candidates < 1) return bad_arg("max_candidates must be >= 1");
    if (o.morton != CCC_MORTON_COMPAT && o.morton != CCC_MORTON_30 && o.morton != CCC_MORTON_63) return bad_arg("unknown Morton mode");
    if (o.ccd_blocks_per_cu < 0 || o.pair_chunk < 0) return bad_arg("launch parameters must be >= 0");
    if (num_meshes < 1) num_meshes = 1;
    std::vector<int64_t> offs;
    if (mesh_face_offsets && num_meshes > 1) {
        offs.assign(mesh_face_offsets, mesh_face_offsets + num_meshes + 1);
        if (offs[0] != 0 || offs[num_meshes] != num_faces) return bad_arg("mesh offsets must run from 0 to num_faces");
        for (int64_t m = 0; m < num_meshes; ++m)
            if (offs[m + 1] <= offs[m]) return bad_arg("every mesh needs at least one face (offsets strictly increasing)");
    } else {
        num_meshes = 1;
    }
    CCC_HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;

    // vertex -> corner index on the host (one time): corners of each vertex in ascending order
    std::vector<int64_t> hf((size_t)(3 * num_faces));
    CCC_HIP_TRY(hipMemcpyAsync(hf.data(), faces, sizeof(int64_t) * hf.size(), hipMemcpyDeviceToHost, st));
    CCC_HIP_TRY(hipStreamSynchronize(st));
    std::vector<int32_t> off((size_t)num_vertices + 1, 0), corner(hf.size());
    for (int64_t v : hf) {
        if (v < 0 || v >= num_vertices) return bad_arg("face vertex index out of range");
        ++off[(size_t)v + 1];
    }
    for (int64_t v = 0; v < num_vertices; ++v) off[(size_t)v + 1] += off[(size_t)v];
    {
        std::vector<int32_t> fill(off.begin(), off.end() - 1);
        for (size_t c = 0; c < hf.size(); ++c) corner[(size_t)fill[(size_t)hf[c]]++] = (int32_t)c;
    }
    std::vector<int32_t> mend;
    if (num_meshes > 1) {
        mend.resize((size_t)num_faces);
        for (int64_t m = 0; m < num_meshes; ++m)
            for (int64_t f = offs[m]; f < offs[m + 1]; ++f) mend[(size_t)f] = (int32_t)offs[m + 1];
    }

    ccc_sim *s = new (std::nothrow) ccc_sim;
    if (!s) return CCC_ERR_NO_MEMORY;
    s->ctx = ctx;
    s->dtype = dtype;
    s->esz = dtype == CCC_FLOAT64 ? 8 : 4;
    s->nv = num_vertices;
    s->nf = num_faces;
    s->meshes = num_meshes;
    s->o = o;
    const size_t vb = s->esz * 3 * (size_t)num_vertices, sb = s->esz * 9 * (size_t)num_faces, rows = 2 * (size_t)num_faces - 1;
    ccc_status r = CCC_OK;
    auto A = [&](void **p, size_t b) { if (r == CCC_OK) r = dmalloc(p, b); };
    A(&s->x0, vb); A(&s->x1, vb); A(&s->x1_save, vb);
    A(&s->masses, s->esz * 3 * (size_t)num_faces);
    A(&s->t0, sb); A(&s->t1, sb); A(&s->dv, sb); A(&s->dx, sb);
    A(&s->bboxes, s->esz * 6 * rows);
    A((void **)&s->faces, sizeof(int64_t) * hf.size());
    A((void **)&s->tree, sizeof(int64_t) * 4 * rows);
    A((void **)&s->cnt, sizeof(int32_t) * 3 * (size_t)num_faces);
    A((void **)&s->csr_off, sizeof(int32_t) * off.size());
    A((void **)&s->csr_corner, sizeof(int32_t) * corner.size());
    A((void **)&s->ctr, sizeof(DevCounters) * (size_t)o.iters);
    if (num_meshes > 1) A((void **)&s->mesh_end, sizeof(int32_t) * mend.size());
    if (o.deterministic) A((void **)&s->det_acc, sizeof(long long) * kDetWordsPerCorner * 3 * (size_t)num_faces);
    if (r == CCC_OK) r = grow_pairs(s, (int64_t)o.max_candidates * num_faces);
    if (r == CCC_OK && hipHostMalloc((void **)&s->h_ctr, sizeof(DevCounters) * (size_t)o.iters) != hipSuccess) {
        set_error("pinned host allocation failed");
        r = CCC_ERR_NO_MEMORY;
    }
    if (r == CCC_OK && hipStreamCreateWithFlags(&s->cs, hipStreamNonBlocking) != hipSuccess) {
        set_error("stream creation failed");
        r = CCC_ERR_HIP;
    }
    if (r != CCC_OK) {
        ccc_sim_destroy(s);
        return r;
    }
    auto C = [&](void *dst, const void *src, size_t b, hipMemcpyKind k) {
        if (r == CCC_OK && hipMemcpyAsync(dst, src, b, k, st) != hipSuccess) {
            set_error("copy into the simulation failed");
            r = CCC_ERR_HIP;
        }
    };
    C(s->x0, x0, vb, hipMemcpyDeviceToDevice);
    C(s->x1, x1, vb, hipMemcpyDeviceToDevice);
    C(s->masses, masses, s->esz * 3 * (size_t)num_faces, hipMemcpyDeviceToDevice);
    C(s->faces, hf.data(), sizeof(int64_t) * hf.size(), hipMemcpyHostToDevice);
    C(s->csr_off, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice);
    C(s->csr_corner, corner.data(), sizeof(int32_t) * corner.size(), hipMemcpyHostToDevice);
    if (num_meshes > 1) C(s->mesh_end, mend.data(), sizeof(int32_t) * mend.size(), hipMemcpyHostToDevice);
    if (r == CCC_OK && hipStreamSynchronize(st) != hipSuccess) r = CCC_ERR_HIP;
    if (r != CCC_OK) {
        ccc_sim_destroy(s);
        return r;
    }
    *out = s;
    return CCC_OK;
}

void ccc_sim_destroy(ccc_sim *s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipDeviceSynchronize();
    drop_graphs(s);
    for (void *p : {s->x0, s->x1, s->x1_save, s->masses, s->t0, s->t1, s->bboxes, s->dv, s->dx, (void *)s->faces, (void *)s->tree,
                    (void *)s->cnt, (void *)s->csr_off, (void *)s->csr_corner, (void *)s->mesh_end, (void *)s->pairs,
                    (void *)s->det_acc, (void *)s->ctr, s->arena})
        dfree(p);
    if (s->h_ctr) (void)hipHostFree(s->h_ctr);
    if (s->cs) (void)hipStreamDestroy(s->cs);
    delete s;
}

size_t ccc_sim_workspace_bytes(const ccc_sim *s) { return s ? arena_need(s) : 0; }

ccc_status ccc_sim_step(ccc_sim *s, void *workspace, size_t workspace_bytes, int32_t advance, void *stream) {
    clear_error();
    if (!s) return bad_arg("simulation is NULL");
    CCC_HIP_TRY(hipSetDevice(s->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    ccc_status r = run_step(s, workspace, workspace_bytes, st);
    if (r != CCC_OK) return r;
    if (advance) {
        if (s->dtype == CCC_FLOAT64)
            hipLaunchKernelGGL((k_advance<double>), dim3(stride_grid(3 * s->nv)), dim3(256), 0, st, (double *)s->x0, (double *)s->x1, 3 * s->nv);
        else
            hipLaunchKernelGGL((k_advance<float>), dim3(stride_grid(3 * s->nv)), dim3(256), 0, st, (float *)s->x0, (float *)s->x1, 3 * s->nv);
        CCC_LAUNCH_CHECK("k_advance");
    }
    return CCC_OK;
}

ccc_status ccc_sim_autotune(ccc_sim *s, void *workspace, size_t workspace_bytes, void *stream) {
    clear_error();
    if (!s) return bad_arg("simulation is NULL");
    CCC_HIP_TRY(hipSetDevice(s->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    const TunedKey key{(int)std::floor(std::log2((double)s->nf) * 2.0), s->dtype, s->o.deterministic, s->meshes > 1 ? 1 : 0};
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tuned.find({s->ctx, key});
        if (it != g_tuned.end()) {
            if (s->o.ccd_blocks_per_cu != it->second.ccd_blocks_per_cu || s->o.morton != it->second.morton) drop_graphs(s);
            s->o.ccd_blocks_per_cu = it->second.ccd_blocks_per_cu;
            s->o.morton = it->second.morton;
            s->tuned_n = 0;
            return CCC_OK;
        }
    }
    // a first step settles the candidate capacity and the stack; every timed step then starts
    // from the same x1
    const size_t vb = s->esz * 3 * (size_t)s->nv;
    void *x1_start = nullptr;
    ccc_status r = dmalloc(&x1_start, vb);
    if (r != CCC_OK) return r;
    CCC_HIP_TRY(hipMemcpyAsync(x1_start, s->x1, vb, hipMemcpyDeviceToDevice, st));
    auto reset = [&]() { return hipMemcpyAsync(s->x1, x1_start, vb, hipMemcpyDeviceToDevice, st); };
    r = run_step(s, workspace, workspace_bytes, st);
    const int grids[3] = {64, 128, 256};
    const int mortons[2] = {CCC_MORTON_COMPAT, CCC_MORTON_30};
    const int nm = s->meshes > 1 ? 1 : 2;
    TunedCfg best{s->o.ccd_blocks_per_cu, s->o.morton};
    double best_ms = 1e300;
    s->tuned_n = 0;
    for (int mi = 0; mi < nm && r == CCC_OK; ++mi) {
        for (int gi = 0; gi < 3 && r == CCC_OK; ++gi) {
            drop_graphs(s);
            s->o.ccd_blocks_per_cu = grids[gi];
            s->o.morton = mortons[mi];
            // with graphs the first run also captures; steps longer than 50 ms are timed once
            // (capture is then a small part), shorter ones once more after the capture
            double ms = 1e300;
            for (int rep = 0; rep < 2 && r == CCC_OK; ++rep) {
                if (reset() != hipSuccess) r = CCC_ERR_HIP;
                if (r == CCC_OK && hipStreamSynchronize(st) != hipSuccess) r = CCC_ERR_HIP;
                const auto t0 = std::chrono::steady_clock::now();
                if (r == CCC_OK) r = run_step(s, workspace, workspace_bytes, st);
                const double e = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (rep > 0 || !s->o.graph || e > 50.0) ms = std::min(ms, e);
                if (rep == 0 && e > 50.0) break;
            }
            if (s->tuned_n < 8) s->tuned_ms[s->tuned_n++] = ms;
            if (r == CCC_OK && ms < best_ms) {
                best_ms = ms;
                best = TunedCfg{grids[gi], mortons[mi]};
            }
        }
    }
    (void)reset();
    (void)hipStreamSynchronize(st);
    dfree(x1_start);
    drop_graphs(s);
    s->o.ccd_blocks_per_cu = best.ccd_blocks_per_cu;
    s->o.morton = best.morton;
    if (r != CCC_OK) return r;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned[{s->ctx, key}] = best;
    return CCC_OK;
}

ccc_status ccc_sim_get_opts(const ccc_sim *s, ccc_sim_opts *o) {
    if (!s || !o) return bad_arg("NULL argument");
    *o = s->o;
    return CCC_OK;
}

int32_t ccc_sim_tuning_times(const ccc_sim *s, double *ms, int32_t max) {
    if (!s) return 0;
    const int n = std::min<int>(s->tuned_n, max);
    for (int i = 0; i < n; ++i) ms[i] = s->tuned_ms[i];
    return s->tuned_n;
}

ccc_status ccc_sim_state(const ccc_sim *s, void **x0, void **x1, int64_t *num_vertices, int64_t *candidate_capacity) {
    if (!s) return bad_arg("simulation is NULL");
    if (x0) *x0 = s->x0;
    if (x1) *x1 = s->x1;
    if (num_vertices) *num_vertices = s->nv;
    if (candidate_capacity) *candidate_capacity = s->cap_pairs;
    return CCC_OK;
}

ccc_status ccc_sim_impulses(const ccc_sim *s, void **dv, void **dx, int32_t **cnt) {
    if (!s) return bad_arg("simulation is NULL");
    if (dv) *dv = s->dv;
    if (dx) *dx = s->dx;
    if (cnt) *cnt = s->cnt;
    return CCC_OK;
}

int32_t ccc_sim_step_counts(const ccc_sim *s, ccc_counts *per_pass, int32_t max_passes) {
    if (!s) return 0;
    const int n = std::min<int>((int)s->counts.size(), max_passes);
    for (int i = 0; i < n; ++i) per_pass[i] = s->counts[(size_t)i];
    return (int32_t)s->counts.size();
}

size_t ccc_sim_footprint(const ccc_sim *s) {
    if (!s) return 0;
    const size_t vb = s->esz * 3 * (size_t)s->nv, sb = s->esz * 9 * (size_t)s->nf, rows = 2 * (size_t)s->nf - 1;
    size_t b = 3 * vb + s->esz * 3 * (size_t)s->nf + 4 * sb + s->esz * 6 * rows + sizeof(int64_t) * 3 * (size_t)s->nf +
               sizeof(int64_t) * 4 * rows + sizeof(int32_t) * (3 * (size_t)s->nf + (size_t)s->nv + 1 + 3 * (size_t)s->nf) +
               sizeof(DevCounters) * (size_t)s->o.iters + sizeof(int2) * (size_t)s->cap_pairs;
    if (s->mesh_end) b += sizeof(int32_t) * (size_t)s->nf;
    if (s->det_acc) b += sizeof(long long) * kDetWordsPerCorner * 3 * (size_t)s->nf;
    return b + s->arena_bytes;
}

}  // extern "C"
