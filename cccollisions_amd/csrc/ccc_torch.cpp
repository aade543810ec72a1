// PyTorch-ROCm extension `cccollisions_amd._cccollisions`: the same 14 entry points,
// argument names, defaults, output shapes and dtypes as the reference bindings
// (src/cccollisions.cpp:102-468), plus compute_impulses_packed for meshes of different
// sizes and a few controls (capacity policy, counts, kernel timing).
// Each function checks its inputs, allocates the outputs and calls the C ABI
// (include/cccollisions_amd.h) on the current HIP stream of the input's device.  Errors
// raise RuntimeError (the reference exits the process); capacity overflows raise a warning.
#include <ATen/DeviceGuard.h>
#include <c10/hip/HIPCachingAllocator.h>
#include <c10/hip/HIPStream.h>
#include <torch/extension.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cccollisions_amd.h"

namespace {

// Reference check macros (src/cccollisions.cpp:98-100) with the same messages.
#define CHECK_CUDA(x) TORCH_CHECK(x.is_cuda(), #x " must be a CUDA tensor")
#define CHECK_CONTIGUOUS(x) TORCH_CHECK(x.is_contiguous(), #x " must be contiguous")
#define CHECK_INPUT(x) \
    CHECK_CUDA(x);     \
    CHECK_CONTIGUOUS(x)

void *torch_alloc(void *, size_t bytes, void *stream) {
    return c10::hip::HIPCachingAllocator::raw_alloc_with_stream(bytes, (hipStream_t)stream);
}
void torch_free(void *, void *ptr, void *) { c10::hip::HIPCachingAllocator::raw_delete(ptr); }

ccc_context *context_for(const at::Tensor &t) {
    static std::mutex mu;
    static std::vector<ccc_context *> ctxs;
    const int dev = t.device().index();
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ctxs.size() <= dev) ctxs.resize(dev + 1, nullptr);
    if (!ctxs[dev]) {
        ccc_context *c = nullptr;
        ccc_status st = ccc_context_create(&c, dev, torch_alloc, torch_free, nullptr);
        TORCH_CHECK(st == CCC_OK, "cccollisions: context creation failed: ", ccc_status_string(st), ": ", ccc_last_error());
        ctxs[dev] = c;
    }
    return ctxs[dev];
}

void *stream_for(const at::Tensor &t) { return (void *)c10::hip::getCurrentHIPStream(t.device().index()).stream(); }

int32_t dtype_of(const at::Tensor &t) {
    if (t.scalar_type() == at::kDouble) return CCC_FLOAT64;
    if (t.scalar_type() == at::kFloat) return CCC_FLOAT32;
    TORCH_CHECK(false, "cccollisions: triangles must be float32 or float64, got ", t.scalar_type());
    return -1;
}

void check_like(const at::Tensor &x, const at::Tensor &ref, const char *name) {
    TORCH_CHECK(x.scalar_type() == ref.scalar_type(), name, " must have dtype ", ref.scalar_type(), ", got ", x.scalar_type());
    TORCH_CHECK(x.device() == ref.device(), name, " must be on ", ref.device());
}

void check_tree(const at::Tensor &bboxes, const at::Tensor &tree, const at::Tensor &triangles) {
    TORCH_CHECK(bboxes.dim() == 3 && bboxes.size(2) == 6, "bboxes must have shape [B, 2N-1, 6]");
    TORCH_CHECK(tree.dim() == 3 && tree.size(2) == 4 && tree.size(1) == bboxes.size(1) && tree.size(0) == bboxes.size(0),
                "tree must have shape [B, 2N-1, 4]");
    TORCH_CHECK(tree.scalar_type() == at::kLong, "tree must be int64");
    TORCH_CHECK(bboxes.size(1) % 2 == 1, "bboxes must hold 2N-1 nodes");
    check_like(bboxes, triangles, "bboxes");
    TORCH_CHECK(tree.device() == triangles.device(), "tree must be on ", triangles.device());
}

void check_soup(const at::Tensor &t, int64_t B, int64_t n, const char *name) {
    TORCH_CHECK(t.dim() == 4 && t.size(0) == B && t.size(1) == n && t.size(2) == 3 && t.size(3) == 3, name,
                " must have shape [", B, ", ", n, ", 3, 3]");
}

thread_local uint32_t g_flags = CCC_FLAG_NONE;
thread_local std::vector<ccc_counts> g_counts;

// Overflow warnings become Python UserWarnings (the reference printed them to stdout).
void warn(const char *msg) {
    if (PyErr_WarnEx(PyExc_UserWarning, msg, 1) != 0) throw py::error_already_set();
}

void finish(ccc_status st, const char *what) {
    TORCH_CHECK((st & CCC_ERR_MASK) == CCC_OK, "cccollisions.", what, " failed: ", ccc_status_string(st), ": ", ccc_last_error());
    if (st & CCC_WARN_CANDIDATES) {
        if (g_flags & CCC_FLAG_FIXED_CAPACITY)
            warn("Number of candidate collisions exceeds maximal allowed number. Some of the candidate collisions are omitted.");
        else
            warn("Number of candidate collisions exceeds max_candidates_per_triangle * N; the candidate buffer was grown to hold all of them.");
    }
    if (st & CCC_WARN_TRUNCATED)
        warn("Number of actual collisions exceeds maximal allowed number. Some of the collisions are omitted.");
}

// The native calls run without the GIL (they touch no Python object), so that host threads
// driving different HIP streams -- a serving loop's worker slots -- overlap on the device.
#define CCC_NOGIL(expr) ([&]() { py::gil_scoped_release nogil_; return (expr); }())

ccc_counts *counts_buffer(int64_t B) {
    g_counts.assign((size_t)B, ccc_counts{0, 0, 0, 0});
    return g_counts.data();
}

// ---- entry points (src/cccollisions.cpp) ------------------------------------------------

std::vector<torch::Tensor> bvh_forward(at::Tensor triangles) {
    CHECK_INPUT(triangles);
    TORCH_CHECK(triangles.dim() == 4 && triangles.size(2) == 3 && triangles.size(3) == 3, "triangles must have shape [B, N, 3, 3]");
    const at::DeviceGuard guard(triangles.device());
    const int64_t B = triangles.size(0), n = triangles.size(1);
    TORCH_CHECK(n >= 1, "triangles must hold at least one triangle");
    auto bboxes = at::empty({B, 2 * n - 1, 6}, triangles.options());
    auto tree = at::empty({B, 2 * n - 1, 4}, triangles.options().dtype(at::kLong));
    finish(CCC_NOGIL(ccc_bvh(context_for(triangles), triangles.data_ptr(), B, n, dtype_of(triangles), bboxes.data_ptr(),
                   tree.data_ptr<int64_t>(), stream_for(triangles))),
           "bvh");
    return {bboxes, tree};
}

std::vector<torch::Tensor> bvh_forward_motion(at::Tensor triangles, at::Tensor triangles_next) {
    CHECK_INPUT(triangles);
    CHECK_INPUT(triangles_next);
    TORCH_CHECK(triangles.dim() == 4 && triangles.size(2) == 3 && triangles.size(3) == 3, "triangles must have shape [B, N, 3, 3]");
    const at::DeviceGuard guard(triangles.device());
    const int64_t B = triangles.size(0), n = triangles.size(1);
    TORCH_CHECK(n >= 1, "triangles must hold at least one triangle");
    check_soup(triangles_next, B, n, "triangles_next");
    check_like(triangles_next, triangles, "triangles_next");
    auto bboxes = at::empty({B, 2 * n - 1, 6}, triangles.options());
    auto tree = at::empty({B, 2 * n - 1, 4}, triangles.options().dtype(at::kLong));
    finish(CCC_NOGIL(ccc_bvh_motion(context_for(triangles), triangles.data_ptr(), triangles_next.data_ptr(), B, n, dtype_of(triangles),
                          bboxes.data_ptr(), tree.data_ptr<int64_t>(), stream_for(triangles))),
           "bvh_motion");
    return {bboxes, tree};
}

at::Tensor find_collisions(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles, float threshold = 0.,
                           int max_collisions = 16) {
    CHECK_INPUT(bboxes);
    CHECK_INPUT(tree);
    CHECK_INPUT(triangles);
    check_tree(bboxes, tree, triangles);
    const at::DeviceGuard guard(triangles.device());
    const int64_t B = bboxes.size(0), n = (bboxes.size(1) + 1) / 2;
    check_soup(triangles, B, n, "triangles");
    TORCH_CHECK(max_collisions >= 1, "max_collisions must be >= 1");
    auto out = at::empty({B, n * max_collisions, 2}, bboxes.options().dtype(at::kLong));
    finish(CCC_NOGIL(ccc_find_collisions(context_for(triangles), bboxes.data_ptr(), tree.data_ptr<int64_t>(), triangles.data_ptr(), B, n,
                               dtype_of(triangles), threshold, max_collisions, out.data_ptr<int64_t>(), counts_buffer(B),
                               g_flags, stream_for(triangles))),
           "find_collisions");
    return out;
}

// Triangle-edge queries (src/cccollisions.cpp:151-218).  `threshold` is accepted and unused,
// as in the reference: traversal and SAT run at threshold 0.
struct TeArgs {
    int64_t B, n;
};

static TeArgs te_check(const at::Tensor &bboxes, const at::Tensor &tree, const at::Tensor &triangles, int max_collisions) {
    CHECK_INPUT(bboxes);
    CHECK_INPUT(tree);
    CHECK_INPUT(triangles);
    check_tree(bboxes, tree, triangles);
    const int64_t B = bboxes.size(0), n = (bboxes.size(1) + 1) / 2;
    check_soup(triangles, B, n, "triangles");
    TORCH_CHECK(max_collisions >= 1, "max_collisions must be >= 1");
    return {B, n};
}

std::vector<torch::Tensor> find_triangle_edge_collisions(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles,
                                                         float threshold = 0., int max_collisions = 16) {
    (void)threshold;
    const TeArgs a = te_check(bboxes, tree, triangles, max_collisions);
    const at::DeviceGuard guard(triangles.device());
    auto out = at::empty({a.B, a.n * max_collisions, 4}, bboxes.options().dtype(at::kLong));
    auto coefs = at::empty({a.B, a.n * max_collisions, 1}, bboxes.options());
    finish(CCC_NOGIL(ccc_find_triangle_edge_collisions(context_for(triangles), bboxes.data_ptr(), tree.data_ptr<int64_t>(),
                                             triangles.data_ptr(), a.B, a.n, dtype_of(triangles), max_collisions,
                                             out.data_ptr<int64_t>(), coefs.data_ptr(), counts_buffer(a.B), g_flags,
                                             stream_for(triangles))),
           "find_triangle_edge_collisions");
    return {out, coefs};
}

torch::Tensor find_triangle_edge_candidates(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles, float threshold = 0.,
                                            int max_collisions = 16) {
    (void)threshold;
    const TeArgs a = te_check(bboxes, tree, triangles, max_collisions);
    const at::DeviceGuard guard(triangles.device());
    auto out = at::empty({a.B, a.n * max_collisions, 4}, bboxes.options().dtype(at::kLong));
    finish(CCC_NOGIL(ccc_find_triangle_edge_candidates(context_for(triangles), bboxes.data_ptr(), tree.data_ptr<int64_t>(),
                                             triangles.data_ptr(), a.B, a.n, dtype_of(triangles), max_collisions,
                                             out.data_ptr<int64_t>(), counts_buffer(a.B), g_flags, stream_for(triangles))),
           "find_triangle_edge_candidates");
    return out;
}

std::vector<torch::Tensor> find_triangle_edge_candidates2(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles,
                                                          float threshold = 0., int max_collisions = 16) {
    (void)threshold;
    const TeArgs a = te_check(bboxes, tree, triangles, max_collisions);
    const at::DeviceGuard guard(triangles.device());
    auto pairs = at::empty({a.B, a.n * max_collisions, 2}, bboxes.options().dtype(at::kLong));
    auto masks = at::empty({a.B, a.n * max_collisions, 2}, bboxes.options().dtype(at::kInt));
    finish(CCC_NOGIL(ccc_find_triangle_edge_candidates2(context_for(triangles), bboxes.data_ptr(), tree.data_ptr<int64_t>(),
                                              triangles.data_ptr(), a.B, a.n, dtype_of(triangles), max_collisions,
                                              pairs.data_ptr<int64_t>(), masks.data_ptr<int32_t>(), counts_buffer(a.B),
                                              g_flags, stream_for(triangles))),
           "find_triangle_edge_candidates2");
    return {pairs, masks};
}

std::vector<torch::Tensor> continuous(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles, at::Tensor triangles_next,
                                      const at::Tensor *to_check, int max_candidates_per_triangle,
                                      int max_collisions_per_triangle, const char *name) {
    CHECK_INPUT(bboxes);
    CHECK_INPUT(tree);
    CHECK_INPUT(triangles);
    CHECK_INPUT(triangles_next);
    check_tree(bboxes, tree, triangles);
    const at::DeviceGuard guard(triangles.device());
    const int64_t B = bboxes.size(0), n = (bboxes.size(1) + 1) / 2;
    check_soup(triangles, B, n, "triangles");
    check_soup(triangles_next, B, n, "triangles_next");
    check_like(triangles_next, triangles, "triangles_next");
    TORCH_CHECK(max_candidates_per_triangle >= 1 && max_collisions_per_triangle >= 1, "capacities must be >= 1");
    const uint8_t *tc = nullptr;
    if (to_check) {
        CHECK_INPUT((*to_check));
        TORCH_CHECK(to_check->scalar_type() == at::kBool && to_check->numel() == B * n, "triangles_to_check must be a [B, N] bool tensor");
        tc = (const uint8_t *)to_check->data_ptr<bool>();
    }
    auto coll = at::empty({B, n * max_collisions_per_triangle, 3}, bboxes.options().dtype(at::kLong));
    auto roots = at::empty({B, n * max_collisions_per_triangle, 1}, bboxes.options());
    ccc_status st = CCC_NOGIL(tc ? ccc_find_collisions_continuous_partial(
                             context_for(triangles), bboxes.data_ptr(), tree.data_ptr<int64_t>(), triangles.data_ptr(),
                             triangles_next.data_ptr(), tc, B, n, dtype_of(triangles), max_candidates_per_triangle,
                             max_collisions_per_triangle, coll.data_ptr<int64_t>(), roots.data_ptr(), counts_buffer(B), g_flags,
                             stream_for(triangles))
                       : ccc_find_collisions_continuous(context_for(triangles), bboxes.data_ptr(), tree.data_ptr<int64_t>(),
                                                        triangles.data_ptr(), triangles_next.data_ptr(), B, n, dtype_of(triangles),
                                                        max_candidates_per_triangle, max_collisions_per_triangle,
                                                        coll.data_ptr<int64_t>(), roots.data_ptr(), counts_buffer(B), g_flags,
                                                        stream_for(triangles)));
    finish(st, name);
    return {coll, roots};
}

std::vector<torch::Tensor> find_collisions_continuous(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles,
                                                      at::Tensor triangles_next, int max_candidates_per_triangle = 16,
                                                      int max_collisions_per_triangle = 16) {
    return continuous(bboxes, tree, triangles, triangles_next, nullptr, max_candidates_per_triangle, max_collisions_per_triangle,
                      "find_collisions_continuous");
}

std::vector<torch::Tensor> find_collisions_continuous_partial(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles,
                                                              at::Tensor triangles_next, at::Tensor triangles_to_check,
                                                              int max_candidates_per_triangle = 16,
                                                              int max_collisions_per_triangle = 16) {
    return continuous(bboxes, tree, triangles, triangles_next, &triangles_to_check, max_candidates_per_triangle,
                      max_collisions_per_triangle, "find_collisions_continuous_partial");
}

std::vector<torch::Tensor> find_proximity(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles, float threshold,
                                          int max_candidates_per_triangle = 16) {
    CHECK_INPUT(bboxes);
    CHECK_INPUT(tree);
    CHECK_INPUT(triangles);
    check_tree(bboxes, tree, triangles);
    const at::DeviceGuard guard(triangles.device());
    const int64_t B = bboxes.size(0), n = (bboxes.size(1) + 1) / 2;
    check_soup(triangles, B, n, "triangles");
    TORCH_CHECK(max_candidates_per_triangle >= 1, "max_candidates_per_triangle must be >= 1");
    auto out = at::empty({B, n * max_candidates_per_triangle, 3}, bboxes.options().dtype(at::kLong));
    finish(CCC_NOGIL(ccc_find_proximity(context_for(triangles), bboxes.data_ptr(), tree.data_ptr<int64_t>(), triangles.data_ptr(), B, n,
                              dtype_of(triangles), threshold, max_candidates_per_triangle, out.data_ptr<int64_t>(),
                              counts_buffer(B), g_flags, stream_for(triangles))),
           "find_proximity");
    return {out};
}

torch::Tensor find_mincol_cand(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles, at::Tensor triangles_next,
                               int max_collisions = 16) {
    CHECK_INPUT(bboxes);
    CHECK_INPUT(tree);
    CHECK_INPUT(triangles);
    CHECK_INPUT(triangles_next);
    check_tree(bboxes, tree, triangles);
    const at::DeviceGuard guard(triangles.device());
    const int64_t B = bboxes.size(0), n = (bboxes.size(1) + 1) / 2;
    check_soup(triangles, B, n, "triangles");
    check_soup(triangles_next, B, n, "triangles_next");
    check_like(triangles_next, triangles, "triangles_next");
    TORCH_CHECK(max_collisions >= 1, "max_collisions must be >= 1");
    auto out = at::empty({B, n * max_collisions * 15, 3}, bboxes.options().dtype(at::kLong));
    finish(CCC_NOGIL(ccc_find_mincol_cand(context_for(triangles), bboxes.data_ptr(), tree.data_ptr<int64_t>(), triangles.data_ptr(),
                                triangles_next.data_ptr(), B, n, dtype_of(triangles), max_collisions, out.data_ptr<int64_t>(),
                                counts_buffer(B), g_flags, stream_for(triangles))),
           "find_mincol_cand");
    return out;
}

std::vector<torch::Tensor> impulses(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles, at::Tensor triangles_next,
                                    at::Tensor triangles_masses, const at::Tensor *to_check, int max_candidates_per_triangle,
                                    int max_collisions_per_triangle, const char *name) {
    CHECK_INPUT(bboxes);
    CHECK_INPUT(tree);
    CHECK_INPUT(triangles);
    CHECK_INPUT(triangles_next);
    CHECK_INPUT(triangles_masses);
    check_tree(bboxes, tree, triangles);
    const at::DeviceGuard guard(triangles.device());
    const int64_t B = bboxes.size(0), n = (bboxes.size(1) + 1) / 2;
    check_soup(triangles, B, n, "triangles");
    check_soup(triangles_next, B, n, "triangles_next");
    check_like(triangles_next, triangles, "triangles_next");
    check_like(triangles_masses, triangles, "triangles_mass");
    TORCH_CHECK(triangles_masses.numel() == B * n * 3, "triangles_mass must have shape [B, N, 3]");
    TORCH_CHECK(max_candidates_per_triangle >= 1, "max_candidates_per_triangle must be >= 1");
    const uint8_t *tc = nullptr;
    if (to_check) {
        CHECK_INPUT((*to_check));
        TORCH_CHECK(to_check->scalar_type() == at::kBool && to_check->numel() == B * n, "triangles_to_check must be a [B, N] bool tensor");
        tc = (const uint8_t *)to_check->data_ptr<bool>();
    }
    auto dv = at::empty({B, n, 3, 3}, bboxes.options());
    auto dx = at::empty({B, n, 3, 3}, bboxes.options());
    auto cnt = at::empty({B, n, 3}, bboxes.options().dtype(at::kInt));
    ccc_status st = CCC_NOGIL(tc ? ccc_compute_impulses_partial(context_for(triangles), bboxes.data_ptr(), tree.data_ptr<int64_t>(),
                                                      triangles.data_ptr(), triangles_next.data_ptr(), triangles_masses.data_ptr(),
                                                      tc, B, n, dtype_of(triangles), max_candidates_per_triangle,
                                                      max_collisions_per_triangle, dv.data_ptr(), dx.data_ptr(),
                                                      cnt.data_ptr<int32_t>(), counts_buffer(B), g_flags, stream_for(triangles))
                       : ccc_compute_impulses(context_for(triangles), bboxes.data_ptr(), tree.data_ptr<int64_t>(),
                                              triangles.data_ptr(), triangles_next.data_ptr(), triangles_masses.data_ptr(), B, n,
                                              dtype_of(triangles), max_candidates_per_triangle, max_collisions_per_triangle,
                                              dv.data_ptr(), dx.data_ptr(), cnt.data_ptr<int32_t>(), counts_buffer(B), g_flags,
                                              stream_for(triangles)));
    finish(st, name);
    return {dv, dx, cnt};
}

// Meshes of different sizes back to back ([N, 3, 3] soups, [N, 3] masses, offsets[m] = first
// triangle of mesh m, offsets[-1] = N): every mesh's compute_impulses after its own
// bvh_motion, in one pass of launches.
std::vector<torch::Tensor> compute_impulses_packed(at::Tensor triangles, at::Tensor triangles_next, at::Tensor triangles_mass,
                                                   std::vector<int64_t> offsets, int max_candidates_per_triangle = 16) {
    CHECK_INPUT(triangles);
    CHECK_INPUT(triangles_next);
    CHECK_INPUT(triangles_mass);
    TORCH_CHECK(triangles.dim() == 3 && triangles.size(1) == 3 && triangles.size(2) == 3, "triangles must have shape [N, 3, 3]");
    TORCH_CHECK(triangles.scalar_type() == at::kDouble || triangles.scalar_type() == at::kFloat,
                "triangles must be float32 or float64");
    check_like(triangles_next, triangles, "triangles_next");
    check_like(triangles_mass, triangles, "triangles_mass");
    const int64_t n = triangles.size(0);
    TORCH_CHECK(triangles_next.sizes() == triangles.sizes(), "triangles_next must have the shape of triangles");
    TORCH_CHECK(triangles_mass.numel() == n * 3, "triangles_mass must have shape [N, 3]");
    TORCH_CHECK(offsets.size() >= 2 && offsets.front() == 0 && offsets.back() == n,
                "offsets must start at 0 and end at the number of triangles");
    TORCH_CHECK(max_candidates_per_triangle >= 1, "max_candidates_per_triangle must be >= 1");
    const at::DeviceGuard guard(triangles.device());
    auto dv = at::empty({n, 3, 3}, triangles.options());
    auto dx = at::empty({n, 3, 3}, triangles.options());
    auto cnt = at::empty({n, 3}, triangles.options().dtype(at::kInt));
    finish(CCC_NOGIL(ccc_compute_impulses_packed(context_for(triangles), triangles.data_ptr(), triangles_next.data_ptr(),
                                       triangles_mass.data_ptr(), offsets.data(), (int64_t)offsets.size() - 1,
                                       dtype_of(triangles), max_candidates_per_triangle, dv.data_ptr(), dx.data_ptr(),
                                       cnt.data_ptr<int32_t>(), counts_buffer(1), g_flags, stream_for(triangles))),
           "compute_impulses_packed");
    return {dv, dx, cnt};
}

std::vector<torch::Tensor> compute_impulses(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles, at::Tensor triangles_next,
                                            at::Tensor triangles_masses, int max_candidates_per_triangle = 16,
                                            int max_collisions_per_triangle = 16) {
    return impulses(bboxes, tree, triangles, triangles_next, triangles_masses, nullptr, max_candidates_per_triangle,
                    max_collisions_per_triangle, "compute_impulses");
}

std::vector<torch::Tensor> compute_impulses_partial(at::Tensor bboxes, at::Tensor tree, at::Tensor triangles,
                                                    at::Tensor triangles_next, at::Tensor triangles_masses,
                                                    at::Tensor triangles_to_check, int max_candidates_per_triangle = 16,
                                                    int max_collisions_per_triangle = 16) {
    return impulses(bboxes, tree, triangles, triangles_next, triangles_masses, &triangles_to_check, max_candidates_per_triangle,
                    max_collisions_per_triangle, "compute_impulses_partial");
}

std::vector<torch::Tensor> compute_impulses_from_collisions(at::Tensor collisions, at::Tensor roots, at::Tensor triangles,
                                                            at::Tensor triangles_next, at::Tensor triangles_masses) {
    CHECK_INPUT(collisions);
    CHECK_INPUT(roots);
    CHECK_INPUT(triangles);
    CHECK_INPUT(triangles_next);
    CHECK_INPUT(triangles_masses);
    TORCH_CHECK(triangles.dim() == 4 && triangles.size(2) == 3 && triangles.size(3) == 3, "triangles must have shape [B, N, 3, 3]");
    const at::DeviceGuard guard(triangles.device());
    const int64_t B = triangles.size(0), n = triangles.size(1);
    check_soup(triangles_next, B, n, "triangles_next");
    check_like(triangles_next, triangles, "triangles_next");
    check_like(triangles_masses, triangles, "triangles_mass");
    TORCH_CHECK(triangles_masses.numel() == B * n * 3, "triangles_mass must have shape [B, N, 3]");
    TORCH_CHECK(collisions.scalar_type() == at::kLong && collisions.dim() == 3 && collisions.size(0) == B && collisions.size(2) == 3,
                "collisions must be an int64 [B, M, 3] tensor");
    const int64_t m = collisions.size(1);
    check_like(roots, triangles, "roots");
    TORCH_CHECK(roots.numel() == B * m, "roots must have shape [B, M, 1]");
    auto dv = at::empty({B, n, 3, 3}, triangles.options());
    auto dx = at::empty({B, n, 3, 3}, triangles.options());
    auto cnt = at::empty({B, n, 3}, triangles.options().dtype(at::kInt));
    finish(CCC_NOGIL(ccc_compute_impulses_from_collisions(context_for(triangles), collisions.data_ptr<int64_t>(), roots.data_ptr(), m,
                                                triangles.data_ptr(), triangles_next.data_ptr(), triangles_masses.data_ptr(), B,
                                                n, dtype_of(triangles), dv.data_ptr(), dx.data_ptr(), cnt.data_ptr<int32_t>(),
                                                g_flags, stream_for(triangles))),
           "compute_impulses_from_collisions");
    return {dv, dx, cnt};
}

// ---- extras (not in the reference) -------------------------------------------------------
void set_fixed_capacity(bool on) { g_flags = on ? (g_flags | CCC_FLAG_FIXED_CAPACITY) : (g_flags & ~CCC_FLAG_FIXED_CAPACITY); }

// Pairs per narrow-phase chunk (0: default) of the current device's context.
void set_pair_chunk(int64_t pairs) {
    const int dev = c10::hip::current_device();
    auto t = at::empty({0}, at::TensorOptions().device(at::Device(at::kCUDA, dev)));
    ccc_status st = ccc_context_set_pair_chunk(context_for(t), pairs);
    TORCH_CHECK(st == CCC_OK, "set_pair_chunk failed: ", ccc_last_error());
}

// Morton mode of the trees bvh / bvh_motion build on the current device (0: the reference's,
// 1: fitted scene box with 30-bit codes, 2: fitted with 63-bit codes).
void set_morton_mode(int mode) {
    const int dev = c10::hip::current_device();
    auto t = at::empty({0}, at::TensorOptions().device(at::Device(at::kCUDA, dev)));
    ccc_status st = ccc_context_set_morton(context_for(t), mode);
    TORCH_CHECK(st == CCC_OK, "set_morton_mode failed: ", ccc_last_error());
}

// Opt-in per-kernel timing with HIP events (device of the current stream).
void set_profiling(bool on) {
    const int dev = c10::hip::current_device();
    auto t = at::empty({0}, at::TensorOptions().device(at::Device(at::kCUDA, dev)));
    ccc_status st = ccc_context_set_profiling(context_for(t), on ? 1 : 0);
    TORCH_CHECK(st == CCC_OK, "set_profiling failed");
}

std::map<std::string, std::pair<double, int64_t>> kernel_times() {
    const int dev = c10::hip::current_device();
    auto t = at::empty({0}, at::TensorOptions().device(at::Device(at::kCUDA, dev)));
    ccc_context *c = context_for(t);
    std::vector<char> buf((size_t)std::max(ccc_context_kernel_names(c, nullptr, 0), 0) + 1);
    ccc_context_kernel_names(c, buf.data(), buf.size());
    std::map<std::string, std::pair<double, int64_t>> out;
    std::string names(buf.data());
    size_t pos = 0;
    while (pos < names.size()) {
        size_t e = names.find(',', pos);
        if (e == std::string::npos) e = names.size();
        std::string k = names.substr(pos, e - pos);
        double ms = 0;
        int64_t n = 0;
        ccc_context_kernel_time(c, k.c_str(), &ms, &n);
        out[k] = {ms, n};
        pos = e + 1;
    }
    return out;
}

std::vector<std::vector<int64_t>> last_counts() {
    std::vector<std::vector<int64_t>> out;
    for (const auto &c : g_counts) out.push_back({c.candidates, c.subprims, c.roots, c.collisions});
    return out;
}

}  // namespace

PYBIND11_MODULE(_cccollisions, m) {
    m.doc() = "MI355X-native CCCollisions entry points (drop-in for the `cccollisions` extension)";
    m.def("bvh", &bvh_forward, "BVH over static triangles", py::arg("triangles"));
    m.def("bvh_motion", &bvh_forward_motion, "BVH over swept triangles", py::arg("triangles"), py::arg("triangles_next"));
    m.def("find_collisions", &find_collisions, "", py::arg("bboxes"), py::arg("tree"), py::arg("triangles"),
          py::arg("threshold") = 0., py::arg("max_collisions") = 16);
    m.def("find_triangle_edge_collisions", &find_triangle_edge_collisions, "", py::arg("bboxes"), py::arg("tree"),
          py::arg("triangles"), py::arg("threshold") = 0., py::arg("max_collisions") = 16);
    m.def("find_triangle_edge_candidates", &find_triangle_edge_candidates, "", py::arg("bboxes"), py::arg("tree"),
          py::arg("triangles"), py::arg("threshold") = 0., py::arg("max_collisions") = 16);
    m.def("find_triangle_edge_candidates2", &find_triangle_edge_candidates2, "", py::arg("bboxes"), py::arg("tree"),
          py::arg("triangles"), py::arg("threshold") = 0., py::arg("max_collisions") = 16);
    m.def("compute_impulses_packed", &compute_impulses_packed,
          "compute_impulses for meshes of different sizes stored back to back (one BVH and one traversal over all)",
          py::arg("triangles"), py::arg("triangles_next"), py::arg("triangles_mass"), py::arg("offsets"),
          py::arg("max_candidates_per_triangle") = 16);
    m.def("find_collisions_continuous", &find_collisions_continuous, "", py::arg("bboxes"), py::arg("tree"), py::arg("triangles"),
          py::arg("triangles_next"), py::arg("max_candidates_per_triangle") = 16, py::arg("max_collisions_per_triangle") = 16);
    m.def("find_collisions_continuous_partial", &find_collisions_continuous_partial, "", py::arg("bboxes"), py::arg("tree"),
          py::arg("triangles"), py::arg("triangles_next"), py::arg("triangles_to_check"),
          py::arg("max_candidates_per_triangle") = 16, py::arg("max_collisions_per_triangle") = 16);
    m.def("find_proximity", &find_proximity, "", py::arg("bboxes"), py::arg("tree"), py::arg("triangles"), py::arg("threshold"),
          py::arg("max_candidates_per_triangle") = 16);
    m.def("find_mincol_cand", &find_mincol_cand, "", py::arg("bboxes"), py::arg("tree"), py::arg("triangles"),
          py::arg("triangles_next"), py::arg("max_collisions") = 16);
    m.def("compute_impulses", &compute_impulses, "", py::arg("bboxes"), py::arg("tree"), py::arg("triangles"),
          py::arg("triangles_next"), py::arg("triangles_mass"), py::arg("max_candidates_per_triangle") = 16,
          py::arg("max_collisions_per_triangle") = 16);
    m.def("compute_impulses_partial", &compute_impulses_partial, "", py::arg("bboxes"), py::arg("tree"), py::arg("triangles"),
          py::arg("triangles_next"), py::arg("triangles_mass"), py::arg("triangles_to_check"),
          py::arg("max_candidates_per_triangle") = 16, py::arg("max_collisions_per_triangle") = 16);
    m.def("compute_impulses_from_collisions", &compute_impulses_from_collisions, "", py::arg("collisions"), py::arg("roots"),
          py::arg("triangles"), py::arg("triangles_next"), py::arg("triangles_mass"));
    m.def("set_fixed_capacity", &set_fixed_capacity,
          "Keep the reference candidate capacity (max_candidates_per_triangle * N) instead of growing it", py::arg("on"));
    m.def("set_pair_chunk", &set_pair_chunk,
          "Candidate pairs per narrow-phase chunk over a call's batch (0: default 2^25); results do not depend on it",
          py::arg("pairs"));
    m.def("set_morton_mode", &set_morton_mode,
          "Morton codes of the trees bvh / bvh_motion build: 0 the reference's (default), 1 fitted scene box with "
          "30-bit codes, 2 fitted with 63-bit codes; queries give the same sets on every mode",
          py::arg("mode"));
    m.def("set_profiling", &set_profiling, "Time every kernel launch with HIP events (clears totals)", py::arg("on"));
    m.def("kernel_times", &kernel_times, "Per-kernel {name: (total_ms, launches)} since profiling was enabled");
    m.def("last_counts", &last_counts, "Per-batch [K, M, V, R] stage counts of the last query on this thread");
}
