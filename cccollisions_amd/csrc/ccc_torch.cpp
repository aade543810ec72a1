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

// ---- resident simulations ------------------------------------------------------------------
// Sim(x0, x1, faces, masses, ...) copies one simulation (or several meshes packed back to
// back) onto the device; step() runs `iters` impulse passes on the current stream, in place
// on the resident x1, as one replayed HIP graph.
class PySim {
  public:
    PySim(at::Tensor x0, at::Tensor x1, at::Tensor faces, at::Tensor masses, std::vector<int64_t> mesh_offsets, int iters,
          int max_candidates, bool deterministic, bool graph, int morton, int ccd_blocks_per_cu, int64_t pair_chunk) {
        CHECK_INPUT(x0);
        CHECK_INPUT(x1);
        CHECK_INPUT(faces);
        CHECK_INPUT(masses);
        TORCH_CHECK(x0.dim() == 2 && x0.size(1) == 3, "x0 must have shape [V, 3]");
        TORCH_CHECK(x1.sizes() == x0.sizes(), "x1 must have the shape of x0");
        TORCH_CHECK(faces.dim() == 2 && faces.size(1) == 3 && faces.scalar_type() == at::kLong, "faces must be int64 [F, 3]");
        TORCH_CHECK(masses.dim() == 2 && masses.size(0) == faces.size(0) && masses.size(1) == 3, "masses must have shape [F, 3]");
        check_like(x1, x0, "x1");
        check_like(masses, x0, "masses");
        TORCH_CHECK(faces.device() == x0.device(), "faces must be on ", x0.device());
        const at::DeviceGuard guard(x0.device());
        opts_ = x0.options();
        nv_ = x0.size(0);
        nf_ = faces.size(0);
        ccc_sim_opts o;
        ccc_sim_default_opts(&o);
        o.iters = iters;
        o.max_candidates = max_candidates;
        o.deterministic = deterministic ? 1 : 0;
        o.graph = graph ? 1 : 0;
        o.morton = morton;
        o.ccd_blocks_per_cu = ccd_blocks_per_cu;
        o.pair_chunk = pair_chunk;
        ccc_status st = CCC_NOGIL(ccc_sim_create(context_for(x0), x0.data_ptr(), x1.data_ptr(), nv_, faces.data_ptr<int64_t>(),
                                                 masses.data_ptr(), faces.size(0), mesh_offsets.empty() ? nullptr : mesh_offsets.data(),
                                                 mesh_offsets.empty() ? 1 : (int64_t)mesh_offsets.size() - 1, dtype_of(x0), &o,
                                                 stream_for(x0), &sim_));
        TORCH_CHECK(st == CCC_OK, "Sim: creation failed: ", ccc_status_string(st), ": ", ccc_last_error());
    }
    ~PySim() {
        if (sim_) ccc_sim_destroy(sim_);
    }
    void step(bool advance, c10::optional<at::Tensor> workspace) {
        const at::DeviceGuard guard(opts_.device());
        void *ws = nullptr;
        size_t bytes = 0;
        if (workspace && workspace->defined()) {
            TORCH_CHECK(workspace->is_cuda() && workspace->is_contiguous(), "workspace must be a contiguous device tensor");
            ws = workspace->data_ptr();
            bytes = (size_t)workspace->nbytes();
        }
        auto stream = (void *)c10::hip::getCurrentHIPStream(opts_.device().index()).stream();
        ccc_status st = CCC_NOGIL(ccc_sim_step(sim_, ws, bytes, advance ? 1 : 0, stream));
        TORCH_CHECK(st == CCC_OK, "Sim.step failed: ", ccc_status_string(st), ": ", ccc_last_error());
    }
    void autotune(c10::optional<at::Tensor> workspace) {
        const at::DeviceGuard guard(opts_.device());
        void *ws = nullptr;
        size_t bytes = 0;
        if (workspace && workspace->defined()) {
            ws = workspace->data_ptr();
            bytes = (size_t)workspace->nbytes();
        }
        auto stream = (void *)c10::hip::getCurrentHIPStream(opts_.device().index()).stream();
        ccc_status st = CCC_NOGIL(ccc_sim_autotune(sim_, ws, bytes, stream));
        TORCH_CHECK(st == CCC_OK, "Sim.autotune failed: ", ccc_status_string(st), ": ", ccc_last_error());
    }
    // Copies of the resident positions (x0, x1), [V, 3].
    std::vector<at::Tensor> positions() {
        const at::DeviceGuard guard(opts_.device());
        void *x0 = nullptr, *x1 = nullptr;
        ccc_sim_state(sim_, &x0, &x1, nullptr, nullptr);
        auto a = at::from_blob(x0, {nv_, 3}, opts_).clone();
        auto b = at::from_blob(x1, {nv_, 3}, opts_).clone();
        return {a, b};
    }
    // Overwrites the resident positions (e.g. with the simulator's next prediction).
    void set_positions(at::Tensor x0, at::Tensor x1) {
        CHECK_INPUT(x0);
        CHECK_INPUT(x1);
        TORCH_CHECK(x0.scalar_type() == opts_.dtype().toScalarType() && x1.scalar_type() == x0.scalar_type(), "dtype mismatch");
        TORCH_CHECK(x0.numel() == nv_ * 3 && x1.numel() == nv_ * 3, "positions must have shape [V, 3]");
        const at::DeviceGuard guard(opts_.device());
        void *d0 = nullptr, *d1 = nullptr;
        ccc_sim_state(sim_, &d0, &d1, nullptr, nullptr);
        at::from_blob(d0, {nv_, 3}, opts_).copy_(x0.reshape({nv_, 3}));
        at::from_blob(d1, {nv_, 3}, opts_).copy_(x1.reshape({nv_, 3}));
    }
    // Copies of the last pass's impulse outputs (dv, dx [F, 3, 3], cnt [F, 3]).
    std::vector<at::Tensor> last_impulses() {
        const at::DeviceGuard guard(opts_.device());
        void *dv = nullptr, *dx = nullptr;
        int32_t *cnt = nullptr;
        ccc_sim_impulses(sim_, &dv, &dx, &cnt);
        const int64_t nf = nf_;
        return {at::from_blob(dv, {nf, 3, 3}, opts_).clone(), at::from_blob(dx, {nf, 3, 3}, opts_).clone(),
                at::from_blob(cnt, {nf, 3}, opts_.dtype(at::kInt)).clone()};
    }
    std::vector<std::vector<int64_t>> counts() const {
        std::vector<ccc_counts> c(64);
        const int n = ccc_sim_step_counts(sim_, c.data(), (int32_t)c.size());
        std::vector<std::vector<int64_t>> out;
        for (int i = 0; i < std::min(n, 64); ++i) out.push_back({c[i].candidates, c[i].subprims, c[i].roots, c[i].collisions});
        return out;
    }
    py::dict config() const {
        ccc_sim_opts o;
        ccc_sim_get_opts(sim_, &o);
        int64_t cap = 0;
        ccc_sim_state(sim_, nullptr, nullptr, nullptr, &cap);
        py::dict d;
        d["iters"] = o.iters;
        d["deterministic"] = (bool)o.deterministic;
        d["graph"] = (bool)o.graph;
        d["morton"] = o.morton;
        d["ccd_blocks_per_cu"] = o.ccd_blocks_per_cu;
        d["pair_chunk"] = o.pair_chunk;
        d["candidate_capacity"] = cap;
        return d;
    }
    std::vector<double> tuning_times() const {
        double ms[8];
        const int n = ccc_sim_tuning_times(sim_, ms, 8);
        return std::vector<double>(ms, ms + std::min(n, 8));
    }
    int64_t footprint() const { return (int64_t)ccc_sim_footprint(sim_); }
    int64_t workspace_bytes() const { return (int64_t)ccc_sim_workspace_bytes(sim_); }

  private:
    ccc_sim *sim_ = nullptr;
    at::TensorOptions opts_;
    int64_t nv_ = 0, nf_ = 0;
};

}  // namespace

PYBIND11_MODULE(_cccollisions, m) {
    py::class_<PySim>(m, "Sim", "A simulation (or packed meshes) resident on the device; step() runs `iters` impulse passes")
        .def(py::init<at::Tensor, at::Tensor, at::Tensor, at::Tensor, std::vector<int64_t>, int, int, bool, bool, int, int, int64_t>(),
             py::arg("x0"), py::arg("x1"), py::arg("faces"), py::arg("masses"), py::arg("mesh_offsets") = std::vector<int64_t>{},
             py::arg("iters") = 10, py::arg("max_candidates") = 64, py::arg("deterministic") = false, py::arg("graph") = true,
             py::arg("morton") = 0, py::arg("ccd_blocks_per_cu") = 128, py::arg("pair_chunk") = 0)
        .def("step", &PySim::step, py::arg("advance") = false, py::arg("workspace") = py::none())
        .def("autotune", &PySim::autotune, py::arg("workspace") = py::none())
        .def("positions", &PySim::positions)
        .def("set_positions", &PySim::set_positions, py::arg("x0"), py::arg("x1"))
        .def("last_impulses", &PySim::last_impulses)
        .def("counts", &PySim::counts)
        .def("config", &PySim::config)
        .def("tuning_times", &PySim::tuning_times)
        .def("footprint", &PySim::footprint)
        .def("workspace_bytes", &PySim::workspace_bytes);
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
