// Context, workspace and error plumbing of the C ABI.
#include "ccc_internal.h"

#include <cstring>
#include <new>

namespace ccc {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

Workspace::~Workspace() {
    if (arena_) return;
    for (void *p : ptrs_) {
        if (ctx_ && ctx_->free_fn) ctx_->free_fn(ctx_->user, p, (void *)s_);
        else (void)hipFreeAsync(p, s_);
    }
}

void Workspace::release(void *p) {
    if (arena_) {
        if (!marks_.empty() && arena_ + marks_.back() == static_cast<char *>(p)) {
            used_ = marks_.back();
            marks_.pop_back();
        }
        return;
    }
    for (size_t i = 0; i < ptrs_.size(); ++i) {
        if (ptrs_[i] != p) continue;
        if (ctx_ && ctx_->free_fn) ctx_->free_fn(ctx_->user, p, (void *)s_);
        else (void)hipFreeAsync(p, s_);
        ptrs_.erase(ptrs_.begin() + (std::ptrdiff_t)i);
        return;
    }
}

void *Workspace::get(size_t bytes) {
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~(size_t)255;
    if (arena_) {
        if (bytes > arena_left()) {
            failed_ = true;
            set_error("the call needs more than the " + std::to_string(arena_bytes_) +
                      "-byte scratch buffer: " + std::to_string(used_) + " bytes in use, " + std::to_string(bytes) +
                      " more requested");
            return nullptr;
        }
        marks_.push_back(used_);
        void *p = arena_ + used_;
        used_ += bytes;
        return p;
    }
    void *p = nullptr;
    if (ctx_ && ctx_->alloc) {
        p = ctx_->alloc(ctx_->user, bytes, (void *)s_);
    } else {
        if (hipMallocAsync(&p, bytes, s_) != hipSuccess) p = nullptr;
    }
    if (!p) {
        failed_ = true;
        set_error("workspace allocation of " + std::to_string(bytes) + " bytes failed");
        return nullptr;
    }
    ptrs_.push_back(p);
    return p;
}

__global__ void k_fill_u64(uint64_t *p, int64_t count, uint64_t value) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}
__global__ void k_fill_u32(uint32_t *p, int64_t count, uint32_t value) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

static unsigned fill_grid(int64_t count) {
    int64_t g = (count + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

ccc_status fill_u64(void *p, int64_t count, uint64_t value, hipStream_t s) {
    if (count <= 0) return CCC_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(fill_grid(count)), dim3(256), 0, s, (uint64_t *)p, count, value);
    CCC_LAUNCH_CHECK("fill_u64");
    return CCC_OK;
}
ccc_status fill_u32(void *p, int64_t count, uint32_t value, hipStream_t s) {
    if (count <= 0) return CCC_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(count)), dim3(256), 0, s, (uint32_t *)p, count, value);
    CCC_LAUNCH_CHECK("fill_u32");
    return CCC_OK;
}

}  // namespace ccc

extern "C" {

ccc_status ccc_context_create(ccc_context **out, int device, ccc_alloc_fn alloc, ccc_free_fn free_fn, void *user) {
    ccc::clear_error();
    if (!out) return ccc::bad_arg("ccc_context_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    CCC_HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return ccc::bad_arg("ccc_context_create: invalid device " + std::to_string(device));
    int cus = 0;
    CCC_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    ccc_context *c = new (std::nothrow) ccc_context;
    if (!c) return CCC_ERR_NO_MEMORY;
    c->device = device;
    c->alloc = alloc;
    c->free_fn = free_fn;
    c->user = user;
    c->num_cus = cus > 0 ? cus : 256;
    ccc_status st = ccc::scratch_address(device, &c->scratch);
    if (st != CCC_OK) {
        delete c;
        return st;
    }
    *out = c;
    return CCC_OK;
}

void ccc_context_destroy(ccc_context *ctx) { delete ctx; }

const char *ccc_status_string(ccc_status status) {
    switch (status & CCC_ERR_MASK) {
        case CCC_OK: return (status & CCC_WARN_TRUNCATED) ? "ok (output truncated)" : (status & CCC_WARN_CANDIDATES) ? "ok (candidate capacity exceeded)" : "ok";
        case CCC_ERR_BAD_ARG: return "invalid argument";
        case CCC_ERR_HIP: return "HIP runtime error";
        case CCC_ERR_NO_MEMORY: return "out of device memory";
        case CCC_ERR_BAD_TREE: return "malformed BVH tree";
        case CCC_ERR_STACK: return "BVH deeper than the traversal stack";
        default: return "unknown status";
    }
}

const char *ccc_last_error(void) { return ccc::g_last_error.c_str(); }

int ccc_abi_version(void) { return CCC_ABI_VERSION; }

ccc_status ccc_context_set_pair_chunk(ccc_context *ctx, int64_t pairs) {
    if (!ctx) return ccc::bad_arg("context is NULL");
    if (pairs < 0) return ccc::bad_arg("pair chunk must be >= 0");
    ctx->pair_chunk = pairs;
    return CCC_OK;
}

ccc_status ccc_context_set_morton(ccc_context *ctx, int mode) {
    if (!ctx) return ccc::bad_arg("context is NULL");
    if (mode != CCC_MORTON_COMPAT && mode != CCC_MORTON_30 && mode != CCC_MORTON_63)
        return ccc::bad_arg("Morton mode must be CCC_MORTON_COMPAT, CCC_MORTON_30 or CCC_MORTON_63");
    ctx->morton = mode;
    return CCC_OK;
}

ccc_status ccc_context_set_profiling(ccc_context *ctx, int on) {
    if (!ctx) return ccc::bad_arg("context is NULL");
    std::lock_guard<std::mutex> lock(ctx->prof_mu);
    ctx->profiling = on != 0;
    ctx->prof.clear();
    return CCC_OK;
}

ccc_status ccc_context_kernel_time(ccc_context *ctx, const char *kernel, double *total_ms, int64_t *launches) {
    if (!ctx || !kernel || !total_ms || !launches) return ccc::bad_arg("NULL argument");
    std::lock_guard<std::mutex> lock(ctx->prof_mu);
    auto it = ctx->prof.find(kernel);
    *total_ms = it == ctx->prof.end() ? 0.0 : it->second.first;
    *launches = it == ctx->prof.end() ? 0 : it->second.second;
    return CCC_OK;
}

int ccc_context_kernel_names(ccc_context *ctx, char *buf, size_t len) {
    if (!ctx) return -1;
    std::string all;
    {
        std::lock_guard<std::mutex> lock(ctx->prof_mu);
        for (auto &kv : ctx->prof) { if (!all.empty()) all += ","; all += kv.first; }
    }
    if (buf && len > 0) {
        size_t n = all.size() < len - 1 ? all.size() : len - 1;
        memcpy(buf, all.data(), n);
        buf[n] = 0;
    }
    return (int)all.size();
}

}  // extern "C"
