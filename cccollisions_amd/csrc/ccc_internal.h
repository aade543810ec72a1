// Host-side plumbing shared by the C-ABI translation units: context, workspace
// allocation, error reporting, launch helpers and wave-level primitives.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cccollisions_amd.h"

struct ccc_context {
    int device;
    ccc_alloc_fn alloc;
    ccc_free_fn free_fn;
    void *user;
    int num_cus;
    // pairs per narrow-phase chunk over the whole batch (0: the default, 2^25)
    int64_t pair_chunk = 0;
    // this device's copy of the allocation-free entry point's scratch array
    void *scratch = nullptr;
    // Morton mode of the trees bvh / bvh_motion build (CCC_MORTON_*)
    int morton = CCC_MORTON_COMPAT;
    // opt-in kernel timing (HIP events on the call's stream), accumulated per kernel name
    bool profiling = false;
    std::mutex prof_mu;
    std::map<std::string, std::pair<double, int64_t>> prof;
};

namespace ccc {

void set_error(const std::string &msg);
void clear_error();

// Per-call device workspace.  Buffers come from the context's allocator (or
// hipMallocAsync on the call's stream) and are released when the Workspace goes out of
// scope; the call has synchronised its stream before that happens.
// Arena mode (arena != NULL): buffers are carved from [arena, arena + arena_bytes) in
// order, 256-byte aligned, and nothing is allocated or freed; a request that does not fit
// fails like an allocation failure.
class Workspace {
  public:
    Workspace(ccc_context *ctx, hipStream_t s, void *arena = nullptr, size_t arena_bytes = 0)
        : ctx_(ctx), s_(s), arena_(static_cast<char *>(arena)), arena_bytes_(arena_bytes) {}
    ~Workspace();
    void *get(size_t bytes);
    template <typename P> P *get_as(size_t count) { return static_cast<P *>(get(count * sizeof(P))); }
    // Returns a buffer obtained from get() before the Workspace ends (stream-ordered).  In
    // arena mode only the most recent buffer can be returned (its bytes become free again).
    void release(void *p);
    bool failed() const { return failed_; }
    bool arena() const { return arena_ != nullptr; }
    // Arena mode: bytes not yet handed out (after 256-byte alignment of the next request).
    size_t arena_left() const { return arena_bytes_ > used_ ? arena_bytes_ - used_ : 0; }

  private:
    ccc_context *ctx_;
    hipStream_t s_;
    std::vector<void *> ptrs_;
    bool failed_ = false;
    char *arena_ = nullptr;
    size_t arena_bytes_ = 0, used_ = 0;
    std::vector<size_t> marks_;   // arena mode: offset of each buffer handed out
};

// Records HIP events around kernel launches when the context profiles; collect() runs
// after the call's stream synchronisation and adds the elapsed times to the context.
class KernelTimer {
  public:
    KernelTimer(ccc_context *ctx, hipStream_t s) : ctx_(ctx), s_(s), on_(ctx && ctx->profiling) {}
    ~KernelTimer() { collect(); }
    void begin(const char *name) {
        if (!on_) return;
        Rec r{name, nullptr, nullptr};
        if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) { on_ = false; return; }
        (void)hipEventRecord(r.a, s_);
        recs_.push_back(r);
    }
    void end() {
        if (!on_ || recs_.empty()) return;
        (void)hipEventRecord(recs_.back().b, s_);
    }
    void collect() {
        if (recs_.empty()) return;
        (void)hipStreamSynchronize(s_);
        std::lock_guard<std::mutex> lock(ctx_->prof_mu);
        for (auto &r : recs_) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                auto &e = ctx_->prof[r.name];
                e.first += ms;
                e.second += 1;
            }
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        recs_.clear();
    }

  private:
    struct Rec { const char *name; hipEvent_t a, b; };
    ccc_context *ctx_;
    hipStream_t s_;
    bool on_;
    std::vector<Rec> recs_;
};

#define CCC_TIMED(timer, name, launch) \
    do {                               \
        (timer).begin(name);           \
        launch;                        \
        (timer).end();                 \
    } while (0)

inline ccc_status hip_fail(hipError_t e, const char *what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return CCC_ERR_HIP;
}

#define CCC_HIP_TRY(expr)                                                                       \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess) return ::ccc::hip_fail(_e, #expr);                                \
    } while (0)

#define CCC_LAUNCH_CHECK(what)                                                                  \
    do {                                                                                        \
        hipError_t _e = hipGetLastError();                                                      \
        if (_e != hipSuccess) return ::ccc::hip_fail(_e, what);                                 \
    } while (0)

inline ccc_status bad_arg(const std::string &msg) {
    set_error(msg);
    return CCC_ERR_BAD_ARG;
}

inline unsigned grid_for(int64_t items, int block) {
    int64_t g = (items + block - 1) / block;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// Device error bits reported by kernels.
// DEV_UNORDERED is informational: the tree's subtrees do not cover contiguous, ordered
// leaf ranges (or are deeper than 64), so the rightmost leaves come from the climb.
enum : unsigned { DEV_BAD_TREE = 1u, DEV_STACK = 2u, DEV_UNORDERED = 4u };

// Per-batch-element counters kept on the device during a call.
struct alignas(64) DevCounters {
    unsigned long long cand;    // K (total, may exceed the candidate capacity)
    unsigned long long items;   // M of the current pair chunk
    unsigned long long roots;   // V
    unsigned long long coll;    // R (total, may exceed the output capacity)
    unsigned long long items_done;  // M of the pair chunks already finished
    unsigned int flags;         // DEV_* bits
    unsigned int pad[5];
};
static_assert(sizeof(DevCounters) == 64, "DevCounters is one 64-byte line");

// Device address of the allocation-free entry point's scratch array on `device` (looked up
// once per device with that device current, then cached).
ccc_status scratch_address(int device, void **out);

// One compute_impulses pass of one mesh (or one packed set of meshes) over caller-owned,
// resident buffers, for the resident simulation.  Nothing waits on the host and nothing is
// allocated: workspace comes from `arena` in a fixed order (so the buffers have the same
// addresses on every call with the same sizes and the pass can be captured in a HIP graph),
// the candidate list keeps at most cap_pairs pairs (more are counted in ctr->cand and must
// be detected by the caller after synchronising), the narrow phase covers [0, cap_pairs)
// in chunks of `chunk` pairs whose kernels stop at the counts on the device, and the
// counters stay in `ctr` (zeroed by the pass).  det_acc != NULL: the impulse sums are
// order-independent (fixed-point limbs, kDetWords int64 per corner) and then written to dv
// / dx, so that a pass gives the same bits however its atomics interleave.
struct ResidentPass {
    const void *bboxes;
    const int64_t *tree;
    const void *t0, *t1, *masses;
    int64_t n;
    int32_t dtype;
    const int32_t *mesh_end;     // packed meshes: per triangle, the end of its mesh; else NULL
    int2 *pairs;
    int64_t cap_pairs;
    int64_t chunk;
    int ccd_blocks_per_cu;       // grid cap of k_ccd (0: 128)
    bool stack_deep;             // traverse with the deep stack from the start
    long long *det_acc;
    void *dv, *dx;
    int32_t *cnt;
    DevCounters *ctr;
};
ccc_status resident_pass(ccc_context *ctx, const ResidentPass &p, void *arena, size_t arena_bytes, hipStream_t s);
size_t resident_pass_arena_bytes(ccc_context *ctx, const ResidentPass &p);
constexpr int kDetWordsPerCorner = 24;

// Fill kernels (used instead of hipMemsetAsync so that the calls stay capturable).
ccc_status fill_u64(void *p, int64_t count, uint64_t value, hipStream_t s);
ccc_status fill_u32(void *p, int64_t count, uint32_t value, hipStream_t s);

// BVH over B soups of n triangles each (t1 = NULL: static boxes) into the wire format, with
// Morton mode `morton` (CCC_MORTON_*); arguments already validated, the device already
// current.
ccc_status build_bvh_device(ccc_context *ctx, const void *t0, const void *t1, int64_t B, int64_t n, int32_t dtype, int morton,
                            void *bboxes, int64_t *tree, hipStream_t s, void *arena = nullptr, size_t arena_bytes = 0);
// Workspace bytes of build_bvh_device (all of its buffers, 256-byte rounded).
size_t build_bvh_arena_bytes(int64_t B, int64_t n, int32_t dtype, int morton);

}  // namespace ccc

// ---- wave primitives (wave64) ---------------------------------------------------------
namespace ccc {
// Number of lanes below the calling lane whose bit is set in `mask`.
__device__ __forceinline__ unsigned lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}
__device__ __forceinline__ unsigned long long readfirstlane_u64(unsigned long long v) {
    unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
// Slot of the calling lane in a list whose length is the LDS counter *counter: every active
// lane appends one entry, the wave's entries go in lane order after one atomic.
__device__ __forceinline__ int wave_append1(int *counter) {
    const unsigned long long act = __ballot(1);
    const unsigned below = lanes_below(act);
    int base = 0;
    if (below == 0) base = atomicAdd(counter, (int)__popcll(act));
    base = __builtin_amdgcn_readfirstlane(base);
    return base + (int)below;
}
// Reserve `cnt` consecutive slots (cnt < 16) per active lane of the wave from `counter`
// with a single atomic; returns this lane's first slot.
__device__ __forceinline__ unsigned long long wave_reserve(unsigned long long *counter, unsigned cnt) {
    unsigned long long b0 = __ballot(cnt & 1u), b1 = __ballot(cnt & 2u), b2 = __ballot(cnt & 4u), b3 = __ballot(cnt & 8u);
    unsigned below = lanes_below(b0) + 2u * lanes_below(b1) + 4u * lanes_below(b2) + 8u * lanes_below(b3);
    unsigned long long total = (unsigned long long)__popcll(b0) + 2ull * __popcll(b1) + 4ull * __popcll(b2) + 8ull * __popcll(b3);
    unsigned long long base = 0;
    if (lanes_below(__ballot(1)) == 0 && total) base = atomicAdd(counter, total);
    base = readfirstlane_u64(base);
    return base + below;
}
}  // namespace ccc
