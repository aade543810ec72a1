/*
 * cccollisions_amd -- C ABI of the MI355X (gfx950) collision hot path.
 *
 * Drop-in replacement for the host functions behind the `cccollisions` PyTorch
 * extension of Dolorousrtur/CCCollisions.  Each entry point below replaces one host
 * orchestrator of the reference (cited per function) and keeps its argument meaning,
 * capacities and output layouts; the torch extension (cccollisions_amd/csrc/ccc_torch.cpp)
 * is the counterpart of src/cccollisions.cpp and calls only these functions.
 *
 * Conventions
 *   - Plain device pointers and sizes; no torch types.  Every tensor is contiguous,
 *     row-major, on the device the context was created for.
 *   - Triangle soups: [batch, n, 3, 3] of `dtype` (vertex-major: tri, corner, xyz).
 *   - BVH wire format: bboxes [batch, 2n-1, 6] (min xyz, max xyz) of `dtype`,
 *     tree [batch, 2n-1, 4] int64 (parent, left, right, triangle id); internal nodes are
 *     rows 0..n-2 (row 0 is the root), leaves are rows n-1..2n-2 in sorted order.
 *   - Outputs are caller-allocated and fully written by the call: index rows and roots
 *     are padded with -1, impulse sums and counts start from zero.
 *   - All work is enqueued on `stream` (a hipStream_t; NULL = default stream).  A call
 *     returns after the results it reports (counts, warnings) are known, i.e. after a
 *     stream synchronisation.
 *   - Returns a ccc_status: low byte = error code (0 = success), upper bits = warnings.
 *   - A context holds no per-call state; calls on one context from several host threads
 *     are safe as long as they use different streams.
 */
#ifndef CCCOLLISIONS_AMD_H
#define CCCOLLISIONS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCC_ABI_VERSION 1

typedef int32_t ccc_status;
#define CCC_OK 0
#define CCC_ERR_BAD_ARG 1          /* invalid size, pointer or dtype                     */
#define CCC_ERR_HIP 2              /* a HIP runtime call or kernel launch failed          */
#define CCC_ERR_NO_MEMORY 3        /* workspace allocation failed                         */
#define CCC_ERR_BAD_TREE 4         /* tree tensor is not a binary tree over n leaves      */
#define CCC_ERR_STACK 5            /* traversal deeper than the stack (tree too deep)     */
#define CCC_ERR_MASK 0xff
#define CCC_WARN_CANDIDATES 0x100  /* more candidate pairs than max_candidates * n        */
#define CCC_WARN_TRUNCATED 0x200   /* more result rows than the output holds; rest dropped */

typedef enum { CCC_FLOAT32 = 0, CCC_FLOAT64 = 1 } ccc_dtype;

/* Call flags. */
#define CCC_FLAG_NONE 0u
/* Keep the reference capacity for candidate pairs (max_candidates * n): pairs beyond it
 * are dropped and CCC_WARN_CANDIDATES is raised.  Without this flag the candidate buffer
 * grows to the exact count and the warning still reports that the capacity was exceeded. */
#define CCC_FLAG_FIXED_CAPACITY 1u

/* Per-batch-element stage counts. */
typedef struct {
    int64_t candidates;  /* K: triangle pairs from the broad phase                        */
    int64_t subprims;    /* M: vertex-face / edge-edge pairs after the 15-way box filter */
    int64_t roots;       /* V: cubic roots in [0, 1]                                     */
    int64_t collisions;  /* R: accepted collisions / close pairs / intersecting pairs    */
} ccc_counts;

/* Device workspace allocator (e.g. a framework's caching allocator).  Both callbacks are
 * optional; without them the library uses hipMallocAsync / hipFreeAsync on the stream. */
typedef void *(*ccc_alloc_fn)(void *user, size_t bytes, void *stream);
typedef void (*ccc_free_fn)(void *user, void *ptr, void *stream);

typedef struct ccc_context ccc_context;

ccc_status ccc_context_create(ccc_context **out, int device, ccc_alloc_fn alloc, ccc_free_fn free_fn,
                              void *user);
void ccc_context_destroy(ccc_context *ctx);
const char *ccc_status_string(ccc_status status);
/* Message of the last failed call made on this host thread (empty if none). */
const char *ccc_last_error(void);
int ccc_abi_version(void);

/* Pairs of candidate triangles per narrow-phase chunk, over all batch elements of a call (no
 * counterpart in the reference, whose narrow phase materialises every candidate's
 * tetrahedra at once, src/collisions_continuous.cu:887-1005).  The sub-primitive items of
 * one chunk (at most 15 per pair, 4 bytes each in the continuous queries, 8 in the static
 * ones) are the largest workspace buffer; 0 restores the default of 2^26 pairs (4 GiB of
 * continuous items), at most 2^28.  Results do not depend on it. */
ccc_status ccc_context_set_pair_chunk(ccc_context *ctx, int64_t pairs);

/* Morton codes of the trees ccc_bvh / ccc_bvh_motion build (no counterpart in the reference,
 * whose scene box reduction starts from [-MAX, +MAX], src/build_bvh.cu:243-244, :505-506, so
 * that every code is 0 and the tree is the radix tree over triangle indices).
 *   CCC_MORTON_COMPAT (default): the reference's tree, bit for bit.
 *   CCC_MORTON_30: scene box fitted to the leaf boxes, 30-bit codes as in
 *                  src/bvh.cu:54-62 (10 bits per axis); ties broken by triangle id.
 *   CCC_MORTON_63: fitted scene box, 63-bit codes (21 bits per axis).
 * The fitted trees are valid wire-format BVHs over the same leaf boxes: every query returns
 * the same pair and collision sets on them (rows in another order), except the partial
 * queries, whose triangles_to_check mask is read by sorted leaf position for the query
 * triangle (src/collisions_continuous.cu:1090), which equals the triangle id only in
 * CCC_MORTON_COMPAT.  Faces stored in no spatial order (meshes from files, shuffled soups)
 * give the compat tree boxes that span the whole mesh; the fitted trees do not. */
#define CCC_MORTON_COMPAT 0
#define CCC_MORTON_30 1
#define CCC_MORTON_63 2
ccc_status ccc_context_set_morton(ccc_context *ctx, int mode);

/* Test hook (no counterpart in the reference): the narrow phase's device arithmetic on n
 * given inputs, so that tests can compare it with the CPU oracle directly.  `x` and `y` are
 * device arrays.  what 0-4: y[i] = cbrt / acos / sin / cos / cbrtf(float) of x[i], the
 * portable functions the cubic solver uses; what 5: x[4i..4i+3] are the cubic coefficients
 * c3 c2 c1 c0 (src/collisions_continuous.cu:511-613 solveCubic), y[4i] the number of roots
 * kept in [0, 1] and y[4i+1..4i+3] those roots in order (0 where absent); what 6 / 7 / 8:
 * y[i] = x[i] / 3, x[i] / 27 and x[i] / 3 for finite or NaN x[i], by the solver's
 * division-free sequences (equal to the division bit for bit); what 9: sqrt(x[i]) without
 * rescaling (x[i] = +-0, >= 2^-767, negative or NaN); what 10 / 11: y[i] = 1 when the
 * edge-edge filter's division-free test accepts the quotient x[i] / x[n + i] in double /
 * float (x holds 2 n values), else 0. */
ccc_status ccc_debug_device_math(ccc_context *ctx, int what, const double *x, double *y, int64_t n, void *stream);

/* Opt-in kernel timing (no counterpart in the reference; used by the benchmark).  While on,
 * every kernel launch of a call is bracketed by HIP events on the call's stream and the
 * elapsed times are accumulated per kernel name.  Enabling or disabling clears the totals. */
ccc_status ccc_context_set_profiling(ccc_context *ctx, int on);
ccc_status ccc_context_kernel_time(ccc_context *ctx, const char *kernel, double *total_ms, int64_t *launches);
/* Comma-separated names of the kernels timed so far; returns the full length. */
int ccc_context_kernel_names(ccc_context *ctx, char *buf, size_t len);

/* Replaces build_bvh(triangles, ...) -- src/build_bvh.cu:381-449 (binding
 * src/cccollisions.cpp:102-116).  Static leaf boxes. */
ccc_status ccc_bvh(ccc_context *ctx, const void *triangles, int64_t batch, int64_t n, int32_t dtype,
                   void *bboxes, int64_t *tree, void *stream);

/* Replaces build_bvh(triangles, triangles_next, ...) -- src/build_bvh.cu:570-648 (binding
 * src/cccollisions.cpp:118-132).  Swept leaf boxes box(t0) U box(t1). */
ccc_status ccc_bvh_motion(ccc_context *ctx, const void *triangles, const void *triangles_next,
                          int64_t batch, int64_t n, int32_t dtype, void *bboxes, int64_t *tree,
                          void *stream);

/* Replaces find_collisions_from_bbox_tree(bboxes, tree, triangles, collisions, threshold,
 * max_collisions) -- src/collisions_static.cu:206-342 (binding src/cccollisions.cpp:134-148).
 * out: [batch, n*max_collisions, 2] int64 intersecting pairs (x < y). */
ccc_status ccc_find_collisions(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                               const void *triangles, int64_t batch, int64_t n, int32_t dtype,
                               float threshold, int32_t max_collisions, int64_t *out,
                               ccc_counts *counts, uint32_t flags, void *stream);

/* Replaces find_proximity_from_bbox_tree -- src/proximity_static.cu:510-695 (binding
 * src/cccollisions.cpp:276-296).  out: [batch, n*max_candidates, 3] int64 (x, y, type). */
ccc_status ccc_find_proximity(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                              const void *triangles, int64_t batch, int64_t n, int32_t dtype,
                              float threshold, int32_t max_candidates, int64_t *out,
                              ccc_counts *counts, uint32_t flags, void *stream);

/* Replaces find_collisions_from_bbox_tree(..., triangles_next, collisions, roots, ...) --
 * src/collisions_continuous.cu:814-1076 (binding src/cccollisions.cpp:220-245).
 * collisions: [batch, n*max_collisions, 3] int64 (x, y, type); roots: [batch, n*max_collisions]. */
ccc_status ccc_find_collisions_continuous(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                                          const void *triangles, const void *triangles_next,
                                          int64_t batch, int64_t n, int32_t dtype,
                                          int32_t max_candidates, int32_t max_collisions,
                                          int64_t *collisions, void *roots, ccc_counts *counts,
                                          uint32_t flags, void *stream);

/* Replaces find_collisions_from_bbox_tree_partial -- src/collisions_continuous.cu:1103-1376
 * (binding src/cccollisions.cpp:247-273).  to_check: [batch, n] bool (1 byte each). */
ccc_status ccc_find_collisions_continuous_partial(ccc_context *ctx, const void *bboxes,
                                                  const int64_t *tree, const void *triangles,
                                                  const void *triangles_next, const uint8_t *to_check,
                                                  int64_t batch, int64_t n, int32_t dtype,
                                                  int32_t max_candidates, int32_t max_collisions,
                                                  int64_t *collisions, void *roots,
                                                  ccc_counts *counts, uint32_t flags, void *stream);

/* Replaces find_triangle_edge_collisions_from_bbox_tree -- src/collisions_static.cu:612-779
 * (binding src/cccollisions.cpp:151-173; kernels :345-500).  Traversal and SAT at threshold 0
 * (the binding's threshold is unused), pairs sharing a vertex rejected, then each of the six
 * edges of a pair tested against the other triangle.  out: [batch, n*max_collisions, 4]
 * int64 rows (x, y, edge id 0..5, loop id -1/0), -1 padded; coefs: [batch, n*max_collisions]
 * of dtype, the edge parameter of the piercing point, -1 padded.  Rows of batch element b
 * start at b * n*max_collisions (the reference's copy uses a stride of 3 there and writes
 * up to 6 rows per pair past the end; here rows beyond the capacity are dropped with
 * CCC_WARN_TRUNCATED). */
ccc_status ccc_find_triangle_edge_collisions(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                                             const void *triangles, int64_t batch, int64_t n, int32_t dtype,
                                             int32_t max_collisions, int64_t *out, void *coefs,
                                             ccc_counts *counts, uint32_t flags, void *stream);

/* Replaces find_triangle_edge_candidates -- src/collisions_static.cu:782-943 (binding
 * src/cccollisions.cpp:176-194; kernel :520-604).  Pairs may share one vertex; a row
 * (x, y, edge id, loop id) for every edge that does not end on a corner of the other
 * triangle, loop id 0 when it starts on one (edge ids 0..2 only).  out: [batch,
 * n*max_collisions, 4] int64, -1 padded, truncated as above. */
ccc_status ccc_find_triangle_edge_candidates(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                                             const void *triangles, int64_t batch, int64_t n, int32_t dtype,
                                             int32_t max_collisions, int64_t *out, ccc_counts *counts,
                                             uint32_t flags, void *stream);

/* Replaces find_triangle_edge_candidates2 -- src/icontour.cu:110-282 (binding
 * src/cccollisions.cpp:197-218; kernel icontour.cu:11-108).  One row per pair that passes
 * the candidate filter: pairs [batch, n*max_collisions, 2] int64 (-1 padded) and masks
 * [batch, n*max_collisions, 2] int32 (0 padded) with bit e of mask.x set for each edge
 * row the candidates query would emit and bit e of mask.y for each loop node. */
ccc_status ccc_find_triangle_edge_candidates2(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                                              const void *triangles, int64_t batch, int64_t n, int32_t dtype,
                                              int32_t max_collisions, int64_t *pairs, int32_t *masks,
                                              ccc_counts *counts, uint32_t flags, void *stream);

/* Replaces find_minimal_candidates_from_bbox_tree -- src/collisions_continuous.cu:1379-1519
 * (binding src/cccollisions.cpp:405-423).  out: [batch, n*max_collisions*15, 3] int64. */
ccc_status ccc_find_mincol_cand(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                                const void *triangles, const void *triangles_next, int64_t batch,
                                int64_t n, int32_t dtype, int32_t max_collisions, int64_t *out,
                                ccc_counts *counts, uint32_t flags, void *stream);

/* Replaces collision_impulses_from_bbox_tree -- src/collisions_impulses.cu:1069-1397
 * (binding src/cccollisions.cpp:299-333).  masses: [batch, n, 3]; dv, dx: [batch, n, 3, 3]
 * of dtype; cnt: [batch, n, 3] int32.  max_collisions is accepted and unused, as in the
 * reference. */
ccc_status ccc_compute_impulses(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                                const void *triangles, const void *triangles_next,
                                const void *masses, int64_t batch, int64_t n, int32_t dtype,
                                int32_t max_candidates, int32_t max_collisions, void *dv, void *dx,
                                int32_t *cnt, ccc_counts *counts, uint32_t flags, void *stream);

/* Packed batch of independent meshes of different sizes: compute_impulses for every mesh
 * in one pass of launches.  No single reference function corresponds: the reference loops
 * over the equal-size meshes of a [batch, n] tensor one at a time
 * (src/collisions_impulses.cu:1069-1397, `for (int bidx = 0; ...)`); this entry point gives
 * each mesh the results of that function (bvh_motion, src/build_bvh.cu:570-648, followed by
 * collision_impulses_from_bbox_tree) for meshes of any sizes.
 * triangles, triangles_next: [N, 3, 3], masses: [N, 3], meshes back to back, mesh m owning
 * triangles [offsets[m], offsets[m+1]) (offsets: HOST array of num_meshes + 1 entries,
 * offsets[0] = 0, strictly increasing, N = offsets[num_meshes] < 2^28).  dv, dx: [N, 3, 3],
 * cnt: [N, 3].  max_candidates bounds the candidate list at max_candidates * N pairs over
 * the whole batch (grown as in ccc_compute_impulses).  counts: one entry, batch totals.
 * Pairs never cross meshes; each mesh's pairs, roots and counts equal its single-mesh call,
 * its impulses within fp64 atomic reordering. */
ccc_status ccc_compute_impulses_packed(ccc_context *ctx, const void *triangles, const void *triangles_next,
                                       const void *masses, const int64_t *offsets, int64_t num_meshes,
                                       int32_t dtype, int32_t max_candidates, void *dv, void *dx, int32_t *cnt,
                                       ccc_counts *counts, uint32_t flags, void *stream);

/* Size of the device scratch array behind ccc_compute_impulses_noalloc (256 MiB per GPU). */
#define CCC_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)

/* ccc_compute_impulses without any device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results and statuses as
 * ccc_compute_impulses; the caller passes every input and output, and the call's workspace
 * (counters, traversal nodes, candidate pairs, sub-primitive items) is carved from one
 * static device array of CCC_NOALLOC_SCRATCH_BYTES per GPU, held from the first context
 * created on that GPU until the process ends.  A call whose workspace does not fit returns
 * CCC_ERR_NO_MEMORY with a message (the candidate pairs alone take max_candidates * n * 8
 * bytes per batch element); the sub-primitive item list shrinks to what is left, which does
 * not change the results.  Every call on one GPU uses that GPU's single array: calls that
 * overlap on one GPU (several streams or host threads) race and must be serialised by the
 * caller.  The call still synchronises its stream, as ccc_compute_impulses does. */
ccc_status ccc_compute_impulses_noalloc(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                                        const void *triangles, const void *triangles_next,
                                        const void *masses, int64_t batch, int64_t n, int32_t dtype,
                                        int32_t max_candidates, int32_t max_collisions, void *dv, void *dx,
                                        int32_t *cnt, ccc_counts *counts, uint32_t flags, void *stream);

/* ---- Resident simulations (no counterpart in the reference) --------------------------------
 * The reference is stateless: its caller (a cloth simulator, not part of the reference) runs
 * bvh_motion + compute_impulses and applies the corrections itself, with a host
 * synchronisation after every launch (src/flags.h:63-77).  A ccc_sim keeps one simulation
 * resident on the device -- positions x0 (start of step) and x1 (predicted end of step)
 * [V, 3], faces [F, 3] int64, corner masses [F, 3], triangle soups, BVH, candidate-pair
 * list, impulse outputs and a vertex -> corner index -- and runs whole steps of `iters`
 * impulse passes; a pass is bvh_motion (src/build_bvh.cu:570-648) + compute_impulses
 * (src/collisions_impulses.cu:1069-1397) + the caller's update: every vertex of x1 moves by
 * the mean of the finite position corrections dx of its triangle corners (corners summed in
 * ascending order; a corner with a non-finite dx is skipped).  Several meshes of different
 * sizes may be packed into one simulation (faces index one vertex array; mesh m owns faces
 * [offsets[m], offsets[m+1])): pairs never cross meshes, one BVH and one traversal cover
 * them all.  A step neither allocates nor waits on the host until its end: it is captured
 * once as a HIP graph (opts.graph) per workspace arena and replayed.  At the end of a step
 * the per-pass counters are read; a pass that found more candidate pairs than the list
 * holds, or whose traversal needed the deep stack, grows the list and the step runs again
 * from the saved x1.  All calls on one simulation must be serialised by the caller. */
typedef struct ccc_sim ccc_sim;
typedef struct {
    int32_t iters;              /* impulse passes per step (default 10)                          */
    int32_t max_candidates;     /* initial candidate capacity per triangle (default 64; grows)  */
    int32_t deterministic;      /* order-independent fixed-point impulse sums: every step gives
                                   the same bits, alone or beside other work (default 0)      */
    int32_t graph;              /* capture each step as one HIP graph (default 1)               */
    int32_t morton;             /* CCC_MORTON_* of the step's trees (one-mesh simulations)      */
    int32_t ccd_blocks_per_cu;  /* grid cap of the narrow-phase kernel (default 128)             */
    int64_t pair_chunk;         /* candidate pairs per narrow-phase chunk (0: 2^26)            */
} ccc_sim_opts;
void ccc_sim_default_opts(ccc_sim_opts *opts);
/* x0, x1: device [V, 3]; faces: device [F, 3] int64; masses: device [F, 3]; all copied into
 * the simulation.  mesh_face_offsets: HOST array of num_meshes + 1 entries, or NULL for one
 * mesh.  The call synchronises `stream`. */
ccc_status ccc_sim_create(ccc_context *ctx, const void *x0, const void *x1, int64_t num_vertices, const int64_t *faces,
                          const void *masses, int64_t num_faces, const int64_t *mesh_face_offsets, int64_t num_meshes,
                          int32_t dtype, const ccc_sim_opts *opts, void *stream, ccc_sim **out);
void ccc_sim_destroy(ccc_sim *sim);
/* Bytes of device workspace one step needs (at the current candidate capacity). */
size_t ccc_sim_workspace_bytes(const ccc_sim *sim);
/* One step on `stream` (synchronised before returning).  workspace: NULL for the
 * simulation's own (allocated when first needed), or a caller-owned device arena of at least
 * ccc_sim_workspace_bytes bytes (e.g. one per serving slot; CCC_ERR_NO_MEMORY if smaller).
 * advance != 0: afterwards x0 <- x1 and x1 <- x1 + (x1 - x0) (the next step's prediction). */
ccc_status ccc_sim_step(ccc_sim *sim, void *workspace, size_t workspace_bytes, int32_t advance, void *stream);
/* Times steps under candidate launch configurations (narrow-phase grid 64 / 128 / 256 blocks
 * per CU; for one-mesh simulations also the reference's and the 30-bit fitted Morton tree),
 * keeps the fastest and remembers it for the context's simulations of the same size bucket
 * (then no timing).  x1 is left as it was. */
ccc_status ccc_sim_autotune(ccc_sim *sim, void *workspace, size_t workspace_bytes, void *stream);
ccc_status ccc_sim_get_opts(const ccc_sim *sim, ccc_sim_opts *opts);
/* Milliseconds per step of each configuration the last autotune timed; returns their number. */
int32_t ccc_sim_tuning_times(const ccc_sim *sim, double *ms, int32_t max);
/* Device addresses of the resident positions ([V, 3] of the simulation's dtype). */
ccc_status ccc_sim_state(const ccc_sim *sim, void **x0, void **x1, int64_t *num_vertices, int64_t *candidate_capacity);
/* Device addresses of the last pass's impulse outputs: dv, dx [F, 3, 3], cnt [F, 3] int32. */
ccc_status ccc_sim_impulses(const ccc_sim *sim, void **dv, void **dx, int32_t **cnt);
/* Stage counts of each pass of the last step; returns the number of passes. */
int32_t ccc_sim_step_counts(const ccc_sim *sim, ccc_counts *per_pass, int32_t max_passes);
/* Device bytes the simulation holds (resident state + own workspace). */
size_t ccc_sim_footprint(const ccc_sim *sim);

/* Replaces collision_impulses_from_bbox_tree_partial -- src/collisions_impulses.cu:1402-1734
 * (binding src/cccollisions.cpp:336-370). */
ccc_status ccc_compute_impulses_partial(ccc_context *ctx, const void *bboxes, const int64_t *tree,
                                        const void *triangles, const void *triangles_next,
                                        const void *masses, const uint8_t *to_check, int64_t batch,
                                        int64_t n, int32_t dtype, int32_t max_candidates,
                                        int32_t max_collisions, void *dv, void *dx, int32_t *cnt,
                                        ccc_counts *counts, uint32_t flags, void *stream);

/* Replaces collision_impulses_from_continuous_collisions -- src/collisions_impulses.cu:1736-1880
 * (binding src/cccollisions.cpp:374-403).  collisions: [batch, m, 3] int64; roots:
 * [batch, m] of dtype.  Rows holding a negative index (the -1 padding) are skipped. */
ccc_status ccc_compute_impulses_from_collisions(ccc_context *ctx, const int64_t *collisions,
                                                const void *roots, int64_t m,
                                                const void *triangles, const void *triangles_next,
                                                const void *masses, int64_t batch, int64_t n,
                                                int32_t dtype, void *dv, void *dx, int32_t *cnt,
                                                uint32_t flags, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CCCOLLISIONS_AMD_H */
