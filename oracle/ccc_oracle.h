/*
 * CPU oracle for the CCCollisions collision hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against; it is never linked into, loaded by or called from
 * the product library (cccollisions_amd).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it.
 *
 * It is a plain-C restatement of the reference algorithm (Dolorousrtur/CCCollisions,
 * files under src/ and include/), written from the reference sources; every
 * function cites the reference file:line it follows.  All arithmetic mirrors the
 * reference expression order.  Where the reference relies on CUDA device
 * semantics the restatement makes them explicit:
 *   - CUDA min/max on floating point are fmin/fmax (NaN-ignoring);
 *   - nvcc's default contraction of the `coef += x*y*z` accumulations in
 *     figures.cu:150-281 is written as fma(x*y, z, coef);
 *   - every other expression is evaluated without contraction
 *     (built with -ffp-contract=off);
 *   - CAMPARY certifiedMul (double-double) is replaced by Dekker/Ogita-Rump-Oishi
 *     TwoProduct + Fast2Sum, whose leading term is the same rounded product;
 *   - rsqrt in normalize() is evaluated as 1/sqrt.
 *
 * Parity status per function is listed in the header comment of
 * ccc_oracle_body.h.
 *
 * Two instantiations are exported: suffix _f64 (double) and _f32 (float),
 * matching the reference's AT_DISPATCH_FLOATING_TYPES instantiations.
 */
#ifndef CCC_ORACLE_H
#define CCC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-stage counts of one batch element (same meaning as the product's). */
typedef struct {
    int64_t candidates;   /* K: triangle pairs from the broad phase            */
    int64_t subprims;     /* M: VF/EE sub-primitive pairs after the box filter */
    int64_t roots;        /* V: cubic roots inside [0,1]                       */
    int64_t collisions;   /* R: accepted collisions / close pairs              */
} orc_counts;

#define ORC_DECL(SFX, REAL)                                                                  \
    int orc_bvh##SFX(const REAL *tris, const REAL *tris_next, int64_t n, int morton_mode,   \
                     REAL *bboxes, int64_t *tree, uint64_t *codes_out, int32_t *ids_out);   \
    int64_t orc_traverse##SFX(const REAL *bboxes, const int64_t *tree, int64_t n,           \
                              REAL threshold, int64_t *pairs, int64_t cap);                 \
    int64_t orc_traverse_partial##SFX(const REAL *bboxes, const int64_t *tree, int64_t n,   \
                                      const uint8_t *to_check, int64_t *pairs,              \
                                      int64_t cap);                                         \
    int64_t orc_bruteforce_pairs##SFX(const REAL *bboxes, const int64_t *tree, int64_t n,   \
                                      REAL threshold, int64_t *pairs, int64_t cap);         \
    int orc_ccd##SFX(const REAL *tris, const REAL *tris_next, int64_t n,                    \
                     const int64_t *pairs, int64_t k, int64_t *coll, REAL *roots,           \
                     int64_t cap, orc_counts *counts);                                      \
    int orc_impulses##SFX(const REAL *tris, const REAL *tris_next, const REAL *masses,      \
                          int64_t n, const int64_t *coll, const REAL *roots, int64_t r,     \
                          REAL *dv, REAL *dx, int32_t *cnt);                                \
    int orc_proximity##SFX(const REAL *tris, int64_t n, const int64_t *pairs, int64_t k,    \
                           REAL threshold, int64_t *out, int64_t cap, orc_counts *counts);  \
    int64_t orc_static_collisions##SFX(const REAL *tris, int64_t n, const int64_t *pairs,   \
                                       int64_t k, REAL threshold, int64_t *out,             \
                                       int64_t cap);                                        \
    int64_t orc_triangle_edge##SFX(const REAL *tris, int64_t n, const int64_t *pairs,       \
                                   int64_t k, int mode, int64_t *rows, REAL *coefs,         \
                                   int32_t *masks, int64_t cap);                            \
    int64_t orc_mincol##SFX(const REAL *tris, const REAL *tris_next, int64_t n,             \
                            const int64_t *pairs, int64_t k, int64_t *out, int64_t cap);    \
    int orc_cubic_coeffs##SFX(const REAL *tet, REAL *coef);                                 \
    int orc_solve_cubic##SFX(const REAL *coef, REAL *roots3);

ORC_DECL(_f64, double)
ORC_DECL(_f32, float)

/* Double-only helpers exposed for pinning tests. */
double orc_dd_cube_lead(double b);
int orc_karras_node(const uint64_t *codes, const int32_t *ids, int64_t n, int64_t i,
                    int64_t *left, int64_t *right);
int orc_karras_node_literal(const uint64_t *codes, const int32_t *ids, int64_t n, int64_t i,
                            int64_t *left, int64_t *right);
uint32_t orc_morton3d_f64(double x, double y, double z);
uint64_t orc_morton63_f64(double x, double y, double z);
void orc_libm_eval(int fn, const double *x, double *y, int64_t n);

/* Whole CCD+impulse pass (bvh_motion -> traverse -> narrow phase -> impulses) used as
 * the CPU baseline; OpenMP over queries and candidate pairs.  Returns 0 on success. */
int orc_impulse_pass_f64(const double *tris, const double *tris_next, const double *masses,
                         int64_t n, int nthreads, int64_t query_stride, double *dv, double *dx,
                         int32_t *cnt, orc_counts *counts);

#ifdef __cplusplus
}
#endif
#endif
