/*
 * CPU oracle for the CCCollisions hot path -- TEST INFRASTRUCTURE ONLY (see ccc_oracle.h).
 * Built by oracle/Makefile with -ffp-contract=off so that every expression is evaluated
 * exactly as written (explicit fma() calls model nvcc's contraction where noted).
 */
#include "ccc_oracle.h"
#include "ccc_libm_ref.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

/* ---- dtype-independent helpers ---------------------------------------------------- */

/* bvh.cu:34-46 expandBits: spread the low 10 bits of v to every third bit. */
static uint32_t orc_expand_bits(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

/* bvh.cu:54-62 morton3D<double>: x = min(max(x*1024, 0), 1023) with fmin/fmax (CUDA
 * min/max overloads), truncated to an integer; code = xx*4 + yy*2 + zz. */
uint32_t orc_morton3d_f64(double x, double y, double z) {
    x = fmin(fmax(x * 1024.0f, 0.0f), 1023.0f);
    y = fmin(fmax(y * 1024.0f, 0.0f), 1023.0f);
    z = fmin(fmax(z * 1024.0f, 0.0f), 1023.0f);
    uint32_t xx = orc_expand_bits((uint32_t)x), yy = orc_expand_bits((uint32_t)y),
             zz = orc_expand_bits((uint32_t)z);
    return xx * 4 + yy * 2 + zz;
}
static uint32_t orc_morton3d_f32(float x, float y, float z) {
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
    z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
    uint32_t xx = orc_expand_bits((uint32_t)x), yy = orc_expand_bits((uint32_t)y),
             zz = orc_expand_bits((uint32_t)z);
    return xx * 4 + yy * 2 + zz;
}

/* 63-bit variant (not in the reference, whose codes have 10 bits per axis): the same
 * quantisation with 21 bits per axis (scale 2^21, clamp to [0, 2^21 - 1], truncate, NaN -> 0)
 * and the same interleave order, x bits at 3b + 2, y at 3b + 1, z at 3b. */
static uint64_t orc_expand_bits21(uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}
uint64_t orc_morton63_f64(double x, double y, double z) {
    x = fmin(fmax(x * 2097152.0, 0.0), 2097151.0);
    y = fmin(fmax(y * 2097152.0, 0.0), 2097151.0);
    z = fmin(fmax(z * 2097152.0, 0.0), 2097151.0);
    return orc_expand_bits21((uint64_t)x) * 4 + orc_expand_bits21((uint64_t)y) * 2 + orc_expand_bits21((uint64_t)z);
}
static uint64_t orc_morton63_f32(float x, float y, float z) {
    x = fminf(fmaxf(x * 2097152.0f, 0.0f), 2097151.0f);
    y = fminf(fmaxf(y * 2097152.0f, 0.0f), 2097151.0f);
    z = fminf(fmaxf(z * 2097152.0f, 0.0f), 2097151.0f);
    return orc_expand_bits21((uint64_t)x) * 4 + orc_expand_bits21((uint64_t)y) * 2 + orc_expand_bits21((uint64_t)z);
}

/* Stable LSD radix sort of (code, id) by code -- the contract of thrust::sort_by_key on
 * unsigned keys (build_bvh.cu:274-275, :536-537); 64-bit keys, 8 bits per pass. */
static void orc_stable_sort_u64(uint64_t *keys, int32_t *vals, int64_t n) {
    uint64_t *k2 = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    int32_t *v2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    for (int pass = 0; pass < 8; ++pass) {
        int sh = 8 * pass;
        int64_t hist[257];
        memset(hist, 0, sizeof(hist));
        for (int64_t i = 0; i < n; ++i) hist[((keys[i] >> sh) & 255u) + 1]++;
        for (int b = 0; b < 256; ++b) hist[b + 1] += hist[b];
        for (int64_t i = 0; i < n; ++i) {
            int64_t o = hist[(keys[i] >> sh) & 255u]++;
            k2[o] = keys[i]; v2[o] = vals[i];
        }
        memcpy(keys, k2, sizeof(uint64_t) * (size_t)n);
        memcpy(vals, v2, sizeof(int32_t) * (size_t)n);
    }
    free(k2); free(v2);
}

/* Edge table of the 15-way sub-primitive types: E = {(0,1),(0,2),(1,2)}
 * (collisions_continuous.cu:46-91). */
static const int ORC_EDGE[3][2] = {{0, 1}, {0, 2}, {1, 2}};

static int orc_clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }
static int orc_clz64(uint64_t x) { return x == 0 ? 64 : __builtin_clzll(x); }

/* bvh.cu:70-89 LongestCommonPrefix: -1 outside [0, n-1]; equal keys add the common
 * prefix of the ids.  Keys are held in 64 bits: for the reference's 30-bit codes every
 * valid delta is then 32 larger than the reference's clz32-based one (32 + clz(id ^ id) on
 * ties becomes 64 + clz(id ^ id)), which changes none of the comparisons Karras makes, so
 * the trees are the same. */
static int orc_delta(const uint64_t *codes, const int32_t *ids, int64_t n, int64_t i, int64_t j) {
    if (i < 0 || i > n - 1 || j < 0 || j > n - 1) return -1;
    uint64_t a = codes[i], b = codes[j];
    if (a == b) return 64 + orc_clz32((uint32_t)(ids[i] ^ ids[j]));
    return orc_clz64(a ^ b);
}

/* Textbook Karras 2012 (Algorithm 1) with the delta above; this is what
 * BuildRadixTree (build_bvh.cu:74-143) computes.  left/right are tree row indices:
 * internal g -> g, leaf p -> n-1+p. */
int orc_karras_node(const uint64_t *codes, const int32_t *ids, int64_t n, int64_t i,
                    int64_t *left, int64_t *right) {
    int d = (orc_delta(codes, ids, n, i, i + 1) - orc_delta(codes, ids, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = orc_delta(codes, ids, n, i, i - d);
    int64_t lmax = 2;
    while (orc_delta(codes, ids, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (orc_delta(codes, ids, n, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = orc_delta(codes, ids, n, i, j);
    int64_t s = 0;
    for (int64_t div = 2;; div *= 2) {
        int64_t t = (l + div - 1) / div;
        if (orc_delta(codes, ids, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    int64_t g = i + s * d + (d < 0 ? d : 0);
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    *left = (lo == g) ? (n - 1 + g) : g;
    *right = (hi == g + 1) ? (n - 1 + g + 1) : (g + 1);
    return 0;
}

/* The two search loops exactly as written in build_bvh.cu:100-123: the first step of
 * each loop is tested twice and the ceil-based split loop keeps testing t = 1 while the
 * 32-bit divider doubles (it is stopped here where the int would overflow). */
int orc_karras_node_literal(const uint64_t *codes, const int32_t *ids, int64_t n, int64_t i,
                            int64_t *left, int64_t *right) {
    int delta_next = orc_delta(codes, ids, n, i, i + 1);
    int delta_last = orc_delta(codes, ids, n, i, i - 1);
    int direction = delta_next - delta_last >= 0 ? 1 : -1;
    int delta_min = orc_delta(codes, ids, n, i, i - direction);
    int64_t lmax = 2;
    while (orc_delta(codes, ids, n, i, i + lmax * direction) > delta_min) lmax *= 2;
    int64_t l = 0;
    int64_t divider = 2;
    for (int64_t t = lmax / divider; t >= 1; divider *= 2) {
        if (orc_delta(codes, ids, n, i, i + (l + t) * direction) > delta_min) l = l + t;
        t = lmax / divider;
    }
    int64_t j = i + l * direction;
    int node_delta = orc_delta(codes, ids, n, i, j);
    int64_t s = 0;
    divider = 2;
    for (int64_t t = (l + (divider - 1)) / divider; t >= 1; divider *= 2) {
        if (orc_delta(codes, ids, n, i, i + (s + t) * direction) > node_delta) s = s + t;
        if (divider >= ((int64_t)1 << 31)) break;
        t = (l + (divider - 1)) / divider;
    }
    int64_t split = i + s * direction + (direction < 0 ? direction : 0);
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    *left = (lo == split) ? (n - 1 + split) : split;
    *right = (hi == split + 1) ? (n - 1 + split + 1) : (split + 1);
    return 0;
}

/* Leading term of the double-double expansion of b*b*b.  Replaces CAMPARY
 * certifiedMul<1,1,2> followed by certifiedMul<2,1,2> (collisions_continuous.cu:477-484,
 * :501-503) with TwoProduct (Dekker 1971, via fma: Ogita, Rump & Oishi 2005) and a final
 * Fast2Sum: b^2 = p + e exactly; p*b = q + f exactly; lead = fl(q + fl(f + fl(e*b))). */
double orc_dd_cube_lead(double b) {
    double p = b * b;
    double e = fma(b, b, -p);
    double q = p * b;
    double f = fma(p, b, -q);
    double t = f + e * b;
    return q + t;
}

/* ---- instantiations ----------------------------------------------------------------- */
#define REAL double
#define SFX _f64
#define RFMIN fmin
#define RFMAX fmax
#define RFMA fma
#define RFABS fabs
#define RSQRT_BASE sqrt
#define RMAXV DBL_MAX
#define MORTON(x, y, z) orc_morton3d_f64((x), (y), (z))
#define MORTON63(x, y, z) orc_morton63_f64((x), (y), (z))
#include "ccc_oracle_body.h"
#undef REAL
#undef SFX
#undef RFMIN
#undef RFMAX
#undef RFMA
#undef RFABS
#undef RSQRT_BASE
#undef RMAXV
#undef MORTON
#undef MORTON63

#define REAL float
#define SFX _f32
#define RFMIN fminf
#define RFMAX fmaxf
#define RFMA fmaf
#define RFABS fabsf
#define RSQRT_BASE sqrtf
#define RMAXV FLT_MAX
#define MORTON(x, y, z) orc_morton3d_f32((x), (y), (z))
#define MORTON63(x, y, z) orc_morton63_f32((x), (y), (z))
#include "ccc_oracle_body.h"
#undef REAL
#undef SFX
#undef RFMIN
#undef RFMAX
#undef RFMA
#undef RFABS
#undef RSQRT_BASE
#undef RMAXV
#undef MORTON
#undef MORTON63

/* ---- CPU baseline: one CCD + impulse pass ------------------------------------------
 * bvh_motion (orc_bvh_f64) -> traversal -> narrow phase -> impulses, parallelised with
 * OpenMP over query leaves.  query_stride > 1 processes only every query_stride-th query
 * leaf (a bounded sample of the pass for timing).  Impulse sums use per-thread
 * partial arrays reduced at the end (sum order differs from the sequential oracle at the
 * ULP level only). */
int orc_impulse_pass_f64(const double *tris, const double *tris_next, const double *masses,
                         int64_t n, int nthreads, int64_t query_stride, double *dv, double *dx,
                         int32_t *cnt, orc_counts *counts) {
    if (query_stride < 1) query_stride = 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t m = 2 * n - 1;
    double *bb = (double *)malloc(sizeof(double) * 6 * (size_t)m);
    int64_t *tree = (int64_t *)malloc(sizeof(int64_t) * 4 * (size_t)m);
    if (!bb || !tree) { free(bb); free(tree); return -2; }
    orc_bvh_f64(tris, tris_next, n, 0, bb, tree, NULL, NULL);
    memset(dv, 0, sizeof(double) * 9 * (size_t)n);
    memset(dx, 0, sizeof(double) * 9 * (size_t)n);
    memset(cnt, 0, sizeof(int32_t) * 3 * (size_t)n);
    int64_t K = 0, M = 0, V = 0, R = 0;
    int64_t *rm = orc_rightmost_f64(tree, n);
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    double **pdv = (double **)calloc((size_t)nt, sizeof(double *));
    double **pdx = (double **)calloc((size_t)nt, sizeof(double *));
    int32_t **pcnt = (int32_t **)calloc((size_t)nt, sizeof(int32_t *));
#pragma omp parallel reduction(+ : K, M, V, R)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        pdv[tid] = (double *)calloc(9 * (size_t)n, sizeof(double));
        pdx[tid] = (double *)calloc(9 * (size_t)n, sizeof(double));
        pcnt[tid] = (int32_t *)calloc(3 * (size_t)n, sizeof(int32_t));
        int64_t stk[64];
#pragma omp for schedule(dynamic, 64)
        for (int64_t p = 0; p < n - 1; p += query_stride) {
            int64_t qrow = n - 1 + p;
            box_f64 q = load_box_f64(bb + 6 * qrow);
            int64_t qid = tree[4 * qrow + 3];
            int sp = 0;
            stk[sp++] = -1;
            int64_t node = 0;
            do {
                int64_t ch[2] = {tree[4 * node + 1], tree[4 * node + 2]};
                int ov[2], trav[2];
                for (int sd = 0; sd < 2; ++sd) {
                    ov[sd] = check_overlap_f64(q, load_box_f64(bb + 6 * ch[sd]), 0.0) && p < rm[ch[sd]];
                    trav[sd] = ov[sd] && ch[sd] < n - 1;
                    if (ov[sd] && ch[sd] >= n - 1) {
                        int64_t oid = tree[4 * ch[sd] + 3];
                        int64_t x = qid < oid ? qid : oid, y = qid < oid ? oid : qid;
                        K++;
                        mtri_f64 a = load_mtri_f64(tris, tris_next, x), b = load_mtri_f64(tris, tris_next, y);
                        if (share_vertex_f64(a.v, b.v)) continue;
                        for (int type = 0; type < 15; ++type) {
                            if (!minimal_overlap_moving_f64(a, b, type)) continue;
                            M++;
                            mtet_f64 t = make_mtet_f64(&a, &b, type);
                            double cf[4], sl[3];
                            cubic_coeffs_f64(&t, cf);
                            solve_cubic_f64(cf, sl);
                            for (int s = 0; s < 3; ++s) {
                                double r = sl[s];
                                if (!(r >= 0 && r <= 1)) continue;
                                V++;
                                if (!filter_collision_f64(&t, type, r)) continue;
                                R++;
                                impulse_one_f64(tris, tris_next, masses, x, y, type, r, pdv[tid], pdx[tid], pcnt[tid]);
                            }
                        }
                    }
                }
                if (!trav[0] && !trav[1]) node = stk[--sp];
                else {
                    node = trav[0] ? ch[0] : ch[1];
                    if (trav[0] && trav[1] && sp < 64) stk[sp++] = ch[1];
                }
            } while (node != -1);
        }
    }
    for (int t = 0; t < nt; ++t) {
        if (!pdv[t]) continue;
        for (int64_t i = 0; i < 9 * n; ++i) { dv[i] += pdv[t][i]; dx[i] += pdx[t][i]; }
        for (int64_t i = 0; i < 3 * n; ++i) cnt[i] += pcnt[t][i];
        free(pdv[t]); free(pdx[t]); free(pcnt[t]);
    }
    free(pdv); free(pdx); free(pcnt); free(rm); free(bb); free(tree);
    if (counts) { counts->candidates = K; counts->subprims = M; counts->roots = V; counts->collisions = R; }
    return 0;
}

/* Vector evaluation of the portable math functions (ccc_libm_ref.h) for the pinning tests:
 * fn 0 cbrt, 1 acos, 2 sin, 3 cos, 4 cbrtf (float argument and result). */
void orc_libm_eval(int fn, const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        switch (fn) {
            case 0: y[i] = orc_cbrt(x[i]); break;
            case 1: y[i] = orc_acos(x[i]); break;
            case 2: y[i] = orc_sin(x[i]); break;
            case 3: y[i] = orc_cos(x[i]); break;
            case 4: y[i] = (double)orc_cbrtf((float)x[i]); break;
            default: y[i] = NAN;
        }
    }
}
