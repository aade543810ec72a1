/*
 * Oracle body, instantiated once per floating type by ccc_oracle.c with
 *   REAL  = double | float
 *   SFX   = _f64   | _f32
 * TEST INFRASTRUCTURE ONLY (see ccc_oracle.h).
 *
 * Parity pinning (what checks each function; the tests live in tests/test_oracle_*.py):
 *   orc_bvh            - compat tree equals the index radix tree rebuilt in Python from the
 *                        textbook Karras definition; Morton codes are all zero (invariant);
 *                        the literal reference search loops equal the textbook search for
 *                        N = 2..3000; boxes equal numpy min/max (exact) and parent boxes equal
 *                        the union of children.  Fitted-box modes (1: 30-bit, 2: 63-bit codes):
 *                        codes equal a bit-by-bit interleave of numpy-quantised centroids,
 *                        the leaf order equals numpy's stable argsort of those codes, and
 *                        every internal node equals the textbook Karras node rebuilt in
 *                        Python over (code, id) keys.
 *   orc_traverse       - equals orc_bruteforce_pairs and an independent numpy O(N^2) pair
 *                        enumeration (exact predicate) as a set.
 *   orc_ccd            - coefficients equal a sympy expansion of the determinant; roots
 *                        agree with mpmath 50-digit roots of the same coefficients; the
 *                        linear-branch drop, quadratic branch, root window (float32 roots
 *                        rounded before the window) and duplicate roots are pinned by hand
 *                        cases; whole pipeline checked on constructed crossings of every
 *                        sub-primitive type at known times of impact.
 *   orc_impulses       - constructed two-triangle scenes for all 15 types, three each,
 *                        with unequal corner masses (tests/golden/ccd_typed_cases.json):
 *                        impulses, corrections and counts from the construction (contact
 *                        weights, normal, normal velocity); a hand-derived vertex-face
 *                        case; EE momentum invariant with equal masses.
 *   orc_proximity      - numpy brute-force distances (float-rounded) on small meshes.
 *   orc_static_collisions - textbook SAT cases (separated, touching, interpenetrating,
 *                        shared vertex) and a numpy SAT re-implementation.
 *   orc_mincol         - numpy re-implementation of the 15-way swept-box filter.
 *   orc_triangle_edge  - hand cases (edge piercing a face at a known parameter, parallel,
 *                        outside, shared vertex loops) and a numpy re-implementation of the
 *                        edge-plane intersection on random soups.
 *   ULP-level behaviour of CUDA libm (cbrt/acos/cos/sin/rsqrt) and the exact rounding of
 *   CAMPARY certifiedMul are parity unpinned (no reference binary or vectors exist); they
 *   are covered by the 1e-9 relative tolerance.
 */

#define V3T CAT(v3, SFX)
#define BOXT CAT(box, SFX)
#define FN(name) CAT(name, SFX)

typedef struct { REAL x, y, z; } V3T;
typedef struct { V3T mn, mx; } BOXT;

/* ---- vector helpers: include/double_vec_ops.h, src/utils.cu:45-93 (same operation order) */
static inline V3T FN(mk)(REAL x, REAL y, REAL z) { V3T r; r.x = x; r.y = y; r.z = z; return r; }
static inline V3T FN(vadd)(V3T a, V3T b) { return FN(mk)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3T FN(vsub)(V3T a, V3T b) { return FN(mk)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3T FN(vmul)(V3T a, REAL s) { return FN(mk)(a.x * s, a.y * s, a.z * s); }
static inline REAL FN(vdot)(V3T a, V3T b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3T FN(vcross)(V3T a, V3T b) {
    return FN(mk)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline REAL FN(vnorm2)(V3T a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
/* utils.cu:91-93 normalize = a * rsqrt(|a|^2); evaluated as 1/sqrt here. */
static inline V3T FN(vnormalize)(V3T a) {
    REAL n2 = a.x * a.x + a.y * a.y + a.z * a.z;
    return FN(vmul)(a, (REAL)1 / RSQRT_BASE(n2));
}
/* utils.cu:59-62 vec_sq_diff returns float. */
static inline float FN(vsqdiff)(V3T a, V3T b) {
    V3T d = FN(vsub)(a, b);
    return (float)FN(vdot)(d, d);
}

static inline V3T FN(load_v3)(const REAL *p) { return FN(mk)(p[0], p[1], p[2]); }

/* ---- AABB: src/aabb.cu:22-69 (CUDA min/max -> fmin/fmax) */
static inline BOXT FN(box_union)(BOXT a, BOXT b) {
    BOXT r;
    r.mn = FN(mk)(RFMIN(a.mn.x, b.mn.x), RFMIN(a.mn.y, b.mn.y), RFMIN(a.mn.z, b.mn.z));
    r.mx = FN(mk)(RFMAX(a.mx.x, b.mx.x), RFMAX(a.mx.y, b.mx.y), RFMAX(a.mx.z, b.mx.z));
    return r;
}
/* src/aabb.cu:101-111 checkOverlap(bbox1, bbox2, threshold) */
static inline int FN(check_overlap)(BOXT b1, BOXT b2, REAL thr) {
    return (b1.mn.x <= (b2.mx.x + thr)) && (b1.mx.x >= (b2.mn.x - thr)) &&
           (b1.mn.y <= (b2.mx.y + thr)) && (b1.mx.y >= (b2.mn.y - thr)) &&
           (b1.mn.z <= (b2.mx.z + thr)) && (b1.mx.z >= (b2.mn.z - thr));
}
static inline BOXT FN(point_box)(V3T v) { BOXT b; b.mn = v; b.mx = v; return b; }
static inline BOXT FN(load_box)(const REAL *p) {
    BOXT b; b.mn = FN(mk)(p[0], p[1], p[2]); b.mx = FN(mk)(p[3], p[4], p[5]); return b;
}

/* ---- Triangle: src/figures.cu:24-70 */
typedef struct { V3T v[3]; } CAT(tri, SFX);
#define TRIT CAT(tri, SFX)
static inline TRIT FN(load_tri)(const REAL *tris, int64_t i) {
    TRIT t; const REAL *p = tris + 9 * i;
    t.v[0] = FN(load_v3)(p); t.v[1] = FN(load_v3)(p + 3); t.v[2] = FN(load_v3)(p + 6);
    return t;
}
/* figures.cu:37-41 Triangle::ComputeBBox */
static inline BOXT FN(tri_box)(TRIT t) {
    BOXT b;
    b.mn = FN(mk)(RFMIN(t.v[0].x, RFMIN(t.v[1].x, t.v[2].x)), RFMIN(t.v[0].y, RFMIN(t.v[1].y, t.v[2].y)),
                  RFMIN(t.v[0].z, RFMIN(t.v[1].z, t.v[2].z)));
    b.mx = FN(mk)(RFMAX(t.v[0].x, RFMAX(t.v[1].x, t.v[2].x)), RFMAX(t.v[0].y, RFMAX(t.v[1].y, t.v[2].y)),
                  RFMAX(t.v[0].z, RFMAX(t.v[1].z, t.v[2].z)));
    return b;
}
/* figures.cu:58-67 static vertex / edge boxes */
static inline BOXT FN(tri_vbox)(TRIT t, int i) { return FN(point_box)(t.v[i]); }
static inline BOXT FN(tri_ebox)(TRIT t, int i, int j) {
    return FN(box_union)(FN(tri_vbox)(t, i), FN(tri_vbox)(t, j));
}

/* ---- MovingTriangle: src/figures.cu:295-366 (dv = v1 - v0; swept boxes use v + dv) */
typedef struct { V3T v[3]; V3T dv[3]; } CAT(mtri, SFX);
#define MTRIT CAT(mtri, SFX)
static inline MTRIT FN(load_mtri)(const REAL *t0, const REAL *t1, int64_t i) {
    MTRIT m; TRIT a = FN(load_tri)(t0, i), b = FN(load_tri)(t1, i);
    for (int k = 0; k < 3; ++k) { m.v[k] = a.v[k]; m.dv[k] = FN(vsub)(b.v[k], a.v[k]); }
    return m;
}
/* figures.cu:346-351 */
static inline BOXT FN(mtri_vbox)(MTRIT t, int i) {
    V3T v = t.v[i], w = FN(vadd)(t.v[i], t.dv[i]);
    BOXT b;
    b.mn = FN(mk)(RFMIN(v.x, w.x), RFMIN(v.y, w.y), RFMIN(v.z, w.z));
    b.mx = FN(mk)(RFMAX(v.x, w.x), RFMAX(v.y, w.y), RFMAX(v.z, w.z));
    return b;
}
/* figures.cu:353-364 */
static inline BOXT FN(mtri_ebox)(MTRIT t, int i, int j) {
    return FN(box_union)(FN(mtri_vbox)(t, i), FN(mtri_vbox)(t, j));
}
static inline BOXT FN(mtri_box)(MTRIT t) {
    return FN(box_union)(FN(box_union)(FN(mtri_vbox)(t, 0), FN(mtri_vbox)(t, 1)), FN(mtri_vbox)(t, 2));
}
static inline int FN(veq)(V3T a, V3T b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
/* figures.cu:527-544 shareVertex (t0 positions, exact equality) */
static inline int FN(share_vertex)(const V3T *a, const V3T *b) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (FN(veq)(a[i], b[j])) return 1;
    return 0;
}
/* figures.cu:547-576 countSharedVerts */
static inline int FN(count_shared)(const V3T *a, const V3T *b) {
    int c = 0;
    for (int i = 0; i < 3; ++i)
        if (FN(veq)(a[i], b[0]) || FN(veq)(a[i], b[1]) || FN(veq)(a[i], b[2])) c++;
    return c;
}

/* ---- 15-way sub-primitive tables: collisions_continuous.cu:46-122, :250-311;
 *      collisions_impulses.cu:30-121, :691-932.  Types 0-8: edge E[t/3] of x vs edge E[t%3]
 *      of y, E = {(0,1),(0,2),(1,2)}; 9-11: face x vs vertex t-9 of y; 12-14: vertex t-12 of x
 *      vs face y.  Tetrahedron corner k comes from (triangle sel[k], corner idx[k]). */
static inline void FN(type_corners)(int type, int *sel, int *idx) {
    if (type < 9) {
        int a = type / 3, b = type % 3;
        sel[0] = 0; idx[0] = ORC_EDGE[a][0]; sel[1] = 0; idx[1] = ORC_EDGE[a][1];
        sel[2] = 1; idx[2] = ORC_EDGE[b][0]; sel[3] = 1; idx[3] = ORC_EDGE[b][1];
    } else if (type < 12) {
        sel[0] = 0; idx[0] = 0; sel[1] = 0; idx[1] = 1; sel[2] = 0; idx[2] = 2;
        sel[3] = 1; idx[3] = type - 9;
    } else {
        sel[0] = 0; idx[0] = type - 12;
        sel[1] = 1; idx[1] = 0; sel[2] = 1; idx[2] = 1; sel[3] = 1; idx[3] = 2;
    }
}

/* collisions_continuous.cu:234-318 check_minimal_bbox_overlap (moving, threshold 0) */
static inline int FN(minimal_overlap_moving)(MTRIT a, MTRIT b, int type) {
    BOXT b0, b1;
    if (type < 9) {
        int ea = type / 3, eb = type % 3;
        b0 = FN(mtri_ebox)(a, ORC_EDGE[ea][0], ORC_EDGE[ea][1]);
        b1 = FN(mtri_ebox)(b, ORC_EDGE[eb][0], ORC_EDGE[eb][1]);
    } else if (type < 12) {
        b0 = FN(mtri_box)(a); b1 = FN(mtri_vbox)(b, type - 9);
    } else {
        b0 = FN(mtri_vbox)(a, type - 12); b1 = FN(mtri_box)(b);
    }
    return FN(check_overlap)(b0, b1, (REAL)0);
}
/* proximity_static.cu:11-98 check_minimal_bbox_overlap (static, threshold) */
static inline int FN(minimal_overlap_static)(TRIT a, TRIT b, int type, REAL thr) {
    BOXT b0, b1;
    if (type < 9) {
        int ea = type / 3, eb = type % 3;
        b0 = FN(tri_ebox)(a, ORC_EDGE[ea][0], ORC_EDGE[ea][1]);
        b1 = FN(tri_ebox)(b, ORC_EDGE[eb][0], ORC_EDGE[eb][1]);
    } else if (type < 12) {
        b0 = FN(tri_box)(a); b1 = FN(tri_vbox)(b, type - 9);
    } else {
        b0 = FN(tri_vbox)(a, type - 12); b1 = FN(tri_box)(b);
    }
    return FN(check_overlap)(b0, b1, thr);
}

/* Moving tetrahedron (figures.cu:74-96), gathered per make_tetrahedron
 * (collisions_continuous.cu:36-124). */
typedef struct { V3T v[4]; V3T dv[4]; } CAT(mtet, SFX);
#define MTETT CAT(mtet, SFX)
static inline MTETT FN(make_mtet)(const MTRIT *a, const MTRIT *b, int type) {
    int sel[4], idx[4];
    FN(type_corners)(type, sel, idx);
    MTETT t;
    for (int k = 0; k < 4; ++k) {
        const MTRIT *s = sel[k] ? b : a;
        t.v[k] = s->v[idx[k]]; t.dv[k] = s->dv[idx[k]];
    }
    return t;
}

/* figures.cu:122-293 MovingTetraMatrix coefficients.  The `coef += x*y*z` accumulations
 * are written as fma(x*y, z, coef), nvcc's default contraction of that statement. */
#define ACC3(c, p, q, r) (c) = RFMA((p) * (q), (r), (c))
static inline void FN(cubic_coeffs)(const MTETT *t, REAL *out) {
    V3T v0 = FN(vsub)(t->v[0], t->v[3]), v1 = FN(vsub)(t->v[1], t->v[3]), v2 = FN(vsub)(t->v[2], t->v[3]);
    V3T dv0 = FN(vsub)(t->dv[0], t->dv[3]), dv1 = FN(vsub)(t->dv[1], t->dv[3]), dv2 = FN(vsub)(t->dv[2], t->dv[3]);
    /* SCALE == 1 (flags.h:57-59): multiplication by 1 is exact and omitted. */
    REAL c3 = 0, c2 = 0, c1 = 0, c0 = 0;
    /* coef3: figures.cu:150-160 */
    ACC3(c3, dv2.z, dv1.y, dv0.x);
    ACC3(c3, -dv1.z, dv2.y, dv0.x);
    ACC3(c3, -dv2.z, dv0.y, dv1.x);
    ACC3(c3, dv0.z, dv2.y, dv1.x);
    ACC3(c3, dv1.z, dv0.y, dv2.x);
    ACC3(c3, -dv0.z, dv1.y, dv2.x);
    /* coef2: figures.cu:226-247 */
    ACC3(c2, v2.z, dv1.y, dv0.x);
    ACC3(c2, dv2.z, v1.y, dv0.x);
    ACC3(c2, dv2.z, dv1.y, v0.x);
    ACC3(c2, -v1.z, dv2.y, dv0.x);
    ACC3(c2, -dv1.z, v2.y, dv0.x);
    ACC3(c2, -dv1.z, dv2.y, v0.x);
    ACC3(c2, -v2.z, dv0.y, dv1.x);
    ACC3(c2, -dv2.z, v0.y, dv1.x);
    ACC3(c2, -dv2.z, dv0.y, v1.x);
    ACC3(c2, v0.z, dv2.y, dv1.x);
    ACC3(c2, dv0.z, v2.y, dv1.x);
    ACC3(c2, dv0.z, dv2.y, v1.x);
    ACC3(c2, v1.z, dv0.y, dv2.x);
    ACC3(c2, dv1.z, v0.y, dv2.x);
    ACC3(c2, dv1.z, dv0.y, v2.x);
    ACC3(c2, -v0.z, dv1.y, dv2.x);
    ACC3(c2, -dv0.z, v1.y, dv2.x);
    ACC3(c2, -dv0.z, dv1.y, v2.x);
    /* coef1: figures.cu:249-270 */
    ACC3(c1, v2.z, v1.y, dv0.x);
    ACC3(c1, v2.z, dv1.y, v0.x);
    ACC3(c1, dv2.z, v1.y, v0.x);
    ACC3(c1, -v1.z, v2.y, dv0.x);
    ACC3(c1, -v1.z, dv2.y, v0.x);
    ACC3(c1, -dv1.z, v2.y, v0.x);
    ACC3(c1, -v2.z, v0.y, dv1.x);
    ACC3(c1, -v2.z, dv0.y, v1.x);
    ACC3(c1, -dv2.z, v0.y, v1.x);
    ACC3(c1, v0.z, v2.y, dv1.x);
    ACC3(c1, v0.z, dv2.y, v1.x);
    ACC3(c1, dv0.z, v2.y, v1.x);
    ACC3(c1, v1.z, v0.y, dv2.x);
    ACC3(c1, v1.z, dv0.y, v2.x);
    ACC3(c1, dv1.z, v0.y, v2.x);
    ACC3(c1, -v0.z, v1.y, dv2.x);
    ACC3(c1, -v0.z, dv1.y, v2.x);
    ACC3(c1, -dv0.z, v1.y, v2.x);
    /* coef0: figures.cu:272-281 */
    ACC3(c0, v2.z, v1.y, v0.x);
    ACC3(c0, -v1.z, v2.y, v0.x);
    ACC3(c0, -v2.z, v0.y, v1.x);
    ACC3(c0, v0.z, v2.y, v1.x);
    ACC3(c0, v1.z, v0.y, v2.x);
    ACC3(c0, -v0.z, v1.y, v2.x);
    out[0] = c3; out[1] = c2; out[2] = c1; out[3] = c0;
}

/* collisions_continuous.cu:448-464 add_root: window [-EPSILON, 1+EPSILON], EPSILON = 1e-16
 * (flags.h:36-38); up to three slots.  The reference's parameter is `T root`, so the double
 * root is rounded to REAL at the call and the window is tested on the rounded value (float:
 * a root in (1, 1 + 2^-24] rounds to 1.0f and is kept). */
static inline void FN(add_root)(REAL *slots, double root, int *n) {
    const REAL r = (REAL)root;
    if ((double)r < -1e-16 || (double)r > 1 + 1e-16) return;
    if (*n < 3) slots[*n] = r;
    (*n)++;
}

/* collisions_continuous.cu:511-613 solveCubic (always double precision); cbrt/acos/sin/cos are
 * the portable versions of ccc_libm_ref.h. */
static void FN(solve_cubic)(const REAL *coef, REAL *slots) {
    double a = coef[0], b = coef[1], c = coef[2], d = coef[3];
    int n = 0;
    slots[0] = slots[1] = slots[2] = (REAL)-1;
    if (a == 0 && b == 0) {
        /* :532-536 the linear root -d/c is computed but never stored. */
        return;
    } else if (fabs(a) < 1e-16) {
        /* :537-547 quadratic branch */
        double D = c * c - 4.0 * b * d;
        if (D >= 0) {
            D = sqrt(D);
            double x1 = (-c + D) / (2.0 * b);
            double x2 = (-c - D) / (2.0 * b);
            FN(add_root)(slots, x1, &n);
            FN(add_root)(slots, x2, &n);
        }
    } else {
        /* :557-561, computeG :472-494, computeH :497-507 */
        double f = ((3.0 * c / a) - ((b * b) / (a * a))) / 3.0;
        double b3 = 2 * orc_dd_cube_lead(b);
        double a2 = a * a;
        double a3 = orc_dd_cube_lead(a);
        double p1 = b3 / a3;
        double p2 = (b * c * 9) / a2;
        double p3 = 27.0 * d / a;
        double g = (p1 - p2 + p3) / 27.0;
        double f3 = orc_dd_cube_lead(f);
        double h = (g * g) / 4.0 + f3 / 27.0;
        if (f == 0 && g == 0 && h == 0) {
            /* :564-575 triple root through a float k */
            float k = (float)(d / a);
            double x1;
            if (k >= 0) x1 = (double)(orc_cbrtf(k) * -1);
            else x1 = (double)orc_cbrtf(-k);
            FN(add_root)(slots, x1, &n);
        } else if (h <= 0) {
            /* :576-591 three real roots, trigonometric form */
            double i = sqrt(((g * g) / 4.0) - h);
            double j = orc_cbrt(i);
            double k = orc_acos(-(g / (2 * i)));
            double L = j * -1;
            double M = orc_cos(k / 3.0);
            double N = sqrt(3.0) * orc_sin(k / 3.0);
            double P = (b / (3.0 * a)) * -1;
            double x1 = 2 * j * orc_cos(k / 3.0) - (b / (3.0 * a));
            double x2 = L * (M + N) + P;
            double x3 = L * (M - N) + P;
            FN(add_root)(slots, x1, &n);
            FN(add_root)(slots, x2, &n);
            FN(add_root)(slots, x3, &n);
        } else {
            /* :592-611 one real root, Cardano */
            double R = -(g / 2.0) + sqrt(h);
            double S, U;
            if (R >= 0) S = orc_cbrt(R); else S = -orc_cbrt(-R);
            double V = -(g / 2.0) - sqrt(h);
            if (V >= 0) U = orc_cbrt(V); else U = -orc_cbrt(-V);
            double x1 = (S + U) - (b / (3.0 * a));
            FN(add_root)(slots, x1, &n);
        }
    }
}

/* collisions_continuous.cu:615-637 is_inside */
static inline int FN(is_inside)(V3T v0, V3T v1, V3T v2, V3T vt) {
    V3T e1 = FN(vsub)(v0, v1), e2 = FN(vsub)(v2, v1), ep = FN(vsub)(vt, v1);
    REAL d00 = FN(vdot)(e1, e1), d01 = FN(vdot)(e1, e2), d11 = FN(vdot)(e2, e2);
    REAL d20 = FN(vdot)(ep, e1), d21 = FN(vdot)(ep, e2);
    REAL denom = d00 * d11 - d01 * d01;
    REAL v = (d11 * d20 - d01 * d21) / denom;
    REAL w = (d00 * d21 - d01 * d20) / denom;
    REAL u = 1 - v - w;
    if (v < 0 || v > 1) return 0;
    if (w < 0 || w > 1) return 0;
    if (u < 0 || u > 1) return 0;
    return 1;   /* NaN passes every test: accepted */
}

/* figures.cu:112-117 Tetrahedron(moving, t) */
static inline void FN(tet_at)(const MTETT *t, REAL root, V3T *p) {
    for (int k = 0; k < 4; ++k) p[k] = FN(vadd)(t->v[k], FN(vmul)(t->dv[k], root));
}

/* collisions_continuous.cu:640-684 filterCollisions */
static inline int FN(filter_collision)(const MTETT *t, int type, REAL root) {
    if (root < -1e-16) return 0;
    V3T p[4];
    FN(tet_at)(t, root, p);
    if (type < 9) {
        V3T da = FN(vsub)(p[1], p[0]), db = FN(vsub)(p[3], p[2]), dc = FN(vsub)(p[2], p[0]);
        V3T cab = FN(vcross)(da, db);
        REAL s = FN(vdot)(FN(vcross)(dc, db), cab) / FN(vnorm2)(cab);
        REAL q = FN(vdot)(FN(vcross)(dc, da), cab) / FN(vnorm2)(cab);
        if (s < 0 || s > 1 || q < 0 || q > 1) return 0;
        return 1;
    } else if (type < 12) {
        return FN(is_inside)(p[0], p[1], p[2], p[3]);
    }
    return FN(is_inside)(p[1], p[2], p[3], p[0]);
}

/* ---------------------------------------------------------------------------------------
 * BVH (src/build_bvh.cu).  Tree rows: internal 0..n-2, leaf n-1+p for sorted position p.
 */
int FN(orc_bvh)(const REAL *tris, const REAL *tris_next, int64_t n, int morton_mode,
                REAL *bboxes, int64_t *tree, uint64_t *codes_out, int32_t *ids_out) {
    if (n < 1) return -1;
    BOXT *box = (BOXT *)malloc(sizeof(BOXT) * (size_t)n);
    uint64_t *codes = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    if (!box || !codes || !ids) { free(box); free(codes); free(ids); return -2; }
    /* ComputeTriBoundingBoxes :212-219 (+ MergeTriBoundingBoxes :452-460 for motion) */
    for (int64_t i = 0; i < n; ++i) {
        box[i] = FN(tri_box)(FN(load_tri)(tris, i));
        if (tris_next) box[i] = FN(box_union)(box[i], FN(tri_box)(FN(load_tri)(tris_next, i)));
    }
    /* scene box :243-244 / :505-506: reduce with init AABB<T>() = [-MAX, +MAX]
     * (aabb.cu:25-33).  morton_mode 1 (30-bit codes) and 2 (63-bit codes), both not the
     * reference, start from an empty box: the scene box fits the leaf boxes. */
    BOXT scene;
    if (morton_mode == 0) {
        scene.mn = FN(mk)(-RMAXV, -RMAXV, -RMAXV); scene.mx = FN(mk)(RMAXV, RMAXV, RMAXV);
    } else {
        scene.mn = FN(mk)(INFINITY, INFINITY, INFINITY); scene.mx = FN(mk)(-INFINITY, -INFINITY, -INFINITY);
    }
    for (int64_t i = 0; i < n; ++i) scene = FN(box_union)(scene, box[i]);
    /* ComputeMortonCodes :26-46 (static) / :48-71 (motion); morton3D bvh.cu:54-62 */
    for (int64_t i = 0; i < n; ++i) {
        TRIT t = FN(load_tri)(tris, i);
        V3T c = FN(mk)((t.v[0].x + t.v[1].x + t.v[2].x) / (REAL)3.0, (t.v[0].y + t.v[1].y + t.v[2].y) / (REAL)3.0,
                       (t.v[0].z + t.v[1].z + t.v[2].z) / (REAL)3.0);
        if (tris_next) {
            TRIT u = FN(load_tri)(tris_next, i);
            V3T c2 = FN(mk)((u.v[0].x + u.v[1].x + u.v[2].x) / (REAL)3.0, (u.v[0].y + u.v[1].y + u.v[2].y) / (REAL)3.0,
                            (u.v[0].z + u.v[1].z + u.v[2].z) / (REAL)3.0);
            c = FN(mk)((c.x + c2.x) / (REAL)2.0, (c.y + c2.y) / (REAL)2.0, (c.z + c2.z) / (REAL)2.0);
        }
        REAL x = (c.x - scene.mn.x) / (scene.mx.x - scene.mn.x);
        REAL y = (c.y - scene.mn.y) / (scene.mx.y - scene.mn.y);
        REAL z = (c.z - scene.mn.z) / (scene.mx.z - scene.mn.z);
        codes[i] = morton_mode == 2 ? MORTON63(x, y, z) : (uint64_t)MORTON(x, y, z);
    }
    /* thrust::sequence + stable sort_by_key :268-279 / :530-541 */
    for (int64_t i = 0; i < n; ++i) ids[i] = (int32_t)i;
    orc_stable_sort_u64(codes, ids, n);
    if (codes_out) memcpy(codes_out, codes, sizeof(uint64_t) * (size_t)n);
    if (ids_out) memcpy(ids_out, ids, sizeof(int32_t) * (size_t)n);

    /* BuildRadixTree :74-143 and flatten :309-378 */
    for (int64_t r = 0; r < 2 * n - 1; ++r) { tree[4 * r + 0] = -1; tree[4 * r + 1] = -1; tree[4 * r + 2] = -1; tree[4 * r + 3] = -1; }
    for (int64_t i = 0; i + 1 < n; ++i) {
        int64_t l, rr;
        orc_karras_node(codes, ids, n, i, &l, &rr);
        tree[4 * i + 1] = l; tree[4 * i + 2] = rr;
        tree[4 * l + 0] = i; tree[4 * rr + 0] = i;
    }
    for (int64_t p = 0; p < n; ++p) tree[4 * (n - 1 + p) + 3] = ids[p];
    /* CreateHierarchy :145-209: leaf box = box[id], parent = union of children.
     * Children of internal i have row indices > i is not guaranteed, so evaluate by
     * post-order with an explicit stack. */
    BOXT *nb = (BOXT *)malloc(sizeof(BOXT) * (size_t)(2 * n - 1));
    for (int64_t p = 0; p < n; ++p) nb[n - 1 + p] = box[ids[p]];
    if (n > 1) {
        int64_t *stk = (int64_t *)malloc(sizeof(int64_t) * (size_t)(4 * n));
        uint8_t *done = (uint8_t *)calloc((size_t)(2 * n - 1), 1);
        int64_t sp = 0;
        stk[sp++] = 0;
        while (sp > 0) {
            int64_t r = stk[sp - 1];
            if (r >= n - 1) { done[r] = 1; sp--; continue; }
            int64_t l = tree[4 * r + 1], rr = tree[4 * r + 2];
            if (!done[l]) { stk[sp++] = l; continue; }
            if (!done[rr]) { stk[sp++] = rr; continue; }
            nb[r] = FN(box_union)(nb[l], nb[rr]);
            done[r] = 1; sp--;
        }
        free(stk); free(done);
    }
    for (int64_t r = 0; r < 2 * n - 1; ++r) {
        REAL *o = bboxes + 6 * r;
        o[0] = nb[r].mn.x; o[1] = nb[r].mn.y; o[2] = nb[r].mn.z;
        o[3] = nb[r].mx.x; o[4] = nb[r].mx.y; o[5] = nb[r].mx.z;
    }
    free(nb); free(box); free(codes); free(ids);
    return 0;
}

/* Reconstruction of rightmost (build_bvh.cu:707-743): rightmost leaf position of each row. */
static int64_t *FN(orc_rightmost)(const int64_t *tree, int64_t n) {
    int64_t m = 2 * n - 1;
    int64_t *rm = (int64_t *)malloc(sizeof(int64_t) * (size_t)m);
    for (int64_t r = 0; r < m; ++r) rm[r] = (r >= n - 1) ? r - (n - 1) : -1;
    /* post-order from the root */
    int64_t *stk = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * m + 2));
    int64_t sp = 0;
    if (n > 1) stk[sp++] = 0;
    while (sp > 0) {
        int64_t r = stk[sp - 1];
        if (r >= n - 1) { sp--; continue; }
        int64_t l = tree[4 * r + 1], rr = tree[4 * r + 2];
        if (rm[l] < 0) { stk[sp++] = l; continue; }
        if (rm[rr] < 0) { stk[sp++] = rr; continue; }
        rm[r] = rm[l] > rm[rr] ? rm[l] : rm[rr];
        sp--;
    }
    free(stk);
    return rm;
}

/* collisions_static.cu:10-120 traverseBVH / findPotentialCollisions (ordering on).
 * Query = leaf at sorted position p with its stored box; a child is skipped when its
 * rightmost leaf position is <= p.  Emits (min id, max id).  Stores the first `cap`
 * pairs (the reference wraps modulo capacity instead) and returns the total. */
int64_t FN(orc_traverse)(const REAL *bboxes, const int64_t *tree, int64_t n, REAL thr,
                         int64_t *pairs, int64_t cap) {
    if (n < 2) return 0;
    int64_t *rm = FN(orc_rightmost)(tree, n);
    int64_t total = 0;
    int64_t stk[64];
    for (int64_t p = 0; p < n; ++p) {
        int64_t qrow = n - 1 + p;
        BOXT q = FN(load_box)(bboxes + 6 * qrow);
        int64_t qid = tree[4 * qrow + 3];
        int sp = 0;
        stk[sp++] = -1;
        int64_t node = 0;
        do {
            int64_t cl = tree[4 * node + 1], cr = tree[4 * node + 2];
            int ol = FN(check_overlap)(q, FN(load_box)(bboxes + 6 * cl), thr);
            int orr = FN(check_overlap)(q, FN(load_box)(bboxes + 6 * cr), thr);
            if (p >= rm[cl]) ol = 0;
            if (p >= rm[cr]) orr = 0;
            int leafl = cl >= n - 1, leafr = cr >= n - 1;
            if (ol && leafl) {
                int64_t oid = tree[4 * cl + 3];
                if (total < cap) { pairs[2 * total] = qid < oid ? qid : oid; pairs[2 * total + 1] = qid < oid ? oid : qid; }
                total++;
            }
            if (orr && leafr) {
                int64_t oid = tree[4 * cr + 3];
                if (total < cap) { pairs[2 * total] = qid < oid ? qid : oid; pairs[2 * total + 1] = qid < oid ? oid : qid; }
                total++;
            }
            int tl = ol && !leafl, tr = orr && !leafr;
            if (!tl && !tr) node = stk[--sp];
            else {
                node = tl ? cl : cr;
                if (tl && tr) {
                    if (sp >= 64) { free(rm); return -1; }
                    stk[sp++] = cr;
                }
            }
        } while (node != -1);
    }
    free(rm);
    return total;
}

/* collisions_continuous.cu:127-230 traverseBVHPartial + :1079-1100: ordering off; only
 * queries whose to_check[position] is set; a leaf child c is reported when
 * !to_check[id(c)] || query id < id(c). */
int64_t FN(orc_traverse_partial)(const REAL *bboxes, const int64_t *tree, int64_t n,
                                 const uint8_t *to_check, int64_t *pairs, int64_t cap) {
    if (n < 2) return 0;
    int64_t total = 0;
    int64_t stk[64];
    for (int64_t p = 0; p < n; ++p) {
        if (!to_check[p]) continue;
        int64_t qrow = n - 1 + p;
        BOXT q = FN(load_box)(bboxes + 6 * qrow);
        int64_t qid = tree[4 * qrow + 3];
        int sp = 0;
        stk[sp++] = -1;
        int64_t node = 0;
        do {
            int64_t cl = tree[4 * node + 1], cr = tree[4 * node + 2];
            int ol = FN(check_overlap)(q, FN(load_box)(bboxes + 6 * cl), (REAL)0);
            int orr = FN(check_overlap)(q, FN(load_box)(bboxes + 6 * cr), (REAL)0);
            int leafl = cl >= n - 1, leafr = cr >= n - 1;
            for (int side = 0; side < 2; ++side) {
                int ov = side ? orr : ol; int lf = side ? leafr : leafl; int64_t ch = side ? cr : cl;
                if (ov && lf) {
                    int64_t oid = tree[4 * ch + 3];
                    if (!to_check[oid] || qid < oid) {
                        if (total < cap) { pairs[2 * total] = qid < oid ? qid : oid; pairs[2 * total + 1] = qid < oid ? oid : qid; }
                        total++;
                    }
                }
            }
            int tl = ol && !leafl, tr = orr && !leafr;
            if (!tl && !tr) node = stk[--sp];
            else {
                node = tl ? cl : cr;
                if (tl && tr) {
                    if (sp >= 64) return -1;
                    stk[sp++] = cr;
                }
            }
        } while (node != -1);
    }
    return total;
}

/* Independent O(N^2) enumeration of the candidate set: pairs of sorted positions p < q
 * whose stored leaf boxes overlap with p's box as the query (aabb.cu:101-111). */
int64_t FN(orc_bruteforce_pairs)(const REAL *bboxes, const int64_t *tree, int64_t n, REAL thr,
                                 int64_t *pairs, int64_t cap) {
    int64_t total = 0;
    for (int64_t p = 0; p < n; ++p) {
        BOXT a = FN(load_box)(bboxes + 6 * (n - 1 + p));
        int64_t ida = tree[4 * (n - 1 + p) + 3];
        for (int64_t q = p + 1; q < n; ++q) {
            if (!FN(check_overlap)(a, FN(load_box)(bboxes + 6 * (n - 1 + q)), thr)) continue;
            int64_t idb = tree[4 * (n - 1 + q) + 3];
            if (total < cap) { pairs[2 * total] = ida < idb ? ida : idb; pairs[2 * total + 1] = ida < idb ? idb : ida; }
            total++;
        }
    }
    return total;
}

int FN(orc_cubic_coeffs)(const REAL *tet, REAL *coef) {
    MTETT t;
    for (int k = 0; k < 4; ++k) { t.v[k] = FN(load_v3)(tet + 3 * k); t.dv[k] = FN(load_v3)(tet + 12 + 3 * k); }
    FN(cubic_coeffs)(&t, coef);
    return 0;
}
int FN(orc_solve_cubic)(const REAL *coef, REAL *roots3) {
    FN(solve_cubic)(coef, roots3);
    return 0;
}

/* Narrow phase of find_collisions_from_bbox_tree (collisions_continuous.cu:887-1063):
 * per candidate pair and type: shareVertex reject + swept sub-primitive box filter
 * (:234-318, :340-396), tetrahedron (:36-124), coefficients + solveCubic (:409-613),
 * flatten of roots in [0,1] (:732-765), filterCollisions (:640-684).  Emits
 * (x, y, type) rows and roots; stores the first `cap` (truncation, :1047-1051). */
int FN(orc_ccd)(const REAL *tris, const REAL *tris_next, int64_t n, const int64_t *pairs,
                int64_t k, int64_t *coll, REAL *roots, int64_t cap, orc_counts *counts) {
    (void)n;
    int64_t M = 0, V = 0, R = 0;
    for (int64_t c = 0; c < k; ++c) {
        int64_t x = pairs[2 * c], y = pairs[2 * c + 1];
        MTRIT a = FN(load_mtri)(tris, tris_next, x), b = FN(load_mtri)(tris, tris_next, y);
        if (FN(share_vertex)(a.v, b.v)) continue;
        for (int type = 0; type < 15; ++type) {
            if (!FN(minimal_overlap_moving)(a, b, type)) continue;
            M++;
            MTETT t = FN(make_mtet)(&a, &b, type);
            REAL cf[4], sl[3];
            FN(cubic_coeffs)(&t, cf);
            FN(solve_cubic)(cf, sl);
            for (int s = 0; s < 3; ++s) {
                REAL r = sl[s];
                if (!(r >= 0 && r <= 1)) continue;
                V++;
                if (!FN(filter_collision)(&t, type, r)) continue;
                if (R < cap) {
                    coll[3 * R] = x; coll[3 * R + 1] = y; coll[3 * R + 2] = type;
                    roots[R] = r;
                }
                R++;
            }
        }
    }
    if (counts) { counts->candidates = k; counts->subprims = M; counts->roots = V; counts->collisions = R; }
    return 0;
}

/* Impulse scatter for one collision: collisions_impulses.cu:24-123 (mass quad),
 * :162-543 (coefficients, normal, normal velocity, mass pair), :609-1067 (extrapolate). */
static void FN(impulse_one)(const REAL *tris, const REAL *tris_next, const REAL *masses,
                            int64_t x, int64_t y, int type, REAL root,
                            REAL *dv, REAL *dx, int32_t *cnt) {
    MTRIT a = FN(load_mtri)(tris, tris_next, x), b = FN(load_mtri)(tris, tris_next, y);
    MTETT t = FN(make_mtet)(&a, &b, type);
    int sel[4], idx[4];
    FN(type_corners)(type, sel, idx);
    REAL q[4];
    for (int kk = 0; kk < 4; ++kk) q[kk] = masses[3 * (sel[kk] ? y : x) + idx[kk]];
    V3T p[4];
    FN(tet_at)(&t, root, p);
    V3T nrm; REAL nvel; REAL mpx, mpy;
    if (type < 9) {
        /* compute_coeficients_edge_edge :306-328 */
        V3T da = FN(vsub)(p[1], p[0]), db = FN(vsub)(p[3], p[2]), dc = FN(vsub)(p[2], p[0]);
        REAL denom = FN(vnorm2)(FN(vcross)(da, db));
        REAL s, tt;
        if (RFABS(denom) < 1e-15) { s = (REAL)0.5; tt = (REAL)0.5; }
        else {
            s = FN(vdot)(FN(vcross)(dc, db), FN(vcross)(da, db)) / FN(vnorm2)(FN(vcross)(da, db));
            tt = FN(vdot)(FN(vcross)(dc, da), FN(vcross)(da, db)) / FN(vnorm2)(FN(vcross)(da, db));
        }
        /* compute_normal_direction_edge_edge :330-372 */
        V3T p0 = FN(vadd)(t.v[0], FN(vmul)(FN(vsub)(t.v[1], t.v[0]), s));
        V3T p1 = FN(vadd)(t.v[2], FN(vmul)(FN(vsub)(t.v[3], t.v[2]), tt));
        V3T dp0 = FN(vadd)(t.dv[0], FN(vmul)(FN(vsub)(t.dv[1], t.dv[0]), s));
        V3T dp1 = FN(vadd)(t.dv[2], FN(vmul)(FN(vsub)(t.dv[3], t.dv[2]), tt));
        V3T c0 = FN(vadd)(t.v[0], FN(vmul)(t.dv[0], root));
        V3T c1 = FN(vadd)(t.v[1], FN(vmul)(t.dv[1], root));
        V3T c2 = FN(vadd)(t.v[2], FN(vmul)(t.dv[2], root));
        nrm = FN(vnormalize)(FN(vcross)(FN(vsub)(c1, c0), FN(vsub)(c2, c0)));
        if (FN(vdot)(FN(vsub)(p1, p0), nrm) < 0) nrm = FN(vmul)(nrm, (REAL)-1);
        nvel = FN(vdot)(FN(vsub)(dp1, dp0), nrm);
        /* compute_masses_edge_edge :449-454 */
        mpx = q[0] * (1 - s) + q[1] * s;
        mpy = q[2] * (1 - tt) + q[3] * tt;
    } else {
        /* vertex-face (9-11) or face-vertex with inverse (12-14): :503-535 */
        int inv = type >= 12;
        int i0 = inv ? 1 : 0, i1 = inv ? 2 : 1, i2 = inv ? 3 : 2, it = inv ? 0 : 3;
        /* compute_coeficients_triangle_vertex :162-193 */
        V3T e1 = FN(vsub)(p[i0], p[i1]), e2 = FN(vsub)(p[i2], p[i1]), ep = FN(vsub)(p[it], p[i1]);
        REAL d00 = FN(vdot)(e1, e1), d01 = FN(vdot)(e1, e2), d11 = FN(vdot)(e2, e2);
        REAL d20 = FN(vdot)(ep, e1), d21 = FN(vdot)(ep, e2);
        REAL denom = d00 * d11 - d01 * d01;
        REAL u, w, v;
        if (RFABS(denom) < 1e-15) { u = (REAL)(1. / 3.); w = (REAL)(1. / 3.); }
        else { u = (d11 * d20 - d01 * d21) / denom; w = (d00 * d21 - d01 * d20) / denom; }
        v = 1 - u - w;
        /* compute_normal_direction_triangle_vertex :196-241 */
        V3T pt = FN(vadd)(FN(vadd)(FN(vmul)(t.v[i0], u), FN(vmul)(t.v[i1], v)), FN(vmul)(t.v[i2], w));
        V3T dpt = FN(vadd)(FN(vadd)(FN(vmul)(t.dv[i0], u), FN(vmul)(t.dv[i1], v)), FN(vmul)(t.dv[i2], w));
        V3T c0 = FN(vadd)(t.v[i0], FN(vmul)(t.dv[i0], root));
        V3T c1 = FN(vadd)(t.v[i1], FN(vmul)(t.dv[i1], root));
        V3T c2 = FN(vadd)(t.v[i2], FN(vmul)(t.dv[i2], root));
        nrm = FN(vnormalize)(FN(vcross)(FN(vsub)(c1, c0), FN(vsub)(c2, c0)));
        if (FN(vdot)(FN(vsub)(t.v[it], pt), nrm) < 0) nrm = FN(vmul)(nrm, (REAL)-1);
        nvel = FN(vdot)(FN(vsub)(t.dv[it], dpt), nrm);
        if (inv) nrm = FN(vmul)(nrm, (REAL)-1);
        /* compute_masses_triangle_vertex :456-468 */
        if (inv) { mpx = q[0]; mpy = q[1] * u + q[2] * v + q[3] * w; }
        else { mpx = q[0] * u + q[1] * v + q[2] * w; mpy = q[3]; }
    }
    /* extrapolate_impulses :936-1050; THICKNESS = 2e-3 (:12-14) */
    REAL offset;
    if (RFABS(nvel) < 1e-15) offset = (REAL)0.1;
    else offset = (REAL)(2 * 2e-3 / -nvel);
    offset = (REAL)fmin(fmax((double)offset, 0.), 1.);
    REAL msum = mpx + mpy;
    REAL mf0, mf1;
    if (RFABS(msum) < 1e-15) { mf0 = (REAL)0.5; mf1 = (REAL)0.5; }
    else { mf0 = mpx / msum; mf1 = 1 - mf0; }
    REAL im0 = nvel * mf1;
    REAL im1 = -nvel * mf0;
    V3T imp[4];
    for (int kk = 0; kk < 4; ++kk) {
        int side;   /* 0 -> magnitude im0, 1 -> im1 (:1017-1032) */
        if (type < 9) side = kk >= 2;
        else if (type < 12) side = kk == 3;
        else side = kk >= 1;
        imp[kk] = FN(vmul)(nrm, side ? im1 : im0);
    }
    /* extrapolate_impulses_dv :609-933 */
    REAL root_offset = (REAL)fmax((double)(root - offset), 0.);
    for (int kk = 0; kk < 4; ++kk) {
        V3T nvk = FN(vadd)(t.dv[kk], imp[kk]);
        V3T idx_dx = FN(vmul)(FN(vsub)(nvk, t.dv[kk]), (1 - root_offset));
        int64_t tri = sel[kk] ? y : x;
        int corner = idx[kk];
        REAL *o = dv + 9 * tri + 3 * corner;
        o[0] += imp[kk].x; o[1] += imp[kk].y; o[2] += imp[kk].z;
        REAL *od = dx + 9 * tri + 3 * corner;
        od[0] += idx_dx.x; od[1] += idx_dx.y; od[2] += idx_dx.z;
        cnt[3 * tri + corner] += 1;
    }
}

/* collision_impulses_from_continuous_collisions (collisions_impulses.cu:1736-1880) on a
 * list of collisions; rows with a negative index are skipped (the reference reads them
 * out of bounds, :1793-1795).  dv, dx, cnt are accumulated (caller zeroes them). */
int FN(orc_impulses)(const REAL *tris, const REAL *tris_next, const REAL *masses, int64_t n,
                     const int64_t *coll, const REAL *roots, int64_t r, REAL *dv, REAL *dx,
                     int32_t *cnt) {
    for (int64_t i = 0; i < r; ++i) {
        int64_t x = coll[3 * i], y = coll[3 * i + 1], type = coll[3 * i + 2];
        if (x < 0 || y < 0 || type < 0 || type > 14 || x >= n || y >= n) continue;
        FN(impulse_one)(tris, tris_next, masses, x, y, (int)type, roots[i], dv, dx, cnt);
    }
    return 0;
}

/* proximity_static.cu:205-228 distance_point_to_segment */
static inline void FN(pt_seg)(V3T p, V3T a, V3T b, REAL *dac, int set_s, REAL second) {
    V3T ab = FN(vsub)(b, a), ap = FN(vsub)(p, a);
    REAL t = FN(vdot)(ap, ab) / FN(vdot)(ab, ab);
    if (t < 0.0) t = 0.0;
    if (t > 1.0) t = 1.0;
    V3T c = FN(mk)(a.x + t * ab.x, a.y + t * ab.y, a.z + t * ab.z);
    REAL nd = (REAL)sqrtf(FN(vsqdiff)(p, c));
    if (nd < dac[0]) {
        dac[0] = nd;
        if (set_s) { dac[1] = t; dac[2] = second; }
        else { dac[2] = t; dac[1] = second; }
    }
}

/* proximity_static.cu:239-332 compute_distance_edge_edge (live code only) */
static inline REAL FN(dist_ee)(V3T v0, V3T v1, V3T v2, V3T v3) {
    V3T da = FN(vsub)(v1, v0), db = FN(vsub)(v3, v2), dc = FN(vsub)(v2, v0);
    V3T cab = FN(vcross)(da, db);
    REAL s = FN(vdot)(FN(vcross)(dc, db), cab) / FN(vnorm2)(cab);
    REAL t = FN(vdot)(FN(vcross)(dc, da), cab) / FN(vnorm2)(cab);
    if (isnan(s) || isnan(t)) {
        REAL dac[3] = {(REAL)FLT_MAX, (REAL)-1., (REAL)-1.};
        /* compute_distance_parallel_edges :230-237 */
        FN(pt_seg)(v0, v2, v3, dac, 0, (REAL)0.);
        FN(pt_seg)(v1, v2, v3, dac, 0, (REAL)1.);
        FN(pt_seg)(v2, v0, v1, dac, 1, (REAL)0.);
        FN(pt_seg)(v3, v0, v1, dac, 1, (REAL)1.);
        s = dac[1]; t = dac[2];
    }
    int sv = (s >= 0 && s <= 1), tv = (t >= 0 && t <= 1);
    if (!sv || !tv) return (REAL)-1;
    return (REAL)sqrtf(FN(vsqdiff)(FN(vadd)(v0, FN(vmul)(da, s)), FN(vadd)(v2, FN(vmul)(db, t))));
}

/* proximity_static.cu:334-433 compute_distance_triangle_node (live code only) */
static inline REAL FN(dist_vf)(V3T v0, V3T v1, V3T v2, V3T vt) {
    V3T e1 = FN(vsub)(v0, v1), e2 = FN(vsub)(v2, v1), ep = FN(vsub)(vt, v1);
    REAL d00 = FN(vdot)(e1, e1), d01 = FN(vdot)(e1, e2), d11 = FN(vdot)(e2, e2);
    REAL d20 = FN(vdot)(ep, e1), d21 = FN(vdot)(ep, e2);
    REAL denom = d00 * d11 - d01 * d01;
    REAL u = (d11 * d20 - d01 * d21) / denom;
    REAL w = (d00 * d21 - d01 * d20) / denom;
    REAL v = 1 - w - u;
    int ok = (v >= 0 && v <= 1) && (w >= 0 && w <= 1) && (u >= 0 && u <= 1);
    if (!ok) return (REAL)-1;
    V3T pt = FN(vadd)(FN(vadd)(FN(vmul)(v0, u), FN(vmul)(v1, v)), FN(vmul)(v2, w));
    return (REAL)sqrtf(FN(vsqdiff)(vt, pt));
}

/* find_proximity_from_bbox_tree narrow phase (proximity_static.cu:590-685):
 * static sub-primitive box filter with threshold, distances, is_close
 * (0 <= d <= threshold, :468), rows (x, y, type). */
int FN(orc_proximity)(const REAL *tris, int64_t n, const int64_t *pairs, int64_t k, REAL thr,
                      int64_t *out, int64_t cap, orc_counts *counts) {
    (void)n;
    int64_t M = 0, R = 0;
    for (int64_t c = 0; c < k; ++c) {
        int64_t x = pairs[2 * c], y = pairs[2 * c + 1];
        TRIT a = FN(load_tri)(tris, x), b = FN(load_tri)(tris, y);
        if (FN(share_vertex)(a.v, b.v)) continue;
        for (int type = 0; type < 15; ++type) {
            if (!FN(minimal_overlap_static)(a, b, type, thr)) continue;
            M++;
            int sel[4], idx[4];
            FN(type_corners)(type, sel, idx);
            V3T p[4];
            for (int kk = 0; kk < 4; ++kk) p[kk] = (sel[kk] ? b : a).v[idx[kk]];
            REAL d;
            if (type < 9) d = FN(dist_ee)(p[0], p[1], p[2], p[3]);
            else if (type < 12) d = FN(dist_vf)(p[0], p[1], p[2], p[3]);
            else d = FN(dist_vf)(p[1], p[2], p[3], p[0]);
            if (d <= thr && d >= 0) {
                if (R < cap) { out[3 * R] = x; out[3 * R + 1] = y; out[3 * R + 2] = type; }
                R++;
            }
        }
    }
    if (counts) { counts->candidates = k; counts->subprims = M; counts->roots = 0; counts->collisions = R; }
    return 0;
}

/* figures.cu:378-427 projections and interval overlap; figures.cu:446-509 17-axis SAT */
static inline void FN(interval)(V3T ax, TRIT t, REAL *lo, REAL *hi) {
    REAL d = FN(vdot)(ax, t.v[0]);
    *lo = d; *hi = d;
    d = FN(vdot)(ax, t.v[1]); *lo = RFMIN(*lo, d); *hi = RFMAX(*hi, d);
    d = FN(vdot)(ax, t.v[2]); *lo = RFMIN(*lo, d); *hi = RFMAX(*hi, d);
}
static inline int FN(sat)(TRIT a, TRIT b, REAL thr) {
    V3T a0 = FN(vsub)(a.v[1], a.v[0]), a1 = FN(vsub)(a.v[2], a.v[0]), a2 = FN(vsub)(a.v[2], a.v[1]);
    V3T an = FN(vcross)(a1, a2);
    V3T b0 = FN(vsub)(b.v[1], b.v[0]), b1 = FN(vsub)(b.v[2], b.v[0]), b2 = FN(vsub)(b.v[2], b.v[1]);
    V3T bn = FN(vcross)(b1, b2);
    V3T ae[3] = {a0, a1, a2}, be[3] = {b0, b1, b2};
    V3T axes[17];
    axes[0] = an; axes[1] = bn;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) axes[2 + 3 * i + j] = FN(vcross)(ae[i], be[j]);
    for (int i = 0; i < 3; ++i) axes[11 + i] = FN(vcross)(an, ae[i]);
    for (int i = 0; i < 3; ++i) axes[14 + i] = FN(vcross)(an, be[i]);
    int flag = 1;
    for (int i = 0; i < 17; ++i) {
        REAL l1, h1, l2, h2;
        FN(interval)(axes[i], a, &l1, &h1);
        FN(interval)(axes[i], b, &l2, &h2);
        flag = flag && ((l1 <= h2 + thr) && (h1 >= l2 - thr));
    }
    return flag;
}

/* collisions_static.cu:123-159 checkTriangleIntersections (allow_shared_vertices = 0) */
int64_t FN(orc_static_collisions)(const REAL *tris, int64_t n, const int64_t *pairs, int64_t k,
                                  REAL thr, int64_t *out, int64_t cap) {
    (void)n;
    int64_t R = 0;
    for (int64_t c = 0; c < k; ++c) {
        int64_t x = pairs[2 * c], y = pairs[2 * c + 1];
        TRIT a = FN(load_tri)(tris, x), b = FN(load_tri)(tris, y);
        int hit = FN(sat)(a, b, thr);
        int shared = FN(count_shared)(a.v, b.v);
        if (hit && !(shared > 0)) {
            if (R < cap) { out[2 * R] = x; out[2 * R + 1] = y; }
            R++;
        }
    }
    return R;
}

/* figures.cu:605-610 hasVertex(tri, v): exact coordinate match with one of the corners */
static inline int FN(has_vertex)(TRIT t, V3T v) {
    return FN(veq)(t.v[0], v) || FN(veq)(t.v[1], v) || FN(veq)(t.v[2], v);
}

/* collisions_static.cu:345-442 addTECollsion: edge (e0, e1) of one triangle against the
 * plane and barycentric interior of `t`.  Returns 1 and sets *loop / *coef when a row is
 * emitted.  Both the parametric t and the barycentric test let NaN through, as in the
 * reference (every comparison with NaN is false). */
static inline int FN(te_intersect)(TRIT t, V3T e0, V3T e1, int cid, int64_t *loop, REAL *coef) {
    V3T edge1 = FN(vsub)(t.v[1], t.v[0]), edge2 = FN(vsub)(t.v[2], t.v[0]);
    V3T nrm = FN(vnormalize)(FN(vcross)(edge1, edge2));
    int s0 = FN(has_vertex)(t, e0), s1 = FN(has_vertex)(t, e1);
    if (s0 || s1) {                         /* :385-405 loop node: only cid < 3 with e0 shared */
        if (cid > 2 || !s0) return 0;
        *loop = 0;
        *coef = (REAL)(float)0.0;
        return 1;
    }
    V3T dir = FN(vsub)(e1, e0);
    REAL en = FN(vdot)(dir, nrm);
    if (en == (REAL)0) return 0;            /* :410-414 parallel */
    REAL d = FN(vdot)(nrm, t.v[0]);
    REAL tp = (d - FN(vdot)(nrm, e0)) / en;
    if (tp < (REAL)0 || tp > (REAL)1) return 0;
    V3T ip = FN(vadd)(e0, FN(vmul)(dir, tp));
    V3T ep = FN(vsub)(ip, t.v[0]);
    REAL d00 = FN(vdot)(edge1, edge1), d01 = FN(vdot)(edge1, edge2), d11 = FN(vdot)(edge2, edge2);
    REAL d20 = FN(vdot)(ep, edge1), d21 = FN(vdot)(ep, edge2);
    REAL denom = d00 * d11 - d01 * d01;
    if (denom < (REAL)1e-15) return 0;
    REAL v = (d11 * d20 - d01 * d21) / denom;
    REAL w = (d00 * d21 - d01 * d20) / denom;
    REAL u = (REAL)1.0 - v - w;
    if (u < (REAL)0 || v < (REAL)0 || w < (REAL)0) return 0;
    *loop = -1;
    *coef = tp;
    return 1;
}

/* Triangle-edge queries over candidate pairs (x, y):
 *   mode 0 find_triangle_edge_collisions  (collisions_static.cu:445-500, :612-779):
 *          pairs that pass SAT with no shared vertex; rows (x, y, cid, loop) + coefficient;
 *   mode 1 find_triangle_edge_candidates  (collisions_static.cu:520-604, :782-943):
 *          pairs that pass SAT with at most one shared vertex; rows (x, y, cid, loop);
 *   mode 2 find_triangle_edge_candidates2 (icontour.cu:11-108, :110-282): the same pairs,
 *          one row (x, y) each with bit masks (mx, my).
 * cid 0..2 are the edges (v0v1, v1v2, v2v0) of x against y, 3..5 those of y against x.
 * The candidate filter's box test compares the second triangle with the box of its own
 * edge (collisions_static.cu:594-596), which always holds for finite input; it is kept as
 * written. */
int64_t FN(orc_triangle_edge)(const REAL *tris, int64_t n, const int64_t *pairs, int64_t k, int mode,
                              int64_t *rows, REAL *coefs, int32_t *masks, int64_t cap) {
    (void)n;
    int64_t R = 0;
    const int allow = mode == 0 ? 0 : 1;
    for (int64_t c = 0; c < k; ++c) {
        int64_t x = pairs[2 * c], y = pairs[2 * c + 1];
        TRIT a = FN(load_tri)(tris, x), b = FN(load_tri)(tris, y);
        if (!(FN(sat)(a, b, (REAL)0) && !(FN(count_shared)(a.v, b.v) > allow))) continue;
        int32_t mx = 0, my = 0;
        for (int cid = 0; cid < 6; ++cid) {
            TRIT t1 = cid < 3 ? a : b, t2 = cid < 3 ? b : a;
            int e = cid % 3;
            V3T e0 = t1.v[e], e1 = t1.v[(e + 1) % 3];
            int64_t loop = -1;
            REAL coef = (REAL)-1;
            int emit = 0;
            if (mode == 0) {
                emit = FN(te_intersect)(t2, e0, e1, cid, &loop, &coef);
            } else {
                int s0 = FN(has_vertex)(t2, e0), s1 = FN(has_vertex)(t2, e1);
                if (s0) {
                    if (cid < 3) { emit = 1; loop = 0; mx |= 1 << cid; my |= 1 << cid; }
                } else if (!s1) {
                    BOXT b0 = FN(tri_box)(t2), b1 = FN(tri_ebox)(t2, e, (e + 1) % 3);
                    if (FN(check_overlap)(b0, b1, (REAL)0)) { emit = 1; mx |= 1 << cid; }
                }
            }
            if (emit && mode != 2) {
                if (R < cap) {
                    rows[4 * R] = x; rows[4 * R + 1] = y; rows[4 * R + 2] = cid; rows[4 * R + 3] = loop;
                    if (coefs) coefs[R] = coef;
                }
                R++;
            }
        }
        if (mode == 2) {
            if (R < cap) { rows[2 * R] = x; rows[2 * R + 1] = y; masks[2 * R] = mx; masks[2 * R + 1] = my; }
            R++;
        }
    }
    return R;
}

/* find_minimal_candidates_from_bbox_tree narrow step (collisions_continuous.cu:321-338,
 * :1379-1519): (x, y, type) rows of the 15-way swept-box filter. */
int64_t FN(orc_mincol)(const REAL *tris, const REAL *tris_next, int64_t n, const int64_t *pairs,
                       int64_t k, int64_t *out, int64_t cap) {
    (void)n;
    int64_t M = 0;
    for (int64_t c = 0; c < k; ++c) {
        int64_t x = pairs[2 * c], y = pairs[2 * c + 1];
        MTRIT a = FN(load_mtri)(tris, tris_next, x), b = FN(load_mtri)(tris, tris_next, y);
        if (FN(share_vertex)(a.v, b.v)) continue;
        for (int type = 0; type < 15; ++type) {
            if (!FN(minimal_overlap_moving)(a, b, type)) continue;
            if (M < cap) { out[3 * M] = x; out[3 * M + 1] = y; out[3 * M + 2] = type; }
            M++;
        }
    }
    return M;
}

#undef V3T
#undef BOXT
#undef FN
#undef TRIT
#undef MTRIT
#undef MTETT
#undef ACC3
