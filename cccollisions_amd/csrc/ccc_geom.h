// Device geometry for the collision hot path (gfx950).
//
// Every expression keeps the operation order of the reference (src/figures.cu,
// src/utils.cu, include/double_vec_ops.h, src/collisions_*.cu); the library is compiled
// with -ffp-contract=off and contractions are written out as fma() only where the
// reference's compiler contracts a `coef += x*y*z` accumulation (figures.cu:150-281).
// CUDA's min/max on floating point are fmin/fmax (NaN-ignoring) and are written so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ccc_libm.h"

namespace ccc {

template <typename T> struct V3 { T x, y, z; };
template <typename T> struct Box { V3<T> mn, mx; };

__device__ __forceinline__ double tmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ double tmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float tmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float tmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double tfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float tfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double tabs(double a) { return fabs(a); }
__device__ __forceinline__ float tabs(float a) { return fabsf(a); }
__device__ __forceinline__ double tsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float tsqrt(float a) { return sqrtf(a); }

template <typename T> __device__ __forceinline__ V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <typename T> __device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T> __device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T> __device__ __forceinline__ V3<T> operator*(V3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <typename T> __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> __device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename T> __device__ __forceinline__ T norm2(V3<T> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
// utils.cu:91-93 normalize(a) = a * rsqrt(|a|^2), evaluated as a * (1 / sqrt(|a|^2)).
template <typename T> __device__ __forceinline__ V3<T> normalize(V3<T> a) {
    T n2 = a.x * a.x + a.y * a.y + a.z * a.z;
    return a * ((T)1 / tsqrt(n2));
}
// utils.cu:59-62 vec_sq_diff: the squared distance is returned as float.
template <typename T> __device__ __forceinline__ float sq_diff(V3<T> a, V3<T> b) {
    V3<T> d = a - b;
    return (float)dot(d, d);
}
template <typename T> __device__ __forceinline__ bool veq(V3<T> a, V3<T> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

template <typename T> __device__ __forceinline__ V3<T> ld3(const T *p) { return mk(p[0], p[1], p[2]); }

// aabb.cu:51-56 union
template <typename T> __device__ __forceinline__ Box<T> unite(Box<T> a, Box<T> b) {
    return Box<T>{mk(tmin(a.mn.x, b.mn.x), tmin(a.mn.y, b.mn.y), tmin(a.mn.z, b.mn.z)),
                  mk(tmax(a.mx.x, b.mx.x), tmax(a.mx.y, b.mx.y), tmax(a.mx.z, b.mx.z))};
}
// aabb.cu:101-111 checkOverlap(bbox1, bbox2, threshold)
template <typename T> __device__ __forceinline__ bool overlap(const Box<T> &b1, const Box<T> &b2, T thr) {
    return (b1.mn.x <= (b2.mx.x + thr)) && (b1.mx.x >= (b2.mn.x - thr)) &&
           (b1.mn.y <= (b2.mx.y + thr)) && (b1.mx.y >= (b2.mn.y - thr)) &&
           (b1.mn.z <= (b2.mx.z + thr)) && (b1.mx.z >= (b2.mn.z - thr));
}
template <typename T> __device__ __forceinline__ Box<T> ld_box(const T *p) {
    return Box<T>{mk(p[0], p[1], p[2]), mk(p[3], p[4], p[5])};
}
template <typename T> __device__ __forceinline__ Box<T> point_box(V3<T> v) { return Box<T>{v, v}; }
// figures.cu:37-41 Triangle::ComputeBBox
template <typename T> __device__ __forceinline__ Box<T> tri_box(V3<T> a, V3<T> b, V3<T> c) {
    return Box<T>{mk(tmin(a.x, tmin(b.x, c.x)), tmin(a.y, tmin(b.y, c.y)), tmin(a.z, tmin(b.z, c.z))),
                  mk(tmax(a.x, tmax(b.x, c.x)), tmax(a.y, tmax(b.y, c.y)), tmax(a.z, tmax(b.z, c.z)))};
}

// figures.cu:527-544 shareVertex: exact coincidence of any t0 vertices.
template <typename T> __device__ __forceinline__ bool share_vertex(const V3<T> *a, const V3<T> *b) {
    bool s = false;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s = s || veq(a[i], b[j]);
    return s;
}
// figures.cu:547-576 countSharedVerts
template <typename T> __device__ __forceinline__ int count_shared(const V3<T> *a, const V3<T> *b) {
    int c = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) c += (veq(a[i], b[0]) || veq(a[i], b[1]) || veq(a[i], b[2])) ? 1 : 0;
    return c;
}

// The 15 sub-primitive types (collisions_continuous.cu:46-122 / :250-311,
// collisions_impulses.cu:30-121 / :691-932).  0-8: edge E[t/3] of x vs edge E[t%3] of y
// with E = {(0,1),(0,2),(1,2)}; 9-11: face of x vs vertex t-9 of y; 12-14: vertex t-12 of
// x vs face of y.  corner(t, k) -> (triangle: 0 = x / 1 = y, corner index).
__device__ __forceinline__ int edge_first(int e) { return e == 2 ? 1 : 0; }
__device__ __forceinline__ int edge_second(int e) { return e == 0 ? 1 : 2; }
__device__ __forceinline__ void type_corner(int type, int k, int &side, int &idx) {
    if (type < 9) {
        int e = k < 2 ? type / 3 : type % 3;
        side = k < 2 ? 0 : 1;
        idx = (k & 1) ? edge_second(e) : edge_first(e);
    } else if (type < 12) {
        side = k < 3 ? 0 : 1;
        idx = k < 3 ? k : type - 9;
    } else {
        side = k < 1 ? 0 : 1;
        idx = k < 1 ? type - 12 : k - 1;
    }
}

// Static sub-primitive boxes (figures.cu:58-67; proximity_static.cu:11-98).
template <typename T> __device__ __forceinline__ Box<T> tri_edge_box(const V3<T> *v, int e) {
    return unite(point_box(v[edge_first(e)]), point_box(v[edge_second(e)]));
}

// Same filter for static triangles with a threshold (proximity_static.cu:11-98).
template <typename T> __device__ __forceinline__ uint32_t static_mask(const V3<T> *a, const V3<T> *b, T thr) {
    Box<T> ea[3], eb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) { ea[i] = tri_edge_box(a, i); eb[i] = tri_edge_box(b, i); }
    Box<T> fa = tri_box(a[0], a[1], a[2]), fb = tri_box(b[0], b[1], b[2]);
    uint32_t m = 0;
#pragma unroll
    for (int t = 0; t < 9; ++t) m |= overlap(ea[t / 3], eb[t % 3], thr) ? (1u << t) : 0u;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        m |= overlap(fa, point_box(b[k]), thr) ? (1u << (9 + k)) : 0u;
        m |= overlap(point_box(a[k]), fb, thr) ? (1u << (12 + k)) : 0u;
    }
    return m;
}

// Moving tetrahedron (figures.cu:74-96) of one (x, y, type) pair, gathered from the soups.
template <typename T> struct MTet { V3<T> v[4]; V3<T> dv[4]; };

template <typename T>
__device__ __forceinline__ MTet<T> gather_mtet(const T *t0, const T *t1, int64_t x, int64_t y, int type) {
    MTet<T> m;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int side, idx;
        type_corner(type, k, side, idx);
        int64_t off = 9 * (side ? y : x) + 3 * idx;
        m.v[k] = ld3(t0 + off);
        m.dv[k] = ld3(t1 + off) - m.v[k];
    }
    return m;
}

// Vertex records of the narrow phase: for each triangle corner its position at the start of
// the step and its motion over the step, [x0 y0 z0 dx dy dz] with d = x1 - x0 (the moving
// tetrahedron's velocity, figures.cu:74-96, rounded once here instead of at every read),
// 48 B for double, 16-byte aligned, so that one corner is one contiguous read (three
// 16-byte loads) instead of two 24-byte reads from the t0 and t1 soups.
template <typename T> struct alignas(2 * sizeof(T)) VRec { T a[6]; };

template <typename T> __device__ __forceinline__ void ld_vrec(const T *tv, int64_t corner, V3<T> &v, V3<T> &dv) {
    const VRec<T> r = reinterpret_cast<const VRec<T> *>(tv)[corner];
    v = mk(r.a[0], r.a[1], r.a[2]);
    dv = mk(r.a[3], r.a[4], r.a[5]);
}

// Sub-primitive filter of a moving pair (collisions_continuous.cu:234-318, threshold 0): the
// 15-bit mask of sub-primitive pairs whose swept boxes overlap (the swept box of a vertex is
// the box of v and v + dv, of an edge or face the union of its vertices' boxes,
// figures.cu:295-314 / :346-364), or 0 when the triangles share a t0 vertex exactly
// (shareVertex, figures.cu:527-544).  Triangles x and y come from the vertex records and are
// processed one coordinate axis at a time.  Every step is exact -- the swept end
// t0 + (t1 - t0), fmin / fmax, the comparisons -- so per-axis results combine to the same
// answer: a shared vertex is equal on all three axes, and boxes overlap when their intervals
// overlap on all three.  The axis loop is unrolled so that the loads of all three axes are in
// flight together (k_subprim per pass on the 130k cloth 1.87 -> 1.65 ms, at 127 VGPRs).
template <typename T> __device__ __forceinline__ uint32_t moving_mask_axes(const T *tv, int64_t x, int64_t y) {
    uint32_t eq = 0x1ffu, m = 0x7fffu;    // bit 3i + j: corner i of x equals corner j of y
    const T *rx = tv + 18 * x, *ry = tv + 18 * y;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T lo[6], hi[6], v0[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const T *r = (k < 3 ? rx : ry) + 6 * (k % 3);
            const T v = r[d], w = v + r[3 + d];
            v0[k] = v;
            lo[k] = tmin(v, w);
            hi[k] = tmax(v, w);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                if (!(v0[i] == v0[3 + j])) eq &= ~(1u << (3 * i + j));
        // edge intervals E = {(0,1),(0,2),(1,2)} and the face interval of each triangle
        T elo[6], ehi[6];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const int a = 3 * s + edge_first(e), c = 3 * s + edge_second(e);
                elo[3 * s + e] = tmin(lo[a], lo[c]);
                ehi[3 * s + e] = tmax(hi[a], hi[c]);
            }
        const T flo0 = tmin(tmin(lo[0], lo[1]), lo[2]), fhi0 = tmax(tmax(hi[0], hi[1]), hi[2]);
        const T flo1 = tmin(tmin(lo[3], lo[4]), lo[5]), fhi1 = tmax(tmax(hi[3], hi[4]), hi[5]);
        auto ov = [](T l1, T h1, T l2, T h2) { return (l1 <= (h2 + (T)0)) && (h1 >= (l2 - (T)0)); };
        uint32_t ma = 0;
#pragma unroll
        for (int t = 0; t < 9; ++t) ma |= ov(elo[t / 3], ehi[t / 3], elo[3 + t % 3], ehi[3 + t % 3]) ? (1u << t) : 0u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ma |= ov(flo0, fhi0, lo[3 + k], hi[3 + k]) ? (1u << (9 + k)) : 0u;
            ma |= ov(lo[k], hi[k], flo1, fhi1) ? (1u << (12 + k)) : 0u;
        }
        m &= ma;
    }
    return eq ? 0u : m;
}

// gather_mtet from vertex records (same values, so the same tetrahedron bit for bit).  The
// record of each tetrahedron corner is picked as type_corner does, with the edge-edge and
// the vertex-face types on separate paths: the narrow phase's waves hold one kind, so a wave
// runs one of the two short address computations.
template <typename T>
__device__ __forceinline__ MTet<T> gather_mtet_rec(const T *tv, int64_t x, int64_t y, int type) {
    // record indices (corner 3 * triangle + i, below 3 * 2^28 as meshes hold fewer than 2^28
    // triangles); integers rather than pointers, so that the loads keep the records' global
    // address space
    const uint32_t rx = 3u * (uint32_t)x, ry = 3u * (uint32_t)y;
    uint32_t c[4];
    if (type < 9) {
        // edge type / 3 of x, edge type % 3 of y
        const int ea = type / 3, eb = type - 3 * ea;
        c[0] = rx + edge_first(ea);
        c[1] = rx + edge_second(ea);
        c[2] = ry + edge_first(eb);
        c[3] = ry + edge_second(eb);
    } else {
        // 9-11: face x, vertex type - 9 of y; 12-14: vertex type - 12 of x, face y
        const bool inv = type >= 12;
        const uint32_t j = (uint32_t)(inv ? type - 12 : type - 9);
        c[0] = inv ? rx + j : rx;
        c[1] = inv ? ry : rx + 1;
        c[2] = inv ? ry + 1 : rx + 2;
        c[3] = inv ? ry + 2 : ry + j;
    }
    MTet<T> m;
#pragma unroll
    for (int k = 0; k < 4; ++k) ld_vrec(tv, c[k], m.v[k], m.dv[k]);
    return m;
}

// figures.cu:122-293 coefficients c3..c0 of det[(v_i - v3) + t (dv_i - dv3)], i = 0..2,
// accumulated in the reference's term order; `coef += p*q*r` is fma(p*q, r, coef).
template <typename T> __device__ __forceinline__ void cubic_coeffs(const MTet<T> &t, T *c) {
    const V3<T> v0 = t.v[0] - t.v[3], v1 = t.v[1] - t.v[3], v2 = t.v[2] - t.v[3];
    const V3<T> dv0 = t.dv[0] - t.dv[3], dv1 = t.dv[1] - t.dv[3], dv2 = t.dv[2] - t.dv[3];
#define CCC_ACC(acc, p, q, r) acc = tfma((p) * (q), (r), acc)
    T c3 = 0, c2 = 0, c1 = 0, c0 = 0;
    CCC_ACC(c3, dv2.z, dv1.y, dv0.x);
    CCC_ACC(c3, -dv1.z, dv2.y, dv0.x);
    CCC_ACC(c3, -dv2.z, dv0.y, dv1.x);
    CCC_ACC(c3, dv0.z, dv2.y, dv1.x);
    CCC_ACC(c3, dv1.z, dv0.y, dv2.x);
    CCC_ACC(c3, -dv0.z, dv1.y, dv2.x);

    CCC_ACC(c2, v2.z, dv1.y, dv0.x);
    CCC_ACC(c2, dv2.z, v1.y, dv0.x);
    CCC_ACC(c2, dv2.z, dv1.y, v0.x);
    CCC_ACC(c2, -v1.z, dv2.y, dv0.x);
    CCC_ACC(c2, -dv1.z, v2.y, dv0.x);
    CCC_ACC(c2, -dv1.z, dv2.y, v0.x);
    CCC_ACC(c2, -v2.z, dv0.y, dv1.x);
    CCC_ACC(c2, -dv2.z, v0.y, dv1.x);
    CCC_ACC(c2, -dv2.z, dv0.y, v1.x);
    CCC_ACC(c2, v0.z, dv2.y, dv1.x);
    CCC_ACC(c2, dv0.z, v2.y, dv1.x);
    CCC_ACC(c2, dv0.z, dv2.y, v1.x);
    CCC_ACC(c2, v1.z, dv0.y, dv2.x);
    CCC_ACC(c2, dv1.z, v0.y, dv2.x);
    CCC_ACC(c2, dv1.z, dv0.y, v2.x);
    CCC_ACC(c2, -v0.z, dv1.y, dv2.x);
    CCC_ACC(c2, -dv0.z, v1.y, dv2.x);
    CCC_ACC(c2, -dv0.z, dv1.y, v2.x);

    CCC_ACC(c1, v2.z, v1.y, dv0.x);
    CCC_ACC(c1, v2.z, dv1.y, v0.x);
    CCC_ACC(c1, dv2.z, v1.y, v0.x);
    CCC_ACC(c1, -v1.z, v2.y, dv0.x);
    CCC_ACC(c1, -v1.z, dv2.y, v0.x);
    CCC_ACC(c1, -dv1.z, v2.y, v0.x);
    CCC_ACC(c1, -v2.z, v0.y, dv1.x);
    CCC_ACC(c1, -v2.z, dv0.y, v1.x);
    CCC_ACC(c1, -dv2.z, v0.y, v1.x);
    CCC_ACC(c1, v0.z, v2.y, dv1.x);
    CCC_ACC(c1, v0.z, dv2.y, v1.x);
    CCC_ACC(c1, dv0.z, v2.y, v1.x);
    CCC_ACC(c1, v1.z, v0.y, dv2.x);
    CCC_ACC(c1, v1.z, dv0.y, v2.x);
    CCC_ACC(c1, dv1.z, v0.y, v2.x);
    CCC_ACC(c1, -v0.z, v1.y, dv2.x);
    CCC_ACC(c1, -v0.z, dv1.y, v2.x);
    CCC_ACC(c1, -dv0.z, v1.y, v2.x);

    CCC_ACC(c0, v2.z, v1.y, v0.x);
    CCC_ACC(c0, -v1.z, v2.y, v0.x);
    CCC_ACC(c0, -v2.z, v0.y, v1.x);
    CCC_ACC(c0, v0.z, v2.y, v1.x);
    CCC_ACC(c0, v1.z, v0.y, v2.x);
    CCC_ACC(c0, -v0.z, v1.y, v2.x);
#undef CCC_ACC
    c[0] = c3; c[1] = c2; c[2] = c1; c[3] = c0;
}

// Leading term of the double-double product b*b*b (replaces CAMPARY certifiedMul,
// collisions_continuous.cu:472-507): TwoProduct via fma (Ogita, Rump & Oishi 2005),
// then Fast2Sum of the high part with the gathered low-order terms.
__device__ __forceinline__ double dd_cube_lead(double b) {
    double p = b * b;
    double e = fma(b, b, -p);
    double q = p * b;
    double f = fma(p, b, -q);
    double t = f + e * b;
    return q + t;
}

// collisions_continuous.cu:511-613 solveCubic, in double precision for both dtypes, with the
// portable cbrt / acos / sin / cos of ccc_libm.h.  Split in two so that a kernel can bin
// items by branch between the halves: cubic_prepare classifies and evaluates f, g, h;
// cubic_candidates and cubic_roots produce the roots.
enum : int { CUBIC_NONE = 0, CUBIC_QUADRATIC = 1, CUBIC_TRIPLE = 2, CUBIC_TRIG = 3, CUBIC_CARDANO = 4 };

struct CubicState { double a, b, c, d, f, g, h; };

__device__ __forceinline__ int cubic_prepare(double a, double b, double c, double d, CubicState &st) {
    st.a = a; st.b = b; st.c = c; st.d = d; st.f = 0; st.g = 0; st.h = 0;
    if (a == 0 && b == 0) return CUBIC_NONE;   // :532-536 linear root computed, never stored
    if (fabs(a) < 1e-16) return CUBIC_QUADRATIC;
    // the divisions by 3 and 27 give the division's bits (div_exact)
    const double f = div_exact<3>((3.0 * c / a) - ((b * b) / (a * a)));
    const double b3 = 2 * dd_cube_lead(b);
    const double a2 = a * a;
    const double a3 = dd_cube_lead(a);
    const double p1 = b3 / a3;
    const double p2 = (b * c * 9) / a2;
    const double p3 = 27.0 * d / a;
    const double g = div_exact<27>(p1 - p2 + p3);
    const double f3 = dd_cube_lead(f);
    const double h = (g * g) / 4.0 + div_exact<27>(f3);
    st.f = f; st.g = g; st.h = h;
    if (f == 0 && g == 0 && h == 0) return CUBIC_TRIPLE;
    return h <= 0 ? CUBIC_TRIG : CUBIC_CARDANO;
}

// The candidate roots of a classified cubic (collisions_continuous.cu:537-611) in the
// reference's order, in x0, x1, x2; returns their number (0 to 3).  The trigonometric and
// Cardano forms share their first square root and cube root: the trigonometric form takes
// sqrt(g^2/4 - h) and its cube root, Cardano sqrt(h) and the cube root of |R| (cbrt(R) for
// R >= 0 and -cbrt(-R) otherwise are +-cbrt(|R|)), so a wave holding both forms evaluates
// them once.
__device__ __forceinline__ int cubic_candidates(int code, const CubicState &st, double &x0, double &x1, double &x2) {
    const double a = st.a, b = st.b, c = st.c, d = st.d, g = st.g, h = st.h;
    if (code == CUBIC_QUADRATIC) {
        // :537-547
        double D = c * c - 4.0 * b * d;
        if (D >= 0) {
            D = sqrt(D);
            x0 = (-c + D) / (2.0 * b);
            x1 = (-c - D) / (2.0 * b);
            return 2;
        }
    } else if (code == CUBIC_TRIPLE) {
        // :564-575 triple root through a float k
        // k >= 0 ? -cbrtf(k) : cbrtf(-k), with one cube root
        float k = (float)(d / a);
        const float ck = pm_cbrtf(k >= 0 ? k : -k);
        x0 = k >= 0 ? (double)(ck * -1) : (double)ck;
        return 1;
    } else if (code == CUBIC_TRIG || code == CUBIC_CARDANO) {
        const bool trig = code == CUBIC_TRIG;
        const double sq = sqrt(trig ? ((g * g) / 4.0) - h : h);
        const double R = -(g / 2.0) + sq;
        const double cb = pm_cbrt(trig ? sq : fabs(R));
        const double b3a = b / (3.0 * a);
        if (trig) {
            // :576-591 three real roots: i = sq, j = cb
            const double k = pm_acos(-(g / (2 * sq)));
            const double L = cb * -1;
            // cos(k / 3) appears twice in the formulas; evaluated once (the same value)
            const double k3 = div_exact_finite<3>(k);      // k in [0, pi] or NaN
            const double M = pm_cos(k3);
            const double N = sqrt(3.0) * pm_sin(k3);
            const double P = b3a * -1;
            x0 = 2 * cb * M - b3a;
            x1 = L * (M + N) + P;
            x2 = L * (M - N) + P;
            return 3;
        }
        // :592-611 one real root
        const double S = R >= 0 ? cb : -cb;
        const double V = -(g / 2.0) - sq;
        // V >= 0 ? cbrt(V) : -cbrt(-V), with one cube root (the same bits, -0 and NaN too)
        const double cv = pm_cbrt(V >= 0 ? V : -V);
        const double U = V >= 0 ? cv : -cv;
        x0 = (S + U) - b3a;
        return 1;
    }
    return 0;
}

// The roots of a classified cubic that survive both of the reference's windows, in order,
// in r0, r1, r2; returns their number.  add_root keeps roots in [-1e-16, 1 + 1e-16]
// (:448-464) and flatten_cubic_roots then keeps those in [0, 1] (:743-763), so a root is
// kept exactly when its value rounded to T lies in [0, 1] (NaN fails both); equal roots are
// kept twice.
template <typename T>
__device__ __forceinline__ int cubic_roots(int code, const CubicState &st, T &r0, T &r1, T &r2) {
    double x0 = 0, x1 = 0, x2 = 0;
    const int nc = code == CUBIC_NONE ? 0 : cubic_candidates(code, st, x0, x1, x2);
    const T c0 = (T)x0, c1 = (T)x1, c2 = (T)x2;
    const bool k0 = nc > 0 && c0 >= (T)0 && c0 <= (T)1;
    const bool k1 = nc > 1 && c1 >= (T)0 && c1 <= (T)1;
    const bool k2 = nc > 2 && c2 >= (T)0 && c2 <= (T)1;
    r0 = k0 ? c0 : k1 ? c1 : c2;
    r1 = k0 && k1 ? c1 : c2;
    r2 = c2;
    return (int)k0 + (int)k1 + (int)k2;
}

// Positions at time t: figures.cu:112-117.
template <typename T> __device__ __forceinline__ void tet_at(const MTet<T> &m, T t, V3<T> *p) {
#pragma unroll
    for (int k = 0; k < 4; ++k) p[k] = m.v[k] + m.dv[k] * t;
}

// Where the correctly rounded quotient RN(n / d) lies, decided without dividing when that
// is certain: -1 below 0, +1 above 1, 0 when it may lie in [0, 1] (or when d is not a
// finite number >= dmin, or n is not finite) -- the caller then divides.  With e = 2^-49
// (double) / 2^-20 (float) and unit roundoff u: d * e is exact for d >= dmin, so
// n < -d e gives n / d < -e and RN(n / d) <= -e < 0; and n > RN(d + d e) >= d (1 + e)(1 - u)
// > d (1 + e / 2) gives RN(n / d) >= 1 + e / 2 > 1 (rounding is monotone).  Used only where
// the quotient feeds nothing but these comparisons, so results are unchanged.
__device__ __forceinline__ int ratio_side(double n, double d) {
    constexpr double e = 0x1p-49, dmin = 1e-280, big = 1.7976931348623157e308;
    if (!(d >= dmin && d <= big && fabs(n) <= big)) return 0;
    const double de = d * e;
    if (n < -de) return -1;
    if (n > d + de) return 1;
    return 0;
}
__device__ __forceinline__ int ratio_side(float n, float d) {
    constexpr float e = 0x1p-20f, dmin = 1e-30f, big = 3.40282347e38f;
    if (!(d >= dmin && d <= big && fabsf(n) <= big)) return 0;
    const float de = d * e;
    if (n < -de) return -1;
    if (n > d + de) return 1;
    return 0;
}

// Whether s = RN(n / d) passes the edge-edge filter's test !(s < 0 || s > 1) -- s in [0, 1],
// -0 included, or NaN -- for d = |c|^2 (+0 or more, or NaN), decided without dividing, with
// the division's answer.  For finite n and d in [dlo, max] (d u exact, u = 2^-53 / 2^-24):
// s >= 0 iff n >= 0 or n / d rounds to -0, i.e. |n| <= d * 2^-1075 / 2^-150 (ties go to the
// even -0), tested as ldexp(|n|, 1075 / 150) <= d (exact, or +inf); s <= 1 iff
// n / d <= 1 + u (ties go to the even 1), i.e. n - d <= d u, where n - d is exact for n in
// [d/2, 2d] (Sterbenz) and has the deciding sign elsewhere.  d = 0: s is NaN for n = 0
// (passes) and +-inf otherwise; d = +inf: s is +-0 or NaN (passes); n = +-inf with finite
// d > 0: +-inf (fails); NaN n or d: NaN (passes).  Only 0 < d < dlo is left to the division
// (the caller's `tiny`).
template <typename T> struct UnitRatioK;
template <> struct UnitRatioK<double> {
    static constexpr double dlo = 0x1p-960, u = 0x1p-53;
    static constexpr int up = 1075;
};
template <> struct UnitRatioK<float> {
    static constexpr float dlo = 0x1p-100f, u = 0x1p-24f;
    static constexpr int up = 150;
};
template <typename T> __device__ __forceinline__ bool unit_ratio(T n, T d) {
    using K = UnitRatioK<T>;
    const bool nan = n != n || d != d;
    const bool ge0 = n >= 0 || ldexp(-n, K::up) <= d;
    const bool le1 = n - d <= d * K::u;
    const bool fin = ge0 && le1;                                 // d in [dlo, max], finite n
    const bool at0 = n == 0;                                     // d = 0
    const bool dinf = d == (T)__builtin_huge_val();
    const bool ninf = tabs(n) == (T)__builtin_huge_val();
    return nan || dinf || (d == 0 ? at0 : !ninf && fin);
}

// collisions_continuous.cu:615-637 is_inside (NaN passes every comparison: accepted).  A
// barycentric coordinate that is certainly outside [0, 1] rejects before any division
// (the reference rejects too, at that check or an earlier one).
template <typename T> __device__ __forceinline__ bool is_inside(V3<T> v0, V3<T> v1, V3<T> v2, V3<T> vt) {
    V3<T> e1 = v0 - v1, e2 = v2 - v1, ep = vt - v1;
    T d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2), d20 = dot(ep, e1), d21 = dot(ep, e2);
    T denom = d00 * d11 - d01 * d01;
    const T nv = d11 * d20 - d01 * d21, nw = d00 * d21 - d01 * d20;
    if (ratio_side(nv, denom) != 0 || ratio_side(nw, denom) != 0) return false;
    T v = nv / denom;
    T w = nw / denom;
    T u = 1 - v - w;
    if (v < 0 || v > 1) return false;
    if (w < 0 || w > 1) return false;
    if (u < 0 || u > 1) return false;
    return true;
}

// collisions_continuous.cu:640-684 filterCollisions.
template <typename T> __device__ __forceinline__ bool filter_collision(const MTet<T> &m, int type, T root) {
    if (root < -1e-16) return false;
    V3<T> p[4];
    tet_at(m, root, p);
    if (type < 9) {
        V3<T> da = p[1] - p[0], db = p[3] - p[2], dc = p[2] - p[0];
        V3<T> cab = cross(da, db);
        const T ns = dot(cross(dc, db), cab), nq = dot(cross(dc, da), cab), den = norm2(cab);
        // s = ns / den and q = nq / den in [0, 1] (or NaN), decided without dividing except
        // for 0 < den < dlo (same result)
        if (den > 0 && den < UnitRatioK<T>::dlo) {
            T s = ns / den;
            T q = nq / den;
            return !(s < 0 || s > 1 || q < 0 || q > 1);
        }
        return unit_ratio(ns, den) && unit_ratio(nq, den);
    }
    // vertex-face: types 9-11 test vertex p[3] against face p[0..2], types 12-14 vertex p[0]
    // against face p[1..3]; one is_inside on the lane's corners (a wave holding both runs it
    // once)
    const bool inv = type >= 12;
    return is_inside(inv ? p[1] : p[0], inv ? p[2] : p[1], inv ? p[3] : p[2], inv ? p[0] : p[3]);
}

// Impulse of one accepted collision: collisions_impulses.cu:24-123 (mass quad),
// :162-543 (coefficients, normal, normal velocity, mass pair), :936-1050 (magnitudes) and
// :609-629 (dx extrapolation).  imp[k] / imp_dx[k] go to tetrahedron corner k.
// What one collision contributes: the unit normal n, the two impulse magnitudes and the
// position factor 1 - max(root - offset, 0).
template <typename T> struct ImpulseParams {
    V3<T> nrm;
    T im0, im1, keep;
};

// The two weights of the collision point: edge-edge (type < 9) the closest-point
// parameters s, t of the two edges; vertex-face the barycentric u, w of the face.
template <typename T> struct ImpulseWeights { T a, b; };

template <typename T>
__device__ __forceinline__ ImpulseWeights<T> impulse_weights(const MTet<T> &m, int type, T root) {
    V3<T> p[4];
    tet_at(m, root, p);
    ImpulseWeights<T> w;
    if (type < 9) {
        V3<T> da = p[1] - p[0], db = p[3] - p[2], dc = p[2] - p[0];
        T denom = norm2(cross(da, db));
        if (tabs(denom) < 1e-15) { w.a = (T)0.5; w.b = (T)0.5; }
        else {
            w.a = dot(cross(dc, db), cross(da, db)) / norm2(cross(da, db));
            w.b = dot(cross(dc, da), cross(da, db)) / norm2(cross(da, db));
        }
    } else {
        const bool inv = type >= 12;
        const int i0 = inv ? 1 : 0, i1 = inv ? 2 : 1, i2 = inv ? 3 : 2, it = inv ? 0 : 3;
        V3<T> e1 = p[i0] - p[i1], e2 = p[i2] - p[i1], ep = p[it] - p[i1];
        T d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2), d20 = dot(ep, e1), d21 = dot(ep, e2);
        T denom = d00 * d11 - d01 * d01;
        if (tabs(denom) < 1e-15) { w.a = (T)(1. / 3.); w.b = (T)(1. / 3.); }
        else { w.a = (d11 * d20 - d01 * d21) / denom; w.b = (d00 * d21 - d01 * d20) / denom; }
    }
    return w;
}

template <typename T>
__device__ __forceinline__ ImpulseParams<T> impulse_params(const MTet<T> &m, int type, T root, ImpulseWeights<T> w,
                                                           const T *q) {
    V3<T> nrm;
    T nvel, mpx, mpy;
    if (type < 9) {
        const T s = w.a, tt = w.b;
        V3<T> p0 = m.v[0] + (m.v[1] - m.v[0]) * s;
        V3<T> p1 = m.v[2] + (m.v[3] - m.v[2]) * tt;
        V3<T> dp0 = m.dv[0] + (m.dv[1] - m.dv[0]) * s;
        V3<T> dp1 = m.dv[2] + (m.dv[3] - m.dv[2]) * tt;
        V3<T> c0 = m.v[0] + m.dv[0] * root, c1 = m.v[1] + m.dv[1] * root, c2 = m.v[2] + m.dv[2] * root;
        nrm = normalize(cross(c1 - c0, c2 - c0));
        if (dot(p1 - p0, nrm) < 0) nrm = nrm * (T)-1;
        nvel = dot(dp1 - dp0, nrm);
        mpx = q[0] * (1 - s) + q[1] * s;
        mpy = q[2] * (1 - tt) + q[3] * tt;
    } else {
        const bool inv = type >= 12;
        const int i0 = inv ? 1 : 0, i1 = inv ? 2 : 1, i2 = inv ? 3 : 2, it = inv ? 0 : 3;
        const T u = w.a, ww = w.b;
        T v = 1 - u - ww;
        V3<T> pt = m.v[i0] * u + m.v[i1] * v + m.v[i2] * ww;
        V3<T> dpt = m.dv[i0] * u + m.dv[i1] * v + m.dv[i2] * ww;
        V3<T> c0 = m.v[i0] + m.dv[i0] * root, c1 = m.v[i1] + m.dv[i1] * root, c2 = m.v[i2] + m.dv[i2] * root;
        nrm = normalize(cross(c1 - c0, c2 - c0));
        if (dot(m.v[it] - pt, nrm) < 0) nrm = nrm * (T)-1;
        nvel = dot(m.dv[it] - dpt, nrm);
        if (inv) nrm = nrm * (T)-1;
        if (inv) { mpx = q[0]; mpy = q[1] * u + q[2] * v + q[3] * ww; }
        else { mpx = q[0] * u + q[1] * v + q[2] * ww; mpy = q[3]; }
    }
    T offset;
    if (tabs(nvel) < 1e-15) offset = (T)0.1;
    else offset = (T)(2 * 2e-3 / -nvel);
    offset = (T)fmin(fmax((double)offset, 0.), 1.);
    T msum = mpx + mpy;
    T mf0, mf1;
    if (tabs(msum) < 1e-15) { mf0 = (T)0.5; mf1 = (T)0.5; }
    else { mf0 = mpx / msum; mf1 = 1 - mf0; }
    ImpulseParams<T> ip;
    ip.nrm = nrm;
    ip.im0 = nvel * mf1;
    ip.im1 = -nvel * mf0;
    const T root_offset = (T)fmax((double)(root - offset), 0.);
    ip.keep = 1 - root_offset;
    return ip;
}

template <typename T>
__device__ __forceinline__ ImpulseParams<T> collision_impulse(const MTet<T> &m, int type, T root, const T *q) {
    return impulse_params(m, type, root, impulse_weights(m, type, root), q);
}

// Impulse of tetrahedron corner k: n * imp0 for the first primitive's corners, n * imp1 for
// the second's (collisions_impulses.cu:1017-1032); its position correction
// dx = ((dv + imp) - dv) * (1 - max(root - offset, 0)) (:618-629).
template <typename T>
__device__ __forceinline__ void corner_impulse(const ImpulseParams<T> &ip, int type, int k, V3<T> dvk, V3<T> &imp, V3<T> &imd) {
    const bool second = type < 9 ? (k >= 2) : (type < 12 ? (k == 3) : (k >= 1));
    imp = ip.nrm * (second ? ip.im1 : ip.im0);
    const V3<T> nvk = dvk + imp;
    imd = (nvk - dvk) * ip.keep;
}

// proximity_static.cu:205-228 distance_point_to_segment
template <typename T>
__device__ __forceinline__ void pt_seg(V3<T> p, V3<T> a, V3<T> b, T *dac, bool set_s, T second) {
    V3<T> ab = b - a, ap = p - a;
    T t = dot(ap, ab) / dot(ab, ab);
    if (t < 0.0) t = 0.0;
    if (t > 1.0) t = 1.0;
    V3<T> c = mk(a.x + t * ab.x, a.y + t * ab.y, a.z + t * ab.z);
    T nd = (T)sqrtf(sq_diff(p, c));
    if (nd < dac[0]) {
        dac[0] = nd;
        if (set_s) { dac[1] = t; dac[2] = second; }
        else { dac[2] = t; dac[1] = second; }
    }
}

// proximity_static.cu:239-332 compute_distance_edge_edge (-1 when the closest points fall
// outside either segment).
template <typename T> __device__ __forceinline__ T dist_ee(V3<T> v0, V3<T> v1, V3<T> v2, V3<T> v3) {
    V3<T> da = v1 - v0, db = v3 - v2, dc = v2 - v0;
    V3<T> cab = cross(da, db);
    T s = dot(cross(dc, db), cab) / norm2(cab);
    T t = dot(cross(dc, da), cab) / norm2(cab);
    if (isnan(s) || isnan(t)) {
        T dac[3] = {(T)3.40282346638528859811704183484516925e+38F, (T)-1., (T)-1.};
        pt_seg(v0, v2, v3, dac, false, (T)0.);
        pt_seg(v1, v2, v3, dac, false, (T)1.);
        pt_seg(v2, v0, v1, dac, true, (T)0.);
        pt_seg(v3, v0, v1, dac, true, (T)1.);
        s = dac[1];
        t = dac[2];
    }
    bool sv = (s >= 0 && s <= 1), tv = (t >= 0 && t <= 1);
    if (!sv || !tv) return (T)-1;
    return (T)sqrtf(sq_diff(v0 + da * s, v2 + db * t));
}

// proximity_static.cu:334-433 compute_distance_triangle_node
template <typename T> __device__ __forceinline__ T dist_vf(V3<T> v0, V3<T> v1, V3<T> v2, V3<T> vt) {
    V3<T> e1 = v0 - v1, e2 = v2 - v1, ep = vt - v1;
    T d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2), d20 = dot(ep, e1), d21 = dot(ep, e2);
    T denom = d00 * d11 - d01 * d01;
    T u = (d11 * d20 - d01 * d21) / denom;
    T w = (d00 * d21 - d01 * d20) / denom;
    T v = 1 - w - u;
    bool ok = (v >= 0 && v <= 1) && (w >= 0 && w <= 1) && (u >= 0 && u <= 1);
    if (!ok) return (T)-1;
    V3<T> pt = v0 * u + v1 * v + v2 * w;
    return (T)sqrtf(sq_diff(vt, pt));
}

// figures.cu:378-427 projection interval of a triangle on an axis.
template <typename T> __device__ __forceinline__ void proj_interval(V3<T> ax, const V3<T> *t, T &lo, T &hi) {
    T d = dot(ax, t[0]);
    lo = d; hi = d;
    d = dot(ax, t[1]); lo = tmin(lo, d); hi = tmax(hi, d);
    d = dot(ax, t[2]); lo = tmin(lo, d); hi = tmax(hi, d);
}
// figures.cu:446-509 17-axis separating-axis test with threshold.
template <typename T> __device__ __forceinline__ bool sat_overlap(const V3<T> *a, const V3<T> *b, T thr) {
    V3<T> ae[3] = {a[1] - a[0], a[2] - a[0], a[2] - a[1]};
    V3<T> be[3] = {b[1] - b[0], b[2] - b[0], b[2] - b[1]};
    V3<T> an = cross(ae[1], ae[2]), bn = cross(be[1], be[2]);
    bool flag = true;
#pragma unroll
    for (int i = 0; i < 17; ++i) {
        V3<T> ax;
        if (i == 0) ax = an;
        else if (i == 1) ax = bn;
        else if (i < 11) ax = cross(ae[(i - 2) / 3], be[(i - 2) % 3]);
        else if (i < 14) ax = cross(an, ae[i - 11]);
        else ax = cross(an, be[i - 14]);
        T l1, h1, l2, h2;
        proj_interval(ax, a, l1, h1);
        proj_interval(ax, b, l2, h2);
        flag = flag && ((l1 <= h2 + thr) && (h1 >= l2 - thr));
    }
    return flag;
}

// figures.cu:605-610 hasVertex: exact coordinate match with one of the corners.
template <typename T> __device__ __forceinline__ bool has_vertex(const V3<T> *t, V3<T> v) {
    return veq(t[0], v) || veq(t[1], v) || veq(t[2], v);
}

// collisions_static.cu:345-442 addTECollsion: does edge (e0, e1) of one triangle pierce
// triangle t?  Returns true with (loop, coef) when the reference emits a row.  An edge
// that starts on a corner of t (cid < 3 only) is a loop node with coefficient 0; NaN
// passes both the parameter and the barycentric tests, as in the reference.
template <typename T>
__device__ __forceinline__ bool te_intersect(const V3<T> *t, V3<T> e0, V3<T> e1, int cid, int64_t &loop, T &coef) {
    const V3<T> edge1 = t[1] - t[0], edge2 = t[2] - t[0];
    const V3<T> nrm = normalize(cross(edge1, edge2));
    const bool s0 = has_vertex(t, e0), s1 = has_vertex(t, e1);
    if (s0 || s1) {
        if (cid > 2 || !s0) return false;
        loop = 0;
        coef = (T)0;
        return true;
    }
    const V3<T> dir = e1 - e0;
    const T en = dot(dir, nrm);
    if (en == (T)0) return false;
    const T d = dot(nrm, t[0]);
    const T tp = (d - dot(nrm, e0)) / en;
    if (tp < (T)0 || tp > (T)1) return false;
    const V3<T> ep = (e0 + dir * tp) - t[0];
    const T d00 = dot(edge1, edge1), d01 = dot(edge1, edge2), d11 = dot(edge2, edge2);
    const T d20 = dot(ep, edge1), d21 = dot(ep, edge2);
    const T denom = d00 * d11 - d01 * d01;
    if (denom < (T)1e-15) return false;
    const T v = (d11 * d20 - d01 * d21) / denom;
    const T w = (d00 * d21 - d01 * d20) / denom;
    const T u = (T)1.0 - v - w;
    if (u < (T)0 || v < (T)0 || w < (T)0) return false;
    loop = -1;
    coef = tp;
    return true;
}

}  // namespace ccc
