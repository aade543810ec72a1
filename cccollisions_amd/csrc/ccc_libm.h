// Portable double-precision cbrt / acos / sin / cos for the cubic solver (gfx950 copy).
//
// The reference calls the CUDA math library, whose last-bit behaviour is unspecified and
// differs from glibc and ROCm's.  Roots sitting at 0 or 1 up to rounding (coplanar
// configurations at t = 0 are common in cloth) are kept or dropped on that last bit, so
// the HIP kernels and the CPU oracle evaluate these functions with one algorithm built only
// from correctly rounded IEEE operations (+ - * / sqrt fma frexp ldexp); results agree bit
// for bit between the two.  Accuracy: cbrt correctly rounded on all sampled inputs, acos,
// sin and cos within 1 ulp (pinned against mpmath on the CPU copy).
//   cbrt : x = m0 2^e, m0 in [1/2, 1), e = 3q + r; w ~ m^(-1/3) for m = m0 2^r from a
//          degree-4 polynomial in m0 times 2^(-r/3), two division-free Newton steps
//          w <- w + w (1 - m w^3) / 3, then y = m w^2 and one Newton correction
//          y <- y - (y^3 - m) w^2 / 3 with the residual y^3 - m formed exactly by fma.
//   acos : |x| <= 1/2: pi/2 - asin(x); |x| > 1/2: 2 asin(sqrt((1-|x|)/2)) (mirrored with
//          pi for x < 0); asin(z) = z + z^3 P(z^2), P of degree 13 on [0, 1/4] evaluated
//          by Horner's scheme; pi and pi/2 as double-double.
//   sin, cos : sin(y) = y + y^3 S(y^2), cos(y) = 1 + y^2 C(y^2), S and C of degree 6 and 7
//          on y^2 in [0, 1.1], Horner's scheme (the solver's arguments lie in [0, pi/3]).
// Horner rather than Estrin: each step is one fma reading its coefficient from SGPRs, while
// Estrin's pairs of two coefficients need one of them copied to VGPRs first; the kernels are
// bound by VALU instruction issue, not by the longer dependency chain.
// The polynomials are Chebyshev interpolants rounded to double (gen_libm_coeffs.py prints
// them), materialised in SGPRs at their point of use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccc {

// Polynomial coefficients, constant term first (constant-evaluated only; see sgpr_const).
struct AsinCoef {
    static constexpr double c[14] = {
        0x1.5555555555555p-3, 0x1.3333333333388p-4, 0x1.6db6db6dac1e0p-5, 0x1.f1c71c7a52ba3p-6,
        0x1.6e8ba123e494cp-6, 0x1.1c4efce23019fp-6, 0x1.c990ad3d8fdcap-7, 0x1.7b027ee1dd585p-7,
        0x1.3b49de7121487p-7, 0x1.31622469ce5adp-7, 0x1.8f193743418ffp-9, 0x1.406192d124629p-6,
        -0x1.3b416bb7d9257p-6, 0x1.e529c6fce9bb4p-6};
};
struct SinCoef {
    static constexpr double c[7] = {
        -0x1.5555555555555p-3, 0x1.11111111110efp-7, -0x1.a01a01a012493p-13, 0x1.71de3a4ae46ddp-19,
        -0x1.ae6448cfef0b7p-26, 0x1.611b3f22095d8p-33, -0x1.a871f0b737686p-41};
};
struct CosCoef {
    static constexpr double c[8] = {
        -0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16bafp-10, 0x1.a01a01a00e3e7p-16,
        -0x1.27e4fb6d0860ep-22, 0x1.1eed84cdda596p-29, -0x1.938c53aa9fa3ap-37, 0x1.a85969626b78bp-45};
};
struct RcbrtCoef {
    static constexpr double c[5] = {
        0x1.0cb878fb9eb85p+1, -0x1.76eb5dac744b0p+1, 0x1.b9a7cd80a5ca6p+1, -0x1.1724277eb1344p+1,
        0x1.1ec0fe5370dbbp-1};
};
#define CCC_CBRT_SCALE1 0x1.965fea53d6e3dp-1   // 2^(-1/3)
#define CCC_CBRT_SCALE2 0x1.428a2f98d728bp-1   // 2^(-2/3)
#define CCC_THIRD 0x1.5555555555555p-2
#define CCC_PI_HI 0x1.921fb54442d18p+1
#define CCC_PI_LO 0x1.1a62633145c07p-53
#define CCC_PIO2_HI 0x1.921fb54442d18p+0
#define CCC_PIO2_LO 0x1.1a62633145c07p-54

// A double constant materialised in an SGPR pair by two s_mov_b32 immediates.  The asm
// statements are volatile, so the compiler keeps them where they are used instead of
// hoisting every coefficient of the persistent kernels' loops into VGPRs.
template <uint64_t B> __device__ __forceinline__ double sgpr_const() {
    uint32_t lo, hi;
    asm volatile("s_mov_b32 %0, %1" : "=s"(lo) : "i"((uint32_t)(B & 0xffffffffull)));
    asm volatile("s_mov_b32 %0, %1" : "=s"(hi) : "i"((uint32_t)(B >> 32)));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint64_t)lo);
}
template <class TAB, int K> __device__ __forceinline__ double coef() {
    return sgpr_const<__builtin_bit_cast(uint64_t, TAB::c[K])>();
}
// fma(a, b, c) with the addend c in an SGPR pair, as one VOP3 v_fma_f64 (the same correctly
// rounded fused multiply-add).  Left to itself the compiler forms v_fmac_f64, whose addend is
// also its destination, and first copies c into VGPRs: two moves per polynomial step.
__device__ __forceinline__ double fma_sc(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}
// RN(x / D) for the constant divisors D = 3 and 27, bit for bit what the division gives,
// in three operations instead of the eleven of a general division (one of them a
// quarter-rate reciprocal).  z = RN(1/D) has D z = 1 - 2^-54 for both, so x z lies within
// half an ulp of x / D and q = RN(x z) is a faithful quotient; then the remainder
// t = D q - x is exact (fma) and RN(q - t z) = RN(x / D) by Markstein's theorem (IBM J. Res.
// Dev. 34(1), 1990: y within half an ulp of 1/D, q faithful).  In the subnormal range the
// operations are exact fixed-point steps and x / D is never a tie.  x = +-0 keeps its sign
// (t = +0, q - t z = q); x = +-inf or NaN makes t NaN, and q is then the quotient.  The
// sequence was also compared with the division on 4*10^8 random and 1.6*10^7 structured
// inputs of every exponent, signed zeros, infinities and NaN payloads: no difference.
// div_exact_finite skips the infinity guard, for an x known to be finite or NaN.
template <int D> __device__ __forceinline__ double div_exact_finite(double x) {
    static_assert(D == 3 || D == 27, "D z = 1 - 2^-54 holds for these divisors");
    constexpr double z = 1.0 / D;
    const double q = x * z;
    const double t = fma(q, (double)D, -x);
    return fma(-t, z, q);
}
template <int D> __device__ __forceinline__ double div_exact(double x) {
    static_assert(D == 3 || D == 27, "D z = 1 - 2^-54 holds for these divisors");
    constexpr double z = 1.0 / D;
    const double q = x * z;
    const double t = fma(q, (double)D, -x);
    const double r = fma(-t, z, q);
    return t != t ? q : r;
}

// sqrt(x) for x = +-0, x >= 2^-767, x < 0 or NaN: the correctly rounded sequence the compiler
// emits for sqrt (reciprocal square root estimate, then Goldschmidt / Newton steps with fma,
// x itself for +-0 and +inf), without its rescaling of arguments below 2^-767, which is the
// identity on these inputs.  Gives sqrt's bits there.
__device__ __forceinline__ double sqrt_no_rescale(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    return (x == 0.0 || x == __builtin_huge_val()) ? x : g;
}

// p <- fma(p, t, c[k]) for k = K, K-1, ..., 0
template <class TAB, int K> __device__ __forceinline__ double horner(double p, double t) {
    if constexpr (K < 0) return p;
    else return horner<TAB, K - 1>(fma_sc(p, t, coef<TAB, K>()), t);
}

// Written without branches (zero, infinities and NaN run through the same steps, whose
// result is then replaced by x; the scale is a select of constants), so that a wave runs it
// as one straight sequence.  e = 3q + r with r in {0, 1, 2} is split as an unsigned division
// of e + 1200 (frexp exponents lie in [-1073, 1024]), the same q and r as the floor division.
__device__ __forceinline__ double pm_cbrt(double x) {
    const bool special = x == 0.0 || !isfinite(x);
    int e;
    const double m0 = frexp(fabs(x), &e);
    const unsigned n = (unsigned)(e + 1200), qb = n / 3u;
    const int r = (int)(n - 3u * qb), q = (int)qb - 400;
    double w = horner<RcbrtCoef, 3>(coef<RcbrtCoef, 4>(), m0);
    const double s1 = sgpr_const<__builtin_bit_cast(uint64_t, CCC_CBRT_SCALE1)>();
    const double s2 = sgpr_const<__builtin_bit_cast(uint64_t, CCC_CBRT_SCALE2)>();
    w = w * (r == 0 ? 1.0 : r == 1 ? s1 : s2);
    const double m = ldexp(m0, r);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double w3 = (w * w) * w;
        const double d = fma(-m, w3, 1.0);
        w = fma(w * d, CCC_THIRD, w);
    }
    const double w2 = w * w;
    double y = m * w2;
    const double y2 = y * y, y2e = fma(y, y, -y2);
    const double y3 = y2 * y, y3e = fma(y2, y, -y3);
    const double res = (y3 - m) + (y3e + y2e * y);
    y = fma(-res, w2 * CCC_THIRD, y);
    y = copysign(ldexp(y, q), x);
    return special ? x : y;
}

__device__ __forceinline__ float pm_cbrtf(float x) { return (float)pm_cbrt((double)x); }

__device__ __forceinline__ double pm_asin_small(double z) {
    const double t = z * z;
    return fma(z * t, horner<AsinCoef, 12>(coef<AsinCoef, 13>(), t), z);
}

// |x| <= 1/2: pi/2 - asin(x); otherwise 2 asin(sqrt((1 - |x|) / 2)), reflected about pi/2
// for x < 0.  The one asin_small evaluation takes the argument of the lane's case, so that
// a wave holding both cases runs the polynomial once.
// Written without branches: both double-double subtractions are formed and the lane's
// result selected (NaN and |x| > 1 give NaN, x itself for NaN).
__device__ __forceinline__ double pm_acos(double x) {
    const double ax = fabs(x);
    const bool small = ax <= 0.5;
    // (1 - |x|) / 2 is 0, at least 2^-54, negative (|x| > 1) or NaN
    const double a = pm_asin_small(small ? x : sqrt_no_rescale((1.0 - ax) * 0.5));
    const double rs = CCC_PIO2_HI - a;
    const double es = (CCC_PIO2_HI - rs) - a;
    const double vs = rs + (es + CCC_PIO2_LO);
    const double s = 2.0 * a;
    const double rl = CCC_PI_HI - s;
    const double el = (CCC_PI_HI - rl) - s;
    const double vl = rl + (el + CCC_PI_LO);
    const double v = small ? vs : x > 0 ? s : vl;
    return x != x ? x : (x > 1.0 || x < -1.0) ? (double)NAN : v;
}

__device__ __forceinline__ double pm_sin(double y) {
    const double t = y * y;
    return fma(y * t, horner<SinCoef, 5>(coef<SinCoef, 6>(), t), y);
}

__device__ __forceinline__ double pm_cos(double y) {
    const double t = y * y;
    return fma(t, horner<CosCoef, 6>(coef<CosCoef, 7>(), t), 1.0);
}

}  // namespace ccc
