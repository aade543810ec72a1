// Portable double-precision cbrt / acos / sin / cos for the cubic solver (gfx950 copy).
//
// The reference calls the CUDA math library, whose last-bit behaviour is unspecified and
// differs from glibc and ROCm's.  Roots sitting at 0 or 1 up to rounding (coplanar
// configurations at t = 0 are common in cloth) are kept or dropped on that last bit, so
// the HIP kernels and the CPU oracle evaluate these functions with one algorithm built only
// from correctly rounded IEEE operations (+ - * / sqrt fma frexp ldexp); results agree bit
// for bit between the two.  Accuracy: cbrt correctly rounded on all sampled inputs, acos,
// sin and cos within 1 ulp (pinned against mpmath on the CPU copy).
//   cbrt : range reduction to m in [0.5, 4), linear start, three Halley steps
//          y <- y (y^3 + 2m) / (2y^3 + m), one Newton correction with an fma residual.
//   acos : |x| <= 1/2: pi/2 - asin(x); |x| > 1/2: 2 asin(sqrt((1-|x|)/2)) (mirrored with
//          pi for x < 0); asin(z) = z + z^3 sum c_n z^(2n-2), c_n = (2n)!/(4^n (n!)^2 (2n+1)),
//          27 terms; pi and pi/2 as double-double.
//   sin, cos : Taylor series, 11 terms, for |y| <= 1.1 (the solver's arguments lie in
//          [0, pi/3]).
// Coefficients are the exact rationals rounded to double (hex literals).  They are read
// through volatile pointers (scalar loads from the constant tables): otherwise the
// compiler hoists all of them out of the persistent item loops into ~100 VGPRs.
#pragma once

#include <hip/hip_runtime.h>

namespace ccc {

__device__ const double CCC_ASIN_C[27] = {
    0x1.5555555555555p-3, 0x1.3333333333333p-4, 0x1.6db6db6db6db7p-5, 0x1.f1c71c71c71c7p-6,
    0x1.6e8ba2e8ba2e9p-6, 0x1.1c4ec4ec4ec4fp-6, 0x1.c99999999999ap-7, 0x1.7a87878787878p-7,
    0x1.3fde50d79435ep-7, 0x1.12ef3cf3cf3cfp-7, 0x1.df3bd37a6f4dfp-8, 0x1.a6863d70a3d71p-8,
    0x1.782dda12f684cp-8, 0x1.51ba308d3dcb1p-8, 0x1.31683bdef7bdfp-8, 0x1.15ee9d45d1746p-8,
    0x1.fcaf8fb6db6dbp-9, 0x1.d3d2a8e0dd67dp-9, 0x1.b026f57b13b14p-9, 0x1.90cb77f60c7cep-9,
    0x1.750de64d7d05fp-9, 0x1.5c5f56efaaaabp-9, 0x1.464c0950f7d47p-9, 0x1.3275586c5f2f0p-9,
    0x1.208d3570ae5a6p-9, 0x1.1052bc5fa960ap-9, 0x1.018f963c229bfp-9};
__device__ const double CCC_SIN_C[11] = {
    -0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13, 0x1.71de3a556c734p-19,
    -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33, -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49,
    -0x1.2f49b46814157p-57, 0x1.71b8ef6dcf572p-66, -0x1.761b41316381ap-75};
__device__ const double CCC_COS_C[11] = {
    -0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-16,
    -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29, -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-45,
    -0x1.6827863b97d97p-53, 0x1.e542ba4020225p-62, -0x1.0ce396db7f853p-70};
#define CCC_PI_HI 0x1.921fb54442d18p+1
#define CCC_PI_LO 0x1.1a62633145c07p-53
#define CCC_PIO2_HI 0x1.921fb54442d18p+0
#define CCC_PIO2_LO 0x1.1a62633145c07p-54

__device__ __forceinline__ double pm_cbrt(double x) {
    if (x == 0.0 || !isfinite(x)) return x;
    const double ax = fabs(x);
    int e;
    double m = frexp(ax, &e);
    int q = e / 3, r = e - 3 * q;
    if (r < 0) { r += 3; q -= 1; }
    m = ldexp(m, r);
    double y = 0.6807 + 0.2266 * m;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double y3 = y * y * y;
        y = y * (y3 + 2.0 * m) / (2.0 * y3 + m);
    }
    const double y2 = y * y, y2e = fma(y, y, -y2);
    const double y3 = y2 * y, y3e = fma(y2, y, -y3);
    const double res = (y3 - m) + (y3e + y2e * y);
    y = y - res / (3.0 * y2);
    y = ldexp(y, q);
    return x < 0 ? -y : y;
}

__device__ __forceinline__ float pm_cbrtf(float x) { return (float)pm_cbrt((double)x); }

__device__ __forceinline__ double pm_asin_small(double z) {
    const double t = z * z;
    const volatile double *C = CCC_ASIN_C;
    double p = C[26];
#pragma unroll
    for (int k = 25; k >= 0; --k) p = fma(p, t, C[k]);
    return fma(z * t, p, z);
}

__device__ __forceinline__ double pm_acos(double x) {
    if (x != x) return x;
    if (x > 1.0 || x < -1.0) return NAN;
    const double ax = fabs(x);
    if (ax <= 0.5) {
        const double a = pm_asin_small(x);
        const double r = CCC_PIO2_HI - a;
        const double err = (CCC_PIO2_HI - r) - a;
        return r + (err + CCC_PIO2_LO);
    }
    const double z = sqrt((1.0 - ax) * 0.5);
    const double s = 2.0 * pm_asin_small(z);
    if (x > 0) return s;
    const double r = CCC_PI_HI - s;
    const double err = (CCC_PI_HI - r) - s;
    return r + (err + CCC_PI_LO);
}

__device__ __forceinline__ double pm_sin(double y) {
    const double t = y * y;
    const volatile double *C = CCC_SIN_C;
    double p = C[10];
#pragma unroll
    for (int k = 9; k >= 0; --k) p = fma(p, t, C[k]);
    return fma(y * t, p, y);
}

__device__ __forceinline__ double pm_cos(double y) {
    const double t = y * y;
    const volatile double *C = CCC_COS_C;
    double p = C[10];
#pragma unroll
    for (int k = 9; k >= 0; --k) p = fma(p, t, C[k]);
    return fma(t, p, 1.0);
}

}  // namespace ccc
