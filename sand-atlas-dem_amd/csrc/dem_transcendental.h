// dem_transcendental.h -- FP32 exp/sin/cos shared by the HIP kernels and their
// host-side accuracy test.  Each is evaluated in FP64 and rounded once.
#pragma once
#include <math.h>
#include <stdint.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DEM_HD __host__ __device__ __forceinline__
#else
#define DEM_HD static inline
#endif

// FP32 exp/sin/cos evaluated in FP64 and rounded once, so the FP32 result is the
// correctly rounded value except within ~1e-16 relative of a rounding midpoint.
// exp: x = k ln2 + r, |r| <= ln2/2, with ln2 split in two parts (Cody & Waite 1980),
// exp(r) by its Taylor series to degree 13 (truncation < 1e-17), scaled by 2^k.
// sin/cos: x = k pi/2 + r with pi/2 in three parts, |r| <= pi/4, Taylor series
// to degree 17/18 (truncation < 1e-19), quadrant selection by k mod 4.
DEM_HD float dem_expf(float xf) {
    double x = xf;
    if (xf > 88.8f) return INFINITY;
    if (xf < -104.0f) return 0.0f;
    if (xf != xf) return xf;
    double k = rint(x * 1.4426950408889634);
    double r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
    // Horner evaluation of sum_{n<=13} r^n / n!
    double s = 1.0 / 6227020800.0;
    s = s * r + 1.0 / 479001600.0;
    s = s * r + 1.0 / 39916800.0;
    s = s * r + 1.0 / 3628800.0;
    s = s * r + 1.0 / 362880.0;
    s = s * r + 1.0 / 40320.0;
    s = s * r + 1.0 / 5040.0;
    s = s * r + 1.0 / 720.0;
    s = s * r + 1.0 / 120.0;
    s = s * r + 1.0 / 24.0;
    s = s * r + 1.0 / 6.0;
    s = s * r + 0.5;
    s = s * r + 1.0;
    s = s * r + 1.0;
    return float(ldexp(s, int(k)));
}
DEM_HD void dem_sincos_reduce(float xf, double& r, int& q) {
    double x = xf;
    double k = rint(x * 0.63661977236758134308);
    r = (x - k * 1.57079632673412561417e+00) - k * 6.07710050650619224932e-11;
    q = int(int64_t(k) & 3);
}
DEM_HD double dem_sin_poly(double r) {
    double r2 = r * r;
    return r + r * r2 * (-1.0 / 6 + r2 * (1.0 / 120 + r2 * (-1.0 / 5040 + r2 * (1.0 / 362880 + r2 * (-1.0 / 39916800 +
           r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0 + r2 * (1.0 / 355687428096000.0))))))));
}
DEM_HD double dem_cos_poly(double r) {
    double r2 = r * r;
    return 1.0 + r2 * (-1.0 / 2 + r2 * (1.0 / 24 + r2 * (-1.0 / 720 + r2 * (1.0 / 40320 + r2 * (-1.0 / 3628800 +
           r2 * (1.0 / 479001600 + r2 * (-1.0 / 87178291200.0 + r2 * (1.0 / 20922789888000.0 +
           r2 * (-1.0 / 6402373705728000.0)))))))));
}
DEM_HD float dem_sinf(float x) {
    if (!isfinite(x)) return x - x;
    if (fabsf(x) < 1e-4f) return float(double(x) - double(x) * double(x) * double(x) / 6.0);
    double r; int q;
    dem_sincos_reduce(x, r, q);
    double v = (q & 1) ? dem_cos_poly(r) : dem_sin_poly(r);
    return float((q & 2) ? -v : v);
}
DEM_HD float dem_cosf(float x) {
    if (!isfinite(x)) return x - x;
    double r; int q;
    dem_sincos_reduce(x, r, q);
    double v = (q & 1) ? dem_sin_poly(r) : dem_cos_poly(r);
    return float(((q + 1) & 2) ? -v : v);
}

