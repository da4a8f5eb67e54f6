/*
 * flowreg_detmath.h -- deterministic elementary functions.
 *
 * The data weights of the level solver need x^b (generalized Charbonnier,
 * psi = a (E + eps)^(a-1)).  libm's pow and the GPU's pow are both accurate
 * to about an ulp but do not round identically, and the optical-flow
 * pyramid amplifies one-ulp differences into visible flow differences.
 * These routines use only IEEE +, -, *, /, fused multiply-add, frexp, ldexp
 * and rint (all correctly rounded or exact), so the
 * same source gives bit-identical results on the host (gcc, x86-64) and on
 * gfx950 (hipcc), provided both are compiled with -ffp-contract=off.
 * Polynomials are evaluated by Horner's rule with explicit fma.
 *
 * det_pow:  x > 0 finite.  log via the atanh series on m in [1/sqrt2, sqrt2),
 *           exp via range reduction by ln2 and a degree-17 Taylor polynomial;
 *           relative error below 4e-15 over the solver's range
 *           (x in [1e-5, 1e4], |b| <= 1).
 * det_powf: float version (series truncated for float precision).
 */
#ifndef FLOWREG_DETMATH_H
#define FLOWREG_DETMATH_H

#ifdef __HIPCC__
#define FRH_HD __host__ __device__ inline
#else
#include <math.h>
#define FRH_HD static inline
#endif

FRH_HD double frh_det_log(double x)
{
    int e;
    double m = frexp(x, &e); /* m in [0.5, 1) */
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    /* log m = 2 s sum_k s^(2k) / (2k+1), |s| <= 0.1716, 14 terms */
    double p = 1.0 / 27.0;
    p = fma(p, s2, 1.0 / 25.0);
    p = fma(p, s2, 1.0 / 23.0);
    p = fma(p, s2, 1.0 / 21.0);
    p = fma(p, s2, 1.0 / 19.0);
    p = fma(p, s2, 1.0 / 17.0);
    p = fma(p, s2, 1.0 / 15.0);
    p = fma(p, s2, 1.0 / 13.0);
    p = fma(p, s2, 1.0 / 11.0);
    p = fma(p, s2, 1.0 / 9.0);
    p = fma(p, s2, 1.0 / 7.0);
    p = fma(p, s2, 1.0 / 5.0);
    p = fma(p, s2, 1.0 / 3.0);
    p = fma(p, s2, 1.0);
    const double ln2_hi = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000 */
    const double ln2_lo = 1.90821492927058770002e-10; /* 0x3dea39ef35793c76 */
    double de = (double)e;
    return de * ln2_hi + (de * ln2_lo + 2.0 * s * p);
}

FRH_HD double frh_det_exp(double y)
{
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double n = rint(y * inv_ln2);
    double r = (y - n * ln2_hi) - n * ln2_lo; /* |r| <= 0.3466 + tiny */
    double p = 1.0 / 355687428096000.0;       /* 1/17! */
    p = fma(p, r, 1.0 / 20922789888000.0);       /* 1/16! */
    p = fma(p, r, 1.0 / 1307674368000.0);
    p = fma(p, r, 1.0 / 87178291200.0);
    p = fma(p, r, 1.0 / 6227020800.0);
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}

FRH_HD double frh_det_pow(double x, double b) { return frh_det_exp(b * frh_det_log(x)); }

FRH_HD float frh_det_logf(float x)
{
    int e;
    float m = frexpf(x, &e);
    if (m < 0.70710678f) {
        m = m * 2.0f;
        e = e - 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 1.0f / 13.0f;
    p = fmaf(p, s2, 1.0f / 11.0f);
    p = fmaf(p, s2, 1.0f / 9.0f);
    p = fmaf(p, s2, 1.0f / 7.0f);
    p = fmaf(p, s2, 1.0f / 5.0f);
    p = fmaf(p, s2, 1.0f / 3.0f);
    p = fmaf(p, s2, 1.0f);
    const float ln2_hi = 6.9314575195e-01f; /* 0x3f317200 */
    const float ln2_lo = 1.4286067653e-06f; /* 0x35bfbe8e */
    float de = (float)e;
    return de * ln2_hi + (de * ln2_lo + 2.0f * s * p);
}

FRH_HD float frh_det_expf(float y)
{
    const float inv_ln2 = 1.4426950216e+00f;
    const float ln2_hi = 6.9314575195e-01f;
    const float ln2_lo = 1.4286067653e-06f;
    float n = rintf(y * inv_ln2);
    float r = (y - n * ln2_hi) - n * ln2_lo;
    float p = 1.0f / 40320.0f;
    p = fmaf(p, r, 1.0f / 5040.0f);
    p = fmaf(p, r, 1.0f / 720.0f);
    p = fmaf(p, r, 1.0f / 120.0f);
    p = fmaf(p, r, 1.0f / 24.0f);
    p = fmaf(p, r, 1.0f / 6.0f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)n);
}

FRH_HD float frh_det_powf(float x, float b) { return frh_det_expf(b * frh_det_logf(x)); }

#endif
