/*
 * fmath.h - the binary32 transcendentals of the path tracer: sine, cosine,
 * arc cosine and the fifth power.
 *
 * One definition, compiled into the HIP kernels and into the CPU oracle, so
 * that both sides evaluate the same sequence of IEEE operations and a path
 * sample takes the same branches on both.  The device libraries and glibc
 * round their sinf/cosf/acosf differently in the last bit, which used to let
 * a few samples in a thousand take another branch at a threshold.
 *
 * Method.  Every function evaluates in binary64 with add, multiply, rint and
 * conversions only (no division, no square root, no fused multiply-add: the
 * callers build with -ffp-contract=off), then rounds once to binary32, so the
 * result is within one binary32 ulp of the true value.
 *  - sin / cos: Cody-Waite reduction by pi/2 (Cody and Waite, "Software
 *    Manual for the Elementary Functions", 1980) with pi/2 split into two
 *    33-bit parts and a tail, so k * P1 and k * P2 are exact for |k| < 2^20
 *    (|x| < 1.6e6; larger arguments keep the same operations, and so the same
 *    bits on both sides, but lose accuracy); then the Taylor series of sin and
 *    cos on |r| <= pi/4, truncated where the next term is below 2^-45.
 *  - acos: acos(x) = pi/2 - asin(x) for |x| <= 1/2, and
 *    acos(+-x) = 2 asin(s) or pi - 2 asin(s) with s = sqrt((1 - |x|) / 2) for
 *    |x| > 1/2 (the half-angle identity), so the series of asin,
 *    sum (2n)! / (4^n (n!)^2 (2n + 1)) t^(2n+1), only sees |t| <= 1/2, where
 *    16 terms leave an error below 2^-40.  s is the correctly rounded binary32
 *    square root refined by one Newton step in binary64.
 *  - pow5: ((x x)(x x)) x in binary32.
 * Coefficients are the exact series terms rounded to binary64.
 */
#ifndef STK_FMATH_H
#define STK_FMATH_H

#include <math.h>

#if defined(__HIPCC__)
#define STK_FM_FN __host__ __device__ static inline
#else
#define STK_FM_FN static inline
#endif

/* pi/2 = P1 + P2 + P3; P1 and P2 have 33 significant bits */
#define STK_FM_P1 1.5707963267341256
#define STK_FM_P2 6.077100506303966e-11
#define STK_FM_P3 2.0222662487959506e-21
#define STK_FM_TWO_OVER_PI 0.6366197723675814
#define STK_FM_PI_2 1.5707963267948966
#define STK_FM_PI 3.141592653589793

/* sin(r) on |r| <= pi/4: r + r z (S1 + z (S2 + ... S7)), z = r^2 */
STK_FM_FN double stk_fm_sin_poly(double r)
{
    const double z = r * r;
    double p = -7.647163731819816e-13;
    p = p * z + 1.6059043836821613e-10;
    p = p * z + -2.505210838544172e-08;
    p = p * z + 2.7557319223985893e-06;
    p = p * z + -0.0001984126984126984;
    p = p * z + 0.008333333333333333;
    p = p * z + -0.16666666666666666;
    return r + (r * z) * p;
}

/* cos(r) on |r| <= pi/4: 1 + z (C1 + z (C2 + ... C7)) */
STK_FM_FN double stk_fm_cos_poly(double r)
{
    const double z = r * r;
    double p = -1.1470745597729725e-11;
    p = p * z + 2.08767569878681e-09;
    p = p * z + -2.755731922398589e-07;
    p = p * z + 2.48015873015873e-05;
    p = p * z + -0.001388888888888889;
    p = p * z + 0.041666666666666664;
    p = p * z + -0.5;
    return 1.0 + z * p;
}

/* x = k pi/2 + r with |r| <= pi/4 (about); returns r, *quadrant = k mod 4 */
STK_FM_FN double stk_fm_reduce(float x, int* quadrant)
{
    const double xd = (double)x;
    const double k = rint(xd * STK_FM_TWO_OVER_PI);
    const double r = ((xd - k * STK_FM_P1) - k * STK_FM_P2) - k * STK_FM_P3;
    /* k is an integer-valued double; its low two bits pick the quadrant */
    const double k4 = k - 4.0 * floor(k * 0.25);
    *quadrant = (int)k4;
    return r;
}

STK_FM_FN float stk_sinf(float x)
{
    if (!(fabsf(x) <= 3.0e38f)) return x - x; /* NaN for NaN and +-inf */
    int q;
    const double r = stk_fm_reduce(x, &q);
    double v;
    switch (q) {
    case 0: v = stk_fm_sin_poly(r); break;
    case 1: v = stk_fm_cos_poly(r); break;
    case 2: v = -stk_fm_sin_poly(r); break;
    default: v = -stk_fm_cos_poly(r); break;
    }
    return (float)v;
}

STK_FM_FN float stk_cosf(float x)
{
    if (!(fabsf(x) <= 3.0e38f)) return x - x;
    int q;
    const double r = stk_fm_reduce(x, &q);
    double v;
    switch (q) {
    case 0: v = stk_fm_cos_poly(r); break;
    case 1: v = -stk_fm_sin_poly(r); break;
    case 2: v = -stk_fm_cos_poly(r); break;
    default: v = stk_fm_sin_poly(r); break;
    }
    return (float)v;
}

/* asin(t) for 0 <= t <= 1/2 */
STK_FM_FN double stk_fm_asin_poly(double t)
{
    const double u = t * t;
    double p = 0.004240907093679363;
    p = p * u + 0.004660143486915096;
    p = p * u + 0.005153309682319905;
    p = p * u + 0.005740037670841924;
    p = p * u + 0.006447210311889649;
    p = p * u + 0.0073125258735988454;
    p = p * u + 0.008390335809616815;
    p = p * u + 0.009761609529194078;
    p = p * u + 0.011551800896139705;
    p = p * u + 0.01396484375;
    p = p * u + 0.017352764423076924;
    p = p * u + 0.022372159090909092;
    p = p * u + 0.030381944444444444;
    p = p * u + 0.044642857142857144;
    p = p * u + 0.075;
    p = p * u + 0.16666666666666666;
    return t + (t * u) * p;
}

STK_FM_FN float stk_acosf(float x)
{
    const float ax = fabsf(x);
    if (!(ax <= 1.0f)) return (x - x) / (x - x); /* NaN outside [-1, 1] and for NaN */
    if (ax <= 0.5f) return (float)(STK_FM_PI_2 - stk_fm_asin_poly((double)x));
    /* y = (1 - |x|) / 2 is exact in binary32 for 1/2 < |x| <= 1 */
    const float y = (1.0f - ax) * 0.5f;
    const float s0 = sqrtf(y);
    double s = 0.0;
    if (s0 > 0.0f) {
        /* y - s0^2 is exact in binary64; one Newton step for sqrt */
        const double e = (double)y - (double)s0 * (double)s0;
        const float rinv = 1.0f / s0;
        s = (double)s0 + (e * 0.5) * (double)rinv;
    }
    const double a = 2.0 * stk_fm_asin_poly(s);
    return (float)(x > 0.0f ? a : STK_FM_PI - a);
}

STK_FM_FN float stk_pow5f(float x)
{
    const float x2 = x * x;
    return (x2 * x2) * x;
}

#endif /* STK_FMATH_H */
