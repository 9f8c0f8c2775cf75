/*
 * wr_math.h -- exp, log and cbrt in IEEE double built only from correctly
 * rounded operations (+ - * /, fma, rint, ldexp), so that the same source
 * gives the same bits on the gfx950 device and on an x86-64 host.
 *
 * The ROS2S advance is an adaptive controller: an ulp of difference in a
 * transcendental can flip an accept/reject decision and move a cell's final
 * state by ~rtol.  Evaluating the network with these functions on both sides
 * makes the GPU path reproduce a CPU build of the same algorithm bit for bit;
 * their distance from the host libm is measured separately (ulp statistics
 * over the network's arguments).
 *
 * Methods (textbook, coefficients are exact rationals rounded to double):
 *   exp:  x = k ln2 + r, |r| <= ln2/2 (Cody-Waite two-part ln2 via fma),
 *         exp(r) by its Taylor series to degree 13 (Horner with fma),
 *         result ldexp(p, k).  Truncation < 2^-56 relative.
 *   log:  x = 2^e m, m in [sqrt(2)/2, sqrt(2)), f = m - 1 (exact),
 *         s = f / (2 + f), log m = f - f^2/2 + s (f^2/2 + R(s^2)) with
 *         R(z) = sum_{k=1..11} 2 z^k / (2k + 1) (the atanh series),
 *         plus e ln2 in two parts.  Truncation < 2^-57 relative.
 *   cbrt: exponent/3 bit estimate, four Newton steps and a fifth with the
 *         residual y^3 - a formed with fma.
 * Accuracy against the correctly rounded result is about one ulp.
 *
 * Usable from C (oracle) and HIP: define WR_MATH_FN before including, or
 * accept the default (static inline, plus __host__ __device__ under HIP).
 */
#ifndef WR_MATH_H
#define WR_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef WR_MATH_FN
#if defined(__HIPCC__)
#define WR_MATH_FN static __host__ __device__ __forceinline__
#else
#define WR_MATH_FN static inline
#endif
#endif

WR_MATH_FN uint64_t wr_bits_of(double x) {
    uint64_t b;
    memcpy(&b, &x, sizeof(b));
    return b;
}

WR_MATH_FN double wr_from_bits(uint64_t b) {
    double x;
    memcpy(&x, &b, sizeof(x));
    return x;
}

WR_MATH_FN double wr_exp(double x) {
    /* Branch-free: the polynomial path runs on x clamped to [-746, 710].
     * Above 709.79 (past ln(DBL_MAX)) ldexp then overflows to +inf, and below
     * -745.2 (past ln of half the least subnormal) it rounds to +0, exactly
     * the values the cut-offs of the definition return (checked for every
     * double within 1e7 ulps of each cut-off and on a stride beyond), so
     * only NaN needs a select. */
    const double xc = fmin(fmax(x, -746.0), 710.0);
    const double kd = rint(xc * 0x1.71547652b82fep+0); /* x / ln2 */
    double r = fma(-kd, 0x1.62e42fefa39efp-1, xc);     /* ln2 high part */
    r = fma(-kd, 0x1.abc9e3b39803fp-56, r);            /* ln2 low part */
    double p = 0x1.6124613a86d09p-33;                  /* 1/13! */
    p = fma(p, r, 0x1.1eed8eff8d898p-29);              /* 1/12! */
    p = fma(p, r, 0x1.ae64567f544e4p-26);              /* 1/11! */
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);              /* 1/10! */
    p = fma(p, r, 0x1.71de3a556c734p-19);              /* 1/9! */
    p = fma(p, r, 0x1.a01a01a01a01ap-16);              /* 1/8! */
    p = fma(p, r, 0x1.a01a01a01a01ap-13);              /* 1/7! */
    p = fma(p, r, 0x1.6c16c16c16c17p-10);              /* 1/6! */
    p = fma(p, r, 0x1.1111111111111p-7);               /* 1/5! */
    p = fma(p, r, 0x1.5555555555555p-5);               /* 1/4! */
    p = fma(p, r, 0x1.5555555555555p-3);               /* 1/3! */
    p = fma(p, r, 0x1.0000000000000p-1);               /* 1/2! */
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    const double res = ldexp(p, (int)kd);
    return (x != x) ? x : res;
}

WR_MATH_FN double wr_log(double x) {
    /* Branch-free: the series runs for every x (subnormals scaled by 2^54
     * first) and the special cases are selected afterwards: NaN -> x,
     * +-0 -> -inf, x < 0 -> NaN, +inf -> +inf. */
    const int sub = x < 0x1p-1022;
    const double xs = sub ? x * 0x1p54 : x;
    const uint64_t b = wr_bits_of(xs);
    int e = (int)((b >> 52) & 0x7ff) - 1023 - (sub ? 54 : 0);
    double m = wr_from_bits((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    const int big = m > 0x1.6a09e667f3bcdp+0; /* sqrt(2) */
    m = big ? m * 0.5 : m;
    e += big;
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.642c8590b2164p-4;   /* 2/23 */
    R = fma(R, z, 0x1.8618618618618p-4); /* 2/21 */
    R = fma(R, z, 0x1.af286bca1af28p-4); /* 2/19 */
    R = fma(R, z, 0x1.e1e1e1e1e1e1ep-4); /* 2/17 */
    R = fma(R, z, 0x1.1111111111111p-3); /* 2/15 */
    R = fma(R, z, 0x1.3b13b13b13b14p-3); /* 2/13 */
    R = fma(R, z, 0x1.745d1745d1746p-3); /* 2/11 */
    R = fma(R, z, 0x1.c71c71c71c71cp-3); /* 2/9 */
    R = fma(R, z, 0x1.2492492492492p-2); /* 2/7 */
    R = fma(R, z, 0x1.999999999999ap-2); /* 2/5 */
    R = fma(R, z, 0x1.5555555555555p-1); /* 2/3 */
    R = R * z;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)e;
    /* ln2 split so that dk * high is exact for |e| < 2^11 */
    double res = dk * 0x1.62e42fefa3800p-1 -
                 ((hfsq - (s * (hfsq + R) + dk * 0x1.ef35793c76730p-45)) - f);
    res = (x == HUGE_VAL) ? x : res;
    res = (x < 0.0) ? (double)NAN : res;
    res = (x == 0.0) ? -HUGE_VAL : res;
    return (x != x) ? x : res;
}

WR_MATH_FN double wr_cbrt(double x) {
    if (x == 0.0 || x != x || x == HUGE_VAL || x == -HUGE_VAL) return x;
    double a = fabs(x);
    double scale = 1.0;
    if (a < 0x1p-1022) {
        a *= 0x1p54;
        scale = 0x1p-18;
    }
    double y = wr_from_bits(wr_bits_of(a) / 3 + 0x2aa0000000000000ULL);
    for (int i = 0; i < 4; ++i) y = y - (y * y * y - a) / (3.0 * y * y);
    /* last Newton step with the residual y^3 - a formed nearly exactly */
    const double y2 = y * y;
    const double e2 = fma(y, y, -y2);
    const double res = fma(y2, y, -a) + e2 * y;
    y = y - res / (3.0 * y2);
    y *= scale;
    return (x < 0.0) ? -y : y;
}

#endif /* WR_MATH_H */
