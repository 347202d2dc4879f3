/*
 * pct_math.h -- deterministic FP64/FP32 elementary functions shared by the
 * host (the CPU oracle, the synthetic generator) and the gfx950 kernels.
 *
 * Why: voxel indices, bin numbers and hull masks must agree bit for bit
 * between the CPU restatement of the reference and the GPU kernels. The
 * reference calls the CUDA device math library on the GPU and glibc on the
 * host, so even the reference's own two paths disagree in the last ulp. Here
 * every transcendental is one implementation, written only with IEEE +,-,*,/
 * (correctly rounded on x86-64 SSE2 and on CDNA4), and every translation unit
 * that includes this header is compiled with -ffp-contract=off, so the same
 * operation sequence runs on both sides.
 *
 * Methods (all from the mathematics, no vendor library consulted):
 *   log      : x = 2^k * m, m in [sqrt(1/2), sqrt(2)), log(m) = 2 atanh(s),
 *              s = f/(2+f), f = m-1, summed as f - h + s(h + R) with
 *              h = f^2/2 (the classical decomposition of Cody & Waite,
 *              "Software Manual for the Elementary Functions", 1980) and
 *              R a truncated Taylor series of the atanh terms; k ln 2 is
 *              split in a 32-bit head and a tail.
 *   sin/cos  : Cody-Waite reduction by pi/2 with a three-part constant
 *              (33+33+53 bits), Taylor polynomials on |r| <= pi/4.
 *   atan     : reduction by the breakpoints atan(1/2), atan(1), atan(3/2)
 *              and pi/2 - atan(1/x); Taylor series of atan on |t| <= 7/16.
 *   atan2    : quadrant logic around atan(y/x).
 *   sqrt     : the IEEE square root (correctly rounded on both targets).
 * Accuracy of each function is pinned against numpy/glibc in tests
 * (a few ulp at most); bit equality between host and device is pinned by the
 * GPU tests.
 */
#ifndef PCT_MATH_H
#define PCT_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define PCT_HD __host__ __device__ inline
#define PCT_HD_MEMBER __host__ __device__
#else
#define PCT_HD static inline
#define PCT_HD_MEMBER
#endif

/* the IEEE quotient a / b: the division pct_log_div takes in pct_log */
struct pct_div_ieee {
    PCT_HD_MEMBER double operator()(double a, double b) const { return a / b; }
};

PCT_HD uint64_t pct_bits(double x) { return __builtin_bit_cast(uint64_t, x); }
PCT_HD double pct_from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }
PCT_HD double pct_fabs(double x) { return pct_from_bits(pct_bits(x) & 0x7fffffffffffffffULL); }
PCT_HD float pct_fabsf(float x) { return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) & 0x7fffffffu); }
PCT_HD double pct_sqrt(double x) { return __builtin_sqrt(x); }
PCT_HD float pct_sqrtf(float x) { return __builtin_sqrtf(x); }

/* round-half-away-from-zero to an integer-valued double, |x| < 2^51 */
PCT_HD double pct_round_int(double x) {
    double a = pct_fabs(x) + 0.5;
    /* truncate a to an integer: add and subtract 2^52 after flooring */
    double t = (double)(int64_t)a;
    return x < 0 ? -t : t;
}

#define PCT_PI      3.141592653589793          /* nearest double to pi (= 4*atan(1.0)) */
#define PCT_PI_LO   1.2246467991473532e-16
#define PCT_PIO2    1.5707963267948966
#define PCT_PIO2_LO 6.123233995736766e-17
#define PCT_PIO2_1  1.5707963267341256         /* 33 leading bits of pi/2 */
#define PCT_PIO2_2  6.077100506303966e-11      /* next 33 bits */
#define PCT_PIO2_3  2.0222662487959506e-21     /* remainder, rounded */
#define PCT_2_OVER_PI 0.6366197723675814
#define PCT_LN2_HI  0.6931471803691238         /* 32 leading bits of ln 2 */
#define PCT_LN2_LO  1.9082149292705877e-10

/* natural logarithm, x > 0 finite (x <= 0 returns -inf / NaN like libm) */
/* natural logarithm; div(a, b) must return the IEEE quotient a / b for
 * a = f in [-0.293, 0.415] (or +0) and b = 2 + f, the one division */
template <class Div>
PCT_HD double pct_log_core(double x, int k, Div div);
template <class Div>
PCT_HD double pct_log_div(double x, Div div) {
    if (!(x > 0.0)) return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
    if (pct_bits(x) >= 0x7ff0000000000000ULL) return x; /* +inf */
    int k = 0;
    if (x < 2.2250738585072014e-308) { x *= 18014398509481984.0; k = -54; } /* 2^54 */
    return pct_log_core(x, k, div);
}
/* pct_log_div for a positive normal finite x: the same arithmetic without
 * the special cases (the caller guarantees the range) */
template <class Div>
PCT_HD double pct_log_normal(double x, Div div) { return pct_log_core(x, 0, div); }
/* log(x) + k ln 2 - k ln 2 of a positive normal x: the reduction and the
 * series (k is the exponent offset of a pre-scaled subnormal) */
template <class Div>
PCT_HD double pct_log_core(double x, int k, Div div) {
    uint64_t u = pct_bits(x);
    k += (int)((u >> 52) & 0x7ff) - 1023;
    double m = pct_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.4142135623730951) { m = m * 0.5; k += 1; }
    double f = m - 1.0;
    double h = 0.5 * f * f;
    double s = div(f, 2.0 + f);
    double z = s * s;
    double R = z * (2.0 / 3.0 + z * (2.0 / 5.0 + z * (2.0 / 7.0 + z * (2.0 / 9.0 + z * (2.0 / 11.0
             + z * (2.0 / 13.0 + z * (2.0 / 15.0 + z * (2.0 / 17.0 + z * (2.0 / 19.0 + z * (2.0 / 21.0))))))))));
    double dk = (double)k;
    return dk * PCT_LN2_HI - ((h - (s * (h + R) + dk * PCT_LN2_LO)) - f);
}
PCT_HD double pct_log(double x) { return pct_log_div(x, pct_div_ieee{}); }

/* sin and cos of r, |r| <= ~pi/4 (Taylor to degree 17 / 18) */
PCT_HD double pct_sin_kernel(double r) {
    double z = r * r;
    double p = -1.0 / 6.0 + z * (1.0 / 120.0 + z * (-1.0 / 5040.0 + z * (1.0 / 362880.0
             + z * (-1.0 / 39916800.0 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0
             + z * (1.0 / 355687428096000.0)))))));
    return r + r * (z * p);
}
PCT_HD double pct_cos_kernel(double r) {
    double z = r * r;
    double p = 1.0 / 24.0 + z * (-1.0 / 720.0 + z * (1.0 / 40320.0 + z * (-1.0 / 3628800.0
             + z * (1.0 / 479001600.0 + z * (-1.0 / 87178291200.0 + z * (1.0 / 20922789888000.0
             + z * (-1.0 / 6402373705728000.0)))))));
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + z * (z * p));
}
/* sin and cos of x, |x| < 2^19 * pi/2 (angles of a few turns here) */
PCT_HD void pct_sincos(double x, double* s, double* c) {
    double kd = pct_round_int(x * PCT_2_OVER_PI);
    double r = ((x - kd * PCT_PIO2_1) - kd * PCT_PIO2_2) - kd * PCT_PIO2_3;
    int q = ((int)(int64_t)kd) & 3;
    double sr = pct_sin_kernel(r), cr = pct_cos_kernel(r);
    switch (q) {
        case 0: *s = sr;  *c = cr;  break;
        case 1: *s = cr;  *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}
PCT_HD double pct_sin(double x) { double s, c; pct_sincos(x, &s, &c); return s; }
PCT_HD double pct_cos(double x) { double s, c; pct_sincos(x, &s, &c); return c; }
PCT_HD double pct_tan(double x) { double s, c; pct_sincos(x, &s, &c); return s / c; }
/* single precision sin/cos: the double routine rounded once to float */
PCT_HD float pct_sinf(float x) { return (float)pct_sin((double)x); }
PCT_HD float pct_cosf(float x) { return (float)pct_cos((double)x); }

/* atan(t) - t for |t| <= 7/16: returns t*z*P(z) so that atan(t) = t - ret */
PCT_HD double pct_atan_tail(double t) {
    double z = t * t;
    double p = 0.0;
    /* 1/3 - z/5 + z^2/7 - ... (24 terms) */
    for (int n = 24; n >= 1; --n) {
        double coef = ((n & 1) ? 1.0 : -1.0) / (double)(2 * n + 1);
        p = coef + z * p;
    }
    return t * (z * p);
}
PCT_HD double pct_atan(double x) {
    double ax = pct_fabs(x);
    if (ax != ax) return x;
    double hi, lo, t;
    int id;
    if (ax >= 2.4375) {
        if (ax > 1.0e17) { double r = PCT_PIO2 + PCT_PIO2_LO; return x < 0 ? -r : r; }
        id = 3; t = -1.0 / ax;
    } else if (ax >= 1.1875) {
        id = 2; t = (ax - 1.5) / (1.0 + 1.5 * ax);
    } else if (ax >= 0.6875) {
        id = 1; t = (ax - 1.0) / (ax + 1.0);
    } else if (ax >= 0.4375) {
        id = 0; t = (2.0 * ax - 1.0) / (2.0 + ax);
    } else {
        double r = ax - pct_atan_tail(ax);
        return x < 0 ? -r : r;
    }
    switch (id) {
        case 0: hi = 0.4636476090008061; lo = 2.2698777452961687e-17; break;
        case 1: hi = 0.7853981633974483; lo = 3.061616997868383e-17; break;
        case 2: hi = 0.982793723247329;  lo = 1.3903311031230998e-17; break;
        default: hi = PCT_PIO2;          lo = PCT_PIO2_LO; break;
    }
    double r = hi - ((pct_atan_tail(t) - lo) - t);
    return x < 0 ? -r : r;
}
PCT_HD double pct_atan2(double y, double x) {
    if (x != x || y != y) return x + y;
    if (x == 0.0) {
        if (y == 0.0) return (pct_bits(x) >> 63) ? ((pct_bits(y) >> 63) ? -PCT_PI : PCT_PI) : y;
        return y > 0 ? PCT_PIO2 : -PCT_PIO2;
    }
    double a = pct_atan(pct_fabs(y / x));
    if (x > 0.0) return (pct_bits(y) >> 63) ? -a : a;
    /* x < 0: pi - a, with the tail of pi */
    double r = PCT_PI - (a - PCT_PI_LO);
    return (pct_bits(y) >> 63) ? -r : r;
}

/* single precision atan2: the double routine rounded once to float */
PCT_HD float pct_atan2f(float y, float x) { return (float)pct_atan2((double)y, (double)x); }

#endif /* PCT_MATH_H */
