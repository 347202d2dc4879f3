/*
 * pct_rdiv.h -- FP64 quotients with one reciprocal per denominator (device).
 * Used by the exact MLP step (pct_mlp.hip) and checked against the IEEE
 * quotient by the math self-test (pct_ctx.hip).
 */
#ifndef PCT_RDIV_H
#define PCT_RDIV_H

#include <hip/hip_runtime.h>

/* FP64 quotients with one reciprocal per denominator. The compiler's
 * correctly rounded num / den is: v_div_scale of both operands (a power-of-two
 * scaling when their exponents are extreme), r = v_rcp(den) refined by two
 * Newton steps, q = num r, e = fma(-den, q, num), v_div_fmas (fma(e, r, q)
 * when nothing was scaled), then v_div_fixup (zeros, infinities, NaNs,
 * overflow). For operands whose magnitudes lie in [2^-300, 2^300] (div_ok)
 * the scaling, the fmas' rescale and the fixup change nothing, so the same
 * arithmetic with r formed once per denominator gives the compiler's quotient
 * bit for bit; with a zero numerator and a positive denominator both give +0. */
struct rdiv {
    double den, r;
};

__device__ __forceinline__ rdiv rdiv_make(double den) {
    double r = __builtin_amdgcn_rcp(den);
    double e = __builtin_fma(-den, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-den, r, 1.0);
    r = __builtin_fma(r, e, r);
    return rdiv{den, r};
}

__device__ __forceinline__ double rdiv_q(const rdiv& d, double num) {
    const double q = num * d.r;
    const double e = __builtin_fma(-d.den, q, num);
    return __builtin_fma(e, d.r, q);
}

__device__ __forceinline__ bool div_ok(double v) {
    const double a = __builtin_fabs(v);
    return (a >= 0x1p-300) & (a <= 0x1p300);
}

/* the logarithm's one division: f / (2 + f) with f in [-0.293, 0.415] or +0,
 * always inside the range above (or the zero case) */
struct div_rcp {
    __device__ double operator()(double a, double b) const { return rdiv_q(rdiv_make(b), a); }
};

#endif
