/*
 * detmath.h - deterministic double-precision exp / log / pow.
 *
 * The SSA solver's only transcendental functions are the grounding-line
 * sigmoid (one exp) and Glen's-law viscosity (one pow).  Library versions of
 * these differ by an ulp between glibc on the host and ROCm's device library,
 * which would make host/device trajectories of the nonlinear solve drift
 * apart.  These routines use only IEEE-754 +, -, *, / (compiled without FMA
 * contraction), floor and integer bit operations, so they return the same
 * bits on an x86-64 host and on gfx950.
 *
 * Methods (textbook, independently derived):
 *   exp : Cody-Waite argument reduction x = k*ln2 + r, |r| <= ln2/2, with ln2
 *         split into a 21-bit head and a tail; degree-13 Taylor polynomial of
 *         e^r in Horner form (truncation < 5e-18 relative); exact scaling by
 *         2^k through the exponent field.
 *   log : x = m * 2^e with m in [sqrt(2)/2, sqrt(2)); f = m - 1 (exact),
 *         s = f / (2 + f), log(1 + f) = 2 atanh(s) written as
 *         f - (f^2/2 - s (f^2/2 + Q(s^2))) with Q the atanh series to s^22.
 *   pow : exp(p * log(x)) for x > 0 (relative error ~ (|p log x| + 2) ulp).
 * Accuracy against libm is checked by the test suite (a few ulp).
 */
#ifndef GLIDE_DETMATH_H
#define GLIDE_DETMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define DM_FN __host__ __device__ static inline
#else
#define DM_FN static inline
#endif

#ifdef __cplusplus
#define DM_FLOOR(x) __builtin_floor(x)
#else
#define DM_FLOOR(x) __builtin_floor(x)
#endif

typedef union { double d; uint64_t u; } dm_bits_t;

DM_FN double dm_from_bits(uint64_t u) { dm_bits_t b; b.u = u; return b.d; }
DM_FN uint64_t dm_to_bits(double d) { dm_bits_t b; b.d = d; return b.u; }

/* 2^k for k in [-1022, 1023] */
DM_FN double dm_pow2i(int k) { return dm_from_bits((uint64_t)(k + 1023) << 52); }

DM_FN double dm_exp(double x)
{
    const double INV_LN2 = 1.4426950408889634;
    const double LN2_HI = 0.6931467056274414;      /* 0x3fe62e4200000000 */
    const double LN2_LO = 4.7493250390316726e-07;  /* ln2 - LN2_HI */
    if (x != x) return x;
    if (x > 709.782712893384) return dm_from_bits(0x7ff0000000000000ULL);
    if (x < -745.1332191019412) return 0.0;
    double kd = DM_FLOOR(x * INV_LN2 + 0.5);
    int k = (int)kd;
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    double p = 1.6059043836821613e-10;          /* 1/13! */
    p = 2.08767569878681e-09 + r * p;           /* 1/12! */
    p = 2.505210838544172e-08 + r * p;          /* 1/11! */
    p = 2.755731922398589e-07 + r * p;          /* 1/10! */
    p = 2.7557319223985893e-06 + r * p;         /* 1/9!  */
    p = 2.48015873015873e-05 + r * p;           /* 1/8!  */
    p = 0.0001984126984126984 + r * p;          /* 1/7!  */
    p = 0.001388888888888889 + r * p;           /* 1/6!  */
    p = 0.008333333333333333 + r * p;           /* 1/5!  */
    p = 0.041666666666666664 + r * p;           /* 1/4!  */
    p = 0.16666666666666666 + r * p;            /* 1/3!  */
    p = 0.5 + r * p;                            /* 1/2!  */
    p = 1.0 + r * p;                            /* 1/1!  */
    p = 1.0 + r * p;
    if (k > 1023) return (p * dm_pow2i(1023)) * dm_pow2i(k - 1023);
    if (k < -1022) return (p * dm_pow2i(k + 1000)) * dm_pow2i(-1000);
    return p * dm_pow2i(k);
}

DM_FN double dm_log(double x)
{
    const double LN2_HI = 0.6931467056274414;
    const double LN2_LO = 4.7493250390316726e-07;
    const double SQRT2 = 1.4142135623730951;
    if (x != x) return x;
    if (x < 0.0) return dm_from_bits(0x7ff8000000000000ULL);
    if (x == 0.0) return dm_from_bits(0xfff0000000000000ULL);
    if (x == dm_from_bits(0x7ff0000000000000ULL)) return x;
    int e = 0;
    uint64_t b = dm_to_bits(x);
    if ((b >> 52) == 0) {                 /* subnormal: scale by 2^54 */
        x = x * 18014398509481984.0;
        b = dm_to_bits(x);
        e = -54;
    }
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = dm_from_bits((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > SQRT2) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double Q = 0.08695652173913043;             /* 2/23 */
    Q = 0.09523809523809523 + z * Q;            /* 2/21 */
    Q = 0.10526315789473684 + z * Q;            /* 2/19 */
    Q = 0.11764705882352941 + z * Q;            /* 2/17 */
    Q = 0.13333333333333333 + z * Q;            /* 2/15 */
    Q = 0.15384615384615385 + z * Q;            /* 2/13 */
    Q = 0.18181818181818182 + z * Q;            /* 2/11 */
    Q = 0.2222222222222222 + z * Q;             /* 2/9  */
    Q = 0.2857142857142857 + z * Q;             /* 2/7  */
    Q = 0.4 + z * Q;                            /* 2/5  */
    Q = 0.6666666666666666 + z * Q;             /* 2/3  */
    Q = z * Q;
    double hfsq = 0.5 * f * f;
    double lm = f - (hfsq - s * (hfsq + Q));
    double ed = (double)e;
    return ed * LN2_HI + (ed * LN2_LO + lm);
}

/* x^p for x >= 0 (x > 0 on every call site of the solver) */
DM_FN double dm_pow(double x, double p)
{
    if (x == 0.0) return p > 0.0 ? 0.0 : dm_from_bits(0x7ff0000000000000ULL);
    return dm_exp(p * dm_log(x));
}

#endif
