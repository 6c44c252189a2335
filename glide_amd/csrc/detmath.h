/*
 * detmath.h - deterministic double-precision exp / log / pow.
 *
 * The SSA solver's only transcendental functions are the grounding-line
 * sigmoid (one exp) and Glen's-law viscosity (one pow).  Library versions of
 * these differ by an ulp between glibc on the host and ROCm's device library,
 * which would make host/device trajectories of the nonlinear solve drift
 * apart.  These routines use only IEEE-754 +, -, *, / and fused multiply-add
 * (each one correctly rounded; the code is compiled without FMA contraction,
 * so only the explicit fma() calls fuse), floor and integer bit operations,
 * so they return the same bits on an x86-64 host and on gfx950.  Each Horner
 * step is one fma: half the dependent operations of a multiply and an add.
 *
 * Methods (textbook, independently derived):
 *   exp : Cody-Waite argument reduction x = k*ln2 + r, |r| <= ln2/2 (up to
 *         an ulp of rounding in k), with ln2 split into a 21-bit head and a
 *         tail; degree-13 Taylor polynomial of e^r in Horner form
 *         (truncation < 5e-18 relative); exact scaling by 2^k through the
 *         exponent field.
 *   log : x = m * 2^e with m in [sqrt(2)/2, sqrt(2)); f = m - 1 (exact),
 *         s = f / (2 + f), log(1 + f) = 2 atanh(s) written as
 *         f - (f^2/2 - s (f^2/2 + Q(s^2))) with Q the atanh series to s^22.
 *   pow : exp(p * log(x)) for x > 0 (relative error ~ (|p log x| + 2) ulp).
 * Accuracy against libm is checked by the test suite (a few ulp).
 */
#ifndef GLIDE_DETMATH_H
#define GLIDE_DETMATH_H

#include <math.h>
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
    double kd = DM_FLOOR(fma(x, INV_LN2, 0.5));
    int k = (int)kd;
    double r = fma(-kd, LN2_LO, fma(-kd, LN2_HI, x));
    double p = 1.6059043836821613e-10;          /* 1/13! */
    p = fma(r, p, 2.08767569878681e-09);           /* 1/12! */
    p = fma(r, p, 2.505210838544172e-08);          /* 1/11! */
    p = fma(r, p, 2.755731922398589e-07);          /* 1/10! */
    p = fma(r, p, 2.7557319223985893e-06);         /* 1/9!  */
    p = fma(r, p, 2.48015873015873e-05);           /* 1/8!  */
    p = fma(r, p, 0.0001984126984126984);          /* 1/7!  */
    p = fma(r, p, 0.001388888888888889);           /* 1/6!  */
    p = fma(r, p, 0.008333333333333333);           /* 1/5!  */
    p = fma(r, p, 0.041666666666666664);           /* 1/4!  */
    p = fma(r, p, 0.16666666666666666);            /* 1/3!  */
    p = fma(r, p, 0.5);                            /* 1/2!  */
    p = fma(r, p, 1.0);                            /* 1/1!  */
    p = fma(r, p, 1.0);
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
    Q = fma(z, Q, 0.09523809523809523);            /* 2/21 */
    Q = fma(z, Q, 0.10526315789473684);            /* 2/19 */
    Q = fma(z, Q, 0.11764705882352941);            /* 2/17 */
    Q = fma(z, Q, 0.13333333333333333);            /* 2/15 */
    Q = fma(z, Q, 0.15384615384615385);            /* 2/13 */
    Q = fma(z, Q, 0.18181818181818182);            /* 2/11 */
    Q = fma(z, Q, 0.2222222222222222);             /* 2/9  */
    Q = fma(z, Q, 0.2857142857142857);             /* 2/7  */
    Q = fma(z, Q, 0.4);                            /* 2/5  */
    Q = fma(z, Q, 0.6666666666666666);             /* 2/3  */
    Q = z * Q;
    double hfsq = 0.5 * f * f;
    double lm = f - fma(-s, hfsq + Q, hfsq);
    double ed = (double)e;
    return fma(ed, LN2_HI, fma(ed, LN2_LO, lm));
}

/* x^p for x >= 0 (x > 0 on every call site of the solver) */
DM_FN double dm_pow(double x, double p)
{
    if (x == 0.0) return p > 0.0 ? 0.0 : dm_from_bits(0x7ff0000000000000ULL);
    return dm_exp(p * dm_log(x));
}

#endif
