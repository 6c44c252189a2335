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
 * so they return the same bits on an x86-64 host and on gfx950.  The
 * polynomials use Estrin's scheme (coefficient pairs, then powers of the
 * argument): four dependent fma levels where Horner's rule needs one per
 * coefficient, for two more operations.
 *
 * Methods (textbook, independently derived):
 *   exp : Cody-Waite argument reduction x = k*ln2 + r, |r| <= ln2/2 (up to
 *         an ulp of rounding in k), with ln2 split into a 21-bit head and a
 *         tail; degree-13 Taylor polynomial of e^r (truncation < 5e-18
 *         relative); exact scaling by 2^k through the exponent field.
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
    /* Estrin's scheme: the coefficient pairs, then r^2, r^4, r^8 steps; four
       dependent levels instead of Horner's fourteen. */
    double r2 = r * r;
    double q0 = fma(r, 1.0, 1.0);                                    /* 1/0!, 1/1! */
    double q1 = fma(r, 0.16666666666666666, 0.5);                    /* 1/2!, 1/3! */
    double q2 = fma(r, 0.008333333333333333, 0.041666666666666664);  /* 1/4!, 1/5! */
    double q3 = fma(r, 0.0001984126984126984, 0.001388888888888889); /* 1/6!, 1/7! */
    double q4 = fma(r, 2.7557319223985893e-06, 2.48015873015873e-05);/* 1/8!, 1/9! */
    double q5 = fma(r, 2.505210838544172e-08, 2.755731922398589e-07);/* 1/10!, 1/11! */
    double q6 = fma(r, 1.6059043836821613e-10, 2.08767569878681e-09);/* 1/12!, 1/13! */
    double r4 = r2 * r2;
    double s0 = fma(r2, q1, q0);
    double s1 = fma(r2, q3, q2);
    double s2 = fma(r2, q5, q4);
    double r8 = r4 * r4;
    double t0 = fma(r4, s1, s0);
    double t1 = fma(r4, q6, s2);
    double p = fma(r8, t1, t0);
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
    /* Q = z (2/3 + 2/5 z + ... + 2/23 z^10) by Estrin's scheme: pairs, then
       z^2, z^4, z^8 steps; four dependent levels instead of eleven. */
    double z2 = z * z;
    double a0 = fma(z, 0.4, 0.6666666666666666);                   /* 2/3, 2/5 */
    double a1 = fma(z, 0.2222222222222222, 0.2857142857142857);    /* 2/7, 2/9 */
    double a2 = fma(z, 0.15384615384615385, 0.18181818181818182);  /* 2/11, 2/13 */
    double a3 = fma(z, 0.11764705882352941, 0.13333333333333333);  /* 2/15, 2/17 */
    double a4 = fma(z, 0.09523809523809523, 0.10526315789473684);  /* 2/19, 2/21 */
    double a5 = 0.08695652173913043;                               /* 2/23 */
    double z4 = z2 * z2;
    double b0 = fma(z2, a1, a0);
    double b1 = fma(z2, a3, a2);
    double b2 = fma(z2, a5, a4);
    double z8 = z4 * z4;
    double c0 = fma(z4, b1, b0);
    double Q = z * fma(z8, b2, c0);
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
