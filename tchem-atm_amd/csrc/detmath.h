/* Portable elementary functions for the rate constants: exp, log, log10 and
 * pow written with IEEE-754 double operations only (+, -, *, /, fma, rint,
 * frexp, ldexp), each correctly rounded on the CPU and on the GPU.  The
 * same source therefore gives the same bits on every device, so rate
 * constants computed on the GPU and on the host agree exactly.  Accuracy is
 * about one unit in the last place (checked against the C library over the
 * ranges the rate constants use).
 *
 * Methods (textbook, e.g. Muller, "Elementary Functions", ch. 6 and 11):
 *   exp: x = k ln2 + r with k = rint(x / ln2), |r| <= ln2 / 2, Cody-Waite
 *        splitting of ln2 (ln2_hi has 21 trailing zero bits, so k ln2_hi is
 *        exact and x - k ln2_hi is exact by Sterbenz); e^r by its Taylor
 *        polynomial to degree 13 (remainder < 5e-18 on |r| <= ln2 / 2),
 *        Horner form with fused multiply-adds; result e^r 2^k (ldexp).
 *   log: x = 2^e m with m in [sqrt(1/2), sqrt(2)); with f = m - 1 (exact)
 *        and s = f / (2 + f), log m = 2 atanh s = 2s (1 + s^2/3 + s^4/5 +
 *        ... + s^20/21) (|s| < 0.172, truncation < 1e-18 relative); log x =
 *        e ln2_hi + (e ln2_lo + log m).
 *   log10 x = log x / ln 10;  pow(x, y) = exp(y log x) for x > 0.
 * Integer powers of concentrations use repeated multiplication (x^2 is
 * the correctly rounded square).
 *
 * Used by the HIP kernels (device code) and by the CPU oracle (plain C),
 * compiled there with -ffp-contract=off so no expression is fused
 * implicitly: every fused multiply-add below is written out.
 */
#ifndef TCHEM_DETMATH_H
#define TCHEM_DETMATH_H

#include <math.h>

#if defined(__HIPCC__)
#define TCHEM_DM_FN static inline __host__ __device__
#else
#define TCHEM_DM_FN static inline
#endif

#define TCHEM_DM_LN2_HI 0.6931471803691238      /* 0x3FE62E42FEE00000 */
#define TCHEM_DM_LN2_LO 1.9082149292705877e-10  /* ln 2 - LN2_HI, rounded */
#define TCHEM_DM_INV_LN2 1.4426950408889634
#define TCHEM_DM_INV_LN10 0.4342944819032518
#define TCHEM_DM_SQRT_HALF 0.7071067811865476

TCHEM_DM_FN double tchem_exp(double x) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  if (!(x == x)) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.1332191019412) return 0.0;
  const double k = rint(x * TCHEM_DM_INV_LN2);
  const double r = (x - k * TCHEM_DM_LN2_HI) - k * TCHEM_DM_LN2_LO;
  double p = 1.6059043836821613e-10; /* 1/13! */
  p = fma(p, r, 2.08767569878681e-09);
  p = fma(p, r, 2.505210838544172e-08);
  p = fma(p, r, 2.755731922398589e-07);
  p = fma(p, r, 2.7557319223985893e-06);
  p = fma(p, r, 2.48015873015873e-05);
  p = fma(p, r, 0.0001984126984126984);
  p = fma(p, r, 0.001388888888888889);
  p = fma(p, r, 0.008333333333333333);
  p = fma(p, r, 0.041666666666666664);
  p = fma(p, r, 0.16666666666666666);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

TCHEM_DM_FN double tchem_log(double x) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : (x - x) / (x - x); /* -inf at 0, NaN below */
  if (x == HUGE_VAL) return x;
  int e = 0;
  double m = frexp(x, &e); /* x = m 2^e, m in [1/2, 1) */
  if (m < TCHEM_DM_SQRT_HALF) {
    m = m * 2.0;
    e = e - 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double q = 0.047619047619047616; /* 1/21 */
  q = fma(q, z, 0.05263157894736842);
  q = fma(q, z, 0.058823529411764705);
  q = fma(q, z, 0.06666666666666667);
  q = fma(q, z, 0.07692307692307693);
  q = fma(q, z, 0.09090909090909091);
  q = fma(q, z, 0.1111111111111111);
  q = fma(q, z, 0.14285714285714285);
  q = fma(q, z, 0.2);
  q = fma(q, z, 0.3333333333333333);
  q = q * z; /* s^2/3 + s^4/5 + ... */
  const double t = 2.0 * s;
  const double logm = fma(t, q, t);
  const double de = (double)e;
  return de * TCHEM_DM_LN2_HI + (de * TCHEM_DM_LN2_LO + logm);
}

TCHEM_DM_FN double tchem_log10(double x) { return tchem_log(x) * TCHEM_DM_INV_LN10; }

/* x^y for x > 0 (x == 0: 0 for y > 0, 1 for y == 0, +inf for y < 0) */
TCHEM_DM_FN double tchem_pow(double x, double y) {
  if (y == 0.0) return 1.0;
  if (x == 0.0) return y > 0.0 ? 0.0 : HUGE_VAL;
  return tchem_exp(y * tchem_log(x));
}

/* x^n for a small integer n >= 0 by repeated multiplication */
TCHEM_DM_FN double tchem_ipow(double x, int n) {
  double r = n >= 1 ? x : 1.0;
  for (int i = 1; i < n; ++i) r = r * x;
  return r;
}

#endif /* TCHEM_DETMATH_H */
