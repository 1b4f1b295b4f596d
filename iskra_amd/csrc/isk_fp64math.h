/*
 * isk_fp64math.h -- FP64 exp, log and sin/cos(2*pi*u) built only from IEEE
 * add/mul/fma, so that the GPU (gfx950) and a CPU compiled from this same file
 * produce bit-identical results.  Used by the Monte Carlo collision kernels,
 * whose accept/reject decisions and scattering angles must agree bit-for-bit
 * with a CPU run that consumes the same random draws.
 *
 * Methods (textbook):
 *  - exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2 (ln2 split in a
 *    hi part with 32 trailing zero bits and a lo part), degree-13 Taylor
 *    polynomial of e^r in Horner form, scaling by 2^k through the exponent.
 *  - log: x = 2^e m with m in [sqrt(1/2), sqrt(2)); s = (m-1)/(m+1),
 *    log m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^21/21).
 *  - sincos(2 pi u), u in [0,1): octant reduction on u (exact in binary),
 *    Taylor polynomials of sin and cos on |t| <= pi/4.
 * Accuracy is pinned against the C library by the test suite (a few ulp).
 */
#ifndef ISK_FP64MATH_H
#define ISK_FP64MATH_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define ISK_HD __host__ __device__ inline
#else
#define ISK_HD static inline
#endif

ISK_HD double isk_fma(double a, double b, double c) {
#if defined(__HIPCC__)
  return __builtin_fma(a, b, c);
#else
  return __builtin_fma(a, b, c);
#endif
}

ISK_HD double isk_bits2d(uint64_t u) {
  double d;
  __builtin_memcpy(&d, &u, 8);
  return d;
}
ISK_HD uint64_t isk_d2bits(double d) {
  uint64_t u;
  __builtin_memcpy(&u, &d, 8);
  return u;
}

/* 2^k for integer k in [-1022, 1023] */
ISK_HD double isk_pow2i(int k) {
  return isk_bits2d((uint64_t)(k + 1023) << 52);
}

ISK_HD double isk_exp(double x) {
  if (x != x) return x;
  if (x > 709.0) return isk_bits2d(0x7ff0000000000000ULL);
  if (x < -745.0) return 0.0;
  const double inv_ln2 = 1.4426950408889634;
  /* ln2 = ln2_hi + ln2_lo, ln2_hi has 32 trailing zero bits -> k*ln2_hi exact */
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  double kd = __builtin_rint(x * inv_ln2);
  int k = (int)kd;
  double r = isk_fma(-kd, ln2_hi, x);
  r = isk_fma(-kd, ln2_lo, r);
  /* Horner of sum_{n=0}^{13} r^n/n! */
  double p = 1.0 / 6227020800.0;              /* 1/13! */
  p = isk_fma(p, r, 1.0 / 479001600.0);        /* 1/12! */
  p = isk_fma(p, r, 1.0 / 39916800.0);         /* 1/11! */
  p = isk_fma(p, r, 1.0 / 3628800.0);          /* 1/10! */
  p = isk_fma(p, r, 1.0 / 362880.0);           /* 1/9!  */
  p = isk_fma(p, r, 1.0 / 40320.0);            /* 1/8!  */
  p = isk_fma(p, r, 1.0 / 5040.0);             /* 1/7!  */
  p = isk_fma(p, r, 1.0 / 720.0);              /* 1/6!  */
  p = isk_fma(p, r, 1.0 / 120.0);              /* 1/5!  */
  p = isk_fma(p, r, 1.0 / 24.0);               /* 1/4!  */
  p = isk_fma(p, r, 1.0 / 6.0);                /* 1/3!  */
  p = isk_fma(p, r, 0.5);                      /* 1/2!  */
  p = isk_fma(p, r, 1.0);
  p = isk_fma(p, r, 1.0);
  /* scale by 2^k in two steps to stay in the normal range */
  if (k < -1000) return p * isk_pow2i(k + 600) * isk_pow2i(-600);
  if (k > 1000) return p * isk_pow2i(k - 600) * isk_pow2i(600);
  return p * isk_pow2i(k);
}

/* natural log for x > 0 (normal or subnormal); returns -inf for 0, nan for x<0 */
ISK_HD double isk_log(double x) {
  if (x != x || x < 0.0) return isk_bits2d(0x7ff8000000000000ULL);
  if (x == 0.0) return isk_bits2d(0xfff0000000000000ULL);
  int e = 0;
  if (x < 2.2250738585072014e-308) { /* subnormal: scale up by 2^54 */
    x *= 18014398509481984.0;
    e = -54;
  }
  uint64_t b = isk_d2bits(x);
  e += (int)((b >> 52) & 0x7ff) - 1023;
  double m = isk_bits2d((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > 1.4142135623730951) {
    m *= 0.5;
    e += 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double q = 1.0 / 21.0;
  q = isk_fma(q, s2, 1.0 / 19.0);
  q = isk_fma(q, s2, 1.0 / 17.0);
  q = isk_fma(q, s2, 1.0 / 15.0);
  q = isk_fma(q, s2, 1.0 / 13.0);
  q = isk_fma(q, s2, 1.0 / 11.0);
  q = isk_fma(q, s2, 1.0 / 9.0);
  q = isk_fma(q, s2, 1.0 / 7.0);
  q = isk_fma(q, s2, 1.0 / 5.0);
  q = isk_fma(q, s2, 1.0 / 3.0);
  double lm = 2.0 * isk_fma(s * s2, q, s);
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  double ed = (double)e;
  return isk_fma(ed, ln2_hi, isk_fma(ed, ln2_lo, lm));
}

/* sin and cos of t for |t| <= pi/4 (Taylor to t^19 / t^18) */
ISK_HD void isk_sincos_small(double t, double *s, double *c) {
  double t2 = t * t;
  double ps = -1.0 / 121645100408832000.0;     /* -1/19! */
  ps = isk_fma(ps, t2, 1.0 / 355687428096000.0); /* 1/17! */
  ps = isk_fma(ps, t2, -1.0 / 1307674368000.0);  /* -1/15! */
  ps = isk_fma(ps, t2, 1.0 / 6227020800.0);      /* 1/13! */
  ps = isk_fma(ps, t2, -1.0 / 39916800.0);       /* -1/11! */
  ps = isk_fma(ps, t2, 1.0 / 362880.0);          /* 1/9! */
  ps = isk_fma(ps, t2, -1.0 / 5040.0);           /* -1/7! */
  ps = isk_fma(ps, t2, 1.0 / 120.0);             /* 1/5! */
  ps = isk_fma(ps, t2, -1.0 / 6.0);              /* -1/3! */
  *s = isk_fma(t * t2, ps, t);
  double pc = 1.0 / 6402373705728000.0;          /* 1/18! */
  pc = isk_fma(pc, t2, -1.0 / 20922789888000.0);   /* -1/16! */
  pc = isk_fma(pc, t2, 1.0 / 87178291200.0);       /* 1/14! */
  pc = isk_fma(pc, t2, -1.0 / 479001600.0);        /* -1/12! */
  pc = isk_fma(pc, t2, 1.0 / 3628800.0);           /* 1/10! */
  pc = isk_fma(pc, t2, -1.0 / 40320.0);            /* -1/8! */
  pc = isk_fma(pc, t2, 1.0 / 720.0);               /* 1/6! */
  pc = isk_fma(pc, t2, -1.0 / 24.0);               /* -1/4! */
  pc = isk_fma(pc, t2, 0.5);                       /* 1/2! */
  *c = isk_fma(-t2, pc, 1.0);
}

/* sin(2 pi u), cos(2 pi u) for u in [0, 1) */
ISK_HD void isk_sincos2pi(double u, double *s, double *c) {
  /* octant reduction: u = (q + f)/8, q integer 0..7, f in [-1/2, 1/2] */
  double u8 = u * 8.0; /* exact */
  double qd = __builtin_rint(u8);
  double f = u8 - qd; /* exact */
  int q = ((int)qd) & 7;
  const double pi_4 = 0.78539816339744830962; /* 2 pi / 8 */
  double t = f * pi_4;
  double st, ct;
  isk_sincos_small(t, &st, &ct);
  const double h = 0.70710678118654752440; /* sqrt(1/2) */
  /* angle = q*pi/4 + t */
  switch (q) {
    case 0: *s = st; *c = ct; break;
    case 1: *s = h * (st + ct); *c = h * (ct - st); break;
    case 2: *s = ct; *c = -st; break;
    case 3: *s = h * (ct - st); *c = -h * (st + ct); break;
    case 4: *s = -st; *c = -ct; break;
    case 5: *s = -h * (st + ct); *c = h * (st - ct); break;
    case 6: *s = -ct; *c = st; break;
    default: *s = h * (st - ct); *c = h * (st + ct); break;
  }
}

/* ---- fixed-point CIC weights ----
 * round-to-nearest-even of h * 2^32 and of a * b * 2^32 for floats in [0, 1),
 * computed from the float bit patterns with integer operations.  The scaled
 * values are exact binary numbers (24 or 48 significant bits), so these equal
 * llrint((double)h * 2^32) and llrint((double)a * (double)b * 2^32)
 * exactly; they replace the FP64 conversions on the particle path. */
ISK_HD long long isk_rne_shift(unsigned long long m, int s) {
  /* round-half-even of m / 2^s, m < 2^48, s >= 1 */
  if (s > 49) return 0;
  unsigned long long q = m >> s;
  unsigned long long rem = m & ((1ull << s) - 1ull);
  unsigned long long half = 1ull << (s - 1);
  if (rem > half || (rem == half && (q & 1ull))) ++q;
  return (long long)q;
}

ISK_HD long long isk_fix32(float h) {
  unsigned int b;
  __builtin_memcpy(&b, &h, 4);
  const int E = (int)((b >> 23) & 0xFFu);
  if (E == 0) return 0; /* zero or subnormal: below 2^-126, rounds to 0 */
  const unsigned long long m = (unsigned long long)((b & 0x7FFFFFu) | 0x800000u);
  const int sh = E - 118; /* h * 2^32 = m * 2^(E - 118) */
  if (sh >= 0) return (long long)(m << sh);
  return isk_rne_shift(m, -sh);
}

ISK_HD long long isk_fixprod32(float a, float b) {
  unsigned int ba, bb;
  __builtin_memcpy(&ba, &a, 4);
  __builtin_memcpy(&bb, &b, 4);
  const int Ea = (int)((ba >> 23) & 0xFFu), Eb = (int)((bb >> 23) & 0xFFu);
  if (Ea == 0 || Eb == 0) return 0;
  const unsigned long long ma = (unsigned long long)((ba & 0x7FFFFFu) | 0x800000u);
  const unsigned long long mb = (unsigned long long)((bb & 0x7FFFFFu) | 0x800000u);
  /* a b 2^32 = ma mb 2^(Ea + Eb - 268); Ea, Eb <= 126 so the shift is >= 16 */
  return isk_rne_shift(ma * mb, 268 - Ea - Eb);
}

#endif /* ISK_FP64MATH_H */
