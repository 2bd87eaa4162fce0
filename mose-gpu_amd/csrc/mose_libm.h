/* mose_pow: x**y with the bits of the host libm the reference links (glibc 2.35 pow on x86-64 with FMA), for the
 * device and, from this same source, for the CPU checker.
 *
 * Why: the adaptive integrators turn one ulp of a step-size factor into a different step sequence.  The reference's
 * controllers take ERR**0.25 (hairerRodas.f:805,816; hairerRadau5.f:969,976), FACCON**0.8 (hairerRadau5.f:918) and
 * QNEWT**(-1/(4+NIT-1-NEWT)) (hairerRadau5.f:942) through libm pow, and its Troe factor 10**x
 * (Lib_Chemistry_falloff.f90:78).  A correctly rounded pow would still differ from glibc's on about 8e-4 of the
 * arguments (glibc's pow is not correctly rounded), so the device reproduces glibc's algorithm instead.
 *
 * The algorithm is the published one of ARM's optimized-routines pow.c (Szabolcs Nagy, 2018), which glibc 2.28+
 * ships: x = 2^k z, log x = k ln2 + log c + log1p(z/c - 1) from a 128-entry table and a degree-8 polynomial in
 * double-double, then exp(y log x) = 2^(j/128) (1 + tmp) from a 128-entry table and a degree-5 polynomial.  The
 * x86-64 glibc build selects a variant compiled with FMA and GCC's default contraction; every product that build
 * fuses is written here as an explicit fma(), and the file is compiled with -ffp-contract=off on both sides, so the
 * result does not depend on a compiler's contraction choices.  The tables come from gen_libm_tables.py.
 *
 * Checked bit for bit against the host's pow on random arguments of every class (tests/test_libm.py) and, on the
 * device, against the same (tests/test_gpu_chem.py).
 */
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MOSE_LIBM_FN static __device__ __forceinline__
#define MOSE_LIBM_TABLE static __device__ const
#else
#include <math.h>
#define MOSE_LIBM_FN static inline
#define MOSE_LIBM_TABLE static const
#endif

#include "mose_libm_tables.h"

MOSE_LIBM_FN unsigned long long ml_asu(double d) {
  unsigned long long u;
  __builtin_memcpy(&u, &d, 8);
  return u;
}
MOSE_LIBM_FN double ml_asd(unsigned long long u) {
  double d;
  __builtin_memcpy(&d, &u, 8);
  return d;
}

/* log(x) as hi + *tail for the bits ix of a positive normal x (relative error about 2^-68) */
MOSE_LIBM_FN double ml_log_dd(unsigned long long ix, double* tail) {
  const double A0 = -0x1p-1, A1 = 0x1.555555555556p-2 * -2, A2 = -0x1.0000000000006p-2 * -2,
               A3 = 0x1.999999959554ep-3 * 4, A4 = -0x1.555555529a47ap-3 * 4, A5 = 0x1.2495b9b4845e9p-3 * -8,
               A6 = -0x1.0002b8b263fc3p-3 * -8;
  const double Ln2hi = 0x1.62e42fefa3800p-1, Ln2lo = 0x1.ef35793c76730p-45;
  const unsigned long long tmp = ix - 0x3fe6955500000000ULL;
  const int i = (int)((tmp >> 45) % 128);
  const int k = (int)((long long)tmp >> 52);
  const unsigned long long iz = ix - (tmp & 0xfffULL << 52);
  const double z = ml_asd(iz), kd = (double)k;
  const double invc = MOSE_POW_LOGT[i][0], logc = MOSE_POW_LOGT[i][1], logctail = MOSE_POW_LOGT[i][2];
  const double r = fma(z, invc, -1.0);            /* exact: 1/c has few bits */
  const double t1 = fma(kd, Ln2hi, logc);         /* exact by the choice of logc */
  const double t2 = t1 + r;
  const double lo1 = fma(kd, Ln2lo, logctail);
  const double lo2 = t1 - t2 + r;
  const double ar = A0 * r, ar2 = r * ar, ar3 = r * ar2;
  const double hi = t2 + ar2;
  const double lo3 = fma(ar, r, -ar2);
  const double lo4 = t2 - hi + ar2;
  const double q = fma(ar2, fma(ar2, fma(r, A6, A5), fma(r, A4, A3)), fma(r, A2, A1));
  const double lo = fma(ar3, q, lo1 + lo2 + lo3 + lo4);
  const double y = hi + lo;
  *tail = hi - y + lo;
  return y;
}

/* 2^1009 or 2^-1022 rescaling when |x| >= 512 (results near overflow or in the subnormal range) */
MOSE_LIBM_FN double ml_exp_special(double tmp, unsigned long long sbits, unsigned long long ki) {
  if ((ki & 0x80000000ULL) == 0) {
    sbits -= 1009ULL << 52;
    const double scale = ml_asd(sbits);
    return 0x1p1009 * fma(scale, tmp, scale);
  }
  sbits += 1022ULL << 52;
  const double scale = ml_asd(sbits);
  const double st = scale * tmp;     /* rounded on its own: the x86-64 build fuses no product used in two blocks */
  double y = scale + st;
  if (fabs(y) < 1.0) {
    /* round to the final precision before scaling into the subnormal range (no double rounding) */
    const double one = y < 0.0 ? -1.0 : 1.0;
    double lo = scale - y + st;
    const double hi = one + y;
    lo = one - hi + y + lo;
    y = (hi + lo) - one;
    if (y == 0) y = ml_asd(sbits & 0x8000000000000000ULL);
  }
  return 0x1p-1022 * y;
}

/* exp(x + xtail), |xtail| small, times -1 when sign_bias is set */
MOSE_LIBM_FN double ml_exp_dd(double x, double xtail, unsigned long long sign_bias) {
  const double InvLn2N = 0x1.71547652b82fep0 * 128, NegLn2hiN = -0x1.62e42fefa0000p-8,
               NegLn2loN = -0x1.cf79abc9e3b3ap-47, Shift = 0x1.8p52;
  const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
               C5 = 0x1.1111167a4d017p-7;
  unsigned int abstop = (unsigned int)(ml_asu(x) >> 52) & 0x7ff;
  if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {          /* |x| < 2^-54 or |x| >= 512 (or nan) */
    if (abstop - 0x3c9u >= 0x80000000u) {              /* tiny: 1 + x rounds correctly */
      const double one = 1.0 + x;
      return sign_bias ? -one : one;
    }
    if (abstop >= 0x409u) {                            /* |x| >= 1024 */
      if (ml_asu(x) >> 63) return sign_bias ? -0x1p-767 * 0x1p-767 : 0x1p-767 * 0x1p-767;
      return sign_bias ? -0x1p769 * 0x1p769 : 0x1p769 * 0x1p769;
    }
    abstop = 0;                                        /* 512 <= |x| < 1024: rescaled below */
  }
  double kd = fma(InvLn2N, x, Shift);
  const unsigned long long ki = ml_asu(kd);
  kd -= Shift;
  double r = fma(kd, NegLn2loN, fma(kd, NegLn2hiN, x));
  r += xtail;
  const unsigned long long idx = 2 * (ki % 128);
  const unsigned long long top = (ki + sign_bias) << 45;
  const double tail = ml_asd(MOSE_POW_EXPT[idx]);
  const unsigned long long sbits = MOSE_POW_EXPT[idx + 1] + top;
  const double r2 = r * r;
  const double tmp = fma(r2 * r2, fma(r, C5, C4), fma(r2, fma(r, C3, C2), tail + r));
  if (abstop == 0) return ml_exp_special(tmp, sbits, ki);
  const double scale = ml_asd(sbits);
  return fma(scale, tmp, scale);
}

/* 0: y not an integer, 1: odd integer, 2: even integer */
MOSE_LIBM_FN int ml_checkint(unsigned long long iy) {
  const int e = (int)(iy >> 52 & 0x7ff);
  if (e < 0x3ff) return 0;
  if (e > 0x3ff + 52) return 2;
  if (iy & ((1ULL << (0x3ff + 52 - e)) - 1)) return 0;
  if (iy & (1ULL << (0x3ff + 52 - e))) return 1;
  return 2;
}

/* a signaling nan (quiet bit clear) */
MOSE_LIBM_FN int ml_issignaling(unsigned long long i) {
  return 2 * (i ^ 0x0008000000000000ULL) > 2 * 0x7ff8000000000000ULL;
}

/* 2 * ix - 1 >= 2 * inf - 1: zero, inf or nan */
MOSE_LIBM_FN int ml_zeroinfnan(unsigned long long i) { return 2 * i - 1 >= 2 * 0x7ff0000000000000ULL - 1; }

MOSE_LIBM_FN double mose_pow(double x, double y) {
  unsigned long long sign_bias = 0;
  unsigned long long ix = ml_asu(x);
  const unsigned long long iy = ml_asu(y);
  unsigned int topx = (unsigned int)(ix >> 52);
  const unsigned int topy = (unsigned int)(iy >> 52);
  if (topx - 0x001u >= 0x7ffu - 0x001u || (topy & 0x7ff) - 0x3beu >= 0x43eu - 0x3beu) {
    /* x subnormal, zero, negative, inf or nan; or |y| < 2^-65, |y| >= 2^63, inf or nan */
    if (ml_zeroinfnan(iy)) {
      if (2 * iy == 0) return ml_issignaling(ix) ? x + y : 1.0;
      if (ix == 0x3ff0000000000000ULL) return ml_issignaling(iy) ? x + y : 1.0;
      if (2 * ix > 2 * 0x7ff0000000000000ULL || 2 * iy > 2 * 0x7ff0000000000000ULL) return x + y;
      if (2 * ix == 2 * 0x3ff0000000000000ULL) return 1.0;
      if ((2 * ix < 2 * 0x3ff0000000000000ULL) == !(iy >> 63)) return 0.0;
      return y * y;
    }
    if (ml_zeroinfnan(ix)) {
      double x2 = x * x;
      if ((ix >> 63) && ml_checkint(iy) == 1) x2 = -x2;
      return (iy >> 63) ? 1.0 / x2 : x2;
    }
    if (ix >> 63) {                                    /* finite x < 0 */
      const int yint = ml_checkint(iy);
      if (yint == 0) return (x - x) / (x - x);         /* nan */
      if (yint == 1) sign_bias = 0x800ULL << 7;
      ix &= 0x7fffffffffffffffULL;
      topx &= 0x7ff;
    }
    if ((topy & 0x7ff) - 0x3beu >= 0x43eu - 0x3beu) {
      if (ix == 0x3ff0000000000000ULL) return 1.0;
      if ((topy & 0x7ff) < 0x3beu) return ix > 0x3ff0000000000000ULL ? 1.0 + y : 1.0 - y;
      return (ix > 0x3ff0000000000000ULL) == (topy < 0x800) ? 0x1p769 * 0x1p769 : 0x1p-767 * 0x1p-767;
    }
    if (topx == 0) {                                   /* subnormal x: normalise */
      ix = ml_asu(x * 0x1p52);
      ix &= 0x7fffffffffffffffULL;
      ix -= 52ULL << 52;
    }
  }
  double lo;
  const double hi = ml_log_dd(ix, &lo);
  const double ehi = y * hi;
  const double elo = fma(y, lo, fma(y, hi, -ehi));
  return ml_exp_dd(ehi, elo, sign_bias);
}
