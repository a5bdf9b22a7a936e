// dd.hpp -- double-double arithmetic on FP64 FMA for the DD form of dot / gemv / gemm.
//
// A binary128 input x = (-1)^s m 2^(e-112) (113-bit m) becomes hi + lo with hi = the top 53 bits
// of m (truncated) and lo = the remaining 60 bits rounded to nearest double, so |x - (hi+lo)| <=
// 2^-107 |x|.  Products use the error-free TwoProd (one FMA), sums the error-free TwoSum
// (Knuth) / FastTwoSum (Dekker); the accumulate step is the accurate double-double addition
// (both the high and the low parts go through TwoSum), whose relative error is O(2^-104) even
// under cancellation of the high parts.  Values outside the double exponent range saturate
// (documented limit of this form).
#pragma once
#include "q128.hpp"

// The error-free transforms below are only exact if every + and * is rounded on its own: no
// contraction of a product into a later addition.
#pragma clang fp contract(off)

namespace qb {

struct dd {
  double hi, lo;
};

QB_HD dd dd_zero() {
  dd r;
  r.hi = 0.0;
  r.lo = 0.0;
  return r;
}

QB_HD double ldexp_i(double v, int e) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ldexp(v, e);
#else
  return __builtin_ldexp(v, e);
#endif
}

// binary128 -> double-double (see header comment).  NaN/Inf map to NaN/Inf in hi.
QB_HD dd q_to_dd(q128 a) {
#if defined(__HIP_DEVICE_COMPILE__)
  // Branch-free, the same bits as the general code below (checked on every exponent with random,
  // tie and carry low parts): hi is the double with the top 53 bits (exponent 52, assembled
  // directly) scaled by ldexp; lo is the low 60 bits as two exact 32-bit halves joined by one
  // rounding FMA (= the correctly rounded conversion), scaled by ldexp; signs are OR-ed in.
  const uint32_t be = q_bexp(a);
  const uint64_t sbit = a.hi & 0x8000000000000000ull;
  const int e = (int)be - QBIAS;
  const uint64_t frac52 = ((a.hi & FRAC_HI_MASK) << 4) | (a.lo >> 60);
  const double top = __builtin_bit_cast(double, ((uint64_t)(1023 + 52) << 52) | frac52);
  const double hi0 = ldexp_i(top, e - 52);
  const uint64_t rest = a.lo & 0x0fffffffffffffffull;
  const double rl = __builtin_fma((double)(uint32_t)(rest >> 32), 4294967296.0, (double)(uint32_t)rest);
  const double lo0 = ldexp_i(rl, e - 112);
  const bool special = be == 0 || be == 0x7fffu;
  const bool nanq = be == 0x7fffu && (((a.hi & FRAC_HI_MASK) | a.lo) != 0);
  dd r;
  r.hi = be == 0x7fffu ? (nanq ? __builtin_nan("") : __builtin_bit_cast(double, sbit | 0x7ff0000000000000ull))
                       : __builtin_bit_cast(double, (be == 0 ? 0ull : __builtin_bit_cast(uint64_t, hi0)) | sbit);
  r.lo = special ? 0.0 : __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, lo0) | sbit);
  return r;
#else
  dd r;
  const uint32_t be = q_bexp(a);
  const double sgn = (a.hi >> 63) ? -1.0 : 1.0;
  if (be == 0x7fffu) {
    r.hi = q_isnan(a) ? __builtin_nan("") : sgn * __builtin_inf();
    r.lo = 0.0;
    return r;
  }
  uint64_t mh = a.hi & FRAC_HI_MASK, ml = a.lo;
  int e = (int)be - QBIAS;
  if (be == 0) {
    // binary128 subnormals are far below the double range
    r.hi = sgn * 0.0;
    r.lo = 0.0;
    return r;
  }
  mh |= HIDDEN;
  // m = mh:ml (113 bits); top53 = m >> 60, rest = m & (2^60 - 1)
  const uint64_t top53 = (mh << 4) | (ml >> 60);
  const uint64_t rest = ml & 0x0fffffffffffffffull;
  r.hi = sgn * ldexp_i((double)top53, e - 52);
  // rest < 2^60: convert with round-to-nearest (the i64->f64 conversion is correctly rounded)
  r.lo = sgn * ldexp_i((double)rest, e - 112);
  return r;
#endif
}

QB_HD void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
QB_HD void fast_two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  e = b - (s - a);
}

// accurate double-double addition
QB_HD dd dd_add(dd a, dd b) {
  double s1, s2, t1, t2;
  two_sum(a.hi, b.hi, s1, s2);
  two_sum(a.lo, b.lo, t1, t2);
  s2 += t1;
  fast_two_sum(s1, s2, s1, s2);
  s2 += t2;
  dd r;
  fast_two_sum(s1, s2, r.hi, r.lo);
  return r;
}

// a*b as a double-double (TwoProd on the high parts, cross terms folded into the low part)
QB_HD dd dd_mul(dd a, dd b) {
  const double p = a.hi * b.hi;
  double e = __builtin_fma(a.hi, b.hi, -p);
  e = __builtin_fma(a.hi, b.lo, e);
  e = __builtin_fma(a.lo, b.hi, e);
  dd r;
  fast_two_sum(p, e, r.hi, r.lo);
  return r;
}

// acc + a*b
QB_HD dd dd_fma(dd a, dd b, dd acc) { return dd_add(acc, dd_mul(a, b)); }

// double-double -> binary128: hi + lo as one correctly rounded binary128 addition
QB_HD q128 dd_to_q(dd v) { return add_packed(from_double(v.hi), from_double(v.lo)); }

}  // namespace qb
