// q128.hpp -- software IEEE-754 binary128 arithmetic for gfx950 (and the host, for unit tests).
//
// Replaces the arithmetic the reference takes from SLEEF's quad library: Sleef_fmaq1_u05,
// Sleef_addq1_u05, Sleef_mulq1_u05 and the 2-lane Sleef_fmaq2_u05sse2
// (include/quadblas/simd/quad_vector.hpp:99-139, include/quadblas/algorithms/level1.hpp:31,70,
// level2.hpp:43,48,65,71,79, level3.hpp:36,40,97,107,170,174,206,210).  Every operation here is
// correctly rounded, round-to-nearest-even, including subnormal results, overflow to infinity,
// signed zeros and NaN/Inf propagation (a NaN result is the canonical quiet NaN).
//
// Representation
//   packed  q128 {lo, hi}: the little-endian memory image of __float128.
//   unpacked uq: a finite value  (-1)^s * m * 2^(e - 112)  with the 113-bit significand m held in
//           two 64-bit words (mh: bits 112..64, ml: bits 63..0) and normalised so bit 112 is set.
//           Zero is m = 0 with e = UQ_EZERO.  Subnormal binary128 inputs unpack to a normalised m
//           with e < -16382; every rounding step re-applies the binary128 subnormal grid, so an
//           unpacked chain holds exactly the values a packed chain would.
//
// The fast path (fma_uq) handles finite operands with a nonzero product and a finite addend
// (zero included); everything else (NaN, Inf, zero products, overflow) goes through
// fma_packed's special-case rules.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define QB_HD __host__ __device__ __forceinline__
#else
#define QB_HD static inline
#endif

namespace qb {

struct q128 {
  uint64_t lo, hi;
};

struct uq {
  uint64_t ml, mh;  // significand: mh holds bits 112..64 (bit 48 of mh is bit 112)
  int32_t e;        // value = m * 2^(e - 112)
  uint32_t s;       // sign bit
};

constexpr int32_t QBIAS = 16383;
constexpr int32_t EMIN = 1 - QBIAS;   // exponent of the smallest normal
constexpr int32_t EMAX = QBIAS;       // exponent of the largest finite
constexpr int32_t UQ_EZERO = -(1 << 24);
constexpr uint64_t FRAC_HI_MASK = 0x0000ffffffffffffull;
constexpr uint64_t HIDDEN = 0x0001000000000000ull;

QB_HD q128 q_make(uint64_t lo, uint64_t hi) {
  q128 r;
  r.lo = lo;
  r.hi = hi;
  return r;
}
QB_HD q128 q_nan() { return q_make(0, 0x7fff800000000000ull); }
QB_HD q128 q_inf(uint32_t s) { return q_make(0, ((uint64_t)s << 63) | 0x7fff000000000000ull); }
QB_HD q128 q_zero(uint32_t s) { return q_make(0, (uint64_t)s << 63); }
QB_HD q128 q_one() { return q_make(0, 0x3fff000000000000ull); }

QB_HD uint32_t q_bexp(q128 a) { return (uint32_t)(a.hi >> 48) & 0x7fffu; }
QB_HD bool q_isnan(q128 a) { return q_bexp(a) == 0x7fffu && ((a.hi & FRAC_HI_MASK) | a.lo) != 0; }
QB_HD bool q_isinf(q128 a) { return q_bexp(a) == 0x7fffu && ((a.hi & FRAC_HI_MASK) | a.lo) == 0; }
QB_HD bool q_iszero(q128 a) { return ((a.hi << 1) | a.lo) == 0; }

QB_HD int clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return x ? __clzll((long long)x) : 64;
#else
  return x ? __builtin_clzll(x) : 64;
#endif
}

// ------------------------------------------------------------------------------------------
// multiword helpers over little-endian uint64 words; all indices are compile-time constants
// ------------------------------------------------------------------------------------------

// x >> s for s in [0, 64): the caller guarantees the range
QB_HD uint64_t shr_fill(uint64_t hi, uint64_t lo, int s) {  // low word of (hi:lo) >> s
  return (lo >> s) | ((hi << 1) << (63 - s));
}
QB_HD uint64_t shl_fill(uint64_t hi, uint64_t lo, int s) {  // high word of (hi:lo) << s
  return (hi << s) | ((lo >> 1) >> (63 - s));
}

// r = x << sh, N words, 0 <= sh < 64*N; bits above the top word are dropped.
template <int N>
QB_HD void mw_shl(uint64_t (&r)[N], const uint64_t (&x)[N], int sh) {
  const int ws = sh >> 6, bs = sh & 63;
  uint64_t t[N];
#pragma unroll
  for (int i = N - 1; i > 0; --i) t[i] = shl_fill(x[i], x[i - 1], bs);
  t[0] = x[0] << bs;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    uint64_t v = 0;
#pragma unroll
    for (int k = 0; k <= i; ++k) v = (ws == k) ? t[i - k] : v;
    r[i] = v;
  }
}

// r = x >> sh (truncating); returns 1 if any nonzero bit was shifted out.  Any sh >= 0.
// Callers fold the returned sticky into a *final* sum (round-to-odd), never into an addend.
template <int N>
QB_HD uint64_t mw_shr_sticky(uint64_t (&r)[N], const uint64_t (&x)[N], int sh) {
  if (sh >= 64 * N) {
    uint64_t any = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) any |= x[i];
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] = 0;
    return any != 0;
  }
  const int ws = sh >> 6, bs = sh & 63;
  uint64_t t[N];
#pragma unroll
  for (int i = 0; i < N - 1; ++i) t[i] = shr_fill(x[i + 1], x[i], bs);
  t[N - 1] = x[N - 1] >> bs;
  // sticky: words entirely below ws, plus the low bs bits of word ws
  uint64_t lost = 0;
  const uint64_t lowmask = (bs == 0) ? 0 : (~0ull >> (64 - bs));
#pragma unroll
  for (int i = 0; i < N; ++i) {
    uint64_t part = (i < ws) ? x[i] : ((i == ws) ? (x[i] & lowmask) : 0);
    lost |= part;
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    uint64_t v = 0;
#pragma unroll
    for (int k = 0; k < N - i; ++k) v = (ws == k) ? t[i + k] : v;
    r[i] = v;
  }
  return lost != 0;
}

// ------------------------------------------------------------------------------------------
// pack / unpack
// ------------------------------------------------------------------------------------------

// Unpack a finite nonzero normal value.  Sets 'special' for zero, subnormal, Inf and NaN.
QB_HD uq unpack_fast(q128 a, bool &special) {
  uq r;
  const uint32_t be = q_bexp(a);
  special = (be == 0) | (be == 0x7fffu);
  r.ml = a.lo;
  r.mh = (a.hi & FRAC_HI_MASK) | HIDDEN;
  r.e = (int32_t)be - QBIAS;
  r.s = (uint32_t)(a.hi >> 63);
  return r;
}

// Unpack any finite value (zero and subnormals included).
QB_HD uq unpack_finite(q128 a) {
  uq r;
  const uint32_t be = q_bexp(a);
  r.s = (uint32_t)(a.hi >> 63);
  uint64_t fh = a.hi & FRAC_HI_MASK, fl = a.lo;
  if (be != 0) {
    r.mh = fh | HIDDEN;
    r.ml = fl;
    r.e = (int32_t)be - QBIAS;
    return r;
  }
  if ((fh | fl) == 0) {
    r.mh = 0;
    r.ml = 0;
    r.e = UQ_EZERO;
    return r;
  }
  // subnormal: value = f * 2^(EMIN - 112); normalise f so that bit 112 is set
  const int lz = (fh != 0) ? (clz64(fh) - 15) : (49 + clz64(fl));  // leading zeros within 113 bits
  if (lz >= 64) {
    r.mh = fl << (lz - 64);
    r.ml = 0;
  } else {
    r.mh = shl_fill(fh, fl, lz);
    r.ml = fl << lz;
  }
  r.e = EMIN - lz;
  return r;
}

// Pack a finite value produced by this file's rounding (normal, subnormal or zero).
QB_HD q128 pack(uq v) {
  if ((v.mh | v.ml) == 0) return q_zero(v.s);
  if (v.e >= EMIN) {
    return q_make(v.ml, ((uint64_t)v.s << 63) | ((uint64_t)(v.e + QBIAS) << 48) | (v.mh & FRAC_HI_MASK));
  }
  // subnormal: m has already been rounded to the subnormal grid, shift it down exactly
  const int sh = EMIN - v.e;  // 1..112
  uint64_t fl, fh;
  if (sh >= 64) {
    fl = v.mh >> (sh - 64);
    fh = 0;
  } else {
    fl = shr_fill(v.mh, v.ml, sh);
    fh = v.mh >> sh;
  }
  return q_make(fl, ((uint64_t)v.s << 63) | fh);
}

// ------------------------------------------------------------------------------------------
// rounding of a 5-word (320-bit) magnitude
// ------------------------------------------------------------------------------------------

// Round (-1)^s * W * 2^lsbexp to binary128 (RNE) and return it unpacked.  W != 0.
// 'ovf' is set when the result overflows (the caller turns it into Inf).
QB_HD uq round_w5(uint32_t s, const uint64_t (&W)[5], int lsbexp, bool &ovf) {
  // position of the leading bit
  int msb;
  if (W[4]) msb = 319 - clz64(W[4]);
  else if (W[3]) msb = 255 - clz64(W[3]);
  else if (W[2]) msb = 191 - clz64(W[2]);
  else if (W[1]) msb = 127 - clz64(W[1]);
  else msb = 63 - clz64(W[0]);
  int e = lsbexp + msb;  // exponent of the leading bit
  // left-normalise so the leading bit sits at bit 319 (or, for results below the normal range,
  // so the subnormal ulp 2^(EMIN-112) sits at bit 207)
  int L = 319 - msb;
  const bool sub = e < EMIN;
  if (sub) L -= (EMIN - e);  // may go negative: then shift right with jam
  uint64_t N[5];
  if (L >= 0) {
    mw_shl<5>(N, W, L);
  } else {
    const uint64_t lost = mw_shr_sticky<5>(N, W, -L);
    N[0] |= lost;  // W is the exact final magnitude: round-to-odd at frame bit 0 is safe
  }
  // significand bits 319..207, guard 206, sticky 205..0
  uint64_t sh = N[4] >> 15;
  uint64_t sl = shr_fill(N[4], N[3], 15);
  const uint64_t guard = (N[3] >> 14) & 1;
  const uint64_t sticky = ((N[3] & 0x3fffull) | N[2] | N[1] | N[0]) != 0;
  const uint64_t inc = guard & (sticky | (sl & 1));
  sl += inc;
  sh += (sl == 0) & inc;
  uq r;
  r.s = s;
  if (!sub) {
    if (sh >> 49) {  // carried into 2^113
      sh >>= 1;      // significand becomes exactly 2^112 (sl is 0)
      e += 1;
    }
    r.mh = sh;
    r.ml = sl;
    r.e = e;
  } else {
    // subnormal grid: the kept field is m_sub = (sh:sl) at ulp 2^(EMIN-112); normalise it again
    if ((sh | sl) == 0) {
      r.mh = 0;
      r.ml = 0;
      r.e = UQ_EZERO;
      ovf = false;
      return r;
    }
    const int lz = (sh != 0) ? (clz64(sh) - 15) : (49 + clz64(sl));
    if (lz >= 64) {
      r.mh = sl << (lz - 64);
      r.ml = 0;
    } else {
      r.mh = shl_fill(sh, sl, lz);
      r.ml = sl << lz;
    }
    r.e = EMIN - lz;
  }
  ovf = r.e > EMAX;
  return r;
}

// ------------------------------------------------------------------------------------------
// fused multiply-add on unpacked operands
// ------------------------------------------------------------------------------------------

// 113x113 -> 226-bit exact product of two significands, as four 64-bit words, using 32-bit
// limb multiplies (v_mad_u64_u32).
QB_HD void mul113(uint64_t (&P)[4], uint64_t ah, uint64_t al, uint64_t bh, uint64_t bl) {
  const uint32_t a[4] = {(uint32_t)al, (uint32_t)(al >> 32), (uint32_t)ah, (uint32_t)(ah >> 32)};
  const uint32_t b[4] = {(uint32_t)bl, (uint32_t)(bl >> 32), (uint32_t)bh, (uint32_t)(bh >> 32)};
  uint32_t p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint64_t carry = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t t = (uint64_t)a[i] * b[j] + p[i + j] + carry;
      p[i + j] = (uint32_t)t;
      carry = t >> 32;
    }
    p[i + 4] = (uint32_t)carry;
  }
  P[0] = ((uint64_t)p[1] << 32) | p[0];
  P[1] = ((uint64_t)p[3] << 32) | p[2];
  P[2] = ((uint64_t)p[5] << 32) | p[4];
  P[3] = ((uint64_t)p[7] << 32) | p[6];
}

// 5-word add / subtract
QB_HD void w5_add(uint64_t (&r)[5], const uint64_t (&a)[5], const uint64_t (&b)[5]) {
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const uint64_t s = a[i] + b[i];
    const uint64_t c1 = s < a[i];
    const uint64_t s2 = s + c;
    const uint64_t c2 = s2 < s;
    r[i] = s2;
    c = c1 | c2;
  }
}
// r = a - b - bin; returns the final borrow (1 if a < b + bin)
QB_HD uint64_t w5_sub(uint64_t (&r)[5], const uint64_t (&a)[5], const uint64_t (&b)[5], uint64_t bin = 0) {
  uint64_t br = bin;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const uint64_t d = a[i] - b[i];
    const uint64_t b1 = a[i] < b[i];
    const uint64_t d2 = d - br;
    const uint64_t b2 = d < br;
    r[i] = d2;
    br = b1 | b2;
  }
  return br;
}
QB_HD void w5_neg(uint64_t (&r)[5]) {
  uint64_t c = 1;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const uint64_t v = ~r[i] + c;
    c = (v == 0) & c;
    r[i] = v;
  }
}

// Frame used by fma_uq: 5 words = 320 bits.
//   The product P = ma*mb (226 bits) sits at bits [0, 226) with weight 2^(ea+eb-224) on bit 0.
//   The addend's significand lands at bit offset d = ec - ea - eb + 112:
//     0 <= d <= 206 : exact placement (its leading bit at <= bit 318), exact sum;
//     d < 0         : the addend lies >= 113 bits below P's leading bit; it is truncated and the
//                     lost bits become a sticky that is applied to the sum (round-to-odd);
//     d > 206       : the addend dominates; the frame moves up by d-206 and P is truncated
//                     instead, with the same sticky treatment.
//   In the two inexact cases the larger operand is known, the difference is formed as
//   larger - truncated - sticky and the sticky is ORed into bit 0: the frame value is then the
//   exact result rounded to odd at frame bit 0, >= 100 bits below the rounding position, so the
//   final round-to-nearest-even is correct.
// Result: correctly rounded a*b + c.  Requires a, b finite nonzero; c finite (zero allowed).
// 'ovf' reports overflow (the caller produces Inf).
QB_HD uq fma_uq(const uq &a, const uq &b, const uq &c, bool &ovf) {
  uint64_t P[4];
  mul113(P, a.mh, a.ml, b.mh, b.ml);
  const uint32_t sp = a.s ^ b.s;
  int lsbexp = a.e + b.e - 224;
  uint64_t X[5] = {P[0], P[1], P[2], P[3], 0};
  uint64_t Y[5] = {c.ml, c.mh, 0, 0, 0};
  const bool czero = (c.mh | c.ml) == 0;
  const int d = czero ? -1000 : (c.e - a.e - b.e + 112);
  uint64_t R[5];
  uint32_t s = sp;
  if (d > 206) {
    // addend dominates: Y = c << 206 (206 = 3*64 + 14), X truncated
    const int sh = d - 206;
    uint64_t T[5];
    const uint64_t lost = mw_shr_sticky<5>(T, X, sh);
    Y[4] = shl_fill(c.mh, c.ml, 14);
    Y[3] = c.ml << 14;
    Y[2] = 0;
    Y[1] = 0;
    Y[0] = 0;
    lsbexp += sh;
    s = c.s;
    if (sp == c.s) w5_add(R, Y, T);
    else w5_sub(R, Y, T, lost);
    R[0] |= lost;
  } else if (d >= 0) {
    uint64_t T[5];
    mw_shl<5>(T, Y, d);
    if (sp == c.s) {
      w5_add(R, X, T);
    } else {
      const uint64_t borrow = w5_sub(R, X, T);
      if (borrow) {
        w5_neg(R);
        s = c.s;
      }
      if ((R[0] | R[1] | R[2] | R[3] | R[4]) == 0) {  // exact cancellation: +0 (round to nearest)
        uq z;
        z.mh = 0;
        z.ml = 0;
        z.e = UQ_EZERO;
        z.s = 0;
        ovf = false;
        return z;
      }
    }
  } else {
    // product dominates: addend truncated
    uint64_t T[5];
    const uint64_t lost = mw_shr_sticky<5>(T, Y, -d);
    if (sp == c.s) w5_add(R, X, T);
    else w5_sub(R, X, T, lost);
    R[0] |= lost;
  }
  return round_w5(s, R, lsbexp, ovf);
}

// ------------------------------------------------------------------------------------------
// full IEEE operations on packed values
// ------------------------------------------------------------------------------------------

QB_HD q128 fma_packed(q128 a, q128 b, q128 c) {
  const uint32_t ea = q_bexp(a), eb = q_bexp(b), ec = q_bexp(c);
  const bool sa = (ea == 0) | (ea == 0x7fffu), sb = (eb == 0) | (eb == 0x7fffu), sc = ec == 0x7fffu;
  const uint32_t sp = (uint32_t)((a.hi ^ b.hi) >> 63);
  if (sa | sb | sc) {
    if (q_isnan(a) | q_isnan(b) | q_isnan(c)) return q_nan();
    const bool ia = q_isinf(a), ib = q_isinf(b);
    const bool za = q_iszero(a), zb = q_iszero(b);
    if (ia | ib) {
      if (za | zb) return q_nan();
      if (q_isinf(c) && (uint32_t)(c.hi >> 63) != sp) return q_nan();
      return q_inf(sp);
    }
    if (q_isinf(c)) return c;
    if (za | zb) {
      if (q_iszero(c)) return q_zero(sp & (uint32_t)(c.hi >> 63));
      return c;
    }
  }
  const uq A = unpack_finite(a), B = unpack_finite(b), C = unpack_finite(c);
  bool ovf = false;
  uq r = fma_uq(A, B, C, ovf);
  if (ovf) return q_inf(r.s);
  // a zero here is either exact cancellation (+0, r.s == 0) or an underflow to zero, which
  // keeps the sign of the exact result (r.s)
  return pack(r);
}

QB_HD q128 add_packed(q128 a, q128 b) { return fma_packed(a, q_one(), b); }
QB_HD q128 mul_packed(q128 a, q128 b) { return fma_packed(a, b, q_zero(1)); }

// Accumulating FMA for chains kept unpacked: acc = a*b + acc with a, b packed.  Falls back to the
// packed rules for zeros, subnormals, Inf and NaN; 'accp' mirrors acc in packed form only when
// the chain has left the finite range (acc_special set), which is sticky.
struct chain {
  uq acc;
  q128 special_val;
  bool special;  // acc is Inf/NaN: the value is special_val
};

QB_HD void chain_init(chain &ch) {
  ch.acc.mh = 0;
  ch.acc.ml = 0;
  ch.acc.e = UQ_EZERO;
  ch.acc.s = 0;
  ch.special = false;
  ch.special_val = q_zero(0);
}
QB_HD void chain_set(chain &ch, q128 v) {
  if (q_bexp(v) == 0x7fffu) {
    ch.special = true;
    ch.special_val = v;
    ch.acc.mh = 0;
    ch.acc.ml = 0;
    ch.acc.e = UQ_EZERO;
    ch.acc.s = 0;
  } else {
    ch.special = false;
    ch.acc = unpack_finite(v);
  }
}
QB_HD q128 chain_get(const chain &ch) {
  if (ch.special) return ch.special_val;
  return pack(ch.acc);
}
// Note on zero accumulators: an unpacked zero keeps its sign in acc.s, so pack() restores -0.
QB_HD void chain_fma(chain &ch, q128 a, q128 b) {
  bool spa, spb;
  const uq A = unpack_fast(a, spa);
  const uq B = unpack_fast(b, spb);
  if (!(spa | spb | ch.special)) {
    bool ovf = false;
    uq r = fma_uq(A, B, ch.acc, ovf);
    if (!ovf) {
      ch.acc = r;
      return;
    }
  }
  chain_set(ch, fma_packed(a, b, chain_get(ch)));
}
// Same with an unpacked, known finite nonzero normal 'b' (e.g. a value staged once).
QB_HD void chain_fma_ub(chain &ch, q128 a, const uq &B, q128 bpacked) {
  bool spa;
  const uq A = unpack_fast(a, spa);
  if (!(spa | ch.special)) {
    bool ovf = false;
    uq r = fma_uq(A, B, ch.acc, ovf);
    if (!ovf) {
      ch.acc = r;
      return;
    }
  }
  chain_set(ch, fma_packed(a, bpacked, chain_get(ch)));
}

// ------------------------------------------------------------------------------------------
// conversions
// ------------------------------------------------------------------------------------------

// double -> binary128 (exact)
QB_HD q128 from_double(double d) {
  uint64_t u;
  __builtin_memcpy(&u, &d, 8);
  const uint64_t s = u >> 63;
  const uint32_t be = (uint32_t)(u >> 52) & 0x7ffu;
  uint64_t f = u & 0x000fffffffffffffull;
  if (be == 0x7ffu) {
    if (f) return q_nan();
    return q_inf((uint32_t)s);
  }
  if (be == 0) {
    if (!f) return q_zero((uint32_t)s);
    // subnormal double: value = f * 2^-1074; normalise into the binary128 normal range
    const int lz = clz64(f) - 11;  // leading zeros within 53 bits
    f <<= lz;                      // bit 52 set
    const int e = -1022 - lz;      // exponent of leading bit
    f &= 0x000fffffffffffffull;
    return q_make(f << 60, (s << 63) | ((uint64_t)(e + QBIAS) << 48) | (f >> 4));
  }
  const int e = (int)be - 1023;
  return q_make(f << 60, (s << 63) | ((uint64_t)(e + QBIAS) << 48) | (f >> 4));
}

// binary128 -> double, round to nearest even (subnormal doubles, overflow to Inf, NaN kept quiet)
QB_HD double to_double(q128 a) {
  const uint64_t s = a.hi >> 63;
  const uint32_t be = q_bexp(a);
  const uint64_t fh = a.hi & FRAC_HI_MASK, fl = a.lo;
  uint64_t bits;
  if (be == 0x7fffu) {
    bits = (s << 63) | 0x7ff0000000000000ull | (((fh | fl) != 0) ? 0x0008000000000000ull : 0);
  } else if (be == 0) {
    bits = s << 63;  // binary128 zero or subnormal: far below the double range
  } else {
    const int e = (int)be - QBIAS;  // value = m * 2^(e-112), m = 2^112 + f
    if (e > 1023) {
      bits = (s << 63) | 0x7ff0000000000000ull;
    } else {
      // kept significand: m >> sh; normal doubles keep 53 bits (sh = 60), subnormals fewer
      int sh = 60;
      if (e < -1022) sh += (-1022 - e);
      const uint64_t mh = fh | HIDDEN, ml = fl;
      uint64_t kept, guard, sticky;
      if (sh >= 114) {
        kept = 0;
        guard = 0;
        sticky = 1;  // m != 0 lies entirely below the guard position
        if (sh == 114) guard = 0;
      } else if (sh >= 64) {
        const int r = sh - 64;  // 0..49
        kept = mh >> r;
        guard = (r == 0) ? (ml >> 63) : ((mh >> (r - 1)) & 1);
        const uint64_t below_mh = (r <= 1) ? 0 : (mh & (~0ull >> (65 - r)));
        const uint64_t below_ml = (r == 0) ? (ml & 0x7fffffffffffffffull) : ml;
        sticky = (below_mh | below_ml) != 0;
      } else {
        kept = shr_fill(mh, ml, sh);  // sh in 60..63 -> fits
        guard = (ml >> (sh - 1)) & 1;
        sticky = (ml & (~0ull >> (65 - sh))) != 0;
      }
      kept += guard & (sticky | (kept & 1));
      if (e >= -1022) {
        int ee = e;
        if (kept >> 53) {
          kept >>= 1;
          ee += 1;
        }
        bits = (ee > 1023) ? ((s << 63) | 0x7ff0000000000000ull)
                           : ((s << 63) | ((uint64_t)(ee + 1023) << 52) | (kept & 0x000fffffffffffffull));
      } else {
        // subnormal double (kept < 2^52), or rounded up into the smallest normal (kept == 2^52)
        bits = (s << 63) | kept;
      }
    }
  }
  double d;
  __builtin_memcpy(&d, &bits, 8);
  return d;
}

// Correctly rounded square root (round to nearest even; Sleef_sqrtq1_u05's contract):
// sqrt(+-0) = +-0, sqrt(+inf) = +inf, negative or NaN -> NaN.  With value = m 2^(e-112),
// X = m 2^t (t = 116 or 117, so e - 112 - t is even and X has 229..230 bits) and the 115-bit
// integer root q = floor(sqrt(X)) from the digit-by-digit method; the significand is q >> 2,
// rounded with guard bit 1 of q and a sticky bit from bit 0 and the remainder X - q^2.
QB_HD q128 sqrt_packed(q128 a) {
  const uint32_t be = q_bexp(a);
  const uint32_t s = (uint32_t)(a.hi >> 63);
  if (be == 0x7fffu) {
    if (q_isnan(a) || s) return q_nan();
    return a;
  }
  if (q_iszero(a)) return a;
  if (s) return q_nan();
  const uq u = unpack_finite(a);
  const int t = 116 + (u.e & 1);
  const int bs = t - 64;  // 52 or 53
  uint64_t X[4];
  X[0] = 0;
  X[1] = u.ml << bs;
  X[2] = (u.mh << bs) | (u.ml >> (64 - bs));
  X[3] = u.mh >> (64 - bs);
  uint64_t r0 = 0, r1 = 0;          // root
  uint64_t m0 = 0, m1 = 0, m2 = 0;  // remainder
  for (int i = 114; i >= 0; --i) {
    const int b = 2 * i;
    const uint64_t two = (X[b >> 6] >> (b & 63)) & 3u;
    m2 = (m2 << 2) | (m1 >> 62);
    m1 = (m1 << 2) | (m0 >> 62);
    m0 = (m0 << 2) | two;
    const uint64_t t2 = r1 >> 62, t1 = (r1 << 2) | (r0 >> 62), t0 = (r0 << 2) | 1u;
    const bool ge = (m2 != t2) ? (m2 > t2) : ((m1 != t1) ? (m1 > t1) : (m0 >= t0));
    r1 = (r1 << 1) | (r0 >> 63);
    r0 = (r0 << 1) | (ge ? 1u : 0u);
    if (ge) {
      const uint64_t d0 = m0 - t0, b0 = m0 < t0;
      const uint64_t d1 = m1 - t1 - b0, b1 = (m1 < t1) || (m1 - t1 < b0);
      m2 = m2 - t2 - b1;
      m1 = d1;
      m0 = d0;
    }
  }
  uint64_t ml = (r0 >> 2) | (r1 << 62), mh = r1 >> 2;
  const uint64_t guard = (r0 >> 1) & 1u;
  const uint64_t sticky = (r0 & 1u) | ((m0 | m1 | m2) != 0);
  const uint64_t inc = guard & (sticky | (ml & 1u));
  ml += inc;
  mh += (ml == 0) & inc;
  int er = 114 + (u.e - 112 - t) / 2;
  if (mh >> 49) {
    mh >>= 1;
    er += 1;
  }
  return q_make(ml, ((uint64_t)(er + QBIAS) << 48) | (mh & FRAC_HI_MASK));
}

}  // namespace qb
