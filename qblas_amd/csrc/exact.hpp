// exact.hpp -- exact (correctly rounded) accumulation of binary128 products.
//
// EXACT form of dot and gemv: the result is the exact value of alpha * sum_i a_i b_i + beta * y
// rounded once to binary128 (round to nearest even) -- the correctly rounded answer, independent of
// summation order and thread count.  The reference rounds after every fused multiply-add
// (level1.hpp:20-32); this form has no intermediate rounding at all.
//
// Stage 1 (streaming, per lane): every product a_i*b_i is formed exactly (226 bits, mul113) and
//   added into a 384-bit two's complement fixed-point window W (12 x u32) whose bit 0 has weight
//   2^A.  A product with lsb exponent e_p lands at bit offset o = e_p - A; it fits when
//   1 <= o <= 126, leaving 32 bits of headroom for carries (<= 2^32 additions per window).  Each
//   lane keeps two such windows, 124 bits apart (win2_add_product), anchored on the centre of the
//   first products of its wave, so products within +-124 binary orders of magnitude of that centre
//   are exact.  A product outside both windows is not added and sets the LOST flag: the dot, row or
//   entry is then recomputed on the full-range path (wide.hpp), which holds every binary128 product.
// Stage 2 (combine): windows with possibly different anchors are summed exactly into a wide
//   accumulator (ACCW words), then alpha*S + beta*y is formed exactly as big integers and rounded
//   once.
#pragma once
#include <climits>

#include "q128.hpp"

namespace qb {

constexpr int WW = 12;          // window words
constexpr int WIN_OMIN = 1;     // valid product offsets [WIN_OMIN, WIN_OMAX]
constexpr int WIN_OMAX = 126;
constexpr int WIN_CENTER = 63;  // anchor = centre exponent - WIN_CENTER

enum : uint32_t { FL_LOST = 1u, FL_NAN = 2u, FL_PINF = 4u, FL_NINF = 8u };

QB_HD uint32_t alignbit32(uint32_t hi, uint32_t lo, uint32_t s) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, s);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> (s & 31));
#endif
}

// 113x113 -> 226-bit product as eight 32-bit words (p[7] holds bits 224..225)
QB_HD void mul113_w(uint32_t (&p)[8], const uq &a, const uq &b) {
  const uint32_t x[4] = {(uint32_t)a.ml, (uint32_t)(a.ml >> 32), (uint32_t)a.mh, (uint32_t)(a.mh >> 32)};
  const uint32_t y[4] = {(uint32_t)b.ml, (uint32_t)(b.ml >> 32), (uint32_t)b.mh, (uint32_t)(b.mh >> 32)};
#pragma unroll
  for (int i = 0; i < 8; ++i) p[i] = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint64_t carry = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t t = (uint64_t)x[i] * y[j] + ((uint64_t)p[i + j] + carry);  // < 2^64
      p[i + j] = (uint32_t)t;
      carry = t >> 32;
    }
    p[i + 4] = (uint32_t)carry;
  }
}

// product lsb exponent: value = P * 2^(a.e + b.e - 224)
QB_HD int prod_lsb(const uq &a, const uq &b) { return a.e + b.e - 224; }

QB_HD uint32_t addc32(uint32_t a, uint32_t b, uint32_t cin, uint32_t &cout) {
#if defined(__HIP_DEVICE_COMPILE__)
  unsigned int co;
  const unsigned int r = __builtin_addc(a, b, cin, &co);
  cout = co;
  return r;
#else
  const uint64_t s = (uint64_t)a + b + cin;
  cout = (uint32_t)(s >> 32);
  return (uint32_t)s;
#endif
}

// The product a.m*b.m shifted to bit offset o (1..126) of a 12-word frame, with the sign applied as
// one's complement (X ^ mask) plus a carry-in (mask & 1): adding the returned words with that
// carry-in adds (-1)^(a.s^b.s) * a.m*b.m * 2^o.
QB_HD void place_product(uint32_t (&X)[WW], uint32_t &mask, int o, const uq &a, const uq &b) {
  uint32_t P[8];
  mul113_w(P, a, b);
  const int w = (o - 1) >> 5;                          // 0..3
  const uint32_t sh = 32u - (uint32_t)(o - (w << 5));  // bit offset 1..32 -> funnel 31..0
  uint32_t T[9];
  T[0] = alignbit32(P[0], 0u, sh);
#pragma unroll
  for (int j = 1; j < 8; ++j) T[j] = alignbit32(P[j], P[j - 1], sh);
  T[8] = alignbit32(0u, P[7], sh);
  mask = 0u - ((a.s ^ b.s) & 1u);
#pragma unroll
  for (int i = 0; i < WW; ++i) {
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int j = i - k;
      if (j >= 0 && j < 9) v = (w == k) ? T[j] : v;
    }
    X[i] = v ^ mask;
  }
}

// W += (-1)^(a.s^b.s) * a.m*b.m * 2^(e_p - A).  Returns false (and adds nothing) if the product
// does not fit the window.
QB_HD bool win_add_product(uint32_t (&W)[WW], int A, const uq &a, const uq &b) {
  const int o = prod_lsb(a, b) - A;
  if (o < WIN_OMIN || o > WIN_OMAX) return false;
  uint32_t X[WW], mask;
  place_product(X, mask, o, a, b);
  uint32_t c = mask & 1u;
#pragma unroll
  for (int i = 0; i < WW; ++i) W[i] = addc32(W[i], X[i], c, c);
  return true;
}

// Two windows per lane: lo anchored at A, hi at A + WIN_BAND.  Together they hold products with
// lsb offsets o = e_p - A in [1, WIN_BAND + WIN_OMAX] = [1, 250] exactly, i.e. a spread of 249
// binary orders of magnitude around the centre (anchor A = centre - WIN2_CENTER).
constexpr int WIN_BAND = 124;
constexpr int WIN2_CENTER = 125;
QB_HD bool win2_add_product(uint32_t (&Wl)[WW], uint32_t (&Wh)[WW], int A, const uq &a, const uq &b) {
  int o = prod_lsb(a, b) - A;
  const bool hi = o > WIN_OMAX;
  if (hi) o -= WIN_BAND;
  if (o < WIN_OMIN || o > WIN_OMAX) return false;
  uint32_t X[WW], mask;
  place_product(X, mask, o, a, b);
  // lo window: always updated, the addend masked off for hi products (both windows stay in
  // registers: no data-dependent choice between arrays)
  const uint32_t keep = hi ? 0u : ~0u;
  uint32_t c = mask & keep & 1u;
#pragma unroll
  for (int i = 0; i < WW; ++i) Wl[i] = addc32(Wl[i], X[i] & keep, c, c);
  if (hi) {  // skipped by every wave whose products all fit the lo window
    uint32_t ch = mask & 1u;
#pragma unroll
    for (int i = 0; i < WW; ++i) Wh[i] = addc32(Wh[i], X[i], ch, ch);
  }
  return true;
}

QB_HD void win_add_win(uint32_t (&W)[WW], const uint32_t (&X)[WW]) {
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < WW; ++i) W[i] = addc32(W[i], X[i], c, c);
}

// Classify a packed operand pair for the streaming loop.  Returns 0 for a regular pair (both
// normal nonzero: ua/ub filled), 1 if the product is exactly zero (skip), 2 for a pair involving
// Inf/NaN (flags updated), 3 if a subnormal is involved (ua/ub filled via the general unpack).
QB_HD int classify_pair(q128 a, q128 b, uq &ua, uq &ub, uint32_t &flags) {
  bool spa, spb;
  ua = unpack_fast(a, spa);
  ub = unpack_fast(b, spb);
  if (!(spa | spb)) return 0;
  if (q_isnan(a) | q_isnan(b)) {
    flags |= FL_NAN;
    return 2;
  }
  const bool ia = q_isinf(a), ib = q_isinf(b);
  if (ia | ib) {
    if (q_iszero(a) | q_iszero(b)) {
      flags |= FL_NAN;  // Inf * 0
    } else {
      flags |= ((a.hi ^ b.hi) >> 63) ? FL_NINF : FL_PINF;
    }
    return 2;
  }
  if (q_iszero(a) | q_iszero(b)) return 1;
  ua = unpack_finite(a);
  ub = unpack_finite(b);
  return 3;
}

// ------------------------------------------------------------------------------------------
// stage 2: wide accumulator and exact epilogue (plain loops; runs once per output)
// ------------------------------------------------------------------------------------------

constexpr int ACCW = 40;  // accumulator words (1280 bits) for combining windows
constexpr int BIGW = 56;  // big magnitude words for alpha*S + beta*y

struct wide_acc {
  uint32_t w[ACCW];  // two's complement, bit 0 weight 2^base
  int base;
  bool any;          // a nonzero window has been added
  bool lost;
};

QB_HD void wide_init(wide_acc &acc) {
  for (int i = 0; i < ACCW; ++i) acc.w[i] = 0;
  acc.base = 0;
  acc.any = false;
  acc.lost = false;
}

QB_HD bool win_is_zero(const uint32_t *W) {
  uint32_t o = 0;
  for (int i = 0; i < WW; ++i) o |= W[i];
  return o == 0;
}

// Add window W (anchor A) exactly.  The first nonzero window fixes base = A - 320 (room for
// windows up to 10 words lower and ~16 words higher).  A window that does not fit sets lost.
QB_HD void wide_add_window(wide_acc &acc, const uint32_t *W, int A) {
  if (win_is_zero(W)) return;
  if (!acc.any) {
    acc.any = true;
    acc.base = A - 320;
  }
  const int off = A - acc.base;
  if (off < 0 || off + WW * 32 + 32 > ACCW * 32) {
    acc.lost = true;
    return;
  }
  const int ow = off >> 5, ob = off & 31;
  const uint32_t sext = (W[WW - 1] >> 31) ? 0xffffffffu : 0u;
  // shifted, sign-extended copy added word by word
  uint32_t carry = 0;
  for (int i = ow; i < ACCW; ++i) {
    const int j = i - ow;  // source word index of the shifted value
    const uint32_t cur = (j < WW) ? W[j] : sext;
    const uint32_t prev = (j == 0) ? 0u : ((j - 1 < WW) ? W[j - 1] : sext);
    const uint32_t v = ob ? ((cur << ob) | (prev >> (32 - ob))) : cur;
    const uint64_t s = (uint64_t)acc.w[i] + v + carry;
    acc.w[i] = (uint32_t)s;
    carry = (uint32_t)(s >> 32);
  }
}

// Big magnitude: value = (-1)^s * m * 2^lsb, m over BIGW words.
struct bigmag {
  uint32_t m[BIGW];
  int lsb;
  uint32_t s;
};

QB_HD void big_zero(bigmag &b) {
  for (int i = 0; i < BIGW; ++i) b.m[i] = 0;
  b.lsb = 0;
  b.s = 0;
}
QB_HD int big_msb(const bigmag &b) {
  for (int i = BIGW - 1; i >= 0; --i)
    if (b.m[i]) return i * 32 + 31 - clz64((uint64_t)b.m[i]) + 32;
  return -1;
}
QB_HD uint32_t big_bit(const bigmag &b, int pos) {
  if (pos < 0 || pos >= BIGW * 32) return 0;
  return (b.m[pos >> 5] >> (pos & 31)) & 1u;
}
QB_HD bool big_any_below(const bigmag &b, int pos) {
  if (pos <= 0) return false;
  const int wi = pos >> 5, bi = pos & 31;
  for (int i = 0; i < wi && i < BIGW; ++i)
    if (b.m[i]) return true;
  if (wi < BIGW && bi && (b.m[wi] & ((1u << bi) - 1u))) return true;
  return false;
}
// 32 bits of b starting at bit pos
QB_HD uint32_t big_get32(const bigmag &b, int pos) {
  const int wi = pos >> 5, bi = pos & 31;
  const uint32_t lo = (wi >= 0 && wi < BIGW) ? b.m[wi] : 0u;
  const uint32_t hi = (wi + 1 >= 0 && wi + 1 < BIGW) ? b.m[wi + 1] : 0u;
  return bi ? ((lo >> bi) | (hi << (32 - bi))) : lo;
}

// Round b to binary128 (round to nearest even), with subnormal and overflow handling.  The
// magnitude is exact, or rounded to odd at bit 0 with bit 0 at least 2 positions below the
// rounding position.
QB_HD q128 big_round(const bigmag &b) {
  const int msb = big_msb(b);
  if (msb < 0) return q_zero(b.s);
  int e = b.lsb + msb;
  int keep = msb - 112;
  if (e < EMIN) keep += EMIN - e;
  // gather 113 bits [keep, keep+112]
  uint64_t sl = ((uint64_t)big_get32(b, keep + 32) << 32) | big_get32(b, keep);
  uint64_t sh = ((uint64_t)big_get32(b, keep + 96) << 32) | big_get32(b, keep + 64);
  sh &= (1ull << 49) - 1;
  if (keep + 112 > msb) {
    // fewer than 113 bits kept (subnormal range): bits above msb are zero already
  }
  const uint32_t guard = big_bit(b, keep - 1);
  const bool sticky = big_any_below(b, keep - 1);
  const uint64_t inc = guard & ((sticky ? 1u : 0u) | (uint32_t)(sl & 1));
  sl += inc;
  sh += (sl == 0) & inc;
  int ulp = b.lsb + keep;
  if (sh >> 49) {
    sh >>= 1;
    ulp += 1;
  }
  if (sh >> 48) {
    const int be = ulp + 112 + QBIAS;
    if (be >= 0x7fff) return q_inf(b.s);
    return q_make(sl, ((uint64_t)b.s << 63) | ((uint64_t)be << 48) | (sh & FRAC_HI_MASK));
  }
  return q_make(sl, ((uint64_t)b.s << 63) | sh);  // subnormal
}

// wide accumulator -> signed big magnitude
QB_HD void wide_to_big(const wide_acc &acc, bigmag &b) {
  big_zero(b);
  b.lsb = acc.base;
  const bool neg = (acc.w[ACCW - 1] >> 31) != 0;
  b.s = neg ? 1u : 0u;
  uint32_t carry = 1;
  for (int i = 0; i < ACCW; ++i) {
    uint32_t v = acc.w[i];
    if (neg) {
      const uint64_t t = (uint64_t)(~v) + carry;
      v = (uint32_t)t;
      carry = (uint32_t)(t >> 32);
    }
    b.m[i] = v;
  }
}

// r = b * a.m (113-bit significand), exponent/sign adjusted; b must fit BIGW-4 words
QB_HD void big_mul_q(bigmag &r, const bigmag &b, const uq &a) {
  const uint32_t x[4] = {(uint32_t)a.ml, (uint32_t)(a.ml >> 32), (uint32_t)a.mh, (uint32_t)(a.mh >> 32)};
  bigmag t;
  big_zero(t);
  for (int i = 0; i < BIGW - 4; ++i) {
    const uint32_t bi = b.m[i];
    if (!bi) continue;
    uint64_t carry = 0;
    for (int j = 0; j < 4; ++j) {
      const uint64_t v = (uint64_t)bi * x[j] + t.m[i + j] + carry;
      t.m[i + j] = (uint32_t)v;
      carry = v >> 32;
    }
    for (int k = i + 4; carry && k < BIGW; ++k) {
      const uint64_t v = (uint64_t)t.m[k] + carry;
      t.m[k] = (uint32_t)v;
      carry = v >> 32;
    }
  }
  t.lsb = b.lsb + a.e - 112;
  t.s = b.s ^ a.s;
  r = t;
}

QB_HD bool big_is_zero(const bigmag &b) {
  for (int i = 0; i < BIGW; ++i)
    if (b.m[i]) return false;
  return true;
}

// dst = src * 2^sh over BIGW words (sh may be negative: bits shifted below bit 0 are dropped and
// reported by the return value).  Bits shifted above the top word are dropped (callers size the
// frame so that never happens).
QB_HD bool big_shift(bigmag &dst, const bigmag &src, int sh) {
  bool lost = false;
  for (int i = 0; i < BIGW; ++i) dst.m[i] = 0;
  if (sh >= 0) {
    const int ws = sh >> 5, bs = sh & 31;
    for (int i = BIGW - 1; i >= ws; --i) {
      const int j = i - ws;
      const uint32_t cur = src.m[j];
      const uint32_t prev = (j > 0) ? src.m[j - 1] : 0u;
      dst.m[i] = bs ? ((cur << bs) | (prev >> (32 - bs))) : cur;
    }
  } else {
    const int r = -sh;
    const int ws = r >> 5, bs = r & 31;
    for (int i = 0; i < BIGW; ++i) {
      const int j = i + ws;
      const uint32_t cur = (j < BIGW) ? src.m[j] : 0u;
      const uint32_t next = (j + 1 < BIGW) ? src.m[j + 1] : 0u;
      dst.m[i] = bs ? ((cur >> bs) | (next << (32 - bs))) : cur;
    }
    for (int i = 0; i < ws && i < BIGW; ++i)
      if (src.m[i]) lost = true;
    if (ws < BIGW && bs && (src.m[ws] & ((1u << bs) - 1u))) lost = true;
  }
  return lost;
}

// Correctly rounded (X + Y) for two exact big magnitudes.  The larger operand is placed with its
// leading bit at frame bit BIGW*32-2 (exactly: its width is at most BIGW*32-2 bits); the smaller
// may lose low bits, which then act as a sticky: the difference is formed as larger - truncated
// smaller - 1 and bit 0 is set (the exact result rounded to odd at bit 0, far below the rounding
// position), which keeps the final round-to-nearest-even correct.
QB_HD q128 big_add_round(const bigmag &X, const bigmag &Y) {
  if (big_is_zero(Y)) return big_round(X);
  if (big_is_zero(X)) return big_round(Y);
  const int mx = big_msb(X), my = big_msb(Y);
  const int tx = X.lsb + mx, ty = Y.lsb + my;
  const bool xl = tx >= ty;
  const bigmag &L = xl ? X : Y;
  const bigmag &S = xl ? Y : X;
  const int ml = xl ? mx : my;
  const int shL = (BIGW * 32 - 2) - ml;
  const int flsb = L.lsb - shL;
  bigmag F, G;
  big_shift(F, L, shL);
  const bool lostS = big_shift(G, S, S.lsb - flsb);
  bigmag R;
  R.lsb = flsb;
  if (L.s == S.s) {
    uint64_t c = 0;
    for (int i = 0; i < BIGW; ++i) {
      const uint64_t v = (uint64_t)F.m[i] + G.m[i] + c;
      R.m[i] = (uint32_t)v;
      c = v >> 32;
    }
    R.s = L.s;
  } else {
    int cmp = 0;
    for (int i = BIGW - 1; i >= 0 && !cmp; --i)
      if (F.m[i] != G.m[i]) cmp = F.m[i] > G.m[i] ? 1 : -1;
    if (cmp == 0 && !lostS) return q_zero(0);  // exact cancellation: +0
    if (cmp < 0) {
      // only possible without loss (the smaller operand lost bits only when far below)
      uint64_t br = 0;
      for (int i = 0; i < BIGW; ++i) {
        const uint64_t v = (uint64_t)G.m[i] - F.m[i] - br;
        R.m[i] = (uint32_t)v;
        br = (v >> 63) & 1u;
      }
      R.s = S.s;
    } else {
      uint64_t br = lostS ? 1u : 0u;
      for (int i = 0; i < BIGW; ++i) {
        const uint64_t v = (uint64_t)F.m[i] - G.m[i] - br;
        R.m[i] = (uint32_t)v;
        br = (v >> 63) & 1u;
      }
      R.s = L.s;
    }
  }
  if (lostS) R.m[0] |= 1u;
  return big_round(R);
}

// value of a sum that met Inf/NaN products (IEEE rules for the exact sum)
QB_HD q128 special_result(uint32_t flags) {
  if ((flags & FL_NAN) || ((flags & FL_PINF) && (flags & FL_NINF))) return q_nan();
  if (flags & FL_PINF) return q_inf(0);
  return q_inf(1);
}

// Combine a row's / the dot's exact sum S with alpha, beta, y:  RN(alpha*S + beta*y).
// has_ab = false means the plain rounded S (dot).
QB_HD q128 finish_value(const wide_acc &acc, uint32_t flags, bool has_ab, q128 alpha, q128 beta, q128 yv) {
  const bool sflag = (flags & (FL_NAN | FL_PINF | FL_NINF)) != 0;
  bigmag S;
  if (acc.any) {
    wide_to_big(acc, S);
  } else {
    big_zero(S);
  }
  if (!has_ab) {
    if (sflag) return special_result(flags);
    return acc.any ? big_round(S) : q_zero(0);
  }
  const bool abspecial = (q_bexp(alpha) == 0x7fffu) || (q_bexp(beta) == 0x7fffu) || (q_bexp(yv) == 0x7fffu);
  if (sflag || abspecial) {
    // IEEE special rules through the packed ops: the result is NaN/Inf whatever the rounding
    const q128 sq = sflag ? special_result(flags) : (acc.any ? big_round(S) : q_zero(0));
    return fma_packed(alpha, sq, mul_packed(beta, yv));
  }
  bigmag T1, T2;
  big_zero(T1);
  big_zero(T2);
  const uq ua = unpack_finite(alpha), ub = unpack_finite(beta), uy = unpack_finite(yv);
  if (acc.any && (ua.mh | ua.ml)) big_mul_q(T1, S, ua);
  if ((ub.mh | ub.ml) && (uy.mh | uy.ml)) {
    bigmag Y;
    big_zero(Y);
    Y.m[0] = (uint32_t)uy.ml;
    Y.m[1] = (uint32_t)(uy.ml >> 32);
    Y.m[2] = (uint32_t)uy.mh;
    Y.m[3] = (uint32_t)(uy.mh >> 32);
    Y.lsb = uy.e - 112;
    Y.s = uy.s;
    big_mul_q(T2, Y, ub);
  }
  const bool z1 = big_is_zero(T1), z2 = big_is_zero(T2);
  if (z1 && z2) {
    // exact zero: IEEE sign of (alpha*S) + (beta*y) with zero terms
    const uint32_t s1 = (alpha.hi >> 63) ^ (acc.any ? S.s : 0u);
    const uint32_t s2 = (uint32_t)((beta.hi ^ yv.hi) >> 63);
    return q_zero(s1 & s2);
  }
  return big_add_round(T1, T2);
}


// ------------------------------------------------------------------------------------------
// register-resident epilogue: fixed-size word arrays, every index a compile-time constant
// (variable shifts are done with selects), so nothing is spilled to scratch memory
// ------------------------------------------------------------------------------------------

// r = x << sh (N words, bits above the top dropped), 0 <= sh
template <int N>
QB_HD void rshl(uint32_t (&r)[N], const uint32_t (&x)[N], int sh) {
  uint32_t t[N];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = x[i];
  const int ws = sh >> 5, bs = sh & 31;
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1) {
    if (step >= N) continue;
    const bool on = (ws & step) != 0;
#pragma unroll
    for (int i = N - 1; i >= 0; --i) t[i] = on ? ((i >= step) ? t[i - step] : 0u) : t[i];
  }
  const bool big = ws >= N;
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    const uint32_t lo = (i > 0) ? t[i - 1] : 0u;
    const uint32_t v = bs ? ((t[i] << bs) | (lo >> (32 - bs))) : t[i];
    r[i] = big ? 0u : v;
  }
}

// r = x >> sh (truncating), returns 1 if a nonzero bit was shifted out; any sh >= 0
template <int N>
QB_HD uint32_t rshr(uint32_t (&r)[N], const uint32_t (&x)[N], int sh) {
  if (sh >= 32 * N) {
    uint32_t any = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      any |= x[i];
      r[i] = 0;
    }
    return any != 0;
  }
  const int ws = sh >> 5, bs = sh & 31;
  uint32_t lost = 0;
  // bits lost: words below ws and the low bs bits of word ws
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t m = (i < ws) ? 0xffffffffu : ((i == ws && bs) ? (0xffffffffu >> (32 - bs)) : 0u);
    lost |= x[i] & m;
  }
  uint32_t t[N];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = x[i];
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1) {
    if (step >= N) continue;
    const bool on = (ws & step) != 0;
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = on ? ((i + step < N) ? t[i + step] : 0u) : t[i];
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t hi = (i + 1 < N) ? t[i + 1] : 0u;
    r[i] = bs ? ((t[i] >> bs) | (hi << (32 - bs))) : t[i];
  }
  return lost != 0;
}

template <int N>
QB_HD int rmsb(const uint32_t (&x)[N]) {
  int m = -1;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (x[i]) m = i * 32 + 31 - (clz64((uint64_t)x[i]) - 32);
  return m;
}

template <int N>
QB_HD bool rzero(const uint32_t (&x)[N]) {
  uint32_t o = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) o |= x[i];
  return o == 0;
}

// Round (-1)^s * M * 2^lsb to binary128 (nearest even, subnormals, overflow).  M is exact, or
// rounded to odd at bit 0 with bit 0 at least two positions below the rounding position.
template <int N>
QB_HD q128 rround(const uint32_t (&M)[N], int lsb, uint32_t s) {
  const int msb = rmsb<N>(M);
  if (msb < 0) return q_zero(s);
  const int e = lsb + msb;
  int keep = msb - 112;
  if (e < EMIN) keep += EMIN - e;
  // bring bit keep-2 down to bit 0: bits 0,1 are round bits, 2..114 the significand
  uint32_t T[N];
  uint32_t sticky = 0;
  if (keep - 2 >= 0) {
    sticky = rshr<N>(T, M, keep - 2);
  } else {
    rshl<N>(T, M, 2 - keep);
  }
  uint64_t lo = ((uint64_t)T[1] << 32) | T[0];
  uint64_t hi = (N > 3) ? (((uint64_t)T[3] << 32) | T[2]) : ((N > 2) ? T[2] : 0);
  // significand = (hi:lo) >> 2 (113 bits), guard = bit 1, sticky = bit 0 | shifted-out bits
  const uint32_t guard = (uint32_t)(lo >> 1) & 1u;
  const uint32_t st = ((uint32_t)lo & 1u) | sticky;
  uint64_t sl = (lo >> 2) | (hi << 62);
  uint64_t sh = (hi >> 2) & ((1ull << 49) - 1);
  const uint64_t inc = guard & (st | (uint32_t)(sl & 1));
  sl += inc;
  sh += (sl == 0) & inc;
  int ulp = lsb + keep;
  if (sh >> 49) {
    sh >>= 1;
    ulp += 1;
  }
  if (sh >> 48) {
    const int be = ulp + 112 + QBIAS;
    if (be >= 0x7fff) return q_inf(s);
    return q_make(sl, ((uint64_t)s << 63) | ((uint64_t)be << 48) | (sh & FRAC_HI_MASK));
  }
  return q_make(sl, ((uint64_t)s << 63) | sh);
}

// r[0..N+3] = x[0..N-1] * a.m (113-bit significand)
template <int N>
QB_HD void rmul_q(uint32_t (&r)[N + 4], const uint32_t (&x)[N], const uq &a) {
  const uint32_t y[4] = {(uint32_t)a.ml, (uint32_t)(a.ml >> 32), (uint32_t)a.mh, (uint32_t)(a.mh >> 32)};
#pragma unroll
  for (int i = 0; i < N + 4; ++i) r[i] = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const uint64_t t = (uint64_t)x[i] * y[j] + ((uint64_t)r[i + j] + carry);
      r[i + j] = (uint32_t)t;
      carry = t >> 32;
    }
    r[N + j] = (uint32_t)carry;
  }
}

constexpr int SW = 24;         // words of |S| handled by the register epilogue
constexpr int FW = SW + 4 + 12;  // combine frame for alpha*S + beta*y

// RN((-1)^s1 A 2^e1 + (-1)^s2 B 2^e2) for nonzero magnitudes A (NA words) and B (NB words), in an
// FWN-word frame: the operand with the higher leading bit is placed with that bit at FWN*32-2, the
// other one aligned to it, bits below the frame folded into a sticky bit (with one unit less for a
// subtraction), so the sum is rounded once as the exact sum would be.
template <int FWN, int NA, int NB>
QB_HD q128 sum_round2(const uint32_t (&A)[NA], int e1, uint32_t s1, const uint32_t (&B)[NB], int e2, uint32_t s2) {
  uint32_t F1[FWN], F2[FWN];
#pragma unroll
  for (int i = 0; i < FWN; ++i) {
    F1[i] = (i < NA) ? A[i] : 0u;
    F2[i] = (i < NB) ? B[i] : 0u;
  }
  const int m1 = rmsb<FWN>(F1), m2 = rmsb<FWN>(F2);
  const bool one_big = (e1 + m1) >= (e2 + m2);
  const int topL = one_big ? (e1 + m1) : (e2 + m2);
  const int flsb = topL - (FWN * 32 - 2);
  uint32_t L[FWN], S[FWN];
  rshl<FWN>(L, one_big ? F1 : F2, (FWN * 32 - 2) - (one_big ? m1 : m2));
  const int es = one_big ? e2 : e1;
  uint32_t lost = 0;
  if (es >= flsb) {
    rshl<FWN>(S, one_big ? F2 : F1, es - flsb);
  } else {
    lost = rshr<FWN>(S, one_big ? F2 : F1, flsb - es);
  }
  const uint32_t sL = one_big ? s1 : s2, sSm = one_big ? s2 : s1;
  uint32_t R[FWN];
  uint32_t sR = sL;
  if (sL == sSm) {
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < FWN; ++i) R[i] = addc32(L[i], S[i], c, c);
  } else {
    // R = L - S - lost (two's complement); a borrow out means |S| > |L| (only without loss)
    uint32_t c = lost ? 0u : 1u;
#pragma unroll
    for (int i = 0; i < FWN; ++i) R[i] = addc32(L[i], ~S[i], c, c);
    if (!c) {
      uint32_t c2 = 1;
#pragma unroll
      for (int i = 0; i < FWN; ++i) R[i] = addc32(~R[i], 0u, c2, c2);
      sR = sSm;
    }
    if (!lost && rzero<FWN>(R)) return q_zero(0);
  }
  R[0] |= lost;
  return rround<FWN>(R, flsb, sR);
}

// Correctly rounded alpha*S + beta*y, with S = (-1)^sS * M * 2^lsb exact (sticky must be 0 when
// alpha/beta are used; for the plain sum, sticky marks a nonzero tail below M, i.e. M is S
// rounded to odd).  flags carry Inf/NaN products.  N is the word count of M; the combine frame
// has N + 16 words (room for alpha*S, N + 4 words, plus 12 words below it).
template <int N>
QB_HD q128 finish_reg_t(const uint32_t (&M)[N], int lsb, uint32_t sS, uint32_t sticky, uint32_t flags, bool has_ab,
                        q128 alpha, q128 beta, q128 yv) {
  constexpr int FWN = N + 16;
  const bool sflag = (flags & (FL_NAN | FL_PINF | FL_NINF)) != 0;
  if (!has_ab) {
    if (sflag) return special_result(flags);
    uint32_t Mx[N];
#pragma unroll
    for (int i = 0; i < N; ++i) Mx[i] = M[i];
    Mx[0] |= sticky;
    return rround<N>(Mx, lsb, sS);
  }
  const bool abspecial = (q_bexp(alpha) == 0x7fffu) || (q_bexp(beta) == 0x7fffu) || (q_bexp(yv) == 0x7fffu);
  if (sflag || abspecial) {
    const q128 sq = sflag ? special_result(flags) : rround<N>(M, lsb, sS);
    return fma_packed(alpha, sq, mul_packed(beta, yv));
  }
  const uq ua = unpack_finite(alpha), ub = unpack_finite(beta), uy = unpack_finite(yv);
  const bool z2 = ((ub.mh | ub.ml) == 0) || ((uy.mh | uy.ml) == 0);
  // alpha = +-1 and beta * y = 0: the result is S rounded once (with alpha's sign)
  if (z2 && ua.ml == 0 && ua.mh == HIDDEN && ua.e == 0 && !rzero<N>(M)) return rround<N>(M, lsb, sS ^ ua.s);
  // T1 = alpha * S  (N + 4 words), T2 = beta * y (8 words)
  uint32_t T1[N + 4];
  const bool z1 = rzero<N>(M) || ((ua.mh | ua.ml) == 0);
  if (!z1) {
    rmul_q<N>(T1, M, ua);
  } else {
#pragma unroll
    for (int i = 0; i < N + 4; ++i) T1[i] = 0;
  }
  const int e1 = lsb + ua.e - 112;
  const uint32_t s1 = sS ^ ua.s;
  uint32_t P2[8];
  if (!z2) mul113_w(P2, ub, uy);
  const int e2 = ub.e + uy.e - 224;
  const uint32_t s2 = ub.s ^ uy.s;
  if (z1 && z2) return q_zero(((uint32_t)(alpha.hi >> 63) ^ sS) & s2);
  if (z2) {
    uint32_t F1[FWN];
#pragma unroll
    for (int i = 0; i < FWN; ++i) F1[i] = (i < N + 4) ? T1[i] : 0u;
    return rround<FWN>(F1, e1, s1);
  }
  if (z1) {
    uint32_t F2[FWN];
#pragma unroll
    for (int i = 0; i < FWN; ++i) F2[i] = (i < 8) ? P2[i] : 0u;
    return rround<FWN>(F2, e2, s2);
  }
  return sum_round2<FWN, N + 4, 8>(T1, e1, s1, P2, e2, s2);
}

// RN(alpha M 2^lsb (-1)^sS + beta y) for finite alpha, beta, y and M != 0 (N words), with M cut to
// its top 128 bits before the product with alpha; the rest is exact (one rounding from a sticky
// bit in a 12-word frame).  terr receives an exponent bounding what the cut changes in the sum
// (|alpha| times the cut-off part of M 2^lsb), or INT_MIN when nothing was cut.
template <int N>
QB_HD q128 finish_reg_top(const uint32_t (&M)[N], int lsb, uint32_t sS, q128 alpha, q128 beta, q128 yv, int &terr) {
  terr = INT_MIN;
  const uq ua = unpack_finite(alpha), ub = unpack_finite(beta), uy = unpack_finite(yv);
  const bool z1 = (ua.mh | ua.ml) == 0;
  const bool z2 = ((ub.mh | ub.ml) == 0) || ((uy.mh | uy.ml) == 0);
  // alpha = +-1 and beta * y = 0: the result is M rounded once (with alpha's sign)
  if (z2 && ua.ml == 0 && ua.mh == HIDDEN && ua.e == 0) return rround<N>(M, lsb, sS ^ ua.s);
  uint32_t P2[8];
  if (!z2) mul113_w(P2, ub, uy);
  const int e2 = ub.e + uy.e - 224;
  const uint32_t s2 = ub.s ^ uy.s;
  if (z1) {
    if (z2) return q_zero(((uint32_t)(alpha.hi >> 63) ^ sS) & s2);
    uint32_t F2[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) F2[i] = (i < 8) ? P2[i] : 0u;
    return rround<12>(F2, e2, s2);
  }
  // H = the top 128 bits of M (leading bit 127)
  const int sh = rmsb<N>(M) - 127;
  uint32_t W[N];
  uint32_t cut = 0;
  if (sh > 0)
    cut = rshr<N>(W, M, sh);
  else
    rshl<N>(W, M, -sh);
  uint32_t H[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) H[i] = (i < N) ? W[i] : 0u;
  if (cut) terr = lsb + sh + ua.e + 1;  // |alpha| < 2^(ua.e + 1), the cut part < 2^(lsb + sh)
  uint32_t T1[8];
  rmul_q<4>(T1, H, ua);
  const int e1 = lsb + sh + ua.e - 112;
  const uint32_t s1 = sS ^ ua.s;
  if (z2) {
    uint32_t F1[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) F1[i] = (i < 8) ? T1[i] : 0u;
    return rround<12>(F1, e1, s1);
  }
  return sum_round2<12, 8, 8>(T1, e1, s1, P2, e2, s2);
}

QB_HD q128 finish_reg(const uint32_t (&M)[SW], int lsb, uint32_t sS, uint32_t sticky, uint32_t flags, bool has_ab,
                      q128 alpha, q128 beta, q128 yv) {
  return finish_reg_t<SW>(M, lsb, sS, sticky, flags, has_ab, alpha, beta, yv);
}

// S of one output from its lo/hi window sums (same anchor A: hi is A + WIN_BAND): the magnitude
// in SW words with lsb exponent A, and its sign.
QB_HD void s_from_windows(uint32_t (&M)[SW], uint32_t &sign, const uint32_t (&lo)[WW], const uint32_t (&hi)[WW]) {
  // T = sext(lo) + (sext(hi) << WIN_BAND) over SW words, two's complement
  uint32_t L[SW], H[SW], Hs[SW];
  const uint32_t el = (lo[WW - 1] >> 31) ? 0xffffffffu : 0u, eh = (hi[WW - 1] >> 31) ? 0xffffffffu : 0u;
#pragma unroll
  for (int i = 0; i < SW; ++i) {
    L[i] = (i < WW) ? lo[i] : el;
    H[i] = (i < WW) ? hi[i] : eh;
  }
  rshl<SW>(Hs, H, WIN_BAND);  // sign bits above stay ones: H was sign-extended to SW words
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < SW; ++i) M[i] = addc32(L[i], Hs[i], c, c);
  sign = M[SW - 1] >> 31;
  if (sign) {
    uint32_t c2 = 1;
#pragma unroll
    for (int i = 0; i < SW; ++i) M[i] = addc32(~M[i], 0u, c2, c2);
  }
}

// Compact a wide accumulator (two's complement, ACCW words at acc.base) to its top SW words of
// magnitude: lsb exponent, sign, and a sticky bit for nonzero words below (round-to-odd).
QB_HD void compact_wide(const wide_acc &acc, uint32_t (&M)[SW], int &lsb, uint32_t &sign, uint32_t &sticky) {
  uint32_t A[ACCW];
  sign = acc.w[ACCW - 1] >> 31;
  uint32_t c = sign;
  for (int i = 0; i < ACCW; ++i) A[i] = sign ? addc32(~acc.w[i], 0u, c, c) : acc.w[i];
  int top = -1;
  for (int i = 0; i < ACCW; ++i)
    if (A[i]) top = i;
  const int lo_w = (top >= SW) ? top - SW + 1 : 0;
  sticky = 0;
  for (int i = 0; i < lo_w; ++i) sticky |= (A[i] != 0);
  for (int i = 0; i < SW; ++i) M[i] = A[lo_w + i < ACCW ? lo_w + i : ACCW - 1] & (lo_w + i < ACCW ? 0xffffffffu : 0u);
  lsb = acc.base + 32 * lo_w;
}
}  // namespace qb
