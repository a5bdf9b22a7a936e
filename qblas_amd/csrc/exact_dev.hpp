// exact_dev.hpp -- device helpers shared by the kernels that accumulate binary128 products
// exactly (kernels_exact.hip: dot and gemv; kernels_ozaki.hip: the exact repair of GEMM entries):
// 128-bit loads and stores, wave reductions of exponents, flags and windows, the per-lane
// two-window accumulate step and the partial-record format.
#pragma once
#include "exact.hpp"

namespace qbk {
using namespace qb;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ q128 ldq_nt(const q128 *p) {
  const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
  return q_make(((uint64_t)v.y << 32) | v.x, ((uint64_t)v.w << 32) | v.z);
}
__device__ __forceinline__ q128 ldq_c(const q128 *p) {
  const uint4 v = *reinterpret_cast<const uint4 *>(p);
  return q_make(((uint64_t)v.y << 32) | v.x, ((uint64_t)v.w << 32) | v.z);
}
__device__ __forceinline__ void stq_c(q128 *p, q128 v) {
  uint4 w;
  w.x = (uint32_t)v.lo;
  w.y = (uint32_t)(v.lo >> 32);
  w.z = (uint32_t)v.hi;
  w.w = (uint32_t)(v.hi >> 32);
  *reinterpret_cast<uint4 *>(p) = w;
}

struct part_t {
  int32_t A;
  uint32_t flags;
  uint32_t W[WW];
  uint32_t pad[2];
};
static_assert(sizeof(part_t) == 64, "partial record is one 64-byte line");

constexpr int EXP_NONE = 0x3fffffff;

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = min(v, __shfl_xor(v, m));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m));
  return v;
}
__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v |= (uint32_t)__shfl_xor((int)v, m);
  return v;
}
__device__ __forceinline__ void wave_sum_win(uint32_t (&W)[WW]) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    uint32_t X[WW];
#pragma unroll
    for (int i = 0; i < WW; ++i) X[i] = (uint32_t)__shfl_xor((int)W[i], m);
    win_add_win(W, X);
  }
}

// product lsb exponent of a pair, or EXP_NONE if the pair is not a regular nonzero finite product
__device__ __forceinline__ int pair_exp(q128 a, q128 b) {
  uq ua, ub;
  uint32_t fl = 0;
  const int c = classify_pair(a, b, ua, ub, fl);
  if (c == 0 || c == 3) return prod_lsb(ua, ub);
  return EXP_NONE;
}

// wave-uniform anchor centred between the smallest and largest first products of the wave
__device__ __forceinline__ int wave_anchor(int e0) {
  const int mn = wave_min(e0);
  const int mx = wave_max(e0 == EXP_NONE ? -EXP_NONE : e0);
  if (mn == EXP_NONE) return 0;
  return ((mn + mx) >> 1) - WIN2_CENTER;
}

__device__ __forceinline__ void accumulate(uint32_t (&Wl)[WW], uint32_t (&Wh)[WW], int A, q128 a, q128 b,
                                           uint32_t &flags) {
  uq ua, ub;
  const int c = classify_pair(a, b, ua, ub, flags);
  if (c == 0 || c == 3) {
    if (!win2_add_product(Wl, Wh, A, ua, ub)) flags |= FL_LOST;
  }
}
__device__ __forceinline__ void put_parts(part_t *p, int A, uint32_t flags, const uint32_t (&Wl)[WW],
                                          const uint32_t (&Wh)[WW]) {
  p[0].A = A;
  p[0].flags = flags;
  p[1].A = A + WIN_BAND;
  p[1].flags = 0;
#pragma unroll
  for (int k = 0; k < WW; ++k) {
    p[0].W[k] = Wl[k];
    p[1].W[k] = Wh[k];
  }
}

// ------------------------------------------------------------------------------------------
// Lazy digits in LDS: the streaming kernels' per-lane accumulator.  Lane l of a wave owns DIG
// 64-bit digits at d[k * 64] (d = the wave's digit array + l), digit k weighing 2^(A + 29 k).  A
// product with lsb offset o = e_p - A in [1, WIN_BAND + WIN_OMAX] (the band of the two register
// windows, so the same products go to the full-range path) is formed at its offset within digit
// w = o / 29: one significand shifted by r = o - 29 w is cut into five 29-bit limbs, the other
// into four, and the column sums of the 20 limb products (each < 2^58, a column < 2^61 with its
// carry in, one 64-bit multiply-add each) are cut into nine 29-bit digits that go to LDS digits
// w .. w + 8 by 64-bit atomic adds -- no carries between LDS digits, no shifting network, no
// window-wide carry chain.  A negative product adds the complements of its digits within 29
// bits, +1 at digit w and -1 (2^64 - 1) at digit w + 9:
//   sum_{k<9} (2^29 - 1 - t_k) 2^(29 k) + 1 - 2^261 = -T.
// A digit takes < 2^29 + 1 per product, so 2^31 products per lane stay far within int64.
// ------------------------------------------------------------------------------------------
constexpr int DIG = 18;     // digits 0..17: products reach digit 8 + 8, the sign term digit 8 + 9
constexpr int DB = 29;      // bits per digit
constexpr uint32_t DMASK = (1u << DB) - 1u;

__device__ __forceinline__ void dig_add(unsigned long long *p, unsigned long long v) {
  (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void dig_clear(unsigned long long *d) {
#pragma unroll
  for (int k = 0; k < DIG; ++k) d[k * 64] = 0;
}

// the product (-1)^s a.m b.m 2^(A + o) into the digits: b given as its four 29-bit limbs
__device__ __forceinline__ void dig_core(unsigned long long *d, int o, uint64_t aml, uint64_t amh, uint32_t B0,
                                         uint32_t B1, uint32_t B2, uint32_t B3, uint32_t s) {
  const uint32_t w = ((uint32_t)o * 565u) >> 14;  // o / 29 for 0 <= o < 300 (checked exhaustively)
  const uint32_t r = (uint32_t)o - w * (uint32_t)DB;  // 0..28
  // a.m << r (at most 141 bits) as five 29-bit limbs
  const uint64_t lo = aml << r;
  const uint64_t mid = (amh << r) | ((aml >> 1) >> (63u - r));
  const uint64_t hi = (amh >> 1) >> (63u - r);
  const uint32_t A0 = (uint32_t)lo & DMASK, A1 = (uint32_t)(lo >> 29) & DMASK,
                 A2 = (uint32_t)((lo >> 58) | (mid << 6)) & DMASK, A3 = (uint32_t)(mid >> 23) & DMASK,
                 A4 = (uint32_t)((mid >> 52) | (hi << 12)) & DMASK;
  uint32_t t[9];
  uint64_t acc = (uint64_t)A0 * B0;
  t[0] = (uint32_t)acc & DMASK;
  acc = (acc >> 29) + (uint64_t)A1 * B0;
  acc += (uint64_t)A0 * B1;
  t[1] = (uint32_t)acc & DMASK;
  acc = (acc >> 29) + (uint64_t)A2 * B0;
  acc += (uint64_t)A1 * B1;
  acc += (uint64_t)A0 * B2;
  t[2] = (uint32_t)acc & DMASK;
  acc = (acc >> 29) + (uint64_t)A3 * B0;
  acc += (uint64_t)A2 * B1;
  acc += (uint64_t)A1 * B2;
  acc += (uint64_t)A0 * B3;
  t[3] = (uint32_t)acc & DMASK;
  acc = (acc >> 29) + (uint64_t)A4 * B0;
  acc += (uint64_t)A3 * B1;
  acc += (uint64_t)A2 * B2;
  acc += (uint64_t)A1 * B3;
  t[4] = (uint32_t)acc & DMASK;
  acc = (acc >> 29) + (uint64_t)A4 * B1;
  acc += (uint64_t)A3 * B2;
  acc += (uint64_t)A2 * B3;
  t[5] = (uint32_t)acc & DMASK;
  acc = (acc >> 29) + (uint64_t)A4 * B2;
  acc += (uint64_t)A3 * B3;
  t[6] = (uint32_t)acc & DMASK;
  acc = (acc >> 29) + (uint64_t)A4 * B3;
  t[7] = (uint32_t)acc & DMASK;
  t[8] = (uint32_t)(acc >> 29);  // < 2^22: the product is below 2^254
  const uint32_t m = 0u - (s & 1u), m29 = m & DMASK;
  unsigned long long *q = d + w * 64;
  dig_add(q, (unsigned long long)((t[0] ^ m29) + (m & 1u)));
#pragma unroll
  for (int k = 1; k < 9; ++k) dig_add(q + k * 64, (unsigned long long)(t[k] ^ m29));
  if (m) dig_add(q + 9 * 64, ~0ull);
}

// b.m (113 bits) as four 29-bit limbs
__device__ __forceinline__ void limbs4(const uq &b, uint32_t &B0, uint32_t &B1, uint32_t &B2, uint32_t &B3) {
  B0 = (uint32_t)b.ml & DMASK;
  B1 = (uint32_t)(b.ml >> 29) & DMASK;
  B2 = (uint32_t)((b.ml >> 58) | (b.mh << 6)) & DMASK;
  B3 = (uint32_t)(b.mh >> 23);
}

__device__ __forceinline__ void dig_product(unsigned long long *d, int A, const uq &a, const uq &b, uint32_t &flags) {
  const int o = prod_lsb(a, b) - A;
  if (o < WIN_OMIN || o > WIN_BAND + WIN_OMAX) {
    flags |= FL_LOST;
    return;
  }
  uint32_t B0, B1, B2, B3;
  limbs4(b, B0, B1, B2, B3);
  dig_core(d, o, a.ml, a.mh, B0, B1, B2, B3, a.s ^ b.s);
}

// A vector operand prepared once per element for the gemv kernels, in two arrays: its four
// 29-bit limbs (uint4, value = limbs * 2^(e - 112)) and a tag e << 3 | class << 1 | sign (class 0:
// nonzero finite, 1: zero, 2: Inf or NaN).  Two arrays keep a wave's loads of consecutive
// elements dense (16 and 4 bytes a lane).
struct xsplit {
  uint32_t B[4];
  int32_t tag;
};

__device__ __forceinline__ xsplit make_xsplit(q128 x) {
  xsplit r;
  const uint32_t be = q_bexp(x);
  const uint32_t cls = (be == 0x7fffu) ? 2u : (q_iszero(x) ? 1u : 0u);
  const uq u = cls == 0 ? unpack_finite(x) : unpack_finite(q_zero(0));
  limbs4(u, r.B[0], r.B[1], r.B[2], r.B[3]);
  r.tag = (int32_t)((uint32_t)u.e << 3) | (int32_t)(cls << 1) | (int32_t)(x.hi >> 63);
  return r;
}
__device__ __forceinline__ xsplit ldx(const uint4 *__restrict__ xl, const int32_t *__restrict__ xt, int64_t j) {
  const uint4 v = xl[j];
  xsplit r;
  r.B[0] = v.x;
  r.B[1] = v.y;
  r.B[2] = v.z;
  r.B[3] = v.w;
  r.tag = xt[j];
  return r;
}

// a (streamed) times a prepared vector element: normal a with a nonzero finite x goes straight to
// the digits; anything else takes the general classification on the raw values (xraw)
__device__ __forceinline__ void dig_accumulate_x(unsigned long long *d, int A, q128 a, const xsplit &x,
                                                 const q128 *xraw, uint32_t &flags) {
  const uint32_t be = q_bexp(a);
  if ((be - 1u) < 0x7ffeu && (x.tag & 6) == 0) {
    const int o = ((int)be - QBIAS) + (x.tag >> 3) - 224 - A;
    if (o < WIN_OMIN || o > WIN_BAND + WIN_OMAX) {
      flags |= FL_LOST;
      return;
    }
    dig_core(d, o, a.lo, (a.hi & FRAC_HI_MASK) | HIDDEN, x.B[0], x.B[1], x.B[2], x.B[3],
             (uint32_t)(a.hi >> 63) ^ (uint32_t)x.tag);
    return;
  }
  uq ua, ub;
  const int c = classify_pair(a, ldq_c(xraw), ua, ub, flags);
  if (c == 0 || c == 3) dig_product(d, A, ua, ub, flags);
}

// regular pairs into the digits, specials into flags (as accumulate())
__device__ __forceinline__ void dig_accumulate(unsigned long long *d, int A, q128 a, q128 b, uint32_t &flags) {
  uq ua, ub;
  const int c = classify_pair(a, b, ua, ub, flags);
  if (c == 0 || c == 3) dig_product(d, A, ua, ub, flags);
}

// This lane's digits as the two windows of anchor A: V = sum_k D_k 2^(29 k) with the carries
// rippled once (two's complement over 32-bit words), split as V = Wl + Wh 2^WIN_BAND with
// Wl = V mod 2^WIN_BAND.
__device__ __forceinline__ void dig_to_windows(const unsigned long long *d, uint32_t (&Wl)[WW], uint32_t (&Wh)[WW]) {
  constexpr int VW = (DIG * DB + 31) / 32 + 1;  // 18 words
  uint32_t nd[DIG];
  int64_t carry = 0;
#pragma unroll
  for (int k = 0; k < DIG; ++k) {
    const int64_t s = (int64_t)d[k * 64] + carry;
    nd[k] = (uint32_t)s & DMASK;
    carry = s >> 29;  // arithmetic
  }
  const uint32_t ext = carry < 0 ? 0xffffffffu : 0u;  // carry is 0 or -1 here
  uint32_t V[VW];
#pragma unroll
  for (int wd = 0; wd < VW; ++wd) {
    // bits 32 wd .. 32 wd + 31 of sum nd_k 2^(29 k) (digits do not overlap) and the sign above
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < DIG; ++k) {
      const int b0 = DB * k - 32 * wd;  // position of digit k relative to word wd
      if (b0 > -DB && b0 < 32) v |= (b0 >= 0) ? (nd[k] << b0) : (nd[k] >> (-b0));
    }
    const int top = DB * DIG - 32 * wd;  // first bit of the sign extension within word wd
    if (top <= 0) v = ext;
    else if (top < 32) v |= ext << top;
    V[wd] = v;
  }
  constexpr int WB = WIN_BAND / 32, BB = WIN_BAND % 32;  // 3, 28
#pragma unroll
  for (int k = 0; k < WW; ++k) {
    const uint32_t lo = (WB + k < VW) ? V[WB + k] : ext;
    const uint32_t hi = (WB + k + 1 < VW) ? V[WB + k + 1] : ext;
    Wh[k] = alignbit32(hi, lo, BB);
    Wl[k] = (k < WB) ? V[k] : ((k == WB) ? (V[WB] & ((1u << BB) - 1u)) : 0u);
  }
}

}  // namespace qbk
