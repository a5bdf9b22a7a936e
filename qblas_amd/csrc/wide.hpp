// wide.hpp -- full-range exact accumulation of binary128 products: the slow path of the EXACT form
// (dot, gemv, batched gemv) and of the OZAKI form's exact repair.
//
// The fast paths keep two 384-bit windows per lane (exact.hpp), which hold products whose least
// significant bit lies within 250 binary orders of magnitude of an anchor.  An item (a dot, a
// gemv row, a GEMM entry) with a product outside them is recomputed here by a whole workgroup:
// products that fit the lane windows still go there, every other product is added exactly into a
// fixed-point accumulator in LDS that spans the whole binary128 range (a Kulisch accumulator):
// KUL_N signed 64-bit limbs, limb i collecting 32-bit digits of weight 2^(KUL_LSB + 32 i) through
// LDS atomics.  With alpha and beta (gemv, gemm) the accumulator receives alpha times each product
// (339 bits) and beta*y, so it holds alpha * sum + beta * y exactly and is rounded once.
//
// Range: a finite binary128 value is m 2^(e-112) with e in [-16494, 16383] (subnormals
// normalised), so alpha*a*b has its lsb at or above -49818 and stays below 2^49152; the windows'
// anchors are at or above -33337 - 16606.  Limb 0 at 2^-50016 and 3104 limbs (24.8 KB) cover that
// with the carries of up to 2^40 terms.  A limb takes at most KUL_SEG deposits of < 2^32 between
// two normalisations, so it never leaves int64.
#pragma once
#include "exact_dev.hpp"

namespace qbk {
using namespace qb;

constexpr int KUL_LSB = -50016;
constexpr int KUL_N = 3104;
constexpr int64_t KUL_SEG = (int64_t)1 << 29;  // elements of one item between two normalisations

struct kul_t {
  unsigned long long L[KUL_N];
  int lo, hi;      // limbs outside [lo, hi] are zero (lo > hi: all zero)
  uint32_t flags;  // FL_NAN / FL_PINF / FL_NINF of the item's products; FL_LOST if a term fell
                   // outside the accumulator (cannot happen for binary128 inputs)
  int neg;         // sign of the value after kul_normalize
  q128 out;        // the item's result, for every thread
};

__device__ __forceinline__ unsigned long long kul_atomic_add(unsigned long long *p, unsigned long long v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// every thread; a barrier must follow before first use
__device__ __forceinline__ void kul_clear_all(kul_t &K) {
  for (int i = threadIdx.x; i < KUL_N; i += blockDim.x) K.L[i] = 0;
  if (threadIdx.x == 0) {
    K.lo = KUL_N;
    K.hi = -1;
    K.flags = 0;
    K.neg = 0;
  }
}

// K += (-1)^neg * X * 2^lsb (X: NW words, least significant first).  tlo/thi track the limbs this
// thread touched.
template <int NW>
__device__ __forceinline__ void kul_add(kul_t &K, const uint32_t (&X)[NW], int lsb, uint32_t neg, int &tlo, int &thi,
                                        uint32_t &flags) {
  const int o = lsb - KUL_LSB;
  if (o < 1) {
    bool any = false;
#pragma unroll
    for (int j = 0; j < NW; ++j) any |= X[j] != 0;
    if (any) flags |= FL_LOST;
    return;
  }
  const int w = (o - 1) >> 5;
  const uint32_t sh = 32u - (uint32_t)(o - (w << 5));  // bit offset 1..32 within word w -> funnel 31..0
  uint32_t T[NW + 1];
  T[0] = alignbit32(X[0], 0u, sh);
#pragma unroll
  for (int j = 1; j < NW; ++j) T[j] = alignbit32(X[j], X[j - 1], sh);
  T[NW] = alignbit32(0u, X[NW - 1], sh);
#pragma unroll
  for (int j = 0; j <= NW; ++j) {
    if (T[j]) {
      const int i = w + j;
      if (i >= KUL_N) {
        flags |= FL_LOST;
        continue;
      }
      kul_atomic_add(&K.L[i], neg ? (0ull - (unsigned long long)T[j]) : (unsigned long long)T[j]);
      tlo = min(tlo, i);
      thi = max(thi, i);
    }
  }
}

// one thread: limbs [lo, hi] -> 32-bit digits, canonical form value = sum_{lo <= i <= t} D_i 2^(32 i)
// - neg 2^(32 (t + 1)) with no sign-extension digits left at the top; the limbs keep that form
// (D_i in [0, 2^32), limb t + 1 = -1 when neg) so that deposits can go on.  Returns t (t < lo:
// the value is 0, or -2^(32 lo) when neg).
__device__ int kul_normalize(kul_t &K) {
  const int lo = K.lo;
  if (lo > K.hi) {
    K.neg = 0;
    return lo - 1;
  }
  const int top = min(K.hi + 2, KUL_N - 1);
  int64_t carry = 0;
  for (int i = lo; i <= top; ++i) {
    const int64_t v = (int64_t)K.L[i] + carry;
    K.L[i] = (uint32_t)v;
    carry = v >> 32;  // arithmetic: the limbs are signed
  }
  const uint32_t ext = carry ? 0xffffffffu : 0u;
  int t = top;
  while (t >= lo && (uint32_t)K.L[t] == ext) {
    K.L[t] = 0;
    --t;
  }
  K.neg = carry != 0;
  if (carry) {
    if (t + 1 < KUL_N) {
      K.L[t + 1] = ~0ull;
      K.hi = t + 1;
    } else {
      K.flags |= FL_LOST;
      K.hi = KUL_N - 1;
    }
  } else {
    K.hi = t;
    if (t < lo) K.lo = KUL_N;  // zero
  }
  return t;
}

// one thread, after kul_normalize returned t: the value rounded to binary128 (nearest even);
// *zero set when it is exactly zero (the result is then +0)
__device__ q128 kul_round(kul_t &K, int t, bool &zero) {
  const int lo = K.lo;
  zero = false;
  if (!K.neg) {
    if (t < lo) {
      zero = true;
      return q_zero(0);
    }
  } else {
    // magnitude = 2^(32(t+1)) - D: two's complement of the digits, a carry out means D = 0
    uint32_t c = 1;
    for (int i = lo; i <= t; ++i) {
      uint32_t co;
      const uint32_t v = addc32(~(uint32_t)K.L[i], 0u, c, co);
      K.L[i] = v;
      c = co;
    }
    if (c) {
      ++t;
      K.L[t] = 1;
    }
    while (t > lo && (uint32_t)K.L[t] == 0) --t;
  }
  const int base = t - (SW - 1);
  uint32_t M[SW];
#pragma unroll
  for (int k = 0; k < SW; ++k) M[k] = (base + k >= lo) ? (uint32_t)K.L[base + k] : 0u;
  uint32_t sticky = 0;
  for (int i = lo; i < base; ++i) sticky |= (uint32_t)K.L[i];
  M[0] |= sticky ? 1u : 0u;  // rounded to odd far below the rounding position
  return rround<SW>(M, KUL_LSB + 32 * base, (uint32_t)K.neg);
}

// two's complement window (12 words, weight 2^A of bit 0) into K, times alpha's significand when
// triple (lsb of alpha = ae - 112, sign as)
__device__ __forceinline__ void kul_add_window(kul_t &K, const uint32_t (&W)[WW], int A, bool triple, const uq &ua,
                                               int &tlo, int &thi, uint32_t &flags) {
  uint32_t o = 0;
#pragma unroll
  for (int k = 0; k < WW; ++k) o |= W[k];
  if (!o) return;
  const uint32_t s = W[WW - 1] >> 31;
  uint32_t X[WW];
  uint32_t c = s;
#pragma unroll
  for (int k = 0; k < WW; ++k) X[k] = s ? addc32(~W[k], 0u, c, c) : W[k];
  if (triple) {
    uint32_t Y[WW + 4];
    rmul_q<WW>(Y, X, ua);
    kul_add<WW + 4>(K, Y, A + ua.e - 112, s ^ ua.s, tlo, thi, flags);
  } else {
    kul_add<WW>(K, X, A, s, tlo, thi, flags);
  }
}

// (-1)^s a.m b.m (times alpha's significand when triple) at its exponent into K
__device__ __forceinline__ void kul_add_product(kul_t &K, const uq &a, const uq &b, bool triple, const uq &ua,
                                                int &tlo, int &thi, uint32_t &flags) {
  uint32_t P[8];
  mul113_w(P, a, b);
  if (triple) {
    uint32_t Y[12];
    rmul_q<8>(Y, P, ua);
    kul_add<12>(K, Y, prod_lsb(a, b) + ua.e - 112, a.s ^ b.s ^ ua.s, tlo, thi, flags);
  } else {
    kul_add<8>(K, P, prod_lsb(a, b), a.s ^ b.s, tlo, thi, flags);
  }
}

// The per-thread part of an item's accumulation over elements [l0, l1) with stride blockDim.x:
// products into the lane windows (anchor An) where they fit, into K otherwise.
__device__ __forceinline__ void kul_accumulate(kul_t &K, int64_t l0, int64_t l1, const q128 *__restrict__ a,
                                               int64_t inca, const q128 *__restrict__ b, int64_t incb, int An,
                                               uint32_t (&Wl)[WW], uint32_t (&Wh)[WW], bool triple, const uq &ua,
                                               int &tlo, int &thi, uint32_t &flags) {
  for (int64_t l = l0 + threadIdx.x; l < l1; l += blockDim.x) {
    uq x, y;
    const int c = classify_pair(ldq_c(a + l * inca), ldq_c(b + l * incb), x, y, flags);
    if (c == 0 || c == 3) {
      if (!win2_add_product(Wl, Wh, An, x, y)) kul_add_product(K, x, y, triple, ua, tlo, thi, flags);
    }
  }
}

// the block's touched limb range into K.lo / K.hi (before a barrier)
__device__ __forceinline__ void kul_publish_range(kul_t &K, int tlo, int thi) {
  if (tlo <= thi) {
    atomicMin(&K.lo, tlo);
    atomicMax(&K.hi, thi);
  }
}

// Exact RN(alpha * sum_{l<n} a_l b_l + beta * yv) (ab) or RN(sum) (!ab), a_l = a[l inca],
// b_l = b[l incb], by the whole workgroup (every thread calls it; blockDim.x a multiple of 64).
// Returns the value to every thread.  K must be clear (kul_clear_all, then a barrier) and is
// clear again on return.  IEEE rules as the fast path's finish (exact.hpp finish_reg_t): Inf/NaN
// products, alpha/beta/y special values, the sign of exact zeros.  status (nullable): bit 0 is set
// if a term fell outside the accumulator (the result is then NaN; binary128 inputs never do that).
__device__ q128 wide_item(kul_t &K, int64_t n, const q128 *__restrict__ a, int64_t inca, const q128 *__restrict__ b,
                          int64_t incb, bool ab, q128 alpha, q128 beta, q128 yv, uint32_t *status) {
  const bool abspecial = ab && ((q_bexp(alpha) == 0x7fffu) || (q_bexp(beta) == 0x7fffu) || (q_bexp(yv) == 0x7fffu));
  const bool azero = ab && !abspecial && q_iszero(alpha);
  const bool triple = ab && !abspecial && !azero;
  uq ua;
  ua.ml = ua.mh = 0;
  ua.e = 0;
  ua.s = 0;
  if (triple) ua = unpack_finite(alpha);
  // wave anchor from the wave's first products (as the fast path)
  int e0 = EXP_NONE;
  if (threadIdx.x < n) e0 = pair_exp(ldq_c(a + (int64_t)threadIdx.x * inca), ldq_c(b + (int64_t)threadIdx.x * incb));
  const int An = wave_anchor(e0);
  uint32_t Wl[WW], Wh[WW];
#pragma unroll
  for (int k = 0; k < WW; ++k) {
    Wl[k] = 0;
    Wh[k] = 0;
  }
  uint32_t flags = 0;
  int tlo = KUL_N, thi = -1;
  for (int64_t s0 = 0; s0 < n; s0 += KUL_SEG) {
    const int64_t s1 = (n - s0 > KUL_SEG) ? s0 + KUL_SEG : n;
    kul_accumulate(K, s0, s1, a, inca, b, incb, An, Wl, Wh, triple, ua, tlo, thi, flags);
    if (s1 < n) {  // uniform: long items are normalised between segments
      kul_publish_range(K, tlo, thi);
      tlo = KUL_N;
      thi = -1;
      __syncthreads();
      if (threadIdx.x == 0) (void)kul_normalize(K);
      __syncthreads();
    }
  }
  wave_sum_win(Wl);
  wave_sum_win(Wh);
  flags = wave_or(flags);
  if ((threadIdx.x & 63) == 0) {
    kul_add_window(K, Wl, An, triple, ua, tlo, thi, flags);
    kul_add_window(K, Wh, An + WIN_BAND, triple, ua, tlo, thi, flags);
    if (flags) atomicOr(&K.flags, flags);
  }
  if (threadIdx.x == 0 && triple && !q_iszero(beta) && !q_iszero(yv)) {
    const uq ub = unpack_finite(beta), uy = unpack_finite(yv);
    uint32_t P[8], f2 = 0;
    mul113_w(P, ub, uy);
    kul_add<8>(K, P, prod_lsb(ub, uy), ub.s ^ uy.s, tlo, thi, f2);
    if (f2) atomicOr(&K.flags, f2);
  }
  kul_publish_range(K, tlo, thi);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = kul_normalize(K);
    const uint32_t fl = K.flags;
    const bool sflag = (fl & (FL_NAN | FL_PINF | FL_NINF)) != 0;
    q128 r;
    if (fl & FL_LOST) {
      r = q_nan();
      if (status) atomicOr(status, 1u);
    } else if (!ab) {
      bool z;
      r = sflag ? special_result(fl) : kul_round(K, t, z);
    } else if (abspecial) {
      bool z;
      const q128 sq = sflag ? special_result(fl) : kul_round(K, t, z);
      r = fma_packed(alpha, sq, mul_packed(beta, yv));
    } else if (sflag) {
      r = fma_packed(alpha, special_result(fl), mul_packed(beta, yv));
    } else if (azero) {
      // 0 * S + beta * y: beta * y rounded once when it is not an exact zero; else the IEEE sign of
      // (+-0) + (+-0)
      if (!q_iszero(beta) && !q_iszero(yv)) {
        r = mul_packed(beta, yv);
      } else {
        bool z;
        (void)kul_round(K, t, z);
        const uint32_t sS = z ? 0u : (uint32_t)K.neg;
        const uint32_t s2 = (uint32_t)((beta.hi ^ yv.hi) >> 63);
        r = q_zero((((uint32_t)(alpha.hi >> 63)) ^ sS) & s2);
      }
    } else {
      bool z;
      r = kul_round(K, t, z);
      if (z) {
        // exact zero: cancellation of nonzero terms gives +0; with beta*y = 0 the sum was an exact
        // zero too (alpha != 0), and the sign is that of (alpha * +0) + (beta * y)
        const bool t2 = !q_iszero(beta) && !q_iszero(yv);
        const uint32_t s2 = (uint32_t)((beta.hi ^ yv.hi) >> 63);
        r = t2 ? q_zero(0) : q_zero(((uint32_t)(alpha.hi >> 63)) & s2);
      }
    }
    K.out = r;
  }
  __syncthreads();
  const q128 r = K.out;
  const int lo = K.lo, hi = K.hi;
  __syncthreads();
  for (int i = lo + (int)threadIdx.x; i <= hi; i += blockDim.x) K.L[i] = 0;
  if (threadIdx.x == 0) {
    K.lo = KUL_N;
    K.hi = -1;
    K.flags = 0;
    K.neg = 0;
  }
  __syncthreads();
  return r;
}

}  // namespace qbk
