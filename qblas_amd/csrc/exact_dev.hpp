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

}  // namespace qbk
