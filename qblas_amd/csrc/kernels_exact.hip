// kernels_exact.hip -- EXACT form of dot and gemv (correctly rounded results).
//
// Computes the same routines as the reference (include/quadblas/algorithms/level1.hpp:80-137,
// level2.hpp:15-99) but returns RN(alpha * sum_j a_j b_j + beta * y), the exact value rounded once
// (exact.hpp).  Kernels:
//   k_dot_exact_partial    grid-stride lanes, block-uniform window anchor, products into lazy LDS
//                          digits (exact_dev.hpp), wave and block reductions into one lo/hi window
//                          pair per block.
//   k_exact_dot_combine    sums the block partials on a common base in registers (column sums per
//                          workgroup); the last workgroup ripples the carries once and rounds.
//   k_vec_split            the gemv vector prepared once: 29-bit limbs and a tag per element.
//   k_gemv_row_exact       one wave per (row, column chunk), wave-uniform anchor, wave reduction.
//   k_gemv_col_exact       one lane per row, blockIdx.y = column chunk, per-lane anchor.
//   k_exact_finish_rows    one thread per row: sums the row's chunk windows (shared anchor),
//                          forms alpha*S + beta*y exactly in registers and rounds into y.
// Loads of the streamed operands are coalesced 128-bit loads.  A dot or row with a product outside
// its lane windows (products more than ~2^124 from the anchor) is recomputed on the full-range
// path (wide.hpp): the dot by the combine kernel's whole grid, rows by k_exact_wide_rows from a
// queue that the row finish fills.  Header word 2 counts the items so recomputed.
#include "wide.hpp"
#include "kernels.h"

using namespace qb;

namespace qbk {

// ------------------------------------------------------------------------------------------
// dot
// ------------------------------------------------------------------------------------------
constexpr int XD_BS = 256;

__global__ void __launch_bounds__(XD_BS) k_dot_exact_partial(const q128 *__restrict__ x, int64_t incx,
                                                             const q128 *__restrict__ y, int64_t incy, int64_t n,
                                                             part_t *__restrict__ parts, uint32_t *__restrict__ hdr) {
  __shared__ int s_mn[XD_BS / 64], s_mx[XD_BS / 64];
  __shared__ uint32_t s_fl[XD_BS / 64];
  __shared__ unsigned long long s_dig[XD_BS / 64][DIG][64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && threadIdx.x < 4) hdr[threadIdx.x] = 0;  // status, counts, ticket of the combine
  unsigned long long *d = &s_dig[wv][0][lane];
  dig_clear(d);
  const int64_t stride = (int64_t)gridDim.x * XD_BS;
  const int64_t i0 = (int64_t)blockIdx.x * XD_BS + threadIdx.x;
  // the first two element pairs are loaded before the anchor is known (they set it) and are the
  // first ones accumulated
  const q128 zq = q_zero(0);
  const bool h0 = i0 < n, h1 = i0 + stride < n;
  const q128 a0 = h0 ? ldq_nt(x + i0 * incx) : zq, b0 = h0 ? ldq_nt(y + i0 * incy) : zq;
  const q128 a1 = h1 ? ldq_nt(x + (i0 + stride) * incx) : zq, b1 = h1 ? ldq_nt(y + (i0 + stride) * incy) : zq;
  // block-uniform anchor, centred between the smallest and largest first products of the block
  const int e0 = h0 ? pair_exp(a0, b0) : EXP_NONE;
  const int mn = wave_min(e0), mx = wave_max(e0 == EXP_NONE ? -EXP_NONE : e0);
  if (lane == 0) {
    s_mn[wv] = mn;
    s_mx[wv] = mx;
  }
  __syncthreads();
  int bmn = s_mn[0], bmx = s_mx[0];
#pragma unroll
  for (int w = 1; w < XD_BS / 64; ++w) {
    bmn = min(bmn, s_mn[w]);
    bmx = max(bmx, s_mx[w]);
  }
  const int A = (bmn == EXP_NONE) ? 0 : ((bmn + bmx) >> 1) - WIN2_CENTER;
  uint32_t flags = 0;
  if (h0) dig_accumulate(d, A, a0, b0, flags);
  if (h1) dig_accumulate(d, A, a1, b1, flags);
  int64_t i = i0 + 2 * stride;
  if (incx == 1 && incy == 1) {
    // two element pairs in flight per step (four: 10% slower)
    for (; i + stride < n; i += 2 * stride) {
      const q128 u0 = ldq_nt(x + i), v0 = ldq_nt(y + i);
      const q128 u1 = ldq_nt(x + i + stride), v1 = ldq_nt(y + i + stride);
      dig_accumulate(d, A, u0, v0, flags);
      dig_accumulate(d, A, u1, v1, flags);
    }
    for (; i < n; i += stride) dig_accumulate(d, A, ldq_nt(x + i), ldq_nt(y + i), flags);
  } else {
    for (; i < n; i += stride) dig_accumulate(d, A, ldq_nt(x + i * incx), ldq_nt(y + i * incy), flags);
  }
  // block reduction in the digit domain: digit k of the block is the int64 sum of the 256 lanes'
  // digit k (each lane's digit takes < 2^29 + 1 per product: no overflow for any practical n);
  // DP threads per digit each add 256 / DP of them, then a shuffle reduction over the DP threads
  constexpr int DP = 8;
  static_assert(DIG * DP <= XD_BS && 64 % DP == 0, "digit sums: DP threads per digit within one wave");
  flags = wave_or(flags);
  if (lane == 0) s_fl[wv] = flags;
  __syncthreads();  // every lane's digits are final
  if (threadIdx.x < DIG * DP) {
    const int k = threadIdx.x / DP, q = threadIdx.x % DP;
    int64_t sum = 0;
#pragma unroll
    for (int w = 0; w < XD_BS / 64; ++w) {
#pragma unroll
      for (int j = 0; j < 64 / DP; ++j) sum += (int64_t)s_dig[w][k][q + DP * j];
    }
#pragma unroll
    for (int m = 1; m < DP; m <<= 1) sum += __shfl_xor(sum, m);
    if (q == 0) s_dig[0][k][0] = (unsigned long long)sum;  // only this thread reads that slot
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t Wl[WW], Wh[WW];
    dig_to_windows(&s_dig[0][0][0], Wl, Wh);
#pragma unroll
    for (int w = 1; w < XD_BS / 64; ++w) flags |= s_fl[w];
    // a zero window is marked by the anchor EXP_NONE, so the combine reads 8 bytes per record
    // to find the range of the nonzero ones
    part_t *p = parts + 2 * (int64_t)blockIdx.x;
    put_parts(p, A, flags, Wl, Wh);
    if (win_is_zero(Wl)) p[0].A = EXP_NONE;
    if (win_is_zero(Wh)) p[1].A = EXP_NONE;
  }
}

// ------------------------------------------------------------------------------------------
// gemv
// ------------------------------------------------------------------------------------------
constexpr int XR_BS = 256;

// The vector operand prepared once per call (exact_dev.hpp: xsplit), so the streaming kernels
// unpack only the matrix element of each product.
__global__ void __launch_bounds__(256) k_vec_split(int64_t n, const q128 *__restrict__ x, int64_t incx,
                                                  uint4 *__restrict__ xl, int32_t *__restrict__ xt) {
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 256) {
    const xsplit v = make_xsplit(ldq_c(x + j * incx));
    xl[j] = make_uint4(v.B[0], v.B[1], v.B[2], v.B[3]);
    xt[j] = v.tag;
  }
}

// RowMajor: wave = (row, chunk).  Lanes stride the chunk's columns (coalesced).
__global__ void __launch_bounds__(XR_BS) k_gemv_row_exact(int64_t m, int64_t n, int64_t chunk, int nch,
                                                          const q128 *__restrict__ A, int64_t lda,
                                                          const q128 *__restrict__ x, int64_t incx,
                                                          const uint4 *__restrict__ xl, const int32_t *__restrict__ xt,
                                                          part_t *__restrict__ parts,
                                                          uint32_t *__restrict__ hdr) {
  __shared__ unsigned long long s_dig[XR_BS / 64][DIG][64];
  const int lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && threadIdx.x < 4) hdr[threadIdx.x] = 0;  // status and the full-range queue count
  unsigned long long *d = &s_dig[threadIdx.x >> 6][0][lane];
  const int64_t nw = (int64_t)gridDim.x * (XR_BS / 64);
  // whole waves stride over the (row, chunk) items
  // the wave index as a scalar: the item's row, chunk and row address stay in scalar registers
  for (int64_t wid = (int64_t)blockIdx.x * (XR_BS / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
       wid < m * nch; wid += nw) {
    const int64_t row = wid / nch;
    const int c = (int)(wid - row * nch);
    const int64_t j0 = (int64_t)c * chunk;
    const int64_t j1 = (j0 + chunk < n) ? j0 + chunk : n;
    const q128 *r = A + row * lda;
    // anchor from the row's first elements, so every column chunk of the row shares it
    int e0 = EXP_NONE;
    if (lane < n) e0 = pair_exp(ldq_c(r + lane), ldq_c(x + (int64_t)lane * incx));
    const int An = wave_anchor(e0);
    dig_clear(d);
    uint32_t flags = 0;
    int64_t j = j0 + lane;
    for (; j + 192 < j1; j += 256) {
      const q128 a0 = ldq_nt(r + j), a1 = ldq_nt(r + j + 64), a2 = ldq_nt(r + j + 128), a3 = ldq_nt(r + j + 192);
      const xsplit b0 = ldx(xl, xt, j), b1 = ldx(xl, xt, j + 64), b2 = ldx(xl, xt, j + 128),
                   b3 = ldx(xl, xt, j + 192);
      dig_accumulate_x(d, An, a0, b0, x + j * incx, flags);
      dig_accumulate_x(d, An, a1, b1, x + (j + 64) * incx, flags);
      dig_accumulate_x(d, An, a2, b2, x + (j + 128) * incx, flags);
      dig_accumulate_x(d, An, a3, b3, x + (j + 192) * incx, flags);
    }
    for (; j < j1; j += 64) dig_accumulate_x(d, An, ldq_nt(r + j), ldx(xl, xt, j), x + j * incx, flags);
    uint32_t Wl[WW], Wh[WW];
    dig_to_windows(d, Wl, Wh);
    wave_sum_win(Wl);
    wave_sum_win(Wh);
    flags = wave_or(flags);
    if (lane == 0) put_parts(parts + 2 * ((int64_t)c * m + row), An, flags, Wl, Wh);
  }
}

// ColMajor: lane = row i, blockIdx.y = column chunk; per-lane anchor.
constexpr int XC_BS = 256;
__global__ void __launch_bounds__(XC_BS) k_gemv_col_exact(int64_t m, int64_t n, int64_t chunk,
                                                          const q128 *__restrict__ A, int64_t lda,
                                                          const q128 *__restrict__ x, int64_t incx,
                                                          const uint4 *__restrict__ xl, const int32_t *__restrict__ xt,
                                                          part_t *__restrict__ parts,
                                                          uint32_t *__restrict__ hdr) {
  __shared__ unsigned long long s_dig[XC_BS / 64][DIG][64];
  const int64_t i = (int64_t)blockIdx.x * XC_BS + threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.y * chunk;
  const int64_t j1 = (j0 + chunk < n) ? j0 + chunk : n;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 4) hdr[threadIdx.x] = 0;
  if (i >= m) return;
  unsigned long long *d = &s_dig[threadIdx.x >> 6][0][threadIdx.x & 63];
  dig_clear(d);
  // lane anchor: centred on the row's first regular product (first columns of the matrix), so
  // every column chunk of the row shares it
  int e0 = EXP_NONE;
  for (int64_t j = 0; j < n && j < 4 && e0 == EXP_NONE; ++j) e0 = pair_exp(ldq_c(A + j * lda + i), ldq_c(x + j * incx));
  const int An = (e0 == EXP_NONE) ? 0 : e0 - WIN2_CENTER;
  uint32_t flags = 0;
  int64_t j = j0;
  for (; j + 3 < j1; j += 4) {
    const q128 a0 = ldq_nt(A + j * lda + i), a1 = ldq_nt(A + (j + 1) * lda + i);
    const q128 a2 = ldq_nt(A + (j + 2) * lda + i), a3 = ldq_nt(A + (j + 3) * lda + i);
    const xsplit b0 = ldx(xl, xt, j), b1 = ldx(xl, xt, j + 1), b2 = ldx(xl, xt, j + 2), b3 = ldx(xl, xt, j + 3);
    dig_accumulate_x(d, An, a0, b0, x + j * incx, flags);
    dig_accumulate_x(d, An, a1, b1, x + (j + 1) * incx, flags);
    dig_accumulate_x(d, An, a2, b2, x + (j + 2) * incx, flags);
    dig_accumulate_x(d, An, a3, b3, x + (j + 3) * incx, flags);
  }
  for (; j < j1; ++j) dig_accumulate_x(d, An, ldq_nt(A + j * lda + i), ldx(xl, xt, j), x + j * incx, flags);
  uint32_t Wl[WW], Wh[WW];
  dig_to_windows(d, Wl, Wh);
  put_parts(parts + 2 * ((int64_t)blockIdx.y * m + i), An, flags, Wl, Wh);
}

// Combining the dot's block partials and finishing, in one launch of XC <= XF_BLOCKS workgroups
// (one thread per record):
//   every workgroup finds the common base (the minimum anchor over nonzero windows, minus 32) and
//   the widest anchor; each thread shifts its record into a register accumulator of NW words (24
//   when the anchors lie within 320 bits, else ACCW; select-based shifts), the workgroup sums each
//   word column over its threads in 64 bits through LDS (no carry chains) into one record per
//   workgroup, and the last workgroup to finish (a ticket in header word 3) adds the records'
//   columns; its wave 0 (one lane per word) ripples the carries, takes the magnitude and finds the
//   top word with ballots, and rounds once.
//   When a product fell outside its block's windows, or the block anchors lie too far apart for the
//   ACCW-word accumulator, the same workgroups recompute the dot on the full-range path instead:
//   workgroup b accumulates its slice of the elements in LDS (wide.hpp), normalises it and writes
//   its digits to record b; the last workgroup adds the records digit by digit and rounds once
//   (header word 2 = 1: the call took the full-range path).
constexpr int XF_BS = 256;
constexpr int XF_BLOCKS = 8;
constexpr int XF_PARTS = 2048;  // most partial records of a dot (two per block, at most 1024 blocks)
static_assert(XF_PARTS <= XF_BLOCKS * XF_BS, "the combine has one thread per partial record");
constexpr int XF_NARROW = 24;   // accumulator words when the anchors lie within 320 bits
static_assert(ACCW + SW <= 64, "wave 0 holds one word per lane");
struct col_rec {
  uint64_t c[ACCW];  // one workgroup's 64-bit word column sums
};
struct dot_rec {
  int lo, t, neg;
  uint32_t flags;
  uint32_t D[KUL_N];
};

__device__ __forceinline__ void parts_range_flags(const part_t *__restrict__ parts, int nparts, int &amin, int &amax,
                                                  uint32_t &fl) {
  __shared__ int s_min[XF_BS / 64], s_max[XF_BS / 64];
  __shared__ uint32_t s_fl[XF_BS / 64];
  // nparts <= XF_PARTS: each thread loads its (up to) XF_PARTS / XF_BS {anchor, flags} pairs at
  // once (indexes clamped into range, surplus loads ignored) instead of one dependent load per step
  uint2 v[XF_PARTS / XF_BS];
#pragma unroll
  for (int u = 0; u < XF_PARTS / XF_BS; ++u) {
    const int p = min((int)threadIdx.x + u * XF_BS, nparts - 1);
    v[u] = *reinterpret_cast<const uint2 *>(&parts[p].A);
  }
  amin = EXP_NONE;
  amax = -EXP_NONE;
  fl = 0;
#pragma unroll
  for (int u = 0; u < XF_PARTS / XF_BS; ++u) {
    if ((int)threadIdx.x + u * XF_BS >= nparts) continue;
    fl |= v[u].y;
    const int a = (int)v[u].x;
    if (a != EXP_NONE) {  // EXP_NONE marks a zero window
      amin = min(amin, a);
      amax = max(amax, a);
    }
  }
  amin = wave_min(amin);
  amax = wave_max(amax);
  fl = wave_or(fl);
  if ((threadIdx.x & 63) == 0) {
    s_min[threadIdx.x >> 6] = amin;
    s_max[threadIdx.x >> 6] = amax;
    s_fl[threadIdx.x >> 6] = fl;
  }
  __syncthreads();
  amin = s_min[0];
  amax = s_max[0];
  fl = s_fl[0];
#pragma unroll
  for (int w = 1; w < XF_BS / 64; ++w) {
    amin = min(amin, s_min[w]);
    amax = max(amax, s_max[w]);
    fl |= s_fl[w];
  }
}

// last workgroup of a ticket (header word 3, cleared by the partial kernel): true for exactly one
// workgroup, after every other one has made its record visible
__device__ __forceinline__ bool last_block(uint32_t *__restrict__ hdr) {
  __shared__ int s_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t ticket = __hip_atomic_fetch_add(hdr + 3, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = ticket == gridDim.x - 1;
  }
  __syncthreads();
  const bool last = s_last != 0;
  if (last) __threadfence();
  return last;
}

// the full-range path of the dot over the whole grid (see above)
__device__ void dot_wide_grid(kul_t &K, const q128 *__restrict__ x, int64_t incx, const q128 *__restrict__ y,
                              int64_t incy, int64_t n, dot_rec *__restrict__ wrec, q128 *__restrict__ out,
                              uint32_t *__restrict__ hdr) {
  kul_clear_all(K);
  __syncthreads();
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t l0 = (int64_t)blockIdx.x * chunk;
  const int64_t l1 = (l0 + chunk < n) ? l0 + chunk : n;
  int e0 = EXP_NONE;
  if (l0 + (int64_t)threadIdx.x < l1) e0 = pair_exp(ldq_c(x + (l0 + threadIdx.x) * incx), ldq_c(y + (l0 + threadIdx.x) * incy));
  const int An = wave_anchor(e0);
  uint32_t Wl[WW], Wh[WW];
#pragma unroll
  for (int k = 0; k < WW; ++k) {
    Wl[k] = 0;
    Wh[k] = 0;
  }
  uint32_t flags = 0;
  int tlo = KUL_N, thi = -1;
  const uq none = unpack_finite(q_zero(0));
  for (int64_t s0 = l0; s0 < l1; s0 += KUL_SEG) {
    const int64_t s1 = (l1 - s0 > KUL_SEG) ? s0 + KUL_SEG : l1;
    kul_accumulate(K, s0, s1, x, incx, y, incy, An, Wl, Wh, false, none, tlo, thi, flags);
    if (s1 < l1) {
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
    kul_add_window(K, Wl, An, false, none, tlo, thi, flags);
    kul_add_window(K, Wh, An + WIN_BAND, false, none, tlo, thi, flags);
    if (flags) atomicOr(&K.flags, flags);
  }
  kul_publish_range(K, tlo, thi);
  __syncthreads();
  dot_rec *my = wrec + blockIdx.x;
  if (threadIdx.x == 0) {
    const int t = kul_normalize(K);
    my->lo = K.lo;
    my->t = t;
    my->neg = K.neg;
    my->flags = K.flags;
  }
  __syncthreads();
  {
    const int lo = K.lo, hi = K.hi;
    for (int i = lo + (int)threadIdx.x; i <= hi; i += blockDim.x) {
      my->D[i] = (uint32_t)K.L[i];
      K.L[i] = 0;
    }
  }
  if (!last_block(hdr)) return;
  // last workgroup: digit sums of all records (each < 2^32 * gridDim.x), then one rounding
  __shared__ int s_lo, s_hi;
  __shared__ uint32_t s_fl;
  if (threadIdx.x == 0) {
    s_lo = KUL_N;
    s_hi = -1;
    s_fl = 0;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
    const dot_rec *r = wrec + b;
    const int top = r->neg ? r->t + 1 : r->t;
    if (r->lo <= top) {
      atomicMin(&s_lo, r->lo);
      atomicMax(&s_hi, top);
    }
    if (r->flags) atomicOr(&s_fl, r->flags);
  }
  __syncthreads();
  const int glo = s_lo, ghi = min(s_hi, KUL_N - 1);
  for (int i = glo + (int)threadIdx.x; i <= ghi; i += blockDim.x) {
    int64_t sum = 0;
    for (int b = 0; b < (int)gridDim.x; ++b) {
      const dot_rec *r = wrec + b;
      if (i >= r->lo && i <= r->t) sum += (int64_t)r->D[i];
      if (r->neg && i == r->t + 1) sum -= 1;
    }
    K.L[i] = (unsigned long long)sum;
  }
  if (threadIdx.x == 0) {
    K.lo = glo;
    K.hi = ghi;
    K.flags = s_fl;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = kul_normalize(K);
    const uint32_t fl = K.flags;
    q128 r;
    if (fl & FL_LOST) {
      r = q_nan();
      atomicOr(hdr, 1u);
    } else if (fl & (FL_NAN | FL_PINF | FL_NINF)) {
      r = special_result(fl);
    } else {
      bool z;
      r = kul_round(K, t, z);
    }
    stq_c(out, r);
    hdr[2] = 1;  // the dot took the full-range path
  }
}

// One workgroup's column sums of its records in NW-word accumulators (base = the lowest anchor
// minus 32; every nonzero record fits: anchor - base + WW * 32 + 32 <= NW * 32).
template <int NW>
__device__ __forceinline__ void dot_combine_cols(const uint4 (&rec)[4], bool have, int base, uint32_t *__restrict__ tile,
                                                 col_rec *__restrict__ mine) {
  uint32_t acc[NW];
  const int pA = (int)rec[0].x;
  if (have && pA != EXP_NONE) {  // EXP_NONE: a zero window
    const uint32_t Wr[WW] = {rec[0].z, rec[0].w, rec[1].x, rec[1].y, rec[1].z, rec[1].w,
                             rec[2].x, rec[2].y, rec[2].z, rec[2].w, rec[3].x, rec[3].y};
    const uint32_t sext = (Wr[WW - 1] >> 31) ? 0xffffffffu : 0u;
    uint32_t W[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) W[k] = (k < WW) ? Wr[k] : sext;
    rshl<NW>(acc, W, pA - base);  // zeros enter from below, the sign extension stays above
  } else {
#pragma unroll
    for (int k = 0; k < NW; ++k) acc[k] = 0;
  }
#pragma unroll
  for (int k = 0; k < NW; ++k) tile[threadIdx.x * (NW + 1) + k] = acc[k];  // the pad spreads the banks
  __syncthreads();
  // G groups of NW threads each sum every G-th row, then NW threads add the G sums
  constexpr int G = XF_BS / NW;
  __shared__ uint64_t gsum[G * NW];
  const int cidx = threadIdx.x % NW, g = threadIdx.x / NW;
  if (g < G) {
    uint64_t sum = 0;
#pragma unroll
    for (int i = 0; i < (XF_BS + G - 1) / G; ++i) {
      const int t = g + i * G;
      if (t < XF_BS) sum += tile[t * (NW + 1) + cidx];
    }
    gsum[g * NW + cidx] = sum;
  }
  __syncthreads();
  if (threadIdx.x < NW) {
    uint64_t sum = 0;
#pragma unroll
    for (int gg = 0; gg < G; ++gg) sum += gsum[gg * NW + threadIdx.x];
    mine->c[threadIdx.x] = sum;
  }
}

// Wave 0 of the last workgroup: lane k < NW adds column k of the nrec records (< 2^43: at most 2048
// words of 32 bits), ripples the carries, and rounds the two's complement sum (bit 0 weighing
// 2^base) once.
template <int NW>
__device__ __forceinline__ void dot_round_wave(const col_rec *__restrict__ recs, int nrec, int base, uint32_t fl,
                                               q128 *__restrict__ out) {
  const int lane = threadIdx.x;
  const bool in = lane < NW;
  uint64_t col = 0;
  if (in) {
    uint64_t v[XF_BLOCKS];  // every record's column loaded at once (surplus loads clamped, ignored)
#pragma unroll
    for (int b = 0; b < XF_BLOCKS; ++b) v[b] = recs[min(b, nrec - 1)].c[lane];
#pragma unroll
    for (int b = 0; b < XF_BLOCKS; ++b) col += (b < nrec) ? v[b] : 0ull;
  }
  // word k = low half of column k + high half of column k - 1; the carries of those 33-bit sums
  // (generate) run through the words equal to 2^32 - 1 (propagate): found with ballots
  const uint32_t hprev = (uint32_t)__shfl_up((int)(uint32_t)(col >> 32), 1);
  const uint64_t s33 = (uint64_t)(uint32_t)col + (lane ? hprev : 0u);
  uint32_t v = (uint32_t)s33;
  const uint64_t gen = __ballot(in && (s33 >> 32) != 0) << 1;
  const uint64_t prop = __ballot(in && v == 0xffffffffu);
  const uint64_t cin = gen | ((prop + gen) ^ prop ^ gen);
  v += (uint32_t)((cin >> lane) & 1u);  // word `lane` of the sum (mod 2^(32 NW))
  const uint32_t sign = (uint32_t)__shfl((int)v, NW - 1) >> 31;
  uint32_t M = v;
  if (sign) {  // magnitude ~W + 1: the +1 carries through the low words that are zero
    const uint64_t zero = __ballot(in && v == 0u);
    const int tz = __builtin_ctzll(~zero);  // lanes >= NW are never in `zero`
    M = ~v + (lane <= tz ? 1u : 0u);
  }
  if (!in) M = 0;
  const uint64_t nz = __ballot(M != 0u);
  q128 r;
  if (nz == 0) {
    r = (fl & (FL_NAN | FL_PINF | FL_NINF)) ? special_result(fl) : q_zero(0);
  } else {
    const int top = 63 - __builtin_clzll(nz);
    const int low = top >= SW ? top - SW + 1 : 0;
    const uint32_t sticky = (nz & ((1ull << low) - 1ull)) ? 1u : 0u;
    uint32_t T[SW];
#pragma unroll
    for (int j = 0; j < SW; ++j) T[j] = (uint32_t)__shfl((int)M, low + j);  // low + j <= max(top, SW - 1) < 64
    r = finish_reg(T, base + 32 * low, sign, sticky, fl, false, q_one(), q_zero(0), q_zero(0));
  }
  if (lane == 0) stq_c(out, r);
}

__global__ void __launch_bounds__(XF_BS) k_exact_dot_combine(const part_t *__restrict__ parts, int nparts,
                                                             col_rec *__restrict__ recs, const q128 *__restrict__ x,
                                                             int64_t incx, const q128 *__restrict__ y, int64_t incy,
                                                             int64_t n, dot_rec *__restrict__ wrec,
                                                             q128 *__restrict__ out, uint32_t *__restrict__ hdr) {
  // one LDS buffer: the column-sum tile of the common path or the full-range accumulator
  constexpr int TILE_WORDS = XF_BS * (ACCW + 1);
  constexpr size_t SMEM = sizeof(kul_t) > TILE_WORDS * 4 ? sizeof(kul_t) : TILE_WORDS * 4;
  __shared__ __align__(16) unsigned char smem[SMEM];
  // this thread's record, loaded before the range scan so that the two sets of loads overlap
  const int pme = blockIdx.x * XF_BS + threadIdx.x;
  uint4 rec[4];
  {
    const uint4 *src = reinterpret_cast<const uint4 *>(parts + min(pme, nparts - 1));
#pragma unroll
    for (int q = 0; q < 4; ++q) rec[q] = src[q];
  }
  int amin, amax;
  uint32_t fl;
  parts_range_flags(parts, nparts, amin, amax, fl);
  const int base = amin - 32;
  const int span = amax - base + WW * 32 + 32;  // bits the accumulator needs
  const bool wide = (fl & FL_LOST) || (amin != EXP_NONE && span > ACCW * 32);
  if (wide) {  // uniform over the grid
    dot_wide_grid(*reinterpret_cast<kul_t *>(smem), x, incx, y, incy, n, wrec, out, hdr);
    return;
  }
  if (amin == EXP_NONE) {  // every window zero: uniform over the grid
    if (blockIdx.x == 0 && threadIdx.x == 0)
      stq_c(out, (fl & (FL_NAN | FL_PINF | FL_NINF)) ? special_result(fl) : q_zero(0));
    return;
  }
  uint32_t *tile = reinterpret_cast<uint32_t *>(smem);
  const bool narrow = span <= XF_NARROW * 32;
  if (narrow)
    dot_combine_cols<XF_NARROW>(rec, pme < nparts, base, tile, recs + blockIdx.x);
  else
    dot_combine_cols<ACCW>(rec, pme < nparts, base, tile, recs + blockIdx.x);
  if (!last_block(hdr) || threadIdx.x >= 64) return;
  if (narrow)
    dot_round_wave<XF_NARROW>(recs, gridDim.x, base, fl, out);
  else
    dot_round_wave<ACCW>(recs, gridDim.x, base, fl, out);
}

// one thread per row: sums the row's chunk windows (they share the row's anchor), then
// RN(alpha*S + beta*y) in registers.  A row with a product outside its windows is queued (header
// word 2 counts the queue) and left for k_exact_wide_rows; its y is not written here.
__global__ void __launch_bounds__(64) k_exact_finish_rows(int64_t m, int nch, const part_t *__restrict__ parts,
                                                         q128 alpha, q128 beta, q128 *__restrict__ y, int64_t incy,
                                                         uint32_t *__restrict__ hdr, int64_t *__restrict__ queue) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    uint32_t lo[WW], hi[WW];
#pragma unroll
    for (int k = 0; k < WW; ++k) {
      lo[k] = 0;
      hi[k] = 0;
    }
    uint32_t fl = 0;
    const int A0 = parts[2 * i].A;
    int c = 0;
    // four chunks' records loaded before they are added: one row per thread leaves few waves, so
    // the loads must overlap (window additions are exact, so their order does not matter)
    for (; c + 3 < nch; c += 4) {
      uint32_t X[4][WW], Y[4][WW];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const part_t *p = parts + 2 * ((int64_t)(c + u) * m + i);
        fl |= p[0].flags | p[1].flags;
        if (p[0].A != A0) fl |= FL_LOST;
#pragma unroll
        for (int k = 0; k < WW; ++k) {
          X[u][k] = p[0].W[k];
          Y[u][k] = p[1].W[k];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        win_add_win(lo, X[u]);
        win_add_win(hi, Y[u]);
      }
    }
    for (; c < nch; ++c) {
      const part_t *p = parts + 2 * ((int64_t)c * m + i);
      fl |= p[0].flags | p[1].flags;
      if (p[0].A != A0) fl |= FL_LOST;  // chunks of a row share its anchor by construction
      uint32_t X[WW], Y[WW];
#pragma unroll
      for (int k = 0; k < WW; ++k) {
        X[k] = p[0].W[k];
        Y[k] = p[1].W[k];
      }
      win_add_win(lo, X);
      win_add_win(hi, Y);
    }
    q128 *yi = y + i * incy;
    if (fl & FL_LOST) {
      const uint32_t slot = atomicAdd(hdr + 2, 1u);
      queue[slot] = i;
    } else {
      uint32_t M[SW], sign;
      s_from_windows(M, sign, lo, hi);
      stq_c(yi, finish_reg(M, A0, sign, 0u, fl, true, alpha, beta, ldq_c(yi)));
    }
  }
}

// The queued rows on the full-range path, one workgroup per row (grid-stride over the queue; the
// whole grid returns at once when the queue is empty).
constexpr int XW_ROW_BLOCKS = 256;
__global__ void __launch_bounds__(256) k_exact_wide_rows(int colmajor, int64_t n, const q128 *__restrict__ A,
                                                        int64_t lda, const q128 *__restrict__ x, int64_t incx,
                                                        q128 alpha, q128 beta, q128 *__restrict__ y, int64_t incy,
                                                        uint32_t *__restrict__ hdr,
                                                        const int64_t *__restrict__ queue) {
  __shared__ kul_t K;
  const uint32_t count = hdr[2];
  if (blockIdx.x >= count) return;
  kul_clear_all(K);
  __syncthreads();
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    const int64_t i = queue[e];
    const q128 *row = colmajor ? A + i : A + i * lda;
    q128 *yi = y + i * incy;
    const q128 r = wide_item(K, n, row, colmajor ? lda : 1, x, incx, true, alpha, beta, ldq_c(yi), hdr);
    if (threadIdx.x == 0) stq_c(yi, r);
  }
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
static int dot_exact_blocks(int64_t n) {
  int64_t b = (n + XD_BS * 8 - 1) / (XD_BS * 8);
  if (b < 1) b = 1;
  if (b > XF_PARTS / 2) b = XF_PARTS / 2;  // 1024; at n = 2^24: 2048 blocks 6% slower, 512 10%
  return (int)b;
}
static void gemv_exact_split(int colmajor, int64_t m, int64_t n, int &nch, int64_t &chunk) {
  int64_t want;
  if (!colmajor) {
    // waves = m * nch >= ~8192, chunk >= 256 columns
    want = (8192 + m - 1) / m;
    const int64_t maxc = (n + 255) / 256;
    if (want > maxc) want = maxc;
  } else {
    const int64_t rowblocks = (m + XC_BS - 1) / XC_BS;
    // 1024 workgroups: at 8192^2 the call takes 264 us against 282 us with 2048 (twice the chunk
    // records to write and to sum in the finish), 272 with 768 or 1280, 300 with 512
    want = (1024 + rowblocks - 1) / rowblocks;
    const int64_t maxc = (n + 63) / 64;
    if (want > maxc) want = maxc;
  }
  if (want < 1) want = 1;
  if (want > 4096) want = 4096;
  chunk = (n + want - 1) / want;
  nch = (int)((n + chunk - 1) / chunk);
}
// workspace of the dot: [header 256 B | partials | column records | full-range records]
static size_t dot_exact_recs_off(int64_t n) {
  return 256 + 2 * (size_t)dot_exact_blocks(n) * sizeof(part_t) + XF_BLOCKS * sizeof(col_rec);
}
size_t dot_exact_ws(int64_t n) { return (dot_exact_recs_off(n) + 255) / 256 * 256 + XF_BLOCKS * sizeof(dot_rec); }
// workspace of gemv: [header 256 B | partials | queue of rows for the full-range path | the
// prepared vector: n limb quadruples (uint4), then n tags (int32), 256-byte aligned]
static size_t gemv_exact_queue_off(int colmajor, int64_t m, int64_t n) {
  int nch;
  int64_t chunk;
  gemv_exact_split(colmajor, m, n, nch, chunk);
  return 256 + 2 * (size_t)nch * (size_t)m * sizeof(part_t);
}
static size_t gemv_exact_xs_off(int colmajor, int64_t m, int64_t n) {
  return (gemv_exact_queue_off(colmajor, m, n) + (size_t)m * sizeof(int64_t) + 255) / 256 * 256;
}
size_t gemv_exact_ws(int colmajor, int64_t m, int64_t n) {
  return gemv_exact_xs_off(colmajor, m, n) + (size_t)n * (sizeof(uint4) + sizeof(int32_t));
}

// workspace: [status (u32) | pad to 256 | partials]
hipError_t launch_dot_exact(const q128 *x, int64_t incx, const q128 *y, int64_t incy, int64_t n, void *ws,
                            q128 *out, hipStream_t st) {
  uint32_t *hdr = (uint32_t *)ws;
  const int blocks = dot_exact_blocks(n);
  part_t *parts = (part_t *)((char *)ws + 256);
  col_rec *recs = (col_rec *)(parts + 2 * blocks);
  dot_rec *wrec = (dot_rec *)((char *)ws + (dot_exact_recs_off(n) + 255) / 256 * 256);
  hipLaunchKernelGGL(k_dot_exact_partial, dim3(blocks), dim3(XD_BS), 0, st, x, incx, y, incy, n, parts, hdr);
  const int nparts = 2 * blocks;
  const int cblocks = (nparts + XF_BS - 1) / XF_BS;  // one thread per record (nparts <= XF_PARTS)
  hipLaunchKernelGGL(k_exact_dot_combine, dim3(cblocks), dim3(XF_BS), 0, st, parts, nparts, recs, x, incx, y, incy, n,
                     wrec, out, hdr);
  return hipGetLastError();
}

hipError_t launch_gemv_exact(int colmajor, int64_t m, int64_t n, q128 alpha, const q128 *A, int64_t lda,
                             const q128 *x, int64_t incx, q128 beta, q128 *y, int64_t incy, void *ws,
                             hipStream_t st) {
  uint32_t *status = (uint32_t *)ws;
  part_t *parts = (part_t *)((char *)ws + 256);
  int nch;
  int64_t chunk;
  gemv_exact_split(colmajor, m, n, nch, chunk);
  uint4 *xl = (uint4 *)((char *)ws + gemv_exact_xs_off(colmajor, m, n));
  int32_t *xt = (int32_t *)(xl + n);
  hipLaunchKernelGGL(k_vec_split, dim3(stride_blocks(n, 256)), dim3(256), 0, st, n, x, incx, xl, xt);
  if (!colmajor) {
    const int64_t waves = m * nch;
    hipLaunchKernelGGL(k_gemv_row_exact, dim3(stride_blocks(waves, XR_BS / 64)), dim3(XR_BS), 0,
                       st, m, n, chunk, nch, A, lda, x, incx, xl, xt, parts, status);
  } else {
    if (!grid_fits((m + XC_BS - 1) / XC_BS, XC_BS)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gemv_col_exact, dim3((unsigned)((m + XC_BS - 1) / XC_BS), (unsigned)nch), dim3(XC_BS), 0,
                       st, m, n, chunk, A, lda, x, incx, xl, xt, parts, status);
  }
  int64_t *queue = (int64_t *)((char *)ws + gemv_exact_queue_off(colmajor, m, n));
  hipLaunchKernelGGL(k_exact_finish_rows, dim3(stride_blocks(m, 64)), dim3(64), 0, st, m, nch, parts, alpha,
                     beta, y, incy, status, queue);
  const unsigned wb = (unsigned)(m < XW_ROW_BLOCKS ? m : XW_ROW_BLOCKS);
  hipLaunchKernelGGL(k_exact_wide_rows, dim3(wb), dim3(256), 0, st, colmajor, n, A, lda, x, incx, alpha, beta, y,
                     incy, status, queue);
  return hipGetLastError();
}

}  // namespace qbk
