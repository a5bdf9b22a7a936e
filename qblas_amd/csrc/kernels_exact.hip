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
  __shared__ uint32_t s_w[XD_BS / 64][2 * WW + 1];
  __shared__ unsigned long long s_dig[XD_BS / 64][DIG][64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && threadIdx.x < 4) hdr[threadIdx.x] = 0;  // status, counts, ticket of the combine
  unsigned long long *d = &s_dig[wv][0][lane];
  dig_clear(d);
  const int64_t stride = (int64_t)gridDim.x * XD_BS;
  const int64_t i0 = (int64_t)blockIdx.x * XD_BS + threadIdx.x;
  // block-uniform anchor, centred between the smallest and largest first products of the block
  int e0 = EXP_NONE;
  if (i0 < n) e0 = pair_exp(ldq_c(x + i0 * incx), ldq_c(y + i0 * incy));
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
  int64_t i = i0;
  if (incx == 1 && incy == 1) {
    for (; i + stride < n; i += 2 * stride) {
      const q128 a0 = ldq_nt(x + i), b0 = ldq_nt(y + i);
      const q128 a1 = ldq_nt(x + i + stride), b1 = ldq_nt(y + i + stride);
      dig_accumulate(d, A, a0, b0, flags);
      dig_accumulate(d, A, a1, b1, flags);
    }
    if (i < n) dig_accumulate(d, A, ldq_nt(x + i), ldq_nt(y + i), flags);
  } else {
    for (; i < n; i += stride) dig_accumulate(d, A, ldq_nt(x + i * incx), ldq_nt(y + i * incy), flags);
  }
  uint32_t Wl[WW], Wh[WW];
  dig_to_windows(d, Wl, Wh);
  wave_sum_win(Wl);
  wave_sum_win(Wh);
  flags = wave_or(flags);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < WW; ++k) {
      s_w[wv][k] = Wl[k];
      s_w[wv][WW + k] = Wh[k];
    }
    s_w[wv][2 * WW] = flags;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 1; w < XD_BS / 64; ++w) {
      uint32_t X[WW], Y[WW];
#pragma unroll
      for (int k = 0; k < WW; ++k) {
        X[k] = s_w[w][k];
        Y[k] = s_w[w][WW + k];
      }
      win_add_win(Wl, X);
      win_add_win(Wh, Y);
      flags |= s_w[w][2 * WW];
    }
    put_parts(parts + 2 * (int64_t)blockIdx.x, A, flags, Wl, Wh);
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
  for (int64_t wid = (int64_t)blockIdx.x * (XR_BS / 64) + (threadIdx.x >> 6); wid < m * nch; wid += nw) {
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

// Combining the dot's block partials and finishing, in one launch of XC workgroups (XC <= XF_BLOCKS):
//   every workgroup finds the common base (the minimum anchor over nonzero windows, minus 32) and
//   the widest anchor; each thread adds its windows into a register-resident ACCW-word accumulator
//   (select-based shifts), lanes 0..ACCW-1 sum one word column each over the block's threads
//   (64-bit column sums, no carry chains) into one record per workgroup, and the last workgroup to
//   finish (a ticket in header word 3) sums the records' columns, ripples the carries once and
//   rounds.
//   When a product fell outside its block's windows, or the block anchors lie too far apart for the
//   ACCW-word accumulator, the same workgroups recompute the dot on the full-range path instead:
//   workgroup b accumulates its slice of the elements in LDS (wide.hpp), normalises it and writes
//   its digits to record b; the last workgroup adds the records digit by digit and rounds once
//   (header word 2 = 1: the call took the full-range path).
constexpr int XF_BS = 256;
constexpr int XF_BLOCKS = 32;
struct col_rec {
  uint64_t c[ACCW];
  uint32_t flags;
  uint32_t pad[3];
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
  amin = EXP_NONE;
  amax = -EXP_NONE;
  fl = 0;
  for (int p = threadIdx.x; p < nparts; p += XF_BS) {
    fl |= parts[p].flags;
    if (!win_is_zero(parts[p].W)) {
      amin = min(amin, parts[p].A);
      amax = max(amax, parts[p].A);
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

// Round a two's complement ACCW-word sum (bit 0 weighing 2^base) to binary128, in registers.
__device__ __forceinline__ q128 round_accw(const uint32_t (&W)[ACCW], int base, uint32_t fl) {
  const uint32_t sign = W[ACCW - 1] >> 31;
  uint32_t M[ACCW];
  uint32_t c = sign;
#pragma unroll
  for (int k = 0; k < ACCW; ++k) M[k] = sign ? addc32(~W[k], 0u, c, c) : W[k];
  int top = -1;
#pragma unroll
  for (int k = 0; k < ACCW; ++k) top = M[k] ? k : top;
  if (top < 0) return (fl & (FL_NAN | FL_PINF | FL_NINF)) ? special_result(fl) : q_zero(0);
  const int low = top >= SW ? top - SW + 1 : 0;
  uint32_t sticky = 0;
#pragma unroll
  for (int k = 0; k < ACCW; ++k) sticky |= (k < low) ? M[k] : 0u;
  uint32_t S[ACCW];
  (void)rshr<ACCW>(S, M, 32 * low);
  uint32_t T[SW];
#pragma unroll
  for (int k = 0; k < SW; ++k) T[k] = S[k];
  return finish_reg(T, base + 32 * low, sign, sticky ? 1u : 0u, fl, false, q_one(), q_zero(0), q_zero(0));
}

__global__ void __launch_bounds__(XF_BS) k_exact_dot_combine(const part_t *__restrict__ parts, int nparts,
                                                             col_rec *__restrict__ recs, const q128 *__restrict__ x,
                                                             int64_t incx, const q128 *__restrict__ y, int64_t incy,
                                                             int64_t n, dot_rec *__restrict__ wrec,
                                                             q128 *__restrict__ out, uint32_t *__restrict__ hdr) {
  __shared__ uint32_t tile[XF_BS][ACCW + 1];
  __shared__ kul_t K;
  __shared__ uint64_t col[ACCW];
  int amin, amax;
  uint32_t fl;
  parts_range_flags(parts, nparts, amin, amax, fl);
  const int base = amin - 32;
  const bool wide = (fl & FL_LOST) || (amin != EXP_NONE && amax - base + WW * 32 + 32 > ACCW * 32);
  if (wide) {  // uniform over the grid
    dot_wide_grid(K, x, incx, y, incy, n, wrec, out, hdr);
    return;
  }
  uint32_t acc[ACCW];
#pragma unroll
  for (int k = 0; k < ACCW; ++k) acc[k] = 0;
  if (amin != EXP_NONE) {
    for (int p = blockIdx.x * XF_BS + threadIdx.x; p < nparts; p += gridDim.x * XF_BS) {
      if (win_is_zero(parts[p].W)) continue;
      const int off = parts[p].A - base;  // 32 .. ACCW * 32 - WW * 32 - 32 (checked above)
      uint32_t W[ACCW], S[ACCW];
      const uint32_t sext = (parts[p].W[WW - 1] >> 31) ? 0xffffffffu : 0u;
#pragma unroll
      for (int k = 0; k < ACCW; ++k) W[k] = (k < WW) ? parts[p].W[k] : sext;
      rshl<ACCW>(S, W, off);  // zeros enter from below, the sign extension stays above
      uint32_t c = 0;
#pragma unroll
      for (int k = 0; k < ACCW; ++k) acc[k] = addc32(acc[k], S[k], c, c);
    }
  }
#pragma unroll
  for (int k = 0; k < ACCW; ++k) tile[threadIdx.x][k] = acc[k];
  __syncthreads();
  // column sums: six groups of ACCW threads each sum a sixth of the rows, then one adds the six
  constexpr int G = XF_BS / ACCW;  // 6
  __shared__ uint64_t part6[G][ACCW];
  if (threadIdx.x < G * ACCW) {
    const int cidx = threadIdx.x % ACCW, g = threadIdx.x / ACCW;
    uint64_t sum = 0;
    for (int t = g; t < XF_BS; t += G) sum += tile[t][cidx];
    part6[g][cidx] = sum;
  }
  __syncthreads();
  if (threadIdx.x < ACCW) {
    uint64_t sum = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) sum += part6[g][threadIdx.x];
    recs[blockIdx.x].c[threadIdx.x] = sum;
  }
  if (!last_block(hdr)) return;
  if (threadIdx.x < ACCW) {
    uint64_t sum = 0;
    for (int b = 0; b < (int)gridDim.x; ++b) sum += recs[b].c[threadIdx.x];  // < 2^45 for <= 32 records
    col[threadIdx.x] = sum;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  q128 r;
  if (amin == EXP_NONE) {
    r = (fl & (FL_NAN | FL_PINF | FL_NINF)) ? special_result(fl) : q_zero(0);
  } else {
    uint32_t W[ACCW];
    uint64_t carry = 0;
#pragma unroll
    for (int k = 0; k < ACCW; ++k) {
      const uint64_t v = col[k] + carry;  // two's complement words: the top carry is dropped
      W[k] = (uint32_t)v;
      carry = v >> 32;
    }
    r = round_accw(W, base, fl);
  }
  stq_c(out, r);
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
  if (b > 1024) b = 1024;  // at n = 2^24: 2048 blocks 6% slower (twice the partials to combine), 512 10%
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
    want = (2048 + rowblocks - 1) / rowblocks;  // 8192^2: 39% of HBM peak; 1024 or 4096 blocks: 37%, 512: 29%
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
  const int cblocks = nparts < XF_BLOCKS * XF_BS ? (nparts + XF_BS - 1) / XF_BS : XF_BLOCKS;
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
