// kernels_dd.hip -- double-double (FP64 FMA) forms of dot and gemv.
//
// Same routines as the reference (include/quadblas/algorithms/level1.hpp:80-137,
// level2.hpp:15-99), evaluated in double-double: inputs are split to hi+lo (106 bits), products
// and sums use error-free transforms (dd.hpp).  The summation order is a fixed tree (per-lane
// strided partial sums, then wave/block/grid reductions in a fixed order), so results are
// deterministic for a given size.  These kernels stream A (and x, y for dot) once with coalesced
// 128-bit loads; they are HBM-bound.
#include "dd.hpp"
#include "kernels.h"

using namespace qb;

namespace qbk {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// streamed operand: read once, no reuse -> nontemporal
__device__ __forceinline__ q128 ldq(const q128 *p) {
  const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
  return q_make(((uint64_t)v.y << 32) | v.x, ((uint64_t)v.w << 32) | v.z);
}
__device__ __forceinline__ q128 ldq_cached(const q128 *p) {
  const uint4 v = *reinterpret_cast<const uint4 *>(p);
  return q_make(((uint64_t)v.y << 32) | v.x, ((uint64_t)v.w << 32) | v.z);
}
__device__ __forceinline__ void stq(q128 *p, q128 v) {
  uint4 w;
  w.x = (uint32_t)v.lo;
  w.y = (uint32_t)(v.lo >> 32);
  w.z = (uint32_t)v.hi;
  w.w = (uint32_t)(v.hi >> 32);
  *reinterpret_cast<uint4 *>(p) = w;
}
__device__ __forceinline__ dd ld_dd(const dd *p) {
  const double2 v = *reinterpret_cast<const double2 *>(p);
  dd r;
  r.hi = v.x;
  r.lo = v.y;
  return r;
}
__device__ __forceinline__ void st_dd(dd *p, dd v) {
  double2 w;
  w.x = v.hi;
  w.y = v.lo;
  *reinterpret_cast<double2 *>(p) = w;
}

__device__ __forceinline__ dd shfl_xor_dd(dd v, int m) {
  dd r;
  r.hi = __shfl_xor(v.hi, m);
  r.lo = __shfl_xor(v.lo, m);
  return r;
}
// fixed-order wave reduction (butterfly): every lane ends with the same value
__device__ __forceinline__ dd wave_sum_dd(dd v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = dd_add(v, shfl_xor_dd(v, m));
  return v;
}
// block of BS threads -> lane 0 of wave 0 holds the block sum
template <int BS>
__device__ __forceinline__ dd block_sum_dd(dd v) {
  __shared__ double2 red[BS / 64];
  v = wave_sum_dd(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = make_double2(v.hi, v.lo);
  __syncthreads();
  dd s = dd_zero();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) {
      dd t;
      t.hi = red[i].x;
      t.lo = red[i].y;
      s = dd_add(s, t);
    }
  }
  return s;
}

// ------------------------------------------------------------------------------------------
// dot
// ------------------------------------------------------------------------------------------
constexpr int DOT_BS = 256;

__global__ void __launch_bounds__(DOT_BS) k_dot_dd_partial(const q128 *__restrict__ x, int64_t incx,
                                                           const q128 *__restrict__ y, int64_t incy, int64_t n,
                                                           dd *__restrict__ part) {
  const int64_t stride = (int64_t)gridDim.x * DOT_BS;
  int64_t i = (int64_t)blockIdx.x * DOT_BS + threadIdx.x;
  dd acc0 = dd_zero(), acc1 = dd_zero();
  if (incx == 1 && incy == 1) {
    for (; i + stride < n; i += 2 * stride) {
      const q128 a0 = ldq(x + i), b0 = ldq(y + i), a1 = ldq(x + i + stride), b1 = ldq(y + i + stride);
      acc0 = dd_fma(q_to_dd(a0), q_to_dd(b0), acc0);
      acc1 = dd_fma(q_to_dd(a1), q_to_dd(b1), acc1);
    }
    if (i < n) acc0 = dd_fma(q_to_dd(ldq(x + i)), q_to_dd(ldq(y + i)), acc0);
  } else {
    for (; i < n; i += stride) acc0 = dd_fma(q_to_dd(ldq(x + i * incx)), q_to_dd(ldq(y + i * incy)), acc0);
  }
  const dd s = block_sum_dd<DOT_BS>(dd_add(acc0, acc1));
  if (threadIdx.x == 0) st_dd(part + blockIdx.x, s);
}

__global__ void __launch_bounds__(256) k_dot_dd_final(const dd *__restrict__ part, int nparts,
                                                      q128 *__restrict__ out) {
  dd acc = dd_zero();
  for (int i = threadIdx.x; i < nparts; i += 256) acc = dd_add(acc, ld_dd(part + i));
  const dd s = block_sum_dd<256>(acc);
  if (threadIdx.x == 0) stq(out, dd_to_q(s));
}

// ------------------------------------------------------------------------------------------
// gemv
// ------------------------------------------------------------------------------------------

// xs[j] = dd(scale * x_j) with the scale applied in binary128 first (ColMajor pre-rounds
// alpha*x_j like level2.hpp:71; RowMajor uses scale = 1).
__global__ void k_vec_to_dd(int64_t n, q128 scale, int use_scale, const q128 *__restrict__ x, int64_t incx,
                            dd *__restrict__ xs) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    q128 v = ldq_cached(x + j * incx);
    if (use_scale) v = mul_packed(scale, v);
    st_dd(xs + j, q_to_dd(v));
  }
}

// RowMajor: R rows per wave (R = 2, and 1 for the last row of an odd m), lanes stride along the
// rows (coalesced), GV_U column steps of 64 per iteration so that 2 * GV_U loads of A are in
// flight per lane and each x element read serves both rows; dd partials per lane, butterfly wave
// sum, then y_i = alpha*sum + beta*y_i in binary128.
constexpr int GV_BS = 256;
constexpr int GV_U = 4;
template <int R>
__device__ __forceinline__ void gemv_rows_dd(int64_t row0, int64_t n, q128 alpha, const q128 *__restrict__ A,
                                             int64_t lda, const dd *__restrict__ xs, q128 beta, q128 *__restrict__ y,
                                             int64_t incy, int lane) {
  dd acc[R][GV_U];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int u = 0; u < GV_U; ++u) acc[r][u] = dd_zero();
  const q128 *rp = A + row0 * lda;
  int64_t j = lane;
  for (; j + 64 * (GV_U - 1) < n; j += 64 * GV_U) {
    dd xv[GV_U];
    q128 a[R][GV_U];
#pragma unroll
    for (int u = 0; u < GV_U; ++u) xv[u] = ld_dd(xs + j + 64 * u);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int u = 0; u < GV_U; ++u) a[r][u] = ldq(rp + r * lda + j + 64 * u);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int u = 0; u < GV_U; ++u) acc[r][u] = dd_fma(q_to_dd(a[r][u]), xv[u], acc[r][u]);
  }
  for (; j < n; j += 64) {
    const dd xv = ld_dd(xs + j);
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r][0] = dd_fma(q_to_dd(ldq(rp + r * lda + j)), xv, acc[r][0]);
  }
  // every lane holds every row's sum after the butterflies; lane r finishes row r, so the R
  // binary128 epilogues run side by side
  dd mine = dd_zero();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const dd s = wave_sum_dd(dd_add(dd_add(acc[r][0], acc[r][1]), dd_add(acc[r][2], acc[r][3])));
    if (lane == r) mine = s;
  }
  if (lane < R) {
    q128 *yi = y + (row0 + lane) * incy;
    stq(yi, fma_packed(alpha, dd_to_q(mine), mul_packed(beta, ldq_cached(yi))));
  }
}
static_assert(GV_U == 4, "the per-row combine above sums four accumulators");
// Waves take row pairs grid-stride (a grid of 3 blocks per CU streams best: the row pairs of
// one wave then follow each other, and the waves of a round end at different times).
constexpr int GV_BLOCKS = 768;
__global__ void __launch_bounds__(GV_BS) k_gemv_row_dd(int64_t m, int64_t n, q128 alpha, const q128 *__restrict__ A,
                                                       int64_t lda, const dd *__restrict__ xs, q128 beta,
                                                       q128 *__restrict__ y, int64_t incy) {
  const int64_t W = (int64_t)gridDim.x * (GV_BS / 64);
  const int lane = threadIdx.x & 63;
  // the wave index as a scalar: the row pair's addresses stay in scalar registers
  for (int64_t row0 = ((int64_t)blockIdx.x * (GV_BS / 64) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))) * 2;
       row0 < m; row0 += 2 * W) {
    if (row0 + 1 < m)  // uniform per wave
      gemv_rows_dd<2>(row0, n, alpha, A, lda, xs, beta, y, incy, lane);
    else
      gemv_rows_dd<1>(row0, n, alpha, A, lda, xs, beta, y, incy, lane);
  }
}

// ColMajor: lane = row, blockIdx.y = column chunk; GC_U columns per iteration (GC_U loads of A in
// flight per lane); partial dd sums per (chunk, row).
constexpr int GC_BS = 256;
constexpr int GC_U = 8;
__global__ void __launch_bounds__(GC_BS) k_gemv_col_dd_partial(int64_t m, int64_t n, int64_t chunk,
                                                               const q128 *__restrict__ A, int64_t lda,
                                                               const dd *__restrict__ axs, dd *__restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * GC_BS + threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.y * chunk;
  const int64_t j1 = (j0 + chunk < n) ? j0 + chunk : n;
  if (i >= m) return;
  dd acc[GC_U];
#pragma unroll
  for (int u = 0; u < GC_U; ++u) acc[u] = dd_zero();
  int64_t j = j0;
  for (; j + GC_U - 1 < j1; j += GC_U) {
    q128 a[GC_U];
#pragma unroll
    for (int u = 0; u < GC_U; ++u) a[u] = ldq(A + (j + u) * lda + i);
#pragma unroll
    for (int u = 0; u < GC_U; ++u) acc[u] = dd_fma(q_to_dd(a[u]), ld_dd(axs + j + u), acc[u]);
  }
  for (; j < j1; ++j) acc[0] = dd_fma(q_to_dd(ldq(A + j * lda + i)), ld_dd(axs + j), acc[0]);
  dd s = acc[0];
#pragma unroll
  for (int u = 1; u < GC_U; ++u) s = dd_add(s, acc[u]);
  st_dd(part + blockIdx.y * m + i, s);
}
// y_i = RN(beta * y_i) + RN(sum of the chunk partials), the partials read 8 at a time
__global__ void k_gemv_col_dd_final(int64_t m, int nchunks, const dd *__restrict__ part, q128 beta,
                                    q128 *__restrict__ y, int64_t incy) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    dd s = dd_zero();
    int c = 0;
    for (; c + 7 < nchunks; c += 8) {
      dd t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = ld_dd(part + (int64_t)(c + u) * m + i);
#pragma unroll
      for (int u = 0; u < 8; ++u) s = dd_add(s, t[u]);
    }
    for (; c < nchunks; ++c) s = dd_add(s, ld_dd(part + (int64_t)c * m + i));
    q128 *yi = y + i * incy;
    stq(yi, add_packed(mul_packed(beta, ldq_cached(yi)), dd_to_q(s)));
  }
}

// ------------------------------------------------------------------------------------------
// gemm (level3.hpp:215-336 semantics, C = alpha * A B + beta * C, RowMajor; a ColMajor call is
// the RowMajor product C^T = B^T A^T on the same storage)
// ------------------------------------------------------------------------------------------
constexpr int GD_T = 64;   // output tile
constexpr int GD_K = 16;   // k step staged in LDS

// out[r * Cp + c] = dd(X[r * ld + c]) for r < R, c < C; zeros in the padding up to Rp x Cp
__global__ void k_rows_to_dd(const q128 *__restrict__ X, int64_t R, int64_t C, int64_t ld, dd *__restrict__ out,
                             int64_t Rp, int64_t Cp) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < Rp * Cp; g += stride) {
    const int64_t r = g / Cp, c = g - r * Cp;
    st_dd(out + g, (r < R && c < C) ? q_to_dd(ldq(X + r * ld + c)) : dd_zero());
  }
}

// C entry = RN(alpha * RN(acc) + RN(beta * c)), the reference's epilogue (level3.hpp:297-300);
// kept out of line so the 16-entry epilogue loop stays unrolled with the accumulators in registers
__device__ __attribute__((noinline)) q128 dd_gemm_out(double hi, double lo, q128 alpha, q128 beta, q128 c) {
  dd v;
  v.hi = hi;
  v.lo = lo;
  return fma_packed(alpha, dd_to_q(v), mul_packed(beta, c));
}

// 256 threads = 16 x 16; thread (tx, ty) owns outputs (ty + 16 r, tx + 16 c), r, c < 4.  Per k
// step the tile A[64 x 16] (stored k-major in LDS) and B[16 x 64] are staged through registers,
// so the next step's global loads overlap this step's 16 x 16 double-double FMAs.
__global__ void __launch_bounds__(256) k_gemm_dd(int64_t M, int64_t N, int64_t Kp, const dd *__restrict__ Ad,
                                                 const dd *__restrict__ Bd, int64_t Np, q128 alpha, q128 beta,
                                                 q128 *__restrict__ C, int64_t ldc) {
  __shared__ dd sA[GD_K][GD_T];
  __shared__ dd sB[GD_K][GD_T];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, t = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.y * GD_T, j0 = (int64_t)blockIdx.x * GD_T;
  // loader roles: A row t/4, k (t%4)*4..+3;  B row t/16, columns (t%16)*4..+3
  const int ar = t >> 2, ak = (t & 3) * 4, br = t >> 4, bc = (t & 15) * 4;
  const dd *ap = Ad + (i0 + ar) * Kp + ak;
  const dd *bp = Bd + (int64_t)br * Np + j0 + bc;
  dd ra[4], rb[4];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ra[q] = ld_dd(ap + k0 + q);
      rb[q] = ld_dd(bp + k0 * Np + q);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      sA[ak + q][ar] = ra[q];
      sB[br][bc + q] = rb[q];
    }
  };
  dd acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = dd_zero();
  gload(0);
  for (int64_t k0 = 0; k0 < Kp; k0 += GD_K) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + GD_K < Kp) gload(k0 + GD_K);
#pragma unroll
    for (int k = 0; k < GD_K; ++k) {
      dd a[4], b[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = sA[k][ty + 16 * q];
        b[q] = sB[k][tx + 16 * q];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = dd_fma(a[r], b[c], acc[r][c]);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t i = i0 + ty + 16 * r, j = j0 + tx + 16 * c;
      if (i < M && j < N) {
        q128 *cp = C + i * ldc + j;
        stq(cp, dd_gemm_out(acc[r][c].hi, acc[r][c].lo, alpha, beta, ldq_cached(cp)));
      }
    }
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
// at most 3 blocks per CU: fewer, longer-running blocks stream faster than a deeper grid here
int dot_dd_blocks(int64_t n) {
  int64_t b = (n + DOT_BS * 8 - 1) / (DOT_BS * 8);
  if (b < 1) b = 1;
  if (b > 768) b = 768;
  return (int)b;
}
size_t dot_dd_ws(int64_t n) { return (size_t)dot_dd_blocks(n) * sizeof(dd); }

hipError_t launch_dot_dd(const q128 *x, int64_t incx, const q128 *y, int64_t incy, int64_t n, void *ws, q128 *out,
                         hipStream_t st) {
  const int blocks = dot_dd_blocks(n);
  dd *part = (dd *)ws;
  hipLaunchKernelGGL(k_dot_dd_partial, dim3(blocks), dim3(DOT_BS), 0, st, x, incx, y, incy, n, part);
  hipLaunchKernelGGL(k_dot_dd_final, dim3(1), dim3(256), 0, st, part, blocks, out);
  return hipGetLastError();
}

int gemv_col_chunks(int64_t m, int64_t n) {
  const int64_t rowblocks = (m + GC_BS - 1) / GC_BS;
  int64_t c = (512 + rowblocks - 1) / rowblocks;  // aim for >= 512 blocks (2 per CU)
  if (c > n) c = n;
  if (c < 1) c = 1;
  if (c > 1024) c = 1024;
  return (int)c;
}
size_t gemv_dd_ws(int colmajor, int64_t m, int64_t n) {
  size_t ws = (size_t)n * sizeof(dd);
  if (colmajor) ws += (size_t)gemv_col_chunks(m, n) * (size_t)m * sizeof(dd);
  return ws;
}

hipError_t launch_gemv_dd(int colmajor, int64_t m, int64_t n, q128 alpha, const q128 *A, int64_t lda,
                          const q128 *x, int64_t incx, q128 beta, q128 *y, int64_t incy, void *ws,
                          hipStream_t st) {
  dd *xs = (dd *)ws;
  if (colmajor && !grid_fits((m + GC_BS - 1) / GC_BS, GC_BS)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_vec_to_dd, dim3(stride_blocks(n, 256)), dim3(256), 0, st, n, alpha, colmajor, x, incx,
                     xs);
  if (!colmajor) {
    const int64_t pairs = (m + 1) / 2;  // two rows per wave and step
    int64_t blocks = (pairs + GV_BS / 64 - 1) / (GV_BS / 64);
    if (blocks > GV_BLOCKS) blocks = GV_BLOCKS;
    hipLaunchKernelGGL(k_gemv_row_dd, dim3((unsigned)blocks), dim3(GV_BS), 0, st, m, n, alpha, A, lda, xs, beta, y,
                       incy);
  } else {
    const int nch = gemv_col_chunks(m, n);
    const int64_t chunk = (n + nch - 1) / nch;
    const int nch_eff = (int)((n + chunk - 1) / chunk);
    dd *part = xs + n;
    hipLaunchKernelGGL(k_gemv_col_dd_partial, dim3((unsigned)((m + GC_BS - 1) / GC_BS), (unsigned)nch_eff),
                       dim3(GC_BS), 0, st, m, n, chunk, A, lda, xs, part);
    hipLaunchKernelGGL(k_gemv_col_dd_final, dim3(stride_blocks(m, 256)), dim3(256), 0, st, m, nch_eff, part,
                       beta, y, incy);
  }
  return hipGetLastError();
}


static int64_t round_up(int64_t v, int64_t q) { return (v + q - 1) / q * q; }

size_t gemm_dd_ws(int64_t m, int64_t n, int64_t k) {
  const int64_t Mp = round_up(m, GD_T), Np = round_up(n, GD_T), Kp = round_up(k, GD_K);
  return (size_t)(Mp * Kp + Kp * Np) * sizeof(dd);
}

hipError_t launch_gemm_dd(int colmajor, int64_t m, int64_t n, int64_t k, q128 alpha, const q128 *A, int64_t lda,
                          const q128 *B, int64_t ldb, q128 beta, q128 *C, int64_t ldc, void *ws, hipStream_t st) {
  if (colmajor) {  // C^T (n x m, row-major, ld = ldc) = B^T (n x k, ld = ldb) A^T (k x m, ld = lda)
    const q128 *t = A;
    A = B;
    B = t;
    const int64_t tl = lda;
    lda = ldb;
    ldb = tl;
    const int64_t tm = m;
    m = n;
    n = tm;
  }
  const int64_t Mp = round_up(m, GD_T), Np = round_up(n, GD_T), Kp = round_up(k, GD_K);
  dd *Ad = (dd *)ws, *Bd = Ad + Mp * Kp;
  if (!grid_fits(Np / GD_T, 256)) return hipErrorInvalidValue;
  k_rows_to_dd<<<stride_blocks(Mp * Kp, 256), 256, 0, st>>>(A, m, k, lda, Ad, Mp, Kp);
  k_rows_to_dd<<<stride_blocks(Kp * Np, 256), 256, 0, st>>>(B, k, n, ldb, Bd, Kp, Np);
  dim3 grid((unsigned)(Np / GD_T), (unsigned)(Mp / GD_T));
  k_gemm_dd<<<grid, 256, 0, st>>>(m, n, Kp, Ad, Bd, Np, alpha, beta, C, ldc);
  return hipGetLastError();
}
}  // namespace qbk
