// kernels_ref.hip -- reference-order ("bit-exact") kernels for dot, gemv and gemm, plus the seeded
// input generators.
//
// Each kernel reproduces the evaluation order of the reference routine it cites, with correctly
// rounded binary128 operations (q128.hpp), so on the same inputs it returns the reference's
// result bit for bit:
//   dot  : include/quadblas/algorithms/level1.hpp:14-137
//   gemv : include/quadblas/algorithms/level2.hpp:15-99
//   gemm : include/quadblas/algorithms/level3.hpp:187-336 (kc from detail/blocking.hpp:17-67)
// Every reference chain is inherently sequential, so one lane owns one chain; these kernels are
// latency-bound by construction.  The other forms (kernels_exact.hip, kernels_dd.hip,
// kernels_ozaki.hip) trade the reference order for exact or wider accumulation.
#include "q128.hpp"
#include "kernels.h"

using namespace qb;

namespace qbk {

__device__ __forceinline__ q128 ld(const q128 *p) {
  const uint4 v = *reinterpret_cast<const uint4 *>(p);
  return q_make(((uint64_t)v.y << 32) | v.x, ((uint64_t)v.w << 32) | v.z);
}
__device__ __forceinline__ void st(q128 *p, q128 v) {
  uint4 w;
  w.x = (uint32_t)v.lo;
  w.y = (uint32_t)(v.lo >> 32);
  w.z = (uint32_t)v.hi;
  w.w = (uint32_t)(v.hi >> 32);
  *reinterpret_cast<uint4 *>(p) = w;
}

// ------------------------------------------------------------------------------------------
// dot (level1.hpp)
// ------------------------------------------------------------------------------------------

// Chunk t of T: [t*chunk, t == T-1 ? n : (t+1)*chunk), chunk = n / T  (level1.hpp:47,58-59).
__device__ __forceinline__ void chunk_range(int64_t n, int T, int t, int64_t &s, int64_t &e) {
  const int64_t chunk = n / T;
  s = (int64_t)t * chunk;
  e = (t == T - 1) ? n : s + chunk;
}

// Contiguous: chunk t, lane p in {0,1} runs dot_kernel_vectorized's lane-p chain over the
// chunk's elements start+p, start+p+2, ... for i < len/2 (level1.hpp:20-25).
// Strided: chunk t runs one serial chain over its elements (level1.hpp:111-114, :130-133).
// part[] receives 2*T (contiguous) or T (strided) packed partial chains.
__global__ void k_dot_ref_chains(const q128 *__restrict__ x, int64_t incx, const q128 *__restrict__ y,
                                 int64_t incy, int64_t n, int T, int contiguous, q128 *__restrict__ part) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  chain ch;
  chain_init(ch);
  if (contiguous) {
    const int t = tid >> 1, p = tid & 1;
    if (t >= T) return;
    int64_t s, e;
    chunk_range(n, T, t, s, e);
    const int64_t pairs = (e > s) ? (e - s) / 2 : 0;
    for (int64_t i = 0; i < pairs; ++i) {
      const int64_t k = s + 2 * i + p;
      chain_fma(ch, ld(x + k), ld(y + k));
    }
    st(part + 2 * t + p, chain_get(ch));
  } else {
    const int t = tid;
    if (t >= T) return;
    int64_t s, e;
    chunk_range(n, T, t, s, e);
    for (int64_t i = s; i < e; ++i) chain_fma(ch, ld(x + i * incx), ld(y + i * incy));
    st(part + t, chain_get(ch));
  }
}

// Serial epilogue (one thread): lanes combined by horizontal_sum (quad_vector.hpp:141-150),
// odd tail fma (level1.hpp:29-32), chunk partials summed from +0 in chunk order (:67-71,
// :117-120).  'direct' marks the n < 500 paths, which return the kernel's value without the
// final "0 + partial" sum (level1.hpp:40-43, :128-134).
__global__ void k_dot_ref_finish(const q128 *__restrict__ x, const q128 *__restrict__ y, int64_t n, int T,
                                 int contiguous, int direct, const q128 *__restrict__ part,
                                 q128 *__restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  q128 r = q_zero(0);
  for (int t = 0; t < T; ++t) {
    q128 pt;
    if (contiguous) {
      int64_t s, e;
      chunk_range(n, T, t, s, e);
      if (e > s) {
        pt = add_packed(ld(part + 2 * t), ld(part + 2 * t + 1));
        if ((e - s) & 1) pt = fma_packed(ld(x + e - 1), ld(y + e - 1), pt);
      } else {
        pt = q_zero(0);
      }
    } else {
      pt = ld(part + t);
    }
    r = direct ? pt : add_packed(r, pt);
  }
  st(out, r);
}

// ------------------------------------------------------------------------------------------
// gemv (level2.hpp)
// ------------------------------------------------------------------------------------------

// RowMajor, incx == 1: row i's two lane chains run on lanes 2i and 2i+1 (dot_kernel_vectorized
// on the row, level2.hpp:35-38); the pair is combined with a lane exchange, the odd tail applied,
// then y_i = fma(alpha, sum, mul(beta, y_i)) (level2.hpp:48).
// RowMajor, incx != 1: one chain per row (level2.hpp:41-44).
__global__ void k_gemv_row_ref(int64_t m, int64_t n, q128 alpha, const q128 *__restrict__ A, int64_t lda,
                               const q128 *__restrict__ x, int64_t incx, q128 beta, q128 *__restrict__ y,
                               int64_t incy) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  chain ch;
  chain_init(ch);
  if (incx == 1) {
    const int64_t i = tid >> 1;
    const int p = (int)(tid & 1);
    const bool live = i < m;
    if (live) {
      const q128 *row = A + i * lda;
      const int64_t pairs = n / 2;
      for (int64_t j = 0; j < pairs; ++j) chain_fma(ch, ld(row + 2 * j + p), ld(x + 2 * j + p));
    }
    const q128 mine = chain_get(ch);
    // exchange with the partner lane (same row); both lanes of a pair are live or dead together
    const uint64_t olo = (uint64_t)__shfl_xor((unsigned long long)mine.lo, 1);
    const uint64_t ohi = (uint64_t)__shfl_xor((unsigned long long)mine.hi, 1);
    if (!live || p != 0) return;
    q128 sum = add_packed(mine, q_make(olo, ohi));
    if (n & 1) sum = fma_packed(ld(A + i * lda + n - 1), ld(x + n - 1), sum);
    q128 *yi = y + i * incy;
    st(yi, fma_packed(alpha, sum, mul_packed(beta, ld(yi))));
  } else {
    const int64_t i = tid;
    if (i >= m) return;
    const q128 *row = A + i * lda;
    for (int64_t j = 0; j < n; ++j) chain_fma(ch, ld(row + j), ld(x + j * incx));
    q128 *yi = y + i * incy;
    st(yi, fma_packed(alpha, chain_get(ch), mul_packed(beta, ld(yi))));
  }
}

// ColMajor: ax[j] = mul(alpha, x_j) once per column (level2.hpp:71).
__global__ void k_scale_vec(int64_t n, q128 alpha, const q128 *__restrict__ x, int64_t incx,
                            q128 *__restrict__ ax) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
    st(ax + j, mul_packed(alpha, ld(x + j * incx)));
}
// ColMajor: y_i = mul(beta, y_i) (level2.hpp:63-66), then y_i = fma(A_ij, ax_j, y_i) for
// j = 0..n-1 (level2.hpp:69-80); lane i owns row i, so loads of A columns coalesce.
__global__ void k_gemv_col_ref(int64_t m, int64_t n, const q128 *__restrict__ A, int64_t lda,
                               const q128 *__restrict__ ax, q128 beta, q128 *__restrict__ y, int64_t incy) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    q128 *yi = y + i * incy;
    chain ch;
    chain_set(ch, mul_packed(beta, ld(yi)));
    for (int64_t j = 0; j < n; ++j) chain_fma(ch, ld(A + j * lda + i), ld(ax + j));
    st(yi, chain_get(ch));
  }
}

// ------------------------------------------------------------------------------------------
// gemm (level3.hpp)
// ------------------------------------------------------------------------------------------

// One lane per C element.  k is processed in blocks of kc; each block is a fresh chain t and
// C = fma(alpha, t, mul(first ? beta : 1, C)) (level3.hpp:297-300 with the micro-kernel's
// epilogue :102-109; gemm_simple :200-210 is the single-block case).  Layout-aware indexing for
// both layouts.  A workgroup computes a 16 x 16 tile of C: per step of GR_KT values of l, its 256
// threads load the tile's 16 x GR_KT elements of A and GR_KT x 16 of B once (coalesced along the
// contiguous dimension), unpack them into LDS, and every lane runs its chain over the step from
// LDS, in the same order l = kk, kk + 1, ... as the reference.  An operand that is zero,
// subnormal, Inf or NaN is marked in LDS and taken from memory again for the packed slow path.
constexpr int GR_KT = 32;
constexpr int32_t GR_SPECIAL = INT32_MIN;  // uq.e of an operand the fast path does not take
__device__ __forceinline__ uq gr_unpack(q128 v) {
  bool sp;
  uq u = unpack_fast(v, sp);
  if (sp) u.e = GR_SPECIAL;
  return u;
}
__global__ void __launch_bounds__(256) k_gemm_ref(int colmajor, int64_t m, int64_t n, int64_t k, int64_t kc,
                                                  q128 alpha, const q128 *__restrict__ A, int64_t lda,
                                                  const q128 *__restrict__ B, int64_t ldb, q128 beta,
                                                  q128 *__restrict__ C, int64_t ldc) {
  __shared__ uq sA[GR_KT][16];
  __shared__ uq sB[GR_KT][16];
  const int ti = threadIdx.x & 15, tj = threadIdx.x >> 4;
  const int64_t i0 = (int64_t)blockIdx.y * 16, j0 = (int64_t)blockIdx.x * 16;
  const int64_t i = i0 + ti, j = j0 + tj;
  const bool live = i < m && j < n;  // every thread takes part in the loads and barriers
  const int64_t as_i = colmajor ? 1 : lda, as_l = colmajor ? lda : 1;
  const int64_t bs_l = colmajor ? 1 : ldb, bs_j = colmajor ? ldb : 1;
  q128 *ci = C + (colmajor ? j * ldc + i : i * ldc + j);
  q128 c = live ? ld(ci) : q_zero(0);
  const q128 *ap = A + i * as_i;
  const q128 *bp = B + j * bs_j;
  for (int64_t kk = 0; kk < k; kk += kc) {
    const int64_t ke = (kk + kc < k) ? kk + kc : k;
    chain ch;
    chain_init(ch);
    for (int64_t l0 = kk; l0 < ke; l0 += GR_KT) {
      const int lt = (int)((ke - l0 < GR_KT) ? ke - l0 : GR_KT);
      __syncthreads();  // the previous step's operands are no longer read
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int e = (int)threadIdx.x + 256 * h;  // 512 = 16 x GR_KT elements of each operand
        // A(i0 + r, l0 + q): l contiguous (RowMajor) -> q fastest; rows contiguous -> r fastest
        const int ra = colmajor ? (e & 15) : (e >> 5), qa = colmajor ? (e >> 4) : (e & 31);
        if (qa < lt && i0 + ra < m) sA[qa][ra] = gr_unpack(ld(A + (i0 + ra) * as_i + (l0 + qa) * as_l));
        // B(l0 + q, j0 + c): columns contiguous (RowMajor) -> c fastest; l contiguous -> q fastest
        const int cb = colmajor ? (e >> 5) : (e & 15), qb = colmajor ? (e & 31) : (e >> 4);
        if (qb < lt && j0 + cb < n) sB[qb][cb] = gr_unpack(ld(B + (l0 + qb) * bs_l + (j0 + cb) * bs_j));
      }
      __syncthreads();
      if (live) {
        for (int q = 0; q < lt; ++q) {
          const uq ua = sA[q][ti], ub = sB[q][tj];
          bool done = false;
          if ((ua.e != GR_SPECIAL) & (ub.e != GR_SPECIAL) & !ch.special) {
            bool ovf = false;
            const uq r = fma_uq(ua, ub, ch.acc, ovf);
            if (!ovf) {
              ch.acc = r;
              done = true;
            }
          }
          if (!done) chain_set(ch, fma_packed(ld(ap + (l0 + q) * as_l), ld(bp + (l0 + q) * bs_l), chain_get(ch)));
        }
      }
    }
    c = fma_packed(alpha, chain_get(ch), mul_packed(kk == 0 ? beta : q_one(), c));
  }
  if (live) st(ci, c);
}

// ------------------------------------------------------------------------------------------
// seeded generators (same definitions as the CPU oracle's or_fill / or_fill_pair_y)
// ------------------------------------------------------------------------------------------

__device__ __forceinline__ uint64_t rand64(uint64_t seed, uint64_t idx) {
  uint64_t z = seed + (idx + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u11(uint64_t r) { return (double)(r >> 11) * 0x1p-52 - 1.0; }
__device__ __forceinline__ q128 full_uniform(uint64_t seed, uint64_t i) {
  const double a = u11(rand64(seed, 2 * i)), b = u11(rand64(seed, 2 * i + 1));
  // b * 2^-60 is exact; the sum is one correctly rounded binary128 add
  return add_packed(from_double(a), mul_packed(from_double(b), from_double(0x1p-60)));
}

// value number i of generator kind (0..4)
__device__ q128 gen_value(int kind, uint64_t seed, int64_t i, int lo_exp, int hi_exp) {
  q128 v;
  if (kind == 0) {
    v = full_uniform(seed, (uint64_t)i);
  } else if (kind == 1) {
    const uint64_t r0 = rand64(seed, 3 * (uint64_t)i), r1 = rand64(seed, 3 * (uint64_t)i + 1),
                   r2 = rand64(seed, 3 * (uint64_t)i + 2);
    const uint32_t span = (uint32_t)(hi_exp - lo_exp);
    const int e = lo_exp + (int)((uint32_t)r0 % span);
    v = q_make(r2, (r0 & (1ull << 63)) | ((uint64_t)(e + QBIAS) << 48) | (r1 >> 16));
  } else if (kind == 2 || kind == 4) {
    const int64_t p = i & ~(int64_t)1;
    if ((p % 8) < 2) {
      const double u = 1.0 + (double)(rand64(seed, (uint64_t)p) >> 12) * 0x1p-52;
      v = mul_packed(from_double(u), from_double(kind == 4 ? 1e22 : 1e20));
      if (i & 1) v.hi ^= 1ull << 63;
    } else {
      v = full_uniform(seed, (uint64_t)i);
    }
  } else {  // kind 3: y vector of the cancellation variant (own stream; equal values on each pair)
    const int64_t p0 = i & ~(int64_t)1;
    const int64_t p = ((p0 % 8) < 2) ? p0 : i;
    v = full_uniform(seed + 1, (uint64_t)p);
  }
  return v;
}

__global__ void k_fill(int kind, uint64_t seed, int64_t n, int lo_exp, int hi_exp, q128 *__restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    st(out + i, gen_value(kind, seed, i, lo_exp, hi_exp));
}

// value number r*cols + c of the generator stored at element (r, c) of a rows x cols matrix
// (RowMajor r*ld + c or ColMajor r + c*ld)
__global__ void k_fill_2d(int kind, uint64_t seed, int64_t rows, int64_t cols, int64_t ld, int colmajor, int lo_exp,
                          int hi_exp, q128 *__restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < rows * cols; g += stride) {
    // walk the storage order so that writes coalesce
    const int64_t r = colmajor ? g % rows : g / cols, c = colmajor ? g / rows : g % cols;
    st(out + (colmajor ? r + c * ld : r * ld + c), gen_value(kind, seed, r * cols + c, lo_exp, hi_exp));
  }
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------

hipError_t launch_dot_ref(const q128 *x, int64_t incx, const q128 *y, int64_t incy, int64_t n, int T,
                          q128 *part, q128 *out, hipStream_t st) {
  const int contiguous = (incx == 1 && incy == 1);
  int Teff = T, direct = 0;
  if (n < 500) {  // level1.hpp:40-43 (contiguous) and :128-134 (strided): one chain set, no sum
    Teff = 1;
    direct = 1;
  }
  const int lanes = contiguous ? 2 * Teff : Teff;
  const int threads = 64;
  const int blocks = (lanes + threads - 1) / threads;
  hipLaunchKernelGGL(k_dot_ref_chains, dim3(blocks), dim3(threads), 0, st, x, incx, y, incy, n, Teff,
                     contiguous, part);
  hipLaunchKernelGGL(k_dot_ref_finish, dim3(1), dim3(64), 0, st, x, y, n, Teff, contiguous, direct, part, out);
  return hipGetLastError();
}

hipError_t launch_gemv_ref(int colmajor, int64_t m, int64_t n, q128 alpha, const q128 *A, int64_t lda,
                           const q128 *x, int64_t incx, q128 beta, q128 *y, int64_t incy, q128 *ax,
                           hipStream_t st) {
  const int threads = 64;
  if (!colmajor) {
    const int64_t lanes = (incx == 1) ? 2 * m : m;
    const int64_t blocks = (lanes + threads - 1) / threads;
    if (!grid_fits(blocks, threads)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gemv_row_ref, dim3((unsigned)blocks), dim3(threads), 0, st, m, n, alpha, A, lda, x,
                       incx, beta, y, incy);
  } else {
    hipLaunchKernelGGL(k_scale_vec, dim3(stride_blocks(n, 256)), dim3(256), 0, st, n, alpha, x, incx, ax);
    hipLaunchKernelGGL(k_gemv_col_ref, dim3(stride_blocks(m, threads)), dim3(threads), 0, st, m, n,
                       A, lda, ax, beta, y, incy);
  }
  return hipGetLastError();
}

hipError_t launch_gemm_ref(int colmajor, int64_t m, int64_t n, int64_t k, int64_t kc, q128 alpha, const q128 *A,
                           int64_t lda, const q128 *B, int64_t ldb, q128 beta, q128 *C, int64_t ldc,
                           hipStream_t st) {
  if (!grid_fits((n + 15) / 16, 256)) return hipErrorInvalidValue;
  dim3 grid((unsigned)((n + 15) / 16), (unsigned)((m + 15) / 16));
  hipLaunchKernelGGL(k_gemm_ref, grid, dim3(256), 0, st, colmajor, m, n, k, kc, alpha, A, lda, B, ldb, beta, C,
                     ldc);
  return hipGetLastError();
}

__global__ void k_zero(q128 *__restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) st(p + i, q_zero(0));
}
hipError_t launch_zero(q128 *p, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero, dim3(stride_blocks(n, 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_fill(int kind, uint64_t seed, int64_t n, int lo_exp, int hi_exp, q128 *out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(stride_blocks(n, 256)), dim3(256), 0, st, kind, seed, n, lo_exp, hi_exp,
                     out);
  return hipGetLastError();
}


hipError_t launch_fill_2d(int kind, uint64_t seed, int64_t rows, int64_t cols, int64_t ld, int colmajor, int lo_exp,
                          int hi_exp, q128 *out, hipStream_t st) {
  const int64_t n = rows * cols;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_2d, dim3(stride_blocks(n, 256)), dim3(256), 0, st, kind, seed, rows, cols, ld,
                     colmajor, lo_exp, hi_exp, out);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// axpy and the square root of nrm2 (level1.hpp:139-223, c_interface.hpp:34-59)
// ------------------------------------------------------------------------------------------
// y_i = fma(alpha, x_i, y_i): one correctly rounded FMA per element, independent of order and
// chunking (axpy_kernel_vectorized's lanes and tail compute the same per-element FMA)
__global__ void k_axpy(int64_t n, q128 alpha, const q128 *__restrict__ x, int64_t incx, q128 *__restrict__ y,
                       int64_t incy) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    st(y + i * incy, fma_packed(alpha, ld(x + i * incx), ld(y + i * incy)));
}
__global__ void k_sqrt1(q128 *p) {
  if (threadIdx.x == 0) st(p, sqrt_packed(ld(p)));
}

hipError_t launch_axpy(int64_t n, q128 alpha, const q128 *x, int64_t incx, q128 *y, int64_t incy, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)blocks), dim3(256), 0, st, n, alpha, x, incx, y, incy);
  return hipGetLastError();
}
hipError_t launch_sqrt1(q128 *p, hipStream_t st) {
  hipLaunchKernelGGL(k_sqrt1, dim3(1), dim3(64), 0, st, p);
  return hipGetLastError();
}
}  // namespace qbk
