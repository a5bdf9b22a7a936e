// kernels_batched.hip -- many independent, unevenly sized qgemm / qgemv problems in one launch.
//
// A device array of descriptors (qblas_gemm_problem / qblas_gemv_problem) is turned into a flat
// work list by k_batch_scan (work items per problem, exclusive prefix sums, all on the device),
// then one grouped persistent kernel walks the list: every workgroup (or wave) takes item (or
// chunk) blockIdx.x (or its wave index) first and later ones from a queue word that the scan
// zeroes, finds its problem by binary search in the prefix sums and computes one tile (gemm) or
// one group of rows (gemv).  Nothing is synchronized or copied between the two launches, so the
// sequence can be captured in a hipGraph.
//
// Per problem the semantics are those of the single-problem routines:
//   REFERENCE gemm  level3.hpp:215-336: per entry, k in blocks of kc (kc = k when m, n, k <= 64,
//                   level3.hpp:224-229), each block a fresh correctly rounded FMA chain t, then
//                   C = fma(alpha, t, mul(beta on the first block / 1 after, C))  -> bit-identical
//                   to a loop of single calls.
//   REFERENCE gemv  level2.hpp:15-99: RowMajor two lane chains per row (one when incx != 1),
//                   y = fma(alpha, sum, mul(beta, y)); ColMajor y = mul(beta, y) then
//                   y = fma(A_ij, mul(alpha, x_j), y) over j.
//   DD gemm / gemv  double-double accumulation (dd.hpp), C = fma(alpha, RN(acc), mul(beta, C)).
//   EXACT gemv      RN(alpha * sum_j A_ij x_j + beta * y_i), the exact value rounded once (exact.hpp);
//                   a row with a product outside its accumulation windows is left to
//                   k_gemv_batched_exact_wide, which finds it again and takes the full-range path
//                   (wide.hpp).
// Tile shapes are template parameters; the host picks one configuration per call (the
// autotuner in capi.hip measures them).
#include "dd.hpp"
#include "wide.hpp"
#include "kernels.h"

using namespace qb;

namespace qbk {

__device__ __forceinline__ q128 bld(const q128 *p) {
  const uint4 v = *reinterpret_cast<const uint4 *>(p);
  return q_make(((uint64_t)v.y << 32) | v.x, ((uint64_t)v.w << 32) | v.z);
}
__device__ __forceinline__ void bst(q128 *p, q128 v) {
  uint4 w;
  w.x = (uint32_t)v.lo;
  w.y = (uint32_t)(v.lo >> 32);
  w.z = (uint32_t)v.hi;
  w.w = (uint32_t)(v.hi >> 32);
  *reinterpret_cast<uint4 *>(p) = w;
}
__device__ __forceinline__ q128 qv(const uint64_t (&w)[2]) { return q_make(w[0], w[1]); }

// largest p in [0, count) with start[p] <= t (start is nondecreasing, start[0] = 0 <= t)
__device__ __forceinline__ int64_t find_problem(const int64_t *__restrict__ start, int64_t count, int64_t t) {
  int64_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (start[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// work items of problem p:
//   kind 0 (gemm): tiles of tm x tn outputs (none when m, n or k is 0: level3.hpp:221-222)
//   kind 1 (gemv, REFERENCE): groups of 64 lanes: 32 rows (RowMajor, incx == 1) or 64 rows
//   kind 2 (gemv, DD): RowMajor 4 rows per wave (lanes along the columns); ColMajor 64 rows per wave
//   kind 3 (gemv, EXACT): RowMajor one row per wave; ColMajor 64 rows per wave
// (gemv has no work when m or n is 0: level2.hpp:21-22, :59-60)
__device__ __forceinline__ int64_t batch_items(const void *P, int kind, int64_t p, int tm, int tn, int colmajor) {
  if (kind == 0) {
    const qblas_gemm_problem &g = reinterpret_cast<const qblas_gemm_problem *>(P)[p];
    if (g.m <= 0 || g.n <= 0 || g.k <= 0) return 0;
    return cdiv(g.m, tm) * cdiv(g.n, tn);
  }
  const qblas_gemv_problem &v = reinterpret_cast<const qblas_gemv_problem *>(P)[p];
  if (v.m <= 0 || v.n <= 0) return 0;
  if (kind == 1) return cdiv(v.m, (!colmajor && v.incx == 1) ? 32 : 64);
  if (kind == 3) return colmajor ? cdiv(v.m, 64) : v.m;
  return colmajor ? cdiv(v.m, 64) : cdiv(v.m, 4);
}

// start[0] = 0, start[p + 1] = items of problems 0..p; one block, chunks of 1024 problems.  A
// non-null ctr (the chunk queue of the kernel that follows) is set to zero.
__global__ void __launch_bounds__(1024) k_batch_scan(const void *__restrict__ P, int kind, int64_t count, int tm,
                                                     int tn, int colmajor, int64_t *__restrict__ start,
                                                     uint32_t *__restrict__ ctr) {
  __shared__ int64_t s[1024];
  int64_t base = 0;
  if (threadIdx.x == 0) {
    start[0] = 0;
    if (ctr) *ctr = 0u;
  }
  for (int64_t b = 0; b < count; b += 1024) {
    const int64_t p = b + threadIdx.x;
    s[threadIdx.x] = (p < count) ? batch_items(P, kind, p, tm, tn, colmajor) : 0;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int64_t t = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (p < count) start[p + 1] = base + s[threadIdx.x];
    base += s[1023];
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// gemm, REFERENCE order: one lane per C entry, tile TM x (256 / TM), rows fastest
// ------------------------------------------------------------------------------------------
// one C entry of a REFERENCE batched gemm problem (level3.hpp:215-336 order)
__device__ __forceinline__ void ref_gemm_entry(const qblas_gemm_problem &g, int64_t i, int64_t j, int colmajor,
                                               int64_t kc_default) {
  const int64_t m = g.m, n = g.n, k = g.k;
  const int64_t kc = (m <= 64 && n <= 64 && k <= 64) ? k : kc_default;
  const q128 *A = (const q128 *)g.A;
  const q128 *B = (const q128 *)g.B;
  q128 *C = (q128 *)g.C;
  const int64_t as_i = colmajor ? 1 : g.lda, as_l = colmajor ? g.lda : 1;
  const int64_t bs_l = colmajor ? 1 : g.ldb, bs_j = colmajor ? g.ldb : 1;
  const q128 alpha = qv(g.alpha), beta = qv(g.beta);
  q128 *ci = C + (colmajor ? j * g.ldc + i : i * g.ldc + j);
  q128 c = bld(ci);
  const q128 *ap = A + i * as_i;
  const q128 *bp = B + j * bs_j;
  for (int64_t kk = 0; kk < k; kk += kc) {
    const int64_t ke = (kk + kc < k) ? kk + kc : k;
    chain ch;
    chain_init(ch);
    for (int64_t l = kk; l < ke; ++l) chain_fma(ch, bld(ap + l * as_l), bld(bp + l * bs_l));
    c = fma_packed(alpha, chain_get(ch), mul_packed(kk == 0 ? beta : q_one(), c));
  }
  bst(ci, c);
}

template <int TM>
__global__ void __launch_bounds__(256) k_gemm_batched_ref(const qblas_gemm_problem *__restrict__ P, int64_t count,
                                                          const int64_t *__restrict__ start, int colmajor,
                                                          int64_t kc_default, uint32_t *__restrict__ queue) {
  constexpr int TN = 256 / TM;
  __shared__ uint32_t s_ticket;
  const int64_t total = start[count];
  // tile blockIdx.x first, then tiles gridDim.x, gridDim.x + 1, ... from the queue (zeroed by the
  // scan): a tile's cost grows with its problem's k
  for (int64_t t = blockIdx.x; t < total;) {
    if (threadIdx.x == 0) s_ticket = atomicAdd(queue, 1u);
    const int64_t p = find_problem(start, count, t);
    const qblas_gemm_problem &g = P[p];
    const int64_t tl = t - start[p];
    const int64_t tiles_n = cdiv(g.n, TN);
    const int64_t ti = tl / tiles_n, tj = tl - ti * tiles_n;
    const int64_t i = ti * TM + (threadIdx.x % TM), j = tj * TN + (threadIdx.x / TM);
    if (i < g.m && j < g.n) ref_gemm_entry(g, i, j, colmajor, kc_default);
    __syncthreads();  // the ticket is written
    t = gridDim.x + (int64_t)s_ticket;
    __syncthreads();  // every thread has read it before the next tile writes it
  }
}

// ------------------------------------------------------------------------------------------
// gemm, DD: tile TM x TN, thread block of (TM/RM) x (TN/RN) = 256 threads, RM x RN entries per
// thread, k steps of 16 staged in LDS as double-doubles (converted on load)
// ------------------------------------------------------------------------------------------
__device__ __attribute__((noinline)) q128 bdd_out(double hi, double lo, q128 alpha, q128 beta, q128 c) {
  dd v;
  v.hi = hi;
  v.lo = lo;
  return fma_packed(alpha, dd_to_q(v), mul_packed(beta, c));
}

template <int TM, int TN, int RM, int RN>
__global__ void __launch_bounds__(256) k_gemm_batched_dd(const qblas_gemm_problem *__restrict__ P, int64_t count,
                                                         const int64_t *__restrict__ start, int colmajor,
                                                         uint32_t *__restrict__ queue) {
  constexpr int KS = 16, CT = TN / RN;
  static_assert((TM / RM) * (TN / RN) == 256, "256 threads");
  __shared__ double sAh[KS][TM], sAl[KS][TM], sBh[KS][TN], sBl[KS][TN];
  __shared__ uint32_t s_ticket;
  const int tr = threadIdx.x / CT, tc = threadIdx.x % CT;
  const int64_t total = start[count];
  // tile blockIdx.x first, then tiles gridDim.x, gridDim.x + 1, ... from the queue (zeroed by the
  // scan): a tile's cost grows with its problem's k
  for (int64_t t = blockIdx.x; t < total;) {
    if (threadIdx.x == 0) s_ticket = atomicAdd(queue, 1u);
    const int64_t p = find_problem(start, count, t);
    const qblas_gemm_problem &g = P[p];
    const int64_t m = g.m, n = g.n, k = g.k;
    const int64_t tl = t - start[p];
    const int64_t tiles_n = cdiv(n, TN);
    const int64_t i0 = (tl / tiles_n) * TM, j0 = (tl - (tl / tiles_n) * tiles_n) * TN;
    const q128 *A = (const q128 *)g.A;
    const q128 *B = (const q128 *)g.B;
    const int64_t as_i = colmajor ? 1 : g.lda, as_l = colmajor ? g.lda : 1;
    const int64_t bs_l = colmajor ? 1 : g.ldb, bs_j = colmajor ? g.ldb : 1;
    dd acc[RM][RN];
#pragma unroll
    for (int r = 0; r < RM; ++r)
#pragma unroll
      for (int c = 0; c < RN; ++c) acc[r][c] = dd_zero();
    for (int64_t k0 = 0; k0 < k; k0 += KS) {
      __syncthreads();
      // stage A[i0.., k0..] and B[k0.., j0..]; the contiguous index runs fastest across threads
      for (int e = threadIdx.x; e < TM * KS; e += 256) {
        const int r = colmajor ? e % TM : e / KS, kk = colmajor ? e / TM : e % KS;
        const int64_t i = i0 + r, l = k0 + kk;
        const dd v = (i < m && l < k) ? q_to_dd(bld(A + i * as_i + l * as_l)) : dd_zero();
        sAh[kk][r] = v.hi;
        sAl[kk][r] = v.lo;
      }
      for (int e = threadIdx.x; e < KS * TN; e += 256) {
        const int c = colmajor ? e / KS : e % TN, kk = colmajor ? e % KS : e / TN;
        const int64_t j = j0 + c, l = k0 + kk;
        const dd v = (j < n && l < k) ? q_to_dd(bld(B + l * bs_l + j * bs_j)) : dd_zero();
        sBh[kk][c] = v.hi;
        sBl[kk][c] = v.lo;
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < KS; ++kk) {
        dd a[RM], b[RN];
#pragma unroll
        for (int r = 0; r < RM; ++r) {
          a[r].hi = sAh[kk][tr + (TM / RM) * r];
          a[r].lo = sAl[kk][tr + (TM / RM) * r];
        }
#pragma unroll
        for (int c = 0; c < RN; ++c) {
          b[c].hi = sBh[kk][tc + CT * c];
          b[c].lo = sBl[kk][tc + CT * c];
        }
#pragma unroll
        for (int r = 0; r < RM; ++r)
#pragma unroll
          for (int c = 0; c < RN; ++c) acc[r][c] = dd_fma(a[r], b[c], acc[r][c]);
      }
    }
    const q128 alpha = qv(g.alpha), beta = qv(g.beta);
    q128 *C = (q128 *)g.C;
#pragma unroll
    for (int r = 0; r < RM; ++r)
#pragma unroll
      for (int c = 0; c < RN; ++c) {
        const int64_t i = i0 + tr + (TM / RM) * r, j = j0 + tc + CT * c;
        if (i < m && j < n) {
          q128 *ci = C + (colmajor ? j * g.ldc + i : i * g.ldc + j);
          bst(ci, bdd_out(acc[r][c].hi, acc[r][c].lo, alpha, beta, bld(ci)));
        }
      }
    __syncthreads();  // the ticket is written
    t = gridDim.x + (int64_t)s_ticket;
    __syncthreads();  // every thread has read it before the next tile writes it
  }
}

// ------------------------------------------------------------------------------------------
// gemv, REFERENCE order: units of 64 lanes (one wave) of one problem
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gemv_batched_ref(const qblas_gemv_problem *__restrict__ P, int64_t count,
                                                          const int64_t *__restrict__ start, int colmajor,
                                                          uint32_t *__restrict__ queue) {
  const int lane = threadIdx.x & 63;
  const int64_t total = start[count];
  const int64_t waves = (int64_t)gridDim.x * 4;
  // item = the wave index first, then items waves, waves + 1, ... from the queue (zeroed by the
  // scan): an item's cost grows with its problem's n
  for (int64_t u = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); u < total;) {
    uint32_t ticket = 0;
    if (lane == 0) ticket = atomicAdd(queue, 1u);
    const int64_t p = find_problem(start, count, u);
    const qblas_gemv_problem &v = P[p];
    const int64_t m = v.m, n = v.n, lda = v.lda, incx = v.incx, incy = v.incy;
    const int64_t ul = u - start[p];
    const q128 *A = (const q128 *)v.A;
    const q128 *x = (const q128 *)v.x;
    q128 *y = (q128 *)v.y;
    const q128 alpha = qv(v.alpha), beta = qv(v.beta);
    chain ch;
    if (!colmajor && incx == 1) {
      // two lane chains per row (level2.hpp:35-38 with dot_kernel_vectorized), rows ul*32 + lane/2
      const int64_t i = ul * 32 + (lane >> 1);
      const int par = lane & 1;
      const bool live = i < m;
      chain_init(ch);
      if (live) {
        const q128 *row = A + i * lda;
        const int64_t pairs = n / 2;
        for (int64_t jj = 0; jj < pairs; ++jj) chain_fma(ch, bld(row + 2 * jj + par), bld(x + 2 * jj + par));
      }
      const q128 mine = chain_get(ch);
      const uint64_t olo = (uint64_t)__shfl_xor((unsigned long long)mine.lo, 1);
      const uint64_t ohi = (uint64_t)__shfl_xor((unsigned long long)mine.hi, 1);
      if (live && par == 0) {
        q128 sum = add_packed(mine, q_make(olo, ohi));
        if (n & 1) sum = fma_packed(bld(A + i * lda + n - 1), bld(x + n - 1), sum);
        q128 *yi = y + i * incy;
        bst(yi, fma_packed(alpha, sum, mul_packed(beta, bld(yi))));
      }
    } else if (!colmajor) {
      const int64_t i = ul * 64 + lane;
      if (i < m) {
        chain_init(ch);
        const q128 *row = A + i * lda;
        for (int64_t j = 0; j < n; ++j) chain_fma(ch, bld(row + j), bld(x + j * incx));
        q128 *yi = y + i * incy;
        bst(yi, fma_packed(alpha, chain_get(ch), mul_packed(beta, bld(yi))));
      }
    } else {
      // level2.hpp:63-80: y_i = mul(beta, y_i), then fma(A_ij, mul(alpha, x_j), y_i) over j
      const int64_t i = ul * 64 + lane;
      if (i < m) {
        q128 *yi = y + i * incy;
        chain_set(ch, mul_packed(beta, bld(yi)));
        for (int64_t j = 0; j < n; ++j) chain_fma(ch, bld(A + j * lda + i), mul_packed(alpha, bld(x + j * incx)));
        bst(yi, chain_get(ch));
      }
    }
    u = waves + (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)ticket);  // lane 0's ticket
  }
}

// ------------------------------------------------------------------------------------------
// gemv, DD: RowMajor 4 rows per wave with the lanes along the columns (butterfly sums over the
// wave); ColMajor one lane per row over the columns (coalesced column reads).
// y = fma(alpha, RN(sum), mul(beta, y)).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ dd bshfl_xor_dd(dd v, int msk) {
  dd r;
  r.hi = __shfl_xor(v.hi, msk);
  r.lo = __shfl_xor(v.lo, msk);
  return r;
}

// Waves take chunks of consecutive work items in order (the first by wave index, later ones from
// the queue), so that the waves running at one time work on neighbouring memory;
// within a chunk the problems are stepped through linearly (one search per chunk).
// RowMajor: an item is 4 rows; lanes stride along the columns, 2 column steps of 64 per iteration
// (8 loads of A in flight per lane, each converted x element serving the 4 rows), dd partials per
// lane, one joint butterfly for the 4 rows; a wave takes 16 items and then finishes their 64 rows,
// one per lane.
// ColMajor: an item is 64 rows, one per lane, over the columns (2 accumulators).
constexpr int BGV_CHUNK = 4;    // ColMajor: work items a wave takes in a row
constexpr int BGV_CHUNK_R = 16;  // RowMajor: 16 items of 4 rows, one row per lane at the end
template <int colmajor>
__global__ void __launch_bounds__(256) k_gemv_batched_dd(const qblas_gemv_problem *__restrict__ P, int64_t count,
                                                         const int64_t *__restrict__ start,
                                                         uint32_t *__restrict__ queue) {
  const int lane = threadIdx.x & 63;
  const int64_t total = start[count];
  // the wave index as a scalar: item and problem searches run in scalar registers
  const int64_t waves = (int64_t)gridDim.x * 4,
                w = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (!colmajor) {
    // RowMajor: lanes stride along the columns of an item's 4 rows; one joint butterfly leaves
    // row r's total in lanes 16 r .. 16 r + 15, lane 4 q + r of the wave keeps the total of row r
    // of the chunk's item q (its y loaded when the item starts), and after the 16 items each lane
    // finishes its row: the binary128 epilogue runs once per 64 rows instead of once per 4.
    // chunk w first, then chunks waves, waves + 1, ... in the order the waves ask for them
    // (queue word zeroed by the scan): the items' costs differ up to 32x with n, and a fixed
    // round-robin share left the kernel waiting on its slowest waves
    for (int64_t cw = w; cw * BGV_CHUNK_R < total;) {
      const int64_t c0 = cw * BGV_CHUNK_R;
      const int64_t c1 = (c0 + BGV_CHUNK_R < total) ? c0 + BGV_CHUNK_R : total;
      uint32_t ticket = 0;
      if (lane == 0) ticket = atomicAdd(queue, 1u);
      int64_t p = find_problem(start, count, c0);
      dd mine = dd_zero();
      q128 yv = q_zero(0);
      int64_t pl = -1;  // problem of this lane's row (-1: none)
      int64_t il = 0;  // its row index
      for (int64_t u = c0; u < c1; ++u) {
        while (start[p + 1] <= u) ++p;  // the problem holding item u (empty problems are skipped)
        const qblas_gemv_problem &v = P[p];
        const int64_t m = v.m, n = v.n, lda = v.lda, incx = v.incx, incy = v.incy;
        const int64_t i0 = (u - start[p]) * 4;
        const int R = (m - i0 < 4) ? (int)(m - i0) : 4;  // rows of this item (wave-uniform)
        const q128 *A = (const q128 *)v.A;
        const q128 *x = (const q128 *)v.x;
        const int q = (int)(u - c0);
        if ((lane >> 2) == q && (lane & 3) < R) {  // this lane will finish row i0 + (lane & 3)
          pl = p;
          il = i0 + (lane & 3);
          yv = bld((const q128 *)v.y + il * incy);
        }
        dd acc[4][2];
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r][0] = acc[r][1] = dd_zero();
        // the next column step's A and x are loaded before the current step is summed, so that
        // each lane keeps two steps of loads in flight (every load index stays below n)
        q128 a0[4], a1[4], xq0, xq1;
        auto load_step = [&](int64_t j, q128(&b0)[4], q128(&b1)[4], q128 &xb0, q128 &xb1) {
          const bool h1 = j + 64 < n;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (r < R) {
              const q128 *row = A + (i0 + r) * lda;
              b0[r] = bld(row + j);
              b1[r] = h1 ? bld(row + j + 64) : q_zero(0);
            }
          }
          xb0 = bld(x + j * incx);
          xb1 = h1 ? bld(x + (j + 64) * incx) : q_zero(0);
        };
        if (lane < n) load_step(lane, a0, a1, xq0, xq1);
        for (int64_t j = lane; j < n; j += 128) {
          q128 b0[4], b1[4], xb0, xb1;
          if (j + 128 < n) load_step(j + 128, b0, b1, xb0, xb1);
          const bool h1 = j + 64 < n;
          const dd x0 = q_to_dd(xq0);
          const dd x1 = h1 ? q_to_dd(xq1) : dd_zero();
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (r < R) {
              acc[r][0] = dd_fma(q_to_dd(a0[r]), x0, acc[r][0]);
              acc[r][1] = dd_fma(q_to_dd(a1[r]), x1, acc[r][1]);
            }
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            a0[r] = b0[r];
            a1[r] = b1[r];
          }
          xq0 = xb0;
          xq1 = xb1;
        }
        dd s[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) s[r] = (r < R) ? dd_add(acc[r][0], acc[r][1]) : dd_zero();
        // mask 32: lanes with bit 5 = b keep rows 2b, 2b + 1 and send the other two
        const bool b = (lane & 32) != 0, c = (lane & 16) != 0;
        const dd t0 = dd_add(b ? s[2] : s[0], bshfl_xor_dd(b ? s[0] : s[2], 32));
        const dd t1 = dd_add(b ? s[3] : s[1], bshfl_xor_dd(b ? s[1] : s[3], 32));
        // mask 16: bit 4 = c keeps row 2b + c
        dd tot = dd_add(c ? t1 : t0, bshfl_xor_dd(c ? t0 : t1, 16));
#pragma unroll
        for (int msk = 8; msk >= 1; msk >>= 1) tot = dd_add(tot, bshfl_xor_dd(tot, msk));
        const int src = 16 * (lane & 3);
        dd got;
        got.hi = __shfl(tot.hi, src);
        got.lo = __shfl(tot.lo, src);
        if ((lane >> 2) == q) mine = got;
      }
      if (pl >= 0) {
        const qblas_gemv_problem &v = P[pl];
        bst((q128 *)v.y + il * v.incy, fma_packed(qv(v.alpha), dd_to_q(mine), mul_packed(qv(v.beta), yv)));
      }
      cw = waves + (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)ticket);  // lane 0's ticket
    }
    return;
  }
  for (int64_t cw = w; cw * BGV_CHUNK < total;) {  // chunk w, then chunks from the queue
    const int64_t c0 = cw * BGV_CHUNK;
    const int64_t c1 = (c0 + BGV_CHUNK < total) ? c0 + BGV_CHUNK : total;
    uint32_t ticket = 0;
    if (lane == 0) ticket = atomicAdd(queue, 1u);
    int64_t p = find_problem(start, count, c0);
    for (int64_t u = c0; u < c1; ++u) {
      while (start[p + 1] <= u) ++p;  // the problem holding item u (empty problems are skipped)
      const qblas_gemv_problem &v = P[p];
      const int64_t m = v.m, n = v.n, lda = v.lda, incx = v.incx, incy = v.incy;
      const int64_t ul = u - start[p];
      const q128 *A = (const q128 *)v.A;
      const q128 *x = (const q128 *)v.x;
      q128 *y = (q128 *)v.y;
      const q128 alpha = qv(v.alpha), beta = qv(v.beta);
      {
        const int64_t i = ul * 64 + lane;
        if (i < m) {
          dd acc0 = dd_zero(), acc1 = dd_zero();
          int64_t j = 0;
          for (; j + 1 < n; j += 2) {
            acc0 = dd_fma(q_to_dd(bld(A + j * lda + i)), q_to_dd(bld(x + j * incx)), acc0);
            acc1 = dd_fma(q_to_dd(bld(A + (j + 1) * lda + i)), q_to_dd(bld(x + (j + 1) * incx)), acc1);
          }
          if (j < n) acc0 = dd_fma(q_to_dd(bld(A + j * lda + i)), q_to_dd(bld(x + j * incx)), acc0);
          q128 *yi = y + i * incy;
          bst(yi, fma_packed(alpha, dd_to_q(dd_add(acc0, acc1)), mul_packed(beta, bld(yi))));
        }
      }
    }
    cw = waves + (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)ticket);  // lane 0's ticket
  }
}

// ------------------------------------------------------------------------------------------
// gemv, EXACT: RowMajor one wave per row (wave-uniform anchor from the row's first products);
// ColMajor one lane per row over the columns (anchor from the row's first regular product)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ q128 exact_row_out(const uint32_t (&Wl)[WW], const uint32_t (&Wh)[WW], int An,
                                              uint32_t flags, q128 alpha, q128 beta, q128 yv) {
  uint32_t Mw[SW];
  uint32_t sign;
  s_from_windows(Mw, sign, Wl, Wh);
  return finish_reg(Mw, An, sign, 0u, flags, true, alpha, beta, yv);
}

// status[1] != 0 after this kernel: some row was left for the full-range path (its y unwritten)
__global__ void __launch_bounds__(256) k_gemv_batched_exact(const qblas_gemv_problem *__restrict__ P, int64_t count,
                                                            const int64_t *__restrict__ start, int colmajor,
                                                            uint32_t *__restrict__ status,
                                                            uint32_t *__restrict__ queue) {
  const int lane = threadIdx.x & 63;
  const int64_t total = start[count];
  const int64_t waves = (int64_t)gridDim.x * 4;
  // chunks of XB_CHUNK items: chunk = the wave index first, then chunks waves, waves + 1, ... from
  // the queue (zeroed by the scan); an item's cost grows with its problem's n.  One ticket per
  // item (a row, RowMajor) made the queue word's atomics the bound.
  constexpr int64_t XB_CHUNK = 16;
  for (int64_t cw = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
       cw * XB_CHUNK < total;) {
    uint32_t ticket = 0;
    if (lane == 0) ticket = atomicAdd(queue, 1u);
    const int64_t cend = ((cw + 1) * XB_CHUNK < total) ? (cw + 1) * XB_CHUNK : total;
    for (int64_t u = cw * XB_CHUNK; u < cend; ++u) {
      const int64_t p = find_problem(start, count, u);
      const qblas_gemv_problem &v = P[p];
      const int64_t m = v.m, n = v.n, lda = v.lda, incx = v.incx, incy = v.incy;
      const int64_t ul = u - start[p];
      const q128 *A = (const q128 *)v.A;
      const q128 *x = (const q128 *)v.x;
      q128 *y = (q128 *)v.y;
      const q128 alpha = qv(v.alpha), beta = qv(v.beta);
      uint32_t Wl[WW], Wh[WW];
  #pragma unroll
      for (int k = 0; k < WW; ++k) {
        Wl[k] = 0;
        Wh[k] = 0;
      }
      uint32_t flags = 0;
      if (!colmajor) {
        const int64_t i = ul;
        const q128 *row = A + i * lda;
        int e0 = EXP_NONE;
        if (lane < n) e0 = pair_exp(ldq_c(row + lane), ldq_c(x + (int64_t)lane * incx));
        const int An = wave_anchor(e0);
        for (int64_t j = lane; j < n; j += 64) accumulate(Wl, Wh, An, ldq_c(row + j), ldq_c(x + j * incx), flags);
        wave_sum_win(Wl);
        wave_sum_win(Wh);
        flags = wave_or(flags);
        if (lane == 0) {
          q128 *yi = y + i * incy;
          if (flags & FL_LOST) {
            atomicOr(status + 1, 1u);
          } else {
            stq_c(yi, exact_row_out(Wl, Wh, An, flags, alpha, beta, ldq_c(yi)));
          }
        }
      } else {
        const int64_t i = ul * 64 + lane;
        if (i < m) {
          int e0 = EXP_NONE;
          for (int64_t j = 0; j < n && j < 4 && e0 == EXP_NONE; ++j) e0 = pair_exp(ldq_c(A + j * lda + i), ldq_c(x + j * incx));
          const int An = (e0 == EXP_NONE) ? 0 : e0 - WIN2_CENTER;
          for (int64_t j = 0; j < n; ++j) accumulate(Wl, Wh, An, ldq_c(A + j * lda + i), ldq_c(x + j * incx), flags);
          q128 *yi = y + i * incy;
          if (flags & FL_LOST) {
            atomicOr(status + 1, 1u);
          } else {
            stq_c(yi, exact_row_out(Wl, Wh, An, flags, alpha, beta, ldq_c(yi)));
          }
        }
      }
    }
    cw = waves + (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)ticket);  // lane 0's ticket
  }
}

// whether a product of the pair misses the two windows of anchor An (the fast path's test)
__device__ __forceinline__ bool pair_misses(q128 a, q128 b, int An) {
  uq x, y;
  uint32_t fl = 0;
  const int c = classify_pair(a, b, x, y, fl);
  if (c != 0 && c != 3) return false;
  const int o = prod_lsb(x, y) - An;
  return o < WIN_OMIN || o > WIN_BAND + WIN_OMAX;
}

// After k_gemv_batched_exact: when it left rows (status[1]), the same walk over the work items
// finds them again (same anchors, same window test), four waves' rows at a time into an LDS list,
// and the workgroup computes each on the full-range path.  Returns at once otherwise.
__global__ void __launch_bounds__(256) k_gemv_batched_exact_wide(const qblas_gemv_problem *__restrict__ P,
                                                                 int64_t count, const int64_t *__restrict__ start,
                                                                 int colmajor, uint32_t *__restrict__ status) {
  __shared__ kul_t K;
  __shared__ int64_t s_p[256], s_i[256];
  __shared__ int s_n;
  if (status[1] == 0) return;
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t total = start[count];
  kul_clear_all(K);
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (int64_t g = blockIdx.x; g * 4 < total; g += gridDim.x) {  // uniform over the workgroup
    const int64_t u = g * 4 + wv;
    if (u < total) {
      const int64_t p = find_problem(start, count, u);
      const qblas_gemv_problem &v = P[p];
      const int64_t m = v.m, n = v.n, lda = v.lda, incx = v.incx;
      const int64_t ul = u - start[p];
      const q128 *A = (const q128 *)v.A;
      const q128 *x = (const q128 *)v.x;
      if (!colmajor) {
        const q128 *row = A + ul * lda;
        int e0 = EXP_NONE;
        if (lane < n) e0 = pair_exp(ldq_c(row + lane), ldq_c(x + (int64_t)lane * incx));
        const int An = wave_anchor(e0);
        bool miss = false;
        for (int64_t j = lane; j < n; j += 64) miss |= pair_misses(ldq_c(row + j), ldq_c(x + j * incx), An);
        if (__any(miss) && lane == 0) {
          const int q = atomicAdd(&s_n, 1);
          s_p[q] = p;
          s_i[q] = ul;
        }
      } else {
        const int64_t i = ul * 64 + lane;
        if (i < m) {
          int e0 = EXP_NONE;
          for (int64_t j = 0; j < n && j < 4 && e0 == EXP_NONE; ++j) e0 = pair_exp(ldq_c(A + j * lda + i), ldq_c(x + j * incx));
          const int An = (e0 == EXP_NONE) ? 0 : e0 - WIN2_CENTER;
          bool miss = false;
          for (int64_t j = 0; j < n && !miss; ++j) miss = pair_misses(ldq_c(A + j * lda + i), ldq_c(x + j * incx), An);
          if (miss) {
            const int q = atomicAdd(&s_n, 1);
            s_p[q] = p;
            s_i[q] = i;
          }
        }
      }
    }
    __syncthreads();
    const int nl = s_n;
    for (int q = 0; q < nl; ++q) {
      const qblas_gemv_problem &v = P[s_p[q]];
      const int64_t i = s_i[q];
      const q128 *A = (const q128 *)v.A;
      q128 *yi = (q128 *)v.y + i * v.incy;
      const q128 r = wide_item(K, v.n, colmajor ? A + i : A + i * v.lda, colmajor ? v.lda : 1, (const q128 *)v.x,
                               v.incx, true, qv(v.alpha), qv(v.beta), ldq_c(yi), status);
      if (threadIdx.x == 0) {
        stq_c(yi, r);
        atomicAdd(status + 2, 1u);  // rows on the full-range path
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
constexpr int BATCH_GRID = 2048;  // persistent workgroups (8 per CU)

struct gemm_cfg {
  int tm, tn;
  const char *name;
};
static const gemm_cfg REF_CFG[] = {{16, 16, "16x16 tile, 1 entry/lane"},
                                   {8, 32, "8x32 tile, 1 entry/lane"},
                                   {32, 8, "32x8 tile, 1 entry/lane"},
                                   {4, 64, "4x64 tile, 1 entry/lane"},
                                   {64, 4, "64x4 tile, 1 entry/lane"}};
static const gemm_cfg DD_CFG[] = {{32, 32, "32x32 tile, 2x2 entries/thread"},
                                  {64, 64, "64x64 tile, 4x4 entries/thread"},
                                  {16, 16, "16x16 tile, 1 entry/thread"},
                                  {64, 32, "64x32 tile, 4x2 entries/thread"}};

int gemm_batched_configs(int dd_mode) { return dd_mode ? 4 : 5; }
const char *gemm_batched_config_name(int dd_mode, int idx) {
  if (idx < 0 || idx >= gemm_batched_configs(dd_mode)) return nullptr;
  return dd_mode ? DD_CFG[idx].name : REF_CFG[idx].name;
}

// prefix sums (count + 1) and, 256 bytes further on, the status word of the EXACT form
size_t batched_ws(int64_t count) { return ((size_t)(count + 1) * sizeof(int64_t) + 255) / 256 * 256 + 256; }
size_t batched_status_offset(int64_t count) { return ((size_t)(count + 1) * sizeof(int64_t) + 255) / 256 * 256; }

__global__ void k_batch_status_zero(uint32_t *s) {
  if (threadIdx.x < 4) s[threadIdx.x] = 0;
}

hipError_t launch_gemm_batched(int dd_mode, int colmajor, int64_t count, const qblas_gemm_problem *P, int cfg,
                               int64_t kc, void *ws, hipStream_t st) {
  if (cfg < 0 || cfg >= gemm_batched_configs(dd_mode)) return hipErrorInvalidValue;
  int64_t *start = (int64_t *)ws;
  const gemm_cfg c = dd_mode ? DD_CFG[cfg] : REF_CFG[cfg];
  // the kernels' tile queue: a word of the status block, past the status words
  uint32_t *queue = (uint32_t *)((char *)ws + batched_status_offset(count)) + 32;
  k_batch_scan<<<1, 1024, 0, st>>>(P, 0, count, c.tm, c.tn, colmajor, start, queue);
  if (!dd_mode) {
    switch (cfg) {
      case 0: k_gemm_batched_ref<16><<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, kc, queue); break;
      case 1: k_gemm_batched_ref<8><<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, kc, queue); break;
      case 2: k_gemm_batched_ref<32><<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, kc, queue); break;
      case 3: k_gemm_batched_ref<4><<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, kc, queue); break;
      default: k_gemm_batched_ref<64><<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, kc, queue); break;
    }
  } else {
    switch (cfg) {
      case 0: k_gemm_batched_dd<32, 32, 2, 2><<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, queue); break;
      case 1: k_gemm_batched_dd<64, 64, 4, 4><<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, queue); break;
      case 2: k_gemm_batched_dd<16, 16, 1, 1><<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, queue); break;
      default: k_gemm_batched_dd<64, 32, 4, 2><<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, queue); break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_gemv_batched(int form, int colmajor, int64_t count, const qblas_gemv_problem *P, void *ws,
                               hipStream_t st) {
  int64_t *start = (int64_t *)ws;
  const int kind = form == 1 ? 2 : (form == 2 ? 3 : 1);
  // the kernels' item queue: a word of the status block, past the status words
  uint32_t *queue = (uint32_t *)((char *)ws + batched_status_offset(count)) + 32;
  k_batch_scan<<<1, 1024, 0, st>>>(P, kind, count, 0, 0, colmajor, start, queue);
  if (form == 0) {
    k_gemv_batched_ref<<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, queue);
  } else if (form == 1) {
    if (colmajor)
      k_gemv_batched_dd<1><<<BATCH_GRID, 256, 0, st>>>(P, count, start, queue);
    else
      k_gemv_batched_dd<0><<<BATCH_GRID, 256, 0, st>>>(P, count, start, queue);
  } else {
    uint32_t *status = (uint32_t *)((char *)ws + batched_status_offset(count));
    k_batch_status_zero<<<1, 64, 0, st>>>(status);
    k_gemv_batched_exact<<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, status, queue);
    k_gemv_batched_exact_wide<<<BATCH_GRID, 256, 0, st>>>(P, count, start, colmajor, status);
  }
  return hipGetLastError();
}

}  // namespace qbk
