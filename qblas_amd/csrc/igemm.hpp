// igemm.hpp -- int8 x int8 -> int32 GEMM on gfx950 MFMA (v_mfma_i32_32x32x32_i8), "NT" layout:
//   C[i][j] (int32, row-major, ldc) = sum_k A[i][k] * B[j][k]
// with A [M x K] and B [N x K] both K-contiguous (lda, ldb in bytes).  This is the plain integer
// GEMM inside the Ozaki form of qgemm: every slice-pair product is one of these.
//
// Tiling: a 256-thread workgroup computes a 128 x 128 tile of C; each of its 4 waves a 64 x 64
// quarter as 2 x 2 blocks of 32 x 32 MFMA results (16 int32 per lane each).  K advances in steps of
// 64 bytes through double-buffered LDS tiles (A and B: 128 rows x 64 bytes each, rows padded to
// 80 bytes against bank conflicts).  MFMA operand layout (checked on the device): lane l holds row
// (l % 32) of the block and 16 consecutive k at offset (l / 32) * 16 of each 32-k step (the k order
// only has to agree between A and B); result element r of lane l is C[8*(r/4) + 4*(l/32) + r%4]
// [l % 32].
// Requirements (checked by the launcher): M, N multiples of 128, K a multiple of 64, all
// addresses 16-byte aligned.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qbk {

constexpr int IG_TM = 128, IG_TN = 128, IG_TK = 64, IG_PAD = 80;  // LDS row pitch in bytes

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

// batch z of the launch: A + z*sA, B + z*sB, C + z*sC, K = K0 + z*dK (bytes / elements)
struct igemm_args {
  const int8_t *A;
  const int8_t *B;
  int32_t *C;
  int64_t lda, ldb, ldc;
  int64_t sA, sB, sC;  // per-batch offsets
  int M, N, K0, dK;
};

__global__ void __launch_bounds__(256) k_igemm_nt(igemm_args a) {
  __shared__ __attribute__((aligned(16))) int8_t sa[2][IG_TM * IG_PAD];
  __shared__ __attribute__((aligned(16))) int8_t sb[2][IG_TN * IG_PAD];
  const int z = blockIdx.z;
  const int K = a.K0 + z * a.dK;
  const int8_t *A = a.A + z * a.sA + (int64_t)blockIdx.y * IG_TM * a.lda;
  const int8_t *B = a.B + z * a.sB + (int64_t)blockIdx.x * IG_TN * a.ldb;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;  // wave's quarter of the tile
  // global -> LDS: 128 rows x 64 bytes per operand = 512 chunks of 16 bytes, 2 per thread
  const int r0 = tid >> 2, c0 = (tid & 3) * 16;  // chunk tid and tid + 256 (rows r0, r0 + 64)
  i32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0;

  i32x4 ga0, ga1, gb0, gb1;
  auto gload = [&](int k0) {
    ga0 = *reinterpret_cast<const i32x4 *>(A + (int64_t)r0 * a.lda + k0 + c0);
    ga1 = *reinterpret_cast<const i32x4 *>(A + (int64_t)(r0 + 64) * a.lda + k0 + c0);
    gb0 = *reinterpret_cast<const i32x4 *>(B + (int64_t)r0 * a.ldb + k0 + c0);
    gb1 = *reinterpret_cast<const i32x4 *>(B + (int64_t)(r0 + 64) * a.ldb + k0 + c0);
  };
  auto sstore = [&](int buf) {
    *reinterpret_cast<i32x4 *>(&sa[buf][r0 * IG_PAD + c0]) = ga0;
    *reinterpret_cast<i32x4 *>(&sa[buf][(r0 + 64) * IG_PAD + c0]) = ga1;
    *reinterpret_cast<i32x4 *>(&sb[buf][r0 * IG_PAD + c0]) = gb0;
    *reinterpret_cast<i32x4 *>(&sb[buf][(r0 + 64) * IG_PAD + c0]) = gb1;
  };
  const int nk = K / IG_TK;
  gload(0);
  sstore(0);
  __syncthreads();
  const int lr = lane & 31, lk = (lane >> 5) * 16;
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) gload((kt + 1) * IG_TK);  // next tile in flight during the MFMAs
#pragma unroll
    for (int ks = 0; ks < IG_TK; ks += 32) {
      i32x4 fa[2], fb[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        fa[i] = *reinterpret_cast<const i32x4 *>(&sa[buf][(wm + i * 32 + lr) * IG_PAD + ks + lk]);
        fb[i] = *reinterpret_cast<const i32x4 *>(&sb[buf][(wn + i * 32 + lr) * IG_PAD + ks + lk]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) {
      sstore(buf ^ 1);
    }
    __syncthreads();
  }
  // epilogue: C rows 8*(r/4) + 4*(lane/32) + r%4 of each block, column lane%32
  int32_t *C = a.C + z * a.sC + (int64_t)blockIdx.y * IG_TM * a.ldc + (int64_t)blockIdx.x * IG_TN;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int col = wn + j * 32 + (lane & 31);
        C[(int64_t)row * a.ldc + col] = acc[i][j][r];
      }
}

}  // namespace qbk
