// kernels_ozaki.hip -- OZAKI form of qgemm: binary128 GEMM on the int8 matrix cores.
//
// Computes C = alpha * A B + beta * C for binary128 operands (the routine of
// include/quadblas/algorithms/level3.hpp:215-336, RowMajor and ColMajor with layout-correct
// indexing), with every output within 2^-104 relative error of the exact value:
//
// 1. Scaling.  Row i of A gets the exponent E_i = max_l ilogb(A_il) + 3 and column j of B the
//    exponent F_j = max_l ilogb(B_lj) + 3, so a_il = A_il 2^-E_i and b_lj = B_lj 2^-F_j lie in
//    (-1/4, 1/4).
// 2. Slicing.  a = sum_{t<S} d_t 256^-(t+1) + r with |r| <= 2^(-8S-1) and balanced digits
//    d_t in [-128, 127]: N = round(a 2^(8S)) (|N| < 2^(8S-2)), then the bytes of N + 0x8080..80
//    XOR 0x80 are the digits (the bias absorbs the borrows of the balanced representation).
//    Digits are stored in MFMA fragment order: for a block of 32 rows and 32 k, 1 KB in which
//    lane l's 16 bytes are row l%32, k 16*(l/32)..+15 -- the operand layout of
//    v_mfma_i32_32x32x32_i8.  Layout [row block][k block][slice][1 KB].
// 3. Products.  P = sum_{t+u<=S-1} 2^(-8(t+u+2)) (D_t^A (D_u^B)^T) exactly: one kernel keeps the
//    S diagonal sums acc_d = sum_{t+u=d} D_t D_u^T of its 32x32 tile in int32 registers (no
//    overflow: (d+1) * 4096 * 2^14 < 2^31 for S <= 31 and k passes of 4096), and writes
//    V = sum_d acc_d 256^(S-1-d) (exact, <= 192 bits) per k pass.
// 4. Finish.  The k passes are summed exactly, alpha*V*2^(E_i+F_j-8(S+1)) + beta*C is rounded
//    once to binary128 (the same register epilogue as the EXACT form), and the entry is
//    certified by the a-priori bound
//        |AB - P 2^(E_i+F_j)| <= K (S+1)/4 2^(E_i+F_j-8S)
//    (representation error of both operands plus the dropped diagonals t+u >= S): an entry is
//    written only if alpha times that bound is <= 2^-106 |result|.
// 5. Repair.  Every other entry (cancellation below the bound, zero sums, Inf/NaN operands) is
//    queued and recomputed by the EXACT form's exact accumulation (correctly rounded).
// So the result of every entry has relative error <= 2^-105.9 or is correctly rounded.
#include "exact_dev.hpp"
#include "kernels.h"

namespace qbk {

typedef int i32x4v __attribute__((ext_vector_type(4)));
typedef int i32x16v __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) void glb_void;

constexpr int OZ_TILE = 64;             // output rows and columns of one workgroup
constexpr int OZ_FRAG = 1024;           // bytes of one operand fragment (32 rows x 32 k)
constexpr int OZ_KPASS_BLOCKS = 128;    // k blocks (of 32) per int32 accumulation pass
constexpr int OZ_PW = 6;                // signed 32-bit words of a pass sum V
constexpr int OZ_PREC = 8;              // words per pass-sum record (32 bytes)
constexpr int OZ_FINW = 8;              // words of |sum of passes| in the finish
constexpr int32_t OZ_EZERO = INT32_MIN; // row/column exponent of an all-zero row/column
constexpr uint32_t OZF_NONFINITE = 1u;
constexpr int OZ_REPAIR_BLOCKS = 1024;
constexpr int OZ_HDR = 256;             // workspace header bytes: status, repair count

struct oz_layout {
  int64_t Mp, Np, Kb, npass;
  size_t eA, fA, eB, fB, sA, sB, part, list, total;
};

static oz_layout oz_plan(int64_t m, int64_t n, int64_t k, int S) {
  oz_layout L;
  L.Mp = (m + OZ_TILE - 1) / OZ_TILE * OZ_TILE;
  L.Np = (n + OZ_TILE - 1) / OZ_TILE * OZ_TILE;
  L.Kb = (k + 31) / 32;
  L.npass = (L.Kb + OZ_KPASS_BLOCKS - 1) / OZ_KPASS_BLOCKS;
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  size_t o = OZ_HDR;
  L.eA = o;
  o = al(o + (size_t)L.Mp * 4);
  L.fA = o;
  o = al(o + (size_t)L.Mp * 4);
  L.eB = o;
  o = al(o + (size_t)L.Np * 4);
  L.fB = o;
  o = al(o + (size_t)L.Np * 4);
  L.sA = o;
  o = al(o + (size_t)(L.Mp / 32) * L.Kb * S * OZ_FRAG);
  L.sB = o;
  o = al(o + (size_t)(L.Np / 32) * L.Kb * S * OZ_FRAG);
  L.part = o;
  o = al(o + (size_t)L.npass * L.Mp * L.Np * OZ_PREC * 4);
  L.list = o;
  o = al(o + (size_t)m * n * 8);
  L.total = o;
  return L;
}

// ------------------------------------------------------------------------------------------
// 1. scaling exponents
// ------------------------------------------------------------------------------------------
__global__ void k_oz_init(int32_t *eA, uint32_t *fA, int64_t Mp, int32_t *eB, uint32_t *fB, int64_t Np,
                          uint32_t *hdr) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < Mp) {
    eA[g] = OZ_EZERO;
    fA[g] = 0;
  }
  if (g < Np) {
    eB[g] = OZ_EZERO;
    fB[g] = 0;
  }
  if (g < OZ_HDR / 4) hdr[g] = 0;
}

__device__ __forceinline__ int oz_ilogb(q128 x, uint32_t &fl) {
  const uint32_t be = q_bexp(x);
  if (be == 0x7fffu) {
    fl |= OZF_NONFINITE;
    return OZ_EZERO;
  }
  if (q_iszero(x)) return OZ_EZERO;
  return unpack_finite(x).e;
}

// max ilogb over c of X[r*sr + c*sc] for r < R (atomic max into E[r]), non-finite flag into F[r].
// contig: sc == 1, one wave per r with lanes along c; otherwise one thread per r.  blockIdx.y
// selects a chunk of c.
__global__ void __launch_bounds__(256) k_oz_scale(const q128 *__restrict__ X, int64_t R, int64_t C, int64_t sr,
                                                  int64_t sc, int contig, int64_t cchunk, int32_t *E, uint32_t *F) {
  const int64_t c0 = (int64_t)blockIdx.y * cchunk;
  const int64_t c1 = (c0 + cchunk < C) ? c0 + cchunk : C;
  int e = OZ_EZERO;
  uint32_t fl = 0;
  if (contig) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= R) return;  // whole waves
    const q128 *row = X + r * sr;
    for (int64_t c = c0 + lane; c < c1; c += 64) e = max(e, oz_ilogb(ldq_c(row + c), fl));
    e = wave_max(e);
    fl = wave_or(fl);
    if (lane == 0) {
      if (e != OZ_EZERO) atomicMax(E + r, e);
      if (fl) atomicOr(F + r, fl);
    }
  } else {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= R) return;
    for (int64_t c = c0; c < c1; ++c) e = max(e, oz_ilogb(ldq_c(X + r * sr + c * sc), fl));
    if (e != OZ_EZERO) atomicMax(E + r, e);
    if (fl) atomicOr(F + r, fl);
  }
}

// ------------------------------------------------------------------------------------------
// 2. slicing
// ------------------------------------------------------------------------------------------
// W (6 words) = N + bias where N = round(x 2^(8S - Es)) (magnitudes rounded half up, then signed)
template <int S>
__device__ __forceinline__ void oz_digits(q128 x, int Es, uint32_t (&W)[OZ_PW]) {
  uint32_t Nw[OZ_PW];
#pragma unroll
  for (int i = 0; i < OZ_PW; ++i) Nw[i] = 0;
  const uint32_t be = q_bexp(x);
  if (be != 0x7fffu && !q_iszero(x)) {
    const uq u = unpack_finite(x);
    const int sh = u.e - 112 - Es + 8 * S;  // N = m * 2^sh;  u.e <= Es - 3 so m 2^sh < 2^(8S-2)
    uint32_t m[OZ_PW] = {(uint32_t)u.ml, (uint32_t)(u.ml >> 32), (uint32_t)u.mh, (uint32_t)(u.mh >> 32), 0u, 0u};
    if (sh >= 0) {
      rshl<OZ_PW>(Nw, m, sh);
    } else if (sh > -115) {
      // add 2^(-sh-1), then shift right: round half up of the magnitude
      const int p = -sh - 1;
      uint32_t c = 0;
#pragma unroll
      for (int i = 0; i < OZ_PW; ++i) m[i] = addc32(m[i], (i == (p >> 5)) ? (1u << (p & 31)) : 0u, c, c);
      rshr<OZ_PW>(Nw, m, -sh);
    }
    if (u.s) {
      uint32_t c = 1;
#pragma unroll
      for (int i = 0; i < OZ_PW; ++i) Nw[i] = addc32(~Nw[i], 0u, c, c);
    }
  }
  // + 0x80 in each of the S low bytes
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < OZ_PW; ++i) {
    const int nb = S - 4 * i;  // bias bytes in this word
    const uint32_t bias = nb >= 4 ? 0x80808080u : (nb == 3 ? 0x808080u : (nb == 2 ? 0x8080u : (nb == 1 ? 0x80u : 0u)));
    W[i] = addc32(Nw[i], bias, c, c);
  }
}

// One thread per (row block, k block, lane): 16 consecutive k of row 32*rb + lane%32.
// X(r, c) = X[r*sr + c*sc] for r < R, c < C; zero digits elsewhere.
template <int S>
__global__ void __launch_bounds__(256) k_oz_slice(const q128 *__restrict__ X, int64_t R, int64_t C, int64_t sr,
                                                  int64_t sc, const int32_t *__restrict__ E, int8_t *__restrict__ out,
                                                  int64_t Kb, int64_t nthreads) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= nthreads) return;
  const int lane = (int)(g & 63);
  const int64_t chunk = g >> 6;  // rb * Kb + kb
  const int64_t rb = chunk / Kb, kb = chunk - rb * Kb;
  const int64_t r = rb * 32 + (lane & 31);
  const int64_t c0 = kb * 32 + (lane >> 5) * 16;
  uint32_t D[S][4];
#pragma unroll
  for (int t = 0; t < S; ++t)
#pragma unroll
    for (int q = 0; q < 4; ++q) D[t][q] = 0;
  const int Er = (r < R) ? E[r] : OZ_EZERO;
  if (Er != OZ_EZERO) {
    const int Es = Er + 3;
    const q128 *row = X + r * sr;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      if (c0 + e < C) {
        uint32_t W[OZ_PW];
        oz_digits<S>(ldq_c(row + (c0 + e) * sc), Es, W);
#pragma unroll
        for (int t = 0; t < S; ++t) {
          const int byte = S - 1 - t;  // slice 0 = most significant byte
          const uint32_t d = ((W[byte >> 2] >> (8 * (byte & 3))) & 0xffu) ^ 0x80u;
          D[t][e >> 2] |= d << (8 * (e & 3));
        }
      }
    }
  }
  int8_t *o = out + chunk * (int64_t)(S * OZ_FRAG) + lane * 16;
#pragma unroll
  for (int t = 0; t < S; ++t) {
    i32x4v v;
    v.x = (int)D[t][0];
    v.y = (int)D[t][1];
    v.z = (int)D[t][2];
    v.w = (int)D[t][3];
    *reinterpret_cast<i32x4v *>(o + t * OZ_FRAG) = v;
  }
}

// ------------------------------------------------------------------------------------------
// 3. slice products on the matrix cores
// ------------------------------------------------------------------------------------------
// Workgroup = 4 waves, 64x64 outputs; wave (wr, wc) owns the 32x32 block (wr, wc).  The k loop
// runs over half-stages h = (k block h/2, slice group h%2): group 0 holds slices [0, G0), group 1
// slices [G0, S).  A half-stage is 4 fragment groups (A row blocks 0,1 and B column blocks 0,1;
// each wave copies one with direct global->LDS loads) in one of 4 LDS slots, so the copies run
// 3 half-stages ahead of the MFMAs.  The MFMAs of group 0 are the slice pairs with t, u < G0;
// those of group 1 the pairs with max(t, u) >= G0 (t + u <= S-1), whose group-0 fragments are
// still in registers.  Within a group, pairs are ordered by max(t, u) so each MFMA needs only
// fragments read before it.
template <int S>
__global__ void __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
k_oz_gemm(const int8_t *__restrict__ As, const int8_t *__restrict__ Bs, int64_t Kb, int64_t Mp, int64_t Np,
          uint32_t *__restrict__ part) {
  constexpr int G0 = (S + 1) / 2, G1 = S - G0;
  constexpr int SLOT = 4 * G0 * OZ_FRAG;
  // s_waitcnt vmcnt(S) lgkmcnt(0): this wave's copies of the two newer half-stages (G0 + G1 = S
  // loads) may still be in flight
  constexpr int WAIT_S = (S & 15) | ((S >> 4) << 14) | (7 << 4);
  __shared__ __attribute__((aligned(1024))) int8_t lds[4 * SLOT];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t tn = blockIdx.x, tm = blockIdx.y, z = blockIdx.z;
  const int64_t kb0 = z * OZ_KPASS_BLOCKS;
  const int64_t nkb = ((kb0 + OZ_KPASS_BLOCKS < Kb) ? kb0 + OZ_KPASS_BLOCKS : Kb) - kb0;
  const int64_t nst = 2 * nkb;
  // this wave's copy source: A row block 2tm + wave (waves 0, 1) or B column block 2tn + wave - 2
  const int8_t *src = (wave < 2) ? As + ((2 * tm + wave) * Kb + kb0) * (int64_t)(S * OZ_FRAG)
                                 : Bs + ((2 * tn + wave - 2) * Kb + kb0) * (int64_t)(S * OZ_FRAG);
  src += lane * 16;
  auto copy = [&](int64_t h) {
    if (h >= nst) return;
    const int8_t *s = src + (h >> 1) * (int64_t)(S * OZ_FRAG) + ((h & 1) ? G0 * OZ_FRAG : 0);
    int8_t *d = &lds[(int)(h & 3) * SLOT + wave * G0 * OZ_FRAG];
    if ((h & 1) == 0) {
#pragma unroll
      for (int t = 0; t < G0; ++t)
        __builtin_amdgcn_global_load_lds((glb_void *)(s + t * OZ_FRAG), (lds_void *)(d + t * OZ_FRAG), 16, 0, 0);
    } else {
#pragma unroll
      for (int t = 0; t < G1; ++t)
        __builtin_amdgcn_global_load_lds((glb_void *)(s + t * OZ_FRAG), (lds_void *)(d + t * OZ_FRAG), 16, 0, 0);
    }
  };
  const int wr = wave >> 1, wc = wave & 1;
  i32x16v acc[S];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[d][r] = 0;
  copy(0);
  copy(1);
  copy(2);
  i32x4v fa[S], fb[S];
  for (int64_t h = 0; h < nst; h += 2) {
    // slice group 0 of k block h/2
    if (h + 2 < nst)
      __builtin_amdgcn_s_waitcnt(WAIT_S);
    else
      __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_s_barrier();
    {
      const int8_t *base = &lds[(int)(h & 3) * SLOT + lane * 16];
#pragma unroll
      for (int t = 0; t < G0; ++t) {
        fa[t] = *reinterpret_cast<const i32x4v *>(base + (wr * G0 + t) * OZ_FRAG);
        fb[t] = *reinterpret_cast<const i32x4v *>(base + ((2 + wc) * G0 + t) * OZ_FRAG);
      }
    }
    copy(h + 3);
#pragma unroll
    for (int mm = 0; mm < G0; ++mm) {
#pragma unroll
      for (int u = 0; u <= mm; ++u) acc[mm + u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[mm], fb[u], acc[mm + u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < mm; ++t) acc[t + mm] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[t], fb[mm], acc[t + mm], 0, 0, 0);
    }
    // slice group 1
    if (h + 3 < nst)
      __builtin_amdgcn_s_waitcnt(WAIT_S);
    else
      __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_s_barrier();
    {
      const int8_t *base = &lds[(int)((h + 1) & 3) * SLOT + lane * 16];
#pragma unroll
      for (int t = 0; t < G1; ++t) {
        fa[G0 + t] = *reinterpret_cast<const i32x4v *>(base + (wr * G0 + t) * OZ_FRAG);
        fb[G0 + t] = *reinterpret_cast<const i32x4v *>(base + ((2 + wc) * G0 + t) * OZ_FRAG);
      }
    }
    copy(h + 4);
#pragma unroll
    for (int mm = G0; mm < S; ++mm) {
#pragma unroll
      for (int u = 0; u <= mm && mm + u < S; ++u)
        acc[mm + u] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[mm], fb[u], acc[mm + u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < mm && t + mm < S; ++t)
        acc[t + mm] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[t], fb[mm], acc[t + mm], 0, 0, 0);
    }
  }
  // V = sum_d acc_d 256^(S-1-d), Horner in 6 signed words; result element r of this lane is
  // row 8(r/4) + 4(lane/32) + r%4, column lane%32 of the wave's block
  const int64_t row0 = tm * OZ_TILE + wr * 32 + 4 * (lane >> 5);
  const int64_t col = tn * OZ_TILE + wc * 32 + (lane & 31);
  uint32_t *pz = part + z * Mp * Np * OZ_PREC;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    uint32_t V[OZ_PW];
#pragma unroll
    for (int i = 0; i < OZ_PW; ++i) V[i] = 0;
#pragma unroll
    for (int d = 0; d < S; ++d) {
#pragma unroll
      for (int i = OZ_PW - 1; i > 0; --i) V[i] = (V[i] << 8) | (V[i - 1] >> 24);
      V[0] <<= 8;
      const uint32_t a = (uint32_t)acc[d][r], ext = (acc[d][r] < 0) ? 0xffffffffu : 0u;
      uint32_t c = 0;
#pragma unroll
      for (int i = 0; i < OZ_PW; ++i) V[i] = addc32(V[i], i == 0 ? a : ext, c, c);
    }
    const int64_t row = row0 + 8 * (r >> 2) + (r & 3);
    uint32_t *p = pz + (row * Np + col) * OZ_PREC;
    i32x4v v0, v1;
    v0.x = (int)V[0];
    v0.y = (int)V[1];
    v0.z = (int)V[2];
    v0.w = (int)V[3];
    v1.x = (int)V[4];
    v1.y = (int)V[5];
    v1.z = 0;
    v1.w = 0;
    *reinterpret_cast<i32x4v *>(p) = v0;
    *reinterpret_cast<i32x4v *>(p + 4) = v1;
  }
}

// ------------------------------------------------------------------------------------------
// 4. finish and certification
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool q_finite_nonzero(q128 x) { return q_bexp(x) != 0x7fffu && !q_iszero(x); }

// One thread per output (i, j), j fastest.  kfac = ceil(log2(K (S+1)/4)).
__global__ void __launch_bounds__(256) k_oz_finish(int64_t M, int64_t N, int S, int kfac, int npass,
                                                   const uint32_t *__restrict__ part, int64_t Mp, int64_t Np,
                                                   const int32_t *__restrict__ eA, const uint32_t *__restrict__ fA,
                                                   const int32_t *__restrict__ eB, const uint32_t *__restrict__ fB,
                                                   q128 alpha, q128 beta, q128 *__restrict__ C, int64_t csr,
                                                   int64_t csc, uint32_t *__restrict__ hdr, uint32_t *__restrict__ list) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= M * N) return;
  const int64_t i = g / N, j = g - i * N;
  q128 *cp = C + i * csr + j * csc;
  const q128 yv = ldq_c(cp);
  const bool ab_ok = (q_bexp(alpha) != 0x7fffu) && (q_bexp(beta) != 0x7fffu) && (q_bexp(yv) != 0x7fffu);
  bool done = false;
  q128 res;
  const bool fin = !((fA[i] | fB[j]) & OZF_NONFINITE);
  if (ab_ok && fin && q_iszero(alpha)) {
    // alpha*AB is a zero; beta*y decides unless it is zero too (then the sign of the exact sum matters)
    res = mul_packed(beta, yv);
    done = !q_iszero(res);
  } else if (ab_ok && fin && eA[i] != OZ_EZERO && eB[j] != OZ_EZERO) {
    uint32_t V[OZ_FINW];
#pragma unroll
    for (int w = 0; w < OZ_FINW; ++w) V[w] = 0;
    for (int z = 0; z < npass; ++z) {
      const uint32_t *p = part + ((int64_t)z * Mp * Np + i * Np + j) * OZ_PREC;
      const i32x4v v0 = *reinterpret_cast<const i32x4v *>(p);
      const i32x4v v1 = *reinterpret_cast<const i32x4v *>(p + 4);
      const uint32_t x[OZ_PW] = {(uint32_t)v0.x, (uint32_t)v0.y, (uint32_t)v0.z, (uint32_t)v0.w, (uint32_t)v1.x,
                                 (uint32_t)v1.y};
      const uint32_t ext = (x[OZ_PW - 1] >> 31) ? 0xffffffffu : 0u;
      uint32_t c = 0;
#pragma unroll
      for (int w = 0; w < OZ_FINW; ++w) V[w] = addc32(V[w], w < OZ_PW ? x[w] : ext, c, c);
    }
    const uint32_t sgn = V[OZ_FINW - 1] >> 31;
    if (sgn) {
      uint32_t c = 1;
#pragma unroll
      for (int w = 0; w < OZ_FINW; ++w) V[w] = addc32(~V[w], 0u, c, c);
    }
    if (!rzero<OZ_FINW>(V)) {
      const int Ei = eA[i] + 3, Fj = eB[j] + 3;
      const int lsb = Ei + Fj - 8 * (S + 1);
      res = finish_reg_t<OZ_FINW>(V, lsb, sgn, 0u, 0u, true, alpha, beta, yv);
      if (q_finite_nonzero(res)) {
        const int ea = unpack_finite(alpha).e;
        const int bexp = ea + 1 + kfac + Ei + Fj - 8 * S;  // |alpha (AB - P)| < 2^bexp
        done = bexp <= unpack_finite(res).e - 106;
      }
    }
  }
  if (done) {
    stq_c(cp, res);
  } else {
    const uint32_t slot = atomicAdd(hdr + 1, 1u);
    list[2 * (int64_t)slot] = (uint32_t)i;
    list[2 * (int64_t)slot + 1] = (uint32_t)j;
  }
}

// ------------------------------------------------------------------------------------------
// 5. exact repair of the queued entries
// ------------------------------------------------------------------------------------------
// One workgroup per queued entry (grid-stride over the queue): exact accumulation of
// sum_l A_il B_lj in two windows per lane with a block-uniform anchor placing the largest
// possible product (from the row and column maxima) at the top of the high window, then the
// exact alpha*S + beta*C rounded once.  A product below the low window sets bit 0 of the
// status word and the entry becomes NaN (as in the EXACT form).
constexpr int OZR_BS = 256;
__global__ void __launch_bounds__(OZR_BS) k_oz_repair(int64_t K, const q128 *__restrict__ A, int64_t asr, int64_t asc,
                                                      const q128 *__restrict__ B, int64_t bsr, int64_t bsc,
                                                      const int32_t *__restrict__ eA, const int32_t *__restrict__ eB,
                                                      q128 alpha, q128 beta, q128 *__restrict__ C, int64_t csr,
                                                      int64_t csc, uint32_t *__restrict__ hdr,
                                                      const uint32_t *__restrict__ list) {
  __shared__ uint32_t s_w[OZR_BS / 64][2 * WW + 1];
  const uint32_t count = hdr[1];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    const int64_t i = list[2 * (int64_t)e], j = list[2 * (int64_t)e + 1];
    const int ei = eA[i], ej = eB[j];
    const int An = (ei == OZ_EZERO || ej == OZ_EZERO) ? 0 : ei + ej - 224 - (WIN_BAND + WIN_OMAX);
    uint32_t Wl[WW], Wh[WW];
#pragma unroll
    for (int k = 0; k < WW; ++k) {
      Wl[k] = 0;
      Wh[k] = 0;
    }
    uint32_t flags = 0;
    const q128 *ar = A + i * asr;
    const q128 *bc = B + j * bsc;
    for (int64_t l = threadIdx.x; l < K; l += OZR_BS) accumulate(Wl, Wh, An, ldq_c(ar + l * asc), ldq_c(bc + l * bsr), flags);
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
      for (int w = 1; w < OZR_BS / 64; ++w) {
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
      q128 *cp = C + i * csr + j * csc;
      q128 out;
      if (flags & FL_LOST) {
        out = q_nan();
        atomicOr(hdr, 1u);
      } else {
        uint32_t Mw[SW];
        uint32_t sign;
        s_from_windows(Mw, sign, Wl, Wh);
        out = finish_reg(Mw, An, sign, 0u, flags, true, alpha, beta, ldq_c(cp));
      }
      stq_c(cp, out);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
size_t gemm_ozaki_ws(int64_t m, int64_t n, int64_t k, int S) { return oz_plan(m, n, k, S).total; }

template <int S>
static void oz_launch_slices(const oz_layout &L, const q128 *A, int64_t m, int64_t k, int64_t asr, int64_t asc,
                             const q128 *B, int64_t n, int64_t bsr, int64_t bsc, char *ws, hipStream_t st) {
  const int64_t ta = (L.Mp / 32) * L.Kb * 64, tb = (L.Np / 32) * L.Kb * 64;
  k_oz_slice<S><<<(unsigned)((ta + 255) / 256), 256, 0, st>>>(A, m, k, asr, asc, (const int32_t *)(ws + L.eA),
                                                               (int8_t *)(ws + L.sA), L.Kb, ta);
  // B as rows j, columns l
  k_oz_slice<S><<<(unsigned)((tb + 255) / 256), 256, 0, st>>>(B, n, k, bsc, bsr, (const int32_t *)(ws + L.eB),
                                                               (int8_t *)(ws + L.sB), L.Kb, tb);
}

template <int S>
static void oz_launch_products(const oz_layout &L, char *ws, hipStream_t st) {
  dim3 grid((unsigned)(L.Np / OZ_TILE), (unsigned)(L.Mp / OZ_TILE), (unsigned)L.npass);
  k_oz_gemm<S><<<grid, 256, 0, st>>>((const int8_t *)(ws + L.sA), (const int8_t *)(ws + L.sB), L.Kb, L.Mp, L.Np,
                                     (uint32_t *)(ws + L.part));
}

static void oz_scale(const q128 *X, int64_t R, int64_t Cn, int64_t sr, int64_t sc, int32_t *E, uint32_t *F,
                     hipStream_t st) {
  const int contig = (sc == 1);
  // enough blocks to fill the chip: split the reduced dimension into chunks
  const int64_t rblocks = contig ? (R + 3) / 4 : (R + 255) / 256;
  int64_t nch = (2048 + rblocks - 1) / rblocks;
  if (nch < 1) nch = 1;
  int64_t cchunk = (Cn + nch - 1) / nch;
  if (cchunk < 64) cchunk = 64;
  nch = (Cn + cchunk - 1) / cchunk;
  dim3 grid((unsigned)rblocks, (unsigned)nch);
  k_oz_scale<<<grid, 256, 0, st>>>(X, R, Cn, sr, sc, contig, cchunk, E, F);
}

hipError_t launch_gemm_ozaki(int colmajor, int64_t m, int64_t n, int64_t k, int S, q128 alpha, const q128 *A,
                             int64_t lda, const q128 *B, int64_t ldb, q128 beta, q128 *C, int64_t ldc, void *wsv,
                             hipStream_t st, hipEvent_t *ev) {
  if (S < OZ_SMIN || S > OZ_SMAX) return hipErrorInvalidValue;
  char *ws = (char *)wsv;
  const oz_layout L = oz_plan(m, n, k, S);
  // element strides: A(i,l), B(l,j), C(i,j)
  const int64_t asr = colmajor ? 1 : lda, asc = colmajor ? lda : 1;
  const int64_t bsr = colmajor ? 1 : ldb, bsc = colmajor ? ldb : 1;
  const int64_t csr = colmajor ? 1 : ldc, csc = colmajor ? ldc : 1;
  int32_t *eA = (int32_t *)(ws + L.eA), *eB = (int32_t *)(ws + L.eB);
  uint32_t *fA = (uint32_t *)(ws + L.fA), *fB = (uint32_t *)(ws + L.fB);
  uint32_t *hdr = (uint32_t *)ws;
  if (ev) (void)hipEventRecord(ev[0], st);
  const int64_t ninit = (L.Mp > L.Np ? L.Mp : L.Np) > OZ_HDR / 4 ? (L.Mp > L.Np ? L.Mp : L.Np) : OZ_HDR / 4;
  k_oz_init<<<(unsigned)((ninit + 255) / 256), 256, 0, st>>>(eA, fA, L.Mp, eB, fB, L.Np, hdr);
  oz_scale(A, m, k, asr, asc, eA, fA, st);
  oz_scale(B, n, k, bsc, bsr, eB, fB, st);  // B columns: r = j, c = l
  switch (S) {
#define QB_OZ_CASE(s) \
  case s: oz_launch_slices<s>(L, A, m, k, asr, asc, B, n, bsr, bsc, ws, st); break;
    QB_OZ_CASE(8) QB_OZ_CASE(9) QB_OZ_CASE(10) QB_OZ_CASE(11) QB_OZ_CASE(12) QB_OZ_CASE(13) QB_OZ_CASE(14)
    QB_OZ_CASE(15) QB_OZ_CASE(16) QB_OZ_CASE(17) QB_OZ_CASE(18) QB_OZ_CASE(19)
#undef QB_OZ_CASE
  default: break;
  }
  if (ev) (void)hipEventRecord(ev[1], st);
  switch (S) {
#define QB_OZ_CASE(s) \
  case s: oz_launch_products<s>(L, ws, st); break;
    QB_OZ_CASE(8) QB_OZ_CASE(9) QB_OZ_CASE(10) QB_OZ_CASE(11) QB_OZ_CASE(12) QB_OZ_CASE(13) QB_OZ_CASE(14)
    QB_OZ_CASE(15) QB_OZ_CASE(16) QB_OZ_CASE(17) QB_OZ_CASE(18) QB_OZ_CASE(19)
#undef QB_OZ_CASE
  default: break;
  }
  if (ev) (void)hipEventRecord(ev[2], st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // kfac = ceil(log2(K (S+1) / 4))
  const double kb = (double)k * (S + 1) / 4.0;
  int kfac = 0;
  while ((double)(1ull << (kfac < 62 ? kfac : 62)) < kb && kfac < 80) ++kfac;
  const int64_t mn = m * n;
  k_oz_finish<<<(unsigned)((mn + 255) / 256), 256, 0, st>>>(m, n, S, kfac, (int)L.npass, (const uint32_t *)(ws + L.part),
                                                            L.Mp, L.Np, eA, fA, eB, fB, alpha, beta, C, csr, csc, hdr,
                                                            (uint32_t *)(ws + L.list));
  if (ev) (void)hipEventRecord(ev[3], st);
  k_oz_repair<<<OZ_REPAIR_BLOCKS, OZR_BS, 0, st>>>(k, A, asr, asc, B, bsr, bsc, eA, eB, alpha, beta, C, csr, csc, hdr,
                                                   (const uint32_t *)(ws + L.list));
  if (ev) (void)hipEventRecord(ev[4], st);
  return hipGetLastError();
}

// number of entries the last call queued for exact repair (header word 1)
hipError_t ozaki_repair_count(const void *ws, uint32_t *out, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(out, (const char *)ws + 4, 4, hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

}  // namespace qbk
