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
//    diagonal sums acc_d = sum_{t+u=d} D_t D_u^T, d in [D0, D1], of its 32x32 tile in int32
//    registers (no overflow: (d+1) * 4096 * 2^14 < 2^31 for S <= 31 and k passes of 4096), and
//    writes V = sum_{d=D0..D1} acc_d 256^(S-1-d) (exact: |V| < 2^(8S+17), 8 signed words for
//    S <= 28) per k pass.  Up to 19
//    slices one launch covers all diagonals; for 20..28 slices two launches split the diagonals
//    so the accumulators of each fit in the register file.
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
constexpr int OZ_PW = 8;                // signed 32-bit words of a pass sum V (and of N + bias)
constexpr int OZ_PREC = 8;              // words per pass-sum record (32 bytes)
constexpr int OZ_FINW = 9;              // words of |sum of passes| in the finish
constexpr int32_t OZ_EZERO = INT32_MIN; // row/column exponent of an all-zero row/column
constexpr uint32_t OZF_NONFINITE = 1u;
constexpr int OZ_REPAIR_BLOCKS = 1024;
constexpr int OZ_HDR = 256;             // workspace header bytes: status, repair count
constexpr int OZ_PROBE_MAX = 1024;      // sampled entries of the slice-count probe

// diagonal ranges: up to 19 slices one range [0, S-1] with 2 slice groups per k block; 20..28
// slices two ranges [0, D-1] and [D, S-1], D = (3S+4)/5 (balancing accumulators against
// fragments in the register file), with 3 slice groups so 4 LDS slots fit in 160 KB
constexpr int OZ_ONE_RANGE_MAX = 19;
__host__ __device__ constexpr int oz_ranges(int S) { return S <= OZ_ONE_RANGE_MAX ? 1 : 2; }
__host__ __device__ constexpr int oz_split(int S) { return (3 * S + 4) / 5; }

struct oz_layout {
  int64_t Mp, Np, Kb, npass, npass_k;
  size_t eA, fA, eB, fB, sA, sB, part, list, bt, total;
};

static oz_layout oz_plan(int64_t m, int64_t n, int64_t k, int S) {
  oz_layout L;
  L.Mp = (m + OZ_TILE - 1) / OZ_TILE * OZ_TILE;
  L.Np = (n + OZ_TILE - 1) / OZ_TILE * OZ_TILE;
  L.Kb = (k + 31) / 32;
  L.npass_k = (L.Kb + OZ_KPASS_BLOCKS - 1) / OZ_KPASS_BLOCKS;
  L.npass = L.npass_k * oz_ranges(S);
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
  L.bt = o;  // B^T (n x k) for the repair of many entries when B's columns are strided
  o = al(o + (size_t)n * k * 16);
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
  if (g < OZ_HDR / 4) hdr[g] = 0u;
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
// W (8 words) = N + bias where N = round(x 2^(8S - Es)) (magnitudes rounded half up, then signed)
template <int S>
__device__ __forceinline__ void oz_digits(q128 x, int Es, uint32_t (&W)[OZ_PW]) {
  uint32_t Nw[OZ_PW];
#pragma unroll
  for (int i = 0; i < OZ_PW; ++i) Nw[i] = 0;
  const uint32_t be = q_bexp(x);
  if (be != 0x7fffu && !q_iszero(x)) {
    const uq u = unpack_finite(x);
    const int sh = u.e - 112 - Es + 8 * S;  // N = m * 2^sh;  u.e <= Es - 3 so m 2^sh < 2^(8S-2)
    uint32_t m[OZ_PW] = {(uint32_t)u.ml, (uint32_t)(u.ml >> 32), (uint32_t)u.mh, (uint32_t)(u.mh >> 32), 0u, 0u, 0u, 0u};
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

// Row-contiguous operands (sc == 1): one workgroup per (32-row block, 4 k blocks).  Threads
// read 4 consecutive entries (64 bytes) each, so a load instruction covers 4 KB of two rows;
// digits are packed 4 per word into an LDS tile [slice][32 rows][128 k], then written out as
// whole 1 KB fragments.  Same output as k_oz_slice.
constexpr int OZS_KB = 4;  // k blocks per workgroup
template <int S>
__global__ void __launch_bounds__(256) k_oz_slice_rows(const q128 *__restrict__ X, int64_t R, int64_t C, int64_t sr,
                                                       const int32_t *__restrict__ E, int8_t *__restrict__ out,
                                                       int64_t Kb) {
  constexpr int KW = OZS_KB * 32;  // k per tile
  __shared__ __attribute__((aligned(16))) uint32_t tile[S][32][KW / 4];
  const int64_t rb = blockIdx.y, kb0 = (int64_t)blockIdx.x * OZS_KB;
  const int t = threadIdx.x;
  // phase 1: 32 x 128 entries, 4 consecutive per thread and pass, 4 passes
#pragma unroll 1
  for (int q = 0; q < (32 * KW) / (4 * 256); ++q) {
    const int e4 = t + 256 * q;        // group of 4 entries
    const int row = e4 / (KW / 4), c4 = e4 % (KW / 4);
    const int64_t r = rb * 32 + row;
    const int64_t c = kb0 * 32 + 4 * c4;
    uint32_t D[S];
#pragma unroll
    for (int sl = 0; sl < S; ++sl) D[sl] = 0;
    const int Er = (r < R) ? E[r] : OZ_EZERO;
    if (Er != OZ_EZERO) {
      const q128 *rowp = X + r * sr;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (c + u < C) {
          uint32_t W[OZ_PW];
          oz_digits<S>(ldq_c(rowp + c + u), Er + 3, W);
#pragma unroll
          for (int sl = 0; sl < S; ++sl) {
            const int byte = S - 1 - sl;
            D[sl] |= (((W[byte >> 2] >> (8 * (byte & 3))) & 0xffu) ^ 0x80u) << (8 * u);
          }
        }
      }
    }
#pragma unroll
    for (int sl = 0; sl < S; ++sl) tile[sl][row][c4] = D[sl];
  }
  __syncthreads();
  // phase 2: fragments (k block kbl, slice sl): lane L holds row L%32, k 16(L/32)..+15
  const int64_t nkb = (kb0 + OZS_KB <= Kb) ? OZS_KB : Kb - kb0;
  const int L = t & 63;
  for (int f = t >> 6; f < OZS_KB * S; f += 4) {
    const int kbl = f / S, sl = f % S;
    if (kbl < nkb) {
      const uint32_t *src = &tile[sl][L & 31][(kbl * 32 + (L >> 5) * 16) / 4];
      i32x4v v;
      v.x = (int)src[0];
      v.y = (int)src[1];
      v.z = (int)src[2];
      v.w = (int)src[3];
      *reinterpret_cast<i32x4v *>(out + ((rb * Kb + kb0 + kbl) * S + sl) * (int64_t)OZ_FRAG + L * 16) = v;
    }
  }
}

// ------------------------------------------------------------------------------------------
// 3. slice products on the matrix cores
// ------------------------------------------------------------------------------------------
// Workgroup = 4 waves, 64x64 outputs; wave (wr, wc) owns the 32x32 block (wr, wc).  The k loop
// runs over stages h = (k block h/NG, slice group h%NG): group g holds slices [gG, min(S, gG+G)),
// G = ceil(S/NG).  A stage is 4 fragment groups (A row blocks 0,1 and B column blocks 0,1; each
// wave copies one with direct global->LDS loads) in one of 4 LDS slots, so the copies run 3
// stages ahead of the MFMAs.  The MFMAs of stage g are the slice pairs (t, u) with max(t, u) in
// group g and D0 <= t + u <= D1; fragments of earlier groups of the same k block are still in
// registers.  Within a stage pairs are ordered by max(t, u), so each MFMA needs only fragments
// read before it.  Pass sums go to record z = zbase + k pass.
template <int S, int D0, int D1, int NG>
__global__ void __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
k_oz_gemm(const int8_t *__restrict__ As, const int8_t *__restrict__ Bs, int64_t Kb, int64_t Mp, int64_t Np,
          int zbase, uint32_t *__restrict__ part) {
  constexpr int G = (S + NG - 1) / NG;
  constexpr int SLOT = 4 * G * OZ_FRAG;
  constexpr int ND = D1 - D0 + 1;
  constexpr int TMAX = (D1 < S - 1) ? D1 : S - 1;  // highest slice any pair of the range uses
  __shared__ __attribute__((aligned(1024))) int8_t lds[4 * SLOT];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int64_t tn = blockIdx.x, tm = blockIdx.y;
  const int64_t kp = blockIdx.z;
  if ((gridDim.x % 16) == 0 && (gridDim.y % 16) == 0) {
    // Rounds of 256 consecutive workgroups cover a 16x16 super-tile (row-major over super-tiles),
    // and the 32 of a round that share an XCD (workgroup id mod 8) a 4x8 block of it: a round's
    // slice panels stay within the infinity cache and each XCD's L2 serves 4 + 8 of them.
    const int64_t L = blockIdx.x + (int64_t)gridDim.x * blockIdx.y;
    const int64_t r = L >> 8, w = L & 255, xcd = w & 7, idx = w >> 3;
    const int64_t SC = gridDim.x >> 4;
    tm = (r / SC) * 16 + (xcd >> 1) * 4 + (idx >> 3);
    tn = (r % SC) * 16 + (xcd & 1) * 8 + (idx & 7);
  }
  const int64_t kb0 = kp * OZ_KPASS_BLOCKS;
  const int64_t nkb = ((kb0 + OZ_KPASS_BLOCKS < Kb) ? kb0 + OZ_KPASS_BLOCKS : Kb) - kb0;
  const int64_t nst = NG * nkb;
  // this wave's copy source: A row block 2tm + wave (waves 0, 1) or B column block 2tn + wave - 2
  const int8_t *src = (wave < 2) ? As + ((2 * tm + wave) * Kb + kb0) * (int64_t)(S * OZ_FRAG)
                                 : Bs + ((2 * tn + wave - 2) * Kb + kb0) * (int64_t)(S * OZ_FRAG);
  src += lane * 16;
  auto copy = [&](int64_t h) {
    if (h >= nst) return;
    const int64_t kb = h / NG;
    const int g = (int)(h - kb * NG);
    const int8_t *sg = src + kb * (int64_t)(S * OZ_FRAG) + g * G * OZ_FRAG;
    int8_t *d = &lds[(int)(h & 3) * SLOT + wave * G * OZ_FRAG];
#pragma unroll
    for (int gg = 0; gg < NG; ++gg) {
      if (g == gg) {
        const int cnt = (S - gg * G < G) ? S - gg * G : G;
#pragma unroll
        for (int t = 0; t < G; ++t)
          if (t < cnt)
            __builtin_amdgcn_global_load_lds((glb_void *)(sg + t * OZ_FRAG), (lds_void *)(d + t * OZ_FRAG), 16, 0,
                                             0);
      }
    }
  };
  const int wr = wave >> 1, wc = wave & 1;
  i32x16v acc[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[d][r] = 0;
  copy(0);
  copy(1);
  copy(2);
  i32x4v fa[S], fb[S];
  for (int64_t kb = 0; kb < nkb; ++kb) {
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const int64_t h = kb * NG + g;
      // this wave's copies of the two newer stages may stay in flight
      const int c1 = (S - ((g + 1) % NG) * G < G) ? S - ((g + 1) % NG) * G : G;
      const int c2 = (S - ((g + 2) % NG) * G < G) ? S - ((g + 2) % NG) * G : G;
      const int pend = c1 + c2;
      if (h + 2 < nst) {
        // s_waitcnt vmcnt(pend) lgkmcnt(0); pend is a compile-time constant after unrolling
        switch (pend) {
#define QB_WAIT_CASE(n) \
  case n: __builtin_amdgcn_s_waitcnt(((n) & 15) | (((n) >> 4) << 14) | (7 << 4)); break;
          QB_WAIT_CASE(2) QB_WAIT_CASE(3) QB_WAIT_CASE(4) QB_WAIT_CASE(5) QB_WAIT_CASE(6) QB_WAIT_CASE(7)
          QB_WAIT_CASE(8) QB_WAIT_CASE(9) QB_WAIT_CASE(10) QB_WAIT_CASE(11) QB_WAIT_CASE(12) QB_WAIT_CASE(13)
          QB_WAIT_CASE(14) QB_WAIT_CASE(15) QB_WAIT_CASE(16) QB_WAIT_CASE(17) QB_WAIT_CASE(18) QB_WAIT_CASE(19)
          QB_WAIT_CASE(20) QB_WAIT_CASE(21) QB_WAIT_CASE(22) QB_WAIT_CASE(23) QB_WAIT_CASE(24) QB_WAIT_CASE(25)
          QB_WAIT_CASE(26) QB_WAIT_CASE(27) QB_WAIT_CASE(28)
#undef QB_WAIT_CASE
          default: __builtin_amdgcn_s_waitcnt(0); break;
        }
      } else {
        __builtin_amdgcn_s_waitcnt(0);
      }
      __builtin_amdgcn_s_barrier();
      {
        const int8_t *base = &lds[(int)(h & 3) * SLOT + lane * 16];
#pragma unroll
        for (int t = 0; t < G; ++t) {
          const int sl = g * G + t;
          if (sl < S && sl <= TMAX) {
            fa[sl] = *reinterpret_cast<const i32x4v *>(base + (wr * G + t) * OZ_FRAG);
            fb[sl] = *reinterpret_cast<const i32x4v *>(base + ((2 + wc) * G + t) * OZ_FRAG);
          }
        }
      }
      copy(h + 3);
#pragma unroll
      for (int t2 = 0; t2 < G; ++t2) {
        const int mm = g * G + t2;
        if (mm < S) {
#pragma unroll
          for (int u = 0; u <= mm; ++u) {
            const int d = mm + u;
            if (d >= D0 && d <= D1) acc[d - D0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[mm], fb[u], acc[d - D0], 0, 0, 0);
          }
#pragma unroll
          for (int t = 0; t < mm; ++t) {
            const int d = t + mm;
            if (d >= D0 && d <= D1) acc[d - D0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[t], fb[mm], acc[d - D0], 0, 0, 0);
          }
        }
      }
    }
  }
  // V = sum_{d=D0..D1} acc_d 256^(S-1-d): Horner over the range in OZ_PW signed words, then a
  // shift by 8 (S-1-D1); result element r of this lane is row 8(r/4) + 4(lane/32) + r%4,
  // column lane%32 of the wave's block
  const int64_t row0 = tm * OZ_TILE + wr * 32 + 4 * (lane >> 5);
  const int64_t col = tn * OZ_TILE + wc * 32 + (lane & 31);
  uint32_t *pz = part + (zbase + kp) * Mp * Np * OZ_PREC;
  constexpr int SH = 8 * (S - 1 - D1);
#pragma clang loop unroll(full)
  for (int r = 0; r < 16; ++r) {
    uint32_t V[OZ_PW];
#pragma unroll
    for (int i = 0; i < OZ_PW; ++i) V[i] = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
#pragma unroll
      for (int i = OZ_PW - 1; i > 0; --i) V[i] = (V[i] << 8) | (V[i - 1] >> 24);
      V[0] <<= 8;
      const uint32_t a = (uint32_t)acc[d][r], ext = (acc[d][r] < 0) ? 0xffffffffu : 0u;
      uint32_t c = 0;
#pragma unroll
      for (int i = 0; i < OZ_PW; ++i) V[i] = addc32(V[i], i == 0 ? a : ext, c, c);
    }
    if (SH > 0) {
      // left shift of the two's complement value by SH bits (SH < 32 * OZ_PW)
      constexpr int WS = SH / 32, BS = SH % 32;
      uint32_t T[OZ_PW];
#pragma unroll
      for (int i = OZ_PW - 1; i >= 0; --i) {
        const uint32_t hi = (i - WS >= 0) ? V[i - WS] : 0u;
        const uint32_t lo = (i - WS - 1 >= 0) ? V[i - WS - 1] : 0u;
        T[i] = BS ? ((hi << BS) | (lo >> (32 - BS))) : hi;
      }
#pragma unroll
      for (int i = 0; i < OZ_PW; ++i) V[i] = T[i];
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
    v1.z = (int)V[6];
    v1.w = (int)V[7];
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
                                 (uint32_t)v1.y, (uint32_t)v1.z, (uint32_t)v1.w};
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

// Exact sum_l A_il B_lj by one workgroup: two windows per lane (anchor An), wave and block
// reductions; thread 0 returns with the block's windows and flags.
__device__ __forceinline__ void oz_block_dot(int64_t K, const q128 *__restrict__ ar, int64_t asc,
                                             const q128 *__restrict__ bc, int64_t bsr, int An, uint32_t (&Wl)[WW],
                                             uint32_t (&Wh)[WW], uint32_t &flags, uint32_t (*s_w)[2 * WW + 1]) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < WW; ++k) {
    Wl[k] = 0;
    Wh[k] = 0;
  }
  flags = 0;
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
  }
}

// window anchor placing the largest possible product (from the row and column maxima) at the top
// of the high window
__device__ __forceinline__ int oz_anchor(int ei, int ej) {
  return (ei == OZ_EZERO || ej == OZ_EZERO) ? 0 : ei + ej - 224 - (WIN_BAND + WIN_OMAX);
}

__global__ void __launch_bounds__(OZR_BS) k_oz_repair(int64_t K, const q128 *__restrict__ A, int64_t asr, int64_t asc,
                                                      const q128 *__restrict__ B, int64_t bsr, int64_t bsc,
                                                      const int32_t *__restrict__ eA, const int32_t *__restrict__ eB,
                                                      q128 alpha, q128 beta, q128 *__restrict__ C, int64_t csr,
                                                      int64_t csc, uint32_t *__restrict__ hdr,
                                                      const uint32_t *__restrict__ list) {
  __shared__ uint32_t s_w[OZR_BS / 64][2 * WW + 1];
  const uint32_t count = hdr[1];
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    const int64_t i = list[2 * (int64_t)e], j = list[2 * (int64_t)e + 1];
    const int An = oz_anchor(eA[i], eB[j]);
    uint32_t Wl[WW], Wh[WW], flags;
    oz_block_dot(K, A + i * asr, asc, B + j * bsc, bsr, An, Wl, Wh, flags, s_w);
    if (threadIdx.x == 0) {
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
// 6. slice-count probe
// ------------------------------------------------------------------------------------------
// For P pseudo-random entries (i, j): the exact sum_l A_il B_lj, and r = ilogb(|sum|) -
// (E_i + F_j) (how far the entry lies below the scale of its row and column), into out[e]
// (INT_MAX for entries that are zero, non-finite, outside the windows or in a zero row/column).
// This drives the choice of the slice count.
__global__ void __launch_bounds__(OZR_BS) k_oz_probe(int64_t M, int64_t N, int64_t K, const q128 *__restrict__ A,
                                                     int64_t asr, int64_t asc, const q128 *__restrict__ B,
                                                     int64_t bsr, int64_t bsc, const int32_t *__restrict__ eA,
                                                     const int32_t *__restrict__ eB, int P, int *__restrict__ out) {
  __shared__ uint32_t s_w[OZR_BS / 64][2 * WW + 1];
  for (int e = blockIdx.x; e < P; e += gridDim.x) {
    // splitmix64 of the sample number picks (i, j)
    uint64_t z = 0x5eed0000ull + (uint64_t)(e + 1) * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    const int64_t i = (int64_t)((z & 0xffffffffull) % (uint64_t)M), j = (int64_t)((z >> 32) % (uint64_t)N);
    const int ei = eA[i], ej = eB[j];
    if (ei == OZ_EZERO || ej == OZ_EZERO) {  // uniform per block
      if (threadIdx.x == 0) out[e] = 0x7fffffff;
      continue;
    }
    const int An = oz_anchor(ei, ej);
    uint32_t Wl[WW], Wh[WW], flags;
    oz_block_dot(K, A + i * asr, asc, B + j * bsc, bsr, An, Wl, Wh, flags, s_w);
    if (threadIdx.x == 0) {
      int r = 0x7fffffff;
      if (!(flags & (FL_LOST | FL_NAN | FL_PINF | FL_NINF))) {
        uint32_t Mw[SW];
        uint32_t sign;
        s_from_windows(Mw, sign, Wl, Wh);
        const int msb = rmsb<SW>(Mw);
        if (msb >= 0) r = An + msb - (ei + 3) - (ej + 3);
      }
      out[e] = r;
    }
    __syncthreads();
  }
}

// out[j * K + l] = B[l * ldb + j] (K x N row-major -> N x K row-major), 32 x 32 tiles via LDS
__global__ void __launch_bounds__(256) k_oz_transpose(const q128 *__restrict__ B, int64_t K, int64_t N, int64_t ldb,
                                                      q128 *__restrict__ out) {
  __shared__ uint4 t[32][33];
  const int64_t l0 = (int64_t)blockIdx.y * 32, j0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int r = ty; r < 32; r += 8) {
    const int64_t l = l0 + r, j = j0 + tx;
    if (l < K && j < N) t[r][tx] = *reinterpret_cast<const uint4 *>(B + l * ldb + j);
  }
  __syncthreads();
#pragma unroll
  for (int r = ty; r < 32; r += 8) {
    const int64_t j = j0 + r, l = l0 + tx;
    if (l < K && j < N) *reinterpret_cast<uint4 *>(out + j * K + l) = t[tx][r];
  }
}

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
size_t gemm_ozaki_ws(int64_t m, int64_t n, int64_t k, int S) { return oz_plan(m, n, k, S).total; }

// one operand X(r, c) = X[r*sr + c*sc], R x C, into slices (rows padded to Rp)
template <int S>
static void oz_slice_operand(const q128 *X, int64_t R, int64_t C, int64_t sr, int64_t sc, int64_t Rp,
                             const int32_t *E, int8_t *out, int64_t Kb, hipStream_t st) {
  if (sc == 1) {
    dim3 grid((unsigned)((Kb + OZS_KB - 1) / OZS_KB), (unsigned)(Rp / 32));
    k_oz_slice_rows<S><<<grid, 256, 0, st>>>(X, R, C, sr, E, out, Kb);
  } else {
    const int64_t nt = (Rp / 32) * Kb * 64;
    k_oz_slice<S><<<(unsigned)((nt + 255) / 256), 256, 0, st>>>(X, R, C, sr, sc, E, out, Kb, nt);
  }
}

template <int S>
static void oz_launch_slices(const oz_layout &L, const q128 *A, int64_t m, int64_t k, int64_t asr, int64_t asc,
                             const q128 *B, int64_t n, int64_t bsr, int64_t bsc, char *ws, hipStream_t st) {
  oz_slice_operand<S>(A, m, k, asr, asc, L.Mp, (const int32_t *)(ws + L.eA), (int8_t *)(ws + L.sA), L.Kb, st);
  // B as rows j, columns l
  oz_slice_operand<S>(B, n, k, bsc, bsr, L.Np, (const int32_t *)(ws + L.eB), (int8_t *)(ws + L.sB), L.Kb, st);
}

template <int S>
static void oz_launch_products(const oz_layout &L, char *ws, hipStream_t st) {
  dim3 grid((unsigned)(L.Np / OZ_TILE), (unsigned)(L.Mp / OZ_TILE), (unsigned)L.npass_k);
  const int8_t *As = (const int8_t *)(ws + L.sA), *Bs = (const int8_t *)(ws + L.sB);
  uint32_t *part = (uint32_t *)(ws + L.part);
  if constexpr (S <= OZ_ONE_RANGE_MAX) {
    k_oz_gemm<S, 0, S - 1, 2><<<grid, 256, 0, st>>>(As, Bs, L.Kb, L.Mp, L.Np, 0, part);
  } else {
    constexpr int D = oz_split(S);
    k_oz_gemm<S, 0, D - 1, 3><<<grid, 256, 0, st>>>(As, Bs, L.Kb, L.Mp, L.Np, 0, part);
    k_oz_gemm<S, D, S - 1, 3><<<grid, 256, 0, st>>>(As, Bs, L.Kb, L.Mp, L.Np, (int)L.npass_k, part);
  }
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
                             hipStream_t st, hipEvent_t *ev, int sync_repair) {
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
    QB_OZ_CASE(15) QB_OZ_CASE(16) QB_OZ_CASE(17) QB_OZ_CASE(18) QB_OZ_CASE(19) QB_OZ_CASE(20) QB_OZ_CASE(21)
    QB_OZ_CASE(22) QB_OZ_CASE(23) QB_OZ_CASE(24) QB_OZ_CASE(25) QB_OZ_CASE(26) QB_OZ_CASE(27) QB_OZ_CASE(28)
#undef QB_OZ_CASE
  default: break;
  }
  if (ev) (void)hipEventRecord(ev[1], st);
  switch (S) {
#define QB_OZ_CASE(s) \
  case s: oz_launch_products<s>(L, ws, st); break;
    QB_OZ_CASE(8) QB_OZ_CASE(9) QB_OZ_CASE(10) QB_OZ_CASE(11) QB_OZ_CASE(12) QB_OZ_CASE(13) QB_OZ_CASE(14)
    QB_OZ_CASE(15) QB_OZ_CASE(16) QB_OZ_CASE(17) QB_OZ_CASE(18) QB_OZ_CASE(19) QB_OZ_CASE(20) QB_OZ_CASE(21)
    QB_OZ_CASE(22) QB_OZ_CASE(23) QB_OZ_CASE(24) QB_OZ_CASE(25) QB_OZ_CASE(26) QB_OZ_CASE(27) QB_OZ_CASE(28)
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
  const q128 *Br = B;
  int64_t rbsr = bsr, rbsc = bsc;
  if (sync_repair && bsr != 1 && n > 1) {
    // the caller accepts a synchronization: read the queue length, and when it is long, read B's
    // columns from a transposed copy (contiguous) instead of gathering them with stride ldb
    uint32_t count = 0;
    hipError_t e2 = hipMemcpyAsync(&count, hdr + 1, 4, hipMemcpyDeviceToHost, st);
    if (e2 == hipSuccess) e2 = hipStreamSynchronize(st);
    if (e2 != hipSuccess) return e2;
    if (count >= 512) {
      q128 *Bt = (q128 *)(ws + L.bt);
      dim3 tg((unsigned)((n + 31) / 32), (unsigned)((k + 31) / 32));
      k_oz_transpose<<<tg, 256, 0, st>>>(B, k, n, bsr, Bt);
      Br = Bt;
      rbsr = 1;
      rbsc = k;
    }
  }
  k_oz_repair<<<OZ_REPAIR_BLOCKS, OZR_BS, 0, st>>>(k, A, asr, asc, Br, rbsr, rbsc, eA, eB, alpha, beta, C, csr, csc,
                                                   hdr, (const uint32_t *)(ws + L.list));
  if (ev) (void)hipEventRecord(ev[4], st);
  return hipGetLastError();
}

size_t ozaki_probe_ws(int64_t m, int64_t n) {
  return (size_t)OZ_HDR + 256 + (size_t)(m + 64) * 8 + (size_t)(n + 64) * 8 + 1024 + OZ_PROBE_MAX * 4;
}
size_t ozaki_probe_out_offset(int64_t m, int64_t n) {
  return (size_t)OZ_HDR + 256 + (size_t)(m + 64) * 8 + (size_t)(n + 64) * 8 + 1024;
}

hipError_t launch_ozaki_probe(int colmajor, int64_t m, int64_t n, int64_t k, const q128 *A, int64_t lda,
                              const q128 *B, int64_t ldb, int samples, void *wsv, hipStream_t st) {
  char *ws = (char *)wsv;
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  uint32_t *hdr = (uint32_t *)ws;
  int32_t *eA = (int32_t *)(ws + OZ_HDR);
  uint32_t *fA = (uint32_t *)(ws + al(OZ_HDR + (size_t)m * 4));
  int32_t *eB = (int32_t *)(ws + al(OZ_HDR + (size_t)m * 4) + al((size_t)m * 4));
  uint32_t *fB = (uint32_t *)((char *)eB + al((size_t)n * 4));
  const int64_t asr = colmajor ? 1 : lda, asc = colmajor ? lda : 1;
  const int64_t bsr = colmajor ? 1 : ldb, bsc = colmajor ? ldb : 1;
  const int64_t ninit = (m > n ? m : n) > OZ_HDR / 4 ? (m > n ? m : n) : OZ_HDR / 4;
  k_oz_init<<<(unsigned)((ninit + 255) / 256), 256, 0, st>>>(eA, fA, m, eB, fB, n, hdr);
  oz_scale(A, m, k, asr, asc, eA, fA, st);
  oz_scale(B, n, k, bsc, bsr, eB, fB, st);
  if (samples > OZ_PROBE_MAX) samples = OZ_PROBE_MAX;
  k_oz_probe<<<(unsigned)(samples < 256 ? samples : 256), OZR_BS, 0, st>>>(
      m, n, k, A, asr, asc, B, bsr, bsc, eA, eB, samples, (int *)(ws + ozaki_probe_out_offset(m, n)));
  return hipGetLastError();
}

// Slice count for a probe value r (sampled entries at or above it lie >= 2^r below
// 2^(E_i+F_j)): the smallest S with 1 + kfac(S) - 8S + 2 <= r - 106, i.e. the certification bound
// (alpha = 1) passes those entries with two bits to spare; clamped to [12, OZ_SMAX].
// r = INT_MAX (no usable sample) gives 17.
int ozaki_slices_for(int64_t k, int r) {
  if (r == 0x7fffffff) return 17;
  for (int S = 12; S <= OZ_SMAX; ++S) {
    const double kb = (double)k * (S + 1) / 4.0;
    int kfac = 0;
    while ((double)(1ull << (kfac < 62 ? kfac : 62)) < kb && kfac < 80) ++kfac;
    if (1 + kfac - 8 * S + 2 <= r - 106) return S;
  }
  return OZ_SMAX;
}

// number of entries the last call queued for exact repair (header word 1)
hipError_t ozaki_repair_count(const void *ws, uint32_t *out, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(out, (const char *)ws + 4, 4, hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

}  // namespace qbk
