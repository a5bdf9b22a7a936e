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
//    queued and recomputed by the EXACT form's exact accumulation (correctly rounded); an entry
//    whose products lie too far apart for the windows takes the full-range path (wide.hpp).
// So the result of every entry has relative error <= 2^-105.9 or is correctly rounded.
#include "wide.hpp"
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
  o = al(o + (size_t)m * n * 12);  // (i, j) pairs, then the full-range queue
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
    const int64_t r = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    if (r >= R) return;  // whole waves
    const q128 *row = X + r * sr;
    int64_t c = c0 + lane;
    for (; c + 192 < c1; c += 256) {  // four loads in flight per lane
      q128 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = ldq_c(row + c + 64 * u);
#pragma unroll
      for (int u = 0; u < 4; ++u) e = max(e, oz_ilogb(v[u], fl));
    }
    for (; c < c1; c += 64) e = max(e, oz_ilogb(ldq_c(row + c), fl));
    e = wave_max(e);
    fl = wave_or(fl);
    if (lane == 0) {
      if (e != OZ_EZERO) atomicMax(E + r, e);
      if (fl) atomicOr(F + r, fl);
    }
  } else {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= R) return;
    int64_t c = c0;
    for (; c + 3 < c1; c += 4) {
      q128 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = ldq_c(X + r * sr + (c + u) * sc);
#pragma unroll
      for (int u = 0; u < 4; ++u) e = max(e, oz_ilogb(v[u], fl));
    }
    for (; c < c1; ++c) e = max(e, oz_ilogb(ldq_c(X + r * sr + c * sc), fl));
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

template <int S>
__device__ __forceinline__ void oz_slice_unit(const q128 *__restrict__ X, int64_t R, int64_t C, int64_t sr,
                                              int64_t sc, const int32_t *__restrict__ E, int8_t *__restrict__ out,
                                              int64_t Kb, int64_t chunk, int lane);
// One thread per (row block, k block, lane): 16 consecutive k of row 32*rb + lane%32.
// X(r, c) = X[r*sr + c*sc] for r < R, c < C; zero digits elsewhere.
template <int S>
__global__ void __launch_bounds__(256) k_oz_slice(const q128 *__restrict__ X, int64_t R, int64_t C, int64_t sr,
                                                  int64_t sc, const int32_t *__restrict__ E, int8_t *__restrict__ out,
                                                  int64_t Kb, int64_t nthreads) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= nthreads) return;
  oz_slice_unit<S>(X, R, C, sr, sc, E, out, Kb, g >> 6, (int)(g & 63));
}
// One lane's 16 elements (row rb*32 + lane%32, k kb*32 + 16(lane/32)..+15) of fragment position
// chunk = rb * Kb + kb: their S digits into the fragment of each slice.
template <int S>
__device__ __forceinline__ void oz_slice_unit(const q128 *__restrict__ X, int64_t R, int64_t C, int64_t sr,
                                              int64_t sc, const int32_t *__restrict__ E, int8_t *__restrict__ out,
                                              int64_t Kb, int64_t chunk, int lane) {
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
// Output tile (tm, tn) of k pass kp; the caller's workgroup (256 threads) runs it whole.
template <int S, int D0, int D1, int NG>
__device__ __forceinline__ void oz_gemm_tile(const int8_t *__restrict__ As, const int8_t *__restrict__ Bs,
                                             int64_t Kb, int64_t Mp, int64_t Np, int64_t tm, int64_t tn,
                                             int64_t kp, int zbase, uint32_t *__restrict__ part) {
  constexpr int G = (S + NG - 1) / NG;
  constexpr int SLOT = 4 * G * OZ_FRAG;
  constexpr int ND = D1 - D0 + 1;
  constexpr int TMAX = (D1 < S - 1) ? D1 : S - 1;  // highest slice any pair of the range uses
  __shared__ __attribute__((aligned(1024))) int8_t lds[4 * SLOT];
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
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
      copy(h + 3);
      // The stage's fragment pairs (A and B of slice g G + t) are read one group ahead: pair 0 now,
      // pair t + 1 right after the first MFMA of group t, so that every wait before a group (the
      // copies in flight make each wait cover all LDS reads issued so far) finds its pair read
      // during the previous group's MFMAs.  The order is pinned with sched_barrier.
      const int8_t *base = &lds[(int)(h & 3) * SLOT + lane * 16];
      auto rd_pair = [&](int t) {
        const int sl = g * G + t;
        if (t < G && sl < S && sl <= TMAX) {
          fa[sl] = *reinterpret_cast<const i32x4v *>(base + (wr * G + t) * OZ_FRAG);
          fb[sl] = *reinterpret_cast<const i32x4v *>(base + ((2 + wc) * G + t) * OZ_FRAG);
        }
      };
      rd_pair(0);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t2 = 0; t2 < G; ++t2) {
        const int mm = g * G + t2;
        if (mm < S) {
          bool next = true;  // pair t2 + 1 still to be read (folds at compile time)
          auto mfma = [&](int d, const i32x4v &x, const i32x4v &y) {
            acc[d - D0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(x, y, acc[d - D0], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
            if (next) {
              rd_pair(t2 + 1);
              next = false;
            }
            __builtin_amdgcn_sched_barrier(0);
          };
#pragma unroll
          for (int u = 0; u <= mm; ++u) {
            const int d = mm + u;
            if (d >= D0 && d <= D1) mfma(d, fa[mm], fb[u]);
          }
#pragma unroll
          for (int t = 0; t < mm; ++t) {
            const int d = t + mm;
            if (d >= D0 && d <= D1) mfma(d, fa[t], fb[mm]);
          }
          if (next) rd_pair(t2 + 1);  // the group had no MFMA in the diagonal range
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
  }
  // V = sum_{d=D0..D1} acc_d 256^(S-1-d) in OZ_PW two's complement words: diagonal d sits at bit
  // off = 8 (S-1-d), in word w = off / 32 shifted by b = off mod 32.  Its low part (acc << b) is
  // added into lo[w] (carries counted in cy[w]) and its high part (acc >> (32 - b), arithmetic;
  // the sign for b = 0) into hi[w + 1]; one signed carry pass over the words then gives V.  Result
  // element r of this lane is row 8(r/4) + 4(lane/32) + r%4, column lane%32 of the wave's block
  const int64_t row0 = tm * OZ_TILE + wr * 32 + 4 * (lane >> 5);
  const int64_t col = tn * OZ_TILE + wc * 32 + (lane & 31);
  uint32_t *pz = part + (zbase + kp) * Mp * Np * OZ_PREC;
  static_assert(8 * (S - 1 - D0) + 31 + 3 < 32 * OZ_PW, "V fits its words");
#pragma clang loop unroll(full)
  for (int r = 0; r < 16; ++r) {
    uint32_t lo[OZ_PW], cy[OZ_PW];
    int32_t hi[OZ_PW];
#pragma unroll
    for (int i = 0; i < OZ_PW; ++i) {
      lo[i] = 0u;
      cy[i] = 0u;
      hi[i] = 0;
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const int off = 8 * (S - 1 - (D0 + d)), w = off >> 5, bs = off & 31;
      const int32_t x = acc[d][r];
      const uint64_t sum = (uint64_t)lo[w] + (uint32_t)(bs ? ((uint32_t)x << bs) : (uint32_t)x);
      lo[w] = (uint32_t)sum;
      cy[w] += (uint32_t)(sum >> 32);
      if (w + 1 < OZ_PW) hi[w + 1] += bs ? (x >> (32 - bs)) : (x >> 31);
    }
    uint32_t V[OZ_PW];
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < OZ_PW; ++i) {
      const int64_t tv = c + (int64_t)lo[i] + ((int64_t)cy[i] << 32) + (int64_t)hi[i];
      V[i] = (uint32_t)tv;
      c = tv >> 32;
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
    __builtin_amdgcn_sched_barrier(0);  // one entry at a time: its word sums stay in few registers
  }
}

template <int S, int D0, int D1, int NG>
__global__ void __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
k_oz_gemm(const int8_t *__restrict__ As, const int8_t *__restrict__ Bs, int64_t Kb, int64_t Mp, int64_t Np,
          int zbase, uint32_t *__restrict__ part) {
  int64_t tn = blockIdx.x, tm = blockIdx.y;
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
  oz_gemm_tile<S, D0, D1, NG>(As, Bs, Kb, Mp, Np, tm, tn, blockIdx.z, zbase, part);
}

// ------------------------------------------------------------------------------------------
// 4. finish and certification
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool q_finite_nonzero(q128 x) { return q_bexp(x) != 0x7fffu && !q_iszero(x); }

// Entry (i, j): res = alpha * (sum of its pass records) 2^(E_i+F_j-8(S+1)) + beta y rounded once;
// true when res is certified (else the entry goes to the exact repair).  kfac = ceil(log2(K (S+1)/4)).
__device__ __forceinline__ bool oz_finish_value(int64_t i, int64_t j, int S, int kfac, int npass,
                                                const uint32_t *__restrict__ part, int64_t Mp, int64_t Np,
                                                const int32_t *__restrict__ eA, const uint32_t *__restrict__ fA,
                                                const int32_t *__restrict__ eB, const uint32_t *__restrict__ fB,
                                                q128 alpha, q128 beta, q128 yv, q128 &res) {
  const bool ab_ok = (q_bexp(alpha) != 0x7fffu) && (q_bexp(beta) != 0x7fffu) && (q_bexp(yv) != 0x7fffu);
  bool done = false;
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
      int terr;
      res = finish_reg_top<OZ_FINW>(V, lsb, sgn, alpha, beta, yv, terr);
      if (q_finite_nonzero(res)) {
        const int ea = unpack_finite(alpha).e;
        const int bexp = ea + 1 + kfac + Ei + Fj - 8 * S;  // |alpha (AB - P)| < 2^bexp
        // with the epilogue's cut of the pass sum the two bounds together stay below 2^(max + 1)
        const int eb = (terr == INT_MIN) ? bexp : max(bexp, terr) + 1;
        done = eb <= unpack_finite(res).e - 106;
      }
    }
  }
  return done;
}

// One thread per output (i, j), j fastest.
__global__ void __launch_bounds__(256) k_oz_finish(int64_t M, int64_t N, int S, int kfac, int npass,
                                                   const uint32_t *__restrict__ part, int64_t Mp, int64_t Np,
                                                   const int32_t *__restrict__ eA, const uint32_t *__restrict__ fA,
                                                   const int32_t *__restrict__ eB, const uint32_t *__restrict__ fB,
                                                   q128 alpha, q128 beta, q128 *__restrict__ C, int64_t csr,
                                                   int64_t csc, uint32_t *__restrict__ hdr, uint32_t *__restrict__ list) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < M * N; g += stride) {
    const int64_t i = g / N, j = g - i * N;
    q128 *cp = C + i * csr + j * csc;
    q128 res;
    if (oz_finish_value(i, j, S, kfac, npass, part, Mp, Np, eA, fA, eB, fB, alpha, beta, ldq_c(cp), res)) {
      stq_c(cp, res);
    } else {
      const uint32_t slot = atomicAdd(hdr + 1, 1u);
      list[2 * (int64_t)slot] = (uint32_t)i;
      list[2 * (int64_t)slot + 1] = (uint32_t)j;
    }
  }
}

// ------------------------------------------------------------------------------------------
// 5. exact repair of the queued entries
// ------------------------------------------------------------------------------------------
// One workgroup per queued entry (grid-stride over the queue): exact accumulation of
// sum_l A_il B_lj in two windows per lane with a block-uniform anchor placing the largest
// possible product (from the row and column maxima) at the top of the high window, then the
// exact alpha*S + beta*C rounded once.  An entry with a product below the low window is
// recomputed on the full-range path by the same workgroup.
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

// One entry by the whole workgroup: *cp = RN(alpha sum_l ar[l asc] bc[l bsr] + beta *cp).  An entry
// with a product outside the windows of anchor An is left unwritten: its queue index e goes to
// the full-range queue wq (header word 2 counts it) for k_oz_wide / k_bo_wide.
__device__ __forceinline__ void oz_repair_entry(int64_t K, const q128 *__restrict__ ar, int64_t asc,
                                                const q128 *__restrict__ bc, int64_t bsr, int An, q128 alpha,
                                                q128 beta, q128 *__restrict__ cp, uint32_t *__restrict__ hdr,
                                                uint32_t (*s_w)[2 * WW + 1], uint32_t e, uint32_t *__restrict__ wq) {
  uint32_t Wl[WW], Wh[WW], flags;
  oz_block_dot(K, ar, asc, bc, bsr, An, Wl, Wh, flags, s_w);
  if (threadIdx.x == 0) {
    if (flags & FL_LOST) {
      wq[atomicAdd(hdr + 2, 1u)] = e;
    } else {
      uint32_t Mw[SW];
      uint32_t sign;
      s_from_windows(Mw, sign, Wl, Wh);
      stq_c(cp, finish_reg(Mw, An, sign, 0u, flags, true, alpha, beta, ldq_c(cp)));
    }
  }
  __syncthreads();
}

// Queued entries at least this many: B's strided columns are read from a transposed copy (one pass
// over B pays off against ~80 ns per gathered column).
constexpr uint32_t OZ_REPAIR_TRANSPOSE = 2048;

// Bt: null, or the transposed copy of B (N x K row-major) that k_oz_transpose writes when the queue
// holds at least OZ_REPAIR_TRANSPOSE entries; the same count selects it here.
__global__ void __launch_bounds__(OZR_BS) k_oz_repair(int64_t K, const q128 *__restrict__ A, int64_t asr, int64_t asc,
                                                      const q128 *__restrict__ B, int64_t bsr, int64_t bsc,
                                                      const q128 *__restrict__ Bt, const int32_t *__restrict__ eA,
                                                      const int32_t *__restrict__ eB, q128 alpha, q128 beta,
                                                      q128 *__restrict__ C, int64_t csr, int64_t csc,
                                                      uint32_t *__restrict__ hdr, const uint32_t *__restrict__ list,
                                                      uint32_t *__restrict__ wq) {
  __shared__ uint32_t s_w[OZR_BS / 64][2 * WW + 1];
  const uint32_t count = hdr[1];
  const bool useT = Bt != nullptr && count >= OZ_REPAIR_TRANSPOSE;
  const q128 *Br = useT ? Bt : B;
  const int64_t rbsr = useT ? 1 : bsr, rbsc = useT ? K : bsc;
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    const int64_t i = list[2 * (int64_t)e], j = list[2 * (int64_t)e + 1];
    oz_repair_entry(K, A + i * asr, asc, Br + j * rbsc, rbsr, oz_anchor(eA[i], eB[j]), alpha, beta,
                    C + i * csr + j * csc, hdr, s_w, e, wq);
  }
}

// Entries whose products lie too far apart for the repair's windows (queue wq of indices into
// list, header word 2), one workgroup each on the full-range path (wide.hpp); the whole grid
// returns at once when there are none.
constexpr int OZ_WIDE_BLOCKS = 256;
__global__ void __launch_bounds__(256) k_oz_wide(int64_t K, const q128 *__restrict__ A, int64_t asr, int64_t asc,
                                                 const q128 *__restrict__ B, int64_t bsr, int64_t bsc,
                                                 const q128 *__restrict__ Bt, q128 alpha, q128 beta,
                                                 q128 *__restrict__ C, int64_t csr, int64_t csc,
                                                 uint32_t *__restrict__ hdr, const uint32_t *__restrict__ list,
                                                 const uint32_t *__restrict__ wq) {
  __shared__ kul_t Kl;
  const uint32_t nw = hdr[2];
  if (blockIdx.x >= nw) return;
  const uint32_t count = hdr[1];
  const bool useT = Bt != nullptr && count >= OZ_REPAIR_TRANSPOSE;
  const q128 *Br = useT ? Bt : B;
  const int64_t rbsr = useT ? 1 : bsr, rbsc = useT ? K : bsc;
  kul_clear_all(Kl);
  __syncthreads();
  for (uint32_t w = blockIdx.x; w < nw; w += gridDim.x) {
    const uint32_t e = wq[w];
    const int64_t i = list[2 * (int64_t)e], j = list[2 * (int64_t)e + 1];
    q128 *cp = C + i * csr + j * csc;
    const q128 r = wide_item(Kl, K, A + i * asr, asc, Br + j * rbsc, rbsr, true, alpha, beta, ldq_c(cp), hdr);
    if (threadIdx.x == 0) stq_c(cp, r);
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

// out[j * K + l] = B[l * ldb + j] (K x N row-major -> N x K row-major), 32 x 32 tiles via LDS; only
// when the repair queue (header word 1) holds at least OZ_REPAIR_TRANSPOSE entries
__global__ void __launch_bounds__(256) k_oz_transpose(const q128 *__restrict__ B, int64_t K, int64_t N, int64_t ldb,
                                                      const uint32_t *__restrict__ hdr, q128 *__restrict__ out) {
  __shared__ uint4 t[32][33];
  if (hdr[1] < OZ_REPAIR_TRANSPOSE) return;  // uniform: the whole grid skips the copy
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
                             const q128 *B, int64_t n, int64_t bsr, int64_t bsc, const int32_t *eA,
                             const int32_t *eB, char *ws, hipStream_t st) {
  oz_slice_operand<S>(A, m, k, asr, asc, L.Mp, eA, (int8_t *)(ws + L.sA), L.Kb, st);
  // B as rows j, columns l
  oz_slice_operand<S>(B, n, k, bsc, bsr, L.Np, eB, (int8_t *)(ws + L.sB), L.Kb, st);
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

// every launch of a call fits the 32-bit grid size: scaling (up to one wave per row), slicing
// (64 threads per 32 x 32 fragment position; row-contiguous: 4 k blocks per workgroup along x),
// products (Np / 64 workgroups along x)
static bool oz_grids_fit(int64_t Mp, int64_t Np, int64_t Kb) {
  const int64_t R = Mp > Np ? Mp : Np;
  return grid_fits((R + 3) / 4, 256) && grid_fits((R / 32) * Kb * 64 / 256 + 1, 256) &&
         grid_fits((Kb + OZS_KB - 1) / OZS_KB, 256);
}

static void oz_scale(const q128 *X, int64_t R, int64_t Cn, int64_t sr, int64_t sc, int32_t *E, uint32_t *F,
                     hipStream_t st) {
  const int contig = (sc == 1);
  // enough blocks to fill the chip: split the reduced dimension into chunks (about 8192 blocks;
  // 2048 blocks of at least 64 columns took 9 us more for the two operands at 4096^2)
  const int64_t rblocks = contig ? (R + 3) / 4 : (R + 255) / 256;
  int64_t nch = (8192 + rblocks - 1) / rblocks;
  if (nch < 1) nch = 1;
  int64_t cchunk = (Cn + nch - 1) / nch;
  if (cchunk < 16) cchunk = 16;
  nch = (Cn + cchunk - 1) / cchunk;
  dim3 grid((unsigned)rblocks, (unsigned)nch);
  k_oz_scale<<<grid, 256, 0, st>>>(X, R, Cn, sr, sc, contig, cchunk, E, F);
}

// Exact recomputation of the queued entries (header word 1 counts them, list holds (i, j)).  When
// B's columns are strided, a transposed copy of B (at Bt) is made first if the queue is long; the
// copy and the repair read the queue length on the device, so nothing is synchronized.
static hipError_t oz_repair_queue(int64_t m, int64_t n, int64_t k, const q128 *A, int64_t asr, int64_t asc,
                                  const q128 *B, int64_t bsr, int64_t bsc, const int32_t *eA, const int32_t *eB,
                                  q128 alpha, q128 beta, q128 *C, int64_t csr, int64_t csc, uint32_t *hdr,
                                  uint32_t *list, q128 *Bt, hipStream_t st, hipEvent_t *ev) {
  uint32_t *wq = list + 2 * m * n;  // full-range queue after the (i, j) list
  const bool strided = bsr != 1 && n > 1;
  if (strided) {
    dim3 tg((unsigned)((n + 31) / 32), (unsigned)((k + 31) / 32));
    k_oz_transpose<<<tg, 256, 0, st>>>(B, k, n, bsr, hdr, Bt);
  }
  k_oz_repair<<<OZ_REPAIR_BLOCKS, OZR_BS, 0, st>>>(k, A, asr, asc, B, bsr, bsc, strided ? Bt : nullptr, eA, eB,
                                                   alpha, beta, C, csr, csc, hdr, list, wq);
  k_oz_wide<<<OZ_WIDE_BLOCKS, 256, 0, st>>>(k, A, asr, asc, B, bsr, bsc, strided ? Bt : nullptr, alpha, beta, C, csr,
                                            csc, hdr, list, wq);
  if (ev) (void)hipEventRecord(ev[4], st);
  return hipGetLastError();
}

void ozaki_probe_scales(const void *wsv, int64_t m, int64_t n, const int32_t **eA, const uint32_t **fA,
                        const int32_t **eB, const uint32_t **fB);
// Header and scaling exponents of a call: the probe's when it ran on the same operands (probe_ws),
// else computed here into the call's own arrays (*eA.. point there on entry).
static void oz_prepare(int64_t Mp, int64_t Np, int64_t m, int64_t n, int64_t k, const q128 *A, int64_t asr,
                       int64_t asc, const q128 *B, int64_t bsr, int64_t bsc, const void *probe_ws, int32_t **eA,
                       uint32_t **fA, int32_t **eB, uint32_t **fB, uint32_t *hdr, hipStream_t st) {
  if (probe_ws) {
    const int32_t *pa, *pb;
    const uint32_t *qa, *qb;
    ozaki_probe_scales(probe_ws, m, n, &pa, &qa, &pb, &qb);
    *eA = const_cast<int32_t *>(pa);
    *fA = const_cast<uint32_t *>(qa);
    *eB = const_cast<int32_t *>(pb);
    *fB = const_cast<uint32_t *>(qb);
    k_oz_init<<<1, 256, 0, st>>>(*eA, *fA, 0, *eB, *fB, 0, hdr);  // header only
    return;
  }
  const int64_t ninit = (Mp > Np ? Mp : Np) > OZ_HDR / 4 ? (Mp > Np ? Mp : Np) : OZ_HDR / 4;
  k_oz_init<<<(unsigned)((ninit + 255) / 256), 256, 0, st>>>(*eA, *fA, Mp, *eB, *fB, Np, hdr);
  oz_scale(A, m, k, asr, asc, *eA, *fA, st);
  oz_scale(B, n, k, bsc, bsr, *eB, *fB, st);  // B columns: r = j, c = l
}

hipError_t launch_gemm_ozaki(int colmajor, int64_t m, int64_t n, int64_t k, int S, q128 alpha, const q128 *A,
                             int64_t lda, const q128 *B, int64_t ldb, q128 beta, q128 *C, int64_t ldc, void *wsv,
                             hipStream_t st, hipEvent_t *ev, const void *probe_ws, const c_gate *gate) {
  if (S < OZ_SMIN || S > OZ_SMAX) return hipErrorInvalidValue;
  char *ws = (char *)wsv;
  const oz_layout L = oz_plan(m, n, k, S);
  if (!oz_grids_fit(L.Mp, L.Np, L.Kb)) return hipErrorInvalidValue;
  // element strides: A(i,l), B(l,j), C(i,j)
  const int64_t asr = colmajor ? 1 : lda, asc = colmajor ? lda : 1;
  const int64_t bsr = colmajor ? 1 : ldb, bsc = colmajor ? ldb : 1;
  const int64_t csr = colmajor ? 1 : ldc, csc = colmajor ? ldc : 1;
  int32_t *eA = (int32_t *)(ws + L.eA), *eB = (int32_t *)(ws + L.eB);
  uint32_t *fA = (uint32_t *)(ws + L.fA), *fB = (uint32_t *)(ws + L.fB);
  uint32_t *hdr = (uint32_t *)ws;
  if (ev) (void)hipEventRecord(ev[0], st);
  oz_prepare(L.Mp, L.Np, m, n, k, A, asr, asc, B, bsr, bsc, probe_ws, &eA, &fA, &eB, &fB, hdr, st);
  switch (S) {
#define QB_OZ_CASE(s) \
  case s: oz_launch_slices<s>(L, A, m, k, asr, asc, B, n, bsr, bsc, eA, eB, ws, st); break;
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
  if (gate && (e = gate->fn(gate->arg, st)) != hipSuccess) return e;
  k_oz_finish<<<stride_blocks(mn, 256), 256, 0, st>>>(m, n, S, kfac, (int)L.npass, (const uint32_t *)(ws + L.part),
                                                            L.Mp, L.Np, eA, fA, eB, fB, alpha, beta, C, csr, csc, hdr,
                                                            (uint32_t *)(ws + L.list));
  if (ev) (void)hipEventRecord(ev[3], st);
  return oz_repair_queue(m, n, k, A, asr, asc, B, bsr, bsc, eA, eB, alpha, beta, C, csr, csc, hdr,
                         (uint32_t *)(ws + L.list), (q128 *)(ws + L.bt), st, ev);
}

size_t ozaki_probe_ws(int64_t m, int64_t n) {
  return (size_t)OZ_HDR + 256 + (size_t)(m + 64) * 8 + (size_t)(n + 64) * 8 + 1024 + OZ_PROBE_MAX * 4;
}
size_t ozaki_probe_out_offset(int64_t m, int64_t n) {
  return (size_t)OZ_HDR + 256 + (size_t)(m + 64) * 8 + (size_t)(n + 64) * 8 + 1024;
}

hipError_t launch_ozaki_probe(int colmajor, int64_t m, int64_t n, int64_t k, const q128 *A, int64_t lda,
                              const q128 *B, int64_t ldb, int samples, void *wsv, hipStream_t st, int *out) {
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
  if (!grid_fits((ninit + 3) / 4, 256)) return hipErrorInvalidValue;
  k_oz_init<<<(unsigned)((ninit + 255) / 256), 256, 0, st>>>(eA, fA, m, eB, fB, n, hdr);
  oz_scale(A, m, k, asr, asc, eA, fA, st);
  oz_scale(B, n, k, bsc, bsr, eB, fB, st);
  if (samples > OZ_PROBE_MAX) samples = OZ_PROBE_MAX;
  k_oz_probe<<<(unsigned)(samples < 256 ? samples : 256), OZR_BS, 0, st>>>(
      m, n, k, A, asr, asc, B, bsr, bsc, eA, eB, samples, out ? out : (int *)(ws + ozaki_probe_out_offset(m, n)));
  return hipGetLastError();
}

// The probe workspace's scaling exponents (m rows of A, n columns of B), valid until its next use.
void ozaki_probe_scales(const void *wsv, int64_t m, int64_t n, const int32_t **eA, const uint32_t **fA,
                        const int32_t **eB, const uint32_t **fB) {
  const char *ws = (const char *)wsv;
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  *eA = (const int32_t *)(ws + OZ_HDR);
  *fA = (const uint32_t *)(ws + al(OZ_HDR + (size_t)m * 4));
  *eB = (const int32_t *)(ws + al(OZ_HDR + (size_t)m * 4) + al((size_t)m * 4));
  *fB = (const uint32_t *)((const char *)*eB + al((size_t)n * 4));
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

// ==========================================================================================
// Batched OZAKI form (slice method): many independent problems (qblas_gemm_problem descriptors
// in device memory), one slice count S (12..19, one diagonal range) for the whole batch.
// k_bo_scan writes, per problem, the exclusive prefix sums of the counts below (BO_F int64 per
// record, count + 1 records after the header); the host reads the totals (one synchronization)
// and lays out the workspace, and every later launch is grouped over all problems: a work item
// finds its problem by binary search in one field of the records.  Per problem the steps and the
// per-entry result are those of the single-problem form (scaling, slicing, products, certified
// finish, exact repair).
// ==========================================================================================
constexpr int BO_F = OZ_BATCHED_TOTALS;
enum {
  BO_ROWS = 0,  // padded rows of A (exponent arrays eA, fA)
  BO_COLS,      // padded columns of B (eB, fB)
  BO_FRAGA,     // fragment positions of A's slices (units of S KB)
  BO_FRAGB,     // and of B's
  BO_PART,      // pass-sum records
  BO_TILES,     // 64x64 product tiles times k passes
  BO_ENT,       // outputs m n
  BO_SCA,       // scaling items of A: rows (RowMajor) or groups of 64 rows (ColMajor)
  BO_SCB        // of B: groups of 64 columns (RowMajor) or columns (ColMajor)
};

struct bo_geom {
  int64_t m, n, k, Mp, Np, Kb, npk;
  int64_t asr, asc, bsr, bsc, csr, csc;  // element strides of A(i,l), B(l,j), C(i,j)
};

// a problem with m, n or k <= 0 has no work (C unchanged), as in the single-problem routine
__device__ __forceinline__ bo_geom bo_geometry(const qblas_gemm_problem &g, int colmajor) {
  bo_geom G;
  const bool live = g.m > 0 && g.n > 0 && g.k > 0;
  G.m = live ? g.m : 0;
  G.n = live ? g.n : 0;
  G.k = live ? g.k : 0;
  G.Mp = (G.m + OZ_TILE - 1) / OZ_TILE * OZ_TILE;
  G.Np = (G.n + OZ_TILE - 1) / OZ_TILE * OZ_TILE;
  G.Kb = (G.k + 31) / 32;
  G.npk = (G.Kb + OZ_KPASS_BLOCKS - 1) / OZ_KPASS_BLOCKS;
  G.asr = colmajor ? 1 : g.lda;
  G.asc = colmajor ? g.lda : 1;
  G.bsr = colmajor ? 1 : g.ldb;
  G.bsc = colmajor ? g.ldb : 1;
  G.csr = colmajor ? 1 : g.ldc;
  G.csc = colmajor ? g.ldc : 1;
  return G;
}

// largest p in [0, count) with off[p][f] <= t
__device__ __forceinline__ int64_t bo_find(const int64_t *__restrict__ off, int64_t count, int f, int64_t t) {
  int64_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid * BO_F + f] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}
// The same for item t >= t0 of a workgroup whose first item is t0: one search for t0 (workgroup
// uniform), then a forward walk (a workgroup's items rarely span more than two problems).
__device__ __forceinline__ int64_t bo_find_from(const int64_t *__restrict__ off, int64_t count, int f, int64_t t0,
                                                int64_t t) {
  int64_t p = bo_find(off, count, f, t0);
  while (p + 1 < count && off[(p + 1) * BO_F + f] <= t) ++p;
  return p;
}

// off[0][*] = 0, off[p + 1][f] = sum of count f over problems 0..p; one block, 1024 problems a chunk
__global__ void __launch_bounds__(1024) k_bo_scan(const qblas_gemm_problem *__restrict__ P, int64_t count,
                                                  int colmajor, int64_t *__restrict__ off) {
  __shared__ int64_t s[1024];
  int64_t base[BO_F];
#pragma unroll
  for (int f = 0; f < BO_F; ++f) base[f] = 0;
  if (threadIdx.x < BO_F) off[threadIdx.x] = 0;
  for (int64_t b = 0; b < count; b += 1024) {
    const int64_t p = b + threadIdx.x;
    int64_t v[BO_F];
#pragma unroll
    for (int f = 0; f < BO_F; ++f) v[f] = 0;
    if (p < count) {
      const bo_geom G = bo_geometry(P[p], colmajor);
      v[BO_ROWS] = G.Mp;
      v[BO_COLS] = G.Np;
      v[BO_FRAGA] = (G.Mp / 32) * G.Kb;
      v[BO_FRAGB] = (G.Np / 32) * G.Kb;
      v[BO_PART] = G.npk * G.Mp * G.Np;
      v[BO_TILES] = (G.Mp / OZ_TILE) * (G.Np / OZ_TILE) * G.npk;
      v[BO_ENT] = G.m * G.n;
      v[BO_SCA] = colmajor ? (G.m + 63) / 64 : G.m;
      v[BO_SCB] = colmajor ? G.n : (G.n + 63) / 64;
    }
#pragma unroll
    for (int f = 0; f < BO_F; ++f) {
      s[threadIdx.x] = v[f];
      __syncthreads();
      for (int o = 1; o < 1024; o <<= 1) {
        const int64_t t = (threadIdx.x >= o) ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
      }
      if (p < count) off[(p + 1) * BO_F + f] = base[f] + s[threadIdx.x];
      base[f] += s[1023];
      __syncthreads();
    }
  }
}

// Scaling exponents of every problem's A rows (which = 0) or B columns (which = 1): one wave per
// contiguous row (lanes along k), or one wave per 64 strided rows (a lane per row).
// Graph replays reuse the totals of the call that was captured: when the problems' shapes now give
// other totals, header word OZ_HDR_STALE and status bit 1 are set and every later kernel of the
// call returns at once (the planned workspace and grids would not fit them).
struct bo_totals {
  int64_t t[BO_F];
};
__global__ void k_bo_check(const int64_t *__restrict__ off, int64_t count, bo_totals plan, uint32_t *__restrict__ hdr) {
  if (threadIdx.x != 0) return;
  bool same = true;
#pragma unroll
  for (int f = 0; f < BO_F; ++f) same = same && off[count * BO_F + f] == plan.t[f];
  if (!same) {
    hdr[OZ_HDR_STALE] = 1u;
    hdr[0] |= 2u;
  }
}

constexpr int64_t BO_SCALE_BLOCKS = 4096;  // k_bo_scale's grid at most: 16384 waves, each a range of items
__global__ void __launch_bounds__(256) k_bo_scale(const qblas_gemm_problem *__restrict__ P, int64_t count,
                                                  const int64_t *__restrict__ off, int colmajor, int which,
                                                  int64_t items, int32_t *__restrict__ E, uint32_t *__restrict__ F,
                                                  const uint32_t *__restrict__ hdr) {
  if (hdr[OZ_HDR_STALE]) return;  // the problems differ from the plan: nothing is computed
  const int lane = threadIdx.x & 63;
  // each wave takes a contiguous range of items: one problem search, then a forward walk
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t w = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t per_w = (items + W - 1) / W, i0 = w * per_w, i1 = (i0 + per_w < items) ? i0 + per_w : items;
  if (i0 >= i1) return;  // whole waves
  const int fs = which ? BO_SCB : BO_SCA;
  int64_t p = bo_find(off, count, fs, i0);
  for (int64_t it = i0; it < i1; ++it) {
    while (p + 1 < count && off[(p + 1) * BO_F + fs] <= it) ++p;
    const bo_geom G = bo_geometry(P[p], colmajor);
    const q128 *X = (const q128 *)(which ? P[p].B : P[p].A);
    const int64_t R = which ? G.n : G.m, sr = which ? G.bsc : G.asr, sc = which ? G.bsr : G.asc;
    const int64_t ro = off[p * BO_F + (which ? BO_COLS : BO_ROWS)];
    const int64_t u = it - off[p * BO_F + fs];
    const bool contig = which ? colmajor : !colmajor;
    int e = OZ_EZERO;
    uint32_t fl = 0;
    // four (contiguous: one row per wave) or eight (one row per lane) loads in flight per lane
    // before their exponents are taken: the rows are short (k <= 256 takes one round per wave)
    if (contig) {
      const q128 *row = X + u * sr;
      for (int64_t c0 = lane; c0 < G.k; c0 += 256) {
        q128 v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = (c0 + 64 * i < G.k) ? ldq_c(row + (c0 + 64 * i) * sc) : q_zero(0);
#pragma unroll
        for (int i = 0; i < 4; ++i) e = max(e, oz_ilogb(v[i], fl));
      }
      e = wave_max(e);
      fl = wave_or(fl);
      if (lane == 0) {
        E[ro + u] = e;
        F[ro + u] = fl;
      }
    } else {
      const int64_t r = u * 64 + lane;
      if (r < R) {
        for (int64_t c0 = 0; c0 < G.k; c0 += 8) {
          q128 v[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) v[i] = (c0 + i < G.k) ? ldq_c(X + r * sr + (c0 + i) * sc) : q_zero(0);
#pragma unroll
          for (int i = 0; i < 8; ++i) e = max(e, oz_ilogb(v[i], fl));
        }
        E[ro + r] = e;
        F[ro + r] = fl;
      }
    }
  }
}

// slices of every problem's A (which = 0) or B (which = 1): one thread per (fragment position, lane)
template <int S>
__global__ void __launch_bounds__(256) k_bo_slice(const qblas_gemm_problem *__restrict__ P, int64_t count,
                                                  const int64_t *__restrict__ off, int colmajor, int which,
                                                  int64_t nthreads, const int32_t *__restrict__ E,
                                                  int8_t *__restrict__ out, const uint32_t *__restrict__ hdr) {
  if (hdr[OZ_HDR_STALE]) return;
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= nthreads) return;
  const int64_t chunk = g >> 6;
  const int fs = which ? BO_FRAGB : BO_FRAGA;
  const int64_t p = bo_find_from(off, count, fs, ((int64_t)blockIdx.x * 256) >> 6, chunk);
  const bo_geom G = bo_geometry(P[p], colmajor);
  const q128 *X = (const q128 *)(which ? P[p].B : P[p].A);
  const int64_t R = which ? G.n : G.m, sr = which ? G.bsc : G.asr, sc = which ? G.bsr : G.asc;
  const int64_t f0 = off[p * BO_F + fs];
  oz_slice_unit<S>(X, R, G.k, sr, sc, E + off[p * BO_F + (which ? BO_COLS : BO_ROWS)], out + f0 * (int64_t)(S * OZ_FRAG),
                   G.Kb, chunk - f0, (int)(g & 63));
}

// one 64x64 output tile of one k pass per workgroup
template <int S>
__global__ void __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
k_bo_gemm(const qblas_gemm_problem *__restrict__ P, int64_t count, const int64_t *__restrict__ off, int colmajor,
          const int8_t *__restrict__ As, const int8_t *__restrict__ Bs, uint32_t *__restrict__ part,
          const uint32_t *__restrict__ hdr) {
  if (hdr[OZ_HDR_STALE]) return;  // uniform: before any barrier
  const int64_t t = blockIdx.x;
  const int64_t p = bo_find(off, count, BO_TILES, t);
  const bo_geom G = bo_geometry(P[p], colmajor);
  const int64_t *o = off + p * BO_F;
  const int64_t u = t - o[BO_TILES];
  const int64_t TN = G.Np / OZ_TILE, TM = G.Mp / OZ_TILE;
  const int64_t tn = u % TN, tm = (u / TN) % TM, kp = u / (TN * TM);
  oz_gemm_tile<S, 0, S - 1, 2>(As + o[BO_FRAGA] * (int64_t)(S * OZ_FRAG), Bs + o[BO_FRAGB] * (int64_t)(S * OZ_FRAG),
                               G.Kb, G.Mp, G.Np, tm, tn, kp, 0, part + o[BO_PART] * OZ_PREC);
}

// one thread per output of every problem; uncertified entries queue (problem, i, j)
__global__ void __launch_bounds__(256) k_bo_finish(const qblas_gemm_problem *__restrict__ P, int64_t count,
                                                   const int64_t *__restrict__ off, int colmajor, int S, int64_t total,
                                                   const uint32_t *__restrict__ part, const int32_t *__restrict__ eA,
                                                   const uint32_t *__restrict__ fA, const int32_t *__restrict__ eB,
                                                   const uint32_t *__restrict__ fB, uint32_t *__restrict__ hdr,
                                                   uint32_t *__restrict__ list) {
  if (hdr[OZ_HDR_STALE]) return;  // nothing queued: the repair has nothing to do either
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= total) return;
  const int64_t p = bo_find_from(off, count, BO_ENT, (int64_t)blockIdx.x * 256, g);
  const qblas_gemm_problem &pr = P[p];
  const bo_geom G = bo_geometry(pr, colmajor);
  const int64_t *o = off + p * BO_F;
  const int64_t u = g - o[BO_ENT];
  const int64_t i = u / G.n, j = u - i * G.n;
  int kfac = 0;  // ceil(log2(K (S+1) / 4))
  while ((4ll << kfac) < G.k * (S + 1)) ++kfac;
  q128 *cp = (q128 *)pr.C + i * G.csr + j * G.csc;
  q128 res;
  if (oz_finish_value(i, j, S, kfac, (int)G.npk, part + o[BO_PART] * OZ_PREC, G.Mp, G.Np, eA + o[BO_ROWS],
                      fA + o[BO_ROWS], eB + o[BO_COLS], fB + o[BO_COLS], q_make(pr.alpha[0], pr.alpha[1]),
                      q_make(pr.beta[0], pr.beta[1]), ldq_c(cp), res)) {
    stq_c(cp, res);
  } else {
    const uint32_t slot = atomicAdd(hdr + 1, 1u);
    list[4 * (int64_t)slot] = (uint32_t)p;
    list[4 * (int64_t)slot + 1] = (uint32_t)i;
    list[4 * (int64_t)slot + 2] = (uint32_t)j;
  }
}

// exact recomputation of the queued entries, one workgroup per entry (grid-stride)
__global__ void __launch_bounds__(OZR_BS) k_bo_repair(const qblas_gemm_problem *__restrict__ P, int64_t count,
                                                      const int64_t *__restrict__ off, int colmajor,
                                                      const int32_t *__restrict__ eA, const int32_t *__restrict__ eB,
                                                      uint32_t *__restrict__ hdr, const uint32_t *__restrict__ list,
                                                      uint32_t *__restrict__ wq) {
  __shared__ uint32_t s_w[OZR_BS / 64][2 * WW + 1];
  const uint32_t nq = hdr[1];
  for (uint32_t e = blockIdx.x; e < nq; e += gridDim.x) {
    const int64_t p = list[4 * (int64_t)e], i = list[4 * (int64_t)e + 1], j = list[4 * (int64_t)e + 2];
    const qblas_gemm_problem &pr = P[p];
    const bo_geom G = bo_geometry(pr, colmajor);
    const int64_t *o = off + p * BO_F;
    oz_repair_entry(G.k, (const q128 *)pr.A + i * G.asr, G.asc, (const q128 *)pr.B + j * G.bsc, G.bsr,
                    oz_anchor(eA[o[BO_ROWS] + i], eB[o[BO_COLS] + j]), q_make(pr.alpha[0], pr.alpha[1]),
                    q_make(pr.beta[0], pr.beta[1]), (q128 *)pr.C + i * G.csr + j * G.csc, hdr, s_w, e, wq);
  }
}

// the batch's entries on the full-range path (queue wq, header word 2), one workgroup each
__global__ void __launch_bounds__(256) k_bo_wide(const qblas_gemm_problem *__restrict__ P, int colmajor,
                                                 uint32_t *__restrict__ hdr, const uint32_t *__restrict__ list,
                                                 const uint32_t *__restrict__ wq) {
  __shared__ kul_t Kl;
  const uint32_t nw = hdr[2];
  if (blockIdx.x >= nw) return;
  kul_clear_all(Kl);
  __syncthreads();
  for (uint32_t w = blockIdx.x; w < nw; w += gridDim.x) {
    const uint32_t e = wq[w];
    const int64_t p = list[4 * (int64_t)e], i = list[4 * (int64_t)e + 1], j = list[4 * (int64_t)e + 2];
    const qblas_gemm_problem &pr = P[p];
    const bo_geom G = bo_geometry(pr, colmajor);
    q128 *cp = (q128 *)pr.C + i * G.csr + j * G.csc;
    const q128 r = wide_item(Kl, G.k, (const q128 *)pr.A + i * G.asr, G.asc, (const q128 *)pr.B + j * G.bsc, G.bsr,
                             true, q_make(pr.alpha[0], pr.alpha[1]), q_make(pr.beta[0], pr.beta[1]), ldq_c(cp), hdr);
    if (threadIdx.x == 0) stq_c(cp, r);
  }
}

struct bo_layout {
  size_t eA, fA, eB, fB, sA, sB, part, list, total;
};
static size_t bo_records_end(int64_t count) { return (OZ_HDR + (size_t)(count + 1) * BO_F * 8 + 255) / 256 * 256; }
static bo_layout bo_plan(int64_t count, const int64_t *tot, int S) {
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  bo_layout L;
  size_t o = bo_records_end(count);
  L.eA = o;
  o = al(o + (size_t)tot[BO_ROWS] * 4);
  L.fA = o;
  o = al(o + (size_t)tot[BO_ROWS] * 4);
  L.eB = o;
  o = al(o + (size_t)tot[BO_COLS] * 4);
  L.fB = o;
  o = al(o + (size_t)tot[BO_COLS] * 4);
  L.sA = o;
  o = al(o + (size_t)tot[BO_FRAGA] * S * OZ_FRAG);
  L.sB = o;
  o = al(o + (size_t)tot[BO_FRAGB] * S * OZ_FRAG);
  L.part = o;
  o = al(o + (size_t)tot[BO_PART] * OZ_PREC * 4);
  L.list = o;
  o = al(o + (size_t)tot[BO_ENT] * 20);  // (problem, i, j) records, then the full-range queue
  L.total = o;
  return L;
}

static const int BO_SLICES[] = {17, 16, 15, 14, 18, 19};
int ozaki_batched_configs() { return (int)(sizeof(BO_SLICES) / sizeof(BO_SLICES[0])); }
int ozaki_batched_slices(int cfg) { return (cfg >= 0 && cfg < ozaki_batched_configs()) ? BO_SLICES[cfg] : 0; }

size_t gemm_ozaki_batched_scan_ws(int64_t count) { return bo_records_end(count); }
size_t gemm_ozaki_batched_totals_offset(int64_t count) { return OZ_HDR + (size_t)count * BO_F * 8; }

hipError_t launch_gemm_ozaki_batched_scan(int colmajor, int64_t count, const qblas_gemm_problem *P, void *ws,
                                          hipStream_t st) {
  k_bo_scan<<<1, 1024, 0, st>>>(P, count, colmajor, (int64_t *)((char *)ws + OZ_HDR));
  return hipGetLastError();
}

size_t gemm_ozaki_batched_ws(int64_t count, const int64_t *tot, int S) { return bo_plan(count, tot, S).total; }

template <int S>
static void bo_launch(int colmajor, int64_t count, const qblas_gemm_problem *P, const int64_t *tot, const int64_t *off,
                      const bo_layout &L, char *ws, hipStream_t st, bool planned) {
  int32_t *eA = (int32_t *)(ws + L.eA), *eB = (int32_t *)(ws + L.eB);
  uint32_t *fA = (uint32_t *)(ws + L.fA), *fB = (uint32_t *)(ws + L.fB);
  int8_t *sA = (int8_t *)(ws + L.sA), *sB = (int8_t *)(ws + L.sB);
  uint32_t *part = (uint32_t *)(ws + L.part), *hdr = (uint32_t *)ws, *list = (uint32_t *)(ws + L.list);
  const int64_t ninit = tot[BO_ROWS] > tot[BO_COLS] ? tot[BO_ROWS] : tot[BO_COLS];
  const int64_t ni = ninit > OZ_HDR / 4 ? ninit : OZ_HDR / 4;
  k_oz_init<<<(unsigned)((ni + 255) / 256), 256, 0, st>>>(eA, fA, tot[BO_ROWS], eB, fB, tot[BO_COLS], hdr);
  if (planned) {
    bo_totals pt;
    for (int f = 0; f < BO_F; ++f) pt.t[f] = tot[f];
    k_bo_check<<<1, 64, 0, st>>>(off, count, pt, hdr);
  }
  if (tot[BO_SCA] > 0)
    k_bo_scale<<<(unsigned)min((tot[BO_SCA] + 3) / 4, (int64_t)BO_SCALE_BLOCKS), 256, 0, st>>>(P, count, off, colmajor, 0,
                                                                                           tot[BO_SCA], eA, fA, hdr);
  if (tot[BO_SCB] > 0)
    k_bo_scale<<<(unsigned)min((tot[BO_SCB] + 3) / 4, (int64_t)BO_SCALE_BLOCKS), 256, 0, st>>>(P, count, off, colmajor, 1,
                                                                                           tot[BO_SCB], eB, fB, hdr);
  const int64_t ta = tot[BO_FRAGA] * 64, tb = tot[BO_FRAGB] * 64;
  k_bo_slice<S><<<(unsigned)((ta + 255) / 256), 256, 0, st>>>(P, count, off, colmajor, 0, ta, eA, sA, hdr);
  k_bo_slice<S><<<(unsigned)((tb + 255) / 256), 256, 0, st>>>(P, count, off, colmajor, 1, tb, eB, sB, hdr);
  k_bo_gemm<S><<<(unsigned)tot[BO_TILES], 256, 0, st>>>(P, count, off, colmajor, sA, sB, part, hdr);
  k_bo_finish<<<(unsigned)((tot[BO_ENT] + 255) / 256), 256, 0, st>>>(P, count, off, colmajor, S, tot[BO_ENT], part,
                                                                       eA, fA, eB, fB, hdr, list);
  uint32_t *wq = list + 4 * tot[BO_ENT];
  k_bo_repair<<<OZ_REPAIR_BLOCKS, OZR_BS, 0, st>>>(P, count, off, colmajor, eA, eB, hdr, list, wq);
  k_bo_wide<<<OZ_WIDE_BLOCKS, 256, 0, st>>>(P, colmajor, hdr, list, wq);
}

// After launch_gemm_ozaki_batched_scan on the same workspace: tot = the totals it wrote (read back
// by the caller), or, when planned, totals of an earlier call that the scan's are checked against
// on the device (no synchronization: graph capture); ws at least gemm_ozaki_batched_ws(count, tot,
// S) bytes.  Header words of ws: 0 status, 1 entries queued for the exact repair, 2 entries of them
// on the full-range path, OZ_HDR_STALE set when planned totals did not match.
hipError_t launch_gemm_ozaki_batched(int colmajor, int64_t count, const qblas_gemm_problem *P, int S,
                                     const int64_t *tot, void *wsv, hipStream_t st, bool planned) {
  if (tot[BO_ENT] == 0) return hipSuccess;
  // grids: one workgroup per tile, one thread per entry and per (fragment position, lane), one
  // wave per scaling item; queued entries are numbered in 32 bits
  const int64_t ni = tot[BO_ROWS] > tot[BO_COLS] ? tot[BO_ROWS] : tot[BO_COLS];
  if (!grid_fits(tot[BO_TILES], 256) || !grid_fits((tot[BO_ENT] + 255) / 256, 256) ||
      !grid_fits(tot[BO_FRAGA] / 4 + 1, 256) || !grid_fits(tot[BO_FRAGB] / 4 + 1, 256) ||
      !grid_fits((tot[BO_SCA] + 3) / 4, 256) || !grid_fits((tot[BO_SCB] + 3) / 4, 256) ||
      !grid_fits((ni + 255) / 256 + 1, 256) || count > 0xffffffffll)
    return hipErrorInvalidValue;
  char *ws = (char *)wsv;
  const bo_layout L = bo_plan(count, tot, S);
  const int64_t *off = (const int64_t *)(ws + OZ_HDR);
  switch (S) {
#define QB_BO_CASE(s) \
  case s: bo_launch<s>(colmajor, count, P, tot, off, L, ws, st, planned); break;
    QB_BO_CASE(14) QB_BO_CASE(15) QB_BO_CASE(16) QB_BO_CASE(17) QB_BO_CASE(18) QB_BO_CASE(19)
#undef QB_BO_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ==========================================================================================
// CRT method (Ozaki scheme II: K. Ozaki, Y. Uchino, T. Imamura, "Ozaki Scheme II: A
// GEMM-oriented emulation of floating-point matrix multiplication using an integer modular
// technique", 2025).  Same scaling as the slice method, but the scaled integers are carried as
// residues modulo nq pairwise coprime moduli p_q <= 256 instead of as base-256 digits:
//
// 1. N_il = round(a_il 2^t) with a_il = A_il 2^-E_i in (-1/4, 1/4) (|N| < 2^(t-2)), likewise
//    N'_lj for B.  t is any bit count (not a multiple of 8).
// 2. Residues.  r_q(N) = N mod p_q in [-p/2, p/2) as int8: from the balanced base-256 digits
//    d_k of N (the slice method's digits with t bits), r_q = sum_k d_k (256^k mod p_q) mod p_q,
//    the sum over digits done on the matrix cores (v_mfma_i32_32x32x32_i8 with the digits as A
//    operand and the powers 256^k mod p_q as B operand), then reduced.  Stored in MFMA fragment
//    order per modulus.
// 3. Products.  One int8 GEMM per modulus, R_q = (r_q(N) r_q(N')) mod p_q (int32 accumulation is
//    exact for k < 2^17; longer k is reduced every 2048 k blocks).
// 4. Reconstruction.  With M = prod p_q and V_q = (M/p_q) ((M/p_q)^-1 mod p_q), X' = sum_q R_q V_q
//    is congruent to X = sum_l N_il N'_lj mod M; M >= 4 K 2^(2t-4) >= 4|X|, so X = X' - kM with
//    k = round(X'/M).  X' is computed exactly on the matrix cores as sum_b 256^b sum_q R_q v_qb
//    (v_qb: balanced bytes of V_q), then carried into 11 words.
// 5. Finish and repair as in the slice method: alpha X 2^(E_i+F_j-2t) + beta C rounded once, and
//    certified by |AB - X 2^(E_i+F_j-2t)| <= K 2^(E_i+F_j-t-2) (representation error of both
//    operands; nothing is dropped), otherwise recomputed exactly.
// The GEMM count is nq (38 to 47 for t = 130..158 and k = 4096) instead of the S(S+1)/2 slice
// pairs of the slice method at the same precision.
// ==========================================================================================
constexpr int CRT_QALL = 54;       // pairwise coprime moduli <= 256 (prime powers)
constexpr int CRT_QMAX = 47;       // moduli used at most: M < 2^330, |X'| < 47 * 128 M < 2^343
constexpr int CRT_TMIN = OZ_CRT_TMIN;  // bits of the scaled operands at least
constexpr int CRT_TMAX = OZ_CRT_TMAX;  // and at most (|N| < 2^156 fits the 20 digits)
constexpr int CRT_ND = 20;         // balanced base-256 digits of a scaled operand (t <= 158)
constexpr int CRT_MW = 11;         // 32-bit words of M and of the reconstructed integer
constexpr int CRT_TILE = 256;      // output rows and columns of one product workgroup
constexpr int CRT_KCHUNK = 2048;   // k blocks between modular reductions of the accumulators

// descending, so that the first nq moduli have the largest product
__device__ const int32_t crt_dev_moduli[CRT_QALL] = {
    256, 251, 243, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191, 181, 179, 173, 169,
    167, 163, 157, 151, 149, 139, 137, 131, 127, 125, 121, 113, 109, 107, 103, 101, 97,  89,
    83,  79,  73,  71,  67,  61,  59,  53,  49,  47,  43,  41,  37,  31,  29,  23,  19,  17};
static const int32_t crt_host_moduli[CRT_QALL] = {
    256, 251, 243, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191, 181, 179, 173, 169,
    167, 163, 157, 151, 149, 139, 137, 131, 127, 125, 121, 113, 109, 107, 103, 101, 97,  89,
    83,  79,  73,  71,  67,  61,  59,  53,  49,  47,  43,  41,  37,  31,  29,  23,  19,  17};

// Per-call constants, built on the device by k_crt_tables (64 columns: moduli beyond nq are 0).
struct crt_tab {
  int32_t p[64];         // moduli
  float pinv[64];        // 1 / p
  double pinvd[64];      // 1 / p
  int8_t cpow[64][32];   // [q][k] = 256^k mod p_q, balanced
  int8_t vt[64][64];     // [b][q] = balanced base-256 digit b of V_q
  uint32_t mdiv[64];     // ceil(2^28 / p): floor(y / p) = (16 y * mdiv) >> 32 for 0 <= y < 2^20
  int32_t half[64];      // (p - 1) / 2, 128 for p = 256
  int32_t off[64];       // ceil(2^19 / p) p + half: x + off is in [0, 2^20) for the residue sums
  uint32_t M[CRT_MW];    // product of the moduli
  double msc;            // 2^(32 w0) / M
  int32_t w0;            // lowest word whose byte sums enter the estimate of X'/M
  int32_t nq;
};
constexpr size_t CRT_TAB_BYTES = (sizeof(crt_tab) + 255) / 256 * 256;

__global__ void __launch_bounds__(64) k_crt_tables(int nq, crt_tab *__restrict__ T) {
  __shared__ uint32_t sM[CRT_MW];
  const int q = threadIdx.x;
  const int32_t p = (q < nq) ? crt_dev_moduli[q] : 0;
  if (q == 0) {
    uint32_t M[CRT_MW];
#pragma unroll
    for (int w = 0; w < CRT_MW; ++w) M[w] = w == 0 ? 1u : 0u;
    for (int i = 0; i < nq; ++i) {
      uint64_t c = 0;
#pragma unroll
      for (int w = 0; w < CRT_MW; ++w) {
        const uint64_t v = (uint64_t)M[w] * (uint32_t)crt_dev_moduli[i] + c;
        M[w] = (uint32_t)v;
        c = v >> 32;
      }
    }
#pragma unroll
    for (int w = 0; w < CRT_MW; ++w) sM[w] = M[w];
  }
  __syncthreads();
  T->p[q] = p;
  T->pinv[q] = p ? 1.0f / (float)p : 0.0f;
  T->pinvd[q] = p ? 1.0 / (double)p : 0.0;
  T->mdiv[q] = p ? (uint32_t)(((1u << 28) + (uint32_t)p - 1) / (uint32_t)p) : 0u;
  const int32_t hf = p == 256 ? 128 : (p - 1) / 2;
  T->half[q] = p ? hf : 0;
  T->off[q] = p ? (((1 << 19) + p - 1) / p) * p + hf : 0;
  const uint32_t up = p ? (uint32_t)p : 1u;
  uint32_t pw = 1u % up;
  for (int k = 0; k < 32; ++k) {
    int v = (p && k < CRT_ND) ? (int)pw : 0;
    if (p && v >= (p + 1) / 2) v -= p;  // balanced: [-p/2, p/2) (p = 256: 256^0 = 1, else 0)
    T->cpow[q][k] = (int8_t)v;
    pw = (pw * 256u) % up;
  }
  uint32_t V[CRT_MW];
#pragma unroll
  for (int w = 0; w < CRT_MW; ++w) V[w] = 0;
  if (p) {
    // Mq = M / p (exact), a = Mq mod p, V = Mq * a^-1 mod p (< M); 32-bit division on 16-bit
    // limbs (remainders < 256 keep every dividend below 2^24)
    uint32_t Mq[CRT_MW];
    uint32_t rem = 0;
    for (int w = CRT_MW - 1; w >= 0; --w) {
      const uint32_t hi = (rem << 16) | (sM[w] >> 16);
      const uint32_t qh = hi / up;
      rem = hi - qh * up;
      const uint32_t lo = (rem << 16) | (sM[w] & 0xffffu);
      const uint32_t ql = lo / up;
      rem = lo - ql * up;
      Mq[w] = (qh << 16) | ql;
    }
    uint32_t a = 0;
    for (int w = CRT_MW - 1; w >= 0; --w) {
      a = ((a << 16) | (Mq[w] >> 16)) % up;
      a = ((a << 16) | (Mq[w] & 0xffffu)) % up;
    }
    // a^-1 mod p by the extended Euclidean algorithm (gcd(a, p) = 1)
    int r0 = p, r1 = (int)a, s0 = 0, s1 = 1;
    while (r1 != 0) {
      const int qq = r0 / r1, r2 = r0 - qq * r1, s2 = s0 - qq * s1;
      r0 = r1;
      r1 = r2;
      s0 = s1;
      s1 = s2;
    }
    const uint32_t inv = (uint32_t)(s0 < 0 ? s0 + p : s0);
    uint64_t c = 0;
    for (int w = 0; w < CRT_MW; ++w) {
      const uint64_t v = (uint64_t)Mq[w] * inv + c;
      V[w] = (uint32_t)v;
      c = v >> 32;
    }
  }
  int carry = 0;
  for (int b = 0; b < 64; ++b) {
    const int byte = (b < 4 * CRT_MW) ? (int)((V[b >> 2] >> (8 * (b & 3))) & 0xffu) : 0;
    int d = byte + carry;
    carry = d >= 128;
    if (carry) d -= 256;
    T->vt[b][q] = (int8_t)(p ? d : 0);
  }
  if (q == 0) {
    // the byte sums below word w0 (|s_b| <= nq 2^14 < 2^20) move X'/M by less than
    // 2^20 2^(32 w0 - 8) / M <= 2^-14
    int bits = 0;
    double md = 0;
    for (int w = 0; w < CRT_MW; ++w) {
      T->M[w] = sM[w];
      if (sM[w]) bits = 32 * w + 32 - __builtin_clz(sM[w]);
      md += __builtin_ldexp((double)sM[w], 32 * w);
    }
    const int w0 = bits > 27 ? (bits - 27) / 32 : 0;
    T->msc = __builtin_ldexp(1.0, 32 * w0) / md;
    T->w0 = w0;
    T->nq = nq;
  }
}

// W = N + sum_{k<20} 128 256^k with N = round(x 2^(t-Es)) (two's complement, 6 words): byte k of W
// XOR 0x80 is the balanced digit d_k of N (|N| < 2^(t-2) <= 2^156 < the bias).  N rounds the
// magnitude half up.  With x = (-1)^s m 2^(e-112) and sh = e - 112 - Es + t <= 43 (e <= Es - 3,
// t <= 158): T = floor(m 2^(sh+1)) is m 2^44 shifted right by 43 - sh (one shift, no branch), and
// |N| = floor(T / 2) + (T & 1); the sign and the bias are added in one carry chain:
// W = (floor(T / 2) XOR -s) + bias + (s XOR (T & 1)).
__device__ __forceinline__ void crt_scaled(q128 x, int Es, int t, uint32_t (&W)[6]) {
  const bool live = q_bexp(x) != 0x7fffu && !q_iszero(x);
  const uq u = unpack_finite(x);
  const uint64_t ml = live ? u.ml : 0ull, mh = live ? u.mh : 0ull;
  const uint32_t s = live ? u.s : 0u;
  int rs = live ? 43 - (u.e - 112 - Es + t) : 0;
  rs = rs < 0 ? 0 : (rs > 191 ? 191 : rs);
  uint32_t b[6] = {0u, (uint32_t)(ml << 12), (uint32_t)(ml >> 20), (uint32_t)(ml >> 52) | (uint32_t)(mh << 12),
                   (uint32_t)(mh >> 20), 0u};  // m 2^44 (157 bits)
  const int ws = rs >> 5, bs = rs & 31;
#pragma unroll
  for (int step = 4; step >= 1; step >>= 1) {
    const bool on = (ws & step) != 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) b[i] = on ? ((i + step < 6) ? b[i + step] : 0u) : b[i];
  }
  uint32_t T[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) T[i] = (uint32_t)(((((uint64_t)((i + 1 < 6) ? b[i + 1] : 0u)) << 32) | b[i]) >> bs);
  const uint32_t g = T[0] & 1u, smask = 0u - s;
  uint32_t c = s ^ g;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const uint32_t th = (uint32_t)(((((uint64_t)((i + 1 < 6) ? T[i + 1] : 0u)) << 32) | T[i]) >> 1);
    W[i] = addc32(th ^ smask, i < 5 ? 0x80808080u : 0u, c, c);
  }
}

// residue of x (|x| < 2^24) modulo p in [-p/2, p/2] (p odd) or [-128, 127] (p = 256), as a byte
__device__ __forceinline__ uint32_t crt_red_f(int x, int p, float ip) {
  if (p == 256) return (uint32_t)x & 0xffu;
  const float xf = (float)x;
  const float r = __builtin_fmaf(-__builtin_rintf(xf * ip), (float)p, xf);
  return (uint32_t)(int)r & 0xffu;
}
// the same for |x| < 2^31
__device__ __forceinline__ uint32_t crt_red_d(int x, int p, double ip) {
  if (p == 256) return (uint32_t)x & 0xffu;
  const double xd = (double)x;
  const double r = __builtin_fma(-__builtin_rint(xd * ip), (double)p, xd);
  return (uint32_t)(int)r & 0xffu;
}

// residue of one sum x (x + off in [0, 2^20)) modulo p as a balanced byte, in four full-rate
// instructions: Y = 16 (x + off) (one shift-add, off16 = 16 off), q = floor(Y mdiv / 2^32) =
// floor((x + off) / p) by a 24-bit multiply-high, r = x - (q p + hmo) with hmo = half - off (mod
// 2^32), so r = x + off - q p - half in [-128, 127]
__device__ __forceinline__ uint32_t crt_red24(uint32_t x, uint32_t p, uint32_t mdiv, uint32_t hmo, uint32_t off16) {
  const uint32_t y16 = (x << 4) + off16;
  const uint32_t q = (uint32_t)(((uint64_t)(y16 & 0xffffffu) * (mdiv & 0xffffffu)) >> 32);
  uint32_t qp;
  // q < 2^20 / p and p <= 256: q p + hmo is one full-rate 24-bit multiply-add (the compiler
  // otherwise picks a quarter-rate 32- or 64-bit multiply)
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(qp) : "v"(q), "v"(p), "v"(hmo));
  return x - qp;
}

// 16 residue sums x of one modulus -> 16 balanced residues as bytes
__device__ __forceinline__ i32x4v crt_pack16(const i32x16v &d, int p, uint32_t mdiv, uint32_t hmo, uint32_t off16) {
  int r[16];
#pragma unroll
  for (int v = 0; v < 16; ++v) r[v] = (int)crt_red24((uint32_t)d[v], (uint32_t)p, mdiv, hmo, off16);
  i32x4v w;
  uint32_t o[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t t01 = __builtin_amdgcn_perm((uint32_t)r[4 * j + 1], (uint32_t)r[4 * j], 0x0c0c0400u);
    const uint32_t t23 = __builtin_amdgcn_perm((uint32_t)r[4 * j + 3], (uint32_t)r[4 * j + 2], 0x0c0c0400u);
    o[j] = __builtin_amdgcn_perm(t23, t01, 0x05040100u);
  }
  w.x = (int)o[0];
  w.y = (int)o[1];
  w.z = (int)o[2];
  w.w = (int)o[3];
  return w;
}

// Residues of the scaled integers of one operand: rows r < R (rows of A, or columns of B), k < K,
// element (r, k) at X[r sr + k sc], row exponent E[r] (+3).  One wave per fragment position
// (rb, kb) of the Rb x Kb grid; plane q < nq at out + q * plane holds [rb][kb][1 KB], lane L's 16
// bytes = row L%32, k 16(L/32)..+15.  The fragment is done in two halves (k 0..15, then 16..31):
// Phase 1: the half's 512 elements (8 per lane, loads coalesced along the contiguous dimension)
// become balanced digits 0..19 in the wave's LDS region (element (r, c) at r*17 + c).
// Phase 2: MFMA row rho of group g is the element (slot 2g + ((rho>>2)&1), byte 4(rho>>3) + (rho&3))
// of the fragment, with its digits 16h..16h+15 as the A operand of lane rho + 32h; the B operand
// holds 256^k mod p for the lane's modulus column, and the accumulator starts at off_q, so the lane
// holding MFMA output rows 8(v/4) + 4h + v%4 holds slot 2g + h, bytes v = 0..15 as x + off.
// Moduli 32.. come from the swapped product (the powers as the A operand, the digits as B): lane
// (c, h) then holds element c's sums for moduli 32 + 8(v/4) + 4h + v%4, so only the moduli below
// nq are reduced (4 or 8 per lane instead of 16); the four lanes of a quad hold four consecutive
// bytes of one slot, and a byte transpose across the quad (DPP broadcasts) gives each lane the
// dword of one modulus.
// Groups 0..15 use slots 0..31 (k < 16), groups 16..31 slots 32..63.
constexpr int CRT_RS = 17;  // LDS row stride (elements) of the residue kernel
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_crt_residues(const q128 *__restrict__ X, int64_t R, int64_t K, int64_t sr,
                                                      int64_t sc, const int32_t *__restrict__ E, int t,
                                                      const crt_tab *__restrict__ T, int64_t Rb, int64_t Kb,
                                                      int64_t plane, int8_t *__restrict__ out) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lo[4][32 * CRT_RS][4];
  __shared__ uint32_t s_hi[4][32 * CRT_RS];
  __shared__ i32x4v s_cst[32];  // moduli 32..63: p, mdiv, half - off, 16 off
  if (threadIdx.x < 32) {
    const int q = 32 + (int)threadIdx.x;
    i32x4v c;
    c.x = T->p[q];
    c.y = (int)T->mdiv[q];
    c.z = T->half[q] - T->off[q];
    c.w = T->off[q] << 4;
    s_cst[threadIdx.x] = c;
  }
  __syncthreads();  // every wave of the block is still here
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63, n = lane & 31,
            h = lane >> 5;
  const int64_t pos = (int64_t)blockIdx.x * 4 + wave;
  if (pos >= Rb * Kb) return;  // whole waves; only wave-local LDS below
  const int64_t rb = pos / Kb, kb = pos - rb * Kb;
  const bool kcontig = (sc == 1);
  const int nq = T->nq;
  const i32x4v bc0 = *reinterpret_cast<const i32x4v *>(&T->cpow[n][16 * h]);
  const i32x4v bc1 = *reinterpret_cast<const i32x4v *>(&T->cpow[32 + n][16 * h]);
  const int p0 = T->p[n];
  const uint32_t m0 = T->mdiv[n];
  const uint32_t hmo0 = (uint32_t)(T->half[n] - T->off[n]), off16_0 = (uint32_t)T->off[n] << 4;
  // quad byte transpose: lane j of a quad takes byte j of each quad lane's word
  const uint32_t qj = (uint32_t)(lane & 3);
  const uint32_t sel_lo = 0x0c0c0000u | ((4u + qj) << 8) | qj, sel_hi = ((4u + qj) << 24) | (qj << 16) | 0x0c0cu;
  const int64_t st_off = pos * OZ_FRAG + 4 * (n >> 3);
  i32x16v zero;
#pragma unroll
  for (int v = 0; v < 16; ++v) zero[v] = 0;
  int8_t *o0 = out + (int64_t)n * plane + pos * OZ_FRAG;
  const int hh = (n >> 2) & 1, ib = 4 * (n >> 3) + (n & 3);
  // element (r, c), c < 16 of a half: k contiguous -> lanes 16 apart read 4 rows of 16; rows
  // contiguous -> lanes read 32 rows of 2 columns.  The second half's elements are loaded before
  // the first half's digit sums run, so their latency is hidden behind the MFMA work.
  q128 v[8];
  int32_t ev[8];
  auto load_half = [&](int half) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int r = kcontig ? 4 * i + (lane >> 4) : n, c = kcontig ? (lane & 15) : 2 * i + h;
      const int64_t row = rb * 32 + r, col = kb * 32 + 16 * half + c;
      const bool in = row < R && col < K;
      v[i] = in ? ldq_c(X + row * sr + col * sc) : q_zero(0);
      ev[i] = in ? E[row] : OZ_EZERO;
    }
  };
  load_half(0);
#pragma nounroll
  for (int half = 0; half < 2; ++half) {
    {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int r = kcontig ? 4 * i + (lane >> 4) : n, c = kcontig ? (lane & 15) : 2 * i + h;
        uint32_t W[6] = {0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u, 0u};
        if (ev[i] != OZ_EZERO) crt_scaled(v[i], ev[i] + 3, t, W);
        const int e = r * CRT_RS + c;
        i32x4v lo;
        lo.x = (int)(W[0] ^ 0x80808080u);
        lo.y = (int)(W[1] ^ 0x80808080u);
        lo.z = (int)(W[2] ^ 0x80808080u);
        lo.w = (int)(W[3] ^ 0x80808080u);
        *reinterpret_cast<i32x4v *>(&s_lo[wave][e][0]) = lo;
        s_hi[wave][e] = W[4] ^ 0x80808080u;
      }
    }
    if (half == 0) load_half(1);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 2  // two groups per iteration: 1.3% faster on this kernel
    for (int gg = 0; gg < 16; ++gg) {
      const int g = 16 * half + gg;
      const int slot = 2 * g + hh;
      const int e = (slot & 31) * CRT_RS + ib;
      // both reads in every lane, then a select (no divergent branch)
      const i32x4v lo = *reinterpret_cast<const i32x4v *>(&s_lo[wave][e][0]);
      const int hi = (int)s_hi[wave][e];
      i32x4v a;
      a.x = h ? hi : lo.x;
      a.y = h ? 0 : lo.y;
      a.z = h ? 0 : lo.z;
      a.w = h ? 0 : lo.w;
      const i32x16v d0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, bc0, zero, 0, 0, 0);
      if (n < nq) *reinterpret_cast<i32x4v *>(o0 + (2 * g + h) * 16) = crt_pack16(d0, p0, m0, hmo0, off16_0);
      if (nq > 32) {  // uniform
        const i32x16v d1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(bc1, a, zero, 0, 0, 0);
        // this lane's element: slot 2g + ((n >> 2) & 1), byte 4 (n >> 3) + (n & 3)
        int8_t *ob = out + st_off + (2 * g + ((n >> 2) & 1)) * 16;
#pragma unroll
        for (int vg = 0; vg < 2; ++vg) {
          if (32 + 8 * vg >= nq) break;  // uniform
          uint32_t r[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const i32x4v c = s_cst[8 * vg + 4 * h + j];
            r[j] = crt_red24((uint32_t)d1[4 * vg + j], (uint32_t)c.x, (uint32_t)c.y, (uint32_t)c.z, (uint32_t)c.w);
          }
          const uint32_t t01 = __builtin_amdgcn_perm(r[1], r[0], 0x0c0c0400u);
          const uint32_t t23 = __builtin_amdgcn_perm(r[3], r[2], 0x0c0c0400u);
          const uint32_t R = __builtin_amdgcn_perm(t23, t01, 0x05040100u);  // moduli +0..3 of this element
          const uint32_t x0 = (uint32_t)__builtin_amdgcn_mov_dpp((int)R, 0x00, 0xf, 0xf, false);
          const uint32_t x1 = (uint32_t)__builtin_amdgcn_mov_dpp((int)R, 0x55, 0xf, 0xf, false);
          const uint32_t x2 = (uint32_t)__builtin_amdgcn_mov_dpp((int)R, 0xaa, 0xf, 0xf, false);
          const uint32_t x3 = (uint32_t)__builtin_amdgcn_mov_dpp((int)R, 0xff, 0xf, 0xf, false);
          const uint32_t word = __builtin_amdgcn_perm(x1, x0, sel_lo) | __builtin_amdgcn_perm(x3, x2, sel_hi);
          const int qi = 32 + 8 * vg + 4 * h + (int)qj;
          if (qi < nq) *reinterpret_cast<uint32_t *>(ob + (int64_t)qi * plane) = word;
        }
      }
    }
    // the next half overwrites the region: every lane's reads of this half are done
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// One int8 GEMM per modulus: RC[q][mb][nb] = (RA[q] RB[q]^T) mod p_q (planes pA, pB, pC bytes apart)
// for a 256 x 256 tile of (mb, nb) blocks, 8 waves of 128 x 64 (4 x 2 MFMA blocks).  Stages of four
// k blocks (4 x (8 A + 8 B) fragments, 64 KB) alternate between two LDS slots: the copies of stage
// s + 1 (global_load_lds, 8 fragments per wave) are issued right after the barrier of stage s and
// land while its 32 MFMAs per wave run.  Every copy issues its loads (k blocks past the end load
// the last block again into a slot that is never read), so the wait before each barrier is the
// same.  The fragments are software-pipelined in two register sets: a wave enters the barrier of
// stage s still holding the last k block of stage s - 1, so after the barrier its MFMAs on those
// run while the first k block of stage s is read from LDS, with the copies of stage s + 1 issued
// between them; the reads of each next k block are issued between the MFMAs on the current one.
// The order is pinned with sched_barrier.  (Products at 4096^3, 39 moduli, on one box: 2.90 ms
// with two-block stages and reads waited for after every barrier and between MFMA pairs; 2.46 ms
// pipelined; 2.32 ms with four-block stages, half the barriers; 2.29 ms persistent.)
// The grid is one workgroup per CU; workgroup w takes tiles w, w + grid, ... in the order
// (modulus, tile row, tile column), so the tiles of one modulus run together (its residue
// planes stay in the caches), and the last stage of a tile copies the first stage of the next, so
// a tile starts without waiting for a load.  Output block
// (q, mb, nb), 1 KB: bytes v of lane L (rows 8(v/4) + 4(L/32) + v%4, column L%32 of the MFMA
// result) at (v/4) * 256 + L * 4 + v%4, so that the reconstruction reads rows of 4 entries of 64
// lanes contiguously.
constexpr int CRT_WR = 2, CRT_WC = 4, CRT_FM = 4, CRT_FN = 2, CRT_KS = 4;
constexpr int CRT_TA = CRT_WR * CRT_FM, CRT_TB = CRT_WC * CRT_FN;
constexpr int CRT_SLOT = (CRT_TA + CRT_TB) * OZ_FRAG * CRT_KS;
constexpr int CRT_PER = (CRT_TA + CRT_TB) * CRT_KS / (CRT_WR * CRT_WC);
constexpr int CRT_CHS = CRT_KCHUNK / CRT_KS;  // stages between modular reductions
constexpr int CRT_NSLOT = 2;                 // LDS ring slots; copies run CRT_NSLOT - 1 stages ahead
static_assert(CRT_TA * 32 == CRT_TILE && CRT_TB * 32 == CRT_TILE, "tile");
static_assert(((CRT_TA + CRT_TB) * CRT_KS) % (CRT_WR * CRT_WC) == 0 && (CRT_NSLOT - 2) * CRT_PER <= 15, "copy split");
static_assert(CRT_KS % 2 == 0 && CRT_PER <= CRT_FM * CRT_FN, "two fragment sets alternate over a stage's k blocks");
static_assert(CRT_PER == CRT_TA && CRT_TA == CRT_TB && CRT_WR * CRT_WC == 2 * CRT_KS, "a wave copies one operand's blocks of one k block");

__global__ void __launch_bounds__(512) k_crt_gemm(const int8_t *__restrict__ RA, const int8_t *__restrict__ RB,
                                                  int64_t Mb, int64_t Nb, int64_t Kb, int64_t pA, int64_t pB,
                                                  int64_t pC, const crt_tab *__restrict__ T,
                                                  int8_t *__restrict__ RC, int64_t ntiles) {
  __shared__ __attribute__((aligned(1024))) int8_t lds[CRT_NSLOT * CRT_SLOT];
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int64_t tilesN = Nb / CRT_TB, per = (Mb / CRT_TA) * tilesN;
  const int64_t nfull = Kb / CRT_KS;            // full stages of a tile
  const int rem = (int)(Kb - nfull * CRT_KS);  // k blocks of its short last stage (0: none)
  const int64_t ns = nfull + (rem ? 1 : 0);
  auto slot = [](int64_t gs) { return (int)((uint32_t)gs % (uint32_t)CRT_NSLOT) * CRT_SLOT; };
  // Wave w copies the 8 row blocks of A (w even) or the 8 column blocks of B (w odd) of k block
  // w / 2 of each stage: fragment c of the stage's k block kk = w / 2 at LDS offset
  // ((kk (TA + TB) + (w odd) TA + c) KB
  const int kkw = wave >> 1;
  const int dstw = (kkw * (CRT_TA + CRT_TB) + (wave & 1) * CRT_TA) * OZ_FRAG;
  // a tile's copy source for this wave: its A row panel or B column panel
  auto tile_src = [&](int64_t t) -> const int8_t * {
    const int64_t q = t / per, tt = t - q * per, tm = tt / tilesN, tn = tt - tm * tilesN;
    return (wave & 1) ? RB + q * pB + tn * CRT_TB * Kb * OZ_FRAG : RA + q * pA + tm * CRT_TA * Kb * OZ_FRAG;
  };
  // copy fragment c of stage st (k blocks past the end load the last block again) into LDS slot d
  auto copy1 = [&](const int8_t *s, int c, int64_t st, int8_t *d) {
    const int64_t kb = min(st * CRT_KS + kkw, Kb - 1);
    __builtin_amdgcn_global_load_lds((glb_void *)(s + ((int64_t)c * Kb + kb) * OZ_FRAG + lane * 16),
                                     (lds_void *)(d + dstw + c * OZ_FRAG), 16, 0, 0);
  };
  // wait for this wave's copies of the stage about to be read and for its LDS reads of the
  // previous stage, whose slot the copy issued after the barrier overwrites; then for every wave's
  auto waitbar = [&]() {
    __builtin_amdgcn_s_waitcnt(((CRT_NSLOT - 2) * CRT_PER) | (7 << 4));  // vmcnt((NSLOT - 2) PER) lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
  };
  const int wr = wave / CRT_WC, wc = wave % CRT_WC;
  const int aoff = wr * CRT_FM * OZ_FRAG + lane * 16, boff = (CRT_TA + wc * CRT_FN) * OZ_FRAG + lane * 16;
  auto rd = [&](int64_t gs, int kk, i32x4v (&fa)[CRT_FM], i32x4v (&fb)[CRT_FN]) {
    const int8_t *base = &lds[slot(gs) + kk * (CRT_TA + CRT_TB) * OZ_FRAG];
#pragma unroll
    for (int a = 0; a < CRT_FM; ++a) fa[a] = *reinterpret_cast<const i32x4v *>(base + aoff + a * OZ_FRAG);
#pragma unroll
    for (int b = 0; b < CRT_FN; ++b) fb[b] = *reinterpret_cast<const i32x4v *>(base + boff + b * OZ_FRAG);
  };
  i32x16v acc[CRT_FM][CRT_FN];
  auto mm = [&](const i32x4v (&fa)[CRT_FM], const i32x4v (&fb)[CRT_FN]) {
#pragma unroll
    for (int a = 0; a < CRT_FM; ++a)
#pragma unroll
      for (int b = 0; b < CRT_FN; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[a], fb[b], acc[a][b], 0, 0, 0);
  };
  // the MFMAs of (fa, fb) with k block kk of global stage gs read into (ga, gb) between them: the
  // reads are issued after this wave's wait for (fa, fb) and land while the MFMAs run
  auto mm_rd = [&](const i32x4v (&fa)[CRT_FM], const i32x4v (&fb)[CRT_FN], int64_t gs, int kk, i32x4v (&ga)[CRT_FM],
                   i32x4v (&gb)[CRT_FN]) {
    const int8_t *base = &lds[slot(gs) + kk * (CRT_TA + CRT_TB) * OZ_FRAG];
#pragma unroll
    for (int a = 0; a < CRT_FM; ++a)
#pragma unroll
      for (int b = 0; b < CRT_FN; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[a], fb[b], acc[a][b], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        const int i = a * CRT_FN + b;  // read i of 6: the A fragments, then the B fragments
        if (i < CRT_FM)
          ga[i] = *reinterpret_cast<const i32x4v *>(base + aoff + i * OZ_FRAG);
        else if (i < CRT_FM + CRT_FN)
          gb[i - CRT_FM] = *reinterpret_cast<const i32x4v *>(base + boff + (i - CRT_FM) * OZ_FRAG);
        __builtin_amdgcn_sched_barrier(0);
      }
  };
  // the MFMAs of (fa, fb) with the copies of stage st of the tile with sources s into slot d
  // issued between them
  auto mm_cp = [&](const i32x4v (&fa)[CRT_FM], const i32x4v (&fb)[CRT_FN], const int8_t *s, int64_t st, int8_t *d) {
#pragma unroll
    for (int a = 0; a < CRT_FM; ++a)
#pragma unroll
      for (int b = 0; b < CRT_FN; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[a], fb[b], acc[a][b], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        const int c = a * CRT_FN + b;
        if (c < CRT_PER) copy1(s, c, st, d);
        __builtin_amdgcn_sched_barrier(0);
      }
  };
  int64_t t = blockIdx.x;
  if (t >= ntiles) return;
  const int8_t *cur = tile_src(t), *nxt, *cs;
  int64_t gs = 0;  // this workgroup's stage count over its tiles: the LDS slot of a stage
#pragma unroll
  for (int c = 0; c < CRT_PER; ++c) copy1(cur, c, 0, &lds[slot(gs)]);
  i32x4v fa0[CRT_FM], fb0[CRT_FN], fa1[CRT_FM], fb1[CRT_FN];
  for (; t < ntiles; t += gridDim.x) {
    // Each stage issues the copies of the next one: stage st + 1 of this tile, or stage 0 of the
    // workgroup's next tile (so its first stage does not wait for a load), or, after the last
    // tile, this tile's last stage again into the free slot (never read).
    const bool has_next = t + gridDim.x < ntiles;
    nxt = tile_src(has_next ? t + gridDim.x : t);
    const int64_t q = t / per, tt = t - q * per, tm = tt / tilesN, tn = tt - tm * tilesN;
    const int p = T->p[q];
    const double ipd = T->pinvd[q];
#pragma unroll
    for (int a = 0; a < CRT_FM; ++a)
#pragma unroll
      for (int b = 0; b < CRT_FN; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[a][b][r] = 0;
    // copy sources and stage for the stage after st
    auto next_of = [&](int64_t st, int64_t &nst) {
      const bool in = st + 1 < ns;
      cs = in ? cur : nxt;
      nst = in ? st + 1 : (has_next ? 0 : ns - 1);
    };
    // one full stage st (global stage gs): (xa, xb) receives its first k block, (ya, yb) holds the
    // previous stage's last block (none if first) and afterwards this stage's last block
    auto stage = [&](int64_t st, bool first, i32x4v (&xa)[CRT_FM], i32x4v (&xb)[CRT_FN], i32x4v (&ya)[CRT_FM],
                     i32x4v (&yb)[CRT_FN]) {
      int64_t nst;
      next_of(st, nst);
      waitbar();
      rd(gs, 0, xa, xb);
      __builtin_amdgcn_sched_barrier(0);
      if (first) {
#pragma unroll
        for (int c = 0; c < CRT_PER; ++c) copy1(cs, c, nst, &lds[slot(gs + 1)]);
      } else {
        mm_cp(ya, yb, cs, nst, &lds[slot(gs + 1)]);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int kk = 0; kk + 1 < CRT_KS; ++kk) {
        if (kk & 1)
          mm_rd(ya, yb, gs, kk + 1, xa, xb);
        else
          mm_rd(xa, xb, gs, kk + 1, ya, yb);
        __builtin_amdgcn_sched_barrier(0);
      }
      ++gs;
    };
    for (int64_t c0 = 0; c0 < nfull; c0 += CRT_CHS) {
      const int64_t c1 = min(nfull, c0 + CRT_CHS);
      stage(c0, true, fa0, fb0, fa1, fb1);
      for (int64_t st = c0 + 1; st < c1; ++st) stage(st, false, fa0, fb0, fa1, fb1);
      mm(fa1, fb1);
      if (c1 == c0 + CRT_CHS && c1 * CRT_KS < Kb) {
        // |acc| <= 2048 * 32 * 2^14 = 2^30: reduce before it can overflow
#pragma unroll
        for (int a = 0; a < CRT_FM; ++a)
#pragma unroll
          for (int b = 0; b < CRT_FN; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = (int)(int8_t)crt_red_d(acc[a][b][r], p, ipd);
      }
    }
    if (rem) {  // the short last stage
      int64_t nst;
      next_of(nfull, nst);
      waitbar();
#pragma unroll
      for (int c = 0; c < CRT_PER; ++c) copy1(cs, c, nst, &lds[slot(gs + 1)]);
      for (int kk = 0; kk < rem; ++kk) {
        rd(gs, kk, fa0, fb0);
        mm(fa0, fb0);
      }
      ++gs;
    }
#pragma unroll
    for (int a = 0; a < CRT_FM; ++a)
#pragma unroll
      for (int b = 0; b < CRT_FN; ++b) {
        uint32_t u[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int v = 0; v < 16; ++v) u[v >> 2] |= crt_red_d(acc[a][b][v], p, ipd) << (8 * (v & 3));
        const int64_t blk = (tm * CRT_TA + wr * CRT_FM + a) * Nb + tn * CRT_TB + wc * CRT_FN + b;
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4)
          *reinterpret_cast<uint32_t *>(RC + q * pC + blk * OZ_FRAG + r4 * 256 + lane * 4) = u[r4];
      }
    cur = nxt;
  }
  __builtin_amdgcn_s_waitcnt(0);  // the last copies land before the workgroup ends
}

static int crt_cu_count() {
  static int cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

// 4 x 4 byte transpose: byte j of o_e = byte e of i_j (v_perm_b32 selects bytes 0-3 from its
// second source and 4-7 from its first)
__device__ __forceinline__ void crt_t4(uint32_t i0, uint32_t i1, uint32_t i2, uint32_t i3, uint32_t &o0, uint32_t &o1,
                                       uint32_t &o2, uint32_t &o3) {
  const uint32_t a = __builtin_amdgcn_perm(i1, i0, 0x05010400u);  // i0.0 i1.0 i0.1 i1.1
  const uint32_t b = __builtin_amdgcn_perm(i1, i0, 0x07030602u);  // i0.2 i1.2 i0.3 i1.3
  const uint32_t c = __builtin_amdgcn_perm(i3, i2, 0x05010400u);  // i2.0 i3.0 i2.1 i3.1
  const uint32_t d = __builtin_amdgcn_perm(i3, i2, 0x07030602u);  // i2.2 i3.2 i2.3 i3.3
  o0 = __builtin_amdgcn_perm(c, a, 0x05040100u);
  o1 = __builtin_amdgcn_perm(c, a, 0x07060302u);
  o2 = __builtin_amdgcn_perm(d, b, 0x05040100u);
  o3 = __builtin_amdgcn_perm(d, b, 0x07060302u);
}

// RN((-1)^s mag 2^lsb) for msb(mag) = msb >= 114 and a normal result (the caller checks the
// biased exponent lsb + msb + QBIAS is in [1, 0x7ffd]); rbe receives the result's biased exponent
template <int N>
__device__ __forceinline__ q128 round_top(const uint32_t (&mag)[N], int msb, int lsb, uint32_t s, int &rbe) {
  uint32_t T[N];
  const uint32_t sticky = rshr<N>(T, mag, msb - 114);  // leading one at bit 114
  const uint64_t lo = ((uint64_t)T[1] << 32) | T[0];
  const uint64_t hi = ((uint64_t)T[3] << 32) | T[2];
  const uint32_t guard = (uint32_t)(lo >> 1) & 1u;
  const uint32_t st = ((uint32_t)lo & 1u) | sticky;
  uint64_t sl = (lo >> 2) | (hi << 62);
  uint64_t sh = (hi >> 2) & ((1ull << 49) - 1);
  const uint64_t inc = guard & (st | (uint32_t)(sl & 1));
  sl += inc;
  sh += (sl == 0) & inc;
  int be = lsb + msb + QBIAS;
  if (sh >> 49) {  // rounded up to 2^113: the significand is 2^112 one binade higher (sl = 0)
    sh >>= 1;
    be += 1;
  }
  rbe = be;
  return q_make(sl, ((uint64_t)s << 63) | ((uint64_t)be << 48) | (sh & FRAC_HI_MASK));
}

// Reconstruction and finish for one 32 x 32 block of C per wave.  The D^T = Vt R product has the
// entries as columns: for the entries (row 8 rq + 4 h' + e, column n) of the block, lane L
// (n = L%32, kh = L/32) holds the residues of moduli 32 kb + 16 kh + 0..15 (B operand, gathered
// from product lane n + 32 h' and transposed in registers), and receives the byte sums
// s_b = sum_q R_q v_qb for b = 8(v/4) + 4 kh + v%4 (+ 32 for the second row block): 4-byte runs
// m = 2(v/4) + kh (+ 8).  Lanes L and L + 32 swap half their runs so that lane kh = 0 holds all
// runs of entry e and lane kh = 1 all runs of entry e + 1.
// the smallest xbits (|X| < 2^xbits) for which the launcher picks XW words for X: XW = 6 below
// 192 bits, 9 below 288, CRT_MW from there
constexpr int crt_xbits_min(int XW) { return XW == 9 ? 192 : XW == CRT_MW ? 288 : 0; }

// XW: words of X = sum_l N_il N'_lj as two's complement (|X| < K 2^(2t-4) < 2^(32 XW - 1)).
// FAST (alpha = +-1, beta = 0): the entry is X 2^(E_i+F_j-2t) rounded once, certified and stored
// here.  Otherwise X is written to xrec (RW words per entry, row-major) for k_crt_finish_ab.
template <int XW, bool FAST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) k_crt_finish(int64_t M, int64_t N, int64_t K, int t, int kc, int64_t Mb,
                                                    int64_t Nb, int64_t plane, const int8_t *__restrict__ RC,
                                                    const crt_tab *__restrict__ T, const int32_t *__restrict__ eA,
                                                    const uint32_t *__restrict__ fA, const int32_t *__restrict__ eB,
                                                    const uint32_t *__restrict__ fB, q128 alpha, q128 beta,
                                                    q128 *__restrict__ C, int64_t csr, int64_t csc,
                                                    uint32_t *__restrict__ hdr, uint32_t *__restrict__ list,
                                                    uint32_t *__restrict__ xrec) {
  constexpr int RW = (XW + 3) & ~3;
  const int64_t blk = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (blk >= Mb * Nb) return;  // whole waves
  const int lane = threadIdx.x & 63, n = lane & 31, kh = lane >> 5;
  const int64_t bm = blk / Nb, bn = blk - bm * Nb;
  const int nq = T->nq;
  i32x4v aV[2][2];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) aV[mb][kb] = *reinterpret_cast<const i32x4v *>(&T->vt[n + 32 * mb][32 * kb + 16 * kh]);
  const int w0 = T->w0;
  const double msc = T->msc;
  uint32_t Mw[XW];  // the low XW words of M (uniform)
#pragma unroll
  for (int w = 0; w < XW; ++w) Mw[w] = T->M[w];
  const uint32_t alpha_sgn = (uint32_t)(alpha.hi >> 63);
  const int64_t col = bn * 32 + n;
  i32x16v zero;
#pragma unroll
  for (int v = 0; v < 16; ++v) zero[v] = 0;
  // The block's residue words, a row quarter rq at a time: the wave loads plane q (q < nq, one
  // 256-byte row per plane: lane L holds column L%32 of rows 8 rq + 4 (L/32) + 0..3, one byte per
  // row), and one swap of the lane halves per pair of words (q, q + 16) gives lane (n, kh) the words
  // of moduli 16 kh + 0..15 (+ 32 for the second MFMA half) for both row halves hp: Q[hp][kb][j].
  // Moduli 48.. are never used (nq <= 47).
  const int8_t *bbase = RC + blk * OZ_FRAG;  // uniform
  // A row quarter's residue words: a plane beyond nq - 1 reads plane nq - 1 again: its column of
  // the digit table (the MFMA's A operand) is zero, so those words add nothing
  uint32_t w[48];
  auto load_rows = [&](int rq) {
    const int8_t *pq = bbase + rq * 256;  // plane min(q, nq - 1), advanced a plane at a time
    int64_t pstep = plane;
    asm volatile("" : "+s"(pstep));  // opaque here: the 48 plane offsets are not kept across rq
#pragma unroll
    for (int q = 0; q < 48; ++q) {
      w[q] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(pq) + lane);
      if (q + 1 < nq) pq += pstep;
    }
  };
  const bool colin = col < N;
  const uint32_t fBc = (FAST && colin) ? fB[col] : 0u;
  const int32_t eBc = (FAST && colin) ? eB[col] : OZ_EZERO;
  // a row quarter's per-entry inputs of the FAST epilogue: row exponent and flags, top word of C
  int32_t eAr[4], eAn[4] = {OZ_EZERO, OZ_EZERO, OZ_EZERO, OZ_EZERO};
  uint32_t fAr[4], chi[4], fAn[4] = {0u, 0u, 0u, 0u}, chn[4] = {0u, 0u, 0u, 0u};
  auto load_entries = [&](int rq, int32_t (&ea)[4], uint32_t (&fa)[4], uint32_t (&ch)[4]) {
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int64_t row = bm * 32 + 8 * rq + 4 * (it >> 1) + 2 * (it & 1) + kh;
      const bool in = FAST && row < M && colin;
      ea[it] = in ? eA[row] : OZ_EZERO;
      fa[it] = in ? fA[row] : 0u;
      ch[it] = in ? reinterpret_cast<const uint32_t *>(C + row * csr + col * csc)[3] : 0u;
    }
  };
  load_entries(0, eAr, fAr, chi);
  load_rows(0);
  // Each quarter waits once, at its top, for everything loaded so far; its swaps free w, and the
  // next quarter's loads are issued before its reconstruction, so they land meanwhile
#pragma nounroll
  for (int rq = 0; rq < 4; ++rq) {
    uint32_t Q[2][2][16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      // lane (n, 0) keeps q = j of row half 0 and takes q = j of half 1 from lane (n, 1), which
      // keeps q = 16 + j of half 1 and takes q = 16 + j of half 0
      const auto s01 = __builtin_amdgcn_permlane32_swap(w[j], w[16 + j], false, false);
      Q[0][0][j] = s01[0];
      Q[1][0][j] = s01[1];
      const auto s2 = __builtin_amdgcn_permlane32_swap(w[32 + j], 0u, false, false);
      Q[0][1][j] = s2[0];
      Q[1][1][j] = s2[1];
    }
    if (rq < 3) {
      load_entries(rq + 1, eAn, fAn, chn);
      load_rows(rq + 1);
    }
#pragma nounroll
    for (int it = 0; it < 4; ++it) {
      const int hp = it >> 1, e = 2 * (it & 1);
      {
        // this pair's residues: bytes e and e + 1 of the words, transposed so that Ec[kb][x][v]
        // holds moduli 32 kb + 16 kh + 4 v + 0..3 of entry e + x
        const uint32_t sel = e ? 0x07030602u : 0x05010400u;
        uint32_t Ec[2][2][4];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            uint32_t P[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) P[i] = hp ? Q[1][kb][4 * v + i] : Q[0][kb][4 * v + i];
            const uint32_t a = __builtin_amdgcn_perm(P[1], P[0], sel);  // i0.e i1.e i0.e+1 i1.e+1
            const uint32_t c = __builtin_amdgcn_perm(P[3], P[2], sel);  // i2.e i3.e i2.e+1 i3.e+1
            Ec[kb][0][v] = __builtin_amdgcn_perm(c, a, 0x05040100u);
            Ec[kb][1][v] = __builtin_amdgcn_perm(c, a, 0x07060302u);
          }
        i32x16v D[2][2];  // [entry e / e + 1][row block]
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int mb = 0; mb < 2; ++mb) {
            i32x4v b0, b1;
            b0.x = (int)Ec[0][x][0];
            b0.y = (int)Ec[0][x][1];
            b0.z = (int)Ec[0][x][2];
            b0.w = (int)Ec[0][x][3];
            b1.x = (int)Ec[1][x][0];
            b1.y = (int)Ec[1][x][1];
            b1.z = (int)Ec[1][x][2];
            b1.w = (int)Ec[1][x][3];
            D[x][mb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(aV[mb][0], b0, zero, 0, 0, 0);
            D[x][mb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(aV[mb][1], b1, D[x][mb], 0, 0, 0);
          }
        // Lanes kh = 0 and 1 hold the byte sums s_b of entries e and e + 1 at complementary
        // positions b = 8 (v/4) + 4 kh + 32 mb + v%4 (D[x][mb][v] for entry e + x).  One swap of
        // the lane halves per register pair gives every lane all sums of its own entry e + kh:
        // then s_b = D[(b/4)%2][b/32][4 ((b%32)/8) + b%4].
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const auto sw = __builtin_amdgcn_permlane32_swap(D[0][mb][v], D[1][mb][v], false, false);
            D[0][mb][v] = (int)sw[0];
            D[1][mb][v] = (int)sw[1];
          }
#define QB_SB(b) D[((b) >> 2) & 1][(b) >> 5][(((b)&31) >> 3) * 4 + ((b)&3)]
        // k = round(X'/M), X' = sum_b s_b 256^b: the sums of words >= WLO, each word's four sums
        // combined exactly in double (below 2^45), scaled by 2^(32 (w - w0)) and the sum by
        // 2^(32 w0)/M, give X'/M within 2^-14
        // (the words below w0 add less than 2^-14 M, so starting at WLO <= w0 only adds
        // precision; each term is below 2^29 M: V_q < M has no digit above byte
        // ceil(log2(M) / 8)); X'/M is within 1/4 of k (|X| <= M/4).  WLO is a lower bound of w0
        // for the XW bucket (a static word range, no per-word select): M >= K 2^(2t-2) gives
        // bits(M) >= xbits + 1, so w0 = (bits(M) - 27) / 32 >= (crt_xbits_min(XW) + 1 - 27) / 32
        constexpr int WLO = (crt_xbits_min(XW) + 1 - 27) / 32;
        double top = 0.0;
#pragma unroll
        for (int w = CRT_MW - 1; w >= WLO; --w) {
          const double r = __builtin_fma((double)QB_SB(4 * w + 3), 16777216.0,
                                         __builtin_fma((double)QB_SB(4 * w + 2), 65536.0,
                                                       __builtin_fma((double)QB_SB(4 * w + 1), 256.0,
                                                                     (double)QB_SB(4 * w))));
          top += __builtin_ldexp(r, 32 * (w - w0));
        }
        const int nk = -(int)__builtin_rint(top * msc);  // -k, |k| <= 128 nq < 2^13
        // X = X' - k M (its low XW words; |X| < 2^(32 XW - 1)), a word at a time: the byte sums
        // (|s_b| < 2^20) in pairs minus k times the 16-bit halves of M's word stay below 2^30 in 32
        // bits; the halves are combined and carried in 64 bits
        uint32_t X[CRT_MW];
        int64_t c = 0;
#pragma unroll
        for (int w = 0; w < XW; ++w) {
          const int lo = __mul24(nk, (int)(Mw[w] & 0xffffu)) + QB_SB(4 * w) + (int)((uint32_t)QB_SB(4 * w + 1) << 8);
          const int hi = __mul24(nk, (int)(Mw[w] >> 16)) + QB_SB(4 * w + 2) + (int)((uint32_t)QB_SB(4 * w + 3) << 8);
          const int64_t v = (int64_t)lo + (int64_t)((uint64_t)(int64_t)hi << 16) + c;
          X[w] = (uint32_t)v;
          c = v >> 32;
        }
#undef QB_SB
        const int64_t row = bm * 32 + 8 * rq + 4 * hp + e + kh;
        if (row < M && col < N) {
          if (!FAST) {
            uint32_t *xr = xrec + (row * N + col) * RW;
#pragma unroll
            for (int w = 0; w < RW; w += 4) {
              i32x4v o;
              o.x = (int)(w < XW ? X[w] : 0u);
              o.y = (int)(w + 1 < XW ? X[w + 1] : 0u);
              o.z = (int)(w + 2 < XW ? X[w + 2] : 0u);
              o.w = (int)(w + 3 < XW ? X[w + 3] : 0u);
              *reinterpret_cast<i32x4v *>(xr + w) = o;
            }
          } else {
            const uint32_t sgn = X[XW - 1] >> 31;
            uint32_t mag[XW];
            uint32_t cc = sgn;
#pragma unroll
            for (int w = 0; w < XW; ++w) mag[w] = addc32(sgn ? ~X[w] : X[w], 0u, cc, cc);
            q128 *cp = C + row * csr + col * csc;
            bool done = false;
            q128 res;
            // beta = 0: beta * y is a zero unless y is Inf or NaN (then the exact repair decides).
            // The entry is certified when |alpha (AB - P)| < 2^bexp, bexp = 1 + kc + Ei + Fj - t - 2
            // (|alpha| = 1), is at most 2^-106 of the result, whose exponent is Ei + Fj - 2t + msb
            // (+1 when rounding carries): that needs msb >= kc + t + 104, so a smaller |X| goes to
            // the repair without rounding, and the rounding never needs a left shift.  A result
            // outside the normal range is left to the repair too.
            const int msb = rmsb<XW>(mag);
            int32_t eAi = eAr[0];
            uint32_t fAi = fAr[0], chii = chi[0];
#pragma unroll
            for (int i = 1; i < 4; ++i)
              if (it == i) {
                eAi = eAr[i];
                fAi = fAr[i];
                chii = chi[i];
              }
            if (!((fAi | fBc) & OZF_NONFINITE) && ((chii >> 16) & 0x7fffu) != 0x7fffu && eAi != OZ_EZERO &&
                eBc != OZ_EZERO && msb >= kc + t + 104) {
              const int Ei = eAi + 3, Fj = eBc + 3;
              const int lsb = Ei + Fj - 2 * t;
              const int be = lsb + msb + QBIAS;  // biased exponent before a rounding carry
              if (be >= 1 && be <= 0x7ffd) {
                int rbe;
                res = round_top<XW>(mag, msb, lsb, sgn ^ alpha_sgn, rbe);
                done = 1 + kc + Ei + Fj - t - 2 <= rbe - QBIAS - 106;
              }
            }
            if (done) {
              stq_c(cp, res);
            } else {
              const uint32_t slot = atomicAdd(hdr + 1, 1u);
              list[2 * (int64_t)slot] = (uint32_t)row;
              list[2 * (int64_t)slot + 1] = (uint32_t)col;
            }
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      eAr[i] = eAn[i];
      fAr[i] = fAn[i];
      chi[i] = chn[i];
    }
  }
}

// General alpha, beta: entry (i, j) = alpha X 2^(E_i+F_j-2t) + beta C rounded once from the record
// of X (two's complement, XW words), certified as in k_oz_finish, else queued for the repair.
template <int XW>
__global__ void __launch_bounds__(256) k_crt_finish_ab(int64_t M, int64_t N, int t, int kc,
                                                       const uint32_t *__restrict__ xrec,
                                                       const int32_t *__restrict__ eA, const uint32_t *__restrict__ fA,
                                                       const int32_t *__restrict__ eB, const uint32_t *__restrict__ fB,
                                                       q128 alpha, q128 beta, q128 *__restrict__ C, int64_t csr,
                                                       int64_t csc, uint32_t *__restrict__ hdr,
                                                       uint32_t *__restrict__ list) {
  constexpr int RW = (XW + 3) & ~3;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < M * N; g += stride) {
    const int64_t i = g / N, j = g - i * N;
    q128 *cp = C + i * csr + j * csc;
    const q128 yv = ldq_c(cp);
    const bool ab_ok = (q_bexp(alpha) != 0x7fffu) && (q_bexp(beta) != 0x7fffu) && (q_bexp(yv) != 0x7fffu);
    bool done = false;
    q128 res;
    const bool fin = !((fA[i] | fB[j]) & OZF_NONFINITE);
    if (ab_ok && fin && q_iszero(alpha)) {
      res = mul_packed(beta, yv);
      done = !q_iszero(res);
    } else if (ab_ok && fin && eA[i] != OZ_EZERO && eB[j] != OZ_EZERO) {
      uint32_t X[XW];
      const uint32_t *xr = xrec + g * RW;
#pragma unroll
      for (int w = 0; w < RW; w += 4) {
        const i32x4v v = *reinterpret_cast<const i32x4v *>(xr + w);
        if (w < XW) X[w] = (uint32_t)v.x;
        if (w + 1 < XW) X[w + 1] = (uint32_t)v.y;
        if (w + 2 < XW) X[w + 2] = (uint32_t)v.z;
        if (w + 3 < XW) X[w + 3] = (uint32_t)v.w;
      }
      const uint32_t sgn = X[XW - 1] >> 31;
      if (sgn) {
        uint32_t c = 1;
#pragma unroll
        for (int w = 0; w < XW; ++w) X[w] = addc32(~X[w], 0u, c, c);
      }
      if (!rzero<XW>(X)) {
        const int Ei = eA[i] + 3, Fj = eB[j] + 3;
        int terr;
        res = finish_reg_top<XW>(X, Ei + Fj - 2 * t, sgn, alpha, beta, yv, terr);
        if (q_finite_nonzero(res)) {
          const int bexp = unpack_finite(alpha).e + 1 + kc + Ei + Fj - t - 2;
          const int eb = (terr == INT_MIN) ? bexp : max(bexp, terr) + 1;  // with the epilogue's cut
          done = eb <= unpack_finite(res).e - 106;
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
}

struct crt_layout {
  int64_t Mp, Np, Kb, Mb, Nb;
  int64_t pA, pB, pC;  // byte strides of the per-modulus planes of RA, RB, RC
  int nq;
  size_t eA, fA, eB, fB, tab, rA, rB, rC, xrec, list, bt, total;
};

static int ceil_log2(int64_t k) {
  int c = 0;
  while (c < 62 && ((int64_t)1 << c) < k) ++c;
  return c;
}

// moduli for t bits at depth k: M >= K 2^(2t-2) (so that |X| <= M/4); 0 when more than CRT_QMAX
int crt_moduli_for(int64_t k, int t) {
  const double need = 2.0 * t - 2.0 + ceil_log2(k < 1 ? 1 : k);
  double bits = 0;
  for (int q = 0; q < CRT_QMAX; ++q) {
    bits += __builtin_log2((double)crt_host_moduli[q]);
    if (bits >= need) return q + 1;
  }
  return 0;
}

constexpr int64_t CRT_PLANE_PAD = 4096;
static crt_layout crt_plan(int64_t m, int64_t n, int64_t k, int nq) {
  crt_layout L;
  L.Mp = (m + CRT_TILE - 1) / CRT_TILE * CRT_TILE;
  L.Np = (n + CRT_TILE - 1) / CRT_TILE * CRT_TILE;
  L.Kb = (k + 31) / 32;
  L.Mb = L.Mp / 32;
  L.Nb = L.Np / 32;
  L.nq = nq;
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
  L.tab = o;
  o = al(o + CRT_TAB_BYTES);
  // Planes a power of two apart (16 MB at 4096^2) put the same offset of every modulus on the same
  // memory channels: the residue kernel's stores (one offset in 32 planes per instruction) ran at
  // 2.0 TB/s that way and at 5.0 TB/s with 4 KB between the planes' starting channels.
  L.pA = L.Mb * L.Kb * OZ_FRAG + CRT_PLANE_PAD;
  L.pB = L.Nb * L.Kb * OZ_FRAG + CRT_PLANE_PAD;
  L.pC = L.Mb * L.Nb * OZ_FRAG + CRT_PLANE_PAD;
  L.rA = o;
  o = al(o + (size_t)nq * L.pA);
  L.rB = o;
  o = al(o + (size_t)nq * L.pB);
  L.rC = o;
  o = al(o + (size_t)nq * L.pC);
  L.xrec = o;  // records of X for general alpha, beta (12 words per entry at most)
  o = al(o + (size_t)m * n * 48);
  L.list = o;
  o = al(o + (size_t)m * n * 12);  // (i, j) pairs, then the full-range queue
  L.bt = o;
  o = al(o + (size_t)n * k * 16);
  L.total = o;
  return L;
}

size_t gemm_crt_ws(int64_t m, int64_t n, int64_t k, int t) {
  const int nq = crt_moduli_for(k, t);
  return nq ? crt_plan(m, n, k, nq).total : 0;
}

// bits t for a probe value r (as ozaki_slices_for): the smallest t with 1 + (kc - 2) - t + 2 <=
// r - 106, raised to the largest t the same moduli carry (precision at no cost: fewer repairs);
// 0 when that needs more than CRT_TMAX bits or CRT_QMAX moduli (the slice method then)
int ozaki_crt_bits_for(int64_t k, int r) {
  const int kc = ceil_log2(k < 1 ? 1 : k);
  int t = (r == 0x7fffffff) ? 136 : kc + 107 - r;
  if (t < CRT_TMIN) t = CRT_TMIN;
  const int nq = t <= CRT_TMAX ? crt_moduli_for(k, t) : 0;
  if (!nq) return 0;
  while (t < CRT_TMAX && crt_moduli_for(k, t + 1) == nq) ++t;
  return t;
}

size_t crt_tab_bytes() { return CRT_TAB_BYTES; }
hipError_t launch_crt_tables(int nq, void *out, hipStream_t st) {
  if (nq < 1 || nq > CRT_QALL) return hipErrorInvalidValue;
  k_crt_tables<<<1, 64, 0, st>>>(nq, (crt_tab *)out);
  return hipGetLastError();
}

hipError_t launch_gemm_crt(int colmajor, int64_t m, int64_t n, int64_t k, int t, q128 alpha, const q128 *A,
                           int64_t lda, const q128 *B, int64_t ldb, q128 beta, q128 *C, int64_t ldc, void *wsv,
                           hipStream_t st, hipEvent_t *ev, const void *probe_ws, const c_gate *gate,
                           const void *tab_in) {
  if (t < CRT_TMIN || t > CRT_TMAX) return hipErrorInvalidValue;
  const int nq = crt_moduli_for(k, t);
  if (!nq) return hipErrorInvalidValue;
  char *ws = (char *)wsv;
  const crt_layout L = crt_plan(m, n, k, nq);
  const int64_t asr = colmajor ? 1 : lda, asc = colmajor ? lda : 1;
  const int64_t bsr = colmajor ? 1 : ldb, bsc = colmajor ? ldb : 1;
  const int64_t csr = colmajor ? 1 : ldc, csc = colmajor ? ldc : 1;
  int32_t *eA = (int32_t *)(ws + L.eA), *eB = (int32_t *)(ws + L.eB);
  uint32_t *fA = (uint32_t *)(ws + L.fA), *fB = (uint32_t *)(ws + L.fB);
  uint32_t *hdr = (uint32_t *)ws;
  crt_tab *tab = tab_in ? (crt_tab *)tab_in : (crt_tab *)(ws + L.tab);
  int8_t *RA = (int8_t *)(ws + L.rA), *RB = (int8_t *)(ws + L.rB), *RC = (int8_t *)(ws + L.rC);
  const int64_t wgs = (int64_t)nq * (L.Mb / CRT_TA) * (L.Nb / CRT_TB);
  if (!oz_grids_fit(L.Mp, L.Np, L.Kb) || !grid_fits((L.Mb * L.Kb + 3) / 4, 256) ||
      !grid_fits((L.Nb * L.Kb + 3) / 4, 256) || !grid_fits(wgs, 512) || !grid_fits((L.Mb * L.Nb + 3) / 4, 256))
    return hipErrorInvalidValue;
  if (ev) (void)hipEventRecord(ev[0], st);
  oz_prepare(L.Mp, L.Np, m, n, k, A, asr, asc, B, bsr, bsc, probe_ws, &eA, &fA, &eB, &fB, hdr, st);
  if (!tab_in) k_crt_tables<<<1, 64, 0, st>>>(nq, tab);
  k_crt_residues<<<(unsigned)((L.Mb * L.Kb + 3) / 4), 256, 0, st>>>(A, m, k, asr, asc, eA, t, tab, L.Mb, L.Kb, L.pA, RA);
  k_crt_residues<<<(unsigned)((L.Nb * L.Kb + 3) / 4), 256, 0, st>>>(B, n, k, bsc, bsr, eB, t, tab, L.Nb, L.Kb, L.pB, RB);
  if (ev) (void)hipEventRecord(ev[1], st);
  k_crt_gemm<<<(unsigned)min(wgs, (int64_t)crt_cu_count()), 512, 0, st>>>(RA, RB, L.Mb, L.Nb, L.Kb, L.pA, L.pB, L.pC, tab,
                                                                     RC, wgs);
  if (ev) (void)hipEventRecord(ev[2], st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int kc = ceil_log2(k < 1 ? 1 : k);
  const int xbits = kc + 2 * t - 3;  // |X| < K 2^(2t-4) < 2^xbits; two's complement needs xbits + 1
  const unsigned fgrid = (unsigned)((L.Mb * L.Nb + 3) / 4);
  uint32_t *list = (uint32_t *)(ws + L.list), *xrec = (uint32_t *)(ws + L.xrec);
  const bool fast = q_iszero(beta) && alpha.lo == 0 && (alpha.hi << 1) == (0x3fffull << 49);
  const int64_t mn = m * n;
  if (gate && (e = gate->fn(gate->arg, st)) != hipSuccess) return e;
#define QB_CRT_FIN(XW)                                                                                              \
  if (fast) {                                                                                                     \
    k_crt_finish<XW, true><<<fgrid, 256, 0, st>>>(m, n, k, t, kc, L.Mb, L.Nb, L.pC, RC, tab, eA, fA, eB, fB, alpha, beta, \
                                                  C, csr, csc, hdr, list, xrec);                                  \
  } else {                                                                                                        \
    k_crt_finish<XW, false><<<fgrid, 256, 0, st>>>(m, n, k, t, kc, L.Mb, L.Nb, L.pC, RC, tab, eA, fA, eB, fB, alpha,     \
                                                   beta, C, csr, csc, hdr, list, xrec);                           \
    k_crt_finish_ab<XW><<<stride_blocks(mn, 256), 256, 0, st>>>(m, n, t, kc, xrec, eA, fA, eB, fB, alpha,    \
                                                                      beta, C, csr, csc, hdr, list);              \
  }
  if (xbits < crt_xbits_min(9)) {
    QB_CRT_FIN(6)
  } else if (xbits < crt_xbits_min(CRT_MW)) {
    QB_CRT_FIN(9)
  } else {
    QB_CRT_FIN(CRT_MW)
  }
#undef QB_CRT_FIN
  if (ev) (void)hipEventRecord(ev[3], st);
  return oz_repair_queue(m, n, k, A, asr, asc, B, bsr, bsc, eA, eB, alpha, beta, C, csr, csc, hdr,
                         (uint32_t *)(ws + L.list), (q128 *)(ws + L.bt), st, ev);
}

}  // namespace qbk
