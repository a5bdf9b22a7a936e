// Host run of the OZAKI finish's epilogue with general alpha, beta (qblas_amd/csrc/exact.hpp):
// finish_reg_top (the sum cut to its top 128 bits before the product with alpha) against
// finish_reg_t (alpha * M + beta * y over all of M, rounded once), on random and cancelling cases.
// When nothing was cut the two must agree bit for bit; otherwise they may differ by at most one
// unit in the last place of the larger plus the reported bound 2^terr (checked in __float128,
// where the difference of the two nearby results is exact).  Usage: epilogue_host_check <seed> <cases>
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../qblas_amd/csrc/exact.hpp"

using namespace qb;

static uint64_t s_state;
static uint64_t rnd() {
  s_state ^= s_state << 13;
  s_state ^= s_state >> 7;
  s_state ^= s_state << 17;
  return s_state;
}
static int rint_in(int lo, int hi) { return lo + (int)(rnd() % (uint64_t)(hi - lo + 1)); }

// a finite binary128 with a random significand and biased exponent in [lo, hi] (0: subnormal)
static q128 rq(int lo, int hi) {
  const uint64_t be = (uint64_t)rint_in(lo, hi);
  const uint64_t hi_w = ((rnd() & 1) << 63) | (be << 48) | (rnd() & 0xffffffffffffull);
  return q_make(rnd(), hi_w);
}
static __float128 f128(q128 v) {
  __float128 f;
  memcpy(&f, &v, 16);
  return f;
}
static q128 from_f128(__float128 f) {
  q128 v;
  memcpy(&v, &f, 16);
  return v;
}

int main(int argc, char **argv) {
  const uint64_t seed = argc > 1 ? strtoull(argv[1], 0, 10) : 1;
  const long cases = argc > 2 ? strtol(argv[2], 0, 10) : 100000;
  s_state = 0x9e3779b97f4a7c15ull * (seed + 1);
  constexpr int N = 9;
  long mism = 0, cut = 0, exact_eq = 0;
  for (long c = 0; c < cases; ++c) {
    uint32_t M[N];
    const int msb = rint_in(0, 32 * N - 1);
    for (int i = 0; i < N; ++i) {
      const int lo = 32 * i;
      uint32_t w = (uint32_t)rnd();
      if (lo > msb) w = 0;
      else if (msb - lo < 31) w &= (2u << (msb - lo)) - 1u;
      if (rnd() % 8 == 0) w = (lo <= msb) ? 0xffffffffu & ((msb - lo < 31) ? (2u << (msb - lo)) - 1u : ~0u) : 0u;
      M[i] = w;
    }
    M[msb >> 5] |= 1u << (msb & 31);
    const int lsb = rint_in(-16300, 16200) - msb;
    const uint32_t sS = (uint32_t)(rnd() & 1);
    const int kind = (int)(rnd() % 6);
    q128 alpha = rq(16383 - 40, 16383 + 40), beta = rq(16383 - 40, 16383 + 40), y = rq(1, 0x7ffe);
    if (kind == 0) beta = q_zero(0);                                             // beta * y = 0
    if (kind == 1) alpha = q_make(0, ((rnd() & 1) << 63) | 0x3fff000000000000ull);  // alpha = +-1
    if (kind == 2) alpha = rq(0, 0);                                             // subnormal alpha
    const q128 s_only = finish_reg_t<N>(M, lsb, sS, 0u, 0u, false, q_one(), q_zero(0), q_zero(0));
    if (kind >= 3 && q_bexp(s_only) != 0x7fffu) {
      // beta * y close to -alpha * S: cancellation in the sum
      const __float128 target = -f128(alpha) * f128(s_only);
      if (kind == 3) {
        beta = q_one();
        y = from_f128(target);
      } else if (kind == 4) {
        beta = from_f128(target / f128(y));
      }
    }
    if (q_bexp(alpha) == 0x7fffu || q_bexp(beta) == 0x7fffu || q_bexp(y) == 0x7fffu) continue;
    if (q_iszero(alpha)) continue;  // the finish handles alpha = 0 before the epilogue
    int terr;
    const q128 r_full = finish_reg_t<N>(M, lsb, sS, 0u, 0u, true, alpha, beta, y);
    const q128 r_top = finish_reg_top<N>(M, lsb, sS, alpha, beta, y, terr);
    if (terr == INT_MIN) {
      ++exact_eq;
      if (r_full.lo != r_top.lo || r_full.hi != r_top.hi) {
        if (++mism <= 10) printf("uncut mismatch case %ld kind %d msb %d lsb %d\n", c, kind, msb, lsb);
      }
      continue;
    }
    ++cut;
    if (q_bexp(r_full) == 0x7fffu || q_bexp(r_top) == 0x7fffu) {
      if (r_full.lo != r_top.lo || r_full.hi != r_top.hi) {
        if (++mism <= 10) printf("overflow mismatch case %ld kind %d\n", c, kind);
      }
      continue;
    }
    const __float128 a = f128(r_full), b = f128(r_top);
    const __float128 d = fabsq(a - b);
    const __float128 big = fmaxq(fabsq(a), fabsq(b));
    int e;
    (void)frexpq(big, &e);  // big = f 2^e, f in [0.5, 1)
    const __float128 ulp = ldexpq(1.0Q, (e - 113 > -16494) ? e - 113 : -16494);
    const __float128 bound = ulp + ldexpq(1.0Q, terr);
    if (!(d <= bound)) {
      if (++mism <= 10) printf("cut mismatch case %ld kind %d terr %d\n", c, kind, terr);
    }
  }
  printf("cases %ld uncut %ld cut %ld mismatches %ld \n", cases, exact_eq, cut, mism);
  return mism ? 1 : 0;
}
