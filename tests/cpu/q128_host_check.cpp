// Host build of the product's binary128 core (qblas_amd/csrc/q128.hpp) checked against
// libquadmath (fmaq, __float128 + and *) on seeded random and edge-case operands.
// Usage: q128_host_check <seed> <count>   -> prints "mismatches <n>" and exits 1 on any mismatch.
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../qblas_amd/csrc/q128.hpp"

using qb::q128;
static __float128 f(q128 a) { __float128 r; memcpy(&r, &a, 16); return r; }
static q128 g(__float128 v) { q128 r; memcpy(&r, &v, 16); return r; }
static uint64_t st;
static uint64_t rnd() {
  uint64_t z = (st += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static q128 rq(int mode, int ec) {
  q128 r = qb::q_make(rnd(), rnd());
  uint64_t s = r.hi & (1ull << 63), fh = r.hi & qb::FRAC_HI_MASK;
  int be;
  switch (mode) {
    case 0: { uint64_t k = rnd() % 16; be = k == 0 ? 0 : (k == 1 ? 0x7fff : (int)(rnd() % 0x7fff)); break; }
    case 1: be = ec + (int)(rnd() % 7) - 3; break;
    case 2: { int nb = (int)(rnd() % 12); be = 16383 + nb; r.lo = 0; fh = (fh >> (48 - nb)) << (48 - nb); break; }
    case 3: { uint64_t k = rnd() % 3; be = k == 0 ? (int)(rnd() % 130) : (k == 1 ? 0x7ffe - (int)(rnd() % 130) : ec); break; }
    default: be = ec + (int)(rnd() % 3) - 1; break;
  }
  if (be < 0) be = 0;
  if (be > 0x7fff) be = 0x7fff;
  if (mode != 0 && be == 0x7fff) be = 0x7ffe;
  r.hi = s | ((uint64_t)be << 48) | fh;
  return r;
}
// structured significands: powers of two, all ones, 1+ulp, trailing runs, short significands
static void sigpat(uint64_t& fh, uint64_t& fl) {
  switch (rnd() % 7) {
    case 0: fh = 0; fl = 0; break;                                  // 2^112
    case 1: fh = qb::FRAC_HI_MASK; fl = ~0ull; break;               // 2^113 - 1
    case 2: fh = 0; fl = 1; break;                                  // 2^112 + 1
    case 3: { int k = (int)(rnd() % 112);                           // 2^113 - 2^k
              fh = qb::FRAC_HI_MASK; fl = ~0ull;
              if (k >= 64) { fl = 0; fh = (fh >> (k - 64)) << (k - 64); } else fl = (fl >> k) << k; break; }
    case 4: { int k = (int)(rnd() % 112);                           // random with low k bits zero
              fh = rnd() & qb::FRAC_HI_MASK; fl = rnd();
              if (k >= 64) { fl = 0; fh = (fh >> (k - 64)) << (k - 64); } else fl = (fl >> k) << k; break; }
    case 5: { fh = rnd() & qb::FRAC_HI_MASK; fl = rnd() | 1; break; }   // odd random
    default: { fh = (rnd() & qb::FRAC_HI_MASK) >> 20 << 20; fl = 0; break; }  // 29-bit
  }
}
static const int kOffsets[] = {-400, -300, -227, -226, -225, -114, -113, -112, -111, -2, -1, 0, 1, 2,
                               93, 94, 110, 111, 112, 113, 114, 115, 200, 205, 206, 207, 208, 225, 226,
                               227, 300, 400};
static q128 structured(int be) {
  uint64_t fh, fl;
  sigpat(fh, fl);
  if (be < 1) be = 1;
  if (be > 0x7ffe) be = 0x7ffe;
  return qb::q_make(fl, (rnd() & (1ull << 63)) | ((uint64_t)be << 48) | fh);
}
static bool same(q128 x, q128 y) {
  if (qb::q_isnan(x) && qb::q_isnan(y)) return true;
  return x.lo == y.lo && x.hi == y.hi;
}
// "sqrt <seed> <count>": print operand and qb::sqrt_packed result as hex words, one pair per line
// (random positive values over the whole exponent range, subnormals included, and exact
// squares); the caller checks correct rounding in exact integer arithmetic
static int sqrt_dump(long count) {
  for (long i = 0; i < count; ++i) {
    q128 a = qb::q_make(rnd(), rnd());
    const uint64_t k = rnd() % 4;
    const int be = k == 0 ? (int)(rnd() % 0x7fff) : (k == 1 ? 0 : 16383 + (int)(rnd() % 200) - 100);
    a.hi = (a.hi & qb::FRAC_HI_MASK) | ((uint64_t)be << 48);
    if (k == 3) {
      const __float128 v = (__float128)(double)(rnd() >> 11) * (__float128)0x1p-20;
      a = g(v * v);
    }
    if (qb::q_iszero(a)) continue;
    const q128 r = qb::sqrt_packed(a);
    printf("%016llx %016llx %016llx %016llx\n", (unsigned long long)a.hi, (unsigned long long)a.lo,
           (unsigned long long)r.hi, (unsigned long long)r.lo);
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc > 1 && !strcmp(argv[1], "sqrt")) {
    st = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
    return sqrt_dump(argc > 3 ? atol(argv[3]) : 10000);
  }
  st = argc > 1 ? strtoull(argv[1], 0, 10) : 1;
  long count = argc > 2 ? atol(argv[2]) : 100000;
  long bad = 0;
  for (long i = 0; i < count; ++i) {
    int mode = (int)(i % 8);
    if (mode >= 6) {
      // mode 6: structured operands, addend exponent at boundary offsets from the product's;
      // mode 7: same with a tiny or huge random third operand (ties broken only by jam bits)
      int e0 = 16383 + (int)(rnd() % 200) - 100;
      q128 a = structured(e0), b = structured(16383 + (int)(rnd() % 200) - 100);
      int ep = (int)qb::q_bexp(a) + (int)qb::q_bexp(b) - 16383;
      int off = kOffsets[rnd() % (sizeof(kOffsets) / sizeof(kOffsets[0]))] - 112;
      q128 c = mode == 6 ? structured(ep + off) : rq(0, 0);
      if (mode == 7) c = structured(ep + off + (int)(rnd() % 5) - 2);
      q128 r1 = qb::fma_packed(a, b, c), e1 = g(fmaq(f(a), f(b), f(c)));
      q128 r2 = qb::add_packed(a, c), e2 = g(f(a) + f(c));
      if (!same(r1, e1) || !same(r2, e2)) {
        if (bad < 5)
          printf("MISMATCH structured a=%016llx:%016llx b=%016llx:%016llx c=%016llx:%016llx\n",
                 (unsigned long long)a.hi, (unsigned long long)a.lo, (unsigned long long)b.hi, (unsigned long long)b.lo,
                 (unsigned long long)c.hi, (unsigned long long)c.lo);
        bad++;
      }
      continue;
    }
    int e0 = 16383 + (int)(rnd() % 240) - 120;
    q128 a = rq(mode == 5 ? 1 : mode, e0), b = rq(mode == 5 ? 1 : mode, e0), c;
    if (mode == 4 || mode == 5) {
      c = g(f(a) * f(b));
      c.hi ^= 1ull << 63;
      c.lo ^= rnd() & (mode == 4 ? 0xffull : 0xffffffffull);
    } else if (mode == 1) {
      int ea = (int)qb::q_bexp(a) - 16383, eb = (int)qb::q_bexp(b) - 16383;
      c = rq(1, 16383 + ea + eb + (int)(rnd() % 240) - 120);
    } else {
      c = rq(mode, e0);
    }
    q128 r1 = qb::fma_packed(a, b, c), e1 = g(fmaq(f(a), f(b), f(c)));
    q128 r2 = qb::add_packed(a, c), e2 = g(f(a) + f(c));
    q128 r3 = qb::mul_packed(a, b), e3 = g(f(a) * f(b));
    if (!same(r1, e1) || !same(r2, e2) || !same(r3, e3)) {
      if (bad < 5) {
        printf("MISMATCH mode %d\n a=%016llx:%016llx b=%016llx:%016llx c=%016llx:%016llx\n", mode,
               (unsigned long long)a.hi, (unsigned long long)a.lo, (unsigned long long)b.hi, (unsigned long long)b.lo,
               (unsigned long long)c.hi, (unsigned long long)c.lo);
        printf(" fma %016llx:%016llx want %016llx:%016llx\n", (unsigned long long)r1.hi, (unsigned long long)r1.lo, (unsigned long long)e1.hi, (unsigned long long)e1.lo);
        printf(" add %016llx:%016llx want %016llx:%016llx\n", (unsigned long long)r2.hi, (unsigned long long)r2.lo, (unsigned long long)e2.hi, (unsigned long long)e2.lo);
        printf(" mul %016llx:%016llx want %016llx:%016llx\n", (unsigned long long)r3.hi, (unsigned long long)r3.lo, (unsigned long long)e3.hi, (unsigned long long)e3.lo);
      }
      bad++;
    }
  }
  // conversions: double -> binary128 exact
  for (long i = 0; i < count / 4; ++i) {
    uint64_t u = rnd();
    double d;
    memcpy(&d, &u, 8);
    q128 r = qb::from_double(d), e = g((__float128)d);
    if (!same(r, e)) { if (bad < 5) printf("from_double mismatch %016llx\n", (unsigned long long)u); bad++; }
  }
  // conversions: binary128 -> double, round to nearest even (incl. subnormal/overflow ranges)
  for (long i = 0; i < count / 4; ++i) {
    q128 a = qb::q_make(rnd(), rnd());
    uint64_t k = rnd() % 4;
    int be = k == 0 ? 16383 + (int)(rnd() % 2200) - 1100 : (k == 1 ? 16383 - 1022 - (int)(rnd() % 60)
                                                           : (k == 2 ? 16383 + 1020 + (int)(rnd() % 8) : (int)(rnd() % 0x8000)));
    if (k == 3 && (rnd() & 1)) a.lo &= ~((1ull << 59) - 1);  // exact ties
    a.hi = (a.hi & ~(0x7fffull << 48)) | ((uint64_t)be << 48);
    double r = qb::to_double(a), e = (double)f(a);
    uint64_t rb, eb;
    memcpy(&rb, &r, 8);
    memcpy(&eb, &e, 8);
    if (rb != eb && !(r != r && e != e)) { if (bad < 5) printf("to_double mismatch %016llx:%016llx\n", (unsigned long long)a.hi, (unsigned long long)a.lo); bad++; }
  }
  printf("mismatches %ld of %ld\n", bad, count);
  return bad ? 1 : 0;
}
