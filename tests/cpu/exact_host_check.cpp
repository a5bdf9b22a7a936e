// Host run of the EXACT form's arithmetic (qblas_amd/csrc/exact.hpp): the kernel structure
// (64-lane waves, wave-uniform anchors, per-wave windows, wide combine, exact alpha*S+beta*y
// epilogue) emulated on the CPU and compared bit-for-bit with the oracle's exact accumulator
// (oracle/qoracle.c or_exact_dot).  Usage: exact_host_check <seed> <cases>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "../../qblas_amd/csrc/exact.hpp"

using namespace qb;
extern "C" {
int or_exact_dot(size_t n, const q128 *x, ptrdiff_t incx, const q128 *y, ptrdiff_t incy, q128 alpha, q128 beta,
                 q128 c, const q128 *cand, size_t ncand, q128 *rounded, double *relerr);
void or_fill(int kind, uint64_t seed, size_t n, int lo_exp, int hi_exp, q128 *out);
void or_fill_pair_y(uint64_t seed, size_t n, q128 *out);
}

static const int EXP_NONE = 0x3fffffff;

static int pair_exp(q128 a, q128 b) {
  uq ua, ub;
  uint32_t fl = 0;
  int c = classify_pair(a, b, ua, ub, fl);
  return (c == 0 || c == 3) ? prod_lsb(ua, ub) : EXP_NONE;
}

// emulate dot over `lanes` lanes grouped in 64-lane waves, grid-stride assignment
static q128 dot_emul(const std::vector<q128> &x, const std::vector<q128> &y, int lanes, bool &lost) {
  const int64_t n = (int64_t)x.size();
  const int waves = lanes / 64;
  std::vector<int> As(waves);
  std::vector<std::vector<uint32_t>> Ws(waves, std::vector<uint32_t>(WW, 0)), Whs(waves, std::vector<uint32_t>(WW, 0));
  uint32_t fl_all = 0;
  for (int w = 0; w < waves; ++w) {
    int mn = EXP_NONE, mx = -EXP_NONE;
    for (int l = 0; l < 64; ++l) {
      int64_t i0 = (int64_t)w * 64 + l;
      int e0 = (i0 < n) ? pair_exp(x[i0], y[i0]) : EXP_NONE;
      mn = e0 < mn ? e0 : mn;
      int ev = e0 == EXP_NONE ? -EXP_NONE : e0;
      mx = ev > mx ? ev : mx;
    }
    const int A = (mn == EXP_NONE) ? 0 : ((mn + mx) >> 1) - WIN2_CENTER;
    As[w] = A;
    uint32_t W[WW] = {0}, Wh[WW] = {0};
    for (int l = 0; l < 64; ++l) {
      uint32_t L[WW] = {0}, Lh[WW] = {0};
      uint32_t fl = 0;
      for (int64_t i = (int64_t)w * 64 + l; i < n; i += lanes) {
        uq ua, ub;
        int c = classify_pair(x[i], y[i], ua, ub, fl);
        if (c == 0 || c == 3)
          if (!win2_add_product(L, Lh, A, ua, ub)) fl |= FL_LOST;
      }
      win_add_win(W, L);
      win_add_win(Wh, Lh);
      fl_all |= fl;
    }
    for (int k = 0; k < WW; ++k) Ws[w][k] = W[k];
    for (int k = 0; k < WW; ++k) Whs[w][k] = Wh[k];
  }
  int amin = EXP_NONE;
  for (int w = 0; w < waves; ++w) {
    if (!win_is_zero(Ws[w].data()) && As[w] < amin) amin = As[w];
    if (!win_is_zero(Whs[w].data()) && As[w] + WIN_BAND < amin) amin = As[w] + WIN_BAND;
  }
  wide_acc acc;
  wide_init(acc);
  if (amin != EXP_NONE) {
    acc.any = true;
    acc.base = amin - 32;
  }
  for (int w = 0; w < waves; ++w) {
    wide_add_window(acc, Ws[w].data(), As[w]);
    wide_add_window(acc, Whs[w].data(), As[w] + WIN_BAND);
  }
  lost = acc.lost || (fl_all & FL_LOST);
  uint32_t M[SW], sign, sticky;
  int lsb;
  compact_wide(acc, M, lsb, sign, sticky);
  if (!acc.any) return (fl_all & (FL_NAN | FL_PINF | FL_NINF)) ? special_result(fl_all) : q_zero(0);
  return finish_reg(M, lsb, sign, sticky, fl_all, false, q_one(), q_zero(0), q_zero(0));
}

// emulate one gemv row: nch column chunks sharing the row anchor (from the row's first 64
// products), window sums, register epilogue
static q128 row_emul(const q128 *r, const q128 *x, int64_t n, int nch, q128 alpha, q128 beta, q128 yv, bool &lost) {
  const int64_t chunk = (n + nch - 1) / nch;
  uint32_t fl_all = 0;
  int mn = EXP_NONE, mx = -EXP_NONE;
  for (int l = 0; l < 64; ++l) {
    int e0 = (l < n) ? pair_exp(r[l], x[l]) : EXP_NONE;
    mn = e0 < mn ? e0 : mn;
    int ev = e0 == EXP_NONE ? -EXP_NONE : e0;
    mx = ev > mx ? ev : mx;
  }
  const int A = (mn == EXP_NONE) ? 0 : ((mn + mx) >> 1) - WIN2_CENTER;
  uint32_t lo[WW] = {0}, hi[WW] = {0};
  for (int c = 0; c < nch; ++c) {
    const int64_t j0 = c * chunk, j1 = (j0 + chunk < n) ? j0 + chunk : n;
    uint32_t W[WW] = {0}, Wh[WW] = {0};
    for (int64_t j = j0; j < j1; ++j) {
      uq ua, ub;
      int k = classify_pair(r[j], x[j], ua, ub, fl_all);
      if (k == 0 || k == 3)
        if (!win2_add_product(W, Wh, A, ua, ub)) fl_all |= FL_LOST;
    }
    win_add_win(lo, W);
    win_add_win(hi, Wh);
  }
  lost = (fl_all & FL_LOST) != 0;
  uint32_t M[SW], sign;
  s_from_windows(M, sign, lo, hi);
  return finish_reg(M, A, sign, 0u, fl_all, true, alpha, beta, yv);
}

static bool same(q128 a, q128 b) { return (q_isnan(a) && q_isnan(b)) || (a.lo == b.lo && a.hi == b.hi); }

int main(int argc, char **argv) {
  uint64_t seed = argc > 1 ? strtoull(argv[1], 0, 10) : 1;
  int cases = argc > 2 ? atoi(argv[2]) : 200;
  long bad = 0, lostc = 0, checked = 0;
  for (int t = 0; t < cases; ++t) {
    const int kind = t % 5;
    const int64_t n = 1 + (int64_t)((seed * 7919 + t * 104729) % 3000);
    std::vector<q128> x(n), y(n);
    if (kind == 0) { or_fill(0, seed + t, n, 0, 0, x.data()); or_fill(0, seed + t + 1000, n, 0, 0, y.data()); }
    else if (kind == 1) { or_fill(1, seed + t, n, -55, 55, x.data()); or_fill(1, seed + t + 1000, n, -55, 55, y.data()); }
    else if (kind == 2) { or_fill(2, seed + t, n, 0, 0, x.data()); or_fill_pair_y(seed + t, n, y.data()); }
    else if (kind == 3) {
      // zeros, signed zeros, subnormal-range and exact cancellation
      or_fill(0, seed + t, n, 0, 0, x.data());
      or_fill(0, seed + t + 1000, n, 0, 0, y.data());
      for (int64_t i = 0; i < n; i += 7) x[i] = q_zero(i & 1);
      for (int64_t i = 3; i + 1 < n; i += 11) { x[i + 1] = x[i]; x[i + 1].hi ^= 1ull << 63; y[i + 1] = y[i]; }
    } else {
      or_fill(1, seed + t, n, -16380, -16300, x.data());  // deep underflow of products
      or_fill(1, seed + t + 1000, n, -40, -30, y.data());
    }
    bool lost = false;
    const int lanes = 64 * (1 + (int)(t % 4));
    q128 got = dot_emul(x, y, lanes, lost);
    q128 want;
    or_exact_dot(n, x.data(), 1, y.data(), 1, q_one(), q_zero(0), q_zero(0), nullptr, 0, &want, nullptr);
    if (lost) { lostc++; continue; }
    checked++;
    if (!same(got, want)) {
      if (bad < 5) printf("dot MISMATCH kind %d n %lld: got %016llx:%016llx want %016llx:%016llx\n", kind, (long long)n,
                          (unsigned long long)got.hi, (unsigned long long)got.lo, (unsigned long long)want.hi, (unsigned long long)want.lo);
      bad++;
    }
    // gemv rows: same vectors as row and x, random alpha/beta/y
    std::vector<q128> ab(3);
    or_fill(0, seed + t + 77, 3, 0, 0, ab.data());
    if (t % 3 == 0) ab[1] = q_zero(0);
    if (t % 7 == 0) { ab[2] = got; ab[2].hi ^= 1ull << 63; ab[0] = q_one(); ab[1] = q_one(); }  // cancels alpha*S
    const int nch = 1 + (int)(t % 5);
    q128 gr = row_emul(x.data(), y.data(), n, nch, ab[0], ab[1], ab[2], lost);
    q128 wr;
    or_exact_dot(n, x.data(), 1, y.data(), 1, ab[0], ab[1], ab[2], nullptr, 0, &wr, nullptr);
    if (lost) { lostc++; continue; }
    checked++;
    if (!same(gr, wr)) {
      if (bad < 5) printf("row MISMATCH kind %d n %lld nch %d: got %016llx:%016llx want %016llx:%016llx\n", kind, (long long)n, nch,
                          (unsigned long long)gr.hi, (unsigned long long)gr.lo, (unsigned long long)wr.hi, (unsigned long long)wr.lo);
      bad++;
    }
  }
  // crafted: ties broken only by a far-away beta*y (sticky), and signs of exact zeros
  {
    const q128 one = q_one(), m113 = q_make(0, (uint64_t)(QBIAS - 113) << 48);  // 2^-113
    for (int sgn_y = 0; sgn_y < 2; ++sgn_y)
      for (int sgn_a = 0; sgn_a < 2; ++sgn_a)
        for (int far = 0; far < 5; ++far) {
          std::vector<q128> x = {one, m113, one}, y = {one, one, q_zero(0)};
          q128 alpha = sgn_a ? q_make(0, one.hi | (1ull << 63)) : one;
          const int ey = far == 0 ? -300 : (far == 1 ? -200 : (far == 2 ? -130 : (far == 3 ? -1500 : -3000)));
          q128 yc = q_make(0, ((uint64_t)sgn_y << 63) | ((uint64_t)(QBIAS + ey) << 48));
          for (int nch = 1; nch <= 2; ++nch) {
            bool lost = false;
            q128 gr = row_emul(x.data(), y.data(), 3, nch, alpha, one, yc, lost);
            q128 wr;
            or_exact_dot(3, x.data(), 1, y.data(), 1, alpha, one, yc, nullptr, 0, &wr, nullptr);
            checked++;
            if (lost || !same(gr, wr)) {
              if (bad < 5) printf("crafted tie MISMATCH sy %d sa %d far %d\n", sgn_y, sgn_a, far);
              bad++;
            }
          }
        }
    // exact zeros: all products zero, beta*y zero, every sign combination
    for (int c = 0; c < 16; ++c) {
      std::vector<q128> x = {q_zero(c & 1), q_zero(0)}, y = {one, one};
      q128 alpha = (c & 2) ? q_make(0, one.hi | (1ull << 63)) : one;
      q128 beta = (c & 4) ? q_zero(1) : q_zero(0);
      q128 yc = (c & 8) ? q_make(0, one.hi | (1ull << 63)) : q_zero(1);
      bool lost = false;
      q128 gr = row_emul(x.data(), y.data(), 2, 1, alpha, beta, yc, lost);
      q128 wr;
      or_exact_dot(2, x.data(), 1, y.data(), 1, alpha, beta, yc, nullptr, 0, &wr, nullptr);
      checked++;
      if (lost || !same(gr, wr)) {
        if (bad < 5) printf("zero-sign MISMATCH case %d got %016llx want %016llx\n", c, (unsigned long long)gr.hi, (unsigned long long)wr.hi);
        bad++;
      }
    }
  }
  // wide accumulators with sparse words far apart (ties broken only by low words): compaction
  // to SW words with a sticky, then rounding; the exact value is rebuilt as a sum of exact
  // binary128 terms w_i * 2^(base + 32 i) for the oracle
  {
    uint64_t st = seed * 0x9e3779b97f4a7c15ull + 17;
    auto rnd = [&]() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; };
    for (int t = 0; t < 3000; ++t) {
      uint32_t mag[ACCW] = {0};
      const int top = 20 + (int)(rnd() % (ACCW - 21));
      mag[top] = (uint32_t)rnd() | 1u;
      if (t % 3 == 0) {  // tie pattern right below the top 113 bits, resolved by a far-away word
        for (int i = top - 4; i < top; ++i) mag[i] = 0;
        const int gbit = 32 * top + 31 - __builtin_clz(mag[top]) - 113;  // bit position of guard
        if (gbit >= 0) mag[gbit >> 5] |= 1u << (gbit & 31);
        if (rnd() & 1) mag[rnd() % (uint64_t)(top > 26 ? top - 26 : 1)] |= 1u << (rnd() % 32);
      } else {
        const int nlow = (int)(rnd() % 4);
        for (int k = 0; k < nlow; ++k) mag[rnd() % (uint64_t)top] = (uint32_t)rnd();
        for (int i = top - 3; i < top; ++i) mag[i] = (uint32_t)rnd();
      }
      const uint32_t sg = (uint32_t)(rnd() & 1);
      wide_acc acc;
      wide_init(acc);
      acc.any = true;
      acc.base = -600 - (int)(rnd() % 200);
      uint32_t c = sg;
      for (int i = 0; i < ACCW; ++i) acc.w[i] = sg ? addc32(~mag[i], 0u, c, c) : mag[i];
      uint32_t M[SW], sign, sticky;
      int lsb;
      compact_wide(acc, M, lsb, sign, sticky);
      q128 got = finish_reg(M, lsb, sign, sticky, 0, false, q_one(), q_zero(0), q_zero(0));
      std::vector<q128> xs, ys;
      for (int i = 0; i < ACCW; ++i) {
        if (!mag[i]) continue;
        q128 w = from_double((double)mag[i]);
        if (sg) w.hi |= 1ull << 63;
        xs.push_back(w);
        ys.push_back(q_make(0, (uint64_t)(QBIAS + acc.base + 32 * i) << 48));
      }
      q128 want;
      or_exact_dot(xs.size(), xs.data(), 1, ys.data(), 1, q_one(), q_zero(0), q_zero(0), nullptr, 0, &want, nullptr);
      checked++;
      if (!same(got, want)) {
        if (bad < 5) printf("wide-acc MISMATCH t %d top %d\n", t, top);
        bad++;
      }
    }
  }
  printf("checked %ld lost %ld mismatches %ld\n", checked, lostc, bad);
  return bad ? 1 : 0;
}
