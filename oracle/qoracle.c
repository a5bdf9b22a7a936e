/*
 * qoracle.c -- CPU oracle for the QuadBLAS binary128 hot path (dot / gemv / gemm).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke check and the
 * CPU-baseline leg of the benchmark load it.  The shipped library never links it, and no
 * product code path falls back to it.
 *
 * What it holds
 *   1. Binary128 arithmetic, round-to-nearest-even:
 *        or_add / or_mul  -- GCC __float128 + and * (libgcc soft-fp, correctly rounded);
 *        or_fma           -- an independent integer implementation of a correctly rounded
 *                            fused multiply-add (unsigned __int128 limbs), pinned bit-for-bit
 *                            against libquadmath fmaq by or_fma_selftest();
 *        or_fmaq          -- libquadmath fmaq itself (slow: ~1 us/op), the pin for or_fma.
 *      The reference does all arithmetic through SLEEF's Sleef_{add,mul,fma}q1_u05 and the
 *      2-lane Sleef_fmaq2_u05sse2 (include/quadblas/simd/quad_vector.hpp:99-139), whose "_u05"
 *      contract is a 0.5-ULP (correctly rounded) result.  SLEEF is not in this image, so the
 *      oracle uses correctly rounded operations; equality with SLEEF's own bits is unpinned.
 *   2. Restatements of the reference's evaluation order (one function per reference routine):
 *        or_dot_kernel   include/quadblas/algorithms/level1.hpp:14-35
 *        or_dot          include/quadblas/algorithms/level1.hpp:38-137 (T = OpenMP thread count)
 *        or_gemv         include/quadblas/algorithms/level2.hpp:15-99
 *        or_gemm         include/quadblas/algorithms/level3.hpp:187-336 with kc from
 *                        include/quadblas/detail/blocking.hpp:17-67 (126 on x86-64 Linux)
 *   3. An exact dot-product accumulator (arbitrary-width two's complement integer) that rounds
 *      the exact value of alpha*sum(x_i*y_i) + beta*c once, and measures the relative error of
 *      a candidate result against the exact value.
 *   4. The seeded synthetic-input generators used by the tests and the benchmark
 *      (counter-based splitmix64 stream; identical definition in the HIP library).
 *
 * Bit layout of a binary128 value: two little-endian uint64 words {lo, hi}; hi bit 63 is the
 * sign, hi bits 62..48 the biased exponent (bias 16383), the remaining 112 bits the fraction.
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>
#include <quadmath.h>

typedef struct { uint64_t lo, hi; } q128;
typedef unsigned __int128 u128;

static inline __float128 q2f(q128 a) { __float128 f; memcpy(&f, &a, 16); return f; }
static inline q128 f2q(__float128 f) { q128 a; memcpy(&a, &f, 16); return a; }

/* ------------------------------------------------------------------------------------------ */
/* 1. arithmetic                                                                               */
/* ------------------------------------------------------------------------------------------ */

q128 or_add(q128 a, q128 b) { return f2q(q2f(a) + q2f(b)); }
q128 or_mul(q128 a, q128 b) { return f2q(q2f(a) * q2f(b)); }
q128 or_fmaq(q128 a, q128 b, q128 c) { return f2q(fmaq(q2f(a), q2f(b), q2f(c))); }
q128 or_from_double(double d) { return f2q((__float128)d); }
double or_to_double(q128 a) { return (double)q2f(a); }

#define EXP_MAX 0x7fff
#define BIAS 16383

static const q128 QNAN = {0, 0x7fff800000000000ull};

/* Decoded finite value: value = sig * 2^(exp - 112), sig normalised to bit 112 unless zero. */
typedef struct { int cls; int sign; int exp; u128 sig; } dec;
enum { C_ZERO, C_FIN, C_INF, C_NAN };

static dec decode(q128 a) {
    dec d;
    d.sign = (int)(a.hi >> 63);
    int be = (int)((a.hi >> 48) & EXP_MAX);
    u128 frac = ((u128)(a.hi & 0xffffffffffffull) << 64) | a.lo;
    if (be == EXP_MAX) { d.cls = frac ? C_NAN : C_INF; d.exp = 0; d.sig = frac; return d; }
    if (be == 0) {
        if (!frac) { d.cls = C_ZERO; d.exp = 0; d.sig = 0; return d; }
        /* subnormal: value = frac * 2^(1 - BIAS - 112); normalise */
        int lz = 0;
        while (!((frac << lz) >> 112 & 1)) lz++;
        d.cls = C_FIN; d.sig = frac << lz; d.exp = 1 - BIAS - lz; return d;
    }
    d.cls = C_FIN; d.sig = frac | ((u128)1 << 112); d.exp = be - BIAS; return d;
}

/* A little-endian multiword magnitude of NW 64-bit words. */
#define NW 7   /* 448 bits: 226-bit product + 113-bit addend + head/guard room */

static int mw_msb(const uint64_t *w) {
    for (int i = NW - 1; i >= 0; --i) if (w[i]) return i * 64 + 63 - __builtin_clzll(w[i]);
    return -1;
}
/* 64 bits of w starting at bit position pos (bits outside [0, NW*64) read as zero) */
static uint64_t mw_get64(const uint64_t *w, int pos) {
    int wi = pos >> 6, b = pos & 63;   /* arithmetic shift: floor for negative pos */
    uint64_t lo = (wi >= 0 && wi < NW) ? w[wi] : 0;
    uint64_t hi = (wi + 1 >= 0 && wi + 1 < NW) ? w[wi + 1] : 0;
    return b ? (lo >> b) | (hi << (64 - b)) : lo;
}
static int mw_bit(const uint64_t *w, int pos) { return (pos >= 0 && pos < NW * 64) ? (int)((w[pos >> 6] >> (pos & 63)) & 1) : 0; }
static int mw_any_below(const uint64_t *w, int pos) {  /* any set bit strictly below pos */
    if (pos <= 0) return 0;
    int wi = pos >> 6, b = pos & 63;
    for (int i = 0; i < wi && i < NW; ++i) if (w[i]) return 1;
    if (wi < NW && b && (w[wi] & ((1ull << b) - 1))) return 1;
    return 0;
}
/* place 'v' (nv words) shifted left by sh (sh may be negative: shift right, OR-ing the lost bits
   into bit 0 -- used only for the smaller operand, whose lost bits lie below every bit of the
   larger one) */
static void mw_place(uint64_t *w, const uint64_t *v, int nv, int sh) {
    memset(w, 0, NW * 8);
    if (sh >= 0) {
        int ws = sh >> 6, bs = sh & 63;
        for (int i = 0; i < nv; ++i) {
            int d = i + ws;
            if (d < NW) w[d] |= v[i] << bs;
            if (bs && d + 1 < NW) w[d + 1] |= v[i] >> (64 - bs);
        }
    } else {
        int s = -sh, ws = s >> 6, bs = s & 63;
        int jam = 0;
        for (int i = 0; i < nv && i < ws; ++i) if (v[i]) jam = 1;
        if (ws < nv && bs && (v[ws] & ((1ull << bs) - 1))) jam = 1;
        for (int i = 0; i < NW; ++i) {
            int lo = i + ws;
            uint64_t a = lo < nv ? v[lo] : 0, b = lo + 1 < nv ? v[lo + 1] : 0;
            w[i] = bs ? (a >> bs) | (b << (64 - bs)) : a;
        }
        w[0] |= (uint64_t)jam;
    }
}
static int mw_cmp(const uint64_t *a, const uint64_t *b) {
    for (int i = NW - 1; i >= 0; --i) { if (a[i] != b[i]) return a[i] > b[i] ? 1 : -1; }
    return 0;
}
static void mw_add(uint64_t *r, const uint64_t *a, const uint64_t *b) {
    u128 c = 0;
    for (int i = 0; i < NW; ++i) { c += (u128)a[i] + b[i]; r[i] = (uint64_t)c; c >>= 64; }
}
static void mw_sub(uint64_t *r, const uint64_t *a, const uint64_t *b) {  /* a >= b */
    uint64_t br = 0;
    for (int i = 0; i < NW; ++i) {
        uint64_t t = a[i] - b[i] - br;
        br = (a[i] < b[i]) || (a[i] == b[i] && br) ? 1 : 0;
        r[i] = t;
    }
}

/* Round the magnitude w * 2^lsbexp (sign s) to binary128, round-to-nearest-even. */
static q128 round_pack(int s, const uint64_t *w, int lsbexp) {
    q128 r;
    int msb = mw_msb(w);
    if (msb < 0) { r.lo = 0; r.hi = (uint64_t)s << 63; return r; }
    int e = lsbexp + msb;            /* exponent of the leading bit */
    int keep_lsb = msb - 112;        /* bit position of the result's ulp */
    if (e < 1 - BIAS) keep_lsb += (1 - BIAS) - e;   /* subnormal: ulp fixed at 2^(1-BIAS-112) */
    /* gather the kept bits: positions [keep_lsb, msb] (at most 113 of them) */
    u128 sig = 0;
    if (keep_lsb <= msb) {
        int nb = msb - keep_lsb + 1;
        u128 lo = mw_get64(w, keep_lsb), hi = mw_get64(w, keep_lsb + 64);
        sig = (hi << 64) | lo;
        if (nb < 128) sig &= (((u128)1) << nb) - 1;
    }
    int guard = mw_bit(w, keep_lsb - 1);
    int sticky = mw_any_below(w, keep_lsb - 1);
    if (guard && (sticky || (sig & 1))) sig += 1;
    int ulpexp = lsbexp + keep_lsb;  /* value = sig * 2^ulpexp */
    /* renormalise after rounding carry */
    if (sig >> 113) { sig >>= 1; ulpexp += 1; }
    int be;
    if (sig >> 112) {
        be = ulpexp + 112 + BIAS;
        if (be >= EXP_MAX) { r.lo = 0; r.hi = ((uint64_t)s << 63) | ((uint64_t)EXP_MAX << 48); return r; }
        sig &= (((u128)1) << 112) - 1;
    } else {
        be = 0;  /* subnormal (or rounded up into the smallest normal handled above) */
    }
    r.lo = (uint64_t)sig;
    r.hi = ((uint64_t)s << 63) | ((uint64_t)be << 48) | (uint64_t)(sig >> 64);
    return r;
}

q128 or_fma(q128 a, q128 b, q128 c) {
    dec A = decode(a), B = decode(b), C = decode(c);
    if (A.cls == C_NAN || B.cls == C_NAN || C.cls == C_NAN) return QNAN;
    int sp = A.sign ^ B.sign;
    if (A.cls == C_INF || B.cls == C_INF) {
        if (A.cls == C_ZERO || B.cls == C_ZERO) return QNAN;
        if (C.cls == C_INF && C.sign != sp) return QNAN;
        q128 r = {0, ((uint64_t)sp << 63) | ((uint64_t)EXP_MAX << 48)};
        return r;
    }
    if (C.cls == C_INF) return c;
    if (A.cls == C_ZERO || B.cls == C_ZERO) {
        if (C.cls == C_ZERO) { q128 r = {0, (uint64_t)(sp & C.sign) << 63}; return r; }
        return c;
    }
    /* exact product: 226-bit integer P, value P * 2^(ea + eb - 224) */
    uint64_t al = (uint64_t)A.sig, ah = (uint64_t)(A.sig >> 64);
    uint64_t bl = (uint64_t)B.sig, bh = (uint64_t)(B.sig >> 64);
    u128 ll = (u128)al * bl, lh = (u128)al * bh, hl = (u128)ah * bl, hh = (u128)ah * bh;
    uint64_t p[4];
    p[0] = (uint64_t)ll;
    u128 mid = (ll >> 64) + (uint64_t)lh + (uint64_t)hl;
    p[1] = (uint64_t)mid;
    u128 upper = (mid >> 64) + (lh >> 64) + (hl >> 64) + (uint64_t)hh;
    p[2] = (uint64_t)upper;
    p[3] = (uint64_t)((upper >> 64) + (hh >> 64));
    int pexp = A.exp + B.exp - 224;  /* weight of P's bit 0 */
    uint64_t W[NW], X[NW], R[NW];
    if (C.cls == C_ZERO) {
        /* anchor the product 64 bits up so rounding never needs bits below the frame */
        mw_place(W, p, 4, 64);
        return round_pack(sp, W, pexp - 64);
    }
    uint64_t cv[2] = {(uint64_t)C.sig, (uint64_t)(C.sig >> 64)};
    int cexp = C.exp - 112;          /* weight of C.sig's bit 0 */
    /* common frame: its top (bit NW*64-2) sits at the larger operand's leading bit + 1 */
    int ptop = pexp + 224 + (int)((p[3] >> 33) & 1);   /* exponent of P's leading bit */
    int ctop = C.exp;
    int top = ptop > ctop ? ptop : ctop;
    int lsb = top + 2 - NW * 64;     /* weight of frame bit 0 */
    mw_place(W, p, 4, pexp - lsb);
    mw_place(X, cv, 2, cexp - lsb);
    int s;
    if (sp == C.sign) { mw_add(R, W, X); s = sp; }
    else {
        int cmp = mw_cmp(W, X);
        if (cmp == 0) { q128 r = {0, 0}; return r; }   /* exact zero: +0 in round-to-nearest */
        if (cmp > 0) { mw_sub(R, W, X); s = sp; } else { mw_sub(R, X, W); s = C.sign; }
    }
    return round_pack(s, R, lsb);
}

/* Compare or_fma with libquadmath fmaq on 'count' pseudo-random triples drawn by 'mode'
   (0: wide random bits incl. specials; 1: exponents near each other (cancellation);
   2: small integers / exact ties; 3: subnormal/overflow edges; 4: c ~ -(a*b)).
   Returns the number of mismatches (NaN results compare equal to any NaN); the first mismatch
   is written to first[0..2] (inputs) and first[3..4] (ours, fmaq). */
static uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static int isnanq_bits(q128 a) { return ((a.hi >> 48) & EXP_MAX) == EXP_MAX && ((a.hi & 0xffffffffffffull) || a.lo); }
static q128 rand_q(uint64_t *st, int mode, int expc) {
    q128 r; r.lo = sm64(st); r.hi = sm64(st);
    uint64_t sgn = r.hi & (1ull << 63);
    uint64_t frac_hi = r.hi & 0xffffffffffffull;
    int be;
    switch (mode) {
    case 0: { uint64_t k = sm64(st) % 16;
              if (k == 0) be = 0; else if (k == 1) be = EXP_MAX; else be = (int)(sm64(st) % 0x7fff); break; }
    case 1: be = expc + (int)(sm64(st) % 5) - 2; break;
    case 2: { /* small integers: fraction has few bits */
              int nb = (int)(sm64(st) % 12);
              be = BIAS + nb; r.lo = 0; frac_hi = (frac_hi >> (48 - nb)) << (48 - nb); break; }
    case 3: { uint64_t k = sm64(st) % 4;
              if (k == 0) be = (int)(sm64(st) % 120);
              else if (k == 1) be = 0x7ffe - (int)(sm64(st) % 120);
              else be = BIAS / 2 + (int)(sm64(st) % 200) - 100 + (k == 2 ? -BIAS / 2 + 60 : BIAS / 2 + 60);
              if (be < 0) be = 0;
              if (be > 0x7ffe) be = 0x7ffe;
              break; }
    default: be = expc + (int)(sm64(st) % 3) - 1; break;
    }
    if (be < 0) be = 0;
    if (be > EXP_MAX) be = EXP_MAX;
    r.hi = sgn | ((uint64_t)be << 48) | frac_hi;
    return r;
}
/* structured significands (powers of two, all ones, 1+ulp, trailing runs, short) at a given
   biased exponent: exact ties and round-up carries that random bits almost never produce */
static q128 struct_q(uint64_t *st, int be) {
    uint64_t fh, fl;
    int k = (int)(sm64(st) % 112);
    switch (sm64(st) % 7) {
    case 0: fh = 0; fl = 0; break;
    case 1: fh = 0xffffffffffffull; fl = ~0ull; break;
    case 2: fh = 0; fl = 1; break;
    case 3: fh = 0xffffffffffffull; fl = ~0ull;
            if (k >= 64) { fl = 0; fh = (fh >> (k - 64)) << (k - 64); } else fl = (fl >> k) << k; break;
    case 4: fh = sm64(st) & 0xffffffffffffull; fl = sm64(st);
            if (k >= 64) { fl = 0; fh = (fh >> (k - 64)) << (k - 64); } else fl = (fl >> k) << k; break;
    case 5: fh = sm64(st) & 0xffffffffffffull; fl = sm64(st) | 1; break;
    default: fh = (sm64(st) & 0xffffffffffffull) >> 20 << 20; fl = 0; break;
    }
    if (be < 1) be = 1;
    if (be > 0x7ffe) be = 0x7ffe;
    q128 r; r.lo = fl; r.hi = (sm64(st) & (1ull << 63)) | ((uint64_t)be << 48) | fh;
    return r;
}
static const int k_offsets[] = {-400, -300, -227, -226, -225, -114, -113, -112, -111, -2, -1, 0, 1, 2,
                                93, 94, 110, 111, 112, 113, 114, 115, 200, 205, 206, 207, 208, 225,
                                226, 227, 300, 400};
uint64_t or_fma_selftest(uint64_t seed, uint64_t count, int mode, q128 *first) {
    uint64_t st = seed, bad = 0;
    for (uint64_t i = 0; i < count; ++i) {
        int e0 = BIAS + (int)(sm64(&st) % 200) - 100;
        q128 a = rand_q(&st, mode, e0), b = rand_q(&st, mode, e0), c;
        if (mode == 5) {
            /* structured operands, addend exponent at boundary offsets from the product's */
            a = struct_q(&st, e0);
            b = struct_q(&st, BIAS + (int)(sm64(&st) % 200) - 100);
            int ep = (int)((a.hi >> 48) & EXP_MAX) + (int)((b.hi >> 48) & EXP_MAX) - BIAS;
            int off = k_offsets[sm64(&st) % (sizeof(k_offsets) / sizeof(k_offsets[0]))] - 112 - 1;
            c = struct_q(&st, ep + off + (int)(sm64(&st) % 5) - 2);
        } else if (mode == 4) {
            /* c = -(a*b) rounded, perturbed in the last bits: deep cancellation */
            c = or_mul(a, b); c.hi ^= 1ull << 63; c.lo ^= sm64(&st) & 0xff;
        } else if (mode == 1) {
            c = rand_q(&st, 1, e0 + ((int)((a.hi >> 48) & EXP_MAX) - BIAS) + ((int)((b.hi >> 48) & EXP_MAX) - BIAS) - (e0 - BIAS));
        } else {
            c = rand_q(&st, mode, e0);
        }
        q128 x = or_fma(a, b, c), y = or_fmaq(a, b, c);
        int same = (x.lo == y.lo && x.hi == y.hi) || (isnanq_bits(x) && isnanq_bits(y));
        if (!same) {
            if (!bad && first) { first[0] = a; first[1] = b; first[2] = c; first[3] = x; first[4] = y; }
            bad++;
        }
    }
    return bad;
}

/* ------------------------------------------------------------------------------------------ */
/* 2. restatements of the reference's evaluation order                                         */
/* ------------------------------------------------------------------------------------------ */

static const q128 QZERO = {0, 0};
static const q128 QONE = {0, 0x3fff000000000000ull};

typedef q128 (*fma_fn)(q128, q128, q128);
static fma_fn g_fma = or_fma;
/* 0: the oracle's own fma (fast); 1: libquadmath fmaq (slow, used to pin the restatements) */
void or_use_fmaq(int on) { g_fma = on ? or_fmaq : or_fma; }

/* level1.hpp:14-35: lane 0 accumulates even i, lane 1 odd i (QuadVector fma, :20-25), lanes
   added (horizontal_sum, quad_vector.hpp:141-150), then a trailing fma for odd n (:29-32). */
q128 or_dot_kernel(const q128 *x, const q128 *y, size_t n) {
    q128 e = QZERO, o = QZERO;
    size_t vn = n / 2;
    for (size_t i = 0; i < vn; ++i) {
        e = g_fma(x[2 * i], y[2 * i], e);
        o = g_fma(x[2 * i + 1], y[2 * i + 1], o);
    }
    q128 r = or_add(e, o);
    if (n & 1) r = g_fma(x[n - 1], y[n - 1], r);
    return r;
}

/* level1.hpp:139-223: y_i = fma(alpha, x_i, y_i), one rounding per element in every path
   (axpy_kernel_vectorized's lanes, its tail and the strided loop compute the same FMA). */
void or_axpy(size_t n, q128 alpha, const q128 *x, size_t incx, q128 *y, size_t incy) {
    for (size_t i = 0; i < n; ++i) y[i * incy] = or_fma(alpha, x[i * incx], y[i * incy]);
}

/* level1.hpp:38-137.  T is omp_get_max_threads() at the call (level1.hpp:46,96). */
q128 or_dot(size_t n, const q128 *x, size_t incx, const q128 *y, size_t incy, int T) {
    if (n == 0) return QZERO;                                  /* :83-84 */
    if (T < 1) T = 1;
    if (incx == 1 && incy == 1) {                              /* :86-89 -> dot_parallel */
        if (n < 500) return or_dot_kernel(x, y, n);            /* :40-43 */
        size_t chunk = n / (size_t)T;                          /* :47 */
        q128 *part = (q128 *)malloc(sizeof(q128) * (size_t)T);
        #pragma omp parallel for schedule(static)
        for (int t = 0; t < T; ++t) {
            size_t start = (size_t)t * chunk;
            size_t end = (t == T - 1) ? n : start + chunk;     /* :58-59 */
            part[t] = (start < end) ? or_dot_kernel(x + start, y + start, end - start) : QZERO;
        }
        q128 r = QZERO;
        for (int t = 0; t < T; ++t) r = or_add(r, part[t]);    /* :67-71 */
        free(part);
        return r;
    }
    q128 r = QZERO;
    if (n >= 500) {                                            /* :93-120 */
        size_t chunk = n / (size_t)T;
        q128 *part = (q128 *)malloc(sizeof(q128) * (size_t)T);
        #pragma omp parallel for schedule(static)
        for (int t = 0; t < T; ++t) {
            size_t start = (size_t)t * chunk;
            size_t end = (t == T - 1) ? n : start + chunk;
            q128 p = QZERO;
            for (size_t i = start; i < end; ++i) p = g_fma(x[i * incx], y[i * incy], p);
            part[t] = p;
        }
        for (int t = 0; t < T; ++t) r = or_add(r, part[t]);
        free(part);
    } else {                                                   /* :128-134 */
        for (size_t i = 0; i < n; ++i) r = g_fma(x[i * incx], y[i * incy], r);
    }
    return r;
}

/* level2.hpp:15-99.  colmajor=0: gemv_row_major (:15-50); colmajor=1: gemv_col_major (:53-82). */
void or_gemv(int colmajor, size_t m, size_t n, q128 alpha, const q128 *A, size_t lda,
             const q128 *x, size_t incx, q128 beta, q128 *y, size_t incy) {
    if (m == 0 || n == 0) return;
    if (!colmajor) {
        #pragma omp parallel for schedule(static)
        for (size_t i = 0; i < m; ++i) {
            const q128 *row = A + i * lda;
            q128 sum = QZERO;
            if (incx == 1) sum = or_dot_kernel(row, x, n);
            else for (size_t j = 0; j < n; ++j) sum = g_fma(row[j], x[j * incx], sum);
            y[i * incy] = g_fma(alpha, sum, or_mul(beta, y[i * incy]));
        }
    } else {
        for (size_t i = 0; i < m; ++i) y[i * incy] = or_mul(beta, y[i * incy]);
        /* The reference forks per column j over rows i; every y_i still sees the columns in
           order j = 0..n-1, so rows are independent chains. */
        q128 *ax = (q128 *)malloc(sizeof(q128) * n);
        for (size_t j = 0; j < n; ++j) ax[j] = or_mul(alpha, x[j * incx]);   /* :71 */
        #pragma omp parallel for schedule(static)
        for (size_t i = 0; i < m; ++i) {
            q128 acc = y[i * incy];
            for (size_t j = 0; j < n; ++j) acc = g_fma(A[j * lda + i], ax[j], acc);   /* :79 */
            y[i * incy] = acc;
        }
        free(ax);
    }
}

/* kc of BlockingParams on x86-64 Linux (blocking.hpp:17-67 with constants.hpp:12-13). */
size_t or_blocking_kc(size_t k) {
    size_t l1_budget = 32768 / (2 * 16);
    size_t kc = (l1_budget - 16) / 8;          /* 126 */
    if (kc > 256) kc = 256;
    if (k > 0 && kc > k) kc = k;
    if (kc < 4) kc = 4;
    return kc;
}

/* level3.hpp:187-336.  Every C element is an independent chain:
     m,n,k <= 64 (gemm_simple :187-213): C = fma(alpha, chain_{l<k}, mul(beta, C))
     otherwise (blocked :231-332): for each k-block of kc: t = fresh chain over the block,
       C = fma(alpha, t, mul(first block ? beta : 1, C))  (:297-300, micro-kernel :102-109).
   colmajor selects column-major indexing for A, B and C.  For the blocked column-major case the
   reference writes C through row-major offsets (level3.hpp:295 with :40,106-107,174) and
   overflows its buffer; this restatement follows the layout-correct semantics instead. */
void or_gemm(int colmajor, size_t m, size_t n, size_t k, q128 alpha, const q128 *A, size_t lda,
             const q128 *B, size_t ldb, q128 beta, q128 *C, size_t ldc, size_t kc_override) {
    if (m == 0 || n == 0 || k == 0) return;
    int simple = (m <= 64 && n <= 64 && k <= 64);
    size_t kc = simple ? k : (kc_override ? kc_override : or_blocking_kc(k));
    #pragma omp parallel for collapse(2) schedule(static)
    for (size_t i = 0; i < m; ++i) {
        for (size_t j = 0; j < n; ++j) {
            size_t ci = colmajor ? j * ldc + i : i * ldc + j;
            q128 c = C[ci];
            for (size_t kk = 0; kk < k; kk += kc) {
                size_t ke = kk + kc < k ? kk + kc : k;
                q128 t = QZERO;
                for (size_t l = kk; l < ke; ++l) {
                    q128 a = colmajor ? A[l * lda + i] : A[i * lda + l];
                    q128 b = colmajor ? B[j * ldb + l] : B[l * ldb + j];
                    t = g_fma(a, b, t);
                }
                c = g_fma(alpha, t, or_mul(kk == 0 ? beta : QONE, c));
            }
            C[ci] = c;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* 3. exact accumulation                                                                       */
/* ------------------------------------------------------------------------------------------ */

/* Arbitrary-width two's complement integer over 64-bit words; bit 0 has weight 2^base. */
typedef struct { uint64_t *w; int nw; int base; } bigacc;

/* add (sign ? -1 : 1) * v * 2^(e) where v is a magnitude of nv words */
static void big_addterm(bigacc *A, int sign, const uint64_t *v, int nv, int e) {
    int sh = e - A->base;   /* >= 0 by construction */
    int ws = sh >> 6, bs = sh & 63;
    /* shifted copy */
    uint64_t tmp[16];
    memset(tmp, 0, sizeof(tmp));
    for (int i = 0; i < nv; ++i) {
        tmp[i] |= v[i] << bs;
        if (bs) tmp[i + 1] |= v[i] >> (64 - bs);
    }
    int nt = nv + 1;
    if (!sign) {
        u128 c = 0;
        int i;
        for (i = 0; i < nt && ws + i < A->nw; ++i) { c += (u128)A->w[ws + i] + tmp[i]; A->w[ws + i] = (uint64_t)c; c >>= 64; }
        for (; c && ws + i < A->nw; ++i) { c += A->w[ws + i]; A->w[ws + i] = (uint64_t)c; c >>= 64; }
    } else {
        uint64_t br = 0;
        int i;
        for (i = 0; i < nt && ws + i < A->nw; ++i) {
            uint64_t a = A->w[ws + i], b = tmp[i];
            uint64_t r = a - b - br;
            br = (a < b) || (a == b && br);
            A->w[ws + i] = r;
        }
        for (; br && ws + i < A->nw; ++i) { uint64_t a = A->w[ws + i]; A->w[ws + i] = a - 1; br = (a == 0); }
    }
}

static void mul_words(uint64_t *r, const uint64_t *a, int na, const uint64_t *b, int nb) {
    memset(r, 0, sizeof(uint64_t) * (size_t)(na + nb));
    for (int i = 0; i < na; ++i) {
        u128 c = 0;
        for (int j = 0; j < nb; ++j) { c += (u128)a[i] * b[j] + r[i + j]; r[i + j] = (uint64_t)c; c >>= 64; }
        r[i + nb] = (uint64_t)c;
    }
}

/* Exact value of alpha * sum_i x[i*incx]*y[i*incy] + beta * c, rounded once to binary128
   (round-to-nearest-even) into *rounded.  For each of the ncand candidates, relerr[j] receives
   |cand_j - exact| / |exact| (inf if the exact value is 0 and the candidate is not; 0 if both
   are 0).  Returns 0, or -1 if any input is Inf/NaN (nothing written).  Inputs must be finite. */
int or_exact_dot(size_t n, const q128 *x, ptrdiff_t incx, const q128 *y, ptrdiff_t incy,
                 q128 alpha, q128 beta, q128 c, const q128 *cand, size_t ncand,
                 q128 *rounded, double *relerr) {
    dec Al = decode(alpha), Be = decode(beta), Cc = decode(c);
    if (Al.cls >= C_INF || Be.cls >= C_INF || Cc.cls >= C_INF) return -1;
    /* terms: alpha*x_i*y_i (339 bits) and beta*c (226 bits); find exponent range */
    int emin = 1 << 30, emax = -(1 << 30);
    for (size_t i = 0; i < n; ++i) {
        dec X = decode(x[i * incx]), Y = decode(y[i * incy]);
        if (X.cls >= C_INF || Y.cls >= C_INF) return -1;
        if (X.cls == C_ZERO || Y.cls == C_ZERO || Al.cls == C_ZERO) continue;
        int e = X.exp + Y.exp + Al.exp - 336;
        if (e < emin) emin = e;
        if (e + 340 > emax) emax = e + 340;
    }
    if (Be.cls == C_FIN && Cc.cls == C_FIN) {
        int e = Be.exp + Cc.exp - 224;
        if (e < emin) emin = e;
        if (e + 227 > emax) emax = e + 227;
    }
    int any_term = emin <= emax;
    for (size_t j = 0; j < ncand; ++j) {
        dec D = decode(cand[j]);
        if (D.cls == C_FIN) { int e = D.exp - 112; if (e < emin) emin = e; if (e + 114 > emax) emax = e + 114; }
    }
    if (!any_term) {  /* every term is zero */
        if (rounded) {
            /* IEEE signed zeros as the reference would produce them: the sum of products starts at
               +0 (level1.hpp:18, level2.hpp:31), alpha*(+0) has alpha's sign, beta*c is a zero
               with sign(beta)^sign(c), and -0 + -0 is the only way to a -0 result */
            int sz = Al.sign & (Be.sign ^ Cc.sign);
            *rounded = QZERO;
            rounded->hi = (uint64_t)sz << 63;
        }
        for (size_t j = 0; j < ncand; ++j) {
            dec D = decode(cand[j]);
            relerr[j] = (D.cls == C_ZERO) ? 0.0 : __builtin_inf();
        }
        return 0;
    }
    int bits = emax - emin + 64 + 2 * 64;   /* headroom for n terms and sign */
    bigacc S;
    S.nw = bits / 64 + 2; S.base = emin;
    S.w = (uint64_t *)calloc((size_t)S.nw, 8);
    uint64_t av[2] = {(uint64_t)Al.sig, (uint64_t)(Al.sig >> 64)};
    for (size_t i = 0; i < n; ++i) {
        dec X = decode(x[i * incx]), Y = decode(y[i * incy]);
        if (X.cls == C_ZERO || Y.cls == C_ZERO || Al.cls == C_ZERO) continue;
        uint64_t xv[2] = {(uint64_t)X.sig, (uint64_t)(X.sig >> 64)};
        uint64_t yv[2] = {(uint64_t)Y.sig, (uint64_t)(Y.sig >> 64)};
        uint64_t p[4], q[6];
        mul_words(p, xv, 2, yv, 2);
        mul_words(q, p, 4, av, 2);
        big_addterm(&S, X.sign ^ Y.sign ^ Al.sign, q, 6, X.exp + Y.exp + Al.exp - 336);
    }
    if (Be.cls == C_FIN && Cc.cls == C_FIN) {
        uint64_t bv[2] = {(uint64_t)Be.sig, (uint64_t)(Be.sig >> 64)};
        uint64_t cv[2] = {(uint64_t)Cc.sig, (uint64_t)(Cc.sig >> 64)};
        uint64_t p[4];
        mul_words(p, bv, 2, cv, 2);
        big_addterm(&S, Be.sign ^ Cc.sign, p, 4, Be.exp + Cc.exp - 224);
    }
    /* sign / magnitude of the exact sum */
    int neg = (int)(S.w[S.nw - 1] >> 63);
    uint64_t *mag = (uint64_t *)calloc((size_t)S.nw, 8);
    if (neg) { uint64_t br = 1; for (int i = 0; i < S.nw; ++i) { uint64_t v = ~S.w[i] + br; br = (br && v == 0); mag[i] = v; } }
    else memcpy(mag, S.w, (size_t)S.nw * 8);
    int msb = -1;
    for (int i = S.nw - 1; i >= 0; --i) if (mag[i]) { msb = i * 64 + 63 - __builtin_clzll(mag[i]); break; }
    if (rounded) {
        if (msb < 0) *rounded = QZERO;
        else {
            /* take the top 320 bits into a local frame, jam the rest, round with round_pack */
            uint64_t fr[NW];
            memset(fr, 0, sizeof(fr));
            int lo = msb - (NW * 64 - 8);   /* frame bit 0 <- mag bit lo */
            int jam = 0;
            for (int p = (lo > 0 ? 0 : 0); p <= msb; ++p) {
                int bit = (int)((mag[p >> 6] >> (p & 63)) & 1);
                if (!bit) continue;
                if (p < lo) { jam = 1; continue; }
                int q = p - lo;
                fr[q >> 6] |= 1ull << (q & 63);
            }
            fr[0] |= (uint64_t)jam;
            *rounded = round_pack(neg, fr, S.base + lo);
        }
    }
    for (size_t j = 0; j < ncand; ++j) {
        dec D = decode(cand[j]);
        if (D.cls >= C_INF) { relerr[j] = __builtin_nan(""); continue; }
        /* diff = cand - exact, exactly */
        bigacc Dd; Dd.nw = S.nw; Dd.base = S.base;
        Dd.w = (uint64_t *)malloc((size_t)S.nw * 8);
        memcpy(Dd.w, S.w, (size_t)S.nw * 8);
        /* negate S */
        { uint64_t br = 1; for (int i = 0; i < Dd.nw; ++i) { uint64_t v = ~Dd.w[i] + br; br = (br && v == 0); Dd.w[i] = v; } }
        if (D.cls == C_FIN) {
            uint64_t dv[2] = {(uint64_t)D.sig, (uint64_t)(D.sig >> 64)};
            big_addterm(&Dd, D.sign, dv, 2, D.exp - 112);
        }
        int dneg = (int)(Dd.w[Dd.nw - 1] >> 63);
        if (dneg) { uint64_t br = 1; for (int i = 0; i < Dd.nw; ++i) { uint64_t v = ~Dd.w[i] + br; br = (br && v == 0); Dd.w[i] = v; } }
        /* ratio from the top two nonzero words of each, scaled by their word offsets */
        int tn = -1, td = -1;
        for (int i = Dd.nw - 1; i >= 0; --i) if (Dd.w[i]) { tn = i; break; }
        for (int i = S.nw - 1; i >= 0; --i) if (mag[i]) { td = i; break; }
        if (td < 0) relerr[j] = (tn < 0) ? 0.0 : __builtin_inf();
        else if (tn < 0) relerr[j] = 0.0;
        else {
            long double num = (long double)Dd.w[tn] * 18446744073709551616.0L + (tn > 0 ? (long double)Dd.w[tn - 1] : 0.0L);
            long double den = (long double)mag[td] * 18446744073709551616.0L + (td > 0 ? (long double)mag[td - 1] : 0.0L);
            relerr[j] = (double)ldexpl(num / den, 64 * (tn - td));
        }
        free(Dd.w);
    }
    free(mag);
    free(S.w);
    return 0;
}

/* Batched sampled entries of a GEMM: for each sample s, the exact value of
   alpha * sum_l A(i_s,l) B(l,j_s) + beta * C0(i_s,j_s), rounded, and the relative error of
   cand[s].  Strides describe either layout: A(i,l) = A[i*ars + l*acs], B(l,j) = B[l*brs + j*bcs]. */
int or_exact_gemm_samples(size_t nsamp, const int64_t *is, const int64_t *js, size_t k,
                          q128 alpha, const q128 *A, int64_t ars, int64_t acs,
                          const q128 *B, int64_t brs, int64_t bcs, q128 beta, const q128 *C0,
                          int64_t crs, int64_t ccs, const q128 *cand, q128 *rounded, double *relerr) {
    int bad = 0;
    #pragma omp parallel for schedule(dynamic, 4) reduction(|:bad)
    for (size_t s = 0; s < nsamp; ++s) {
        const q128 *arow = A + is[s] * ars;
        const q128 *bcol = B + js[s] * bcs;
        q128 c0 = C0 ? C0[is[s] * crs + js[s] * ccs] : QZERO;
        if (or_exact_dot(k, arow, acs, bcol, brs, alpha, beta, c0, cand ? &cand[s] : NULL,
                         cand ? 1 : 0, rounded ? &rounded[s] : NULL, relerr ? &relerr[s] : NULL) != 0) bad = 1;
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------------------------ */
/* 4. seeded input generators (same definitions as the HIP library's fill kernels)             */
/* ------------------------------------------------------------------------------------------ */

/* splitmix64 of (seed, index): the counter-based stream used everywhere */
uint64_t or_rand64(uint64_t seed, uint64_t idx) {
    uint64_t z = seed + (idx + 1) * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
/* uniform double in [-1, 1) with 53 random bits: ((r >> 11) * 2^-52) - 1 */
static double u11(uint64_t r) { return (double)(r >> 11) * 0x1p-52 - 1.0; }

/* kind 0 ("full mantissa uniform"): v_i = a + b * 2^-60 summed in binary128 (round to nearest
   even), a = u11(rand(seed, 2i)), b = u11(rand(seed, 2i+1)).
   kind 1 ("log-uniform"): sign from bit 63 of rand(seed, 3i); exponent uniform in
   [lo_exp, hi_exp) from rand(seed, 3i) bits 0..31; fraction 112 bits from rand(seed, 3i+1)
   (48 bits) and rand(seed, 3i+2) (64 bits).
   kind 2 ("cancellation pairs", x vector of the qdot cancellation variant): for even i,
   v_i = (1e20 * u) and v_{i+1} = -(1e20 * u) with u = 1 + (rand(seed, i) >> 12) * 2^-52
   (u in [1, 2), exact in double), formed by an exact binary128 multiply (53 x 47 bits); entries with
   i % 8 >= 2 are instead kind-0 values (the O(1) residue).  The matching y vector repeats
   y_{i+1} = y_i on the pairs so that the pair products cancel exactly.
   kind 4: kind 2 with 1e22 in place of 1e20 (sums that cancel to ~1e-20 of their largest
   products: the ill-conditioned qgemm workload). */
void or_fill_range(int kind, uint64_t seed, size_t start, size_t n, int lo_exp, int hi_exp, q128 *out) {
    #pragma omp parallel for schedule(static)
    for (size_t t = 0; t < n; ++t) {
        const size_t i = start + t;
        q128 *o = out + t;
        if (kind == 0) {
            double a = u11(or_rand64(seed, 2 * i)), b = u11(or_rand64(seed, 2 * i + 1));
            *o = f2q((__float128)a + (__float128)b * 0x1p-60Q);
        } else if (kind == 1) {
            uint64_t r0 = or_rand64(seed, 3 * i), r1 = or_rand64(seed, 3 * i + 1), r2 = or_rand64(seed, 3 * i + 2);
            uint32_t span = (uint32_t)(hi_exp - lo_exp);
            int e = lo_exp + (int)((uint32_t)r0 % span);
            q128 v; v.lo = r2; v.hi = (r0 & (1ull << 63)) | ((uint64_t)(e + BIAS) << 48) | (r1 >> 16);
            *o = v;
        } else {
            size_t p = i & ~(size_t)1;
            if ((p % 8) < 2) {
                double u = 1.0 + (double)(or_rand64(seed, p) >> 12) * 0x1p-52;
                __float128 v = (__float128)u * (__float128)(kind == 4 ? 1e22 : 1e20);
                *o = f2q((i & 1) ? -v : v);
            } else {
                double a = u11(or_rand64(seed, 2 * i)), b = u11(or_rand64(seed, 2 * i + 1));
                *o = f2q((__float128)a + (__float128)b * 0x1p-60Q);
            }
        }
    }
}
/* the first n values of a sequence (or_fill_range writes values start .. start + n - 1) */
void or_fill(int kind, uint64_t seed, size_t n, int lo_exp, int hi_exp, q128 *out) {
    or_fill_range(kind, seed, 0, n, lo_exp, hi_exp, out);
}
/* y vector for the cancellation variant: kind-0 values from the stream seed + 1, with
   y_{i+1} = y_i on each pair (p % 8 < 2) */
void or_fill_pair_y(uint64_t seed, size_t n, q128 *out) {
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
        size_t p = ((i & ~(size_t)1) % 8 < 2) ? (i & ~(size_t)1) : i;
        double a = u11(or_rand64(seed + 1, 2 * p)), b = u11(or_rand64(seed + 1, 2 * p + 1));
        out[i] = f2q((__float128)a + (__float128)b * 0x1p-60Q);
    }
}

/* conversions used by the C ABI check: binary128 -> double, round to nearest even
   (Sleef_cast_to_doubleq1 at c_interface.hpp:30) */
double or_q2d(q128 a) { return (double)q2f(a); }

/* Exact gemv: y_i <- RN(alpha * sum_j A(i,j) x_j + beta * y_i) for every row (colmajor selects
   A(i,j) = A[j*lda + i], else A[i*lda + j]).  Returns -1 if an input is Inf/NaN. */
int or_exact_gemv(int colmajor, size_t m, size_t n, q128 alpha, const q128 *A, size_t lda, const q128 *x,
                  size_t incx, q128 beta, q128 *y, size_t incy) {
    int bad = 0;
    #pragma omp parallel for schedule(dynamic, 8) reduction(|:bad)
    for (size_t i = 0; i < m; ++i) {
        const q128 *row = colmajor ? A + i : A + i * lda;
        ptrdiff_t rs = colmajor ? (ptrdiff_t)lda : 1;
        q128 r;
        if (or_exact_dot(n, row, rs, x, (ptrdiff_t)incx, alpha, beta, y[i * incy], NULL, 0, &r, NULL) != 0) bad = 1;
        else y[i * incy] = r;
    }
    return bad ? -1 : 0;
}
