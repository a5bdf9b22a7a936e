"""Shared test helpers: exact rational view of binary128 values, and GPU buffer plumbing."""
from fractions import Fraction

import numpy as np

BIAS = 16383


def to_fraction(lo, hi):
    """Exact value of a finite binary128 given as its (lo, hi) words."""
    lo, hi = int(lo), int(hi)
    s = hi >> 63
    be = (hi >> 48) & 0x7FFF
    f = ((hi & ((1 << 48) - 1)) << 64) | lo
    assert be != 0x7FFF, "finite values only"
    if be == 0:
        v = Fraction(f, 1) * Fraction(2) ** (1 - BIAS - 112)
    else:
        v = Fraction(f | (1 << 112), 1) * Fraction(2) ** (be - BIAS - 112)
    return -v if s else v


def round_fraction(v):
    """Round an exact rational to binary128, round-to-nearest-even (normal/subnormal range;
    overflow -> inf).  Returns (lo, hi).  -0 is not produced (exact zero -> +0)."""
    if v == 0:
        return (0, 0)
    s = 1 if v < 0 else 0
    a = -v if s else v
    # exponent of the leading bit: 2^e <= a < 2^(e+1)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    elif Fraction(2) ** (e + 1) <= a:
        e += 1
    emin = 1 - BIAS
    q = max(e, emin) - 112  # exponent of the ulp
    scaled = a / (Fraction(2) ** q)
    m = scaled.numerator // scaled.denominator
    rem = scaled - m
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and (m & 1)):
        m += 1
    if m >> 113:
        m >>= 1
        q += 1
    if m >> 112:
        be = q + 112 + BIAS
        if be >= 0x7FFF:
            return (0, (s << 63) | (0x7FFF << 48))
        frac = m & ((1 << 112) - 1)
    else:
        be = 0
        frac = m
    return (frac & ((1 << 64) - 1), (s << 63) | (be << 48) | (frac >> 64))


def sqrt_rn(lo, hi):
    """Correctly rounded (nearest even) binary128 square root of a finite value >= 0, computed
    exactly with integers: returns (lo, hi)."""
    import math
    v = to_fraction(lo, hi)
    assert v >= 0
    if v == 0:
        return (int(lo), int(hi))
    # v = num / den; sqrt(v) scaled by 2^k with k large enough for > 120 significant bits
    k = 256 - (v.numerator.bit_length() - v.denominator.bit_length()) // 2
    X = (v.numerator << (2 * k)) // v.denominator if k >= 0 else v.numerator // (v.denominator << (-2 * k))
    exact_div = ((v.numerator << (2 * k)) % v.denominator == 0) if k >= 0 else True
    q = math.isqrt(X)
    exact = exact_div and q * q == X
    # sqrt(v) = (q + f) 2^-k, 0 <= f < 1 (f = 0 iff exact); value to round: (q + f) 2^-k
    return round_fraction(Fraction(q, 1) / (Fraction(2) ** k) + (0 if exact else Fraction(1, 2 ** (k + 400))))


def pair(a):
    a = np.asarray(a, dtype=np.uint64).reshape(2)
    return (int(a[0]), int(a[1]))


def rows(a):
    """(n, 2) uint64 array -> list of (lo, hi) tuples."""
    a = np.asarray(a, dtype=np.uint64).reshape(-1, 2)
    return [(int(r[0]), int(r[1])) for r in a]


# ------------------------------------------------------------------------------------------------
# device buffers (torch is only plumbing for device memory here)
# ------------------------------------------------------------------------------------------------

def to_dev(a):
    import torch
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return torch.from_numpy(a.view(np.int64)).to("cuda")


def from_dev(t):
    return t.cpu().numpy().view(np.uint64)


def dev_empty(n):
    import torch
    return torch.empty((n, 2), dtype=torch.int64, device="cuda")
