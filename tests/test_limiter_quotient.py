"""The FP32 limiter quotient of the sweeps (2pm/(p+m): an FP64 reciprocal
estimate, one Newton step, one remainder correction, then rounding to FP32)
against the IEEE double division the reference performs (tvd1.c:71-81,
tvdb.c:72-76).  The corrected double quotient itself equals the IEEE one
(N/D is never within 2^-78 of a double rounding midpoint: N has 48
significant bits, D 24), which a dropped correction step breaks at once.  The reciprocal estimate is modelled with a planted relative
error of up to 2^-24 (the instruction's measured bound is 2^-24.39; the GPU
test test_rcp_f64_bound_of_the_limiter_quotient rechecks it on the device),
including both extremes; fma is evaluated exactly with fractions.
"""
from fractions import Fraction

import numpy as np


def fma(a, b, c):
    if not all(np.isfinite(x) for x in (a, b, c)):
        return a * b + c  # infinities and NaN: no rounding to model
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def ieee(p, m):
    p, m = np.float32(p), np.float32(m)
    if not np.float32(p * m) > 0:
        return np.float32(0)
    s = np.float32(p + m)
    return np.float32(2.0 * float(p) * float(m) / float(s))


def fast_double(p, m, delta, clamp=False):
    """the double quotient before the final rounding to FP32; clamp: the
    branch-free form, an infinite p + m entering as +-2^1000"""
    s = np.float32(p + m)
    N, D = 2.0 * float(p) * float(m), float(s)
    if clamp:
        D = min(max(D, -2.0 ** 1000), 2.0 ** 1000)
    y = float(Fraction(1) / Fraction(D) * (1 + delta))  # planted estimate error
    e = fma(-D, y, 1.0)
    y = fma(y, e, y)
    q = N * y
    r = fma(-D, q, N)
    return fma(r, y, q)


def fast(p, m, delta, clamp=False):
    p, m = np.float32(p), np.float32(m)
    if not np.float32(p * m) > 0:
        return np.float32(0)
    return np.float32(fast_double(p, m, delta, clamp))


def same32(a, b):
    return a.view(np.uint32) == b.view(np.uint32) or (np.isnan(a) and np.isnan(b))


def test_limiter_quotient_equals_ieee_with_planted_reciprocal_error():
    rng = np.random.default_rng(11)
    n = 6000
    # float bit patterns: random exponents across the range, plus pairs near
    # each other (sums close to 2p) and pairs with equal exponents
    pb = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    mb = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    mb[::3] = (mb[::3] & np.uint32(0x807FFFFF)) | (pb[::3] & np.uint32(0x7F800000))
    mb[1::3] = (pb[1::3] + rng.integers(-40, 40, len(mb[1::3])).astype(np.uint32)) & np.uint32(0xFFFFFFFF)
    p = pb.view(np.float32)
    m = mb.view(np.float32)
    eps = Fraction(1, 1 << 24)
    deltas = [eps, -eps, Fraction(0)]
    checked = 0
    with np.errstate(all="ignore"):
        for i in range(n):
            if not (np.isfinite(p[i]) and np.isfinite(m[i])) or not np.isfinite(np.float32(p[i] + m[i])):
                continue
            want = ieee(p[i], m[i])
            d = deltas[i % 3] if i % 4 else Fraction(int(rng.integers(-(1 << 20), 1 << 20)), 1 << 44)
            got = fast(p[i], m[i], d)
            assert want.view(np.uint32) == got.view(np.uint32), (p[i], m[i], d)
            if np.float32(p[i] * m[i]) > 0:
                s32 = np.float32(p[i] + m[i])
                assert fast_double(p[i], m[i], d) == 2.0 * float(p[i]) * float(m[i]) / float(s32)
                checked += 1
    assert checked > n // 4


def test_limiter_quotient_exact_quotients():
    # quotients that are exact doubles
    cases = [(1.0, 1.0), (3.0, 1.0), (0.75, 0.25), (2.0 ** -140, 2.0 ** -140), (1.5, 0.5), (6.0, 2.0),
             (float(np.float32(1 + 2 ** -23)), 1.0), (2.0 ** 100, 2.0 ** 100), (-5.0, -3.0)]
    eps = Fraction(1, 1 << 24)
    with np.errstate(all="ignore"):
        for p, m in cases:
            for d in (eps, -eps, Fraction(0)):
                assert ieee(p, m).view(np.uint32) == fast(p, m, d).view(np.uint32), (p, m, d)


def test_branch_free_quotient_on_overflowing_sums():
    """the branch-free form (streamed tvdb): every pair whose FP32 sum
    overflows or holds an infinite operand gives the IEEE result, a signed
    zero or NaN; a clamp bound small enough to leave a nonzero float (2^200)
    fails"""
    big = [np.float32(x) for x in (2.0 ** 127, 3.0e38, 1.7e38, np.finfo(np.float32).max)]
    rng = np.random.default_rng(5)
    pairs = [(a, b) for a in big for b in big]
    pairs += [(np.float32(2.0 ** 127 * (1 + rng.random())), np.float32(2.0 ** 127 * (1 + rng.random())))
              for _ in range(200)]
    pairs += [(np.float32(np.inf), np.float32(x)) for x in (1.0, 2.0 ** -149, 3.0e38)]
    pairs += [(np.float32(np.inf), np.float32(np.inf))]
    pairs += [(-a, -b) for a, b in pairs]
    eps = Fraction(1, 1 << 24)
    with np.errstate(all="ignore"):
        pairs = [(p, m) for p, m in pairs if not np.isfinite(np.float32(p + m))]
        assert len(pairs) > 300
        for p, m in pairs:
            want = ieee(p, m)
            for d in (eps, -eps, Fraction(0)):
                assert same32(want, fast(p, m, d, clamp=True)), (p, m, d)
        # a bound that is too small leaves 2pm/D representable: the check
        # above is not vacuous
        p, m = big[0], big[0]
        N, D = 2.0 * float(p) * float(m), 2.0 ** 200
        assert np.float32(N / D) != 0


def test_limiter_test_from_the_fp64_numerator():
    """p*m > 0 in FP32 (tvd1.c:71, tvdb.c:72) equals 2.0*(double)p*(double)m
    > 2^-149 (the sweeps form that exact product anyway): the FP32 product is
    positive exactly when the real one exceeds 2^-150, the tie at 2^-150
    rounding to zero.  Random bit patterns, products around the underflow
    threshold (both sides of 2^-150 and the tie itself), FP32 overflow,
    infinities, zeros and NaN.  A threshold 2^-30 off either way fails."""
    rng = np.random.default_rng(17)
    bits = rng.integers(0, 1 << 32, size=(20000, 2), dtype=np.uint64).astype(np.uint32)
    pairs = [(bits[i, 0].view(np.float32), bits[i, 1].view(np.float32)) for i in range(len(bits))]
    for _ in range(4000):
        e = int(rng.integers(-149, 0))
        p = np.float32(np.ldexp(1 + rng.random(), e))
        m = np.float32(np.ldexp(1 + rng.random(), -150 - e + int(rng.integers(-2, 2))))
        pairs.append((p, m))
    tie = [(np.float32(2.0 ** -75), np.float32(2.0 ** -75)), (np.float32(2.0 ** -149), np.float32(0.5)),
           (np.float32(2.0 ** -149), np.float32(0.75)), (np.float32(2.0 ** -149), np.float32(0.5000001)),
           (np.float32(3.0e38), np.float32(2.0)), (np.float32(np.inf), np.float32(1.0)),
           (np.float32(np.inf), np.float32(0.0)), (np.float32(-0.0), np.float32(-0.0)),
           (np.float32(np.nan), np.float32(1.0)), (np.float32(-np.inf), np.float32(-2.0 ** -149))]
    # just above the tie: for each p the smallest m with p*m > 2^-150, so that
    # the product exceeds 2^-150 by as little as two floats allow
    for x in rng.random(3000):
        p = np.float32(np.ldexp(1 + x, -75))
        m = np.float32(2.0 ** -150 / float(p))
        while float(p) * float(m) <= 2.0 ** -150:
            m = np.nextafter(m, np.float32(1))
        tie.append((p, m))
    pairs += tie
    lo, hi = 2.0 ** -149 * (1 - 2.0 ** -30), 2.0 ** -149 * (1 + 2.0 ** -30)
    miss_lo = miss_hi = 0
    with np.errstate(all="ignore"):
        for p, m in pairs:
            want = bool(np.float32(p) * np.float32(m) > 0)
            N = 2.0 * float(p) * float(m)
            assert (N > 2.0 ** -149) == want, (p, m)
            miss_lo += (N > lo) != want
            miss_hi += (N > hi) != want
    assert miss_lo > 0 and miss_hi > 0
