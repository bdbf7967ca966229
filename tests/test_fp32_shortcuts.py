"""The one-operation FP32 forms the sweeps use for FP64 sub-expressions of
the reference's FP32 code, against the reference expression evaluated in
double as the C code does (numpy float64 is IEEE on the host), on float
pairs that cover every exponent gap, cancellation, the subnormal range,
overflow and the special values:

* (float)((double)a - (double)x / 2.0)   (tvdb.c:41 b1, mhdflux.c:41 p = ps - B/2.0)
  equals the FP32 fma(x, -0.5, a), i.e. a - x/2 rounded once to float;
* (float)((double)x / 2.0)               (tvdb.c:59-66 wp, wm; mhdflux.c:37)
  equals the FP32 product x * 0.5.

fma is modelled exactly: the real value, then one correct rounding to
float32 (ties to even, subnormals kept, overflow to inf), by integer
arithmetic, so no double rounding enters the model itself.
"""
from fractions import Fraction

import numpy as np

F32 = np.float32


def round_f32(z):
    """the float32 nearest to the rational z (ties to even)"""
    if z == 0:
        return 0.0
    s = -1.0 if z < 0 else 1.0
    z = abs(z)
    e = z.numerator.bit_length() - z.denominator.bit_length()
    if Fraction(2) ** e > z:
        e -= 1
    q = Fraction(2) ** (max(e, -126) - 23)  # quantum of the binade (subnormals: 2^-149)
    n, r = divmod(z / q, 1)
    n = int(n)
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and n % 2 == 1):
        n += 1
    v = n * q
    if v >= Fraction(2) ** 128:
        return s * np.inf
    return s * float(v)


def fma_f32(a, b, c):
    """IEEE float32 fma of float32 operands"""
    a, b, c = float(a), float(b), float(c)
    if not all(np.isfinite(t) for t in (a, b, c)):
        return F32(a * b + c)  # infinities and NaN: nothing to round
    z = Fraction(a) * Fraction(b) + Fraction(c)
    if z == 0:
        # exact zero: -0 only when both addends are -0 (round to nearest)
        p = a * b
        neg = (np.signbit(p) if p == 0 else p < 0) and np.signbit(c) and c == 0
        return F32(-0.0) if neg else F32(0.0)
    return F32(round_f32(z))


def same(x, y):
    x, y = F32(x), F32(y)
    if np.isnan(x) or np.isnan(y):
        return bool(np.isnan(x) and np.isnan(y))
    return x.view(np.uint32) == y.view(np.uint32)


def pairs(rng, n):
    """float32 pairs: random bit patterns, then exponent gaps 0..40 with
    random significands and signs, near-cancellations, subnormals and values
    near the overflow threshold"""
    out = []
    bits = rng.integers(0, 1 << 32, size=(n, 2), dtype=np.uint64).astype(np.uint32)
    out += [(F32(bits[i, 0].view(np.float32)), F32(bits[i, 1].view(np.float32))) for i in range(n)]
    for _ in range(n):
        ea = int(rng.integers(-149, 128))
        gap = int(rng.integers(-40, 41))
        eb = min(max(ea - gap, -150), 128)
        a = F32(np.ldexp(1 + rng.random(), ea) * rng.choice([-1, 1]))
        x = F32(np.ldexp(1 + rng.random(), eb) * rng.choice([-1, 1]))
        out.append((a, x))
    for _ in range(n // 4):
        a = F32(np.ldexp(1 + rng.random(), int(rng.integers(-140, 120))))
        x = F32(2 * a * F32(1 + rng.integers(-50, 50) * 2.0 ** -23))  # x/2 close to a
        out.append((a, x))
    big = np.finfo(np.float32).max
    out += [(F32(big), F32(-big)), (F32(big), F32(-2.0 ** 104)), (F32(-big), F32(2.0 ** 103)),
            (F32(2.0 ** -149), F32(2.0 ** -149)), (F32(-2.0 ** -149), F32(3 * 2.0 ** -149)),
            (F32(0.0), F32(-0.0)), (F32(-0.0), F32(0.0)), (F32(-0.0), F32(-0.0)),
            (F32(np.inf), F32(1.0)), (F32(1.0), F32(np.inf)), (F32(np.inf), F32(np.inf)),
            (F32(np.nan), F32(1.0)), (F32(1.0), F32(-np.inf))]
    return out


def test_round_f32_is_single_rounding():
    # 1 + 2^-24 + 2^-60 rounds up to 1 + 2^-23 in one step, but to 1 through
    # a double (the 2^-60 is lost, the tie goes to even): the model must not
    # round twice
    z = 1 + Fraction(1, 1 << 24) + Fraction(1, 1 << 60)
    assert round_f32(z) == 1 + 2.0 ** -23
    assert F32(float(z)) == F32(1.0)
    assert round_f32(Fraction(2) ** 128) == np.inf
    assert round_f32(Fraction(3, 2 ** 150)) == 2.0 ** -148  # 1.5 quanta: tie to even (2)
    assert round_f32(Fraction(3, 2 ** 151)) == 2.0 ** -149  # 0.75 quanta
    assert round_f32(Fraction(1, 2 ** 150)) == 0.0  # half the smallest subnormal: tie to 0


def test_difference_with_half_is_one_fma():
    rng = np.random.default_rng(3)
    checked = 0
    with np.errstate(all="ignore"):
        for a, x in pairs(rng, 4000):
            ref = F32(np.float64(a) - np.float64(x) / 2.0)
            got = fma_f32(x, F32(-0.5), a)
            assert same(ref, got), (a, x, ref, got)
            checked += 1
    assert checked > 9000


def test_halving_is_one_product():
    rng = np.random.default_rng(4)
    bits = rng.integers(0, 1 << 32, size=20000, dtype=np.uint64).astype(np.uint32)
    xs = list(bits.view(np.float32)) + [F32(2.0 ** -149), F32(3 * 2.0 ** -149), F32(-5 * 2.0 ** -149)]
    with np.errstate(all="ignore"):
        for x in xs:
            ref = F32(np.float64(x) / 2.0)
            got = F32(x) * F32(0.5)
            assert same(ref, got), x
