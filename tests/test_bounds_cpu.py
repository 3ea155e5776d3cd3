"""The start threshold's float32 upper bound of a squared distance, checked against exact rational arithmetic.

The device computes a = sum of fma(t, t, a) over t = fl32(q_i - x_i) in float32 and certifies
S <= (a + (d + 1) 2^-125)(1 + (d + 4) 2^-23), evaluated in float64 and rounded up to float32. This test
recomputes a in float32 (an fma emulated through float64: t*t is exact there, and the float64 sum rounded to
float32 differs from a single rounding by at most a few float32 units in the last place, which the bound's slack
covers), forms the same certified value, and compares it with the exact S of the float32 inputs as a fraction.
"""
from fractions import Fraction

import numpy as np


def certified(q, x):
    d = q.shape[0]
    acc = np.float32(0.0)
    with np.errstate(over="ignore", under="ignore", invalid="ignore"):
        for qi, xi in zip(q, x):
            t = np.float32(qi - xi)
            acc = np.float32(np.float64(t) * np.float64(t) + np.float64(acc))
    upper = (float(acc) + (d + 1) * 2.0 ** -125) * (1.0 + (d + 4) * 2.0 ** -23)
    return np.nextafter(np.float32(upper), np.float32(np.inf)) if np.isfinite(upper) else np.float32(np.inf)


def exact(q, x):
    return sum((Fraction(float(a)) - Fraction(float(b))) ** 2 for a, b in zip(q, x))


def test_certified_bound_covers_exact_distance():
    rng = np.random.default_rng(77)
    cases = []
    for d in (1, 3, 32, 96, 128, 784):
        cases.append((rng.normal(size=d), rng.normal(size=d)))
        base = rng.normal(size=d) * 1e4  # cancellation: q and x agree in their leading bits
        cases.append((base + rng.normal(size=d) * 1e-3, base + rng.normal(size=d) * 1e-3))
        cases.append((rng.normal(size=d) * 1e-40, rng.normal(size=d) * 1e-40))  # subnormal differences
        cases.append((rng.normal(size=d) * 1e-19, np.zeros(d)))  # squares below the normal range
        cases.append((rng.normal(size=d) * 1e18, rng.normal(size=d) * 1e18))  # sums near float32 overflow
    for q, x in cases:
        q, x = q.astype(np.float32), x.astype(np.float32)
        u = certified(q, x)
        s = exact(q, x)
        if np.isinf(u):
            continue  # an overflowed bound is +inf, which bounds everything
        assert Fraction(float(u)) >= s, (q.shape[0], float(u), float(s))
        if s > Fraction(2.0 ** -100):
            assert Fraction(float(u)) <= s * Fraction(1 + 2 ** -10)  # and it is not loose


def test_bound_is_tight_enough_to_matter():
    """On ordinary data the certified value exceeds the exact distance by about d * 2^-23 relative."""
    rng = np.random.default_rng(5)
    q, x = rng.normal(size=128).astype(np.float32), rng.normal(size=128).astype(np.float32)
    rel = float(Fraction(float(certified(q, x))) / exact(q, x)) - 1.0
    assert 0.0 <= rel < 200 * 2.0 ** -23
