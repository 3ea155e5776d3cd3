"""Adversarial radius inputs, generated here from seeds.

Restates the case families of the reference's radius oracle
(tests/_radius_oracle.py:740-750: lattice, duplicates, permuted, offset,
exponent, r2 sweep, hubs, label-sorted, gaussian) with this file's own
constructions and seeds. Each case stresses one way a rounded comparison
with R2 = fl64(r * r) goes wrong: points exactly on the shell, float64
summation order, coarse float32 grids far from zero, subnormal and huge
coordinates, and radii at the edges of the float64 range.
"""
from __future__ import annotations

import math
from fractions import Fraction
from typing import NamedTuple

import numpy as np

F32_MAX = float(np.finfo(np.float32).max)
TINY = 2.0 ** -149


class Case(NamedTuple):
    X: np.ndarray
    Q: np.ndarray
    radius: float
    name: str
    y: np.ndarray | None = None


def exact32(values) -> np.ndarray:
    """float32 array of values that must be representable exactly."""
    v = np.asarray(values, np.float64)
    out = v.astype(np.float32)
    assert np.array_equal(out.astype(np.float64), v), "value not exact in float32"
    return out


def exact_s(q, x) -> Fraction:
    return sum((Fraction(float(a)) - Fraction(float(b))) ** 2 for a, b in zip(q, x))


def radius_near(S: Fraction, side: str) -> float:
    """A float64 r whose fl64(r*r) is the largest value < S ("below"),
    the smallest value >= S ("above"), or exactly S ("equal")."""
    r0 = math.sqrt(float(S))
    cands = []
    for steps in range(-6, 7):
        r = r0
        for _ in range(abs(steps)):
            r = math.nextafter(r, math.inf if steps > 0 else -math.inf)
        cands.append((Fraction(r * r), r))
    cands.sort()
    if side == "below":
        return max((c for c in cands if c[0] < S), key=lambda c: c[0])[1]
    if side == "above":
        return min((c for c in cands if c[0] >= S), key=lambda c: c[0])[1]
    for R2, r in cands:
        if R2 == S:
            return r
    raise ValueError("no radius squares exactly to S")


def typical_radius(X, Q, count):
    """Median distance to the count-th neighbour over a query sample."""
    qs = Q[:: max(1, len(Q) // 48)].astype(np.float64)
    s = ((qs[:, None, :] - X.astype(np.float64)[None]) ** 2).sum(-1)
    s.sort(axis=1)
    return float(math.sqrt(float(np.median(s[:, min(count, s.shape[1]) - 1]))))


def grid(values, d):
    axes = np.meshgrid(*([np.asarray(values, np.float64)] * d), indexing="ij")
    return np.stack(axes, -1).reshape(-1, d)


def lattice():
    r = np.random.default_rng(4101)
    out = []
    X, Q = exact32(grid(range(-5, 6), 2)), exact32(grid(range(-6, 7), 2))
    out += [Case(X, Q, 2.0, "z2-r2"), Case(X, Q, math.sqrt(5.0), "z2-sqrt5"), Case(X, Q, 3.0, "z2-r3")]
    X, Q = exact32(grid(range(-3, 4), 3)), exact32(r.integers(-4, 5, (120, 3)))
    out += [Case(X, Q, math.sqrt(v), f"z3-sqrt{v}") for v in (2.0, 3.0, 6.0)]
    X, Q = exact32(r.integers(-2, 3, (1400, 6))), exact32(r.integers(-3, 4, (130, 6)))
    out.append(Case(X, Q, 2.0, "z6-r2"))
    # dyadic lattice: (3, 4, 0) * 2^-10 lies on the shell of radius 5 * 2^-10
    X = exact32(r.integers(-20, 21, (800, 3)) * 2.0 ** -10)
    Q = exact32(r.integers(-22, 23, (120, 3)) * 2.0 ** -10)
    out.append(Case(X, Q, 5 * 2.0 ** -10, "dyadic-r5"))
    # integers just below 2^24 (float32 still exact): 3-4-5 triangles
    X = exact32(2.0 ** 24 - 64 + r.integers(-25, 26, (700, 2)))
    Q = exact32(2.0 ** 24 - 64 + r.integers(-27, 28, (120, 2)))
    out.append(Case(X, Q, 5.0, "near-2^24-r5"))
    return out


def duplicates():
    r = np.random.default_rng(4201)
    base = r.normal(size=(120, 6)).astype(np.float32)
    X = np.repeat(base, r.integers(1, 7, 120), axis=0)
    X = X[r.permutation(len(X))]
    Q = np.concatenate([base[:50], r.normal(size=(30, 6)).astype(np.float32)])
    out = [Case(X, Q, 0.0, "dup-r0"), Case(X, Q, 0.9, "dup-r0.9")]
    row = exact32([[2.5, -1.75, 4.0, 0.5]])
    X = np.repeat(row, 256, axis=0)
    bumped = row.copy()
    bumped[0, 1] = np.nextafter(np.float32(-1.75), np.float32(0))
    Q = np.concatenate([X[:3], bumped, np.zeros((1, 4), np.float32), r.normal(size=(6, 4)).astype(np.float32)])
    out += [Case(X, Q, 0.0, "same-r0"), Case(X, Q, 2.0, "same-r2")]
    X = np.zeros((180, 5), np.float32)
    Q = exact32([[0] * 5, [TINY, 0, 0, 0, 0], [0, -TINY, TINY, 0, 0], [1, 0, 1, 0, 1]])
    out += [Case(X, Q, 0.0, "zeros-r0"), Case(X, Q, TINY, "zeros-rtiny")]
    return out


def signed_perms(delta, count, r):
    delta = np.asarray(delta, np.float64)
    rows = np.empty((count, len(delta)))
    for i in range(count):
        rows[i] = delta[r.permutation(len(delta))] * r.choice([-1.0, 1.0], len(delta))
    return rows


def permuted():
    """Rows at signed permutations of one offset vector: identical exact S,
    float64 sums that differ by summation order."""
    r = np.random.default_rng(4301)
    out = []

    def build(delta, radius, name):
        d = len(delta)
        unit = max(delta) * 2.0 ** -23  # float32 spacing at the big coordinate
        rows, qs = [], []
        for off in (np.zeros(d), unit * r.integers(-15, 16, d)):
            qs.append(off)
            rows.append(off + signed_perms(delta, 50, r))
            lone = np.zeros(d)
            lone[r.integers(d)] = max(delta)
            rows.append((off + lone)[None])
        out.append(Case(exact32(np.concatenate(rows)), exact32(np.stack(qs)), radius, name))

    # S = 2^58 + 100 > R2 = 2^58; adding the big square first loses every 25
    build([2.0 ** 29, 5, 5, 5, 5], 2.0 ** 29, "perm-false-in")
    # S = 2^60 + 432 <= R2 = 2^60 + 512; the big-first float64 sum is 2^60 + 768
    build([2.0 ** 30, 12, 12, 12], 2.0 ** 30 + 2.0 ** -22, "perm-false-out")
    # S == R2 == 2^60 + 512 exactly; the big-first sum is again 2^60 + 768
    build([2.0 ** 30, 12, 12, 12, 8, 4], 2.0 ** 30 + 2.0 ** -22, "perm-tie")
    mant = r.integers(2 ** 23, 2 ** 24, 12).astype(np.float64)
    delta = mant * 2.0 ** (r.integers(-18, 19, 12) - 23)
    X = exact32(signed_perms(delta, 250, r))
    S = exact_s(np.zeros(12), X[0])
    Q = np.zeros((3, 12), np.float32)
    Q[1], Q[2] = X[0], X[1]
    out += [Case(X, Q, radius_near(S, "below"), "perm-mixed-below"),
            Case(X, Q, radius_near(S, "above"), "perm-mixed-above")]
    return out


def offsets():
    r = np.random.default_rng(4401)
    out = []
    for off, sig, tag in [(3e3, 0.02, "3e3"), (2e6, 3.0, "2e6"), (5e11, 1e6, "5e11")]:
        for d in (4, 9):
            X = (off + sig * r.normal(size=(900, d))).astype(np.float32)
            Q = (off + sig * r.normal(size=(120, d))).astype(np.float32)
            out.append(Case(X, Q, typical_radius(X, Q, 12), f"offset-{tag}-d{d}"))
    # at 5e11 the float32 spacing is 2^15; a 5-step radius has (3, 4, 0) on the shell
    step = 2.0 ** 15
    X = (5e11 + 2 * step * r.normal(size=(400, 3))).astype(np.float32)
    Q = (5e11 + 2 * step * r.normal(size=(90, 3))).astype(np.float32)
    out.append(Case(X, Q, 5 * step, "offset-5e11-tie"))
    return out


def exponents():
    r = np.random.default_rng(4501)
    out = []
    X = exact32(r.integers(-10, 11, (500, 3)) * TINY)
    Q = exact32(r.integers(-12, 13, (100, 3)) * TINY)
    out.append(Case(X, Q, 5 * TINY, "subnormal-r5"))

    def rand32(shape, lo, hi):
        mant = r.integers(2 ** 23, 2 ** 24, shape).astype(np.float64)
        with np.errstate(under="ignore"):
            return (r.choice([-1.0, 1.0], shape) * mant * 2.0 ** (r.integers(lo, hi + 1, shape) - 23)).astype(np.float32)

    X = rand32((400, 4), -149, -112)
    Q = np.concatenate([X[:30], rand32((50, 4), -149, -112)])
    Q[:30, 1] = np.nextafter(Q[:30, 1], np.float32(np.inf))
    out.append(Case(X, Q, typical_radius(X, Q, 3), "tiny-mixed"))
    base = 2.0 ** -126
    X = exact32(base * (1 + r.integers(-2 ** 21, 2 ** 21, (350, 2)) * 2.0 ** -22))
    Q = exact32(base * (1 + r.integers(-2 ** 21, 2 ** 21, (90, 2)) * 2.0 ** -22))
    out.append(Case(X, Q, typical_radius(X, Q, 5), "min-normal"))
    X = (r.choice([-1.0, 1.0], (350, 3)) * F32_MAX * r.uniform(0.3, 1.0, (350, 3))).astype(np.float32)
    X[:3] = exact32([[F32_MAX, F32_MAX, -F32_MAX], [-F32_MAX, -F32_MAX, F32_MAX], [0, 0, 0]])
    Q = np.concatenate([X[:25], (r.choice([-1.0, 1.0], (40, 3)) * F32_MAX
                                 * r.uniform(0.3, 1.0, (40, 3))).astype(np.float32)])
    out += [Case(X, Q, 2e38, "huge-r2e38"), Case(X, Q, 5e38, "huge-r5e38")]
    X = rand32((450, 8), -138, 120)
    Q = X[:50].copy()
    col = r.integers(0, 8, 50)
    Q[np.arange(50), col] = np.nextafter(Q[np.arange(50), col], np.float32(-np.inf))
    gap = np.abs(Q[np.arange(50), col].astype(np.float64) - X[np.arange(50), col].astype(np.float64))
    Q = np.concatenate([Q, rand32((30, 8), -138, 120)])
    out.append(Case(X, Q, float(np.median(gap)), "mixed-exponents"))
    return out


def r2_sweep():
    r = np.random.default_rng(4601)
    X = np.concatenate([np.zeros((3, 3)), r.integers(-2, 3, (30, 3)) * TINY, r.integers(-3, 4, (25, 3)) * 2.0 ** -63,
                        r.normal(size=(50, 3)), r.normal(size=(25, 3)) * 1e32]).astype(np.float32)
    X = np.concatenate([X, X[::5]])
    X = X[r.permutation(len(X))]
    Q = np.concatenate([X[::4], r.normal(size=(8, 3)).astype(np.float32)])
    radii = [(0.0, "zero"), (2.0 ** -63, "2^-126"), (TINY, "2^-298"), (3e-165, "subnormal-r2"), (5e-324, "underflow"),
             (1.25, "normal"), (1e32, "huge"), (1e45, "1e90"), (1e150, "1e300"), (1e180, "overflow"),
             (math.inf, "inf")]
    out = [Case(X, Q, v, f"r2-{tag}") for v, tag in radii]
    # safe rows and queries beside rows outside the screen's safe range (their
    # +-1e30 values cancel in the centre): at r = 1e45 every row is a member,
    # the unsafe ones exactly once
    huge = (r.normal(size=(6, 3)) * 1e30).astype(np.float32)
    X = np.concatenate([r.normal(size=(60, 3)).astype(np.float32), huge, -huge])
    Q = r.normal(size=(20, 3)).astype(np.float32)
    out += [Case(X, Q, 1e45, "unsafe-rows-r1e45"), Case(X, Q, 2.0, "unsafe-rows-r2")]
    return out


def hubs():
    r = np.random.default_rng(4701)
    out = []
    for d in (2, 6):
        k = 30
        w = 1.0 / (1.0 + np.arange(k)) ** 1.4
        sizes = np.maximum(1, np.round(1400 * w / w.sum())).astype(int)
        spread = 10.0 ** r.uniform(-3, 0, k)
        centres = r.uniform(-8, 8, (k, d))
        X = np.concatenate([centres[c] + spread[c] * r.normal(size=(sizes[c], d)) for c in range(k)]).astype(np.float32)
        X = X[r.permutation(len(X))]
        Q = np.concatenate([centres[r.integers(0, 4, 80)] + 0.01 * r.normal(size=(80, d)),
                            r.uniform(-8, 8, (40, d))]).astype(np.float32)
        out.append(Case(X, Q, 0.04 if d == 2 else 0.15, f"hubs-d{d}"))
    return out


def blobs(r, n, m, d, classes=5, spread=3.0):
    centres = r.normal(scale=spread, size=(classes, d))
    y = r.integers(0, classes, n)
    X = (centres[y] + r.normal(size=(n, d))).astype(np.float32)
    Q = (centres[r.integers(0, classes, m)] + r.normal(size=(m, d))).astype(np.float32)
    return X, Q, y


def label_sorted():
    r = np.random.default_rng(4801)
    X, Q, y = blobs(r, 1100, 150, 4)
    order = np.argsort(y, kind="stable")
    return [Case(X[order], Q, typical_radius(X, Q, 12), "label-sorted", y[order])]


def gaussian():
    out = []
    for d in (1, 2, 3, 8, 16, 64, 100, 128):
        r = np.random.default_rng(4900 + d)
        X = r.normal(size=(1300, d)).astype(np.float32)
        Q = r.normal(size=(160, d)).astype(np.float32)
        out.append(Case(X, Q, typical_radius(X, Q, 10), f"gauss-d{d}"))
        Xb, Qb, yb = blobs(r, 1300, 160, d)
        out.append(Case(Xb, Qb, typical_radius(Xb, Qb, 12), f"blobs-d{d}", yb))
        if d in (3, 16, 64):
            s = ((Q[0].astype(np.float64) - X.astype(np.float64)) ** 2).sum(1)
            j = int(np.argsort(s)[8])
            S = exact_s(Q[0], X[j])
            out.append(Case(X, Q, radius_near(S, "below"), f"gauss-d{d}-edge-below"))
            out.append(Case(X, Q, radius_near(S, "above"), f"gauss-d{d}-edge-above"))
    return out


FAMILIES = [lattice, duplicates, permuted, offsets, exponents, r2_sweep, hubs, label_sorted, gaussian]


def all_cases() -> list[Case]:
    cases = []
    for fam in FAMILIES:
        cases += fam()
    names = [c.name for c in cases]
    assert len(set(names)) == len(names)
    return cases
