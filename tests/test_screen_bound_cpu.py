"""The screen's separable rounding bound, checked against exact rational arithmetic.

The device keeps a row for a query when acc >= thr, with
  acc = fl32(-a_x + sum q~_i x~_i)                      (MFMA accumulation, any order)
  a_x = rd32(screen_row_bound(nx, |x~|^2, |x - x~|^2))
  thr = rd32((qa - tau (1 + 2^-46)) / 2 - f0),  qa = screen_qa(nq, |q~|^2, |q - q~|^2)
(mlk_screen.h). Soundness: every row with sigma^2 S <= tau must be kept, for any
accumulation error within gamma (|a_x| + |q~||x~|) + f0. This test restates the
float64 formulas, takes the smallest accumulator the error bound allows, and
checks with exact fractions that it still reaches the threshold of
tau = sigma^2 S (the row's own exact distance) -- on random data and on
adversarial data where every Cauchy-Schwarz step is tight (constant vectors,
rounding residuals aligned with the other vector). Planted mistakes (a dropped
residual cross term, a zero splitting parameter on an inexact side) must fail.
"""
import math
from fractions import Fraction as F

import numpy as np
import pytest

torch = pytest.importorskip("torch")

BF16_U = 2.0 ** -8 + 2.0 ** -23
FP16_U = 2.0 ** -11 + 2.0 ** -23


def to_screen(v64, prec):
    f32 = torch.from_numpy(np.asarray(v64, np.float64).astype(np.float32))
    dt = torch.bfloat16 if prec == "bf16" else torch.float16
    return f32.to(dt).to(torch.float64).numpy()


def consts(prec, dp, d, rows_exact, queries_exact, drop_cross=False):
    kp = dp + 1
    gamma = kp * 2.0 ** -23 / (1.0 - kp * 2.0 ** -23)
    t = 0.5 * (BF16_U if prec == "bf16" else FP16_U)
    return dict(gamma=gamma, f0=(dp + 2) * 2.0 ** -126, eps=(d + 8) * 2.0 ** -50,
                t1=0.0 if rows_exact else t, t2=0.0 if queries_exact else t, drop_cross=drop_cross)


def alpha(nt, d2, c):
    T2, D2 = nt * (1 + c["eps"]), d2 * (1 + c["eps"])
    a = (c["gamma"] + c["t1"]) * T2 * 0.5
    if D2 > 0:
        extra = 0.0 if c["drop_cross"] else 1.0
        a += (1.0 / c["t2"] + extra) * D2 * 0.5 if c["t2"] > 0 else math.inf
    return a * (1 + 2.0 ** -50)


def beta(nt, d2, c):
    T2, D2 = nt * (1 + c["eps"]), d2 * (1 + c["eps"])
    b = (c["gamma"] + c["t2"]) * T2 * 0.5
    if D2 > 0:
        extra = 0.0 if c["drop_cross"] else 1.0
        b += (1.0 / c["t1"] + extra) * D2 * 0.5 if c["t1"] > 0 else math.inf
    return b * (1 + 2.0 ** -50)


def qa_of(nq, nt, d2, c):
    al = alpha(nt, d2, c)
    return nq * (1 - c["eps"]) - 2 * al - 2.0 ** -50 * (abs(nq) + 2 * al)


def row_bound(nx, nt, d2, c):
    be = beta(nt, d2, c)
    lo = nx * (1 - c["eps"]) - 2 * be
    lo -= 2.0 ** -50 * (abs(nx) + 2 * be)
    if lo >= 0:
        return lo * 0.5 / (1 + c["gamma"]) * (1 - 2.0 ** -50)
    return lo * 0.5 / (1 - c["gamma"]) * (1 + 2.0 ** -50)


def rd32(v):
    f = np.float32(v)
    return f if float(f) <= v else np.nextafter(f, np.float32(-np.inf))


def mirror(v32, centre, scale, prec):
    v = (v32.astype(np.float64) - centre.astype(np.float64)) * scale
    vt = to_screen(v, prec)
    return v, vt, float(np.sum(v * v)), float(np.sum(vt * vt)), float(np.sum((v - vt) ** 2))


def kept(q32, x32, centre, scale, prec, c):
    """Does the smallest admissible accumulator of the pair reach the threshold of tau = sigma^2 S?"""
    _, qt, nq, ntq, d2q = mirror(q32, centre, scale, prec)
    _, xt, nx, ntx, d2x = mirror(x32, centre, scale, prec)
    a_x = float(rd32(row_bound(nx, ntx, d2x, c)))
    S = sum((F(float(a)) - F(float(b))) ** 2 for a, b in zip(q32, x32)) * F(scale) ** 2
    p_hat = sum(F(float(a)) * F(float(b)) for a, b in zip(qt, xt))
    norms = F(math.sqrt(ntq)) * F(math.sqrt(ntx)) * F(1 + 2.0 ** -40)
    acc_min = p_hat - F(a_x) - F(c["gamma"]) * (abs(F(a_x)) + norms) - F(c["f0"])
    tau = float(S) * (1 + 2.0 ** -52)  # an upper bound of S in float64
    qa = qa_of(nq, ntq, d2q, c)
    thr = rd32((qa - tau * (1 + 2.0 ** -46)) * 0.5 - c["f0"])
    return acc_min >= F(float(thr))


def cases(rng, d):
    out = []
    base = rng.normal(size=d).astype(np.float32)
    out.append((rng.normal(size=d).astype(np.float32), rng.normal(size=d).astype(np.float32)))
    out.append((base, (base + rng.normal(size=d) * 1e-3).astype(np.float32)))   # near duplicates
    # constant vectors: every Cauchy-Schwarz step tight, residuals aligned
    for v in (1.0 + 2.0 ** -9, 1.0 + 3 * 2.0 ** -10, 3.0 + 2.0 ** -7, 1.0 - 2.0 ** -10):
        out.append((np.full(d, v, np.float32), np.full(d, v * 1.003, np.float32)))
        out.append((np.full(d, v, np.float32), np.full(d, -v, np.float32)))
    # large offsets with cancellation and mixed magnitudes
    out.append(((base * 1e4).astype(np.float32), (base * 1e4 + 1).astype(np.float32)))
    mixed = (rng.normal(size=d) * 2.0 ** rng.integers(-20, 20, d)).astype(np.float32)
    out.append((mixed, (mixed * 1.01).astype(np.float32)))
    return out


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("d", [1, 7, 32, 128])
def test_screen_keeps_every_row_within_tau(prec, d):
    rng = np.random.default_rng(d)
    dp = (d + 31) // 32 * 32
    centre = np.zeros(d, np.float32)
    c = consts(prec, dp, d, rows_exact=False, queries_exact=False)
    for q, x in cases(rng, d):
        scale = 1.0
        if prec == "fp16":   # the mirror's power-of-two scale into fp16's range
            top = float(max(np.abs(q).max(), np.abs(x).max(), 1e-30))
            scale = 2.0 ** (14 - math.frexp(top)[1])
        assert kept(q, x, centre, scale, prec, c), (prec, d, q[:3], x[:3])


def test_exact_rows_and_queries():
    """Integer pixels: every screen value exact (t1 = t2 = 0); rows exact, queries not (t1 = 0)."""
    rng = np.random.default_rng(3)
    d, dp = 784, 800
    X = rng.integers(0, 256, size=(4, d)).astype(np.float32)
    centre = np.full(d, 128, np.float32)
    c = consts("bf16", dp, d, rows_exact=True, queries_exact=True)
    assert all(kept(X[0], X[i], centre, 1.0, "bf16", c) for i in range(1, 4))
    q = (X[0] + 0.3).astype(np.float32)                 # inexact query against exact rows
    c = consts("bf16", dp, d, rows_exact=True, queries_exact=False)
    assert all(kept(q, X[i], centre, 1.0, "bf16", c) for i in range(4))


def test_planted_mistakes_fail():
    """Dropping the residual cross term |dq||dx|, or declaring an inexact side exact, loses a row
    that lies within tau. The exposing pair is a duplicate of a constant row whose bf16 residual is
    half the unit roundoff, where every Cauchy-Schwarz and splitting step of the bound is tight."""
    d = 128
    centre = np.zeros(d, np.float32)
    q = np.full(d, 1.0 + 2.0 ** -9, np.float32)        # rounds to 1.0 in bf16: residual ratio 2^-9
    good = consts("bf16", 128, d, rows_exact=False, queries_exact=False)
    bad_cross = consts("bf16", 128, d, rows_exact=False, queries_exact=False, drop_cross=True)
    assert kept(q, q, centre, 1.0, "bf16", good)
    assert not kept(q, q, centre, 1.0, "bf16", bad_cross)
    # a query declared exact while its residual is not: its alpha term is infinite, so no
    # threshold can be formed; the device never pairs t2 = 0 with an inexact query
    bad_exact = consts("bf16", 128, d, rows_exact=False, queries_exact=True)
    _, _, nq, nt, d2 = mirror(q, centre, 1.0, "bf16")
    assert math.isinf(alpha(nt, d2, bad_exact)) and math.isfinite(alpha(nt, d2, good))


def test_bound_is_tighter_than_the_worst_case_constant():
    """On Gaussian rows the separable bound is about 0.4 of the unit-roundoff bound eta (|q|^2 + |x|^2)."""
    rng = np.random.default_rng(1)
    d = 128
    c = consts("bf16", 128, d, rows_exact=False, queries_exact=False)
    ratios = []
    for _ in range(50):
        q = rng.normal(size=d).astype(np.float32) * 3
        _, _, n, nt, d2 = mirror(q, np.zeros(d, np.float32), 1.0, "bf16")
        ratios.append((alpha(nt, d2, c) + beta(nt, d2, c)) / (BF16_U * 2 * n))
    assert 0.3 < float(np.median(ratios)) < 0.6
