"""The oracle is pinned before it is trusted.

* against golden vectors written by the reference's own native engines
  (``tests/golden/reference_native.npz``, made by ``tests/golden/make_golden.py``);
* against the reference engines themselves when ``oracle/_ref`` is built;
* against Python big-integer arithmetic (neighbors_cpu.py:39-47, 174-181) on
  adversarial bit patterns;
* against closed forms (lattices with known distances, exact ties);
* with planted mistakes that each check must catch.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_native.npz")
REF_DIR = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("name", ["gauss24", "ties8", "offset4_self", "scales3"])
def test_knn_matches_reference_native_golden(golden, name):
    X, Q = golden[f"knn_{name}_X"], golden[f"knn_{name}_Q"]
    k, ex = int(golden[f"knn_{name}_k"]), bool(golden[f"knn_{name}_exclude"])
    d, i = oracle.knn(X, Q, k, exclude_self=ex)
    np.testing.assert_array_equal(i, golden[f"knn_{name}_idx"])
    np.testing.assert_allclose(d, golden[f"knn_{name}_dist"], rtol=1e-15, atol=0)


@pytest.mark.parametrize("key", ["uniform6", "int4_self"])
@pytest.mark.parametrize("sort", [False, True])
def test_radius_matches_reference_native_golden(golden, key, sort):
    X, Q = golden[f"rad_{key}_X"], golden[f"rad_{key}_Q"]
    R2, ex = float(golden[f"rad_{key}_R2"]), bool(golden[f"rad_{key}_exclude"])
    tag = f"rad_{key}_{'sorted' if sort else 'plain'}"
    ip, ind, dist = oracle.radius(X, Q, R2, exclude_self=ex, sort_results=sort)
    np.testing.assert_array_equal(ip, golden[f"{tag}_indptr"])
    np.testing.assert_array_equal(ind, golden[f"{tag}_indices"])
    np.testing.assert_array_equal(dist, golden[f"{tag}_dist"])  # canonical sum: bit-identical


def _ref_modules():
    if not os.path.isdir(REF_DIR):
        pytest.skip("reference native engines not built (oracle/_ref)")
    sys.path.insert(0, REF_DIR)
    try:
        import _neighbors_cpu_native as rn
        import _radius_cpu_native as rr
    except ImportError:
        pytest.skip("reference native engines not importable")
    return rn, rr


def test_knn_matches_reference_engine_fresh(rng):
    rn, _ = _ref_modules()
    for d in (1, 5, 17, 32):
        X = rng.normal(size=(2500, d)).astype(np.float32)
        X[::97] = X[3]
        Q = rng.normal(size=(33, d)).astype(np.float32)
        dist = np.empty((33, 7))
        idx = np.empty((33, 7), np.int64)
        rn.query(X, Q, 7, False, 0, dist, idx)
        np.testing.assert_array_equal(oracle.knn(X, Q, 7)[1], idx)


def test_big_integer_keys_across_exponent_range():
    r = np.random.default_rng(9021)
    bits = r.integers(0, 2**32, size=(200, 6), dtype=np.uint32)
    bits = (bits & np.uint32(0x807FFFFF)) | (r.integers(0, 255, size=bits.shape, dtype=np.uint32) << np.uint32(23))
    arr = bits.view(np.float32)
    for a, b in zip(arr[:100], arr[100:]):
        assert oracle.exact_key_c(a, b) == oracle.integer_key(a, b)


def test_exact_compare_on_adjacent_floats():
    q = np.zeros(3, np.float32)
    a = np.array([1.0, 2.0, 3.0], np.float32)
    b = a.copy()
    b[2] = np.nextafter(b[2], np.float32(np.inf))
    assert oracle.exact_compare(q, a, b) == -1
    assert oracle.exact_compare(q, b, a) == 1
    assert oracle.exact_compare(q, a, a.copy()) == 0
    # a genuine tie with different coordinates: (-1, 0) vs (1, 0) from the origin
    assert oracle.exact_compare(np.zeros(2, np.float32), np.array([-1, 0], np.float32),
                                np.array([1, 0], np.float32)) == 0


def test_integer_knn_agrees_on_adversarial_rows():
    r = np.random.default_rng(5)
    X = (r.normal(size=(80, 4)) * 2.0 ** r.integers(-40, 40, (80, 1))).astype(np.float32)
    X[10:14] = X[2]
    Q = X[:12] * np.float32(0.999)
    np.testing.assert_array_equal(oracle.knn(X, Q, 9)[1], oracle.integer_knn(X, Q, 9))


def test_lattice_closed_form():
    """Points on an integer lattice: S is a known integer; ties break by index."""
    g = np.stack(np.meshgrid(np.arange(5), np.arange(5), indexing="ij"), -1).reshape(-1, 2).astype(np.float32)
    d, i = oracle.knn(g, np.array([[2.0, 2.0]], np.float32), 5)
    np.testing.assert_array_equal(i[0], [12, 7, 11, 13, 17])  # centre, then the four at distance 1 by index
    np.testing.assert_array_equal(d[0], [0, 1, 1, 1, 1])


def test_radius_band_and_exact_membership():
    # a point exactly on the radius is a member; the next float outside is not
    q = np.zeros(1, np.float32)
    x = np.array([3.0], np.float32)
    assert oracle.exact_le(q, x, 9.0)
    assert not oracle.exact_le(q, x, math.nextafter(9.0, 0.0))
    lo, hi = oracle.radius_band(4, 2.0)
    assert lo < 2.0 < hi


TINY = 2.0 ** -1074
BAND_R2 = [TINY, 3 * TINY, 2.0 ** -1022, 1e-300, 1e-30, 0.5, 1.0, 2.0, 3.3, 1e30, 1e300, 2.0 ** 1023,
           1.7976931348623157e308]


@pytest.mark.parametrize("d", [1, 2, 4, 32, 128, 784, 4096])
def test_radius_band_against_exact_rationals(d):
    """The band of the membership test, pinned by exact rational arithmetic.

    The canonical float64 sum s of d squared differences satisfies
    |s - S| <= gamma * S with gamma = (d+2)u / (1 - (d+2)u), u = 2^-53
    (radius_contract.py:129-153), so s <= lo certifies S <= R2 only if
    lo <= R2 (1 - 2 gamma), and s > hi certifies S > R2 only if
    hi >= R2 (1 + 2 gamma). Both are checked as exact fractions (soundness),
    and each end must lie within four units in the last place of that exact
    value (tightness), so a dropped factor, d in place of d+2 or rounding
    towards the wrong side fails.
    """
    from fractions import Fraction as F
    u = F(1, 2 ** 53)
    g = (d + 2) * u / (1 - (d + 2) * u)
    for R2 in BAND_R2:
        lo, hi = oracle.radius_band(d, R2)
        want_lo = F(R2) * (1 - 2 * g)
        want_hi = F(R2) * (1 + 2 * g)
        assert lo >= 0.0 and F(lo) <= want_lo, (d, R2, lo)
        # tightness: four ulps of the exact value, or four subnormal steps near zero
        slack = max(want_lo * 4 * u * 2, F(4) * F(TINY))
        assert want_lo - F(lo) <= slack, (d, R2, lo)
        if math.isinf(hi):
            assert want_hi > F(1.7976931348623157e308), (d, R2)
        else:
            assert F(hi) >= want_hi, (d, R2, hi)
            assert F(hi) - want_hi <= max(want_hi * 4 * u * 2, F(4) * F(TINY)), (d, R2, hi)
    assert oracle.radius_band(d, 0.0) == (0.0, 0.0)
    assert oracle.radius_band(d, math.inf) == (math.inf, math.inf)


def test_radius_band_planted_mistakes_fail():
    """The rational check above rejects plausible wrong bands (planted on the CPU)."""
    from fractions import Fraction as F
    u = F(1, 2 ** 53)

    def sound(d, R2, lo, hi):
        g = (d + 2) * u / (1 - (d + 2) * u)
        return F(lo) <= F(R2) * (1 - 2 * g) and F(hi) >= F(R2) * (1 + 2 * g)

    d, R2 = 3, 2.0
    g_wrong = d * 2.0 ** -53 / (1 - d * 2.0 ** -53)        # d instead of d + 2
    assert not sound(d, R2, R2 * (1 - 2 * g_wrong), R2 * (1 + 2 * g_wrong))
    g = (d + 2) * 2.0 ** -53 / (1 - (d + 2) * 2.0 ** -53)
    assert not sound(d, R2, R2 * (1 - g), R2 * (1 + g))     # dropped factor 2
    lo, hi = oracle.radius_band(d, R2)
    assert sound(d, R2, lo, hi)


def test_kmeans_assign_lowest_index_on_ties():
    """tests/unit/test_e_means.py:115-126: identical centres -> label 0."""
    r = np.random.default_rng(747)
    x = r.normal(size=(41, 17)).astype(np.float32)
    centres = np.tile(x[0], (67, 1))
    labels, resid = oracle.kmeans_assign(x, centres)
    np.testing.assert_array_equal(labels, 0)
    np.testing.assert_allclose(resid, ((x.astype(np.float64) - centres[0]) ** 2).sum(1), rtol=1e-15)
    assert resid[0] == 0


def test_kmeans_update_sums():
    r = np.random.default_rng(1)
    X = r.normal(size=(500, 7)).astype(np.float32)
    labels = r.integers(0, 9, 500)
    sums, counts = oracle.kmeans_update(X, labels, 9)
    for c in range(9):
        np.testing.assert_allclose(sums[c], X[labels == c].astype(np.float64).sum(0), rtol=1e-12)
    assert counts.sum() == 500


def test_planted_mistakes_are_caught(rng):
    """Each pin above must fail for a plausible oracle bug (dropped term, wrong tie order, sign)."""
    X = rng.normal(size=(300, 5)).astype(np.float32)
    X[40:45] = X[4]
    Q = X[:6].copy()
    d, i = oracle.knn(X, Q, 6)
    want = oracle.integer_knn(X, Q, 6)
    # dropped last feature
    _, i_drop = oracle.knn(X[:, :4], Q[:, :4], 6)
    assert not np.array_equal(i_drop, want)
    # reversed tie order among the duplicate block
    rev = i.copy()
    row = 4
    tied = np.flatnonzero(np.isin(rev[row], np.arange(40, 45)))
    rev[row, tied] = rev[row, tied][::-1]
    assert not np.array_equal(rev, want)
    # sign error in one coordinate
    _, i_sign = oracle.knn(X * np.array([1, 1, 1, 1, -1], np.float32), Q, 6)
    assert not np.array_equal(i_sign, want)
    np.testing.assert_array_equal(i, want)


def test_generators_are_seeded_and_shaped():
    X, Q = oracle.gen_c1(seed=1, n=100, m=10)
    X2, _ = oracle.gen_c1(seed=1, n=100, m=10)
    assert X.shape == (100, 64) and Q.shape == (10, 64) and np.array_equal(X, X2)
    X, Q = oracle.gen_c2(seed=0, n=20_000, m=50)
    assert X.dtype == np.float32 and X.shape == (20_000, 128)
    assert len(np.unique(X.view(np.uint8).reshape(20_000, -1), axis=0)) <= 20_000 - 150
    X, y, Q = oracle.gen_c3(seed=0, n=500, m=20)
    assert X.shape == (500, 784) and set(np.unique(X)) <= set(range(256))
    zeros = (X == 0).mean()
    assert 0.7 < zeros < 0.9
    assert set(np.unique(y)) <= set(range(10))


def _small_radius_cases():
    import radius_cases
    return [c for c in radius_cases.all_cases() if c.X.shape[0] * c.Q.shape[0] <= 60_000]


@pytest.mark.parametrize("case", _small_radius_cases(), ids=lambda c: c.name)
def test_radius_oracle_matches_integer_brute_force(case):
    """Membership S * 2^298 <= floor(R2 * 2^298) with Python integers
    (radius_contract.py:169-217), ascending index rows, (distance, index) when sorted."""
    from fractions import Fraction
    X = np.ascontiguousarray(case.X, np.float32)
    Q = np.ascontiguousarray(case.Q, np.float32)
    R2 = float(case.radius) * float(case.radius)
    limit = math.floor(Fraction(R2) * 2 ** 298) if math.isfinite(R2) else None
    xs = [oracle.integer_row(x) for x in X]
    indptr, indices, dist = oracle.radius(X, Q, R2)
    for r, q in enumerate(Q):
        qi = oracle.integer_row(q)
        keys = [sum((a - b) ** 2 for a, b in zip(qi, xi)) for xi in xs]
        want = [i for i, s in enumerate(keys) if limit is None or s <= limit]
        got = indices[indptr[r]:indptr[r + 1]]
        assert list(got) == want, f"{case.name} row {r}"
        for i, dv in zip(got, dist[indptr[r]:indptr[r + 1]]):
            exact = keys[i] / 2.0 ** 298 if keys[i] < 2 ** 1300 else math.inf
            assert math.isclose(dv, math.sqrt(min(exact, R2)), rel_tol=1e-14, abs_tol=0) or dv == math.sqrt(R2)
    sp, si, sd = oracle.radius(X, Q, R2, sort_results=True)
    np.testing.assert_array_equal(sp, indptr)
    for r in range(len(Q)):
        seg = slice(sp[r], sp[r + 1])
        pairs = list(zip(sd[seg], si[seg]))
        assert pairs == sorted(pairs)
        assert sorted(si[seg]) == list(indices[seg])


def test_kmeans_assign_is_exact_nearest_centre():
    """Labels = the exact 1-NN among the centres (integer brute force), residual = canonical s."""
    r = np.random.default_rng(31)
    X = (r.normal(size=(60, 3)) * 2.0 ** r.integers(-20, 20, (60, 1))).astype(np.float32)
    C = np.concatenate([X[:5], X[:5] * np.float32(1.0000001), r.normal(size=(10, 3)).astype(np.float32)])
    labels, resid = oracle.kmeans_assign(X, C)
    np.testing.assert_array_equal(labels, oracle.integer_knn(C, X, 1)[:, 0])
    for i in range(60):
        assert resid[i] == oracle.canonical_sq(X[i], C[labels[i]])


def test_cpu_baseline_port_routes_are_exact(rng, monkeypatch):
    """Every route of the CPU-path restatement (the timed CPU baseline) returns the
    oracle's neighbours: the Gram proposal with exact repair over the k+2
    candidates, the direct scan of uncertain rows, the direct scan with its own
    exact repair, and the integer Gram of small integer data."""
    calls = {"exact": 0, "direct": 0, "integer": 0}
    for name, key in (("_exact_order", "exact"), ("_direct", "direct"), ("_integer_gram", "integer")):
        orig = getattr(oracle.GramPort, name)

        def wrapped(self, *a, _orig=orig, _key=key, **kw):
            calls[_key] += 1
            return _orig(self, *a, **kw)
        monkeypatch.setattr(oracle.GramPort, name, wrapped)
    X = rng.normal(size=(2000, 24)).astype(np.float32)
    X[100:112] = X[5]                                    # 12 exact copies: uncertain barrier for k = 7
    base = rng.normal(size=24).astype(np.float32)
    for j in range(8):                                   # near-ties: one float32 step apart
        X[300 + j] = base
        X[300 + j, j] = np.nextafter(base[j], np.float32(np.inf))
    Q = np.concatenate([X[[5, 300, 301]], base[None], rng.normal(size=(40, 24)).astype(np.float32)])
    port = oracle.GramPort(X)
    d, i = port.query(Q, 7)
    wd, wi = oracle.knn(X, Q, 7)
    np.testing.assert_array_equal(i, wi)
    np.testing.assert_allclose(d, wd, rtol=1e-12)
    assert calls["direct"] >= 1 and calls["exact"] >= 1
    # short query block: the direct route with its own repair
    d, i = port.query(Q[:4], 7)
    np.testing.assert_array_equal(i, wi[:4])
    # self query without the own row
    d, i = port.query(X[:50], 5, exclude_self=True)
    np.testing.assert_array_equal(i, oracle.knn(X, X[:50], 5, exclude_self=True)[1])
    # small integer data: the float32 integer Gram
    Xi = rng.integers(0, 8, size=(1000, 40)).astype(np.float32)
    Qi = rng.integers(0, 8, size=(64, 40)).astype(np.float32)
    d, i = oracle.GramPort(Xi).query(Qi, 9)
    wd, wi = oracle.knn(Xi, Qi, 9)
    np.testing.assert_array_equal(i, wi)
    np.testing.assert_array_equal(d, wd)
    assert calls["integer"] >= 1


def test_cpu_baseline_port_is_exact(rng):
    """The timed CPU baseline (CPUIndex._gram restatement) returns the oracle's neighbours."""
    X = rng.normal(size=(3000, 16)).astype(np.float32)
    X[100:110] = X[5]
    Q = np.concatenate([X[:20], rng.normal(size=(50, 16)).astype(np.float32)])
    d, i = oracle.GramPort(X).query(Q, 7, block=32)
    wd, wi = oracle.knn(X, Q, 7)
    np.testing.assert_array_equal(i, wi)
    np.testing.assert_allclose(d, wd, rtol=1e-12)


def test_kmeans_cpu_port_labels_are_exact(rng):
    """The timed k-means CPU baseline (float32 argmin + float64 check) returns the exact labels,
    lowest index on ties (duplicate centres, rows on centres, near-equidistant rows)."""
    X = rng.normal(size=(4000, 24)).astype(np.float32)
    C = X[:300].copy()
    C[7] = C[3]                                   # duplicate centres: label 3, never 7
    mid = ((C[10].astype(np.float64) + C[11]) / 2).astype(np.float32)
    X[-5:] = mid                                  # near-equidistant rows
    labels, resid = oracle.KMeansPort(C).assign(X, block=512)
    want, wres = oracle.kmeans_assign(X, C)
    np.testing.assert_array_equal(labels, want)
    np.testing.assert_allclose(resid, wres, rtol=1e-13, atol=0)
    assert not np.any(labels == 7)


def test_reference_radius_engine_matches_oracle(rng):
    """The reference's compiled radius engine, called as the radius CPU baseline calls it,
    gives the oracle's CSR (sorted and unsorted)."""
    if oracle.reference_radius_engine() is None:
        pytest.skip("reference native engines not built (oracle/_ref)")
    X = rng.random((20000, 8), dtype=np.float32)
    X[100:140] = X[5]
    Q = np.concatenate([X[:3], rng.random((60, 8), dtype=np.float32)])
    for sort in (False, True):
        got = oracle.reference_radius(X, Q, 0.09, sort_results=sort, threads=4)
        want = oracle.radius(X, Q, 0.09, sort_results=sort)
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w)
