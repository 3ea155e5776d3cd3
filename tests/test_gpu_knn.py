"""k-NN parity on the GPU: HIP path vs the exact oracle, through the C ABI.

Indices must be bit-identical to the oracle (exact (S, index) order);
distances are sqrt of the same canonical float64 sum and must agree to 1e-12
relative (the contract asks for 1e-6).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

PRECISIONS = ["float32", "bfloat16", "float16"]


@pytest.fixture(scope="module")
def mlk():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mlxlearn_amd
    from mlxlearn_amd import _lib
    _lib.load()
    return mlxlearn_amd


def gpu_knn(mlk, X, Q, k, precision="float32", exclude_self=False, screened=None):
    """Query through the C ABI; `screened=True` also asserts no query needed the exact fallback."""
    from mlxlearn_amd import _lib
    idx = mlk.DeviceNeighborIndex(X)
    dist, ind = idx.query(Q, k, exclude_self=exclude_self, precision=precision)
    out = (dist.cpu().numpy(), ind.cpu().numpy())
    st = _lib.stats()
    idx.close()
    if screened:
        assert st["fallback_queries"] == 0, f"screen route expected, stats {st}"
        assert st["rounds"] >= 1
    return out


def assert_parity(got, want, rtol=1e-12):
    gd, gi = got
    wd, wi = want
    bad = np.flatnonzero(np.any(gi != wi, axis=1))
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}: got {gi[bad[:2]]} want {wi[bad[:2]]}"
    np.testing.assert_allclose(gd, wd, rtol=rtol, atol=0)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_plain_gaussian_matches_oracle(mlk, precision):
    """Workload c1 shape (10,000 x 64, 1,000 queries, k=5)."""
    X, Q = oracle.gen_c1(seed=11)
    got = gpu_knn(mlk, X, Q, 5, precision, screened=True)
    assert_parity(got, oracle.knn(X, Q, 5))


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("n,d,m,k", [(512, 16, 31, 7), (5000, 24, 96, 9), (300, 12, 16, 80), (40, 3, 40, 1),
                                     (3000, 96, 200, 10), (4097, 128, 333, 13), (777, 33, 65, 64)])
def test_shapes_match_oracle(mlk, precision, n, d, m, k, rng):
    """Reference kernel-level shapes (tests/parity/test_neighbors.py:33-45) plus ragged ones."""
    X = rng.normal(size=(n, d)).astype(np.float32)
    Q = rng.normal(size=(m, d)).astype(np.float32)
    assert_parity(gpu_knn(mlk, X, Q, k, precision, screened=k <= 64), oracle.knn(X, Q, k))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_mixture_with_duplicates(mlk, precision):
    """Workload c2 generator at reduced size: 1% exact duplicate rows force ties."""
    X, Q = oracle.gen_c2(seed=5, n=60_000, m=512)
    Q[:8] = X[:8]  # queries sitting on training rows
    got = gpu_knn(mlk, X, Q, 10, precision, screened=True)
    assert_parity(got, oracle.knn(X, Q, 10))


def test_ties_break_on_training_index(mlk):
    """tests/parity/test_neighbors.py:48-54: [[1],[1],[1],[2]], query 0, k=3 -> [0, 1, 2]."""
    X = np.array([[1.0], [1.0], [1.0], [2.0]], np.float32)
    for p in PRECISIONS:
        d, i = gpu_knn(mlk, X, np.zeros((1, 1), np.float32), 3, p)
        np.testing.assert_array_equal(i, [[0, 1, 2]])
        np.testing.assert_allclose(d, [[1.0, 1.0, 1.0]])


def test_duplicate_rows_are_ordered_by_index(mlk):
    """tests/parity/test_neighbors_fp32.py:102-107: 50 identical rows -> arange(10)."""
    X = np.repeat(np.array([[3.0]], np.float32), 50, axis=0)
    for p in PRECISIONS:
        _, i = gpu_knn(mlk, X, np.zeros((1, 1), np.float32), 10, p)
        np.testing.assert_array_equal(i[0], np.arange(10))


def test_decoy_column(mlk):
    """tests/parity/test_neighbors_fp32.py:21-39: 120,000-row decoy; the true five are the last rows."""
    n = 120_000
    X = np.empty((n, 1), np.float64)
    X[: n - 5, 0] = 1.0004 + np.arange(n - 5) * 1e-8
    X[n - 5:, 0] = 1.0000 + np.arange(5) * 1e-8
    X = X.astype(np.float32)
    for p in PRECISIONS:
        _, i = gpu_knn(mlk, X, np.zeros((1, 1), np.float32), 5, p)
        np.testing.assert_array_equal(np.sort(i[0]), np.arange(n - 5, n))
        assert_parity(gpu_knn(mlk, X, np.zeros((1, 1), np.float32), 5, p), oracle.knn(X, np.zeros((1, 1)), 5))


def test_large_offset(mlk, rng):
    """tests/parity/test_neighbors_fp32.py:42-63: coordinates at 1e4, separations 5e-2."""
    X = (1e4 + rng.normal(scale=0.05, size=(50_000, 4))).astype(np.float32)
    Q = (1e4 + rng.normal(scale=0.05, size=(200, 4))).astype(np.float32)
    for p in PRECISIONS:
        assert_parity(gpu_knn(mlk, X, Q, 7, p, screened=True), oracle.knn(X, Q, 7))


def test_offset_seed251_row79(mlk):
    """tests/unit/test_neighbors_precision.py:42-49: recorded ordering failure (row 79)."""
    r = np.random.default_rng(251)
    train = (r.normal(size=(50_000, 32)) + 10_000).astype(np.float32)
    queries = (r.normal(size=(256, 32)) + 10_000).astype(np.float32)
    q = queries[79:80]
    for p in PRECISIONS:
        for k in (1, 2, 10):
            assert_parity(gpu_knn(mlk, train, q, k, p), oracle.knn(train, q, k))
        two = train[[1902, 43153]]
        _, i = gpu_knn(mlk, two, q, 1, p)
        assert i[0, 0] == 1  # distinct distances must not become index ties


@pytest.mark.parametrize("precision", PRECISIONS)
def test_self_query_excludes_own_row(mlk, precision, rng):
    """tests/parity/test_neighbors.py:106-150 (own index dropped, duplicates kept)."""
    X = rng.normal(size=(1500, 4)).astype(np.float32)
    X[10:14] = X[9]
    d, i = gpu_knn(mlk, X, X, 3, precision, exclude_self=True)
    assert not np.any(i == np.arange(1500)[:, None])
    assert_parity((d, i), oracle.knn(X, X, 3, exclude_self=True))
    Xd = np.array([[0.0], [0.0], [0.0], [0.0], [5.0]], np.float32)
    _, i2 = gpu_knn(mlk, Xd, Xd, 2, precision, exclude_self=True)
    assert not np.any(i2 == np.arange(5)[:, None])
    np.testing.assert_array_equal(i2, oracle.knn(Xd, Xd, 2, exclude_self=True)[1])


def test_integer_pixels_exact_screen(mlk):
    """Workload c3 generator at reduced size: a 96-pixel band through the blobs (integer data)."""
    X, y, Q = oracle.gen_c3(seed=3, n=6000, m=300)
    X, Q = X[:, 336:432].copy(), Q[:, 336:432].copy()
    for p in PRECISIONS:
        assert_parity(gpu_knn(mlk, X, Q, 3, p, screened=True), oracle.knn(X, Q, 3))


def test_integer_tie_storm(mlk):
    """Top rows of the pixel images: thousands of all-zero rows tie exactly; the exact fallback resolves them."""
    X, y, Q = oracle.gen_c3(seed=3, n=6000, m=300)
    X, Q = X[:, :96].copy(), Q[:, :96].copy()
    for p in PRECISIONS:
        assert_parity(gpu_knn(mlk, X, Q, 3, p), oracle.knn(X, Q, 3))


def test_wide_features_route(mlk, rng):
    """D=784 and D=1024 (tests/parity/test_neighbors_fp32.py:110-121)."""
    for d in (784, 1024):
        X = rng.normal(size=(2000, d)).astype(np.float32)
        Q = rng.normal(size=(16, d)).astype(np.float32)
        for p in PRECISIONS:
            assert_parity(gpu_knn(mlk, X, Q, 5, p), oracle.knn(X, Q, 5))


def test_forced_fallback_matches(mlk, rng):
    """Host-forced fallback of every query gives identical results (fallback route coverage)."""
    from mlxlearn_amd import _lib
    X = rng.normal(size=(20_000, 32)).astype(np.float32)
    Q = rng.normal(size=(50, 32)).astype(np.float32)
    want = oracle.knn(X, Q, 8)
    _lib.set_test_hooks(_lib.HOOK_FORCE_FALLBACK)
    got = gpu_knn(mlk, X, Q, 8)
    assert _lib.stats()["fallback_queries"] == 50
    _lib.set_test_hooks(0)
    assert_parity(got, want)


def test_withdrawn_certificates_and_tiny_capacity(mlk, rng):
    """tests/unit/test_neighbors_screen_fallback.py:107-160 analogue: withdrawn rows and overflowing lists."""
    from mlxlearn_amd import _lib
    X = rng.normal(size=(30_000, 64)).astype(np.float32)
    Q = rng.normal(size=(300, 64)).astype(np.float32)
    want = oracle.knn(X, Q, 10)
    _lib.set_test_hooks(_lib.HOOK_FORCE_UNCERTIFIED)
    _lib.set_forced_rows([0, 7, 299])
    got = gpu_knn(mlk, X, Q, 10, "bfloat16")
    assert _lib.stats()["fallback_queries"] >= 3
    assert_parity(got, want)
    _lib.set_test_hooks(_lib.HOOK_CAPACITY, 2)
    got = gpu_knn(mlk, X, Q, 10, "bfloat16")
    assert _lib.stats()["fallback_queries"] > 0
    _lib.set_test_hooks(0)
    assert_parity(got, want)


@pytest.mark.parametrize("splits,rounds", [(1, [1.0]), (3, [0.5]), (8, [0.01, 0.1]), (64, [0.25, 0.5, 0.75])])
def test_tiling_is_invariant(mlk, rng, splits, rounds):
    """Blocking invariance (tests/parity/test_neighbors.py:57-75): splits and rounds never change results."""
    from mlxlearn_amd import _lib
    X = rng.normal(size=(9000, 40)).astype(np.float32)
    Q = rng.normal(size=(137, 40)).astype(np.float32)
    want = oracle.knn(X, Q, 6)
    _lib.set_tuning(splits=splits, rounds=rounds)
    for p in PRECISIONS:
        assert_parity(gpu_knn(mlk, X, Q, 6, p), want)
    _lib.set_tuning()


def test_unsafe_rows_and_queries(mlk, rng):
    """Rows and queries outside the screen's safe range (huge / tiny magnitudes) stay exact."""
    X = rng.normal(size=(4000, 16)).astype(np.float32)
    X[5, 3] = np.float32(1e30)
    X[6, 0] = np.float32(2**-140)
    X[7, :] = np.float32(3e-39)
    Q = rng.normal(size=(40, 16)).astype(np.float32)
    Q[0, 0] = np.float32(1e25)
    Q[1, 2] = np.float32(1e-30)
    for p in PRECISIONS:
        assert_parity(gpu_knn(mlk, X, Q, 5, p), oracle.knn(X, Q, 5))


def _clustered(rng, n, m, d, centres=12):
    c = rng.normal(size=(centres, d)) * 4
    X = (c[rng.integers(0, centres, n)] + rng.normal(size=(n, d))).astype(np.float32)
    Q = (c[rng.integers(0, centres, m)] + rng.normal(size=(m, d))).astype(np.float32)
    Q[:5] = X[:5]  # exact duplicates of database rows
    return X, Q


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("k", [1, 5, 32])
def test_regrouped_queries(mlk, precision, k, rng):
    """Queries regrouped by locality (forced at this size): the start thresholds sample the
    queries' own neighbourhoods and the results return to the caller's rows."""
    from mlxlearn_amd import _lib
    X, Q = _clustered(rng, 6000, 700, 24)
    _lib.set_test_hooks(_lib.HOOK_FORCE_ORDER)
    try:
        got = gpu_knn(mlk, X, Q, k, precision)
        idx = mlk.DeviceNeighborIndex(X)
        sq, si = idx.query(Q, k, precision=precision, squared=True)
        _, ni = idx.query(Q, k, precision=precision, return_distance=False)
        idx.close()
    finally:
        _lib.set_test_hooks(0)
    want = oracle.knn(X, Q, k)
    assert_parity(got, want)
    np.testing.assert_array_equal(si.cpu().numpy(), want[1])
    np.testing.assert_array_equal(ni.cpu().numpy(), want[1])
    np.testing.assert_allclose(np.sqrt(sq.cpu().numpy()), want[0], rtol=1e-12, atol=0)


@pytest.mark.parametrize("hook", ["fallback", "capacity"])
def test_regrouped_queries_exact_fallback(mlk, hook, rng):
    """Regrouped queries through the exact fallback (all of them, or those whose candidate
    lists overflow a capacity of one slot)."""
    from mlxlearn_amd import _lib
    X, Q = _clustered(rng, 5000, 300, 16)
    extra = _lib.HOOK_FORCE_FALLBACK if hook == "fallback" else _lib.HOOK_CAPACITY
    _lib.set_test_hooks(_lib.HOOK_FORCE_ORDER | extra, 1)
    try:
        got = gpu_knn(mlk, X, Q, 7, "bfloat16")
        st = _lib.stats()
    finally:
        _lib.set_test_hooks(0)
    assert st["fallback_queries"] > 0
    assert_parity(got, oracle.knn(X, Q, 7))


def test_unsafe_rows_with_unbounded_tau(mlk):
    """Fewer safe rows than k: a query's tau stays infinite, so its screen passes every safe row.
    The unsafe rows (outside the screen's safe range) must still arrive once each, through the
    unsafe-row list, never a second time through the screen."""
    r = np.random.default_rng(77)
    X = r.normal(size=(12, 4)).astype(np.float32)
    huge = (r.normal(size=(4, 4)) * 1e30).astype(np.float32)
    X[:8] = np.concatenate([huge, -huge])  # they cancel in the column means: the other rows stay safe
    Q = r.normal(size=(5, 4)).astype(np.float32)
    for p in PRECISIONS:
        got = gpu_knn(mlk, X, Q, 6, p)
        assert all(len(set(row)) == 6 for row in got[1].tolist())
        assert_parity(got, oracle.knn(X, Q, 6))


def test_exact_order_across_float32_range(mlk):
    """tests/unit/test_neighbors_precision.py:145-186: random bit patterns over the whole exponent range."""
    r = np.random.default_rng(9021)
    bits = r.integers(0, 2**32, size=(300, 8), dtype=np.uint32)
    bits = (bits & np.uint32(0x807FFFFF)) | (r.integers(0, 255, size=bits.shape, dtype=np.uint32) << np.uint32(23))
    arr = bits.view(np.float32)
    Q, X = arr[:40].copy(), arr[40:].copy()
    Q[:4] = 0
    X[:8] = 0
    X[0, 0], X[1, 0] = np.float32(2**-140), np.float32(2**-141)
    X[2:4, 0] = np.float32(1e30)
    X[2, 1], X[3, 1] = 2, 1
    X[4:6, 0] = 1
    X[4, 1], X[5, 1] = np.float32(0.0002), np.float32(0.0001)
    X[6, 0], X[7, 0] = -1, 1
    want = oracle.knn(X, Q, 6)
    np.testing.assert_array_equal(want[1], oracle.integer_knn(X, Q, 6))
    for p in PRECISIONS:
        _, i = gpu_knn(mlk, X, Q, 6, p)
        np.testing.assert_array_equal(i, want[1])


@pytest.mark.parametrize("d", [32, 1000])
def test_batched_equals_separate(mlk, rng, d):
    X = rng.normal(size=(8000, d)).astype(np.float32)
    sizes = [1, 17, 300, 5, 64]
    batches = [rng.normal(size=(s, d)).astype(np.float32) for s in sizes]
    index = mlk.DeviceNeighborIndex(X)
    res = index.query_batched(batches, 7, precision="bfloat16")
    for b, (dd, i) in zip(batches, res):
        assert_parity((dd.cpu().numpy(), i.cpu().numpy()), oracle.knn(X, b, 7))
    index.close()


def test_estimator_kneighbors(mlk, rng):
    X = rng.normal(size=(3000, 8))
    q = rng.normal(size=(25, 8))
    nn = mlk.NearestNeighbors(n_neighbors=4).fit(X)
    d, i = nn.kneighbors(q)
    assert d.dtype == np.float64 and i.dtype == np.int64
    assert_parity((d, i), oracle.knn(X.astype(np.float32), q.astype(np.float32), 4))
    d2, i2 = nn.kneighbors()
    assert_parity((d2, i2), oracle.knn(X.astype(np.float32), X.astype(np.float32), 4, exclude_self=True))
    g = nn.kneighbors_graph(q, mode="distance")
    assert g.shape == (25, 3000)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("d", [129, 200, 384, 500, 640, 784, 800])
def test_wide_screen_matches_oracle(mlk, precision, d):
    """Rows wider than 128 features: the wide bf16 / fp16 MFMA screen, and the float32
    screen in 128-feature chunks; every request runs the precision it names."""
    from mlxlearn_amd import _lib
    r = np.random.default_rng(7000 + d)
    X = r.normal(size=(3000, d)).astype(np.float32)
    X[50:55] = X[49]
    Q = np.concatenate([X[:3], r.normal(size=(157, d)).astype(np.float32)])
    assert_parity(gpu_knn(mlk, X, Q, 7, precision, screened=True), oracle.knn(X, Q, 7))
    assert _lib.stats()["screen_precision"] == _lib.PRECISIONS[precision]


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("d", [801, 1024, 1100, 1536])
def test_deep_screen_matches_oracle(mlk, precision, d):
    """Rows wider than 800 features: the deep MFMA screen walks each row tile in 256-feature chunks
    (bf16 / fp16) or 128-feature chunks (float32). Ragged database (not a multiple of the 64-row tile),
    duplicated rows, queries equal to rows, and no query may need the exact fallback."""
    from mlxlearn_amd import _lib
    r = np.random.default_rng(9000 + d)
    X = r.normal(size=(2999, d)).astype(np.float32)
    X[50:55] = X[49]
    Q = np.concatenate([X[:3], X[50:51], r.normal(size=(150, d)).astype(np.float32)])
    assert_parity(gpu_knn(mlk, X, Q, 7, precision, screened=True), oracle.knn(X, Q, 7))
    assert _lib.stats()["screen_precision"] == _lib.PRECISIONS[precision]


@pytest.mark.parametrize("precision", PRECISIONS)
def test_deep_screen_nearest_and_self_query(mlk, precision):
    """k = 1 and the self query (own row excluded) on 1,280-feature rows, through the screen."""
    r = np.random.default_rng(9101)
    X = r.normal(size=(4100, 1280)).astype(np.float32)
    X[7] = X[3]
    Q = X[:300]
    assert_parity(gpu_knn(mlk, X, Q, 1, precision, screened=True), oracle.knn(X, Q, 1))
    assert_parity(gpu_knn(mlk, X, Q, 4, precision, exclude_self=True, screened=True),
                  oracle.knn(X, Q, 4, exclude_self=True))


def test_deep_pixels_exact_screen(mlk):
    """Integer pixels on 1,024-feature rows: exact in fp16, so the screen's band is the accumulation's alone."""
    r = np.random.default_rng(9202)
    X = r.integers(0, 256, size=(3000, 1024)).astype(np.float32)
    Q = np.concatenate([X[:4], r.integers(0, 256, size=(96, 1024)).astype(np.float32)])
    for p in PRECISIONS:
        assert_parity(gpu_knn(mlk, X, Q, 3, p, screened=True), oracle.knn(X, Q, 3))


def test_wide_pixels_exact_screen(mlk):
    """Workload c3 generator (784 integer pixels) through the wide fp16 / bf16 screens and the
    chunked float32 screen (exact for integer pixels), each request in the precision it names."""
    from mlxlearn_amd import _lib
    X, y, Q = oracle.gen_c3(seed=13, n=8000, m=400)
    for p in PRECISIONS:
        assert_parity(gpu_knn(mlk, X, Q, 3, p), oracle.knn(X, Q, 3))
        assert _lib.stats()["screen_precision"] == _lib.PRECISIONS[p]


@pytest.mark.parametrize("d", [4, 64, 100, 200, 256])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_nearest_row_with_ties_at_group_widths(mlk, precision, d):
    """k = 1 with rows duplicated and queries equal to rows: the lowest index of a tied set wins at every row
    width the one-query-per-lane-group rank handles (d % 4 == 0 up to 256), and just inside its limits."""
    r = np.random.default_rng(4100 + d)
    X = r.normal(size=(3000, d)).astype(np.float32)
    X[2000:2010] = X[17]          # ten later copies of row 17
    X[2500] = X[1999]              # one later copy of row 1999
    Q = r.normal(size=(700, d)).astype(np.float32)
    Q[:5] = X[17]
    Q[5:9] = X[1999]
    Q[9:20] = X[2000:2011]
    want = oracle.knn(X, Q, 1)
    got = gpu_knn(mlk, X, Q, 1, precision)
    assert_parity(got, want)
    assert (got[1][:5, 0] == 17).all() and (got[1][5:9, 0] == 1999).all()


def test_merge_with_more_new_bounds_than_it_holds(mlk, rng):
    """A query with about 6,000 rows inside the screen's error band: one merge receives more new upper bounds than
    it holds in LDS (it keeps its threshold), and the rank's list overflows into the exact fallback; every other
    query is unaffected. Results equal the oracle's."""
    from mlxlearn_amd import _lib
    Q = rng.normal(size=(40, 32)).astype(np.float32)
    X = rng.normal(size=(20_000, 32)).astype(np.float32)
    X[:6000] = Q[0] + rng.normal(scale=0.01, size=(6000, 32)).astype(np.float32)
    want = oracle.knn(X, Q, 10)
    _lib.set_tuning(splits=64, capacity=64, rounds=[0.5])
    try:
        for p in PRECISIONS:
            assert_parity(gpu_knn(mlk, X, Q, 10, p), want)
    finally:
        _lib.set_tuning()


@pytest.mark.parametrize("near", [1500, 3900, 5000])
def test_rank_with_more_survivors_than_one_wave_holds(mlk, rng, near):
    """`near` rows packed around the first queries, far inside the bf16 screen's error band, with lists large enough
    to hold them all: more rows survive the bound filter than one wave's 512-row LDS list. Up to 4,096 survivors the
    query is ranked again with the larger list (no exact fallback); beyond that it takes the exact fallback.
    Results equal the oracle's either way."""
    from mlxlearn_amd import _lib
    X = (rng.normal(size=(near + 1500, 48)) + 10.0).astype(np.float32)
    base = rng.normal(size=48).astype(np.float32)
    X[:near] = base + rng.normal(scale=1e-4, size=(near, 48)).astype(np.float32)
    Q = np.concatenate([base + rng.normal(scale=1e-4, size=(3, 48)), rng.normal(size=(61, 48)) + 10.0]).astype(np.float32)
    want = oracle.knn(X, Q, 5)
    _lib.set_tuning(splits=8, capacity=1024)
    try:
        for p in PRECISIONS:
            assert_parity(gpu_knn(mlk, X, Q, 5, p), want)
        assert_parity(gpu_knn(mlk, X, Q, 5, "bfloat16"), want)
        st = _lib.stats()
        assert st["fallback_queries"] == (0 if near <= 3900 else 3), st
    finally:
        _lib.set_tuning()


def _canonical(Q, X, idx):
    """sum_j (double(q_j) - double(x_j))^2 left to right, each product and sum
    rounded on its own (NumPy's float64 array operations round separately)."""
    s = np.zeros(idx.shape)
    for j in range(Q.shape[1]):
        t = Q[:, None, j].astype(np.float64) - X[idx, j].astype(np.float64)
        s = s + t * t
    return s


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("d,k", [(37, 9), (96, 1), (128, 10), (784, 3)])
def test_squared_distances_are_canonical_sums(mlk, precision, d, k):
    """Reported squared distances are the canonical float64 sums bit for bit
    (no fused multiply-add): the rank kernels, the k = 1 group rank and the
    fallback all write the same value."""
    r = np.random.default_rng(d)
    X = r.normal(size=(4000, d)).astype(np.float32)
    Q = r.normal(size=(300, d)).astype(np.float32)
    index = mlk.DeviceNeighborIndex(X)
    dist, ind = index.query(Q, k, precision=precision, squared=True)
    ind = ind.cpu().numpy()
    np.testing.assert_array_equal(ind, oracle.knn(X, Q, k)[1])
    np.testing.assert_array_equal(dist.cpu().numpy(), _canonical(Q, X, ind))
    index.close()
