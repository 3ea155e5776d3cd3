"""Screen plans and the autotuner on the GPU.

Tile planning is performance-only in the reference (resolve_blocks,
_neighbors_plan.py:74-126; blocking invariance, tests/parity/test_neighbors.py:57-75):
every plan -- each built screen shape, any split count, round schedule or
tuned plan -- must return the oracle's indices bit for bit.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

PRECISIONS = ["float32", "bfloat16", "float16"]
SHAPES = {"float32": [(1, 2), (1, 3)], "bfloat16": [(1, 2), (2, 2), (2, 3), (2, 4)],
          "float16": [(1, 2), (2, 2), (2, 3), (2, 4)]}


@pytest.fixture(scope="module")
def mlk():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mlxlearn_amd
    from mlxlearn_amd import _lib
    _lib.load()
    yield mlxlearn_amd
    _lib.set_tuning()
    _lib.clear_plans()


def knn(index, Q, k, precision):
    d, i = index.query(Q, k, precision=precision)
    return d.cpu().numpy(), i.cpu().numpy()


def assert_parity(got, want):
    np.testing.assert_array_equal(got[1], want[1])
    np.testing.assert_allclose(got[0], want[0], rtol=1e-12, atol=0)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_every_screen_shape_matches_oracle(mlk, precision):
    """Each built (query tiles, stages) shape, narrow widths 32..128, ragged query counts."""
    from mlxlearn_amd import _lib
    r = np.random.default_rng(4400)
    for d in (20, 64, 128):
        X, _ = oracle.gen_c2(seed=d, n=20_000, m=1)
        X = np.ascontiguousarray(X[:, :d])
        Q = np.concatenate([X[:3], r.normal(size=(297, d)).astype(np.float32) * 2])
        want = oracle.knn(X, Q, 7)
        index = mlk.DeviceNeighborIndex(X)
        try:
            for qt, ns in SHAPES[precision]:
                _lib.set_tuning(query_tiles=qt, stages=ns)
                assert_parity(knn(index, Q, 7, precision), want)
                st = _lib.stats()
                assert st["fallback_queries"] == 0, (d, qt, ns)
        finally:
            _lib.set_tuning()
            index.close()


@pytest.mark.parametrize("precision", ["bfloat16", "float16"])
def test_wide_shapes_match_oracle(mlk, precision):
    """Rows wider than 128 features in both wide shapes: whole rows per LDS stage (three stages) and
    rows split in two chunks whose query fragments all stay in registers (two stages), at every wide
    width, with ragged row and query counts, duplicated rows and queries equal to rows."""
    from mlxlearn_amd import _lib
    r = np.random.default_rng(4500)
    for d in (129, 300, 384, 500, 640, 784, 800):
        X = r.normal(size=(3001, d)).astype(np.float32)
        X[50:55] = X[49]
        Q = np.concatenate([X[:3], X[50:51], r.normal(size=(293, d)).astype(np.float32)])
        want = oracle.knn(X, Q, 7)
        want1 = oracle.knn(X, Q, 1)
        index = mlk.DeviceNeighborIndex(X)
        try:
            for ns in (2, 3):
                _lib.set_tuning(stages=ns)
                assert index.plan(Q.shape[0], 7, precision=precision)["stages"] == ns
                assert_parity(knn(index, Q, 7, precision), want)
                assert _lib.stats()["fallback_queries"] == 0, (d, ns)
                assert_parity(knn(index, Q, 1, precision), want1)
        finally:
            _lib.set_tuning()
            index.close()


@pytest.mark.parametrize("precision", ["float32", "bfloat16"])
def test_autotuned_knn_plan_is_used_and_exact(mlk, precision):
    from mlxlearn_amd import _lib
    _lib.clear_plans()
    X, Q = oracle.gen_c2(seed=31, n=50_000, m=1500)
    want = oracle.knn(X, Q, 10)
    index = mlk.DeviceNeighborIndex(X)
    try:
        before = index.plan(1500, 10, precision=precision)
        assert not before["tuned"] and before["rounds"][-1] == 1.0
        assert_parity(knn(index, Q, 10, precision), want)
        plan = index.autotune(Q, 10, precision=precision, max_trials=12)
        assert plan["tuned"] and 2 <= plan["trials"] <= 12
        assert 0 < plan["tuned_ms"] <= plan["default_ms"]
        after = index.plan(1500, 10, precision=precision)
        assert after["tuned"] and after["splits"] == plan["splits"] and after["capacity"] == plan["capacity"]
        assert (after["query_tiles"], after["stages"]) == (plan["query_tiles"], plan["stages"])
        # the same query-count bucket uses the plan; another k does not
        assert index.plan(1400, 10, precision=precision)["tuned"]
        assert not index.plan(1500, 11, precision=precision)["tuned"]
        got = knn(index, Q, 10, precision)
        assert _lib.stats()["splits"] == plan["splits"]
        assert_parity(got, want)
        text = _lib.export_plans()
        _lib.clear_plans()
        assert not index.plan(1500, 10, precision=precision)["tuned"]
        _lib.import_plans(text)  # as a later process would
        again = index.plan(1500, 10, precision=precision)
        assert again["tuned"] and again["splits"] == plan["splits"] and again["tuned_ms"] == plan["tuned_ms"]
        assert_parity(knn(index, Q, 10, precision), want)
        _lib.clear_plans()
    finally:
        index.close()


def test_autotune_rejects_plans_that_fall_back_widely(mlk):
    """Under a one-slot list capacity (test hook) every plan's lists overflow and most queries take the
    exact fallback: the tuner times its starting plan in full and rejects every alternative as soon as
    its fallback count is known (without scanning for it), so the starting plan stays; calls stay exact."""
    from mlxlearn_amd import _lib
    _lib.clear_plans()
    X, Q = oracle.gen_c2(seed=5, n=20_000, m=600)
    want = oracle.knn(X, Q, 5)
    index = mlk.DeviceNeighborIndex(X)
    try:
        _lib.set_test_hooks(_lib.HOOK_CAPACITY, 1)
        start = index.plan(600, 5, precision="bfloat16")
        plan = index.autotune(Q, 5, precision="bfloat16", max_trials=8)
        assert plan["tuned"] and plan["trials"] >= 2
        assert (plan["splits"], plan["query_tiles"], plan["stages"]) == \
            (start["splits"], start["query_tiles"], start["stages"])
        assert plan["tuned_ms"] == plan["default_ms"]
        got = knn(index, Q, 5, "bfloat16")
        assert _lib.stats()["fallback_queries"] > 600 // 256 + 16
        assert_parity(got, want)
    finally:
        _lib.set_test_hooks(0, 0)
        _lib.clear_plans()
        index.close()


@pytest.mark.parametrize("precision", ["bfloat16", "float16"])
def test_radius_shapes_match_oracle(mlk, precision):
    """Radius pairs screened with one and with two query tiles per wave (each LDS fragment feeding two
    MFMAs), sorted rows, at narrow widths, clumped rows and queries equal to rows."""
    from mlxlearn_amd import _lib
    r = np.random.default_rng(4600)
    for d in (20, 32, 64, 128):
        X = r.uniform(size=(12_000, d)).astype(np.float32)
        X[:2000] = X[:2000] * 0.05 + 0.5
        X[3000:3004] = X[17]
        Q = np.concatenate([X[:300], r.uniform(size=(213, d)).astype(np.float32)])
        R2 = 0.9 * d / 12  # about the clump centre's distance from a uniform query: rows of 0 to ~5,000
        want = oracle.radius(X, Q, R2, sort_results=True)
        index = mlk.DeviceNeighborIndex(X)
        try:
            for qt in (1, 2):
                _lib.set_tuning(query_tiles=qt, stages=2)
                assert index.plan(Q.shape[0], radius=True, precision=precision)["query_tiles"] == qt
                indptr, indices, dist = index.radius(Q, R2, precision=precision, sort_results=True)
                np.testing.assert_array_equal(indptr.cpu().numpy(), want[0])
                np.testing.assert_array_equal(indices.cpu().numpy(), want[1])
                np.testing.assert_allclose(dist.cpu().numpy(), want[2], rtol=1e-12, atol=0)
        finally:
            _lib.set_tuning()
            index.close()


def test_autotuned_radius_plan_is_exact(mlk):
    from mlxlearn_amd import _lib
    _lib.clear_plans()
    X, Q, sizes, radius = oracle.gen_c4(seed=3, n=100_000, n_batches=12, max_batch=400)
    m = Q.shape[0]
    R2 = radius * radius
    want = oracle.radius(X, Q, R2, sort_results=True)
    index = mlk.DeviceNeighborIndex(X)
    try:
        plan = index.autotune(Q, radius_sq=R2, sort_results=True, max_trials=8)
        assert plan["tuned"] and plan["rounds"] == [] and plan["tuned_ms"] <= plan["default_ms"]
        assert index.plan(m, radius=True)["splits"] == plan["splits"]
        indptr, indices, dist = index.radius(Q, R2, sort_results=True)
        assert _lib.stats()["splits"] == plan["splits"]
        np.testing.assert_array_equal(indptr.cpu().numpy(), want[0])
        np.testing.assert_array_equal(indices.cpu().numpy(), want[1])
        np.testing.assert_allclose(dist.cpu().numpy(), want[2], rtol=1e-12, atol=0)
    finally:
        _lib.clear_plans()
        index.close()


def test_kmeans_assignment_uses_tuned_plan(mlk):
    """A plan tuned on a centres index applies to mlk_kmeans_assign of the same shape."""
    from mlxlearn_amd import _lib
    from mlxlearn_amd.cluster._k_means import assign_device
    _lib.clear_plans()
    r = np.random.default_rng(77)
    X = r.normal(size=(40_000, 24)).astype(np.float32)
    C = X[r.choice(40_000, 300, replace=False)] + np.float32(0.01)
    index = mlk.DeviceNeighborIndex(C)
    try:
        plan = index.autotune(X, 1, precision="bfloat16", max_trials=10)
    finally:
        index.close()
    labels, resid = assign_device(X, C, precision="bfloat16")
    st = _lib.stats()
    assert st["splits"] == plan["splits"] and st["capacity"] == plan["capacity"]
    want_l, want_r = oracle.kmeans_assign(X, C)
    np.testing.assert_array_equal(labels.cpu().numpy(), want_l)
    np.testing.assert_allclose(resid.cpu().numpy(), want_r, rtol=1e-12, atol=0)
    _lib.clear_plans()


def test_mfma_probe_reports_a_plausible_rate(mlk):
    """Diagnostics: the register-only MFMA loop runs and reports a rate below the dense bf16 peak."""
    import ctypes
    torch = pytest.importorskip("torch")
    from mlxlearn_amd import _lib
    lib = _lib.load()
    out = ctypes.c_double()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.mlk_mfma_probe(256, 2000, ctypes.byref(out), s))
    assert 50.0 < out.value < 2600.0
    with pytest.raises(ValueError):
        _lib.check(lib.mlk_mfma_probe(0, 10, ctypes.byref(out), s))


def test_concurrent_queries_on_one_index(mlk):
    """Calls on one index from several host threads serialise on its lock; tuning knobs
    changed concurrently from another thread never change a result."""
    import threading
    torch = pytest.importorskip("torch")
    from mlxlearn_amd import _lib
    X, Q = oracle.gen_c2(seed=77, n=30_000, m=600, d=48)
    want = oracle.knn(X, Q, 6)
    index = mlk.DeviceNeighborIndex(X)
    errors, done = [], threading.Event()

    def worker(precision):
        try:
            with torch.cuda.stream(torch.cuda.Stream()):
                for _ in range(6):
                    d, i = index.query(Q, 6, precision=precision)
                    torch.cuda.current_stream().synchronize()
                    np.testing.assert_array_equal(i.cpu().numpy(), want[1])
        except Exception as e:  # reported below
            errors.append(e)

    def knobs():
        s = 0
        while not done.is_set():
            _lib.set_tuning(splits=(s % 5) * 8, rounds=[[0.5], [0.1, 0.5], []][s % 3])
            s += 1
        _lib.set_tuning()

    threads = [threading.Thread(target=worker, args=(p,)) for p in PRECISIONS]
    tk = threading.Thread(target=knobs)
    tk.start()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    done.set()
    tk.join()
    index.close()
    assert not errors, errors[0]
