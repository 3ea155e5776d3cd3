"""Backend events carry the native statistics of the search that served each call, and tuning overrides reach
the native plan without changing any result."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _data(n=3000, m=200, d=48, seed=11):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, d)).astype(np.float32)
    X[500:520] = X[3]  # exact ties
    Q = rng.normal(size=(m, d)).astype(np.float32)
    return X, Q


def test_events_carry_search_statistics():
    import mlxlearn_amd as m
    X, Q = _data()
    y = np.arange(X.shape[0]) % 5
    m.clear_backend_diagnostics()
    nn = m.NearestNeighbors(n_neighbors=7, precision="bfloat16").fit(X)
    nn.kneighbors(Q)
    event = m.get_last_backend_event()
    assert (event.estimator, event.operation, event.shape) == ("NearestNeighbors", "kneighbors", Q.shape)
    info = event.extra
    assert info["k"] == 7 and info["queries"] == Q.shape[0] and info["screen_precision"] == "bfloat16"
    assert info["candidates"] >= 7 * Q.shape[0] and 7 * Q.shape[0] <= info["reranked"] <= info["candidates"]
    assert info["rounds"] >= 1 and info["splits"] >= 1 and info["total_ms"] > 0
    nn.radius_neighbors(Q[:20], radius=8.0)
    event = m.get_last_backend_event()
    assert event.operation == "radius_neighbors" and event.extra["queries"] == 20
    assert event.extra["nnz"] == sum(len(r) for r in nn.radius_neighbors(Q[:20], radius=8.0,
                                                                           return_distance=False))
    m.KNeighborsClassifier(n_neighbors=3).fit(X, y).predict(Q)
    assert m.get_last_backend_event().operation == "predict"
    km = m.KMeans(n_clusters=16, init="random", n_init=1, max_iter=3, random_state=0).fit(X)
    event = m.get_last_backend_event()
    assert event.operation == "fit" and event.extra["n_iter"] == km.n_iter_
    summary = m.diagnostics_summary()
    assert summary["NearestNeighbors.kneighbors -> gpu"] == 1 and summary["KMeans.fit -> gpu"] == 1
    with m.config_context(diagnostics=False):
        nn.kneighbors(Q[:5])
    assert m.get_last_backend_event() is event


def test_tuning_context_reaches_the_plan_and_not_the_results():
    import mlxlearn_amd as m
    from mlxlearn_amd import _config
    X, Q = _data(seed=12)
    want_d, want_i = oracle.knn(X, Q, 10)
    nn = m.NearestNeighbors(n_neighbors=10).fit(X)
    with _config.tuning_context(screen_splits=3, candidate_capacity=64, screen_rounds="0.25,1"):
        d, i = nn.kneighbors(Q)
        info = m.get_last_backend_event().extra
        assert info["splits"] == 3 and info["capacity"] == 64 and info["rounds"] <= 2
    np.testing.assert_array_equal(i, want_i)
    np.testing.assert_allclose(d, want_d, rtol=1e-12, atol=0)
    nn.kneighbors(Q)
    assert m.get_last_backend_event().extra["splits"] != 3 or m.get_last_backend_event().extra["capacity"] != 64
