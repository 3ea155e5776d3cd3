"""scikit-learn's estimator check suite against every estimator of the package.

The reference runs ``check_estimator`` over its neighbour estimators
(tests/test_sklearn_compliance.py:63-124) with no exemptions of its own. The
same holds here for the five neighbour estimators. KMeans is checked as well,
which the reference does not do; its only exemptions are the sample-weight
checks, because KMeans rejects ``sample_weight`` exactly as the reference's
does (src/mlxlearn/cluster/_k_means.py:253-254). Every exemption must still
fail, so one cannot outlive the behaviour that justified it.
"""
import warnings

import pytest

pytestmark = pytest.mark.gpu

NEIGHBORS = ["NearestNeighbors", "KNeighborsClassifier", "KNeighborsRegressor", "RadiusNeighborsClassifier",
             "RadiusNeighborsRegressor"]

# (estimator, check) -> reason; each must fail, and nothing else may.
EXEMPT = {("KMeans", name): "sample_weight is rejected, as in the reference (cluster/_k_means.py:253-254)"
          for name in ("check_sample_weights_pandas_series", "check_sample_weights_not_an_array",
                       "check_sample_weights_list", "check_sample_weights_shape",
                       "check_sample_weights_not_overwritten", "check_sample_weight_equivalence_on_dense_data")}


@pytest.mark.parametrize("name", NEIGHBORS + ["KMeans"])
def test_passes_sklearn_checks(name):
    from sklearn.utils.estimator_checks import check_estimator

    import mlxlearn_amd

    estimator = getattr(mlxlearn_amd, name)()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        results = check_estimator(estimator, on_fail=None)
    assert len(results) > 30, f"only {len(results)} checks ran for {name}"
    failed = {r["check_name"]: r["exception"] for r in results if r["status"] == "failed"}
    unexpected = {c: e for c, e in failed.items() if (name, c) not in EXEMPT}
    assert not unexpected, f"{name} failed {len(unexpected)} checks:\n" + "\n".join(
        f"  {c}: {type(e).__name__}: {str(e)[:300]}" for c, e in unexpected.items())
    stale = [c for (n, c) in EXEMPT if n == name and c not in failed]
    assert not stale, f"exempted checks now pass for {name}; remove the exemptions: {stale}"


def test_inference_keeps_estimator_attributes_and_reports_backend():
    """predict rebinds no attribute; last_query_backend_ names the GPU after a query; pickling drops the
    device caches and the unpickled estimator predicts the same labels (reference neighbors/_base.py:161-175,
    515-538)."""
    import pickle

    import numpy as np

    import mlxlearn_amd

    rng = np.random.default_rng(3)
    X = rng.normal(size=(300, 8)).astype(np.float32)
    y = (X[:, 0] > 0).astype(np.int64)
    for est in (mlxlearn_amd.KNeighborsClassifier(n_neighbors=3).fit(X, y),
                mlxlearn_amd.RadiusNeighborsRegressor(radius=2.0).fit(X, y.astype(np.float64))):
        with pytest.raises(AttributeError):
            est.last_query_backend_
        before = {key: id(value) for key, value in est.__dict__.items()}
        first = est.predict(X[:50])
        assert {key: id(value) for key, value in est.__dict__.items()} == before
        assert est.last_query_backend_ == "gpu"
        clone = pickle.loads(pickle.dumps(est))
        assert set(clone._query_state_) == {"backend"}
        np.testing.assert_array_equal(clone.predict(X[:50]), first)
