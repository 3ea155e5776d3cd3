"""CPU-side checks of the native library and the host logic (no device compute)."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mlxlearn_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mlk_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from mlxlearn_amd import _lib
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), f"{name} declared in include/mlxlearn_amd.h but not exported"
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures must cover exactly the header"
    assert lib.mlk_version() == 1


def test_library_validation_without_device():
    """Argument validation runs before any device call and returns MLK_EINVAL with a message."""
    import ctypes
    from mlxlearn_amd import _lib
    lib = _lib.load()
    out = ctypes.c_void_p()
    rc = lib.mlk_index_create(None, 10, 3, 0, None, ctypes.byref(out))
    assert rc == 1
    assert b"device pointer" in lib.mlk_last_error()
    rc = lib.mlk_knn_query(None, None, 0, 1, 0, 0, None, None, None)
    assert rc == 1 and b"index is NULL" in lib.mlk_last_error()
    assert lib.mlk_set_tuning(-1, 0, 0, 0, 0, None) == 1
    assert lib.mlk_set_tuning(0, 0, 0, 1, 0, None) == 1  # a one-stage ring does not exist
    plan = _lib.Plan()
    assert lib.mlk_autotune(None, None, 10, 5, 0.0, 0, 0, ctypes.byref(plan), None) == 1
    assert b"index is NULL" in lib.mlk_last_error()
    assert lib.mlk_autotune(None, None, 10, 5, 0.0, 0, 0, None, None) == 1
    assert lib.mlk_get_plan(None, 10, 5, 0, ctypes.byref(plan)) == 1
    assert lib.mlk_clear_plans() == 0
    with pytest.raises(ValueError):
        _lib.check(1)
    with pytest.raises(MemoryError):
        _lib.check(-1)
    from mlxlearn_amd.exceptions import BackendExecutionError
    with pytest.raises(BackendExecutionError):
        _lib.check(-2)


def test_estimators_fail_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from mlxlearn_amd import NearestNeighbors
    from mlxlearn_amd.exceptions import DeviceError
    nn = NearestNeighbors(n_neighbors=2).fit(np.eye(3))
    with pytest.raises(DeviceError):
        nn.kneighbors(np.eye(3))


@pytest.mark.parametrize(
    "kwargs, parameter",
    [({"algorithm": "kd_tree"}, "algorithm"), ({"algorithm": "ball_tree"}, "algorithm"),
     ({"metric": "manhattan"}, "metric"), ({"metric": "cosine"}, "metric"),
     ({"metric": "minkowski", "p": 1}, "p"), ({"metric": "minkowski", "p": 3}, "p"),
     ({"metric_params": {"w": [1.0]}}, "metric_params"), ({"weights": lambda d: d}, "weights")])
def test_unsupported_parameters_raise_precisely(kwargs, parameter):
    """tests/parity/test_neighbors.py:308-336."""
    from mlxlearn_amd import KNeighborsClassifier
    from mlxlearn_amd.exceptions import UnsupportedParameterError
    X = np.random.default_rng(0).normal(size=(20, 5))
    y = np.arange(20) % 2
    with pytest.raises(UnsupportedParameterError) as excinfo:
        KNeighborsClassifier(n_neighbors=3, **kwargs).fit(X, y)
    assert excinfo.value.parameter == parameter


def test_sparse_and_bad_inputs():
    import scipy.sparse as sp
    from mlxlearn_amd import KNeighborsClassifier, NearestNeighbors
    from mlxlearn_amd.exceptions import SparseInputError, UnsupportedParameterError
    X = np.random.default_rng(0).normal(size=(20, 5))
    y = np.arange(20) % 2
    with pytest.raises(SparseInputError) as e:
        KNeighborsClassifier(n_neighbors=3).fit(sp.csr_matrix(X), y)
    assert isinstance(e.value, TypeError) and "sparse" in str(e.value).lower()
    with pytest.raises(UnsupportedParameterError, match="single-output"):
        KNeighborsClassifier(n_neighbors=3).fit(X, np.stack([y, y], 1))
    Xn = X.copy()
    Xn[0, 0] = np.nan
    with pytest.raises(ValueError):
        NearestNeighbors(n_neighbors=2).fit(Xn)
    with pytest.raises(ValueError, match="float32"):
        NearestNeighbors(n_neighbors=2).fit(np.array([[1e300], [1.0], [2.0]]))


def test_n_neighbors_messages():
    """sklearn's messages (reference _base.py:466-488)."""
    from mlxlearn_amd import NearestNeighbors
    nn = NearestNeighbors(n_neighbors=2).fit(np.eye(6))
    with pytest.raises(ValueError, match="Expected n_neighbors > 0. Got 0"):
        nn.kneighbors(np.eye(6), n_neighbors=0)
    with pytest.raises(ValueError, match="n_neighbors <= n_samples_fit"):
        nn.kneighbors(np.eye(6), n_neighbors=7)
    with pytest.raises(ValueError, match="n_neighbors < n_samples_fit"):
        nn.kneighbors(None, n_neighbors=6)
    with pytest.raises(TypeError, match="does not take"):
        nn.kneighbors(np.eye(6), n_neighbors=2.5)


def test_radius_validation_and_csr_shapes():
    from mlxlearn_amd._index import RadiusCSR, validate_radius
    assert validate_radius(-0.0) == 0.0
    with pytest.raises(ValueError):
        validate_radius(float("nan"))
    with pytest.raises(ValueError):
        validate_radius(-1)
    with pytest.raises(TypeError):
        validate_radius(True)
    with pytest.raises(TypeError):
        validate_radius(np.ones(2))
    csr = RadiusCSR(np.array([0, 2, 2, 3]), np.array([4, 7, 1]), np.array([0.5, 0.75, 0.0]))
    d, i = csr.to_ragged()
    assert [list(r) for r in i] == [[4, 7], [], [1]]
    assert [list(r) for r in d] == [[0.5, 0.75], [], [0.0]]


def test_compute_weights_rules():
    """reference _base.py:752-785: 1/d, coincident rows become indicators, all-zero rows raise."""
    from mlxlearn_amd.neighbors import compute_weights
    assert compute_weights(None, "uniform") is None
    w = compute_weights(np.array([[1.0, 2.0], [0.0, 3.0]]), "distance")
    np.testing.assert_array_equal(w, [[1.0, 0.5], [1.0, 0.0]])
    with pytest.raises(ValueError):
        compute_weights(np.array([[1.0]]), "bogus")


def test_kmeans_parameter_validation():
    from mlxlearn_amd import KMeans
    for kw in ({"n_clusters": 0}, {"max_iter": 0}, {"tol": -1}, {"algorithm": "elkan"}, {"init": "bogus"}):
        with pytest.raises(ValueError):
            KMeans(**kw)._validate_parameters()


def test_pickle_drops_device_state():
    import pickle
    from mlxlearn_amd import NearestNeighbors
    nn = NearestNeighbors(n_neighbors=2).fit(np.eye(4))
    nn2 = pickle.loads(pickle.dumps(nn))
    np.testing.assert_array_equal(nn2._fit_X, np.eye(4, dtype=np.float32))
    assert nn2._index_._device_index is None


def test_radius_estimators_host_logic():
    """RadiusNeighborsClassifier outlier-label rules (reference neighbors/_radius.py:341-380) and loud failure
    without a device."""
    from mlxlearn_amd import RadiusNeighborsClassifier, RadiusNeighborsRegressor
    from mlxlearn_amd.exceptions import DeviceError
    from mlxlearn_amd.neighbors._estimators import _resolve_outlier_label
    classes = np.array([1, 2, 5])
    assert _resolve_outlier_label(None, classes, np.array([0])) is None
    assert _resolve_outlier_label("most_frequent", classes, np.array([0, 2, 2, 1])) == [5]
    assert _resolve_outlier_label(7, classes, np.array([0])) == [7]
    assert _resolve_outlier_label([7], classes, np.array([0])) == [7]
    with pytest.raises(ValueError, match="inconsistent with the output length"):
        _resolve_outlier_label([7, 8], classes, np.array([0]))
    with pytest.raises(TypeError, match="supposed to be a scalar"):
        _resolve_outlier_label([[7]], classes, np.array([0]))
    with pytest.raises(TypeError, match="inconsistent with classes"):
        _resolve_outlier_label("x", classes, np.array([0]))
    X, y = np.array([[0.0], [1.0], [2.0], [3.0]]), np.array([0, 0, 1, 1])
    clf = RadiusNeighborsClassifier(radius=1.0, outlier_label="most_frequent").fit(X, y)
    assert clf.outlier_label_ == [0] and list(clf.classes_) == [0, 1]
    reg = RadiusNeighborsRegressor(radius=1.0).fit(X, y.astype(np.float32))
    for est in (clf, reg):
        if not _gpu():
            with pytest.raises(DeviceError):
                est.predict([[1.5]])


def _gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_plan_cache_text_round_trip():
    """Exported plans import back unchanged; a plan naming a shape that is not built is rejected whole."""
    from mlxlearn_amd import _lib
    _lib.clear_plans()
    assert _lib.export_plans() == ""
    line = "plan 0 1 128 10 40 27 1 2 24 48 5 0x1p-8 0x1p-5 0x1p-3 0x1p-1 0x1p+0 0x1.dp+1 0x1.fp+1 16\n"
    radius = "plan 1 0 32 0 42 34 1 2 8 0 0 0x1.cp+7 0x1.1p+8 5\n"
    _lib.import_plans(line + radius)
    text = _lib.export_plans()
    assert sorted(text.splitlines()) == sorted([line.strip(), radius.strip()])
    for bad in ("plan 0 1 128 10 40 27 3 2 24 48 5 0x1p-8 0x1p-5 0x1p-3 0x1p-1 0x1p+0 1 2 3\n",  # qt 3
                "plan 0 1 128 10 40 27 1 2 24 48 2 0x1p-1 0x1p-2 1 2 3\n",                          # decreasing
                "plan 1 0 32 0 42 34 1 2 8 0 1 0x1p+0 1 2 3\n",                                   # radius rounds
                "nonsense\n"):
        with pytest.raises(ValueError, match="not a valid plan"):
            _lib.import_plans(line + bad)
    assert sorted(_lib.export_plans().splitlines()) == sorted([line.strip(), radius.strip()])
    _lib.clear_plans()


def test_tags_narrow_to_what_is_supported():
    """Dense, single-output, no NaN (reference tests/test_sklearn_compliance.py:166-178)."""
    from sklearn.utils import get_tags
    import mlxlearn_amd
    for name in ("NearestNeighbors", "KNeighborsClassifier", "KNeighborsRegressor", "RadiusNeighborsClassifier",
                 "RadiusNeighborsRegressor", "KMeans"):
        tags = get_tags(getattr(mlxlearn_amd, name)())
        assert tags.input_tags.sparse is False and tags.input_tags.allow_nan is False, name
        if tags.target_tags is not None:
            assert tags.target_tags.multi_output is False, name


def test_kmeans_validates_input_before_the_device():
    """Sparse, empty, non-finite and wrong-width inputs raise scikit-learn's errors with or without a device."""
    import scipy.sparse as sp
    from mlxlearn_amd import KMeans
    from mlxlearn_amd.exceptions import SparseInputError
    X = np.random.default_rng(0).normal(size=(20, 3))
    with pytest.raises(SparseInputError, match="sparse"):
        KMeans(n_clusters=2).fit(sp.csr_matrix(X))
    with pytest.raises(ValueError, match="0 sample"):
        KMeans(n_clusters=2).fit(np.empty((0, 3)))
    Xn = X.copy()
    Xn[3, 1] = np.inf
    with pytest.raises(ValueError, match="infinity"):
        KMeans(n_clusters=2).fit(Xn)
    with pytest.raises(ValueError, match="float32"):
        KMeans(n_clusters=2).fit(np.array([[1e300], [1.0], [2.0]]))
    with pytest.raises(ValueError, match="sample_weight"):
        KMeans(n_clusters=2).fit(X, sample_weight=np.ones(20))


def test_config_surface_and_scoping():
    """Six public options, thread-local and scoped (reference _common/config.py:49-117, 278-330)."""
    import threading
    import mlxlearn_amd as m
    from mlxlearn_amd._config import PUBLIC_CONFIG_FIELDS
    assert PUBLIC_CONFIG_FIELDS == ("device", "fallback_policy", "output_type", "deterministic", "random_state",
                                    "diagnostics")
    with pytest.raises(TypeError, match="unknown configuration option"):
        m.set_config(bogus=1)
    with pytest.raises(ValueError, match="no CPU route"):
        m.set_config(device="cpu")
    with pytest.raises(ValueError, match="not implemented"):
        m.set_config(output_type="mlx")
    with pytest.raises(TypeError):
        m.set_config(diagnostics=1)
    assert m.get_config().device == "auto"
    seen = []
    with pytest.raises(RuntimeError):
        with m.config_context(device="gpu", random_state=7):
            assert m.get_config().random_state == 7
            t = threading.Thread(target=lambda: seen.append(m.get_config().random_state))
            t.start()
            t.join()
            raise RuntimeError("body fails")
    assert seen == [None] and m.get_config().random_state is None and m.get_config().device == "auto"


def test_tuning_environment_overrides(monkeypatch, tmp_path):
    """MLXLEARN_* variables override tuning; malformed values raise (reference config.py:251-275); a plan file is
    imported into the native plan cache."""
    from mlxlearn_amd import _config, _lib
    try:
        monkeypatch.setenv("MLXLEARN_SCREEN_SPLITS", "96")
        monkeypatch.setenv("MLXLEARN_SCREEN_ROUNDS", "0.125,0.5,1")
        _config.reset_config()
        t = _config.get_tuning()
        assert t.screen_splits == 96 and t.rounds() == [0.125, 0.5, 1.0]
        monkeypatch.setenv("MLXLEARN_SCREEN_SPLITS", "many")
        with pytest.raises(ValueError, match="MLXLEARN_SCREEN_SPLITS"):
            _config.reset_config()
        monkeypatch.setenv("MLXLEARN_SCREEN_SPLITS", "0")
        monkeypatch.setenv("MLXLEARN_SCREEN_ROUNDS", "0.5,0.25")
        with pytest.raises(ValueError, match="screen_rounds"):
            _config.reset_config()
        monkeypatch.delenv("MLXLEARN_SCREEN_ROUNDS")
        monkeypatch.setenv("MLXLEARN_QUERY_TILES", "3")
        with pytest.raises(ValueError, match="query_tiles"):
            _config.reset_config()
    finally:
        for key in ("MLXLEARN_SCREEN_SPLITS", "MLXLEARN_SCREEN_ROUNDS", "MLXLEARN_QUERY_TILES"):
            monkeypatch.delenv(key, raising=False)
        _config.reset_config()
    _lib.load()
    _lib.clear_plans()
    line = "plan 0 1 128 10 40 27 1 2 24 48 5 0x1p-8 0x1p-5 0x1p-3 0x1p-1 0x1p+0 0x1.dp+1 0x1.fp+1 16\n"
    path = tmp_path / "plans.txt"
    path.write_text(line)
    with _config.tuning_context(plan_file=str(path)):
        assert _lib.export_plans().strip() == line.strip()
    _lib.clear_plans()
    with pytest.raises(TypeError, match="unknown tuning"):
        with _config.tuning_context(bogus=1):
            pass


def test_diagnostics_ring_records_fit_and_is_bounded():
    """Events are recorded while diagnostics are on, in a ring bounded by diagnostics_capacity
    (reference _common/diagnostics.py:82-200)."""
    import mlxlearn_amd as m
    from mlxlearn_amd import _config
    m.clear_backend_diagnostics()
    X = np.random.default_rng(1).normal(size=(30, 4))
    m.NearestNeighbors(n_neighbors=2).fit(X)
    event = m.get_last_backend_event()
    assert (event.estimator, event.operation, event.backend, event.shape) == ("NearestNeighbors", "fit", "gpu",
                                                                             (30, 4))
    with m.config_context(diagnostics=False):
        m.NearestNeighbors(n_neighbors=2).fit(X)
    assert len(m.get_backend_diagnostics()) == 1
    with _config.tuning_context(diagnostics_capacity=3):
        for _ in range(5):
            m.NearestNeighbors(n_neighbors=2).fit(X)
        assert len(m.get_backend_diagnostics()) == 3
    assert m.diagnostics_summary() == {"NearestNeighbors.fit -> gpu": 3}
    m.clear_backend_diagnostics()
    assert m.get_last_backend_event() is None


def test_bad_environment_tuning_fails_every_load(monkeypatch):
    """A malformed MLXLEARN_* value is an error at each library load, never a silent default."""
    from mlxlearn_amd import _config, _lib
    monkeypatch.setenv("MLXLEARN_SCREEN_STAGES", "1")
    saved = _lib._LIB
    try:
        _lib._LIB = None
        with pytest.raises(ValueError, match="screen_stages"):
            _config.reset_config()
        for _ in range(2):
            with pytest.raises(ValueError, match="screen_stages"):
                _lib.load()
            assert _lib._LIB is None
    finally:
        monkeypatch.delenv("MLXLEARN_SCREEN_STAGES")
        _config.reset_config()
        _lib._LIB = saved


def test_reference_binding_loads_without_device():
    """The committed reference-side binding imports and binds every entry point it calls."""
    import sys
    from mlxlearn_amd import _lib
    os.environ["MLXLEARN_AMD_LIB"] = _lib.LIB_PATH
    sys.path.insert(0, os.path.join(ROOT, "mlx-learn_amd", "integration"))
    import _mi355x
    lib = _mi355x.library()
    for name in ("mlk_index_create", "mlk_knn_query", "mlk_radius_query", "mlk_radius_fill", "mlk_kmeans_assign"):
        assert hasattr(lib, name)
    assert _mi355x.SQUARED_DIST == _lib.SQUARED_DIST and _mi355x.PRECISION == _lib.PRECISIONS
    with pytest.raises(ValueError):
        _mi355x._check(1)
    with pytest.raises(MemoryError):
        _mi355x._check(-1)
