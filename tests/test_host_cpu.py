"""Host-side logic checked on the CPU: argument checks that must fire before any
native call, and scikit-learn's outlier-label contract of the radius classifier."""
import warnings

import numpy as np
import pytest


class _NoNative:
    """Stands in for the ctypes library: any native call fails the test."""

    def __getattr__(self, name):
        raise AssertionError(f"native entry point {name} reached before argument checks")


def _index(n_features=4):
    from mlxlearn_amd._index import DeviceNeighborIndex
    index = object.__new__(DeviceNeighborIndex)
    index._lib = _NoNative()
    index._handle = None
    index.n_samples, index.n_features = 100, n_features
    index._device = "cpu"
    return index


@pytest.mark.parametrize("Q", [np.zeros((3, 5), np.float32), np.zeros(4, np.float32), np.zeros((2, 2, 4), np.float32)])
def test_query_width_checked_before_native_calls(Q):
    index = _index()
    with pytest.raises(ValueError, match="4 features"):
        index.query(Q, 2)
    with pytest.raises(ValueError, match="4 features"):
        index.radius(Q, 1.0)
    with pytest.raises(ValueError, match="4 features"):
        index.radius_reduce(Q, 1.0, targets=np.zeros(100))
    with pytest.raises(ValueError, match="4 features"):
        index.query_batched([np.zeros((2, 4), np.float32), Q], 2)
    with pytest.raises(ValueError, match="4 features"):
        index.autotune(Q, 3)


def test_outlier_label_contract():
    from mlxlearn_amd.neighbors._estimators import _resolve_outlier_label as resolve
    classes = np.array([0, 1, 2])
    y = np.array([0, 1, 1, 2, 2])
    assert resolve(None, classes, y) is None
    assert resolve("most_frequent", classes, y) == [1]          # tie between 1 and 2 -> smaller class
    assert resolve(2, classes, y) == [2]
    assert resolve([2], classes, y) == [2]
    assert resolve(np.array([7]), classes, y) == [7]             # outside the classes is allowed
    with pytest.raises(ValueError, match="inconsistent with the output length"):
        resolve([1, 2], classes, y)
    with pytest.raises(TypeError, match="supposed to be a scalar"):
        resolve([[1]], classes, y)
    with pytest.raises(TypeError, match="dtype of outlier_label"):
        resolve(1.5, classes, y)
    strings = np.array(["a", "b"])
    assert resolve("b", strings, np.array([0, 1])) == ["b"]
    with pytest.raises(TypeError, match="dtype of outlier_label"):
        resolve("long", strings, np.array([0, 1]))


def test_outlier_votes():
    from mlxlearn_amd.neighbors._estimators import _fill_outliers as fill
    classes = np.array([3, 5, 9])
    votes = np.array([[1.0, 0, 0], [0, 0, 0], [0, 2.0, 0]])
    out = fill(votes.copy(), np.array([1, 0, 2]), [9], classes)
    np.testing.assert_array_equal(out[1], [0, 0, 1])              # one-hot for a known label
    np.testing.assert_array_equal(out[[0, 2]], votes[[0, 2]])
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        out = fill(votes.copy(), np.array([1, 0, 2]), [4], classes)
    assert any("not in training classes" in str(w.message) for w in caught)
    np.testing.assert_array_equal(out[1], 0)                      # unknown label: all zero
    with pytest.raises(ValueError, match="No neighbors found for test samples"):
        fill(votes.copy(), np.array([1, 0, 2]), None, classes)
    same = fill(votes.copy(), np.array([1, 1, 2]), None, classes)  # no outlier: untouched
    np.testing.assert_array_equal(same, votes)
