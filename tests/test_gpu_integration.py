"""The reference-side binding (``mlx-learn_amd/integration/_mi355x.py``) called
exactly as the reference calls its device seam, checked against the oracle.

``NeighborIndex._ensure_device`` builds the index with ``device=``
(``distance.py:173-176``); ``query_device`` passes the planner keywords and
``_squared`` (``distance.py:228-240``); ``NeighborIndex.query`` then returns
``np.sqrt(np.asarray(distances, float64))`` and ``np.asarray(indices, int64)``
(``distance.py:262-270``).
"""
import os
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def binding():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mlxlearn_amd import _lib
    os.environ["MLXLEARN_AMD_LIB"] = _lib.LIB_PATH
    sys.path.insert(0, os.path.join(ROOT, "mlx-learn_amd", "integration"))
    import _mi355x
    return _mi355x


def reference_query(index, Q, k, exclude_self=False):
    """The reference's NeighborIndex.query around the seam (distance.py:211-270)."""
    distances, indices = index.query(
        Q, k, return_distance=True, exclude_self=exclude_self, query_block_size=64, train_block_size=None,
        _validated=True, _squared=True, _auto_train_tiling=True, _gram_tile=None)
    return np.sqrt(np.asarray(distances, dtype=np.float64)), np.asarray(indices, dtype=np.int64)


@pytest.mark.parametrize("precision", ["float32", "bfloat16"])
def test_binding_query_as_the_reference_calls_it(binding, precision):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((6000, 40), dtype=np.float32)
    X[50:60] = X[3]
    Q = np.concatenate([X[:8], rng.standard_normal((120, 40), dtype=np.float32)])
    index = binding.DeviceNeighborIndex(X, device="gpu", precision=precision)
    d, i = reference_query(index, Q, 9)
    wd, wi = oracle.knn(X, Q, 9)
    np.testing.assert_array_equal(i, wi)
    np.testing.assert_allclose(d, wd, rtol=1e-12, atol=0)
    # _squared=False gives Euclidean distances directly
    d2, i2 = index.query(Q, 9, _squared=False)
    np.testing.assert_array_equal(i2, wi)
    np.testing.assert_allclose(d2, wd, rtol=1e-12, atol=0)
    # self query without the own row, as kneighbors(X=None) asks
    d3, i3 = reference_query(index, X, 4, exclude_self=True)
    w3 = oracle.knn(X, X, 4, exclude_self=True)
    np.testing.assert_array_equal(i3, w3[1])
    # return_distance=False
    none, i4 = index.query(Q, 3, return_distance=False)
    assert none is None
    np.testing.assert_array_equal(i4, wi[:, :3])
    index.close()


def test_binding_fit_neighbors_radius_and_assign(binding):
    rng = np.random.default_rng(8)
    X = rng.random((4000, 12), dtype=np.float32)
    Q = rng.random((90, 12), dtype=np.float32)
    index = binding.fit_neighbors(X, "gpu")
    r = 0.45
    indptr, ind, dist = index.radius_csr(Q, r, need_distance=True, sort_results=True)
    wp, wi, wd = oracle.radius(X, Q, r * r, sort_results=True)
    np.testing.assert_array_equal(indptr, wp)
    np.testing.assert_array_equal(ind, wi)
    np.testing.assert_allclose(dist, wd, rtol=1e-12, atol=0)
    indptr, ind, dist = index.radius_csr(Q, r, need_distance=False)
    assert dist is None
    np.testing.assert_array_equal(indptr, wp)
    index.close()
    C = X[:37].copy()
    labels, resid = binding.assign(X, C)
    wl, wr = oracle.kmeans_assign(X, C)
    assert labels.dtype == np.uint32 and resid.dtype == np.float32
    np.testing.assert_array_equal(labels, wl)
    np.testing.assert_array_equal(resid, wr.astype(np.float32))


def test_binding_errors(binding):
    X = np.ones((10, 3), np.float32)
    index = binding.DeviceNeighborIndex(X)
    with pytest.raises(ValueError):
        index.query(np.ones((2, 4), np.float32), 3)
    with pytest.raises(ValueError):
        index.query(np.ones((2, 3), np.float32), 11)
    with pytest.raises(ValueError):
        binding.assign(X, X, spherical=True)
    index.close()
