"""Reference fixtures of the screen and selection stages, restated, on the GPU.

* The monotone ``arange(8192 * 32)`` fixture and the clustered fixtures of
  ``tests/unit/test_neighbors_gram.py:19-47`` (whose soundness test checks that
  no screen lower bound exceeds the exact distance): here the screen's bound is
  exercised end to end, at every precision, and the neighbours must equal the
  exact oracle's. An unsound bound drops a true neighbour and fails.
* The selection invariants of ``tests/unit/test_neighbors_general_selection.py:91-156``
  (selection equals a stable sort by (value, index) for random, descending,
  tied, single-lane and front-loaded orders, k = 17..33): here the values are
  the squared distances of integer rows to a query at the origin, so the
  expected neighbours are the stable argsort of the values.
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


def _monotone():
    train = np.arange(8192 * 32, dtype=np.float32).reshape(8192, 32)
    return train, train[[0, 3, 4, 4096, 8191]].copy()


def _clustered(features):
    rng = np.random.default_rng(872)
    n = 32768 if features == 32 else 4096
    centers = rng.normal(size=(4, features)) * 8
    train = centers[rng.integers(0, 4, n)] + 0.2 * rng.normal(size=(n, features))
    queries = centers[rng.integers(0, 4, 65)] + 0.2 * rng.normal(size=(65, features))
    train, queries = train.astype(np.float32), queries.astype(np.float32)
    queries[:2] = train[:2]
    return train, queries


FIXTURES = {"monotone": _monotone, "clustered32": lambda: _clustered(32), "clustered128": lambda: _clustered(128)}


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("fixture", list(FIXTURES))
@pytest.mark.parametrize("k", [1, 10, 33])
def test_gram_fixtures_exact_neighbours(mlk, fixture, precision, k):
    X, Q = FIXTURES[fixture]()
    index = mlk.DeviceNeighborIndex(X)
    d, i = index.query(Q, k, precision=precision)
    wd, wi = oracle.knn(X, Q, k)
    np.testing.assert_array_equal(i.cpu().numpy(), wi)
    np.testing.assert_allclose(d.cpu().numpy(), wd, rtol=1e-12, atol=0)
    # the self queries of the clustered fixtures: the row itself first, at distance 0
    if fixture != "monotone":
        assert list(wi[:2, 0]) == [0, 1] and np.all(wd[:2, 0] == 0)
    index.close()


def _orders(rng, width):
    base = rng.integers(0, 150, size=width)
    yield "random", base
    yield "descending", -np.sort(-base) + 150
    yield "all_equal", np.full(width, 3)
    lane = base + 200
    lane[::32] = np.sort(base[::32])[::-1]            # every winner in one lane residue
    yield "one_lane", lane
    front = base + 200
    front[:2048] = np.sort(base[:2048])[::-1]          # every winner in the first tiles
    yield "front", front


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("width", [100, 2048, 4097, 9001])
def test_selection_is_a_stable_sort(mlk, precision, width):
    rng = np.random.default_rng(width)
    for name, values in _orders(rng, width):
        X = np.zeros((width, 4), np.float32)
        X[:, 0] = values                                # S(0, x_i) = values[i]^2, values >= 0
        index = mlk.DeviceNeighborIndex(X)
        q = np.zeros((1, 4), np.float32)
        expected = np.argsort(values.astype(np.int64) ** 2, kind="stable")
        for k in (1, 2, 16, 17, 20, 24, 31, 32, 33, 41, 64):
            if k > width:
                continue
            d, i = index.query(q, k, precision=precision)
            np.testing.assert_array_equal(i.cpu().numpy()[0], expected[:k], err_msg=f"{name} k={k}")
            np.testing.assert_array_equal(d.cpu().numpy()[0], np.sort(values.astype(np.float64))[:k],
                                          err_msg=f"{name} k={k}")
        index.close()
