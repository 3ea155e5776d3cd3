"""Shared estimator logic: capability checks, validation, ownership, query shaping.

Restates the behaviour of the reference's ``MLXNeighborsMixin`` /
``MLXRadiusMixin`` (``src/mlxlearn/neighbors/_base.py:195-785``): the same
capability errors before validation, scikit-learn's validation messages, an
owned float32 copy of the training data, float64 distances and int64 indices,
``X=None`` as a self-query that drops each row's own index, and
``compute_weights`` for distance voting. All searches run on the GPU through
the native library; there is no CPU route.
"""
from __future__ import annotations

from numbers import Integral

import numpy as np
from sklearn.utils.validation import check_is_fitted, validate_data

from .._diagnostics import record_event
from .._index import NeighborIndex, RadiusCSR
from ..exceptions import SparseInputError, UnsupportedParameterError

SUPPORTED_METRICS = ("euclidean", "l2", "minkowski")
SUPPORTED_ALGORITHMS = ("auto", "brute")
_FLT_MAX = float(np.finfo(np.float32).max)


def _is_sparse(X) -> bool:
    try:
        import scipy.sparse as sp
        return sp.issparse(X)
    except Exception:  # pragma: no cover
        return False


def require_dense(X, estimator: str, operation: str = "fit"):
    if _is_sparse(X):
        raise SparseInputError(f"{estimator}.{operation} does not accept sparse input; "
                               "pass a dense array (sparse matrices are not supported).")


def prepare_input(X):
    """Reject float values that overflow float32 before validation (_base.py:706-726)."""
    if isinstance(X, np.ndarray) and X.dtype.kind == "f" and X.dtype.itemsize > 4:
        with np.errstate(invalid="ignore"):
            finite = np.isfinite(X)
            if finite.all() and X.size and np.max(np.abs(X)) > _FLT_MAX:
                raise ValueError("Input contains values too large for float32.")
    return X


class NeighborsMixin:
    """fit / kneighbors / kneighbors_graph on the exact GPU index."""

    def _check_capability(self, X, y=None):
        name = type(self).__name__
        require_dense(X, name)
        if self.algorithm not in SUPPORTED_ALGORITHMS:
            raise UnsupportedParameterError(
                f"{name} supports algorithm in {SUPPORTED_ALGORITHMS} (exact brute force); got {self.algorithm!r}.",
                parameter="algorithm", value=self.algorithm)
        if self.metric not in SUPPORTED_METRICS:
            raise UnsupportedParameterError(
                f"{name} implements the Euclidean metric only; got metric={self.metric!r}.",
                parameter="metric", value=self.metric)
        if self.metric == "minkowski" and self.p != 2:
            raise UnsupportedParameterError(
                f"{name} implements metric='minkowski' only at p=2 (Euclidean); got p={self.p!r}.",
                parameter="p", value=self.p)
        if getattr(self, "metric_params", None):
            raise UnsupportedParameterError(
                f"{name} does not implement metric_params; got {self.metric_params!r}. It will not ignore it "
                "and return unweighted Euclidean distances.", parameter="metric_params", value=self.metric_params)
        weights = getattr(self, "weights", "uniform")
        if callable(weights):
            raise UnsupportedParameterError(f"{name} does not implement a callable weights function.",
                                            parameter="weights", value=weights)
        if y is not None:
            y_arr = np.asarray(y)
            if y_arr.ndim > 1 and y_arr.shape[1] > 1:
                raise UnsupportedParameterError(
                    f"{name} supports single-output targets only; got y with shape {y_arr.shape}.",
                    parameter="y", value=y_arr.shape)
        return prepare_input(X)

    def _fit_neighbors(self, X, y=None, *, target=None):
        X = self._check_capability(X, y)
        if target is None:
            X_checked = validate_data(self, X, accept_sparse=False, dtype=np.float32, order="C",
                                      ensure_min_samples=1, reset=True)
        else:
            X_checked, y_checked = validate_data(self, X, y, accept_sparse=False, dtype=np.float32, order="C",
                                                 ensure_min_samples=1, multi_output=False,
                                                 y_numeric=target == "regression", reset=True)
            if target == "classification":
                from sklearn.utils.multiclass import check_classification_targets
                check_classification_targets(y_checked)
            self._y_checked_ = y_checked
        if not np.all(np.isfinite(X_checked)):
            raise ValueError("Input X contains infinity or a value too large for dtype('float32').")
        self._index_ = NeighborIndex(X_checked, precision=getattr(self, "precision", "float32"))
        # Inference writes only inside this fitted container (the backend of the last query and lazily
        # uploaded label / target tensors), so predict never adds or rebinds an estimator attribute.
        self._query_state_ = {"backend": None}
        self.n_samples_fit_ = int(X_checked.shape[0])
        self.effective_metric_ = "euclidean"
        self.effective_metric_params_ = {}
        self._fit_method = "brute"
        record_event(type(self).__name__, "fit", shape=X_checked.shape, search=False,
                     detail="training rows owned; the device index is built on the first query")
        return self

    @property
    def _fit_X(self):
        return self._index_.X

    @property
    def last_query_backend_(self):
        """Backend of the last query (always ``"gpu"``); AttributeError before the first query."""
        backend = self.__dict__.get("_query_state_", {}).get("backend")
        if backend is None:
            raise AttributeError("last_query_backend_")
        return backend

    def _set_query_backend(self, backend):
        self._query_state_["backend"] = backend

    def _device_cache(self, key, make):
        """A device tensor kept for the fitted state's lifetime (dropped on pickling)."""
        state = self._query_state_
        if state.get(key) is None:
            state[key] = make()
        return state[key]

    def _resolve_n_neighbors(self, n_neighbors, query_is_train: bool) -> int:
        """scikit-learn's messages (_base.py:466-488)."""
        if n_neighbors is None:
            n_neighbors = self.n_neighbors
        if not isinstance(n_neighbors, Integral) or isinstance(n_neighbors, bool):
            raise TypeError(f"n_neighbors does not take {type(n_neighbors)} value, enter integer value")
        if n_neighbors <= 0:
            raise ValueError(f"Expected n_neighbors > 0. Got {n_neighbors}")
        limit = self.n_samples_fit_ - int(query_is_train)
        if n_neighbors > limit:
            inequality = "n_neighbors < n_samples_fit" if query_is_train else "n_neighbors <= n_samples_fit"
            raise ValueError(f"Expected {inequality}, but n_neighbors = {n_neighbors}, "
                             f"n_samples_fit = {self.n_samples_fit_}")
        return int(n_neighbors)

    def _validate_query(self, X, operation="kneighbors"):
        require_dense(X, type(self).__name__, operation)
        X = prepare_input(X)
        Q = validate_data(self, X, accept_sparse=False, dtype=np.float32, order="C", reset=False)
        if not np.all(np.isfinite(Q)):
            raise ValueError("Input X contains infinity or a value too large for dtype('float32').")
        return Q

    def _prepare_neighbor_query(self, X, n_neighbors=None):
        query_is_train = X is None
        k = self._resolve_n_neighbors(n_neighbors, query_is_train)
        Q = self._index_.X if query_is_train else self._validate_query(X)
        return Q, k, query_is_train

    def kneighbors(self, X=None, n_neighbors=None, return_distance=True):
        """k nearest neighbours: float64 distances ascending, int64 indices by (distance, index)."""
        check_is_fitted(self, "_index_")
        Q, k, query_is_train = self._prepare_neighbor_query(X, n_neighbors)
        dist, idx = self._index_.query(Q, k, return_distance=return_distance, exclude_self=query_is_train)
        self._set_query_backend(self._index_.last_query_backend)
        record_event(type(self).__name__, "kneighbors", shape=Q.shape, search=Q.shape[0] > 0, k=k)
        return (dist, idx) if return_distance else idx

    def kneighbors_graph(self, X=None, n_neighbors=None, mode="connectivity"):
        """CSR graph of the selected neighbours (X=None excludes self) (_base.py:426-444)."""
        check_is_fitted(self, "_index_")
        if mode not in ("connectivity", "distance"):
            raise ValueError('Unsupported mode, must be one of "connectivity", '
                             f'or "distance" but got "{mode}" instead')
        if mode == "distance":
            distances, indices = self.kneighbors(X, n_neighbors, return_distance=True)
            values = distances.ravel()
        else:
            indices = self.kneighbors(X, n_neighbors, return_distance=False)
            values = np.ones(indices.size)
        from scipy import sparse
        count, k = indices.shape
        indptr = np.arange(0, indices.size + 1, k)
        return sparse.csr_matrix((values, indices.ravel(), indptr), shape=(count, self.n_samples_fit_))

    def __getstate__(self):
        state = dict(self.__dict__)
        if "_query_state_" in state:
            state["_query_state_"] = {"backend": state["_query_state_"].get("backend")}
        return state


class RadiusMixin:
    """radius_neighbors / radius_neighbors_graph with exact membership S <= fl64(r**2)."""

    def _radius_csr(self, X, radius, *, need_distance, sort_results, operation) -> RadiusCSR:
        check_is_fitted(self, "_index_")
        query_is_train = X is None
        Q = self._index_.X if query_is_train else self._validate_query(X, "radius_neighbors")
        radius = self.radius if radius is None else radius
        csr = self._index_.query_radius_csr(Q, radius, need_distance=need_distance or sort_results,
                                            exclude_self=query_is_train, sort_results=sort_results)
        self._set_query_backend(self._index_.last_query_backend)
        record_event(type(self).__name__, operation, shape=Q.shape, search=Q.shape[0] > 0, nnz=csr.nnz)
        return csr

    def radius_neighbors(self, X=None, radius=None, return_distance=True, sort_results=False):
        """Rows within an inclusive radius; object arrays as in scikit-learn (_base.py:634-675)."""
        if sort_results and not return_distance:
            raise ValueError("return_distance must be True if sort_results is True.")
        csr = self._radius_csr(X, radius, need_distance=return_distance, sort_results=sort_results,
                               operation="radius_neighbors")
        return csr.to_ragged(return_distance=return_distance)

    def radius_neighbors_graph(self, X=None, radius=None, mode="connectivity", sort_results=False):
        if mode not in ("connectivity", "distance"):
            raise ValueError('Unsupported mode, must be one of "connectivity", '
                             f'or "distance" but got "{mode}" instead')
        csr = self._radius_csr(X, radius, need_distance=mode == "distance", sort_results=sort_results,
                               operation="radius_neighbors_graph")
        data = np.ones(csr.nnz) if mode == "connectivity" else csr.distances
        from scipy import sparse
        return sparse.csr_matrix((data, csr.indices, csr.indptr), shape=(csr.n_query, self.n_samples_fit_))


def compute_weights(dist, weights):
    """Per-neighbour weights (restates _base.py:752-785)."""
    if weights in (None, "uniform"):
        return None
    if weights != "distance":
        raise ValueError(f"weights not recognized: should be 'uniform' or 'distance', got {weights!r}")
    if dist is None:
        raise ValueError("weights='distance' requires neighbor distances.")
    with np.errstate(divide="ignore"):
        w = 1.0 / np.asarray(dist, dtype=np.float64)
    coincident = np.isinf(w)
    rows = np.any(coincident, axis=1)
    if np.any(rows):
        w[rows] = coincident[rows].astype(np.float64)
    if np.any(np.all(w == 0, axis=1)):
        raise ValueError("All neighbors of some sample are getting zero weights. Please modify "
                         "'weights' to avoid this case if you are using a user-defined function.")
    return w
