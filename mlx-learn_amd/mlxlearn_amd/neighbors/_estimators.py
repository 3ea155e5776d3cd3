"""NearestNeighbors, KNeighborsClassifier and KNeighborsRegressor.

Constructor signatures and defaults follow the reference
(``src/mlxlearn/neighbors/_unsupervised.py:84-93``,
``_classification.py:71-80``) plus one extension, ``precision``: the screen
precision of the GPU search (``"float32"`` default, the drop-in;
``"bfloat16"`` / ``"float16"`` run a reduced-precision MFMA screen whose
certified candidates are re-ranked in float64, with identical results).
"""
from __future__ import annotations

import numpy as np
from sklearn.base import BaseEstimator, ClassifierMixin, RegressorMixin
from sklearn.utils.validation import check_is_fitted

from .. import _lib
from ._base import NeighborsMixin, RadiusMixin, compute_weights

__all__ = ["NearestNeighbors", "KNeighborsClassifier", "KNeighborsRegressor"]


class NearestNeighbors(RadiusMixin, NeighborsMixin, BaseEstimator):
    """Exact Euclidean nearest neighbours on an MI355X GPU."""

    def __init__(self, n_neighbors=5, *, radius=1.0, algorithm="auto", leaf_size=30, metric="minkowski", p=2,
                 metric_params=None, n_jobs=None, precision="float32"):
        self.n_neighbors = n_neighbors
        self.radius = radius
        self.algorithm = algorithm
        self.leaf_size = leaf_size
        self.metric = metric
        self.p = p
        self.metric_params = metric_params
        self.n_jobs = n_jobs
        self.precision = precision

    def fit(self, X, y=None):
        return self._fit_neighbors(X)


class KNeighborsClassifier(NeighborsMixin, ClassifierMixin, BaseEstimator):
    """k-NN classification; votes are counted on the GPU (smallest class code wins ties)."""

    def __init__(self, n_neighbors=5, *, weights="uniform", algorithm="auto", leaf_size=30, p=2,
                 metric="minkowski", metric_params=None, n_jobs=None, precision="float32"):
        self.n_neighbors = n_neighbors
        self.weights = weights
        self.algorithm = algorithm
        self.leaf_size = leaf_size
        self.p = p
        self.metric = metric
        self.metric_params = metric_params
        self.n_jobs = n_jobs
        self.precision = precision

    def fit(self, X, y):
        self._fit_neighbors(X, y, target="classification")
        y = self._y_checked_
        del self._y_checked_
        self.classes_, encoded = np.unique(y, return_inverse=True)
        self._y = encoded.astype(np.intp, copy=False)
        self.outputs_2d_ = False
        self._codes_device = None
        return self

    def _codes(self):
        if getattr(self, "_codes_device", None) is None:
            import torch
            self._codes_device = torch.as_tensor(self._y.astype(np.int32)).cuda()
        return self._codes_device

    def __getstate__(self):
        state = dict(self.__dict__)
        state["_codes_device"] = None
        return state

    def _votes(self, X, probabilities):
        import ctypes
        import torch
        if self.weights not in (None, "uniform", "distance"):
            raise ValueError("weights not recognized: should be 'uniform' or 'distance', "
                             f"got {self.weights!r}")
        check_is_fitted(self, "_index_")
        Q, k, query_is_train = self._prepare_neighbor_query(X)
        need_dist = self.weights == "distance"
        index = self._index_
        dist, idx = index.device_index.query(Q, k, return_distance=need_dist, exclude_self=query_is_train,
                                             precision=index.precision)
        m = int(idx.shape[0])
        n_classes = len(self.classes_)
        pred = torch.empty(m, dtype=torch.int64, device=idx.device)
        proba = torch.empty((m, n_classes), dtype=torch.float64, device=idx.device) if probabilities else None
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else ctypes.c_void_p(0)  # noqa
        lib = _lib.load()
        _lib.check(lib.mlk_knn_votes(ptr(idx), ptr(dist), m, k, ptr(self._codes()), n_classes,
                                     _lib.WEIGHTS_DISTANCE if need_dist else _lib.WEIGHTS_UNIFORM, ptr(pred),
                                     ptr(proba), stream), "mlk_knn_votes")
        self.last_query_backend_ = "gpu"
        if probabilities:
            out = proba.cpu().numpy()
            if not np.isfinite(out).all():
                raise ValueError("All neighbors of some sample are getting zero weights.")
            return out
        out = pred.cpu().numpy()
        if np.any(out < 0):
            raise ValueError("All neighbors of some sample are getting zero weights.")
        return self.classes_.take(out)

    def predict(self, X):
        """Labels from ``classes_`` by (weighted) majority vote."""
        return self._votes(X, probabilities=False)

    def predict_proba(self, X):
        """Vote shares per class, float64 rows summing to one."""
        return self._votes(X, probabilities=True)


class KNeighborsRegressor(NeighborsMixin, RegressorMixin, BaseEstimator):
    """k-NN regression: (weighted) mean of the exact neighbours' targets."""

    def __init__(self, n_neighbors=5, *, weights="uniform", algorithm="auto", leaf_size=30, p=2,
                 metric="minkowski", metric_params=None, n_jobs=None, precision="float32"):
        self.n_neighbors = n_neighbors
        self.weights = weights
        self.algorithm = algorithm
        self.leaf_size = leaf_size
        self.p = p
        self.metric = metric
        self.metric_params = metric_params
        self.n_jobs = n_jobs
        self.precision = precision

    def fit(self, X, y):
        self._fit_neighbors(X, y, target="regression")
        self._y = np.asarray(self._y_checked_, dtype=np.float64)
        del self._y_checked_
        return self

    def predict(self, X):
        need = self.weights == "distance"
        out = self.kneighbors(X, return_distance=need)
        dist, idx = out if need else (None, out)
        w = compute_weights(dist, self.weights)
        vals = self._y[idx]
        if w is None:
            return vals.mean(axis=1)
        return (vals * w).sum(axis=1) / w.sum(axis=1)
