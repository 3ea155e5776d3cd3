"""NearestNeighbors, KNeighbors{Classifier,Regressor} and RadiusNeighbors{Classifier,Regressor}.

Constructor signatures and defaults follow the reference
(``src/mlxlearn/neighbors/_unsupervised.py:84-93``,
``_classification.py:71-80``) plus one extension, ``precision``: the screen
precision of the GPU search (``"float32"`` default, the drop-in;
``"bfloat16"`` / ``"float16"`` run a reduced-precision MFMA screen whose
certified candidates are re-ranked in float64, with identical results).
"""
from __future__ import annotations

import warnings

import numpy as np
from sklearn.base import BaseEstimator, ClassifierMixin, RegressorMixin
from sklearn.utils.validation import _is_arraylike_not_scalar, check_is_fitted

from .. import _lib
from .._diagnostics import record_event
from .._index import require_gpu, validate_radius
from ._base import NeighborsMixin, RadiusMixin, compute_weights

__all__ = ["NearestNeighbors", "KNeighborsClassifier", "KNeighborsRegressor", "RadiusNeighborsClassifier",
           "RadiusNeighborsRegressor"]


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
        return self

    def _codes(self):
        import torch
        return self._device_cache("codes", lambda: torch.as_tensor(self._y.astype(np.int32)).cuda())

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
        self._set_query_backend("gpu")
        record_event(type(self).__name__, "predict_proba" if probabilities else "predict", shape=Q.shape, k=k)
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


def _check_weights(weights):
    if weights not in (None, "uniform", "distance"):
        raise ValueError(f"weights not recognized: should be 'uniform' or 'distance', got {weights!r}")
    return weights == "distance"


def _radius_query(est, X):
    """(query rows, exclude_self, R2) of a radius estimator's prediction."""
    check_is_fitted(est, "_index_")
    query_is_train = X is None
    Q = est._index_.X if query_is_train else est._validate_query(X, "predict")
    r = validate_radius(est.radius)
    require_gpu()
    return Q, query_is_train, r * r


class RadiusNeighborsClassifier(NeighborsMixin, RadiusMixin, ClassifierMixin, BaseEstimator):
    """Every training row within ``radius`` votes (reference neighbors/_radius.py:16-216).

    Votes are reduced on the GPU from the exact radius neighbourhoods, in
    ascending training index per query; a tied vote goes to the smallest class
    code. Queries with no neighbour take ``outlier_label`` (``None`` raises).
    """

    n_neighbors = None

    def __init__(self, radius=1.0, *, weights="uniform", algorithm="auto", leaf_size=30, p=2,
                 metric="minkowski", outlier_label=None, metric_params=None, n_jobs=None, precision="float32"):
        self.radius = radius
        self.weights = weights
        self.algorithm = algorithm
        self.leaf_size = leaf_size
        self.p = p
        self.metric = metric
        self.outlier_label = outlier_label
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
        self.outlier_label_ = _resolve_outlier_label(self.outlier_label, self.classes_, self._y)
        return self

    _codes = KNeighborsClassifier._codes

    def _proba(self, X):
        weighted = _check_weights(self.weights)
        Q, query_is_train, R2 = _radius_query(self, X)
        if Q.shape[0] == 0:
            return np.empty((0, len(self.classes_)))
        votes, counts = self._index_.device_index.radius_reduce(
            Q, R2, codes=self._codes(), n_classes=len(self.classes_), weighted=weighted,
            exclude_self=query_is_train, precision=self._index_.precision)
        self._set_query_backend("gpu")
        record_event(type(self).__name__, "predict", shape=Q.shape, weights=self.weights)
        votes = _fill_outliers(votes.cpu().numpy(), counts.cpu().numpy(), self.outlier_label_, self.classes_)
        normalizer = votes.sum(axis=1)[:, None]
        normalizer[normalizer == 0.0] = 1.0
        return votes / normalizer

    def predict(self, X):
        """Labels from ``classes_``; an empty neighbourhood takes the outlier label."""
        proba = self._proba(X)
        labels = self.classes_.take(proba.argmax(axis=1)) if proba.size else self.classes_[:0]
        empty = np.all(proba == 0.0, axis=1)
        if np.any(empty):
            labels = np.array(labels, copy=True)
            labels[empty] = self.outlier_label_[0]
        return labels

    def predict_proba(self, X):
        """Vote shares per class (float64 rows summing to one; outliers one-hot or zero)."""
        return self._proba(X)


class RadiusNeighborsRegressor(NeighborsMixin, RadiusMixin, RegressorMixin, BaseEstimator):
    """(Weighted) mean target of the training rows within ``radius``
    (reference neighbors/_radius.py:219-338); NaN, with a warning, for a query
    without neighbours."""

    n_neighbors = None

    def __init__(self, radius=1.0, *, weights="uniform", algorithm="auto", leaf_size=30, p=2,
                 metric="minkowski", metric_params=None, n_jobs=None, precision="float32"):
        self.radius = radius
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
        checked = self._y_checked_
        del self._y_checked_
        self._y = np.array(checked, dtype=np.float64, copy=True)
        self._y_out_dtype_ = np.float32 if np.asarray(checked).dtype == np.float32 else np.float64
        return self

    def predict(self, X):
        """float64 (float32 for float32 targets) means; NaN where no neighbour lies inside ``radius``."""
        import torch
        weighted = _check_weights(self.weights)
        Q, query_is_train, R2 = _radius_query(self, X)
        if Q.shape[0] == 0:
            return np.empty(0, dtype=self._y_out_dtype_)
        targets = self._device_cache("targets", lambda: torch.as_tensor(self._y).cuda())
        pred, counts = self._index_.device_index.radius_reduce(
            Q, R2, targets=targets, weighted=weighted, exclude_self=query_is_train,
            precision=self._index_.precision)
        self._set_query_backend("gpu")
        record_event(type(self).__name__, "predict", shape=Q.shape, weights=self.weights)
        pred = pred.cpu().numpy()
        if np.any(counts.cpu().numpy() == 0):
            warnings.warn("One or more samples have no neighbors within specified radius; predicting NaN.",
                          stacklevel=2)
        return pred.astype(self._y_out_dtype_, copy=False)


def _resolve_outlier_label(outlier_label, classes, encoded_y):
    """``outlier_label_`` of a single-output classifier: None or a one-element list.

    scikit-learn's documented contract for ``outlier_label``: ``None`` makes an
    empty neighbourhood an error at prediction; ``"most_frequent"`` takes the
    training class with the largest count (argmax of the class counts, so the
    smallest class on a tie); anything else is a manual label, a scalar or a
    one-element list of scalars (one entry per output), whose type must not
    widen the classes' dtype. Bad values raise scikit-learn's errors.
    """
    if outlier_label is None:
        return None
    if isinstance(outlier_label, str) and outlier_label == "most_frequent":
        counts = np.bincount(encoded_y, minlength=len(classes))
        return [classes[int(np.argmax(counts))]]
    given = list(outlier_label) if _is_arraylike_not_scalar(outlier_label) else [outlier_label]
    if len(given) != 1:
        raise ValueError(f"The length of outlier_label: {outlier_label} is inconsistent with the output "
                         "length: 1")
    label = given[0]
    if _is_arraylike_not_scalar(label):
        raise TypeError(f"The outlier_label of classes {classes} is supposed to be a scalar, got {label}.")
    if np.promote_types(classes.dtype, np.asarray(label).dtype) != classes.dtype:
        raise TypeError(f"The dtype of outlier_label {label} is inconsistent with classes {classes} in y.")
    return [label]


def _fill_outliers(votes, counts, outlier_label_, classes):
    """Votes of queries without neighbours, in place (scikit-learn's contract).

    No outlier label: an error naming the empty queries. A label among the
    classes: a one-hot vote for it. A label outside the classes: all-zero
    votes (every class probability 0) and a warning.
    """
    empty = np.flatnonzero(counts == 0)
    if empty.size == 0:
        return votes
    if outlier_label_ is None:
        raise ValueError(f"No neighbors found for test samples {empty!r}, you can try using larger radius, "
                         "giving a label for outliers, or considering removing them from your dataset.")
    label = outlier_label_[0]
    column = int(np.searchsorted(classes, label)) if classes.size else 0
    if column < classes.size and classes[column] == label:
        votes[empty, column] = 1.0
    else:
        warnings.warn(f"Outlier label {label} is not in training classes. All class probabilities of "
                      "outliers will be assigned with 0.", stacklevel=4)
    return votes
