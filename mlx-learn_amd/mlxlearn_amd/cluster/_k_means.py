"""Lloyd k-means with exact GPU assignment and float64 deterministic updates.

Surface of the reference's ``KMeans`` (``src/mlxlearn/cluster/_k_means.py:178-500``):
the same constructor parameters, ``fit`` / ``predict`` / ``transform`` /
``score`` / ``fit_predict``, tolerance ``tol * mean(feature variances)``,
stopping on unchanged labels or small centre shift, farthest-row relocation of
empty clusters and a fresh final assignment whose residuals give
``inertia_``. Differences by design: labels are the exact nearest centre
(lowest index on ties) rather than float32 argmin labels, and centroid
statistics are float64 sums in a fixed order instead of float32 atomics.
"""
from __future__ import annotations

import ctypes
import math
import warnings
from numbers import Integral, Real

import numpy as np
from sklearn.base import BaseEstimator, ClusterMixin, TransformerMixin
from sklearn.utils import check_random_state
from sklearn.utils.validation import check_is_fitted, validate_data

from .. import _lib
from .._config import get_config
from .._diagnostics import record_event
from .._index import precision_flag, require_gpu, to_device_f32
from ..neighbors._base import prepare_input, require_dense

__all__ = ["KMeans", "assign_device", "update_device"]


def _torch():
    import torch
    return torch


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else ctypes.c_void_p(0)


def assign_device(X, C, *, precision="float32", previous=None, out_reused=None):
    """Exact nearest centre of every row: (int64 labels, float64 squared residuals) CUDA tensors.

    With ``previous`` labels (a Lloyd iteration after the first), rows whose
    label a certificate proves unchanged keep it without a search
    (``reusable_labels``, reference lloyd_bounds.py:24-76); the result is the
    same. ``out_reused`` (a list) receives the number of such rows.
    """
    lib = require_gpu()
    torch = _torch()
    Xd, Cd = to_device_f32(X), to_device_f32(C)
    n, d = Xd.shape
    labels = torch.empty(n, dtype=torch.int64, device=Xd.device)
    residual = torch.empty(n, dtype=torch.float64, device=Xd.device)
    if previous is None:
        _lib.check(lib.mlk_kmeans_assign(_ptr(Xd), n, d, _ptr(Cd), Cd.shape[0], precision_flag(precision),
                                         _ptr(labels), _ptr(residual), _stream()), "mlk_kmeans_assign")
        return labels, residual
    prev = torch.as_tensor(previous, device=Xd.device).to(torch.int64).contiguous()
    if prev.shape != (n,):
        raise ValueError(f"previous labels must have shape ({n},)")
    reused = ctypes.c_int64()
    _lib.check(lib.mlk_kmeans_assign_reuse(_ptr(Xd), n, d, _ptr(Cd), Cd.shape[0], _ptr(prev), precision_flag(precision),
                                           _ptr(labels), _ptr(residual), ctypes.byref(reused), _stream()),
               "mlk_kmeans_assign_reuse")
    if out_reused is not None:
        out_reused.append(int(reused.value))
    return labels, residual


def update_device(X, labels, n_clusters, residual=None):
    """float64 per-cluster sums (k x d), int64 counts (k) and inertia (0-d float64) CUDA tensors."""
    lib = require_gpu()
    torch = _torch()
    Xd = to_device_f32(X)
    n, d = Xd.shape
    labels = labels.to(device=Xd.device, dtype=torch.int64).contiguous()
    sums = torch.empty((n_clusters, d), dtype=torch.float64, device=Xd.device)
    counts = torch.empty(n_clusters, dtype=torch.int64, device=Xd.device)
    inertia = torch.empty((), dtype=torch.float64, device=Xd.device) if residual is not None else None
    _lib.check(lib.mlk_kmeans_update(_ptr(Xd), n, d, _ptr(labels), n_clusters,
                                     _ptr(residual) if residual is not None else ctypes.c_void_p(0), _ptr(sums),
                                     _ptr(counts), _ptr(inertia) if inertia is not None else ctypes.c_void_p(0),
                                     _stream()), "mlk_kmeans_update")
    return sums, counts, inertia


def _positive_integer(value, name, *, optional=False):
    if optional and value is None:
        return
    if isinstance(value, bool) or not isinstance(value, Integral) or value < 1:
        raise ValueError(f"{name} must be a positive integer")


class KMeans(ClusterMixin, TransformerMixin, BaseEstimator):
    """Full-batch Lloyd k-means with exact nearest-centre labels on an MI355X GPU."""

    def __init__(self, n_clusters=8, *, init="k-means++", n_init="auto", max_iter=300, tol=1e-4, verbose=False,
                 random_state=None, copy_x=True, algorithm="lloyd", precision="float32"):
        self.n_clusters = n_clusters
        self.init = init
        self.n_init = n_init
        self.max_iter = max_iter
        self.tol = tol
        self.verbose = verbose
        self.random_state = random_state
        self.copy_x = copy_x
        self.algorithm = algorithm
        self.precision = precision

    def _validate_parameters(self):
        _positive_integer(self.n_clusters, "n_clusters")
        _positive_integer(self.max_iter, "max_iter")
        if self.n_init != "auto":
            _positive_integer(self.n_init, "n_init")
        if isinstance(self.tol, bool) or not isinstance(self.tol, Real) or not math.isfinite(self.tol) or self.tol < 0:
            raise ValueError("tol must be finite and nonnegative")
        if self.algorithm != "lloyd":
            raise ValueError("only algorithm='lloyd' is supported")
        if (isinstance(self.init, str) and self.init not in {"random", "k-means++"}) or callable(self.init):
            raise ValueError("init must be 'random', 'k-means++', or an initial center matrix")

    @staticmethod
    def _matrix(X, name="X"):
        arr = np.asarray(X)
        if arr.ndim != 2 or min(arr.shape) < 1:
            raise ValueError(f"{name} must be a nonempty two-dimensional dense matrix")
        if np.iscomplexobj(arr):
            raise ValueError(f"{name} must contain real values")
        with np.errstate(over="ignore", invalid="ignore"):
            arr = np.ascontiguousarray(arr, dtype=np.float32)
        if not np.all(np.isfinite(arr)):
            raise ValueError(f"{name} must contain finite values representable in float32")
        return arr

    def _checked(self, X, *, reset):
        """Dense, finite, nonempty float32 rows with scikit-learn's messages; n_features_in_ is set on fit
        and checked afterwards. Runs before any device call."""
        require_dense(X, type(self).__name__, "fit" if reset else "predict")
        X = prepare_input(X)
        return validate_data(self, X, accept_sparse=False, dtype=np.float32, order="C", reset=reset)

    def _initial_centres(self, X, rng):
        n = X.shape[0]
        k = int(self.n_clusters)
        if not isinstance(self.init, str):
            C = self._matrix(self.init, "init")
            if C.shape != (k, X.shape[1]):
                raise ValueError(f"init must have shape {(k, X.shape[1])}")
            return [C]
        restarts = self.n_init
        if restarts == "auto":
            restarts = 10 if self.init == "random" else 1
        out = []
        for _ in range(int(restarts)):
            if self.init == "random":
                out.append(X[rng.choice(n, size=k, replace=False)].copy())
            else:
                from sklearn.cluster import kmeans_plusplus
                centres, _ = kmeans_plusplus(X, k, random_state=rng)
                out.append(np.ascontiguousarray(centres, dtype=np.float32))
        return out

    def fit(self, X, y=None, sample_weight=None):
        if sample_weight is not None:
            raise ValueError("sample_weight is not supported")
        self._validate_parameters()
        Xh = self._checked(X, reset=True)
        require_gpu()
        torch = _torch()
        n, d = Xh.shape
        k = int(self.n_clusters)
        if k > n:
            raise ValueError(f"n_samples={n} must be >= n_clusters={k}")
        rng = check_random_state(self.random_state if self.random_state is not None else get_config().random_state)
        explicit = not isinstance(self.init, str)
        if explicit and self.n_init not in ("auto", 1):
            warnings.warn("Explicit initial center position passed: performing only one init instead of "
                          f"n_init={self.n_init}.", RuntimeWarning, stacklevel=2)
        initials = self._initial_centres(Xh, rng)
        Xd = to_device_f32(Xh)
        tolerance = 0.0
        if self.tol:
            var = Xd.to(torch.float64).var(dim=0, unbiased=False).mean().item()
            tolerance = var * float(self.tol)
        best = None
        reused = []  # rows per assignment whose label was certified without a search
        for restart, C0 in enumerate(initials):
            centres = to_device_f32(C0)
            previous = None
            reason = "max_iter"
            it = 0
            for it in range(int(self.max_iter)):
                labels, residual = assign_device(Xd, centres, precision=self.precision, previous=previous,
                                                 out_reused=reused)
                sums, counts, _ = update_device(Xd, labels, k, residual)
                sums, counts, empty = self._relocate_empty(Xd, labels, residual, sums, counts)
                new = (sums / counts.clamp(min=1).to(torch.float64)[:, None])
                if empty:
                    largest = int(torch.argmax(counts).item())
                    new[counts == 0] = new[largest].clone()  # a row of `new` itself: copy before the write
                new32 = new.to(torch.float32)
                shift = float(((new32.to(torch.float64) - centres.to(torch.float64)) ** 2).sum().item())
                unchanged = previous is not None and bool(torch.equal(labels, previous))
                if self.verbose:
                    print(f"KMeans restart {restart + 1}, iteration {it + 1}: "
                          f"inertia={float(residual.sum().item()):.8g}, shift={shift:.8g}", flush=True)
                centres = new32
                if unchanged:
                    reason = "labels"
                    break
                if shift <= tolerance:
                    reason = "tolerance"
                    break
                previous = labels
            labels, residual = assign_device(Xd, centres, precision=self.precision, previous=labels,
                                             out_reused=reused)
            _, _, inertia = update_device(Xd, labels, k, residual)
            value = float(inertia.item())
            if best is None or value < best[0]:
                best = (value, centres, labels, it + 1, reason)
        value, centres, labels, n_iter, reason = best
        self.cluster_centers_ = centres.cpu().numpy()
        self.labels_ = labels.cpu().numpy().astype(np.int64)
        self.inertia_ = value
        self.n_iter_ = n_iter
        self.convergence_reason_ = reason
        self.reused_labels_ = int(sum(reused))
        self._n_features_out = k
        record_event(type(self).__name__, "fit", shape=Xh.shape, n_clusters=k, n_iter=n_iter,
                     restarts=len(initials), reused_labels=self.reused_labels_, convergence=reason)
        distinct = np.unique(self.labels_).size
        if distinct < k:
            from sklearn.exceptions import ConvergenceWarning
            warnings.warn(f"Number of distinct clusters ({distinct}) found smaller than n_clusters ({k}). "
                          "Possibly due to duplicate points in X.", ConvergenceWarning, stacklevel=2)
        return self

    @staticmethod
    def _relocate_empty(Xd, labels, residual, sums, counts):
        """Move the farthest rows into empty clusters (reference _k_means.py:116-142)."""
        torch = _torch()
        empty_ids = torch.nonzero(counts == 0).flatten()
        n_empty = int(empty_ids.numel())
        if not n_empty or float(residual.max().item()) <= 0:
            return sums, counts, n_empty
        far = torch.argsort(-residual, stable=True)[:n_empty]
        sums = sums.clone()
        counts = counts.clone()
        for e, r in zip(empty_ids.tolist(), far.tolist()):
            donor = int(labels[r].item())
            row = Xd[r].to(torch.float64)
            sums[donor] -= row
            counts[donor] -= 1
            sums[e] = row
            counts[e] = 1
        return sums, counts, int((counts == 0).sum().item())

    def _inference(self, X):
        check_is_fitted(self, "cluster_centers_")
        Xh = self._checked(X, reset=False)
        require_gpu()
        return Xh

    def predict(self, X):
        """Exact nearest final centre of every row."""
        Xh = self._inference(X)
        labels, _ = assign_device(Xh, self.cluster_centers_, precision=self.precision)
        record_event(type(self).__name__, "predict", shape=Xh.shape)
        return labels.cpu().numpy().astype(np.int64)

    def transform(self, X):
        """Euclidean distances to every centre (float64 direct differences on the GPU, returned as float32)."""
        torch = _torch()
        Xh = self._inference(X)
        Xd = to_device_f32(Xh).to(torch.float64)
        Cd = to_device_f32(self.cluster_centers_).to(torch.float64)
        out = np.empty((Xh.shape[0], Cd.shape[0]), np.float32)
        rows = max(1, (64 << 20) // (8 * Cd.shape[0] * max(1, Xh.shape[1])))
        for s in range(0, Xh.shape[0], rows):
            diff = Xd[s:s + rows, None, :] - Cd[None, :, :]
            out[s:s + rows] = torch.sqrt((diff * diff).sum(-1)).to(torch.float32).cpu().numpy()
        record_event(type(self).__name__, "transform", shape=Xh.shape, search=False)
        return out

    def score(self, X, y=None, sample_weight=None):
        """Negative inertia of a fresh exact assignment."""
        if sample_weight is not None:
            raise ValueError("sample_weight is not supported")
        Xh = self._inference(X)
        labels, residual = assign_device(Xh, self.cluster_centers_, precision=self.precision)
        _, _, inertia = update_device(Xh, labels, self.cluster_centers_.shape[0], residual)
        record_event(type(self).__name__, "score", shape=Xh.shape)
        return -float(inertia.item())

    def fit_predict(self, X, y=None, sample_weight=None):
        return self.fit(X, sample_weight=sample_weight).labels_
