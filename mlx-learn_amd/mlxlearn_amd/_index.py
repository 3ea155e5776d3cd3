"""Device index and the NumPy boundary of the exact neighbour search.

``DeviceNeighborIndex`` is the resident index (torch tensors in, torch tensors
out), the counterpart of the reference's NumPy-free device API
``mlxlearn_device.DeviceNeighborIndex`` / ``fit_neighbors``
(``src/mlxlearn_device/_neighbors.py:243-466``). ``NeighborIndex`` is the
compatibility boundary used by the estimators, the counterpart of
``src/mlxlearn/_kernels/distance.py:66-392``: it owns a float32 host copy,
builds the device index lazily, and converts results to float64 distances and
int64 indices.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .exceptions import DeviceError

__all__ = ["DeviceNeighborIndex", "NeighborIndex", "RadiusCSR", "build_index", "gpu_available"]


def _torch():
    import torch
    return torch


def gpu_available() -> bool:
    try:
        return bool(_torch().cuda.is_available())
    except Exception:  # pragma: no cover
        return False


def require_gpu():
    torch = _torch()
    if not torch.cuda.is_available():
        raise DeviceError("no HIP device is available; the exact search runs on an MI355X GPU only")
    return _lib.load()


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else ctypes.c_void_p(0)


def to_device_f32(A):
    """A C-contiguous float32 CUDA tensor holding A (copying host arrays)."""
    torch = _torch()
    if isinstance(A, torch.Tensor):
        t = A.detach()
        if t.dtype != torch.float32:
            t = t.to(torch.float32)
        if not t.is_cuda:
            t = t.cuda()
        return t.contiguous()
    arr = np.ascontiguousarray(A, dtype=np.float32)
    return torch.from_numpy(arr).cuda()


def precision_flag(precision: str) -> int:
    try:
        return _lib.PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(_lib.PRECISIONS)}, got {precision!r}") from None


class DeviceNeighborIndex:
    """Exact brute-force index resident in HBM (owned float32 copy of X)."""

    def __init__(self, X):
        lib = require_gpu()
        torch = _torch()
        Xd = to_device_f32(X)
        if Xd.ndim != 2 or 0 in Xd.shape:
            raise ValueError("X must be a nonempty two-dimensional matrix")
        if Xd.shape[0] >= 2**31:
            raise ValueError("training row count must fit int32 indices")
        handle = ctypes.c_void_p()
        _lib.check(lib.mlk_index_create(_ptr(Xd), Xd.shape[0], Xd.shape[1], 0, _stream(), ctypes.byref(handle)),
                   "mlk_index_create")
        self._handle = handle
        self._lib = lib
        self.n_samples, self.n_features = int(Xd.shape[0]), int(Xd.shape[1])
        self._device = Xd.device
        torch.cuda.current_stream().synchronize()

    def close(self):
        if getattr(self, "_handle", None):
            self._lib.mlk_index_destroy(self._handle)
            self._handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def _check_width(self, Q):
        if hasattr(Q, "ndim") and (Q.ndim != 2 or Q.shape[1] != self.n_features):
            raise ValueError(f"queries must be a two-dimensional matrix with {self.n_features} features, "
                             f"got shape {tuple(Q.shape)}")

    def _queries(self, Q):
        """Q as a float32 device matrix of n_features columns; the C ABI cannot check the width."""
        self._check_width(Q)
        Qd = to_device_f32(Q)
        if Qd.ndim != 2 or Qd.shape[1] != self.n_features:
            raise ValueError(f"queries must be a two-dimensional matrix with {self.n_features} features, "
                             f"got shape {tuple(Qd.shape)}")
        return Qd

    def query(self, Q, k, *, return_distance=True, exclude_self=False, q_offset=0, precision="float32",
              squared=False, graph=False, out=None):
        """(float64 distances | None, int64 indices) as CUDA tensors, ordered by (S, index).

        ``graph=True`` captures the call's device chain in a HIP graph and
        replays it on later calls with the same query tensor, outputs and
        shape (pass ``out=(dist, idx)`` to keep the outputs fixed); results
        are identical either way.
        """
        torch = _torch()
        Qd = self._queries(Q)
        m, k = int(Qd.shape[0]), int(k)
        if out is not None:
            dist, idx = out
            if idx.shape != (m, k) or idx.dtype != torch.int64 or not idx.is_cuda or not idx.is_contiguous():
                raise ValueError(f"out indices must be a contiguous int64 CUDA tensor of shape {(m, k)}")
            if return_distance and (dist is None or dist.shape != (m, k) or dist.dtype != torch.float64
                                    or not dist.is_cuda or not dist.is_contiguous()):
                raise ValueError(f"out distances must be a contiguous float64 CUDA tensor of shape {(m, k)}")
            dist = dist if return_distance else None
        else:
            idx = torch.empty((m, k), dtype=torch.int64, device=self._device)
            dist = torch.empty((m, k), dtype=torch.float64, device=self._device) if return_distance else None
        flags = precision_flag(precision) | (_lib.EXCLUDE_SELF if exclude_self else 0)
        flags |= (_lib.SQUARED_DIST if squared else 0) | (_lib.GRAPH_REPLAY if graph else 0)
        _lib.check(self._lib.mlk_knn_query(self._handle, _ptr(Qd), m, k, int(q_offset), flags, _ptr(dist), _ptr(idx),
                                           _stream()), "mlk_knn_query")
        return dist, idx

    def query_batched(self, batches, k, *, return_distance=True, precision="float32", graph=False):
        """Several query sets answered by one screen launch per round; list of results."""
        torch = _torch()
        for b in batches:
            self._check_width(b)
        parts = [self._queries(b) for b in batches]
        sizes = [int(p.shape[0]) for p in parts]
        offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        Qd = torch.cat(parts) if parts else torch.empty((0, self.n_features), device=self._device)
        m = int(offsets[-1])
        idx = torch.empty((m, k), dtype=torch.int64, device=self._device)
        dist = torch.empty((m, k), dtype=torch.float64, device=self._device) if return_distance else None
        _lib.check(self._lib.mlk_knn_query_batched(
            self._handle, _ptr(Qd), offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(parts), int(k),
            precision_flag(precision) | (_lib.GRAPH_REPLAY if graph else 0), _ptr(dist), _ptr(idx), _stream()),
            "mlk_knn_query_batched")
        out = []
        for b in range(len(parts)):
            lo, hi = offsets[b], offsets[b + 1]
            out.append((None if dist is None else dist[lo:hi], idx[lo:hi]))
        return out

    def radius(self, Q, R2, *, exclude_self=False, q_offset=0, sort_results=False, need_distance=True,
               precision="float32"):
        """Exact radius CSR as CUDA tensors: (indptr int64, indices int64, distances float64 | None)."""
        torch = _torch()
        Qd = self._queries(Q)
        m = int(Qd.shape[0])
        indptr = torch.empty(m + 1, dtype=torch.int64, device=self._device)
        flags = precision_flag(precision) | (_lib.EXCLUDE_SELF if exclude_self else 0)
        flags |= (_lib.SORT_RESULTS if sort_results else 0) | (_lib.NEED_DISTANCE if need_distance else 0)
        nnz = ctypes.c_int64()
        res = ctypes.c_void_p()
        _lib.check(self._lib.mlk_radius_query(self._handle, _ptr(Qd), m, float(R2), int(q_offset), flags, _ptr(indptr),
                                              ctypes.byref(nnz), ctypes.byref(res), _stream()), "mlk_radius_query")
        try:
            indices = torch.empty(nnz.value, dtype=torch.int64, device=self._device)
            dist = (torch.empty(nnz.value, dtype=torch.float64, device=self._device)
                    if (need_distance or sort_results) else None)
            _lib.check(self._lib.mlk_radius_fill(res, _ptr(indices), _ptr(dist), _stream()), "mlk_radius_fill")
        finally:
            self._lib.mlk_radius_free(res)
        return indptr, indices, (dist if need_distance else None)

    def radius_batched(self, batches, R2, *, sort_results=False, need_distance=True, precision="float32"):
        """Radius neighbourhoods of several query sets in one call (their rows
        concatenated): a list of (indptr, indices, distances | None) per batch,
        each indptr starting at 0."""
        torch = _torch()
        for b in batches:
            self._check_width(b)
        parts = [self._queries(b) for b in batches]
        sizes = [int(p.shape[0]) for p in parts]
        Qd = torch.cat(parts) if parts else torch.empty((0, self.n_features), device=self._device)
        indptr, indices, dist = self.radius(Qd, R2, sort_results=sort_results, need_distance=need_distance,
                                            precision=precision)
        # every batch's first and last CSR position in one device-to-host read
        bounds = torch.tensor([0] + sizes, dtype=torch.int64).cumsum(0)
        ends = indptr[bounds.to(indptr.device)].cpu().tolist()
        out = []
        for i, sz in enumerate(sizes):
            lo, a, b = int(bounds[i]), ends[i], ends[i + 1]
            out.append((indptr[lo:lo + sz + 1] - a, indices[a:b], None if dist is None else dist[a:b]))
        return out

    def radius_reduce(self, Q, R2, *, codes=None, n_classes=None, targets=None, weighted=False,
                      exclude_self=False, q_offset=0, precision="float32"):
        """Radius neighbourhoods reduced on the device without materialising the CSR.

        With ``codes`` (int32 class code per training row, device): float64 votes
        (m x n_classes); with ``targets`` (float64 per training row, device): the
        (weighted) mean per query, NaN for an empty row. Returns (result, counts).
        """
        torch = _torch()
        if (codes is None) == (targets is None):
            raise ValueError("give exactly one of codes and targets")
        Qd = self._queries(Q)
        m = int(Qd.shape[0])
        indptr = torch.empty(m + 1, dtype=torch.int64, device=self._device)
        flags = precision_flag(precision) | (_lib.EXCLUDE_SELF if exclude_self else 0)
        flags |= _lib.NEED_DISTANCE if weighted else 0
        weights = _lib.WEIGHTS_DISTANCE if weighted else _lib.WEIGHTS_UNIFORM
        nnz = ctypes.c_int64()
        res = ctypes.c_void_p()
        _lib.check(self._lib.mlk_radius_query(self._handle, _ptr(Qd), m, float(R2), int(q_offset), flags, _ptr(indptr),
                                              ctypes.byref(nnz), ctypes.byref(res), _stream()), "mlk_radius_query")
        counts = torch.empty(m, dtype=torch.int64, device=self._device)
        try:
            if codes is not None:
                out = torch.empty((m, int(n_classes)), dtype=torch.float64, device=self._device)
                _lib.check(self._lib.mlk_radius_votes(res, _ptr(indptr), _ptr(codes), int(n_classes), weights,
                                                      _ptr(out), _ptr(counts), _stream()), "mlk_radius_votes")
            else:
                out = torch.empty(m, dtype=torch.float64, device=self._device)
                _lib.check(self._lib.mlk_radius_mean(res, _ptr(indptr), _ptr(targets), weights, _ptr(out),
                                                     _ptr(counts), _stream()), "mlk_radius_mean")
        finally:
            self._lib.mlk_radius_free(res)
        return out, counts

    def autotune(self, Q, k=None, *, radius_sq=None, precision="float32", sort_results=False, max_trials=24):
        """Time screen plans on these queries and cache the fastest for calls of this shape.

        k-NN with ``k``; radius search with ``radius_sq`` (R2) instead. Returns the
        chosen plan with its time and the untuned plan's time (milliseconds).
        """
        Qd = self._queries(Q)
        if (k is None) == (radius_sq is None):
            raise ValueError("give exactly one of k and radius_sq")
        flags = precision_flag(precision)
        if radius_sq is not None:
            flags |= _lib.TUNE_RADIUS | _lib.NEED_DISTANCE | (_lib.SORT_RESULTS if sort_results else 0)
        plan = _lib.Plan()
        _lib.check(self._lib.mlk_autotune(self._handle, _ptr(Qd), int(Qd.shape[0]), int(k or 0),
                                          float(radius_sq or 0.0), flags, int(max_trials), ctypes.byref(plan),
                                          _stream()), "mlk_autotune")
        return plan.to_dict()

    def plan(self, m, k=None, *, radius=False, precision="float32"):
        """The plan a call with m queries (k neighbours, or a radius call) would use."""
        flags = precision_flag(precision) | (_lib.TUNE_RADIUS if radius else 0)
        plan = _lib.Plan()
        _lib.check(self._lib.mlk_get_plan(self._handle, int(m), int(k or 0), flags, ctypes.byref(plan)),
                   "mlk_get_plan")
        return plan.to_dict()


@dataclass
class RadiusCSR:
    """CSR neighbourhoods (reference ``radius_result.RadiusCSR``, radius_result.py:34-55)."""

    indptr: np.ndarray
    indices: np.ndarray
    distances: np.ndarray | None

    @property
    def n_query(self) -> int:
        return len(self.indptr) - 1

    @property
    def nnz(self) -> int:
        return int(self.indptr[-1])

    def to_ragged(self, return_distance=True):
        """sklearn's object arrays of 1-d int64 / float64 rows (radius_result.py:109-127)."""
        rows = np.empty(self.n_query, dtype=object)
        drows = np.empty(self.n_query, dtype=object)
        for i in range(self.n_query):
            a, b = self.indptr[i], self.indptr[i + 1]
            rows[i] = self.indices[a:b]
            if self.distances is not None:
                drows[i] = self.distances[a:b]
        return (drows, rows) if return_distance else rows


def validate_radius(radius) -> float:
    """Restates radius_contract.validate_radius (radius_contract.py:78-98)."""
    if isinstance(radius, np.ndarray):
        if radius.ndim != 0:
            raise TypeError(f"radius must be a scalar, got an array of shape {radius.shape}")
        radius = radius[()]
    from numbers import Real
    if isinstance(radius, (bool, np.bool_)) or not isinstance(radius, (Real, np.floating, np.integer)):
        raise TypeError(f"radius must be a real number, got {type(radius).__name__}")
    value = float(radius)
    if math.isnan(value):
        raise ValueError("radius must not be NaN")
    if value < 0:
        raise ValueError(f"radius must be non-negative, got {value}")
    return value + 0.0


class NeighborIndex:
    """Owned float32 training copy + lazily built device index (distance.py:66-364)."""

    def __init__(self, X, *, precision="float32"):
        X = np.array(X, dtype=np.float32, order="C", copy=True)
        if X.ndim != 2 or 0 in X.shape:
            raise ValueError("X must be a nonempty two-dimensional matrix")
        if X.shape[0] >= 2**31:
            raise ValueError("training row count must fit int32 indices")
        self.X = X
        self.precision = precision
        self._device_index = None
        self.last_query_backend = None

    @property
    def n_samples(self):
        return int(self.X.shape[0])

    @property
    def n_features(self):
        return int(self.X.shape[1])

    @property
    def device_index(self) -> DeviceNeighborIndex:
        if self._device_index is None:
            self._device_index = DeviceNeighborIndex(self.X)
        return self._device_index

    def __getstate__(self):
        return {"X": self.X, "precision": self.precision}

    def __setstate__(self, state):
        self.X = state["X"]
        self.precision = state.get("precision", "float32")
        self._device_index = None
        self.last_query_backend = None

    def query(self, Q, k, *, return_distance=True, exclude_self=False, precision=None):
        """float64 distances and int64 indices at the NumPy boundary (distance.py:242-270)."""
        Q = np.ascontiguousarray(Q, dtype=np.float32)
        k = int(k)
        limit = self.n_samples - int(bool(exclude_self))
        if k > limit:
            raise ValueError(f"n_neighbors must be <= {limit}, got {k}")
        if exclude_self and Q.shape[0] != self.n_samples:
            raise ValueError("exclude_self requires the query matrix to be the training matrix: "
                             f"got {Q.shape[0]} query rows and {self.n_samples} training rows.")
        if Q.shape[0] == 0:
            return (np.empty((0, k), np.float64) if return_distance else None), np.empty((0, k), np.int64)
        dist, idx = self.device_index.query(Q, k, return_distance=return_distance, exclude_self=exclude_self,
                                            precision=precision or self.precision)
        self.last_query_backend = "gpu"
        return (None if dist is None else dist.cpu().numpy()), idx.cpu().numpy()

    def query_radius_csr(self, Q, radius, *, need_distance=True, exclude_self=False, sort_results=False,
                         precision=None) -> RadiusCSR:
        """Exact membership S <= fl64(radius**2) (distance.py:285-336)."""
        radius = validate_radius(radius)
        Q = np.ascontiguousarray(Q, dtype=np.float32)
        if exclude_self and Q.shape[0] != self.n_samples:
            raise ValueError("exclude_self requires the query matrix to be the training matrix: "
                             f"got {Q.shape[0]} query rows and {self.n_samples} training rows.")
        R2 = radius * radius
        if Q.shape[0] == 0:
            return RadiusCSR(np.zeros(1, np.int64), np.empty(0, np.int64),
                             np.empty(0, np.float64) if need_distance else None)
        indptr, indices, dist = self.device_index.radius(
            Q, R2, exclude_self=exclude_self, sort_results=sort_results, need_distance=need_distance,
            precision=precision or self.precision)
        self.last_query_backend = "gpu"
        return RadiusCSR(indptr.cpu().numpy(), indices.cpu().numpy(), None if dist is None else dist.cpu().numpy())


def build_index(X, *, precision="float32") -> NeighborIndex:
    """Own the training storage; the device index is built on the first query (distance.py:366-392)."""
    return NeighborIndex(X, precision=precision)
