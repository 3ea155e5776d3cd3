"""Reference-side binding of ``libmlxlearn_amd.so`` for tabulai/mlx-learn.

A maintainer of the reference copies this file to
``src/mlxlearn_device/_mi355x.py`` and constructs its ``DeviceNeighborIndex``
where ``NeighborIndex._ensure_device`` (``src/mlxlearn/_kernels/distance.py:157-181``)
builds ``mlxlearn_device.DeviceNeighborIndex`` today. Nothing else in the
reference changes: ``NeighborIndex.query_device`` (``distance.py:211-240``)
calls ``query`` with the same keywords, and ``NeighborIndex.query``
(``distance.py:242-270``) takes ``np.sqrt(np.asarray(distances, float64))`` and
``np.asarray(indices, int64)`` of what it returns.

The module depends on ctypes, NumPy and PyTorch-ROCm (device buffers and the
current HIP stream) only; it does not import the build's own host package.
The library is found through ``MLXLEARN_AMD_LIB`` (a path), else by the
dynamic loader's search for ``libmlxlearn_amd.so``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_p, _i64, _i32, _u32, _dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32, ctypes.c_double

# flags of include/mlxlearn_amd.h
EXCLUDE_SELF = 0x1
PRECISION = {"float32": 0x00, "bfloat16": 0x10, "float16": 0x20}
SORT_RESULTS = 0x100
NEED_DISTANCE = 0x200
SQUARED_DIST = 0x400

_lib = None


def library():
    """The loaded C library with every signature this binding calls."""
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(os.environ.get("MLXLEARN_AMD_LIB", "libmlxlearn_amd.so"))
        lib.mlk_last_error.restype = ctypes.c_char_p
        lib.mlk_index_create.argtypes = [_p, _i64, _i32, _u32, _p, ctypes.POINTER(_p)]
        lib.mlk_index_destroy.argtypes = [_p]
        lib.mlk_knn_query.argtypes = [_p, _p, _i64, _i32, _i64, _u32, _p, _p, _p]
        lib.mlk_radius_query.argtypes = [_p, _p, _i64, _dbl, _i64, _u32, _p, ctypes.POINTER(_i64),
                                         ctypes.POINTER(_p), _p]
        lib.mlk_radius_fill.argtypes = [_p, _p, _p, _p]
        lib.mlk_radius_free.argtypes = [_p]
        lib.mlk_kmeans_assign.argtypes = [_p, _i64, _i32, _p, _i64, _u32, _p, _p, _p]
        _lib = lib
    return _lib


def _check(rc):
    """Status convention of the reference's native modules (_neighbors_cpu_native.cpp:28-65):
    input errors raise ValueError, allocation failures MemoryError, device failures
    the reference's BackendExecutionError (RuntimeError outside the reference)."""
    if rc == 0:
        return
    msg = library().mlk_last_error().decode()
    if rc > 0:
        raise ValueError(msg)
    if rc == -1:
        raise MemoryError(msg)
    try:
        from mlxlearn.exceptions import BackendExecutionError
    except ImportError:  # outside the reference tree
        BackendExecutionError = RuntimeError
    raise BackendExecutionError(msg)


def _torch():
    import torch
    return torch


def _device_f32(A):
    """A C-contiguous float32 device tensor of A (NumPy arrays, MLX arrays via
    NumPy, or tensors)."""
    torch = _torch()
    if isinstance(A, torch.Tensor):
        return A.detach().to(device="cuda", dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(A), dtype=np.float32)).cuda()


def _ptr(t):
    return _p(t.data_ptr()) if t is not None and t.numel() else _p(0)


def _stream():
    return _p(_torch().cuda.current_stream().cuda_stream)


class DeviceNeighborIndex:
    """Drop-in for ``mlxlearn_device.DeviceNeighborIndex`` (``_neighbors.py:243-443``).

    ``device`` is accepted for the reference's call (``distance.py:173-176``);
    the index lives on the current HIP device whatever it says, since this
    build has no CPU route. ``precision`` selects the screen ("float32" is the
    drop-in; results are identical at every precision).
    """

    def __init__(self, X, *, device=None, precision="float32"):
        del device
        if precision not in PRECISION:
            raise ValueError(f"precision must be one of {sorted(PRECISION)}")
        Xd = _device_f32(X)
        if Xd.ndim != 2 or 0 in Xd.shape:
            raise ValueError("X must be a nonempty two-dimensional matrix")
        if Xd.shape[0] >= 2 ** 31:
            raise ValueError("training row count must fit int32 indices")
        self.n_samples, self.n_features = int(Xd.shape[0]), int(Xd.shape[1])
        self.precision = precision
        self._h = _p()
        _check(library().mlk_index_create(_ptr(Xd), Xd.shape[0], Xd.shape[1], 0, _stream(), ctypes.byref(self._h)))

    def query(self, Q, k, *, return_distance=True, exclude_self=False, query_block_size=None,
              train_block_size=None, _validated=False, _squared=False, _auto_train_tiling=True, _gram_tile=None):
        """(distances | None, indices) as NumPy arrays: float64 squared distances
        when ``_squared`` (what ``NeighborIndex.query`` square-roots), else
        float64 Euclidean distances; int64 indices ordered by (exact distance,
        index). The block and tiling keywords of the reference planner are
        accepted and ignored: the library plans its own tiles, and results do
        not depend on them."""
        del query_block_size, train_block_size, _validated, _auto_train_tiling, _gram_tile
        torch = _torch()
        Qd = _device_f32(Q)
        if Qd.ndim != 2 or Qd.shape[1] != self.n_features:
            raise ValueError(f"queries must be a two-dimensional matrix with {self.n_features} features")
        m, k = int(Qd.shape[0]), int(k)
        if not 1 <= k <= self.n_samples - int(bool(exclude_self)):
            raise ValueError(f"n_neighbors must be in [1, {self.n_samples - int(bool(exclude_self))}], got {k}")
        if exclude_self and m != self.n_samples:
            raise ValueError("exclude_self requires the query matrix to be the training matrix")
        idx = torch.empty((m, k), dtype=torch.int64, device="cuda")
        dist = torch.empty((m, k), dtype=torch.float64, device="cuda") if return_distance else None
        flags = PRECISION[self.precision] | (EXCLUDE_SELF if exclude_self else 0) | (SQUARED_DIST if _squared else 0)
        if m:
            _check(library().mlk_knn_query(self._h, _ptr(Qd), m, k, 0, flags, _ptr(dist), _ptr(idx), _stream()))
        torch.cuda.current_stream().synchronize()
        return (None if dist is None else dist.cpu().numpy()), idx.cpu().numpy()

    def radius_csr(self, Q, radius, *, need_distance=True, exclude_self=False, sort_results=False):
        """One CSR batch of ``query_radius_csr`` (``distance.py:285-336``):
        (indptr int64[m+1], indices int64[nnz], distances float64[nnz] | None),
        membership exact S <= fl64(radius * radius)."""
        torch = _torch()
        Qd = _device_f32(Q)
        if Qd.ndim != 2 or Qd.shape[1] != self.n_features:
            raise ValueError(f"queries must be a two-dimensional matrix with {self.n_features} features")
        m = int(Qd.shape[0])
        indptr = torch.empty(m + 1, dtype=torch.int64, device="cuda")
        flags = PRECISION[self.precision] | (EXCLUDE_SELF if exclude_self else 0)
        flags |= (SORT_RESULTS if sort_results else 0) | (NEED_DISTANCE if need_distance else 0)
        nnz, res = _i64(), _p()
        r = float(radius)
        _check(library().mlk_radius_query(self._h, _ptr(Qd), m, r * r, 0, flags, _ptr(indptr), ctypes.byref(nnz),
                                          ctypes.byref(res), _stream()))
        try:
            ind = torch.empty(nnz.value, dtype=torch.int64, device="cuda")
            dst = (torch.empty(nnz.value, dtype=torch.float64, device="cuda")
                   if (need_distance or sort_results) else None)
            _check(library().mlk_radius_fill(res, _ptr(ind), _ptr(dst), _stream()))
        finally:
            library().mlk_radius_free(res)
        torch.cuda.current_stream().synchronize()
        return (indptr.cpu().numpy(), ind.cpu().numpy(),
                None if (dst is None or not need_distance) else dst.cpu().numpy())

    def close(self):
        if getattr(self, "_h", None):
            library().mlk_index_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fit_neighbors(X, device=None):
    """Drop-in for ``mlxlearn_device.fit_neighbors`` (``_neighbors.py:447-466``)."""
    return DeviceNeighborIndex(X, device=device)


def assign(x, centers, *, batch_size=None, spherical=False, method="auto", precision="float32"):
    """Drop-in for ``e_means.assign`` (``e_means.py:245-345``): (uint32 labels,
    float32 squared residuals), as the reference returns them. Labels are the
    exact nearest centre (lowest index on ties); residuals are the canonical
    float64 values rounded to float32. ``batch_size`` and ``method`` are
    accepted and ignored (the library tiles itself); spherical assignment is
    not supported."""
    del batch_size, method
    if spherical:
        raise ValueError("spherical assignment is not supported by the MI355X library")
    if precision not in PRECISION:
        raise ValueError("precision must be float32, float16, or bfloat16")
    torch = _torch()
    X, C = _device_f32(x), _device_f32(centers)
    if C.ndim != 2 or C.shape[0] < 1 or X.ndim != 2 or C.shape[1] != X.shape[1]:
        raise ValueError("centers must be nonempty and have the same feature dimension")
    n = int(X.shape[0])
    if n == 0:
        return np.zeros(0, np.uint32), np.zeros(0, np.float32)
    labels = torch.empty(n, dtype=torch.int64, device="cuda")
    resid = torch.empty(n, dtype=torch.float64, device="cuda")
    _check(library().mlk_kmeans_assign(_ptr(X), n, X.shape[1], _ptr(C), C.shape[0], PRECISION[precision],
                                       _ptr(labels), _ptr(resid), _stream()))
    torch.cuda.current_stream().synchronize()
    return labels.cpu().numpy().astype(np.uint32), resid.cpu().numpy().astype(np.float32)
