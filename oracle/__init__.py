"""Exact neighbour oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or, for
``cpu_baseline``, as the timed CPU restatement). The product package never
imports it.

Contents
--------
* ``knn`` / ``radius`` / ``kmeans_assign`` -- ctypes bindings of the C oracle
  (``exact_oracle.c``): exact order by S = sum (q - x)**2 of the float32
  coordinates, lower index on exact ties.
* ``integer_key`` / ``integer_knn`` -- pure-Python big-integer restatement of
  ``src/mlxlearn/_kernels/neighbors_cpu.py:39-47, 174-181`` for tiny cases.
* ``radius_band`` -- the outward-rounded float64 band of
  ``radius_contract.thresholds`` (``radius_contract.py:129-153``), via the C
  oracle.
* ``KMeansPort`` -- float32 Gram argmin + float64 check (the k-means CPU
  baseline, ``e_means.py:50-69``).
* ``reference_radius`` -- the reference's own radius engine compiled into
  ``oracle/_ref`` (the radius CPU baseline), when it is built.
* ``GramPort`` -- numpy restatement of the reference CPU path
  ``CPUIndex._gram`` / ``_direct`` / ``_exact_order``
  (``neighbors_cpu.py:174-181, 213-262, 314-364``), used as the timed CPU
  baseline.
* ``gen_*`` -- seeded synthetic generators for the five benchmark workloads.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


def default_threads() -> int:
    try:
        return max(1, min(len(os.sched_getaffinity(0)), 16))
    except AttributeError:  # pragma: no cover
        return max(1, min(os.cpu_count() or 1, 16))


def lib():
    """Load (building if needed) the C oracle."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(HERE, "_oracle.so")
    src = os.path.join(HERE, "exact_oracle.c")
    if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(src):
        from .build_oracle import build_oracle
        build_oracle()
    L = ctypes.CDLL(path)
    L.oracle_knn.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, _f32p, ctypes.c_int64, ctypes.c_int,
                             ctypes.c_int, ctypes.c_int64, _i64p, ctypes.c_void_p, ctypes.c_int]
    L.oracle_knn.restype = ctypes.c_int
    L.oracle_exact_compare.argtypes = [_f32p, _f32p, _f32p, ctypes.c_int]
    L.oracle_exact_compare.restype = ctypes.c_int
    L.oracle_exact_key.argtypes = [_f32p, _f32p, ctypes.c_int, _u32p]
    L.oracle_exact_key.restype = None
    L.oracle_canonical_sq.argtypes = [_f32p, _f32p, ctypes.c_int]
    L.oracle_canonical_sq.restype = ctypes.c_double
    L.oracle_radius_band.argtypes = [ctypes.c_int, ctypes.c_double,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.oracle_radius_band.restype = None
    L.oracle_radius.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, _f32p, ctypes.c_int64, ctypes.c_double,
                                ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    L.oracle_radius.restype = ctypes.c_int
    L.oracle_exact_le.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_double]
    L.oracle_exact_le.restype = ctypes.c_int
    L.oracle_kmeans_assign.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, _f32p, ctypes.c_int64,
                                       _i64p, ctypes.c_void_p, ctypes.c_int]
    L.oracle_kmeans_assign.restype = ctypes.c_int
    _LIB = L
    return L


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def knn(X, Q, k, *, exclude_self=False, q_offset=0, threads=None):
    """Exact k-NN: (float64 distances, int64 indices) ordered by (S, index)."""
    X, Q = _c32(X), _c32(Q)
    m, d = Q.shape
    idx = np.empty((m, k), np.int64)
    dist = np.empty((m, k), np.float64)
    if m:
        rc = lib().oracle_knn(X, X.shape[0], d, Q, m, int(k), int(bool(exclude_self)), int(q_offset),
                              idx, _ptr(dist), threads or default_threads())
        if rc:
            raise RuntimeError(f"oracle_knn failed with status {rc}")
    return dist, idx


def exact_compare(q, a, b):
    """sign(S(q, a) - S(q, b)) exactly."""
    q, a, b = _c32(q), _c32(a), _c32(b)
    return int(lib().oracle_exact_compare(q, a, b, q.shape[-1]))


def exact_key_c(q, x):
    """S(q, x) * 2**298 as a Python int, via the C oracle."""
    q, x = _c32(q), _c32(x)
    out = np.zeros(20, np.uint32)
    lib().oracle_exact_key(q, x, q.shape[-1], out)
    return sum(int(w) << (32 * i) for i, w in enumerate(out))


def canonical_sq(q, x):
    q, x = _c32(q), _c32(x)
    return float(lib().oracle_canonical_sq(q, x, q.shape[-1]))


def radius_band(d, R2):
    lo, hi = ctypes.c_double(), ctypes.c_double()
    lib().oracle_radius_band(int(d), float(R2), ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def radius(X, Q, R2, *, exclude_self=False, q_offset=0, threads=None, need_distance=True,
           sort_results=False):
    """Exact radius CSR: (indptr int64, indices int64, distances float64|None)."""
    X, Q = _c32(X), _c32(Q)
    m, d = Q.shape
    threads = threads or default_threads()
    counts = np.zeros(m, np.int64)
    L = lib()
    if m:
        L.oracle_radius(X, X.shape[0], d, Q, m, float(R2), int(bool(exclude_self)), int(q_offset), 1,
                        _ptr(counts), None, None, None, threads)
    indptr = np.zeros(m + 1, np.int64)
    np.cumsum(counts, out=indptr[1:])
    indices = np.empty(int(indptr[-1]), np.int64)
    distances = np.empty(int(indptr[-1]), np.float64)
    if m and indices.size:
        L.oracle_radius(X, X.shape[0], d, Q, m, float(R2), int(bool(exclude_self)), int(q_offset), 2,
                        None, _ptr(indptr), _ptr(indices), _ptr(distances), threads)
    if sort_results:
        for r in range(m):
            a, b = indptr[r], indptr[r + 1]
            order = np.lexsort((indices[a:b], distances[a:b]))
            indices[a:b] = indices[a:b][order]
            distances[a:b] = distances[a:b][order]
    return indptr, indices, (distances if need_distance or sort_results else None)


def exact_le(q, x, R2):
    q, x = _c32(q), _c32(x)
    return bool(lib().oracle_exact_le(q, x, q.shape[-1], float(R2)))


def kmeans_assign(X, C, *, threads=None):
    """Exact nearest centre (lowest index on ties) and canonical float64 residuals."""
    X, C = _c32(X), _c32(C)
    labels = np.empty(X.shape[0], np.int64)
    residual = np.empty(X.shape[0], np.float64)
    rc = lib().oracle_kmeans_assign(X, X.shape[0], X.shape[1], C, C.shape[0], labels, _ptr(residual),
                                    threads or default_threads())
    if rc:
        raise RuntimeError(f"oracle_kmeans_assign failed with status {rc}")
    return labels, residual


def kmeans_update(X, labels, n_clusters):
    """Float64 centroid sums and counts (exact order-independent reference via math.fsum)."""
    X = _c32(X)
    d = X.shape[1]
    sums = np.zeros((n_clusters, d), np.float64)
    counts = np.bincount(labels, minlength=n_clusters).astype(np.int64)
    order = np.argsort(labels, kind="stable")
    bounds = np.concatenate([[0], np.cumsum(counts)])
    for c in range(n_clusters):
        rows = X[order[bounds[c]:bounds[c + 1]]].astype(np.float64)
        if len(rows):
            sums[c] = [math.fsum(col) for col in rows.T]
    return sums, counts


# ------------------------------------------------------------------ pure Python


def integer_row(row):
    """float32 values as integers in units of 2**-149 (neighbors_cpu.py:39-47)."""
    out = []
    for raw in np.asarray(row, np.float32).view(np.uint32).tolist():
        exponent = (raw >> 23) & 255
        mantissa = (raw & 0x7FFFFF) | (0x800000 if exponent else 0)
        value = mantissa << max(0, exponent - 1)
        out.append(-value if raw >> 31 else value)
    return out


def integer_key(q, x):
    """S(q, x) * 2**298 as a Python int."""
    return sum((a - b) ** 2 for a, b in zip(integer_row(q), integer_row(x), strict=True))


def integer_knn(X, Q, k, *, exclude_self=False, q_offset=0):
    """Brute-force exact order with Python integers (neighbors_cpu.py:174-181); tiny inputs only."""
    X, Q = _c32(X), _c32(Q)
    rows = [integer_row(x) for x in X]
    out = np.empty((len(Q), k), np.int64)
    for r, q in enumerate(Q):
        qi = integer_row(q)
        keys = []
        for i, xi in enumerate(rows):
            if exclude_self and i == q_offset + r:
                continue
            keys.append((sum((a - b) ** 2 for a, b in zip(qi, xi)), i))
        keys.sort()
        out[r] = [i for _, i in keys[:k]]
    return out


# ------------------------------------------------------- reference CPU path port


def _stable_minima(work, take):
    """The `take` smallest entries per row, lower column first on equal values,
    by repeated row argmin (restates ``_extract``, neighbors_cpu.py:49-59)."""
    rows = np.arange(len(work))
    ids = np.empty((len(work), take), np.intp)
    values = np.empty((len(work), take), work.dtype)
    for slot in range(take):
        best = np.argmin(work, axis=1)
        ids[:, slot] = best
        values[:, slot] = work[rows, best]
        work[rows, best] = np.inf
    return values, ids


def _partition_minima(work, take):
    """Partition-based prefix, stable minima for rows whose prefix ends at zero
    (restates ``_extract_partition``, neighbors_cpu.py:62-77)."""
    rows = np.arange(len(work))[:, None]
    ids = np.argpartition(work, take - 1, axis=1)[:, :take]
    values = work[rows, ids]
    order = np.lexsort((ids, values), axis=1)
    values, ids = values[rows, order], ids[rows, order]
    zero = np.flatnonzero(values[:, -1] == 0)
    if len(zero):
        values[zero], ids[zero] = _stable_minima(work[zero].copy(), take)
    work[rows, ids] = np.inf
    return values, ids


class GramPort:
    """Numpy restatement of the reference CPU path ``CPUIndex`` (neighbors_cpu.py).

    ``query`` follows ``CPUIndex.query`` (:402-448) for the methods the benchmark
    shapes reach, in query blocks of 256 rows:

    * ``_gram`` (:314-364): float64 BLAS Gram lower bound
      ``(1 - 64(D+1)eps)(|q|^2 + |x|^2) - 2 q.x``, the k+2 smallest as candidates
      plus the next as barrier, float64 direct refinement, interval
      ``16(D+1)eps``; ambiguous certified rows are ordered exactly over their
      k+2 candidates only (``_exact_order``); uncertain rows (barrier not above
      the k-th upper value) rerun through ``_direct`` alone.
    * ``_direct`` (:213-262): float64 squared distances to every row, the k+1
      smallest, exact repair of overlapping intervals over the retained prefix
      plus every row whose lower value reaches the k-th upper one.
    * ``_integer_gram`` (:264-287): float32 Gram arithmetic for small integer
      data (every term exact), used when the reference would use it.
    * ``_exact_order`` (:174-181): exact order by Python-integer keys
      (coordinates in units of 2^-149).

    Not restated: the compiled D <= 32 scan (``_neighbors_cpu_native.query``,
    reached only for D <= 32 and n <= 16384), the k-d tree proposal of
    ``_tree_candidates`` (n <= 2048), explicit train tiling (``_tiled``) and the
    batched int64 dyadic order (``_native_exact_order``, a faster route to the
    same exact order as ``_exact_order``). Timed as the CPU baseline.
    """

    QUERY_BLOCK = 256

    def __init__(self, X):
        self.X = _c32(X)
        self.train64 = self.X.astype(np.float64)
        self.norm64 = np.einsum("ij,ij->i", self.train64, self.train64)
        self._norm32 = None
        self._integral = None

    # -------------------------------------------------------------- routing
    def _integer_query(self, Q):
        """Small integer data whose float32 Gram arithmetic is exact (:148-172)."""
        n, features = self.X.shape
        if n > 2048 or features > 64:
            return False
        if self._integral is None:
            top = float(np.max(np.abs(self.X)))
            self._top = top
            self._integral = bool(4 * features * top * top <= 2 ** 24
                                  and np.equal(self.X.view(np.uint32), np.trunc(self.X).view(np.uint32)).all())
        if not self._integral:
            return False
        top = max(self._top, float(np.max(np.abs(Q))))
        return bool(4 * features * top * top <= 2 ** 24
                    and np.equal(Q.view(np.uint32), np.trunc(Q).view(np.uint32)).all())

    def query(self, Q, k, block=None, *, exclude_self=False):
        """(float64 distances, int64 indices) ordered by (S, index)."""
        Q = _c32(Q)
        block = block or self.QUERY_BLOCK
        exact = self._integer_query(Q) if len(Q) else False
        out_d = np.empty((len(Q), k))
        out_i = np.empty((len(Q), k), np.int64)
        for s in range(0, len(Q), block):
            part = Q[s:s + block]
            if exact and len(Q) >= 16 and self.X.shape[1] >= 32 and len(self.X) >= 512:
                d, i = self._integer_gram(part, k, exclude_self, s)
            elif len(Q) >= 16 and len(self.X) >= 256:
                d, i = self._gram(part, k, exclude_self, s)
            else:
                d, i = self._direct(part, k, exclude_self, s, exact=exact)
            out_d[s:s + block] = np.sqrt(d)
            out_i[s:s + block] = i
        return out_d, out_i

    # -------------------------------------------------------------- pieces
    def _exact_order(self, query, candidates, k):
        keys = sorted((integer_key(query, self.X[c]), int(c)) for c in candidates)
        return np.asarray([c for _, c in keys[:k]], np.intp)

    def _direct(self, Q, k, exclude_self=False, offset=0, exact=False):
        from scipy.spatial.distance import cdist
        train, query = self.train64, Q.astype(np.float64)
        work = cdist(query, train, "sqeuclidean")
        if exclude_self:
            work[np.arange(len(Q)), np.arange(offset, offset + len(Q))] = np.inf
        take = min(k + 1, len(train) - int(exclude_self))
        values, ids = (_partition_minima(work, take) if not exact and len(train) <= 256
                       else _stable_minima(work, take))
        if not exact:
            error = 16 * (self.X.shape[1] + 1) * np.finfo(np.float64).eps
            lower, upper = values * (1 - error), values * (1 + error)
            ambiguous = np.any(lower[:, 1:] <= upper[:, :-1], axis=1) & (values[:, k - 1] != 0)
            for row in np.flatnonzero(ambiguous):
                if take > k and lower[row, k] <= upper[row, k - 1]:
                    cutoff = np.nextafter(upper[row, k - 1] / (1 - error), np.inf)
                    cands = np.concatenate((ids[row], np.flatnonzero(work[row] <= cutoff)))
                else:
                    cands = ids[row, :k]
                chosen = self._exact_order(Q[row], cands, k)
                ids[row, :k] = chosen
                values[row, :k] = np.sum((train[chosen] - query[row]) ** 2, axis=1)
        return values[:, :k], ids[:, :k]

    def _integer_gram(self, Q, k, exclude_self=False, offset=0):
        if self._norm32 is None:
            self._norm32 = np.einsum("ij,ij->i", self.X, self.X)
        work = Q @ self.X.T
        work *= -2
        work += self._norm32[None, :]
        work += np.einsum("ij,ij->i", Q, Q)[:, None]
        if exclude_self:
            work[np.arange(len(Q)), np.arange(offset, offset + len(Q))] = np.inf
        values, ids = _stable_minima(work, k)
        return values.astype(np.float64), ids

    def _gram(self, Q, k, exclude_self=False, offset=0):
        train, query = self.train64, Q.astype(np.float64)
        D = self.X.shape[1]
        lower = query @ train.T
        lower *= -2
        factor = 64 * (D + 1) * np.finfo(np.float64).eps
        lower += ((1 - factor) * self.norm64)[None, :]
        lower += ((1 - factor) * np.einsum("ij,ij->i", query, query))[:, None]
        np.maximum(lower, 0, out=lower)
        if exclude_self:
            lower[np.arange(len(Q)), np.arange(offset, offset + len(Q))] = np.inf
        take = min(k + 2, len(train) - int(exclude_self))
        if take < len(train):
            part = np.argpartition(lower, take, axis=1)[:, :take + 1]
            cand = part[:, :take]
            barrier = lower[np.arange(len(Q)), part[:, take]]
        else:
            cand = np.broadcast_to(np.arange(len(train)), lower.shape)
            barrier = np.full(len(Q), np.inf)
        delta = train[cand] - query[:, None, :]
        values = np.sum(delta * delta, axis=2)
        order = np.lexsort((cand, values), axis=1)
        rows = np.arange(len(Q))[:, None]
        values, cand = values[rows, order], cand[rows, order]
        error = 16 * (D + 1) * np.finfo(np.float64).eps
        upper = values[:, k - 1] * (1 + error)
        uncertain = ~(barrier > upper)
        if take == len(train) - int(exclude_self):
            uncertain[:] = False
        ambiguous = np.any(values[:, 1:] * (1 - error) <= values[:, :-1] * (1 + error), axis=1)
        for row in np.flatnonzero(ambiguous & ~uncertain):
            chosen = self._exact_order(Q[row], cand[row], k)
            cand[row, :k] = chosen
            values[row, :k] = np.sum((train[chosen] - query[row]) ** 2, axis=1)
        for row in np.flatnonzero(uncertain):
            sq, ids = self._direct(Q[row:row + 1], k, exclude_self, offset + row)
            cand[row, :k], values[row, :k] = ids[0], sq[0]
        return values[:, :k], cand[:, :k]


class KMeansPort:
    """The reference's k-means assignment on the CPU: float32 Gram argmin plus a
    float64 check (the timed CPU baseline of the 10M x 96 workload).

    Restates ``e_means.assign``'s scoring (``e_means.py:50-69``): per row the
    float32 score ``|c|^2 - 2 x.c`` over every centre (one BLAS sgemm per row
    block, all host cores), the smallest score with the lowest index on ties.
    The float64 check recomputes the two best centres' squared distances in
    float64; a row whose two best are within the float32 scores' error is
    resolved by float64 distances to every centre (``direct``). The reference
    itself stops at the float32 argmin; the check is the certification the
    exact contract needs.
    """

    def __init__(self, C):
        self.C = _c32(C)
        self.C64 = self.C.astype(np.float64)
        self.norms = np.einsum("ij,ij->i", self.C64, self.C64).astype(np.float32)

    def assign(self, X, block=2048):
        X = _c32(X)
        labels = np.empty(len(X), np.int64)
        resid = np.empty(len(X), np.float64)
        d = X.shape[1]
        for s in range(0, len(X), block):
            xb = X[s:s + block]
            scores = self.norms[None, :] - np.float32(2.0) * (xb @ self.C.T)
            two = np.argpartition(scores, 1, axis=1)[:, :2]
            first = np.argmin(scores, axis=1)                      # lowest index among equal minima
            rows = np.arange(len(xb))
            other = np.where(two[:, 0] == first, two[:, 1], two[:, 0])
            x64 = xb.astype(np.float64)
            d_first = np.sum((x64 - self.C64[first]) ** 2, axis=1)
            d_other = np.sum((x64 - self.C64[other]) ** 2, axis=1)
            # float32 score error: about (d + 2) 2^-24 (|x|^2 + |c|^2) per score
            slack = (d + 2) * 2.0 ** -23 * (np.sum(x64 * x64, axis=1) + self.norms[first] + self.norms[other])
            close = np.flatnonzero(np.abs(d_first - d_other) <= slack)
            for r in close:
                full = np.sum((x64[r] - self.C64) ** 2, axis=1)
                first[r] = int(np.argmin(full))
                d_first[r] = full[first[r]]
            labels[s:s + block] = first
            resid[s:s + block] = d_first
            del rows
        return labels, resid


def reference_radius_engine():
    """The reference's own radius engine (``_radius_cpu_native``), compiled from
    its sources into ``oracle/_ref`` by ``build_oracle.py``; None when absent."""
    ref = os.path.join(HERE, "_ref")
    if not os.path.isdir(ref):
        return None
    import sys
    if ref not in sys.path:
        sys.path.insert(0, ref)
    try:
        import _radius_cpu_native
    except ImportError:
        return None
    return _radius_cpu_native


def reference_thresholds(d, R2):
    """The threshold tuple the reference engine takes (R2, f64_in, f64_out,
    prune_above, bulk_below), built as ``radius_contract.thresholds`` builds
    it (radius_contract.py:129-153); the band pair is ``radius_band``'s."""
    if R2 == 0.0:
        return (0.0, 0.0, 0.0, 0.0, 0.0)
    lo, hi = radius_band(d, R2)
    u = 2.0 ** -53
    g = (d + 2) * u / (1.0 - (d + 2) * u)
    dn = lambda v: math.nextafter(v, -math.inf)  # noqa: E731
    up = lambda v: math.nextafter(v, math.inf)  # noqa: E731
    return (float(R2), lo, hi, up(up(R2 * up(1.0 + 4.0 * g))), max(dn(dn(R2 * dn(1.0 - 4.0 * g))), 0.0))


def reference_radius(X, Q, R2, *, sort_results=False, threads=None):
    """CSR of the reference engine's ``brute_query`` + ``result_fill_csr``
    (``_radius_cpu_native.cpp:219-310, 692-800``)."""
    rr = reference_radius_engine()
    if rr is None:
        raise RuntimeError("the reference radius engine is not built (oracle/_ref)")
    X, Q = _c32(X), _c32(Q)
    mode = rr.NEED_DISTANCE | (rr.SORT if sort_results else 0)
    res = rr.brute_query(X, Q, reference_thresholds(X.shape[1], float(R2)), 0, False, mode,
                         threads or default_threads())
    nq, nnz = rr.result_sizes(res)
    indptr = np.empty(nq + 1, np.int64)
    indices = np.empty(nnz, np.int64)
    dist = np.empty(nnz, np.float64)
    rr.result_fill_csr(res, indptr, indices, dist)
    return indptr, indices, dist


# ------------------------------------------------------------------ generators


def gen_c1(seed=0, n=10_000, m=1_000, d=64):
    """Plain exact kNN workload: N(0,1) float32, k=5."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d), dtype=np.float32)
    Q = rng.standard_normal((m, d), dtype=np.float32)
    return X, Q


def gen_c2(seed=0, n=1_000_000, m=10_000, d=128, n_centres=256, dup_frac=0.01):
    """Mixture of isotropic Gaussians: centres N(0, 4I), sigma=1, 1% exact duplicate rows."""
    rng = np.random.default_rng(seed)
    centres = (2.0 * rng.standard_normal((n_centres, d))).astype(np.float32)
    X = np.empty((n, d), np.float32)
    step = 1 << 18
    for s in range(0, n, step):
        e = min(n, s + step)
        lab = rng.integers(0, n_centres, e - s)
        X[s:e] = centres[lab] + rng.standard_normal((e - s, d), dtype=np.float32)
    n_dup = int(round(dup_frac * n))
    if n_dup:
        dst = rng.choice(n, n_dup, replace=False)
        src = rng.integers(0, n, n_dup)
        X[dst] = X[src]
    lab = rng.integers(0, n_centres, m)
    Q = centres[lab] + rng.standard_normal((m, d), dtype=np.float32)
    return X, Q.astype(np.float32)


def _blob_images(rng, count, side=28, zero_frac=0.8):
    yy, xx = np.mgrid[0:side, 0:side].astype(np.float32)
    area = (1.0 - zero_frac) * side * side
    out = np.zeros((count, side * side), np.float32)
    step = 8192
    for s in range(0, count, step):
        e = min(count, s + step)
        c = e - s
        aspect = rng.uniform(0.5, 2.0, c).astype(np.float32)
        ry = np.sqrt(area / np.pi * aspect)
        rx = np.sqrt(area / np.pi / aspect)
        cy = rng.uniform(ry, side - ry).astype(np.float32)
        cx = rng.uniform(rx, side - rx).astype(np.float32)
        inside = (((yy[None] - cy[:, None, None]) / ry[:, None, None]) ** 2
                  + ((xx[None] - cx[:, None, None]) / rx[:, None, None]) ** 2) <= 1.0
        vals = rng.integers(1, 256, (c, side, side)).astype(np.float32)
        out[s:e] = np.where(inside, vals, 0.0).reshape(c, -1)
    return out


def gen_c3(seed=0, n=60_000, m=10_000, n_classes=10):
    """Pixel-like integer rows 0..255 (about 80% zeros, contiguous blobs), 10 labels."""
    rng = np.random.default_rng(seed)
    X = _blob_images(rng, n)
    y = rng.integers(0, n_classes, n).astype(np.int64)
    Q = _blob_images(rng, m)
    return X, y, Q


def gen_c4_points(rng, count, d, clump_centres, clump_frac=0.2, sigma=0.02):
    out = rng.random((count, d), dtype=np.float32)
    n_clump = int(round(clump_frac * count))
    if n_clump:
        pick = rng.choice(count, n_clump, replace=False)
        which = rng.integers(0, len(clump_centres), n_clump)
        out[pick] = (clump_centres[which]
                     + sigma * rng.standard_normal((n_clump, d), dtype=np.float32)).astype(np.float32)
    return out


def gen_c4(seed=0, n=2_000_000, d=32, n_batches=256, max_batch=4096, target_median=50):
    """Uniform [0,1]^32 + 64 Gaussian clumps (20%), log-uniform batch sizes, calibrated radius."""
    rng = np.random.default_rng(seed)
    clumps = rng.random((64, d), dtype=np.float32)
    X = gen_c4_points(rng, n, d, clumps)
    sizes = np.maximum(1, np.rint(np.exp(rng.uniform(0, math.log(max_batch), n_batches)))).astype(np.int64)
    Q = gen_c4_points(rng, int(sizes.sum()), d, clumps)
    radius = calibrate_radius(X, Q, target_median, rng)
    return X, Q, sizes, radius


def calibrate_radius(X, Q, target_median, rng, n_q=128, n_x=131_072):
    """Radius whose median neighbour count is about target_median (sampled, deterministic)."""
    qs = Q[rng.choice(len(Q), min(n_q, len(Q)), replace=False)].astype(np.float64)
    xs = X[rng.choice(len(X), min(n_x, len(X)), replace=False)].astype(np.float64)
    frac = len(xs) / len(X)
    kth = max(1, int(round(target_median * frac)))
    vals = []
    for q in qs:
        s = np.sum((xs - q) ** 2, axis=1)
        vals.append(np.partition(s, kth - 1)[kth - 1])
    r = math.sqrt(float(np.median(vals)))
    return float(f"{r:.4g}")


def gen_c5(seed=0, n=10_000_000, d=96, n_components=4096, n_clusters=65_536, sigma=0.3):
    """Gaussian mixture rows (centres N(0,I), sigma 0.3) and centroids sampled from rows."""
    rng = np.random.default_rng(seed)
    centres = rng.standard_normal((n_components, d), dtype=np.float32)
    X = np.empty((n, d), np.float32)
    step = 1 << 20
    for s in range(0, n, step):
        e = min(n, s + step)
        lab = rng.integers(0, n_components, e - s)
        X[s:e] = centres[lab] + np.float32(sigma) * rng.standard_normal((e - s, d), dtype=np.float32)
    pick = rng.choice(n, n_clusters, replace=False)
    C = X[pick].copy()
    return X, C
