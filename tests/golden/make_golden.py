"""Write golden vectors from the reference's own native CPU engines.

Inputs are seeded synthetic arrays; expected outputs come from the reference's
exact engines compiled from their sources into ``oracle/_ref`` by
``oracle/build_oracle.py``:

* ``_neighbors_cpu_native.query`` (src/mlxlearn/_kernels/_neighbors_cpu_native.cpp:301-415):
  exact k-NN for d <= 32, n <= 16384, ties to the lower index.
* ``_radius_cpu_native.brute_query`` + ``result_fill_csr``
  (src/mlxlearn/_kernels/_radius_cpu_native.cpp:219-310, 692-800): exact radius
  CSR, ascending indices or (distance, index) order, distances sqrt(min(s, R2)).

Run from the repository root: ``python tests/golden/make_golden.py``.
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle", "_ref"))


def thresholds_tuple(d, R2):
    """(R2, f64_in, f64_out, prune_above, bulk_below) as radius_contract.thresholds builds them."""
    u = 2.0 ** -53
    g = (d + 2) * u / (1.0 - (d + 2) * u)
    dn = lambda v: math.nextafter(v, -math.inf)  # noqa: E731
    up = lambda v: math.nextafter(v, math.inf)  # noqa: E731
    if R2 == 0.0:
        return (0.0, 0.0, 0.0, 0.0, 0.0)
    return (R2, max(dn(dn(R2 * dn(1.0 - 2.0 * g))), 0.0), up(up(R2 * up(1.0 + 2.0 * g))),
            up(up(R2 * up(1.0 + 4.0 * g))), max(dn(dn(R2 * dn(1.0 - 4.0 * g))), 0.0))


def knn_cases():
    rng = np.random.default_rng(20260726)
    cases = {}
    X = rng.normal(size=(3000, 24)).astype(np.float32)
    X[100:120] = X[7]
    Q = rng.normal(size=(48, 24)).astype(np.float32)
    Q[0] = X[7]
    cases["gauss24"] = (X, Q, 9, False)
    Xi = rng.integers(0, 4, size=(2048, 8)).astype(np.float32)
    Qi = rng.integers(0, 4, size=(64, 8)).astype(np.float32)
    cases["ties8"] = (Xi, Qi, 12, False)
    Xo = (rng.normal(scale=0.05, size=(1500, 4)) + 1e4).astype(np.float32)
    cases["offset4_self"] = (Xo, Xo, 5, True)
    Xr = rng.normal(size=(900, 3)).astype(np.float32) * np.float32(2.0) ** rng.integers(-30, 30, (900, 1))
    Xr = Xr.astype(np.float32)
    Qr = Xr[:30] * np.float32(1.5)
    cases["scales3"] = (Xr, Qr.astype(np.float32), 4, False)
    return cases


def radius_cases():
    rng = np.random.default_rng(4242)
    X = rng.random((4000, 6), dtype=np.float32)
    X[10:20] = X[3]
    Q = rng.random((50, 6), dtype=np.float32)
    Q[0] = X[3]
    cases = {"uniform6": (X, Q, 0.2 ** 2, False)}
    Xi = rng.integers(0, 5, size=(1000, 4)).astype(np.float32)
    cases["int4_self"] = (Xi, Xi, 2.0, True)  # many distances exactly on the radius
    return cases


def main():
    import _neighbors_cpu_native as rn
    import _radius_cpu_native as rr
    out = {}
    for name, (X, Q, k, excl) in knn_cases().items():
        dist = np.empty((len(Q), k))
        idx = np.empty((len(Q), k), np.int64)
        rn.query(X, Q, k, excl, 0, dist, idx)
        out[f"knn_{name}_X"], out[f"knn_{name}_Q"] = X, Q
        out[f"knn_{name}_k"], out[f"knn_{name}_exclude"] = np.int64(k), np.bool_(excl)
        out[f"knn_{name}_idx"], out[f"knn_{name}_dist"] = idx, dist
    for name, (X, Q, R2, excl) in radius_cases().items():
        for sort in (False, True):
            mode = rr.NEED_DISTANCE | (rr.SORT if sort else 0)
            res = rr.brute_query(X, Q, thresholds_tuple(X.shape[1], R2), 0, excl, mode, 4)
            nq, nnz = rr.result_sizes(res)
            indptr = np.empty(nq + 1, np.int64)
            indices = np.empty(nnz, np.int64)
            dist = np.empty(nnz, np.float64)
            rr.result_fill_csr(res, indptr, indices, dist)
            tag = f"rad_{name}_{'sorted' if sort else 'plain'}"
            out[f"{tag}_indptr"], out[f"{tag}_indices"], out[f"{tag}_dist"] = indptr, indices, dist
        out[f"rad_{name}_X"], out[f"rad_{name}_Q"] = X, Q
        out[f"rad_{name}_R2"], out[f"rad_{name}_exclude"] = np.float64(R2), np.bool_(excl)
    path = os.path.join(HERE, "reference_native.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
