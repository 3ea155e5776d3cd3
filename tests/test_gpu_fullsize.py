"""Parity at the full workload sizes, on the GPU, through the C ABI.

* k-NN, 1,000,000 x 128 mixture with 1% duplicate rows, 10,000 queries, k = 10,
  every screen precision: all 10,000 rows identical across precisions, 256
  queries spread over the batch identical to the oracle.
* Radius, 2,000,000 x 32 cube with clumps, 256 log-uniform batches in one call,
  rows sorted by (distance, index): 64 rows spread over the batches identical
  to the oracle, row pointers consistent with the batches.
* k-means, 10,000,000 x 96 mixture, 65,536 centroids sampled from the rows:
  4,096 rows spread over the data identical to the oracle (labels and
  residuals), inertia within 1e-9 relative of math.fsum of the residuals.

Each step prints what it runs, so a failing run names its case.
"""
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mlk():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mlxlearn_amd
    from mlxlearn_amd import _lib
    _lib.load()
    return mlxlearn_amd


def test_knn_full_size_all_precisions(mlk):
    import torch
    from mlxlearn_amd import _lib
    X, Q = oracle.gen_c2(seed=3)
    index = mlk.DeviceNeighborIndex(torch.from_numpy(X).cuda())
    Qd = torch.from_numpy(Q).cuda()
    got = {}
    for precision in ("float32", "bfloat16", "float16"):
        print(f"full-size k-NN 1M x 128, 10,000 queries, {precision}", flush=True)
        d, i = index.query(Qd, 10, precision=precision)
        st = _lib.stats()
        want_prec = _lib.PRECISIONS[precision]
        assert st["screen_precision"] == want_prec, st
        got[precision] = (d.cpu().numpy(), i.cpu().numpy())
    index.close()
    ref_d, ref_i = got["float32"]
    for precision in ("bfloat16", "float16"):
        np.testing.assert_array_equal(got[precision][1], ref_i)
        np.testing.assert_array_equal(got[precision][0], ref_d)
    picks = np.linspace(0, len(Q) - 1, 256).astype(np.int64)
    wd, wi = oracle.knn(X, Q[picks], 10)
    np.testing.assert_array_equal(ref_i[picks], wi)
    np.testing.assert_allclose(ref_d[picks], wd, rtol=1e-12, atol=0)


def test_radius_full_size_batches_in_one_call(mlk):
    import torch
    X, Q, sizes, radius = oracle.gen_c4(seed=4)
    assert len(sizes) == 256
    R2 = radius * radius
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    index = mlk.DeviceNeighborIndex(torch.from_numpy(X).cuda())
    for precision in ("float32", "bfloat16"):
        print(f"full-size radius 2M x 32, {len(Q)} queries in 256 batches, sorted, {precision}", flush=True)
        indptr, indices, dist = index.radius(torch.from_numpy(Q).cuda(), R2, sort_results=True,
                                             precision=precision)
        indptr = indptr.cpu().numpy()
        ind, dst = indices.cpu().numpy(), dist.cpu().numpy()
        assert indptr[0] == 0 and np.all(np.diff(indptr) >= 0) and indptr[-1] == len(ind)
        picks = [int(offsets[b] + (b * 7919) % sizes[b]) for b in range(0, 256, 4)]
        for q in picks:
            wp, wi, wd = oracle.radius(X, Q[q:q + 1], R2, sort_results=True)
            np.testing.assert_array_equal(ind[indptr[q]:indptr[q + 1]], wi, err_msg=f"query {q}")
            np.testing.assert_allclose(dst[indptr[q]:indptr[q + 1]], wd, rtol=1e-12, atol=0, err_msg=f"query {q}")
        # every row sorted by (distance, index)
        rows = np.repeat(np.arange(len(Q)), np.diff(indptr))
        order = np.lexsort((ind, dst, rows))
        assert np.array_equal(order, np.arange(len(ind)))
    index.close()


def test_kmeans_full_size(mlk):
    import torch
    from mlxlearn_amd.cluster._k_means import assign_device, update_device
    X, C = oracle.gen_c5(seed=5)
    assert X.shape == (10_000_000, 96) and C.shape == (65_536, 96)
    Xd, Cd = torch.from_numpy(X).cuda(), torch.from_numpy(C).cuda()
    print("full-size k-means 10M x 96, 65,536 centroids, bfloat16 screen", flush=True)
    labels, resid = assign_device(Xd, Cd, precision="bfloat16")
    sums, counts, inertia = update_device(Xd, labels, C.shape[0], resid)
    labels, resid = labels.cpu().numpy(), resid.cpu().numpy()
    picks = np.linspace(0, len(X) - 1, 4096).astype(np.int64)
    wl, wr = oracle.kmeans_assign(X[picks], C)
    np.testing.assert_array_equal(labels[picks], wl)
    np.testing.assert_array_equal(resid[picks], wr)
    assert math.isclose(float(inertia.item()), math.fsum(resid.tolist()), rel_tol=1e-9)
    np.testing.assert_array_equal(counts.cpu().numpy(), np.bincount(labels, minlength=C.shape[0]))
    del Xd, Cd, sums
    torch.cuda.empty_cache()
