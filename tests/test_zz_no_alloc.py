"""mlk_knn_query_noalloc (temporaries in the library's fixed device array)
against mlk_knn_query: bit-identical indices and distances, on two calls in a
row (the second with fewer queries), at every screen precision, with chunked
queries, self queries and the exact fallback, plus its two refusals."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mlxlearn_amd import _lib
    return _lib.load()


def _call(lib, name, index, Q, k, flags=0, q_offset=0):
    import torch
    from mlxlearn_amd import _lib
    from mlxlearn_amd._index import _ptr, _stream
    m = int(Q.shape[0])
    idx = torch.full((m, k), -7, dtype=torch.int64, device="cuda")
    dist = torch.full((m, k), -7.0, dtype=torch.float64, device="cuda")
    print(f"{name}: m={m} k={k} flags={flags:#x} q_offset={q_offset}", flush=True)
    rc = getattr(lib, name)(index._handle, _ptr(Q), m, k, q_offset, flags, _ptr(dist), _ptr(idx), _stream())
    _lib.check(rc, name)
    torch.cuda.synchronize()
    return dist.cpu().numpy(), idx.cpu().numpy()


def _same(a, b):
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[0], b[0])


@pytest.mark.parametrize("precision", ["float32", "bfloat16", "float16"])
def test_noalloc_matches_query_twice(lib, precision):
    import torch
    import mlxlearn_amd
    from mlxlearn_amd import _lib
    r = np.random.default_rng(41)
    X = r.normal(size=(20_000, 64)).astype(np.float32)
    X[100:120] = X[3]
    Xd = torch.from_numpy(X).cuda()
    index = mlxlearn_amd.DeviceNeighborIndex(Xd)
    flags = _lib.PRECISIONS[precision]
    Q1 = torch.from_numpy(np.concatenate([X[:50], r.normal(size=(2950, 64)).astype(np.float32)])).cuda()
    Q2 = torch.from_numpy(r.normal(size=(333, 64)).astype(np.float32)).cuda()
    want1 = _call(lib, "mlk_knn_query", index, Q1, 10, flags)       # also prepares the mirror
    want2 = _call(lib, "mlk_knn_query", index, Q2, 10, flags)
    _same(_call(lib, "mlk_knn_query_noalloc", index, Q1, 10, flags), want1)
    _same(_call(lib, "mlk_knn_query_noalloc", index, Q2, 10, flags), want2)   # smaller, same scratch
    # self queries of a slice of the training rows
    Qs = Xd[1000:1600]
    want = _call(lib, "mlk_knn_query", index, Qs, 5, flags | _lib.EXCLUDE_SELF, 1000)
    _same(_call(lib, "mlk_knn_query_noalloc", index, Qs, 5, flags | _lib.EXCLUDE_SELF, 1000), want)
    index.close()


def test_noalloc_chunks_and_fallback(lib):
    import torch
    import mlxlearn_amd
    from mlxlearn_amd import _lib
    r = np.random.default_rng(42)
    X = r.normal(size=(30_000, 32)).astype(np.float32)
    index = mlxlearn_amd.DeviceNeighborIndex(torch.from_numpy(X).cuda())
    Q = torch.from_numpy(r.normal(size=(4000, 32)).astype(np.float32)).cuda()
    flags = _lib.PREC_BF16
    want = _call(lib, "mlk_knn_query", index, Q, 7, flags)
    # large candidate lists: the queries no longer fit one chunk of the scratch
    _lib.set_test_hooks(_lib.HOOK_CAPACITY, 4000)
    _same(_call(lib, "mlk_knn_query_noalloc", index, Q, 7, flags), want)
    # every query through the exact fallback, in batches that fit the scratch
    _lib.set_test_hooks(_lib.HOOK_FORCE_FALLBACK, 0)
    _same(_call(lib, "mlk_knn_query_noalloc", index, Q[:300], 7, flags), (want[0][:300], want[1][:300]))
    _lib.set_test_hooks(0, 0)
    _same(_call(lib, "mlk_knn_query_noalloc", index, Q[:100], 7, flags), (want[0][:100], want[1][:100]))
    index.close()


def test_noalloc_refusals(lib):
    import torch
    import mlxlearn_amd
    from mlxlearn_amd import _lib
    r = np.random.default_rng(43)
    X = r.normal(size=(5000, 16)).astype(np.float32)
    index = mlxlearn_amd.DeviceNeighborIndex(torch.from_numpy(X).cuda())
    Q = torch.from_numpy(X[:10]).cuda()
    with pytest.raises(ValueError, match="no screen mirror"):
        _call(lib, "mlk_knn_query_noalloc", index, Q, 3, _lib.PREC_FP16)
    index.close()
    big = mlxlearn_amd.DeviceNeighborIndex(torch.zeros((3_500_000, 2), dtype=torch.float32, device="cuda")
                                           + torch.arange(3_500_000, device="cuda", dtype=torch.float32)[:, None])
    Qb = torch.zeros((4, 2), dtype=torch.float32, device="cuda")
    _call(lib, "mlk_knn_query", big, Qb, 2, 0)
    with pytest.raises(ValueError, match="at most"):
        _call(lib, "mlk_knn_query_noalloc", big, Qb, 2, 0)
    big.close()
