"""k-means assignment (k = 1 search over the centroids) and update on the GPU vs the oracle.

Labels are the exact nearest centre with the lowest index on ties, so they
must equal the oracle's bit for bit; residuals are the canonical float64
squared distances (rtol 1e-12); centroid sums are float64 sums of float32
rows in row order (within count * 2^-52 * sum|x| of the exactly rounded
sums), counts exact, inertia within 1e-12 relative of math.fsum.
"""
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

PRECISIONS = ["float32", "bfloat16", "float16"]


@pytest.fixture(scope="module")
def mlk():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mlxlearn_amd
    from mlxlearn_amd import _lib
    _lib.load()
    return mlxlearn_amd


def assign(X, C, precision="float32"):
    from mlxlearn_amd.cluster._k_means import assign_device
    labels, residual = assign_device(X, C, precision=precision)
    return labels.cpu().numpy(), residual.cpu().numpy()


def check_update(X, labels, kc):
    import torch
    from mlxlearn_amd.cluster._k_means import update_device
    _, residual = oracle.kmeans_assign(X, X[:1])  # any residual vector for the inertia check
    res_d = torch.as_tensor(residual).cuda()
    sums, counts, inertia = update_device(X, torch.as_tensor(labels).cuda(), kc, res_d)
    want_sums, want_counts = oracle.kmeans_update(X, labels, kc)
    np.testing.assert_array_equal(counts.cpu().numpy(), want_counts)
    absum = np.zeros_like(want_sums)
    np.add.at(absum, labels, np.abs(X.astype(np.float64)))
    bound = np.maximum(want_counts, 1)[:, None] * 2.0 ** -52 * absum
    assert np.all(np.abs(sums.cpu().numpy() - want_sums) <= bound)
    assert math.isclose(float(inertia.item()), math.fsum(residual), rel_tol=1e-12)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("n,k,d", [(7, 3, 1), (33, 17, 9), (129, 65, 31), (83, 11, 128), (5000, 700, 96)])
def test_assignment_matches_oracle(mlk, precision, n, k, d):
    """Shapes of tests/unit/test_e_means.py:101-113 plus a wider one."""
    r = np.random.default_rng(2600 + d)
    X = r.normal(size=(n, d)).astype(np.float32)
    C = r.normal(size=(k, d)).astype(np.float32)
    want_l, want_r = oracle.kmeans_assign(X, C)
    got_l, got_r = assign(X, C, precision)
    np.testing.assert_array_equal(got_l, want_l)
    np.testing.assert_allclose(got_r, want_r, rtol=1e-12, atol=0)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_duplicate_centroids_choose_lowest_index(mlk, precision):
    """tests/unit/test_e_means.py:115-126: 67 copies of one centre -> every label is 0."""
    r = np.random.default_rng(747)
    X = r.normal(size=(41, 17)).astype(np.float32)
    C = np.tile(X[0], (67, 1))
    labels, residual = assign(X, C, precision)
    np.testing.assert_array_equal(labels, np.zeros(41, np.int64))
    assert residual[0] == 0.0
    np.testing.assert_allclose(residual, ((X.astype(np.float64) - C[0]) ** 2).sum(1), rtol=1e-12)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_mixture_workload_reduced(mlk, precision):
    """Workload c5 generator at reduced size (200,000 x 96 rows, 2,048 centroids from the rows)."""
    X, C = oracle.gen_c5(seed=3, n=200_000, n_components=512, n_clusters=2048)
    want_l, want_r = oracle.kmeans_assign(X, C)
    got_l, got_r = assign(X, C, precision)
    np.testing.assert_array_equal(got_l, want_l)
    np.testing.assert_allclose(got_r, want_r, rtol=1e-12, atol=0)
    check_update(X, want_l, 2048)


def test_update_sums_counts_inertia(mlk, rng):
    X = rng.normal(size=(20_000, 33)).astype(np.float32)
    labels = rng.integers(0, 300, 20_000)
    labels[labels == 17] = 18  # an empty cluster
    check_update(X, labels, 300)


def test_update_rejects_labels_out_of_range(mlk, rng):
    import torch
    from mlxlearn_amd.cluster._k_means import update_device
    X = rng.normal(size=(100, 4)).astype(np.float32)
    for bad in (-1, 10):
        labels = torch.zeros(100, dtype=torch.int64).cuda()
        labels[37] = bad
        with pytest.raises(ValueError, match="outside"):
            update_device(X, labels, 10)


def test_kmeans_one_lloyd_step(mlk):
    """KMeans(init=C0, max_iter=1): centres = float32 means of the exact assignment,
    labels_ and inertia_ from a fresh exact assignment to those centres (_k_means.py:363-423)."""
    r = np.random.default_rng(77)
    X, _ = oracle.gen_c5(seed=9, n=30_000, d=24, n_components=64, n_clusters=16)
    C0 = X[r.choice(len(X), 40, replace=False)].copy()
    km = mlk.KMeans(n_clusters=40, init=C0, max_iter=1, tol=0).fit(X)
    l0, _ = oracle.kmeans_assign(X, C0)
    sums, counts = oracle.kmeans_update(X, l0, 40)
    assert counts.min() > 0
    C1 = (sums / counts[:, None]).astype(np.float32)
    np.testing.assert_array_equal(km.cluster_centers_, C1)
    l1, r1 = oracle.kmeans_assign(X, C1)
    np.testing.assert_array_equal(km.labels_, l1)
    assert math.isclose(km.inertia_, math.fsum(r1), rel_tol=1e-12)
    np.testing.assert_array_equal(km.predict(X[:500]), l1[:500])
    assert km.n_iter_ == 1


def test_kmeans_converges_on_separated_blobs(mlk):
    r = np.random.default_rng(5)
    centres = r.normal(scale=20, size=(6, 5))
    X = (centres[r.integers(0, 6, 3000)] + r.normal(size=(3000, 5))).astype(np.float32)
    km = mlk.KMeans(n_clusters=6, n_init=3, random_state=0, init="random").fit(X)
    l, res = oracle.kmeans_assign(X, km.cluster_centers_)
    np.testing.assert_array_equal(km.labels_, l)
    assert math.isclose(km.inertia_, math.fsum(res), rel_tol=1e-12)
    assert km.convergence_reason_ in ("labels", "tolerance")
