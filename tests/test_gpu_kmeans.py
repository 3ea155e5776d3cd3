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
@pytest.mark.parametrize("n,k,d", [(7, 3, 1), (33, 17, 9), (129, 65, 31), (83, 11, 128), (5000, 700, 96),
                                   (3000, 300, 784), (2000, 257, 1100)])
def test_assignment_matches_oracle(mlk, precision, n, k, d):
    """Shapes of tests/unit/test_e_means.py:101-113 plus wider ones (the wide and the deep screens)."""
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


def _blobs(seed, n, d, k, scale=10.0):
    r = np.random.default_rng(seed)
    centres = r.normal(scale=scale, size=(k, d))
    X = (centres[r.integers(0, k, n)] + r.normal(size=(n, d))).astype(np.float32)
    return X, (centres + r.normal(scale=0.5, size=centres.shape)).astype(np.float32)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_label_reuse_equals_exact_assignment(mlk, precision):
    """Certified reuse of the previous labels (reference lloyd_bounds.py:24-76): the
    labels and residuals are those of a full exact assignment, bit for bit."""
    torch = pytest.importorskip("torch")
    from mlxlearn_amd.cluster._k_means import assign_device
    X, C0 = _blobs(21, 40_000, 32, 64)
    l0, _ = oracle.kmeans_assign(X, C0)
    sums, counts = oracle.kmeans_update(X, l0, 64)
    C1 = np.where(counts[:, None] > 0, sums / np.maximum(counts, 1)[:, None], C0).astype(np.float32)
    reused = []
    labels, resid = assign_device(X, C1, precision=precision, previous=torch.as_tensor(l0).cuda(),
                                  out_reused=reused)
    want_l, want_r = oracle.kmeans_assign(X, C1)
    np.testing.assert_array_equal(labels.cpu().numpy(), want_l)
    np.testing.assert_allclose(resid.cpu().numpy(), want_r, rtol=1e-12, atol=0)
    full_l, full_r = assign_device(X, C1, precision=precision)
    assert torch.equal(labels, full_l) and torch.equal(resid, full_r)
    assert reused[0] > len(X) // 2, reused  # separated blobs: most rows keep their label
    # previous labels that are all wrong: nothing is certified wrongly
    bad_prev = (torch.as_tensor(want_l).cuda() + 1) % 64
    labels, resid = assign_device(X, C1, precision=precision, previous=bad_prev)
    np.testing.assert_array_equal(labels.cpu().numpy(), want_l)
    assert torch.equal(resid, full_r)
    # dense centres (the separation test fails on the sample): the plain assignment
    Xd, _ = oracle.gen_c5(seed=21, n=20_000, d=32, n_components=16, n_clusters=16)
    Cd = Xd[np.random.default_rng(2).choice(len(Xd), 512, replace=False)].copy()
    want_l, want_r = oracle.kmeans_assign(Xd, Cd)
    reused = []
    labels, resid = assign_device(Xd, Cd, precision=precision, previous=torch.as_tensor(want_l).cuda(),
                                  out_reused=reused)
    np.testing.assert_array_equal(labels.cpu().numpy(), want_l)
    np.testing.assert_allclose(resid.cpu().numpy(), want_r, rtol=1e-12, atol=0)


def test_label_reuse_edge_cases(mlk):
    import ctypes
    torch = pytest.importorskip("torch")
    from mlxlearn_amd import _lib
    from mlxlearn_amd.cluster._k_means import assign_device
    X, C = _blobs(5, 6000, 12, 20)
    # duplicate centres: separation 0, never certified; the lowest index wins
    C[7] = C[3]
    C[15] = C[3]
    want_l, want_r = oracle.kmeans_assign(X, C)
    prev = torch.as_tensor(want_l).cuda()
    prev[torch.as_tensor(want_l == 3).cuda()] = 15
    reused = []
    labels, resid = assign_device(X, C, previous=prev, out_reused=reused)
    np.testing.assert_array_equal(labels.cpu().numpy(), want_l)
    np.testing.assert_allclose(resid.cpu().numpy(), want_r, rtol=1e-12, atol=0)
    assert 0 < reused[0] < 6000
    # one centre: every row keeps label 0
    reused = []
    labels, _ = assign_device(X, C[:1], previous=torch.zeros(6000, dtype=torch.int64).cuda(), out_reused=reused)
    assert reused == [6000] and int(labels.abs().sum().item()) == 0
    # labels outside [0, n_clusters) are an input error, whether or not the gate passes
    for Cx in (C, np.zeros_like(C)):
        bad = prev.clone()
        bad[11] = 20
        with pytest.raises(ValueError, match="outside"):
            assign_device(X, Cx, previous=bad)
    # in place: labels written over the previous labels
    Xd, Cd = torch.as_tensor(X).cuda(), torch.as_tensor(C).cuda()
    inplace = torch.as_tensor(want_l).cuda()
    resid = torch.empty(6000, dtype=torch.float64).cuda()
    n_reused = ctypes.c_int64()
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.mlk_kmeans_assign_reuse(ctypes.c_void_p(Xd.data_ptr()), 6000, 12, ctypes.c_void_p(Cd.data_ptr()), 20,
                                           ctypes.c_void_p(inplace.data_ptr()), 0, ctypes.c_void_p(inplace.data_ptr()),
                                           ctypes.c_void_p(resid.data_ptr()), ctypes.byref(n_reused), s))
    np.testing.assert_array_equal(inplace.cpu().numpy(), want_l)
    assert 0 < n_reused.value < 6000


def test_kmeans_fit_reuses_labels(mlk):
    """The Lloyd loop reuses certified labels after its first iteration; results stay exact."""
    X, _ = _blobs(4, 20_000, 16, 24)
    km = mlk.KMeans(n_clusters=24, init="k-means++", random_state=3, max_iter=20, tol=0).fit(X)
    l, res = oracle.kmeans_assign(X, km.cluster_centers_)
    np.testing.assert_array_equal(km.labels_, l)
    assert math.isclose(km.inertia_, math.fsum(res), rel_tol=1e-12)
    assert km.n_iter_ > 1 and km.reused_labels_ > 0


def test_misaligned_rows_assignment(mlk, rng):
    """Rows one float into a buffer (not 16-byte aligned): same labels and residuals."""
    torch = pytest.importorskip("torch")
    from mlxlearn_amd.cluster._k_means import assign_device
    X = rng.normal(size=(9000, 96)).astype(np.float32)
    C = X[:300] + np.float32(0.05)
    want_l, want_r = oracle.kmeans_assign(X, C)
    buf = torch.zeros(X.size + 1, dtype=torch.float32, device="cuda")
    Xm = buf[1:].view(X.shape)
    Xm.copy_(torch.as_tensor(X))
    assert Xm.data_ptr() % 16 != 0
    for prec in PRECISIONS:
        labels, resid = assign_device(Xm, C, precision=prec)
        np.testing.assert_array_equal(labels.cpu().numpy(), want_l)
        np.testing.assert_allclose(resid.cpu().numpy(), want_r, rtol=1e-12, atol=0)


def _relocated_centres(X, labels, residual, k):
    """NumPy restatement of the reference's empty-cluster relocation
    (_k_means.py:116-142): the farthest rows (largest residual, lower row on a
    tie) leave their clusters for the empty ones in ascending id order; a
    cluster left empty takes the largest cluster's mean."""
    sums, counts = oracle.kmeans_update(X, labels, k)
    empty = np.flatnonzero(counts == 0)
    far = np.argsort(-residual, kind="stable")[:empty.size]
    if empty.size and residual.max() > 0:
        for e, r in zip(empty, far):
            sums[labels[r]] -= X[r]
            counts[labels[r]] -= 1
            sums[e] = X[r]
            counts[e] = 1
    means = sums / np.maximum(counts, 1)[:, None]
    means[counts == 0] = means[np.argmax(counts)]
    return means.astype(np.float32), empty, far


def test_kmeans_relocates_empty_clusters(mlk):
    """Initial centres far from every row leave clusters empty; one Lloyd step
    moves the farthest rows into them (and refills a donor that empties)."""
    r = np.random.default_rng(123)
    X = (r.normal(size=(3000, 5)) + 10 * r.integers(0, 4, (3000, 1))).astype(np.float32)
    X[17] = [400.0, 0, 0, 0, 0]                         # an outlier alone in its cluster
    C0 = np.concatenate([X[[0, 1, 2, 3, 4, 5]], [[200.0, 0, 0, 0, 0]],  # centre 6 holds only the outlier
                         np.full((3, 5), 1e4, np.float32)]).astype(np.float32)
    labels, resid = oracle.kmeans_assign(X, C0)
    assert np.bincount(labels, minlength=10)[6] == 1 and np.bincount(labels, minlength=10)[7:].sum() == 0
    want, empty, far = _relocated_centres(X, labels, resid, 10)
    assert list(empty) == [7, 8, 9] and far[0] == 17    # the outlier moves first and empties centre 6
    km = mlk.KMeans(n_clusters=10, init=C0, max_iter=1, tol=0).fit(X)
    got = km.cluster_centers_
    for e, row in zip(empty, far):
        np.testing.assert_array_equal(got[e], X[row])   # a relocated cluster is its one row, exactly
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)
    l1, r1 = oracle.kmeans_assign(X, got)
    np.testing.assert_array_equal(km.labels_, l1)
