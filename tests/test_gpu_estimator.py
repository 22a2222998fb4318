"""GPU tests of kmeans_dpcpp_amd.KMeans (the sklearn-style caller of
kmeans_lloyd_driver) against scikit-learn's KMeans(algorithm="lloyd") on the
same data, initialisation and random stream.

Parity bar: on well-separated blobs the labels are identical; centres and
inertia agree to 1e-9 relative in float64 and 1e-5 in float32 (the FP64
oracle's bar for the Lloyd path).  The truth for a float32 fit is
scikit-learn's float64 fit of the same values: scikit-learn's own float32
Lloyd can drift from it (one label flipped by float32 centre rounding
cascades), while this library's exact integer centroid sums track it.
n_iter_ may exceed scikit-learn's by one: scikit-learn also stops when the
labels did not change, the reference driver's loop only on the centre shift
(src/kmeans_lloyd_driver.hpp:83)."""
import warnings

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
skc = pytest.importorskip("sklearn.cluster")


def _blobs(seed, n, d, k, dtype, spread=20.0):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-spread, spread, size=(k, d))
    X = centers[rng.integers(0, k, n)] + rng.normal(size=(n, d)) * rng.uniform(0.5, 1.5, k)[rng.integers(0, k, n)][:, None]
    return X.astype(dtype)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("dtype,rtol", [(np.float64, 1e-9), (np.float32, 1e-5)])
def test_fit_from_array_init_matches_sklearn(kd, dtype, rtol):
    X = _blobs(1, 20000, 9, 12, dtype)
    init = X[np.random.default_rng(2).choice(len(X), 12, replace=False)]
    ours = kd.KMeans(12, init=init, n_init=1, max_iter=100, tol=1e-6).fit(X)
    ref = skc.KMeans(12, init=init.astype(np.float64), n_init=1, max_iter=100,
                     tol=1e-6, algorithm="lloyd").fit(X.astype(np.float64))
    assert_array_equal(ours.labels_, ref.labels_)
    assert _rel(ours.cluster_centers_, ref.cluster_centers_) < rtol
    assert ours.inertia_ == pytest.approx(ref.inertia_, rel=10 * rtol)
    if dtype == np.float64:
        assert ref.n_iter_ <= ours.n_iter_ <= ref.n_iter_ + 1


def test_fit_kmeanspp_matches_sklearn(kd):
    """Same random_state: the same k-means++ seeds (GPU FP64 seeding equals
    scikit-learn's float64 seeding), then the same Lloyd fit."""
    X = _blobs(3, 30000, 16, 40, np.float64)
    ours = kd.KMeans(40, random_state=7, max_iter=200).fit(X)
    ref = skc.KMeans(40, random_state=7, max_iter=200, algorithm="lloyd").fit(X)
    assert_array_equal(ours.labels_, ref.labels_)
    assert _rel(ours.cluster_centers_, ref.cluster_centers_) < 1e-9


def test_fit_random_init_n_init_is_best_of_single_fits(kd):
    """n_init = 4 random inits in float32 run as one batched call.  The seeds
    are scikit-learn's (the same RandomState.choice calls), and the kept fit
    is bit-identical to the best of the four fits run one at a time."""
    X = _blobs(4, 15000, 8, 10, np.float32, spread=8.0)
    ours = kd.KMeans(10, init="random", n_init=4, random_state=11, max_iter=300).fit(X)
    rs = np.random.RandomState(11)
    w = np.ones(len(X), dtype=np.float32)
    singles = []
    for _ in range(4):
        seeds = rs.choice(len(X), size=10, replace=False, p=w / w.sum())
        singles.append(kd.KMeans(10, init=X[seeds], n_init=1, max_iter=300).fit(X))
    best = None
    for f in singles:           # scikit-learn's selection rule
        if best is None or (f.inertia_ < best.inertia_ and
                            len(np.unique(f.labels_.astype(np.int64) * 10 + best.labels_)) !=
                            len(np.unique(f.labels_))):
            best = f
    assert_array_equal(ours.labels_, best.labels_)
    assert_array_equal(ours.cluster_centers_, best.cluster_centers_)
    assert ours.inertia_ == best.inertia_


def test_weighted_fit_and_device_tensors(kd):
    X = _blobs(5, 12000, 6, 8, np.float64)
    w = np.random.default_rng(6).uniform(0.5, 2.0, len(X))
    Xd, wd = torch.from_numpy(X).cuda(), torch.from_numpy(w).cuda()
    ours = kd.KMeans(8, random_state=1).fit(Xd, sample_weight=wd)
    ref = skc.KMeans(8, random_state=1, algorithm="lloyd").fit(X, sample_weight=w)
    assert isinstance(ours.labels_, torch.Tensor) and ours.labels_.is_cuda
    assert_array_equal(ours.labels_.cpu().numpy(), ref.labels_)
    assert _rel(ours.cluster_centers_.cpu().numpy(), ref.cluster_centers_) < 1e-9


def test_predict_transform_score(kd):
    X = _blobs(7, 10000, 5, 6, np.float64)
    Xn = _blobs(8, 3000, 5, 6, np.float64)
    ours = kd.KMeans(6, random_state=3).fit(X)
    ref = skc.KMeans(6, random_state=3, algorithm="lloyd").fit(X)
    assert_array_equal(ours.predict(Xn), ref.predict(Xn))
    assert_allclose(ours.transform(Xn), ref.transform(Xn), rtol=1e-9, atol=1e-9)
    assert ours.score(Xn) == pytest.approx(ref.score(Xn), rel=1e-9)
    assert_array_equal(ours.fit_predict(X), ours.labels_)
    assert ours.fit_transform(Xn).shape == (3000, 6)


def test_precision_modes_give_identical_fits(kd):
    X = _blobs(9, 40000, 24, 64, np.float32)
    base = kd.KMeans(64, random_state=2, tol=0, max_iter=15).fit(X)
    for prec in ("bf16x3", "bf16"):
        other = kd.KMeans(64, random_state=2, tol=0, max_iter=15, precision=prec).fit(X)
        assert_array_equal(other.labels_, base.labels_)
        assert_array_equal(other.cluster_centers_, base.cluster_centers_)


def test_tolerance_is_scaled_by_feature_variance(kd):
    from sklearn.cluster._kmeans import _tolerance
    X = _blobs(10, 5000, 4, 5, np.float64) * np.array([1.0, 10.0, 100.0, 0.1])
    est = kd.KMeans(5, tol=1e-3, random_state=0).fit(X)
    assert est._tol == pytest.approx(_tolerance(X, 1e-3), rel=1e-12)


@pytest.mark.parametrize("n,d,k", [(7, 1, 1), (5, 3, 5), (300, 1, 4), (64, 2, 64)])
def test_edge_shapes_match_sklearn(kd, n, d, k):
    """k = 1, d = 1, n = k: same labels and centres as scikit-learn from the
    same k-means++ seeds."""
    rng = np.random.default_rng(n * 10 + d)
    X = rng.normal(size=(n, d)) * 3 + rng.integers(0, 4, size=(n, 1)) * 10
    ours = kd.KMeans(k, random_state=4, max_iter=50).fit(X)
    ref = skc.KMeans(k, random_state=4, max_iter=50, algorithm="lloyd").fit(X)
    assert_array_equal(ours.labels_, ref.labels_)
    assert _rel(ours.cluster_centers_, ref.cluster_centers_) < 1e-9
    assert ours.inertia_ == pytest.approx(ref.inertia_, rel=1e-9, abs=1e-9)


def test_bad_parameters(kd):
    X = np.zeros((5, 2))
    with pytest.raises(ValueError):
        kd.KMeans(6).fit(X)
    with pytest.raises(ValueError):
        kd.KMeans(2, init="nope").fit(X)
    with pytest.raises(ValueError):
        kd.KMeans(2, precision="fp4").fit(X)
    with pytest.raises(ValueError):
        kd.KMeans(2, init=np.zeros((3, 2))).fit(X + np.arange(5)[:, None])
    with warnings.catch_warnings(record=True):
        warnings.simplefilter("always")
        est = kd.KMeans(2, init=np.array([[0.0, 0.0], [4.0, 4.0]]), n_init=3).fit(
            np.arange(10, dtype=float).reshape(5, 2))
    assert est.n_iter_ >= 1


def test_callable_init_matches_sklearn(kd):
    """A callable init sees the centred X and its result is used in centred
    coordinates (scikit-learn's rule); n_init='auto' runs it 10 times.  The
    callable halves the chosen samples, so it does not commute with the
    centring: passing uncentred X would give other seeds."""
    X = _blobs(11, 12000, 6, 9, np.float64) + 50.0

    def half_samples(Xc, k, random_state):
        return np.asarray(Xc)[random_state.choice(len(Xc), k, replace=False)] * 0.5

    for n_init in (1, "auto"):
        ours = kd.KMeans(9, init=half_samples, n_init=n_init, max_iter=100,
                         tol=1e-6, random_state=5).fit(X)
        ref = skc.KMeans(9, init=half_samples, n_init=n_init, max_iter=100,
                         tol=1e-6, random_state=5, algorithm="lloyd").fit(X)
        assert_array_equal(ours.labels_, ref.labels_)
        assert _rel(ours.cluster_centers_, ref.cluster_centers_) < 1e-9
        assert ours.inertia_ == pytest.approx(ref.inertia_, rel=1e-9)
