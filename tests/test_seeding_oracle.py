"""Pins of the seeding oracle (oracle/seeding_oracle.py) on the CPU.

The oracle restates scikit-learn 1.7.2's greedy k-means++
(sklearn/cluster/_kmeans.py:174-277) with the random draws passed in; these
tests check it against scikit-learn itself (same RandomState -> same
indices, same position in the random stream afterwards), numpy's weighted
choice rule for the first center, and scikit-learn's dataset-scaled
tolerance.  A planted mistake in the oracle (the candidate search's side,
the trial count) is shown to fail the scikit-learn comparison.
"""
import numpy as np
import pytest

from oracle import seeding_oracle as so

sk_cluster = pytest.importorskip("sklearn.cluster")


def _blobs(rng, n, d, dtype):
    centers = rng.normal(size=(max(2, n // 50), d)) * 6
    X = centers[rng.integers(0, len(centers), n)] + rng.normal(size=(n, d))
    return X.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("weighted", [False, True])
def test_oracle_kmeanspp_equals_sklearn(dtype, weighted):
    rng = np.random.default_rng(11 + weighted)
    for case in range(8):
        n = int(rng.integers(40, 2500))
        d = int(rng.integers(1, 24))
        k = int(rng.integers(1, min(48, n)))
        X = _blobs(rng, n, d, dtype)
        w = (rng.uniform(0.2, 3.0, n) if weighted else np.ones(n)).astype(dtype)
        seed = int(rng.integers(1 << 30))
        T = so.default_local_trials(k)
        C_ref, i_ref = sk_cluster.kmeans_plusplus(X, k, sample_weight=w,
                                                  random_state=seed)
        u = so.kmeanspp_draws(np.random.RandomState(seed), k, T)
        C, idx = so.kmeans_plusplus(X, w, u)
        assert np.array_equal(idx, i_ref), (case, n, d, k)
        assert np.array_equal(C, C_ref)


def test_draws_leave_the_stream_where_sklearn_does():
    X = _blobs(np.random.default_rng(3), 600, 5, np.float64)
    for k in (1, 2, 9, 30):
        rs_a, rs_b = np.random.RandomState(7), np.random.RandomState(7)
        sk_cluster.kmeans_plusplus(X, k, random_state=rs_a)
        so.kmeanspp_draws(rs_b, k, so.default_local_trials(k))
        assert rs_a.random_sample() == rs_b.random_sample()


def test_first_center_is_numpy_weighted_choice():
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = int(rng.integers(1, 300))
        w = rng.uniform(0, 2, n) * (rng.uniform(size=n) > 0.2)
        if w.sum() == 0:
            w[0] = 1.0
        seed = int(rng.integers(1 << 30))
        want = np.random.RandomState(seed).choice(n, p=w / w.sum())
        u = np.zeros((1, 1))
        u[0, 0] = np.random.RandomState(seed).random_sample()
        X = np.arange(n, dtype=np.float64)[:, None]
        _, idx = so.kmeans_plusplus(X, w, u)
        assert idx[0] == want


def test_greedy_choice_by_brute_force():
    """Tiny case: the winner is the candidate with the smallest potential,
    recomputed here by enumeration."""
    X = np.array([[0.0], [1.0], [10.0], [11.0], [30.0]])
    w = np.ones(5)
    u = np.array([[0.1, 0.0, 0.0], [0.05, 0.5, 0.97]])
    _, idx = so.kmeans_plusplus(X, w, u)
    first = 0                                         # cdf 0.2 > 0.1
    d2 = (X[:, 0] - X[first, 0]) ** 2                 # [0, 1, 100, 121, 900]
    cums = np.cumsum(d2)                              # pot = 1122
    cands = [int(np.argmax(cums >= r)) for r in u[1] * d2.sum()]
    pots = [np.minimum(d2, (X[:, 0] - X[c, 0]) ** 2).sum() for c in cands]
    assert idx[0] == first and idx[1] == cands[int(np.argmin(pots))]


def test_planted_mistakes_fail_the_sklearn_comparison():
    rng = np.random.default_rng(21)
    X = _blobs(rng, 3000, 4, np.float64)
    w = np.ones(3000)
    k, seed = 24, 1234
    _, i_ref = sk_cluster.kmeans_plusplus(X, k, random_state=seed)
    # wrong number of local trials
    u_bad = so.kmeanspp_draws(np.random.RandomState(seed), k, so.default_local_trials(k) + 1)
    assert not np.array_equal(so.kmeans_plusplus(X, w, u_bad)[1], i_ref)
    # potential maximised instead of minimised
    u = so.kmeanspp_draws(np.random.RandomState(seed), k, so.default_local_trials(k))
    orig = np.argmin
    try:
        so.np.argmin = np.argmax
        bad = so.kmeans_plusplus(X, w, u)[1]
    finally:
        so.np.argmin = orig
    assert not np.array_equal(bad, i_ref)


def test_scaled_tolerance_and_moments_match_sklearn():
    from sklearn.cluster._kmeans import _tolerance
    rng = np.random.default_rng(8)
    X = rng.normal(size=(5000, 7)) * rng.uniform(0.1, 20, 7) + 3
    mean, var = so.column_mean_var(X)
    np.testing.assert_allclose(mean, X.mean(axis=0), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(var, X.var(axis=0), rtol=1e-12)
    assert so.scaled_tolerance(X, 1e-4) == pytest.approx(_tolerance(X, 1e-4), rel=1e-12)
    assert so.scaled_tolerance(X, 0) == 0.0
