"""NumPy restatement of the estimator front-end's seeding and preparation.

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.

The reference has no seeding (its driver takes the initial centroids from the
caller, python_api/_kmeans_lloyd.cpp:1437).  The algorithm restated here is
scikit-learn 1.7.2's greedy k-means++ (sklearn/cluster/_kmeans.py:174-277,
the published algorithm of Arthur & Vassilvitskii with 2 + floor(ln k) local
trials) and its dataset-scaled tolerance (sklearn/cluster/_kmeans.py:279-287),
with the random draws taken as an explicit argument so the GPU path and this
oracle consume the same numbers.  Distances, potentials and prefix sums are
FP64 with direct differences (sum_f (x_f - c_f)^2).
"""
from __future__ import annotations

import numpy as np


def kmeanspp_draws(random_state, n_clusters: int, n_local_trials: int) -> np.ndarray:
    """The uniforms scikit-learn's k-means++ consumes, in its order.

    Row 0 column 0: the draw behind ``random_state.choice(n, p=w / w.sum())``
    (numpy's legacy choice draws one ``random_sample()`` and takes the first
    index whose normalised cdf exceeds it); row c >= 1: the
    ``random_state.uniform(size=n_local_trials)`` of center c
    (sklearn/cluster/_kmeans.py:225, 243).
    """
    u = np.zeros((n_clusters, n_local_trials), dtype=np.float64)
    u[0, 0] = random_state.random_sample()
    for c in range(1, n_clusters):
        u[c] = random_state.uniform(size=n_local_trials)
    return u


def default_local_trials(n_clusters: int) -> int:
    """2 + int(log(k)) (sklearn/cluster/_kmeans.py:218-222)."""
    return 2 + int(np.log(n_clusters))


def kmeans_plusplus(X, sample_weight, uniforms):
    """Greedy k-means++ over sample-major X (n, d) with the given draws.

    Returns (centers (k, d) in X's dtype, indices (k,) int64).
    Follows sklearn/cluster/_kmeans.py:213-269 step by step:
      first center: first i with cdf_i > u[0, 0], cdf = cumsum(w) / sum;
      center c: candidates = searchsorted(cumsum(w D^2), u[c] * pot, 'left')
      clipped to n - 1; D_t^2 = min(D^2, |x - cand_t|^2); winner = argmin_t
      sum w D_t^2 (first on ties); D^2 <- D_winner^2.
    """
    X = np.asarray(X)
    w = np.asarray(sample_weight, dtype=np.float64)
    u = np.asarray(uniforms, dtype=np.float64)
    n, d = X.shape
    k, T = u.shape
    Xd = X.astype(np.float64)
    cdf = np.cumsum(w)
    first = int(np.searchsorted(cdf / cdf[-1], u[0, 0], side="right"))
    first = min(first, n - 1)
    indices = np.full(k, -1, dtype=np.int64)
    indices[0] = first

    def sqdist(j):
        diff = Xd - Xd[j][None, :]
        return np.einsum("ij,ij->i", diff, diff)

    closest = sqdist(first)
    pot = float(closest @ w)
    for c in range(1, k):
        r = u[c] * pot
        cand = np.searchsorted(np.cumsum(w * closest), r, side="left")
        cand = np.minimum(cand, n - 1)
        dist = np.stack([np.minimum(closest, sqdist(j)) for j in cand])
        pots = dist @ w
        best = int(np.argmin(pots))
        pot = float(pots[best])
        closest = dist[best]
        indices[c] = cand[best]
    return X[indices].copy(), indices


def column_mean_var(X):
    """Per-feature mean and variance (np.var: mean squared deviation), FP64."""
    Xd = np.asarray(X, dtype=np.float64)
    mean = Xd.mean(axis=0)
    var = ((Xd - mean) ** 2).mean(axis=0)
    return mean, var


def scaled_tolerance(X, tol: float) -> float:
    """tol x mean per-feature variance, 0 when tol == 0
    (sklearn/cluster/_kmeans.py:279-287)."""
    if tol == 0:
        return 0.0
    return float(np.mean(column_mean_var(X)[1]) * tol)
