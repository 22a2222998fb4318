"""NumPy restatement of the reference's Lloyd entry points (test oracle).

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.

Every function states the reference lines it restates.  Layouts are the
reference's: X_t (n_features, n_samples), centroids_t (n_features,
n_clusters), private copies (n_copies, n_features, n_clusters), distances
(n_clusters, n_samples).  Functions return new arrays instead of writing into
caller buffers; "accumulate-into" and "caller-initialised counter" semantics
are modelled with explicit input arguments.

Arithmetic: in the data type of the inputs unless a function says otherwise.
Sequential sums over features are written as explicit loops so the order of
accumulation is the reference's (feature index ascending).
"""
from __future__ import annotations

import math

import numpy as np

WINDOW_N_CENTROIDS = 32  # preferred_work_group_size_multiple (8) x width multiplier (4),
                         # python_api/_kmeans_lloyd.cpp:20-21, src/lloyd_single_step.hpp:58-60
PRIVATIZATION_GRANULE = 8  # src/lloyd_single_step.hpp:174-176


# ---------------------------------------------------------------------------
# small kernels
# ---------------------------------------------------------------------------

def broadcast_divide(X, y):
    """X[r, c] / y[c].  src/util_kernels.hpp:9-41; the binding passes
    X.shape[0] as n_features and X.shape[1] as n_clusters
    (python_api/_kmeans_lloyd.cpp:91-92)."""
    X = np.asarray(X)
    return (X / np.asarray(y, dtype=X.dtype)[None, :]).astype(X.dtype)


def half_l2_norm_squared(C_t):
    """hn[c] = (sum_f C[f, c]^2) / 2, f ascending.  src/util_kernels.hpp:47-82."""
    C_t = np.asarray(C_t)
    acc = np.zeros(C_t.shape[1], dtype=C_t.dtype)
    for f in range(C_t.shape[0]):
        acc = acc + C_t[f] * C_t[f]
    return (acc / C_t.dtype.type(2)).astype(C_t.dtype)


def reduce_centroids_data(sizes_copies, cent_copies, empty_list_in, n_empty_in):
    """Sum private copies (copy index ascending) and list empty clusters.

    src/util_kernels.hpp:84-161.  The reference appends empty clusters with an
    atomic counter starting at the caller's n_empty value (order
    nondeterministic, :143-152); the restatement uses ascending cluster order,
    one of the reference's possible outcomes.  Slots not written keep their
    value (tests/test_kernels.py:121-128 canary)."""
    sizes_copies = np.asarray(sizes_copies)
    cent_copies = np.asarray(cent_copies)
    dt = sizes_copies.dtype
    sizes = np.zeros(sizes_copies.shape[1], dtype=dt)
    cent = np.zeros(cent_copies.shape[1:], dtype=dt)
    for c in range(sizes_copies.shape[0]):
        sizes = sizes + sizes_copies[c]
        cent = cent + cent_copies[c]
    lst = np.array(empty_list_in, copy=True)
    n_empty = int(n_empty_in)
    for c in np.flatnonzero(sizes == 0):
        if 0 <= n_empty < lst.shape[0]:
            lst[n_empty] = c
        n_empty += 1
    return sizes, cent, lst, n_empty


def compute_centroid_shifts_squared(old_t, new_t):
    """shift[c] = sum_f (old - new)^2, f ascending.
    src/util_kernels.hpp:478-516."""
    old_t = np.asarray(old_t)
    new_t = np.asarray(new_t)
    acc = np.zeros(old_t.shape[1], dtype=old_t.dtype)
    for f in range(old_t.shape[0]):
        diff = old_t[f] - new_t[f]
        acc = acc + diff * diff
    return acc


def compute_inertia(X_t, sample_weight, C_t, labels):
    """I[s] = w[s] * sum_f (X[f, s] - C[f, l_s])^2 (weight applied after the
    sum).  src/compute_inertia.hpp:11-53; with sample_weight None the
    unweighted form, src/compute_inertia.hpp:58-99."""
    X_t = np.asarray(X_t)
    C_t = np.asarray(C_t)
    labels = np.asarray(labels).astype(np.int64)
    acc = np.zeros(X_t.shape[1], dtype=X_t.dtype)
    for f in range(X_t.shape[0]):
        diff = X_t[f] - C_t[f, labels]
        acc = acc + diff * diff
    if sample_weight is None:
        return acc
    return (acc * np.asarray(sample_weight, dtype=X_t.dtype)).astype(X_t.dtype)


def reduce_vector(data):
    """Sum of a vector.  src/compute_inertia.hpp:104-135 (reduction order
    unspecified in the reference); computed in FP64."""
    return float(np.sum(np.asarray(data, dtype=np.float64)))


def compute_number_of_private_copies(cache_bytes, itemsize, n_samples,
                                     n_features, n_clusters, occupancy,
                                     work_group_size):
    """min(floor(cache * occ / (sizeof(T) k (d+1))), ceil(n/wg)*wg/8),
    clamped to >= 1 (the reference can return 0).
    src/lloyd_single_step.hpp:12-35."""
    n_bytes = itemsize * n_clusters * (n_features + 1)
    copies = int(math.floor(cache_bytes * occupancy / n_bytes))
    global_size = -(-n_samples // work_group_size) * work_group_size
    copies = min(global_size // PRIVATIZATION_GRANULE, copies)
    return max(1, copies)


# ---------------------------------------------------------------------------
# sort / threshold / selection / relocation
# ---------------------------------------------------------------------------

def _sort_key(x):
    """Total order of the sort: -0 == +0, NaN after every number."""
    x = np.asarray(x)
    return np.where(np.isnan(x), np.inf, x), np.isnan(x)


def iterative_merge_sort(data):
    """Stable ascending sort (std::less), NaN last.
    src/iterative_merge_sort.hpp:204-368 (merges take the first run on ties:
    lower_bound for first-run items, upper_bound for second-run items,
    :108 and :158)."""
    data = np.asarray(data)
    val, isnan = _sort_key(data)
    # lexsort: last key is primary; stable for equal keys
    order = np.lexsort((np.arange(data.shape[0]), val, isnan))
    return data[order]


def compute_threshold(data, topk):
    """topk == 1: max; topk >= 2: ascending_sort(data)[n - topk]; topk == 0:
    None (no-op).  src/util_kernels.hpp:163-206,
    python_api/_kmeans_lloyd.cpp:312-316."""
    data = np.asarray(data)
    if topk == 0:
        return None
    if topk == 1:
        finite = data[~np.isnan(data)]
        return data.dtype.type(finite.max() if finite.size else -np.inf)
    return iterative_merge_sort(data)[data.shape[0] - topk]


def select_samples_far_from_centroid(n_selected, dist, threshold, selected_in,
                                     n_gt_in, n_eq_in):
    """Indices with d > T at the front (ascending index), indices with
    d == T at the back at n - n_eq (n_eq growing from the caller's value),
    until n_selected entries are selected.  src/util_kernels.hpp:208-297
    (the reference's order within each group and its early exits are
    race-dependent, :255-260, :279)."""
    dist = np.asarray(dist)
    n = dist.shape[0]
    sel = np.array(selected_in, copy=True)
    gt = np.flatnonzero(dist > threshold)
    eq = np.flatnonzero(dist == threshold)
    n_gt = int(n_gt_in)
    for i in gt:
        if 0 <= n_gt < n:
            sel[n_gt] = i
        n_gt += 1
    eq_limit = max(0, n_selected - gt.shape[0])
    n_eq = int(n_eq_in)
    for i in eq[:eq_limit]:
        pos = n - n_eq
        if 0 <= pos < n:
            sel[pos] = i
        n_eq += 1
    return sel, n_gt, n_eq


def farthest_samples(sq_dist, m):
    """The m farthest samples ordered by (sq_dist descending, index
    ascending): the relocation order this build defines for the reference's
    src/util_kernels.hpp:353 (which is UB for j >= n_gt)."""
    sq_dist = np.asarray(sq_dist)
    order = np.lexsort((np.arange(sq_dist.shape[0]), -sq_dist.astype(np.float64)))
    return order[:m]


def relocate_empty_clusters(n_empty, X_t, sample_weight, labels, empty_list,
                            sq_dist, C_t, sizes, per_sample_inertia):
    """Move the j-th farthest sample into the j-th empty cluster, on
    unnormalised sums.  src/util_kernels.hpp:302-390 (kernel) and :392-476
    (host): sum[:, prev] -= w x; sum[:, e] = w x; size[prev] -= w;
    size[e] = w; inertia[s] = 0."""
    X_t = np.asarray(X_t)
    C = np.array(C_t, copy=True)
    sz = np.array(sizes, copy=True)
    inert = np.array(per_sample_inertia, copy=True)
    labels = np.asarray(labels).astype(np.int64)
    w = np.asarray(sample_weight)
    m = min(int(n_empty), X_t.shape[1])
    for j, s in enumerate(farthest_samples(sq_dist, m)):
        p = labels[s]
        e = int(empty_list[j])
        add = X_t[:, s] * w[s]
        C[:, p] = C[:, p] - add
        C[:, e] = add
        sz[p] = sz[p] - w[s]
        sz[e] = w[s]
        inert[s] = 0
    return C, sz, inert


# ---------------------------------------------------------------------------
# assignment, distances, fused step
# ---------------------------------------------------------------------------

def dot_products(X_t, C_t):
    """dot[c, s] = sum_f C[f, c] X[f, s] accumulated with f ascending in the
    input dtype (src/device_functions.hpp:120-136)."""
    X_t = np.asarray(X_t)
    C_t = np.asarray(C_t)
    acc = np.zeros((C_t.shape[1], X_t.shape[1]), dtype=X_t.dtype)
    for f in range(X_t.shape[0]):
        acc = acc + C_t[f][:, None] * X_t[f][None, :]
    return acc


def assignment_from_pseudo(pseudo):
    """argmin over centroids with the reference rule: a centroid replaces the
    running best iff its value is strictly smaller, scanning centroids in
    ascending order from (index 0, +inf) (src/device_functions.hpp:139-162,
    src/lloyd_single_step.hpp:91-92).  NaN never wins; all-NaN/inf -> 0."""
    pseudo = np.asarray(pseudo)
    k, n = pseudo.shape
    best = np.full(n, np.inf, dtype=pseudo.dtype)
    idx = np.zeros(n, dtype=np.int64)
    for c in range(k):
        upd = pseudo[c] < best
        best = np.where(upd, pseudo[c], best)
        idx = np.where(upd, c, idx)
    return idx


def assignment(X_t, C_t, hn):
    """labels = argmin_c (hn[c] - dot[c, s]).  src/assignment.hpp:11-146."""
    hn = np.asarray(hn)
    pseudo = hn[:, None] - dot_products(X_t, C_t)
    return assignment_from_pseudo(pseudo)


def compute_distances(X_t, C_t):
    """D[c, s] = sqrt(sum_f (C[f, c] - X[f, s])^2), correct for every shape
    (the reference's window loops are swapped,
    src/compute_euclidean_distance.hpp:56,64)."""
    X_t = np.asarray(X_t)
    C_t = np.asarray(C_t)
    acc = np.zeros((C_t.shape[1], X_t.shape[1]), dtype=X_t.dtype)
    for f in range(X_t.shape[0]):
        diff = C_t[f][:, None] - X_t[f][None, :]
        acc = acc + diff * diff
    return np.sqrt(acc)


def fused_lloyd_single_step(X_t, sample_weight, C_t, hn, n_copies):
    """Labels as assignment(); then accumulate w_s x_s into copy
    (s / 8) % n_copies and w_s into the size copy.
    src/lloyd_single_step.hpp:37-209.  Returns (labels, copies, size copies)
    starting from zeroed copies."""
    X_t = np.asarray(X_t)
    w = np.asarray(sample_weight, dtype=X_t.dtype)
    d, n = X_t.shape
    k = np.asarray(C_t).shape[1]
    labels = assignment(X_t, C_t, hn)
    copies = np.zeros((n_copies, d, k), dtype=X_t.dtype)
    sizes = np.zeros((n_copies, k), dtype=X_t.dtype)
    priv = (np.arange(n) // PRIVATIZATION_GRANULE) % n_copies
    for s in range(n):
        copies[priv[s], :, labels[s]] += X_t[:, s] * w[s]
        sizes[priv[s], labels[s]] += w[s]
    return labels, copies, sizes


# ---------------------------------------------------------------------------
# the driver
# ---------------------------------------------------------------------------

def kmeans_lloyd(X_t, sample_weight, init_t, tol, max_iter, dtype=np.float64):
    """Full fit with the driver's loop semantics
    (src/kmeans_lloyd_driver.hpp:21-431):

      while it < max_iter and sum(shift) > tol:        (:83)
          hn; labels; sums/sizes                      (:86-170)
          if n_empty > 0: sq_dist (unweighted) with current centroids,
                          relocate on the sums         (:207-312)
          centroids = sums / sizes; shift; swap        (:317-353)
      final labels from the final centroids, weighted inertia  (:356-418)

    Computed in `dtype` (float64 = the accuracy truth).  Returns
    (labels, centroids_t, n_iter, inertia)."""
    X_t = np.asarray(X_t, dtype=dtype)
    w = np.asarray(sample_weight, dtype=dtype)
    C = np.array(init_t, dtype=dtype, copy=True)
    d, n = X_t.shape
    k = C.shape[1]
    tol_t = dtype(tol)
    it = 0
    shift = dtype(np.inf)
    while it < max_iter and shift > tol_t:
        hn = (np.einsum("fc,fc->c", C, C) / 2).astype(dtype)
        pseudo = hn[:, None] - (C.T @ X_t)
        labels = assignment_from_pseudo(pseudo)
        sums = np.zeros((d, k), dtype=dtype)
        for f in range(d):
            sums[f] = np.bincount(labels, weights=X_t[f] * w, minlength=k)
        sizes = np.bincount(labels, weights=w, minlength=k).astype(dtype)
        empty = np.flatnonzero(sizes == 0)
        if empty.size:
            sq = compute_inertia(X_t, None, C, labels)
            sums, sizes, _ = relocate_empty_clusters(
                empty.size, X_t, w, labels, empty, sq, sums, sizes, sq)
        new = (sums / sizes[None, :]).astype(dtype)
        shift = dtype(np.sum((C - new) ** 2))
        C = new
        it += 1
    hn = (np.einsum("fc,fc->c", C, C) / 2).astype(dtype)
    labels = assignment_from_pseudo(hn[:, None] - (C.T @ X_t))
    inertia = float(np.sum(compute_inertia(X_t, w, C, labels).astype(np.float64)))
    return labels, C, it, inertia


def pseudo_margin(X_t, C_t):
    """FP64 best and second-best pseudo-inertia gap per sample, used to tell
    genuine label disagreements from near-ties."""
    X = np.asarray(X_t, dtype=np.float64)
    C = np.asarray(C_t, dtype=np.float64)
    hn = np.einsum("fc,fc->c", C, C) / 2
    pseudo = hn[:, None] - C.T @ X
    part = np.partition(pseudo, 1, axis=0) if pseudo.shape[0] > 1 else pseudo
    return part[1] - part[0] if pseudo.shape[0] > 1 else np.full(X.shape[1], np.inf)


# ---------------------------------------------------------------------------
# synthetic data of the benchmark configurations
# ---------------------------------------------------------------------------

def gaussian_blobs(n, d, k, seed, center_box=10.0, std=1.0, std_range=None,
                   anisotropic=False, dtype=np.float32):
    """Blob samples in the reference's transposed layout X_t (d, n) plus a
    random-sample init (d, k).  Centers uniform in [-box, box]^d; per-blob
    std either fixed or uniform in std_range; anisotropic blobs use a random
    orthogonal rotation and log-uniform axis scales in [0.1, 3]."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-center_box, center_box, size=(k, d))
    blob = rng.integers(0, k, size=n)
    if anisotropic:
        X = np.empty((n, d), dtype=np.float64)
        q, _ = np.linalg.qr(rng.standard_normal((d, d)))
        scales = np.exp(rng.uniform(np.log(0.1), np.log(3.0), size=(k, d)))
        z = rng.standard_normal((n, d))
        for c in range(k):
            m = blob == c
            X[m] = centers[c] + (z[m] * scales[c]) @ q.T
    else:
        if std_range is not None:
            stds = rng.uniform(std_range[0], std_range[1], size=k)
        else:
            stds = np.full(k, std)
        X = centers[blob] + rng.standard_normal((n, d)) * stds[blob][:, None]
    X_t = np.ascontiguousarray(X.T.astype(dtype))
    init_idx = rng.permutation(n)[:k]
    init_t = np.ascontiguousarray(X_t[:, init_idx])
    return X_t, init_t
