"""scikit-learn style estimator over the library's kernels.

The reference's README (README.md:3-4) names the intended caller of
``kmeans_lloyd_driver``: an sklearn-style ``KMeans`` estimator holding
sample-major data.  This module is that caller for ROCm devices:

* ``fit`` takes X (n_samples, n_features) -- a torch tensor on a ROCm device,
  any DLPack ROCm array, or a host array (copied to the device) -- computes
  the per-feature means and variances on the GPU, writes the centred,
  feature-major X_t the Lloyd kernels read (one transpose kernel), scales
  ``tol`` by the mean feature variance (scikit-learn's ``_tolerance``),
  seeds with GPU k-means++ (scikit-learn's greedy algorithm, same random
  stream) or random samples or a given array, and runs ``n_init`` fits with
  ``kmeans_lloyd_driver`` (several fits at once with the batched engine when
  the shape allows), keeping the lowest inertia as scikit-learn does.
* ``predict``, ``score``, ``transform`` (and ``fit_predict``,
  ``fit_transform``) reuse the assignment, inertia and distance kernels.

Differences from scikit-learn's KMeans: the Lloyd loop stops when the summed
squared centre shift is <= the scaled tol (the reference driver's rule,
src/kmeans_lloyd_driver.hpp:83); it does not stop early on unchanged labels,
so ``n_iter_`` can be one more than scikit-learn's for the same result.
Only ``algorithm="lloyd"`` exists.  Results are device tensors when X was a
device array, NumPy arrays when X was a host array.
"""
from __future__ import annotations

import numbers
import warnings

import numpy as np

from . import _kmeans_dpcpp_amd as _native

__all__ = ["KMeans", "kmeans_plusplus"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("kmeans_dpcpp_amd.KMeans needs a ROCm device")
    return torch


def _check_random_state(seed):
    """None -> numpy's global RandomState; int -> RandomState(seed); a
    RandomState passes through (sklearn.utils.check_random_state)."""
    if seed is None or seed is np.random:
        return np.random.mtrand._rand
    if isinstance(seed, numbers.Integral):
        return np.random.RandomState(seed)
    if isinstance(seed, np.random.RandomState):
        return seed
    raise ValueError(f"{seed!r} cannot be used to seed a numpy.random.RandomState")


def _as_device(X, device, dtype=None):
    """(tensor on the device, was_host).  Keeps float32 / float64; other
    dtypes become float64 (as scikit-learn's check_array does)."""
    torch = _torch()
    if isinstance(X, torch.Tensor):
        t, host = X, X.device.type != "cuda"
    elif hasattr(X, "__dlpack__") and hasattr(X, "__dlpack_device__") and \
            X.__dlpack_device__()[0] not in (1, 3):      # not CPU / pinned CPU
        t, host = torch.from_dlpack(X), False
    else:
        t, host = torch.as_tensor(np.asarray(X)), True
    if dtype is None:
        dtype = t.dtype if t.dtype in (torch.float32, torch.float64) else torch.float64
    t = t.to(device=device, dtype=dtype).contiguous()
    return t, host


def _n_local_trials(n_clusters):
    return 2 + int(np.log(n_clusters))


def kmeans_plusplus(X, n_clusters, *, sample_weight=None, x_squared_norms=None,
                    random_state=None, n_local_trials=None, device=None):
    """GPU k-means++ seeding with scikit-learn's signature
    (sklearn.cluster.kmeans_plusplus).  Returns (centers (k, d), indices (k,)).
    The same random_state gives scikit-learn's draws; the indices equal its
    indices except where a draw lands within rounding of a cumulative-sum
    step (FP64 here).  x_squared_norms is accepted for compatibility and not
    used (the kernels compute the norms they need)."""
    torch = _torch()
    device = device or torch.device("cuda", torch.cuda.current_device())
    Xd, host = _as_device(X, device)
    n, d = Xd.shape
    if n < n_clusters:
        raise ValueError(f"n_samples={n} should be >= n_clusters={n_clusters}.")
    w = (torch.ones(n, dtype=Xd.dtype, device=device) if sample_weight is None
         else _as_device(sample_weight, device, Xd.dtype)[0])
    T = _n_local_trials(n_clusters) if n_local_trials is None else int(n_local_trials)
    rs = _check_random_state(random_state)
    u = _draws(rs, n_clusters, T)
    X_t = torch.empty((d, n), dtype=Xd.dtype, device=device)
    q = torch.cuda.current_stream(device)
    _native.transpose(Xd, X_t, None, q)
    centers_t = torch.empty((d, n_clusters), dtype=Xd.dtype, device=device)
    idx = torch.empty(n_clusters, dtype=torch.int64, device=device)
    _native.kmeans_plusplus_t(X_t, w, u, centers_t, idx, q)
    centers = centers_t.t().contiguous()
    if host:
        return centers.cpu().numpy(), idx.cpu().numpy()
    return centers, idx


def _draws(rs, n_clusters, T):
    """Uniforms in the order scikit-learn's k-means++ consumes them: the
    first center's choice() draw, then uniform(size=T) per later center."""
    u = np.zeros((n_clusters, T), dtype=np.float64)
    u[0, 0] = rs.random_sample()
    for c in range(1, n_clusters):
        u[c] = rs.uniform(size=T)
    return u


class KMeans:
    """K-means clustering (Lloyd) on an MI355X; scikit-learn's KMeans API.

    Parameters follow sklearn.cluster.KMeans: n_clusters, init
    ("k-means++", "random", an array (n_clusters, n_features) or a callable
    (X, n_clusters, random_state) -> array, called with the centred X and
    its result taken in centred coordinates, as scikit-learn does), n_init
    ("auto": 1 for k-means++ / array, 10 for random / callable), max_iter,
    tol (scaled by
    the mean feature variance), verbose, random_state, copy_x (X is never
    modified here).  Extra: precision ("fp32", "bf16x3", "bf16", "fp8": the cross
    term of the assignment; labels equal the fp32 ones in every mode),
    device (default: the current ROCm device).
    """

    def __init__(self, n_clusters=8, *, init="k-means++", n_init="auto",
                 max_iter=300, tol=1e-4, verbose=0, random_state=None,
                 copy_x=True, algorithm="lloyd", precision="fp32", device=None):
        self.n_clusters = n_clusters
        self.init = init
        self.n_init = n_init
        self.max_iter = max_iter
        self.tol = tol
        self.verbose = verbose
        self.random_state = random_state
        self.copy_x = copy_x
        self.algorithm = algorithm
        self.precision = precision
        self.device = device

    # -- helpers ------------------------------------------------------------
    def _dev(self):
        torch = _torch()
        return self.device or torch.device("cuda", torch.cuda.current_device())

    def _check_params(self, n):
        if not (isinstance(self.n_clusters, numbers.Integral) and self.n_clusters >= 1):
            raise ValueError("n_clusters must be a positive integer")
        if n < self.n_clusters:
            raise ValueError(f"n_samples={n} should be >= n_clusters={self.n_clusters}.")
        if self.algorithm != "lloyd":
            raise ValueError("only algorithm='lloyd' is implemented")
        if self.precision not in ("fp32", "bf16x3", "bf16", "fp8"):
            raise ValueError("precision must be 'fp32', 'bf16x3', 'bf16' or 'fp8'")
        if not (isinstance(self.max_iter, numbers.Integral) and self.max_iter >= 1):
            raise ValueError("max_iter must be a positive integer")
        if self.tol < 0:
            raise ValueError("tol must be >= 0")
        init_is_array = not isinstance(self.init, str) and not callable(self.init)
        if isinstance(self.init, str) and self.init not in ("k-means++", "random"):
            raise ValueError("init must be 'k-means++', 'random', an array or a callable")
        if self.n_init == "auto":
            # scikit-learn 1.7's rule: 10 runs for 'random' and for a callable,
            # one for 'k-means++' and for an explicit array
            n_init = 10 if ((isinstance(self.init, str) and self.init == "random")
                            or callable(self.init)) else 1
        else:
            n_init = int(self.n_init)
            if n_init < 1:
                raise ValueError("n_init must be >= 1 or 'auto'")
        if init_is_array and n_init != 1:
            warnings.warn("Explicit initial center position passed: performing only"
                          f" one init in KMeans instead of n_init={n_init}.",
                          RuntimeWarning, stacklevel=3)
            n_init = 1
        return n_init

    def _init_centers(self, X_t, Xd, w, mean, rs, q):
        """Initial centres (d, k) in the centred coordinates of X_t."""
        torch = _torch()
        d, n = X_t.shape
        k = self.n_clusters
        if isinstance(self.init, str) and self.init == "k-means++":
            T = _n_local_trials(k)
            u = _draws(rs, k, T)
            C = torch.empty((d, k), dtype=X_t.dtype, device=X_t.device)
            idx = torch.empty(k, dtype=torch.int64, device=X_t.device)
            _native.kmeans_plusplus_t(X_t, w, u, C, idx, q)
            return C
        if isinstance(self.init, str) and self.init == "random":
            wh = w.cpu().numpy()
            seeds = rs.choice(n, size=k, replace=False, p=wh / wh.sum())
            idx = torch.as_tensor(seeds.astype(np.int64), device=X_t.device)
            C = torch.empty((d, k), dtype=X_t.dtype, device=X_t.device)
            _native.gather_columns(X_t, idx, C, q)
            return C
        if callable(self.init):
            # as scikit-learn: the callable sees the centred data and its
            # result is taken in centred coordinates
            Xc = X_t.t()
            Xc = Xc.cpu().numpy() if self._host else Xc.contiguous()
            C0, _ = _as_device(self.init(Xc, k, random_state=rs), X_t.device,
                               X_t.dtype)
            if tuple(C0.shape) != (k, d):
                raise ValueError(f"The shape of the initial centers {tuple(C0.shape)}"
                                 f" does not match the number of clusters {k} / "
                                 f"features {d}.")
            C = torch.empty((d, k), dtype=X_t.dtype, device=X_t.device)
            _native.transpose(C0.contiguous(), C, None, q)
            return C
        C0, _ = _as_device(self.init, X_t.device, X_t.dtype)
        if tuple(C0.shape) != (k, d):
            raise ValueError(f"The shape of the initial centers {tuple(C0.shape)} does"
                             f" not match the number of clusters {k} / features {d}.")
        C = torch.empty((d, k), dtype=X_t.dtype, device=X_t.device)
        _native.transpose(C0, C, mean, q)   # (k, d)^T - mean per feature
        return C

    @staticmethod
    def _same_clustering(l1, l2, k):
        """sklearn's _is_same_clustering: l1 -> l2 is a consistent mapping."""
        torch = _torch()
        pairs = torch.unique(l1.to(torch.int64) * k + l2.to(torch.int64))
        return pairs.numel() == torch.unique(l1).numel()

    # -- API --------------------------------------------------------------------
    def fit(self, X, y=None, sample_weight=None):
        torch = _torch()
        dev = self._dev()
        Xd, host = _as_device(X, dev)
        if Xd.ndim != 2:
            raise ValueError("X must be a 2-D array (n_samples, n_features)")
        n, d = Xd.shape
        n_init = self._check_params(n)
        self._host = host
        w = (torch.ones(n, dtype=Xd.dtype, device=dev) if sample_weight is None
             else _as_device(sample_weight, dev, Xd.dtype)[0])
        if w.shape != (n,):
            raise ValueError("sample_weight must have shape (n_samples,)")
        q = torch.cuda.current_stream(dev)
        mean = torch.empty(d, dtype=torch.float64, device=dev)
        var = torch.empty(d, dtype=torch.float64, device=dev)
        _native.column_mean_var(Xd, mean, var, q)
        X_t = torch.empty((d, n), dtype=Xd.dtype, device=dev)
        _native.transpose(Xd, X_t, mean, q)            # centred, feature-major
        tol = 0.0 if self.tol == 0 else float(var.mean().item()) * float(self.tol)
        self._tol = tol
        rs = _check_random_state(self.random_state)
        k = self.n_clusters
        inits = [self._init_centers(X_t, Xd, w, mean, rs, q) for _ in range(n_init)]
        if self.verbose:
            print("Initialization complete")
        runs = self._run_lloyd(X_t, w, inits, tol, q)
        best = None
        for labels, res, n_iter, inertia in runs:
            if best is None or (inertia < best[3] and
                                not self._same_clustering(labels, best[0], k)):
                best = (labels, res, n_iter, inertia)
        labels, res, n_iter, inertia = best
        centers = (res.t() + mean.to(res.dtype)[None, :]).contiguous()   # (k, d)
        distinct = torch.unique(labels).numel()
        if distinct < k:
            warnings.warn(f"Number of distinct clusters ({distinct}) found smaller than "
                          f"n_clusters ({k}). Possibly due to duplicate points in X.",
                          RuntimeWarning, stacklevel=2)
        self.cluster_centers_ = centers.cpu().numpy() if host else centers
        self.labels_ = labels.cpu().numpy() if host else labels
        self.inertia_ = float(inertia)
        self.n_iter_ = int(n_iter)
        self.n_features_in_ = d
        self._centers_dev = centers
        return self

    def _run_lloyd(self, X_t, w, inits, tol, q):
        torch = _torch()
        d, n = X_t.shape
        k = self.n_clusters
        out = []
        batched = (len(inits) > 1 and self.precision in ("fp32", "bf16") and
                   X_t.dtype == torch.float32 and not self.verbose)
        if batched:
            fits = []
            for C in inits:
                lab = torch.empty(n, dtype=torch.int32, device=X_t.device)
                res = torch.empty_like(C)
                fits.append((X_t, w, C, lab, res, tol, self.max_iter))
            # shapes outside the batched kernels run fit by fit in the driver
            results, _ = _native.kmeans_lloyd_batched(fits, self.precision, q)
            for (_, _, _, lab, res, _, _), (it, inertia) in zip(fits, results):
                out.append((lab, res, it, float(inertia)))
            return out
        for C in inits:
            lab = torch.empty(n, dtype=torch.int32, device=X_t.device)
            res = torch.empty_like(C)
            it, inertia, _ = _native.kmeans_lloyd_driver_ex(
                X_t, w, C, lab, res, tol, bool(self.verbose), self.max_iter,
                self.precision, q)
            out.append((lab, res, it, float(inertia)))
        return out

    def _prep_predict(self, X):
        torch = _torch()
        if not hasattr(self, "_centers_dev"):
            raise RuntimeError("This KMeans instance is not fitted yet")
        dev = self._dev()
        C = self._centers_dev
        Xd, host = _as_device(X, dev, C.dtype)
        if Xd.ndim != 2 or Xd.shape[1] != C.shape[1]:
            raise ValueError(f"X has {Xd.shape[-1]} features, but KMeans is expecting "
                             f"{C.shape[1]} features as input.")
        q = torch.cuda.current_stream(dev)
        n, d = Xd.shape
        X_t = torch.empty((d, n), dtype=Xd.dtype, device=dev)
        _native.transpose(Xd, X_t, None, q)
        C_t = C.t().contiguous()
        return X_t, C_t, host, q

    def _labels(self, X_t, C_t, q):
        torch = _torch()
        k = C_t.shape[1]
        hn = torch.empty(k, dtype=C_t.dtype, device=C_t.device)
        _native.half_l2_norm_squared(C_t, hn, 128, q)
        lab = torch.empty(X_t.shape[1], dtype=torch.int32, device=C_t.device)
        _native.assignment(X_t, C_t, hn, lab, 32, 128, q)
        return lab

    def predict(self, X, sample_weight=None):
        X_t, C_t, host, q = self._prep_predict(X)
        lab = self._labels(X_t, C_t, q)
        return lab.cpu().numpy() if host else lab

    def fit_predict(self, X, y=None, sample_weight=None):
        return self.fit(X, sample_weight=sample_weight).labels_

    def score(self, X, y=None, sample_weight=None):
        torch = _torch()
        X_t, C_t, host, q = self._prep_predict(X)
        n = X_t.shape[1]
        w = (torch.ones(n, dtype=X_t.dtype, device=X_t.device) if sample_weight is None
             else _as_device(sample_weight, X_t.device, X_t.dtype)[0])
        lab = self._labels(X_t, C_t, q)
        per = torch.empty(n, dtype=X_t.dtype, device=X_t.device)
        _native.compute_inertia(X_t, w, C_t, lab, per, 128, q)
        return -float(_native.reduce_vector_blocking(per, q))

    def transform(self, X):
        torch = _torch()
        X_t, C_t, host, q = self._prep_predict(X)
        d, n = X_t.shape
        k = C_t.shape[1]
        D = torch.empty((k, n), dtype=X_t.dtype, device=X_t.device)
        _native.compute_centroid_to_sample_distances(X_t, C_t, D, 128, 32, q)
        out = torch.empty((n, k), dtype=X_t.dtype, device=X_t.device)
        _native.transpose(D, out, None, q)
        return out.cpu().numpy() if host else out

    def fit_transform(self, X, y=None, sample_weight=None):
        return self.fit(X, sample_weight=sample_weight).transform(X)
