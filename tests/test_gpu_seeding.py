"""GPU tests of the estimator front-end's kernels through the C ABI's Python
binding: per-feature mean / variance, centred transpose, column gather and
k-means++ seeding, against the seeding oracle (oracle/seeding_oracle.py,
itself pinned to scikit-learn 1.7.2) and NumPy.

k-means++ parity bar: identical chosen indices (so identical centers, which
are copies of samples) for the same uniforms; FP64 prefix sums differ from
the oracle's sequential ones only by rounding, which could move a draw that
lands within an ulp of a cumulative-sum step -- none of the seeded cases
here does."""
import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

from oracle import seeding_oracle as so
from tests.conftest import to_dev, to_np

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def q():
    return torch.cuda.current_stream()


def _blobs(rng, n, d, dtype, n_centers=None):
    n_centers = n_centers or max(2, min(64, n // 40))
    centers = rng.uniform(-10, 10, size=(n_centers, d))
    X = centers[rng.integers(0, n_centers, n)] + rng.normal(size=(n, d))
    return X.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n,d", [(1, 1), (1000, 3), (4097, 32), (3000, 300), (257, 1024)])
def test_column_mean_var(kd, dtype, n, d):
    rng = np.random.default_rng(n + d)
    X = (rng.normal(size=(n, d)) * rng.uniform(0.1, 50, d) + rng.normal(size=d) * 100).astype(dtype)
    mean = torch.empty(d, dtype=torch.float64, device="cuda")
    var = torch.empty(d, dtype=torch.float64, device="cuda")
    kd.column_mean_var(to_dev(X), mean, var, q())
    m_ref, v_ref = so.column_mean_var(X)
    assert_allclose(to_np(mean), m_ref, rtol=1e-12, atol=1e-12)
    assert_allclose(to_np(var), v_ref, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("rows,cols", [(1, 1), (63, 65), (1000, 32), (130, 7), (5, 2000)])
@pytest.mark.parametrize("shifted", [False, True])
def test_transpose_shift_is_exact(kd, dtype, rows, cols, shifted):
    rng = np.random.default_rng(rows * cols)
    A = rng.normal(size=(rows, cols)).astype(dtype)
    shift = rng.normal(size=cols)
    out = torch.empty((cols, rows), dtype=to_dev(A).dtype, device="cuda")
    kd.transpose(to_dev(A), out, to_dev(shift) if shifted else None, q())
    want = (A - shift.astype(dtype)[None, :]) if shifted else A
    assert_array_equal(to_np(out), want.T)


def test_gather_columns(kd):
    rng = np.random.default_rng(2)
    Xt = rng.normal(size=(9, 500)).astype(np.float32)
    idx = np.array([0, 499, 17, 17, 250], dtype=np.int64)
    out = torch.empty((9, 5), dtype=torch.float32, device="cuda")
    kd.gather_columns(to_dev(Xt), to_dev(idx), out, q())
    assert_array_equal(to_np(out), Xt[:, idx])


def _gpu_kpp(kd, X, w, u):
    n, d = X.shape
    k = u.shape[0]
    Xt = to_dev(np.ascontiguousarray(X.T))
    C = torch.empty((d, k), dtype=Xt.dtype, device="cuda")
    idx = torch.empty(k, dtype=torch.int64, device="cuda")
    kd.kmeans_plusplus_t(Xt, to_dev(w), u, C, idx, q())
    return to_np(C), to_np(idx)


KPP_CASES = [
    # n, d, k, trials, dtype, weighted
    (1000, 3, 10, None, np.float64, False),
    (5003, 17, 40, 5, np.float32, False),
    (20000, 32, 64, None, np.float32, True),
    (4096, 300, 16, 16, np.float64, False),   # features staged in two chunks
    (50, 2, 50, 3, np.float64, False),        # k == n
    (3001, 8, 33, 1, np.float32, False),      # one trial
    (1025, 5, 7, 2, np.float64, True),        # two segments, ragged tail
    (6000, 200, 24, None, np.float32, False), # LDS-staged rows, 4 feature chunks
    (4100, 64, 12, 8, np.float32, True),      # LDS-staged rows, ragged last segment
]


@pytest.mark.parametrize("n,d,k,T,dtype,weighted", KPP_CASES)
def test_kmeans_plusplus_matches_oracle(kd, n, d, k, T, dtype, weighted):
    rng = np.random.default_rng(n * 7 + k)
    X = _blobs(rng, n, d, dtype)
    w = (rng.uniform(0.1, 4.0, n) if weighted else np.ones(n)).astype(dtype)
    T = T or so.default_local_trials(k)
    u = so.kmeanspp_draws(np.random.RandomState(n + k), k, T)
    print(f"kmeans++ n={n} d={d} k={k} T={T} {np.dtype(dtype).name}", flush=True)
    C, idx = _gpu_kpp(kd, X, w, u)
    C_ref, idx_ref = so.kmeans_plusplus(X, w, u)
    assert_array_equal(idx, idx_ref)
    assert_array_equal(C, C_ref.T)


def test_kmeans_plusplus_matches_sklearn(kd):
    skc = pytest.importorskip("sklearn.cluster")
    rng = np.random.default_rng(77)
    X = _blobs(rng, 30000, 12, np.float64)
    C_ref, idx_ref = skc.kmeans_plusplus(X, 100, random_state=5)
    C, idx = kd.kmeans_plusplus(X, 100, random_state=5)      # host in -> host out
    assert_array_equal(idx, idx_ref)
    assert_array_equal(C, C_ref)


def test_kmeans_plusplus_large_is_deterministic_and_valid(kd):
    """n = 2,000,000, d = 32, k = 256: two runs give the same indices, all
    distinct, and the chosen set matches the oracle on its first centers."""
    rng = np.random.default_rng(9)
    n, d, k = 2_000_000, 32, 256
    Xt = torch.randn(d, n, device="cuda", generator=None)
    w = torch.ones(n, device="cuda")
    T = so.default_local_trials(k)
    u = so.kmeanspp_draws(np.random.RandomState(4), k, T)
    outs = []
    for _ in range(2):
        C = torch.empty((d, k), device="cuda")
        idx = torch.empty(k, dtype=torch.int64, device="cuda")
        kd.kmeans_plusplus_t(Xt, w, u, C, idx, q())
        outs.append((to_np(C), to_np(idx)))
    assert_array_equal(outs[0][1], outs[1][1])
    assert_array_equal(outs[0][0], outs[1][0])
    idx = outs[0][1]
    assert len(np.unique(idx)) == k
    assert_array_equal(outs[0][0], to_np(Xt[:, torch.as_tensor(idx, device="cuda")]))
    # the first centers against the oracle on the same data (host copy)
    X = to_np(Xt).T.astype(np.float64)
    _, idx_ref = so.kmeans_plusplus(X, np.ones(n), u[:4])
    assert_array_equal(idx[:4], idx_ref)


def test_kmeans_plusplus_rejects_bad_arguments(kd):
    Xt = torch.zeros((3, 10), device="cuda")
    w = torch.ones(10, device="cuda")
    C = torch.empty((3, 4), device="cuda")
    idx = torch.empty(4, dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError):
        kd.kmeans_plusplus_t(Xt, w, np.full((4, 17), 0.5), C, idx, q())  # 17 trials
    with pytest.raises(ValueError):
        kd.kmeans_plusplus_t(Xt, w, np.full((4, 2), 1.5), C, idx, q())   # u >= 1
    with pytest.raises(ValueError):
        kd.kmeans_plusplus_t(Xt, w, np.full((3, 2), 0.5), C, idx, q())   # rows != k
    C11 = torch.empty((3, 11), device="cuda")
    idx11 = torch.empty(11, dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError):
        kd.kmeans_plusplus_t(Xt, w, np.full((11, 2), 0.5), C11, idx11, q())  # k > n
