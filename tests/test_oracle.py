"""Pins the oracle (oracle/lloyd_oracle.py, oracle/restate.c) before it is
trusted: against the reference test-suite's known answers
(tests/golden/kats.npz, from the reference's tests/test_kernels.py), against
independent library routines (scipy cdist, np.sort, np.partition,
scikit-learn KMeans), brute force on tiny inputs, and invariants.  A planted
mistake in the oracle must fail these checks (see test_planted_*)."""
import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal
from scipy.spatial.distance import cdist

from oracle import c_oracle as co
from oracle import lloyd_oracle as lo
from tests.conftest import CORNERS

DTYPES = [np.float32, np.float64]


def res(dt):
    return np.finfo(dt).resolution


@pytest.mark.parametrize("dt", DTYPES)
def test_broadcast_divide_kat(golden, dt):
    out = lo.broadcast_divide(golden["bdiv_X"].astype(dt), golden["bdiv_y"].astype(dt))
    assert_allclose(out, golden["bdiv_expected"], rtol=res(dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_half_l2_norm_kat(golden, dt):
    assert_allclose(lo.half_l2_norm_squared(golden["hn_X"].astype(dt)),
                    golden["hn_expected"], rtol=res(dt))
    assert_allclose(co.half_norm(golden["hn_X"].astype(dt)), golden["hn_expected"],
                    rtol=res(dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_reduce_centroids_kat(golden, dt):
    for pre, n_empty_exp in (("rcd", 0), ("rcde", 1)):
        sc = golden[f"{pre}_sizes_copies"].astype(dt)
        cc = golden[f"{pre}_cent_copies"].astype(dt)
        canary = int(golden["rcd_canary"])
        sizes, cent, lst, n_empty = lo.reduce_centroids_data(
            sc, cc, np.full(4, canary), 0)
        assert n_empty == n_empty_exp
        assert_allclose(cent, cc.astype(np.float64).sum(axis=0), rtol=res(dt))
        assert_allclose(sizes, sc.sum(axis=0), rtol=res(dt))
        if pre == "rcd":
            assert_array_equal(lst, np.full(4, canary))
        else:
            assert_array_equal(lst, golden["rcde_expected_list"])


@pytest.mark.parametrize("dt", DTYPES)
def test_shifts_kat(golden, dt):
    assert_allclose(lo.compute_centroid_shifts_squared(golden["shift_a"].astype(dt),
                                                       golden["shift_b"].astype(dt)),
                    golden["shift_expected"], rtol=res(dt))


def test_reduce_vector_kat(golden):
    assert lo.reduce_vector(golden["rv_data"].astype(np.float32)) == float(golden["rv_expected"])


@pytest.mark.parametrize("dt", DTYPES)
def test_threshold_kat(golden, dt):
    data = golden["thr_data"].astype(dt)
    n = data.shape[0]
    assert float(lo.compute_threshold(data, 1)) == float(max(data))
    assert float(lo.compute_threshold(data, 2)) == float(np.partition(data, n - 2)[n - 2])
    assert lo.compute_threshold(data, 0) is None


@pytest.mark.parametrize("dt", DTYPES)
def test_select_samples_kat(golden, dt):
    dist = golden["sel_dist"].astype(dt)
    n, m = dist.shape[0], int(golden["sel_n_empty"])
    thr = lo.compute_threshold(dist, m)
    sel, ngt, neq = lo.select_samples_far_from_centroid(
        m, dist, thr, np.full(n, 172), 0, 1)
    # invariants asserted by tests/test_kernels.py:360-362
    assert ngt + neq - 1 == m
    assert np.all(dist[sel[:ngt]] > thr)
    assert np.all(dist[sel[1 - neq:]] == thr)


@pytest.mark.parametrize("dt", DTYPES)
def test_relocate_kat(golden, dt):
    X = golden["rel_X"].astype(dt)
    Xt = np.ascontiguousarray(X.T)
    C = golden["rel_C"].astype(dt)
    Ct = np.ascontiguousarray(C.T)
    sq = np.min(np.square(X[:, None, :] - C[None, :, :]).sum(axis=-1), axis=-1)
    Cn, sz, inert = lo.relocate_empty_clusters(
        1, Xt, np.ones(8, dt), np.zeros(8, np.int64), golden["rel_empty_list"],
        sq, Ct, golden["rel_sizes"].astype(dt), sq)
    exp_inert = sq.copy()
    exp_inert[int(golden["rel_moved_sample"])] = 0
    assert_allclose(sz, golden["rel_expected_sizes"], rtol=res(dt))
    assert_allclose(Cn, golden["rel_expected_centroid_t"], rtol=res(dt))
    assert_allclose(inert, exp_inert, rtol=res(dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_distances_vs_cdist(golden, dt):
    X = golden["dist_X"].astype(dt)
    D = lo.compute_distances(np.ascontiguousarray(X.T), np.ascontiguousarray(CORNERS.T.astype(dt)))
    f = 64 if dt == np.float32 else 1   # tests/test_kernels.py:598,638
    assert_allclose(D, cdist(CORNERS, X), rtol=f * res(dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_assignment_kat_and_brute_force(golden, dt):
    for key, cloud in (("assign_X_1734", 1734), ("assign_X_32", 32)):
        X = golden[key].astype(dt)
        Xt = np.ascontiguousarray(X.T)
        Ct = np.ascontiguousarray(CORNERS.T.astype(dt))
        exp = np.repeat(np.arange(8), cloud)
        assert_array_equal(lo.assignment(Xt, Ct, lo.half_l2_norm_squared(Ct)), exp)
        assert_array_equal(co.assign(Xt, Ct), exp)
    rng = np.random.default_rng(7)
    for _ in range(20):
        n, d, k = rng.integers(1, 40), rng.integers(1, 9), rng.integers(1, 70)
        X = rng.standard_normal((n, d)).astype(dt)
        C = rng.standard_normal((k, d)).astype(dt)
        brute = np.argmin(cdist(X.astype(np.float64), C.astype(np.float64)), axis=1)
        Xt, Ct = np.ascontiguousarray(X.T), np.ascontiguousarray(C.T)
        assert_array_equal(lo.assignment(Xt, Ct, lo.half_l2_norm_squared(Ct)), brute)
        assert_array_equal(co.assign(Xt, Ct), brute)


def test_assignment_tie_and_nan_rules():
    # duplicated centroids: the lowest index wins (strict <, ascending scan)
    C = np.array([[0.0, 1.0, 0.0, 1.0]], dtype=np.float32)   # d=1, k=4
    X = np.array([[0.0, 1.0, 0.4, 0.6]], dtype=np.float32)
    assert_array_equal(lo.assignment(X, C, lo.half_l2_norm_squared(C)), [0, 1, 0, 1])
    assert_array_equal(co.assign(X, C), [0, 1, 0, 1])
    # NaN sample -> no strict improvement over +inf -> label 0
    Xn = np.array([[np.nan, 2.0]], dtype=np.float32)
    C2 = np.array([[5.0, 2.0]], dtype=np.float32)
    assert_array_equal(lo.assignment(Xn, C2, lo.half_l2_norm_squared(C2)), [0, 1])
    assert_array_equal(co.assign(Xn, C2), [0, 1])


@pytest.mark.parametrize("dt", DTYPES)
def test_inertia_vs_cdist(golden, dt):
    X = golden["inertia_X"].astype(dt)
    ids = np.repeat(np.arange(8), 32)
    w = np.linspace(0.5, 2.0, X.shape[0]).astype(dt)
    out = lo.compute_inertia(np.ascontiguousarray(X.T), w, np.ascontiguousarray(CORNERS.T.astype(dt)), ids)
    exp = cdist(X, CORNERS, "sqeuclidean")[np.arange(X.shape[0]), ids] * w
    assert_allclose(out, exp, rtol=8 * res(dt))
    out_u = lo.compute_inertia(np.ascontiguousarray(X.T), None, np.ascontiguousarray(CORNERS.T.astype(dt)), ids)
    assert_allclose(out_u, exp / w, rtol=8 * res(dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_fused_step_kat(golden, dt):
    X = golden["step_X"].astype(dt)
    Xt = np.ascontiguousarray(X.T)
    Ct = np.ascontiguousarray(CORNERS.T.astype(dt))
    labels, copies, sizes = lo.fused_lloyd_single_step(
        Xt, np.ones(X.shape[0], dt), Ct, lo.half_l2_norm_squared(Ct),
        int(golden["step_n_copies"]))
    assert_array_equal(labels, np.repeat(np.arange(8), 16))
    assert_allclose(sizes.sum(axis=0), np.full(8, 16), rtol=res(dt))
    # closed form of tests/test_kernels.py:944-951
    exp = np.reshape(Xt, (3, 8, 16)).sum(axis=-1)
    assert_allclose(copies.sum(axis=0), exp, rtol=res(dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_sort_vs_numpy_stable(dt):
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 100, 5000):
        x = rng.integers(-20, 20, size=n).astype(dt) / dt(4)
        assert_array_equal(lo.iterative_merge_sort(x), np.sort(x, kind="stable"))
    # signed zeros keep their input order (they compare equal under std::less)
    x = np.array([0.0, -0.0, 1.0, -0.0, 0.0, -1.0], dtype=dt)
    out = lo.iterative_merge_sort(x)
    assert_array_equal(np.signbit(out), [True, False, True, True, False, False])
    # NaN after every number
    x = np.array([np.nan, 1.0, -np.inf, np.inf, np.nan, 0.5], dtype=dt)
    out = lo.iterative_merge_sort(x)
    assert_array_equal(out[:4], [-np.inf, 0.5, 1.0, np.inf])
    assert np.isnan(out[4:]).all()


@pytest.mark.parametrize("dt", DTYPES)
def test_driver_kat(golden, dt):
    X = golden[f"drv_X_{'f32' if dt == np.float32 else 'f64'}"].astype(dt)
    Xt = np.ascontiguousarray(X.T)
    Ct = np.ascontiguousarray(CORNERS.T.astype(dt))
    w = np.ones(X.shape[0], dt)
    exp = np.repeat(np.arange(8), 32)
    tol, mi = float(golden["drv_tol"]), int(golden["drv_max_iter"])
    for labels, C, it, inertia in (lo.kmeans_lloyd(Xt, w, Ct, tol, mi, dtype=dt),
                                   co.lloyd(Xt, w, Ct, tol, mi)):
        assert_array_equal(labels, exp)
        assert it == int(golden["drv_expected_n_iter"])
        # closed form: the final centroids are the cloud means
        assert_allclose(C, Xt.reshape(3, 8, 32).mean(axis=-1), rtol=4 * res(dt))


def test_driver_vs_sklearn():
    """Independent cross-check: scikit-learn's Lloyd from the same init."""
    from sklearn.cluster import KMeans
    Xt, init = lo.gaussian_blobs(4000, 6, 12, seed=11, dtype=np.float64)
    w = np.ones(Xt.shape[1])
    labels, C, it, inertia = lo.kmeans_lloyd(Xt, w, init, 0.0, 300, dtype=np.float64)
    km = KMeans(n_clusters=12, init=init.T, n_init=1, algorithm="lloyd", tol=0.0,
                max_iter=300).fit(Xt.T)
    assert_array_equal(labels, km.labels_)
    assert_allclose(C.T, km.cluster_centers_, rtol=1e-9, atol=1e-9)
    assert_allclose(inertia, km.inertia_, rtol=1e-9)
    lc, Cc, itc, inc = co.lloyd(Xt, w, init, 0.0, 300)
    assert_array_equal(lc, km.labels_)
    assert_allclose(Cc.T, km.cluster_centers_, rtol=1e-9, atol=1e-9)


def test_driver_empty_cluster_relocation_two_empties():
    """n_empty >= 2: the j-th empty cluster (ascending) takes the j-th farthest
    sample (sq-dist descending, index ascending); checked by brute force."""
    X = np.array([[0.0, 0.1, 0.2, 5.0, 5.2, 9.0, -7.0]], dtype=np.float64)
    init = np.array([[0.1, 100.0, 200.0, 5.1]])   # clusters 1, 2 are empty
    w = np.ones(7)
    labels0 = lo.assignment(X, init, lo.half_l2_norm_squared(init))
    sq = (X[0] - init[0, labels0]) ** 2
    far = np.argsort(-sq, kind="stable")[:2]
    assert list(far) == [6, 5]        # -7.0 (d^2 = 50.41) then 9.0 (d^2 = 15.21)
    labels, C, it, _ = lo.kmeans_lloyd(X, w, init, 0.0, 1)
    # after one iteration: cluster 1 holds sample 6 alone, cluster 2 sample 5
    assert_allclose(C[0, 1], -7.0)
    assert_allclose(C[0, 2], 9.0)
    lc, Cc, itc, _ = co.lloyd(X, w, init, 0.0, 1)
    assert_allclose(Cc, C, rtol=1e-12)
    assert it == itc == 1


def test_c_restatement_matches_numpy_oracle_c0():
    Xt, init = lo.gaussian_blobs(50000, 14, 127, seed=0)
    w = np.ones(Xt.shape[1], np.float32)
    hn = lo.half_l2_norm_squared(init)
    assert_array_equal(co.assign(Xt, init, hn), lo.assignment(Xt, init, hn))
    lc, Cc, itc, inc = co.lloyd(Xt, w, init, 1e-4, 100)
    ln, Cn, itn, inn = lo.kmeans_lloyd(Xt, w, init, 1e-4, 100, dtype=np.float64)
    assert itc == itn
    assert (lc == ln).mean() > 0.9999
    assert_allclose(Cc, Cn, rtol=1e-4, atol=1e-4)
    assert_allclose(inc, inn, rtol=1e-5)


def test_number_of_private_copies_formula():
    # 4 MiB cache, occupancy 0.5, float32, k=256, d=32: 2 MiB / (4*256*33)
    assert lo.compute_number_of_private_copies(4 << 20, 4, 10**7, 32, 256, 0.5, 256) == 62
    # capped by the number of 8-sample groups, and never 0
    assert lo.compute_number_of_private_copies(4 << 20, 4, 20, 32, 256, 0.5, 256) == 32
    assert lo.compute_number_of_private_copies(1024, 8, 10**6, 128, 1024, 0.1, 256) == 1


# ---- planted mistakes: each is planted IN THE ORACLE (monkeypatched on the
# CPU) and must make the pins above fail ---------------------------------

def _pins_fail(fn, *args):
    try:
        fn(*args)
    except AssertionError:
        return True
    return False


def test_planted_sign_error_is_caught(golden, monkeypatch):
    good = lo.dot_products
    monkeypatch.setattr(lo, "dot_products", lambda X_t, C_t: -good(X_t, C_t))
    assert _pins_fail(test_assignment_kat_and_brute_force, golden, np.float64)


def test_planted_transposed_operand_is_caught(golden, monkeypatch):
    good = lo.reduce_centroids_data

    def transposed(sizes_copies, cent_copies, lst, n_empty):
        # each copy's (d, k) table read as if it were stored (k, d)
        cc = np.asarray(cent_copies)
        c, d, k = cc.shape
        return good(sizes_copies, cc.reshape(c, k, d).transpose(0, 2, 1), lst, n_empty)
    monkeypatch.setattr(lo, "reduce_centroids_data", transposed)
    assert _pins_fail(test_reduce_centroids_kat, golden, np.float64)


def test_planted_transposed_dot_is_caught(golden, monkeypatch):
    def transposed(X_t, C_t):
        # features of the centroid table taken in reverse order
        return np.asarray(C_t)[::-1].T @ np.asarray(X_t)
    monkeypatch.setattr(lo, "dot_products", transposed)
    assert _pins_fail(test_assignment_kat_and_brute_force, golden, np.float64)


def test_planted_dropped_term_is_caught(golden, monkeypatch):
    good = lo.compute_distances
    monkeypatch.setattr(lo, "compute_distances",
                        lambda X_t, C_t: good(np.asarray(X_t)[:-1], np.asarray(C_t)[:-1]))
    assert _pins_fail(test_distances_vs_cdist, golden, np.float64)


def test_planted_missing_half_is_caught(golden, monkeypatch):
    good = lo.half_l2_norm_squared
    monkeypatch.setattr(lo, "half_l2_norm_squared", lambda C_t: 2 * good(C_t))
    assert _pins_fail(test_half_l2_norm_kat, golden, np.float64)


def test_planted_weight_before_sum_dropped_is_caught(golden, monkeypatch):
    good = lo.relocate_empty_clusters

    def no_weight(n_empty, X_t, w, *rest):
        return good(n_empty, X_t, np.ones_like(w) * 2, *rest)
    monkeypatch.setattr(lo, "relocate_empty_clusters", no_weight)
    assert _pins_fail(test_relocate_kat, golden, np.float64)


def test_planted_tie_rule_is_caught(monkeypatch):
    def last_index_wins(pseudo):
        pseudo = np.asarray(pseudo)
        k = pseudo.shape[0]
        return (k - 1 - np.argmin(pseudo[::-1], axis=0)).astype(np.int64)
    monkeypatch.setattr(lo, "assignment_from_pseudo", last_index_wins)
    assert _pins_fail(test_assignment_tie_and_nan_rules)
