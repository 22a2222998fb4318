"""GPU parity tests of the entry points (pybind11 -> C ABI -> HIP kernels),
ported from the reference's tests/test_kernels.py known answers and extended
with oracle parity at larger shapes, int64 labels, ragged n, multi-window
d/k, ties and the reference's broken cases (distance window swap, n_empty >= 2).

Tolerances: rtol = finfo(dtype).resolution as in the reference tests unless a
test states otherwise; integer outputs (labels, counts, indices, sorted
values) are compared bit-exactly."""
import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal
from scipy.spatial.distance import cdist

from oracle import c_oracle as co
from oracle import lloyd_oracle as lo
from tests.conftest import CORNERS, check_label_mismatches, scal, to_dev, to_np

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FDT = [np.float32, np.float64]
IDT = [np.int32, np.int64]


def res(dt):
    return np.finfo(dt).resolution


def q():
    return torch.cuda.current_stream()


def sync():
    torch.cuda.synchronize()


@pytest.mark.parametrize("dt", FDT)
def test_broadcast_divide(kd, golden, torch_dev, dt):
    X = to_dev(golden["bdiv_X"], dt)
    y = to_dev(golden["bdiv_y"], dt)
    ht, _ = kd.broadcast_divide(X, y, work_group_size=32, sycl_queue=q())
    ht.wait()
    assert_allclose(to_np(X), golden["bdiv_expected"], rtol=res(dt))


@pytest.mark.parametrize("dt", FDT)
def test_half_l2_norm_squared(kd, golden, torch_dev, dt):
    X = to_dev(golden["hn_X"], dt)
    y = torch.empty(X.shape[1], dtype=X.dtype, device=torch_dev)
    ht, _ = kd.half_l2_norm_squared(X, y, work_group_size=32, sycl_queue=q())
    ht.wait()
    assert_allclose(to_np(y), golden["hn_expected"], rtol=res(dt))
    # bit-exact against the C restatement on a larger matrix
    C = np.random.default_rng(1).standard_normal((37, 300)).astype(dt)
    y2 = torch.empty(300, dtype=X.dtype, device=torch_dev)
    kd.half_l2_norm_squared(to_dev(C), y2, 256, q())[0].wait()
    assert_array_equal(to_np(y2), co.half_norm(C))


@pytest.mark.parametrize("dt", FDT)
@pytest.mark.parametrize("it", IDT)
def test_reduce_centroids_data(kd, golden, torch_dev, dt, it):
    canary = int(golden["rcd_canary"])
    for pre, exp_n in (("rcd", 0), ("rcde", 1)):
        sc = golden[f"{pre}_sizes_copies"].astype(dt)
        cc = golden[f"{pre}_cent_copies"].astype(dt)
        os_ = torch.empty(4, dtype=to_dev(sc).dtype, device=torch_dev)
        oc = torch.empty((2, 4), dtype=os_.dtype, device=torch_dev)
        lst = to_dev(np.full(4, canary, dtype=it))
        ne = to_dev(np.zeros(1, dtype=it))
        ht, ev = kd.reduce_centroids_data(to_dev(sc), to_dev(cc), os_, oc, lst, ne,
                                          work_group_size=128, sycl_queue=q())
        ht.wait()
        assert int(to_np(ne)[0]) == exp_n
        assert_allclose(to_np(oc), cc.astype(np.float64).sum(axis=0), rtol=res(dt))
        assert_allclose(to_np(os_), sc.sum(axis=0), rtol=res(dt))
        if pre == "rcd":
            assert_array_equal(to_np(lst), np.full(4, canary))
        else:
            assert_array_equal(to_np(lst), golden["rcde_expected_list"])


def test_reduce_centroids_many_empty_accumulates(kd, torch_dev):
    """Counter starts at the caller's value; empties listed in ascending order;
    large k spans several chunks of the list kernel."""
    rng = np.random.default_rng(5)
    n_copies, d, k = 5, 3, 3000
    sc = rng.integers(0, 3, size=(n_copies, k)).astype(np.float32)
    sc[:, rng.choice(k, 700, replace=False)] = 0
    cc = rng.standard_normal((n_copies, d, k)).astype(np.float32)
    sizes, cent, lst, n_empty = lo.reduce_centroids_data(sc, cc, np.full(k, -1), 2)
    os_ = torch.empty(k, device=torch_dev)
    oc = torch.empty((d, k), device=torch_dev)
    lst_d = to_dev(np.full(k, -1, dtype=np.int32))
    ne = to_dev(np.array([2], dtype=np.int32))
    kd.reduce_centroids_data(to_dev(sc), to_dev(cc), os_, oc, lst_d, ne, 256, q())[0].wait()
    assert int(to_np(ne)[0]) == n_empty
    assert_array_equal(to_np(lst_d), lst)
    assert_array_equal(to_np(os_), sizes)     # integer-valued sums: exact
    assert_allclose(to_np(oc), cent, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dt", FDT)
def test_compute_threshold(kd, golden, torch_dev, dt):
    data = golden["thr_data"].astype(dt)
    n = data.shape[0]
    d = to_dev(data)
    for topk, exp in ((1, max(data)), (2, np.partition(data, n - 2)[n - 2]),
                      (17, np.sort(data)[n - 17])):
        thr = to_dev(np.full((), -127, dtype=dt))
        ht, _ = kd.compute_threshold(d, topk, thr, sycl_queue=q())
        ht.wait()
        assert float(scal(thr)) == float(exp)
    thr = to_dev(np.full((), -127, dtype=dt))
    ht, ev = kd.compute_threshold(d, 0, thr, sycl_queue=q())   # no-op, null events
    ht.wait()
    assert float(scal(thr)) == -127


@pytest.mark.parametrize("dt", FDT)
@pytest.mark.parametrize("n", [2, 3, 2049, 100_003, 1_000_000])
def test_compute_threshold_is_the_stable_sort_element(kd, torch_dev, dt, n):
    """topk >= 2: the element at n - topk of the stable ascending sort (NaN
    last, -0 and +0 equal and kept in input order), bit for bit -- found by
    radix select, checked against np.sort(kind="stable") at positions in the
    middle, at both ends of the run of zeros and inside the NaNs."""
    rng = np.random.default_rng(n + 7)
    x = (rng.integers(-50, 50, size=n) / 4).astype(dt)          # heavy ties
    z = rng.choice(n, min(n, 40), replace=False)
    x[z] = np.where(rng.random(len(z)) < 0.5, -0.0, 0.0).astype(dt)
    if n > 10:
        x[rng.choice(n, 3, replace=False)] = np.array([np.inf, -np.inf, np.nan], dtype=dt)
        nan2 = np.array([np.nan], dtype=dt)
        nan2.view(np.uint32 if dt == np.float32 else np.uint64)[0] ^= 1  # another payload
        x[rng.choice(n, 1)] = nan2[0]
    srt = np.sort(x, kind="stable")
    ut = np.uint32 if dt == np.float32 else np.uint64
    zpos = np.nonzero(srt == 0)[0]
    pos = {0, n - 2, n // 2, int(zpos[0]), int(zpos[-1]), int(zpos[len(zpos) // 2])}
    if n > 10:
        pos |= {n - 2, n - 3}                                      # the NaNs / +inf
    xd = to_dev(x)
    for p in sorted(q for q in pos if 0 <= q <= n - 2):
        thr = to_dev(np.full((), -127, dtype=dt))
        kd.compute_threshold(xd, n - p, thr, sycl_queue=q())[0].wait()
        got = to_np(thr).reshape(()).view(ut)
        assert got == srt[p].view(ut), (n, p, srt[p])


@pytest.mark.parametrize("dt", FDT)
@pytest.mark.parametrize("n", [1, 5, 2047, 2048, 2049, 4096 + 3, 100_000, 1_000_003])
def test_iterative_merge_sort_bit_exact(kd, torch_dev, dt, n):
    rng = np.random.default_rng(n)
    x = (rng.integers(-1000, 1000, size=n) / 8).astype(dt)       # many ties
    if n > 10:
        x[rng.choice(n, 5, replace=False)] = np.array([np.inf, -np.inf, -0.0, 0.0, -0.0], dtype=dt)
    out = torch.empty_like(to_dev(x))
    kd.iterative_merge_sort(to_dev(x), out, q())[0].wait()
    got = to_np(out)
    exp = np.sort(x, kind="stable")
    assert_array_equal(got.view(np.uint32 if dt == np.float32 else np.uint64),
                       lo.iterative_merge_sort(x).view(np.uint32 if dt == np.float32 else np.uint64))
    assert_array_equal(got, exp)


def test_merge_sort_nan_last(kd, torch_dev):
    x = np.random.default_rng(0).standard_normal(5000).astype(np.float32)
    x[[3, 700, 4999]] = np.nan
    out = torch.empty_like(to_dev(x))
    kd.iterative_merge_sort(to_dev(x), out, q())[0].wait()
    got = to_np(out)
    assert np.isnan(got[-3:]).all()
    assert_array_equal(got[:-3], np.sort(x[~np.isnan(x)]))


@pytest.mark.parametrize("dt", FDT)
@pytest.mark.parametrize("it", IDT)
def test_select_samples_far_from_centroid(kd, golden, torch_dev, dt, it):
    dist_np = golden["sel_dist"].astype(dt)
    n, m = dist_np.shape[0], int(golden["sel_n_empty"])
    dist = to_dev(dist_np)
    thr = to_dev(np.empty((), dtype=dt))
    sel = to_dev(np.full(n, 172, dtype=it))
    ngt = to_dev(np.zeros((), dtype=it))
    neq = to_dev(np.ones((), dtype=it))
    ht_ev, c_ev = kd.compute_threshold(dist, m, thr, sycl_queue=q())
    ht_ev2, _ = kd.select_samples_far_from_centroid(m, dist, thr, sel, ngt, neq,
                                                    work_group_size=256, sycl_queue=q(),
                                                    depends=[c_ev])
    ht_ev2.wait()
    ht_ev.wait()
    g, e, t, s = int(scal(ngt)), int(scal(neq)), float(scal(thr)), to_np(sel)
    # tests/test_kernels.py:360-362
    assert g + e - 1 == m
    assert np.all(dist_np[s[:g]] > t)
    assert np.all(dist_np[s[1 - e:]] == t)
    # deterministic layout: exactly the oracle's
    so, go, eo = lo.select_samples_far_from_centroid(m, dist_np, dist_np.dtype.type(t),
                                                     np.full(n, 172), 0, 1)
    assert (g, e) == (go, eo)
    assert_array_equal(s, so)


def test_select_with_ties_at_threshold(kd, torch_dev):
    dist_np = np.array([1, 5, 3, 5, 5, 2, 5, 0], dtype=np.float32)
    n, m = 8, 3
    thr = to_dev(np.array(5.0, dtype=np.float32))
    sel = to_dev(np.full(n, -1, dtype=np.int32))
    ngt, neq = to_dev(np.zeros((), np.int32)), to_dev(np.ones((), np.int32))
    kd.select_samples_far_from_centroid(m, to_dev(dist_np), thr, sel, ngt, neq, 64, q())[0].wait()
    assert int(scal(ngt)) == 0 and int(scal(neq)) == 4
    # eq indices 1, 3, 4 (ascending) land at n-1, n-2, n-3
    assert_array_equal(to_np(sel)[-3:], [4, 3, 1])
    so, go, eo = lo.select_samples_far_from_centroid(m, dist_np, np.float32(5.0),
                                                     np.full(n, -1), 0, 1)
    assert_array_equal(to_np(sel), so)


@pytest.mark.parametrize("dt", FDT)
@pytest.mark.parametrize("it", IDT)
def test_relocate_empty_clusters(kd, golden, torch_dev, dt, it):
    X = golden["rel_X"].astype(dt)
    C = golden["rel_C"].astype(dt)
    sq = np.min(np.square(X[:, None, :] - C[None, :, :]).sum(axis=-1), axis=-1).astype(dt)
    Xt = to_dev(np.ascontiguousarray(X.T))
    Ct = to_dev(np.ascontiguousarray(C.T))
    w = to_dev(np.ones(8, dtype=dt))
    lst = to_dev(golden["rel_empty_list"].astype(it))
    sq_d = to_dev(sq)
    sizes = to_dev(golden["rel_sizes"].astype(dt))
    lab = to_dev(np.zeros(8, dtype=it))
    ht, _ = kd.relocate_empty_clusters(1, Xt, w, lab, lst, sq_d, Ct, sizes, sq_d,
                                       work_group_size=256, sycl_queue=q())
    ht.wait()
    exp_inert = sq.copy()
    exp_inert[int(golden["rel_moved_sample"])] = 0
    assert_allclose(to_np(sizes), golden["rel_expected_sizes"], rtol=res(dt))
    assert_allclose(to_np(Ct), golden["rel_expected_centroid_t"], rtol=res(dt))
    assert_allclose(to_np(sq_d), exp_inert, rtol=res(dt))


def test_relocate_three_empties_matches_oracle(kd, torch_dev):
    rng = np.random.default_rng(9)
    n, d, k = 5000, 5, 10
    Xt = rng.standard_normal((d, n)).astype(np.float32)
    lab = rng.integers(0, 7, size=n)          # clusters 7, 8, 9 empty
    w = rng.uniform(0.5, 2.0, size=n).astype(np.float32)
    C = rng.standard_normal((d, k)).astype(np.float32)
    sq = lo.compute_inertia(Xt, None, C, lab)
    sq[[10, 20]] = sq.max()                   # a tie at the top: index order decides
    sums = np.stack([np.bincount(lab, Xt[f] * w, minlength=k) for f in range(d)]).astype(np.float32)
    sizes = np.bincount(lab, w, minlength=k).astype(np.float32)
    lst = np.array([7, 8, 9, 0, 0, 0, 0, 0, 0, 0])
    Ce, se, ie = lo.relocate_empty_clusters(3, Xt, w, lab, lst, sq, sums, sizes, sq)
    Cd, sd, idd = to_dev(sums), to_dev(sizes), to_dev(sq)
    kd.relocate_empty_clusters(3, to_dev(Xt), to_dev(w), to_dev(lab.astype(np.int64)),
                               to_dev(lst.astype(np.int64)), to_dev(sq), Cd, sd, idd, 256,
                               q())[0].wait()
    assert_allclose(to_np(Cd), Ce, rtol=1e-6, atol=1e-5)
    assert_allclose(to_np(sd), se, rtol=1e-6)
    assert_array_equal(to_np(idd), ie)
    assert_array_equal(np.flatnonzero(to_np(idd) == 0), np.flatnonzero(ie == 0))


@pytest.mark.parametrize("dt", FDT)
def test_centroid_shifts(kd, golden, torch_dev, dt):
    a, b = to_dev(golden["shift_a"], dt), to_dev(golden["shift_b"], dt)
    cs = torch.empty(2, dtype=a.dtype, device=torch_dev)
    ht, _ = kd.compute_centroid_shifts_squared(a, b, cs, sycl_queue=q())
    ht.wait()
    assert_allclose(to_np(cs), golden["shift_expected"], rtol=res(dt))


@pytest.mark.parametrize("dt", FDT)
def test_distances(kd, golden, torch_dev, dt):
    X = golden["dist_X"].astype(dt)
    Xt = to_dev(np.ascontiguousarray(X.T))
    Ct = to_dev(np.ascontiguousarray(CORNERS.T), dt)
    dm = torch.empty((8, X.shape[0]), dtype=Xt.dtype, device=torch_dev)
    ht, _ = kd.compute_centroid_to_sample_distances(Xt, Ct, dm, 256, 8, sycl_queue=q())
    ht.wait()
    f = 64 if dt == np.float32 else 1          # tests/test_kernels.py:598,638
    assert_allclose(to_np(dm), cdist(CORNERS, X), rtol=f * res(dt))


@pytest.mark.parametrize("shape", [(35, 70, 1000), (1, 1, 3), (130, 33, 513), (16, 256, 4099)])
def test_distances_all_shapes(kd, torch_dev, shape):
    """The reference is only right when ceil(d/h) == ceil(k/32); this build is
    right for every shape."""
    d, k, n = shape
    rng = np.random.default_rng(d * k)
    X = rng.standard_normal((n, d)).astype(np.float32)
    C = rng.standard_normal((k, d)).astype(np.float32)
    dm = torch.empty((k, n), device=torch_dev)
    kd.compute_centroid_to_sample_distances(to_dev(np.ascontiguousarray(X.T)),
                                            to_dev(np.ascontiguousarray(C.T)), dm, 128, 8,
                                            q())[0].wait()
    assert_allclose(to_np(dm), cdist(C.astype(np.float64), X.astype(np.float64)),
                    rtol=64 * res(np.float32), atol=1e-5)


@pytest.mark.parametrize("dt,key,cloud", [(np.float32, "assign_X_1734", 1734),
                                          (np.float64, "assign_X_32", 32)])
@pytest.mark.parametrize("it", IDT)
def test_assignment(kd, golden, torch_dev, dt, key, cloud, it):
    X = golden[key].astype(dt)
    Xt = to_dev(np.ascontiguousarray(X.T))
    Ct = to_dev(np.ascontiguousarray(CORNERS.T), dt)
    hl2n = torch.empty(8, dtype=Xt.dtype, device=torch_dev)
    ids = to_dev(np.empty(X.shape[0], dtype=it))
    ht1, e_hl2n = kd.half_l2_norm_squared(Ct, hl2n, work_group_size=256, sycl_queue=q())
    ht2, e_assign = kd.assignment(Xt, Ct, hl2n, ids, centroids_window_height=8,
                                  work_group_size=256, sycl_queue=q(), depends=[e_hl2n])
    ht1.wait()
    ht2.wait()
    assert_array_equal(to_np(ids), np.repeat(np.arange(8), cloud))


@pytest.mark.parametrize("dt", FDT)
@pytest.mark.parametrize("shape", [(50_000, 14, 127), (10_001, 3, 1), (7_777, 70, 33),
                                   (3_000, 32, 256), (20_000, 128, 300), (1, 5, 40)])
def test_assignment_matches_c_restatement(kd, torch_dev, dt, shape):
    """Labels of the GPU assignment equal the C restatement's, except samples
    whose FP64 best/second-best margin is within rounding (counted)."""
    n, d, k = shape
    Xt, init = lo.gaussian_blobs(n, d, k, seed=n + d + k, dtype=dt)
    hn = co.half_norm(init)
    exp = co.assign(Xt, init, hn)
    ids = to_dev(np.empty(n, dtype=np.int32))
    kd.assignment(to_dev(Xt), to_dev(init), to_dev(hn), ids, 8, 256, q())[0].wait()
    got = to_np(ids)
    diff = np.flatnonzero(got != exp)
    if diff.size:
        margin = lo.pseudo_margin(Xt[:, diff], init)
        scale = np.abs(hn).max() + np.abs(Xt[:, diff]).max() * np.abs(init).max() * d
        assert np.all(margin <= 64 * res(dt) * scale), (diff.size, margin.max())
    assert diff.size <= max(1, n // 10000)


def test_assignment_ties_pick_lowest_index(kd, torch_dev):
    C = np.array([[0.0, 1.0, 0.0, 1.0] * 20], dtype=np.float32)   # k=80 with duplicates
    X = np.array([[0.0, 1.0, 0.4, 0.6, np.nan]], dtype=np.float32)
    hn = lo.half_l2_norm_squared(C)
    ids = to_dev(np.full(5, -1, dtype=np.int64))
    kd.assignment(to_dev(X), to_dev(C), to_dev(hn), ids, 8, 256, q())[0].wait()
    assert_array_equal(to_np(ids), [0, 1, 0, 1, 0])


@pytest.mark.parametrize("dt", FDT)
def test_compute_inertia(kd, golden, torch_dev, dt):
    X = golden["inertia_X"].astype(dt)
    ids = np.repeat(np.arange(8), 32).astype(np.int32)
    out = torch.empty(X.shape[0], dtype=to_dev(X).dtype, device=torch_dev)
    ht, _ = kd.compute_inertia(to_dev(np.ascontiguousarray(X.T)), to_dev(np.ones(X.shape[0], dt)),
                               to_dev(np.ascontiguousarray(CORNERS.T), dt), to_dev(ids), out,
                               work_group_size=256, sycl_queue=q())
    ht.wait()
    exp = np.sum(np.square(X.T - np.take_along_axis(CORNERS.T, ids[None, :], axis=1)), axis=0)
    assert_allclose(to_np(out), exp, rtol=res(dt))


@pytest.mark.parametrize("n,d,k,dt", [
    (50_000, 32, 256, np.float32),    # centroids staged in LDS (33 KB)
    (50_000, 32, 256, np.float64),    # 67.6 KB: read from global memory
    (20_000, 64, 300, np.float32),    # 78 KB: global memory
    (4_099, 3, 4_000, np.float32),    # many narrow centroids: global memory
    (777, 7, 9, np.float64),          # LDS, ragged n
])
def test_compute_inertia_both_centroid_paths(kd, torch_dev, n, d, k, dt):
    """Weighted per-sample inertia w_s * sum_f (x - c_l)^2 at sizes on both
    sides of the 64 KB LDS threshold for the centroid table."""
    rng = np.random.default_rng(n + d + k)
    X = rng.normal(size=(n, d)).astype(dt) * 3
    C = rng.normal(size=(k, d)).astype(dt) * 3
    w = rng.uniform(0.5, 2.0, n).astype(dt)
    ids = rng.integers(0, k, n).astype(np.int32)
    out = torch.empty(n, dtype=to_dev(X).dtype, device=torch_dev)
    ht, _ = kd.compute_inertia(to_dev(np.ascontiguousarray(X.T)), to_dev(w),
                               to_dev(np.ascontiguousarray(C.T)), to_dev(ids), out,
                               work_group_size=256, sycl_queue=q())
    ht.wait()
    Xd, Cd = X.astype(np.float64), C.astype(np.float64)
    exp = np.sum(np.square(Xd - Cd[ids]), axis=1) * w.astype(np.float64)
    assert_allclose(to_np(out), exp, rtol=1e-5 if dt == np.float32 else 1e-12)


@pytest.mark.parametrize("dt", FDT)
def test_reduce_vector_blocking(kd, golden, torch_dev, dt):
    vec = to_dev(golden["rv_data"], dt)
    assert kd.reduce_vector_blocking(vec, sycl_queue=q()) == 50 * 99
    big = np.random.default_rng(2).standard_normal(3_000_001).astype(dt)
    got = kd.reduce_vector_blocking(to_dev(big), q())
    assert_allclose(got, np.sum(big.astype(np.float64)), rtol=1e-5 if dt == np.float32 else 1e-12)
    assert kd.reduce_vector_blocking(to_dev(np.zeros(0, dt)), q()) == 0.0


@pytest.mark.parametrize("dt", FDT)
@pytest.mark.parametrize("it", IDT)
def test_fused_lloyd_single_step(kd, golden, torch_dev, dt, it):
    X = golden["step_X"].astype(dt)
    Xt_np = np.ascontiguousarray(X.T)
    n_features, n_samples = Xt_np.shape
    Ct = np.ascontiguousarray(CORNERS.T.astype(dt))
    hn = to_dev(np.sum(np.square(Ct), axis=0) / 2)
    ids = to_dev(np.empty(n_samples, dtype=it))
    n_copies = int(golden["step_n_copies"])
    cp = torch.zeros((n_copies, n_features, 8), dtype=hn.dtype, device=torch_dev)
    scp = torch.zeros((n_copies, 8), dtype=hn.dtype, device=torch_dev)
    ht, _ = kd.fused_lloyd_single_step(to_dev(Xt_np), to_dev(np.ones(n_samples, dt)), to_dev(Ct),
                                       hn, ids, cp, scp, 8, 256, q())
    ht.wait()
    assert_array_equal(to_np(ids), np.repeat(np.arange(8), 16))
    assert_allclose(to_np(scp).sum(axis=0), np.full(8, 16), rtol=res(dt))
    exp = np.reshape(Xt_np, (n_features, 8, 16)).sum(axis=-1)
    assert_allclose(to_np(cp).sum(axis=0), exp, rtol=res(dt))


@pytest.mark.parametrize("dt", FDT)
@pytest.mark.parametrize("shape,n_copies", [((50_000, 14, 127), 7), ((100_003, 32, 256), 16),
                                            ((20_000, 100, 200), 3), ((4_096, 7, 1), 1)])
def test_fused_step_matches_oracle(kd, torch_dev, dt, shape, n_copies):
    """Accumulate-into semantics (copies start non-zero), bit-exact unit-weight
    counts, sums within rtol 1e-5 of an FP64 sum over the GPU labels."""
    n, d, k = shape
    Xt, init = lo.gaussian_blobs(n, d, k, seed=3 * n + k, dtype=dt)
    hn = co.half_norm(init)
    w = np.ones(n, dt)
    ids = to_dev(np.empty(n, dtype=np.int32))
    base = np.full((n_copies, d, k), 0.5, dtype=dt)
    cp = to_dev(base)
    scp = to_dev(np.full((n_copies, k), 2.0, dtype=dt))
    kd.fused_lloyd_single_step(to_dev(Xt), to_dev(w), to_dev(init), to_dev(hn), ids, cp, scp,
                               8, 256, q())[0].wait()
    lab = to_np(ids).astype(np.int64)
    check_label_mismatches(Xt, init, lab, co.assign(Xt, init, hn), limit=max(1, n // 10000))
    sizes = to_np(scp).sum(axis=0) - 2.0 * n_copies
    assert_array_equal(sizes, np.bincount(lab, minlength=k))
    sums = to_np(cp).astype(np.float64).sum(axis=0) - 0.5 * n_copies
    exp_sums = np.stack([np.bincount(lab, Xt[f].astype(np.float64), minlength=k) for f in range(d)])
    scale = np.abs(exp_sums).max()
    assert np.abs(sums - exp_sums).max() <= 1e-5 * scale


def test_fused_step_values_outside_the_sum_window(kd, torch_dev):
    """The float32 fused step sums exactly in binade windows set from the
    centroids and adds values outside them (outliers far above every
    centroid coordinate, tiny values, infinities and NaNs) with FP32 atomics
    as the reference does: every cluster's sums stay within FP32 rounding of
    an FP64 sum over the GPU labels, and a NaN reaches its cluster's sum."""
    n, d, k = 60_000, 8, 16
    Xt, init = lo.gaussian_blobs(n, d, k, seed=11, dtype=np.float32)
    rng = np.random.default_rng(5)
    Xt[0, rng.choice(n, 40, replace=False)] = 3e6        # far above the window
    Xt[1, rng.choice(n, 300, replace=False)] = 1e-12      # below it
    Xt[2, rng.choice(n, 300, replace=False)] *= 1e-7      # partly below it
    Xt[5] = rng.uniform(-1e-12, 1e-12, n)                 # a whole feature below it
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    w[rng.choice(n, 20, replace=False)] = 3e7             # weights beyond the window
    hn = co.half_norm(init)
    n_copies = 4
    ids = to_dev(np.empty(n, dtype=np.int32))
    cp = to_dev(np.zeros((n_copies, d, k), np.float32))
    scp = to_dev(np.zeros((n_copies, k), np.float32))
    kd.fused_lloyd_single_step(to_dev(Xt), to_dev(w), to_dev(init), to_dev(hn), ids, cp, scp,
                               8, 128, q())[0].wait()
    lab = to_np(ids).astype(np.int64)
    ref_ids = to_dev(np.empty(n, dtype=np.int32))
    kd.assignment(to_dev(Xt), to_dev(init), to_dev(hn), ref_ids, 8, 128, q())
    sync()
    assert_array_equal(lab, to_np(ref_ids))
    sums = to_np(cp).astype(np.float64).sum(axis=0)
    sizes = to_np(scp).astype(np.float64).sum(axis=0)
    wx = Xt.astype(np.float64) * w.astype(np.float64)
    exp = np.stack([np.bincount(lab, wx[f], minlength=k) for f in range(d)])
    mag = np.stack([np.bincount(lab, np.abs(wx[f]), minlength=k) for f in range(d)])
    assert np.all(np.abs(sums - exp) <= 1e-5 * mag + 1e-30)
    exp_w = np.bincount(lab, w.astype(np.float64), minlength=k)
    assert np.all(np.abs(sizes - exp_w) <= 1e-5 * exp_w)
    # non-finite values reach their cluster's sums
    Xt2 = Xt.copy()
    Xt2[3, 7] = np.nan
    Xt2[4, 9] = np.inf
    cp2 = to_dev(np.zeros((n_copies, d, k), np.float32))
    scp2 = to_dev(np.zeros((n_copies, k), np.float32))
    ids2 = to_dev(np.empty(n, dtype=np.int32))
    kd.fused_lloyd_single_step(to_dev(Xt2), to_dev(w), to_dev(init), to_dev(hn), ids2, cp2,
                               scp2, 8, 128, q())[0].wait()
    lab2 = to_np(ids2)
    s2 = to_np(cp2).astype(np.float64).sum(axis=0)
    assert np.isnan(s2[3, lab2[7]])
    assert np.isinf(s2[4, lab2[9]]) or np.isnan(s2[4, lab2[9]])


def test_fused_step_padded_rows_with_infinite_weight(kd, torch_dev):
    """d = 5: the few-centroid kernel holds the features in rows padded to 8
    (x = 0 there).  A sample of infinite weight makes its own cluster's sums
    and size non-finite, as the reference's FP32 atomics would, and no other
    cluster's: the padded rows' 0 x inf has no element of the caller's
    copies to go to."""
    n, d, k = 20_000, 5, 8
    Xt, init = lo.gaussian_blobs(n, d, k, seed=13, dtype=np.float32)
    w = np.ones(n, np.float32)
    w[123] = np.inf
    hn = co.half_norm(init)
    n_copies = 4
    ids = to_dev(np.empty(n, dtype=np.int32))
    cp = to_dev(np.zeros((n_copies, d, k), np.float32))
    scp = to_dev(np.zeros((n_copies, k), np.float32))
    kd.fused_lloyd_single_step(to_dev(Xt), to_dev(w), to_dev(init), to_dev(hn), ids, cp, scp,
                               8, 128, q())[0].wait()
    lab = to_np(ids).astype(np.int64)
    ref_ids = to_dev(np.empty(n, dtype=np.int32))
    kd.assignment(to_dev(Xt), to_dev(init), to_dev(hn), ref_ids, 8, 128, q())
    sync()
    assert_array_equal(lab, to_np(ref_ids))
    c = lab[123]
    sums = to_np(cp).astype(np.float64).sum(axis=0)
    sizes = to_np(scp).astype(np.float64).sum(axis=0)
    others = np.arange(k) != c
    assert np.isfinite(sums[:, others]).all() and np.isfinite(sizes[others]).all()
    assert np.isinf(sizes[c])
    assert (~np.isfinite(sums[:, c]) | (Xt[:, 123] == 0)).all()


def test_compute_number_of_private_copies(kd, torch_dev):
    arr = torch.empty(1, device=torch_dev)
    l2 = torch.cuda.get_device_properties(0).L2_cache_size
    for n, d, k, occ, wg in ((10**7, 32, 256, 0.7, 128), (100, 3, 8, 0.5, 128), (10**6, 512, 4096, 0.1, 256)):
        got = kd.compute_number_of_private_copies(arr, n, d, k, occ, wg)
        assert got == lo.compute_number_of_private_copies(l2, 4, n, d, k, occ, wg)
    with pytest.raises(ValueError, match="strictly between"):
        kd.compute_number_of_private_copies(arr, 10, 3, 4, 1.0, 32)
