"""GPU tests of kmeans_lloyd_driver (src/kmeans_lloyd_driver.hpp:21-431
semantics): the reference's known answer (n_iter == 2, exact labels),
parity with the C restatement and the FP64 oracle on the CPU-runnable
benchmark configuration (n=50,000, d=14, k=127), empty-cluster relocation
with one and several empty clusters, max_iter = 0, verbose output and
stream/depends ordering."""
import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

from oracle import c_oracle as co
from oracle import lloyd_oracle as lo
from tests.conftest import CORNERS, check_label_mismatches, fp64_update, to_dev, to_np

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def q():
    return torch.cuda.current_stream()


def run_driver(kd, Xt, w, init, tol, max_iter, itype=np.int32, verbose=False,
               h=8, wg=128, occ=0.7):
    dev = torch.device("cuda", 0)
    X_d, w_d = to_dev(Xt), to_dev(w)
    init_d = to_dev(init)
    res_d = torch.empty_like(init_d)
    lab_d = torch.empty(Xt.shape[1], dtype=torch.int32 if itype == np.int32 else torch.int64,
                        device=dev)
    n_iter, inertia = kd.kmeans_lloyd_driver(X_d, w_d, init_d, lab_d, res_d, tol, verbose,
                                             max_iter, h, wg, occ, sycl_queue=q())
    return to_np(lab_d).astype(np.int64), to_np(res_d), int(n_iter), inertia


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("it", [np.int32, np.int64])
def test_driver_kat(kd, golden, torch_dev, dt, it):
    X = golden[f"drv_X_{'f32' if dt == np.float32 else 'f64'}"].astype(dt)
    Xt = np.ascontiguousarray(X.T)
    Ct = np.ascontiguousarray(CORNERS.T.astype(dt))
    lab, C, n_iter, inertia = run_driver(kd, Xt, np.ones(X.shape[0], dt), Ct,
                                         float(golden["drv_tol"]), int(golden["drv_max_iter"]),
                                         itype=it)
    assert_array_equal(lab, np.repeat(np.arange(8), 32))
    assert n_iter < int(golden["drv_max_iter"])
    assert n_iter == int(golden["drv_expected_n_iter"])
    assert inertia.shape == (1,) and inertia.dtype == dt
    assert_allclose(C, Xt.reshape(3, 8, 32).mean(axis=-1), rtol=4 * np.finfo(dt).resolution)
    _, _, _, inert_o = lo.kmeans_lloyd(Xt, np.ones(X.shape[0]), Ct, 1e-6, 255, dtype=np.float64)
    assert_allclose(inertia[0], inert_o, rtol=1e-5)


def test_driver_c0_config_parity(kd, torch_dev):
    """n=50,000, d=14, k=127 blobs (std 1, centers in [-10,10]), random-sample
    init, max_iter=100, tol=1e-4."""
    Xt, init = lo.gaussian_blobs(50_000, 14, 127, seed=1000, center_box=10.0, std=1.0)
    w = np.ones(Xt.shape[1], np.float32)
    lab, C, n_iter, inertia = run_driver(kd, Xt, w, init, 1e-4, 100)
    lc, Cc, itc, inc = co.lloyd(Xt, w, init, 1e-4, 100)
    l64, C64, it64, in64 = lo.kmeans_lloyd(Xt, w, init, 1e-4, 100, dtype=np.float64)
    assert abs(n_iter - itc) <= 1
    assert (lab != lc).sum() <= 5
    scale = np.abs(C64).max()
    assert np.abs(C - C64).max() <= 1e-5 * scale
    assert_allclose(inertia[0], in64, rtol=1e-5)


def test_driver_relocates_empty_clusters(kd, torch_dev):
    """Init centroids far from the data leave clusters empty in the first
    iteration; the fit must follow the oracle's relocation semantics."""
    Xt, init = lo.gaussian_blobs(20_000, 8, 16, seed=77, dtype=np.float64)
    init = init.copy()
    init[:, 3] = 1e4
    init[:, 7] = -1e4
    init[:, 11] = 5e3
    w = np.linspace(0.5, 1.5, Xt.shape[1])
    lab, C, n_iter, inertia = run_driver(kd, Xt, w, init, 1e-9, 50)
    l64, C64, it64, in64 = lo.kmeans_lloyd(Xt, w, init, 1e-9, 50, dtype=np.float64)
    assert n_iter == it64
    assert_array_equal(lab, l64)
    assert_allclose(C, C64, rtol=1e-9, atol=1e-9)
    assert_allclose(inertia[0], in64, rtol=1e-9)


def test_driver_max_iter_zero(kd, torch_dev):
    Xt, init = lo.gaussian_blobs(3000, 5, 7, seed=5)
    w = np.ones(3000, np.float32)
    lab, C, n_iter, inertia = run_driver(kd, Xt, w, init, 0.0, 0)
    assert n_iter == 0
    assert_array_equal(C, init)
    assert_array_equal(lab, co.assign(Xt, init))
    exp = np.sum(lo.compute_inertia(Xt.astype(np.float64), w, init.astype(np.float64), lab))
    assert_allclose(inertia[0], exp, rtol=1e-5)


def test_driver_verbose_prints_each_iteration(kd, torch_dev, capfd):
    Xt, init = lo.gaussian_blobs(2000, 4, 5, seed=6)
    lab, C, n_iter, inertia = run_driver(kd, Xt, np.ones(2000, np.float32), init, 0.0, 3,
                                         verbose=True)
    out = capfd.readouterr().out
    assert out.count("Iteration:") == n_iter
    assert "Inertia:" in out


def test_driver_other_stream_and_depends(kd, torch_dev):
    """Work enqueued on a library stream after an event of torch's stream."""
    Xt, init = lo.gaussian_blobs(10_000, 6, 9, seed=8)
    s = kd.Stream()
    X_d = to_dev(Xt)
    C_d = to_dev(init)
    hn = torch.empty(9, device=torch_dev)
    ev_torch = kd.record_event(q())
    ht, ev = kd.half_l2_norm_squared(C_d, hn, 64, s, depends=[ev_torch])
    ids = torch.empty(10_000, dtype=torch.int32, device=torch_dev)
    ht2, ev2 = kd.assignment(X_d, C_d, hn, ids, 8, 256, s, depends=[ev])
    ht2.wait()
    ht.wait()
    assert ev2.is_complete
    assert_array_equal(to_np(ids), co.assign(Xt, init))
    s.synchronize()
    s.close()


def test_driver_ragged_and_multiwindow_shapes(kd, torch_dev):
    for n, d, k in ((1, 3, 1), (257, 33, 65), (4099, 70, 3), (12_345, 17, 300)):
        Xt, init = lo.gaussian_blobs(n, d, k if k <= n else n, seed=n, dtype=np.float64)
        kk = init.shape[1]
        w = np.ones(n)
        lab, C, n_iter, inertia = run_driver(kd, Xt, w, init, 1e-12, 20, itype=np.int64)
        l64, C64, it64, in64 = lo.kmeans_lloyd(Xt, w, init, 1e-12, 20, dtype=np.float64)
        assert n_iter == it64, (n, d, kk)
        assert_array_equal(lab, l64)
        assert_allclose(C, C64, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("shape", [(100_003, 32, 256), (60_000, 64, 100), (30_001, 16, 33),
                                   (10_000, 5, 7), (20_000, 100, 40)])
def test_float64_fit_matches_fp64_oracle(kd, torch_dev, shape):
    """Float64 fits (the FP64 matrix-core kernels where the centroids and
    FP64 sums fit LDS, d <= 64; the generic kernel beyond) against the FP64
    oracle: same labels and iteration count, centroids and inertia to a few
    FP64 ulps of accumulated rounding."""
    n, d, k = shape
    print(f"float64 fit n={n} d={d} k={k}", flush=True)
    Xt, init = lo.gaussian_blobs(n, d, k, seed=d * k, dtype=np.float64)
    w = np.random.default_rng(k).uniform(0.5, 2.0, n)
    lab, C, n_iter, inertia = run_driver(kd, Xt, w, init, 1e-10, 12)
    l64, C64, it64, in64 = lo.kmeans_lloyd(Xt, w, init, 1e-10, 12, dtype=np.float64)
    assert n_iter == it64
    assert_array_equal(lab, l64)
    assert np.abs(C - C64).max() <= 1e-12 * np.abs(C64).max()
    assert abs(inertia - in64) <= 1e-12 * abs(in64)


def separated_blobs(n, d, k, seed):
    """Blobs far apart relative to their spread, init near the true centers:
    no sample is within rounding of a tie, so FP32 and FP64 fits follow the
    same trajectory and can be compared after several iterations."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-10, 10, size=(k, d))
    blob = rng.integers(0, k, size=n)
    X = centers[blob] + 0.3 * rng.standard_normal((n, d))
    init = centers + 0.3 * rng.standard_normal((k, d))
    return (np.ascontiguousarray(X.T.astype(np.float32)),
            np.ascontiguousarray(init.T.astype(np.float32)))


def test_driver_fp32_is_deterministic(kd, torch_dev):
    """Centroid sums of the FP32 path are exact fixed-point integers, so two
    runs give bit-identical centroids, labels and inertia; centroids stay
    within 1e-5 relative of the FP64 oracle."""
    Xt, init = separated_blobs(200_003, 32, 64, seed=21)
    w = np.random.default_rng(4).uniform(0.5, 2.0, 200_003).astype(np.float32)
    r1 = run_driver(kd, Xt, w, init, 0.0, 7)
    r2 = run_driver(kd, Xt, w, init, 0.0, 7)
    assert_array_equal(r1[0], r2[0])
    assert_array_equal(r1[1].view(np.uint32), r2[1].view(np.uint32))
    assert r1[2] == r2[2]
    assert r1[3][0] == r2[3][0]
    l64, C64, it64, in64 = lo.kmeans_lloyd(Xt, w, init, 0.0, 7, dtype=np.float64)
    assert r1[2] == it64          # converges exactly (shift 0) like the oracle
    assert_array_equal(r1[0], l64)
    assert np.abs(r1[1] - C64).max() <= 1e-5 * np.abs(C64).max()
    assert_allclose(r1[3][0], in64, rtol=1e-5)


def test_driver_nonfinite_input_propagates(kd, torch_dev):
    """A NaN sample makes the fixed-point pre-pass fall back to FP32 sums;
    the NaN reaches its cluster's centroid as in the reference."""
    Xt, init = lo.gaussian_blobs(4096, 8, 4, seed=2)
    Xt = Xt.copy()
    Xt[3, 100] = np.nan
    lab, C, n_iter, inertia = run_driver(kd, Xt, np.ones(4096, np.float32), init, 0.0, 1)
    assert np.isnan(C).any()


def run_driver_ex(kd, Xt, w, init, tol, max_iter, precision, verbose=False):
    dev = torch.device("cuda", 0)
    X_d, w_d = to_dev(Xt), to_dev(w)
    init_d = to_dev(init)
    res_d = torch.empty_like(init_d)
    lab_d = torch.empty(Xt.shape[1], dtype=torch.int32, device=dev)
    n_iter, inertia, stats = kd.kmeans_lloyd_driver_ex(X_d, w_d, init_d, lab_d, res_d, tol,
                                                       verbose, max_iter, precision, q())
    return to_np(lab_d).astype(np.int64), to_np(res_d), int(n_iter), inertia, stats


@pytest.mark.parametrize("tiled", [False, True])
@pytest.mark.parametrize("prec", ["fp32", "bf16", "bf16x3"])
def test_relocation_with_next_iteration_enqueued_ahead(kd, torch_dev, monkeypatch, capfd,
                                                       prec, tiled):
    """FP32 data whose init leaves five clusters empty in the first iteration:
    the driver has already enqueued the second iteration when it learns of
    them, so that iteration's main kernel must skip itself and be redone after
    the relocation.  The fit equals the verbose one (which enqueues nothing
    ahead), is the same on every path and precision, and follows the FP64
    oracle."""
    n, d, k = 40_000, 16, 24
    Xt, init = lo.gaussian_blobs(n, d, k, seed=51, std=1.0)
    init = init.copy()
    init[:, 3] = 1e4
    init[:, 7] = -1e4
    init[:, 11:14] = 5e3
    w = np.ones(n, np.float32)
    if tiled:
        monkeypatch.setenv("KMD_FORCE_TILED", "1")
    elif prec != "fp32":
        # the LDS-resident mixed kernels (for k <= 32 every precision runs
        # the FP32 few-centroid kernel otherwise; the fp32 case covers it)
        monkeypatch.setenv("KMD_NO_SMALLK", "1")
    got = run_driver_ex(kd, Xt, w, init, 1e-7, 30, prec)
    assert got[4]["path"] == ("tiled-" if tiled else "") + (
        "fp32" + ("" if tiled else "-mfma") if prec == "fp32" else f"{prec}+rescore")
    seq = run_driver_ex(kd, Xt, w, init, 1e-7, 30, prec, verbose=True)
    capfd.readouterr()
    same_fit(got, seq)
    monkeypatch.delenv("KMD_FORCE_TILED", raising=False)
    ref = run_driver_ex(kd, Xt, w, init, 1e-7, 30, "fp32")
    same_fit(got, ref)
    # FP32 against FP64: a sample within rounding of a tie may go the other
    # way (then the trajectories can differ by an iteration)
    l64, C64, it64, in64 = lo.kmeans_lloyd(Xt, w, init, 1e-7, 30, dtype=np.float64)
    assert abs(got[2] - it64) <= 1
    assert (got[0] != l64).sum() <= 5
    if got[2] == it64:
        assert np.abs(got[1] - C64).max() <= 1e-5 * np.abs(C64).max()


@pytest.mark.parametrize("shape,seed,std", [((50_000, 14, 127), 1000, 1.0),
                                            ((300_001, 32, 256), 7, 2.0),
                                            ((100_000, 64, 100), 3, 1.5),
                                            ((65_537, 5, 33), 9, 3.0)])
@pytest.mark.parametrize("rescore", ["in-kernel", "candidate-tiles", "gather"])
def test_mixed_precision_matches_fp32_bit_exactly(kd, torch_dev, monkeypatch, shape, seed,
                                                  std, rescore):
    """BF16 (+ certified-margin FP32 re-score) and BF16x3 fits give the FP32
    path's labels; with exact fixed-point sums the centroids are then
    bit-identical too, iteration for iteration.  The undecided samples are
    re-scored inside the BF16 kernel (the default where it fits), or by a
    second kernel over each queued sample's candidate tiles (read from a
    sample-major copy) with a separate exact update of the queue, or by a
    second kernel that gathers every queued sample over all tiles."""
    if rescore == "candidate-tiles":
        monkeypatch.setenv("KMD_MIXED_CAND", "1")
    elif rescore == "gather":
        monkeypatch.setenv("KMD_MIXED_GATHER", "1")
    n, d, k = shape
    Xt, init = lo.gaussian_blobs(n, d, k, seed=seed, std=std)
    w = np.ones(n, np.float32)
    ref = run_driver_ex(kd, Xt, w, init, 1e-4, 25, "fp32")
    assert ref[4]["path"] == "fp32-mfma"
    rescored = {}
    for prec, path in (("bf16", "bf16+rescore"), ("bf16x3", "bf16x3+rescore")):
        got = run_driver_ex(kd, Xt, w, init, 1e-4, 25, prec)
        assert got[4]["path"] == path
        assert got[2] == ref[2]
        assert_array_equal(got[0], ref[0])
        assert_array_equal(got[1].view(np.uint32), ref[1].view(np.uint32))
        assert got[3] == ref[3]
        rescored[prec] = got[4]["n_rescored"]
    assert rescored["bf16x3"] <= rescored["bf16"]
    # the last update is within 1e-5 (relative to the centroid scale) of an
    # FP64 update from the labels of the centroids one iteration earlier, and
    # the returned labels are those of the returned centroids
    prev = run_driver_ex(kd, Xt, w, init, 1e-4, ref[2] - 1, "bf16")
    C64, cnt = fp64_update(Xt, w, prev[0], k)
    ok = cnt > 0
    assert np.abs(ref[1][:, ok] - C64[:, ok]).max() <= 1e-5 * np.abs(C64[:, ok]).max()
    check_label_mismatches(Xt, ref[1], ref[0], co.assign(Xt, ref[1]), limit=max(1, n // 10000))


def test_mixed_precision_falls_back_on_nonfinite(kd, torch_dev):
    Xt, init = lo.gaussian_blobs(4096, 8, 4, seed=2)
    Xt = Xt.copy()
    Xt[3, 100] = np.inf
    got = run_driver_ex(kd, Xt, np.ones(4096, np.float32), init, 0.0, 1, "bf16")
    assert got[4]["path"] == "fp32-mfma"


# ---- tiled path (centroid tiles streamed through LDS, separate exact update)

def same_fit(a, b):
    assert a[2] == b[2]
    assert_array_equal(a[0], b[0])
    assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert a[3] == b[3]


@pytest.mark.parametrize("shape", [(30_000, 24, 40), (20_001, 64, 100), (8_193, 7, 33),
                                   (30_000, 24, 8)])
def test_tiled_fp32_equals_lds_resident_fp32(kd, torch_dev, monkeypatch, shape):
    """Forced onto the tiled path, a shape the LDS-resident kernel serves
    gives the same labels, centroids (bit for bit), iteration count and
    inertia: the same FP32 MFMA arithmetic, and int64 sums either way."""
    n, d, k = shape
    Xt, init = lo.gaussian_blobs(n, d, k, seed=11, std=1.5)
    w = np.random.default_rng(5).uniform(0.5, 2.0, n).astype(np.float32)
    ref = run_driver_ex(kd, Xt, w, init, 1e-4, 20, "fp32")
    assert ref[4]["path"] == "fp32-mfma"
    monkeypatch.setenv("KMD_FORCE_TILED", "1")
    got = run_driver_ex(kd, Xt, w, init, 1e-4, 20, "fp32")
    assert got[4]["path"] == "tiled-fp32"
    same_fit(got, ref)


# ---- few centroids: the one-sample-per-lane 4x4x1 kernel

def assign_labels(kd, Xt, C, itype=torch.int32):
    hn = lo.half_l2_norm_squared(C)
    lab_d = torch.empty(Xt.shape[1], dtype=itype, device="cuda")
    _, ev = kd.assignment(to_dev(Xt), to_dev(C), to_dev(hn), lab_d, 8, 128, q())
    ev.wait()
    return to_np(lab_d).astype(np.int64)


@pytest.mark.parametrize("shape", [(100_003, 32, 8), (70_001, 7, 3), (50_000, 16, 16),
                                   (40_001, 33, 5), (30_000, 64, 13), (20_011, 1, 1),
                                   (9_999, 48, 9), (65, 2, 4), (3, 5, 2), (20_000, 32, 10),
                                   (12_345, 40, 12), (8_000, 64, 11), (60_000, 32, 32),
                                   (30_011, 24, 17), (20_000, 64, 28), (15_000, 9, 24),
                                   (40_000, 16, 20)])
def test_small_k_kernel_equals_tile_kernel(kd, torch_dev, monkeypatch, shape):
    """k <= 32 runs on the kernel that gives each lane one sample and all its
    centroids (v_mfma_f32_4x4x1f32); KMD_NO_SMALLK=1 runs the same fit on
    the 32-row tile kernel.  Both chains of FP32 MFMAs round like fmaf over
    ascending features from -hn, the argmin rule is the same and the sums
    are the same int64 fixed-point integers, so fits and assignments are
    equal bit for bit -- on blobs and on a grid full of exact ties between
    duplicated centroids (lowest index wins in both)."""
    n, d, k = shape
    Xt, init = lo.gaussian_blobs(n, d, k, seed=17, std=1.5)
    w = np.random.default_rng(6).uniform(0.5, 2.0, n).astype(np.float32)
    rng = np.random.default_rng(7)
    Xg = rng.integers(0, 3, size=(d, n)).astype(np.float32)
    Cg = rng.integers(0, 3, size=(d, k)).astype(np.float32)
    if k > 2:
        Cg[:, k - 1] = Cg[:, 1]
    got = run_driver_ex(kd, Xt, w, init, 1e-4, 15, "fp32")
    got_a = [assign_labels(kd, Xg, Cg, it) for it in (torch.int32, torch.int64)]
    monkeypatch.setenv("KMD_NO_SMALLK", "1")
    ref = run_driver_ex(kd, Xt, w, init, 1e-4, 15, "fp32")
    ref_a = [assign_labels(kd, Xg, Cg, it) for it in (torch.int32, torch.int64)]
    same_fit(got, ref)
    for g, r in zip(got_a, ref_a):
        assert_array_equal(g, r)
    assert_array_equal(got_a[0], co.assign(Xg, Cg))   # integer data: exact scores


def test_small_k_runs_fp32_for_every_precision(kd, torch_dev, monkeypatch):
    """With k <= 32 the step is bound by reading X: every precision runs the
    FP32 few-centroid kernel (path "fp32-mfma").  With KMD_NO_SMALLK=1 the
    BF16, BF16x3 and FP8 kernels run instead, and every one of these fits
    equals the FP32 fit bit for bit."""
    n, d, k = 60_000, 16, 8
    Xt, init = lo.gaussian_blobs(n, d, k, seed=23, std=1.5)
    w = np.random.default_rng(2).uniform(0.5, 2.0, n).astype(np.float32)
    ref = run_driver_ex(kd, Xt, w, init, 1e-4, 12, "fp32")
    assert ref[4]["path"] == "fp32-mfma"
    for prec in ("bf16", "bf16x3", "fp8"):
        got = run_driver_ex(kd, Xt, w, init, 1e-4, 12, prec)
        assert got[4]["path"] == "fp32-mfma"
        same_fit(got, ref)
    monkeypatch.setenv("KMD_NO_SMALLK", "1")
    for prec, path in (("bf16", "bf16+rescore"), ("bf16x3", "bf16x3+rescore"),
                       ("fp8", "tiled-fp8+rescore")):
        got = run_driver_ex(kd, Xt, w, init, 1e-4, 12, prec)
        assert got[4]["path"] == path
        same_fit(got, ref)


def test_small_k_labels_past_one_stage(kd, torch_dev, monkeypatch):
    """The few-centroid kernel keeps each wave's labels in LDS for up to 64
    of its super-blocks; n = 17,000,011 gives the waves 77 or more, so the
    stage is written out mid-run as well as at the end.  Fit and assignment
    equal the tile kernel's bit for bit."""
    n, d, k = 17_000_011, 4, 5
    rng = np.random.default_rng(12)
    centers = rng.uniform(-5, 5, size=(k, d))
    Xt = np.ascontiguousarray(
        (centers[rng.integers(0, k, n)] + rng.standard_normal((n, d))).T.astype(np.float32))
    init = np.ascontiguousarray(centers.T.astype(np.float32))
    w = np.ones(n, np.float32)
    got = run_driver_ex(kd, Xt, w, init, 0.0, 3, "fp32")
    got_a = assign_labels(kd, Xt, init, torch.int64)
    monkeypatch.setenv("KMD_NO_SMALLK", "1")
    ref = run_driver_ex(kd, Xt, w, init, 0.0, 3, "fp32")
    ref_a = assign_labels(kd, Xt, init, torch.int64)
    same_fit(got, ref)
    assert_array_equal(got_a, ref_a)


@pytest.mark.parametrize("k,nan_cols", [(8, [0, 1, 2, 3]),
                                        (20, [0, 1, 2, 3, 8, 9, 10, 11, 16, 17, 18, 19])])
@pytest.mark.parametrize("smallk", ["1", "0"])
def test_nan_centroids_are_passed_over(kd, torch_dev, monkeypatch, k, nan_cols, smallk):
    """Centroids with NaN coordinates score NaN for every sample; the
    reference's scan (strict <) never takes a NaN, so each sample goes to its
    best finite centroid.  On the 32-row tile kernel (KMD_NO_SMALLK=1) k = 20
    puts the NaN centroids on every row that one half of the tile holds (rows
    0-3, 8-11, 16-19), the case in which that half has no candidate at all;
    both k run on the few-centroid kernel too."""
    monkeypatch.setenv("KMD_NO_SMALLK", "0" if smallk == "1" else "1")
    rng = np.random.default_rng(3)
    d, n = 6, 4096
    Xt = rng.standard_normal((d, n)).astype(np.float32)
    C = (2 * rng.standard_normal((d, k))).astype(np.float32)
    C[:, nan_cols] = np.nan
    got = assign_labels(kd, Xt, C)
    ref = co.assign(Xt, C)
    finite = np.setdiff1d(np.arange(k), nan_cols)
    assert np.isin(ref, finite).all()
    margin = lo.pseudo_margin(Xt.astype(np.float64), C[:, finite].astype(np.float64))
    bad = got != ref
    assert (margin[bad] < 1e-4).all(), margin[bad]
    assert bad.sum() <= 2


def test_tiled_fp32_wide_shape_vs_oracle(kd, torch_dev):
    """d = 100, k = 300 (beyond the LDS-resident kernels): separated blobs
    with init near the centers, so no sample is within rounding of a tie;
    labels and iterations equal the FP64 oracle's, centroids within 1e-5."""
    Xt, init = separated_blobs(40_000, 100, 300, seed=31)
    w = np.ones(40_000, np.float32)
    got = run_driver_ex(kd, Xt, w, init, 0.0, 6, "fp32")
    assert got[4]["path"] == "tiled-fp32"
    l64, C64, it64, in64 = lo.kmeans_lloyd(Xt, w, init, 0.0, 6, dtype=np.float64)
    assert got[2] == it64
    assert_array_equal(got[0], l64)
    assert np.abs(got[1] - C64).max() <= 1e-5 * np.abs(C64).max()
    assert_allclose(got[3], in64, rtol=1e-5)


@pytest.mark.parametrize("shape,seed,std", [((60_000, 100, 300), 41, 2.0),
                                            ((50_001, 128, 1024), 43, 1.0),
                                            ((30_000, 48, 2000), 45, 1.5),
                                            ((40_000, 96, 512), 47, 1.0),
                                            ((25_001, 70, 200), 49, 2.0)])
def test_tiled_mixed_matches_tiled_fp32_bit_exactly(kd, torch_dev, shape, seed, std):
    """BF16, BF16x3 and FP8 (e4m3, per-tile scales) on the tiled path, with
    the certified margin and the FP32 re-score, give the tiled FP32 fit bit
    for bit."""
    n, d, k = shape
    Xt, init = lo.gaussian_blobs(n, d, k, seed=seed, std=std)
    w = np.ones(n, np.float32)
    ref = run_driver_ex(kd, Xt, w, init, 1e-4, 8, "fp32")
    assert ref[4]["path"] == "tiled-fp32"
    for prec in ("bf16", "bf16x3", "fp8"):
        got = run_driver_ex(kd, Xt, w, init, 1e-4, 8, prec)
        assert got[4]["path"] == f"tiled-{prec}+rescore"
        same_fit(got, ref)


def test_assignment_wide_shape_matches_c_restatement(kd, torch_dev):
    """kmd_assignment at d = 128, k = 1024 (tiled FP32 kernel) against the C
    restatement: labels equal except where the oracle's margin is below the
    FP32 rounding bound."""
    rng = np.random.default_rng(8)
    Xt = rng.standard_normal((128, 5000)).astype(np.float32)
    C = rng.standard_normal((128, 1024)).astype(np.float32)
    hn = lo.half_l2_norm_squared(C)
    lab_d = torch.empty(5000, dtype=torch.int32, device="cuda")
    _, ev = kd.assignment(to_dev(Xt), to_dev(C), to_dev(hn), lab_d, 8, 128, q())
    ev.wait()
    got = to_np(lab_d)
    ref = co.assign(Xt, C, hn)
    margin = lo.pseudo_margin(Xt.astype(np.float64), C.astype(np.float64))
    bad = got != ref
    assert (margin[bad] < 1e-3).all(), margin[bad]
    assert bad.mean() < 1e-3


# ---- launch plans and the autotuner

def test_launch_plans_do_not_change_results(kd, torch_dev):
    """Every plan the autotuner can pick -- LDS-resident or tiled centroids,
    each launch shape of the LDS-resident step (waves per block x column
    blocks per wave), any cap on persistent blocks per CU -- gives the same
    fit bit for bit, on an even and an odd sample count (the two-column
    shapes load pairs of samples when n is even)."""
    for n, d in ((50_000, 32), (30_001, 14)):
        Xt, init = lo.gaussian_blobs(n, d, 64, seed=61, std=1.5)
        w = np.random.default_rng(2).uniform(0.5, 2.0, n).astype(np.float32)
        try:
            for prec in ("fp32", "bf16x3"):
                runs = []
                plans = ((1, 0, 0), (1, 1, 0), (1, 0, 1), (1, 0, 2), (1, 2, 2),
                         (2, 0, 0), (2, 2, 0))
                for path, cap, shape in plans:
                    kd.set_tuning(path, cap, shape)
                    runs.append(run_driver_ex(kd, Xt, w, init, 1e-4, 12, prec))
                assert runs[0][4]["path"].startswith(("fp32", "bf16"))
                assert runs[5][4]["path"].startswith("tiled")
                for r in runs[1:]:
                    same_fit(r, runs[0])
        finally:
            kd.set_tuning(-1)


def test_autotune_records_a_plan_the_driver_uses(kd, torch_dev):
    Xt, init = lo.gaussian_blobs(40_000, 20, 50, seed=62, std=1.0)
    w = np.ones(40_000, np.float32)
    r = kd.autotune(to_dev(Xt), to_dev(w), to_dev(init), "fp32", 2, q())
    assert r["path"] in ("lds-resident", "tiled")
    # LDS-resident: 3 launch shapes x 4 block caps; tiled: 4 block caps
    assert len(r["candidates"]) == 16 and all(c[3] > 0 for c in r["candidates"])
    assert {c[2] for c in r["candidates"] if c[0] == "lds-resident"} == {0, 1, 2}
    assert r["ms"] == min(c[3] for c in r["candidates"])
    got = run_driver_ex(kd, Xt, w, init, 1e-4, 5, "fp32")
    assert got[4]["path"] == ("fp32-mfma" if r["path"] == "lds-resident" else "tiled-fp32")


def _wide_range_data(seed=21):
    """Four ordinary blobs, a cluster of tiny values U(0, 1e-3) and one outlier
    sample at 1e6 in feature 0 (its own cluster); init = one member of each
    group, so every fit keeps the groups apart."""
    rng = np.random.default_rng(seed)
    d = 4
    centers = np.array([[10, 0, 0, 0], [0, 10, 0, 0], [0, 0, 10, 0], [0, 0, 0, 10]],
                       np.float64)
    big = centers[rng.integers(0, 4, 8000)] + rng.normal(size=(8000, d))
    tiny = rng.uniform(0, 1e-3, size=(2000, d))
    X = np.concatenate([big, tiny, [[1e6, 0.5, 0.5, 0.5]]]).astype(np.float32)
    perm = rng.permutation(len(X))
    X = X[perm]
    groups = np.concatenate([np.argmax(big, axis=1), np.full(2000, 4), [5]])[perm]
    init_idx = [int(np.flatnonzero(groups == g)[0]) for g in range(6)]
    return np.ascontiguousarray(X.T), np.ascontiguousarray(X[init_idx].T), groups


def test_wide_range_data_keeps_small_clusters_accurate(kd, torch_dev):
    """One outlier of 1e6 in a feature would make fixed-point steps of 2^-10
    there; the pre-pass sees the range and the sums go through floating point,
    so the tiny cluster's centre is accurate against its own magnitude."""
    Xt, init, groups = _wide_range_data()
    w = np.ones(Xt.shape[1], np.float32)
    for name in ("fp32", "bf16"):
        dev = torch.device("cuda", 0)
        lab_d = torch.empty(Xt.shape[1], dtype=torch.int32, device=dev)
        init_d, res_d = to_dev(init), torch.empty(init.shape, device=dev)
        n_iter, inertia, _ = kd.kmeans_lloyd_driver_ex(to_dev(Xt), to_dev(w), init_d,
                                                       lab_d, res_d, 0.0, False, 6, name,
                                                       q())
        lab, C = to_np(lab_d), to_np(res_d)
        assert_array_equal(lab, groups)
        for c in range(6):
            mean = Xt[:, lab == c].astype(np.float64).mean(axis=1)
            assert np.abs(C[:, c] - mean).max() <= 1e-5 * np.abs(mean).max(), (name, c)


def test_tiny_weight_cluster_is_not_empty(kd, torch_dev):
    """A cluster whose only member has weight 1e-10 next to unit weights: its
    count must not round to zero (which would relocate it as empty)."""
    Xt, init, groups = _wide_range_data(seed=22)
    Xt = Xt.copy()
    Xt[0, groups == 5] = 60.0            # no value outlier: only the weight is extreme
    init[0, 5] = 60.0
    w = np.ones(Xt.shape[1], np.float32)
    w[groups == 5] = 1e-10
    lab, C, n_iter, inertia = run_driver(kd, Xt, w, init, 0.0, 5)
    assert_array_equal(lab, groups)
    assert_allclose(C[:, 5], Xt[:, groups == 5][:, 0], rtol=1e-6)
    for c in range(5):
        mean = Xt[:, lab == c].astype(np.float64).mean(axis=1)
        assert np.abs(C[:, c] - mean).max() <= 1e-5 * np.abs(mean).max()


def test_ordinary_data_keeps_exact_sums(kd, torch_dev):
    """C1-like blobs (centres in [-20, 20], no outliers) stay on the exact
    fixed-point sums: two runs are bitwise identical and equal the mixed
    path's (which uses the same integers)."""
    Xt, init = lo.gaussian_blobs(60_000, 32, 64, seed=31, center_box=20.0,
                                 std_range=(0.5, 2.0))
    w = np.ones(Xt.shape[1], np.float32)
    a = run_driver(kd, Xt, w, init, 0.0, 4)
    b = run_driver(kd, Xt, w, init, 0.0, 4)
    assert_array_equal(a[0], b[0])
    assert a[1].tobytes() == b[1].tobytes()


def test_fp8_conversion_rounds_to_nearest_even(kd, torch_dev):
    """The e4m3 conversion the FP8 kernels use rounds to nearest (ties to
    even) with 3 mantissa bits: |e4m3(v) - v| <= 2^-4 |v| for normal values
    and <= 2^-10 below them -- the error bound's assumptions."""
    rng = np.random.default_rng(17)
    v = np.concatenate([
        np.exp(rng.uniform(np.log(2.0 ** -14), np.log(224.0), 200_000)) *
        rng.choice([-1.0, 1.0], 200_000),
        [0.0, 1.0, 1.0625, 1.1875, 1.0625 + 2.0 ** -10, 0.5, 224.0, -224.0, 2.0 ** -6,
         2.0 ** -9, 3 * 2.0 ** -10]]).astype(np.float32)
    out = torch.empty(v.shape[0], dtype=torch.float32, device="cuda")
    kd.fp8_roundtrip(to_dev(v), out, q())
    r = to_np(out).astype(np.float64)
    x = v.astype(np.float64)
    err = np.abs(r - x)
    normal = np.abs(x) >= 2.0 ** -6
    assert np.all(err[normal] <= 2.0 ** -4 * np.abs(x[normal]))
    assert np.all(err[~normal] <= 2.0 ** -10)
    # ties go to the even mantissa; exact values stay exact
    tail = r[-11:]
    assert tail[0] == 0.0 and tail[1] == 1.0
    assert tail[2] == 1.0           # 1 + 2^-4: halfway between 1 and 1.125
    assert tail[3] == 1.25          # 1 + 3 2^-4: halfway between 1.125 and 1.25
    assert tail[4] == 1.125         # just above the tie
    assert tail[5] == 0.5 and tail[6] == 224.0 and tail[7] == -224.0
    assert tail[8] == 2.0 ** -6 and tail[9] == 2.0 ** -9
